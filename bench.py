#!/usr/bin/env python3
"""Benchmark of the MI355X even-odd Wilson-clover path.

Headline (``value``): EO Dhop (Meooe, odd -> even) bandwidth in GB/s on a
32^3 x 64 lattice, FP64 fields and 18-real links, with links
U = exp(i 0.3 H) (H seeded random traceless Hermitian) and complex-Gaussian
spinors; bytes are the compulsory traffic per output site 8*L + 2*S
(1536 B in FP64), so ``roofline.frac`` is that over the 8 TB/s HBM peak.
One step = one Meooe call; 4 (in, out) pairs rotate so the 256 MB last-level
cache cannot hold the spinors between calls.  Extra fields report the FP32
and 12-real-link variants, a copy-kernel bandwidth, the CG solve rate, and
the CPU baseline (the C/OpenMP restatement of the same hop on the host).

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--lattice 32,32,32,64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lattice", default="32,32,32,64")
    ap.add_argument("--pairs", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-cg", action="store_true", help="skip the CG solve measurement")
    ap.add_argument("--only", default=None, help="run one variant: fp64_18, fp64_12, fp32_18, fp32_12")
    return ap.parse_args()


def su3_generators(torch, dev):
    gens = []
    for i in range(3):
        for j in range(i + 1, 3):
            a = torch.zeros((3, 3), dtype=torch.complex128)
            a[i, j] = a[j, i] = 1.0
            gens.append(a)
            a = torch.zeros((3, 3), dtype=torch.complex128)
            a[i, j] = -1j
            a[j, i] = 1j
            gens.append(a)
            if j == i + 1:
                a = torch.zeros((3, 3), dtype=torch.complex128)
                for k in range(j):
                    a[k, k] = 1.0
                a[j, j] = -j
                gens.append(a)
    return torch.stack([g / np.sqrt(2.0 * torch.trace(g @ g).real.item()) for g in gens]).to(dev)


def smooth_links_device(torch, fd, eps, seed, dev):
    """U_mu(x) = exp(i eps H), H = sum_a h_a T_a, h_a ~ N(0,1): synthetic input made on the GPU."""
    import gpt_amd as g

    V = int(np.prod(fd))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    T = su3_generators(torch, dev)
    U = []
    chunk = 1 << 20
    for mu in range(4):
        out = torch.empty((V, 3, 3), dtype=torch.complex128, device=dev)
        for s in range(0, V, chunk):
            n = min(chunk, V - s)
            h = torch.randn((n, 8), generator=gen, dtype=torch.float64, device=dev)
            H = torch.einsum("na,aij->nij", h.to(torch.complex128), T)
            out[s:s + n] = torch.linalg.matrix_exp(1j * eps * H)
        U.append(g.qcd.gauge.link_field(g.grid(fd, g.double), out))
    return U


def gaussian_half(torch, g, grid_eo, cb, seed, dev):
    x = g.lattice(grid_eo)
    x.checkerboard(cb)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    x.data.copy_(torch.randn(x.data.shape, generator=gen, dtype=torch.float64, device=dev)
                 .to(x.data.dtype) / np.sqrt(2.0))
    return x


def time_meooe(torch, g, w, pairs, steps, warmup):
    stream = torch.cuda.current_stream()
    for i in range(warmup):
        s, d = pairs[i % len(pairs)]
        w.Meooe.mat(d, s)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    t0 = time.perf_counter()
    for i in range(steps):
        s, d = pairs[i % len(pairs)]
        w.Meooe.mat(d, s)
    e1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return e0.elapsed_time(e1) / steps * 1e-3, (t1 - t0) / steps


def copy_bandwidth(torch, dev, nbytes=2 << 30, reps=20):
    a = torch.empty(nbytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / reps * 1e-3
    del a, b
    return 2 * nbytes / t / 1e9


def cpu_baseline(budget_s=12.0):
    """The oracle's C/OpenMP EO hop on the host cores, 16^3 x 32, FP64."""
    from oracle import cpu as C
    from oracle import gauge as G
    from oracle import wilson_clover as W

    fd = [16, 16, 16, 32]
    U = G.smooth_links(fd, 0.3, 2001)
    w = W.wilson_clover(U, mass=0.08, csw_r=0.0, csw_t=0.0, boundary_phases=[1, 1, 1, -1])
    wc = C.wilson_clover_cpu(w)
    rng = np.random.default_rng(2002)
    src = (rng.standard_normal((wc.Vh, 4, 3)) + 1j * rng.standard_normal((wc.Vh, 4, 3))) / np.sqrt(2)
    wc.meooe(src, 1)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        wc.meooe(src, 1)
        n += 1
    dt = (time.perf_counter() - t0) / n
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(1536.0 * wc.Vh / dt / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port",
            "sample": f"EO Dhop (Meooe odd->even), 16^3x32 FP64 18-real links, {n} calls in "
                      f"{budget_s:.0f} s, C/OpenMP restatement (oracle/wc_cpu.c) standing in for "
                      f"Grid's CPU path; {dt * 1e3:.2f} ms per call"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; use --gpus 1"}))
        sys.exit(2)
    import torch

    import gpt_amd as g

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    fd = [int(x) for x in args.lattice.split(",")]
    Vh = int(np.prod(fd)) // 2
    log(f"bench: lattice {fd}, Vh = {Vh}, steps {args.steps}, warmup {args.warmup}")
    t = time.perf_counter()
    U = smooth_links_device(torch, fd, 0.3, 2001, dev)
    torch.cuda.synchronize()
    log(f"bench: links generated in {time.perf_counter() - t:.1f} s")

    variants = [("fp64_18", g.double, False), ("fp64_12", g.double, True),
                ("fp32_18", g.single, False), ("fp32_12", g.single, True)]
    if args.only:
        variants = [v for v in variants if v[0] == args.only]
    res = {}
    for name, prec, comp in variants:
        w = g.qcd.fermion.wilson_clover(U, mass=0.08, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0,
                                        boundary_phases=[1, 1, 1, -1], precision=prec,
                                        compress_links=comp)
        pairs = []
        for k in range(args.pairs):
            s = gaussian_half(torch, g, w.F_grid_eo, g.odd, 2002 + k, dev)
            d = g.lattice(w.F_grid_eo)
            d.checkerboard(g.even)
            pairs.append((s, d))
        t_ev, t_wall = time_meooe(torch, g, w, pairs, args.steps, args.warmup)
        real_b = prec.nbytes
        nl = 12 if comp else 18
        bytes_site = (8 * nl + 2 * 24) * real_b
        res[name] = {"us_per_call": round(t_ev * 1e6, 2), "wall_us_per_call": round(t_wall * 1e6, 2),
                     "GBps": round(bytes_site * Vh / t_ev / 1e9, 2),
                     "GFLOPps": round(1320.0 * Vh / t_ev / 1e9, 2), "bytes_per_site": bytes_site,
                     "reference_effective_GBps": round(2880.0 / 8 * real_b * Vh / t_ev / 1e9, 2)}
        log(f"bench: {name}: {res[name]}")
        del w, pairs
        torch.cuda.empty_cache()

    copy_gbs = copy_bandwidth(torch, dev)
    log(f"bench: copy kernel {copy_gbs:.0f} GB/s")

    cg_res = None
    if not args.no_cg:
        cg_res = cg_measure(torch, g, dev)
        log(f"bench: cg {cg_res}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()
        log(f"bench: cpu baseline {cpu}")

    head = res.get("fp64_18") or next(iter(res.values()))
    achieved = head["GBps"]
    out = {
        "metric": "EO Dhop GB/s (% HBM peak); mixed-prec CG solves/s + p99 latency at 1e-10",
        "value": achieved,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["us_per_call"] / 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: U = exp(i 0.3 H) seeded, complex-Gaussian spinors",
        "config": {"workload": f"EO Dhop (Meooe odd->even) {'x'.join(map(str, fd))}, FP64, 18-real links, "
                               f"{args.pairs} rotating in/out pairs",
                   "lattice": fd, "half_sites": Vh},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "variants": res,
        "copy_kernel_GBps": round(copy_gbs, 1),
        "cg": cg_res,
    }
    print(json.dumps(out), flush=True)


def cg_measure(torch, g, dev, fd=(16, 16, 16, 32)):
    """FP64 eo2_ne + CG to 1e-10 on one Gaussian source (solves/s)."""
    fd = list(fd)
    U = smooth_links_device(torch, fd, 0.3, 4001, dev)
    w = g.qcd.fermion.wilson_clover(U, mass=0.08, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])
    inv, pc = g.algorithms.inverter, g.qcd.fermion.preconditioner
    cg = inv.cg(eps=1e-10, maxiter=5000)
    slv = w.propagator(inv.preconditioned(pc.eo2_ne(), cg))
    src = g.lattice(w.F_grid)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4100)
    src.data.copy_(torch.randn(src.data.shape, generator=gen, dtype=torch.float64, device=dev))
    g(slv * src)
    torch.cuda.synchronize()
    lat = []
    its = []
    for _ in range(3):
        t0 = time.perf_counter()
        g(slv * src)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        its.append(len(cg.history))
    return {"lattice": fd, "precision": "fp64", "tol": 1e-10, "iterations": its[-1],
            "solves_per_s": round(1.0 / np.median(lat), 3), "ms_per_solve": round(np.median(lat) * 1e3, 2),
            "ms_per_iteration": round(np.median(lat) * 1e3 / its[-1], 4)}


if __name__ == "__main__":
    main()
