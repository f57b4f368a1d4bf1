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


def profiled_traffic(shape):
    """HBM bytes per launch of the headline hop from the committed rocprofv3
    summary (FETCH_SIZE and WRITE_SIZE passes, copy-calibrated; lines
    ``traffic_bytes_per_launch[rb,tpb,occ,seg,ls]: bytes``), when the launch
    shape in use is a profiled one; else None"""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hop_fp64_18_32x32x32x64.md")
    key = "traffic_bytes_per_launch[" + ",".join(str(int(x)) for x in shape) + "]: "
    try:
        for line in open(path).read().splitlines():
            if line.startswith(key):
                return int(line[len(key):])
    except (OSError, ValueError):
        pass
    return None


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
    ap.add_argument("--only", default=None,
                    help="comma-separated variants to run: fp64_18, fp64_12, fp32_18, fp32_12, "
                         "fp64_18_rhs12, fp32_18_rhs12, fp16_rhs (the FP16-storage multi-RHS hop on the "
                         "matrix cores and on the vector ALUs)")
    ap.add_argument("--cg-lattice", default="24,24,24,48")
    ap.add_argument("--no-dhop", action="store_true", help="skip the Dhop measurement")
    ap.add_argument("--nrhs", type=int, default=12)
    ap.add_argument("--serve", type=int, default=12,
                    help="serving stream: requests of the capacity burst (0: skip serving)")
    ap.add_argument("--serve-poisson", type=int, default=200, help="serving stream: Poisson-arrival requests")
    ap.add_argument("--multi-mass", default="all",
                    help="multi-mass batch cases after the serving stream: all, none, or one of "
                         "one, other, three, twelve (that case alone, without the warm-up; also with --serve 0)")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="Poisson arrival rate as a fraction of the measured capacity")
    ap.add_argument("--serve-streams", type=int, default=4, help="HIP streams (worker threads) of the server")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="Poisson arrival rate in requests/s (default: --serve-load times the capacity)")
    ap.add_argument("--serve-min-slots", type=int, default=1, help="smallest batch a serving group opens with")
    ap.add_argument("--no-c1", action="store_true", help="skip the 8^4 point-source propagator")
    ap.add_argument("--c4", default="16x16x16x32:12,16x16x16x32:48,32x32x32x64:12,32x32x32x64:48",
                    help="multi-RHS cases LXxLYxLZxLT:NRHS, comma-separated ('' to skip)")
    ap.add_argument("--slab-overlap", action="store_true",
                    help="also time the slab share's hop with its face exchange on a second stream "
                         "(a device copy standing in for the transfer), graph-captured")
    ap.add_argument("--shape", default=None,
                    help="hop launch shape rb,tpb,occ,seg,ls for the Dhop variants instead of tuning")
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


def haar_links_device(torch, fd, seed, dev):
    """Haar-random SU(3) links made on the GPU: QR of a complex Gaussian
    matrix with the phases of R's diagonal moved into Q, divided by a cube
    root of its determinant."""
    import gpt_amd as g

    V = int(np.prod(fd))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    U = []
    for mu in range(4):
        z = torch.complex(torch.randn((V, 3, 3), generator=gen, dtype=torch.float64, device=dev),
                          torch.randn((V, 3, 3), generator=gen, dtype=torch.float64, device=dev))
        q, r = torch.linalg.qr(z)
        d = torch.diagonal(r, dim1=-2, dim2=-1)
        q = q * (d / d.abs()).unsqueeze(-2)
        q = q / torch.pow(torch.linalg.det(q), 1.0 / 3.0)[:, None, None]
        U.append(g.qcd.gauge.link_field(g.grid(fd, g.double), q.contiguous()))
    return U


def gaussian_half(torch, g, grid_eo, cb, seed, dev):
    x = g.lattice(grid_eo)
    x.checkerboard(cb)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    x.data.copy_(torch.randn(x.data.shape, generator=gen, dtype=torch.float64, device=dev)
                 .to(x.data.dtype) / np.sqrt(2.0))
    return x


def call_stats(times):
    """median, mean, min and max (seconds) of per-call times, and whether
    the mean is more than 5 % away from the median (a stall inside the timed
    calls: the median is the call's time, the mean is not)"""
    t = np.asarray(times, dtype=np.float64)
    med = float(np.median(t))
    return {"median_s": med, "mean_s": float(t.mean()), "min_s": float(t.min()), "max_s": float(t.max()),
            "mean_off_median": bool(abs(t.mean() - med) > 0.05 * med)}


def time_meooe(torch, g, w, pairs, steps, warmup):
    """Meooe calls on rotating (src, dst) pairs: `warmup` untimed, then
    `steps` calls each between its own pair of HIP events on the launch
    stream (consecutive calls share an event, so the events add nothing
    between calls).  Returns call_stats of the per-call times and the wall
    time per call of the whole timed loop."""
    stream = torch.cuda.current_stream()
    for i in range(warmup):
        s, d = pairs[i % len(pairs)]
        w.Meooe.mat(d, s)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(steps):
        s, d = pairs[i % len(pairs)]
        w.Meooe.mat(d, s)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    per_call = [ev[i].elapsed_time(ev[i + 1]) * 1e-3 for i in range(steps)]
    return call_stats(per_call), (t1 - t0) / steps


def hop16_measure(torch, g, dev, U, fd, nrhs=12, steps=20, warmup=5):
    """The FP16-storage Meooe of nrhs right-hand sides, as the 16-bit inner
    solve runs it (FP32 operator with 12-real links, FP16 spinors in and
    out), on the matrix cores (k_hop_mma) and on the vector ALUs (k_hop):
    per-call HIP events, median.  Compulsory bytes per output site: 8 links
    of 12 FP32 reals and nrhs spinors of 24 FP16 reals in and out."""
    Vh = int(np.prod(fd)) // 2
    w = g.qcd.fermion.wilson_clover(U, mass=0.05, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0,
                                    boundary_phases=[1, 1, 1, -1], precision=g.single, compress_links=True,
                                    n_rhs=nrhs)
    w.tune_hop(nrhs, reps=5)  # the FP32 shape the vector form of the 16-bit hop uses
    pairs = []
    for k in range(2):
        s = gaussian_half(torch, g, w.F_grid_eo, g.odd, 2102 + k, dev)
        pairs.append((s.data.to(torch.float16), torch.empty(s.data.shape, dtype=torch.float16, device=dev)))
    bytes_site = 8 * 12 * 4 + nrhs * 2 * 24 * 2
    out = {"nrhs": nrhs, "bytes_per_site": bytes_site, "lattice": fd}
    stream = torch.cuda.current_stream()
    for form, mma in (("mma", True), ("vector", False)):
        w.set_hop_mma(mma)
        for i in range(warmup):
            s16, d16 = pairs[i % 2]
            w.hop_s16(s16, 1, dst=d16)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        ev[0].record(stream)
        for i in range(steps):
            s16, d16 = pairs[i % 2]
            w.hop_s16(s16, 1, dst=d16)
            ev[i + 1].record(stream)
        torch.cuda.synchronize()
        st = call_stats([ev[i].elapsed_time(ev[i + 1]) * 1e-3 for i in range(steps)])
        t = st["median_s"]
        out[form] = {"us_per_call": round(t * 1e6, 2), "us_per_call_min": round(st["min_s"] * 1e6, 2),
                     "GBps": round(bytes_site * Vh / t / 1e9, 2),
                     "frac_hbm_peak": round(bytes_site * Vh / t / 1e9 / HBM_PEAK_GBS, 4),
                     "GFLOPps": round(1320.0 * nrhs * Vh / t / 1e9, 2)}
    w.set_hop_mma("auto")
    out["mma_speedup"] = round(out["vector"]["us_per_call"] / out["mma"]["us_per_call"], 3)
    w.release()
    return out


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


def probe_bandwidth(torch, dev, nbytes=2 << 30, reps=20):
    """The library's streaming probe (wc_bandwidth_probe): GB/s of a plain
    16-byte-per-lane copy (bytes read + written) and of a read-only stream,
    the practical HBM ceiling next to the 8 TB/s peak"""
    from gpt_amd import _lib
    from gpt_amd import core

    lib = _lib.load()
    a = torch.empty(nbytes, dtype=torch.uint8, device=dev).fill_(1)
    b = torch.empty_like(a)
    out = {}
    for kind, name, moved in [(0, "copy_GBps", 2 * nbytes), (1, "read_GBps", nbytes)]:
        for _ in range(3):
            _lib.check(lib.wc_bandwidth_probe(kind, a.data_ptr(), b.data_ptr(), nbytes, core.stream_ptr()))
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        stream = torch.cuda.current_stream()
        e0.record(stream)
        for _ in range(reps):
            _lib.check(lib.wc_bandwidth_probe(kind, a.data_ptr(), b.data_ptr(), nbytes, core.stream_ptr()))
        e1.record(stream)
        torch.cuda.synchronize()
        out[name] = round(moved / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9, 1)
    del a, b
    torch.cuda.empty_cache()
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(U_dev, budget_s=15.0):
    """The oracle's C/OpenMP EO hop on the host cores, on the headline's own
    lattice and links (copied from the GPU), FP64, 18-real links: the same
    workload as the headline, timed for about budget_s seconds."""
    from gpt_amd.qcd import gauge as GA
    from oracle import cpu as C
    from oracle import wilson_clover as W

    U = GA.to_numpy(U_dev)
    fd = list(reversed(U[0].shape[:4]))
    w = W.wilson_clover(U, mass=0.08, csw_r=0.0, csw_t=0.0, boundary_phases=[1, 1, 1, -1])
    wc = C.wilson_clover_cpu(w)
    del w, U
    rng = np.random.default_rng(2002)
    src = (rng.standard_normal((wc.Vh, 4, 3)) + 1j * rng.standard_normal((wc.Vh, 4, 3))) / np.sqrt(2)
    wc.meooe(src, 1)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        wc.meooe(src, 1)
        n += 1
    dt = (time.perf_counter() - t0) / n
    threads = C.threads()
    out = {"value": round(1536.0 * wc.Vh / dt / 1e9, 3), "unit": "GB/s", "cores": threads,
           "kind": "port", "cpu_model": cpu_model(), "omp_threads": threads,
           "host_cpus": os.cpu_count(),
           "sample": f"EO Dhop (Meooe odd->even), {'x'.join(map(str, fd))} FP64 18-real links (the headline's "
                     f"lattice and links), {n} calls in {budget_s:.0f} s on {threads} OpenMP threads, "
                     f"C/OpenMP restatement (oracle/wc_cpu.c) standing in for Grid's CPU path; "
                     f"{dt * 1e3:.2f} ms per call"}
    del wc, src
    return out


def slab_share(torch, g, dev, fd, args):
    """The FP64 18-real hop on one GPU's share of a T-slab decomposition over 8
    GPUs (Lx x Ly x Lz x Lt/8, periodic in t standing in for the halo), and the
    bytes each hop would exchange: per t-face one parity of projected half
    spinors, 6 complex per face site"""
    sd = [fd[0], fd[1], fd[2], fd[3] // 8]
    Vh = int(np.prod(sd)) // 2
    Us = smooth_links_device(torch, sd, 0.3, 2001, dev)
    w = g.qcd.fermion.wilson_clover(Us, mass=0.08, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0,
                                    boundary_phases=[1, 1, 1, -1])
    pairs = []
    for k in range(args.pairs):
        src = gaussian_half(torch, g, w.F_grid_eo, g.odd, 2002 + k, dev)
        d = g.lattice(w.F_grid_eo)
        d.checkerboard(g.even)
        pairs.append((src, d))
    shape, _ = w.tune_hop(1, reps=3)
    st, _ = time_meooe(torch, g, w, pairs, args.steps, args.warmup)
    t_ev = st["median_s"]
    face_sites = fd[0] * fd[1] * fd[2] // 2
    out = {"local_lattice": sd, "us_per_call": round(t_ev * 1e6, 2), "GBps": round(1536 * Vh / t_ev / 1e9, 2),
           "halo_bytes_per_hop": 2 * face_sites * 6 * 16,
           "halo_us_at_153GBps_per_link": round(face_sites * 6 * 16 / 153e9 * 1e6, 2),
           "shape": list(shape)}
    # the halo kernels of the exchange (wc_halo_pack / wc_halo_unpack); bytes
    # per face site: pack reads 2 spinors and 1 link and writes 2 half
    # spinors, unpack reads 2 half spinors and 1 link and updates 2 spinors
    # a call takes a few microseconds, less than its host-side launch, so
    # `reps` calls are captured in one graph and the replay is timed
    src, dst = pairs[0]
    lo, hi = w.halo_pack(src)
    reps = 100
    for name, fn, nbytes in [("halo_pack", lambda: w.halo_pack(src, out=(lo, hi)),
                              (2 * 192 + 144 + 2 * 96) * face_sites),
                             ("halo_unpack", lambda: w.halo_unpack(lo, hi, dst),
                              (2 * 96 + 144 + 4 * 192) * face_sites),
                             ("interior_hop", lambda: w.halo_interior(src, dst=dst), 1536 * Vh),
                             ("periodic_hop_graph", lambda: w.Meooe.mat(dst, src), 1536 * Vh)]:
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for _ in range(reps):
                fn()
        graph.replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        stream = torch.cuda.current_stream()
        e0.record(stream)
        for _ in range(5):
            graph.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / (5 * reps) * 1e-3
        out[name] = {"us_per_call": round(t * 1e6, 2), "bytes_per_call": nbytes, "GBps": round(nbytes / t / 1e9, 1)}
        del graph
    # the rank's whole hop without the transfer: interior, then pack and unpack
    out["interior_pack_unpack_us"] = round(sum(out[k]["us_per_call"] for k in
                                               ("interior_hop", "halo_pack", "halo_unpack")), 2)
    if args.slab_overlap:
        # pack, the faces copied on a second stream (standing in for the
        # transfer) while the interior runs, unpack: one rank's schedule
        comm = torch.cuda.Stream()
        lo_in, hi_in = torch.empty_like(lo), torch.empty_like(hi)

        def exchange(a, b):
            lo_in.copy_(a)
            hi_in.copy_(b)
            return lo_in, hi_in

        step = lambda: w.halo_hop(src, dst, exchange, comm_stream=comm)  # noqa: E731
        for _ in range(3):
            step()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for _ in range(reps):
                step()
        graph.replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        stream = torch.cuda.current_stream()
        e0.record(stream)
        for _ in range(5):
            graph.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        out["overlapped_halo_hop_us"] = round(e0.elapsed_time(e1) / (5 * reps) * 1e3, 2)
        del graph
    del w, pairs, Us, lo, hi
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; use --gpus 1"}))
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

    variants = [("fp64_18", g.double, False, 1), ("fp64_12", g.double, True, 1),
                ("fp32_18", g.single, False, 1), ("fp32_12", g.single, True, 1),
                ("fp64_18_rhs12", g.double, False, 12), ("fp32_18_rhs12", g.single, False, 12)]
    if args.only:
        keep = args.only.split(",")
        variants = [v for v in variants if v[0] in keep]
    if args.no_dhop:
        variants = []
    res = {}
    for name, prec, comp, nrhs in variants:
        w = g.qcd.fermion.wilson_clover(U, mass=0.08, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0,
                                        boundary_phases=[1, 1, 1, -1], precision=prec,
                                        compress_links=comp, n_rhs=nrhs)
        pairs = []
        npairs = args.pairs if nrhs == 1 else 2
        for k in range(npairs):
            s = gaussian_half(torch, g, w.F_grid_eo, g.odd, 2002 + k, dev)
            d = g.lattice(w.F_grid_eo)
            d.checkerboard(g.even)
            pairs.append((s, d))
        if args.shape:
            w.set_hop_shape(tuple(int(x) for x in args.shape.split(",")))
            shape = w.hop_shape()
        else:
            # 10 calls per shape: the best shapes of the single-RHS hops differ by ~1-4 %
            shape, _ = w.tune_hop(nrhs, reps=10)
        st, t_wall = time_meooe(torch, g, w, pairs, args.steps, args.warmup)
        t_ev = st["median_s"]
        real_b = prec.nbytes
        nl = 12 if comp else 18
        # compulsory bytes per output site: links once, nrhs spinors in and out
        bytes_site = (8 * nl + nrhs * 2 * 24) * real_b
        res[name] = {"us_per_call": round(t_ev * 1e6, 2), "wall_us_per_call": round(t_wall * 1e6, 2),
                     "us_per_call_mean": round(st["mean_s"] * 1e6, 2),
                     "us_per_call_min": round(st["min_s"] * 1e6, 2),
                     "us_per_call_max": round(st["max_s"] * 1e6, 2),
                     "mean_off_median": st["mean_off_median"],
                     "GBps": round(bytes_site * Vh / t_ev / 1e9, 2), "nrhs": nrhs,
                     "GFLOPps": round(1320.0 * nrhs * Vh / t_ev / 1e9, 2), "bytes_per_site": bytes_site,
                     "shape": {"rhs_per_thread": shape[0], "threads_per_block": 64 * shape[1],
                               "max_blocks_per_cu": shape[2], "t_runs": shape[3], "lds_staging": shape[4]},
                     "reference_effective_GBps": round(2880.0 / 8 * real_b * nrhs * Vh / t_ev / 1e9, 2)}
        log(f"bench: {name}: {res[name]}")
        del w, pairs
        torch.cuda.empty_cache()
    if "fp64_18" in res and fd[3] % 8 == 0:
        res["fp64_18_slab_share"] = slab_share(torch, g, dev, fd, args)
        log(f"bench: fp64_18_slab_share: {res['fp64_18_slab_share']}")
    hop16 = []
    if not args.no_dhop and (not args.only or "fp16_rhs" in args.only.split(",")):
        for dims in ([int(x) for x in args.cg_lattice.split(",")], fd):
            Ud = U if dims == fd else smooth_links_device(torch, dims, 0.3, 3001, dev)
            hop16.append(hop16_measure(torch, g, dev, Ud, dims, nrhs=args.nrhs))
            log(f"bench: fp16 hop {dims} N={args.nrhs}: {hop16[-1]}")
            del Ud
            torch.cuda.empty_cache()

    copy_gbs = copy_bandwidth(torch, dev) if not args.no_dhop else 0.0
    log(f"bench: copy kernel {copy_gbs:.0f} GB/s")
    probe = probe_bandwidth(torch, dev) if not args.no_dhop else {}
    log(f"bench: streaming probe {probe}")

    # serving runs first after the Dhop section, as a serving process starts on a settled device: right
    # after the multi-RHS section has freed ~200 GB, a hipMalloc of a new serving group can block the
    # device for ~2 s (HIP API trace), which lands in the Poisson p99; 15 s later it does not
    serve_res = None
    if args.serve > 0:
        serve_res = serve_measure(torch, g, dev, args.serve, npoisson=args.serve_poisson, load=args.serve_load,
                                  streams=args.serve_streams, rate=args.serve_rate, min_slots=args.serve_min_slots)
    if args.multi_mass != "none" and (args.serve > 0 or args.multi_mass != "all"):
        mm = multi_mass_measure(torch, g, dev, only=None if args.multi_mass == "all" else args.multi_mass)
        serve_res = serve_res if serve_res is not None else {}
        serve_res["multi_mass"] = mm

    cg_res = None
    if not args.no_cg:
        cg_res = cg_measure(torch, g, dev, fd=[int(x) for x in args.cg_lattice.split(",")], nrhs=args.nrhs)

    c1_res = None
    if not args.no_c1:
        c1_res = c1_measure(torch, g, dev)

    c4_res = []
    for case in [c for c in args.c4.split(",") if c]:
        dims, nr = case.split(":")
        c4_res.append(c4_measure(torch, g, dev, [int(x) for x in dims.split("x")], int(nr)))
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(U)
        log(f"bench: cpu baseline {cpu}")

    head = res.get("fp64_18") or next(iter(res.values()), {"GBps": 0.0, "us_per_call": 0.0})
    achieved = head["GBps"]
    traffic = None
    if "fp64_18" in res and [int(x) for x in fd] == [32, 32, 32, 64]:
        sh = head["shape"]
        traffic = profiled_traffic((sh["rhs_per_thread"], sh["threads_per_block"] // 64, sh["max_blocks_per_cu"],
                                    sh["t_runs"], sh["lds_staging"]))
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        # the headline's per-call spread at the top level: value and ms_per_step are the median call
        "headline": {k: head.get(k) for k in ("us_per_call", "us_per_call_mean", "us_per_call_min",
                                               "us_per_call_max", "mean_off_median", "wall_us_per_call",
                                               "shape", "bytes_per_site")},
        "timing": "HIP events around every call on the launch stream; value and ms_per_step from the median call",
        "cpu_baseline": cpu,
        "variants": res,
        "fp16_hop_multi_rhs": hop16,
        "copy_kernel_GBps": round(copy_gbs, 1),
        "hbm_probe": probe,
        "cg": cg_res,
        "c1_point_source_8x8x8x8": c1_res,
        "multi_rhs": c4_res,
        "serving": serve_res,
    }
    print(json.dumps(out), flush=True)


def true_residual(g, w, x, src):
    """|src - M x| / |src| with the FP64 operator (per right-hand side)"""
    r = g(src - w * x)
    n2r, n2s = g.norm2(r), g.norm2(src)
    if not isinstance(n2r, list):
        n2r, n2s = [n2r], [n2s]
    return [float(np.sqrt(a / b)) for a, b in zip(n2r, n2s)]


def cg_measure(torch, g, dev, fd=(24, 24, 24, 48), mass=0.05, nrhs=12, tol=1e-10, reps=3):
    """Solves/s to a true relative residual tol: plain FP64 eo2_ne CG one RHS
    at a time, the mixed-precision reliable-update solve (FP32 inner) one RHS
    at a time, and the same on nrhs right-hand sides at once."""
    from gpt_amd.qcd.fermion.operator import pack

    fd = list(fd)
    U = smooth_links_device(torch, fd, 0.3, 3001, dev)
    prm = dict(mass=mass, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0, boundary_phases=[1, 1, 1, -1])
    w = g.qcd.fermion.wilson_clover(U, prm)
    wN = g.qcd.fermion.wilson_clover(U, prm, n_rhs=nrhs)
    inv, pc = g.algorithms.inverter, g.qcd.fermion.preconditioner
    gen = torch.Generator(device=dev)
    gen.manual_seed(3002)
    srcs = []
    for _ in range(nrhs):
        x = g.lattice(w.F_grid)
        x.data.copy_(torch.randn(x.data.shape, generator=gen, dtype=torch.float64, device=dev) / np.sqrt(2))
        srcs.append(x)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    out = {"lattice": fd, "mass": mass, "csw": 1.0, "tol_true_residual": tol, "nrhs": nrhs}
    # (a) plain FP64 CG on the normal equation, eps tightened until the true residual meets tol
    eps = tol
    for _ in range(6):
        cg = inv.cg(eps=eps, maxiter=20000)
        slv = w.propagator(inv.preconditioned(pc.eo2_ne(), cg))
        x, _t = timed(lambda: g(slv * srcs[0]))
        if true_residual(g, w, x, srcs[0])[0] <= tol:
            break
        eps /= 2.0
    lat, res, its = [], [], []
    for i in range(reps):
        x, t = timed(lambda: g(slv * srcs[i]))
        lat.append(t)
        res.append(true_residual(g, w, x, srcs[i])[0])
        its.append(len(cg.history))
    x64 = [g(slv * srcs[i]) for i in range(2)]
    out["fp64_cg"] = {"eps_normal_equation": eps, "iterations": its, "true_residual": max(res),
                      "s_per_solve": float(np.median(lat)), "solves_per_s": float(1.0 / np.median(lat))}
    log(f"bench: cg fp64 {out['fp64_cg']}")
    # (b) mixed precision, one RHS at a time
    mcg = inv.cg_mixed(eps=tol, maxiter=20000)
    slvm = w.propagator(inv.preconditioned(pc.eo2_ne(), mcg))
    g(slvm * srcs[0])
    lat, res, its, err = [], [], [], []
    for i in range(reps):
        x, t = timed(lambda: g(slvm * srcs[i]))
        lat.append(t)
        res.append(true_residual(g, w, x, srcs[i])[0])
        its.append(mcg.iterations[0])
        if i < 2:
            err.append(float(np.sqrt(g.norm2(g(x - x64[i])) / g.norm2(x64[i]))))
    out["mixed_1rhs"] = {"iterations": its, "reliable_updates": mcg.reliable_updates, "true_residual": max(res),
                         "rel_error_vs_fp64_solution": max(err), "s_per_solve": float(np.median(lat)),
                         "solves_per_s": float(1.0 / np.median(lat))}
    log(f"bench: cg mixed 1 {out['mixed_1rhs']}")
    # (c) mixed precision, nrhs at once
    mN = inv.cg_mixed(eps=tol, maxiter=20000)
    slvN = wN.propagator(inv.preconditioned(pc.eo2_ne(), mN))
    S = pack(srcs)
    g(slvN * S)
    lat = []
    for _ in range(2):
        X, t = timed(lambda: g(slvN * S))
        lat.append(t)
    res = true_residual(g, wN, X, S)
    err = [float(np.sqrt(g.norm2(g(X.rhs(i) - x64[i])) / g.norm2(x64[i]))) for i in range(2)]
    tN = float(np.median(lat))
    out["mixed_multi"] = {"iterations": mN.iterations, "reliable_updates": mN.reliable_updates,
                          "true_residual": max(res), "rel_error_vs_fp64_solution": max(err),
                          "s_per_batch": tN, "solves_per_s": float(nrhs / tN)}
    log(f"bench: cg mixed {nrhs} {out['mixed_multi']}")
    # (d) the same with 16-bit storage (FP16, then BF16) of the inner search
    # direction and Schur intermediates
    for key, op_, S_, nr, st in [("mixed_1rhs_half", w, srcs[0], 1, "half"), ("mixed_multi_half", wN, S, nrhs, "half"),
                                 ("mixed_1rhs_bf16", w, srcs[0], 1, "bf16"), ("mixed_multi_bf16", wN, S, nrhs, "bf16")]:
        mh = inv.cg_mixed(eps=tol, maxiter=20000, inner_storage=st)
        slvh = op_.propagator(inv.preconditioned(pc.eo2_ne(), mh))
        g(slvh * S_)
        lat = []
        for _ in range(2):
            X, t = timed(lambda: g(slvh * S_))
            lat.append(t)
        res = true_residual(g, op_, X, S_)
        xs = [X] if nr == 1 else [X.rhs(i) for i in range(2)]
        err = [float(np.sqrt(g.norm2(g(xs[i] - x64[i])) / g.norm2(x64[i]))) for i in range(len(xs))]
        th = float(np.median(lat))
        out[key] = {"iterations": mh.iterations, "reliable_updates": mh.reliable_updates,
                    "true_residual": max(res), "rel_error_vs_fp64_solution": max(err),
                    "s_per_batch": th, "solves_per_s": float(nr / th),
                    "inner_hop_on_matrix_cores": mh.inner_operator(op_).hop_mma(nr)}
        log(f"bench: cg {key} {out[key]}")
    # (d') the 12-RHS FP16 solve with the inner hops forced onto each form
    for key, form in (("mixed_multi_half_matrix_cores", True), ("mixed_multi_half_vector_alu", False)):
        mh = inv.cg_mixed(eps=tol, maxiter=20000, inner_storage="half", mma=form)
        slvh = wN.propagator(inv.preconditioned(pc.eo2_ne(), mh))
        g(slvh * S)
        lat = []
        for _ in range(2):
            X, t = timed(lambda: g(slvh * S))
            lat.append(t)
        th = float(np.median(lat))
        out[key] = {"iterations": mh.iterations, "reliable_updates": mh.reliable_updates,
                    "max_true_residual": max(true_residual(g, wN, X, S)), "s_per_batch": th,
                    "solves_per_s": float(nrhs / th)}
        log(f"bench: cg {key} {out[key]}")
    # (e) FP16 storage on 48 right-hand sides at once (the 12 above and 36 more from the same generator)
    n48 = 48
    extra = []
    for _ in range(n48 - nrhs):
        x = g.lattice(w.F_grid)
        x.data.copy_(torch.randn(x.data.shape, generator=gen, dtype=torch.float64, device=dev) / np.sqrt(2))
        extra.append(x)
    w48 = g.qcd.fermion.wilson_clover(U, prm, n_rhs=n48)
    S48 = pack(srcs + extra)
    m48 = inv.cg_mixed(eps=tol, maxiter=20000, inner_storage="half")
    slv48 = w48.propagator(inv.preconditioned(pc.eo2_ne(), m48))
    g(slv48 * S48)
    lat = []
    for _ in range(2):
        X, t = timed(lambda: g(slv48 * S48))
        lat.append(t)
    res = true_residual(g, w48, X, S48)
    t48 = float(np.median(lat))
    out["mixed_multi48_half"] = {"nrhs": n48, "iterations_max": max(m48.iterations),
                                 "reliable_updates": m48.reliable_updates, "max_true_residual": max(res),
                                 "s_per_batch": t48, "solves_per_s": float(n48 / t48)}
    log(f"bench: cg mixed_multi48_half {out['mixed_multi48_half']}")
    del X, S48, slv48, w48, extra
    out["speedup_mixed_multi_vs_fp64"] = out["mixed_multi"]["solves_per_s"] / out["fp64_cg"]["solves_per_s"]
    out["speedup_mixed_multi_half_vs_fp64"] = (out["mixed_multi_half"]["solves_per_s"]
                                               / out["fp64_cg"]["solves_per_s"])
    out["speedup_mixed_multi_bf16_vs_fp64"] = (out["mixed_multi_bf16"]["solves_per_s"]
                                               / out["fp64_cg"]["solves_per_s"])
    out["speedup_mixed_multi48_half_vs_fp64"] = (out["mixed_multi48_half"]["solves_per_s"]
                                                 / out["fp64_cg"]["solves_per_s"])
    return out


def _timed(torch, fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def c1_measure(torch, g, dev, reps=3):
    """8^4 point-source propagator: Haar-random links (seed 1001), mass 0.1,
    csw 1, boundary phases [1,1,1,-1]; the 12 spin-colour right-hand sides
    solved with eo2_ne + CG to 1e-8, (a) one after another as gpt does and
    (b) as one 12-RHS solve; each with the iterations run one at a time
    (host test after each) and in captured graphs."""
    from gpt_amd.qcd.fermion.operator import pack

    fd = [8, 8, 8, 8]
    U = haar_links_device(torch, fd, 1001, dev)
    prm = dict(mass=0.1, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])
    inv, pc = g.algorithms.inverter, g.qcd.fermion.preconditioner
    grid = g.grid(fd, g.double)
    srcs = []
    for s in range(4):
        for c in range(3):
            a = np.zeros(tuple(reversed(fd)) + (4, 3), complex)
            a[0, 0, 0, 0, s, c] = 1.0
            x = g.lattice(grid)
            x.from_numpy(a)
            srcs.append(x)
    out = {"lattice": fd, "mass": 0.1, "eps": 1e-8}
    for mode, chunk in [("one_at_a_time", 1), ("graphs", 0)]:
        w = g.qcd.fermion.wilson_clover(U, prm)
        w.set_iteration_chunk(chunk)
        cg = inv.cg(eps=1e-8, maxiter=5000)
        slv = w.propagator(inv.preconditioned(pc.eo2_ne(), cg))
        its = []

        def seq():
            its.clear()
            for x in srcs:
                g(slv * x)
                its.append(len(cg.history))

        seq()
        lat = [_timed(torch, seq)[1] for _ in range(reps)]
        w12 = g.qcd.fermion.wilson_clover(U, prm, n_rhs=12)
        w12.set_iteration_chunk(chunk)
        cg12 = inv.cg(eps=1e-8, maxiter=5000)
        slv12 = w12.propagator(inv.preconditioned(pc.eo2_ne(), cg12))
        S = pack(srcs)
        g(slv12 * S)
        lat12 = [_timed(torch, lambda: g(slv12 * S))[1] for _ in range(reps)]
        out[mode] = {"sequential_ms_per_propagator": round(1e3 * float(np.median(lat)), 3),
                     "sequential_iterations": its[:],
                     "rhs12_ms_per_propagator": round(1e3 * float(np.median(lat12)), 3),
                     "rhs12_iterations": list(cg12.iterations),
                     "graph_runs": w12.iteration_runs()[0]}
        log(f"bench: c1 {mode} {out[mode]}")
        w.release()
        w12.release()
    return out


def release_device_caches(torch):
    """return the blocks torch and the library keep cached, and the pool of
    parked operators, to the device, so that the free memory reported is what
    a new allocation can get"""
    import ctypes
    import gc

    from gpt_amd import _lib
    from gpt_amd.qcd.fermion import operator as fop

    gc.collect()  # operators dropped earlier park their handles in the pool
    torch.cuda.synchronize()
    fop.release_pool()
    n = ctypes.c_long()
    _lib.check(_lib.load().wc_release_cached_memory(ctypes.byref(n)))
    torch.cuda.empty_cache()


def c4_batch(Vh, nrhs, lt, free_bytes):
    """the largest batch that divides nrhs such that a t-slice of an FP64 half
    field of that many right-hand sides fits one buffer descriptor (4 GB; the
    hop addresses larger fields per t-slice) and about 20 FP64 half fields of
    it (the mixed solve's 13, the sources, solutions and block CG's) fit in
    85 % of the free device memory"""
    field = Vh * 12 * 16
    return max(d for d in range(1, nrhs + 1)
               if nrhs % d == 0 and field // lt * d < (1 << 32) and 20 * field * d <= 0.85 * free_bytes)


def c4_measure(torch, g, dev, fd, nrhs, tol=1e-10, seed=4001):
    """Multi-RHS: smoothed links (seed 4001), mass 0.08, csw 1; nrhs
    complex-Gaussian sources (seeds 4100 + j) to a true relative residual
    tol.  FP64 eo2_ne CG one RHS at a time against the mixed-precision
    multi-RHS solve with 16-bit inner storage, in equal batches as large as
    the device memory allows (c4_batch; a batch that fails to allocate is
    run again as two)."""
    from gpt_amd.qcd.fermion.operator import pack

    U = smooth_links_device(torch, fd, 0.3, seed, dev)
    prm = dict(mass=0.08, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])
    inv, pc = g.algorithms.inverter, g.qcd.fermion.preconditioner
    grid = g.grid(fd, g.double)
    Vh = int(np.prod(fd)) // 2
    release_device_caches(torch)
    nb = c4_batch(Vh, nrhs, fd[3], torch.cuda.mem_get_info()[0])
    srcs = []
    for j in range(nrhs):
        gen = torch.Generator(device=dev)
        gen.manual_seed(4100 + j)
        x = g.lattice(grid)
        x.data.copy_(torch.randn(x.data.shape, generator=gen, dtype=torch.float64, device=dev) / np.sqrt(2))
        srcs.append(x)
    out = {"lattice": fd, "nrhs": nrhs, "mass": 0.08, "tol_true_residual": tol, "batch": nb}
    w = g.qcd.fermion.wilson_clover(U, prm)
    eps = tol
    for _ in range(6):
        cg = inv.cg(eps=eps, maxiter=20000)
        slv = w.propagator(inv.preconditioned(pc.eo2_ne(), cg))
        x0 = g(slv * srcs[0])
        if true_residual(g, w, x0, srcs[0])[0] <= tol:
            break
        eps /= 2.0
    lat, its = [], []
    for i in range(2):
        x, t = _timed(torch, lambda: g(slv * srcs[i]))
        lat.append(t)
        its.append(len(cg.history))
    t64 = float(np.median(lat))
    out["fp64_sequential"] = {"eps_normal_equation": eps, "iterations": its, "s_per_solve": t64,
                              "solves_per_s": 1.0 / t64}
    def batched(nb):
        """the 16-bit mixed multi-RHS solve and FP64 block CG in batches of nb"""
        res_out = {}
        wN = g.qcd.fermion.wilson_clover(U, prm, n_rhs=nb)
        mh = inv.cg_mixed(eps=tol, maxiter=20000, inner_storage="half")
        try:
            slvh = wN.propagator(inv.preconditioned(pc.eo2_ne(), mh))
            batches = [pack(srcs[i:i + nb]) for i in range(0, nrhs, nb)]
            g(slvh * batches[0])  # warm-up (tuning, graphs)
            res_out["inner_hop_on_matrix_cores"] = mh.inner_operator(wN).hop_mma(nb)

            def run():
                res, itr = [], []
                for S in batches:
                    X = g(slvh * S)
                    res += true_residual(g, wN, X, S)
                    itr += list(mh.iterations)
                return res, itr

            (res, itr), t = _timed(torch, run)
            res_out["mixed_multi_half"] = {"iterations_max": max(itr), "reliable_updates": mh.reliable_updates,
                                           "max_true_residual": max(res), "s_total": t,
                                           "solves_per_s": nrhs / t}
            res_out["speedup_vs_fp64_sequential"] = res_out["mixed_multi_half"]["solves_per_s"] * t64
            # the mixed solve's operators and scratch are not needed by block CG
            mh.inner_operator(wN).release()
            wN.release()
            torch.cuda.empty_cache()
            # FP64 block CG (block_cg.py) on the same batches, eps as the FP64 CG's
            wB = g.qcd.fermion.wilson_clover(U, prm, n_rhs=nb)
            bcg = inv.block_cg(eps=eps, maxiter=20000)
            slvb = wB.propagator(inv.preconditioned(pc.eo2_ne(), bcg))
            g(slvb * batches[0])

            def run_block():
                res, itr = [], []
                for S in batches:
                    X = g(slvb * S)
                    res += true_residual(g, wB, X, S)
                    itr.append(bcg.iterations)
                return res, itr

            (res, itr), t = _timed(torch, run_block)
            res_out["block_cg_fp64"] = {"iterations": itr, "max_true_residual": max(res), "s_total": t,
                                        "solves_per_s": nrhs / t}
            wB.release()
        finally:
            for op in (wN, mh.inner_operator(wN)):
                op.release()
            release_device_caches(torch)
        return res_out

    while True:
        try:
            out.update(batched(nb))
            break
        except RuntimeError as e:
            if nb == 1 or "hipMalloc" not in str(e) and "out of memory" not in str(e):
                raise
            log(f"bench: multi-rhs {fd} N={nrhs}: batch {nb} does not fit ({e}); halving")
            nb = max(d for d in range(1, nb // 2 + 1) if nrhs % d == 0)
            out["batch"] = nb
            out["batch_fallback"] = True
    log(f"bench: multi-rhs {fd} N={nrhs}: {out}")
    w.release()
    del w, srcs, U
    return out


def serve_measure(torch, g, dev, nreq, seed=5001, nslots=12, load=0.8, tol=1e-10, iteration_chunk=None,
                  npoisson=200, streams=4, rate=None, min_slots=1):
    """Serving stream: requests over volumes {8^4, 16^3x32, 24^3x48, 32^3x64},
    3 resident fields per volume (seeds 5101..5103), mass U[0.02, 0.2], 1..12
    Gaussian right-hand sides, tolerance tol on the true residual.  First nreq
    requests at once (capacity, right-hand sides per second), then npoisson
    new ones with Poisson arrivals at load x that capacity: p50/p99 latency
    and solves/s.  One group per field serves every mass (clover inverses
    per slot, one set per mass); the time to set a group up (operators and batch) is reported
    apart from the latencies, which include it.  Sources come from a pool of
    12 Gaussian fields per volume (a request takes distinct ones), and
    results are dropped once checked, so that hundreds of requests fit in
    memory."""
    from gpt_amd import serve

    vols = [[8, 8, 8, 8], [16, 16, 16, 32], [24, 24, 24, 48], [32, 32, 32, 64]]
    prm = dict(csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0, boundary_phases=[1, 1, 1, -1])
    srv = serve.server(prm, nslots=nslots, streams=streams, iteration_chunk=iteration_chunk, min_slots=min_slots)
    for vi, fd in enumerate(vols):
        for k in range(3):
            srv.add_field((vi, k), smooth_links_device(torch, fd, 0.3, 5101 + k, dev))
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    pool = []
    for fd in vols:
        grid = g.grid(fd, g.double)
        fs = []
        for _ in range(12):
            x = g.lattice(grid)
            x.data.copy_(torch.randn(x.data.shape, generator=gen, dtype=torch.float64, device=dev) / np.sqrt(2))
            fs.append(x)
        pool.append(fs)

    def make(n):
        reqs = []
        for _ in range(n):
            vi, k = int(rng.integers(4)), int(rng.integers(3))
            m, nr = float(rng.uniform(0.02, 0.2)), int(rng.integers(1, 13))
            pick = rng.permutation(12)[:nr]
            reqs.append(serve.request((vi, k), m, [pool[vi][int(i)] for i in pick], tol=tol, maxiter=50000))
        return reqs

    def summary(done, T):
        nrhs = sum(len(r.sources) for r in done)
        st = serve.latency_stats(done)
        return {"requests": len(done), "rhs": nrhs, "wall_s": T, "solves_per_s": nrhs / T,
                "p50_latency_s": st["p50_s"], "p99_latency_s": st["p99_s"],
                "max_true_residual": max(max(r.true_residual) for r in done),
                "all_converged": all(all(r.converged) for r in done)}

    def progress(r):
        if r.id % 20 == 0:
            log(f"bench: serving request {r.id} done: {len(r.sources)} rhs, field {r.field}, mass {r.mass:.3f}, "
                f"latency {r.latency:.3f} s, iterations {max(r.iterations)}")
        r.results = [None] * len(r.results)  # checked (true residual, converged): free the fields

    def setups(ts):
        ts = np.asarray(ts) if len(ts) else np.zeros(1)
        return {"groups": int(len(ts)), "median_s": float(np.median(ts)), "max_s": float(ts.max()),
                "total_s": float(ts.sum())}

    reqs = make(nreq)
    t0 = time.perf_counter()
    done = srv.run([(0.0, r) for r in reqs], on_done=progress)
    cap = summary(done, time.perf_counter() - t0)
    cap["group_setup"] = setups(srv.setup_times)
    log(f"bench: serving capacity {cap}")
    del reqs, done
    srv.setup_times = []
    peak_cap = srv.peak_groups
    srv.peak_groups = 0
    reqs = make(npoisson)
    mean_rhs = float(np.mean([len(r.sources) for r in reqs]))
    lam = rate if rate else load * cap["solves_per_s"] / mean_rhs
    arrivals = np.cumsum(rng.exponential(1.0 / lam, npoisson))
    t0 = time.perf_counter()
    done = srv.run(list(zip(arrivals.tolist(), reqs)), on_done=progress)
    pois = summary(done, time.perf_counter() - t0)
    pois["arrival_rate_per_s"] = lam
    pois["group_setup"] = setups(srv.setup_times)
    log(f"bench: serving poisson {pois}")
    out = {"nslots": nslots, "streams": len(srv.streams), "load": load, "capacity": cap, "poisson": pois,
           "peak_groups_capacity": peak_cap, "peak_groups_poisson": srv.peak_groups,
           "resident_fields": len(srv.fields), "grouping": "one group per resident field, any mass",
           "hbm_budget_gb": srv.hbm_budget / 2**30, "scheduler_passes_waiting_for_hbm": srv.waited_for_memory}
    srv.close()
    return out


def multi_mass_measure(torch, g, dev, fd=(24, 24, 24, 48), nslots=12, tol=1e-10, only=None):
    """Cost of several masses in one serving batch: the same 12 Gaussian
    sources in a 12-slot batch (FP64 outer, 12-real FP32 inner operator at
    mass 0.1, csw 1) with one mass (the operators'), one other mass, three
    masses of 4 slots each, and twelve masses; seconds from the first
    admission until all 12 solves are extracted (tolerance tol on the true
    residual).  Slots of one mass share a clover-inverse set, and a batch
    whose occupied slots all use one set stages it per site, so the cost
    grows with the number of distinct masses, not with the slots.  only: one
    case (one, other, three, twelve) without the warm-up, for profiling."""
    from gpt_amd import serve

    fd = list(fd)
    U = smooth_links_device(torch, fd, 0.3, 5101, dev)
    prm = dict(mass=0.1, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0, boundary_phases=[1, 1, 1, -1])
    op = g.qcd.fermion.wilson_clover(U, prm)
    inner = g.qcd.fermion.wilson_clover(U, dict(prm, compress_links=True), precision=g.single)
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    grid = g.grid(fd, g.double)
    srcs = []
    for _ in range(nslots):
        x = g.lattice(grid)
        x.data.copy_(torch.randn(x.data.shape, generator=gen, dtype=torch.float64, device=dev) / np.sqrt(2))
        srcs.append(x)
    third = nslots // 3
    cases = [("warm-up", [None] * nslots), ("one mass", [None] * nslots), ("one other mass", [0.08] * nslots),
             ("three masses", [0.08] * third + [None] * third + [0.12] * (nslots - 2 * third)),
             ("twelve masses", [0.1 + 0.005 * (j - 6) if j != 6 else None for j in range(nslots)])]
    if only is not None:
        names = {"one": "one mass", "other": "one other mass", "three": "three masses", "twelve": "twelve masses"}
        if only not in names:
            raise SystemExit(f"--multi-mass {only}: not one of all, none, {', '.join(names)}")
        cases = [c for c in cases if c[0] == names[only]]
    out = {"lattice": "x".join(map(str, fd)), "slots": nslots, "tol": tol}
    dst = g.lattice(op.F_grid)
    for label, masses in cases:
        b = serve.batch(op, inner, nslots)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for x, m in zip(srcs, masses):
            b.admit(x, tol, 20000, mass=m)
        done, worst, conv = 0, 0.0, True
        while done < nslots:
            for sl in b.step(25):
                it, tres, c = b.extract(sl, dst)
                worst, conv, done = max(worst, tres), conv and bool(c), done + 1
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        b.close()
        if label != "warm-up":
            out[label.replace(" ", "_") + "_s"] = t
            out["max_true_residual"] = max(out.get("max_true_residual", 0.0), worst)
            out["all_converged"] = out.get("all_converged", True) and conv
        log(f"bench: multi-mass batch, {label}: {t:.3f} s for {nslots} solves")
    op.release()
    inner.release()
    return out


if __name__ == "__main__":
    main()
