"""The hop at the benchmarked geometry: 32^3 x 64, the headline's links
(U = exp(i 0.3 H), the benchmark's own generator on the GPU), against the C
restatement of the oracle per site.  This covers what the small lattices
cannot: the 2.4 GB double-stored link buffers, the XCD tiling over 16,384
workgroups, the spinor-block staging with 16-half-site rows (ls = 2) and the
tuner's pick on the shape that is timed.

Bars: FP64 <= 1e-13 relative per site (18- and 12-real links), FP32 <= 2e-6."""
import os
import sys

import numpy as np
import pytest

from oracle import cpu as C
from oracle import wilson_clover as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd.qcd import gauge as GA  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402

FD = [32, 32, 32, 64]
P = dict(mass=0.08, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0, boundary_phases=[1, 1, 1, -1])


def site_rel_err_half(got, ref):
    """max over half-lattice sites of |got - ref|_site / |ref|_site"""
    d = np.sqrt((np.abs(got - ref) ** 2).reshape(len(ref), -1).sum(-1))
    r = np.sqrt((np.abs(ref) ** 2).reshape(len(ref), -1).sum(-1))
    return float((d / np.where(r > 0, r, 1.0)).max())


@pytest.fixture(scope="module")
def big():
    dev = torch.device("cuda", 0)
    print("fullsize: generating 32^3x64 links on the GPU", flush=True)
    U = bench.smooth_links_device(torch, FD, 0.3, 2001, dev)
    Un = GA.to_numpy(U)
    # the hop does not involve the clover term: the oracle's operator without it
    ref = C.wilson_clover_cpu(W.wilson_clover(Un, mass=P["mass"], boundary_phases=P["boundary_phases"]))
    del Un
    rng = np.random.default_rng(2002)
    Vh = ref.Vh
    srcs = {p: (rng.standard_normal((Vh, 4, 3)) + 1j * rng.standard_normal((Vh, 4, 3))) / np.sqrt(2.0)
            for p in (0, 1)}
    return U, ref, srcs


def _half_in(w, arr_half, p, dtype):
    """a half field of parity p from a [Vh, 4, 3] array in the oracle's half order"""
    full = np.zeros(tuple(reversed(FD)) + (4, 3), dtype)
    flat = full.reshape(-1, 4, 3)
    flat[C.half_indices(full.shape[:4], p)] = arr_half.astype(dtype)
    x = g.lattice(w.F_grid_eo)
    x.checkerboard(g.odd if p else g.even)
    x.from_numpy(full)
    return x


def _half_out(x, p):
    return C.half(x.to_numpy().astype(np.complex128), p)


@pytest.mark.parametrize("prec,compress", [(g.double, False), (g.double, True), (g.single, False),
                                           (g.single, True)])
def test_meooe_32x32x32x64_per_site(big, prec, compress):
    U, ref, srcs = big
    tol = 1e-13 if prec is g.double else 2e-6
    w = g.qcd.fermion.wilson_clover(U, P, precision=prec, compress_links=compress)
    tuned, _ = w.tune_hop(1, reps=3)
    print(f"fullsize: {prec.__name__} {'12' if compress else '18'}-real tuned shape {tuned}", flush=True)
    dtype = np.complex128 if prec is g.double else np.complex64
    exp = {(pin, dag): ref.meooe(srcs[pin], pin, dag=dag) for pin in (0, 1) for dag in (False, True)}
    for shape in [tuned, (1, 1, 4, 1, 2), (1, 1, 4, 1, 0), (1, 4, 0, 1, 0)]:
        w.set_hop_shape(shape)
        for pin in (0, 1):
            s = _half_in(w, srcs[pin], pin, dtype)
            for dag in (False, True):
                op = w.Meooe.adj() if dag else w.Meooe
                err = site_rel_err_half(_half_out(op(s), 1 - pin), exp[(pin, dag)])
                print(f"fullsize: shape {shape} pin {pin} dag {dag}: max site rel err {err:.2e}", flush=True)
                assert err < tol, (shape, pin, dag, err)
    w.release()


@pytest.mark.parametrize("prec", [g.double, g.single])
def test_fields_beyond_the_4gb_window(big, prec):
    """32^3 x 64 with 24 right-hand sides in FP64 (a 4.8 GB half field) and 48
    in FP32 (4.8 GB): the hop addresses such fields per t-slice.  Meooe and
    the normal operator (clover, axpy and inner-product epilogues) of the
    packed field equal the single-RHS results for the first, a middle and the
    last right-hand side, bitwise or to 1e-14 (FP32: 1e-6) of the field"""
    U, ref, srcs = big
    N = 24 if prec is g.double else 48
    dev = torch.device("cuda", 0)
    wN = g.qcd.fermion.wilson_clover(U, P, precision=prec, n_rhs=N)
    w1 = g.qcd.fermion.wilson_clover(U, P, precision=prec)
    assert 12 * (2 ** 20) * N * prec.nbytes * 2 > 2 ** 32  # Vh = 2^20 half-sites, complex = 2 reals
    X = g.lattice(wN.F_grid_eo)
    X.checkerboard(g.odd)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2003)
    X.data.copy_(torch.randn(X.data.shape, generator=gen, dtype=torch.float64, device=dev).to(X.data.dtype))
    Y = wN.Meooe(X)
    pc = g.qcd.fermion.preconditioner
    Z = pc.eo2_ne()(wN).Mpc(X)
    # the t-marching order (4 runs per time slice) addresses the same slices
    # in another order: the same bits
    sh = wN.hop_shape(N)
    wN.set_hop_shape((1, 4, 4, 4, 1))
    assert torch.equal(wN.Meooe(X).data, Y.data)
    wN.set_hop_shape(None)
    assert wN.hop_shape(N) == sh
    torch.cuda.synchronize()
    tol = 1e-14 if prec is g.double else 1e-6
    for r in (0, N // 2, N - 1):
        x1 = X.rhs(r)
        y1 = w1.Meooe(x1)
        z1 = pc.eo2_ne()(w1).Mpc(x1)
        for a, b in ((Y.rhs(r), y1), (Z.rhs(r), z1)):
            d = float((a.data - b.data).abs().max() / b.data.abs().max())
            print(f"beyond 4 GB: {prec.__name__} N={N} rhs {r}: max rel diff {d:.1e}", flush=True)
            assert d <= tol, (r, d)
    wN.release()
    w1.release()
