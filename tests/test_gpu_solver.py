"""CG / eo2_ne solves on the GPU against the oracle: iteration counts within
+-1 of the reference recursion, solutions, and the reference's correlator
known answer (tests/qcd/fermion_operators.py:150-218) run in single precision
as the reference runs it."""
import cmath

import numpy as np
import pytest

from oracle import cpu as C
from oracle import gauge as G
from oracle import rng as R
from oracle import wilson_clover as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402

inv = g.algorithms.inverter
pc = g.qcd.fermion.preconditioner

BASE = dict(mass=0.1, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])


def to_dev(arr, grid):
    x = g.lattice(grid)
    x.from_numpy(arr)
    return x


@pytest.fixture(scope="module")
def config1():
    """8^4, Haar-random links (seed 1001), m = 0.1, csw = 1, BC [1,1,1,-1]"""
    fd = [8, 8, 8, 8]
    U = G.haar_links(fd, 1001)
    return fd, U


def test_point_source_propagator_matches_oracle(config1):
    fd, U = config1
    ref = W.wilson_clover(U, **BASE)
    refc = C.wilson_clover_cpu(ref)
    w = g.qcd.fermion.wilson_clover(U, BASE)
    cg = inv.cg(eps=1e-8, maxiter=2000)
    slv = w.propagator(inv.preconditioned(pc.eo2_ne(), cg))
    shape = tuple(reversed(fd))
    for s in range(4):
        for c in range(3):
            src = np.zeros(shape + (4, 3), complex)
            src[0, 0, 0, 0, s, c] = 1.0
            x_ref, it_ref = refc.solve(src, eps=1e-8, maxiter=2000)
            dst = g(slv * to_dev(src, w.F_grid))
            it = len(cg.history)
            assert cg.converged and abs(it - it_ref) <= 1, (s, c, it, it_ref)
            x = dst.to_numpy()
            assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-6
            # true residual of the full operator
            r = ref.M(x) - src
            assert np.linalg.norm(r) / np.linalg.norm(src) < 1e-6


def test_cg_history_matches_reference_recursion(config1):
    fd, U = config1
    ref = W.wilson_clover(U, **BASE)
    refc = C.wilson_clover_cpu(ref)
    rng = np.random.default_rng(7)
    shape = tuple(reversed(fd))
    b = rng.standard_normal(shape + (4, 3)) + 1j * rng.standard_normal(shape + (4, 3))
    b[ref.parity == 0] = 0.0
    xh, it_ref, hist_ref = refc.cg_ne(C.half(b, 1), 1, eps=1e-10, maxiter=3000, history=True)
    w = g.qcd.fermion.wilson_clover(U, BASE)
    p = pc.eo2_ne()(w)
    B = g.lattice(w.F_grid_eo)
    B.checkerboard(g.odd)
    B.from_numpy(b)
    cg = inv.cg(eps=1e-10, maxiter=3000)
    X = cg(p.Mpc)(B)
    assert abs(len(cg.history) - it_ref) <= 1
    n = min(len(cg.history), it_ref) // 2
    assert np.allclose(cg.history[:n], hist_ref[:n], rtol=1e-6)
    x = C.half(X.to_numpy(), 1)
    assert np.linalg.norm(x - xh) / np.linalg.norm(xh) < 1e-8
    # the generic (python-driven) CG on the same operator follows the same recursion
    cg2 = inv.cg(eps=1e-10, maxiter=3000)
    mpc_generic = g.matrix_operator(p.Mpc.mat, make_dst=p.Mpc.make_dst)
    X2 = cg2(mpc_generic)(B)
    assert abs(len(cg2.history) - len(cg.history)) <= 1
    assert np.linalg.norm(C.half(X2.to_numpy(), 1) - x) / np.linalg.norm(x) < 1e-8


def test_correlator_known_answer_single_precision(kats):
    """The reference runs this test with U converted to single precision."""
    k = kats["correlator"]
    gr = R.grid(k["fdims"])
    rng = R.random(k["seed"])
    U = G.random_links(gr, rng)
    p = k["params"]
    w = g.qcd.fermion.wilson_clover(
        U, kappa=p["kappa"], csw_r=0.0, csw_t=0.0, xi_0=1.0, nu=1.0, isAnisotropic=False,
        boundary_phases=[cmath.exp(1j * a) for a in p["boundary_phase_angles"]], precision=g.single)
    cg = inv.cg({"eps": 1e-6, "maxiter": 1000})
    slv = w.propagator(inv.preconditioned(pc.eo2_ne(), cg))
    x0, y0, z0, t0 = k["point_source_xyzt"]
    corr = np.zeros(gr.shape[0])
    for s in range(4):
        for c in range(3):
            src = np.zeros(gr.shape + (4, 3), np.complex64)
            src[t0, z0, y0, x0, s, c] = 1.0
            dst = g(slv * to_dev(src, w.F_grid))
            assert cg.converged
            corr += (np.abs(dst.to_numpy().astype(np.complex128)) ** 2).sum(axis=(1, 2, 3, 4, 5))
    assert np.linalg.norm(corr - np.array(k["reference"])) < k["tol_norm"]


def test_mixed_precision_defect_correction(config1):
    """defect_correcting(mixed_precision(preconditioned(eo2_ne, cg))) as in
    tests/algorithms/solvers.py:111-118, against the FP64 solve."""
    fd, U = config1
    w = g.qcd.fermion.wilson_clover(U, BASE)
    rng = np.random.default_rng(11)
    shape = tuple(reversed(fd))
    src = rng.standard_normal(shape + (4, 3)) + 1j * rng.standard_normal(shape + (4, 3))
    S = to_dev(src, w.F_grid)
    cg64 = inv.cg(eps=1e-12, maxiter=3000)
    x64 = g(w.propagator(inv.preconditioned(pc.eo2_ne(), cg64)) * S)
    dc = inv.defect_correcting(
        inv.mixed_precision(inv.preconditioned(pc.eo2_ne(), inv.cg(eps=1e-5, maxiter=300)), g.single,
                            g.double), eps=1e-10, maxiter=40)
    x = g(dc(w) * S)
    assert dc.history[-1] < 1e-10
    a, b = x.to_numpy(), x64.to_numpy()
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-8


def test_point_source_propagator_as_mspincolor(kats):
    """The reference's propagator test written with its own API
    (tests/qcd/fermion_operators.py:135-218): g.mspincolor, g.create.point,
    dst @= slv * src over the 12 spin-colour columns, EO1 against EO2, the
    true residual w * dst - src, and the correlator g.slice(g.trace(dst *
    g.adj(dst)), 3) against its known answer; a column equals the
    vspincolor solve of the same point source bitwise."""
    k = kats["correlator"]
    gr = R.grid(k["fdims"])
    rng = R.random(k["seed"])
    U = G.random_links(gr, rng)
    p = k["params"]
    w = g.qcd.fermion.wilson_clover(
        U, kappa=p["kappa"], csw_r=0.0, csw_t=0.0, xi_0=1.0, nu=1.0, isAnisotropic=False,
        boundary_phases=[cmath.exp(1j * a) for a in p["boundary_phase_angles"]], precision=g.single)
    grid = w.F_grid
    src = g.mspincolor(grid)
    g.create.point(src, k["point_source_xyzt"])
    cg = inv.cg({"eps": 1e-6, "maxiter": 1000})
    slv_eo1 = w.propagator(inv.preconditioned(pc.eo1_ne(), cg))
    slv_eo2 = w.propagator(inv.preconditioned(pc.eo2_ne(), cg))
    dst_eo1 = g.mspincolor(grid)
    dst_eo2 = g.mspincolor(grid)
    dst_eo1 @= slv_eo1 * src
    dst_eo2 @= slv_eo2 * src
    eps2 = g.norm2(dst_eo1 - dst_eo2) / g.norm2(dst_eo1)
    print(f"EO1 versus EO2: eps2 = {eps2}", flush=True)
    assert eps2 < 1e-12
    eps2 = g.norm2(w * dst_eo1 - src) / g.norm2(src)
    assert eps2 < 1e-10
    correlator = g.slice(g.trace(dst_eo2 * g.adj(dst_eo2)), 3)
    eps = np.linalg.norm(np.array(correlator) - np.array(k["reference"]))
    print(f"correlator eps = {eps}", flush=True)
    assert eps < k["tol_norm"]
    # column (s, c) = 2, 1 is the solve of the point source in spin 2, colour 1
    x0, y0, z0, t0 = k["point_source_xyzt"]
    one = np.zeros(gr.shape + (4, 3), np.complex64)
    one[t0, z0, y0, x0, 2, 1] = 1.0
    col = g(slv_eo2 * to_dev(one, grid)).to_numpy()
    assert np.array_equal(dst_eo2.to_numpy()[..., :, 2, :, 1], col)
