"""Block CG (block_cg.py:36-149) on the GPU against the oracle's block CG:
the same iteration count within +-1 and the same solution to 1e-9 on
identical inputs; one right-hand side reproduces CG; the block needs fewer
iterations than CG needs for any of its right-hand sides; FP32; the
reference's behaviour on a zero block and on a rank-deficient block."""
import numpy as np
import pytest

from oracle import gauge as G
from oracle import rng as R
from oracle import solvers as S
from oracle import wilson_clover as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd.qcd.fermion.operator import pack, unpack  # noqa: E402

inv = g.algorithms.inverter
pc = g.qcd.fermion.preconditioner
P = dict(mass=0.1, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])
FD = [4, 4, 4, 8]


@pytest.fixture(scope="module")
def setup():
    gr = R.grid(FD)
    rng = R.random("block_cg")
    U = G.random_links(gr, rng, scale=0.5)
    srcs = [rng.vspincolor(gr) for _ in range(4)]
    return gr, U, srcs


def _dev(a, grid):
    x = g.lattice(grid)
    x.from_numpy(a)
    return x


def test_block_cg_matches_oracle(setup):
    gr, U, srcs = setup
    N, eps = 3, 1e-10
    ref = W.wilson_clover(U, **P)
    opc = S.eo2_ne(ref, parity=1)
    hist_ref = []
    Xr, it_ref, conv_ref = S.block_cg(opc.Mpc_ne, [opc.R_ne(s) for s in srcs[:N]], eps=eps, maxiter=500,
                                      history=hist_ref)
    x_ref = [opc.L(x) + opc.S(s) for x, s in zip(Xr, srcs[:N])]
    w = g.qcd.fermion.wilson_clover(U, P, n_rhs=N)
    bcg = inv.block_cg(eps=eps, maxiter=500)
    X = g(w.propagator(inv.preconditioned(pc.eo2_ne(), bcg)) * pack([_dev(s, g.grid(FD, g.double)) for s in srcs[:N]]))
    print(f"block CG: GPU {bcg.iterations} iterations, oracle {it_ref}", flush=True)
    assert conv_ref and bcg.converged and abs(bcg.iterations - it_ref) <= 1
    assert len(bcg.history) == bcg.iterations
    k = min(len(hist_ref), len(bcg.history)) - 1
    assert abs(np.log(bcg.history[k] / hist_ref[k])) < 0.5  # same convergence curve
    for a, b, s in zip(unpack(X), x_ref, srcs[:N]):
        a = a.to_numpy()
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-9
        assert np.linalg.norm(ref.M(a) - s) / np.linalg.norm(s) < 1e-8


def test_one_rhs_is_cg_and_block_beats_cg(setup):
    gr, U, srcs = setup
    w1 = g.qcd.fermion.wilson_clover(U, P)
    b1 = inv.block_cg(eps=1e-12, maxiter=500)
    c1 = inv.cg(eps=1e-12, maxiter=500)
    xb = g(w1.propagator(inv.preconditioned(pc.eo2_ne(), b1)) * _dev(srcs[0], w1.F_grid)).to_numpy()
    xc = g(w1.propagator(inv.preconditioned(pc.eo2_ne(), c1)) * _dev(srcs[0], w1.F_grid)).to_numpy()
    assert abs(b1.iterations - len(c1.history)) <= 1
    assert np.linalg.norm(xb - xc) / np.linalg.norm(xc) < 1e-10
    w4 = g.qcd.fermion.wilson_clover(U, P, n_rhs=4)
    b4 = inv.block_cg(eps=1e-10, maxiter=500)
    c4 = inv.cg(eps=1e-10, maxiter=500)
    S4 = pack([_dev(s, g.grid(FD, g.double)) for s in srcs])
    g(w4.propagator(inv.preconditioned(pc.eo2_ne(), b4)) * S4)
    g(w4.propagator(inv.preconditioned(pc.eo2_ne(), c4)) * S4)
    assert b4.converged and b4.iterations < min(c4.iterations)


def test_block_cg_single_precision(setup):
    gr, U, srcs = setup
    ref = W.wilson_clover(U, **P)
    w = g.qcd.fermion.wilson_clover(U, P, n_rhs=2, precision=g.single)
    b = inv.block_cg(eps=1e-5, maxiter=500)
    grs = g.grid(FD, g.single)
    X = g(w.propagator(inv.preconditioned(pc.eo2_ne(), b)) * pack([_dev(s.astype(np.complex64), grs)
                                                                   for s in srcs[:2]]))
    assert b.converged
    for a, s in zip(unpack(X), srcs[:2]):
        assert np.linalg.norm(ref.M(a.to_numpy().astype(complex)) - s) / np.linalg.norm(s) < 1e-4


def test_zero_and_rank_deficient_blocks(setup):
    gr, U, srcs = setup
    w = g.qcd.fermion.wilson_clover(U, P, n_rhs=2)
    z = np.zeros_like(srcs[0])
    b = inv.block_cg(eps=1e-10, maxiter=100)
    X = g(w.propagator(inv.preconditioned(pc.eo2_ne(), b)) * pack([_dev(z, g.grid(FD, g.double)), _dev(z, g.grid(FD, g.double))]))
    assert b.converged and b.iterations == 0 and float(X.data.abs().max()) == 0.0
    # a zero right-hand side beside a non-zero one: the Gram matrix of the
    # residual block is singular, as in the reference (Cholesky fails)
    with pytest.raises(RuntimeError, match="full rank"):
        g(w.propagator(inv.preconditioned(pc.eo2_ne(), inv.block_cg(eps=1e-10, maxiter=100)))
          * pack([_dev(srcs[0], g.grid(FD, g.double)), _dev(z, g.grid(FD, g.double))]))


@pytest.mark.parametrize("N", [12, 20, 40])
def test_block_cg_on_matrix_cores_matches_oracle(N):
    """FP64 blocks of 8..48 right-hand sides run the block recursion's Gram
    matrices and field updates on the matrix cores (v_mfma_f64_16x16x4f64,
    1, 2 and 3 tiles of 16 per dimension, the last ones padded): the same
    iteration count as the oracle's block CG within +-1, the same solutions
    to 1e-9, and each true residual <= 1e-8 through the oracle's M"""
    gr = R.grid(FD)
    rng = R.random(f"block_cg_mfma_{N}")
    U = G.random_links(gr, rng, scale=0.5)
    srcs = [rng.vspincolor(gr) for _ in range(N)]
    eps = 1e-10
    ref = W.wilson_clover(U, **P)
    opc = S.eo2_ne(ref, parity=1)
    Xr, it_ref, conv_ref = S.block_cg(opc.Mpc_ne, [opc.R_ne(s) for s in srcs], eps=eps, maxiter=500)
    x_ref = [opc.L(x) + opc.S(s) for x, s in zip(Xr, srcs)]
    w = g.qcd.fermion.wilson_clover(U, P, n_rhs=N)
    bcg = inv.block_cg(eps=eps, maxiter=500)
    X = g(w.propagator(inv.preconditioned(pc.eo2_ne(), bcg)) * pack([_dev(s, g.grid(FD, g.double)) for s in srcs]))
    print(f"block CG N={N} on matrix cores: GPU {bcg.iterations} iterations, oracle {it_ref}", flush=True)
    assert conv_ref and bcg.converged and abs(bcg.iterations - it_ref) <= 1
    for a, b, s in zip(unpack(X), x_ref, srcs):
        a = a.to_numpy()
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-9
        assert np.linalg.norm(ref.M(a) - s) / np.linalg.norm(s) < 1e-8
