"""The multi-RHS configuration (one smoothed field, mass 0.08, csw 1,
complex-Gaussian sources) at the benchmarked right-hand-side counts 12 and 48,
on an 8^3 x 16 lattice the oracle handles in seconds:

* the packed n_rhs operator equals the single-RHS one per right-hand side to
  the reference's criterion eps^2 < 1e-15 (tests/qcd/fermion_operators.py:97-132);
* the batched FP64 eo2_ne CG gives every right-hand side the iteration count
  of the oracle's C restatement within +-1 and its solution;
* the 16-bit (FP16 storage) mixed multi-RHS solve reaches a true relative
  residual <= 1.05 tol for every right-hand side, computed with the oracle's
  FP64 operator M.
"""
import numpy as np
import pytest

from oracle import cpu as C
from oracle import gauge as G
from oracle import wilson_clover as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd.qcd.fermion.operator import pack, unpack  # noqa: E402

inv = g.algorithms.inverter
pc = g.qcd.fermion.preconditioner
FD = [8, 8, 8, 16]
P = dict(mass=0.08, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0, boundary_phases=[1, 1, 1, -1])


@pytest.fixture(scope="module")
def c4():
    U = G.smooth_links(FD, 0.3, 4001)
    shape = tuple(reversed(FD)) + (4, 3)
    srcs = []
    for j in range(48):
        r = np.random.default_rng(4100 + j)
        srcs.append((r.standard_normal(shape) + 1j * r.standard_normal(shape)) / np.sqrt(2.0))
    ref = W.wilson_clover(U, **P)
    return U, srcs, ref


def _fields(arrs, grid):
    out = []
    for a in arrs:
        x = g.lattice(grid)
        x.from_numpy(a)
        out.append(x)
    return out


@pytest.mark.parametrize("N", [12, 48])
def test_packed_operator_n12_n48(c4, N):
    U, srcs, ref = c4
    w1 = g.qcd.fermion.wilson_clover(U, P)
    wN = g.qcd.fermion.wilson_clover(U, P, n_rhs=N)
    test = _fields(srcs[:N], w1.F_grid)
    one = [w1(t) for t in test]
    many = wN.packed()(test)
    for i in range(N):
        eps2 = g.norm2(g(one[i] - many[i])) / g.norm2(one[i])
        assert eps2 < 1e-15, (N, i, eps2)
    # the even-odd pieces and the normal operator on the n_rhs field directly
    half = [g.pick_checkerboard(g.odd, t) for t in test]
    H = pack(half)
    for name in ["Meooe", "Mooee"]:
        a = [getattr(w1, name)(h) for h in half]
        b = unpack(getattr(wN, name)(H))
        for i in range(N):
            assert np.abs(a[i].to_numpy() - b[i].to_numpy()).max() <= 1e-14 * np.abs(a[i].to_numpy()).max()
    m1, mN = pc.eo2_ne()(w1).Mpc, pc.eo2_ne()(wN).Mpc
    a = [m1(h) for h in half]
    b = unpack(mN(H))
    for i in range(N):
        assert torch.equal(a[i].data, b[i].data) or (
            np.abs(a[i].to_numpy() - b[i].to_numpy()).max() <= 1e-14 * np.abs(a[i].to_numpy()).max())
    w1.release()
    wN.release()


@pytest.mark.parametrize("N", [12, 48])
def test_batched_fp64_cg_against_c_oracle(c4, N):
    U, srcs, ref = c4
    eps = 1e-10
    wN = g.qcd.fermion.wilson_clover(U, P, n_rhs=N)
    cgN = inv.cg(eps=eps, maxiter=5000)
    X = unpack(g(wN.propagator(inv.preconditioned(pc.eo2_ne(), cgN)) * pack(_fields(srcs[:N], g.grid(FD, g.double)))))
    assert cgN.converged
    cpu = C.wilson_clover_cpu(ref)
    for i in range(N):
        xc, itc = cpu.solve(srcs[i], pd=1, eps=eps, maxiter=5000)
        assert abs(cgN.iterations[i] - itc) <= 1, (i, cgN.iterations[i], itc)
        x = X[i].to_numpy()
        assert np.linalg.norm(x - xc) / np.linalg.norm(xc) < 1e-8, i
    wN.release()


@pytest.mark.parametrize("N", [12, 48])
def test_fp16_mixed_multi_rhs_true_residual(c4, N):
    U, srcs, ref = c4
    tol = 1e-10
    wN = g.qcd.fermion.wilson_clover(U, P, n_rhs=N)
    mh = inv.cg_mixed(eps=tol, maxiter=5000, inner_storage="half")
    X = unpack(g(wN.propagator(inv.preconditioned(pc.eo2_ne(), mh)) * pack(_fields(srcs[:N], g.grid(FD, g.double)))))
    assert mh.converged
    print(f"c4 fp16 mixed N={N}: iterations {max(mh.iterations)}, reliable updates {mh.reliable_updates}",
          flush=True)
    for i in range(N):
        x = X[i].to_numpy()
        true = np.linalg.norm(ref.M(x) - srcs[i]) / np.linalg.norm(srcs[i])
        assert true <= 1.05 * tol, (i, true)
    mh.inner_operator(wN).release()
    wN.release()
