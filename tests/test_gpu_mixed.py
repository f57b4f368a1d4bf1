"""Mixed-precision reliable-update solve (FP32 inner CG, FP64 solution and
true residuals), with the inner state in FP32 or in 16-bit storage (FP16 or BF16): the true
residual of the full system, recomputed here with the oracle's FP64
operator, meets the requested tolerance; the solution agrees with the FP64
CG one to 1e-8; a right-hand side gives the same result alone and inside a
batch (to 1e-8: both are solves to 1e-10)."""
import numpy as np
import pytest

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
P = dict(mass=0.05, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])


@pytest.fixture(scope="module")
def setup():
    fd = [8, 8, 8, 16]
    U = G.smooth_links(fd, 0.3, 3001)
    rng = np.random.default_rng(3002)
    shape = tuple(reversed(fd))
    srcs = [(rng.standard_normal(shape + (4, 3)) + 1j * rng.standard_normal(shape + (4, 3))) / np.sqrt(2)
            for _ in range(3)]
    return fd, U, srcs


def _dev(a, grid):
    x = g.lattice(grid)
    x.from_numpy(a)
    return x


STORAGE = [None, "half", "bf16"]


@pytest.mark.parametrize("storage", STORAGE)
def test_true_residual_and_fp64_agreement(setup, storage):
    fd, U, srcs = setup
    ref = W.wilson_clover(U, **P)
    w = g.qcd.fermion.wilson_clover(U, P)
    tol = 1e-10
    mcg = inv.cg_mixed(eps=tol, maxiter=5000, inner_storage=storage)
    slv = w.propagator(inv.preconditioned(pc.eo2_ne(), mcg))
    x = g(slv * _dev(srcs[0], w.F_grid)).to_numpy()
    assert mcg.converged and mcg.reliable_updates >= 2
    true = np.linalg.norm(ref.M(x) - srcs[0]) / np.linalg.norm(srcs[0])
    assert true <= 1.05 * tol, true
    # the solver's own FP64 true residual is the same quantity
    assert abs(mcg.true_residual[0] - true) <= 0.05 * true + 1e-14
    cg64 = inv.cg(eps=1e-12, maxiter=5000)
    x64 = g(w.propagator(inv.preconditioned(pc.eo2_ne(), cg64)) * _dev(srcs[0], w.F_grid)).to_numpy()
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) < 1e-8


@pytest.mark.parametrize("storage", STORAGE)
def test_alone_equals_in_batch(setup, storage):
    fd, U, srcs = setup
    ref = W.wilson_clover(U, **P)
    tol = 1e-10
    w1 = g.qcd.fermion.wilson_clover(U, P)
    w3 = g.qcd.fermion.wilson_clover(U, P, n_rhs=3)
    m1 = inv.cg_mixed(eps=tol, maxiter=5000, inner_storage=storage)
    s1 = w1.propagator(inv.preconditioned(pc.eo2_ne(), m1))
    alone = [g(s1 * _dev(s, w1.F_grid)).to_numpy() for s in srcs]
    m3 = inv.cg_mixed(eps=tol, maxiter=5000, inner_storage=storage)
    s3 = w3.propagator(inv.preconditioned(pc.eo2_ne(), m3))
    batch = unpack(g(s3 * pack([_dev(s, w1.F_grid) for s in srcs])))
    assert m3.converged
    for i, s in enumerate(srcs):
        b = batch[i].to_numpy()
        true = np.linalg.norm(ref.M(b) - s) / np.linalg.norm(s)
        assert true <= 1.05 * tol
        # both meet the tolerance; they agree to the tolerance-level accuracy
        assert np.linalg.norm(b - alone[i]) / np.linalg.norm(alone[i]) < 1e-8


@pytest.mark.parametrize("storage", ["half", "bf16"])
def test_half_storage_point_source_and_errors(setup, storage):
    """a point source (wide dynamic range across sites) converges with 16-bit
    inner storage; 16-bit storage needs the FP32 inner operator"""
    fd, U, srcs = setup
    ref = W.wilson_clover(U, **P)
    w = g.qcd.fermion.wilson_clover(U, P)
    pt = np.zeros_like(srcs[0])
    pt[0, 0, 0, 0, 0, 0] = 1.0
    mcg = inv.cg_mixed(eps=1e-10, maxiter=5000, inner_storage=storage)
    x = g(w.propagator(inv.preconditioned(pc.eo2_ne(), mcg)) * _dev(pt, w.F_grid)).to_numpy()
    assert mcg.converged
    assert np.linalg.norm(ref.M(x) - pt) / np.linalg.norm(pt) <= 1.05e-10
    with pytest.raises(ValueError):
        inv.cg_mixed(eps=1e-10, inner_precision=g.double, inner_storage=storage)
    with pytest.raises(ValueError):
        inv.cg_mixed(eps=1e-10, inner_storage="fp8")


def test_16bit_inner_operator_tracks_fp32(setup):
    """the 16-bit inner state (FP16 or BF16 storage, FP16 spin projection)
    changes the inner operator only at its rounding level: the solve needs
    about the iterations of FP32 storage, and every reliable update finds the
    true residual close to what the recursion predicted (a wrong 16-bit hop or
    update would show as extra iterations or updates)"""
    fd, U, srcs = setup
    w = g.qcd.fermion.wilson_clover(U, P)
    src = _dev(srcs[1], w.F_grid)
    runs = {}
    for storage in (None, "half", "bf16"):
        m = inv.cg_mixed(eps=1e-10, maxiter=5000, inner_storage=storage)
        g(w.propagator(inv.preconditioned(pc.eo2_ne(), m)) * src)
        assert m.converged
        runs[storage] = (m.iterations[0], m.reliable_updates)
        print(f"storage {storage}: iterations {m.iterations[0]}, reliable updates {m.reliable_updates}", flush=True)
    # 8^3x16, m 0.05: FP32 72 iterations / 6 updates, FP16 73 / 6, BF16 73 / 11
    # (BF16 updates at least every 0.1 of residual reduction, the others every 0.01)
    it32, up32 = runs[None]
    for storage, max_up in (("half", up32 + 2), ("bf16", 2 * up32 + 2)):
        it, up = runs[storage]
        assert it <= it32 + 3, (storage, it, it32)
        assert up <= max_up, (storage, up, up32)
