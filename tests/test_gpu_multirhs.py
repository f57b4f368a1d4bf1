"""Multiple right-hand sides (gpt's n_rhs operator, lib/gpt/qcd/fermion/wilson.py:82-86)
against the single-RHS path: the packed operator equals the single-RHS one per
RHS (tests/qcd/fermion_operators.py:97-132, eps^2 < 1e-15), and the batched
CG gives each RHS the iteration count and solution of its own solve."""
import numpy as np
import pytest

from oracle import gauge as G
from oracle import rng as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402

inv = g.algorithms.inverter
pc = g.qcd.fermion.preconditioner
P = dict(mass=0.123, csw_r=0.6, csw_t=0.6, cF=1.0, xi_0=1.0, nu=1.0, isAnisotropic=False,
         boundary_phases=[1, 1, 1, -1])


@pytest.fixture(scope="module")
def setup():
    gr = R.grid([8, 8, 8, 16])
    rng = R.random("multi_rhs")
    U = G.random_links(gr, rng, scale=0.5)
    srcs = [rng.vspincolor(gr) for _ in range(5)]
    return gr, U, srcs


def _fields(arrs, grid):
    out = []
    for a in arrs:
        x = g.lattice(grid)
        x.from_numpy(a)
        out.append(x)
    return out


@pytest.mark.parametrize("prec", ["double", "single"])
def test_packed_operator_equals_single_rhs(setup, prec):
    gr, U, srcs = setup
    pr = g.double if prec == "double" else g.single
    w1 = g.qcd.fermion.wilson_clover(U, P, precision=pr)
    w4 = g.qcd.fermion.wilson_clover(U, P, precision=pr, n_rhs=4)
    assert w4.F_grid.nrhs == 4
    test = _fields([s.astype(pr.complex_dtype) for s in srcs[:4]], w1.F_grid)
    test0 = [w1(t) for t in test]
    test1 = w4.packed()(test)
    tol = 1e-15 if prec == "double" else 1e-12
    for i in range(4):
        eps2 = g.norm2(g(test0[i] - test1[i])) / g.norm2(test0[i])
        assert eps2 < tol, (i, eps2)
    # the even-odd pieces on an n_rhs field directly
    from gpt_amd.qcd.fermion.operator import pack, unpack

    half = [g.pick_checkerboard(g.odd, t) for t in test]
    H = pack(half)
    for name in ["Meooe", "Mooee"]:
        op1, op4 = getattr(w1, name), getattr(w4, name)
        one = [op1(h) for h in half]
        many = unpack(op4(H))
        for i in range(4):
            d = np.abs(one[i].to_numpy() - many[i].to_numpy()).max()
            assert d <= (1e-14 if prec == "double" else 1e-5) * np.abs(one[i].to_numpy()).max(), (name, i)
    inv_one = [w1.Mooee.inv()(h) for h in half]
    inv_many = unpack(w4.Mooee.inv()(H))
    for i in range(4):
        assert np.abs(inv_one[i].to_numpy() - inv_many[i].to_numpy()).max() <= 1e-14 * np.abs(
            inv_one[i].to_numpy()).max() + (0 if prec == "double" else 1e-6)


def test_batched_cg_matches_single_solves(setup):
    gr, U, srcs = setup
    N = 5
    w1 = g.qcd.fermion.wilson_clover(U, P)
    wN = g.qcd.fermion.wilson_clover(U, P, n_rhs=N)
    arrs = [s.copy() for s in srcs]
    arrs[2][:] = 0.0  # a zero source inside the batch: its solution is zero, as gpt's cg gives
    single = _fields(arrs, w1.F_grid)
    cg1 = inv.cg(eps=1e-10, maxiter=2000)
    slv1 = w1.propagator(inv.preconditioned(pc.eo2_ne(), cg1))
    x1, it1 = [], []
    for s in single:
        x1.append(g(slv1 * s))
        it1.append(len(cg1.history))
    from gpt_amd.qcd.fermion.operator import pack, unpack

    cgN = inv.cg(eps=1e-10, maxiter=2000)
    slvN = wN.propagator(inv.preconditioned(pc.eo2_ne(), cgN))
    XN = unpack(g(slvN * pack(single)))
    assert cgN.converged
    for i in range(N):
        assert abs(cgN.iterations[i] - it1[i]) <= 1, (i, cgN.iterations, it1)
        a, b = XN[i].to_numpy(), x1[i].to_numpy()
        if i == 2:
            assert np.abs(a).max() == 0.0 and cgN.iterations[i] == 0
        else:
            assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-10


def test_every_hop_shape_gives_identical_bits(setup):
    """Launch shapes (right-hand sides per thread, threads per block) change
    only the schedule: the hop, the fused normal operator with its
    inner-product partials and a whole CG agree bit for bit."""
    from gpt_amd.qcd.fermion.operator import pack

    gr, U, srcs = setup
    N = 5  # not a multiple of any G * rb > 1: padded right-hand sides included
    for pr in (g.double, g.single):
        w = g.qcd.fermion.wilson_clover(U, P, precision=pr, n_rhs=N)
        w1 = g.qcd.fermion.wilson_clover(U, P, precision=pr)
        single = _fields([s.astype(pr.complex_dtype) for s in srcs[:N]], w1.F_grid)
        H = pack([g.pick_checkerboard(g.odd, x) for x in single])
        mpc = pc.eo2_ne()(w).Mpc
        rbs = (1, 2) if pr is g.single else (1,)
        shapes = [(rb, tpb, occ, seg, ls) for rb in rbs for tpb in (1, 3, 8) for occ in (0, 4) for seg in (1, 2, 4)
                  for ls in (0, 1)]
        ref = None
        staged = 0
        for sh in shapes:
            try:
                w.set_hop_shape(sh)
            except RuntimeError as e:  # a staging area above the LDS limit
                assert sh[4] == 1 and "staging area" in str(e), (sh, str(e))
                continue
            staged += sh[4]
            assert w.hop_shape() == sh
            out = (w.Meooe(H).data.clone(), mpc(H).data.clone())
            if ref is None:
                ref = out
            else:
                assert torch.equal(out[0], ref[0]) and torch.equal(out[1], ref[1]), (pr.__name__, sh)
        assert staged >= 6
        # a CG (hop partials feed alpha) is identical under two shapes
        res = []
        for sh in [(1, 1, 0, 1, 0), (rbs[-1], 3, 8, 4, 1)]:
            w.set_hop_shape(sh)
            cgN = inv.cg(eps=1e-6 if pr is g.single else 1e-10, maxiter=500)
            X = g(w.propagator(inv.preconditioned(pc.eo2_ne(), cgN)) * pack(single))
            res.append((X.data.clone(), list(cgN.iterations)))
        assert res[0][1] == res[1][1] and torch.equal(res[0][0], res[1][0])
        w.set_hop_shape(None)
        with pytest.raises(RuntimeError, match="invalid"):
            w.set_hop_shape((3, 1, 0, 1, 0))
        with pytest.raises(RuntimeError, match="invalid"):
            w.set_hop_shape((1, 9, 0, 1, 0))
        with pytest.raises(RuntimeError, match="invalid"):
            w.set_hop_shape((1, 1, 0, 5, 0))  # 8^3 x 16: 4 blocks per time slice
        with pytest.raises(RuntimeError, match="invalid"):
            w.set_hop_shape((1, 1, 0, 1, 3))


@pytest.mark.parametrize("fd", [[8, 8, 8, 16], [12, 8, 8, 8]])
def test_single_rhs_spinor_staging_gives_identical_bits(fd):
    """ls = 2 (one right-hand side: the workgroup's own input spinors in LDS,
    the +-x and most +-y gathers read from there) changes only where values
    are read from: the hop, its adjoint, the fused normal operator and a CG
    agree bit for bit with ls = 0 for every workgroup size.  Lx = 12 puts
    rows of 6 half-sites across the 64-site blocks, so some x and y
    neighbours of a workgroup's sites lie outside its range."""
    rng = R.random("staging")
    U = G.random_links(R.grid(fd), rng, scale=0.5)
    src = rng.vspincolor(R.grid(fd))
    for pr in (g.double, g.single):
        for compress in (False, True):
            w = g.qcd.fermion.wilson_clover(U, P, precision=pr, compress_links=compress)
            s = g.lattice(w.F_grid)
            s.from_numpy(src.astype(pr.complex_dtype))
            H = g.pick_checkerboard(g.odd, s)
            pre = pc.eo2_ne()(w)
            ref = None
            for tpb in (1, 2, 3, 4, 8):
                for occ in (0, 4):
                    for ls in (0, 2):
                        try:
                            w.set_hop_shape((1, tpb, occ, 1, ls))
                        except RuntimeError as e:  # FP64 at 512 threads: 96 KB of staging
                            assert ls == 2 and pr is g.double and tpb == 8 and "staging area" in str(e)
                            continue
                        out = [w.Meooe(H).data.clone(), w.Meooe.adj()(H).data.clone(),
                               pre.Mpc(H).data.clone(), pre.Mpc.adj()(H).data.clone()]
                        if ref is None:
                            ref = out
                        else:
                            for a, b in zip(out, ref):
                                assert torch.equal(a, b), (pr.__name__, compress, tpb, occ, ls)
            res = []
            for sh in [(1, 1, 0, 1, 0), (1, 2, 4, 1, 2)]:
                w.set_hop_shape(sh)
                cg1 = inv.cg(eps=1e-5 if pr is g.single else 1e-10, maxiter=500)
                X = g(w.propagator(inv.preconditioned(pc.eo2_ne(), cg1)) * s)
                res.append((X.data.clone(), cg1.iterations))
            assert res[0][1] == res[1][1] and torch.equal(res[0][0], res[1][0])
            w.set_hop_shape(None)


def test_spinor_staging_shape_on_a_16bit_inner_operator():
    """An ls = 2 shape on the FP32 inner operator of the mixed solve: its
    16-bit hops do not stage (the shape's other fields still apply) and the
    solve is the same bit for bit."""
    fd = [8, 8, 8, 16]
    U = G.smooth_links(fd, 0.3, 3001)
    rng = np.random.default_rng(3002)
    a = rng.standard_normal(tuple(reversed(fd)) + (4, 3)) + 1j * rng.standard_normal(tuple(reversed(fd)) + (4, 3))
    Pm = dict(mass=0.05, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])
    w = g.qcd.fermion.wilson_clover(U, Pm)
    s = g.lattice(w.F_grid)
    s.from_numpy(a)
    out = []
    for ls in (0, 2):
        for storage in ("half", None):
            mcg = inv.cg_mixed(eps=1e-10, maxiter=5000, inner_storage=storage)
            mcg.inner_operator(w).set_hop_shape((1, 2, 4, 1, ls), n_rhs=1)
            X = g(w.propagator(inv.preconditioned(pc.eo2_ne(), mcg)) * s)
            assert mcg.converged
            out.append((X.data.clone(), mcg.iterations, mcg.reliable_updates))
    for k in (0, 1):
        assert torch.equal(out[k][0], out[k + 2][0]) and out[k][1:] == out[k + 2][1:]


def test_tuner_picks_a_valid_shape(setup):
    gr, U, srcs = setup
    w = g.qcd.fermion.wilson_clover(U, P, precision=g.single, n_rhs=4)
    sh, ms = w.tune_hop(reps=2)
    assert ms > 0 and sh[0] in (1, 2) and sh[1] in (1, 2, 4, 8) and sh[2] in (0, 2, 4, 8, 16) and sh[3] == 1
    assert sh[4] in (0, 1)
    assert w.hop_shape() == sh
