"""Solvers of the oracle: the reference's correlator KAT
(tests/qcd/fermion_operators.py:150-218), brute force on tiny lattices, and
the C restatement against the numpy restatement."""
import cmath

import numpy as np
import pytest

from oracle import conventions as cv
from oracle import cpu as C
from oracle import gauge as G
from oracle import rng as R
from oracle import solvers as S
from oracle import wilson_clover as W


def _dense(fn, shape):
    n = int(np.prod(shape)) * 12
    M = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1.0
        M[:, j] = fn(e.reshape(tuple(shape) + (4, 3))).reshape(-1)
    return M


@pytest.fixture(scope="module")
def tiny_op():
    g = R.grid([2, 2, 2, 4])
    rng = R.random("tiny")
    U = G.random_links(g, rng, scale=0.7)
    return W.wilson_clover(U, mass=0.2, csw_r=1.0, csw_t=1.2, boundary_phases=[1, 1, 1, -1]), g


def test_eo2_ne_solve_matches_dense_inverse(tiny_op):
    w, g = tiny_op
    M = _dense(w.M, g.shape)
    b = R.random("rhs").vspincolor(g)
    x_exact = np.linalg.solve(M, b.reshape(-1)).reshape(b.shape)
    x, it, conv = S.preconditioned(w, b, eps=1e-13, maxiter=500)
    assert conv
    assert np.linalg.norm(x - x_exact) / np.linalg.norm(x_exact) < 1e-11
    # the C restatement gives the same solution
    xc, itc = C.wilson_clover_cpu(w).solve(b, eps=1e-13, maxiter=500)
    assert np.linalg.norm(xc - x_exact) / np.linalg.norm(x_exact) < 1e-11
    assert abs(itc - it) <= 1


def test_schur_determinant_identity(tiny_op):
    """det M = det DD det CC det Mpc with D = odd (schur_complement_two.py:24-36)."""
    w, g = tiny_op
    pc = S.eo2_ne(w, parity=1)
    par = cv.parity(g.shape).reshape(-1)
    sel = {p: np.repeat(par == p, 12) for p in (0, 1)}
    M = _dense(w.M, g.shape)
    Mpc = _dense(pc.Mpc, g.shape)[np.ix_(sel[1], sel[1])]
    DD = M[np.ix_(sel[1], sel[1])]
    CC = M[np.ix_(sel[0], sel[0])]
    lhs = np.linalg.slogdet(M)
    rhs = [np.linalg.slogdet(a) for a in (DD, CC, Mpc)]
    assert abs(lhs[1] - sum(r[1] for r in rhs)) < 1e-10
    assert abs(lhs[0] - np.prod([r[0] for r in rhs])) < 1e-10


def test_cg_recursion_numpy_vs_c():
    g = R.grid([4, 4, 4, 8])
    rng = R.random("cg")
    U = G.random_links(g, rng, scale=0.5)
    w = W.wilson_clover(U, mass=0.1, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])
    pc = S.eo2_ne(w, parity=1)
    b = pc.R_ne(rng.vspincolor(g))
    hist = []
    x, it, conv = S.cg(pc.Mpc_ne, b, eps=1e-10, maxiter=1000, history=hist)
    wc = C.wilson_clover_cpu(w)
    xc, itc, hc = wc.cg_ne(C.half(b, 1), 1, eps=1e-10, maxiter=1000, history=True)
    assert conv and itc == it
    assert np.allclose(hc, hist, rtol=1e-8)
    assert np.linalg.norm(C.half(x, 1) - xc) / np.linalg.norm(xc) < 1e-10


@pytest.mark.parametrize("case", ["aniso", "open", "twisted"])
def test_c_operator_matches_numpy_per_site(case):
    g = R.grid([4, 4, 4, 8])
    rng = R.random("percase")
    U = G.random_links(g, rng)
    src = rng.vspincolor(g)
    params = {
        "aniso": dict(kappa=0.135, csw_r=1.5, csw_t=1.951, xi_0=1.33111, nu=2.61,
                      boundary_phases=[1.0, -1.0, 1.0, -1.0]),
        "open": dict(kappa=0.135, csw_r=1.978, csw_t=1.978, cF=1.3, boundary_phases=[1, 1, 1, 0.0]),
        "twisted": dict(kappa=0.137, boundary_phases=[cmath.exp(1j * k) for k in (1, 2, 3, 4)]),
    }[case]
    w = W.wilson_clover(U, **params)
    wc = C.wilson_clover_cpu(w)
    for pin in (0, 1):
        for dag in (False, True):
            ref = C.half((w.MeooeDag if dag else w.Meooe)(src, pin), 1 - pin)
            got = wc.meooe(C.half(src, pin), pin, dag)
            err = np.abs(got - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2)).clip(1e-300)
            assert err[np.abs(ref).max(axis=(1, 2)) > 0].max() < 1e-13
        for inv in (False, True):
            ref = C.half((w.MooeeInv if inv else w.Mooee)(src, pin), pin)
            got = wc.mooee(C.half(src, pin), pin, inv)
            assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def test_correlator_known_answer(kats):
    k = kats["correlator"]
    g = R.grid(k["fdims"])
    rng = R.random(k["seed"])
    U = [u.astype(np.complex64).astype(np.complex128) for u in G.random_links(g, rng)]
    p = k["params"]
    w = W.wilson_clover(U, kappa=p["kappa"], csw_r=0.0, csw_t=0.0, xi_0=1.0, nu=1.0,
                        boundary_phases=[cmath.exp(1j * a) for a in p["boundary_phase_angles"]])
    wc = C.wilson_clover_cpu(w)
    x0, y0, z0, t0 = k["point_source_xyzt"]
    corr = np.zeros(g.shape[0])
    for s in range(4):
        for c in range(3):
            src = np.zeros(g.shape + (4, 3), complex)
            src[t0, z0, y0, x0, s, c] = 1.0
            dst, it = wc.solve(src, eps=1e-6, maxiter=1000)
            assert it > 0
            corr += (np.abs(dst) ** 2).sum(axis=(1, 2, 3, 4, 5))
    assert np.linalg.norm(corr - np.array(k["reference"])) < k["tol_norm"]


def test_block_cg_pinned(tiny_op):
    """block_cg of the oracle: one right-hand side is CG (same iterations +-1,
    same solution); on the tiny lattice four right-hand sides match the dense
    solve of the normal operator, within fewer iterations than CG needs for
    any of them (block Krylov space)."""
    w, g = tiny_op
    pc = S.eo2_ne(w, parity=1)
    par = cv.parity(g.shape).reshape(-1)
    selD = np.repeat(par == 1, 12)
    Mne = _dense(pc.Mpc_ne, g.shape)[np.ix_(selD, selD)]
    rng = R.random("block")
    srcs = [pc.R_ne(rng.vspincolor(g)) for _ in range(4)]
    x1, it1, conv1 = S.cg(pc.Mpc_ne, srcs[0], eps=1e-12, maxiter=500)
    (xb,), itb, convb = S.block_cg(pc.Mpc_ne, srcs[:1], eps=1e-12, maxiter=500)
    assert conv1 and convb and abs(it1 - itb) <= 1
    assert np.linalg.norm(xb - x1) / np.linalg.norm(x1) < 1e-10
    hist = []
    X, it4, conv4 = S.block_cg(pc.Mpc_ne, srcs, eps=1e-12, maxiter=500, history=hist)
    assert conv4 and len(hist) == it4 and hist[-1] <= 1e-24
    its = []
    for s_, x in zip(srcs, X):
        exact = np.linalg.solve(Mne, s_.reshape(-1)[selD])
        assert np.linalg.norm(x.reshape(-1)[selD] - exact) / np.linalg.norm(exact) < 1e-10
        assert np.abs(x.reshape(-1)[~selD]).max() == 0.0
        its.append(S.cg(pc.Mpc_ne, s_, eps=1e-12, maxiter=500)[1])
    assert it4 < min(its)
    # a zero block is solved at once
    Z, itz, convz = S.block_cg(pc.Mpc_ne, [np.zeros_like(srcs[0])] * 2, eps=1e-12)
    assert convz and itz == 0 and all(np.abs(z).max() == 0 for z in Z)


def test_eo1_ne_and_eo2_kappa_ne_pinned(tiny_op):
    """eo1_ne (Schur complement one, normal equation) and eo2_kappa_ne of the
    oracle: det M = det DD det CC det N for N = 1 - DD^-1 DC CC^-1 CD, both
    preconditioned solves equal the dense solve, and the kappa-scaled eo2
    needs the iterations of eo2_ne (+-1: only the scale of its recursion
    differs)."""
    w, g = tiny_op
    par = cv.parity(g.shape).reshape(-1)
    sel = {p: np.repeat(par == p, 12) for p in (0, 1)}
    M = _dense(w.M, g.shape)
    N1 = _dense(S.eo1_ne(w, parity=1).Mpc, g.shape)[np.ix_(sel[1], sel[1])]
    lhs = np.linalg.slogdet(M)
    rhs = [np.linalg.slogdet(a) for a in (M[np.ix_(sel[1], sel[1])], M[np.ix_(sel[0], sel[0])], N1)]
    assert abs(lhs[1] - sum(r[1] for r in rhs)) < 1e-10
    b = R.random("rhs1").vspincolor(g)
    x_exact = np.linalg.solve(M, b.reshape(-1)).reshape(b.shape)
    its = {}
    for name, cls in [("eo1_ne", S.eo1_ne), ("eo2_ne", S.eo2_ne), ("eo2_kappa_ne", S.eo2_kappa_ne)]:
        x, it, conv = S.preconditioned(w, b, pc_class=cls, eps=1e-13, maxiter=500)
        assert conv, name
        assert np.linalg.norm(x - x_exact) / np.linalg.norm(x_exact) < 1e-11, name
        its[name] = it
    assert abs(its["eo2_kappa_ne"] - its["eo2_ne"]) <= 1


def _dense_small():
    g = R.grid([2, 2, 2, 4])
    U = G.random_links(g, R.random("dc"), scale=0.7)
    w = W.wilson_clover(U, mass=0.2, csw_r=1.0, csw_t=1.2, boundary_phases=[1, 1, 1, -1])
    M = _dense(w.M, g.shape)
    b = R.random("dc_rhs").vspincolor(g)
    return g, M, b


def _mat(M, shape):
    return lambda x: (M @ x.reshape(-1)).reshape(x.shape)


def test_defect_correcting_pinned():
    """defect_correcting (lib/gpt/algorithms/inverter/defect_correcting.py:74-142)
    pinned by closed forms: an exact inner inverse converges in one outer
    step; an inner inverse that returns half the exact correction leaves a
    true residual of exactly 2^-i after i steps, so the loop stops at the first
    i with 2^-i < eps; a zero source returns the zero guess at once."""
    g, M, b = _dense_small()
    A = _mat(M, g.shape)
    x_exact = np.linalg.solve(M, b.reshape(-1)).reshape(b.shape)
    exact = lambda s: np.linalg.solve(M, s.reshape(-1)).reshape(s.shape)  # noqa: E731
    x, it, conv = S.defect_correcting(A, exact, b, eps=1e-12)
    assert conv and it == 1
    assert np.linalg.norm(x - x_exact) / np.linalg.norm(x_exact) < 1e-13
    x, it, conv = S.defect_correcting(A, lambda s: 0.5 * exact(s), b, eps=1e-6)
    assert conv and it == 20  # 2^-19 = 1.9e-6 >= 1e-6 > 2^-20
    res = np.linalg.norm(b - A(x)) / np.linalg.norm(b)
    assert abs(res / 2.0**-20 - 1.0) < 1e-6
    x, it, conv = S.defect_correcting(A, exact, np.zeros_like(b), eps=1e-12)
    assert conv and it == 0 and np.abs(x).max() == 0.0
    x, it, conv = S.defect_correcting(A, lambda s: 0.5 * exact(s), b, eps=1e-12, maxiter=5)
    assert not conv and it == 5


def _single_precision_inner(M):
    M32 = M.astype(np.complex64)
    return lambda s: np.linalg.solve(M32, s.reshape(-1).astype(np.complex64)).astype(np.complex128).reshape(s.shape)


def test_defect_correcting_normalises_the_inner_source():
    """the defect is divided by |src| before the inner solve and the
    correction multiplied back (defect_correcting.py:124-134): with a source
    of magnitude 1e-45, below the single-precision range, a single-precision
    inner solve still converges to 1e-12 in a few outer steps"""
    g, M, b = _dense_small()
    A = _mat(M, g.shape)
    tiny = b * 1e-45
    x, it, conv = S.defect_correcting(A, _single_precision_inner(M), tiny, eps=1e-12)
    assert conv and 2 <= it <= 6
    assert np.linalg.norm(tiny - A(x)) / np.linalg.norm(tiny) < 1e-12


def test_defect_correcting_planted_normalisation_mistakes_fail():
    """the normalisation test has teeth: the same loop with the division by
    |src| dropped, or the multiplication back dropped, does not converge"""
    import inspect
    import textwrap

    src_code = textwrap.dedent(inspect.getsource(S.defect_correcting))
    g, M, b = _dense_small()
    A = _mat(M, g.shape)
    tiny = b * 1e-45
    for old, new in [("s = s / n2src**0.5", "s = s"),
                     ("inner_inv(s) * n2src**0.5", "inner_inv(s)")]:
        assert old in src_code
        ns = dict(S.__dict__)
        exec(src_code.replace(old, new), ns)
        with np.errstate(all="ignore"):  # the broken loops overflow single precision
            x, it, conv = ns["defect_correcting"](A, _single_precision_inner(M), tiny, eps=1e-12, maxiter=20)
        assert not conv, old
