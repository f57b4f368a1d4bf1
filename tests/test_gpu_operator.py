"""Parity of the HIP operator (through the C ABI) with the oracle.

FP64 bar: <= 1e-13 relative per site against the numpy oracle (whose
operator reproduces Grid's fingerprints); the fingerprints themselves to the
reference's tolerance 1e-13 (tests/qcd/fermion_operators.py:527,906-908).
"""
import cmath

import numpy as np
import pytest

from oracle import conventions as cv
from oracle import gauge as G
from oracle import rng as R
from oracle import wilson_clover as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - collected on CPU only with -m gpu
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402


def site_rel_err(got, ref):
    """max over sites of |got - ref|_site / |ref|_site (sites with ref = 0 compared absolutely)."""
    d = np.abs(got - ref).reshape(got.shape[:4] + (-1,))
    r = np.abs(ref).reshape(ref.shape[:4] + (-1,))
    dn = np.sqrt((d**2).sum(-1))
    rn = np.sqrt((r**2).sum(-1))
    scale = np.where(rn > 0, rn, 1.0)
    return float((dn / scale).max())


def to_dev(arr, grid):
    x = g.lattice(grid)
    x.from_numpy(arr)
    return x


def half_dev(arr, grid_eo, cb):
    x = g.lattice(grid_eo)
    x.checkerboard(cb)
    x.from_numpy(arr)
    return x


CASES = {
    "baseline_clover": dict(mass=0.1, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1]),
    "aniso_clover": dict(kappa=0.135, csw_r=1.5, csw_t=1.951, xi_0=1.33111, nu=2.61,
                         isAnisotropic=True, boundary_phases=[1.0, -1.0, 1.0, -1.0]),
    "open_clover": dict(kappa=0.135, csw_r=1.978, csw_t=1.978, cF=1.3, xi_0=1, nu=1,
                        boundary_phases=[1.0, 1.0, 1.0, 0.0]),
    "wilson_twisted": dict(kappa=0.137, csw_r=0.0, csw_t=0.0,
                           boundary_phases=[cmath.exp(1j * k) for k in (1, 2, 3, 4)]),
}


@pytest.fixture(scope="module")
def small():
    gr = R.grid([4, 4, 4, 8])
    rng = R.random("gpu_parity")
    U = G.random_links(gr, rng, scale=0.8)
    src = rng.vspincolor(gr)
    return gr, U, src


def _ref_params(p):
    return {k: v for k, v in p.items() if k not in ("compress_links", "precision")}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("compress", [False, True])
def test_every_opcode_matches_oracle_fp64(small, case, compress):
    gr, U, src = small
    p = CASES[case]
    ref = W.wilson_clover(U, **_ref_params(p))
    w = g.qcd.fermion.wilson_clover(U, p, compress_links=compress)
    if p["boundary_phases"][-1] == 0.0:
        src = src.copy()
        src[0] = 0.0
        src[-1] = 0.0
    s = to_dev(src, w.F_grid)
    tol = 1e-13
    for name, fn in [("M", ref.M), ("Mdag", ref.Mdag), ("Dhop", ref.Dhop), ("DhopDag", ref.DhopDag),
                     ("Mdiag", ref.Mdiag)]:
        op = {"M": w, "Mdag": w.adj(), "Dhop": w.Dhop, "DhopDag": w.Dhop.adj(), "Mdiag": w.Mdiag}[name]
        got = op(s).to_numpy()
        assert site_rel_err(got, fn(src)) < tol, (case, name)
    for pin, cb in [(0, g.even), (1, g.odd)]:
        sh = half_dev(src, w.F_grid_eo, cb)
        masked = ref._mask(src, pin)
        for name, opx, fn in [
            ("Meooe", w.Meooe, lambda x: ref.Meooe(x, pin)),
            ("MeooeDag", w.Meooe.adj(), lambda x: ref.MeooeDag(x, pin)),
            ("Mooee", w.Mooee, lambda x: ref.Mooee(x, pin)),
            ("MooeeInv", w.Mooee.inv(), lambda x: ref.MooeeInv(x, pin)),
            ("MooeeDag", w.Mooee.adj(), lambda x: ref.MooeeDag(x, pin)),
            ("MooeeInvDag", w.Mooee.adj().inv(), lambda x: ref.MooeeInvDag(x, pin)),
        ]:
            out = opx(sh)
            expect_cb = cb.inv() if name.startswith("Meooe") else cb
            assert out.checkerboard() is expect_cb
            assert site_rel_err(out.to_numpy(), fn(masked)) < tol, (case, name, pin)


@pytest.mark.parametrize("compress", [False, True])
def test_fp32_operator_matches_oracle(small, compress):
    """FP32 with 18-real links and with 12-real links (the inner operator of
    every mixed solve: third row rebuilt in FP32)"""
    gr, U, src = small
    p = CASES["baseline_clover"]
    ref = W.wilson_clover(U, **p)
    w = g.qcd.fermion.wilson_clover(U, p, precision=g.single, compress_links=compress)
    s = to_dev(src.astype(np.complex64), w.F_grid)
    assert site_rel_err(w(s).to_numpy().astype(np.complex128), ref.M(src)) < 2e-6
    sh = half_dev(src.astype(np.complex64), w.F_grid_eo, g.odd)
    got = w.Meooe(sh).to_numpy().astype(np.complex128)
    assert site_rel_err(got, ref.Meooe(ref._mask(src, 1), 1)) < 2e-6
    got = w.Meooe.adj()(sh).to_numpy().astype(np.complex128)
    assert site_rel_err(got, ref.MeooeDag(ref._mask(src, 1), 1)) < 2e-6


def test_grid_fingerprints_on_gpu(kats):
    k = kats["fingerprints"]
    gr = R.grid(k["fdims"])
    rng = R.random(k["seed"])
    U = G.random_links(gr, rng)
    src = rng.vspincolor(gr)
    dst = rng.vspincolor(gr)
    for name, op in k["operators"].items():
        p = dict(op["params"])
        s, d = src.copy(), dst.copy()
        if p["boundary_phases"][-1] == 0.0:
            for f in (s, d):
                f[0] = 0.0
                f[-1] = 0.0
        w = g.qcd.fermion.wilson_clover(U, p)
        S = to_dev(s, w.F_grid)
        D = to_dev(d, w.F_grid)
        for tag, opx in [("M", w), ("Mdiag", w.Mdiag)]:
            X = g.inner_product(D, opx(S))
            ref = complex(*op[tag])
            assert abs(X - ref) / abs(X) < k["tol_rel"], (name, tag, X, ref)


def test_schur_pieces_match_oracle(small):
    from oracle import solvers as S

    gr, U, src = small
    p = CASES["baseline_clover"]
    ref = W.wilson_clover(U, **p)
    pcr = S.eo2_ne(ref, parity=1)
    w = g.qcd.fermion.wilson_clover(U, p)
    pc = g.qcd.fermion.preconditioner.eo2_ne()(w)
    x = ref._mask(src, 1)
    xd = half_dev(src, w.F_grid_eo, g.odd)
    for name, opx, fn in [("Mpc", pc.Mpc_schur, pcr.Mpc), ("MpcDag", pc.Mpc_schur.adj(), pcr.MpcDag),
                          ("Mpc_ne", pc.Mpc, pcr.Mpc_ne)]:
        got = opx(xd).to_numpy()
        assert site_rel_err(got, ref._mask(fn(x), 1)) < 1e-13, name
    # R_ne, L, S on full fields
    sfull = to_dev(src, w.F_grid)
    got = pc.R(sfull).to_numpy()
    assert site_rel_err(got, pcr.R_ne(src)) < 1e-13
    got = pc.L(xd).to_numpy()
    assert site_rel_err(got, pcr.L(x)) < 1e-13
    got = pc.S(sfull).to_numpy()
    assert site_rel_err(got, pcr.S(src)) < 1e-13


def test_blas_kernels():
    gr = g.grid([4, 4, 4, 8], g.double)
    rng = np.random.default_rng(5)
    a = rng.standard_normal((8, 4, 4, 4, 4, 3)) + 1j * rng.standard_normal((8, 4, 4, 4, 4, 3))
    b = rng.standard_normal((8, 4, 4, 4, 4, 3)) + 1j * rng.standard_normal((8, 4, 4, 4, 4, 3))
    A, B = to_dev(a, gr), to_dev(b, gr)
    assert abs(g.norm2(A) - np.vdot(a, a).real) < 1e-12 * np.vdot(a, a).real
    ip = g.inner_product(A, B)
    assert abs(ip - np.vdot(a, b)) < 1e-12 * abs(np.vdot(a, b))
    Z = g.lattice(A)
    c = 0.3 - 1.7j
    g.axpy(Z, c, A, B)
    assert np.abs(Z.to_numpy() - (c * a + b)).max() < 1e-14 * np.abs(c * a + b).max()
    n2 = g.axpy_norm2(Z, -c, A, B)
    assert abs(n2 - np.vdot(b - c * a, b - c * a).real) < 1e-12 * n2
    # reductions are deterministic (fixed grid, fixed order)
    assert all(g.norm2(A) == g.norm2(A) for _ in range(3))
    # expressions: psi += a * p ; g(x - y)
    P = g.lattice(A)
    g.copy(P, A)
    P += 2.0 * B
    assert np.abs(P.to_numpy() - (a + 2.0 * b)).max() < 1e-13
    D = g(A - B)
    assert np.abs(D.to_numpy() - (a - b)).max() < 1e-13
    # lists: the matrix [<a_i, b_j>] (transform.py:101-106), inner_product_norm2
    M = g.inner_product([A, B], [B, A, D])
    ref = np.array([[np.vdot(x, y) for y in (b, a, a - b)] for x in (a, b)])
    assert M.shape == (2, 3) and np.abs(M - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.array_equal(g.rank_inner_product([A], [B]), g.inner_product([A], [B]))
    ip2, n2a = g.inner_product_norm2(A, B)
    assert ip2 == ip and n2a == g.norm2(A)


def test_errors_raise():
    gr = R.grid([4, 4, 4, 4])
    U = G.unit_links(gr.shape)
    w = g.qcd.fermion.wilson_clover(U, mass=0.1, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])
    x = g.lattice(w.F_grid_eo)
    x.checkerboard(g.even)
    y = g.lattice(w.F_grid_eo)
    with pytest.raises(RuntimeError, match="DhopEO"):
        w.DhopEO.mat(y, x)
    with pytest.raises(RuntimeError, match="Unknown opcode"):
        w.interface.apply_unary_operator(9999, y, x)
    with pytest.raises(ValueError):
        w.Meooe.mat(x, x)
    with pytest.raises(ValueError):
        g.qcd.fermion.wilson_clover(U, mass=0.1, boundary_phases=[1, 1, 1, -1], cF=1.2)


def test_operator_pool_reuse_is_exact(small):
    import gc

    from gpt_amd.qcd.fermion import operator as O

    gr, U, src = small
    p = dict(CASES["baseline_clover"])
    w1 = g.qcd.fermion.wilson_clover(U, p)
    tag = O.operator_tag[w1.handle]
    del w1
    gc.collect()
    parked = len(O.operator_limbo.get(tag, []))
    assert parked >= 1
    # same tag with another mass: a parked handle is reused (set_mass, then update)
    p2 = dict(p, mass=0.23)
    w2 = g.qcd.fermion.wilson_clover(U, p2)
    assert len(O.operator_limbo.get(tag, [])) == parked - 1
    assert O.operator_tag[w2.handle] == tag
    ref = W.wilson_clover(U, **p2)
    got = w2(to_dev(src, w2.F_grid)).to_numpy()
    assert site_rel_err(got, ref.M(src)) < 1e-13


@pytest.mark.parametrize("case", ["baseline_clover", "open_clover", "wilson_twisted"])
def test_gauge_covariance(small, case):
    """M[U^V] (V psi) = V (M[U] psi) for a random SU(3) gauge transformation V,
    U^V_mu(x) = V(x) U_mu(x) V(x + mu)^dag (the reference's gauge
    transformation check, fermion_operators.py:258-270, at the operator
    level): pins the link directions, the double storage and the clover
    leaves independently of the oracle"""
    gr, U, src = small
    V = G.random_links(gr, R.random("gauge_transform"), scale=1.0)[0]
    Ut = [V @ U[mu] @ cv.dagger(cv.shift(V, mu, 1)) for mu in range(4)]
    for compress in (False, True):
        p = dict(CASES[case], compress_links=compress)
        w = g.qcd.fermion.wilson_clover(U, p)
        wt = g.qcd.fermion.wilson_clover(Ut, p)
        a = g(w * to_dev(src, w.F_grid)).to_numpy()
        b = g(wt * to_dev(cv.colour_mul(V, src), wt.F_grid)).to_numpy()
        assert site_rel_err(b, cv.colour_mul(V, a)) < 1e-13, compress
