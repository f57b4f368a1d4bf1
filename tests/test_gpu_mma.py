"""The FP16-storage multi-RHS hop on the matrix cores (k_hop_mma), through
the C ABI, against the oracle and against the vector-ALU form of the same hop.

Inputs are complex-Gaussian spinors rounded to FP16 (so the 16-bit field
holds them exactly) on an 8^3 x 16 lattice with smoothed SU(3) links, for
4, 12 and 20 right-hand sides (20: a chunk of 16 and one of 4), with the
links stored as 18 or 12 reals, hop and adjoint, and three parameter sets
(isotropic clover; anisotropic clover with complex boundary phases; open
boundaries).  Bars:

* Meooe vs the oracle's FP64 Meooe of the same (widened) input: per site
  <= 3e-3 relative.  The hop projects in FP16 (one rounding, 2^-11, per
  projected value, as the vector form does) and stores FP16 (2^-11); a
  wrong sign, spin, colour or direction gives O(1).
* Meooe vs the vector form: both project identically and store FP16; they
  differ by the order of their FP32 sums (and the 2^-22 split of the
  links), so the stored values differ by at most one FP16 unit in the last
  place, and only where the exact value lies near a rounding boundary; plus
  an absolute 2^-13 for projected spinor values below FP16's normal range
  (6.1e-5), which the matrix cores flush to zero (at most a few per site,
  each times a link entry <= 1/2).
* Mpc^dag Mpc with 16-bit intermediates (clover, axpy and inner-product
  epilogues) vs the vector form (<= 2e-4 of the field norm) and vs the
  oracle's FP64 normal operator (<= 5e-3 per site).
* The FP16 mixed solve with the matrix-core hops: true relative residual
  <= 1.05 tol through the oracle's M, iterations within 3 of the vector
  form, solutions agreeing to 1e-8.
"""
import numpy as np
import pytest

from oracle import gauge as G
from oracle import solvers as SV
from oracle import wilson_clover as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd.qcd.fermion.operator import pack, unpack  # noqa: E402

FD = [8, 8, 8, 16]
CASES = {
    "clover": dict(mass=0.05, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0, boundary_phases=[1, 1, 1, -1]),
    "aniso_phases": dict(mass=0.1, csw_r=1.5, csw_t=1.9, xi_0=1.3, nu=2.6, isAnisotropic=True,
                         boundary_phases=[np.exp(0.3j), -1.0, np.exp(1.1j), np.exp(2.2j)]),
    "open": dict(mass=0.1, csw_r=1.9, csw_t=1.9, cF=1.3, xi_0=1.0, nu=1.0, boundary_phases=[1, 1, 1, 0]),
}


@pytest.fixture(scope="module")
def links():
    return G.smooth_links(FD, 0.3, 7001)


def fp16_sources(n, seed):
    shape = tuple(reversed(FD)) + (4, 3)
    r = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        z = (r.standard_normal(shape) + 1j * r.standard_normal(shape)) / np.sqrt(2.0)
        out.append(z.real.astype(np.float16).astype(np.float64) + 1j * z.imag.astype(np.float16).astype(np.float64))
    return out


def half_field(arrs, w, cb):
    """the N right-hand sides as one FP32 half field of parity cb"""
    fs = []
    for a in arrs:
        x = g.lattice(w.F_grid_eo.single_rhs())
        x.checkerboard(cb)
        x.from_numpy(a)
        fs.append(x)
    return pack(fs)


def from16(t16, like, cb):
    """a 16-bit tensor in the layout of `like` -> FP32 half field of parity cb"""
    out = g.lattice(like)
    out.data.copy_(t16.float().view_as(out.data))
    out.checkerboard(cb)
    return out


def site_rel_err(got, ref):
    d = np.abs(got - ref).reshape(got.shape[:4] + (-1,))
    r = np.abs(ref).reshape(ref.shape[:4] + (-1,))
    dn = np.sqrt((d**2).sum(-1))
    rn = np.sqrt((r**2).sum(-1))
    return float((dn / np.where(rn > 0, rn, 1.0)).max())


def inner_op(U, p, N, compress):
    """the FP32 operator of N right-hand sides, as a mixed solve's inner one"""
    return g.qcd.fermion.wilson_clover(U, p, n_rhs=N, precision=g.single, compress_links=compress)


@pytest.mark.parametrize("case,compress,dag,N", [
    ("clover", True, False, 4), ("clover", True, False, 12), ("clover", True, True, 12),
    ("clover", False, False, 20), ("clover", False, True, 4), ("clover", True, True, 20),
    ("aniso_phases", True, False, 12), ("aniso_phases", False, True, 12), ("open", True, False, 12),
])
def test_hop_matches_oracle_and_vector_form(links, case, compress, dag, N):
    p = CASES[case]
    ref = W.wilson_clover(links, **p)
    srcs = fp16_sources(N, 7100 + N)
    w = inner_op(links, p, N, compress)
    for cb_in in (g.odd, g.even):
        pin = 1 if cb_in is g.odd else 0
        H = half_field(srcs, w, cb_in)
        s16 = H.data.to(torch.float16)
        assert torch.equal(s16.float(), H.data)  # the FP16 field holds the inputs exactly
        w.set_hop_mma(True)
        o_mma = w.hop_s16(s16, pin, dag=dag, n_rhs=N)
        o_mma2 = w.hop_s16(s16, pin, dag=dag, n_rhs=N)
        w.set_hop_mma(False)
        o_vec = w.hop_s16(s16, pin, dag=dag, n_rhs=N)
        torch.cuda.synchronize()
        assert torch.equal(o_mma, o_mma2), "the matrix-core hop is not deterministic"
        a, b = o_mma.float(), o_vec.float()
        # one FP16 unit in the last place of the larger value at most, plus the
        # flushed FP16 subnormals
        ulp = torch.maximum(a.abs(), b.abs()) * 2.0 ** -10 + 2.0 ** -13
        assert bool(((a - b).abs() <= ulp).all()), (case, compress, dag, N, float((a - b).abs().max()))
        frac = float(((a - b) != 0).float().mean())
        assert frac < 0.02, frac
        got = unpack(from16(o_mma, H, cb_in.inv()))
        for r in range(N):
            x = srcs[r]
            exp = ref.MeooeDag(ref._mask(x, pin), pin) if dag else ref.Meooe(ref._mask(x, pin), pin)
            err = site_rel_err(got[r].to_numpy().astype(np.complex128), exp)
            assert err < 3e-3, (case, compress, dag, N, pin, r, err)


@pytest.mark.parametrize("N,compress", [(12, True), (20, False)])
def test_normal_operator_16bit(links, N, compress):
    """Mpc^dag Mpc as the inner CG applies it (16-bit in and intermediates,
    FP32 out): the epilogues CLOV, AXPY and CLOV|AXPY|DOT of the MFMA kernel"""
    p = CASES["clover"]
    ref = W.wilson_clover(links, **p)
    pc_ref = SV.eo2_ne(ref, 1)
    srcs = fp16_sources(N, 7300 + N)
    w = inner_op(links, p, N, compress)
    H = half_field(srcs, w, g.odd)
    s16 = H.data.to(torch.float16)
    w.set_hop_mma(True)
    m = w.mpc_ne_s16(s16, 1, n_rhs=N)
    w.set_hop_mma(False)
    v = w.mpc_ne_s16(s16, 1, n_rhs=N)
    torch.cuda.synchronize()
    rel = float((m - v).norm() / v.norm())
    assert rel < 2e-4, rel
    out = g.lattice(H)
    out.data.copy_(m.view_as(out.data))
    got = unpack(out)
    for r in range(N):
        exp = pc_ref.Mpc_ne(ref._mask(srcs[r], 1))
        err = site_rel_err(got[r].to_numpy().astype(np.complex128), exp)
        assert err < 5e-3, (N, r, err)


@pytest.mark.parametrize("N", [4, 12])
def test_fp16_mixed_solve_on_matrix_cores(links, N):
    p = CASES["clover"]
    ref = W.wilson_clover(links, **p)
    inv, pc = g.algorithms.inverter, g.qcd.fermion.preconditioner
    srcs = fp16_sources(N, 7500 + N)
    wN = g.qcd.fermion.wilson_clover(links, p, n_rhs=N)
    full = []
    for a in srcs:
        x = g.lattice(wN.F_grid.single_rhs())
        x.from_numpy(a)
        full.append(x)
    S = pack(full)
    tol = 1e-10
    res = {}
    for mma in (True, False):
        m = inv.cg_mixed(eps=tol, maxiter=5000, inner_storage="half", mma=mma)
        dst = g(wN.propagator(inv.preconditioned(pc.eo2_ne(), m)) * S)
        res[mma] = (unpack(dst), list(m.iterations), m.reliable_updates)
        print(f"N={N} mma={mma}: iterations {m.iterations}, reliable updates {m.reliable_updates}", flush=True)
    sol, its, _ = res[True]
    sol_v, its_v, _ = res[False]
    for r in range(N):
        x = sol[r].to_numpy()
        tres = np.linalg.norm(srcs[r] - ref.M(x)) / np.linalg.norm(srcs[r])
        assert tres <= 1.05 * tol, (r, tres)
        assert abs(its[r] - its_v[r]) <= 3, (r, its[r], its_v[r])
        xv = sol_v[r].to_numpy()
        assert np.linalg.norm(x - xv) / np.linalg.norm(xv) < 1e-8


def test_bf16_and_odd_counts_stay_on_vector_form(links):
    """BF16 storage and right-hand-side counts that are not a multiple of 4
    take the vector form (the switch changes nothing for them)"""
    p = CASES["clover"]
    ref = W.wilson_clover(links, **p)
    for storage, N, tol in (("bf16", 12, 2.5e-2), ("half", 3, 3e-3)):
        srcs = fp16_sources(N, 7700 + N)
        w = inner_op(links, p, N, True)
        H = half_field(srcs, w, g.odd)
        dt = torch.bfloat16 if storage == "bf16" else torch.float16
        s16 = H.data.to(dt)
        w.set_hop_mma(True)
        a = w.hop_s16(s16, 1, storage=storage, n_rhs=N)
        w.set_hop_mma(False)
        b = w.hop_s16(s16, 1, storage=storage, n_rhs=N)
        torch.cuda.synchronize()
        assert torch.equal(a, b), storage
        got = unpack(from16(a, H, g.even))
        stored = unpack(from16(s16, H, g.odd))  # the inputs as the 16-bit field holds them
        for r in range(N):
            inp = stored[r].to_numpy().astype(np.complex128)
            exp = ref.Meooe(ref._mask(inp, 1), 1)
            err = site_rel_err(got[r].to_numpy().astype(np.complex128), exp)
            assert err < tol, (storage, N, r, err)


def test_automatic_form_on_a_tuned_lattice():
    """16^3 x 32 with 4 right-hand sides: the solvers tune there (>= 2^17
    site-RHS pairs), the automatic form is the one timed faster, and the
    solve meets its tolerance through the oracle's M either way"""
    fd = [16, 16, 16, 32]
    U = G.smooth_links(fd, 0.3, 7901)
    p = CASES["clover"]
    ref = W.wilson_clover(U, **p)
    inv, pc = g.algorithms.inverter, g.qcd.fermion.preconditioner
    N = 4
    wN = g.qcd.fermion.wilson_clover(U, p, n_rhs=N)
    shape = tuple(reversed(fd)) + (4, 3)
    rng = np.random.default_rng(7902)
    srcs = [(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0) for _ in range(N)]
    full = []
    for a in srcs:
        x = g.lattice(wN.F_grid.single_rhs())
        x.from_numpy(a)
        full.append(x)
    m = inv.cg_mixed(eps=1e-10, maxiter=5000, inner_storage="half")
    assert m.mma == "auto"
    dst = unpack(g(wN.propagator(inv.preconditioned(pc.eo2_ne(), m)) * pack(full)))
    form = m.inner_operator(wN).hop_mma(N)
    print(f"16^3x32, 4 RHS: automatic form {'matrix cores' if form else 'vector ALUs'}, "
          f"iterations {m.iterations}", flush=True)
    assert isinstance(form, bool)
    assert not m.inner_operator(wN).hop_mma(3)  # 3 right-hand sides never take the matrix-core form
    for r in range(N):
        tres = np.linalg.norm(srcs[r] - ref.M(dst[r].to_numpy())) / np.linalg.norm(srcs[r])
        assert tres <= 1.05e-10, (r, tres)


def test_vector_16bit_hop_same_bits_for_every_shape(links):
    """the 16-bit hops take their own launch shape (timed when the solvers
    tune); a caller's set_hop_shape applies to them too, and every vector
    shape gives the same bits, so the choice changes speed only"""
    p = CASES["clover"]
    N = 8
    srcs = fp16_sources(N, 7800)
    w = inner_op(links, p, N, True)
    H = half_field(srcs, w, g.odd)
    s16 = H.data.to(torch.float16)
    w.set_hop_mma(False)
    ref = None
    for sh in [(1, 1, 0, 1, 0), (1, 2, 4, 1, 1), (1, 8, 0, 1, 1), (1, 4, 8, 1, 0), (1, 4, 4, 2, 1)]:
        w.set_hop_shape(sh)
        out = w.hop_s16(s16, 1, n_rhs=N).clone()
        Y = w.mpc_ne_s16(s16, 1, n_rhs=N).clone()
        torch.cuda.synchronize()
        if ref is None:
            ref = (out, Y)
        else:
            assert torch.equal(out, ref[0]) and torch.equal(Y, ref[1]), sh
    w.set_hop_shape(None)
    w.set_hop_mma("auto")
