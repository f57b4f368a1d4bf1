"""Halo of a T-slab decomposition (wc_halo_pack / wc_halo_unpack) against the
oracle's hop.

On one GPU with periodic t the rank above and below a slab is the slab
itself, so the faces that pack writes are the ones unpack reads.  What unpack
adds must be exactly the hop's terms through the wrapping t-links: the
oracle's Meooe with the links U_t(x, t = Lt - 1) set to zero differs from the
full Meooe by those terms alone (the t-terms of the output slices Lt - 1 and 0).
So, per case (boundary phase, link storage, hop or adjoint, parity):
  unpack(pack(psi)) into a zero field == Meooe_U psi - Meooe_U' psi   (oracle)
  GPU Meooe_U' psi + unpack(pack(psi)) == Meooe_U psi                  (oracle)
FP64 bar: 1e-13 relative per site (the differences of two oracle hops are
compared relative to the full hop's site norm).  Multi-RHS faces hold each
right-hand side's single-RHS face bit for bit.
"""
import cmath

import numpy as np
import pytest

from oracle import gauge as G
from oracle import rng as R
from oracle import wilson_clover as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - collected on CPU only with -m gpu
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd.qcd.fermion.operator import pack  # noqa: E402

CASES = {
    "antiperiodic_clover": dict(mass=0.1, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1]),
    "twisted_wilson": dict(kappa=0.137, csw_r=0.0, csw_t=0.0,
                           boundary_phases=[cmath.exp(1j * k) for k in (1, 2, 3, 4)]),
}


@pytest.fixture(scope="module")
def slab():
    # Lx, Ly, Lz, Lt = 4, 8, 6, 4: unequal extents, nf = 96 face sites
    gr = R.grid([4, 8, 6, 4])
    rng = R.random("halo")
    U = G.random_links(gr, rng, scale=0.8)
    Ucut = [u.copy() for u in U]
    Ucut[3][-1] = 0.0  # U_t on the last slice: the links that wrap around in t
    src = rng.vspincolor(gr)
    return gr, U, Ucut, src


def half_dev(arr, grid_eo, cb):
    x = g.lattice(grid_eo)
    x.checkerboard(cb)
    x.from_numpy(arr)
    return x


def zero_half(grid_eo, cb):
    x = g.lattice(grid_eo)
    x.checkerboard(cb)
    x[:] = 0
    return x


def rel_err_sites(got, ref, scale_ref):
    d = np.sqrt((np.abs(got - ref) ** 2).reshape(got.shape[:4] + (-1,)).sum(-1))
    s = np.sqrt((np.abs(scale_ref) ** 2).reshape(got.shape[:4] + (-1,)).sum(-1))
    return float((d / np.where(s > 0, s, 1.0)).max())


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("compress", [False, True])
@pytest.mark.parametrize("dag", [0, 1])
def test_unpack_of_pack_is_the_wrapping_t_terms(slab, case, compress, dag):
    gr, U, Ucut, src = slab
    p = CASES[case]
    ref = W.wilson_clover(U, **p)
    ref_cut = W.wilson_clover(Ucut, **p)
    w = g.qcd.fermion.wilson_clover(U, p, compress_links=compress)
    w_cut = g.qcd.fermion.wilson_clover(Ucut, p, compress_links=compress)
    Lt = src.shape[0]
    for pin, cb in [(0, g.even), (1, g.odd)]:
        masked = ref._mask(src, pin)
        hop = (lambda r: r.MeooeDag(masked, pin)) if dag else (lambda r: r.Meooe(masked, pin))
        full, cut = hop(ref), hop(ref_cut)
        sh = half_dev(src, w.F_grid_eo, cb)
        lo, hi = w.halo_pack(sh, dag=dag)
        assert lo.numel() == 6 * (4 * 8 * 6 // 2) and lo.dtype == torch.complex128
        # the t-terms alone
        d = zero_half(w.F_grid_eo, cb.inv())
        w.halo_unpack(lo, hi, d, dag=dag)
        got = d.to_numpy()
        assert rel_err_sites(got, full - cut, full) < 1e-13, (case, compress, dag, pin)
        assert not np.any(got[1:Lt - 1]), "unpack touched a slice that is not on a face"
        assert np.any(got[0]) and np.any(got[Lt - 1])
        # the hop without the wrapping links, completed by the halo
        op_cut = w_cut.Meooe.adj() if dag else w_cut.Meooe
        out = op_cut(sh)
        w.halo_unpack(lo, hi, out, dag=dag)
        assert rel_err_sites(out.to_numpy(), full, full) < 1e-13, (case, compress, dag, pin)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("compress", [False, True])
@pytest.mark.parametrize("dag", [0, 1])
def test_interior_plus_halo_is_the_hop(slab, case, compress, dag):
    """A T-slab rank's hop split the multi-GPU way: wc_halo_interior (every
    term but the wrapping t-terms) equals the hop with the wrapping t-links
    cut, value for value, and interior + unpack(pack) equals the oracle's
    hop to 1e-13 per site (3 right-hand sides: each equals its 1-RHS split)."""
    gr, U, Ucut, src = slab
    p = CASES[case]
    ref = W.wilson_clover(U, **p)
    w = g.qcd.fermion.wilson_clover(U, p, compress_links=compress)
    w_cut = g.qcd.fermion.wilson_clover(Ucut, p, compress_links=compress)
    for pin, cb in [(0, g.even), (1, g.odd)]:
        masked = ref._mask(src, pin)
        full = ref.MeooeDag(masked, pin) if dag else ref.Meooe(masked, pin)
        sh = half_dev(src, w.F_grid_eo, cb)
        inner = w.halo_interior(sh, dag=dag)
        assert inner.checkerboard() is cb.inv()
        cut = (w_cut.Meooe.adj() if dag else w_cut.Meooe)(sh)
        assert torch.equal(inner.data, cut.data), (case, compress, dag, pin)
        lo, hi = w.halo_pack(sh, dag=dag)
        w.halo_unpack(lo, hi, inner, dag=dag)
        assert rel_err_sites(inner.to_numpy(), full, full) < 1e-13, (case, compress, dag, pin)


def test_interior_multi_rhs(slab):
    gr, U, Ucut, src = slab
    p = CASES["antiperiodic_clover"]
    rng = np.random.default_rng(11)
    srcs = [src] + [(rng.standard_normal(src.shape) + 1j * rng.standard_normal(src.shape)) for _ in range(2)]
    w1 = g.qcd.fermion.wilson_clover(U, p)
    w3 = g.qcd.fermion.wilson_clover(U, p, n_rhs=3)
    halves = [half_dev(s, w1.F_grid_eo, g.odd) for s in srcs]
    m3 = w3.halo_interior(pack(halves))
    m1 = pack([w1.halo_interior(h) for h in halves])
    assert torch.equal(m3.data, m1.data)


def test_overlapped_schedule_equals_the_serial_one(slab):
    """halo_hop (faces exchanged on a second stream while the interior runs)
    gives bit for bit the serial interior + unpack(pack), and the oracle's hop
    to 1e-13; repeated calls on one comm stream give the same bits."""
    gr, U, Ucut, src = slab
    p = CASES["twisted_wilson"]
    ref = W.wilson_clover(U, **p)
    w = g.qcd.fermion.wilson_clover(U, p)
    sh = half_dev(src, w.F_grid_eo, g.odd)
    serial = w.halo_interior(sh)
    lo, hi = w.halo_pack(sh)
    w.halo_unpack(lo, hi, serial)
    comm = torch.cuda.Stream()
    for _ in range(3):
        d = zero_half(w.F_grid_eo, g.even)
        w.halo_hop(sh, d, lambda a, b: (a.clone(), b.clone()), comm_stream=comm)
        torch.cuda.synchronize()
        assert torch.equal(d.data, serial.data)
    full = ref.Meooe(ref._mask(src, 1), 1)
    assert rel_err_sites(d.to_numpy(), full, full) < 1e-13


def test_fp32_halo(slab):
    gr, U, Ucut, src = slab
    p = CASES["antiperiodic_clover"]
    ref = W.wilson_clover(U, **p)
    w_cut = g.qcd.fermion.wilson_clover(Ucut, p, precision=g.single)
    w = g.qcd.fermion.wilson_clover(U, p, precision=g.single)
    sh = half_dev(src.astype(np.complex64), w.F_grid_eo, g.odd)
    lo, hi = w.halo_pack(sh)
    assert lo.dtype == torch.complex64
    out = w_cut.Meooe(sh)
    w.halo_unpack(lo, hi, out)
    full = ref.Meooe(ref._mask(src, 1), 1)
    assert rel_err_sites(out.to_numpy().astype(np.complex128), full, full) < 2e-6


def test_multi_rhs_faces_are_the_single_rhs_faces(slab):
    gr, U, Ucut, src = slab
    p = CASES["twisted_wilson"]
    rng = np.random.default_rng(7)
    srcs = [src] + [(rng.standard_normal(src.shape) + 1j * rng.standard_normal(src.shape)) for _ in range(2)]
    w1 = g.qcd.fermion.wilson_clover(U, p)
    w3 = g.qcd.fermion.wilson_clover(U, p, n_rhs=3)
    halves = [half_dev(s, w1.F_grid_eo, g.even) for s in srcs]
    multi = pack(halves)
    assert multi.grid.nrhs == 3
    lo3, hi3 = w3.halo_pack(multi)
    nf = 4 * 8 * 6 // 2
    for r, h in enumerate(halves):
        lo1, hi1 = w1.halo_pack(h)
        assert torch.equal(lo3.view(6, nf, 3)[:, :, r], lo1.view(6, nf))
        assert torch.equal(hi3.view(6, nf, 3)[:, :, r], hi1.view(6, nf))
    # unpack on 3 right-hand sides equals unpack per right-hand side
    outs = [zero_half(w1.F_grid_eo, g.odd) for _ in range(3)]
    for r, o in enumerate(outs):
        lo1, hi1 = w1.halo_pack(halves[r])
        w1.halo_unpack(lo1, hi1, o)
    d3 = pack(outs)
    d3[:] = 0
    w3.halo_unpack(lo3, hi3, d3)
    assert torch.equal(d3.data, pack(outs).data)


def test_bad_arguments_are_refused(slab):
    gr, U, Ucut, src = slab
    w = g.qcd.fermion.wilson_clover(U, CASES["antiperiodic_clover"])
    sh = half_dev(src, w.F_grid_eo, g.odd)
    lo, hi = w.halo_pack(sh)
    d = zero_half(w.F_grid_eo, g.even)
    with pytest.raises(ValueError):
        w.halo_unpack(lo[:-1], hi, d)
    with pytest.raises(ValueError):
        w.halo_unpack(lo.to(torch.complex64), hi, d)
    with pytest.raises(RuntimeError, match="alias"):
        w.halo_unpack(lo, lo, d)
    w_open = g.qcd.fermion.wilson_clover(U, dict(kappa=0.135, csw_r=1.0, csw_t=1.0, cF=1.0, xi_0=1, nu=1,
                                                 boundary_phases=[1.0, 1.0, 1.0, 0.0]))
    with pytest.raises(RuntimeError, match="open boundaries"):
        w_open.halo_pack(sh)
    with pytest.raises(RuntimeError, match="open boundaries"):
        w_open.halo_interior(sh)
    with pytest.raises(ValueError):
        w.halo_interior(sh, dst=sh)
