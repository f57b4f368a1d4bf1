"""The even-odd preconditioners beside eo2_ne (even_odd_sites.py:23-89) on
the GPU against the oracle: eo1 / eo1_ne (Schur complement one), eo2 and
eo2_kappa_ne.  Each piece (Mpc, its adjoint, R, L, L^-1, S) agrees per
site to 1e-13 with the oracle on the same inputs, and CG through each
preconditioner needs the oracle's iterations (+-1) and reaches the oracle's
solution."""
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

inv = g.algorithms.inverter
pc = g.qcd.fermion.preconditioner
P = dict(mass=0.1, csw_r=1.0, csw_t=1.2, boundary_phases=[1, 1, 1, -1])
FD = [4, 4, 4, 8]


@pytest.fixture(scope="module")
def setup():
    gr = R.grid(FD)
    rng = R.random("precond")
    U = G.random_links(gr, rng, scale=0.5)
    ref = W.wilson_clover(U, **P)
    src = rng.vspincolor(gr)
    return gr, U, ref, src


def _full(a):
    x = g.lattice(g.grid(FD, g.double))
    x.from_numpy(a)
    return x


def _half(a, cb):
    return g.pick_checkerboard(cb, _full(a))


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("name", ["eo1", "eo1_ne", "eo2", "eo2_kappa_ne"])
def test_pieces_match_oracle(setup, name):
    gr, U, ref, src = setup
    w = g.qcd.fermion.wilson_clover(U, P)
    m = getattr(pc, name)()(w)
    base = {"eo1": S.eo1_ne, "eo1_ne": S.eo1_ne, "eo2": S.eo2_ne, "eo2_kappa_ne": S.eo2_kappa_ne}[name]
    o = base(ref, parity=1)
    ne = name.endswith("_ne")
    d = ref._mask(src, 1)
    mpc = (o.Mpc_ne if ne else o.Mpc)(d)
    assert _rel(m.Mpc(_half(d, g.odd)).to_numpy(), mpc) < 1e-13
    if not ne:
        assert _rel(m.Mpc.adj()(_half(d, g.odd)).to_numpy(), o.MpcDag(d)) < 1e-13
    r = o.R_ne(src) if ne else (o.R(src) if name == "eo1" else ref._mask(o.R(src), 1))
    assert _rel(m.R(_full(src)).to_numpy(), r) < 1e-13
    lo = g.lattice(g.grid(FD, g.double))
    m.L.mat(lo, _half(d, g.odd))
    assert _rel(lo.to_numpy(), o.L(d)) < 1e-13
    li = m.L.inv_mat_make_dst(None)
    m.L.inv_mat(li, _full(src))
    assert _rel(li.to_numpy(), ref._mask(o.L_inv(src), 1)) < 1e-13
    assert _rel(m.S(_full(src)).to_numpy(), o.S(src)) < 1e-13


@pytest.mark.parametrize("name", ["eo1_ne", "eo2_kappa_ne"])
def test_cg_through_preconditioner(setup, name):
    gr, U, ref, src = setup
    cls = {"eo1_ne": S.eo1_ne, "eo2_kappa_ne": S.eo2_kappa_ne}[name]
    x_ref, it_ref, conv = S.preconditioned(ref, src, pc_class=cls, eps=1e-10, maxiter=1000)
    assert conv
    w = g.qcd.fermion.wilson_clover(U, P)
    cg = inv.cg(eps=1e-10, maxiter=1000)
    x = g(w.propagator(inv.preconditioned(getattr(pc, name)(), cg)) * _full(src)).to_numpy()
    print(f"{name}: GPU {len(cg.history)} iterations, oracle {it_ref}", flush=True)
    assert cg.converged and abs(len(cg.history) - it_ref) <= 1
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-9
    assert np.linalg.norm(ref.M(x) - src) / np.linalg.norm(src) < 1e-8


def test_g5m_ne_against_eo1_ne(setup):
    """g5m_ne (g5m.py: Mpc = (gamma5 M)^2 on the full lattice, R = G5M gamma5,
    L = 1, S = 0) gives the solution of the even-odd solve, as the
    reference's propagator test checks (fermion_operators.py:170-175:
    eps2 < 1e-10), and meets the true residual via the oracle's M"""
    gr, U, ref, src = setup
    w = g.qcd.fermion.wilson_clover(U, P)
    cg = inv.cg(eps=1e-12, maxiter=4000)
    x_g5 = g(w.propagator(inv.preconditioned(pc.g5m_ne(), cg)) * _full(src)).to_numpy()
    assert cg.converged
    x_eo = g(w.propagator(inv.preconditioned(pc.eo1_ne(), inv.cg(eps=1e-12, maxiter=4000))) * _full(src)).to_numpy()
    eps2 = np.linalg.norm(x_g5 - x_eo) ** 2 / np.linalg.norm(x_eo) ** 2
    print(f"g5m_ne: {len(cg.history)} iterations, eps2 vs eo1_ne {eps2}", flush=True)
    assert eps2 < 1e-10
    assert np.linalg.norm(ref.M(x_g5) - src) / np.linalg.norm(src) < 1e-9


def test_inverse_of_the_daggered_operator(setup):
    """the reference's check of inv(dag(w)) (fermion_operators.py:177-187):
    <v2, M^-1 v1> = conj(<v1, (M^dag)^-1 v2>), with eo1_ne on w and on w.adj()"""
    gr, U, ref, src = setup
    w = g.qcd.fermion.wilson_clover(U, P)
    rng = np.random.default_rng(17)
    shape = tuple(reversed(FD)) + (4, 3)
    v1 = _full(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    v2 = _full(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    cg = inv.cg(eps=1e-10, maxiter=4000)
    w_v1 = g(inv.preconditioned(pc.eo1_ne(), cg)(w) * v1)
    wadj_v2 = g(inv.preconditioned(pc.eo1_ne(), cg)(w.adj()) * v2)
    eps = abs(g.inner_product(v2, w_v1) / g.inner_product(v1, wadj_v2).conjugate() - 1.0)
    print(f"inv(dag(w)): {eps}", flush=True)
    assert eps < 1e-8
    # and (M^dag)^-1 v2 solves M^dag x = v2 (the oracle's adjoint)
    x = wadj_v2.to_numpy()
    assert np.linalg.norm(ref.Mdag(x) - v2.to_numpy()) / np.linalg.norm(v2.to_numpy()) < 1e-8
