"""Even-odd preconditioners (lib/gpt/qcd/fermion/preconditioner/even_odd_sites.py:61-72).

``eo2_ne(parity=odd)`` builds the normal equation of the Schur complement two
(lib/gpt/algorithms/preconditioner/schur_complement_two.py:58-152,
normal_equation.py:35-65) with D = ``parity`` sites:
  Mpc    = 1 - DC CC^-1 CD DD^-1          Mpc_ne = Mpc^dag Mpc
  R_ne   = Mpc^dag (P_D - DC CC^-1 P_C)   L = [DD^-1 ; -CC^-1 CD DD^-1]
  S      = CC^-1 on C                     L^-1 (pseudo) = DD P_D
Mpc and Mpc_ne run as fused HIP kernels (wc_apply_mpc); the CG inverter
recognises this operator and runs its native fused loop.

g5m_ne (g5m.py) solves on the full lattice with (gamma5 M)^2; it and the
neighbouring preconditioners of even_odd_sites.py are compositions of
the operator's HIP pieces (Meooe, Mooee and their inverses and adjoints) and
the HIP BLAS-1 kernels, solved by the generic CG recursion (single
right-hand-side fields):
  eo1 / eo1_ne    schur_complement_one.py: N = 1 - DD^-1 DC CC^-1 CD on D,
                  R = DD^-1 (P_D - DC CC^-1 P_C), L = [1 ; -CC^-1 CD],
                  L^-1 (pseudo) = P_D, S = CC^-1 on C (even_odd_sites.py:23-46)
  eo2             schur_complement_two.py without the normal equation
  eo2_kappa_ne    schur complement two under similarity_transformation.py with
                  the scalar V = kappa (R -> kappa R, L -> L / kappa,
                  L^-1 -> kappa L^-1, Mpc unchanged), then the normal equation
                  (even_odd_sites.py:75-89)
"""
from ... import _lib
from ... import core
from ...core import lattice, matrix_operator, full


class _eo2_ne:
    def __init__(self, op, parity):
        self.op = op
        self.parity = parity
        self.D = parity
        self.C = parity.inv()
        g_eo = op.F_grid_eo

        def half(cb):
            def mk(src=None):
                x = lattice(g_eo)
                x.checkerboard(cb)
                return x
            return mk

        self._halfD = half(self.D)
        self._halfC = half(self.C)

        def mpc(which):
            def f(dst, src):
                if src.checkerboard() is not self.D:
                    raise ValueError("Mpc acts on fields of the D parity")
                _lib.check(_lib.load().wc_apply_mpc_block(op.handle, self.D.tag, which, src.grid.nrhs,
                                                          dst.ptr(), src.ptr(), core.stream_ptr()))
                dst.checkerboard(self.D)
            return f

        # the Schur complement itself and its normal equation
        self.Mpc_schur = matrix_operator(mpc(0), mpc(1), make_dst=lambda s: self._halfD())
        self.Mpc = matrix_operator(mpc(2), mpc(2), make_dst=lambda s: self._halfD())
        self.Mpc.eo2_ne = self  # marks the native fused path for the CG inverter

        Mooee, Meooe = op.Mooee, op.Meooe

        def _R(o_d, i):
            # Mpc^dag (i_D - DC CC^-1 i_C)
            iC = core.pick_checkerboard(self.C, i)
            t = Mooee.inv()(iC)
            r = Meooe(t)
            iD = core.pick_checkerboard(self.D, i)
            core.axpy(r, -1.0, r, iD)
            self.Mpc_schur.adj_mat(o_d, r)

        def _L(o, i_d):
            d = Mooee.inv()(i_d)
            c = Mooee.inv()(Meooe(d))
            core.scale(c, -1.0, c)
            core.set_checkerboard(o, d)
            core.set_checkerboard(o, c)

        def _L_pseudo_inverse(o_d, i):
            iD = core.pick_checkerboard(self.D, i)
            Mooee.mat(o_d, iD)

        def _S(o, i):
            iC = core.pick_checkerboard(self.C, i)
            c = Mooee.inv()(iC)
            z = lattice(c)
            z.data.zero_()
            z.checkerboard(self.D)
            core.set_checkerboard(o, c)
            core.set_checkerboard(o, z)

        full_dst = lambda s: lattice(op.F_grid)  # noqa: E731
        self.R = matrix_operator(_R, make_dst=lambda s: self._halfD())
        self.L = matrix_operator(_L, inv_mat=_L_pseudo_inverse, make_dst=full_dst)
        self.L.inv_mat_make_dst = lambda s: self._halfD()
        self.S = matrix_operator(_S, make_dst=full_dst)


def _half_maker(op, cb):
    def mk(src=None):
        x = lattice(op.F_grid_eo)
        x.checkerboard(cb)
        return x
    return mk


def _set_full(o, parts):
    for x in parts:
        core.set_checkerboard(o, x)


class _schur_one:
    """schur_complement_one.py with D = parity, on the HIP operator pieces"""

    def __init__(self, op, parity):
        self.op, self.D, self.C = op, parity, parity.inv()
        D, C = self.D, self.C
        Meooe, Mooee = op.Meooe, op.Mooee
        Minv, Minv_dag, Meooe_dag = Mooee.inv(), Mooee.adj().inv(), Meooe.adj()
        halfD = _half_maker(op, D)

        def _N(o_d, i_d):
            t = Minv(Meooe(Minv(Meooe(i_d))))  # DD^-1 DC CC^-1 CD i_d
            core.axpy(o_d, -1.0, t, i_d)

        def _N_dag(o_d, i_d):
            t = Meooe_dag(Minv_dag(Meooe_dag(Minv_dag(i_d))))  # CD^dag CC^-dag DC^dag DD^-dag i_d
            core.axpy(o_d, -1.0, t, i_d)

        def _L(o, i_d):
            c = Minv(Meooe(i_d))
            core.scale(c, -1.0, c)
            _set_full(o, [i_d, c])

        def _L_pseudo_inverse(o_d, i):
            core.pick_checkerboard(D, o_d, i)

        def _R(o_d, i):
            t = Meooe(Minv(core.pick_checkerboard(C, i)))
            core.axpy(t, -1.0, t, core.pick_checkerboard(D, i))
            Minv.mat(o_d, t)

        def _S(o, i):
            c = Minv(core.pick_checkerboard(C, i))
            z = lattice(c)
            z.data.zero_()
            z.checkerboard(D)
            _set_full(o, [c, z])

        full_dst = lambda s: lattice(op.F_grid)  # noqa: E731
        self.Mpc = matrix_operator(_N, _N_dag, make_dst=lambda s: halfD())
        self.R = matrix_operator(_R, make_dst=lambda s: halfD())
        self.L = matrix_operator(_L, inv_mat=_L_pseudo_inverse, make_dst=full_dst)
        self.L.inv_mat_make_dst = lambda s: halfD()
        self.S = matrix_operator(_S, make_dst=full_dst)


class _schur_two:
    """schur_complement_two.py with D = parity: the eo2_ne pieces without the
    normal equation (Mpc = 1 - DC CC^-1 CD DD^-1, R = P_D - DC CC^-1 P_C)"""

    def __init__(self, op, parity):
        ne = _eo2_ne(op, parity)
        self.op, self.D, self.C = op, ne.D, ne.C
        Meooe, Minv = op.Meooe, op.Mooee.inv()
        self.Mpc = ne.Mpc_schur

        def _R(o_d, i):
            t = Meooe(Minv(core.pick_checkerboard(self.C, i)))
            core.axpy(o_d, -1.0, t, core.pick_checkerboard(self.D, i))

        self.R = matrix_operator(_R, make_dst=lambda s: ne._halfD())
        self.L, self.S = ne.L, ne.S


class _normal_equation:
    """normal_equation.py: Mpc^dag Mpc, R -> Mpc^dag R, L and S kept"""

    def __init__(self, pc):
        self.op, self.D, self.C = pc.op, pc.D, pc.C
        Mpc, R = pc.Mpc, pc.R
        Mpc_dag = Mpc.adj()

        def _NdN(o_d, i_d):
            Mpc_dag.mat(o_d, Mpc(i_d))

        def _R(o_d, i):
            Mpc_dag.mat(o_d, R(i))

        self.Mpc = matrix_operator(_NdN, _NdN, make_dst=Mpc.make_dst)
        self.R = matrix_operator(_R, make_dst=R.make_dst)
        self.L, self.S = pc.L, pc.S


class _similarity_scalar:
    """similarity_transformation.py with a scalar V: Mpc unchanged, R -> V R,
    L -> L V^-1, L^-1 -> V L^-1"""

    def __init__(self, pc, v):
        self.op, self.D, self.C = pc.op, pc.D, pc.C
        self.Mpc, self.S = pc.Mpc, pc.S
        R, L = pc.R, pc.L

        def _R(o_d, i):
            R.mat(o_d, i)
            core.scale(o_d, v, o_d)

        def _L(o, i_d):
            t = lattice(i_d)
            core.scale(t, 1.0 / v, i_d)
            t.checkerboard(i_d.checkerboard())
            L.mat(o, t)

        def _L_inv(o_d, i):
            L.inv_mat(o_d, i)
            core.scale(o_d, v, o_d)

        self.R = matrix_operator(_R, make_dst=R.make_dst)
        self.L = matrix_operator(_L, inv_mat=_L_inv, make_dst=L.make_dst)
        self.L.inv_mat_make_dst = L.inv_mat_make_dst


def _factory(kind, build):
    def make(params=None, parity=core.odd):
        if isinstance(params, dict):
            parity = params.get("parity", parity)

        def instantiate(op):
            return build(op, parity)

        instantiate.parity = parity
        instantiate.kind = kind
        return instantiate

    return make


eo1 = _factory("eo1", _schur_one)
eo1_ne = _factory("eo1_ne", lambda op, p: _normal_equation(_schur_one(op, p)))
eo2 = _factory("eo2", _schur_two)
eo2_kappa_ne = _factory("eo2_kappa_ne",
                        lambda op, p: _normal_equation(_similarity_scalar(_schur_two(op, p), op.kappa())))


def _gamma5(dst, src):
    """dst = gamma5 src; gamma5 = diag(1, 1, -1, -1) in gpt's chiral basis
    (lib/gpt/core/gamma.py:28-53): the lower spins (components 6..11) negated"""
    core.copy(dst, src)
    d = dst.data
    d.view(-1, 12, d.shape[-3] * d.shape[-2] * 2)[:, 6:12].neg_()


class _g5m_ne:
    """g5m_ne (lib/gpt/qcd/fermion/preconditioner/g5m.py:28-60): the full
    lattice with the Hermitian gamma5 M: Mpc = (G5M)^2, L = 1, R = G5M gamma5,
    S = 0, so that L Mpc^-1 R = M^-1"""

    def __init__(self, op):
        self.op = op
        self.F_grid = op.F_grid
        G5M = op.G5M

        def make(s):
            return lattice(op.F_grid)

        def mpc(dst, src):
            G5M.mat(dst, G5M(src))

        def ident(dst, src):
            core.copy(dst, src)

        def r(dst, src):
            t = lattice(op.F_grid)
            _gamma5(t, src)
            G5M.mat(dst, t)

        def zero(dst, src):
            dst.data.zero_()

        self.Mpc = matrix_operator(mpc, mpc, make_dst=make)
        self.L = matrix_operator(ident, inv_mat=ident, make_dst=make)
        self.R = matrix_operator(r, make_dst=make)
        self.S = matrix_operator(zero, make_dst=make)


def g5m_ne(params=None):
    """pc.g5m_ne() (g5m.py:63-69)"""

    def instantiate(op):
        return _g5m_ne(op)

    instantiate.kind = "g5m_ne"
    return instantiate


def eo2_ne(params=None, parity=core.odd):
    if isinstance(params, dict):
        parity = params.get("parity", parity)

    def instantiate(op):
        return _eo2_ne(op, parity)

    instantiate.parity = parity
    instantiate.kind = "eo2_ne"
    return instantiate
