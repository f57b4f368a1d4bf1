"""Even-odd preconditioners (lib/gpt/qcd/fermion/preconditioner/even_odd_sites.py:61-72).

``eo2_ne(parity=odd)`` builds the normal equation of the Schur complement two
(lib/gpt/algorithms/preconditioner/schur_complement_two.py:58-152,
normal_equation.py:35-65) with D = ``parity`` sites:
  Mpc    = 1 - DC CC^-1 CD DD^-1          Mpc_ne = Mpc^dag Mpc
  R_ne   = Mpc^dag (P_D - DC CC^-1 P_C)   L = [DD^-1 ; -CC^-1 CD DD^-1]
  S      = CC^-1 on C                     L^-1 (pseudo) = DD P_D
Mpc and Mpc_ne run as fused HIP kernels (wc_apply_mpc); the CG inverter
recognises this operator and runs its native fused loop.
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


def eo2_ne(params=None, parity=core.odd):
    if isinstance(params, dict):
        parity = params.get("parity", parity)

    def instantiate(op):
        return _eo2_ne(op, parity)

    instantiate.parity = parity
    instantiate.kind = "eo2_ne"
    return instantiate
