"""Inverters of gpt_amd (lib/gpt/algorithms/inverter/__init__.py).

* ``cg``                -- lib/gpt/algorithms/inverter/cg.py:24-117.  On the
  eo2_ne normal operator of a gpt_amd fermion it runs the native fused loop
  (wc_cg_eo2_ne: hop kernels with fused clover/axpy/dot epilogues, device
  scalars); on any other matrix_operator the same recursion with the HIP
  BLAS-1 kernels.  ``history`` holds |r|^2 per iteration as gpt's
  base_iterative.log_convergence does (lib/gpt/algorithms/base.py:76-106).
* ``block_cg``          -- lib/gpt/algorithms/inverter/block_cg.py:36-149: the
  right-hand sides of an ``n_rhs`` field solved together over their block
  Krylov space (wc_block_cg_eo2_ne); on the eo2_ne normal operator of a
  gpt_amd fermion, alone or inside ``preconditioned``.  ``history`` holds the
  largest relative |r|^2 of the block per iteration.
* ``preconditioned``    -- lib/gpt/algorithms/inverter/preconditioned.py:24-54
* ``defect_correcting`` -- lib/gpt/algorithms/inverter/defect_correcting.py:74-142
* ``mixed_precision``   -- lib/gpt/algorithms/inverter/mixed_precision.py:23-32
* ``cg_mixed``          -- no gpt counterpart: CG with reliable updates whose
  recursion runs in ``inner_precision`` (FP32 by default) while FP64 keeps the
  solution and the true residuals; used as
  ``preconditioned(eo2_ne(), cg_mixed(eps=1e-10))`` it stops on the full
  system's true relative residual |src - M dst| / |src| <= eps.
  ``inner_storage="half"`` (IEEE FP16) or ``"bf16"`` (bfloat16) stores the
  inner search direction and the Schur intermediates as 16-bit floats (FP32
  arithmetic); ``inner_compress_links``
  (default True) gives the inner operator 12-real links.  The inner operator
  only drives the recursion between reliable updates -- every true residual
  and the stopping test use the exact FP64 operator -- so approximations in
  it change the speed, never the result.
"""
import ctypes

import numpy as np

from .. import _lib
from .. import core
from ..core import lattice, matrix_operator


class cg:
    def __init__(self, params=None, **kw):
        p = {"eps": 1e-15, "maxiter": 1000000, "eps_abs": None, "miniter": 0, "prec": None,
             "fail_if_not_converged": False}
        p.update(params or {})
        p.update(kw)
        self.params = p
        self.eps = p["eps"]
        self.eps_abs = p["eps_abs"]
        self.maxiter = p["maxiter"]
        self.miniter = p["miniter"]
        self.fail_if_not_converged = p["fail_if_not_converged"]
        self.history = []
        self.converged = None

    def modified(self, **params):
        return cg({**self.params, **params})

    def _native_ok(self, mat):
        return (getattr(mat, "eo2_ne", None) is not None and self.eps_abs is None
                and self.miniter == 0 and self.params["prec"] is None)

    def _finish(self, it, conv, rsq_last, rsq):
        self.converged = conv
        if not conv and self.fail_if_not_converged:
            raise ValueError("FATAL error: CG not converged")

    def _native_result(self, N, hist, iters, conv):
        """history: |r|^2 per iteration (of the slowest right-hand side when
        there are several); iterations: per right-hand side"""
        it = [int(x) for x in iters]
        self.iterations = it
        kmax = max(it) if it else 0
        worst = int(np.argmax(it)) if it else 0
        h = hist[: kmax * N].reshape(kmax, N)[:, worst] if kmax else hist[:0]
        self.history = list(h)
        self._finish(kmax, bool(all(conv)), h[-1] if kmax else 0.0, 0.0)

    def __call__(self, mat):
        if self._native_ok(mat):
            pc = mat.eo2_ne

            def inv_native(psi, src):
                N = src.grid.nrhs
                hist = np.zeros(max(self.maxiter, 1) * N)
                iters = np.zeros(N, dtype=np.int64)
                conv = np.zeros(N, dtype=np.int32)
                _lib.check(_lib.load().wc_cg_eo2_ne_block(
                    pc.op.handle, pc.D.tag, N, psi.ptr(), src.ptr(), float(self.eps), int(self.maxiter),
                    hist.ctypes.data, iters.ctypes.data, conv.ctypes.data, None, None, core.stream_ptr()))
                psi.checkerboard(pc.D)
                self._native_result(N, hist, iters, conv)

            return matrix_operator(inv_native, inv_mat=mat.mat, make_dst=mat.make_dst,
                                   accept_guess=(True, False))

        def inv(psi, src):
            if src is psi:
                raise ValueError("src must differ from psi")
            p, mmp, r = lattice(src), lattice(src), lattice(src)
            mat.mat(mmp, psi)
            core.axpy(r, -1.0, mmp, src)
            core.copy(p, r)
            cp = core.norm2(p)
            ssq = core.norm2(src)
            self.history = []
            if ssq == 0.0:
                psi[:] = 0
                self._finish(0, True, 0.0, 0.0)
                return
            rsq = self.eps**2 * ssq
            for k in range(self.maxiter):
                c = cp
                mat.mat(mmp, p)
                d = core.inner_product(p, mmp).real
                a = c / d
                cp = core.axpy_norm2(r, -a, mmp, r)
                b = cp / c
                core.axpy(psi, a, p, psi)
                core.axpy(p, b, p, r)
                res = abs(cp)
                self.history.append(res)
                if k + 1 >= self.miniter:
                    if self.eps_abs is not None and res <= self.eps_abs**2:
                        self._finish(k + 1, True, res, rsq)
                        return
                    if res <= rsq:
                        self._finish(k + 1, True, res, rsq)
                        return
            self._finish(self.maxiter, False, res, rsq)

        return matrix_operator(inv, inv_mat=mat.mat, make_dst=mat.make_dst, accept_guess=(True, False))


class cg_mixed:
    """Reliable-update CG in a lower inner precision (see the module notes).
    Only the native eo2_ne path exists: it needs the fermion operator, so use
    it inside ``preconditioned(eo2_ne(), cg_mixed(...))``."""

    def __init__(self, params=None, **kw):
        from .. import core as _core

        p = {"eps": 1e-10, "maxiter": 100000, "inner_precision": _core.single, "delta": 0.01,
             "inner_storage": None, "inner_compress_links": True, "fail_if_not_converged": False,
             "mma": "auto"}
        p.update(params or {})
        p.update(kw)
        self.params = p
        self.eps = p["eps"]
        self.maxiter = p["maxiter"]
        self.inner_precision = p["inner_precision"]
        self.delta = p["delta"]
        if p["inner_storage"] not in _lib.STORAGE_CODES:
            raise ValueError("inner_storage must be None, 'half' or 'bf16'")
        if p["inner_storage"] is not None and self.inner_precision is not _core.single:
            raise ValueError("16-bit inner storage needs inner_precision=single")
        self.inner_storage = p["inner_storage"]
        self.inner_compress_links = bool(p["inner_compress_links"])
        # FP16-storage inner hops of 4k right-hand sides on the matrix cores:
        # True, False or "auto" (the faster, timed at first use)
        self.mma = p["mma"] if p["mma"] == "auto" else bool(p["mma"])
        self.fail_if_not_converged = p["fail_if_not_converged"]
        self.iterations = []
        self.true_residual = []
        self.reliable_updates = 0
        self.converged = None
        self._inner = {}

    def inner_operator(self, op):
        key = id(op)
        if key not in self._inner:
            if self.inner_compress_links:
                params = {**op.params_constructor, "compress_links": True}
                inner = type(op)(op.name, op.U, params, op.daggered, self.inner_precision)
            else:
                inner = op.converted(self.inner_precision)
            inner.set_hop_mma(self.mma)
            self._inner[key] = (op, inner)
        return self._inner[key][1]

    def solve(self, op, parity, dst, src):
        N = src.grid.nrhs
        inner = self.inner_operator(op)
        iters = np.zeros(N, dtype=np.int64)
        conv = np.zeros(N, dtype=np.int32)
        tres = np.zeros(N)
        nupd = np.zeros(1, dtype=np.int64)
        _lib.check(_lib.load().wc_solve_eo2_ne_mixed(
            op.handle, inner.handle, parity.tag, N, dst.ptr(), src.ptr(), float(self.eps),
            int(self.maxiter), float(self.delta),
            _lib.STORAGE_CODES[self.inner_storage],
            iters.ctypes.data, conv.ctypes.data, tres.ctypes.data, nupd.ctypes.data, core.stream_ptr()))
        self.iterations = [int(x) for x in iters]
        self.true_residual = [float(x) for x in tres]
        self.reliable_updates = int(nupd[0])
        self.converged = bool(all(conv))
        if not self.converged and self.fail_if_not_converged:
            raise ValueError("FATAL error: mixed-precision CG not converged")

    def __call__(self, mat):
        raise TypeError("cg_mixed runs inside preconditioned(eo2_ne(), cg_mixed(...))")


class block_cg:
    """Block CG (block_cg.py:36-149) with the reference's parameters eps,
    maxiter, eps_abs and miniter."""

    def __init__(self, params=None, **kw):
        p = {"eps": 1e-15, "maxiter": 1000000, "eps_abs": None, "miniter": 0, "fail_if_not_converged": False}
        p.update(params or {})
        p.update(kw)
        self.params = p
        self.eps = p["eps"]
        self.eps_abs = p["eps_abs"]
        self.maxiter = p["maxiter"]
        self.miniter = p["miniter"]
        self.fail_if_not_converged = p["fail_if_not_converged"]
        self.history = []
        self.iterations = 0
        self.converged = None

    def modified(self, **params):
        return block_cg({**self.params, **params})

    def _call(self, fn, op, parity, dst, src):
        N = src.grid.nrhs
        hist = np.zeros(max(self.maxiter, 1))
        it = np.zeros(1, dtype=np.int64)
        conv = np.zeros(1, dtype=np.int32)
        _lib.check(fn(op.handle, parity.tag, N, dst.ptr(), src.ptr(), float(self.eps),
                      float(self.eps_abs or 0.0), int(self.miniter), int(self.maxiter), it.ctypes.data,
                      conv.ctypes.data, hist.ctypes.data, core.stream_ptr()))
        self.iterations = int(it[0])
        self.history = list(hist[: self.iterations])
        self.converged = bool(conv[0])
        if not self.converged and self.fail_if_not_converged:
            raise ValueError("FATAL error: block CG not converged")

    def solve(self, op, parity, dst, src):
        """dst = M^-1 src through eo2_ne (full fields)"""
        self._call(_lib.load().wc_solve_eo2_ne_block_cg, op, parity, dst, src)

    def __call__(self, mat):
        pc = getattr(mat, "eo2_ne", None)
        if pc is None:
            raise TypeError("block_cg runs on the eo2_ne normal operator of a gpt_amd fermion")

        def inv(psi, src):
            if src is psi:
                raise ValueError("src must differ from psi")
            self._call(_lib.load().wc_block_cg_eo2_ne, pc.op, pc.D, psi, src)
            psi.checkerboard(pc.D)

        return matrix_operator(inv, inv_mat=mat.mat, make_dst=mat.make_dst, accept_guess=(True, False))


class preconditioned:
    """M^-1 = L Mpc^-1 R + S (preconditioned.py:24-54)."""

    def __init__(self, preconditioner, inverter, params=None):
        self.preconditioner = preconditioner
        self.inverter = inverter

    def __call__(self, mat):
        matrix = self.preconditioner(mat)
        if isinstance(self.inverter, block_cg):
            if getattr(matrix.Mpc, "eo2_ne", None) is None:
                raise TypeError("block_cg needs the eo2_ne preconditioner of a gpt_amd fermion")

            def inv_block(dst, src):
                self.inverter.solve(matrix.op, matrix.D, dst, src)

            return matrix_operator(inv_block, inv_mat=mat.mat, make_dst=mat.make_dst,
                                   accept_guess=(True, False))
        if isinstance(self.inverter, cg_mixed):
            if getattr(matrix.Mpc, "eo2_ne", None) is None:
                raise TypeError("cg_mixed needs the eo2_ne preconditioner of a gpt_amd fermion")

            def inv_mixed(dst, src):
                self.inverter.solve(matrix.op, matrix.D, dst, src)

            return matrix_operator(inv_mixed, inv_mat=mat.mat, make_dst=mat.make_dst,
                                   accept_guess=(True, False))
        inv_mat = self.inverter(matrix.Mpc)
        native = (isinstance(self.inverter, cg) and self.inverter._native_ok(matrix.Mpc))

        def inv(dst, src):
            if native:
                # one fused call: R_ne, the guess L^-1 dst, CG, L x + S src
                it = self.inverter
                N = src.grid.nrhs
                hist = np.zeros(max(it.maxiter, 1) * N)
                iters = np.zeros(N, dtype=np.int64)
                conv = np.zeros(N, dtype=np.int32)
                _lib.check(_lib.load().wc_solve_eo2_ne_block(
                    matrix.op.handle, matrix.D.tag, N, dst.ptr(), src.ptr(), float(it.eps),
                    int(it.maxiter), hist.ctypes.data, iters.ctypes.data, conv.ctypes.data, None, None,
                    core.stream_ptr()))
                it._native_result(N, hist, iters, conv)
                return
            pc_src = matrix.R(src)
            pc_dst = lattice(pc_src)
            matrix.L.inv_mat(pc_dst, dst)
            inv_mat(pc_dst, pc_src)
            a = matrix.L(pc_dst)
            b = matrix.S(src)
            core.axpy(dst, 1.0, a, b)

        return matrix_operator(inv, inv_mat=mat.mat, make_dst=mat.make_dst, accept_guess=(True, False))


class defect_correcting:
    """FP64 outer true-residual loop around an inner inverter (defect_correcting.py:74-142)."""

    def __init__(self, inner_inverter, params=None, **kw):
        p = {"eps": 1e-15, "maxiter": 1000000}
        p.update(params or {})
        p.update(kw)
        self.eps = p["eps"]
        self.maxiter = p["maxiter"]
        self.inner_inverter = inner_inverter
        self.history = []

    def __call__(self, outer_mat):
        inner_inv_mat = self.inner_inverter(outer_mat)

        def inv(psi, src):
            n2src = core.norm2(src)
            if n2src == 0.0:
                n2src = core.norm2(outer_mat(psi)) or 1.0
            self.history = []
            s = lattice(src)
            for i in range(self.maxiter):
                outer_mat.mat(s, psi)
                core.axpy(s, -1.0, s, src)
                eps = (core.norm2(s) / n2src) ** 0.5
                self.history.append(eps)
                if eps < self.eps:
                    return
                core.scale(s, 1.0 / n2src**0.5, s)
                d = inner_inv_mat(s)
                core.axpy(psi, n2src**0.5, d, psi)

        return matrix_operator(inv, inv_mat=outer_mat.mat, make_dst=outer_mat.make_dst,
                               accept_guess=(True, False))


class mixed_precision:
    """Run an inverter on the operator converted to inner_precision, converting
    fields at the boundary (mixed_precision.py:23-32, matrix_operator.py:139-178)."""

    def __init__(self, inverter, inner_precision, outer_precision):
        self.inverter = inverter
        self.inner_precision = inner_precision
        self.outer_precision = outer_precision

    def __call__(self, mat):
        inner = mat.converted(self.inner_precision)
        inner_inv = self.inverter(inner)

        def inv(dst, src):
            s = core.convert(src, self.inner_precision)
            d = core.convert(dst, self.inner_precision)
            inner_inv(d, s)
            core.convert(dst, d)

        return matrix_operator(inv, inv_mat=mat.mat, make_dst=mat.make_dst, accept_guess=(True, False))
