"""ctypes front end of the C/OpenMP restatement ``wc_cpu.c`` (oracle and CPU
baseline; test infrastructure only).

``half(field, p)`` / ``full(halves)`` convert between the numpy oracle's
natural layout [t, z, y, x, 4, 3] and half-lattice arrays [Vh, 4, 3] whose
index is (((t Lz + z) Ly + y) Lx + x) // 2, the checkerboarded order of gpt
(lib/gpt/core/checkerboard.py:23-35).
"""
import ctypes
import os

import numpy as np

from . import conventions as cv

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libwc_cpu.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle CPU library missing: build it with `make -C {_HERE}`")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.wcc_create.restype = P
        lib.wcc_create.argtypes = [P, P, P, P, P, P, ctypes.c_double, ctypes.c_int]
        lib.wcc_destroy.argtypes = [P]
        lib.wcc_meooe.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int]
        lib.wcc_mooee.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int]
        lib.wcc_mpc_ne.argtypes = [P, P, P, ctypes.c_int, P]
        lib.wcc_cg_ne.restype = ctypes.c_long
        lib.wcc_cg_ne.argtypes = [P, P, P, ctypes.c_int, ctypes.c_double, ctypes.c_long, P]
        lib.wcc_threads.restype = ctypes.c_int
        lib.wcc_threads.argtypes = []
        _LIB = lib
    return _LIB


def threads():
    """size of the OpenMP team the C restatement's loops run with"""
    return int(_lib().wcc_threads())


def half_indices(shape, p):
    return np.nonzero(cv.parity(shape).reshape(-1) == p)[0]


def half(field, p):
    shape = field.shape[:4]
    flat = field.reshape((-1,) + field.shape[4:])
    return np.ascontiguousarray(flat[half_indices(shape, p)])


def full(shape, even, odd):
    V = int(np.prod(shape))
    out = np.zeros((V,) + even.shape[1:], dtype=np.result_type(even, odd))
    out[half_indices(shape, 0)] = even
    out[half_indices(shape, 1)] = odd
    return out.reshape(tuple(shape) + even.shape[1:])


class wilson_clover_cpu:
    """The EO operator of a numpy-oracle ``wilson_clover`` object, in C."""

    def __init__(self, op):
        self.op = op
        self.shape = op.shape
        L = (ctypes.c_int * 4)(*op.L)
        U = np.ascontiguousarray(np.stack(op.U_raw), dtype=np.complex128)
        ph = np.array(op.boundary_phases, dtype=np.complex128)
        c = np.array([op.nu / op.xi_0] * 3 + [1.0])
        self._keep = [U, ph, c]
        blocks = iblocks = None
        if op.clover is not None:
            up, dn = op.clover_blocks()
            iup, idn = op.clover_blocks(inverse=True)
            blocks = np.ascontiguousarray(np.stack([up, dn], axis=-3).reshape(-1, 2, 6, 6))
            iblocks = np.ascontiguousarray(np.stack([iup, idn], axis=-3).reshape(-1, 2, 6, 6))
            self._keep += [blocks, iblocks]
        self.h = _lib().wcc_create(L, U.ctypes.data, ph.ctypes.data, c.ctypes.data,
                                   None if blocks is None else blocks.ctypes.data,
                                   None if iblocks is None else iblocks.ctypes.data,
                                   0.5 / op.kappa, int(op.open_bc))
        self.Vh = int(np.prod(self.shape)) // 2

    def __del__(self):
        if _LIB is not None and getattr(self, "h", None):
            _LIB.wcc_destroy(self.h)

    def _new(self):
        return np.zeros((self.Vh, 4, 3), np.complex128)

    def meooe(self, src_half, pin, dag=False):
        out = self._new()
        src = np.ascontiguousarray(src_half, dtype=np.complex128)
        _lib().wcc_meooe(self.h, out.ctypes.data, src.ctypes.data, pin, int(dag))
        return out

    def mooee(self, src_half, p, inv=False):
        out = self._new()
        src = np.ascontiguousarray(src_half, dtype=np.complex128)
        _lib().wcc_mooee(self.h, out.ctypes.data, src.ctypes.data, p, int(inv))
        return out

    def mpc_ne(self, src_half, pd=1):
        out = self._new()
        scr = np.zeros((3 * self.Vh, 4, 3), np.complex128)
        src = np.ascontiguousarray(src_half, dtype=np.complex128)
        _lib().wcc_mpc_ne(self.h, out.ctypes.data, src.ctypes.data, pd, scr.ctypes.data)
        return out

    def cg_ne(self, b_half, pd=1, eps=1e-8, maxiter=10000, x0=None, history=False):
        x = self._new() if x0 is None else np.array(x0, dtype=np.complex128, copy=True)
        b = np.ascontiguousarray(b_half, dtype=np.complex128)
        hist = np.zeros(maxiter) if history else None
        it = _lib().wcc_cg_ne(self.h, x.ctypes.data, b.ctypes.data, pd, eps, maxiter,
                              None if hist is None else hist.ctypes.data)
        if history:
            return x, int(it), hist[: abs(int(it))]
        return x, int(it)

    def solve(self, src, pd=1, eps=1e-8, maxiter=10000):
        """dst = M^-1 src through eo2_ne + CG on natural-layout fields."""
        pc = 1 - pd
        s_d, s_c = half(src, pd), half(src, pc)
        # R = s_D - Meooe A_C^-1 s_C ; R_ne = Mpc^dag R
        r = s_d - self.meooe(self.mooee(s_c, pc, inv=True), pc)
        # Mpc^dag r
        t = self.meooe(r, pd, dag=True)
        t = self.mooee(t, pc, inv=True)
        t = self.meooe(t, pc, dag=True)
        t = self.mooee(t, pd, inv=True)
        rne = r - t
        x, it = self.cg_ne(rne, pd, eps, maxiter)
        d_d = self.mooee(x, pd, inv=True)
        d_c = -self.mooee(self.meooe(d_d, pd), pc, inv=True) + self.mooee(s_c, pc, inv=True)
        ev, od = (d_d, d_c) if pd == 0 else (d_c, d_d)
        return full(self.shape, ev, od), it
