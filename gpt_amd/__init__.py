"""gpt_amd: an MI355X-native even-odd Wilson-clover operator and CG solve
behind gpt's API (g.qcd.fermion.wilson_clover, pc.eo2_ne(), inv.cg,
inv.preconditioned).

    import gpt_amd as g
    w = g.qcd.fermion.wilson_clover(U, mass=0.1, csw_r=1.0, csw_t=1.0,
                                    boundary_phases=[1, 1, 1, -1])
    inv, pc = g.algorithms.inverter, g.qcd.fermion.preconditioner
    slv = w.propagator(inv.preconditioned(pc.eo2_ne(), inv.cg(eps=1e-8)))
    dst = g(slv * src)

The kernels are hand-written HIP for gfx950 behind the C ABI in
include/gpt_amd_wc.h; there is no CPU fallback.
"""
import sys as _sys
import types as _types

from . import _lib  # noqa: F401
from .core import (  # noqa: F401
    single, double, even, odd, none, full, redblack, grid, lattice, vspincolor, copy, axpy, scale,
    norm2, inner_product, rank_inner_product, inner_product_norm2, axpy_norm2, convert, pick_checkerboard,
    set_checkerboard, expression,
    eval, matrix_operator, mspincolor, adj, trace, slice_ as slice,
)
from . import create  # noqa: F401
from . import qcd  # noqa: F401
from . import algorithms  # noqa: F401
from .io import format, load, save  # noqa: F401


class _callable_module(_types.ModuleType):
    """g(expr) evaluates an expression, as in gpt (lib/gpt/__init__.py)."""

    def __call__(self, first, second=None):
        return eval(first, second)


_sys.modules[__name__].__class__ = _callable_module
