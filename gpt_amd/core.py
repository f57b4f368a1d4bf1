"""Grids, lattices, expressions and BLAS-1 of gpt_amd.

Mirrors the parts of gpt's core that the Wilson-clover path uses
(lib/gpt/core/grid.py, lattice.py, checkerboard.py, transform.py,
operator/matrix_operator.py).  Field data lives in HBM as torch tensors in the
even-odd SoA layout of include/gpt_amd_wc.h; every arithmetic operation goes
through the HIP library (there is no CPU fallback).
"""
import numbers

import numpy as np
import torch

from . import _lib

# ------------------------------------------------------------- precisions


class precision:
    pass


class single(precision):
    """lib/gpt/core/precision.py:28-38"""
    nbytes = 4
    eps = 1e-7
    code = _lib.WC_SINGLE
    real_dtype = torch.float32
    complex_dtype = np.complex64


class double(precision):
    """lib/gpt/core/precision.py:41-51"""
    nbytes = 8
    eps = 1e-15
    code = _lib.WC_DOUBLE
    real_dtype = torch.float64
    complex_dtype = np.complex128


# ----------------------------------------------------------- checkerboards


class even:
    """sum of coordinates % 2 == 0 (lib/gpt/core/checkerboard.py:23-30)"""
    tag = 0

    @staticmethod
    def inv():
        return odd


class odd:
    tag = 1

    @staticmethod
    def inv():
        return even


class none:
    tag = None

    @staticmethod
    def inv():
        return none


class full:
    pass


class redblack:
    pass


# ------------------------------------------------------------------- grids

_device = None


def device():
    global _device
    if _device is None:
        if not torch.cuda.is_available():
            raise RuntimeError("gpt_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        _device = torch.device("cuda", torch.cuda.current_device())
    return _device


def stream_ptr():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(x):
    return None if x is None else int(x)


class grid:
    """A 4-d lattice geometry (lib/gpt/core/grid.py:98-284): ``fdimensions``
    [Lx, Ly, Lz, Lt], a precision, and full or red-black (even-odd) layout."""

    def __init__(self, fdimensions, precision, cb=full, nrhs=1):
        self.fdimensions = [int(x) for x in fdimensions]
        self.nrhs = int(nrhs)  # right-hand sides per site (gpt's 5-d n_rhs grid)
        if not 1 <= self.nrhs <= 64:
            raise ValueError("nrhs must be in [1, 64]")
        assert len(self.fdimensions) == 4, "gpt_amd supports 4-d grids"
        for L in self.fdimensions:
            if L < 2 or L % 2:
                raise ValueError(f"grid dimensions must be even and >= 2: {self.fdimensions}")
        self.precision = precision
        self.cb = cb
        self.gsites = int(np.prod(self.fdimensions))
        self.Vh = self.gsites // 2
        if self.Vh % 64:
            raise ValueError(f"the half volume {self.Vh} of {self.fdimensions} must be a multiple of 64")
        self.nd = 4

    def checkerboarded(self, cb):
        return grid(self.fdimensions, self.precision, cb, self.nrhs)

    def converted(self, prec):
        return grid(self.fdimensions, prec, self.cb, self.nrhs)

    def inserted_dimension(self, n_rhs):
        """the n_rhs grid of this geometry (lib/gpt/core/grid.py inserted_dimension)"""
        return grid(self.fdimensions, self.precision, self.cb, n_rhs)

    def single_rhs(self):
        return grid(self.fdimensions, self.precision, self.cb, 1)

    def __eq__(self, o):
        return (isinstance(o, grid) and self.fdimensions == o.fdimensions and self.nrhs == o.nrhs
                and self.precision is o.precision and self.cb is o.cb)

    def __hash__(self):
        return hash((tuple(self.fdimensions), self.nrhs, self.precision.__name__, self.cb.__name__))

    def __repr__(self):
        r = f", nrhs={self.nrhs}" if self.nrhs > 1 else ""
        return f"grid({self.fdimensions}, {self.precision.__name__}, {self.cb.__name__}{r})"


# ----------------------------------------------------------------- lattices


class lattice:
    """A spin-colour vector field (gpt's vspincolor) in HBM.

    Full grids hold both parities, red-black grids one parity tagged by
    ``checkerboard()``.  Storage is AoSoA in blocks of 64 half-sites with the
    right-hand side innermost:
    ``data[parity][h // 64][component][h % 64][rhs][re/im]``."""

    def __init__(self, first, second=None):
        if isinstance(first, lattice):
            g = first.grid
            cb = first._cb
        else:
            g = first
            cb = even if g.cb is redblack else none
        self.grid = g
        self._cb = cb
        blocks = (g.Vh // 64, 12, 64, g.nrhs, 2)
        shape = (2,) + blocks if g.cb is full else blocks
        self.data = torch.empty(shape, dtype=g.precision.real_dtype, device=device())

    # gpt API
    def checkerboard(self, cb=None):
        if cb is None:
            return self._cb
        if self.grid.cb is full and cb is not none:
            raise ValueError("full-grid lattice has no checkerboard")
        self._cb = cb
        return self

    @property
    def nentries(self):
        """number of complex entries"""
        return self.data.numel() // 2

    def ptr(self):
        return self.data.data_ptr()

    def __setitem__(self, key, value):
        if not (isinstance(key, slice) and key == slice(None)):
            raise NotImplementedError("only field[:] = scalar is supported")
        if value != 0:
            raise NotImplementedError("only field[:] = 0 is supported")
        self.data.zero_()

    def __imatmul__(self, expr):
        eval(self, expr)
        return self

    def __iadd__(self, other):
        eval(self, expression.of(self) + expression.of(other))
        return self

    def __isub__(self, other):
        eval(self, expression.of(self) - expression.of(other))
        return self

    def __itruediv__(self, s):
        scale(self, 1.0 / s, self)
        return self

    def __add__(self, o):
        return expression.of(self) + o

    def __radd__(self, o):
        return expression.of(o) + self

    def __sub__(self, o):
        return expression.of(self) - o

    def __rsub__(self, o):
        return expression.of(o) - self

    def __mul__(self, s):
        return expression.of(self) * s

    __rmul__ = __mul__

    def __neg__(self):
        return expression.of(self) * -1.0

    def __truediv__(self, s):
        return expression.of(self) * (1.0 / s)

    @property
    def nrhs(self):
        return self.grid.nrhs

    def rhs(self, r):
        """a single-RHS copy of right-hand side r"""
        out = lattice(self.grid.single_rhs())
        out._cb = self._cb
        out.data.copy_(self.data[..., r:r + 1, :])
        return out

    def set_rhs(self, r, field):
        if field.grid.nrhs != 1 or field.grid.Vh != self.grid.Vh or field.grid.cb is not self.grid.cb:
            raise ValueError("set_rhs needs a single-RHS field of the same geometry")
        self.data[..., r:r + 1, :].copy_(field.data)
        return self

    # host transfer (layout plumbing): natural numpy [t, z, y, x, 4, 3]
    # (a leading [nrhs] axis for multi-RHS fields)
    def to_numpy(self):
        g = self.grid
        if g.nrhs > 1:
            return np.stack([self.rhs(r).to_numpy() for r in range(g.nrhs)])
        if g.cb is full:
            nat = torch.empty((g.gsites, 12, 2), dtype=self.data.dtype, device=self.data.device)
            fd = (ctypes_int4)(*g.fdimensions)
            _lib.check(_lib.load().wc_field_export(fd, g.precision.code, self.ptr(), nat.data_ptr(),
                                                   stream_ptr()))
            a = nat.cpu().numpy()
        else:
            full_l = lattice(g.checkerboarded(full))
            full_l.data.zero_()
            full_l.data[self._cb.tag].copy_(self.data)
            return full_l.to_numpy()
        c = (a[..., 0] + 1j * a[..., 1]).astype(g.precision.complex_dtype)
        return c.reshape(tuple(reversed(g.fdimensions)) + (4, 3))

    def from_numpy(self, arr):
        g = self.grid
        arr = np.asarray(arr)
        if g.nrhs > 1:
            if arr.shape[0] != g.nrhs:
                raise ValueError(f"expected a leading axis of {g.nrhs} right-hand sides")
            for r in range(g.nrhs):
                one = lattice(g.single_rhs())
                one._cb = self._cb
                one.from_numpy(arr[r])
                self.set_rhs(r, one)
            return self
        shape = tuple(reversed(g.fdimensions)) + (4, 3)
        if arr.shape != shape:
            raise ValueError(f"expected natural layout {shape}, got {arr.shape}")
        if g.cb is not full:
            full_l = lattice(g.checkerboarded(full))
            full_l.from_numpy(arr)
            self.data.copy_(full_l.data[self._cb.tag])
            return self
        a = np.ascontiguousarray(arr.reshape(g.gsites, 12)).astype(np.complex128)
        ri = np.stack([a.real, a.imag], axis=-1)
        nat = torch.from_numpy(ri).to(device=self.data.device, dtype=self.data.dtype)
        fd = (ctypes_int4)(*g.fdimensions)
        _lib.check(_lib.load().wc_field_import(fd, g.precision.code, nat.data_ptr(), self.ptr(),
                                               stream_ptr()))
        torch.cuda.current_stream().synchronize()
        return self


import ctypes as _ct  # noqa: E402

ctypes_int4 = _ct.c_int * 4


class mspincolor_lattice:
    """gpt's mspincolor: a 12 x 12 spin-colour matrix per site (a propagator),
    held in HBM in natural layout ``data[site][s1][s2][c1][c2]`` (site
    ((t Lz + z) Ly + y) Lx + x).  It is a container: operators act on it column
    by column, column (s, c) being the vspincolor right-hand side c + 3 s
    (lib/gpt/core/object_type/container.py:330-357)."""

    def __init__(self, first):
        g = first.grid if isinstance(first, (lattice, mspincolor_lattice)) else first
        if g.cb is not full or g.nrhs != 1:
            raise ValueError("an mspincolor lives on a full single-RHS grid")
        self.grid = g
        self.data = torch.zeros((g.gsites, 4, 4, 3, 3), dtype=_complex_dtype(g.precision), device=device())

    def checkerboard(self, cb=None):
        if cb is None or cb is none:
            return none
        raise ValueError("full-grid lattice has no checkerboard")

    # columns as vspincolor fields
    def column(self, s, c):
        g = self.grid
        col = self.data[:, :, s, :, c].reshape(g.gsites, 12)
        nat = torch.view_as_real(col.contiguous()).to(g.precision.real_dtype).contiguous()
        out = lattice(g)
        fd = (ctypes_int4)(*g.fdimensions)
        _lib.check(_lib.load().wc_field_import(fd, g.precision.code, nat.data_ptr(), out.ptr(), stream_ptr()))
        torch.cuda.current_stream().synchronize()
        return out

    def set_column(self, s, c, x):
        g = self.grid
        if x.grid.cb is not full or x.grid.gsites != g.gsites or x.grid.nrhs != 1:
            raise ValueError("set_column needs a full single-RHS vspincolor of the same lattice")
        nat = torch.empty((g.gsites, 12, 2), dtype=x.data.dtype, device=x.data.device)
        fd = (ctypes_int4)(*g.fdimensions)
        _lib.check(_lib.load().wc_field_export(fd, g.precision.code, x.ptr(), nat.data_ptr(), stream_ptr()))
        self.data[:, :, s, :, c] = torch.view_as_complex(nat.contiguous()).reshape(g.gsites, 4, 3)
        return self

    def __setitem__(self, key, value):
        """src[:] = 0, and src[(x, y, z, t)] = identity (gpt's create.point)"""
        if isinstance(key, slice) and key == slice(None):
            if value != 0:
                raise NotImplementedError("only field[:] = 0 is supported")
            self.data.zero_()
            return
        x, y, z, t = key
        L = self.grid.fdimensions
        site = ((t * L[2] + z) * L[1] + y) * L[0] + x
        self.data[site] = torch.as_tensor(np.asarray(value), dtype=self.data.dtype, device=self.data.device)

    def identity(self):
        """the 12 x 12 unit matrix in [s1][s2][c1][c2] order"""
        return np.einsum("ab,cd->abcd", np.eye(4), np.eye(3))

    # host transfer: natural numpy [t, z, y, x, 4, 4, 3, 3]
    def to_numpy(self):
        return self.data.cpu().numpy().reshape(tuple(reversed(self.grid.fdimensions)) + (4, 4, 3, 3))

    def from_numpy(self, arr):
        arr = np.asarray(arr).reshape(self.grid.gsites, 4, 4, 3, 3)
        self.data.copy_(torch.from_numpy(np.ascontiguousarray(arr)).to(self.data.dtype))
        return self

    def __imatmul__(self, expr):
        eval(self, expr)
        return self

    def __add__(self, o):
        return expression.of(self) + o

    def __sub__(self, o):
        return expression.of(self) - o

    def __mul__(self, o):
        if isinstance(o, _adjoint):
            return _site_product(self, o.of, True)
        return expression.of(self) * o

    def __rmul__(self, o):
        return expression.of(self) * o

    def __neg__(self):
        return expression.of(self) * -1.0


def _complex_dtype(prec):
    return torch.complex128 if prec is double else torch.complex64


def mspincolor(g):
    """g.mspincolor(grid): a zero propagator field"""
    return mspincolor_lattice(g)


class _adjoint:
    def __init__(self, of):
        self.of = of


def adj(x):
    """g.adj of an mspincolor, for site-local contractions such as
    g.trace(prop * g.adj(prop))"""
    if isinstance(x, mspincolor_lattice):
        return _adjoint(x)
    raise NotImplementedError("g.adj is provided for mspincolor fields (operators have .adj())")


class _site_product:
    """the site-wise matrix product A B^dag of two mspincolor fields (lazy)"""

    def __init__(self, a, b, dagger):
        self.a, self.b, self.dagger = a, b, dagger


class site_scalar:
    """a complex number per site, natural order, in HBM (the result of g.trace)"""

    def __init__(self, grid, data):
        self.grid, self.data = grid, data

    def to_numpy(self):
        return self.data.cpu().numpy().reshape(tuple(reversed(self.grid.fdimensions)))


def trace(x):
    """g.trace of an mspincolor (the 12 x 12 trace per site) or of a site
    product A * g.adj(B): tr(A B^dag) = sum_ij A_ij conj(B_ij), elementwise"""
    # sums in complex128 whatever the field's precision
    if isinstance(x, _site_product):
        a, b = x.a.data.to(torch.complex128), x.b.data.to(torch.complex128)
        return site_scalar(x.a.grid, (a * b.conj()).reshape(a.shape[0], -1).sum(dim=1))
    if isinstance(x, mspincolor_lattice):
        d = x.data.to(torch.complex128)
        return site_scalar(x.grid, torch.diagonal(torch.diagonal(d, dim1=1, dim2=2), dim1=1, dim2=2).sum(dim=(1, 2)))
    raise NotImplementedError("g.trace of this object")


def slice_(x, dim):
    """g.slice(field, dim): the sum over the sites of each coordinate value of
    dimension dim (0 x, 1 y, 2 z, 3 t), a list of complex numbers"""
    if not isinstance(x, site_scalar):
        raise NotImplementedError("g.slice is provided for g.trace results")
    L = x.grid.fdimensions
    v = x.data.reshape(L[3], L[2], L[1], L[0])
    other = tuple(d for d in range(4) if d != 3 - dim)
    return [complex(c) for c in v.sum(dim=other).cpu().numpy()]


def vspincolor(g):
    return lattice(g)


def copy(dst, src=None):
    """g.copy (lib/gpt/core/transform.py)"""
    if src is None:
        src = dst
        dst = lattice(src)
    dst.checkerboard(src.checkerboard()) if dst.grid.cb is not full else None
    dst.data.copy_(src.data)
    return dst


# -------------------------------------------------------------------- BLAS-1


def _same(a, b):
    if a.grid.precision is not b.grid.precision or a.data.shape != b.data.shape:
        raise ValueError("fields of different precision or shape")


def axpy(z, a, x, y):
    """z = a x + y (lib/gpt/core/transform.py:134-148)"""
    _same(x, y)
    _same(x, z)
    a = complex(a)
    _lib.check(_lib.load().wc_axpy(x.grid.precision.code, x.nentries, a.real, a.imag, x.ptr(),
                                   y.ptr(), z.ptr(), stream_ptr()))
    if z.grid.cb is not full:
        z.checkerboard(x.checkerboard())


def scale(z, a, x):
    _same(x, z)
    a = complex(a)
    _lib.check(_lib.load().wc_scale(x.grid.precision.code, x.nentries, a.real, a.imag, x.ptr(),
                                    z.ptr(), stream_ptr()))


def norm2(x):
    """|x|^2 accumulated in FP64 (lib/gpt/core/transform.py:109-114); a list
    (one per right-hand side) for multi-RHS fields"""
    if isinstance(x, (list, tuple)):
        return [norm2(y) for y in x]
    if isinstance(x, expression):
        x = eval(x)
    if isinstance(x, mspincolor_lattice):
        return float(torch.view_as_real(x.data).to(torch.float64).square().sum())
    if x.nrhs > 1:
        return [norm2(x.rhs(r)) for r in range(x.nrhs)]
    r = _ct.c_double()
    _lib.check(_lib.load().wc_norm2(x.grid.precision.code, x.nentries, x.ptr(), _ct.byref(r),
                                    stream_ptr()))
    return r.value


def inner_product(a, b, n_block=1, use_accelerator=True):
    """<a, b> = sum conj(a) b, FP64 accumulation (lib/gpt/core/transform.py:101-106);
    with a list on either side, the matrix [<a_i, b_j>] as gpt returns it
    (n_block and use_accelerator are accepted for gpt's signature)"""
    if isinstance(a, (list, tuple)) or isinstance(b, (list, tuple)):
        la = [eval(x) if isinstance(x, expression) else x for x in (a if isinstance(a, (list, tuple)) else [a])]
        lb = [eval(x) if isinstance(x, expression) else x for x in (b if isinstance(b, (list, tuple)) else [b])]
        return np.array([[inner_product(x, y) for y in lb] for x in la], dtype=np.complex128)
    if isinstance(a, expression):
        a = eval(a)
    if isinstance(b, expression):
        b = eval(b)
    _same(a, b)
    r = (_ct.c_double * 2)()
    _lib.check(_lib.load().wc_inner_product(a.grid.precision.code, a.nentries, a.ptr(), b.ptr(), r,
                                            stream_ptr()))
    return complex(r[0], r[1])


def rank_inner_product(a, b, n_block=1, use_accelerator=True):
    """gpt's rank_inner_product: on one device the same as inner_product"""
    return inner_product(a, b, n_block, use_accelerator)


def inner_product_norm2(a, b):
    """(<a, b>, |a|^2) (lib/gpt/core/transform.py:117-130)"""
    return inner_product(a, b), norm2(a)


def axpy_norm2(z, a, x, y):
    """z = a x + y, returns |z|^2 (lib/gpt/core/transform.py:151-153)"""
    _same(x, y)
    _same(x, z)
    a = complex(a)
    r = _ct.c_double()
    _lib.check(_lib.load().wc_axpy_norm2(x.grid.precision.code, x.nentries, a.real, a.imag, x.ptr(),
                                         y.ptr(), z.ptr(), _ct.byref(r), stream_ptr()))
    return r.value


def convert(first, second):
    """g.convert: convert(dst, src) copies across precisions; convert(src,
    precision) returns a new field (lib/gpt/core/foundation/lattice/__init__.py:168-205)."""
    if isinstance(second, type) and issubclass(second, precision):
        src = first
        dst = lattice(src.grid.converted(second))
        if dst.grid.cb is not full:
            dst.checkerboard(src.checkerboard())
    else:
        dst, src = first, second
    if dst.nentries != src.nentries:
        raise ValueError("convert: shapes differ")
    _lib.check(_lib.load().wc_convert(dst.grid.precision.code, dst.ptr(), src.grid.precision.code,
                                      src.ptr(), src.nentries, stream_ptr()))
    if dst.grid.cb is not full:
        dst.checkerboard(src.checkerboard())
    return dst


def pick_checkerboard(cb, dst, src=None):
    """lib/gpt/core/checkerboard.py:38-56"""
    if src is None:
        src = dst
        dst = lattice(src.grid.checkerboarded(redblack))
    dst.data.copy_(src.data[cb.tag])
    dst.checkerboard(cb)
    return dst


def set_checkerboard(dst, src):
    """lib/gpt/core/checkerboard.py:59-67"""
    dst.data[src.checkerboard().tag].copy_(src.data)
    return dst


# ------------------------------------------------------------- expressions


class expression:
    """A linear combination sum_i c_i (op_i1 op_i2 ... field_i), evaluated
    eagerly by ``eval`` (a minimal form of lib/gpt/core/expr.py)."""

    def __init__(self, terms):
        self.terms = terms  # list of (coef, [ops], lattice)

    @staticmethod
    def of(x):
        if isinstance(x, expression):
            return x
        if isinstance(x, (lattice, mspincolor_lattice)):
            return expression([(1.0, [], x)])
        raise TypeError(f"cannot form an expression from {type(x)}")

    def __add__(self, o):
        return expression(self.terms + expression.of(o).terms)

    def __radd__(self, o):
        return expression.of(o) + self

    def __sub__(self, o):
        return expression(self.terms + [(-c, ops, f) for c, ops, f in expression.of(o).terms])

    def __rsub__(self, o):
        return expression.of(o) - self

    def __mul__(self, s):
        if isinstance(s, numbers.Number):
            return expression([(c * s, ops, f) for c, ops, f in self.terms])
        return NotImplemented

    __rmul__ = __mul__

    def __neg__(self):
        return self * -1.0

    def __truediv__(self, s):
        return self * (1.0 / s)


def _apply_ops(ops, f):
    for op in reversed(ops):
        f = op(f)
    return f


def eval(first, second=None):
    """g.eval(dst, expr) / g.eval(expr) / g(expr)."""
    if second is None:
        expr, dst = first, None
    else:
        dst, expr = first, second
    if isinstance(expr, list):
        return [eval(x) for x in expr] if dst is None else [eval(d, x) for d, x in zip(dst, expr)]
    expr = expression.of(expr)
    fields = [(c, _apply_ops(ops, f)) for c, ops, f in expr.terms]
    if isinstance(fields[0][1], mspincolor_lattice):
        # a propagator container: linear combinations on its natural layout
        acc = sum(c * f.data for c, f in fields)
        if dst is None:
            dst = mspincolor_lattice(fields[0][1])
        dst.data.copy_(acc)
        return dst
    if dst is None:
        dst = lattice(fields[0][1])
    c0, f0 = fields[0]
    if dst.grid.cb is not full:
        dst.checkerboard(f0.checkerboard())
    # combine into a temporary when dst is also an input (e.g. psi += a*p)
    aliased = any(f is dst for _, f in fields[1:])
    out = lattice(dst) if aliased else dst
    if c0 == 1.0:
        if f0 is not out:
            out.data.copy_(f0.data)
    else:
        scale(out, c0, f0)
    for c, f in fields[1:]:
        axpy(out, c, f, out)
    if out is not dst:
        dst.data.copy_(out.data)
    return dst


# --------------------------------------------------------- matrix operators


class matrix_operator:
    """Linear operator with optional adjoint and inverse
    (lib/gpt/core/operator/matrix_operator.py:35-239).  ``op * field`` forms
    an expression, ``op(src)`` returns a new field, ``op(dst, src)`` applies in
    place, ``op.mat(dst, src)`` is the raw application."""

    def __init__(self, mat, adj_mat=None, inv_mat=None, adj_inv_mat=None, make_dst=None,
                 accept_guess=(False, False), accept_list=False):
        self.mat = mat
        self.adj_mat = adj_mat
        self.inv_mat = inv_mat
        self.adj_inv_mat = adj_inv_mat
        self.make_dst = make_dst
        self.accept_guess = accept_guess
        # accept_list: mat(dst_list, src_list) takes whole lists (matrix_operator.py:288-295)
        self.accept_list = accept_list

    def adj(self):
        return matrix_operator(self.adj_mat, self.mat, self.adj_inv_mat, self.inv_mat, self.make_dst)

    def inv(self):
        return matrix_operator(self.inv_mat, self.adj_inv_mat, self.mat, self.adj_mat, self.make_dst)

    def _new_dst(self, src):
        d = self.make_dst(src) if self.make_dst else lattice(src)
        if self.accept_guess[0]:
            d.data.zero_()
        return d

    def _apply(self, src):
        if isinstance(src, mspincolor_lattice):
            return self._apply_columns(src)
        dst = self._new_dst(src)
        self.mat(dst, src)
        return dst

    def _apply_columns(self, src):
        """an mspincolor operand: its 12 columns as right-hand sides c + 3 s
        (container.py:330-357), through the list path, then reassembled"""
        cols = [src.column(s, c) for s in range(4) for c in range(3)]
        out = self(cols)
        res = mspincolor_lattice(out[0].grid)
        for s in range(4):
            for c in range(3):
                res.set_column(s, c, out[c + 3 * s])
        return res

    def __call__(self, first, second=None):
        if second is None:
            if isinstance(first, list):
                if self.accept_list:
                    dst = [self._new_dst(x) for x in first]
                    self.mat(dst, first)
                    return dst
                return [self(x) for x in first]
            return self._apply(eval(first) if isinstance(first, expression) else first)
        if isinstance(second, mspincolor_lattice):
            first.data.copy_(self._apply_columns(second).data)
            return first
        if isinstance(first, list):
            if self.accept_list:
                self.mat(first, second)
                return first
            for d, s in zip(first, second):
                self.mat(d, s)
            return first
        self.mat(first, second)
        return first

    def __mul__(self, other):
        if isinstance(other, matrix_operator):
            return matrix_operator(lambda d, s: self.mat(d, other(s)),
                                   make_dst=self.make_dst)
        if isinstance(other, list):
            return [self * x for x in other]
        e = expression.of(other)
        return expression([(c, [self._apply] + ops, f) for c, ops, f in e.terms])
