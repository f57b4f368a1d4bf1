"""The Wilson-clover fermion operator object of gpt_amd.

Mirrors gpt's operator stack for this path:
  lib/gpt/qcd/fermion/wilson.py:63-96           wilson_clover(U, params) factory
  lib/gpt/qcd/fermion/operator/base.py:28-306   matrix_operator attributes,
                                                converted / updated / adj /
                                                propagator / even_odd_sites_decomposed
  lib/gpt/qcd/fermion/operator/interface.py:23-103  handle set-up, the operator
                                                pool (suspend / reuse via
                                                set_mass then update)
  lib/gpt/qcd/fermion/register.py:2-34          opcode numbers
The arithmetic happens in the HIP library behind include/gpt_amd_wc.h.
"""
import copy as _copy
import ctypes

import torch

from ... import _lib
from ... import core
from ...core import even, odd, none, full, redblack, lattice, matrix_operator
from .. import gauge

OPCODES = {
    "M": 2001, "Mdag": 2002, "Meooe": 2003, "MeooeDag": 2004, "Mooee": 2005, "MooeeDag": 2006,
    "MooeeInv": 2007, "MooeeInvDag": 2008, "Mdiag": 2009, "Dminus": 2010, "DminusDag": 2011,
    "ImportPhysicalFermionSource": 2012, "ImportUnphysicalFermion": 2013,
    "ExportPhysicalFermionSolution": 2014, "ExportPhysicalFermionSource": 2015,
    "Dhop": 3001, "DhopDag": 4001, "DhopEO": 3002, "DhopEODag": 4002,
}
FULL_FIELD = {2001, 2002, 2009, 2010, 2011, 2012, 2013, 2014, 2015, 3001, 4001}

# pool of suspended operator handles by tag (interface.py:23-24)
operator_tag = {}
operator_limbo = {}


def release_pool():
    """destroy every parked operator handle (their links and clover); the
    pool is refilled as operators are dropped.  Returns how many there were."""
    n = 0
    for tag, handles in operator_limbo.items():
        while handles:
            h = handles.pop()
            operator_tag.pop(h, None)
            _lib.check(_lib.load().wc_destroy(h))
            n += 1
    return n


def _phases(bp):
    out = []
    for b in bp:
        c = complex(b)
        out.append((c.real, c.imag))
    return out


def _iteration_runs(handle):
    """the C operator's cumulative (graph, direct) iteration-run counters"""
    g_, d_ = ctypes.c_long(), ctypes.c_long()
    _lib.check(_lib.load().wc_iteration_runs(handle, ctypes.addressof(g_), ctypes.addressof(d_)))
    return int(g_.value), int(d_.value)


class interface:
    """Owns one wc_operator handle (lib/gpt/qcd/fermion/operator/interface.py:29-103)."""

    def __init__(self):
        self.obj = None

    def _setup(self, grid, params, U):
        assert self.obj is None
        lib = _lib.load()
        tag_params = {k: params[k] for k in params if k not in ["mass"]}
        tag = f"wilson_clover_{grid.precision.__name__}_{tuple(grid.fdimensions)}_{tag_params}"
        Ut = gauge.stacked(U)
        stream = core.stream_ptr()
        if tag in operator_limbo and operator_limbo[tag]:
            self.obj = operator_limbo[tag].pop()
            # set mass needs to precede update for clover-type fermions (interface.py:43-45)
            _lib.check(lib.wc_set_mass(self.obj, float(params["mass"]), stream))
            _lib.check(lib.wc_update(self.obj, Ut.data_ptr(), stream))
            _lib.check(lib.wc_set_iteration_chunk(self.obj, 0))  # a new operator's default schedule
            _lib.check(lib.wc_set_hop_mma(self.obj, 2))  # and its default (automatic) hop form
        else:
            p = _lib.wc_params()
            for i in range(4):
                p.fdims[i] = grid.fdimensions[i]
            p.mass = float(params["mass"])
            p.csw_r = float(params["csw_r"])
            p.csw_t = float(params["csw_t"])
            p.cF = float(params.get("cF", 1.0))
            p.xi_0 = float(params["xi_0"])
            p.nu = float(params["nu"])
            p.isAnisotropic = int(bool(params.get("isAnisotropic", False)))
            for mu, (re, im) in enumerate(_phases(params["boundary_phases"])):
                p.boundary_phases[mu][0] = re
                p.boundary_phases[mu][1] = im
            p.precision = grid.precision.code
            p.compress_links = int(bool(params.get("compress_links", False)))
            h = ctypes.c_void_p()
            _lib.check(lib.wc_create(ctypes.byref(p), Ut.data_ptr(), stream, ctypes.byref(h)))
            self.obj = h.value
            operator_tag[self.obj] = tag
        # a pooled handle keeps its run counters: count from here
        self._runs0 = _iteration_runs(self.obj)
        core.torch.cuda.current_stream().synchronize()

    def setup(self, grid, params, U):
        self.setup_arguments = (grid, params, U)
        self._setup(grid, params, U)

    def __del__(self):
        self.suspend()

    def suspend(self):
        if self.obj is not None:
            # a parked operator keeps its links and clover, not its solver scratch
            try:
                _lib.check(_lib.load().wc_trim(self.obj))
            except Exception:  # interpreter shutdown: the runtime may be gone
                pass
            operator_limbo.setdefault(operator_tag[self.obj], []).append(self.obj)
            self.obj = None

    def release(self):
        """free the native operator now instead of parking it in the pool"""
        if self.obj is not None:
            operator_tag.pop(self.obj, None)
            _lib.check(_lib.load().wc_destroy(self.obj))
            self.obj = None

    def update(self, params, U):
        if self.obj is None:
            self._setup(*self.setup_arguments)
        _lib.check(_lib.load().wc_update(self.obj, gauge.stacked(U).data_ptr(), core.stream_ptr()))

    def apply_unary_operator(self, opcode, dst, src):
        assert self.obj is not None
        if dst is src:
            raise ValueError("dst must differ from src")
        if dst.grid.nrhs != src.grid.nrhs:
            raise ValueError("dst and src hold different numbers of right-hand sides")
        cb = -1 if opcode in FULL_FIELD else src.checkerboard().tag
        _lib.check(_lib.load().wc_apply_block(self.obj, opcode, src.grid.nrhs, dst.ptr(), src.ptr(), cb,
                                              core.stream_ptr()))


class fine_operator(matrix_operator):
    """A gpt-style fermion operator object (operator/base.py:28-306)."""

    def __init__(self, name, U, params, daggered=False, precision=None):
        self.name = name
        self.U = U
        self.params_constructor = params
        self.daggered = daggered
        prec = precision if precision is not None else params.get("precision", core.double)
        self.precision = prec
        fd = U[0].grid.fdimensions
        self.n_rhs = int(params.get("n_rhs", 1))
        self.U_grid = core.grid(fd, prec, full)
        self.U_grid_eo = self.U_grid.checkerboarded(redblack)
        # n_rhs > 1: fields carry n_rhs right-hand sides (wilson.py:82-86, base.py:47-60)
        self.F_grid = self.U_grid.inserted_dimension(self.n_rhs)
        self.F_grid_eo = self.F_grid.checkerboarded(redblack)
        self.params = {k: v for k, v in params.items() if k != "n_rhs"}
        self.interface = interface()
        self.interface.setup(self.U_grid, self.params, U)
        iface = self.interface

        def reg(opcode, out_cb=None):
            def f(dst, src):
                if isinstance(src, list):
                    return packed_apply(f, dst, src)
                if src.grid.precision is not prec or dst.grid.precision is not prec:
                    raise ValueError(f"field precision differs from the operator's ({prec.__name__})")
                iface.apply_unary_operator(opcode, dst, src)
                if dst.grid.cb is not full:
                    dst.checkerboard(out_cb(src.checkerboard()))
            return f

        same = lambda cb: cb  # noqa: E731
        flip = lambda cb: cb.inv()  # noqa: E731

        def make_full(src):
            return lattice(self.F_grid)

        def make_like(src):
            return lattice(src)

        def OP(x):
            return x if not daggered else x.adj()

        M = reg(OPCODES["M"])
        Mdag = reg(OPCODES["Mdag"])
        super().__init__(mat=M if not daggered else Mdag, adj_mat=Mdag if not daggered else M,
                         make_dst=make_full)
        self.Meooe = OP(matrix_operator(reg(OPCODES["Meooe"], flip), reg(OPCODES["MeooeDag"], flip),
                                        make_dst=make_like))
        self.Mooee = OP(matrix_operator(reg(OPCODES["Mooee"], same), reg(OPCODES["MooeeDag"], same),
                                        reg(OPCODES["MooeeInv"], same), reg(OPCODES["MooeeInvDag"], same),
                                        make_dst=make_like))
        self.DhopEO = OP(matrix_operator(reg(OPCODES["DhopEO"], flip), reg(OPCODES["DhopEODag"], flip),
                                         make_dst=make_like))
        self.Dhop = OP(matrix_operator(reg(OPCODES["Dhop"]), reg(OPCODES["DhopDag"]), make_dst=make_full))
        self.Mdiag = OP(matrix_operator(reg(OPCODES["Mdiag"]), make_dst=make_full))
        self.Dminus = OP(matrix_operator(reg(OPCODES["Dminus"]), reg(OPCODES["DminusDag"]),
                                         make_dst=make_full))
        for nm in ["ImportPhysicalFermionSource", "ImportUnphysicalFermion",
                   "ExportPhysicalFermionSolution", "ExportPhysicalFermionSource"]:
            setattr(self, nm, OP(matrix_operator(reg(OPCODES[nm]), make_dst=make_full)))

        def _G5M(dst, src):
            # gamma5 = diag(1, 1, -1, -1): negate the lower spins (components 6..11)
            self.mat(dst, src)
            d = dst.data
            d.view(-1, 12, d.shape[-3] * d.shape[-2] * 2)[:, 6:12].neg_()

        self.G5M = matrix_operator(_G5M, make_dst=make_full)

    def set_mass(self, mass):
        """change the mass in place: wc_set_mass rebuilds the clover blocks
        and their inverses (gpt re-masses a pooled operator the same way,
        lib/gpt/qcd/fermion/operator/interface.py:41-48)"""
        self.params["mass"] = float(mass)
        self.params_constructor = dict(self.params_constructor, mass=float(mass))
        _lib.check(_lib.load().wc_set_mass(self.handle, float(mass), core.stream_ptr()))

    # native handle for the fused Schur / CG kernels
    @property
    def handle(self):
        if self.interface.obj is None:
            self.interface._setup(*self.interface.setup_arguments)
        return self.interface.obj

    # hop launch shapes (rb, tpb, occ, seg, ls): rb right-hand sides per
    # thread, 64 * tpb threads per workgroup, at most occ workgroups per CU (0:
    # no cap), seg t-marching runs per time slice (1: natural order), ls 1: a
    # workgroup stages its sites' links and clover blocks in LDS, ls 2 (one
    # right-hand side): it stages the input spinors of its own site range
    # (wc_tune_hop / wc_set_hop_shape / wc_get_hop_shape); every shape gives
    # bitwise the same result
    def tune_hop(self, n_rhs=None, reps=5):
        """time every launch shape of the hop for n_rhs right-hand sides and keep
        the fastest; returns (shape, milliseconds per call)"""
        n = self.n_rhs if n_rhs is None else int(n_rhs)
        sh = (ctypes.c_int * 5)()
        ms = ctypes.c_float()
        _lib.check(_lib.load().wc_tune_hop(self.handle, n, int(reps), ctypes.addressof(sh),
                                           ctypes.addressof(ms), core.stream_ptr()))
        return tuple(sh), float(ms.value)

    def hop_shape(self, n_rhs=None):
        n = self.n_rhs if n_rhs is None else int(n_rhs)
        sh = (ctypes.c_int * 5)()
        _lib.check(_lib.load().wc_get_hop_shape(self.handle, n, ctypes.addressof(sh)))
        return tuple(sh)

    def set_hop_shape(self, shape, n_rhs=None):
        """shape = (rb, tpb, occ, seg, ls), or None for automatic"""
        n = self.n_rhs if n_rhs is None else int(n_rhs)
        if shape is None:
            _lib.check(_lib.load().wc_set_hop_shape(self.handle, n, None))
            return
        sh = (ctypes.c_int * 5)(*[int(x) for x in shape])
        _lib.check(_lib.load().wc_set_hop_shape(self.handle, n, ctypes.addressof(sh)))

    # FP16-storage hops of 4k right-hand sides on the matrix cores
    # (wc_set_hop_mma): True, False, or "auto" (the default: the faster form,
    # timed at first use on lattices where the solvers tune)
    def set_hop_mma(self, mode):
        code = 2 if mode == "auto" else int(bool(mode))
        _lib.check(_lib.load().wc_set_hop_mma(self.handle, code))

    def hop_mma(self, n_rhs=None):
        """True if FP16-storage hops of n_rhs right-hand sides run on the matrix cores now"""
        n = self.n_rhs if n_rhs is None else int(n_rhs)
        f = ctypes.c_int()
        _lib.check(_lib.load().wc_get_hop_mma(self.handle, n, ctypes.byref(f)))
        return bool(f.value)

    # the inner solve's pieces with 16-bit storage (wc_apply_block_s16,
    # wc_apply_mpc_block_s16), on torch tensors holding (re, im) pairs of
    # float16 (storage "half") or bfloat16 ("bf16") in a half field's layout
    def hop_s16(self, src, src_cb, dag=False, storage="half", n_rhs=None, dst=None):
        """Meooe (MeooeDag) of the 16-bit half field src of parity src_cb"""
        n = self.n_rhs if n_rhs is None else int(n_rhs)
        dst = torch.empty_like(src) if dst is None else dst
        _lib.check(_lib.load().wc_apply_block_s16(self.handle, 2004 if dag else 2003, n,
                                                  _lib.STORAGE_CODES[storage], dst.data_ptr(), src.data_ptr(),
                                                  int(src_cb), core.stream_ptr()))
        return dst

    def mpc_ne_s16(self, src, parity, storage="half", n_rhs=None):
        """Mpc^dag Mpc (D = parity) of the 16-bit src with 16-bit Schur
        intermediates, as an FP32 tensor of the same layout"""
        n = self.n_rhs if n_rhs is None else int(n_rhs)
        dst = torch.empty(src.shape, dtype=torch.float32, device=src.device)
        _lib.check(_lib.load().wc_apply_mpc_block_s16(self.handle, int(parity), n, _lib.STORAGE_CODES[storage],
                                                      dst.data_ptr(), src.data_ptr(), core.stream_ptr()))
        return dst

    # halo of a T-slab decomposition (wc_halo_pack / wc_halo_unpack): with
    # this operator's lattice as one rank's slab, the spin-projected t-faces
    # a hop would send to the ranks below and above, and the t-terms the
    # received faces add to the two boundary slices
    def halo_pack(self, src, dag=False, out=None):
        """(lo, hi) face tensors of the half field src for the hop into parity
        1 - src's: lo the projected t = 0 slice, hi U_t^dag times the projected
        t = Lt - 1 slice; 6 * nf * n_rhs complex values each, entry
        (3 * spin + colour, face site, rhs).  out: (lo, hi) to write into."""
        self._check_halo_field(src)
        fd = src.grid.fdimensions
        nf = fd[0] * fd[1] * fd[2] // 2
        n = src.grid.nrhs
        dt = core._complex_dtype(src.grid.precision)
        if out is None:
            lo = core.torch.empty(6 * nf * n, dtype=dt, device=core.device())
            hi = core.torch.empty_like(lo)
        else:
            lo, hi = out
            self._check_faces(lo, hi, dt, 6 * nf * n)
        pout = 1 - src.checkerboard().tag
        _lib.check(_lib.load().wc_halo_pack(self.handle, pout, int(bool(dag)), n, src.ptr(), lo.data_ptr(),
                                            hi.data_ptr(), core.stream_ptr()))
        return lo, hi

    def halo_unpack(self, lo, hi, dst, dag=False):
        """dst (a half field) += the t-terms of its t = Lt - 1 slice from lo
        (the t = 0 face of the rank above) and of its t = 0 slice from hi (the
        t = Lt - 1 face of the rank below)"""
        self._check_halo_field(dst)
        fd = dst.grid.fdimensions
        n = dst.grid.nrhs
        self._check_faces(lo, hi, core._complex_dtype(dst.grid.precision), 6 * (fd[0] * fd[1] * fd[2] // 2) * n)
        _lib.check(_lib.load().wc_halo_unpack(self.handle, dst.checkerboard().tag, int(bool(dag)), n,
                                              lo.data_ptr(), hi.data_ptr(), dst.ptr(), core.stream_ptr()))

    def halo_interior(self, src, dag=False, dst=None):
        """the hop of the half field src into parity 1 - src's without the
        t-terms that the halo brings (wc_halo_interior); halo_unpack of
        halo_pack(src) into the result completes the hop"""
        self._check_halo_field(src)
        if dst is None:
            dst = lattice(src)
        else:
            self._check_halo_field(dst)
            if dst.grid.nrhs != src.grid.nrhs or dst is src:
                raise ValueError("dst must be a different field of the same right-hand-side count")
        dst.checkerboard(src.checkerboard().inv())
        _lib.check(_lib.load().wc_halo_interior(self.handle, 1 - src.checkerboard().tag, int(bool(dag)),
                                                src.grid.nrhs, dst.ptr(), src.ptr(), core.stream_ptr()))
        return dst

    def halo_hop(self, src, dst, exchange, comm_stream=None, dag=False):
        """one T-slab rank's hop with its face exchange overlapped: pack the
        t-faces, run exchange(lo, hi) -> (lo_in, hi_in) on comm_stream (a
        second stream at the default priority) while the interior hop runs
        on the current stream, then unpack the received faces into dst.
        exchange stands for the transfer to and from the neighbouring ranks
        (on one GPU, with periodic t, a copy of the rank's own faces).
        dst=None returns a new field."""
        torch = core.torch
        main = torch.cuda.current_stream()
        lo, hi = self.halo_pack(src, dag=dag)
        comm = comm_stream if comm_stream is not None else torch.cuda.Stream()
        comm.wait_stream(main)
        with torch.cuda.stream(comm):
            lo_in, hi_in = exchange(lo, hi)
        lo.record_stream(comm)
        hi.record_stream(comm)
        dst = self.halo_interior(src, dag=dag, dst=dst)
        main.wait_stream(comm)
        lo_in.record_stream(main)
        hi_in.record_stream(main)
        self.halo_unpack(lo_in, hi_in, dst, dag=dag)
        return dst

    def _check_halo_field(self, x):
        if x.grid.cb is not core.redblack:
            raise ValueError("the halo calls take a half (checkerboarded) field")
        if x.grid.precision is not self.precision or list(x.grid.fdimensions) != list(self.U_grid.fdimensions):
            raise ValueError("field precision or lattice differs from the operator's")

    @staticmethod
    def _check_faces(lo, hi, dt, size):
        for f in (lo, hi):
            if f.dtype != dt or f.numel() != size or not f.is_contiguous() or f.device != core.device():
                raise ValueError(f"a halo face must be a contiguous {dt} device tensor of {size} values")

    # CG iteration schedule (wc_set_iteration_chunk): iterations run in
    # captured graphs of up to `chunk` iterations with the stopping test on the
    # device; 0 automatic, 1 one at a time (host test after each iteration)
    def set_iteration_chunk(self, chunk):
        _lib.check(_lib.load().wc_set_iteration_chunk(self.handle, int(chunk)))

    def iteration_runs(self):
        """(runs replayed from graphs, runs of one iteration) of this
        operator's CG solves, and of the mixed solves it is the inner operator
        of, since it was set up"""
        g_, d_ = _iteration_runs(self.handle)
        g0, d0 = self.interface._runs0
        return g_ - g0, d_ - d0

    def kappa(self):
        p = self.params
        return 1.0 / (2.0 * (p["mass"] + 1.0 + 3.0 * p["nu"] / p["xi_0"]))

    def adj(self):
        return type(self)(self.name, self.U, self.params_constructor, not self.daggered, self.precision)

    def updated(self, U):
        return type(self)(self.name, U, self.params_constructor, self.daggered, self.precision)

    def modified(self, **params):
        return type(self)(self.name, self.U, {**self.params_constructor, **params}, self.daggered,
                          self.precision)

    def converted(self, dst_precision):
        """a second operator in another precision (operator/base.py:230-231)"""
        return type(self)(self.name, self.U, self.params_constructor, self.daggered, dst_precision)

    def update(self, U):
        self.U = U
        self.interface.update(self.params, U)

    def suspend(self):
        self.interface.suspend()

    def release(self):
        """free the operator's device memory now (it is not parked for reuse);
        it is set up again if used afterwards"""
        self.interface.release()

    def arguments(self):
        return self.U

    def packed(self):
        """the operator on lists of n_rhs single-RHS fields, applied as one
        n_rhs field (lib/gpt/core/operator/matrix_operator.py:200-239)"""
        op = self

        def mat(dst, src):
            packed_apply(op.mat, dst, src)

        return matrix_operator(mat, make_dst=lambda s: lattice(s), accept_list=True)

    def propagator(self, solver):
        """exp * inv * imp (operator/base.py:270-286); Import/Export are the identity
        for Wilson-clover."""
        inv_matrix = solver(self)

        def prop(dst, src):
            inv_matrix(dst, src)

        return matrix_operator(prop, make_dst=lambda s: lattice(self.F_grid), accept_guess=(True, False))

    def even_odd_sites_decomposed(self, parity):
        """operator/base.py:288-306: D = parity sites, C = the other parity."""
        op = self

        class even_odd_sites:
            D_parity = parity
            C_parity = parity.inv()
            DD = op.Mooee
            CC = op.Mooee
            CD = op.Meooe
            DC = op.Meooe
            fermion = op

        return even_odd_sites()


def wilson_clover(U, params=None, **kwargs):
    """g.qcd.fermion.wilson_clover (lib/gpt/qcd/fermion/wilson.py:63-96).

    U: list of four link fields (gpt_amd.qcd.gauge) or natural-layout numpy
    arrays; params as gpt: kappa or mass, csw_r, csw_t, cF, xi_0, nu,
    isAnisotropic, boundary_phases, plus gpt_amd's ``precision``
    (core.double / core.single) and ``compress_links`` (12-real links)."""
    p = {"kappa": None, "mass": None, "cF": 1.0, "use_legacy": False, "boundary_phases": None,
         "isAnisotropic": None, "csw_r": None, "csw_t": None, "nu": None, "xi_0": None, "n_rhs": 1}
    p.update(_copy.deepcopy(params or {}))
    p.update(kwargs)
    if p["kappa"] is not None:
        if p["mass"] is not None:
            raise ValueError("give kappa or mass, not both")
        p["mass"] = 1.0 / p["kappa"] / 2.0 - 4.0
    del p["kappa"]
    if p["mass"] is None:
        raise ValueError("kappa or mass is required")
    if p["boundary_phases"] is None:
        raise ValueError("boundary_phases is required")
    if complex(p["boundary_phases"][-1]) != 0.0 and p["cF"] != 1.0:
        # wilson.py:89-90 forbids cF without an open temporal boundary
        raise ValueError("cF != 1 needs an open temporal boundary")
    for k, dflt in [("csw_r", 0.0), ("csw_t", 0.0), ("nu", 1.0), ("xi_0", 1.0)]:
        if p[k] is None:
            p[k] = dflt
    if not 1 <= int(p["n_rhs"]) <= 64:
        raise ValueError("n_rhs must be in [1, 64]")
    if not isinstance(U[0], gauge.link_field):
        fd = list(reversed(U[0].shape[:4]))
        U = gauge.from_numpy(core.grid(fd, core.double), U)
    return fine_operator("wilson_clover", U, p)


def pack(fields, g=None):
    """list of single-RHS fields -> one multi-RHS field (RHS innermost)"""
    n = len(fields)
    g = (g or fields[0].grid).inserted_dimension(n)
    out = lattice(g)
    out._cb = fields[0].checkerboard()
    for r, f in enumerate(fields):
        out.set_rhs(r, f)
    return out


def unpack(field, dst=None):
    n = field.grid.nrhs
    if dst is None:
        return [field.rhs(r) for r in range(n)]
    for r in range(n):
        dst[r].data.copy_(field.data[..., r:r + 1, :])
        if dst[r].grid.cb is not full:
            dst[r].checkerboard(field.checkerboard())
    return dst


def packed_apply(mat, dst, src):
    P = pack(src)
    Q = lattice(P)
    mat(Q, P)
    unpack(Q, dst)
    return dst


def apply_open_boundaries(field):
    """zero the first and last time slice (lib/gpt/qcd/fermion/boundary_conditions.py:23-31)"""
    if isinstance(field, list):
        return [apply_open_boundaries(f) for f in field]
    a = field.to_numpy()
    a[0] = 0.0
    a[-1] = 0.0
    field.from_numpy(a)
    return field
