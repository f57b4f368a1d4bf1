"""gpt.random restated for the oracle (test infrastructure only).

Mirrors lib/gpt/core/random.py:23-148: one engine per seed string, one set of
per-block parallel streams per grid (lib/cgpt/lib/random/engine.h:48-97),
lattice fills of ``uniform_real``, ``normal`` and ``cnormal``, and
``element`` for SU(3) links (random.py:110-148).  The bit-level generator is
``gpt_rng.c``.

Pinned by the reference's known answers (tests/random/simple.py:18-31 and
:69-130, reproduced in tests/test_oracle_rng.py): the plaquette of
``gauge.random`` at two scales and ten normal draws at five sites after
prescribed sequences.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ENGINES = {"vectorized_ranlux24_389_64": 389, "vectorized_ranlux24_24_64": 24}


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libgpt_rng.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle RNG library missing: build it with `make -C {_HERE}`")
        lib = ctypes.CDLL(path)
        lib.gr_create.restype = ctypes.c_void_p
        lib.gr_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_long),
                                  ctypes.POINTER(ctypes.c_long), ctypes.c_int, ctypes.c_int]
        lib.gr_destroy.argtypes = [ctypes.c_void_p]
        lib.gr_fill.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int, ctypes.c_void_p]
        lib.gr_raw.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_long, ctypes.c_void_p]
        lib.gr_sha256_words.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p]
        _LIB = lib
    return _LIB


class grid:
    """A lattice geometry; distinct grid objects own distinct RNG streams
    (gpt keys its parallel streams by grid object, engine.h:76-84)."""

    def __init__(self, fdims, precision="double"):
        self.fdimensions = list(fdims)
        self.gdimensions = list(fdims)
        self.precision = precision
        self.shape = tuple(reversed(self.fdimensions))
        self.gsites = int(np.prod(self.fdimensions))


class random:
    def __init__(self, seed, engine="vectorized_ranlux24_389_64"):
        self.seed = seed
        self.lux = ENGINES[engine]
        self._streams = {}

    def __del__(self):
        lib = _LIB
        if lib is None:
            return
        for h in self._streams.values():
            lib.gr_destroy(h[1])

    def _handle(self, g):
        key = id(g)
        if key not in self._streams:
            nd = len(g.fdimensions)
            fd = (ctypes.c_long * nd)(*g.fdimensions)
            gd = (ctypes.c_long * nd)(*g.gdimensions)
            h = _lib().gr_create(self.seed.encode(), fd, gd, nd, self.lux)
            if not h:
                raise ValueError("grid dimensions must be even for the block streams")
            self._streams[key] = (g, h)  # keep g alive so id() stays unique
        return self._streams[key][1]

    def _fill(self, g, dist, lo, hi, n_per_site):
        out = np.empty((g.gsites, n_per_site), dtype=np.complex128)
        _lib().gr_fill(self._handle(g), dist, lo, hi, n_per_site, out.ctypes.data)
        out = out.reshape(g.shape + (n_per_site,))
        if g.precision == "single":
            out = out.astype(np.complex64)
        return out

    def uniform_real(self, g, inner_shape=(), min=0.0, max=1.0):
        n = int(np.prod(inner_shape)) if inner_shape else 1
        return self._fill(g, 0, min, max, n).reshape(g.shape + tuple(inner_shape))

    def normal(self, g, inner_shape=(), mu=0.0, sigma=1.0):
        n = int(np.prod(inner_shape)) if inner_shape else 1
        return self._fill(g, 1, mu, sigma, n).reshape(g.shape + tuple(inner_shape))

    def cnormal(self, g, inner_shape=(), mu=0.0, sigma=1.0):
        n = int(np.prod(inner_shape)) if inner_shape else 1
        return self._fill(g, 2, mu, sigma, n).reshape(g.shape + tuple(inner_shape))

    def vspincolor(self, g):
        """rng.cnormal(g.vspincolor(grid)): 12 complex per site, spin-major."""
        return self.cnormal(g, (4, 3))

    def element(self, g, scale=1.0):
        """One SU(3) link field: exp(i sum_a scale u_a T_a), u_a ~ U[-1/2, 1/2)
        drawn one generator at a time (random.py:110-148)."""
        from . import gauge

        alg = np.zeros(g.shape + (3, 3), dtype=np.complex128)
        for ta in gauge.su3_generators():
            ca = self.uniform_real(g, (), -0.5, 0.5)
            alg += (scale * ca.astype(np.complex128))[..., None, None] * ta
        u = gauge.gpt_matrix_exp(1j * alg)
        if g.precision == "single":
            u = u.astype(np.complex64)
        return u


def sha256_words(data: bytes):
    out = (ctypes.c_uint32 * 8)()
    _lib().gr_sha256_words(data, len(data), out)
    return list(out)
