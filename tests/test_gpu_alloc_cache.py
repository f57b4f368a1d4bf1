"""The library's device blocks are cached by size when freed
(wc_release_cached_memory): an operator released and created again on the
same lattice takes its blocks back from the cache, allocating no new device
memory, and computes the same bits; flushing the cache returns the blocks
to the runtime (the device's free memory grows by what was reported)."""
import ctypes

import numpy as np
import pytest

from oracle import gauge as G
from oracle import rng as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd import _lib  # noqa: E402

FD = [8, 8, 8, 16]
P = dict(mass=0.1, csw_r=1.0, csw_t=1.0, xi_0=1.0, nu=1.0, boundary_phases=[1, 1, 1, -1])


def release_cache():
    n = ctypes.c_long()
    _lib.check(_lib.load().wc_release_cached_memory(ctypes.byref(n)))
    torch.cuda.synchronize()
    return n.value


def meooe(U, src):
    w = g.qcd.fermion.wilson_clover(U, P)
    s = g.lattice(w.F_grid)
    s.from_numpy(src)
    h = g.pick_checkerboard(g.odd, s)
    out = w.Meooe(h).data.clone()
    torch.cuda.synchronize()
    return w, out


def test_released_operator_blocks_are_reused_then_flushed():
    rng = R.random("alloc_cache")
    U = G.random_links(R.grid(FD), rng, scale=0.5)
    src = rng.vspincolor(R.grid(FD))
    release_cache()
    w, ref = meooe(U, src)
    w.release()  # its blocks go to the cache, not to the runtime
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    w, out = meooe(U, src)  # the same sizes again: served from the cache
    free1, _ = torch.cuda.mem_get_info()
    assert torch.equal(out, ref)
    links_bytes = 2 * (int(np.prod(FD)) // 2) * 8 * 9 * 16
    assert free0 - free1 < links_bytes, (free0, free1)
    w.release()
    torch.cuda.synchronize()
    free2, _ = torch.cuda.mem_get_info()
    released = release_cache()
    free3, _ = torch.cuda.mem_get_info()
    assert released >= links_bytes
    assert free3 - free2 >= links_bytes
    assert release_cache() == 0


def test_parked_operators_are_destroyed_by_release_pool():
    """an operator dropped without release() parks its handle for reuse
    (gpt's operator pool, interface.py:23-80); release_pool destroys the
    parked handles, and a new operator then computes the same bits"""
    import gc

    from gpt_amd.qcd.fermion import operator as fop

    rng = R.random("alloc_pool")
    U = G.random_links(R.grid(FD), rng, scale=0.5)
    src = rng.vspincolor(R.grid(FD))
    w, ref = meooe(U, src)
    del w
    gc.collect()
    assert fop.release_pool() >= 1
    assert fop.release_pool() == 0
    w, out = meooe(U, src)
    assert torch.equal(out, ref)
    w.release()
