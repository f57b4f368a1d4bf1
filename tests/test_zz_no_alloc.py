"""wc_apply_mpc_block_noalloc, the entry point that allocates nothing: the
Schur pieces Mpc, Mpc^dag and Mpc^dag Mpc with their scratch in the library's
fixed per-GPU array give bitwise what wc_apply_mpc_block gives (same
kernels), on two calls in a row of which the second has the smaller input,
in FP64 and FP32; a size beyond the array is refused with an error before
anything runs."""
import numpy as np
import pytest

from oracle import gauge as G

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd as g  # noqa: E402
from gpt_amd import _lib, core  # noqa: E402
from gpt_amd.qcd.fermion.operator import pack  # noqa: E402

P = dict(mass=0.1, csw_r=1.0, csw_t=1.0, boundary_phases=[1, 1, 1, -1])


def sources(fd, n, seed, w):
    rng = np.random.default_rng(seed)
    shape = tuple(reversed(fd)) + (4, 3)
    fs = []
    for _ in range(n):
        x = g.lattice(w.F_grid_eo.single_rhs())
        x.checkerboard(g.odd)
        x.from_numpy(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        fs.append(x)
    return pack(fs)


def noalloc(w, which, n, dst, src):
    return _lib.load().wc_apply_mpc_block_noalloc(w.handle, 1, which, n, dst.data_ptr(), src.data_ptr(),
                                                  core.stream_ptr())


def with_alloc(w, which, n, dst, src):
    _lib.check(_lib.load().wc_apply_mpc_block(w.handle, 1, which, n, dst.data_ptr(), src.data_ptr(),
                                              core.stream_ptr()))


@pytest.mark.parametrize("prec", ["double", "single"])
def test_noalloc_equals_mpc_block(prec):
    fd = [8, 8, 8, 16]
    U = G.smooth_links(fd, 0.3, 9001)
    precision = g.double if prec == "double" else g.single
    for which in (2, 0, 1):
        # 3 right-hand sides, then 1 (the smaller input second)
        for n in (3, 1):
            w = g.qcd.fermion.wilson_clover(U, P, n_rhs=n, precision=precision)
            S = sources(fd, n, 9100 + n, w)
            a = torch.full_like(S.data, float("nan"))
            rc = noalloc(w, which, n, a, S.data)
            if rc:
                _lib.check(rc)
            b = torch.full_like(S.data, float("nan"))
            with_alloc(w, which, n, b, S.data)
            torch.cuda.synchronize()
            assert bool(torch.isfinite(a).all())
            assert torch.equal(a, b), (prec, which, n, float((a - b).abs().max()))
        print(f"{prec} which={which}: no-alloc equals wc_apply_mpc_block for 3 then 1 right-hand sides",
              flush=True)


def test_noalloc_refuses_sizes_beyond_its_array():
    # 16^3 x 32 FP64: two scratch fields of 3 right-hand sides need 75 MB > 64 MiB
    fd = [16, 16, 16, 32]
    U = G.smooth_links(fd, 0.3, 9201)
    w = g.qcd.fermion.wilson_clover(U, P, n_rhs=3)
    S = sources(fd, 3, 9300, w)
    d = torch.empty_like(S.data)
    rc = noalloc(w, 2, 3, d, S.data)
    assert rc == _lib.WC_ERR_ARG
    msg = _lib.load().wc_last_error().decode()
    assert "bytes of scratch needed" in msg, msg
    # Mpc^dag needs one field (37.7 MB): accepted, and equal to the allocating call
    a = torch.full_like(S.data, float("nan"))
    _lib.check(noalloc(w, 1, 3, a, S.data))
    b = torch.full_like(S.data, float("nan"))
    with_alloc(w, 1, 3, b, S.data)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


