"""The HBM bandwidth probe (wc_bandwidth_probe) that bench.py reports beside
the hop: the copy moves every byte, including a tail that is not a whole
number of 4-access strides, the read probe leaves its source alone and
writes the XOR fold of the source, and bad arguments are refused."""
import functools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import gpt_amd  # noqa: E402,F401
from gpt_amd import _lib, core  # noqa: E402


def probe(kind, a, b, nbytes):
    _lib.check(_lib.load().wc_bandwidth_probe(kind, a.data_ptr(), b.data_ptr(), nbytes, core.stream_ptr()))
    torch.cuda.synchronize()


@pytest.mark.parametrize("nbytes", [16, 16 * 1000003, 64 << 20])
def test_copy_moves_every_byte(nbytes):
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    a = torch.randint(0, 256, (nbytes,), dtype=torch.uint8, device="cuda", generator=g)
    b = torch.zeros_like(a)
    probe(0, a, b, nbytes)
    assert torch.equal(a, b)


def test_read_folds_the_source():
    nbytes = 16 * 1000003
    a = torch.randint(0, 256, (nbytes,), dtype=torch.uint8, device="cuda")
    keep = a.clone()
    b = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
    probe(1, a, b, nbytes)
    assert torch.equal(a, keep)
    words = b[:4 * 8192].cpu().numpy().view(np.uint32)
    total = functools.reduce(np.bitwise_xor, a.cpu().numpy().view(np.uint32))
    assert functools.reduce(np.bitwise_xor, words) == total


def test_bad_arguments_are_refused():
    a = torch.zeros(1 << 16, dtype=torch.uint8, device="cuda")
    b = torch.zeros_like(a)
    lib = _lib.load()
    s = core.stream_ptr()
    with pytest.raises(RuntimeError, match="multiple of 16"):
        _lib.check(lib.wc_bandwidth_probe(0, a.data_ptr(), b.data_ptr(), 15, s))
    with pytest.raises(RuntimeError, match="kind"):
        _lib.check(lib.wc_bandwidth_probe(2, a.data_ptr(), b.data_ptr(), 1 << 16, s))
    with pytest.raises(RuntimeError, match="differ"):
        _lib.check(lib.wc_bandwidth_probe(0, a.data_ptr(), a.data_ptr(), 1 << 16, s))
    with pytest.raises(RuntimeError, match="aligned"):
        _lib.check(lib.wc_bandwidth_probe(0, a.data_ptr() + 8, b.data_ptr(), 1 << 10, s))
    with pytest.raises(RuntimeError, match="at least"):
        _lib.check(lib.wc_bandwidth_probe(1, a.data_ptr(), b.data_ptr(), 1024, s))
