"""NERSC configuration IO pieces that run on the host (lib/gpt/core/io/nersc_io.py,
lib/cgpt/lib/checksums/nersc.h, lib/cgpt/lib/munge.h): the checksum against a
word-by-word sum and closed forms, the third-row reconstruction of SU(3)
links, and header parsing."""
import struct

import numpy as np

from gpt_amd import io
from oracle import gauge as G
from oracle import rng as R


def test_checksum():
    assert io.checksum(np.zeros(16)) == 0
    # 1.0 is 0x3FF0000000000000: words 0x00000000 and 0x3FF00000 (little-endian host)
    assert io.checksum(np.array([1.0])) == 0x3FF00000
    x = np.random.default_rng(1).standard_normal(1001)
    words = struct.unpack(f"<{2 * x.size}I", x.astype("<f8").tobytes())
    assert io.checksum(x) == sum(words) % (1 << 32)
    y = x.astype(np.float32)
    words = struct.unpack(f"<{y.size}I", y.astype("<f4").tobytes())
    assert io.checksum(y) == sum(words) % (1 << 32)


def test_third_row_of_su3():
    U = G.random_links(R.grid([4, 4, 4, 4]), R.random("io"), scale=1.0)[0].reshape(-1, 3, 3)
    V = io.reconstruct_third_row(U[:, :2, :])
    assert np.abs(V - U).max() < 1e-14


def test_header(tmp_path):
    p = tmp_path / "cfg"
    p.write_bytes(b"BEGIN_HEADER\nDIMENSION_1 = 4\nPLAQUETTE  = 0.5\nCHECKSUM =   1f\nEND_HEADER\n" + bytes(8))
    meta, off = io.read_header(str(p))
    assert meta["DIMENSION_1"] == "4" and meta["PLAQUETTE"] == "0.5" and meta["CHECKSUM"] == "1f"
    assert off == len(p.read_bytes()) - 8
    q = tmp_path / "other"
    q.write_bytes(b"\x00\x01binary")
    assert io.read_header(str(q)) is None
    assert io.read_header(str(tmp_path / "missing")) is None
