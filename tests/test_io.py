"""Configuration IO pieces that run on the host (lib/gpt/core/io/nersc_io.py,
lib/gpt/core/io/lime_io/, lib/cgpt/lib/checksums/nersc.h, lib/cgpt/lib/munge.h,
lib/cgpt/lib/io/openQCD.h): the NERSC checksum against a word-by-word sum and
closed forms, the third-row reconstruction of SU(3) links, NERSC header
parsing, the SciDAC checksums and LIME record layout, and openQCD's point
order and header test."""
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


def test_crc_and_scidac_checksums():
    """the CRC is CRC-32 (zlib / ISO-HDLC: check value of "123456789" is
    0xCBF43926); one site is not rotated; site i is rotated left by i mod 29
    (suma) and i mod 31 (sumb), so the checksums see where each site sits"""
    import zlib

    assert zlib.crc32(b"123456789") == 0xCBF43926
    site = bytes(range(64))
    c = zlib.crc32(site)
    assert io.scidac_checksums(site, 64) == (c, c)
    a, b = io.scidac_checksums(site * 2, 64)
    rot1 = ((c << 1) | (c >> 31)) & 0xFFFFFFFF
    assert a == b == c ^ rot1
    # site 29: suma unrotated, sumb rotated by 29 (= right by 3)
    data = bytes(64) * 29 + site
    z = zlib.crc32(bytes(64))
    za = zb = 0
    for i in range(29):
        za ^= ((z << (i % 29)) | (z >> (32 - i % 29))) & 0xFFFFFFFF if i % 29 else z
        zb ^= ((z << (i % 31)) | (z >> (32 - i % 31))) & 0xFFFFFFFF if i % 31 else z
    a, b = io.scidac_checksums(data, 64)
    assert a == za ^ c
    assert b == zb ^ (((c >> 3) | (c << 29)) & 0xFFFFFFFF)
    # swapping two different sites changes the checksums
    assert io.scidac_checksums(site + bytes(64), 64) != io.scidac_checksums(bytes(64) + site, 64)


def test_lime_records(tmp_path):
    fn = str(tmp_path / "f.lime")
    w = io.lime_writer(fn)
    w.text("first-record", "hello")
    w.record("payload-binary-data", bytes(range(13)))
    w.xml("meta", "metaRoot", {"a": 1, "b": "x"})
    w.close()
    raw = open(fn, "rb").read()
    # record header: magic, version 1, flags, 64-bit size, 128-byte tag; payloads padded to 8 bytes
    assert raw[:4] == bytes([0x45, 0x67, 0x89, 0xAB])
    assert struct.unpack(">H", raw[4:6])[0] == 1
    assert struct.unpack(">Q", raw[8:16])[0] == 5
    assert raw[16:16 + 13] == b"first-record\0"
    assert len(raw) % 8 == 0
    r = io.lime_reader(fn)
    assert list(r.index) == ["first-record", "payload-binary-data", "meta"]
    assert r.index["first-record"] == (144, 5)
    assert r.index["payload-binary-data"] == (144 + 8 + 144, 13)
    assert r.text("first-record") == "hello"
    assert r.read("payload-binary-data") == bytes(range(13))
    assert r.xml_dict("meta") == {"a": "1", "b": "x"}
    assert io.lime_reader.is_file(fn) and not io.lime_reader.is_file(str(tmp_path / "missing"))


def test_openqcd_point_order_and_header(tmp_path):
    fd = [4, 6, 2, 8]  # X, Y, Z, T
    x, y, z, t = io.openqcd_odd_points(fd)
    assert len(x) == np.prod(fd) // 2
    assert ((x + y + z + t) % 2 == 1).all()
    # x3 (z) fastest, then x2 (y), x1 (x), x0 (t) slowest
    assert (x[0], y[0], z[0], t[0]) == (0, 0, 1, 0)
    assert (x[1], y[1], z[1], t[1]) == (0, 1, 0, 0)
    key = ((t * fd[0] + x) * fd[1] + y) * fd[2] + z
    assert (np.diff(key) > 0).all()
    fn = str(tmp_path / "cfg")
    open(fn, "wb").write(struct.pack("<iiiid", 8, 4, 6, 2, 1.8) + bytes(8))
    assert io.read_openqcd_header(fn) == ([4, 6, 2, 8], 1.8)
    open(fn, "wb").write(b"BEGIN_HEADER\nDIMENSION_1 = 4\nEND_HEADER\n")
    assert io.read_openqcd_header(fn) is None
