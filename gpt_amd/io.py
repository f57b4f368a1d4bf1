"""Gauge-configuration IO: the NERSC, LIME (ILDG and gpt's own gauge
records) and openQCD formats.

Mirrors gpt's ``g.load(fn)`` / ``g.save(fn, U, fmt)``
(lib/gpt/core/io/load.py:50-73, save.py:23-73).

NERSC (nersc_io.py):

* header: ``BEGIN_HEADER`` / ``key = value`` lines / ``END_HEADER``;
  DIMENSION_1..4, DATATYPE (``4D_SU3_GAUGE_3x3``, or ``4D_SU3_GAUGE`` with
  two rows stored and the third rebuilt as conj(row0 x row1)), FLOATING_POINT
  (IEEE64BIG / IEEE64LITTLE / IEEE64 / IEEE32BIG / IEEE32LITTLE / IEEE32),
  CHECKSUM, PLAQUETTE, LINK_TRACE;
* data: per site (x fastest, then y, z, t) the 4 links mu = 0..3, each
  row-major 3x3 (or 2x3) complex;
* CHECKSUM: the sum modulo 2^32 of the data's 32-bit words in host byte
  order after conversion to host floats (lib/cgpt/lib/checksums/nersc.h);
* after reading, the plaquette and link trace are compared with the header
  to the printed digits (at least 100 x the precision's epsilon), as the
  reference does; the writer stores FP64 big-endian 3x3 with those values
  (computed on the GPU, wc_gauge_observables).

LIME (lime_io/load.py, lime_io/save.py, lime_io/scidac.py): a sequence of
records, each a 144-byte big-endian header (magic 0x456789AB, version 1,
flags, payload size, 128-byte tag) and its payload padded to 8 bytes.
* ILDG gauge files: an ``ildg-format`` XML record (lx, ly, lz, lt,
  precision 64 or 32, field su3gauge) and an ``ildg-binary-data`` record;
* gpt's own files (``g.format.lime()``): a ``gpt-format`` record (grid
  description, number of fields, object type) and ``gpt-binary-data``;
* data: per site (x fastest) the 4 links, each row-major 3x3 complex,
  big-endian; as the reference does, the byte order is reversed on 8-byte
  words for either precision (cgpt munge_byte_order(data, data, 8));
* SciDAC checksum: per site the CRC-32 of the site's stored bytes, rotated
  left by (global site index mod 29) for suma and (mod 31) for sumb, XORed
  over all sites.

openQCD (lib/cgpt/lib/io/openQCD.h, which hands the data to Grid's
OpenQcdIO): a 24-byte little-endian header (N0 = T, N1 = X, N2 = Y, N3 = Z as
int32, the average plaquette as a double normalised to 3), then for every odd
point x = (x0, x1, x2, x3), x3 running fastest, the 8 links U(x, 0),
U(x - 0, 0), U(x, 1), U(x - 1, 1), ... U(x - 3, 3) as row-major 3x3 complex
doubles; openQCD's direction 0 is time (gpt's mu = 3) and 1, 2, 3 are x, y,
z.  The header plaquette / 3 is checked against the field's.

File parsing is host work by nature; the links go to HBM as FP64 link fields.
"""
import ctypes
import datetime
import getpass
import os
import socket
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np
import torch

from . import _lib
from . import core
from .qcd import gauge

_FLOATS = {"IEEE64BIG": (">f8", 8), "IEEE64LITTLE": ("<f8", 8), "IEEE64": ("<f8", 8),
           "IEEE32BIG": (">f4", 4), "IEEE32LITTLE": ("<f4", 4), "IEEE32": ("<f4", 4)}


class format:
    class nersc:
        """g.format.nersc(label="", id="gpt", sequence_number=1) (save.py:43-47)"""

        def __init__(self, params=None, **kw):
            p = {"label": "", "id": "gpt", "sequence_number": 1}
            p.update(params or {})
            p.update(kw)
            self.params = p

    class lime:
        """g.format.lime(binary_data_tag="gpt-binary-data", tags={}, tag_order=[])
        (save.py:31-35): text records tags[t] for t in tag_order, then the
        binary record, then scidac-checksum and gpt-format"""

        def __init__(self, params=None, **kw):
            p = {"binary_data_tag": "gpt-binary-data", "tags": {}, "tag_order": []}
            p.update(params or {})
            p.update(kw)
            self.params = p


_GRID_FORMAT = (
    "<FieldMetaData><nd>0</nd><dimension/><boundary/><data_start>0</data_start><hdr_version/><storage_format/>"
    "<link_trace>0</link_trace><plaquette>0</plaquette><checksum>0</checksum>"
    "<scidac_checksuma>0</scidac_checksuma><scidac_checksumb>0</scidac_checksumb>"
    "<sequence_number>0</sequence_number><data_type/><ensemble_id/><ensemble_label/><ildg_lfn/><creator/>"
    "<creator_hardware/><creation_date/><archive_date/><floating_point/></FieldMetaData>"
)
# g.format.grid_scidac (save.py:48-60): the records Grid's SciDAC reader expects
format.grid_scidac = format.lime(
    binary_data_tag="ildg-binary-data",
    tag_order=["grid-format", "scidac-record-xml", "scidac-private-record-xml"],
    tags={
        "scidac-private-record-xml": "<scidacRecord><version>1</version><date/><recordtype>0</recordtype>"
        "<datatype>0</datatype><precision>0</precision><colors>0</colors><spins>0</spins>"
        "<typesize>0</typesize><datacount>0</datacount></scidacRecord>",
        "scidac-record-xml": "<emptyUserRecord><dummy>0</dummy></emptyUserRecord>",
        "grid-format": _GRID_FORMAT,
    },
)


def read_header(path):
    """(metadata dict, byte offset of the data), or None if not a NERSC file"""
    if not os.path.isfile(path):
        return None
    meta = {}
    with open(path, "rb") as f:
        try:
            line = f.readline().decode("utf-8").strip()
            if line != "BEGIN_HEADER":
                return None
            while True:
                raw = f.readline()
                if not raw:
                    return None
                line = raw.decode("utf-8").strip()
                if line == "END_HEADER":
                    break
                if "=" in line:
                    k, v = line.split("=", 1)
                    meta[k.strip()] = v.strip()
        except UnicodeDecodeError:
            return None
        return meta, f.tell()


def checksum(native):
    """NERSC checksum of a host-order float array: sum of its 32-bit words mod 2^32"""
    words = np.ascontiguousarray(native).view(np.uint32)
    return int(words.sum(dtype=np.uint64) & 0xFFFFFFFF)


def reconstruct_third_row(two_rows):
    """(..., 2, 3) complex -> (..., 3, 3) with row 2 = conj(row0 x row1)"""
    r0, r1 = two_rows[..., 0, :], two_rows[..., 1, :]
    r2 = np.conj(np.cross(r0, r1))
    return np.stack([r0, r1, r2], axis=-2)


def gauge_observables(U):
    """(plaquette, link trace) of 4 FP64 link fields, on the GPU"""
    fd = (ctypes.c_int * 4)(*U[0].grid.fdimensions)
    st = gauge.stacked(U)
    p, l = ctypes.c_double(), ctypes.c_double()
    _lib.check(_lib.load().wc_gauge_observables(fd, st.data_ptr(), ctypes.byref(p), ctypes.byref(l),
                                                core.stream_ptr()))
    return p.value, l.value


def _digits_threshold(text, eps):
    frac = text.split(".")[1] if "." in text else ""
    digits = len(frac.lower().split("e")[0])
    return max(10.0 ** (-digits + 2), 1e2 * eps)


def load_nersc(path):
    hdr = read_header(path)
    if hdr is None:
        raise NotImplementedError(f"{path} is not a NERSC file")
    meta, offset = hdr
    fd = [int(meta[f"DIMENSION_{i + 1}"]) for i in range(4)]
    fp = meta["FLOATING_POINT"]
    if fp not in _FLOATS:
        raise NotImplementedError(f"unknown FLOATING_POINT {fp}")
    dt, nbytes = _FLOATS[fp]
    kind = meta["DATATYPE"]
    rows = {"4D_SU3_GAUGE_3x3": 3, "4D_SU3_GAUGE": 2}.get(kind)
    if rows is None:
        raise NotImplementedError(f"unknown DATATYPE {kind}")
    V = int(np.prod(fd))
    floats = V * 4 * rows * 3 * 2
    data = np.fromfile(path, dtype=dt, count=floats, offset=offset)
    if data.size != floats or os.path.getsize(path) - offset != floats * nbytes:
        raise ValueError(f"{path}: {os.path.getsize(path) - offset} data bytes, expected {floats * nbytes}")
    native = data.astype(np.float64 if nbytes == 8 else np.float32)
    cs = checksum(native)
    cs_exp = int(meta["CHECKSUM"], 16)
    if cs != cs_exp:
        raise ValueError(f"{path}: checksum {cs:x}, header says {cs_exp:x}")
    m = native.astype(np.float64).view(np.complex128).reshape(V, 4, rows, 3)
    if rows == 2:
        m = reconstruct_third_row(m)
    grid = core.grid(fd, core.double)
    U = [gauge.link_field(grid, torch.from_numpy(np.ascontiguousarray(m[:, mu])).to(core.device()))
         for mu in range(4)]
    eps = np.finfo(np.float64 if nbytes == 8 else np.float32).eps
    P, L = gauge_observables(U)
    if "PLAQUETTE" in meta and abs(P - float(meta["PLAQUETTE"])) >= _digits_threshold(meta["PLAQUETTE"], eps):
        raise ValueError(f"{path}: plaquette {P!r}, header says {meta['PLAQUETTE']}")
    if "LINK_TRACE" in meta and abs(L - float(meta["LINK_TRACE"])) >= _digits_threshold(meta["LINK_TRACE"], eps):
        raise ValueError(f"{path}: link trace {L!r}, header says {meta['LINK_TRACE']}")
    for u in U:
        u.metadata = meta
    return U


def save_nersc(path, U, params):
    """FP64 big-endian 4D_SU3_GAUGE_3x3 (nersc_io.py:254-380)"""
    if len(U) != 4:
        raise ValueError("a NERSC configuration holds 4 link fields")
    fd = list(U[0].grid.fdimensions)
    m = torch.stack([u.data for u in U], dim=1).cpu().numpy()  # [V][4][3][3]
    native = np.ascontiguousarray(m).view(np.float64).reshape(-1)
    cs = checksum(native)
    P, L = gauge_observables(U)
    now = datetime.datetime.utcnow().strftime("%c %Z")
    header = (
        "BEGIN_HEADER\n"
        "HDR_VERSION = 1.0\n"
        "DATATYPE = 4D_SU3_GAUGE_3x3\n"
        "STORAGE_FORMAT =\n"
        + "".join(f"DIMENSION_{i + 1} = {fd[i]}\n" for i in range(4))
        + f"LINK_TRACE = {L:.15g}\n"
        f"PLAQUETTE  = {P:.15g}\n"
        + "".join(f"BOUNDARY_{i + 1} = PERIODIC\n" for i in range(4))
        + f"CHECKSUM =   {cs:x}\n"
        "SCIDAC_CHECKSUMA =          0\n"
        "SCIDAC_CHECKSUMB =          0\n"
        f"ENSEMBLE_ID = {params['id']}\n"
        f"ENSEMBLE_LABEL = {params['label']}\n"
        f"SEQUENCE_NUMBER = {params['sequence_number']}\n"
        f"CREATOR = {getpass.getuser()}\n"
        f"CREATOR_HARDWARE = {socket.gethostname()}\n"
        f"CREATION_DATE = {now}\n"
        f"ARCHIVE_DATE = {now}\n"
        "FLOATING_POINT = IEEE64BIG\n"
        "END_HEADER\n"
    )
    with open(path, "wb") as f:
        f.write(header.encode("utf-8"))
        f.write(native.astype(">f8").tobytes())


# ------------------------------------------------------------------ LIME
LIME_MAGIC = 0x456789AB
_LIME_HEADER = struct.Struct(">LHHQ128s")  # magic, version, flags, payload bytes, tag
_SU3 = "ot_matrix_su_n_fundamental_group(3)"


class lime_reader:
    """record index of a LIME file: tag -> (payload offset, payload bytes)
    (lime_io/load.py:27-79)"""

    @staticmethod
    def is_file(path):
        if not os.path.isfile(path):
            return False
        with open(path, "rb") as f:
            m = f.read(4)
        return len(m) == 4 and struct.unpack(">L", m)[0] == LIME_MAGIC

    def __init__(self, path):
        self.path = path
        self.index = {}
        end = os.path.getsize(path)
        with open(path, "rb") as f:
            pos = 0
            while pos < end:
                head = f.read(_LIME_HEADER.size)
                if len(head) != _LIME_HEADER.size:
                    raise ValueError(f"{path}: truncated LIME record header at byte {pos}")
                magic, version, _, size, tag = _LIME_HEADER.unpack(head)
                if magic != LIME_MAGIC or version != 1:
                    raise ValueError(f"{path}: bad LIME record at byte {pos}")
                offset = pos + _LIME_HEADER.size
                self.index[tag.decode("utf-8").strip("\0")] = (offset, size)
                pos = offset + size + (-size) % 8
                f.seek(pos)
            if pos > end:
                raise ValueError(f"{path}: LIME payload runs past the end of the file")

    def read(self, tag):
        offset, size = self.index[tag]
        with open(self.path, "rb") as f:
            f.seek(offset)
            return f.read(size)

    def text(self, tag):
        return self.read(tag).decode("utf-8").strip("\0")

    def xml_dict(self, tag):
        out = {}
        for child in ET.fromstring(self.text(tag)):
            t = child.tag
            out[t[t.index("}") + 1:] if t[0] == "{" else t] = child.text
        return out


class lime_writer:
    """appends LIME records (lime_io/save.py:27-118); flags 0xFFFF as gpt writes them"""

    def __init__(self, path):
        self.f = open(path, "wb")

    def record(self, tag, payload):
        t = (tag + "\0").encode("utf-8")
        if len(t) > 128:
            raise ValueError(f"LIME tag {tag!r} is longer than 127 bytes")
        self.f.write(_LIME_HEADER.pack(LIME_MAGIC, 1, 0xFFFF, len(payload), t))
        self.f.write(payload)
        self.f.write(bytes((-len(payload)) % 8))

    def text(self, tag, text):
        self.record(tag, text.encode("utf-8"))

    def xml(self, tag, root, fields):
        r = ET.Element(root)
        for k, v in fields.items():
            ET.SubElement(r, k).text = str(v)
        self.text(tag, ET.tostring(r, encoding="utf-8", xml_declaration=True).decode("utf-8"))

    def close(self):
        self.f.close()


def _rotl32(c, k):
    return ((c << k) | (c >> (32 - k))) & np.uint64(0xFFFFFFFF)


def scidac_checksums(data, site_bytes):
    """(suma, sumb) of stored bytes laid out site by site in lexicographic
    order (lime_io/scidac.py:21-31): CRC-32 of each site's bytes rotated left
    by (site mod 29) resp. (site mod 31), XORed over the sites"""
    mv = memoryview(data)
    n = len(mv) // site_bytes
    crc = np.fromiter((zlib.crc32(mv[i * site_bytes:(i + 1) * site_bytes]) for i in range(n)), np.uint64, n)
    site = np.arange(n, dtype=np.uint64)
    a = np.bitwise_xor.reduce(_rotl32(crc, site % np.uint64(29))) if n else 0
    b = np.bitwise_xor.reduce(_rotl32(crc, site % np.uint64(31))) if n else 0
    return int(a), int(b)


def _links_from_lime(raw, V, nbytes):
    """site-major stored links -> (V, 4, 3, 3) complex128: 8-byte words byte-reversed (munge.cc:54-81)"""
    w = np.frombuffer(raw, np.uint8).reshape(-1, 8)[:, ::-1].copy()
    vals = w.view("<f8").reshape(-1) if nbytes == 8 else w.view("<f4").reshape(-1).astype(np.float64)
    return vals.view(np.complex128).reshape(V, 4, 3, 3)


def load_lime(path):
    """ILDG gauge files and gpt's own LIME files of 4 SU(3) link fields (lime_io/load.py:127-234)"""
    if not lime_reader.is_file(path):
        raise NotImplementedError(f"{path} is not a LIME file")
    r = lime_reader(path)
    if "ildg-format" in r.index:
        d = r.xml_dict("ildg-format")
        if d.get("field") != "su3gauge":
            raise NotImplementedError(f"{path}: ILDG field {d.get('field')!r}, expected su3gauge")
        fd = [int(d["lx"]), int(d["ly"]), int(d["lz"]), int(d["lt"])]
        prec = d["precision"]
    elif "gpt-format" in r.index:
        d = r.xml_dict("gpt-format")
        p = d["grid"].split(";")
        if len(p) != 3 or p[2] != "full" or d["otype"] != _SU3 + ";none" or int(d["n"]) != 4:
            raise NotImplementedError(f"{path}: gpt_amd reads gauge configurations (4 full SU(3) fields), "
                                      f"not {d['n']} x {d['otype']} on {d['grid']}")
        fd = [int(x) for x in p[0].strip("[]").split(",")]
        prec = {"double": "64", "single": "32"}.get(p[1], p[1])
    else:
        raise NotImplementedError(f"{path}: unknown LIME content {sorted(r.index)}")
    if prec not in ("64", "32") or len(fd) != 4:
        raise NotImplementedError(f"{path}: precision {prec} on a {len(fd)}-d grid")
    nbytes = 8 if prec == "64" else 4
    tags = [t for t in r.index if t.endswith("-binary-data")]
    if len(tags) != 1:
        raise ValueError(f"{path}: expected one *-binary-data record, found {tags}")
    V = int(np.prod(fd))
    site_bytes = 4 * 9 * 2 * nbytes
    if r.index[tags[0]][1] != V * site_bytes:
        raise ValueError(f"{path}: {r.index[tags[0]][1]} data bytes, expected {V * site_bytes}")
    raw = r.read(tags[0])
    if "scidac-checksum" in r.index:
        cs = r.xml_dict("scidac-checksum")
        a, b = scidac_checksums(raw, site_bytes)
        if a != int(cs["suma"], 16) or b != int(cs["sumb"], 16):
            raise ValueError(f"{path}: SciDAC checksums {a:x} {b:x}, file says {cs['suma'].strip()} "
                             f"{cs['sumb'].strip()}")
    m = _links_from_lime(raw, V, nbytes)
    grid = core.grid(fd, core.double)
    return [gauge.link_field(grid, torch.from_numpy(np.ascontiguousarray(m[:, mu])).to(core.device()))
            for mu in range(4)]


def save_lime(path, U, params):
    """FP64 gauge links as gpt writes them with g.format.lime(...) (lime_io/save.py:121-223)"""
    if len(U) != 4:
        raise ValueError("a gauge configuration holds 4 link fields")
    fd = list(U[0].grid.fdimensions)
    m = torch.stack([u.data for u in U], dim=1).cpu().numpy()  # [V][4][3][3]
    raw = np.ascontiguousarray(m).view(np.float64).astype(">f8").tobytes()
    a, b = scidac_checksums(raw, 4 * 9 * 2 * 8)
    w = lime_writer(path)
    try:
        for t in params["tag_order"]:
            w.text(t, params["tags"][t])
        w.record(params["binary_data_tag"], raw)
        w.xml("scidac-checksum", "scidacChecksum", {"version": "1.0", "suma": f"{a:8x}", "sumb": f"{b:8x}"})
        w.xml("gpt-format", "gptFormat", {"grid": str(fd).replace(" ", "") + ";double;full", "n": 4,
                                         "otype": _SU3 + ";none"})
    finally:
        w.close()


# --------------------------------------------------------------- openQCD
_OQ_HEADER = struct.Struct("<iiiid")  # N0 (T), N1 (X), N2 (Y), N3 (Z), plaquette


def read_openqcd_header(path):
    """(dims [X, Y, Z, T], plaquette) or None; as loose a test as the
    reference's (openQCD.h:19-58): extents in 1..10000, |plaquette| <= 100"""
    if not os.path.isfile(path) or os.path.getsize(path) < _OQ_HEADER.size:
        return None
    with open(path, "rb") as f:
        n0, n1, n2, n3, plaq = _OQ_HEADER.unpack(f.read(_OQ_HEADER.size))
    if any(not (1 <= n <= 10000) for n in (n0, n1, n2, n3)) or not (-100.0 <= plaq <= 100.0):
        return None
    return [n1, n2, n3, n0], plaq


def openqcd_odd_points(fd):
    """the odd points (x0 + x1 + x2 + x3 odd) in openQCD's file order, x3
    fastest, as gpt coordinates (x, y, z, t) arrays"""
    X, Y, Z, T = fd
    t, x, y, z = (a.reshape(-1) for a in np.indices((T, X, Y, Z)))
    odd = (t + x + y + z) % 2 == 1
    return x[odd], y[odd], z[odd], t[odd]


def load_openqcd(path):
    hdr = read_openqcd_header(path)
    if hdr is None:
        raise NotImplementedError(f"{path} is not an openQCD file")
    fd, plaq = hdr
    V = int(np.prod(fd))
    if V % 2 or os.path.getsize(path) != _OQ_HEADER.size + V * 4 * 144:
        raise NotImplementedError(f"{path}: {os.path.getsize(path)} bytes, not an openQCD configuration of {fd}")
    data = np.fromfile(path, dtype="<c16", offset=_OQ_HEADER.size).reshape(V // 2, 8, 3, 3)
    X, Y, Z, T = fd
    c = openqcd_odd_points(fd)
    lex = lambda x, y, z, t: x + X * (y + Y * (z + Z * t))  # noqa: E731
    m = np.empty((4, V, 3, 3), np.complex128)
    filled = np.zeros((4, V), bool)
    for d, mu in enumerate((3, 0, 1, 2)):  # openQCD direction d is gpt's mu
        step = [0, 0, 0, 0]
        step[mu] = 1
        back = [(c[i] - step[i]) % fd[i] for i in range(4)]
        m[mu, lex(*c)] = data[:, 2 * d]
        m[mu, lex(*back)] = data[:, 2 * d + 1]
        filled[mu, lex(*c)] = True
        filled[mu, lex(*back)] = True
    if not filled.all():
        raise ValueError(f"{path}: the odd-point links do not cover the lattice {fd}")
    grid = core.grid(fd, core.double)
    U = [gauge.link_field(grid, torch.from_numpy(np.ascontiguousarray(m[mu])).to(core.device())) for mu in range(4)]
    P, _ = gauge_observables(U)
    if abs(P - plaq / 3.0) > 1e-10 * max(1.0, abs(P)):
        raise ValueError(f"{path}: plaquette {P!r}, header says {plaq / 3.0!r} (openQCD's {plaq!r} / 3)")
    meta = {f"DIMENSION_{i + 1}": str(fd[i]) for i in range(4)}
    meta["PLAQUETTE"] = f"{plaq:.15f}"
    for u in U:
        u.metadata = meta
    return U


# ---------------------------------------------------------------- g.load
def load(path, **params):
    """g.load (load.py:50-73): NERSC, then LIME (ILDG / gpt gauge records), then openQCD"""
    for fmt in (load_nersc, load_lime, load_openqcd):
        try:
            return fmt(path)
        except NotImplementedError:
            pass
    raise NotImplementedError(f"{path}: not a NERSC, LIME or openQCD gauge configuration")


def save(path, U, fmt=None):
    """g.save(path, U, fmt) with g.format.nersc(...) or g.format.lime(...) / g.format.grid_scidac"""
    fmt = fmt or format.nersc()
    if isinstance(fmt, format.nersc):
        save_nersc(path, U, fmt.params)
    elif isinstance(fmt, format.lime):
        save_lime(path, U, fmt.params)
    else:
        raise NotImplementedError("gpt_amd writes the NERSC and LIME formats")


__all__ = ["format", "load", "save", "read_header", "checksum", "reconstruct_third_row", "gauge_observables",
           "lime_reader", "lime_writer", "scidac_checksums", "read_openqcd_header", "openqcd_odd_points"]
