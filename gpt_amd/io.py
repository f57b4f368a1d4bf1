"""Gauge-configuration IO: the NERSC format.

Mirrors gpt's ``g.load(fn)`` / ``g.save(fn, U, g.format.nersc(...))``
(lib/gpt/core/io/load.py:50-70, save.py:43-47, nersc_io.py):

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

File parsing is host work by nature; the links go to HBM as FP64 link fields.
"""
import ctypes
import datetime
import getpass
import os
import socket

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


def load(path, **params):
    """g.load: the formats gpt_amd reads (NERSC)"""
    return load_nersc(path)


def save(path, U, fmt=None):
    """g.save(path, U, g.format.nersc(...))"""
    fmt = fmt or format.nersc()
    if not isinstance(fmt, format.nersc):
        raise NotImplementedError("gpt_amd writes the NERSC format only")
    save_nersc(path, U, fmt.params)


__all__ = ["format", "load", "save", "read_header", "checksum", "reconstruct_third_row", "gauge_observables"]
