"""Reader/writer of the reference's .field checkpoint format.

Unformatted sequential Fortran records with 4-byte length markers
(read_write.f90:16-40): three header records
  (time, nnx, nny, nnz, xl, yl, zl), (dt, z0, utau),
  (divgls, fcor, ugtop, ugbot, vgtop, vgbot)
then one record per level iz = 1..nnz holding u, v, w, e, p(1:nnx,1:nny).
Reals are REAL(4) in the shipped reference and REAL(8) in a promoted build
and in the files les_gpu writes; ``read_field`` detects the kind.
"""
from __future__ import annotations

import struct

import numpy as np


def _records(buf):
    pos = 0
    while pos < len(buf):
        (n,) = struct.unpack_from("<i", buf, pos)
        yield buf[pos + 4: pos + 4 + n]
        pos += 8 + n


def read_field(path):
    buf = open(path, "rb").read()
    recs = list(_records(buf))
    rk = 4 if len(recs[0]) == 4 * 4 + 3 * 4 else 8
    f = "<f4" if rk == 4 else "<f8"
    h0 = recs[0]
    t = np.frombuffer(h0[:rk], f)[0]
    nnx, nny, nnz = struct.unpack_from("<iii", h0, rk)
    xl, yl, zl = np.frombuffer(h0[rk + 12:], f)
    dt, z0, utau = np.frombuffer(recs[1], f)
    extra = np.frombuffer(recs[2], f)
    lv = [np.frombuffer(r, f).astype(np.float64).reshape(5, nny, nnx) for r in recs[3: 3 + nnz]]
    arr = np.stack(lv, axis=1)  # (5, nnz, nny, nnx)
    hdr = dict(time=float(t), nnx=nnx, nny=nny, nnz=nnz, xl=float(xl), yl=float(yl), zl=float(zl),
               dt=float(dt), z0=float(z0), utau=float(utau), extra=extra.astype(float), kind=rk)
    return hdr, dict(zip("uvwep", arr))


def write_field(path, hdr, fields, kind=8):
    f = "<f4" if kind == 4 else "<f8"

    def rec(b):
        return struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))
    out = rec(np.array([hdr["time"]], f).tobytes() + struct.pack("<iii", hdr["nnx"], hdr["nny"], hdr["nnz"])
              + np.array([hdr["xl"], hdr["yl"], hdr["zl"]], f).tobytes())
    out += rec(np.array([hdr["dt"], hdr["z0"], hdr.get("utau", 0.0)], f).tobytes())
    out += rec(np.array(hdr.get("extra", [0.0, 1e-4, 15.0, 15.0, 0.0, 0.0]), f).tobytes())
    for iz in range(hdr["nnz"]):
        out += rec(b"".join(np.ascontiguousarray(fields[k][iz], f).tobytes() for k in "uvwep"))
    open(path, "wb").write(out)
