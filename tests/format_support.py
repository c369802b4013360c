"""TEST INFRASTRUCTURE: the reference's on-disk formats restated in numpy.

* Fortran sequential unformatted records (4-byte little-endian length,
  data, length) -- SRC/save_raw.F90:44-65, SRC/save_vel.F90;
* the restart file: time step, then a01..a19 over (0:l+1, 0:m+1, 0:n+1);
* the legacy VTK planes with big-endian FP32 vectors
  (SRC/vtk_xy.binary.F90:59-97, vtk_xz, vtk_yz, SRC/vtk_om.xy.binary.F90),
  with the reference's mixed-kind arithmetic (sums in the working kind
  rounded to FP32, FP32 division), computed from an oracle state.

Used to pin those formats against the bytes the reference's own writers
produce, and to digest files for the fixtures.
"""
from __future__ import annotations

import hashlib
import struct

import numpy as np

KIND = {"sp": np.float32, "dp": np.float64}


def sha(b) -> str:
    return hashlib.sha256(bytes(b)).hexdigest()


def parse_records(b: bytes):
    out, pos = [], 0
    while pos < len(b):
        (n,) = struct.unpack_from("<I", b, pos)
        data = b[pos + 4: pos + 4 + n]
        (n2,) = struct.unpack_from("<I", b, pos + 4 + n)
        assert n2 == n, "damaged record"
        out.append(data)
        pos += 8 + n
    return out


def records(chunks) -> bytes:
    return b"".join(struct.pack("<I", len(c)) + c + struct.pack("<I", len(c)) for c in chunks)


def restart_interior_digest(b: bytes, l, m, n, prec):
    recs = parse_records(b)
    assert len(recs) == 20, len(recs)
    (itime,) = struct.unpack("<i", recs[0])
    dt = KIND[prec]
    parts = []
    for r in recs[1:]:
        a = np.frombuffer(r, dtype=dt).reshape(n + 2, m + 2, l + 2)
        parts.append(np.ascontiguousarray(a[1:-1, 1:-1, 1:-1]).tobytes())
    return {"itime": itime, "record_bytes": [len(r) for r in recs], "sha256_interior": sha(b"".join(parts))}


def restart_bytes(itime, A, l, m, n):
    """A: (19, nz, ny, nx) oracle lattice (nx, ny may include padding)."""
    return records([struct.pack("<i", itime)] +
                   [np.ascontiguousarray(A[d, :n + 2, :m + 2, :l + 2]).tobytes() for d in range(19)])


def vtk_header(b: bytes) -> str:
    end = b.index(b"float\n", b.index(b"VECTORS")) + len(b"float\n")
    return b[:end].decode()


def file_digests(files, l, m, n, prec):
    out = {}
    for name, b in files.items():
        if name.startswith("save."):
            out["save_raw"] = restart_interior_digest(b, l, m, n, prec)
        elif name.startswith("save_vel."):
            out["save_vel"] = {"sha256": sha(b), "bytes": len(b)}
        elif name.startswith("tec_"):
            key = {"tec_xy": "vtk_xy", "tec_xz": "vtk_xz", "tec_yz": "vtk_yz", "tec_om": "vtk_om_xy"}[name[:6]]
            out[key] = {"sha256": sha(b), "bytes": len(b), "header": vtk_header(b)}
    return out


# ---- values the writers compute --------------------------------------------

def _lsum(vals):
    acc = vals[0]
    for v in vals[1:]:
        acc = acc + v
    return acc


def _moments(x):
    """u, w (y), v (z) momentum sums and the plain population sum, in the
    arrays' kind, the reference's left-to-right order."""
    u = _lsum([x[0], x[1], x[2], x[3], x[4], -x[9], -x[10], -x[11], -x[12], -x[13]])
    w = _lsum([x[2], x[6], x[7], x[8], x[11], -x[0], -x[9], -x[15], -x[16], -x[17]])
    v = _lsum([x[3], x[5], x[6], x[12], x[17], -x[1], -x[8], -x[10], -x[14], -x[15]])
    return u, w, v, _lsum(list(x))


def vtk_plane_values(A, l, m, n, plane, idx, prec):
    """(npts, 3) FP32 of SRC/vtk_{xy,xz,yz}.binary.F90 at index idx."""
    if plane == 0:
        x = [A[d, idx, 1:m + 1, 1:l + 1] for d in range(19)]
    elif plane == 1:
        x = [A[d, 1:n + 1, idx, 1:l + 1] for d in range(19)]
    else:
        x = [A[d, 1:n + 1, 1:m + 1, idx] for d in range(19)]
    kind = KIND[prec]
    x = [np.asarray(a, dtype=kind) for a in x]
    u, w, v, s = _moments(x)
    den = (s + kind(np.float32(1.0))).astype(np.float32)
    u, w, v = u.astype(np.float32), w.astype(np.float32), v.astype(np.float32)
    return np.stack([(u / den).ravel(), (w / den).ravel(), (v / den).ravel()], axis=1)


def _velocity_sp(A, k, l, m, prec):
    kind = KIND[prec]
    x = [A[d, k, 1:m + 1, 1:l + 1].astype(np.float32) for d in range(19)]
    rho = _lsum(x) + np.float32(1.0)
    rhoinv = (kind(1) / rho.astype(kind)).astype(np.float32)
    u = _lsum([x[0], x[1], x[2], x[3], x[4], -x[9], -x[10], -x[11], -x[12], -x[13]]) * rhoinv
    v = _lsum([x[2], x[6], x[7], x[8], x[11], -x[0], -x[9], -x[15], -x[16], -x[17]]) * rhoinv
    w = _lsum([x[3], x[5], x[6], x[12], x[17], -x[1], -x[8], -x[10], -x[14], -x[15]]) * rhoinv
    return u, v, w  # arrays (m, l) over j = 1..m, i = 1..l


def vorticity_xy_values(A, l, m, n, k0, prec):
    """(npts, 3) FP32 of SRC/vtk_om.xy.binary.F90 at k0 (i = 2..l-1 fastest)."""
    U = [_velocity_sp(A, k0 + dk, l, m, prec) for dk in (-1, 0, 1)]
    h = np.float32(0.5)
    (um, vm, wm), (u0, v0, w0), (up, vp, wp) = U
    # interior of the plane: j = 2..m-1 -> rows 1..m-2, i = 2..l-1 -> cols 1..l-2
    c = (slice(1, m - 1), slice(1, l - 1))
    jp, jm = (slice(2, m), slice(1, l - 1)), (slice(0, m - 2), slice(1, l - 1))
    ip, im = (slice(1, m - 1), slice(2, l)), (slice(1, m - 1), slice(0, l - 2))
    dudy = h * (u0[jp] - u0[jm])
    dudz = h * (up[c] - um[c])
    dvdx = h * (v0[ip] - v0[im])
    dvdz = h * (vp[c] - vm[c])
    dwdx = h * (w0[ip] - w0[im])
    dwdy = h * (w0[jp] - w0[jm])
    return np.stack([(dvdz - dwdy).ravel(), (dwdx - dudz).ravel(), (dudy - dvdx).ravel()], axis=1)


def vtk_bytes(kind, l, m, n, idx, values):
    dims = {0: (l, m, 1), 1: (l, 1, n), 2: (1, m, n), 3: (l - 2, m - 2, 1)}[kind]
    origin = {0: (1, 1, idx), 1: (1, idx, 1), 2: (idx, 1, 1), 3: (2, 2, idx)}[kind]
    npts = dims[0] * dims[1] * dims[2]
    vec = " VECTORS vorticity float" if kind == 3 else " VECTORS velocity float"
    head = ("# vtk DataFile Version 2.0\nCampo\nBINARY\nDATASET STRUCTURED_POINTS\n"
            f"DIMENSIONS {dims[0]:10d} {dims[1]:10d} {dims[2]:10d}\n"
            f"ORIGIN {origin[0]:10d} {origin[1]:10d} {origin[2]:10d}\n"
            f"SPACING {1:10d} {1:10d} {1:10d}\n"
            f"POINT_DATA{npts:10d}\n{vec}\n")
    return head.encode() + np.asarray(values, dtype=">f4").tobytes()


def save_vel_bytes(itime, A, l, m, n, prec):
    kind = KIND[prec]
    x = [A[d, 1:n + 1, 1:m + 1, 1:l + 1].astype(kind) for d in range(19)]
    rho = _lsum(x) + kind(1)
    rhoinv = kind(1) / rho
    u = _lsum([x[0], x[1], x[2], x[3], x[4], -x[9], -x[10], -x[11], -x[12], -x[13]]) * rhoinv
    v = _lsum([x[2], x[6], x[7], x[8], x[11], -x[0], -x[9], -x[15], -x[16], -x[17]]) * rhoinv
    w = _lsum([x[3], x[5], x[6], x[12], x[17], -x[1], -x[8], -x[10], -x[14], -x[15]]) * rhoinv
    trip = np.stack([u.ravel(), v.ravel(), w.ravel()], axis=1).astype(kind)
    return records([struct.pack("<i", itime)] + [row.tobytes() for row in trip])
