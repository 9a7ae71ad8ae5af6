"""Field files and run configuration in the reference's formats.

Field files (``u1_/u2_/u3_/p_NNNNxNNNNxNNNN_tTTTTT.dat``) are Fortran sequential-unformatted files:

* record 1: ``t, Re, alp, bet, mpgx`` (f64) then ``nband, iter, dummint(88)`` (i32)
  -- written at FOU3D.f90:2069, read at start.f90:1576;
* record 2: ``N(4, 0:nband+1)`` (i32, column-major) -- FOU3D.f90:2070, start.f90:1598;
* record 3: ``y, ddthetavi, dthdy`` (f64) -- FOU3D.f90:2071;
* one record per plane: ``j, id, nx, nz`` (i32), ``y_j`` (f64), ``buff(nx, nz)`` (f64) with
  ``nx = Nx+2`` interleaved (Re, Im) pairs for kx = 0..Nx/2 and ``nz = Nz`` z-modes in the reference's
  column order (kz = 0..Nz/2-1, then -Nz/2..-1) -- FOU3D.f90:2077, u_to_buff FOU3D.f90:2266-2293.

The plane buffer's (k, i) order is the reference's global column order (start.f90:963-981), so a file
plane is one y-row of a column field.  Column fields here are ``complex128[ncol, nrows]`` with
``ncol = (Nx/2+1)*Nz`` and rows ordered as the reference's ``x(j, column)`` (j fastest):
u1/u3 rows j=0..nyu+1, u2 rows j=0..nyv+1, p rows j=1..nyu.
"""
from __future__ import annotations

import dataclasses
import re
import struct

import numpy as np


@dataclasses.dataclass
class FieldFile:
    t: float
    Re: float
    alp: float
    bet: float
    mpgx: float
    nband: int
    iter: int
    N: np.ndarray          # int32 [4, nband+2] (Fortran N(4,0:nband+1))
    y: np.ndarray
    ddthetavi: float
    dthdy: np.ndarray
    j: np.ndarray          # plane indices as stored
    planes: np.ndarray     # complex128 [nplanes, Nz, Kx]
    plane_y: np.ndarray


def _records(buf: bytes):
    off = 0
    n = len(buf)
    while off < n:
        (ln,) = struct.unpack_from("<i", buf, off)
        start = off + 4
        end = start + ln
        (ln2,) = struct.unpack_from("<i", buf, end)
        if ln2 != ln:
            raise ValueError("corrupt Fortran record at byte %d" % off)
        yield buf[start:end]
        off = end + 4


def read_field(path: str) -> FieldFile:
    """Read one reference field file (start.f90:1545-1997 reads the same records)."""
    with open(path, "rb") as f:
        buf = f.read()
    recs = list(_records(buf))
    if len(recs) < 4:
        raise ValueError("%s: too few records" % path)
    t, Re, alp, bet, mpgx = struct.unpack_from("<5d", recs[0], 0)
    nband, it = struct.unpack_from("<2i", recs[0], 40)
    Narr = np.frombuffer(recs[1], dtype="<i4")
    N = Narr.reshape(-1, 4).T.copy()  # N(4, 0:nband+1), column-major
    r3 = np.frombuffer(recs[2], dtype="<f8")
    ny_pts = (len(r3) - 1) // 2
    y = r3[:ny_pts].copy()
    ddthetavi = float(r3[ny_pts])
    dthdy = r3[ny_pts + 1:].copy()
    js, ys, planes = [], [], []
    for r in recs[3:]:
        j, _id, nx, nz = struct.unpack_from("<4i", r, 0)
        (yj,) = struct.unpack_from("<d", r, 16)
        b = np.frombuffer(r, dtype="<f8", offset=24, count=nx * nz).reshape(nz, nx)
        js.append(j)
        ys.append(yj)
        planes.append(b[:, 0::2] + 1j * b[:, 1::2])
    return FieldFile(t, Re, alp, bet, mpgx, nband, it, N, y, ddthetavi, dthdy,
                     np.array(js), np.array(planes), np.array(ys))


def write_field(path: str, ff: FieldFile, dummint: np.ndarray | None = None, grid_id: int = 1) -> None:
    """Write a field file with the record layout of record_out (FOU3D.f90:2063-2161); grid_id is the
    second integer of each plane record (record_out writes 1, 2, 3, 4 for u1, u2, u3, p)."""
    def rec(payload: bytes) -> bytes:
        return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))

    if dummint is None:
        dummint = np.zeros(88, dtype="<i4")
    out = [rec(struct.pack("<5d2i", ff.t, ff.Re, ff.alp, ff.bet, ff.mpgx, ff.nband, ff.iter)
               + np.asarray(dummint, "<i4").tobytes())]
    out.append(rec(np.asarray(ff.N, "<i4").T.tobytes()))
    out.append(rec(np.asarray(ff.y, "<f8").tobytes() + struct.pack("<d", ff.ddthetavi)
                   + np.asarray(ff.dthdy, "<f8").tobytes()))
    for jj, yj, pl in zip(ff.j, ff.plane_y, ff.planes):
        nz, kx = pl.shape
        b = np.empty((nz, 2 * kx), dtype="<f8")
        b[:, 0::2] = pl.real
        b[:, 1::2] = pl.imag
        out.append(rec(struct.pack("<4id", int(jj), grid_id, 2 * kx, nz, float(yj)) + b.tobytes()))
    with open(path, "wb") as f:
        f.write(b"".join(out))


# ---- history file hist_NNNNxNNNNxNNNN.dat (start.f90:273-284, FOU3D.f90:2196-2200) -------------------
# record 1: t, Re, alp, bet, mpgx (f64), nband (i32); record 2: N(4, 0:nband+1) (i32); then per
# record_out: flag_ctpress, iter (i32), t, dtv, dtc, dt, err, Qx, [QxT,] mpgx, [dgx,] Umax, Uslip (f64),
# the bracketed pair only when flag_ctpress = 0.
HIST_FIELDS = {0: ("t", "dtv", "dtc", "dt", "err", "Qx", "QxT", "mpgx", "dgx", "Umax", "Uslip"),
               1: ("t", "dtv", "dtc", "dt", "err", "Qx", "mpgx", "Umax", "Uslip")}


def _frec(payload: bytes) -> bytes:
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def write_hist(path: str, head: dict, N: np.ndarray, records: list, append: bool = False) -> None:
    """Create (or append to) a history file; records are dicts with flag_ctpress, iter and the fields."""
    out = []
    if not append:
        out.append(_frec(struct.pack("<5di", head["t"], head["Re"], head["alp"], head["bet"], head["mpgx"],
                                      head.get("nband", 3))))
        out.append(_frec(np.asarray(N, "<i4").T.tobytes()))
    for r in records:
        f = int(r["flag_ctpress"])
        vals = [float(r[k]) for k in HIST_FIELDS[f]]
        out.append(_frec(struct.pack("<2i%dd" % len(vals), f, int(r["iter"]), *vals)))
    with open(path, "ab" if append else "wb") as fh:
        fh.write(b"".join(out))


def read_hist(path: str):
    """(header dict, N int32[4, nband+2], list of record dicts)."""
    with open(path, "rb") as fh:
        recs = list(_records(fh.read()))
    t, Re, alp, bet, mpgx = struct.unpack_from("<5d", recs[0], 0)
    (nband,) = struct.unpack_from("<i", recs[0], 40)
    N = np.frombuffer(recs[1], "<i4").reshape(-1, 4).T.copy()
    rows = []
    for r in recs[2:]:
        f, it = struct.unpack_from("<2i", r, 0)
        names = HIST_FIELDS[f]
        vals = struct.unpack_from("<%dd" % len(names), r, 8)
        d = dict(zip(names, vals))
        d.update(flag_ctpress=f, iter=it)
        rows.append(d)
    return dict(t=t, Re=Re, alp=alp, bet=bet, mpgx=mpgx, nband=nband), N, rows


# ---- statistics and spectra files (stats.f90:340-362, spectra.f90:167-186) ----------------------------
# stats_NNNNxNNNNxNNNN_TTTTT.dat: record 1 as a field file (t, Re, alp, bet, mpgx, nband, iter,
# dummint(88)); record 2: the physical grid sizes in N's layout (the reference names it Ngal); records
# 3, 4: yu, dthetai, dthdyu and yv, dthetai, dthdyv; record 5: istat (i32); then one f64 record per
# statistic in STATS_FILE_ORDER over its grid's planes (u: 0..nyu+1, v: 0..nyv+1, p: 1..nyu).
# sp{U,V,W,UV,P}_...: the same first five records (N of the field files in record 2), then buffSP.
STATS_FILE_ORDER = ("Um", "U2m", "Vm", "V2m", "Wm", "W2m", "Pm", "P2m", "UVm", "UWm", "VWm", "wxm", "wx2m")
_STAT_GRID = {"Vm": "v", "V2m": "v", "VWm": "v", "wxm": "v", "wx2m": "v", "Pm": "p", "P2m": "p"}


def _stats_head(recs):
    t, Re, alp, bet, mpgx = struct.unpack_from("<5d", recs[0], 0)
    nband, it = struct.unpack_from("<2i", recs[0], 40)
    N = np.frombuffer(recs[1], "<i4").reshape(-1, 4).T.copy()
    (istat,) = struct.unpack("<i", recs[4])
    return dict(t=t, Re=Re, alp=alp, bet=bet, mpgx=mpgx, nband=nband, iter=it, N=N, istat=istat)


def read_stats(path: str) -> dict:
    """Header fields, istat and the 13 statistics (arrays over the planes written)."""
    with open(path, "rb") as fh:
        recs = list(_records(fh.read()))
    out = _stats_head(recs)
    for name, r in zip(STATS_FILE_ORDER, recs[5:]):
        out[name] = np.frombuffer(r, "<f8").copy()
    return out


def fold_spectrum(sp: np.ndarray, Nx: int, Nz: int, j0: int = 0) -> np.ndarray:
    """recvwrspec's buffSP(0:Nx/2, 1:Nz/2+1, rows) from sp[column, row] (spectra.f90:142-165): column
    (i, k) goes to (i, kk), kk = k for k <= Nz/2 and Nz+2-k above (k 1-based), +kz and -kz adding; entries
    i = 1..Nx/2, kk = 1..Nz/2 doubled.  Returned as [rows, Nz/2+1, Nx/2+1] (the file's order reversed)."""
    Kx = Nx // 2 + 1
    ncol, nr = sp.shape
    buff = np.zeros((nr, Nz // 2 + 1, Kx))
    for col in range(ncol):
        i, k = col % Kx, col // Kx + 1
        kk = k if k <= Nz // 2 else Nz + 2 - k
        buff[:, kk - 1, i] += sp[col]
    buff[:, :Nz // 2, 1:] *= 2.0
    return buff


def read_spectrum(path: str, Nx: int, Nz: int) -> dict:
    """Header fields, istat and buffSP as [rows, Nz/2+1, Nx/2+1]."""
    with open(path, "rb") as fh:
        recs = list(_records(fh.read()))
    out = _stats_head(recs)
    out["sp"] = np.frombuffer(recs[5], "<f8").reshape(-1, Nz // 2 + 1, Nx // 2 + 1).copy()
    return out


def planes_to_columns(planes: np.ndarray) -> np.ndarray:
    """[nplanes, Nz, Kx] -> column field [ncol, nplanes] (column = k*Kx + i)."""
    npl, nz, kx = planes.shape
    return np.ascontiguousarray(planes.reshape(npl, nz * kx).T)


def columns_to_planes(cols: np.ndarray, Nz: int) -> np.ndarray:
    ncol, nrows = cols.shape
    return np.ascontiguousarray(cols.T.reshape(nrows, Nz, ncol // Nz))


def load_state(prefix_dir: str, tag: str, ny: int):
    """Load u1, u2, u3, p of one snapshot ``<dir>/{u1,u2,u3,p}<tag>`` as reference column fields.

    The rows are placed by the plane index j stored in each record, so files written with the
    legacy multi-band header read the same way (start.f90:1611-1620 only uses N(:,0) and N(:,3)).
    Returns (u1, u2, u3, p, header FieldFile of u1).
    """
    nyv = ny - 2
    nyu = nyv + 1
    rows = {"u1": (0, nyu + 2), "u3": (0, nyu + 2), "u2": (0, nyv + 2), "p": (1, nyu)}
    out = {}
    head = None
    for name, (j0, nr) in rows.items():
        ff = read_field("%s/%s%s" % (prefix_dir, name, tag))
        if head is None or name == "u1":
            head = ff if name == "u1" else head
        nz, kx = ff.planes.shape[1:]
        cols = np.zeros((nz * kx, nr), dtype=np.complex128)
        for jj, pl in zip(ff.j, ff.planes):
            r = int(jj) - j0
            if 0 <= r < nr:
                cols[:, r] = pl.reshape(-1)
        out[name] = cols
    return out["u1"], out["u2"], out["u3"], out["p"], head


# ---------------------------------------------------------------------------------------------
# input.in : positional text file, one value per line at fixed columns, read with the FORMATs
# (7X,D10.1), (7X,I10), (12X,I5), (8X,A50) (start.f90:86-252, 485-488).  Line numbers (1-based):
_LINES = {"Re": (10, "r"), "alp": (13, "r"), "bet": (14, "r"), "Mx": (15, "i"), "Mz": (17, "i"),
          "ny": (19, "i"), "dyq": (20, "r"), "ppp": (21, "i"), "flag_init": (24, "i"), "fnmeb": (28, "a"),
          "dirin": (30, "a"), "dirout": (31, "a"), "CFL": (34, "r"), "nwrite": (35, "i"),
          "flag_ctpress": (36, "i5"), "mpgx": (38, "r"), "maxt": (44, "r")}


def _field(line: str, kind: str):
    if kind == "a":
        return line[8:58].strip()
    txt = line[12:17] if kind == "i5" else line[7:17]
    txt = txt.strip()
    if kind in ("i", "i5"):
        return int(txt) if txt else 0
    return float(txt.replace("d", "e").replace("D", "e")) if txt else 0.0


def read_input(path: str) -> dict:
    """Parse an input.in the way the reference's set-up does: by line position and column."""
    with open(path) as f:
        lines = f.read().split("\n")
    cfg = {k: _field(lines[n - 1], kind) for k, (n, kind) in _LINES.items()}
    if cfg["flag_ctpress"] not in (0, 1):
        raise ValueError("Wrong presure flag")  # start.f90:203-204
    if cfg["flag_ctpress"] == 0:
        cfg["mpgx"] = 0.0
    return cfg
