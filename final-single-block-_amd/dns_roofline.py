"""Algorithmic (compulsory) HBM bytes of each kernel of one RK3 substep, and of the whole step.

Every kernel is charged the bytes it must read and write at least once: each input row read once, each
output row written once, no re-reads, no cache effects.  Spectral rows are complex FP64 (16 B per
retained mode, ncol = (Nx/2+1)*Nz modes per row).  Plane-stage intermediates hold Kx x Mz complex values
per field in the plane precision (16 B FP64, 8 B mixed).

Rows per field (device layout):
  u1, u3: 0..nyu+1     u2: 0..nyv+1     p, psi, Nu1, Nu3, Nxz1, Nxz3: 1..nyu    Nu2, Nxz2: 1..nyv
  uv_f, wv_f: 0..nyv+1 vv_c: 0..nyu+1   w3u: nyu+2 reals, w3v: nyv+2 reals, cP: nyu reals
"""
from __future__ import annotations

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def kernel_bytes(Mx: int, Mz: int, ny: int, fp32: bool = False) -> dict:
    Nx, Nz = 2 * Mx // 3, 2 * Mz // 3
    Kx = Nx // 2 + 1
    C = Kx * Nz
    nyv = ny - 2
    nyu = nyv + 1
    S = 16 * C                 # one spectral row (FP64 complex)
    Sf = (8 if fp32 else 16) * C   # one row of Nu, Nxz or a flux: stored in the plane precision
    R = 8 * C                  # one row of real per-column scratch
    I = Kx * Mz * (8 if fp32 else 16)   # one plane-stage field
    npl = 2 * nyu              # u-planes 1..nyu, v-planes 1..nyv+1
    b = {}
    # z-inverse: u1, u3 rows 1..nyu+1, u2 rows 0..nyv+1 read once; 3 fields per plane written
    b["zinv"] = 3 * (nyu + 1) * S + 3 * npl * I
    # x-pass: 3 fields per plane in; 4 products per u-plane, 2 per v-plane out
    b["xpass"] = 3 * npl * I + (4 * nyu + 2 * nyu) * I
    # z-forward: products in; Nxz1, Nxz3, vv_c per u-plane; Nxz2 (j <= nyv), uv_f, wv_f per v-plane
    b["zfwd"] = (4 * nyu + 2 * nyu) * I + (3 * nyu + 2 * nyu + nyv) * Sf
    # column sweep up: reads u1,u3 (nyu+2), u2 (nyv+2), p (nyu); Nu1/3 old, Nxz1/3, vv_c (nyu each), Nu2
    # old, Nxz2 (nyv each), uv_f, wv_f (nyu+1 each); writes u1,u3,u2 eliminated, Nu1/2/3 new, w3u, w3v
    rd = (2 * (nyu + 2) + (nyv + 2) + nyu) * S + (5 * nyu + 2 * nyv + 2 * (nyu + 1)) * Sf
    wr = (2 * (nyu + 2) + (nyv + 2)) * S + (2 * nyu + nyv) * Sf + ((nyu + 2) + (nyv + 2)) * R
    b["col_up"] = rd + wr
    # column sweep down: reads u1,u3,u2 (eliminated) and w3u, w3v; writes u1,u3 (0..nyu), u2 (0..nyv),
    # psi (1..nyu), cP
    b["col_down"] = ((2 * (nyu + 2) + (nyv + 2)) * S + ((nyu + 2) + (nyv + 2)) * R
                     + (2 * (nyu + 1) + (nyv + 1) + nyu) * S + nyu * R)
    # final sweep: reads psi, cP, p (1..nyu), u1, u3 (1..nyu), u2 (1..nyv); writes p, u1, u3, u2 (all rows)
    b["col_final"] = ((nyu + nyu + 2 * nyu + nyv) * S + nyu * R
                      + (nyu + 2 * (nyu + 2) + (nyv + 2)) * S)
    # one-pass column stage: every field read once and written once.  Reads u1, u3 (0..nyu+1), u2
    # (0..nyv+1), p, the old Nu, Nxz, uv_f, wv_f (0..nyv+1), vv_c (1..nyu+1); writes u1, u3, u2, p, Nu.
    # Nu, Nxz and the fluxes are stored in the plane precision.
    vel = (2 * (nyu + 2) + (nyv + 2) + nyu) * S
    b["col_pass"] = (2 * vel + 2 * (2 * nyu + nyv) * Sf + (2 * (nyv + 2) + (nyu + 1)) * Sf
                     + (2 * nyu + nyv) * Sf)
    b["substep_sweeps"] = sum(b[k] for k in ("zinv", "xpass", "zfwd", "col_up", "col_down", "col_final"))
    b["substep"] = b["zinv"] + b["xpass"] + b["zfwd"] + b["col_pass"]
    b["step"] = 3 * b["substep"]
    return b


def roofline(kernel: str, Mx: int, Mz: int, ny: int, fp32: bool, avg_ms: float) -> dict:
    """Achieved bandwidth of one kernel from its algorithmic bytes and its measured average duration."""
    by = kernel_bytes(Mx, Mz, ny, fp32)[kernel]
    gbs = by / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "bytes_per_launch": by}


def reference_design_substep_classes(Mx: int, Mz: int, ny: int, fp32: bool = False) -> dict:
    """The reference-design bytes of one substep split by kernel class: "fft_and_transposes" (the 12
    transposes, z-inverse, x-pass and z-forward terms) and "column" (21.5 Cb); see
    reference_design_step_bytes."""
    Nx, Nz = 2 * Mx // 3, 2 * Mz // 3
    Kx = Nx // 2 + 1
    sf = 4 if fp32 else 8
    S, I = Kx * Nz * 2 * sf, Kx * Mz * 2 * sf
    Cb = Kx * Nz * ny * 16
    nu, nv = ny - 1, ny - 2
    fft = (2 * (3 * nu + 3 * nv) * 2 * S + (3 * nu + 3 * nv) * (S + I) + nu * 7 * I + nv * 5 * I
           + nu * (4 * I + 3 * S) + nv * (2 * I + 3 * S))
    return {"fft_and_transposes": float(fft), "column": 21.5 * Cb}


def reference_design_step_bytes(Mx: int, Mz: int, ny: int, fp32: bool = False) -> float:
    """Compulsory HBM bytes of one RK3 step of the reference's own decomposition (columns -> 12 transposes
    -> plane FFTs -> column solves), per substep:
      transposes (3nu + 3nv) 2S twice, z-inverse (3nu + 3nv)(S + I), fused x-pass nu 7I + nv 5I,
      z-forward nu (4I + 3S) + nv (2I + 3S), column stage 21.5 Cb,
    with S = Kx Nz 2s_f, I = Kx Mz 2s_f (s_f = 8 FP64, 4 mixed), Cb = Kx Nz ny 16, nu = ny - 1, nv = ny - 2.
    The device layout here needs no transposes, so its own compulsory bytes (kernel_bytes) are lower;
    this figure prices a step the way the reference's data movement would."""
    c = reference_design_substep_classes(Mx, Mz, ny, fp32)
    return 3.0 * (c["fft_and_transposes"] + c["column"])
