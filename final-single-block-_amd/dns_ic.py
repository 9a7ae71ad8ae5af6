"""Seeded initial conditions in the reference's column layout.

The reference's only working initial condition reads a field file (start.f90:1251-1282); its parabolic
option is a stub (start.f90:1287).  This generator builds a laminar Poiseuille profile plus a seeded
perturbation that is divergence-free for the reference's own discrete operator
(divergence, littleharsh.f90:242-278):

  div_j = i kx u1_j + i kz u3_j + (v_j - v_{j-1}) dthdyu_j = 0,

by choosing v on the faces (zero at both walls), D_j = (v_j - v_{j-1}) dthdyu_j, and
  u1 = (i kx / k^2) D + i kz phi,   u3 = (i kz / k^2) D - i kx phi
for any phi(y) (the second part is solenoidal by itself).  Ghost rows follow v_corr's wall rule
u(0) = gbc u(1), u(nyu+1) = gbc u(nyu) (littleharsh.f90:935-984).  Modes with kx = 0 are made
Hermitian (mode -kz = conjugate of +kz) so the physical field is real; the kz = -Nz/2 row is left zero.
The fields can be written as reference field files (dns_io.write_field) so that the reference reads the
same initial state with flag_init = 1.
"""
from __future__ import annotations

import numpy as np

import dnship


def seeded_channel(Mx, Mz, ny, seed=1, amp=0.1, nmodes=8, dyq=10.0, ppp=5, bulk=1.0, broadband=False):
    """Returns (u1, u2, u3, p) complex128 arrays [ncol, nrows] in the reference layout.

    ``nmodes``: perturb |kx index| <= nmodes and |kz index| <= nmodes (ignored when ``broadband``: then
    every retained mode is perturbed with a k^-2 amplitude spectrum)."""
    Nx, Nz = 2 * Mx // 3, 2 * Mz // 3
    Kx = Nx // 2 + 1
    C = Kx * Nz
    nyv = ny - 2
    nyu = nyv + 1
    g = dnship.ygrid(ny, dyq, ppp)
    yu, yv, dthdyu = g["yu"], g["yv"], g["dthdyu"]
    gbc = (yu[0] - yv[0]) / (yu[1] - yv[0])
    rng = np.random.default_rng(seed)
    u1 = np.zeros((C, ny + 1), np.complex128)
    u3 = np.zeros((C, ny + 1), np.complex128)
    u2 = np.zeros((C, ny), np.complex128)
    p = np.zeros((C, ny - 1), np.complex128)
    # laminar Poiseuille profile with bulk velocity `bulk`, on the u grid interior
    u1[0, 1:nyu + 1] = 1.5 * bulk * (1.0 - yu[1:nyu + 1] ** 2)
    prof_v = (1.0 - yv ** 2) ** 2
    prof_v[0] = prof_v[-1] = 0.0
    prof_u = 1.0 - yu ** 2

    def kz_of(k):
        return k if k < Nz // 2 else k - Nz

    modes = []
    for k in range(Nz):
        kz = kz_of(k)
        if kz == -Nz // 2:
            continue
        for i in range(Kx):
            if i == 0 and kz <= 0:
                continue  # kx = 0: kz > 0 drawn, kz < 0 set by symmetry; (0,0) is the mean
            if not broadband and (i > nmodes or abs(kz) > nmodes):
                continue
            modes.append((i, k, kz))
    nm = max(len(modes), 1)
    for (i, k, kz) in modes:
        kx_w, kz_w = 1.0 * i, 2.0 * kz  # alp=1, bet=2 wavenumbers only shape the draw
        scale = amp / np.sqrt(nm)
        if broadband:
            scale = amp * 4.0 / (1.0 + kx_w ** 2 + kz_w ** 2) / np.sqrt(nm) * np.sqrt(nm / 10.0)
        a = scale * (rng.standard_normal() + 1j * rng.standard_normal())
        b = scale * (rng.standard_normal() + 1j * rng.standard_normal())
        col = k * Kx + i
        v = a * prof_v
        D = np.zeros(ny + 1, np.complex128)
        D[1:nyu + 1] = (v[1:nyu + 1] - v[0:nyu]) * dthdyu[1:nyu + 1]
        kx, kzz = kx_w, kz_w
        k2 = kx * kx + kzz * kzz
        phi = b * prof_u
        uu = (1j * kx / k2) * D + 1j * kzz * phi
        ww = (1j * kzz / k2) * D - 1j * kx * phi
        uu[0] = gbc * uu[1]
        uu[nyu + 1] = gbc * uu[nyu]
        ww[0] = gbc * ww[1]
        ww[nyu + 1] = gbc * ww[nyu]
        u2[col] = v
        u1[col] = uu
        u3[col] = ww
        if i == 0:
            kneg = (Nz - kz) % Nz
            cneg = kneg * Kx
            u2[cneg] = np.conj(v)
            u1[cneg] = np.conj(uu)
            u3[cneg] = np.conj(ww)
    u1[0, 0] = gbc * u1[0, 1]
    u1[0, nyu + 1] = gbc * u1[0, nyu]
    return u1, u2, u3, p


def divergence(u1, u2, u3, Nx, Nz, ny, dyq=10.0, ppp=5, alp=1.0, bet=2.0):
    """Discrete divergence of littleharsh.f90:242-278 at centres j = 1..nyu (for checks)."""
    Kx = Nx // 2 + 1
    g = dnship.ygrid(ny, dyq, ppp)
    nyu = ny - 1
    i = np.arange(u1.shape[0]) % Kx
    k = np.arange(u1.shape[0]) // Kx
    kz = np.where(k < Nz // 2, k, k - Nz)
    ikx = 1j * alp * i
    ikz = 1j * bet * kz
    return (ikx[:, None] * u1[:, 1:nyu + 1] + ikz[:, None] * u3[:, 1:nyu + 1]
            + g["dthdyu"][None, 1:nyu + 1] * (u2[:, 1:nyu + 1] - u2[:, 0:nyu]))
