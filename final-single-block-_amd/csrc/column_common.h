// Per-column helpers shared by the column-stage kernels (column_kernels.hip, column_pass.hip).
#pragma once
#include "dns_kernels.h"

namespace dnsk {

struct Mode {
  double kx, kz, k2x, k2z;
};

// def_k : start.f90:838-862 (k1F = i*k stored as its imaginary part, k2F = -k^2)
__device__ __forceinline__ Mode col_mode(const ColArgs &a, int col) {
  const int i = col % a.Kx, k = col / a.Kx;
  Mode m;
  m.kx = a.alp * (double)i;
  m.k2x = -((a.alp * (double)i) * (a.alp * (double)i));
  if (k < a.Nz / 2) {
    m.kz = a.bet * (double)k;
    m.k2z = -((a.bet * (double)k) * (a.bet * (double)k));
  } else {
    const double q = (double)(a.Nz - k);
    m.kz = -(a.bet * q);
    m.k2z = -((a.bet * q) * (a.bet * q));
  }
  return m;
}

#define D2U(c, j) a.dyu2i[3 * (j) + (c)-1]
#define D2V(c, j) a.dyv2i[3 * (j) + (c)-1]

// dt for this substep; at kRK = 1 the CFL rule of FOU3D.f90:290-303 with dtc_calc's maxima.
// dthetavi is never assigned in the reference (declaration.f90:64), so its v term is 1/(u2m*0) = Inf.
// With a fixed dt (a.fixed) every step takes a.fixed_dt and dtc stays the CFL diagnostic.
__device__ __forceinline__ double substep_dt(const ColArgs &a, int col) {
  if (a.kRK != 1) return a.st->dt;
  const double u1m = a.st->cfl[0], u3m = a.st->cfl[1];
  const double d1 = 1.0 / (a.alp * (double)(a.Nx / 2) * u1m), d3 = 1.0 / (a.bet * (double)(a.Nz / 2) * u3m);
  const double dtc = fmin(d1, d3);
  const double dt = a.fixed ? a.fixed_dt : fmin(2.5 * a.dtv, a.CFL * dtc);
  if (col == 0) {
    a.st->dtc = dtc;
    a.st->dt = dt;
    a.st->dti = 1.0 / dt;
    a.st->t = a.st->t + dt;
    if (!(isfinite(u1m) && isfinite(u3m))) a.st->nonfinite = 1.0;
  }
  return dt;
}

// pressure matrix rows (LU_buildP, tridLU_3D.f90:242-280); rows j = 1..nyu
__device__ __forceinline__ void dg_row(const ColArgs &a, int col, double k2x, double k2z, int j, double &a1,
                                       double &a2, double &a3) {
  const int e = a.nyu;
  if (col == 0 && j == 1) {
    a1 = 0.0;
    a2 = 1.0 / (a.dyp * a.dyp);
    a3 = 0.0;
    return;
  }
  if (j == 1) {
    a1 = 0.0;
    a2 = (a.dthdyu[1] * (-a.dthdyv[1])) + k2x + k2z;
    a3 = a.dthdyu[1] * a.dthdyv[2 - 1];
  } else if (j == e) {
    a1 = (-a.dthdyu[e]) * (-a.dthdyv[e - 1]);
    a2 = ((-a.dthdyu[e]) * a.dthdyv[e - 1]) + k2x + k2z;
    a3 = 0.0;
  } else {
    a1 = (-a.dthdyu[j]) * (-a.dthdyv[j - 1]);
    a2 = ((-a.dthdyu[j]) * a.dthdyv[j - 1]) + (a.dthdyu[j] * (-a.dthdyv[j])) + k2x + k2z;
    a3 = a.dthdyu[j] * a.dthdyv[j];
  }
}

}  // namespace dnsk
