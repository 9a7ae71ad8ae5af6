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

// meanflow_ctU for the one-pass column stage, which has already run solveP and v_corr.  For mode (0,0)
// (kx = kz = 0) neither the divergence nor v_corr's interior update reads u1 (both multiply it by kx = 0),
// so the correction commutes with them: the interior rows get + g u11 exactly as flowrate_corr adds it
// (littleharsh.f90:515-548), and the ghost rows are then v_corr's wall rows of the corrected interior
// (littleharsh.f90:935-943), which overwrite what flowrate_corr added there.  The block stages the
// matrix rows, yv and u1's mode (0,0) in LDS; one thread runs LUsol0's sweeps in the reference's order
// (tridLU_3D.f90:134-160, 299-325) and the two flow-rate sums.
constexpr int MEANFLOW_MAXN = 516;  // nyu + 2 rows, nyu <= 512 (the one-pass column stage's range)
// The block's LDS: matrix rows, yv, u1's mode (0,0), the solution and the elimination's w3, and g.
struct MeanflowLds {
  double c1[MEANFLOW_MAXN], c2[MEANFLOW_MAXN], c3[MEANFLOW_MAXN], yv[MEANFLOW_MAXN], ur[MEANFLOW_MAXN],
      u11[MEANFLOW_MAXN], w3[MEANFLOW_MAXN];
  double g;
};
// Every thread of the block calls it (blockDim.x threads; it has three barriers).
__device__ __forceinline__ void meanflow_post_body(const ColArgs &a_, int member, MeanflowLds &L) {
  const ColArgs a = col_view(a_, member);
  double *c1 = L.c1, *c2 = L.c2, *c3 = L.c3, *yv = L.yv, *ur = L.ur, *u11 = L.u11, *w3 = L.w3;
  double &gsh = L.g;
  const int n = a.nyu + 2, C = a.C;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    if (j >= 1 && j < n - 1) {
      c1[j] = D2U(1, j);
      c2[j] = D2U(2, j);
      c3[j] = D2U(3, j);
    }
    if (j <= a.nyv + 1) yv[j] = a.yv[j];
    ur[j] = a.u1[(size_t)j * C].x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double dt = a.st->dt;
    const double beta = a.bRK * dt / a.Re;
    const double bdt = -a.dRK * dt;
    double w2 = 1.0 / 1.0, y = 0.0 * w2, w3p = a.gw1 * w2;
    u11[0] = y;
    w3[0] = w3p;
    for (int j = 1; j < n; j++) {
      double a1, a2, a3, r;
      if (j < n - 1) {
        a1 = -beta * c1[j];
        a2 = 1.0 - beta * c2[j];
        a3 = -beta * c3[j];
        r = bdt;
      } else {
        a1 = a.gw2; a2 = 1.0; a3 = 0.0; r = 0.0;
      }
      w2 = 1.0 / (a2 - a1 * w3p);
      y = (r - a1 * y) * w2;
      w3p = a3 * w2;
      u11[j] = y;
      w3[j] = w3p;
    }
    double x = y;
    for (int j = n - 2; j >= 0; j--) {
      x = u11[j] - w3[j] * x;
      u11[j] = x;
    }
    double Qx1 = 0.0, Qx0 = 0.0;
    // flowrateRe's dummy u(1:nn+2) starts at u11(0) (littleharsh.f90:505-511 with start.f90:261)
    for (int j = 1; j <= a.nyv + 1; j++) Qx1 = Qx1 + (yv[j] - yv[j - 1]) * u11[j - 1];
    for (int j = 1; j <= a.nyv + 1; j++) Qx0 = Qx0 + (yv[j] - yv[j - 1]) * ur[j];
    const double g = (a.st->QxT - Qx0) / Qx1;
    a.st->mpgx = a.st->mpgx + g;
    a.st->dgx = g;
    gsh = g;
  }
  __syncthreads();
  const double g = gsh;
  for (int j = 1 + threadIdx.x; j <= a.nyu; j += blockDim.x) {
    const double v = ur[j] + g * u11[j];
    ur[j] = v;
    a.u1[(size_t)j * C].x = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // v_corr's wall rows of the corrected interior
    a.u1[0] = rmul(a.gbc, d2(ur[1], a.u1[C].y));
    a.u1[(size_t)(a.nyu + 1) * C] = rmul(a.gbc, d2(ur[a.nyu], a.u1[(size_t)a.nyu * C].y));
  }
}

}  // namespace dnsk
