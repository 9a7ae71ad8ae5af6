// Column stage: everything of the RK3 substep that is local to one (kx, kz) column.
//
// Thread per column; consecutive threads own consecutive columns, so every y-row access of the
// plane-major device layout is a coalesced run of the columns' values.  Two paths:
//  * staged: one kernel per reference subroutine, same arithmetic order as the reference
//      rhs0_kernel    RHS0_u1/u2/u3 + laplacian_U/V  littleharsh.f90:281-368, 682-856
//      nl_col_kernel  der_yu_h/der_yv_h + sum + du -= g N, dt at kRK=1   FOU3D.f90:268-303, 359-368
//      solveU_kernel  solveU x3 -> LU_build/LU_dec/LUsolU   littleharsh.f90:863-894, tridLU_3D.f90:38-221
//      meanflow_kernel meanflow_ctU/flowrate_corr/LUsol0    littleharsh.f90:515-607, tridLU_3D.f90:134-325
//      solveP_kernel  divergence + LU_buildP/LU_decP/LUsolP + p += psi   littleharsh.f90:242-278, 626-680
//      vcorr_kernel   v_corr + divergence                 littleharsh.f90:896-993
//  * fused: three sweeps per column for a whole substep
//      fused_up    (j up)  : RHS0 + nonlinear column part + du + forward elimination of the 3 solves
//      fused_down  (j down): back substitution, divergence, pressure elimination (top-down)
//      fused_final (j up)  : pressure back substitution, p += psi, v_corr, wall rows, max|div|
//    The pressure system is eliminated top-down (UL) so that it can start while the velocity back
//    substitution runs downward; it has the same unique solution as the reference's LU order.
#include "column_common.h"

namespace dnsk {

// ================================================================================================
// staged path
__global__ void rhs0_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.C) return;
  const int C = a.C;
  const Mode m = col_mode(a, col);
  const double C1 = a.aRK / a.Re, C2 = -a.cRK;
  const double c3x = -(a.dRK * m.kx), c3z = -(a.dRK * m.kz);
  const double k2 = m.k2x + m.k2z;
  for (int j = 1; j <= a.nyu; j++) {
    const double d1 = D2U(1, j), d2 = k2 + D2U(2, j), d3 = D2U(3, j);
    const double2 am = a.u1[(size_t)(j - 1) * C + col], a0 = a.u1[(size_t)j * C + col], ap = a.u1[(size_t)(j + 1) * C + col];
    const double2 bm = a.u3[(size_t)(j - 1) * C + col], b0 = a.u3[(size_t)j * C + col], bp = a.u3[(size_t)(j + 1) * C + col];
    const double2 L1 = cadd(cadd(rmul(d1, am), rmul(d2, a0)), rmul(d3, ap));
    const double2 L3 = cadd(cadd(rmul(d1, bm), rmul(d2, b0)), rmul(d3, bp));
    const double2 pj = a.p[(size_t)(j - 1) * C + col];
    double2 r1 = cadd(cadd(rmul(C1, L1), rmul(C2, a.Nu1[(size_t)(j - 1) * C + col])), imul(c3x, pj));
    const double2 r3 = cadd(cadd(rmul(C1, L3), rmul(C2, a.Nu3[(size_t)(j - 1) * C + col])), imul(c3z, pj));
    if (col == 0) r1.x = r1.x + (-(a.dRK * a.st->mpgx));  // littleharsh.f90:727-733
    a.du1[(size_t)j * C + col] = r1;
    a.du3[(size_t)j * C + col] = r3;
  }
  for (int j = 1; j <= a.nyv; j++) {
    const double d1 = D2V(1, j), d2 = k2 + D2V(2, j), d3 = D2V(3, j);
    const double2 vm = a.u2[(size_t)(j - 1) * C + col], v0 = a.u2[(size_t)j * C + col], vp = a.u2[(size_t)(j + 1) * C + col];
    const double2 L2 = cadd(cadd(rmul(d1, vm), rmul(d2, v0)), rmul(d3, vp));
    const double C3 = -((a.dRK * 1.0) * a.dthdyv[j]);
    const double2 dp = csub(a.p[(size_t)j * C + col], a.p[(size_t)(j - 1) * C + col]);
    a.du2[(size_t)j * C + col] = cadd(cadd(rmul(C1, L2), rmul(C2, a.Nu2[(size_t)(j - 1) * C + col])), rmul(C3, dp));
  }
}

__global__ void nl_col_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.C) return;
  const int C = a.C;
  (void)substep_dt(a, col);
  const double Cg = -a.gRK;
  for (int j = 1; j <= a.nyu; j++) {
    const double2 dy1 = rmul(a.dthdyu[j], csub(a.uvf[(size_t)j * C + col], a.uvf[(size_t)(j - 1) * C + col]));
    const double2 dy3 = rmul(a.dthdyu[j], csub(a.wvf[(size_t)j * C + col], a.wvf[(size_t)(j - 1) * C + col]));
    const double2 N1 = cadd(a.Nxz1[(size_t)(j - 1) * C + col], dy1);
    const double2 N3 = cadd(a.Nxz3[(size_t)(j - 1) * C + col], dy3);
    a.Nu1.set((size_t)(j - 1) * C + col, N1);
    a.Nu3.set((size_t)(j - 1) * C + col, N3);
    a.du1[(size_t)j * C + col] = cadd(a.du1[(size_t)j * C + col], rmul(Cg, N1));
    a.du3[(size_t)j * C + col] = cadd(a.du3[(size_t)j * C + col], rmul(Cg, N3));
  }
  for (int j = 1; j <= a.nyv; j++) {
    const double2 dy2 = rmul(a.dthdyv[j], csub(a.vvc[(size_t)(j + 1) * C + col], a.vvc[(size_t)j * C + col]));
    const double2 N2 = cadd(a.Nxz2[(size_t)(j - 1) * C + col], dy2);
    a.Nu2.set((size_t)(j - 1) * C + col, N2);
    a.du2[(size_t)j * C + col] = cadd(a.du2[(size_t)j * C + col], rmul(Cg, N2));
  }
}

// LU_build (tridLU_3D.f90:163-221) + LU_dec (:327-364) + LUsolU (:70-83) for one column of one grid.
// rows 0..n-1; u holds the field on entry (rhs = u + dt*du) and the solution on exit.
__device__ __forceinline__ void solve_column(const ColArgs &a, bool ugrid, double dt, double beta, double k2x,
                                             double k2z, double2 *u, const double2 *du, double *w3, int col,
                                             int n) {
  const int C = a.C;
  // forward: LU_dec fused with the forward substitution (identical operations per row)
  double w2 = 1.0 / 1.0;
  double2 y = rmul(w2, d2(0.0, 0.0));  // rhs(0) = 0 (littleharsh.f90:882)
  double w3p = (ugrid ? a.gw1 : 0.0) * w2;
  u[col] = y;
  w3[col] = w3p;
  for (int j = 1; j < n; j++) {
    double a1, a2, a3;
    double2 rhs;
    if (j < n - 1) {
      const double c1 = ugrid ? D2U(1, j) : D2V(1, j), c2 = ugrid ? D2U(2, j) : D2V(2, j),
                   c3 = ugrid ? D2U(3, j) : D2V(3, j);
      a1 = -beta * c1;
      a2 = 1.0 - beta * (c2 + k2x + k2z);
      a3 = -beta * c3;
      const double2 uj = u[(size_t)j * C + col], dj = du[(size_t)j * C + col];
      rhs = d2(uj.x + dt * dj.x, uj.y + dt * dj.y);
    } else {
      a1 = ugrid ? a.gw2 : 0.0;
      a2 = 1.0;
      a3 = 0.0;
      rhs = d2(0.0, 0.0);
    }
    w2 = 1.0 / (a2 - a1 * w3p);
    y = d2((rhs.x - a1 * y.x) * w2, (rhs.y - a1 * y.y) * w2);
    w3p = a3 * w2;
    u[(size_t)j * C + col] = y;
    w3[(size_t)j * C + col] = w3p;
  }
  // back substitution
  double2 x = y;
  for (int j = n - 2; j >= 0; j--) {
    const double2 yj = u[(size_t)j * C + col];
    const double c = w3[(size_t)j * C + col];
    x = d2(yj.x - c * x.x, yj.y - c * x.y);
    u[(size_t)j * C + col] = x;
  }
}

__global__ void solveU_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.C) return;
  const Mode m = col_mode(a, col);
  const double dt = a.st->dt;
  const double beta = a.bRK * dt / a.Re;
  solve_column(a, true, dt, beta, m.k2x, m.k2z, a.u1, a.du1, a.w3u, col, a.nyu + 2);
  solve_column(a, false, dt, beta, m.k2x, m.k2z, a.u2, a.du2, a.w3v, col, a.nyv + 2);
  solve_column(a, true, dt, beta, m.k2x, m.k2z, a.u3, a.du3, a.w3u, col, a.nyu + 2);
}

// flowrate_corr (littleharsh.f90:515-548) with LUsol0/LU_build0 (tridLU_3D.f90:134-160, 299-325);
// one thread: mode (0,0) is column 0.
__global__ void meanflow_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int n = a.nyu + 2, C = a.C;
  const double dt = a.st->dt;
  const double beta = a.bRK * dt / a.Re;
  const double bdt = -a.dRK * dt;
  double *u11 = a.mscr, *w3 = a.mscr + n;
  double w2 = 1.0 / 1.0, y = 0.0 * w2, w3p = a.gw1 * w2;
  u11[0] = y;
  w3[0] = w3p;
  for (int j = 1; j < n; j++) {
    double a1, a2, a3, r;
    if (j < n - 1) {
      a1 = -beta * D2U(1, j);
      a2 = 1.0 - beta * D2U(2, j);
      a3 = -beta * D2U(3, j);
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
  // flowrateRe's dummy u(1:nn+2) starts at u11(0) (littleharsh.f90:505-511 with start.f90:261): u(j) = u11(j-1)
  for (int j = 1; j <= a.nyv + 1; j++) Qx1 = Qx1 + (a.yv[j] - a.yv[j - 1]) * u11[j - 1];
  for (int j = 1; j <= a.nyv + 1; j++) Qx0 = Qx0 + (a.yv[j] - a.yv[j - 1]) * a.u1[(size_t)j * C].x;
  const double g = (a.st->QxT - Qx0) / Qx1;
  a.st->mpgx = a.st->mpgx + g;
  a.st->dgx = g;
  for (int j = 0; j < n; j++) a.u1[(size_t)j * C].x = a.u1[(size_t)j * C].x + g * u11[j];
}

__global__ void __launch_bounds__(256) meanflow_post_kernel(ColArgs a) {
  __shared__ MeanflowLds L;
  meanflow_post_body(a, blockIdx.y, L);
}

// one block per member of the groups with flags[g] = 1 (constant flow rate)
__global__ void __launch_bounds__(256) ens_meanflow_post_kernel(EnsColLaunch E, int4 flags) {
  const int f[4] = {flags.x, flags.y, flags.z, flags.w};
  int m = blockIdx.x, g = 0;
  for (; g < E.ng; g++) {
    const int n = f[g] ? E.a[g].nm : 0;
    if (m < n) break;
    m -= n;
  }
  __shared__ MeanflowLds L;
  if (g < E.ng) meanflow_post_body(E.a[g], m, L);
}
void launch_ens_meanflow_post(const EnsColLaunch &E, const int *flags, hipStream_t st) {
  static_assert(ENS_MAXG == 4, "flags travel as an int4");
  int total = 0;
  for (int g = 0; g < E.ng; g++) total += flags[g] ? E.a[g].nm : 0;
  if (total == 0) return;
  const int4 f = make_int4(flags[0], E.ng > 1 ? flags[1] : 0, E.ng > 2 ? flags[2] : 0, E.ng > 3 ? flags[3] : 0);
  ens_meanflow_post_kernel<<<total, 256, 0, st>>>(E, f);
}

// flowrateIm of mode (0,0) : littleharsh.f90:473-492, meanflow_ctP :572-589
__global__ void qx_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double Q = 0.0;
  for (int j = 1; j <= a.nyv + 1; j++) Q = Q + (a.yv[j] - a.yv[j - 1]) * a.u1[(size_t)j * a.C].x;
  a.st->Qx = Q;
}

__global__ void solveP_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.C) return;
  const int C = a.C, e = a.nyu;
  const Mode m = col_mode(a, col);
  const double dtirk = a.st->dti / (a.aRK + a.bRK);
  // divergence (littleharsh.f90:264-276), psi = dtirk*div, pin, and the LU forward sweep
  double w2 = 0.0, w3p = 0.0;
  double2 x = d2(0.0, 0.0);
  for (int j = 1; j <= e; j++) {
    const double2 dv = cadd(cadd(imul(m.kx, a.u1[(size_t)j * C + col]), imul(m.kz, a.u3[(size_t)j * C + col])),
                            rmul(1.0 * a.dthdyu[j], csub(a.u2[(size_t)j * C + col], a.u2[(size_t)(j - 1) * C + col])));
    a.div[(size_t)(j - 1) * C + col] = dv;
    double2 r = rmul(dtirk, dv);
    if (col == 0 && j == 1) r = d2(0.0, 0.0);
    double a1, a2, a3;
    dg_row(a, col, m.k2x, m.k2z, j, a1, a2, a3);
    if (j == 1) {
      w2 = 1.0 / a2;
      x = rmul(w2, r);
    } else {
      const double w3prev = a.cP[(size_t)(j - 2) * C + col];
      w2 = 1.0 / (a2 - a1 * w3prev);
      x = d2((r.x - a1 * x.x) * w2, (r.y - a1 * x.y) * w2);
    }
    a.cP[(size_t)(j - 1) * C + col] = a3 * w2;
    a.psi[(size_t)(j - 1) * C + col] = x;
  }
  for (int j = e - 1; j >= 1; j--) {
    const double2 xj = a.psi[(size_t)(j - 1) * C + col];
    const double c = a.cP[(size_t)(j - 1) * C + col];
    x = d2(xj.x - c * x.x, xj.y - c * x.y);
    a.psi[(size_t)(j - 1) * C + col] = x;
  }
  for (int j = 1; j <= e; j++) {
    const size_t o = (size_t)(j - 1) * C + col;
    a.p[o] = cadd(a.p[o], a.psi[o]);
  }
}

template <int NTH>
__global__ void __launch_bounds__(NTH) vcorr_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  __shared__ double red[NTH / 64];
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  double dmax = 0.0;
  if (col < a.C) {
    const int C = a.C;
    const Mode m = col_mode(a, col);
    const double dtrk = a.st->dt * (a.aRK + a.bRK);
    const double kxd = dtrk * m.kx, kzd = dtrk * m.kz;
    for (int j = 1; j <= a.nyu; j++) {
      const double2 ps = a.psi[(size_t)(j - 1) * C + col];
      a.u1[(size_t)j * C + col] = csub(a.u1[(size_t)j * C + col], imul(kxd, ps));
      a.u3[(size_t)j * C + col] = csub(a.u3[(size_t)j * C + col], imul(kzd, ps));
    }
    a.u1[col] = rmul(a.gbc, a.u1[(size_t)C + col]);
    a.u1[(size_t)(a.nyu + 1) * C + col] = rmul(a.gbc, a.u1[(size_t)a.nyu * C + col]);
    a.u3[col] = rmul(a.gbc, a.u3[(size_t)C + col]);
    a.u3[(size_t)(a.nyu + 1) * C + col] = rmul(a.gbc, a.u3[(size_t)a.nyu * C + col]);
    for (int j = 1; j <= a.nyv; j++) {
      const double2 dps = csub(a.psi[(size_t)j * C + col], a.psi[(size_t)(j - 1) * C + col]);
      a.u2[(size_t)j * C + col] = csub(a.u2[(size_t)j * C + col], rmul(dtrk * 1.0 * a.dthdyv[j], dps));
    }
    a.u2[col] = d2(0.0, 0.0);
    a.u2[(size_t)(a.nyv + 1) * C + col] = d2(0.0, 0.0);
    for (int j = 1; j <= a.nyu; j++) {
      const double2 dv = cadd(cadd(imul(m.kx, a.u1[(size_t)j * C + col]), imul(m.kz, a.u3[(size_t)j * C + col])),
                              rmul(1.0 * a.dthdyu[j], csub(a.u2[(size_t)j * C + col], a.u2[(size_t)(j - 1) * C + col])));
      a.div[(size_t)(j - 1) * C + col] = dv;
      dmax = fmax(dmax, hypot(dv.x, dv.y));
    }
  }
  dmax = block_max<NTH>(dmax, red);
  if (threadIdx.x == 0) atomic_max_nonneg(&a.st->divmax, dmax);
}

// ================================================================================================
// fused path
__global__ void __launch_bounds__(256) fused_up_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.C) return;
  const int C = a.C, nyu = a.nyu, nyv = a.nyv;
  const Mode m = col_mode(a, col);
  const double dt = substep_dt(a, col);
  const double C1 = a.aRK / a.Re, C2 = -a.cRK, Cg = -a.gRK;
  const double beta = a.bRK * dt / a.Re;
  const double c3x = -(a.dRK * m.kx), c3z = -(a.dRK * m.kz);
  const double Cm = col == 0 ? -(a.dRK * a.st->mpgx) : 0.0;
  const double k2 = m.k2x + m.k2z;
  // sliding windows of the old fields
  double2 u1m = a.u1[col], u1c = a.u1[(size_t)C + col];
  double2 u3m = a.u3[col], u3c = a.u3[(size_t)C + col];
  double2 u2m = a.u2[col], u2c = a.u2[(size_t)C + col];
  double2 uvm = a.uvf[col], wvm = a.wvf[col];
  double2 vvc_c = a.vvc[(size_t)C + col];
  double2 pc = a.p[col];
  // row 0 of both grids: rhs = 0, (a1, a2, a3) = (0, 1, gw1) or (0, 1, 0)
  double w2 = 1.0 / 1.0;
  double2 y1 = rmul(w2, d2(0.0, 0.0)), y3 = y1, y2 = y1;
  double w3u = a.gw1 * w2, w3v = 0.0 * w2;
  a.u1[col] = y1;
  a.u3[col] = y3;
  a.u2[col] = y2;
  a.w3u[col] = w3u;
  a.w3v[col] = w3v;
  for (int j = 1; j <= nyu; j++) {
    // ---------------- u-grid row j (interior) ----------------
    const double2 u1p = a.u1[(size_t)(j + 1) * C + col], u3p = a.u3[(size_t)(j + 1) * C + col];
    const double2 uvc = a.uvf[(size_t)j * C + col], wvc = a.wvf[(size_t)j * C + col];
    const double d1 = D2U(1, j), dd2 = k2 + D2U(2, j), d3 = D2U(3, j);
    const double2 L1 = cadd(cadd(rmul(d1, u1m), rmul(dd2, u1c)), rmul(d3, u1p));
    const double2 L3 = cadd(cadd(rmul(d1, u3m), rmul(dd2, u3c)), rmul(d3, u3p));
    double2 r1 = cadd(cadd(rmul(C1, L1), rmul(C2, a.Nu1[(size_t)(j - 1) * C + col])), imul(c3x, pc));
    double2 r3 = cadd(cadd(rmul(C1, L3), rmul(C2, a.Nu3[(size_t)(j - 1) * C + col])), imul(c3z, pc));
    r1.x = r1.x + Cm;
    const double2 N1 = cadd(a.Nxz1[(size_t)(j - 1) * C + col], rmul(a.dthdyu[j], csub(uvc, uvm)));
    const double2 N3 = cadd(a.Nxz3[(size_t)(j - 1) * C + col], rmul(a.dthdyu[j], csub(wvc, wvm)));
    r1 = cadd(r1, rmul(Cg, N1));
    r3 = cadd(r3, rmul(Cg, N3));
    a.Nu1.set((size_t)(j - 1) * C + col, N1);
    a.Nu3.set((size_t)(j - 1) * C + col, N3);
    const double2 rhs1 = d2(u1c.x + dt * r1.x, u1c.y + dt * r1.y);
    const double2 rhs3 = d2(u3c.x + dt * r3.x, u3c.y + dt * r3.y);
    {
      const double a1 = -beta * D2U(1, j), a2 = 1.0 - beta * (D2U(2, j) + m.k2x + m.k2z), a3 = -beta * D2U(3, j);
      w2 = 1.0 / (a2 - a1 * w3u);
      y1 = d2((rhs1.x - a1 * y1.x) * w2, (rhs1.y - a1 * y1.y) * w2);
      y3 = d2((rhs3.x - a1 * y3.x) * w2, (rhs3.y - a1 * y3.y) * w2);
      w3u = a3 * w2;
    }
    a.u1[(size_t)j * C + col] = y1;
    a.u3[(size_t)j * C + col] = y3;
    a.w3u[(size_t)j * C + col] = w3u;
    u1m = u1c; u1c = u1p;
    u3m = u3c; u3c = u3p;
    uvm = uvc; wvm = wvc;
    // ---------------- v-grid row j ----------------
    const double2 u2p = (j <= nyv) ? a.u2[(size_t)(j + 1) * C + col] : d2(0.0, 0.0);
    if (j <= nyv) {
      const double2 pn = a.p[(size_t)j * C + col];
      const double2 vvn = a.vvc[(size_t)(j + 1) * C + col];
      const double e1 = D2V(1, j), e2 = k2 + D2V(2, j), e3 = D2V(3, j);
      const double2 L2 = cadd(cadd(rmul(e1, u2m), rmul(e2, u2c)), rmul(e3, u2p));
      const double C3 = -((a.dRK * 1.0) * a.dthdyv[j]);
      double2 r2 = cadd(cadd(rmul(C1, L2), rmul(C2, a.Nu2[(size_t)(j - 1) * C + col])), rmul(C3, csub(pn, pc)));
      const double2 N2 = cadd(a.Nxz2[(size_t)(j - 1) * C + col], rmul(a.dthdyv[j], csub(vvn, vvc_c)));
      r2 = cadd(r2, rmul(Cg, N2));
      a.Nu2.set((size_t)(j - 1) * C + col, N2);
      const double2 rhs2 = d2(u2c.x + dt * r2.x, u2c.y + dt * r2.y);
      const double a1 = -beta * D2V(1, j), a2 = 1.0 - beta * (D2V(2, j) + m.k2x + m.k2z), a3 = -beta * D2V(3, j);
      w2 = 1.0 / (a2 - a1 * w3v);
      y2 = d2((rhs2.x - a1 * y2.x) * w2, (rhs2.y - a1 * y2.y) * w2);
      w3v = a3 * w2;
      pc = pn;
      vvc_c = vvn;
    } else {  // top wall row of v: (0, 1, 0), rhs 0
      w2 = 1.0 / (1.0 - 0.0 * w3v);
      y2 = d2((0.0 - 0.0 * y2.x) * w2, (0.0 - 0.0 * y2.y) * w2);
      w3v = 0.0 * w2;
    }
    a.u2[(size_t)j * C + col] = y2;
    a.w3v[(size_t)j * C + col] = w3v;
    u2m = u2c; u2c = u2p;
  }
  // u-grid top ghost row: (gw2, 1, 0), rhs 0
  w2 = 1.0 / (1.0 - a.gw2 * w3u);
  y1 = d2((0.0 - a.gw2 * y1.x) * w2, (0.0 - a.gw2 * y1.y) * w2);
  y3 = d2((0.0 - a.gw2 * y3.x) * w2, (0.0 - a.gw2 * y3.y) * w2);
  a.u1[(size_t)(nyu + 1) * C + col] = y1;
  a.u3[(size_t)(nyu + 1) * C + col] = y3;
  a.w3u[(size_t)(nyu + 1) * C + col] = 0.0 * w2;
}

// The same sweep with the three solves on three wavefronts of one workgroup (wave 0: u1, wave 1: u3,
// wave 2: u2; 64 columns per workgroup): three times the wavefronts of the thread-per-column form, and
// each one streams a third of the rows, so more memory requests are in flight.  With PF the next row's
// inputs are loaded before the current row is computed.  Arithmetic identical to fused_up_kernel.
template <bool PF>
__global__ void __launch_bounds__(192) fused_up3_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  const int f = threadIdx.x >> 6;
  const int col = blockIdx.x * 64 + (threadIdx.x & 63);
  if (col >= a.C) return;
  const int C = a.C, nyu = a.nyu, nyv = a.nyv;
  const Mode m = col_mode(a, col);
  const double dt = substep_dt(a, f == 0 ? col : -1);  // the state's dt and t are written once
  const double C1 = a.aRK / a.Re, C2 = -a.cRK, Cg = -a.gRK;
  const double beta = a.bRK * dt / a.Re;
  const double k2 = m.k2x + m.k2z;
  double w2 = 1.0 / 1.0;
  if (f < 2) {
    double2 *U = f == 0 ? a.u1 : a.u3;
    const NArr Nu = f == 0 ? a.Nu1 : a.Nu3;
    const NArr Nxz = f == 0 ? a.Nxz1 : a.Nxz3;
    const NArr fl = f == 0 ? a.uvf : a.wvf;
    const double c3 = f == 0 ? -(a.dRK * m.kx) : -(a.dRK * m.kz);
    const double Cm = col == 0 ? -(a.dRK * a.st->mpgx) : 0.0;
    double2 um = U[col], uc = U[(size_t)C + col];
    double2 flm = fl[col];
    double2 pc = a.p[col];
    double2 y = rmul(w2, d2(0.0, 0.0));
    double w3 = a.gw1 * w2;
    U[col] = y;
    if (f == 0) a.w3u[col] = w3;
    struct Row { double2 up, flc, nu, nx, pn; };
    auto load = [&](int j) {
      Row r;
      r.up = U[(size_t)(j + 1) * C + col];
      r.flc = fl[(size_t)j * C + col];
      r.nu = Nu[(size_t)(j - 1) * C + col];
      r.nx = Nxz[(size_t)(j - 1) * C + col];
      r.pn = (j <= nyv) ? a.p[(size_t)j * C + col] : d2(0.0, 0.0);
      return r;
    };
    Row cur = load(1);
    for (int j = 1; j <= nyu; j++) {
      Row nxt;
      if (PF && j < nyu) nxt = load(j + 1);
      const double d1 = D2U(1, j), dd2 = k2 + D2U(2, j), d3 = D2U(3, j);
      const double2 L = cadd(cadd(rmul(d1, um), rmul(dd2, uc)), rmul(d3, cur.up));
      double2 r = cadd(cadd(rmul(C1, L), rmul(C2, cur.nu)), imul(c3, pc));
      if (f == 0) r.x = r.x + Cm;
      const double2 N = cadd(cur.nx, rmul(a.dthdyu[j], csub(cur.flc, flm)));
      r = cadd(r, rmul(Cg, N));
      Nu.set((size_t)(j - 1) * C + col, N);
      const double2 rhs = d2(uc.x + dt * r.x, uc.y + dt * r.y);
      const double a1 = -beta * D2U(1, j), a2 = 1.0 - beta * (D2U(2, j) + m.k2x + m.k2z), a3 = -beta * D2U(3, j);
      w2 = 1.0 / (a2 - a1 * w3);
      y = d2((rhs.x - a1 * y.x) * w2, (rhs.y - a1 * y.y) * w2);
      w3 = a3 * w2;
      U[(size_t)j * C + col] = y;
      if (f == 0) a.w3u[(size_t)j * C + col] = w3;
      um = uc;
      uc = cur.up;
      flm = cur.flc;
      if (j <= nyv) pc = cur.pn;
      if (!PF && j < nyu) nxt = load(j + 1);
      cur = nxt;
    }
    w2 = 1.0 / (1.0 - a.gw2 * w3);
    y = d2((0.0 - a.gw2 * y.x) * w2, (0.0 - a.gw2 * y.y) * w2);
    U[(size_t)(nyu + 1) * C + col] = y;
    if (f == 0) a.w3u[(size_t)(nyu + 1) * C + col] = 0.0 * w2;
  } else {
    double2 u2m = a.u2[col], u2c = a.u2[(size_t)C + col];
    double2 vvc_c = a.vvc[(size_t)C + col];
    double2 pc = a.p[col];
    double2 y2 = rmul(w2, d2(0.0, 0.0));
    double w3v = 0.0 * w2;
    a.u2[col] = y2;
    a.w3v[col] = w3v;
    struct Row { double2 u2p, pn, vvn, nu, nx; };
    auto load = [&](int j) {
      Row r;
      const bool in = j <= nyv;
      r.u2p = in ? a.u2[(size_t)(j + 1) * C + col] : d2(0.0, 0.0);
      r.pn = in ? a.p[(size_t)j * C + col] : d2(0.0, 0.0);
      r.vvn = in ? a.vvc[(size_t)(j + 1) * C + col] : d2(0.0, 0.0);
      r.nu = in ? a.Nu2[(size_t)(j - 1) * C + col] : d2(0.0, 0.0);
      r.nx = in ? a.Nxz2[(size_t)(j - 1) * C + col] : d2(0.0, 0.0);
      return r;
    };
    Row cur = load(1);
    for (int j = 1; j <= nyu; j++) {
      Row nxt;
      if (PF && j < nyu) nxt = load(j + 1);
      if (j <= nyv) {
        const double e1 = D2V(1, j), e2 = k2 + D2V(2, j), e3 = D2V(3, j);
        const double2 L2 = cadd(cadd(rmul(e1, u2m), rmul(e2, u2c)), rmul(e3, cur.u2p));
        const double C3 = -((a.dRK * 1.0) * a.dthdyv[j]);
        double2 r2 = cadd(cadd(rmul(C1, L2), rmul(C2, cur.nu)), rmul(C3, csub(cur.pn, pc)));
        const double2 N2 = cadd(cur.nx, rmul(a.dthdyv[j], csub(cur.vvn, vvc_c)));
        r2 = cadd(r2, rmul(Cg, N2));
        a.Nu2.set((size_t)(j - 1) * C + col, N2);
        const double2 rhs2 = d2(u2c.x + dt * r2.x, u2c.y + dt * r2.y);
        const double a1 = -beta * D2V(1, j), a2 = 1.0 - beta * (D2V(2, j) + m.k2x + m.k2z), a3 = -beta * D2V(3, j);
        w2 = 1.0 / (a2 - a1 * w3v);
        y2 = d2((rhs2.x - a1 * y2.x) * w2, (rhs2.y - a1 * y2.y) * w2);
        w3v = a3 * w2;
        pc = cur.pn;
        vvc_c = cur.vvn;
      } else {  // top wall row of v: (0, 1, 0), rhs 0
        w2 = 1.0 / (1.0 - 0.0 * w3v);
        y2 = d2((0.0 - 0.0 * y2.x) * w2, (0.0 - 0.0 * y2.y) * w2);
        w3v = 0.0 * w2;
      }
      a.u2[(size_t)j * C + col] = y2;
      a.w3v[(size_t)j * C + col] = w3v;
      u2m = u2c;
      u2c = cur.u2p;
      if (!PF && j < nyu) nxt = load(j + 1);
      cur = nxt;
    }
  }
}

__global__ void __launch_bounds__(256) fused_down_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.C) return;
  const int C = a.C, nyu = a.nyu;
  const Mode m = col_mode(a, col);
  const double dtirk = a.st->dti / (a.aRK + a.bRK);
  double2 U1 = a.u1[(size_t)(nyu + 1) * C + col], U3 = a.u3[(size_t)(nyu + 1) * C + col];
  double2 V = a.u2[(size_t)nyu * C + col];  // v row nyv+1
  double2 xp = d2(0.0, 0.0);
  double cprev = 0.0;
  for (int j = nyu; j >= 0; j--) {
    // velocity back substitution, u row j and v row j-1
    {
      const double2 y1 = a.u1[(size_t)j * C + col], y3 = a.u3[(size_t)j * C + col];
      const double c = a.w3u[(size_t)j * C + col];
      U1 = d2(y1.x - c * U1.x, y1.y - c * U1.y);
      U3 = d2(y3.x - c * U3.x, y3.y - c * U3.y);
      a.u1[(size_t)j * C + col] = U1;
      a.u3[(size_t)j * C + col] = U3;
    }
    if (j == 0) break;
    const double2 Vj = V;
    {
      const double2 y2 = a.u2[(size_t)(j - 1) * C + col];
      const double c = a.w3v[(size_t)(j - 1) * C + col];
      V = d2(y2.x - c * V.x, y2.y - c * V.y);
      a.u2[(size_t)(j - 1) * C + col] = V;
    }
    // divergence at centre j, psi right-hand side, top-down elimination of the pressure system
    const double2 dv = cadd(cadd(imul(m.kx, U1), imul(m.kz, U3)), rmul(1.0 * a.dthdyu[j], csub(Vj, V)));
    double2 r = rmul(dtirk, dv);
    if (col == 0 && j == 1) r = d2(0.0, 0.0);
    double a1, a2, a3;
    dg_row(a, col, m.k2x, m.k2z, j, a1, a2, a3);
    const double piv = (j == nyu) ? a2 : a2 - a3 * cprev;
    const double ip = 1.0 / piv;
    xp = (j == nyu) ? rmul(ip, r) : d2((r.x - a3 * xp.x) * ip, (r.y - a3 * xp.y) * ip);
    cprev = a1 * ip;
    a.psi[(size_t)(j - 1) * C + col] = xp;
    a.cP[(size_t)(j - 1) * C + col] = cprev;
  }
}

// The down sweep on four wavefronts of one workgroup (64 columns): waves 0, 1, 2 back-substitute u1, u3,
// u2 and hand each chunk of K rows to wave 3 through double-buffered LDS; wave 3 forms the divergence and
// eliminates the pressure system one chunk behind.  Arithmetic identical to fused_down_kernel.
template <int K>
__global__ void __launch_bounds__(256) fused_down4_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  __shared__ double2 buf[2][3][K][64];
  const int f = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = blockIdx.x * 64 + lane;
  const bool act = col < a.C;
  const int C = a.C, nyu = a.nyu;
  const int nch = (nyu + K - 1) / K;  // chunk c: rows j = nyu - c K - k, k < K, j >= 1
  double2 U = d2(0.0, 0.0);
  double2 *F = f == 0 ? a.u1 : f == 1 ? a.u3 : a.u2;
  if (act && f < 3) U = F[(size_t)(f < 2 ? nyu + 1 : nyu) * C + col];  // u rows nyu+1; v row nyv+1
  // pressure wave state
  Mode m = col_mode(a, act ? col : 0);
  const double dtirk = a.st->dti / (a.aRK + a.bRK);
  double2 xp = d2(0.0, 0.0), Vj = act ? a.u2[(size_t)nyu * C + col] : d2(0.0, 0.0);
  double cprev = 0.0;
  for (int c = 0; c <= nch; c++) {
    if (f < 3 && c < nch && act) {
      for (int k = 0; k < K; k++) {
        const int j = nyu - c * K - k;
        if (j < 1) break;
        const int row = f < 2 ? j : j - 1;  // u row j, v row j-1
        const double2 y = F[(size_t)row * C + col];
        const double cc = (f < 2 ? a.w3u : a.w3v)[(size_t)row * C + col];
        U = d2(y.x - cc * U.x, y.y - cc * U.y);
        F[(size_t)row * C + col] = U;
        buf[c & 1][f][k][lane] = U;
      }
    }
    if (f == 3 && c > 0 && act) {
      const int cc = c - 1;
      for (int k = 0; k < K; k++) {
        const int j = nyu - cc * K - k;
        if (j < 1) break;
        const double2 U1 = buf[cc & 1][0][k][lane], U3 = buf[cc & 1][1][k][lane], V = buf[cc & 1][2][k][lane];
        const double2 dv = cadd(cadd(imul(m.kx, U1), imul(m.kz, U3)), rmul(1.0 * a.dthdyu[j], csub(Vj, V)));
        double2 r = rmul(dtirk, dv);
        if (col == 0 && j == 1) r = d2(0.0, 0.0);
        double a1, a2, a3;
        dg_row(a, col, m.k2x, m.k2z, j, a1, a2, a3);
        const double piv = (j == nyu) ? a2 : a2 - a3 * cprev;
        const double ip = 1.0 / piv;
        xp = (j == nyu) ? rmul(ip, r) : d2((r.x - a3 * xp.x) * ip, (r.y - a3 * xp.y) * ip);
        cprev = a1 * ip;
        a.psi[(size_t)(j - 1) * C + col] = xp;
        a.cP[(size_t)(j - 1) * C + col] = cprev;
        Vj = V;
      }
    }
    __syncthreads();
  }
  if (f < 2 && act) {  // u row 0
    const double2 y = F[col];
    const double cc = a.w3u[col];
    U = d2(y.x - cc * U.x, y.y - cc * U.y);
    F[col] = U;
  }
}

// The final sweep on four wavefronts (64 columns): wave 0 back-substitutes the pressure correction
// and updates p; waves 1, 2, 3 apply the projection to u1, u3, u2 one chunk of K rows behind; wave 0
// forms the divergence of the projected velocities one further chunk behind.  Arithmetic identical to
// fused_final_kernel.
template <int K>
__global__ void __launch_bounds__(256) fused_final4_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  __shared__ double2 psb[2][K][64];     // ps of rows j (chunk parity)
  __shared__ double2 vb[2][3][K][64];   // projected u1(j), u3(j), v(j-1)
  __shared__ double red[4];
  const int f = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = blockIdx.x * 64 + lane;
  const bool act = col < a.C;
  const int C = a.C, nyu = a.nyu, nyv = a.nyv;
  const int nch = (nyu + K - 1) / K;  // chunk c: rows j = 1 + c K + k <= nyu
  const Mode m = col_mode(a, act ? col : 0);
  const double dtrk = a.st->dt * (a.aRK + a.bRK);
  const double kxd = dtrk * m.kx, kzd = dtrk * m.kz;
  double dmax = 0.0;
  double2 psm = d2(0.0, 0.0);                      // wave 0: previous ps; wave 3: ps(j-1)
  double2 last = d2(0.0, 0.0);                     // waves 1, 2: last projected u row
  double Q = 0.0;
  const bool qx = (col == 0) && (a.kRK == 3) && (a.flag_ctpress == 1);
  double2 u1prev = d2(0.0, 0.0), u3prev = d2(0.0, 0.0), v_m2 = d2(0.0, 0.0);  // wave 0, divergence
  for (int c = 0; c <= nch + 1; c++) {
    if (f == 0 && act) {
      if (c < nch) {  // pressure back substitution of chunk c
        for (int k = 0; k < K; k++) {
          const int j = 1 + c * K + k;
          if (j > nyu) break;
          const size_t o = (size_t)(j - 1) * C + col;
          const double2 xj = a.psi[o];
          const double2 ps = (j == 1) ? xj : csub(xj, rmul(a.cP[o], psm));
          a.p[o] = cadd(a.p[o], ps);
          psb[c & 1][k][lane] = ps;
          psm = ps;
        }
      }
      if (c >= 2) {  // divergence of chunk c-2 (centres j-1)
        const int cc = c - 2;
        for (int k = 0; k < K; k++) {
          const int j = 1 + cc * K + k;
          if (j > nyu) break;
          const double2 u1n = vb[cc & 1][0][k][lane], u3n = vb[cc & 1][1][k][lane];
          if (j >= 2) {
            const double2 v_m1 = vb[cc & 1][2][k][lane];
            const double2 dv = cadd(cadd(imul(m.kx, u1prev), imul(m.kz, u3prev)),
                                    rmul(1.0 * a.dthdyu[j - 1], csub(v_m1, v_m2)));
            dmax = fmax(dmax, hypot(dv.x, dv.y));
            v_m2 = v_m1;
          }
          u1prev = u1n;
          u3prev = u3n;
        }
      }
    }
    if (f > 0 && act && c >= 1 && c <= nch) {  // projection of chunk c-1
      const int cc = c - 1;
      for (int k = 0; k < K; k++) {
        const int j = 1 + cc * K + k;
        if (j > nyu) break;
        const double2 ps = psb[cc & 1][k][lane];
        if (f == 1 || f == 2) {
          double2 *U = f == 1 ? a.u1 : a.u3;
          const double kd = f == 1 ? kxd : kzd;
          const double2 un = csub(U[(size_t)j * C + col], imul(kd, ps));
          U[(size_t)j * C + col] = un;
          if (j == 1) U[col] = rmul(a.gbc, un);
          if (f == 1 && qx) Q = Q + (a.yv[j] - a.yv[j - 1]) * un.x;
          vb[cc & 1][f - 1][k][lane] = un;
          last = un;
        } else {
          double2 vn = d2(0.0, 0.0);
          if (j == 1) {
            a.u2[col] = d2(0.0, 0.0);
          } else {  // v row j-1 in 1..nyv
            vn = csub(a.u2[(size_t)(j - 1) * C + col], rmul(dtrk * 1.0 * a.dthdyv[j - 1], csub(ps, psm)));
            a.u2[(size_t)(j - 1) * C + col] = vn;
          }
          vb[cc & 1][2][k][lane] = vn;
          psm = ps;
        }
      }
    }
    __syncthreads();
  }
  if (act) {
    if (f == 1) {
      a.u1[(size_t)(nyu + 1) * C + col] = rmul(a.gbc, last);
      if (qx) a.st->Qx = Q;
      if (col == 0 && a.kRK == 3) a.st->iter = a.st->iter + 1.0;
    } else if (f == 2) {
      a.u3[(size_t)(nyu + 1) * C + col] = rmul(a.gbc, last);
    } else if (f == 3) {
      a.u2[(size_t)(nyv + 1) * C + col] = d2(0.0, 0.0);
    } else {
      const double2 dv = cadd(cadd(imul(m.kx, u1prev), imul(m.kz, u3prev)),
                              rmul(1.0 * a.dthdyu[nyu], csub(d2(0.0, 0.0), v_m2)));
      dmax = fmax(dmax, hypot(dv.x, dv.y));
    }
  }
  // block max over wave 0's lanes (the other waves hold 0)
  for (int off = 32; off > 0; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
  if (lane == 0) red[f] = dmax;
  __syncthreads();
  if (threadIdx.x == 0) atomic_max_nonneg(&a.st->divmax, fmax(fmax(red[0], red[1]), fmax(red[2], red[3])));
}

template <int NTH>
__global__ void __launch_bounds__(NTH) fused_final_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  __shared__ double red[NTH / 64];
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  double dmax = 0.0;
  if (col < a.C) {
    const int C = a.C, nyu = a.nyu, nyv = a.nyv;
    const Mode m = col_mode(a, col);
    const double dtrk = a.st->dt * (a.aRK + a.bRK);
    const double kxd = dtrk * m.kx, kzd = dtrk * m.kz;
    const bool qx = (col == 0) && (a.kRK == 3) && (a.flag_ctpress == 1);
    double Q = 0.0;
    double2 psm = d2(0.0, 0.0);
    double2 u1prev = d2(0.0, 0.0), u3prev = d2(0.0, 0.0), v_m1 = d2(0.0, 0.0), v_m2 = d2(0.0, 0.0);
    for (int j = 1; j <= nyu; j++) {
      const size_t o = (size_t)(j - 1) * C + col;
      const double2 xj = a.psi[o];
      const double2 ps = (j == 1) ? xj : csub(xj, rmul(a.cP[o], psm));
      a.p[o] = cadd(a.p[o], ps);
      const double2 u1n = csub(a.u1[(size_t)j * C + col], imul(kxd, ps));
      const double2 u3n = csub(a.u3[(size_t)j * C + col], imul(kzd, ps));
      a.u1[(size_t)j * C + col] = u1n;
      a.u3[(size_t)j * C + col] = u3n;
      if (j == 1) {
        a.u1[col] = rmul(a.gbc, u1n);
        a.u3[col] = rmul(a.gbc, u3n);
        a.u2[col] = d2(0.0, 0.0);
        v_m2 = d2(0.0, 0.0);
      } else {  // v row j-1 in 1..nyv
        const double2 vn = csub(a.u2[(size_t)(j - 1) * C + col], rmul(dtrk * 1.0 * a.dthdyv[j - 1], csub(ps, psm)));
        a.u2[(size_t)(j - 1) * C + col] = vn;
        v_m1 = vn;
        // divergence at centre j-1 (all its neighbours are final now)
        const double2 dv = cadd(cadd(imul(m.kx, u1prev), imul(m.kz, u3prev)),
                                rmul(1.0 * a.dthdyu[j - 1], csub(v_m1, v_m2)));
        dmax = fmax(dmax, hypot(dv.x, dv.y));
        v_m2 = v_m1;
      }
      if (qx) Q = Q + (a.yv[j] - a.yv[j - 1]) * u1n.x;
      u1prev = u1n;
      u3prev = u3n;
      psm = ps;
    }
    a.u2[(size_t)(nyv + 1) * C + col] = d2(0.0, 0.0);
    a.u1[(size_t)(nyu + 1) * C + col] = rmul(a.gbc, u1prev);
    a.u3[(size_t)(nyu + 1) * C + col] = rmul(a.gbc, u3prev);
    {
      const double2 dv = cadd(cadd(imul(m.kx, u1prev), imul(m.kz, u3prev)),
                              rmul(1.0 * a.dthdyu[nyu], csub(d2(0.0, 0.0), v_m2)));
      dmax = fmax(dmax, hypot(dv.x, dv.y));
    }
    if (qx) a.st->Qx = Q;
    if (col == 0 && a.kRK == 3) a.st->iter = a.st->iter + 1.0;
  }
  dmax = block_max<NTH>(dmax, red);
  if (threadIdx.x == 0) atomic_max_nonneg(&a.st->divmax, dmax);
}

// The final sweep, thread per column, with the inputs of the next D rows in flight while a row is
// computed (one wavefront per SIMD streams the fields, so its own loads must cover the latency).
// Arithmetic identical to fused_final_kernel.
template <int NTH, int D>
__global__ void __launch_bounds__(NTH) fused_final_pf_kernel(ColArgs a_) {
  const ColArgs a = col_view(a_, blockIdx.y);  // ensemble member
  __shared__ double red[NTH / 64];
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  double dmax = 0.0;
  if (col < a.C) {
    const int C = a.C, nyu = a.nyu, nyv = a.nyv;
    const Mode m = col_mode(a, col);
    const double dtrk = a.st->dt * (a.aRK + a.bRK);
    const double kxd = dtrk * m.kx, kzd = dtrk * m.kz;
    const bool qx = (col == 0) && (a.kRK == 3) && (a.flag_ctpress == 1);
    double Q = 0.0;
    double2 psm = d2(0.0, 0.0);
    double2 u1prev = d2(0.0, 0.0), u3prev = d2(0.0, 0.0), v_m1 = d2(0.0, 0.0), v_m2 = d2(0.0, 0.0);
    struct Row { double2 x, p, u1, u3, u2; double c; };
    auto load = [&](int j) {
      Row r;
      const size_t o = (size_t)(j - 1) * C + col;
      r.x = a.psi[o];
      r.c = a.cP[o];
      r.p = a.p[o];
      r.u1 = a.u1[(size_t)j * C + col];
      r.u3 = a.u3[(size_t)j * C + col];
      r.u2 = j >= 2 ? a.u2[(size_t)(j - 1) * C + col] : d2(0.0, 0.0);
      return r;
    };
    Row q[D + 1];
#pragma unroll
    for (int k = 0; k < D; k++) if (1 + k <= nyu) q[k] = load(1 + k);
    for (int j0 = 1; j0 <= nyu; j0 += D + 1) {
#pragma unroll
      for (int k = 0; k <= D; k++) {
        const int j = j0 + k;
        if (j > nyu) break;
        // issue the loads of row j + D into the slot freed by row j - 1
        if (j + D <= nyu) q[(k + D) % (D + 1)] = load(j + D);
        const Row &r = q[k];
        const size_t o = (size_t)(j - 1) * C + col;
        const double2 ps = (j == 1) ? r.x : csub(r.x, rmul(r.c, psm));
        a.p[o] = cadd(r.p, ps);
        const double2 u1n = csub(r.u1, imul(kxd, ps));
        const double2 u3n = csub(r.u3, imul(kzd, ps));
        a.u1[(size_t)j * C + col] = u1n;
        a.u3[(size_t)j * C + col] = u3n;
        if (j == 1) {
          a.u1[col] = rmul(a.gbc, u1n);
          a.u3[col] = rmul(a.gbc, u3n);
          a.u2[col] = d2(0.0, 0.0);
          v_m2 = d2(0.0, 0.0);
        } else {  // v row j-1 in 1..nyv
          const double2 vn = csub(r.u2, rmul(dtrk * 1.0 * a.dthdyv[j - 1], csub(ps, psm)));
          a.u2[(size_t)(j - 1) * C + col] = vn;
          v_m1 = vn;
          const double2 dv = cadd(cadd(imul(m.kx, u1prev), imul(m.kz, u3prev)),
                                  rmul(1.0 * a.dthdyu[j - 1], csub(v_m1, v_m2)));
          dmax = fmax(dmax, hypot(dv.x, dv.y));
          v_m2 = v_m1;
        }
        if (qx) Q = Q + (a.yv[j] - a.yv[j - 1]) * u1n.x;
        u1prev = u1n;
        u3prev = u3n;
        psm = ps;
      }
    }
    a.u2[(size_t)(nyv + 1) * C + col] = d2(0.0, 0.0);
    a.u1[(size_t)(nyu + 1) * C + col] = rmul(a.gbc, u1prev);
    a.u3[(size_t)(nyu + 1) * C + col] = rmul(a.gbc, u3prev);
    {
      const double2 dv = cadd(cadd(imul(m.kx, u1prev), imul(m.kz, u3prev)),
                              rmul(1.0 * a.dthdyu[nyu], csub(d2(0.0, 0.0), v_m2)));
      dmax = fmax(dmax, hypot(dv.x, dv.y));
    }
    if (qx) a.st->Qx = Q;
    if (col == 0 && a.kRK == 3) a.st->iter = a.st->iter + 1.0;
  }
  dmax = block_max<NTH>(dmax, red);
  if (threadIdx.x == 0) atomic_max_nonneg(&a.st->divmax, dmax);
}

// ================================================================================================
// layout: reference column layout [col][row] <-> device layout [row][col], 32x32 LDS tiles
__global__ void transpose_kernel(const double2 *in, double2 *out, int rows_in, int cols_in) {
  __shared__ double2 tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int y = by + r, x = bx + threadIdx.x;
    if (y < rows_in && x < cols_in) tile[r][threadIdx.x] = in[(size_t)y * cols_in + x];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int y = bx + r, x = by + threadIdx.x;  // out is cols_in x rows_in
    if (y < cols_in && x < rows_in) out[(size_t)y * rows_in + x] = tile[threadIdx.x][r];
  }
}

// modes_to_planes_UVP (FOU3D.f90:752-759) from reference-layout columns; zero padding included
__global__ void modes_to_planes_kernel(const double2 *cols, double *planes, int C, int nrows, int Kx, int Nz,
                                       int Mx, int Mz) {
  const size_t n = (size_t)(Mx + 2) * Mz;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (e >= n) return;
  const int z = (int)(e / (Mx + 2)), x = (int)(e % (Mx + 2));
  double v = 0.0;
  const int i = x / 2;
  const bool kept = (z < Nz / 2) || (z >= Mz - Nz / 2);
  if (kept && i < Kx) {
    const int k = z < Nz / 2 ? z : z - (Mz - Nz);
    const double2 c = cols[((size_t)k * Kx + i) * nrows + row];
    v = (x & 1) ? c.y : c.x;
  }
  planes[(size_t)row * n + e] = v;
}

// planes_to_modes_UVP (FOU3D.f90:1747-1753)
__global__ void planes_to_modes_kernel(const double *planes, double2 *cols, int C, int nrows, int Kx, int Nz,
                                       int Mx, int Mz) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = blockIdx.y;
  if (col >= C) return;
  const int i = col % Kx, k = col / Kx;
  const int z = k < Nz / 2 ? k : k + Mz - Nz;
  const double *pl = planes + (size_t)row * (Mx + 2) * Mz + (size_t)z * (Mx + 2);
  cols[(size_t)col * nrows + row] = d2(pl[2 * i], pl[2 * i + 1]);
}

// ================================================================================================
static inline int nblk(int n, int b) { return (n + b - 1) / b; }

void launch_rhs0(const ColArgs &a, hipStream_t st) { rhs0_kernel<<<dim3(nblk(a.C, 256), a.nm), 256, 0, st>>>(a); }
void launch_nl_col(const ColArgs &a, hipStream_t st) { nl_col_kernel<<<dim3(nblk(a.C, 256), a.nm), 256, 0, st>>>(a); }
void launch_solveU(const ColArgs &a, hipStream_t st) { solveU_kernel<<<dim3(nblk(a.C, 256), a.nm), 256, 0, st>>>(a); }
void launch_meanflow(const ColArgs &a, hipStream_t st) { meanflow_kernel<<<dim3(1, a.nm), 64, 0, st>>>(a); }
int launch_meanflow_post(const ColArgs &a, hipStream_t st) {
  if (a.nyu + 2 > MEANFLOW_MAXN) return -1;
  meanflow_post_kernel<<<dim3(1, a.nm), 256, 0, st>>>(a);
  return 0;
}
void launch_qx(const ColArgs &a, hipStream_t st) { qx_kernel<<<dim3(1, a.nm), 64, 0, st>>>(a); }
void launch_solveP(const ColArgs &a, hipStream_t st) { solveP_kernel<<<dim3(nblk(a.C, 256), a.nm), 256, 0, st>>>(a); }
void launch_vcorr(const ColArgs &a, hipStream_t st) { vcorr_kernel<256><<<dim3(nblk(a.C, 256), a.nm), 256, 0, st>>>(a); }
void launch_fused_up(const ColArgs &a, int variant, hipStream_t st) {
  if (variant == 1) fused_up3_kernel<false><<<dim3(nblk(a.C, 64), a.nm), 192, 0, st>>>(a);
  else if (variant == 2) fused_up3_kernel<true><<<dim3(nblk(a.C, 64), a.nm), 192, 0, st>>>(a);
  else fused_up_kernel<<<dim3(nblk(a.C, 256), a.nm), 256, 0, st>>>(a);
}
void launch_fused_down(const ColArgs &a, int variant, hipStream_t st) {
  if (variant == 1) fused_down4_kernel<4><<<dim3(nblk(a.C, 64), a.nm), 256, 0, st>>>(a);
  else if (variant == 2) fused_down4_kernel<8><<<dim3(nblk(a.C, 64), a.nm), 256, 0, st>>>(a);
  else fused_down_kernel<<<dim3(nblk(a.C, 256), a.nm), 256, 0, st>>>(a);
}
void launch_fused_final(const ColArgs &a, int variant, hipStream_t st) {
  if (variant == 1) fused_final_pf_kernel<256, 2><<<dim3(nblk(a.C, 256), a.nm), 256, 0, st>>>(a);
  else if (variant == 2) fused_final4_kernel<4><<<dim3(nblk(a.C, 64), a.nm), 256, 0, st>>>(a);
  else fused_final_kernel<256><<<dim3(nblk(a.C, 256), a.nm), 256, 0, st>>>(a);
}

__global__ void widen_kernel(const float2 *in, double2 *out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = d2(in[i].x, in[i].y);
}
__global__ void narrow_kernel(const double2 *in, float2 *out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float2 v;
    v.x = (float)in[i].x;
    v.y = (float)in[i].y;
    out[i] = v;
  }
}
void launch_widen(const float2 *in, double2 *out, size_t n, hipStream_t st) { widen_kernel<<<1024, 256, 0, st>>>(in, out, n); }
void launch_narrow(const double2 *in, float2 *out, size_t n, hipStream_t st) { narrow_kernel<<<1024, 256, 0, st>>>(in, out, n); }

void launch_ref_to_dev(const double2 *ref, double2 *dev, int C, int nrows, hipStream_t st) {
  // ref is [C][nrows] (rows_in = C, cols_in = nrows) -> dev [nrows][C]
  dim3 g(nblk(nrows, 32), nblk(C, 32)), b(32, 8);
  transpose_kernel<<<g, b, 0, st>>>(ref, dev, C, nrows);
}
void launch_dev_to_ref(const double2 *dev, double2 *ref, int C, int nrows, hipStream_t st) {
  dim3 g(nblk(C, 32), nblk(nrows, 32)), b(32, 8);
  transpose_kernel<<<g, b, 0, st>>>(dev, ref, nrows, C);
}
void launch_modes_to_planes(const double2 *refcols, double *planes, int C, int nrows, int Kx, int Nz, int Mx,
                            int Mz, hipStream_t st) {
  dim3 g(nblk((Mx + 2) * Mz, 256), nrows);
  modes_to_planes_kernel<<<g, 256, 0, st>>>(refcols, planes, C, nrows, Kx, Nz, Mx, Mz);
}
void launch_planes_to_modes(const double *planes, double2 *refcols, int C, int nrows, int Kx, int Nz, int Mx,
                            int Mz, hipStream_t st) {
  dim3 g(nblk(C, 256), nrows);
  planes_to_modes_kernel<<<g, 256, 0, st>>>(planes, refcols, C, nrows, Kx, Nz, Mx, Mz);
}

}  // namespace dnsk
