// One-pass column stage: the whole column part of an RK3 substep for a block of NC adjacent columns in
// one kernel, with the columns held on chip (registers + LDS), so that every field is read once and
// written once per substep.
//
// The same substep as the three sweeps of column_kernels.hip (fused_up + fused_down + fused_final):
//   RHS0 + laplacian          littleharsh.f90:281-368, 682-856
//   nonlinear column part     FOU3D.f90:268-303, 359-368 (der_yu_h / der_yv_h, du -= g N, CFL dt at kRK=1)
//   solveU x3                 littleharsh.f90:863-894, tridLU_3D.f90:38-221
//   solveP                    littleharsh.f90:626-680, tridLU_3D.f90:98-132, 224-297
//   v_corr (+ max|div|)       littleharsh.f90:896-993
//   meanflow_ctP (kRK = 3)    littleharsh.f90:572-589
// With a constant flow rate (flag_ctpress = 0) the correction of meanflow_ctU follows this kernel in
// meanflow_post_kernel (column_kernels.hip): it changes only mode (0,0) of u1, which the pressure solve
// and v_corr do not read, so applying it after them gives the reference's result.
//
// Layout of the work.  Rows 1..P*M of every column are split into P segments of M rows; thread
// (c, s) = (tid % NC, tid / NC) owns rows 1 + s*M .. (s+1)*M of column c, so NC consecutive threads
// read NC consecutive columns (one 128-byte line of a row for NC = 8).  The ghost rows (u/w at 0 and
// nyu+1) and the walls of v (0 and nyv+1) are eliminated from the systems exactly as the reference's
// forward elimination does for its first row (pivot b_1 - a_1*gw(1)); rows above nyu are identity
// rows.  Each of the three tridiagonal systems (u1 and u3 share one matrix; u2; the pressure) is solved
// by the partition method:
//   1. each thread eliminates its segment so that every interior row reads
//        x_r = d_r - a_r x_first - c_r x_last,
//      leaving two equations per segment in the segment's first and last unknowns;
//   2. the last unknown of every segment is eliminated with the neighbouring segment's last equation
//      (one exchange), which leaves a tridiagonal system in the P first unknowns of a column, one
//      equation per thread;
//   3. that system is solved by parallel cyclic reduction over the column's P threads (log2 P steps
//      through double-buffered LDS, one barrier each);
//   4. every thread recovers its last unknown (one more exchange) and substitutes both into its rows.
// All three systems are diagonally dominant (the pressure matrix of mode (0,0) weakly, with its pin
// row), so the segment eliminations and the reduced solve need no pivoting, as the reference's LU.
#pragma once
#include "column_common.h"

namespace dnsk {

// Segment of M rows with NR right-hand sides: a[r] x_{r-1} + b[r] x_r + c[r] x_{r+1} = d[k][r].
template <int M, int NR>
struct Segment {
  double a[M], c[M];
  double2 d[NR][M];
  double a0, b0, c0;

  // Row r (rows arrive in order).  Rows 1..M-1 are eliminated downwards against x_first:
  //   a[r] x_first + x_r + c[r] x_{r+1} = d[r].
  __device__ __forceinline__ void row(int r, double ar, double br, double cr, const double2 *dr) {
    if (r == 0) {
      a0 = ar; b0 = br; c0 = cr;
#pragma unroll
      for (int k = 0; k < NR; k++) d[k][0] = dr[k];
    } else if (r == 1) {
      const double ip = 1.0 / br;
      a[1] = ar * ip;
      c[1] = cr * ip;
#pragma unroll
      for (int k = 0; k < NR; k++) d[k][1] = d2(dr[k].x * ip, dr[k].y * ip);
    } else {
      const double ip = 1.0 / (br - ar * c[r - 1]);
      a[r] = -(ar * a[r - 1]) * ip;
      c[r] = cr * ip;
#pragma unroll
      for (int k = 0; k < NR; k++)
        d[k][r] = d2((dr[k].x - ar * d[k][r - 1].x) * ip, (dr[k].y - ar * d[k][r - 1].y) * ip);
    }
  }

  // Upward elimination of rows M-3..1 against x_last, then the two reduced equations:
  //   first: A x_prev + B x_first + C x_last = D        last: aL x_first + x_last + cL x_next = dL
  __device__ __forceinline__ void reduce(double &A, double &B, double &C, double2 *D, double &aL, double &cL,
                                         double2 *dL) {
#pragma unroll
    for (int r = M - 3; r >= 1; r--) {
      a[r] = a[r] - c[r] * a[r + 1];
#pragma unroll
      for (int k = 0; k < NR; k++)
        d[k][r] = d2(d[k][r].x - c[r] * d[k][r + 1].x, d[k][r].y - c[r] * d[k][r + 1].y);
      c[r] = -(c[r] * c[r + 1]);
    }
    A = a0;
    B = b0 - c0 * a[1];
    C = -(c0 * c[1]);
#pragma unroll
    for (int k = 0; k < NR; k++) D[k] = d2(d[k][0].x - c0 * d[k][1].x, d[k][0].y - c0 * d[k][1].y);
    aL = a[M - 1];
    cL = c[M - 1];
#pragma unroll
    for (int k = 0; k < NR; k++) dL[k] = d[k][M - 1];
  }

  // Back substitution of the segment's first and last unknowns; d[k][r] = x_r afterwards.
  __device__ __forceinline__ void expand(const double2 *xf, const double2 *xl) {
#pragma unroll
    for (int k = 0; k < NR; k++) {
#pragma unroll
      for (int r = 1; r <= M - 2; r++)
        d[k][r] = d2(d[k][r].x - a[r] * xf[k].x - c[r] * xl[k].x, d[k][r].y - a[r] * xf[k].y - c[r] * xl[k].y);
      d[k][0] = xf[k];
      d[k][M - 1] = xl[k];
    }
  }
};

// One equation of a reduced system: a x_{s-1} + b x_s + c x_{s+1} = d[k], ib = 1/b.
template <int NR>
struct Eq {
  double a, b, c, ib;
  double2 d[NR];
};

template <int NR>
__device__ __forceinline__ Eq<NR> eq_zero() {
  Eq<NR> e;
  e.a = 0.0; e.b = 1.0; e.c = 0.0; e.ib = 1.0;
#pragma unroll
  for (int k = 0; k < NR; k++) e.d[k] = d2(0.0, 0.0);
  return e;
}

// N equations in LDS as separate arrays (consecutive threads touch consecutive words: no bank
// conflicts); b is not stored, the neighbours' steps use ib.
template <int NR, int N>
struct EqBuf {
  double a[N], c[N], ib[N];
  double2 d[NR][N];
  __device__ __forceinline__ Eq<NR> get(int i) const {
    Eq<NR> e;
    e.a = a[i]; e.c = c[i]; e.ib = ib[i]; e.b = 0.0;
#pragma unroll
    for (int k = 0; k < NR; k++) e.d[k] = d[k][i];
    return e;
  }
  __device__ __forceinline__ void put(int i, const Eq<NR> &e) {
    a[i] = e.a; c[i] = e.c; ib[i] = e.ib;
#pragma unroll
    for (int k = 0; k < NR; k++) d[k][i] = e.d[k];
  }
};

// Partition solve of one tridiagonal system across the P threads of a column (thread s owns segment s;
// index s*NC + c in the LDS buffers U[0], U[1]).  The calls are made by every thread of the workgroup
// in the order first(), form(), pcr() x log2(P), expand(), with a barrier after each call.
template <int NC, int P, int M, int NR>
struct Part {
  using Buf = EqBuf<NR, NC * P>;
  Segment<M, NR> sg;
  Eq<NR> F, L, e;
  int idx, s;

  __device__ __forceinline__ void first(Buf *U, int c, int s_) {
    s = s_;
    idx = s * NC + c;
    sg.reduce(F.a, F.b, F.c, F.d, L.a, L.c, L.d);
    L.b = 1.0;
    L.ib = 1.0;
    U[1].put(idx, L);
  }
  // F_s with x_last(s-1) and x_last(s) eliminated by L_{s-1} and L_s:
  //   -F.a Lp.a x_F(s-1) + (F.b - F.a Lp.c - F.c L.a) x_F(s) - F.c L.c x_F(s+1) = F.d - F.a Lp.d - F.c L.d
  __device__ __forceinline__ void form(Buf *U) {
    const Eq<NR> Lp = s > 0 ? U[1].get(idx - NC) : eq_zero<NR>();
    e.a = -(F.a * Lp.a);
    e.b = F.b - F.a * Lp.c - F.c * L.a;
    e.c = -(F.c * L.c);
#pragma unroll
    for (int k = 0; k < NR; k++)
      e.d[k] = d2(F.d[k].x - F.a * Lp.d[k].x - F.c * L.d[k].x, F.d[k].y - F.a * Lp.d[k].y - F.c * L.d[k].y);
    e.ib = 1.0 / e.b;
    U[0].put(idx, e);
  }
  // cyclic reduction step of span h: reads buffer cur, writes buffer cur ^ 1 (the last step writes the
  // solution x_F = d/b instead)
  __device__ __forceinline__ void pcr(Buf *U, int h, int cur) {
    const bool hl = s - h >= 0, hh = s + h < P;
    const Eq<NR> lo = hl ? U[cur].get(idx - h * NC) : eq_zero<NR>();
    const Eq<NR> hi = hh ? U[cur].get(idx + h * NC) : eq_zero<NR>();
    const double al = -(e.a * lo.ib), ga = -(e.c * hi.ib);
    Eq<NR> n;
    n.a = al * lo.a;
    n.c = ga * hi.c;
    n.b = e.b + al * lo.c + ga * hi.a;
    n.ib = 1.0 / n.b;
#pragma unroll
    for (int k = 0; k < NR; k++)
      n.d[k] = d2(e.d[k].x + al * lo.d[k].x + ga * hi.d[k].x, e.d[k].y + al * lo.d[k].y + ga * hi.d[k].y);
    e = n;
    if (2 * h >= P) {
#pragma unroll
      for (int k = 0; k < NR; k++) e.d[k] = d2(e.d[k].x * e.ib, e.d[k].y * e.ib);
    }
    U[cur ^ 1].put(idx, e);
  }
  // x_L(s) = L.d - L.a x_F(s) - L.c x_F(s+1), then the rows of the segment
  __device__ __forceinline__ void expand(Buf *U, int cur) {
    double2 xf[NR], xl[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const double2 nx = s + 1 < P ? U[cur].d[k][idx + NC] : d2(0.0, 0.0);
      xf[k] = e.d[k];
      xl[k] = d2(L.d[k].x - L.a * xf[k].x - L.c * nx.x, L.d[k].y - L.a * xf[k].y - L.c * nx.y);
    }
    sg.expand(xf, xl);
  }
};

template <int NC, int P>
struct PassShared {
  EqBuf<2, NC * P> U[2];                 // u1/u3 system
  EqBuf<1, NC * P> V[2];                 // u2 system, then the pressure
  double2 x0[P * NC], x1[P * NC];        // neighbour exchanges between segments
  double red[NC * P / 64 > 0 ? NC * P / 64 : 1];
};

// Copy of x the optimiser cannot see through: index arithmetic of a later phase is recomputed there
// instead of being kept live (and spilled) from an earlier one.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Nonlinear-term / flux array in its storage precision T (double, or float in mixed precision); the
// type is a template parameter so that no load sits behind a run-time branch.
template <typename T>
struct NA {
  void *p;
  __device__ __forceinline__ double2 operator[](size_t i) const { return to_d2(static_cast<const cx<T> *>(p)[i]); }
  __device__ __forceinline__ void set(size_t i, double2 v) const { static_cast<cx<T> *>(p)[i] = from_d2<T>(v); }
};

// The kernel is a body (LDS layout + run(), block given by the caller) and a thin __global__ wrapper; the
// ensemble's grouped launch (ensemble.hip) runs bodies of several row counts in one launch.
template <int NC, int P, int M>
struct PassLds {
  // The segment's p rows from the first phase, kept in LDS for p += psi at the end instead of being read
  // from HBM again: with 64 segments (one workgroup per CU) where LDS has room; with fewer segments the
  // stash made the compiler spill.
  static constexpr bool PST = P >= 64 && sizeof(PassShared<NC, P>) + sizeof(double) * P * M +
                                             sizeof(double2) * M * NC * P <= 160 * 1024;
  PassShared<NC, P> sh;
  double u1q[P * M];  // column 0's u1 for the flow rate
  double2 pst[PST ? M * NC * P : 1];
};

template <int NC, int P, int M, typename NT>
struct ColumnPassBody {
  static constexpr int NTH = NC * P;
  static constexpr size_t LDS = sizeof(PassLds<NC, P, M>);
  // b: column block of member my (nblk column blocks per member); tid: thread index in the body's NTH
  // threads (a grouped launch runs several bodies side by side in one workgroup, all of them through the
  // same barriers); active = false: a padding body that computes and stores nothing
  __device__ __forceinline__ static void run(const ColArgs &a_, int b, int nblk, int my, int tid,
                                             unsigned char *lds, bool active = true);
};

template <int NC, int P, int M, typename NT>
__device__ __forceinline__ void ColumnPassBody<NC, P, M, NT>::run(const ColArgs &a_, int b, int nblk, int my,
                                                                  int tid, unsigned char *lds, bool active) {
  PassLds<NC, P, M> &L = *reinterpret_cast<PassLds<NC, P, M> *>(lds);
  PassShared<NC, P> &sh = L.sh;
  double *u1q = L.u1q;
  double2 *pst = L.pst;
  constexpr bool PST = PassLds<NC, P, M>::PST;
  const ColArgs a = col_view(a_, my);  // ensemble member
  const NA<NT> Nu1{a.Nu1.p}, Nu2{a.Nu2.p}, Nu3{a.Nu3.p}, Nxz1{a.Nxz1.p}, Nxz2{a.Nxz2.p}, Nxz3{a.Nxz3.p};
  const NA<NT> uvf{a.uvf.p}, vvc{a.vvc.p}, wvf{a.wvf.p};
  const int c = tid % NC, s = tid / NC;
  const int C = a.C, nyu = a.nyu, nyv = a.nyv;
  // Workgroups are dealt to the 8 XCDs in turn; give each XCD a contiguous run of column blocks, so
  // that blocks sharing a 128-byte line of a row (NC = 4) are read through the same L2.
  const int cb = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
  const int col0 = cb * NC + c;
  const bool live = active && col0 < C;
  const int col = live ? col0 : C - 1;  // threads past the last column redo its work and store nothing
  const Mode m = col_mode(a, col);
  const double dt = substep_dt(a, (live && s == 0) ? col : -1);
  const double dti = a.kRK == 1 ? 1.0 / dt : a.st->dti;  // the kRK = 1 value is being written by column 0
  const double C1 = a.aRK / a.Re, C2 = -a.cRK, Cg = -a.gRK;
  const double beta = a.bRK * dt / a.Re;
  const double c3x = -(a.dRK * m.kx), c3z = -(a.dRK * m.kz);
  const double Cm = col == 0 ? -(a.dRK * a.st->mpgx) : 0.0;
  const double k2 = m.k2x + m.k2z;
  const int j0 = 1 + s * M;
  auto at = [&](int j) { return (size_t)j * C + col; };
  // row clamps: every load is issued unconditionally and in bounds; rows outside a grid's range are
  // masked in the arithmetic
  auto cu = [&](int j) { return j < 0 ? 0 : j > nyu + 1 ? nyu + 1 : j; };
  auto cv = [&](int j) { return j < 0 ? 0 : j > nyv + 1 ? nyv + 1 : j; };
  auto cp = [&](int j) { return j < 1 ? 1 : j > nyu ? nyu : j; };
  auto cq = [&](int j) { return j < 1 ? 1 : j > nyv ? nyv : j; };
  const double2 z2 = d2(0.0, 0.0);

  // ---- RHS0 + nonlinear column terms + du; segment elimination of the velocity systems ----------
  Part<NC, P, M, 2> pu;  // u1, u3
  Part<NC, P, M, 1> pv;  // u2
  Segment<M, 2> &su = pu.sg;
  Segment<M, 1> &sv = pv.sg;
  {
    // Both grids' rows in one loop: every load is in bounds (clamped rows are masked in the
    // arithmetic), so a row's loads of both grids are issued together.
    double2 u1m = a.u1[at(cu(j0 - 1))], u1c = a.u1[at(cu(j0))];
    double2 u3m = a.u3[at(cu(j0 - 1))], u3c = a.u3[at(cu(j0))];
    double2 uvm = uvf[at(cv(j0 - 1))], wvm = wvf[at(cv(j0 - 1))];
    double2 u2m = a.u2[at(cv(j0 - 1))], u2c = a.u2[at(cv(j0))];
    double2 vvc_c = vvc[at(cu(j0))];
    double2 pc = a.p[at(cp(j0) - 1)];
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int j = j0 + r;
      const bool ur = j <= nyu, vr = j <= nyv;
      const int ju = cp(j), jv = cq(j);
      if constexpr (PST) pst[r * NTH + tid] = pc;  // p row cp(j)
      // u-grid row j
      const double2 u1p = a.u1[at(cu(j + 1))], u3p = a.u3[at(cu(j + 1))];
      const double2 uvc = uvf[at(cv(j))], wvc = wvf[at(cv(j))];
      const size_t oq = at(ju - 1);
      const double2 nu1 = Nu1[oq], nu3 = Nu3[oq], nx1 = Nxz1[oq], nx3 = Nxz3[oq];
      const double g1 = D2U(1, ju), g2 = D2U(2, ju), g3 = D2U(3, ju), hu = a.dthdyu[ju];
      double2 N1, N2, N3;
      {
        const double dd2 = k2 + g2;
        const double2 L1 = cadd(cadd(rmul(g1, u1m), rmul(dd2, u1c)), rmul(g3, u1p));
        const double2 L3 = cadd(cadd(rmul(g1, u3m), rmul(dd2, u3c)), rmul(g3, u3p));
        double2 r1 = cadd(cadd(rmul(C1, L1), rmul(C2, nu1)), imul(c3x, pc));
        double2 r3 = cadd(cadd(rmul(C1, L3), rmul(C2, nu3)), imul(c3z, pc));
        r1.x = r1.x + Cm;
        N1 = cadd(nx1, rmul(hu, csub(uvc, uvm)));
        N3 = cadd(nx3, rmul(hu, csub(wvc, wvm)));
        r1 = cadd(r1, rmul(Cg, N1));
        r3 = cadd(r3, rmul(Cg, N3));
        double au = -beta * g1, bu = 1.0 - beta * (g2 + m.k2x + m.k2z), cc = -beta * g3;
        // ghost rows folded in: u_0 = -gw(1) u_1 into row 1, u_{nyu+1} = -gw(2) u_nyu into row nyu
        const double bf = bu - au * a.gw1;
        bu = j == 1 ? bf : bu;
        au = j == 1 ? 0.0 : au;
        const double bl = bu - cc * a.gw2;
        bu = j == nyu ? bl : bu;
        cc = j == nyu ? 0.0 : cc;
        const double2 rr[2] = {ur ? d2(u1c.x + dt * r1.x, u1c.y + dt * r1.y) : z2,
                               ur ? d2(u3c.x + dt * r3.x, u3c.y + dt * r3.y) : z2};
        su.row(r, ur ? au : 0.0, ur ? bu : 1.0, ur ? cc : 0.0, rr);
      }
      u1m = u1c; u1c = u1p;
      u3m = u3c; u3c = u3p;
      uvm = uvc; wvm = wvc;
      // v-grid row j
      const double2 u2p = a.u2[at(cv(j + 1))];
      const double2 pn = a.p[at(cp(j + 1) - 1)];
      const double2 vvn = vvc[at(cu(j + 1))];
      const size_t ov = at(jv - 1);
      const double2 nu2 = Nu2[ov], nx2 = Nxz2[ov];
      const double e1 = D2V(1, jv), e2 = D2V(2, jv), e3 = D2V(3, jv), hv = a.dthdyv[jv];
      {
        const double2 L2 = cadd(cadd(rmul(e1, u2m), rmul(k2 + e2, u2c)), rmul(e3, u2p));
        const double C3 = -((a.dRK * 1.0) * hv);
        double2 r2 = cadd(cadd(rmul(C1, L2), rmul(C2, nu2)), rmul(C3, csub(pn, pc)));
        N2 = cadd(nx2, rmul(hv, csub(vvn, vvc_c)));
        r2 = cadd(r2, rmul(Cg, N2));
        const double av = j == 1 ? 0.0 : -beta * e1;    // wall v_0 = 0
        const double bv = 1.0 - beta * (e2 + m.k2x + m.k2z);
        const double cw = j == nyv ? 0.0 : -beta * e3;  // wall v_{nyv+1} = 0
        const double2 rv = vr ? d2(u2c.x + dt * r2.x, u2c.y + dt * r2.y) : z2;
        sv.row(r, vr ? av : 0.0, vr ? bv : 1.0, vr ? cw : 0.0, &rv);
      }
      // stores after both grids' loads of the row (a store behind a branch ends the batch of loads)
      if (live && ur) {
        Nu1.set(oq, N1);
        Nu3.set(oq, N3);
      }
      if (live && vr) Nu2.set(ov, N2);
      pc = pn;
      vvc_c = vvn;
      u2m = u2c; u2c = u2p;
    }
  }
  pu.first(sh.U, c, s);
  pv.first(sh.V, c, s);
  __syncthreads();
  pu.form(sh.U);
  pv.form(sh.V);
  __syncthreads();
  {
    int cur = 0;
#pragma unroll
    for (int h = 1; h < P; h *= 2, cur ^= 1) {
      pu.pcr(sh.U, h, cur);
      pv.pcr(sh.V, h, cur);
      __syncthreads();
    }
    pu.expand(sh.U, cur);
    pv.expand(sh.V, cur);
  }
  // su.d[0][r], su.d[1][r], sv.d[0][r]: u1, u3 (row j0+r), u2 (row j0+r) after solveU
  sh.x0[s * NC + c] = sv.d[0][M - 1];
  __syncthreads();

  // ---- divergence, psi right-hand side, segment elimination of the pressure system -------------
  // rows of LU_buildP (tridLU_3D.f90:242-280) in the operand order of dg_row
  const double dtirk = dti / (a.aRK + a.bRK);
  Part<NC, P, M, 1> pp;
  Segment<M, 1> &sp = pp.sg;
  {
    double2 vm = s == 0 ? z2 : sh.x0[(s - 1) * NC + c];
    const int jb = opaque(j0);
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int j = jb + r;
      const int jp = cp(j);
      const double hu = a.dthdyu[jp], hvl = a.dthdyv[jp - 1], hvu = a.dthdyv[jp < nyu ? jp : nyu - 1];
      const bool pr = j <= nyu;
      const double2 dv = cadd(cadd(imul(m.kx, su.d[0][r]), imul(m.kz, su.d[1][r])),
                              rmul(1.0 * hu, csub(sv.d[0][r], vm)));
      double2 rp = rmul(dtirk, dv);
      if (!pr || (col == 0 && j == 1)) rp = z2;
      const double lower = (-hu) * hvl, upper = hu * (-hvu);
      double a1 = j == 1 ? 0.0 : (-hu) * (-hvl);
      double a2 = (j == 1 ? upper : j == nyu ? lower : lower + upper) + m.k2x + m.k2z;
      double a3 = j == nyu ? 0.0 : hu * hvu;
      if (col == 0 && j == 1) {  // pressure pin of mode (0,0)
        a1 = 0.0;
        a2 = 1.0 / (a.dyp * a.dyp);
        a3 = 0.0;
      }
      sp.row(r, pr ? a1 : 0.0, pr ? a2 : 1.0, pr ? a3 : 0.0, &rp);
      vm = sv.d[0][r];
    }
  }
  pp.first(sh.V, c, s);  // the u2 system's buffers are free again (read before the last barrier)
  __syncthreads();
  pp.form(sh.V);
  __syncthreads();
  {
    int cur = 0;
#pragma unroll
    for (int h = 1; h < P; h *= 2, cur ^= 1) {
      pp.pcr(sh.V, h, cur);
      __syncthreads();
    }
    pp.expand(sh.V, cur);
  }
  sh.x1[s * NC + c] = sp.d[0][0];
  __syncthreads();

  // ---- p += psi, v_corr, wall rows; corrected values stay in su / sv for max|div| --------------
  const double dtrk = dt * (a.aRK + a.bRK);
  const double kxd = dtrk * m.kx, kzd = dtrk * m.kz;
  const bool qx = active && cb == 0 && a.kRK == 3 && a.flag_ctpress == 1;
  {
    const double2 psn_top = s == P - 1 ? z2 : sh.x1[(s + 1) * NC + c];
    const int jb = opaque(j0), cl = opaque(col);
    auto at = [&](int j) { return (size_t)j * C + cl; };
    double2 pold[M];
#pragma unroll
    for (int r = 0; r < M; r++) {
      if constexpr (PST) pold[r] = pst[r * NTH + tid];
      else pold[r] = a.p[at(cp(jb + r) - 1)];
    }
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int j = jb + r;
      const double hv = a.dthdyv[cq(j)];
      const double2 ps = sp.d[0][r];
      const double2 u1n = csub(su.d[0][r], imul(kxd, ps));
      const double2 u3n = csub(su.d[1][r], imul(kzd, ps));
      const double2 psn = r == M - 1 ? psn_top : sp.d[0][r + 1];
      const double2 vn = j <= nyv ? csub(sv.d[0][r], rmul(dtrk * 1.0 * hv, csub(psn, ps))) : z2;
      su.d[0][r] = u1n;
      su.d[1][r] = u3n;
      sv.d[0][r] = vn;
      if (j <= nyu) {
        if (live) {
          a.p[at(j - 1)] = cadd(pold[r], ps);
          a.u1[at(j)] = u1n;
          a.u3[at(j)] = u3n;
          a.u2[at(j)] = vn;  // row nyv+1 = nyu is the wall: 0
          if (j == 1) {
            a.u1[at(0)] = rmul(a.gbc, u1n);
            a.u3[at(0)] = rmul(a.gbc, u3n);
            a.u2[at(0)] = z2;
          }
          if (j == nyu) {
            a.u1[at(nyu + 1)] = rmul(a.gbc, u1n);
            a.u3[at(nyu + 1)] = rmul(a.gbc, u3n);
          }
        }
        if (qx && c == 0) u1q[j - 1] = u1n.x;
      }
    }
  }
  sh.x0[s * NC + c] = sv.d[0][M - 1];
  __syncthreads();
  double dmax = 0.0;
  {
    double2 vm = s == 0 ? z2 : sh.x0[(s - 1) * NC + c];
    const int jb = opaque(j0);
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int j = jb + r;
      const double hu = a.dthdyu[cp(j)];
      const double2 dv = cadd(cadd(imul(m.kx, su.d[0][r]), imul(m.kz, su.d[1][r])),
                              rmul(1.0 * hu, csub(sv.d[0][r], vm)));
      if (j <= nyu) dmax = fmax(dmax, hypot(dv.x, dv.y));
      vm = sv.d[0][r];
    }
  }
  if (!live) dmax = 0.0;
  dmax = block_max_sub<NTH>(dmax, sh.red, tid);
  if (tid == 0) {
    atomic_max_nonneg(&a.st->divmax, dmax);
    if (qx) {  // flowrateIm of mode (0,0) in the reference's summation order (littleharsh.f90:473-492)
      double Q = 0.0;
      for (int j = 1; j <= nyv + 1; j++) Q = Q + (a.yv[j] - a.yv[j - 1]) * u1q[j - 1];
      a.st->Qx = Q;
    }
    if (active && cb == 0 && a.kRK == 3) a.st->iter = a.st->iter + 1.0;
  }
}

}  // namespace dnsk
