// One-pass column stage with the three tridiagonal systems eliminated one after another (u1/u3, then
// u2, then the pressure), for long row segments.  The same substep as ColumnPassBody (column_pass.h) and
// the same per-row arithmetic (the same reference lines), with a different schedule of the row loop:
//   phase U : the u-grid rows of the segment (RHS0 + N1, N3 + du for u1, u3), eliminated into the u1/u3
//             segment and solved (with 8-row segments one field after the other);
//   phase V : the v-grid rows (RHS0 + N2 + du for u2), eliminated into the u2 segment; its solve;
//   phase P : divergence, psi right-hand side, the pressure system; then p += psi, v_corr, wall rows.
// Only one system's segment state is live at a time, so a thread can hold M = 8 rows (a 256-thread
// workgroup of 8 columns x 32 segments covers 256 rows) within 256 registers, and the workgroup needs
// about 70 KB of LDS: two workgroups share a CU, and one's solves (LDS and barriers, no HBM traffic) run
// while the other streams its rows.  The p rows of the segment are read once (phase U) and kept in LDS
// for phase V and for p += psi.
#pragma once
#include "column_pass.h"

namespace dnsk {

template <int NC, int P, int M>
struct SeqLds {
  static constexpr int NTH = NC * P;
  EqBuf<2, NTH> U[2];      // the system being solved (u1/u3 with two right-hand sides; u2 and the
                           // pressure use the same space as EqBuf<1, NTH>)
  double2 x0[NTH], x1[NTH];
  double red[NTH / 64 > 0 ? NTH / 64 : 1];
  double u1q[P * M];       // column 0's u1 for the flow rate
  double2 pst[M * NTH];    // the segment's p rows
  double yc[4][P * M];     // y coefficients of rows j = 1..P M (index j - 1): dthdyv(cq j), dthdyu(cp j),
                           // dthdyv(cp j - 1), dthdyv(min(cp j, nyu - 1))
};

template <int NC, int P, int M, typename NT>
struct ColumnSeqBody {
  static constexpr int NTH = NC * P;
  static constexpr bool SPLITU = M >= 8;  // phase U one field at a time (see run)
  static constexpr size_t LDS = sizeof(SeqLds<NC, P, M>);
  __device__ __forceinline__ static void run(const ColArgs &a_, int b, int nblk, int my, int tid,
                                             unsigned char *lds, bool active = true);
};

// Partition solve of one system (Part's calls, each followed by a barrier); buffers U[0], U[1].
template <int NC, int P, int M, int NR>
__device__ __forceinline__ void seq_solve(Part<NC, P, M, NR> &pt, EqBuf<NR, NC * P> *U, int c, int s) {
  pt.first(U, c, s);
  __syncthreads();
  pt.form(U);
  __syncthreads();
  int cur = 0;
#pragma unroll
  for (int h = 1; h < P; h *= 2, cur ^= 1) {
    pt.pcr(U, h, cur);
    __syncthreads();
  }
  pt.expand(U, cur);
  __syncthreads();  // the buffers are read by expand; the next system may overwrite them after this
}

template <int NC, int P, int M, typename NT>
__device__ __forceinline__ void ColumnSeqBody<NC, P, M, NT>::run(const ColArgs &a_, int b, int nblk, int my,
                                                                 int tid, unsigned char *lds, bool active) {
  SeqLds<NC, P, M> &L = *reinterpret_cast<SeqLds<NC, P, M> *>(lds);
  EqBuf<2, NTH> *U2 = L.U;
  EqBuf<1, NTH> *U1 = reinterpret_cast<EqBuf<1, NTH> *>(L.U);
  static_assert(2 * sizeof(EqBuf<1, NTH>) <= sizeof(L.U), "one-rhs systems share the two-rhs buffers");
  double *u1q = L.u1q;
  double2 *pst = L.pst;
  const ColArgs a = col_view(a_, my);  // ensemble member
  const NA<NT> Nu1{a.Nu1.p}, Nu2{a.Nu2.p}, Nu3{a.Nu3.p}, Nxz1{a.Nxz1.p}, Nxz2{a.Nxz2.p}, Nxz3{a.Nxz3.p};
  const NA<NT> uvf{a.uvf.p}, vvc{a.vvc.p}, wvf{a.wvf.p};
  const int c = tid % NC, s = tid / NC;
  const int C = a.C, nyu = a.nyu, nyv = a.nyv;
  const int cb = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;  // XCD-aware, as ColumnPassBody
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
  auto cu = [&](int j) { return j < 0 ? 0 : j > nyu + 1 ? nyu + 1 : j; };
  auto cv = [&](int j) { return j < 0 ? 0 : j > nyv + 1 ? nyv + 1 : j; };
  auto cp = [&](int j) { return j < 1 ? 1 : j > nyu ? nyu : j; };
  auto cq = [&](int j) { return j < 1 ? 1 : j > nyv ? nyv : j; };
  const double2 z2 = d2(0.0, 0.0);
  // The y coefficients that phases V and P and the final row loop read, copied to LDS before the first
  // barrier (in phase U's solve): read from global memory in the final loop, each row's read waited
  // behind the previous row's stores.
  for (int i = tid; i < P * M; i += NTH) {
    const int j = 1 + i, jp = cp(j);
    L.yc[0][i] = a.dthdyv[cq(j)];
    L.yc[1][i] = a.dthdyu[jp];
    L.yc[2][i] = a.dthdyv[jp - 1];
    L.yc[3][i] = a.dthdyv[jp < nyu ? jp : nyu - 1];
  }

  // ---- phase U: u-grid rows.  With SPLITU (8-row segments) one field's segment at a time, u1 then u3
  // (the same matrix: eliminating one right-hand side at a time gives the same bits as both at once, and
  // keeps an 8-row segment within the registers); otherwise u1 and u3 together ----------------------
  double2 xu[2][M];  // u1, u3 of the segment's rows after solveU
  if constexpr (!SPLITU) {
    Part<NC, P, M, 2> pu;
    Segment<M, 2> &su2 = pu.sg;
    auto at = [&](int j) { return (size_t)j * C + col; };
    double2 u1m = a.u1[at(cu(j0 - 1))], u1c = a.u1[at(cu(j0))];
    double2 u3m = a.u3[at(cu(j0 - 1))], u3c = a.u3[at(cu(j0))];
    double2 uvm = uvf[at(cv(j0 - 1))], wvm = wvf[at(cv(j0 - 1))];
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int j = j0 + r;
      const bool ur = j <= nyu;
      const int ju = cp(j);
      const double2 pc = a.p[at(ju - 1)];
      pst[r * NTH + tid] = pc;  // p row cp(j)
      const double2 u1p = a.u1[at(cu(j + 1))], u3p = a.u3[at(cu(j + 1))];
      const double2 uvc = uvf[at(cv(j))], wvc = wvf[at(cv(j))];
      const size_t oq = at(ju - 1);
      const double2 nu1 = Nu1[oq], nu3 = Nu3[oq], nx1 = Nxz1[oq], nx3 = Nxz3[oq];
      const double g1 = D2U(1, ju), g2 = D2U(2, ju), g3 = D2U(3, ju), hu = a.dthdyu[ju];
      const double dd2 = k2 + g2;
      const double2 L1 = cadd(cadd(rmul(g1, u1m), rmul(dd2, u1c)), rmul(g3, u1p));
      const double2 L3 = cadd(cadd(rmul(g1, u3m), rmul(dd2, u3c)), rmul(g3, u3p));
      double2 r1 = cadd(cadd(rmul(C1, L1), rmul(C2, nu1)), imul(c3x, pc));
      double2 r3 = cadd(cadd(rmul(C1, L3), rmul(C2, nu3)), imul(c3z, pc));
      r1.x = r1.x + Cm;
      const double2 N1 = cadd(nx1, rmul(hu, csub(uvc, uvm)));
      const double2 N3 = cadd(nx3, rmul(hu, csub(wvc, wvm)));
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
      su2.row(r, ur ? au : 0.0, ur ? bu : 1.0, ur ? cc : 0.0, rr);
      if (live && ur) {
        Nu1.set(oq, N1);
        Nu3.set(oq, N3);
      }
      u1m = u1c; u1c = u1p;
      u3m = u3c; u3c = u3p;
      uvm = uvc; wvm = wvc;
    }
    seq_solve<NC, P, M, 2>(pu, U2, c, s);
#pragma unroll
    for (int r = 0; r < M; r++) {
      xu[0][r] = su2.d[0][r];
      xu[1][r] = su2.d[1][r];
    }
  } else {
    // field F = 0 (u1: flux uv, Nu1, Nxz1, the kx pressure term and the mean pressure gradient) or
    // F = 1 (u3: flux wv, Nu3, Nxz3, the kz pressure term); the same operations per field as above
    auto uphase = [&](auto fi) {
      constexpr int F = decltype(fi)::value;
      Part<NC, P, M, 1> pf;
      Segment<M, 1> &sf = pf.sg;
      const double2 *uf = F == 0 ? a.u1 : a.u3;
      const NA<NT> &fl = F == 0 ? uvf : wvf;
      const NA<NT> &nuo = F == 0 ? Nu1 : Nu3;
      const NA<NT> &nxz = F == 0 ? Nxz1 : Nxz3;
      const double c3 = F == 0 ? c3x : c3z;
      const int jb = F == 0 ? j0 : opaque(j0), cl = F == 0 ? col : opaque(col);
      auto at = [&](int j) { return (size_t)j * C + cl; };
      double2 um = uf[at(cu(jb - 1))], uc = uf[at(cu(jb))];
      double2 fm = fl[at(cv(jb - 1))];
      // the global loads of a row are issued one row ahead, before the previous row's arithmetic and
      // store, so a wavefront has two rows of loads in flight instead of one
      struct RowIn {
        double2 pc, up, fc, nu, nx;
        double g1, g2, g3, hu;
      };
      auto load_row = [&](int r) {
        RowIn q;
        const int j = jb + r, ju = cp(j);
        const size_t oq = at(ju - 1);
        if constexpr (F == 0) q.pc = a.p[oq];
        q.up = uf[at(cu(j + 1))];
        q.fc = fl[at(cv(j))];
        q.nu = nuo[oq];
        q.nx = nxz[oq];
        q.g1 = D2U(1, ju);
        q.g2 = D2U(2, ju);
        q.g3 = D2U(3, ju);
        q.hu = a.dthdyu[ju];
        return q;
      };
      RowIn nxt = load_row(0);
#pragma unroll
      for (int r = 0; r < M; r++) {
        const RowIn q = nxt;
        if (r + 1 < M) nxt = load_row(r + 1);
        const int j = jb + r;
        const bool ur = j <= nyu;
        const int ju = cp(j);
        double2 pc;
        if constexpr (F == 0) {
          pc = q.pc;
          pst[r * NTH + tid] = pc;  // p row cp(j)
        } else {
          pc = pst[r * NTH + tid];
        }
        const double2 up = q.up, fc = q.fc, nu = q.nu, nx = q.nx;
        const size_t oq = at(ju - 1);
        const double g1 = q.g1, g2 = q.g2, g3 = q.g3, hu = q.hu;
        const double dd2 = k2 + g2;
        const double2 Lf = cadd(cadd(rmul(g1, um), rmul(dd2, uc)), rmul(g3, up));
        double2 rf = cadd(cadd(rmul(C1, Lf), rmul(C2, nu)), imul(c3, pc));
        if constexpr (F == 0) rf.x = rf.x + Cm;
        const double2 Nf = cadd(nx, rmul(hu, csub(fc, fm)));
        rf = cadd(rf, rmul(Cg, Nf));
        double au = -beta * g1, bu = 1.0 - beta * (g2 + m.k2x + m.k2z), cc = -beta * g3;
        const double bf = bu - au * a.gw1;
        bu = j == 1 ? bf : bu;
        au = j == 1 ? 0.0 : au;
        const double bl = bu - cc * a.gw2;
        bu = j == nyu ? bl : bu;
        cc = j == nyu ? 0.0 : cc;
        const double2 rr = ur ? d2(uc.x + dt * rf.x, uc.y + dt * rf.y) : z2;
        sf.row(r, ur ? au : 0.0, ur ? bu : 1.0, ur ? cc : 0.0, &rr);
        if (live && ur) nuo.set(oq, Nf);
        um = uc; uc = up;
        fm = fc;
      }
      seq_solve<NC, P, M, 1>(pf, U1, c, s);
#pragma unroll
      for (int r = 0; r < M; r++) xu[F][r] = sf.d[0][r];
    };
    uphase(std::integral_constant<int, 0>{});
    uphase(std::integral_constant<int, 1>{});
  }
  struct {
    double2 d[2][M];
  } su;
#pragma unroll
  for (int r = 0; r < M; r++) {
    su.d[0][r] = xu[0][r];
    su.d[1][r] = xu[1][r];
  }
  // su.d[0][r], su.d[1][r]: u1, u3 of row j0+r after solveU

  // ---- phase V: v-grid rows, u2 segment ------------------------------------------------------------
  Part<NC, P, M, 1> pv;
  Segment<M, 1> &sv = pv.sg;
  {
    const int jb = opaque(j0), cl = opaque(col);
    auto at = [&](int j) { return (size_t)j * C + cl; };
    double2 u2m = a.u2[at(cv(jb - 1))], u2c = a.u2[at(cv(jb))];
    double2 vvc_c = vvc[at(cu(jb))];
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int j = jb + r;
      const bool vr = j <= nyv;
      const int jv = cq(j);
      const double2 pc = pst[r * NTH + tid];
      const double2 pn = r + 1 < M ? pst[(r + 1) * NTH + tid] : a.p[at(cp(j + 1) - 1)];
      const double2 u2p = a.u2[at(cv(j + 1))];
      const double2 vvn = vvc[at(cu(j + 1))];
      const size_t ov = at(jv - 1);
      const double2 nu2 = Nu2[ov], nx2 = Nxz2[ov];
      const double e1 = D2V(1, jv), e2 = D2V(2, jv), e3 = D2V(3, jv), hv = L.yc[0][j - 1];
      const double2 L2 = cadd(cadd(rmul(e1, u2m), rmul(k2 + e2, u2c)), rmul(e3, u2p));
      const double C3 = -((a.dRK * 1.0) * hv);
      double2 r2 = cadd(cadd(rmul(C1, L2), rmul(C2, nu2)), rmul(C3, csub(pn, pc)));
      const double2 N2 = cadd(nx2, rmul(hv, csub(vvn, vvc_c)));
      r2 = cadd(r2, rmul(Cg, N2));
      const double av = j == 1 ? 0.0 : -beta * e1;    // wall v_0 = 0
      const double bv = 1.0 - beta * (e2 + m.k2x + m.k2z);
      const double cw = j == nyv ? 0.0 : -beta * e3;  // wall v_{nyv+1} = 0
      const double2 rv = vr ? d2(u2c.x + dt * r2.x, u2c.y + dt * r2.y) : z2;
      sv.row(r, vr ? av : 0.0, vr ? bv : 1.0, vr ? cw : 0.0, &rv);
      if (live && vr) Nu2.set(ov, N2);
      vvc_c = vvn;
      u2m = u2c; u2c = u2p;
    }
  }
  seq_solve<NC, P, M, 1>(pv, U1, c, s);
  // sv.d[0][r]: u2 of row j0+r after solveU
  L.x0[s * NC + c] = sv.d[0][M - 1];
  __syncthreads();

  // ---- phase P: divergence, psi right-hand side, the pressure system -------------------------------
  // rows of LU_buildP (tridLU_3D.f90:242-280) in the operand order of dg_row
  const double dtirk = dti / (a.aRK + a.bRK);
  Part<NC, P, M, 1> pp;
  Segment<M, 1> &sp = pp.sg;
  {
    double2 vm = s == 0 ? z2 : L.x0[(s - 1) * NC + c];
    const int jb = opaque(j0);
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int j = jb + r;
      const int jp = cp(j);
      const double hu = L.yc[1][j - 1], hvl = L.yc[2][j - 1], hvu = L.yc[3][j - 1];
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
  seq_solve<NC, P, M, 1>(pp, U1, c, s);
  L.x1[s * NC + c] = sp.d[0][0];
  __syncthreads();

  // ---- p += psi, v_corr, wall rows; corrected values stay in su / sv for max|div| --------------
  const double dtrk = dt * (a.aRK + a.bRK);
  const double kxd = dtrk * m.kx, kzd = dtrk * m.kz;
  const bool qx = active && cb == 0 && a.kRK == 3 && a.flag_ctpress == 1;
  {
    const double2 psn_top = s == P - 1 ? z2 : L.x1[(s + 1) * NC + c];
    const int jb = opaque(j0), cl = opaque(col);
    auto at = [&](int j) { return (size_t)j * C + cl; };
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int j = jb + r;
      const double hv = L.yc[0][j - 1];
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
          a.p[at(j - 1)] = cadd(pst[r * NTH + tid], ps);
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
  L.x0[s * NC + c] = sv.d[0][M - 1];
  __syncthreads();
  double dmax = 0.0;
  {
    double2 vm = s == 0 ? z2 : L.x0[(s - 1) * NC + c];
    const int jb = opaque(j0);
#pragma unroll
    for (int r = 0; r < M; r++) {
      const int j = jb + r;
      const double hu = L.yc[1][j - 1];  // from LDS: a global read here would wait for the stores above
      const double2 dv = cadd(cadd(imul(m.kx, su.d[0][r]), imul(m.kz, su.d[1][r])),
                              rmul(1.0 * hu, csub(sv.d[0][r], vm)));
      if (j <= nyu) dmax = fmax(dmax, hypot(dv.x, dv.y));
      vm = sv.d[0][r];
    }
  }
  if (!live) dmax = 0.0;
  dmax = block_max_sub<NTH>(dmax, L.red, tid);
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
