// Plane-stage kernels of the nonlinear term (FOU3D.f90:51-371, ops_in_planes2 FOU3D.f90:1514-1652).
//
// Per y-plane the reference runs: zero-padded inverse FFT in z on the Kx retained x-modes (cft, pruned),
// inverse real FFT in x on all Mz rows (rft), pointwise products, forward real FFT in x, forward FFT in z
// on the Kx x-modes, the spectral derivatives der_x/der_z with their dealias mask, and keeps vv, uv, wv
// for the y-flux.  Here that work is three kernels over a batch of planes:
//   zinv_kernel : spectral rows of one field -> I[plane][field][z][i]                    (z-inverse)
//                 A workgroup walks the y-rows of a chunk of planes, transforms every field row once
//                 and writes both the plain row (u1, u3 of u-plane j; u2 of v-plane j) and, from the
//                 row and the previous one kept in registers, the y-interpolated row (interp_u/interp_v,
//                 FOU3D.f90:373-418): the z-transform is linear, so interpolating after it equals
//                 interpolating before it.
//   xreg_kernel : I rows -> physical rows -> products -> P[plane][product][z][i]          (x-pass,
//                 xpass_reg.h); also the CFL maxima of dtc_calc (FOU3D.f90:613-651) at kRK = 1.
//   zfwd_kernel : P -> forward z FFT -> der_x + der_z (masked) and raw fluxes in device layout.
// Planes: u-plane j (u1, u2_itp, u3 -> uu, uw, vv, ww) and v-plane j (u1_itp, u2, u3_itp -> uv, wv) for
// j = 1..nyu.  v-planes run to j = nyv+1 as the reference does on one rank (start.f90:1128-1180 gives
// the single rank v-planes 1..nyu).  A batch holds planes j0..j0+nbj-1 of both kinds: I and P plane
// index jj (u) and nbj + jj (v), jj = j - j0.
#pragma once
#include "dns_kernels.h"

namespace dnsk {

// stride N / first radix of plan PL for length N
template <class PL, int N> constexpr int PL_M = N / plan_info<PL>::first;

// ------------------------------------------------------------------------------------------------
// Grid: (column tiles, field, chunk of ZCH planes); ZCH + 1 field rows per workgroup.  Field rows walked: u1, u3 rows jc0..jc1+1 (plain
// -> u-plane r, interpolated with the row below -> v-plane r-1); u2 rows jc0-1..jc1 (plain -> v-plane r,
// interpolated with the row below -> u-plane r).
// Every plane kernel is a body (a struct with the LDS it needs and run()) that takes its logical block
// index and an LDS buffer, and a thin __global__ wrapper; the ensemble's grouped launches
// (ensemble.hip) run the bodies of several transform lengths in one launch.
template <typename R, int MZ, int T, int NTH_, class PL, int ZCH, int PS = 4>
struct ZinvBody {
  static constexpr int NTH = NTH_;
  using L = Lay<MZ, T, T, 1, 1, PS, false>;  // T adjacent columns, column fastest across threads
  static constexpr size_t LDS = sizeof(cx<R>) * L::SIZE;
  // grid (gx column tiles, gy = 3 fields, chunks x members); lb = logical block, tile fastest
  __device__ __forceinline__ static void run(const PlaneArgs &a_, int lb, int gx, int gy, unsigned char *lds);
};
template <typename R, int MZ, int T, int NTH_, class PL, int ZCH, int PS>
__device__ __forceinline__ void ZinvBody<R, MZ, T, NTH_, PL, ZCH, PS>::run(const PlaneArgs &a_, int lb, int gx,
                                                                           int gy, unsigned char *lds) {
  cx<R> *s = reinterpret_cast<cx<R> *>(lds);
  constexpr int KS = last_slots<L, NTH, PL>();
  const int nch = (a_.nbj + ZCH - 1) / ZCH;  // grid z = chunks x ensemble members
  const int bx = lb % gx, by = (lb / gx) % gy, bz = lb / (gx * gy);
  const PlaneArgs a = plane_view<R>(a_, bz / nch);
  const int f = by;
  const int jc0 = a.j0 + (bz % nch) * ZCH, jc1 = min(jc0 + ZCH - 1, a.j0 + a.nbj - 1);
  const int i0 = bx * T;
  const int Kx = a.Kx, C = a.C, Kp = row_pitch<R>(Kx);
  const size_t fsz = (size_t)MZ * Kp;
  const cx<R> *tw = reinterpret_cast<const cx<R> *>(a.twz);
  const double2 *fld = f == 0 ? a.u1 : f == 1 ? a.u2 : a.u3;
  cx<R> *I = reinterpret_cast<cx<R> *>(a.I) + (size_t)f * fsz + i0;
  const int rbeg = f == 1 ? jc0 - 1 : jc0, rend = f == 1 ? jc1 : jc1 + 1;
  cx<R> prev[KS];
  for (int r = rbeg; r <= rend; r++) {
    const double2 *row = fld + (size_t)r * C;
    // plain output: u-plane r (u1, u3) or v-plane r (u2); interpolated output with row r-1:
    // v-plane r-1 (u1, u3) or u-plane r (u2)
    const bool plain = r >= jc0 && r <= jc1;
    const int jp = f == 1 ? r : r - 1;
    const bool itp = r > rbeg;
    cx<R> *Ip = I + (size_t)((f == 1 ? a.nbj : 0) + (plain ? r - a.j0 : 0)) * 3 * fsz;
    cx<R> *Ii = I + (size_t)((f == 1 ? 0 : a.nbj) + (itp ? jp - a.j0 : 0)) * 3 * fsz;
    double ca_ = 0.0, cb_ = 0.0;  // weight of row r (upper) and row r-1 (lower)
    if (itp) {
      if (f == 1) {  // interp_v at yu[jp] between yv[jp-1] and yv[jp]
        const double dd = a.yv[jp] - a.yv[jp - 1];
        ca_ = (a.yu[jp] - a.yv[jp - 1]) / dd;
        cb_ = (a.yv[jp] - a.yu[jp]) / dd;
      } else {       // interp_u at yv[jp] between yu[jp] and yu[jp+1]
        const double dd = a.yu[jp + 1] - a.yu[jp];
        ca_ = (a.yv[jp] - a.yu[jp]) / dd;
        cb_ = (a.yu[jp + 1] - a.yv[jp]) / dd;
      }
    }
    const R ca = (R)ca_, cb = (R)cb_;
    // position i + rM of column i0 + t; the rows of the 3/2 padding band (MZ/3 <= z < 2 MZ/3) are
    // zero.  When the first pass's stride divides MZ/3 the band is known per rM at compile time.
    auto ld = [&](int i, auto rM, int t) -> cx<R> {
      constexpr int zr = decltype(rM)::value, B0 = MZ / 3, B1 = 2 * MZ / 3;
      constexpr int M = PL_M<PL, MZ>;
      constexpr bool exact = (B0 % M) == 0;
      const int z = i + zr;
      const bool kept = exact ? (zr < B0 || zr >= B1) : (z < B0 || z >= B1);
      if (exact && !kept) return mk<R>(0, 0);
      const int ic = min(i0 + t, Kx - 1);             // clamped: loads stay in bounds, masked below
      const int zc = (exact ? zr < B0 : z < B0) ? z : z - B0;
      double2 v = row[zc * Kx + ic];
      if (!kept || i0 + t >= Kx) v = d2(0.0, 0.0);
      return from_d2<R>(v);
    };
    auto st = [&](int z, int t, int k, const cx<R> &v) {
      if (i0 + t < Kx) {
        if (plain) Ip[z * Kp + t] = v;
        if (itp) Ii[z * Kp + t] = mk<R>(ca * v.x + cb * prev[k].x, ca * v.y + cb * prev[k].y);
      }
      prev[k] = v;
    };
    fft_mem<R, L, NTH, +1, PL>(s, tw, ld, st);
    __syncthreads();
  }
}

template <typename R, int MZ, int T, int NTH, class PL, int ZCH, int WPE = 1, int PS = 4>
__global__ void __launch_bounds__(NTH, WPE) zinv_kernel(PlaneArgs a) {
  using B = ZinvBody<R, MZ, T, NTH, PL, ZCH, PS>;
  __shared__ __align__(16) unsigned char lds[B::LDS];
  // logical block (tile fastest, then field, then chunk), XCD-aware: neighbouring tiles of one row run
  // on one XCD and share the lines at their edges (a k-row of Kx modes is not line-aligned)
  const int gx = gridDim.x, gy = gridDim.y;
  const int lb = xcd_block(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z);
  B::run(a, lb, gx, gy, lds);
}

// ------------------------------------------------------------------------------------------------
// The product fields are transformed one at a time, ordered so that at most two outputs are live in
// registers: u-planes uw (starts der_z(uw) and der_x(uw)), uu (completes Nxz1 = der_x(uu) + der_z(uw)),
// ww (completes Nxz3 = der_x(uw) + der_z(ww)), vv (flux); v-planes uv (flux, starts der_x(uv)),
// wv (flux, completes Nxz2 = der_x(uv) + der_z(wv)).  Sums are formed in the reference's operand order
// (ops_in_planes2, FOU3D.f90:1590-1640).  The kz = -Nz/2 row is inside der_x/der_z's zero band; the raw
// fluxes keep it.
template <typename R, int MZ, int T, int NTH_, bool VPL, class PL, int PS = 4>
struct ZfwdBody {
  static constexpr int NTH = NTH_;
  using L = Lay<MZ, T, T, 1, 1, PS, false>;  // T adjacent columns, column fastest across threads
  static constexpr size_t LDS = sizeof(cx<R>) * L::SIZE;
  // grid (gx column tiles, planes x members); lb = logical block, tile fastest
  __device__ __forceinline__ static void run(const PlaneArgs &a_, int lb, int gx, int gy, unsigned char *lds);
};
template <typename R, int MZ, int T, int NTH_, bool VPL, class PL, int PS>
__device__ __forceinline__ void ZfwdBody<R, MZ, T, NTH_, VPL, PL, PS>::run(const PlaneArgs &a_, int lb, int gx,
                                                                           int gy, unsigned char *lds) {
  (void)gy;
  const int bx = lb % gx, by = lb / gx;
  const PlaneArgs a = plane_view<R>(a_, by / a_.nbj);  // grid y = planes x ensemble members
  cx<R> *s = reinterpret_cast<cx<R> *>(lds);
  constexpr int PERMAX = ((2 * MZ / 3) * T + NTH - 1) / NTH;
  constexpr int NP = VPL ? 2 : 4;
  const int jj = by % a.nbj, j = a.j0 + jj;
  const int i0 = bx * T;
  constexpr int Nz = 2 * MZ / 3;    // retained z-modes (the launcher checks a.Nz)
  const int Kx = a.Kx, C = a.C, Kp = row_pitch<R>(Kx);
  const cx<R> *tw = reinterpret_cast<const cx<R> *>(a.twz);
  const cx<R> *P = reinterpret_cast<const cx<R> *>(a.P) + (size_t)((VPL ? a.nbj : 0) + jj) * 4 * MZ * Kp;
  const double inv = 1.0 / (double)MZ;
  const size_t ju = (size_t)(j - 1) * C, jf = (size_t)j * C;
  // outputs are stored in the plane precision (FP32 in mixed mode)
  auto put = [](void *base, size_t i, double2 v) { reinterpret_cast<cx<R> *>(base)[i] = from_d2<R>(v); };
  double2 accA[PERMAX], accB[PERMAX];
#pragma unroll
  for (int step = 0; step < NP; step++) {
    // product slot of this step: u-planes uw, uu, ww, vv = 1, 0, 3, 2; v-planes uv, wv = 0, 1
    const int p = VPL ? step : (step == 0 ? 1 : step == 1 ? 0 : step == 2 ? 3 : 2);
    auto ld = [&](int i, auto rM, int t) -> cx<R> {
      const int z = i + decltype(rM)::value;
      const int ic = min(i0 + t, Kx - 1);             // clamped: loads stay in bounds, masked below
      const cx<R> v = P[((size_t)p * MZ + z) * Kp + ic];
      return i0 + t < Kx ? v : mk<R>(0, 0);
    };
    fft_from_mem<R, L, NTH, -1, PL>(s, tw, ld);
#pragma unroll
    for (int q = 0; q < PERMAX; q++) {
      const int idx = body_tid<NTH>() + q * NTH;
      if (idx < Nz * T) {
        const int k = idx / T, t = idx - k * T, i = i0 + t;
        const int zp = k < Nz / 2 ? k : k + MZ - Nz;
        double2 v = to_d2(s[L::addr(zp, t)]);
        v = d2(inv * v.x, inv * v.y);
        const double kx = a.alp * (double)i;
        const double kz = k < Nz / 2 ? a.bet * (double)k : -(a.bet * (double)(Nz - k));
        const bool st = i < Kx, masked = (k == Nz / 2);
        const size_t col = (size_t)k * Kx + i;
        if (!VPL) {
          if (step == 0) { accA[q] = imul(kz, v); accB[q] = imul(kx, v); }
          else if (step == 1) { if (st) put(a.Nxz1, ju + col, masked ? d2(0.0, 0.0) : cadd(imul(kx, v), accA[q])); }
          else if (step == 2) { if (st) put(a.Nxz3, ju + col, masked ? d2(0.0, 0.0) : cadd(accB[q], imul(kz, v))); }
          else { if (st) put(a.vvc, jf + col, v); }
        } else {
          if (step == 0) { accA[q] = imul(kx, v); if (st) put(a.uvf, jf + col, v); }
          else {
            if (st) {
              if (j <= a.nyv) put(a.Nxz2, ju + col, masked ? d2(0.0, 0.0) : cadd(accA[q], imul(kz, v)));
              put(a.wvf, jf + col, v);
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

template <typename R, int MZ, int T, int NTH, bool VPL, class PL, int WPE = 1, int PS = 4>
__global__ void __launch_bounds__(NTH, WPE) zfwd_kernel(PlaneArgs a) {
  using B = ZfwdBody<R, MZ, T, NTH, VPL, PL, PS>;
  __shared__ __align__(16) unsigned char lds[B::LDS];
  // logical block (tile fastest), XCD-aware: neighbouring tiles share the output lines at their edges
  const int gx = gridDim.x;
  const int lb = xcd_block(blockIdx.x + gx * blockIdx.y, gx * gridDim.y);
  B::run(a, lb, gx, gridDim.y, lds);
}

}  // namespace dnsk
