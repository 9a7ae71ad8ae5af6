// The x-pass (the x-direction half of four_to_phys_u / phys_to_four_du and the products of
// ops_in_planes2, FOU3D.f90:1514-1652, rft_buff.f:62-84), register-resident, for rows whose packed
// length H = MX/2 is RR * L with RR = 1, 3 or 9 and L = 2..64 a power of two (every supported Mx).
//
// One row transform per group of L lanes, RR points per lane, no LDS between the transforms' passes (the
// single-channel and grouped launches keep their twiddle tables in LDS, TWL below):
//   inverse (sign +1), input Z[L n1 + t] in lane t, register n1 (the row's memory order, so the loads
//     are coalesced): RR-point DFT in registers, twiddle e^{+2 pi i t m1/H}, then an L-point DFT across
//     the lanes as radix-2 decimation-in-frequency stages (partner lane t ^ h, values exchanged with DPP
//     for h = 1, 2 and ds_swizzle / ds_bpermute above).  Lane t ends with z[m1 + RR brev(t)]: bit-reversed blocks of RR physical points.
//   forward (sign -1), input in that same layout: radix-2 decimation-in-time stages across the lanes
//     (bit-reversed input, natural output k2 = t), twiddle e^{-2 pi i m1 t/H}, RR-point DFT in registers.
//     Lane t ends with Z[L k1 + t], the inverse's input layout, so the outputs are stored coalesced.
// The products are pointwise and every field of a row has the same layout, so a lane forms them from
// its own registers.  The only cross-lane traffic besides the butterflies is the Hermitian partners
// (X[H - n] of the inverse's pre-processing, Z[H - k] of the forward's post-processing), in lane
// (L - t) mod L (ds_bpermute).
// Pre- and post-processing are the packed real transform (rft_buff.f's semantics): inverse
//   Z_n = (X_n + conj X_{H-n}) + i e^{+2 pi i n/MX} (X_n - conj X_{H-n}),  z = IFFT_H(Z)
// with Im c_0 and Im c_{MX/2} ignored (rft_buff.f:76-84), and forward
//   X_k = [ (Z_k + conj Z_{H-k}) - i e^{-2 pi i k/MX} (Z_k - conj Z_{H-k}) ] / (2 MX),  Z = FFT_H(z)
// scaled 1/MX with Im c_0 = 0 (rft_buff.f:62-74); z packs a real row x as z_m = x_{2m} + i x_{2m+1}.
// The products are formed pointwise: u-planes uu, uw, vv, ww (P slots 0..3), v-planes uv, wv (0..1).
#pragma once
#include "plane_kernels.h"

namespace dnsk {

template <int H> struct XregShape {
  static constexpr int L = H & -H;  // largest power of two dividing H
  static constexpr int RR = H / L;
  static constexpr bool ok = L >= 2 && L <= 64 && (RR == 1 || RR == 3 || RR == 9);
};
// value of lane (lane ^ M) of the wavefront, M < 64 a power of two
template <int M> __device__ __forceinline__ int lane_xor_i(int v) {
  if constexpr (M == 1) {
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);   // quad_perm [1,0,3,2]
  } else if constexpr (M == 2) {
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);   // quad_perm [2,3,0,1]
  } else if constexpr (M == 4 || M == 8 || M == 16) {
    // ds_swizzle bit mode within 32 lanes (and 0x1f, or 0, xor M): the exchange runs on the LDS unit,
    // which the transforms otherwise leave idle, instead of taking VALU issue slots
    return __builtin_amdgcn_ds_swizzle(v, 0x1F | (M << 10));
  } else {
    static_assert(M == 32, "lane exchange");
    return __builtin_amdgcn_ds_bpermute((int)(((threadIdx.x & 63) ^ 32) << 2), v);
  }
}
template <int M> __device__ __forceinline__ double lane_xor(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_xor_i<M>((int)(unsigned)(b & 0xffffffffLL)), hi = lane_xor_i<M>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int M> __device__ __forceinline__ float lane_xor(float v) {
  return __int_as_float(lane_xor_i<M>(__float_as_int(v)));
}
template <int M, typename T> __device__ __forceinline__ T lane_xor_c(T v) {
  T c;
  c.x = lane_xor<M>(v.x);
  c.y = lane_xor<M>(v.y);
  return c;
}
// value of lane src (byte address src * 4) of the wavefront
__device__ __forceinline__ double lane_read(double v, int addr) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float lane_read(float v, int addr) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

template <typename R, int MX, int NTH_, bool VPL, bool TWL = false>
struct XregBody {
  static constexpr int NTH = NTH_;
  static constexpr int H = MX / 2;
  using S = XregShape<H>;
  static constexpr int L = S::L, RR = S::RR;
  static_assert(S::ok, "register x-pass: H = RR L, RR in {1, 3, 9}, L in 2..64");
  static constexpr int GPW = 64 / L;            // rows per wavefront
  static constexpr int RB = (NTH / 64) * GPW;   // rows per workgroup
  static constexpr int ROWS = RB;
  static constexpr int LOGL = L == 2 ? 1 : L == 4 ? 2 : L == 8 ? 3 : L == 16 ? 4 : L == 32 ? 5 : 6;
  static constexpr int NSTG = LOGL - 1;         // butterfly stages with twiddles (half-size h >= 2)
  // TWB: the twiddles of a transform are read into registers before its first arithmetic, the stage
  // twiddles come from the LDS table, and the stashed field is reloaded (FP64; in FP32, where the kernel
  // holds more wavefronts per SIMD, reading them where they are used was faster)
  static constexpr bool TWB = TWL && sizeof(R) == 8;
  // u-planes keep u's physical values in LDS meanwhile; with TWB the v-planes keep v's there too, and
  // both reload them where they are next used instead of holding them in registers (fewer live registers
  // while the next field is loaded and transformed)
  static constexpr int NSTASH = (VPL && !TWB) ? 0 : RR;
  static constexpr int Kx = MX / 3 + 1;
  // TWL: the pre/post-processing twiddles (H entries) and the inverse's/forward's twiddles ((RR - 1) L
  // entries) are copied to LDS at the start, so that their reads inside the transforms do not wait on
  // the vector memory counter behind the row's stores
  static constexpr int NTW = TWL ? H + (RR - 1) * L : 0;
  static constexpr size_t LDS = sizeof(cx<R>) * (NSTASH * NTH + NTW) + sizeof(double) * (NTH / 64);
  static constexpr int KOUT = (Kx + L - 1) / L;  // registers holding outputs k < Kx
  __device__ __forceinline__ static void run(const PlaneArgs &a_, int lb, int gx, int gy, unsigned char *lds);
};

template <typename R, int MX, int NTH_, bool VPL, bool TWL>
__device__ __forceinline__ void XregBody<R, MX, NTH_, VPL, TWL>::run(const PlaneArgs &a_, int lb, int gx, int gy,
                                                               unsigned char *lds) {
  (void)gy;
  const int bx = lb % gx, by = lb / gx;
  const PlaneArgs a = plane_view<R>(a_, by / a_.nbj);  // grid y = planes x ensemble members
  const int jj = by % a.nbj, j = a.j0 + jj;
  const int ltid = body_tid<NTH>();  // this body's thread index (a workgroup may run several bodies)
  const int lane = ltid & 63, wv = ltid >> 6;
  const int t = lane & (L - 1), g = lane / L;
  const int MZ = a.Mz;
  // this lane group's row; groups past the last row (Mz not a multiple of RB) redo the last row and store
  // nothing (its lanes still take part in the wavefront's exchanges)
  const int zr = bx * RB + wv * GPW + g;
  const bool zok = zr < MZ;
  const int z = zok ? zr : MZ - 1;
  constexpr int Kp = row_pitch<R>(Kx);
  const size_t fsz = (size_t)MZ * Kp;
  const size_t pl = (size_t)((VPL ? a.nbj : 0) + jj);
  const cx<R> *I = reinterpret_cast<const cx<R> *>(a.I) + pl * 3 * fsz + (size_t)z * Kp;
  cx<R> *P = reinterpret_cast<cx<R> *>(a.P) + pl * 4 * fsz + (size_t)z * Kp;
  const cx<R> *twx = reinterpret_cast<const cx<R> *>(a.twx);  // e^{-2 pi i m/MX}
  const cx<R> *twh = reinterpret_cast<const cx<R> *>(a.twh);  // e^{-2 pi i m/H}, then twr
  const cx<R> *twr = twh + H;                                  // twr[(m1-1) L + t] = e^{-2 pi i t m1/H}
  cx<R> *stash = reinterpret_cast<cx<R> *>(lds);
  double *red = reinterpret_cast<double *>(lds + sizeof(cx<R>) * (NSTASH * NTH + NTW));
  if constexpr (TWL) {
    cx<R> *ltw = stash + NSTASH * NTH;
    for (int i = ltid; i < H; i += NTH) ltw[i] = twx[i];
    for (int i = ltid; i < (RR - 1) * L; i += NTH) ltw[H + i] = twr[i];
    __syncthreads();
    twx = ltw;
    twr = ltw + H;
  }

  // butterfly twiddles of this lane: stage s (half-size h = 2 << s), e^{-2 pi i (t mod h)/(2h)} on the
  // upper lane of each pair, 1 on the lower, held in registers
  // (with TWB they are read from the LDS copy of twx at each stage instead: twh[m] = twx[2 m], the same
  // bits, since both tables hold the rounded e^{-2 pi i a} of the same a)
  cx<R> ws[NSTG > 0 ? NSTG : 1];
  if constexpr (!TWB) {
#pragma unroll
    for (int s = 0; s < NSTG; s++) {
      const int h = 2 << s;
      ws[s] = (t & h) ? twh[(t & (h - 1)) * (H / (2 * h))] : mk<R>(1, 0);
    }
  }
  auto stage_tw = [&](auto hc) -> cx<R> {
    constexpr int h = decltype(hc)::value;
    constexpr int s = (h == 2 ? 0 : h == 4 ? 1 : h == 8 ? 2 : h == 16 ? 3 : 4);
    if constexpr (TWB) {
      cx<R> w = twx[(t & (h - 1)) * (H / h)];
      if (!(t & h)) w = mk<R>(1, 0);
      return w;
    } else {
      return ws[s];
    }
  };

  // lane (L - t) mod L of this lane's group, as a ds_bpermute address: it holds the Hermitian partners
  const int paddr = ((lane & ~(L - 1)) | ((L - t) & (L - 1))) << 2;
  // packed half-length spectrum of field f: v[n1] = Z[L n1 + t].  In FP32 the row's modes are loaded once
  // (X[L n1 + t], zero from Kx on) and the partner X[H - n] = X[L (RR-1-n1) + (L-t)] is taken from
  // register RR-1-n1 of lane L - t (for t = 0: X[L (RR - n1)], this lane's register RR - n1); in FP64,
  // where the exchange costs two ds_bpermute per component, it is loaded again.  Same values either way.
  auto load_inverse = [&](int f, cx<R> *v) {
    const cx<R> *row = I + f * fsz;
    if constexpr (sizeof(R) == 4) {
      cx<R> xr[RR];
#pragma unroll
      for (int n1 = 0; n1 < RR; n1++) {
        const int n = L * n1 + t;
        xr[n1] = n < Kx ? row[n] : mk<R>(0, 0);
      }
#pragma unroll
      for (int n1 = 0; n1 < RR; n1++) {
        const int n = L * n1 + t;
        cx<R> xb;
        xb.x = lane_read(xr[RR - 1 - n1].x, paddr);
        xb.y = lane_read(xr[RR - 1 - n1].y, paddr);
        if (t == 0) xb = n1 == 0 ? mk<R>(0, 0) : xr[(RR - n1) % RR];
        cx<R> xa = xr[n1];
        if (n == 0) xa.y = 0;
        const cx<R> sum = cadd(xa, cconj(xb)), dif = csub(xa, cconj(xb));
        const cx<R> e = cconj(twx[n]);  // e^{+2 pi i n/MX}
        v[n1] = cadd(sum, cmul_si<1>(cmul(dif, e)));
      }
    } else {
      cx<R> xa[RR], xb[RR], e[RR];
#pragma unroll
      for (int n1 = 0; n1 < RR; n1++) {
        const int n = L * n1 + t;
        xa[n1] = mk<R>(0, 0);
        xb[n1] = mk<R>(0, 0);
        if (n < Kx) xa[n1] = row[n];
        if (n > 0 && H - n < Kx) xb[n1] = row[H - n];
      }
#pragma unroll
      for (int n1 = 0; n1 < RR; n1++) e[n1] = cconj(twx[L * n1 + t]);  // e^{+2 pi i n/MX}
      if constexpr (TWB) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int n1 = 0; n1 < RR; n1++) {
        const int n = L * n1 + t;
        if (n == 0) xa[n1].y = 0;
        const cx<R> sum = cadd(xa[n1], cconj(xb[n1])), dif = csub(xa[n1], cconj(xb[n1]));
        v[n1] = cadd(sum, cmul_si<1>(cmul(dif, e[n1])));
      }
    }
  };
  // inverse transform: in-register DFT, twiddles, cross-lane DIF; lane t ends with z[m1 + RR brev(t)]
  auto inverse = [&](cx<R> *v) {
    if constexpr (RR > 1) {
      cx<R> tw[RR - 1];  // with TWB read before the DFT, which hides their latency
      if constexpr (TWB) {
#pragma unroll
        for (int m1 = 1; m1 < RR; m1++) tw[m1 - 1] = twr[(m1 - 1) * L + t];
        __builtin_amdgcn_sched_barrier(0);
      }
      dft<R, RR, +1>::run(v);
#pragma unroll
      for (int m1 = 1; m1 < RR; m1++) v[m1] = cmulc(v[m1], TWB ? tw[m1 - 1] : twr[(m1 - 1) * L + t]);
    }
    sfor<0, LOGL>([&](auto si) {
      constexpr int h = L >> (decltype(si)::value + 1);
      const R sg = (t & h) ? (R)-1 : (R)1;
      // every exchange of the stage is issued before the first is used, so the LDS unit's latency is
      // paid once per stage rather than once per few values
      cx<R> pv[RR], w = mk<R>(1, 0);
      if constexpr (h >= 2) w = stage_tw(std::integral_constant<int, h>{});
#pragma unroll
      for (int i = 0; i < RR; i++) pv[i] = lane_xor_c<h>(v[i]);
      if constexpr (h >= 4) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < RR; i++) {
        const cx<R> p = pv[i];
        cx<R> d = mk<R>(fma(sg, v[i].x, p.x), fma(sg, v[i].y, p.y));  // lower: v + p, upper: p - v
        if constexpr (h >= 2) d = cmulc(d, w);
        v[i] = d;
      }
    });
  };
  // forward transform: cross-lane DIT, twiddles, in-register DFT; lane t ends with Z[L k1 + t]
  auto forward = [&](cx<R> *v) {
    cx<R> tw[RR > 1 ? RR - 1 : 1];  // with TWB read before the cross-lane stages, which hide their latency
    if constexpr (TWB) {
#pragma unroll
      for (int m1 = 1; m1 < RR; m1++) tw[m1 - 1] = twr[(m1 - 1) * L + t];
      __builtin_amdgcn_sched_barrier(0);
    }
    sfor<0, LOGL>([&](auto si) {
      constexpr int h = 1 << decltype(si)::value;
      const R sg = (t & h) ? (R)-1 : (R)1;
      cx<R> pv[RR], w = mk<R>(1, 0);
      if constexpr (h >= 2) w = stage_tw(std::integral_constant<int, h>{});
#pragma unroll
      for (int i = 0; i < RR; i++) {
        if constexpr (h >= 2) v[i] = cmul(v[i], w);
        pv[i] = lane_xor_c<h>(v[i]);
      }
      if constexpr (h >= 4) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < RR; i++)
        v[i] = mk<R>(fma(sg, v[i].x, pv[i].x), fma(sg, v[i].y, pv[i].y));  // lower: m + p, upper: p - m
    });
    if constexpr (RR > 1) {
#pragma unroll
      for (int m1 = 1; m1 < RR; m1++) v[m1] = cmul(v[m1], TWB ? tw[m1 - 1] : twr[(m1 - 1) * L + t]);
      dft<R, RR, -1>::run(v);
    }
  };
  // real-to-complex post-processing and store of the retained modes k < Kx of product slot p
  const R sc = (R)(0.5 / (double)MX);
  auto store = [&](int p, const cx<R> *v) {
    cx<R> *out = P + (size_t)p * MZ * Kp;
    cx<R> qv[KOUT], e[KOUT];  // every partner exchange (and with TWB twiddle) is read before the first use
#pragma unroll
    for (int k1 = 0; k1 < KOUT; k1++) {
      qv[k1].x = lane_read(v[RR - 1 - k1].x, paddr);
      qv[k1].y = lane_read(v[RR - 1 - k1].y, paddr);
      if constexpr (TWB) e[k1] = twx[L * k1 + t];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k1 = 0; k1 < KOUT; k1++) {
      cx<R> q = qv[k1];
      if (t == 0) q = v[(RR - k1) % RR];  // Z[H - L k1] is this lane's own
      const int k = L * k1 + t;
      if (k < Kx && zok) {
        const cx<R> zk = v[k1], zm = cconj(q);
        const cx<R> sum = cadd(zk, zm), dif = csub(zk, zm);
        cx<R> x = csub(sum, cmul_si<1>(cmul(dif, TWB ? e[k1] : twx[k])));
        x = cscale(x, sc);
        if (k == 0) x.y = 0;
        out[k] = x;
      }
    }
  };
  auto sq = [&](cx<R> *v) {
#pragma unroll
    for (int i = 0; i < RR; i++) v[i] = mk<R>(v[i].x * v[i].x, v[i].y * v[i].y);
  };
  auto times = [&](cx<R> *v, const cx<R> *w) {
#pragma unroll
    for (int i = 0; i < RR; i++) v[i] = mk<R>(v[i].x * w[i].x, v[i].y * w[i].y);
  };

  cx<R> A[RR], B[RR];
  if (!VPL) {
    const bool cfl = (a.kRK == 1) && (j <= a.nyu - 1);
    R m1 = 0, m3 = 0;
    load_inverse(1, A);  // v: vv
    inverse(A);
    sq(A);
    forward(A);
    store(2, A);
    if constexpr (TWB) {
      // one field row in registers at a time: w waits in LDS while ww is transformed, then u w takes its
      // place there while uu is transformed
      load_inverse(2, A);  // w
      inverse(A);
#pragma unroll
      for (int i = 0; i < RR; i++) {
        if (cfl) m3 = fmax(m3, fmax(fabs(A[i].x), fabs(A[i].y)));
        stash[i * NTH + ltid] = A[i];
      }
      // the maxima are pinned here: sunk to their use at the end, they kept w's and u's physical values
      // live through the later transforms (60 more VGPRs)
      asm volatile("" : "+v"(m3));
      __builtin_amdgcn_sched_barrier(0);
      sq(A);  // ww
      forward(A);
      store(3, A);
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
      load_inverse(0, A);  // u
      inverse(A);
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < RR; i++) {
        if (cfl) m1 = fmax(m1, fmax(fabs(A[i].x), fabs(A[i].y)));
        const cx<R> w = stash[i * NTH + ltid];
        stash[i * NTH + ltid] = mk<R>(A[i].x * w.x, A[i].y * w.y);  // uw
      }
      asm volatile("" : "+v"(m1));
      __builtin_amdgcn_sched_barrier(0);
      sq(A);  // uu
      forward(A);
      store(0, A);
      asm volatile("" ::: "memory");
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < RR; i++) A[i] = stash[i * NTH + ltid];
      forward(A);
      store(1, A);
    } else {
      load_inverse(0, A);  // u, kept in LDS (each lane its own slots)
      inverse(A);
#pragma unroll
      for (int i = 0; i < RR; i++) {
        if (cfl) m1 = fmax(m1, fmax(fabs(A[i].x), fabs(A[i].y)));
        stash[i * NTH + ltid] = A[i];
      }
      asm volatile("" : "+v"(m1));
      load_inverse(2, B);  // w
      inverse(B);
#pragma unroll
      for (int i = 0; i < RR; i++) {
        if (cfl) m3 = fmax(m3, fmax(fabs(B[i].x), fabs(B[i].y)));
        A[i] = mk<R>(A[i].x * B[i].x, A[i].y * B[i].y);  // uw
      }
      asm volatile("" : "+v"(m3));
      forward(A);
      store(1, A);
      sq(B);  // ww
      forward(B);
      store(3, B);
#pragma unroll
      for (int i = 0; i < RR; i++) A[i] = stash[i * NTH + ltid];
      sq(A);  // uu
      forward(A);
      store(0, A);
    }
    if (a.kRK == 1) {  // uniform over the workgroup (m1 = m3 = 0 on planes past nyu - 1)
      const double b1 = block_max_sub<NTH>((double)m1, red, ltid);
      const double b3 = block_max_sub<NTH>((double)m3, red, ltid);
      if (ltid == 0 && cfl) {  // spread over CFL_SLOTS addresses: one address for every block serialises
        const int slot = (bx + by) & (CFL_SLOTS - 1);
        atomic_max_nonneg(&a.cfl[slot], b1);
        atomic_max_nonneg(&a.cfl[CFL_SLOTS + slot], b3);
      }
    }
  } else if constexpr (TWB) {
    load_inverse(1, A);  // v, kept in LDS
    inverse(A);
#pragma unroll
    for (int i = 0; i < RR; i++) stash[i * NTH + ltid] = A[i];
#pragma unroll
    for (int f = 0; f < 2; f++) {  // u: uv, then w: wv
      asm volatile("" ::: "memory");
      load_inverse(f == 0 ? 0 : 2, B);
      inverse(B);
      asm volatile("" ::: "memory");
#pragma unroll
      for (int i = 0; i < RR; i++) A[i] = stash[i * NTH + ltid];
      times(B, A);
      forward(B);
      store(f, B);
    }
  } else {
    load_inverse(1, A);  // v
    inverse(A);
    load_inverse(0, B);  // u: uv
    inverse(B);
    times(B, A);
    forward(B);
    store(0, B);
    load_inverse(2, B);  // w: wv
    inverse(B);
    times(B, A);
    forward(B);
    store(1, B);
  }
}

template <typename R, int MX, int NTH, bool VPL, int WPE = 1>
__global__ void __launch_bounds__(NTH, WPE) xreg_kernel(PlaneArgs a) {
  using B = XregBody<R, MX, NTH, VPL, true>;
  __shared__ __align__(16) unsigned char lds[B::LDS];
  B::run(a, blockIdx.x + gridDim.x * blockIdx.y, gridDim.x, gridDim.y, lds);
}

// twr[(m1 - 1) L + t] = e^{-2 pi i t m1/H} (m1 = 1..RR-1, t < L) after the table e^{-2 pi i m/H}: copies of
// table entries, so both give the same bits
template <typename R, int H>
__global__ void xreg_twiddles_kernel(cx<R> *tw) {
  using S = XregShape<H>;
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < H; m += gridDim.x * blockDim.x) {
    const int m1 = m / S::L + 1, t = m % S::L;
    tw[H + m] = m1 < S::RR ? tw[(t * m1) % H] : mk<R>(0, 0);
  }
}

}  // namespace dnsk
