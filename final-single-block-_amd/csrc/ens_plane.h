// Grouped plane launches of the ensemble form: channels of different sizes (one dns_handle per size,
// each holding its same-size members) advance together, one launch per plane kernel class per RK3
// substep for all members of all sizes.  Each launch is a 1-D grid whose workgroup ranges belong to the
// groups in turn; a workgroup finds its group and runs the bodies (plane_kernels.h, xpass_reg.h) of the
// group's transform length and tile variant.  Every member runs the reference's single-channel RK3 loop
// (littleharsh.f90:163-233).
//
// Tile shapes: variant v of a length is a body of 64 << v threads (z-kernels: 16/32/32 columns at 48
// points, 8 << v at 96, 4 << v at 192, about one butterfly of the first radix pass per thread; the
// register x-pass: 1, 2 or 4 wavefronts of rows).  A launch has one variant and one workgroup size WG
// (128 or 256 threads), and a workgroup runs WG / (64 << v) bodies of its group side by side, each on the
// next logical block: five shapes per class, (WG, v) = (128, 0), (128, 1), (256, 0), (256, 1), (256, 2),
// each its own kernel (so each keeps the registers of its own bodies).  The ensemble autotuner
// (dns_ensemble_autotune) times the five on the whole launch and keeps the fastest per class; every shape
// computes the same bits (the radix plan is fixed per length), so a member's result equals the same
// channel run alone.
#pragma once
#include "plane_launch.h"

namespace dnsk {

__host__ __device__ constexpr size_t al16(size_t x) { return (x + 15) / 16 * 16; }

template <int N, int V> struct EnsZT {  // z-kernel tile of variant V at length N
  static constexpr int NTH = 64 << V;
  static constexpr int T = N == 48 ? (V == 0 ? 16 : 32) : N == 96 ? (8 << V) : (4 << V);
  static constexpr int ZCH = 2;  // planes per z-inverse workgroup
};
// which: 0 z-inverse, 1/2 x-pass u/v-planes, 3/4 z-forward u/v-planes; body of length N, variant V
template <typename R, int WHICH, int N, int V> struct EnsBodyV;
template <typename R, int N, int V> struct EnsBodyV<R, 0, N, V> {
  using Z = EnsZT<N, V>;
  using B = ZinvBody<R, N, Z::T, Z::NTH, typename ZPlanT<N>::type, Z::ZCH>;
};
// (the grouped launches' x-pass bodies take their twiddles from LDS, as the single-channel kernel does)
template <typename R, int N, int V> struct EnsBodyV<R, 1, N, V> { using B = XregBody<R, N, 64 << V, false, true>; };
template <typename R, int N, int V> struct EnsBodyV<R, 2, N, V> { using B = XregBody<R, N, 64 << V, true, true>; };
template <typename R, int N, int V> struct EnsBodyV<R, 3, N, V> {
  using Z = EnsZT<N, V>;
  using B = ZfwdBody<R, N, Z::T, Z::NTH, false, typename ZPlanT<N>::type>;
};
template <typename R, int N, int V> struct EnsBodyV<R, 4, N, V> {
  using Z = EnsZT<N, V>;
  using B = ZfwdBody<R, N, Z::T, Z::NTH, true, typename ZPlanT<N>::type>;
};
// the 128-thread variant (the persistent launch's plane bodies; its x-pass bodies read the twiddles from
// global memory and hold no barrier of their own)
template <typename R, int WHICH, int N> struct EnsBody { using B = typename EnsBodyV<R, WHICH, N, 1>::B; };
template <typename R, int N> struct EnsBody<R, 1, N> { using B = XregBody<R, N, 128, false, false>; };
template <typename R, int N> struct EnsBody<R, 2, N> { using B = XregBody<R, N, 128, true, false>; };

// LDS of a WG-thread workgroup of variant V: WG / NTH bodies of the largest length
template <typename R, int WHICH, int WG, int V> constexpr size_t ens_plane_lds() {
  constexpr int NTH = 64 << V;
  constexpr size_t a = EnsBodyV<R, WHICH, 48, V>::B::LDS, b = EnsBodyV<R, WHICH, 96, V>::B::LDS,
                   c = EnsBodyV<R, WHICH, 192, V>::B::LDS;
  return (WG / NTH) * al16(a > b ? (a > c ? a : c) : (b > c ? b : c));
}

__device__ __forceinline__ int ens_group(const int *off, int ng, int lb) {
  int g = 0;
  while (g + 1 < ng && lb >= off[g + 1]) g++;
  return g;
}

// Workgroup lw of a group runs logical blocks lw * SUB .. lw * SUB + SUB - 1, one per body.  A body past
// the group's last block computes that block again: every barrier is reached, and it stores the same
// values as the body that owns the block (no plane kernel reads what it writes).
template <typename R, int WHICH, int N, int V, int WG>
__device__ __forceinline__ void ens_plane_run(const PlaneArgs &a, int lw, int gx, int gy, int nblk,
                                              unsigned char *lds) {
  using B = typename EnsBodyV<R, WHICH, N, V>::B;
  if constexpr (B::NTH <= WG) {
    constexpr int SUB = WG / B::NTH;
    const int sub = SUB == 1 ? 0 : (int)threadIdx.x / B::NTH;
    const int lb = min(lw * SUB + sub, nblk - 1);
    B::run(a, lb, gx, gy, lds + sub * al16(B::LDS));
  }
}

template <typename R, int WHICH, int WG, int V>
__global__ void __launch_bounds__(WG) ens_plane_kernel(EnsPlaneLaunch E) {
  __shared__ __align__(16) unsigned char lds[ens_plane_lds<R, WHICH, WG, V>()];
  const int g = ens_group(E.off, E.ng, blockIdx.x);
  int lw = blockIdx.x - E.off[g];
  if constexpr (WHICH == 0 || WHICH >= 3) {
    // z-kernels: each group's workgroups start at a multiple of 8 (the padding workgroups leave at once),
    // and their order inside the group is XCD-aware: each XCD runs a contiguous run of the group's
    // workgroups (neighbouring tiles of a row share their edge lines in one L2), and every XCD gets its
    // share of every group (the groups' tiles differ in cost)
    if (lw >= E.nwg[g]) return;
    lw = xcd_block(lw, E.nwg[g]);
  }
  const PlaneArgs &a = E.a[g];
  const int gx = E.gx[g], gy = E.gy[g], nb = E.nblk[g];
  switch (E.len[g]) {
    case 48: ens_plane_run<R, WHICH, 48, V, WG>(a, lw, gx, gy, nb, lds); break;
    case 96: ens_plane_run<R, WHICH, 96, V, WG>(a, lw, gx, gy, nb, lds); break;
    default: ens_plane_run<R, WHICH, 192, V, WG>(a, lw, gx, gy, nb, lds); break;
  }
}

// Logical grid of group g for kernel class which and tile variant v (the single-channel launchers'
// grids): gx tiles fastest, then gy, nblk logical blocks in all.
inline void ens_plane_grid(const PlaneArgs &a, int which, int v, int &gx, int &gy, int &nblk) {
  const int n = (which == 1 || which == 2) ? a.Mx : a.Mz;
  const int nth = 64 << v;
  const int T = n == 48 ? (v == 0 ? 16 : 32) : n == 96 ? (8 << v) : (4 << v);
  const int ZCH = 2;
  if (which == 0) {
    gx = (a.Kx + T - 1) / T;
    gy = 3;
    nblk = gx * gy * ((a.nbj + ZCH - 1) / ZCH) * a.nm;
  } else if (which <= 2) {
    const int L = n == 48 ? XregShape<24>::L : n == 96 ? XregShape<48>::L : XregShape<96>::L;  // lanes per row
    const int RB = (nth / 64) * (64 / L);                                                        // rows per body
    gx = (a.Mz + RB - 1) / RB;
    gy = a.nbj * a.nm;
    nblk = gx * gy;
  } else {
    gx = (a.Kx + T - 1) / T;
    gy = a.nbj * a.nm;
    nblk = gx * gy;
  }
}

// Whether group a takes shape (wg, v) of class which: the body fits the workgroup, and several z-inverse
// bodies in one workgroup walk chunks of the same length (each row of a chunk is a barrier, so the planes
// must split into whole chunks).
inline bool ens_plane_fits(const PlaneArgs &a, int which, int v, int wg) {
  const int nth = 64 << v;
  if (v < 0 || v >= ENS_NVAR || nth > wg || (wg != 128 && wg != 256)) return false;
  return !(which == 0 && wg > nth && a.nbj % EnsZT<48, 0>::ZCH != 0);
}

// Grouped launch of class which, shape (wg, v); -1 when a group does not take the shape
template <typename R, int WHICH>
int launch_ens_plane_t(EnsPlaneLaunch E, int wg, int v, hipStream_t st) {
  int total = 0;
  for (int g = 0; g < E.ng; g++) {
    if (!ens_plane_fits(E.a[g], WHICH, v, wg)) return -1;
    const int sub = wg / (64 << v);
    int nb = 0;
    ens_plane_grid(E.a[g], WHICH, v, E.gx[g], E.gy[g], nb);
    E.len[g] = (WHICH == 1 || WHICH == 2) ? E.a[g].Mx : E.a[g].Mz;
    E.nblk[g] = nb;
    E.nwg[g] = (nb + sub - 1) / sub;
    E.off[g] = total;
    total += (WHICH == 1 || WHICH == 2) ? E.nwg[g] : (E.nwg[g] + 7) / 8 * 8;
  }
  E.off[E.ng] = total;
  if (wg == 128 && v == 0) ens_plane_kernel<R, WHICH, 128, 0><<<total, 128, 0, st>>>(E);
  else if (wg == 128) ens_plane_kernel<R, WHICH, 128, 1><<<total, 128, 0, st>>>(E);
  else if (v == 0) ens_plane_kernel<R, WHICH, 256, 0><<<total, 256, 0, st>>>(E);
  else if (v == 1) ens_plane_kernel<R, WHICH, 256, 1><<<total, 256, 0, st>>>(E);
  else ens_plane_kernel<R, WHICH, 256, 2><<<total, 256, 0, st>>>(E);
  return 0;
}

}  // namespace dnsk
