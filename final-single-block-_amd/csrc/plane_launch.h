// Tile shapes and launch templates of the plane kernels (plane_kernels.h), shared by the translation
// units that instantiate them (plane_zinv.hip, plane_xpass_f64.hip, plane_xpass_f32.hip, plane_zfwd.hip:
// one unit per kernel class and precision, so they compile in parallel).
#pragma once
#include "plane_kernels.h"
#include "xpass_reg.h"

namespace dnsk {

// ------------------------------------------------------------------------------------------------
// launchers.  The radix plan is fixed per transform length (so every tile variant of a kernel gives the
// same bits); the tile variants differ in columns/rows per workgroup and threads, and the autotuner
// picks among them.  Plans and tiles of the tuned lengths were chosen by timing on MI355X:
//   z-transforms of 576: 12.12.4 (one butterfly per thread in every pass with 48 threads per column),
//     16 columns per workgroup (256-byte row segments in FP64);
//   x-pass: the register x-pass (xpass_reg.h) at every length, 3 or 9 points per lane (9 at 576 and
//     1152, with 32 or 64 lanes per row).
#define DNS_SIZES(X) X(12) X(24) X(48) X(96) X(144) X(192) X(288) X(384) X(576) X(1152)
#define DNS_TUNED(N) ((N) == 192 || (N) == 576 || (N) == 1152)


constexpr size_t LDS_MAX = 160 * 1024;

// z-kernel tiles (T columns, NTH threads, ZCH planes per z-inverse workgroup, which reads ZCH + 1 rows
// of each field); 12.12.x plans: 48 threads per column (one butterfly each), at most 768 threads.
// 576: variant 1 is variant 0 with 4 planes per z-inverse workgroup (fewer rows read twice); variant 2
// takes 32 columns (two butterflies per thread; 512-byte FP32 row segments, an LDS tile that only FP32
// fits); in FP64 it is variant 0 with 8 planes per z-inverse workgroup instead (tile 3).
// Other lengths: 16, 8 or 32 columns, with threads for about one butterfly of the 16-point first pass
// each (64 to 256).
template <int N, int V> struct ZTile {
  static constexpr bool p12 = (N == 576 || N == 1152);
  static constexpr int T = N == 576 ? (V == 2 ? 32 : 16)
                         : N == 1152 ? (V == 0 ? 8 : V == 1 ? 4 : 16) : (V == 0 ? 16 : V == 1 ? 8 : 32);
  static constexpr int NB = (T * N / 16 + 63) / 64 * 64;  // first-pass butterflies, whole wavefronts
  static constexpr int NTH = p12 ? (48 * T < 768 ? 48 * T : 768) : (NB < 64 ? 64 : NB > 256 ? 256 : NB);
  static constexpr int ZCH = (N == 576 && V == 1) ? 4 : (N == 576 && V == 3) ? 8 : 2;
};
template <typename R, int MZ, int V>
inline void zinv_v(const PlaneArgs &a, int nw, hipStream_t st) {
  using Z = ZTile<MZ, V>;
  if constexpr (V > 0 && (size_t)Z::T * MZ * sizeof(cx<R>) > LDS_MAX) {
    zinv_v<R, MZ, (MZ == 576 && V == 2) ? 3 : 0>(a, nw, st);
  } else {
    (void)nw;
    dim3 g((a.Kx + Z::T - 1) / Z::T, 3, ((a.nbj + Z::ZCH - 1) / Z::ZCH) * a.nm);
    zinv_kernel<R, MZ, Z::T, Z::NTH, typename ZPlanT<MZ>::type, Z::ZCH, 1, 0><<<g, Z::NTH, 0, st>>>(a);
  }
}
template <typename R, int MZ, int V, bool VPL>
inline void zfwd_v(const PlaneArgs &a, int nw, hipStream_t st) {
  using Z = ZTile<MZ, V>;
  if constexpr (V > 0 && (size_t)Z::T * MZ * sizeof(cx<R>) > LDS_MAX) {
    zfwd_v<R, MZ, 0, VPL>(a, nw, st);
  } else {
    dim3 g((a.Kx + Z::T - 1) / Z::T, nw * a.nm);
    zfwd_kernel<R, MZ, Z::T, Z::NTH, VPL, typename ZPlanT<MZ>::type, 1, 0><<<g, Z::NTH, 0, st>>>(a);
  }
}
// register x-pass (xpass_reg.h): variants 0, 1, 2 = 4, 2, 1 wavefronts per workgroup
template <typename R, int MX, int V, bool VPL>
inline void xreg_v(const PlaneArgs &a, int nw, hipStream_t st) {
  constexpr int NTH = V == 0 ? 256 : V == 1 ? 128 : 64;
  using B = XregBody<R, MX, NTH, VPL, true>;
  xreg_kernel<R, MX, NTH, VPL><<<dim3((a.Mz + B::RB - 1) / B::RB, nw * a.nm), NTH, 0, st>>>(a);
}

template <typename R, int MX, int V, bool VPL>
inline void xpass_v(const PlaneArgs &a, int nw, hipStream_t st) {
  xreg_v<R, MX, V, VPL>(a, nw, st);
}

template <typename R, int N>
inline void zinv_sized(const PlaneArgs &a, int nw, int v, hipStream_t st) {
  if constexpr (DNS_TUNED(N)) {
    if (v == 1) return zinv_v<R, N, 1>(a, nw, st);
    if (v == 2) return zinv_v<R, N, 2>(a, nw, st);
  }
  zinv_v<R, N, 0>(a, nw, st);
}
template <typename R, int N, bool VPL>
inline void zfwd_sized(const PlaneArgs &a, int nw, int v, hipStream_t st) {
  if constexpr (DNS_TUNED(N)) {
    if (v == 1) return zfwd_v<R, N, 1, VPL>(a, nw, st);
    if (v == 2) return zfwd_v<R, N, 2, VPL>(a, nw, st);
  }
  zfwd_v<R, N, 0, VPL>(a, nw, st);
}
template <typename R, int N, bool VPL>
inline void xpass_sized(const PlaneArgs &a, int nw, int v, hipStream_t st) {
  if constexpr (DNS_TUNED(N)) {
    if (v == 1) return xpass_v<R, N, 1, VPL>(a, nw, st);
    if (v == 2) return xpass_v<R, N, 2, VPL>(a, nw, st);
  }
  xpass_v<R, N, 0, VPL>(a, nw, st);
}


// per-class launchers (one translation unit each); 0, or -1 when the length is not instantiated
int launch_zinv(const PlaneArgs &a, int nw, bool fp32, int v, hipStream_t st);
int launch_xpass_f64(const PlaneArgs &a, int nw, bool vpl, int v, hipStream_t st);
int launch_xpass_f32(const PlaneArgs &a, int nw, bool vpl, int v, hipStream_t st);
int launch_zfwd(const PlaneArgs &a, int nw, bool fp32, bool vpl, int v, hipStream_t st);

}  // namespace dnsk
