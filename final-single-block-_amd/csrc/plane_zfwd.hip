// z-forward kernels (zfwd_kernel) of every supported length, both plane kinds and precisions.
#include "plane_launch.h"

namespace dnsk {

template <typename R, bool VPL>
static int zfwd_t(const PlaneArgs &a, int nw, int v, hipStream_t st) {
#define CASE(N) if (a.Mz == N) { zfwd_sized<R, N, VPL>(a, nw, v, st); return 0; }
  DNS_SIZES(CASE)
#undef CASE
  return -1;
}

int launch_zfwd(const PlaneArgs &a, int nw, bool fp32, bool vpl, int v, hipStream_t st) {
  if (fp32) return vpl ? zfwd_t<float, true>(a, nw, v, st) : zfwd_t<float, false>(a, nw, v, st);
  return vpl ? zfwd_t<double, true>(a, nw, v, st) : zfwd_t<double, false>(a, nw, v, st);
}

}  // namespace dnsk
