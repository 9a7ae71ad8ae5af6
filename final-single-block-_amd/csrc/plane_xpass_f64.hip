// x-pass kernels (xreg_kernel) of every supported length, both plane kinds, in FP64.
#include "plane_launch.h"

namespace dnsk {

template <bool VPL>
static int xpass_t(const PlaneArgs &a, int nw, int v, hipStream_t st) {
#define CASE(N) if (a.Mx == N) { xpass_sized<double, N, VPL>(a, nw, v, st); return 0; }
  DNS_SIZES(CASE)
#undef CASE
  return -1;
}

int launch_xpass_f64(const PlaneArgs &a, int nw, bool vpl, int v, hipStream_t st) {
  return vpl ? xpass_t<true>(a, nw, v, st) : xpass_t<false>(a, nw, v, st);
}

}  // namespace dnsk
