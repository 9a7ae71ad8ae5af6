// z-inverse kernels (zinv_kernel) of every supported length, both precisions.
#include "plane_launch.h"

namespace dnsk {

template <typename R>
static int zinv_t(const PlaneArgs &a, int nw, int v, hipStream_t st) {
#define CASE(N) if (a.Mz == N) { zinv_sized<R, N>(a, nw, v, st); return 0; }
  DNS_SIZES(CASE)
#undef CASE
  return -1;
}

int launch_zinv(const PlaneArgs &a, int nw, bool fp32, int v, hipStream_t st) {
  return fp32 ? zinv_t<float>(a, nw, v, st) : zinv_t<double>(a, nw, v, st);
}

}  // namespace dnsk
