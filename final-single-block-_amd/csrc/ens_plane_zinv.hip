// Grouped z-inverse launches of the ensemble (ens_plane.h), both precisions.
#include "ens_plane.h"

namespace dnsk {

int launch_ens_zinv(const EnsPlaneLaunch &E, bool fp32, int wg, int v, hipStream_t st) {
  return fp32 ? launch_ens_plane_t<float, 0>(E, wg, v, st) : launch_ens_plane_t<double, 0>(E, wg, v, st);
}

}  // namespace dnsk
