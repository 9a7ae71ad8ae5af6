// Grouped z-forward launches of the ensemble (ens_plane.h), FP64 planes.
#include "ens_plane.h"

namespace dnsk {

int launch_ens_zfwd_f64(const EnsPlaneLaunch &E, bool vpl, int wg, int v, hipStream_t st) {
  return vpl ? launch_ens_plane_t<double, 4>(E, wg, v, st) : launch_ens_plane_t<double, 3>(E, wg, v, st);
}

}  // namespace dnsk
