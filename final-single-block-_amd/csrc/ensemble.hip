// Grouped launches of the ensemble form: channels of different sizes (one dns_handle per size, each
// holding its same-size members) advance together, one launch per kernel class per RK3 substep for all
// members of all sizes.  Each launch is a 1-D grid whose block ranges belong to the groups in turn; a
// block finds its group, takes its logical index inside the group's grid, and runs the body of the
// group's transform length (plane kernels) or row count (column pass) from plane_kernels.h /
// column_pass.hip.  Every member runs the reference's single-channel RK3 loop (littleharsh.f90:163-233).
//
// Tiles: 128-thread workgroups for the plane kernels at every length (z-kernels: 32, 16, 8 columns at
// 48, 96, 192 points; the register x-pass: 16, 8, 4 rows), so that one launch has one block size.  All
// tile variants of a kernel compute the same bits, so a member's result equals the same channel run alone.
#include "plane_launch.h"

namespace dnsk {

// ---- plane kernels -------------------------------------------------------------------------------
template <int N> struct EnsZ {   // z-kernel tile: ZTile variant with 128 threads
  static constexpr int V = N == 48 ? 2 : N == 96 ? 0 : 1;
  using Z = ZTile<N, V>;
  static_assert(Z::NTH == ENS_PLANE_NTH, "one block size per launch");
};
template <typename R, int N> using EZinv = ZinvBody<R, N, EnsZ<N>::Z::T, ENS_PLANE_NTH, typename ZPlanT<N>::type,
                                                    EnsZ<N>::Z::ZCH>;
template <typename R, int N, bool VPL>
using EZfwd = ZfwdBody<R, N, EnsZ<N>::Z::T, ENS_PLANE_NTH, VPL, typename ZPlanT<N>::type>;
// which: 0 z-inverse, 1/2 x-pass u/v-planes, 3/4 z-forward u/v-planes; body of length N
template <typename R, int WHICH, int N> struct EnsBody;
template <typename R, int N> struct EnsBody<R, 0, N> { using B = EZinv<R, N>; };
// x-pass: the register x-pass body of the single-channel launches, 128 threads
template <typename R, int N, bool VPL> using EX = XregBody<R, N, ENS_PLANE_NTH, VPL>;
template <typename R, int N> struct EnsBody<R, 1, N> { using B = EX<R, N, false>; };
template <typename R, int N> struct EnsBody<R, 2, N> { using B = EX<R, N, true>; };
template <typename R, int N> struct EnsBody<R, 3, N> { using B = EZfwd<R, N, false>; };
template <typename R, int N> struct EnsBody<R, 4, N> { using B = EZfwd<R, N, true>; };

template <typename R, int WHICH> constexpr size_t ens_plane_lds() {
  constexpr size_t a = EnsBody<R, WHICH, 48>::B::LDS, b = EnsBody<R, WHICH, 96>::B::LDS,
                   c = EnsBody<R, WHICH, 192>::B::LDS;
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

__device__ __forceinline__ int ens_group(const int *off, int ng, int lb) {
  int g = 0;
  while (g + 1 < ng && lb >= off[g + 1]) g++;
  return g;
}

template <typename R, int WHICH>
__global__ void __launch_bounds__(ENS_PLANE_NTH) ens_plane_kernel(EnsPlaneLaunch E) {
  __shared__ __align__(16) unsigned char lds[ens_plane_lds<R, WHICH>()];
  // the z-kernels' logical order is XCD-aware over the whole launch (each XCD runs a contiguous run of
  // logical blocks, so neighbouring tiles of a row share their edge lines in one L2)
  const int lbg = (WHICH == 1 || WHICH == 2) ? (int)blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
  const int g = ens_group(E.off, E.ng, lbg);
  const int lb = lbg - E.off[g];
  const PlaneArgs &a = E.a[g];
  switch (E.len[g]) {
    case 48: EnsBody<R, WHICH, 48>::B::run(a, lb, E.gx[g], E.gy[g], lds); break;
    case 96: EnsBody<R, WHICH, 96>::B::run(a, lb, E.gx[g], E.gy[g], lds); break;
    default: EnsBody<R, WHICH, 192>::B::run(a, lb, E.gx[g], E.gy[g], lds); break;
  }
}

bool ens_plane_length_supported(int n) { return n == 48 || n == 96 || n == 192; }

// grid of group g for kernel class which (the same grids as the single-channel launchers)
static void ens_plane_grid(const PlaneArgs &a, int which, int &gx, int &gy, int &nblk) {
  const int n = (which == 1 || which == 2) ? a.Mx : a.Mz;
  int T = 0, ZCH = 1, RB = 1;
  switch (n) {
    case 48: T = EnsZ<48>::Z::T; ZCH = EnsZ<48>::Z::ZCH; RB = EX<double, 48, false>::ROWS; break;
    case 96: T = EnsZ<96>::Z::T; ZCH = EnsZ<96>::Z::ZCH; RB = EX<double, 96, false>::ROWS; break;
    default: T = EnsZ<192>::Z::T; ZCH = EnsZ<192>::Z::ZCH; RB = EX<double, 192, false>::ROWS; break;
  }
  if (which == 0) {
    gx = (a.Kx + T - 1) / T;
    gy = 3;
    nblk = gx * gy * ((a.nbj + ZCH - 1) / ZCH) * a.nm;
  } else if (which <= 2) {
    gx = (a.Mz + RB - 1) / RB;
    gy = a.nbj * a.nm;
    nblk = gx * gy;
  } else {
    gx = (a.Kx + T - 1) / T;
    gy = a.nbj * a.nm;
    nblk = gx * gy;
  }
}

int ens_plane_rows_ok(int Mx, int Mz) {  // the register x-pass takes any Mz
  (void)Mx;
  (void)Mz;
  return 1;
}

int launch_ens_plane(EnsPlaneLaunch E, bool fp32, int which, hipStream_t st) {
  int total = 0;
  for (int g = 0; g < E.ng; g++) {
    const PlaneArgs &a = E.a[g];
    const int n = (which == 1 || which == 2) ? a.Mx : a.Mz;
    if (!ens_plane_length_supported(a.Mx) || !ens_plane_length_supported(a.Mz) || a.Kx != a.Mx / 3 + 1 ||
        a.Nz != 2 * a.Mz / 3 || !ens_plane_rows_ok(a.Mx, a.Mz))
      return -1;
    int nb = 0;
    ens_plane_grid(a, which, E.gx[g], E.gy[g], nb);
    E.len[g] = n;
    E.off[g] = total;
    total += nb;
  }
  E.off[E.ng] = total;
#define GO(R, W) ens_plane_kernel<R, W><<<total, ENS_PLANE_NTH, 0, st>>>(E)
  if (fp32) {
    switch (which) {
      case 0: GO(float, 0); break; case 1: GO(float, 1); break; case 2: GO(float, 2); break;
      case 3: GO(float, 3); break; default: GO(float, 4); break;
    }
  } else {
    switch (which) {
      case 0: GO(double, 0); break; case 1: GO(double, 1); break; case 2: GO(double, 2); break;
      case 3: GO(double, 3); break; default: GO(double, 4); break;
    }
  }
#undef GO
  return 0;
}

// ---- per-substep helpers of all groups ------------------------------------------------------------
// kRK = 1: every member's CFL maxima, max|div| and CFL slots to zero; kRK > 1: max|div| only.
__global__ void ens_reset_kernel(EnsSmall E, int kRK) {
  for (int g = 0; g < E.ng; g++) {
    const int nm = E.nm[g];
    for (int i = threadIdx.x; i < nm; i += blockDim.x) {
      if (kRK == 1) {
        E.st[g][i].cfl[0] = 0.0;
        E.st[g][i].cfl[1] = 0.0;
      }
      E.st[g][i].divmax = 0.0;
    }
    if (kRK == 1)
      for (int i = threadIdx.x; i < nm * 2 * CFL_SLOTS; i += blockDim.x) E.slots[g][i] = 0.0;
  }
}

// ghost plane 0 of u1/u3 into the CFL maxima (cfl_ghost_kernel for every member of every group)
__global__ void __launch_bounds__(256) ens_ghost_kernel(EnsSmall E) {
  __shared__ double red[4];
  const int g = ens_group(E.off, E.ng, blockIdx.x);
  const int lb = blockIdx.x - E.off[g];
  const int nbx = E.gx[g], bx = lb % nbx, m = lb / nbx;
  const int C = E.C[g];
  const double2 *u1 = E.u1[g] + (size_t)m * E.ms[g], *u3 = E.u3[g] + (size_t)m * E.ms[g];
  double *cfl = E.slots[g] + (size_t)m * 2 * CFL_SLOTS;
  double m1 = 0.0, m3 = 0.0;
  for (int c = bx * blockDim.x + threadIdx.x; c < C; c += nbx * blockDim.x) {
    const double2 a = u1[c], b = u3[c];
    m1 = fmax(m1, fmax(fabs(a.x), fabs(a.y)));
    m3 = fmax(m3, fmax(fabs(b.x), fabs(b.y)));
  }
  m1 = block_max<256>(m1, red);
  m3 = block_max<256>(m3, red);
  if (threadIdx.x == 0) {
    const int slot = bx & (CFL_SLOTS - 1);
    atomic_max_nonneg(&cfl[slot], m1);
    atomic_max_nonneg(&cfl[CFL_SLOTS + slot], m3);
  }
}

// max over each member's CFL slots (one 64-thread block per member of every group)
__global__ void ens_cfl_reduce_kernel(EnsSmall E) {
  int g = 0, m = blockIdx.x;
  while (g + 1 < E.ng && m >= E.nm[g]) m -= E.nm[g++];
  const double *slots = E.slots[g] + (size_t)m * 2 * CFL_SLOTS;
  const int l = threadIdx.x;
  double m1 = slots[l], m3 = slots[CFL_SLOTS + l];
  static_assert(CFL_SLOTS == 64, "one wavefront");
  for (int off = 32; off > 0; off >>= 1) {
    m1 = fmax(m1, __shfl_xor(m1, off, 64));
    m3 = fmax(m3, __shfl_xor(m3, off, 64));
  }
  if (l == 0) {
    E.st[g][m].cfl[0] = m1;
    E.st[g][m].cfl[1] = m3;
  }
}

void launch_ens_reset(const EnsSmall &E, int kRK, hipStream_t st) { ens_reset_kernel<<<1, 256, 0, st>>>(E, kRK); }

void launch_ens_ghost(EnsSmall E, hipStream_t st) {
  int total = 0;
  for (int g = 0; g < E.ng; g++) {
    int nb = (E.C[g] + 255) / 256;
    if (nb > 1024) nb = 1024;
    E.gx[g] = nb;
    E.off[g] = total;
    total += nb * E.nm[g];
  }
  E.off[E.ng] = total;
  ens_ghost_kernel<<<total, 256, 0, st>>>(E);
}

void launch_ens_cfl_reduce(const EnsSmall &E, hipStream_t st) {
  int total = 0;
  for (int g = 0; g < E.ng; g++) total += E.nm[g];
  ens_cfl_reduce_kernel<<<total, CFL_SLOTS, 0, st>>>(E);
}

}  // namespace dnsk
