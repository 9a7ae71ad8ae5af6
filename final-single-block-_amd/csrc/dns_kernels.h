// Kernel argument blocks and launcher declarations (device code in plane_kernels.hip and
// column_kernels.hip, host code in dns_api.hip).
#pragma once
#include "dns_device.h"

namespace dnsk {

// Device layout of every spectral field: plane-major, row j of the field at [row][column] with
// column = k*Kx + i (the reference's global column order).  Row offsets:
//   u1, u3, du1, du3, vvc : rows 0..nyu+1 at j*C        u2, du2, uvf, wvf : rows 0..nyv+1 at j*C
//   p, psi, div, Nu1, Nu3, Nxz1, Nxz3 : rows 1..nyu at (j-1)*C
//   Nu2, Nxz2 : rows 1..nyv at (j-1)*C
constexpr int CFL_SLOTS = 64;  // addresses the CFL maxima are spread over

// Row pitch (complex elements of precision R) of the plane intermediates I and P: Kx rounded up to
// whole 128-byte lines, so that the tile segments of a row start on a line.
template <typename R> __host__ __device__ constexpr int row_pitch(int kx) {
  constexpr int e = (int)(128 / (2 * sizeof(R)));
  return (kx + e - 1) / e * e;
}

struct PlaneArgs {
  int Mx, Mz, Nx, Nz, Kx, C, nyu, nyv;
  int j0, nbj;                     // planes j0..j0+nbj-1 (of both kinds) in this batch
  int kRK;
  double alp, bet;
  const double2 *u1, *u2, *u3;
  const double *yu, *yv;
  void *I, *P;                     // plane intermediates, precision of the plane stage
  const void *twz, *twx, *twh;     // twiddles e^{-2 pi i m/N} for N = Mz, Mx, Mx/2
  void *Nxz1, *Nxz2, *Nxz3, *uvf, *vvc, *wvf;  // complex of the plane precision
  double *cfl;                     // CFL_SLOTS partial max|u1|, then CFL_SLOTS partial max|u3|
};

struct DnsState {
  double t, dt, dti, mpgx, Qx, dgx, dtc, QxT;
  double cfl[2];
  double divmax;
  double iter;
  double nonfinite;                // set to 1 when a CFL maximum is NaN or Inf
  double pad;
};

// Complex array of nonlinear terms or fluxes: FP64, or FP32 storage in mixed precision (f32 is uniform,
// so the test costs a scalar branch).
struct NArr {
  void *p;
  int f32;
  __device__ __forceinline__ double2 operator[](size_t i) const {
    if (f32) {
      const float2 v = static_cast<const float2 *>(p)[i];
      double2 r;
      r.x = v.x;
      r.y = v.y;
      return r;
    }
    return static_cast<const double2 *>(p)[i];
  }
  __device__ __forceinline__ void set(size_t i, double2 v) const {
    if (f32) {
      float2 w;
      w.x = (float)v.x;
      w.y = (float)v.y;
      static_cast<float2 *>(p)[i] = w;
    } else {
      static_cast<double2 *>(p)[i] = v;
    }
  }
};

struct ColArgs {
  int Kx, Nx, Nz, C, nyu, nyv, kRK, flag_ctpress;
  int fixed;                       // 1: dt is fixed (no CFL update)
  double Re, alp, bet, CFL, dtv, fixed_dt;
  double aRK, bRK, gRK, cRK, dRK;
  double gw1, gw2, gbc, gi1, gi2, dyp;   // dyp = yu(2) - yu(1) (pressure pin row)
  const double *yu, *yv, *dthdyu, *dthdyv, *dyu2i, *dyv2i;
  double2 *u1, *u2, *u3, *p, *psi, *div;
  double2 *du1, *du2, *du3;
  NArr Nu1, Nu2, Nu3;              // nonlinear terms of the substep (old ones on entry)
  NArr Nxz1, Nxz2, Nxz3, uvf, vvc, wvf;  // plane-stage outputs
  double *w3u, *w3v, *cP;          // per-row scratch of the eliminations
  double *mscr;                    // 2*(nyu+2) doubles for the mean-mode solve
  DnsState *st;
};

bool plane_size_supported(int n);
int xpass_rows_per_block(int Mx);  // rows of one x-pass block (largest variant); must divide Mz
int plane_variants(int n);          // tile variants instantiated for transform length n
// which: 0 z-inverse (nw = nbj + 1 field rows), 1/2 x-pass u/v-planes, 3/4 z-forward u/v-planes (nw = nbj)
int launch_plane_kernel(const PlaneArgs &a, int nw, bool fp32, int which, int variant, hipStream_t st);
void launch_cfl_reduce(const double *slots, double *out, hipStream_t st);
void launch_widen(const float2 *in, double2 *out, size_t n, hipStream_t st);
void launch_narrow(const double2 *in, float2 *out, size_t n, hipStream_t st);
void launch_cfl_ghost(const double2 *u1, const double2 *u3, int C, double *cfl, hipStream_t st);
void launch_zero(double *p, int n, hipStream_t st);
int launch_plane_roundtrip(int Mx, int Mz, const double *in, double *phys, double *back, int n,
                           const double2 *twx, const double2 *twh, const double2 *twz, hipStream_t st);

// staged path (one kernel per reference subroutine)
void launch_rhs0(const ColArgs &a, hipStream_t st);
void launch_nl_col(const ColArgs &a, hipStream_t st);
void launch_solveU(const ColArgs &a, hipStream_t st);
void launch_meanflow(const ColArgs &a, hipStream_t st);
void launch_solveP(const ColArgs &a, hipStream_t st);
void launch_vcorr(const ColArgs &a, hipStream_t st);
void launch_qx(const ColArgs &a, hipStream_t st);
// fused path
void launch_fused_up(const ColArgs &a, int variant, hipStream_t st);  // 0 thread/column, 1-2 wave/solve
void launch_fused_down(const ColArgs &a, int variant, hipStream_t st);   // 0 thread/column, 1-2 four waves
void launch_fused_final(const ColArgs &a, int variant, hipStream_t st);  // 0 thread/column, 1-2 four waves
// layout
void launch_ref_to_dev(const double2 *ref, double2 *dev, int C, int nrows, hipStream_t st);
void launch_dev_to_ref(const double2 *dev, double2 *ref, int C, int nrows, hipStream_t st);
void launch_modes_to_planes(const double2 *refcols, double *planes, int C, int nrows, int Kx, int Nz, int Mx,
                            int Mz, hipStream_t st);
void launch_planes_to_modes(const double *planes, double2 *refcols, int C, int nrows, int Kx, int Nz, int Mx,
                            int Mz, hipStream_t st);

}  // namespace dnsk
