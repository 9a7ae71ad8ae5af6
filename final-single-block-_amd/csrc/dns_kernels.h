// Kernel argument blocks and launcher declarations (device code in plane_kernels.hip and
// column_kernels.hip, host code in dns_api.hip).
#pragma once
#include "dns_device.h"

#include <vector>

namespace dnsk {

// Device layout of every spectral field: plane-major, row j of the field at [row][column] with
// column = k*Kx + i (the reference's global column order).  Row offsets:
//   u1, u3, du1, du3, vvc : rows 0..nyu+1 at j*C        u2, du2, uvf, wvf : rows 0..nyv+1 at j*C
//   p, psi, div, Nu1, Nu3, Nxz1, Nxz3 : rows 1..nyu at (j-1)*C
//   Nu2, Nxz2 : rows 1..nyv at (j-1)*C
// An ensemble of nm members of one size stores the members one after another in every array (member
// strides in member_strides); kernels take the member from the last grid dimension.
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
  int nm, bjcap;                   // ensemble members; plane-index capacity of one member's I/P
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
  double Re, dtv;                  // this member's Reynolds number and viscous step limit
};

// Elements per member of the three row families (u-grid with ghosts, v-grid, p-grid).
struct MemberStrides {
  size_t ru, rv, rp;
};
__host__ __device__ __forceinline__ MemberStrides member_strides(int C, int nyu, int nyv) {
  return {(size_t)(nyu + 2) * C, (size_t)(nyv + 2) * C, (size_t)nyu * C};
}
// plane-stage intermediates of one member (complex elements of the plane precision)
__host__ __device__ __forceinline__ size_t member_istride(int bjcap, int Mz, int Kp) {
  return (size_t)2 * bjcap * 3 * Mz * Kp;
}
__host__ __device__ __forceinline__ size_t member_pstride(int bjcap, int Mz, int Kp) {
  return (size_t)2 * bjcap * 4 * Mz * Kp;
}

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
  int nm;                          // ensemble members (grid dimension y)
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

// Member m's view of the arguments: every array pointer moved to the member's part.
template <typename R>
__device__ __forceinline__ PlaneArgs plane_view(PlaneArgs a, int m) {
  if (m == 0) return a;
  const MemberStrides s = member_strides(a.C, a.nyu, a.nyv);
  const int Kp = row_pitch<R>(a.Kx);
  a.u1 += m * s.ru; a.u3 += m * s.ru; a.u2 += m * s.rv;
  a.I = reinterpret_cast<cx<R> *>(a.I) + m * member_istride(a.bjcap, a.Mz, Kp);
  a.P = reinterpret_cast<cx<R> *>(a.P) + m * member_pstride(a.bjcap, a.Mz, Kp);
  a.Nxz1 = reinterpret_cast<cx<R> *>(a.Nxz1) + m * s.rp;
  a.Nxz2 = reinterpret_cast<cx<R> *>(a.Nxz2) + m * s.rp;
  a.Nxz3 = reinterpret_cast<cx<R> *>(a.Nxz3) + m * s.rp;
  a.uvf = reinterpret_cast<cx<R> *>(a.uvf) + m * s.rv;
  a.wvf = reinterpret_cast<cx<R> *>(a.wvf) + m * s.rv;
  a.vvc = reinterpret_cast<cx<R> *>(a.vvc) + m * s.ru;
  a.cfl += (size_t)m * 2 * CFL_SLOTS;
  return a;
}

__device__ __forceinline__ NArr narr_at(NArr x, size_t off) {
  x.p = x.f32 ? (void *)(static_cast<float2 *>(x.p) + off) : (void *)(static_cast<double2 *>(x.p) + off);
  return x;
}

// Member m's view of the column arguments; its Re and dtv come from its device state.
__device__ __forceinline__ ColArgs col_view(ColArgs a, int m) {
  const MemberStrides s = member_strides(a.C, a.nyu, a.nyv);
  if (m > 0) {
    a.u1 += m * s.ru; a.u3 += m * s.ru; a.du1 += m * s.ru; a.du3 += m * s.ru; a.w3u += m * s.ru;
    a.u2 += m * s.rv; a.du2 += m * s.rv; a.w3v += m * s.rv;
    a.p += m * s.rp; a.psi += m * s.rp; a.div += m * s.rp; a.cP += m * s.rp;
    a.Nu1 = narr_at(a.Nu1, m * s.rp); a.Nu2 = narr_at(a.Nu2, m * s.rp); a.Nu3 = narr_at(a.Nu3, m * s.rp);
    a.Nxz1 = narr_at(a.Nxz1, m * s.rp); a.Nxz2 = narr_at(a.Nxz2, m * s.rp); a.Nxz3 = narr_at(a.Nxz3, m * s.rp);
    a.uvf = narr_at(a.uvf, m * s.rv); a.wvf = narr_at(a.wvf, m * s.rv); a.vvc = narr_at(a.vvc, m * s.ru);
    a.mscr += (size_t)m * 2 * (a.nyu + 2);
    a.st += m;
  }
  a.Re = a.st->Re;
  a.dtv = a.st->dtv;
  return a;
}

bool plane_size_supported(int n);
int xpass_rows_per_block(int Mx);  // rows an x-pass launch needs Mz to be a multiple of (1: any Mz)
int plane_variants(int n);          // tile variants instantiated for transform length n
// which: 0 z-inverse (nw = nbj + 1 field rows), 1/2 x-pass u/v-planes, 3/4 z-forward u/v-planes (nw = nbj)
int launch_plane_kernel(const PlaneArgs &a, int nw, bool fp32, int which, int variant, hipStream_t st);
// fills the second half of the x-pass twiddle table twh (2 * Mx/2 entries, plane precision): the
// register x-pass's lane twiddles
int build_xpass_twiddles(void *twh, int Mx, bool fp32, hipStream_t st);
void launch_cfl_reduce(const double *slots, DnsState *out, int nm, hipStream_t st);
void launch_cfl_reset(DnsState *st, double *slots, int nm, hipStream_t s);
void launch_widen(const float2 *in, double2 *out, size_t n, hipStream_t st);
void launch_narrow(const double2 *in, float2 *out, size_t n, hipStream_t st);
void launch_cfl_ghost(const double2 *u1, const double2 *u3, int C, size_t ms, int nm, double *cfl, hipStream_t st);
void launch_zero(double *p, int n, hipStream_t st);
void launch_divmax_reset(DnsState *st, int nm, hipStream_t s);  // every member's max|div| = 0
int launch_plane_roundtrip(int Mx, int Mz, const double *in, double *phys, double *back, int n,
                           const double2 *twx, const double2 *twh, const double2 *twz, hipStream_t st);

// staged path (one kernel per reference subroutine)
void launch_rhs0(const ColArgs &a, hipStream_t st);
void launch_nl_col(const ColArgs &a, hipStream_t st);
void launch_solveU(const ColArgs &a, hipStream_t st);
void launch_meanflow(const ColArgs &a, hipStream_t st);
// meanflow_ctU after the one-pass column stage (solveP and v_corr already done); -1 when nyu > 512
int launch_meanflow_post(const ColArgs &a, hipStream_t st);
void launch_solveP(const ColArgs &a, hipStream_t st);
void launch_vcorr(const ColArgs &a, hipStream_t st);
void launch_qx(const ColArgs &a, hipStream_t st);
// fused path
void launch_fused_up(const ColArgs &a, int variant, hipStream_t st);  // 0 thread/column, 1-2 wave/solve
void launch_fused_down(const ColArgs &a, int variant, hipStream_t st);   // 0 thread/column, 1-2 four waves
void launch_fused_final(const ColArgs &a, int variant, hipStream_t st);  // 0 thread/column, 1 + prefetch, 2 four waves
// statistics and spectra (stats_kernels.hip)
constexpr int STAT_NQ = 13;       // Um U2m Wm W2m UVm UWm Vm V2m VWm wxm wx2m Pm P2m

struct StatsArgs {
  int Kx, Nx, Nz, Mx, Mz, C, nyu, nyv, nm, nchunk;
  double bet, gi1, gi2;
  const double *yu, *yv, *dthdyv;
  const double2 *u1, *u2, *u3, *p;
  const double2 *pm;  // pressure summed into Pm, P2m: p, or the copy ppPL holds (dns_refresh_stats_pressure)
  double *partial;  // [m][j][chunk][q]
  double *acc;      // [m][q][j], j = 0..nyu+1
  double *sp;       // spectra [m]: spU (nyu+2 rows), spV (nyv+2), spW (nyu+2), spUV (nyu+2), spP (nyu), rows x C
};

int stats_nchunk(int C);
void launch_stats(const StatsArgs &a, hipStream_t st);

// four_to_phys post-processing (post_kernels.hip): n spectral planes (Nx+2) x Nz -> physical (Mx+2) x Mz
struct FtpArgs {
  int Nx, Nz, Mx, Mz, n;
  int in_pitch, mid_pitch, out_pitch;  // row pitches (doubles, complex, doubles): whole 128-byte lines
  const double *in;  // [plane][Nz][in_pitch]
  double2 *mid;      // z-inverted columns [plane][Mz][mid_pitch]
  double *out;       // [plane][Mz][out_pitch]
  const double2 *twz, *twx, *twh;  // e^{-2 pi i m/N} for N = Mz, Mx, Mx/2
};
bool four_to_phys_supported(int Mx, int Mz);
int launch_four_to_phys(const FtpArgs &a, hipStream_t st);
// the allocation-free entry point's workspace (post_kernels.hip): a static device array of this size;
// *ptr = its address on the current device (0), or -1 when the address cannot be had
constexpr size_t FTP_SCRATCH_BYTES = (size_t)16 << 20;
int four_to_phys_scratch(void **ptr);

// one-pass column stage (column_pass.hip): 8 <= nyu <= 512 (with flag_ctpress = 0 followed by
// launch_meanflow_post)
bool column_pass_supported(int nyu);
int column_pass_variants(int nyu);  // workgroup shapes of nyu's bucket (same arithmetic, same bits)
int launch_column_pass(const ColArgs &a, int variant, hipStream_t st);  // 0, or 1 when nyu is not covered
// ---- ensemble of several sizes: grouped launches (ens_plane*.hip, column_pass.hip, ensemble.hip) ----
constexpr int ENS_MAXG = 4;          // size groups per ensemble
constexpr int ENS_PLANE_NTH = 128;   // threads of a plane body of the persistent launch
constexpr int ENS_NVAR = 3;          // tile variants per plane class and length, per column row bucket
// one grouped plane launch: group g's arguments; the launcher fills the transform length of the class,
// the logical grid (gx, gy; nblk blocks), the group's workgroups (nwg) and its first workgroup (off)
struct EnsPlaneLaunch {
  PlaneArgs a[ENS_MAXG];
  int len[ENS_MAXG], gx[ENS_MAXG], gy[ENS_MAXG], nblk[ENS_MAXG], nwg[ENS_MAXG], off[ENS_MAXG + 1];
  int ng;
};
// the small per-substep kernels of every group
struct EnsSmall {
  DnsState *st[ENS_MAXG];
  double *slots[ENS_MAXG];
  const double2 *u1[ENS_MAXG], *u3[ENS_MAXG];
  size_t ms[ENS_MAXG];               // member stride of u1/u3
  int nm[ENS_MAXG], C[ENS_MAXG], gx[ENS_MAXG], off[ENS_MAXG + 1];
  int ng;
};
// one grouped column-pass launch (shape by row count, up to 160 rows, and the launch's variant; 256-thread
// workgroups); the launcher fills nblk and off
struct EnsColLaunch {
  ColArgs a[ENS_MAXG];
  int nblk[ENS_MAXG], off[ENS_MAXG + 1];
  int ng;
};
bool ens_plane_length_supported(int n);
int ens_plane_rows_ok(int Mx, int Mz);
// which: 0 z-inverse, 1/2 x-pass u/v-planes, 3/4 z-forward u/v-planes; every group's j0 = 1, nbj = nyu;
// tile shape: wg = 128 or 256 threads per workgroup, bodies of 64 << v threads.  -1 when a group's size
// or the shape is not supported.
int launch_ens_plane(EnsPlaneLaunch E, bool fp32, int which, int wg, int v, hipStream_t st);
// whether the group with arguments a (j0 = 1, nbj = nyu) takes shape (wg, v) of class which
bool ens_plane_variant_ok(const PlaneArgs &a, int which, int v, int wg);
void launch_ens_reset(const EnsSmall &E, int kRK, hipStream_t st);
void launch_ens_ghost(EnsSmall E, hipStream_t st);
void launch_ens_cfl_reduce(const EnsSmall &E, hipStream_t st);
bool ens_column_rows_supported(int nyu);
int launch_ens_column_pass(EnsColLaunch E, bool fp32, int v, hipStream_t st);  // v: variant 0..2
// meanflow_ctU after the column pass for the groups with a constant flow rate (flags[g] = 1)
void launch_ens_meanflow_post(const EnsColLaunch &E, const int *flags, hipStream_t st);

// ---- persistent launch of the ensemble (ensemble.hip): one launch per RK3 substep for every member of
// every group.  Its workgroups take tasks from a queue in device memory; a task is a run of logical
// blocks of one kernel class for one member, and waits for the tasks it depends on through per-member
// completion counters.
constexpr int PERSIST_NTH = 256;   // threads per workgroup: two 128-thread plane bodies side by side
// task classes, in queue order
enum : int { PC_RESET = 0, PC_GHOST, PC_ZINV, PC_XU, PC_XV, PC_CFLR, PC_ZU, PC_ZV, PC_COL, PC_MF, PC_N };
// Units u0..u1-1 of class cls for member m of group g (plane classes: a unit is two logical blocks of
// the member's grid, one per body; column stage: one workgroup of the group's column-pass shape); it may
// start once counter dep[i] has reached need[i] (dep[i] < 0: none; the dependencies fill dep[] from 0) and
// adds one to counter done; rel = 1 when a later task of the launch waits on that counter.
struct PTask {
  int cls, g, m, u0, u1, done;
  int dep[3], need[3];
  int rel;
};
// ctr[]: queue head, workgroups finished, dependency-wait timeout flag, then the completion counters
constexpr int PCTR_HEAD = 0, PCTR_EXIT = 1, PCTR_ERR = 2, PCTR0 = 4;
struct PersistGroup {
  int gxz, gxx, gxf;   // tiles per row of the z-inverse, x-pass and z-forward grids
  int nbz, nbx, nbf;   // their logical blocks per member
  int ncb, cshape;     // column blocks per member, column-pass shape (rows bucket 32, 64, 96, 128, 160)
  int ctp0;            // 1: constant flow rate (meanflow_ctU after the column pass)
};
struct PersistArgs {
  PlaneArgs pa[ENS_MAXG];
  ColArgs ca[ENS_MAXG];
  PersistGroup q[ENS_MAXG];
  const PTask *task;
  unsigned *ctr;
  int ntask, nctr, kRK, ng;   // tasks, completion counters
};
// Grids of group g (its PlaneArgs with j0 = 1, nbj = nyu, and its column rows); -1 when not supported.
int ens_persist_group(const PlaneArgs &a, int ctp0, PersistGroup &q);
// The queue of one substep: tasks in order and the number of completion counters; every dependency is
// checked to be met by earlier tasks (so the queue cannot deadlock).  units[c]: units per task of class c.
int ens_persist_tasks(const PersistGroup *q, const int *nm, int ng, int kRK, const int *units,
                      std::vector<PTask> &tasks, int &ncounters);
int launch_ens_persist(const PersistArgs *E, bool fp32, int nwg, hipStream_t st);  // E in device memory
int ens_persist_max_blocks(bool fp32);   // workgroups of the persistent kernel resident per CU

// layout
void launch_ref_to_dev(const double2 *ref, double2 *dev, int C, int nrows, hipStream_t st);
void launch_dev_to_ref(const double2 *dev, double2 *ref, int C, int nrows, hipStream_t st);
void launch_modes_to_planes(const double2 *refcols, double *planes, int C, int nrows, int Kx, int Nz, int Mx,
                            int Mz, hipStream_t st);
void launch_planes_to_modes(const double *planes, double2 *refcols, int C, int nrows, int Kx, int Nz, int Mx,
                            int Mz, hipStream_t st);

}  // namespace dnsk
