// One-pass column stage: kernels and launchers of the body in column_pass.h (single channel, and the
// ensemble's grouped launch).
#include "column_pass.h"
#include "column_seq.h"

namespace dnsk {

// The sequential-systems form (column_seq.h): 512 threads per CU (two workgroups of 256).
template <int NC, int P, int M, typename NT>
__global__ void __launch_bounds__(NC * P, NC * P >= 512 ? 1 : 512 / (NC * P)) column_seq_kernel(ColArgs a) {
  using B = ColumnSeqBody<NC, P, M, NT>;
  __shared__ __align__(16) unsigned char lds[B::LDS];
  B::run(a, blockIdx.x, gridDim.x, blockIdx.y, threadIdx.x, lds);
}
template <int NC, int P, int M>
static void launch_seq(const ColArgs &a, hipStream_t st) {
  const dim3 grid((a.C + NC - 1) / NC, a.nm);
  if (a.Nu1.f32) column_seq_kernel<NC, P, M, float><<<grid, NC * P, 0, st>>>(a);
  else column_seq_kernel<NC, P, M, double><<<grid, NC * P, 0, st>>>(a);
}

template <int NC, int P, int M, typename NT>
__global__ void __launch_bounds__(NC * P, M <= 4 ? 2 : 1) column_pass_kernel(ColArgs a) {
  using B = ColumnPassBody<NC, P, M, NT>;
  __shared__ __align__(16) unsigned char lds[B::LDS];
  B::run(a, blockIdx.x, gridDim.x, blockIdx.y, threadIdx.x, lds);
}

// (NC, P, M) by the rows to cover (nyu <= P*M, M >= 3: a segment needs a first, an interior and a last
// row).  With both velocity systems live, M = 4 keeps a thread within 256 registers (two wavefronts per
// SIMD) and longer segments spill; the sequential form (column_seq.h) holds one system at a time, which
// lets 257-point grids take 8-row segments in 256-thread workgroups, two per CU.  Each bucket's shape
// was timed against alternatives (3-row segments at 96 rows; at the 153-point grids 32 x 5 rows, in the
// sequential form 0.135 ms against 0.170 for 16 simultaneous columns at 128 x 128; at 257 points 8 x 32 x 8
// sequential against 8 x 64 x 4: 1.61 against 1.80 ms FP64 at 384 x 384).
__host__ __device__ int column_pass_rows(int nyu) {
  return nyu <= 32 ? 32 : nyu <= 64 ? 64 : nyu <= 96 ? 96 : nyu <= 128 ? 128 : nyu <= 160 ? 160
         : nyu <= 192 ? 192 : nyu <= 256 ? 256 : nyu <= 384 ? 384 : nyu <= 512 ? 512 : 0;
}

bool column_pass_supported(int nyu) { return nyu >= 8 && column_pass_rows(nyu) > 0; }

template <int NC, int P, int M>
static void launch_pass(const ColArgs &a, hipStream_t st) {
  const dim3 grid((a.C + NC - 1) / NC, a.nm);
  if (a.Nu1.f32) column_pass_kernel<NC, P, M, float><<<grid, NC * P, 0, st>>>(a);
  else column_pass_kernel<NC, P, M, double><<<grid, NC * P, 0, st>>>(a);
}

// Shapes of a bucket for the autotuner: the same P and M (so the same arithmetic per column, the same
// bits) with NC = the default, then the other two columns-per-workgroup counts that keep NC * P within
// 64..512 threads (at 385 points: the sequential form and two simultaneous shapes).
int column_pass_variants(int nyu) {
  switch (column_pass_rows(nyu)) {
    case 32: case 64: case 96: case 128: case 160: case 256: case 384: return 3;
    case 192: case 512: return 2;
    default: return 0;
  }
}

int launch_column_pass(const ColArgs &a, int v, hipStream_t st) {
  switch (column_pass_rows(a.nyu)) {
    case 32: if (v == 1) launch_pass<16, 8, 4>(a, st); else if (v == 2) launch_pass<32, 8, 4>(a, st);
             else launch_pass<8, 8, 4>(a, st); break;
    case 64: if (v == 1) launch_pass<16, 16, 4>(a, st); else if (v == 2) launch_pass<4, 16, 4>(a, st);
             else launch_pass<8, 16, 4>(a, st); break;
    case 96: if (v == 1) launch_pass<16, 32, 3>(a, st); else if (v == 2) launch_pass<4, 32, 3>(a, st);
             else launch_pass<8, 32, 3>(a, st); break;
    case 128: if (v == 1) launch_pass<16, 32, 4>(a, st); else if (v == 2) launch_pass<4, 32, 4>(a, st);
              else launch_pass<8, 32, 4>(a, st); break;
    case 160: if (v == 1) launch_pass<16, 32, 5>(a, st); else if (v == 2) launch_seq<16, 32, 5>(a, st);
              else launch_seq<8, 32, 5>(a, st); break;  // 153-point grids: 32 x 5 rows instead of 64 x 4
    case 192: if (v == 1) launch_pass<4, 64, 3>(a, st); else launch_pass<8, 64, 3>(a, st); break;
    // 257-point grids: the three systems one after another, 32 segments of 8 rows (2 workgroups per CU)
    case 256: if (v == 1) launch_seq<16, 32, 8>(a, st); else if (v == 2) launch_seq<4, 32, 8>(a, st);
              else launch_seq<8, 32, 8>(a, st); break;
    // 385-point grids: the sequential form (one system's segment live at a time: 16 spilled registers
    // instead of 140) and the simultaneous one, same segments, same bits
    case 384: if (v == 1) launch_pass<8, 64, 6>(a, st); else if (v == 2) launch_pass<4, 64, 6>(a, st);
              else launch_seq<8, 64, 6>(a, st); break;
    case 512: if (v == 1) launch_pass<2, 64, 8>(a, st); else launch_pass<4, 64, 8>(a, st); break;
    default: return 1;
  }
  return 0;
}

// ---- grouped launch of the ensemble: every member of every size group in one launch -----------------
// A 256-thread workgroup (two per CU) runs 256/NTH bodies of the group's shape side by side (each with
// its own slice of LDS and its own column block), so every member gets exactly the arithmetic of the
// single-channel launch.  Shapes by row bucket (8 <= nyu <= 160), three variants with the same segments
// (P, M) in each bucket, so the same bits; a launch takes one variant for all groups (each variant its
// own kernel, with the registers of its own bodies), chosen by the ensemble autotuner (0 by default):
//    32 rows:  8, 16 or 32 columns x 8 segments x 4 rows        (64, 128, 256 threads)
//    64 rows:  8, 16 or 4 columns x 16 x 4
//    96 rows:  8, 4 or 2 columns x 32 x 3
//   128 rows:  8 or 4 columns x 32 x 4, or 8 x 32 x 4 in the sequential form
//   160 rows:  8 or 4 columns x 32 x 5 in the sequential form, or 8 x 32 x 5 simultaneous
// 512-thread workgroups (one per CU) took 1.32 against 1.24 ms per substep at the 48-member configuration
// (FP64).
constexpr int ENS_COL_NTH = 256;
template <int ROWS, int V, typename NT> struct EnsColB;
template <int V, typename NT> struct EnsColB<32, V, NT> { using B = ColumnPassBody<8 << V, 8, 4, NT>; };
template <int V, typename NT> struct EnsColB<64, V, NT> {
  using B = ColumnPassBody<V == 0 ? 8 : V == 1 ? 16 : 4, 16, 4, NT>;
};
template <int V, typename NT> struct EnsColB<96, V, NT> { using B = ColumnPassBody<(8 >> V), 32, 3, NT>; };
template <int V, typename NT> struct EnsColB<128, V, NT> {
  using B = typename std::conditional<V == 2, ColumnSeqBody<8, 32, 4, NT>, ColumnPassBody<(8 >> V), 32, 4, NT>>::type;
};
template <int V, typename NT> struct EnsColB<160, V, NT> {
  using B = typename std::conditional<V == 2, ColumnPassBody<8, 32, 5, NT>, ColumnSeqBody<(8 >> V), 32, 5, NT>>::type;
};
template <int ROWS, int V, typename NT> struct EnsColShape {
  using B = typename EnsColB<ROWS, V, NT>::B;
  static_assert(ENS_COL_NTH % B::NTH == 0, "whole bodies per workgroup");
  static constexpr int SUB = ENS_COL_NTH / B::NTH;   // bodies per workgroup
  static constexpr size_t LDS = SUB * ((B::LDS + 15) / 16 * 16);
};
template <int V, typename NT> constexpr size_t ens_col_lds() {
  size_t m = 0;
  for (size_t v : {EnsColShape<32, V, NT>::LDS, EnsColShape<64, V, NT>::LDS, EnsColShape<96, V, NT>::LDS,
                   EnsColShape<128, V, NT>::LDS, EnsColShape<160, V, NT>::LDS})
    m = v > m ? v : m;
  return m;
}
static int ens_col_nc(int rows, int v) {  // columns per body
  switch (rows) {
    case 32: return 8 << v;
    case 64: return v == 0 ? 8 : v == 1 ? 16 : 4;
    case 96: return 8 >> v;
    default: return v == 1 ? 4 : 8;
  }
}
static int ens_col_nth(int rows, int v) { return ens_col_nc(rows, v) * (rows == 32 ? 8 : rows == 64 ? 16 : 32); }

bool ens_column_rows_supported(int nyu) { return nyu >= 8 && nyu <= 160; }

template <int ROWS, int V, typename NT>
__device__ __forceinline__ void ens_col_run(const ColArgs &a, int lb, int nb, unsigned char *lds) {
  using S = EnsColShape<ROWS, V, NT>;
  using B = typename S::B;
  const int sub = threadIdx.x / B::NTH, tid = threadIdx.x % B::NTH;
  // workgroup lb holds column blocks lb*SUB .. lb*SUB + SUB-1 of the group (member-major); a body past the
  // group's last block computes the last block again (every barrier is reached) and stores nothing
  const int gb = lb * S::SUB + sub;
  const int total = nb * a.nm;
  const bool active = gb < total;
  const int b = active ? gb : total - 1;
  B::run(a, b % nb, nb, b / nb, tid, lds + sub * ((B::LDS + 15) / 16 * 16), active);
}

template <int V, typename NT>
__global__ void __launch_bounds__(ENS_COL_NTH, 2) ens_column_pass_kernel(EnsColLaunch E) {
  __shared__ __align__(16) unsigned char lds[ens_col_lds<V, NT>()];
  int g = 0;
  const int b = blockIdx.x;
  while (g + 1 < E.ng && b >= E.off[g + 1]) g++;
  const int lb = b - E.off[g], nb = E.nblk[g];
  switch (column_pass_rows(E.a[g].nyu)) {
    case 32: ens_col_run<32, V, NT>(E.a[g], lb, nb, lds); break;
    case 64: ens_col_run<64, V, NT>(E.a[g], lb, nb, lds); break;
    case 96: ens_col_run<96, V, NT>(E.a[g], lb, nb, lds); break;
    case 128: ens_col_run<128, V, NT>(E.a[g], lb, nb, lds); break;
    default: ens_col_run<160, V, NT>(E.a[g], lb, nb, lds); break;
  }
}

int launch_ens_column_pass(EnsColLaunch E, bool fp32, int v, hipStream_t st) {
  if (v < 0 || v >= ENS_NVAR) return -1;
  int total = 0;
  for (int g = 0; g < E.ng; g++) {
    const int nyu = E.a[g].nyu;
    if (!ens_column_rows_supported(nyu)) return -1;
    const int rows = column_pass_rows(nyu);
    const int nc = ens_col_nc(rows, v);
    const int sub = ENS_COL_NTH / ens_col_nth(rows, v);                   // bodies per workgroup
    E.nblk[g] = (E.a[g].C + nc - 1) / nc;                                  // column blocks per member
    E.off[g] = total;
    total += (E.nblk[g] * E.a[g].nm + sub - 1) / sub;
  }
  E.off[E.ng] = total;
#define GO(V)                                                           \
  if (fp32) ens_column_pass_kernel<V, float><<<total, ENS_COL_NTH, 0, st>>>(E); \
  else ens_column_pass_kernel<V, double><<<total, ENS_COL_NTH, 0, st>>>(E);
  if (v == 0) { GO(0) }
  else if (v == 1) { GO(1) }
  else { GO(2) }
#undef GO
  return 0;
}

}  // namespace dnsk
