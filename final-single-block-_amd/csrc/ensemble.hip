// Ensemble form: the grouped launches' dispatch and small per-substep kernels, and the persistent launch.
// The grouped plane kernels are in ens_plane.h (instantiated per class in ens_plane_*.hip), the grouped
// column pass in column_pass.hip.
#include "ens_plane.h"
#include "column_pass.h"
#include "column_seq.h"

namespace dnsk {

// per-class launchers (ens_plane_zinv.hip, ens_plane_xpass_f64/f32.hip, ens_plane_zfwd_f64/f32.hip)
int launch_ens_zinv(const EnsPlaneLaunch &E, bool fp32, int wg, int v, hipStream_t st);
int launch_ens_xpass_f64(const EnsPlaneLaunch &E, bool vpl, int wg, int v, hipStream_t st);
int launch_ens_xpass_f32(const EnsPlaneLaunch &E, bool vpl, int wg, int v, hipStream_t st);
int launch_ens_zfwd_f64(const EnsPlaneLaunch &E, bool vpl, int wg, int v, hipStream_t st);
int launch_ens_zfwd_f32(const EnsPlaneLaunch &E, bool vpl, int wg, int v, hipStream_t st);

bool ens_plane_length_supported(int n) { return n == 48 || n == 96 || n == 192; }

int ens_plane_rows_ok(int Mx, int Mz) {  // the register x-pass takes any Mz
  (void)Mx;
  (void)Mz;
  return 1;
}

bool ens_plane_variant_ok(const PlaneArgs &a, int which, int v, int wg) { return ens_plane_fits(a, which, v, wg); }

int launch_ens_plane(EnsPlaneLaunch E, bool fp32, int which, int wg, int v, hipStream_t st) {
  for (int g = 0; g < E.ng; g++) {
    const PlaneArgs &a = E.a[g];
    if (!ens_plane_length_supported(a.Mx) || !ens_plane_length_supported(a.Mz) || a.Kx != a.Mx / 3 + 1 ||
        a.Nz != 2 * a.Mz / 3 || !ens_plane_rows_ok(a.Mx, a.Mz))
      return -1;
  }
  switch (which) {
    case 0: return launch_ens_zinv(E, fp32, wg, v, st);
    case 1: case 2:
      return fp32 ? launch_ens_xpass_f32(E, which == 2, wg, v, st) : launch_ens_xpass_f64(E, which == 2, wg, v, st);
    case 3: case 4:
      return fp32 ? launch_ens_zfwd_f32(E, which == 4, wg, v, st) : launch_ens_zfwd_f64(E, which == 4, wg, v, st);
    default: return -1;
  }
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


// ==== persistent launch ==================================================================================
// One launch per RK3 substep runs every kernel class for every member of every group.  Its workgroups
// (PERSIST_NTH threads, two per CU) take tasks from a queue built by ens_persist_tasks, in order.  A task
// runs the bodies of the grouped launches (two 128-thread plane bodies side by side, or the group's
// column-pass shape), so every member gets the same bits as the same channel run alone.
// Ordering between workgroups, which may run on different XCDs with L2 caches of their own: a task that
// has finished waits for its stores, writes back the caches (agent-scope release fence) and then adds one
// to its completion counter; a task first waits until its counters have reached their targets, then
// invalidates its caches (agent-scope acquire fence) before the workgroup reads anything.  The queue puts
// every task after all the tasks it waits for, and a workgroup only takes a task once it runs, so a wait
// only ever waits on tasks that running workgroups hold: the queue cannot deadlock.  The waits are
// bounded all the same: one that runs out sets a flag (dns_ensemble_sync reports it) and every later wait
// stops at once, so the kernel always ends.  The last workgroup to finish sets the queue head and every
// counter back to zero for the next launch.
// Registers: the plane bodies, the x-pass and the column bodies each need most of the 256 registers of
// two wavefronts per SIMD, and inlined into one function they spilled 550-700 registers per lane.  So
// every class runs in a function of its own (not inlined: its own register allocation), which keeps
// taking tasks while they are of its class; a workgroup calls about one such function per class per
// substep.  The price: pointers reach those functions as generic pointers, so their global memory
// accesses are flat instructions.

#ifndef PERSIST_FN
#define PERSIST_FN __attribute__((noinline))
#endif
constexpr int PERSIST_SPIN_LIMIT = 1 << 20;
using lds_ptr = __attribute__((address_space(3))) unsigned char *;

// LDS of one 128-thread plane body (the largest of every class and length)
template <typename R> constexpr size_t persist_sub_lds() {
  size_t m = 0;
  const size_t l[] = {EnsBody<R, 0, 48>::B::LDS, EnsBody<R, 0, 96>::B::LDS, EnsBody<R, 0, 192>::B::LDS,
                      EnsBody<R, 1, 48>::B::LDS, EnsBody<R, 1, 96>::B::LDS, EnsBody<R, 1, 192>::B::LDS,
                      EnsBody<R, 2, 48>::B::LDS, EnsBody<R, 2, 96>::B::LDS, EnsBody<R, 2, 192>::B::LDS,
                      EnsBody<R, 3, 48>::B::LDS, EnsBody<R, 3, 96>::B::LDS, EnsBody<R, 3, 192>::B::LDS,
                      EnsBody<R, 4, 48>::B::LDS, EnsBody<R, 4, 96>::B::LDS, EnsBody<R, 4, 192>::B::LDS};
  for (size_t v : l) m = al16(v) > m ? al16(v) : m;
  return m;
}
// column-pass shape of a rows bucket: 8 columns x P segments x M rows, PERSIST_NTH / (8 P) bodies
template <int ROWS> struct PCol;
template <> struct PCol<32> { static constexpr int P = 8, M = 4; };
template <> struct PCol<64> { static constexpr int P = 16, M = 4; };
template <> struct PCol<96> { static constexpr int P = 32, M = 3; };
template <> struct PCol<128> { static constexpr int P = 32, M = 4; };
template <> struct PCol<160> { static constexpr int P = 32, M = 5; };
template <int ROWS, typename NT> struct PColB {
  // the sequential form at 160 rows (as the single-channel and grouped launches: the same bits)
  using B = typename std::conditional<ROWS == 160, ColumnSeqBody<8, PCol<ROWS>::P, PCol<ROWS>::M, NT>,
                                      ColumnPassBody<8, PCol<ROWS>::P, PCol<ROWS>::M, NT>>::type;
  static constexpr int SUB = PERSIST_NTH / B::NTH;
  static constexpr size_t LDS = SUB * al16(B::LDS);
};
template <typename R> constexpr size_t persist_lds() {
  size_t m = 2 * persist_sub_lds<R>();
  const size_t l[] = {PColB<32, R>::LDS, PColB<64, R>::LDS, PColB<96, R>::LDS, PColB<128, R>::LDS,
                      PColB<160, R>::LDS, sizeof(MeanflowLds)};
  for (size_t v : l) m = v > m ? v : m;
  return m;
}
static int persist_col_sub(int rows) { return rows == 32 ? 4 : rows == 64 ? 2 : 1; }
static int persist_rows(int nyu) {
  return nyu <= 32 ? 32 : nyu <= 64 ? 64 : nyu <= 96 ? 96 : nyu <= 128 ? 128 : nyu <= 160 ? 160 : 0;
}

// Next ticket of the queue, the same on every thread of the workgroup.
__device__ __forceinline__ int persist_next(const PersistArgs &E, int *tk) {
  if (threadIdx.x == 0) *tk = (int)atomicAdd(E.ctr + PCTR_HEAD, 1u);
  __syncthreads();
  const int t = __builtin_amdgcn_readfirstlane(*tk);
  __syncthreads();  // every thread has read the ticket before thread 0 writes the next one
  return t;
}

// Thread 0 waits for task T's counters, then the acquire fence; the workgroup starts after the barrier.
__device__ __forceinline__ void persist_begin(const PersistArgs &E, const PTask &T) {
  if (threadIdx.x == 0) {
    for (int d = 0; d < 3; d++) {
      if (T.dep[d] < 0) continue;
      unsigned *c = E.ctr + PCTR0 + T.dep[d];
      for (int spin = 0;; spin++) {
        if ((int)__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= T.need[d]) break;
        if (__hip_atomic_load(E.ctr + PCTR_ERR, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
        if (spin >= PERSIST_SPIN_LIMIT) {
          atomicOr(E.ctr + PCTR_ERR, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(8);
      }
    }
    // acquire (agent scope): invalidates this CU's caches, so the reads see other XCDs' results
    if (T.dep[0] >= 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Every wavefront waits for its stores; then thread 0 writes back (release fence) and counts the task.
__device__ __forceinline__ void persist_end(const PersistArgs &E, const PTask &T) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    // release (agent scope), only where a later task of this launch reads the results
    if (T.rel) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    atomicAdd(E.ctr + PCTR0 + T.done, 1u);
  }
}

template <typename R, int W>
__device__ __forceinline__ void persist_plane_body(const PlaneArgs &a, int lb, int gx, int gy, int n,
                                                   unsigned char *l) {
  switch (n) {
    case 48: EnsBody<R, W, 48>::B::run(a, lb, gx, gy, l); break;
    case 96: EnsBody<R, W, 96>::B::run(a, lb, gx, gy, l); break;
    default: EnsBody<R, W, 192>::B::run(a, lb, gx, gy, l); break;
  }
}

// Tasks of plane class W (0 z-inverse, 1/2 x-pass u/v-planes, 3/4 z-forward u/v-planes) from ticket t on,
// while the queue holds that class; returns the first ticket of another class (or past the end).  Unit u
// of member m is the member's logical blocks 2u (body 0) and 2u + 1 (body 1); a body past the member's
// last block repeats that block (the same values written twice).
template <typename R, int W>
__device__ PERSIST_FN int persist_plane(const PersistArgs &E, int t, int cls, lds_ptr lds3,
                                                       int *tk) {
  unsigned char *lds = (unsigned char *)lds3;
  for (;;) {
    const PTask T = E.task[t];
    persist_begin(E, T);
    for (int u = T.u0; u < T.u1; u++) {
      // the group index is opaque to the optimiser in every unit, so nothing of a body is hoisted out of
      // this loop and kept live (in registers) across the units
      int g = __builtin_amdgcn_readfirstlane(T.g);
      asm volatile("" : "+s"(g));
      const PlaneArgs &a = E.pa[g];
      const PersistGroup &q = E.q[g];
      const int sub = (int)threadIdx.x / ENS_PLANE_NTH;
      unsigned char *l = lds + sub * persist_sub_lds<R>();
      const int gx = W == 0 ? q.gxz : (W <= 2 ? q.gxx : q.gxf);
      const int nbm = W == 0 ? q.nbz : (W <= 2 ? q.nbx : q.nbf);
      const int gy = W == 0 ? 3 : a.nbj * a.nm;
      const int n = (W == 1 || W == 2) ? a.Mx : a.Mz;
      const int lbm = min(2 * u + sub, nbm - 1);
      persist_plane_body<R, W>(a, T.m * nbm + lbm, gx, gy, n, l);
      __syncthreads();
    }
    persist_end(E, T);
    t = persist_next(E, tk);
    if (t >= E.ntask || E.task[t].cls != cls) return t;
  }
}

// Column stage: unit u of member m is column blocks u*SUB .. u*SUB+SUB-1 (one per body); a body past the
// last block computes it again and stores nothing.
template <int ROWS, typename NT>
__device__ __forceinline__ void persist_col_rows(const ColArgs &a, const PTask &T, int ncb, unsigned char *lds) {
  using S = PColB<ROWS, NT>;
  const int sub = (int)threadIdx.x / S::B::NTH, tid = (int)threadIdx.x % S::B::NTH;
  unsigned char *l = lds + sub * al16(S::B::LDS);
  for (int u = T.u0; u < T.u1; u++) {
    const int b = u * S::SUB + sub;
    const bool active = b < ncb;
    int m = __builtin_amdgcn_readfirstlane(T.m);  // opaque in every unit, as in persist_plane
    asm volatile("" : "+s"(m));
    S::B::run(a, active ? b : ncb - 1, ncb, m, tid, l, active);
    __syncthreads();
  }
}
template <int ROWS, typename NT>
__device__ PERSIST_FN int persist_col(const PersistArgs &E, int t, lds_ptr lds3, int *tk) {
  unsigned char *lds = (unsigned char *)lds3;
  for (;;) {
    const PTask T = E.task[t];
    persist_begin(E, T);
    persist_col_rows<ROWS, NT>(E.ca[T.g], T, E.q[T.g].ncb, lds);
    persist_end(E, T);
    t = persist_next(E, tk);
    if (t >= E.ntask || E.task[t].cls != PC_COL || E.q[E.task[t].g].cshape != ROWS) return t;
  }
}

// kRK = 1: every member's CFL maxima and slots to zero; every kRK: max|div| to zero (ens_reset_kernel).
__device__ __forceinline__ void persist_reset(const PersistArgs &E) {
  for (int g = 0; g < E.ng; g++) {
    const int nm = E.pa[g].nm;
    DnsState *st = E.ca[g].st;
    for (int i = threadIdx.x; i < nm; i += PERSIST_NTH) {
      if (E.kRK == 1) {
        st[i].cfl[0] = 0.0;
        st[i].cfl[1] = 0.0;
      }
      st[i].divmax = 0.0;
    }
    if (E.kRK == 1)
      for (int i = threadIdx.x; i < nm * 2 * CFL_SLOTS; i += PERSIST_NTH) E.pa[g].cfl[i] = 0.0;
  }
}

// ghost plane 0 of u1/u3 of member m into its CFL maxima (ens_ghost_kernel over every column)
__device__ __forceinline__ void persist_ghost(const PersistArgs &E, const PTask &T, double *red) {
  const PlaneArgs &a = E.pa[T.g];
  const size_t ms = member_strides(a.C, a.nyu, a.nyv).ru;
  const double2 *u1 = a.u1 + (size_t)T.m * ms, *u3 = a.u3 + (size_t)T.m * ms;
  double m1 = 0.0, m3 = 0.0;
  for (int c = threadIdx.x; c < a.C; c += PERSIST_NTH) {
    const double2 x = u1[c], y = u3[c];
    m1 = fmax(m1, fmax(fabs(x.x), fabs(x.y)));
    m3 = fmax(m3, fmax(fabs(y.x), fabs(y.y)));
  }
  m1 = block_max_sub<PERSIST_NTH>(m1, red, threadIdx.x);
  m3 = block_max_sub<PERSIST_NTH>(m3, red, threadIdx.x);
  if (threadIdx.x == 0) {
    double *cfl = a.cfl + (size_t)T.m * 2 * CFL_SLOTS;
    atomic_max_nonneg(&cfl[0], m1);
    atomic_max_nonneg(&cfl[CFL_SLOTS], m3);
  }
}

// max over member m's CFL slots (ens_cfl_reduce_kernel), first wavefront
__device__ __forceinline__ void persist_cflr(const PersistArgs &E, const PTask &T) {
  static_assert(CFL_SLOTS == 64, "one wavefront");
  if (threadIdx.x >= 64) return;
  const double *slots = E.pa[T.g].cfl + (size_t)T.m * 2 * CFL_SLOTS;
  const int l = threadIdx.x;
  double m1 = slots[l], m3 = slots[CFL_SLOTS + l];
  for (int off = 32; off > 0; off >>= 1) {
    m1 = fmax(m1, __shfl_xor(m1, off, 64));
    m3 = fmax(m3, __shfl_xor(m3, off, 64));
  }
  if (l == 0) {
    E.ca[T.g].st[T.m].cfl[0] = m1;
    E.ca[T.g].st[T.m].cfl[1] = m3;
  }
}

// the small classes (reset, ghost plane, CFL reduction, meanflow_ctU), one task at a time
__device__ PERSIST_FN int persist_small(const PersistArgs &E, int t, lds_ptr lds3, int *tk,
                                                       double *red) {
  unsigned char *lds = (unsigned char *)lds3;
  const PTask T = E.task[t];
  persist_begin(E, T);
  switch (T.cls) {
    case PC_RESET: persist_reset(E); break;
    case PC_GHOST: persist_ghost(E, T, red); break;
    case PC_CFLR: persist_cflr(E, T); break;
    default: meanflow_post_body(E.ca[T.g], T.m, *reinterpret_cast<MeanflowLds *>(lds)); break;
  }
  persist_end(E, T);
  return persist_next(E, tk);
}

template <typename R>
__global__ void __launch_bounds__(PERSIST_NTH, 2) ens_persist_kernel(const PersistArgs *__restrict__ Ep) {
  __shared__ __align__(16) unsigned char lds[persist_lds<R>()];
  __shared__ double red[PERSIST_NTH / 64];
  __shared__ int tk, last;
  const PersistArgs &E = *Ep;
  const lds_ptr l3 = (lds_ptr)lds;
  int t = persist_next(E, &tk);
  while (t < E.ntask) {
    const int cls = E.task[t].cls;
    switch (cls) {
      case PC_ZINV: t = persist_plane<R, 0>(E, t, cls, l3, &tk); break;
      case PC_XU: t = persist_plane<R, 1>(E, t, cls, l3, &tk); break;
      case PC_XV: t = persist_plane<R, 2>(E, t, cls, l3, &tk); break;
      case PC_ZU: t = persist_plane<R, 3>(E, t, cls, l3, &tk); break;
      case PC_ZV: t = persist_plane<R, 4>(E, t, cls, l3, &tk); break;
      case PC_COL:
        switch (E.q[E.task[t].g].cshape) {
          case 32: t = persist_col<32, R>(E, t, l3, &tk); break;
          case 64: t = persist_col<64, R>(E, t, l3, &tk); break;
          case 96: t = persist_col<96, R>(E, t, l3, &tk); break;
          case 128: t = persist_col<128, R>(E, t, l3, &tk); break;
          default: t = persist_col<160, R>(E, t, l3, &tk); break;
        }
        break;
      default: t = persist_small(E, t, l3, &tk, red); break;
    }
  }
  if (threadIdx.x == 0) last = atomicAdd(E.ctr + PCTR_EXIT, 1u) == gridDim.x - 1;
  __syncthreads();
  if (last) {  // every other workgroup has left: clear the counters for the next launch
    __threadfence();
    for (int i = threadIdx.x; i < E.nctr; i += PERSIST_NTH) E.ctr[PCTR0 + i] = 0u;
    if (threadIdx.x == 0) {
      E.ctr[PCTR_HEAD] = 0u;
      E.ctr[PCTR_EXIT] = 0u;
    }
  }
}

int ens_persist_group(const PlaneArgs &a, int ctp0, PersistGroup &q) {
  if (!ens_plane_length_supported(a.Mx) || !ens_plane_length_supported(a.Mz) || a.Kx != a.Mx / 3 + 1 ||
      a.Nz != 2 * a.Mz / 3 || persist_rows(a.nyu) == 0 || a.nyu < 8)
    return -1;
  int gy = 0, nb = 0;
  ens_plane_grid(a, 0, 1, q.gxz, gy, nb);
  q.nbz = nb / a.nm;
  // a z-inverse workgroup walks ZCH + 1 field rows; both bodies of a task must take the same number of
  // barriers, so every chunk of planes must be whole
  const int zch = EnsZT<48, 1>::ZCH;
  if (a.nbj % zch != 0) return -1;
  ens_plane_grid(a, 1, 1, q.gxx, gy, nb);
  q.nbx = nb / a.nm;
  ens_plane_grid(a, 3, 1, q.gxf, gy, nb);
  q.nbf = nb / a.nm;
  q.cshape = persist_rows(a.nyu);
  q.ncb = (a.C + 7) / 8;
  q.ctp0 = ctp0;
  return 0;
}

int ens_persist_tasks(const PersistGroup *q, const int *nm, int ng, int kRK, const int *units,
                      std::vector<PTask> &tasks, int &ncounters) {
  int nmx = 1;
  for (int g = 0; g < ng; g++) nmx = nm[g] > nmx ? nm[g] : nmx;
  ncounters = ng * nmx * PC_N;
  auto ctr = [&](int g, int m, int c) { return (g * nmx + m) * PC_N + c; };
  auto nunits = [&](int g, int c) -> int {
    switch (c) {
      case PC_ZINV: return (q[g].nbz + 1) / 2;
      case PC_XU: case PC_XV: return (q[g].nbx + 1) / 2;
      case PC_ZU: case PC_ZV: return (q[g].nbf + 1) / 2;
      case PC_COL: { const int s = persist_col_sub(q[g].cshape); return (q[g].ncb + s - 1) / s; }
      default: return 1;
    }
  };
  auto per = [&](int c) { return units && units[c] > 0 ? units[c] : 1; };
  auto ntask = [&](int g, int c) { return (nunits(g, c) + per(c) - 1) / per(c); };
  tasks.clear();
  auto add = [&](int c, int g, int m, int d0, int n0, int d1, int n1, int d2, int n2) {
    const int nu = nunits(g, c), pc = per(c);
    for (int u0 = 0; u0 < nu; u0 += pc) {
      PTask t;
      t.cls = c; t.g = g; t.m = m; t.u0 = u0; t.u1 = u0 + pc < nu ? u0 + pc : nu;
      t.done = ctr(g, m, c);
      t.dep[0] = d0; t.need[0] = n0; t.dep[1] = d1; t.need[1] = n1; t.dep[2] = d2; t.need[2] = n2;
      t.rel = 1;
      tasks.push_back(t);
    }
  };
  const int reset = ctr(0, 0, PC_RESET);
  add(PC_RESET, 0, 0, -1, 0, -1, 0, -1, 0);
  if (kRK == 1)
    for (int g = 0; g < ng; g++)
      for (int m = 0; m < nm[g]; m++) add(PC_GHOST, g, m, reset, 1, -1, 0, -1, 0);
  for (int g = 0; g < ng; g++)
    for (int m = 0; m < nm[g]; m++) add(PC_ZINV, g, m, -1, 0, -1, 0, -1, 0);
  for (int g = 0; g < ng; g++)
    for (int m = 0; m < nm[g]; m++)
      add(PC_XU, g, m, ctr(g, m, PC_ZINV), ntask(g, PC_ZINV), kRK == 1 ? reset : -1, 1, -1, 0);
  for (int g = 0; g < ng; g++)
    for (int m = 0; m < nm[g]; m++) add(PC_XV, g, m, ctr(g, m, PC_ZINV), ntask(g, PC_ZINV), -1, 0, -1, 0);
  if (kRK == 1)
    for (int g = 0; g < ng; g++)
      for (int m = 0; m < nm[g]; m++)
        add(PC_CFLR, g, m, ctr(g, m, PC_GHOST), 1, ctr(g, m, PC_XU), ntask(g, PC_XU), -1, 0);
  for (int g = 0; g < ng; g++)
    for (int m = 0; m < nm[g]; m++) add(PC_ZU, g, m, ctr(g, m, PC_XU), ntask(g, PC_XU), -1, 0, -1, 0);
  for (int g = 0; g < ng; g++)
    for (int m = 0; m < nm[g]; m++) add(PC_ZV, g, m, ctr(g, m, PC_XV), ntask(g, PC_XV), -1, 0, -1, 0);
  for (int g = 0; g < ng; g++)
    for (int m = 0; m < nm[g]; m++)
      add(PC_COL, g, m, ctr(g, m, PC_ZU), ntask(g, PC_ZU), ctr(g, m, PC_ZV), ntask(g, PC_ZV),
          kRK == 1 ? ctr(g, m, PC_CFLR) : reset, 1);
  for (int g = 0; g < ng; g++)
    if (q[g].ctp0)
      for (int m = 0; m < nm[g]; m++) add(PC_MF, g, m, ctr(g, m, PC_COL), ntask(g, PC_COL), -1, 0, -1, 0);
  // every wait must be met by earlier tasks alone (no deadlock) and by all the tasks it names (none of
  // them may still be running when the waiting task starts)
  std::vector<int> before(ncounters, 0), total(ncounters, 0), waited(ncounters, 0);
  for (const PTask &t : tasks) {
    total[t.done]++;
    for (int d = 0; d < 3; d++)
      if (t.dep[d] >= 0) waited[t.dep[d]] = 1;
  }
  for (PTask &t : tasks) t.rel = waited[t.done];
  for (const PTask &t : tasks) {
    for (int d = 0; d < 3; d++)
      if (t.dep[d] >= 0 && (before[t.dep[d]] < t.need[d] || total[t.dep[d]] != t.need[d])) return -1;
    before[t.done]++;
  }
  return 0;
}

int launch_ens_persist(const PersistArgs *E, bool fp32, int nwg, hipStream_t st) {
  if (fp32) ens_persist_kernel<float><<<nwg, PERSIST_NTH, 0, st>>>(E);
  else ens_persist_kernel<double><<<nwg, PERSIST_NTH, 0, st>>>(E);
  return 0;
}

int ens_persist_max_blocks(bool fp32) {
  int n = 0;
  hipError_t r = fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ens_persist_kernel<float>, PERSIST_NTH, 0)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ens_persist_kernel<double>, PERSIST_NTH, 0);
  return r == hipSuccess ? n : 0;
}

}  // namespace dnsk
