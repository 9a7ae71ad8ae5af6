// Device helpers shared by the plane-stage and column-stage kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

// ------------------------------------------------------------------------------------------------
// complex arithmetic on double2 / float2 (x = re, y = im)
template <typename R> struct cvec;
template <> struct cvec<double> { using T = double2; };
template <> struct cvec<float> { using T = float2; };
template <typename R> using cx = typename cvec<R>::T;

template <typename R> __device__ __forceinline__ cx<R> mk(R a, R b) { cx<R> c; c.x = a; c.y = b; return c; }
template <typename T> __device__ __forceinline__ T cadd(T a, T b) { T c; c.x = a.x + b.x; c.y = a.y + b.y; return c; }
template <typename T> __device__ __forceinline__ T csub(T a, T b) { T c; c.x = a.x - b.x; c.y = a.y - b.y; return c; }
// The library is compiled without floating-point contraction, so every kernel that shares these
// helpers rounds identically whatever its tiling; fused multiply-adds are written out where wanted.
template <typename T> __device__ __forceinline__ T cmul(T a, T b) {
  T c; c.x = fma(a.x, b.x, -(a.y * b.y)); c.y = fma(a.x, b.y, a.y * b.x); return c;
}
template <typename T> __device__ __forceinline__ T cmulc(T a, T b) {  // a * conj(b)
  T c; c.x = fma(a.x, b.x, a.y * b.y); c.y = fma(a.y, b.x, -(a.x * b.y)); return c;
}
template <typename T> __device__ __forceinline__ T cconj(T a) { a.y = -a.y; return a; }
template <typename T, typename R> __device__ __forceinline__ T cscale(T a, R s) { a.x *= s; a.y *= s; return a; }
// multiply by s*i, s = +1 or -1
template <int S, typename T> __device__ __forceinline__ T cmul_si(T a) {
  T c; if (S > 0) { c.x = -a.y; c.y = a.x; } else { c.x = a.y; c.y = -a.x; } return c;
}
// (0 + i k) * a, exactly as the reference's complex multiply with a purely imaginary factor
__device__ __forceinline__ double2 imul(double k, double2 a) {
  double2 c; c.x = -(k * a.y); c.y = k * a.x; return c;
}
__device__ __forceinline__ double2 d2(double a, double b) { double2 c; c.x = a; c.y = b; return c; }
__device__ __forceinline__ double2 rmul(double r, double2 a) { a.x *= r; a.y *= r; return a; }

template <typename T> __device__ __forceinline__ double2 to_d2(T a) { return d2((double)a.x, (double)a.y); }
template <typename R> __device__ __forceinline__ cx<R> from_d2(double2 a);
template <> __device__ __forceinline__ double2 from_d2<double>(double2 a) { return a; }
template <> __device__ __forceinline__ float2 from_d2<float>(double2 a) {
  float2 c; c.x = (float)a.x; c.y = (float)a.y; return c;
}

// ------------------------------------------------------------------------------------------------
// Compile-time cos / sin of 2 pi num/den: reduced to [0, pi/4] by the quadrant and octant symmetries
// (exact integer arithmetic), then a Taylor series (error below 1 ulp of double on that range).
constexpr double taylor_sin(double x) {
  double t = x, s = x;
  for (int n = 1; n < 12; n++) { t *= -x * x / ((2.0 * n) * (2.0 * n + 1.0)); s += t; }
  return s;
}
constexpr double taylor_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int n = 1; n < 12; n++) { t *= -x * x / ((2.0 * n - 1.0) * (2.0 * n)); s += t; }
  return s;
}
struct cs2 { double c, s; };
constexpr cs2 cossin2pi(long num, long den) {
  num %= den;
  if (num < 0) num += den;
  const long q = (4 * num) / den, r = 4 * num - q * den;  // angle = (q + r/den) * pi/2
  const double hp = 1.57079632679489661923;
  double c, s;
  if (r == 0) { c = 1.0; s = 0.0; }
  else if (2 * r <= den) { const double ph = hp * (double)r / (double)den; c = taylor_cos(ph); s = taylor_sin(ph); }
  else { const double ph = hp * (double)(den - r) / (double)den; c = taylor_sin(ph); s = taylor_cos(ph); }
  if (q == 0) return {c, s};
  if (q == 1) return {-s, c};
  if (q == 2) return {-c, -s};
  return {s, -c};
}

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, class F>
__device__ __forceinline__ void sfor(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// v * e^{S 2 pi i E/N} with E, N known at compile time; exact for multiples of a quarter turn
template <typename R, int S, long E, long N>
__device__ __forceinline__ cx<R> ctw(cx<R> v) {
  constexpr long e = ((E % N) + N) % N;
  if constexpr (e == 0) return v;
  else if constexpr (4 * e == N) return cmul_si<S>(v);
  else if constexpr (2 * e == N) return mk<R>(-v.x, -v.y);
  else if constexpr (4 * e == 3 * N) return cmul_si<-S>(v);
  else {
    constexpr cs2 w = cossin2pi(e, N);
    return cmul(v, mk<R>((R)w.c, (R)(S * w.s)));
  }
}

// ------------------------------------------------------------------------------------------------
// Small DFTs in registers.  S = -1: forward (e^{-2 pi i rs/R}); S = +1: backward.
// Lengths without a hand-written kernel are split N = A * B (A the largest hand-written length that
// divides N): x[B a + b] -> A-point DFTs over a, twiddles e^{S 2 pi i b k1/N}, B-point DFTs over b,
// output k1 + A k2.
__host__ __device__ constexpr int dft_split(int n) {
  return (n % 16 == 0 && n > 16) ? 16 : (n % 9 == 0 && n > 9) ? 9 : (n % 8 == 0 && n > 8) ? 8
       : (n % 4 == 0 && n > 4) ? 4 : (n % 3 == 0 && n > 3) ? 3 : 2;
}
template <typename R, int RAD, int S> struct dft {
  static constexpr int A = dft_split(RAD), B = RAD / A;
  static_assert(A * B == RAD && B > 1, "DFT length must be 2^a 3^b");
  __device__ __forceinline__ static void run(cx<R> *u) {
    cx<R> y[B][A];
    sfor<0, B>([&](auto bi) {
      constexpr int b = decltype(bi)::value;
#pragma unroll
      for (int a = 0; a < A; a++) y[b][a] = u[B * a + b];
      dft<R, A, S>::run(y[b]);
      sfor<1, A>([&](auto ki) {
        constexpr int k1 = decltype(ki)::value;
        y[b][k1] = ctw<R, S, (long)b * k1, RAD>(y[b][k1]);
      });
    });
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) {
      cx<R> z[B];
#pragma unroll
      for (int b = 0; b < B; b++) z[b] = y[b][k1];
      dft<R, B, S>::run(z);
#pragma unroll
      for (int k2 = 0; k2 < B; k2++) u[k1 + A * k2] = z[k2];
    }
  }
};

template <typename R, int S> struct dft<R, 2, S> {
  __device__ __forceinline__ static void run(cx<R> *u) {
    cx<R> a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
  }
};

template <typename R, int S> struct dft<R, 3, S> {
  __device__ __forceinline__ static void run(cx<R> *u) {
    const R h = (R)0.5, s3 = (R)0.86602540378443864676372317075294;
    cx<R> t1 = cadd(u[1], u[2]);
    cx<R> t2 = mk<R>(fma(-h, t1.x, u[0].x), fma(-h, t1.y, u[0].y));
    cx<R> t3 = cmul_si<S>(cscale(csub(u[1], u[2]), s3));
    u[0] = cadd(u[0], t1);
    u[1] = cadd(t2, t3);
    u[2] = csub(t2, t3);
  }
};

template <typename R, int S> struct dft<R, 4, S> {
  __device__ __forceinline__ static void run(cx<R> *u) {
    cx<R> t0 = cadd(u[0], u[2]), t1 = csub(u[0], u[2]);
    cx<R> t2 = cadd(u[1], u[3]), t3 = cmul_si<S>(csub(u[1], u[3]));
    u[0] = cadd(t0, t2);
    u[1] = cadd(t1, t3);
    u[2] = csub(t0, t2);
    u[3] = csub(t1, t3);
  }
};

template <typename R, int S> struct dft<R, 8, S> {
  __device__ __forceinline__ static void run(cx<R> *u) {
    const R r2 = (R)0.70710678118654752440084436210485;
    cx<R> a[4] = {u[0], u[2], u[4], u[6]}, b[4] = {u[1], u[3], u[5], u[7]};
    dft<R, 4, S>::run(a);
    dft<R, 4, S>::run(b);
    // W8^m = e^{S 2 pi i m / 8}
    cx<R> w1 = mk<R>(r2, (R)S * r2), w3 = mk<R>(-r2, (R)S * r2);
    cx<R> b1 = cmul(b[1], w1), b2 = cmul_si<S>(b[2]), b3 = cmul(b[3], w3);
    u[0] = cadd(a[0], b[0]); u[4] = csub(a[0], b[0]);
    u[1] = cadd(a[1], b1);   u[5] = csub(a[1], b1);
    u[2] = cadd(a[2], b2);   u[6] = csub(a[2], b2);
    u[3] = cadd(a[3], b3);   u[7] = csub(a[3], b3);
  }
};

// radix 9 = 3 x 3: n = 3a + b, k = c + 3d, X_{c+3d} = sum_b W3^{bd} W9^{bc} sum_a x_{3a+b} W3^{ac}
template <typename R, int S> struct dft<R, 9, S> {
  __device__ __forceinline__ static void run(cx<R> *u) {
    const R c1 = (R)0.76604444311897803520239265055541667, s1 = (R)0.64278760968653932632264340990726343;  // 2pi/9
    const R c2 = (R)0.17364817766693034885171662676931479, s2 = (R)0.98480775301220805936674302458952301;  // 4pi/9
    const R c4 = (R)-0.93969262078590838405410927732473147, s4 = (R)0.34202014332566873304409961468225958; // 8pi/9
    cx<R> y[3][3];
#pragma unroll
    for (int b = 0; b < 3; b++) {
      y[b][0] = u[b]; y[b][1] = u[3 + b]; y[b][2] = u[6 + b];
      dft<R, 3, S>::run(y[b]);
    }
    const cx<R> w1 = mk<R>(c1, (R)S * s1), w2 = mk<R>(c2, (R)S * s2), w4 = mk<R>(c4, (R)S * s4);
    y[1][1] = cmul(y[1][1], w1);
    y[1][2] = cmul(y[1][2], w2);
    y[2][1] = cmul(y[2][1], w2);
    y[2][2] = cmul(y[2][2], w4);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      cx<R> z[3] = {y[0][c], y[1][c], y[2][c]};
      dft<R, 3, S>::run(z);
      u[c] = z[0]; u[c + 3] = z[1]; u[c + 6] = z[2];
    }
  }
};

// radix 16 = 4 x 4: n = 4a + b, k = c + 4d
template <typename R, int S> struct dft<R, 16, S> {
  __device__ __forceinline__ static void run(cx<R> *u) {
    const R r2 = (R)0.70710678118654752440084436210485;
    const R ca = (R)0.92387953251128675612818318939678829, sa = (R)0.38268343236508977172845998403039886;  // pi/8
    cx<R> y[4][4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
      y[b][0] = u[b]; y[b][1] = u[4 + b]; y[b][2] = u[8 + b]; y[b][3] = u[12 + b];
      dft<R, 4, S>::run(y[b]);
    }
    // W16^{bc}: (1,1)=w1 (1,2)=w2 (1,3)=w3 (2,1)=w2 (2,2)=w4 (2,3)=w6 (3,1)=w3 (3,2)=w6 (3,3)=w9
    const cx<R> w1 = mk<R>(ca, (R)S * sa), w2 = mk<R>(r2, (R)S * r2), w3 = mk<R>(sa, (R)S * ca);
    const cx<R> w6 = mk<R>(-r2, (R)S * r2), w9 = mk<R>(-ca, -(R)S * sa);
    y[1][1] = cmul(y[1][1], w1); y[1][2] = cmul(y[1][2], w2); y[1][3] = cmul(y[1][3], w3);
    y[2][1] = cmul(y[2][1], w2); y[2][2] = cmul_si<S>(y[2][2]); y[2][3] = cmul(y[2][3], w6);
    y[3][1] = cmul(y[3][1], w3); y[3][2] = cmul(y[3][2], w6); y[3][3] = cmul(y[3][3], w9);
#pragma unroll
    for (int c = 0; c < 4; c++) {
      cx<R> z[4] = {y[0][c], y[1][c], y[2][c], y[3][c]};
      dft<R, 4, S>::run(z);
      u[c] = z[0]; u[c + 4] = z[1]; u[c + 8] = z[2]; u[c + 12] = z[3];
    }
  }
};

// ------------------------------------------------------------------------------------------------
// Stockham autosort FFT of NTR transforms of length N held in LDS as s[n*NTR + t].
// Pass with radix RAD and span P: butterfly i (0..N/RAD-1) of transform t reads s[(i + r*N/RAD)],
// multiplies by w^(r*k), k = i mod P, w = e^{S 2 pi i/(P*RAD)}, and writes the radix-RAD DFT to
// (i-k)*RAD + k + q*P.  tw[m] = e^{-2 pi i m/N} (conjugated for S = +1).  All loads of a pass are
// issued before the barrier and all stores after it, so the pass works in place.
// Radix plan: 16 while it divides, then 9, 8, 4, 2, 3 -- at most 3 LDS passes for every 2^a 3^b
// length used here (576 = 16.9.4, 288 = 16.9.2, 192 = 16.4.3, 1152 = 16.9.8).
__host__ __device__ constexpr int next_radix(int rem) {
  return rem % 16 == 0 ? 16 : rem % 9 == 0 ? 9 : rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2
                                                                                  : rem % 3 == 0 ? 3 : 0;
}

// ---- LDS layouts of a batch of transforms ------------------------------------------------------
// N: length; NTR slots; NACT active transforms, active transform a in slot (a / GA) * GS + a % GA.
// PS > 0 inserts one padding row after every 2^PS positions, which breaks the power-of-two strides of
// a first pass's stores across LDS banks.  BF = false: s[row(n) * NTR + slot] and consecutive threads
// take consecutive transforms (memory-coalesced when the transforms are adjacent columns);
// BF = true: s[slot * LR + row(n)] and consecutive threads take consecutive butterflies (for transforms
// that are contiguous rows).
// TR = true maps active transform a to slot (a % GA) * GS + a / GA instead (transform-major order of the
// active transforms, e.g. all rows of field 0 first).
// OPQ = false lets the optimiser keep the passes' index arithmetic in registers (for kernels whose
// occupancy LDS limits, so that registers are not the scarce resource).
template <int N_, int NTR_, int NACT_ = NTR_, int GA_ = 1, int GS_ = 1, int PS_ = 0, bool BF_ = false,
          bool TR_ = false, bool OPQ_ = true>
struct Lay {
  static constexpr int N = N_, NTR = NTR_, NACT = NACT_, PS = PS_;
  static constexpr bool BF = BF_;
  static constexpr int LR = PS > 0 ? N + (N >> PS) : N;  // rows per slot
  static constexpr int SIZE = NTR * LR;                   // elements
  __device__ __forceinline__ static int row(int n) { return PS > 0 ? n + (n >> PS) : n; }
  __device__ __forceinline__ static int addr(int n, int t) { return BF ? t * LR + row(n) : row(n) * NTR + t; }
  __device__ __forceinline__ static int slot(int a) {
    return TR_ ? (a % GA_) * GS_ + a / GA_ : (a / GA_) * GS_ + a % GA_;
  }
  template <int NTH> __device__ __forceinline__ static int tid();
  // thread work item b of a pass with M butterflies per transform -> (active transform a, butterfly i)
  template <int M>
  __device__ __forceinline__ static void split(int b, int &a, int &i) {
    if (BF) { i = b % M; a = b / M; } else { a = b % NACT; i = b / NACT; }
  }
};

// Workgroup b of nblk (linear order) as the hardware deals them out: to the 8 XCDs in turn, b % 8.  The
// returned logical block gives each XCD a contiguous run of logical blocks (the first nblk % 8 XCDs one
// more), so workgroups that run together on one XCD are neighbours in logical order and share the
// 128-byte lines at their edges through that XCD's L2.  A bijection on 0..nblk-1.
__device__ __forceinline__ int xcd_block(int b, int nblk) {
  const int q = nblk / 8, r = nblk % 8, x = b % 8, i = b / 8;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Thread index within a body of NTH threads.  A workgroup may run several bodies of a power-of-two
// NTH side by side (the ensemble's persistent launch), body b on threads b*NTH .. b*NTH+NTH-1; in a
// launch whose workgroup is one body the mask is known to be a no-op (the launch bounds give
// threadIdx.x < NTH) and disappears.
template <int NTH> __device__ __forceinline__ int body_tid() {
  return (NTH & (NTH - 1)) == 0 ? (int)(threadIdx.x & (NTH - 1)) : (int)threadIdx.x;
}

// Thread index, opaque to the optimiser: index arithmetic derived from it is recomputed where it is used
// instead of being hoisted out of outer loops and kept live in registers across them.
template <int NTH> __device__ __forceinline__ int tidx() {
  int t = body_tid<NTH>();
  asm volatile("" : "+v"(t));
  return t;
}

template <int N_, int NTR_, int NACT_, int GA_, int GS_, int PS_, bool BF_, bool TR_, bool OPQ_>
template <int NTH>
__device__ __forceinline__ int Lay<N_, NTR_, NACT_, GA_, GS_, PS_, BF_, TR_, OPQ_>::tid() {
  return OPQ_ ? tidx<NTH>() : body_tid<NTH>();
}

// Butterfly of one pass on registers u[0..RAD): twiddle (P > 1), DFT.
template <typename R, int N, int S, int P, int RAD>
__device__ __forceinline__ void butterfly(cx<R> *u, int k, const cx<R> *__restrict__ tw) {
  constexpr int TSTEP = N / (P * RAD);
  if (P > 1) {
#pragma unroll
    for (int r = 1; r < RAD; r++) {
      const cx<R> w = tw[r * k * TSTEP];
      u[r] = (S < 0) ? cmul(u[r], w) : cmulc(u[r], w);
    }
  }
  dft<R, RAD, S>::run(u);
}

// One LDS-to-LDS pass (in place: all loads before the barrier, all stores after, barrier at the end).
template <typename R, class L, int NTH, int S, int P, int RAD>
__device__ __forceinline__ void fft_pass(cx<R> *s, const cx<R> *__restrict__ tw) {
  constexpr int M = L::N / RAD;
  constexpr int NB = M * L::NACT;
  constexpr int PER = (NB + NTH - 1) / NTH;
  cx<R> u[PER][RAD];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = L::template tid<NTH>() + q * NTH;
    if ((NB % NTH) == 0 || b < NB) {
      int a, i;
      L::template split<M>(b, a, i);
      const int t = L::slot(a);
#pragma unroll
      for (int r = 0; r < RAD; r++) u[q][r] = s[L::addr(i + r * M, t)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = L::template tid<NTH>() + q * NTH;
    if ((NB % NTH) == 0 || b < NB) {
      int a, i;
      L::template split<M>(b, a, i);
      const int t = L::slot(a);
      const int k = i % P;
      butterfly<R, L::N, S, P, RAD>(u[q], k, tw);
      const int j0 = (i - k) * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; r++) s[L::addr(j0 + r * P, t)] = u[q][r];
    }
  }
  __syncthreads();
}

// ---- radix plans --------------------------------------------------------------------------------
// Plan<R0, R1, ...>: the radices of the passes in order (product = transform length).
template <int... Rs> struct Plan {};
template <int R0, class PL> struct plan_prepend;
template <int R0, int... Rs> struct plan_prepend<R0, Plan<Rs...>> { using type = Plan<R0, Rs...>; };
template <class PL> struct plan_info;
template <int R0> struct plan_info<Plan<R0>> {
  static constexpr int first = R0, last = R0, prod = R0, len = 1;
  using init = Plan<>;   // all but the last radix
  using tail = Plan<>;   // all but the first radix
};
template <int R0, int R1, int... Rs> struct plan_info<Plan<R0, R1, Rs...>> {
  using nx = plan_info<Plan<R1, Rs...>>;
  static constexpr int first = R0, last = nx::last, prod = R0 * nx::prod, len = 1 + nx::len;
  using init = typename plan_prepend<R0, typename nx::init>::type;
  using tail = Plan<R1, Rs...>;
};
// greedy plan: 16 while it divides, then 9, 8, 4, 2, 3
template <int REM, class Acc> struct greedy_plan;
template <int REM, int... Rs> struct greedy_plan<REM, Plan<Rs...>> {
  using type = typename greedy_plan<REM / next_radix(REM), Plan<Rs..., next_radix(REM)>>::type;
};
template <int... Rs> struct greedy_plan<1, Plan<Rs...>> { using type = Plan<Rs...>; };
template <int N> using GreedyPlan = typename greedy_plan<N, Plan<>>::type;

// Plans of the tuned lengths (chosen by timing on MI355X): z-transforms of 576 and 1152 as 12.12.x (48
// threads per column give one butterfly per thread in every pass), the x half-lengths 288 and 576 in two
// passes.  Other lengths take the greedy plan.
template <int N> struct ZPlanT { using type = GreedyPlan<N>; };
template <> struct ZPlanT<576> { using type = Plan<12, 12, 4>; };
template <> struct ZPlanT<1152> { using type = Plan<12, 12, 8>; };
template <> struct ZPlanT<192> { using type = Plan<16, 12>; };
template <> struct ZPlanT<96> { using type = Plan<8, 12>; };
template <int H> struct XPlanT { using type = GreedyPlan<H>; };
template <> struct XPlanT<288> { using type = Plan<16, 18>; };
template <> struct XPlanT<96> { using type = Plan<8, 12>; };
template <> struct XPlanT<48> { using type = Plan<8, 6>; };
template <> struct XPlanT<24> { using type = Plan<4, 6>; };
template <> struct XPlanT<576> { using type = Plan<8, 8, 9>; };

// Runs the passes of plan PL starting at span P (LDS to LDS).
template <typename R, class L, int NTH, int S, int P, class PL> struct fft_passes;
template <typename R, class L, int NTH, int S, int P> struct fft_passes<R, L, NTH, S, P, Plan<>> {
  __device__ __forceinline__ static void run(cx<R> *, const cx<R> *__restrict__) {}
};
template <typename R, class L, int NTH, int S, int P, int R0, int... Rs>
struct fft_passes<R, L, NTH, S, P, Plan<R0, Rs...>> {
  __device__ __forceinline__ static void run(cx<R> *s, const cx<R> *__restrict__ tw) {
    fft_pass<R, L, NTH, S, P, R0>(s, tw);
    fft_passes<R, L, NTH, S, P * R0, Plan<Rs...>>::run(s, tw);
  }
};

// Full transform of every active slot in LDS; expects the data in s and a barrier before the call.
template <typename R, class L, int NTH, int S, class PL = GreedyPlan<L::N>>
__device__ __forceinline__ void fft_lds(cx<R> *s, const cx<R> *__restrict__ tw) {
  static_assert(plan_info<PL>::prod == L::N, "plan does not match the length");
  fft_passes<R, L, NTH, S, 1, PL>::run(s, tw);
}

// First pass (radix RAD) straight from memory: ld(i, rM, a) gives position i + rM of active transform a,
// rM = r*N/RAD passed as std::integral_constant (compile-time), 0 <= i < N/RAD.  Leaves the pass's
// output in s and ends with a barrier (no staging of the input through LDS).
template <typename R, class L, int NTH, int S, int RAD, typename LD>
__device__ __forceinline__ void fft_first(cx<R> *s, LD ld) {
  constexpr int M = L::N / RAD;
  constexpr int NB = M * L::NACT;
  constexpr int PER = (NB + NTH - 1) / NTH;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = L::template tid<NTH>() + q * NTH;
    if ((NB % NTH) == 0 || b < NB) {
      int a, i;
      L::template split<M>(b, a, i);
      const int t = L::slot(a);
      cx<R> u[RAD];
      sfor<0, RAD>([&](auto ri) {
        constexpr int r = decltype(ri)::value;
        u[r] = ld(i, std::integral_constant<int, r * M>{}, a);
      });
      dft<R, RAD, S>::run(u);
#pragma unroll
      for (int r = 0; r < RAD; r++) s[L::addr(i * RAD + r, t)] = u[r];
    }
  }
  __syncthreads();
}

// Last pass (radix RAD, span P = N/RAD) straight to memory: st(n, a, k, value) receives position n of
// active transform a; k = q*RAD + r (q-th work item of this thread, r-th output) is a compile-time index
// after unrolling, the same for the same (n, a) on every call, so a caller can keep per-output state in
// registers across transforms.  Only reads s: the caller needs a barrier before s is written again.
template <typename R, class L, int NTH, int S, int RAD, typename ST>
__device__ __forceinline__ void fft_last(const cx<R> *s, const cx<R> *__restrict__ tw, ST st) {
  constexpr int P = L::N / RAD;
  constexpr int NB = P * L::NACT;
  constexpr int PER = (NB + NTH - 1) / NTH;
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = L::template tid<NTH>() + q * NTH;
    if ((NB % NTH) == 0 || b < NB) {
      int a, i;
      L::template split<P>(b, a, i);
      const int t = L::slot(a);
      cx<R> u[RAD];
#pragma unroll
      for (int r = 0; r < RAD; r++) u[r] = s[L::addr(i + r * P, t)];
      butterfly<R, L::N, S, P, RAD>(u, i, tw);
#pragma unroll
      for (int r = 0; r < RAD; r++) st(i + r * P, a, q * RAD + r, u[r]);
    }
  }
}

// Register slots a thread needs for fft_last's outputs (PER work items x RAD outputs).
template <class L, int NTH, class PL>
constexpr int last_slots() {
  constexpr int RAD = plan_info<PL>::last, P = L::N / RAD;
  return ((P * L::NACT + NTH - 1) / NTH) * RAD;
}

// Memory -> LDS -> memory: first pass from memory, middle passes in LDS, last pass to memory.
template <typename R, class L, int NTH, int S, class PL, typename LD, typename ST>
__device__ __forceinline__ void fft_mem(cx<R> *s, const cx<R> *__restrict__ tw, LD ld, ST st) {
  using I = plan_info<PL>;
  static_assert(I::prod == L::N && I::len >= 2, "plan does not match the length / fewer than two passes");
  fft_first<R, L, NTH, S, I::first>(s, ld);
  fft_passes<R, L, NTH, S, I::first, typename plan_info<typename I::tail>::init>::run(s, tw);
  fft_last<R, L, NTH, S, I::last>(s, tw, st);
}

// Memory -> LDS: first pass from memory, the other passes in LDS (natural order in s, barrier done).
template <typename R, class L, int NTH, int S, class PL, typename LD>
__device__ __forceinline__ void fft_from_mem(cx<R> *s, const cx<R> *__restrict__ tw, LD ld) {
  using I = plan_info<PL>;
  static_assert(I::prod == L::N && I::len >= 2, "plan does not match the length / fewer than two passes");
  fft_first<R, L, NTH, S, I::first>(s, ld);
  fft_passes<R, L, NTH, S, I::first, typename I::tail>::run(s, tw);
}

// ------------------------------------------------------------------------------------------------
// Non-negative double max via the ordering of IEEE bit patterns.
__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

// block_max over the NTH threads tid = 0..NTH-1 of one body of a workgroup that may run several bodies
// side by side (each with its own red); every thread of the workgroup must call it.  Valid on tid 0.
template <int NTH>
__device__ __forceinline__ double block_max_sub(double v, double *red, int tid) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int w = tid >> 6, l = tid & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  if (tid == 0) {
    for (int q = 1; q < NTH / 64; q++) v = fmax(v, red[q]);
  }
  return v;
}

template <int NTH>
__device__ __forceinline__ double block_max(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < NTH / 64; q++) v = fmax(v, red[q]);
  }
  return v;  // valid on thread 0
}
