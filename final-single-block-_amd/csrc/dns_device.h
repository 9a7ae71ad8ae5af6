// Device helpers shared by the plane-stage and column-stage kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

// ------------------------------------------------------------------------------------------------
// complex arithmetic on double2 / float2 (x = re, y = im)
template <typename R> struct cvec;
template <> struct cvec<double> { using T = double2; };
template <> struct cvec<float> { using T = float2; };
template <typename R> using cx = typename cvec<R>::T;

template <typename R> __device__ __forceinline__ cx<R> mk(R a, R b) { cx<R> c; c.x = a; c.y = b; return c; }
template <typename T> __device__ __forceinline__ T cadd(T a, T b) { T c; c.x = a.x + b.x; c.y = a.y + b.y; return c; }
template <typename T> __device__ __forceinline__ T csub(T a, T b) { T c; c.x = a.x - b.x; c.y = a.y - b.y; return c; }
template <typename T> __device__ __forceinline__ T cmul(T a, T b) {
  T c; c.x = a.x * b.x - a.y * b.y; c.y = a.x * b.y + a.y * b.x; return c;
}
template <typename T> __device__ __forceinline__ T cmulc(T a, T b) {  // a * conj(b)
  T c; c.x = a.x * b.x + a.y * b.y; c.y = a.y * b.x - a.x * b.y; return c;
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
// Small DFTs in registers.  S = -1: forward (e^{-2 pi i rs/R}); S = +1: backward.
template <typename R, int RAD, int S> struct dft;

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
    cx<R> t2 = csub(u[0], cscale(t1, h));
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

// ------------------------------------------------------------------------------------------------
// Stockham autosort FFT of NTR transforms of length N held in LDS as s[n*NTR + t].
// Pass with radix RAD and span P: butterfly i (0..N/RAD-1) of transform t reads s[(i + r*N/RAD)],
// multiplies by w^(r*k), k = i mod P, w = e^{S 2 pi i/(P*RAD)}, and writes the radix-RAD DFT to
// (i-k)*RAD + k + q*P.  tw[m] = e^{-2 pi i m/N} (conjugated for S = +1).  All loads of a pass are
// issued before the barrier and all stores after it, so the pass works in place.
__host__ __device__ constexpr int next_radix(int rem) {
  return rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 3 == 0 ? 3 : 0;
}

template <typename R, int N, int NTR, int NTH, int S, int P, int RAD>
__device__ __forceinline__ void fft_pass(cx<R> *s, const cx<R> *__restrict__ tw) {
  constexpr int M = N / RAD;
  constexpr int NB = M * NTR;
  constexpr int PER = (NB + NTH - 1) / NTH;
  constexpr int TSTEP = N / (P * RAD);
  cx<R> u[PER][RAD];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = threadIdx.x + q * NTH;
    if ((NB % NTH) == 0 || b < NB) {
      const int t = b % NTR, i = b / NTR;
#pragma unroll
      for (int r = 0; r < RAD; r++) u[q][r] = s[(i + r * M) * NTR + t];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = threadIdx.x + q * NTH;
    if ((NB % NTH) == 0 || b < NB) {
      const int t = b % NTR, i = b / NTR;
      const int k = i % P;
      if (P > 1) {
#pragma unroll
        for (int r = 1; r < RAD; r++) {
          const cx<R> w = tw[r * k * TSTEP];
          u[q][r] = (S < 0) ? cmul(u[q][r], w) : cmulc(u[q][r], w);
        }
      }
      dft<R, RAD, S>::run(u[q]);
      const int j0 = (i - k) * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; r++) s[(j0 + r * P) * NTR + t] = u[q][r];
    }
  }
  __syncthreads();
}

template <typename R, int N, int NTR, int NTH, int S, int P>
struct fft_passes {
  __device__ __forceinline__ static void run(cx<R> *s, const cx<R> *__restrict__ tw) {
    constexpr int RAD = next_radix(N / P);
    static_assert(RAD > 0, "FFT length must be 2^a 3^b");
    fft_pass<R, N, NTR, NTH, S, P, RAD>(s, tw);
    fft_passes<R, N, NTR, NTH, S, P * RAD>::run(s, tw);
  }
};
template <typename R, int N, int NTR, int NTH, int S>
struct fft_passes<R, N, NTR, NTH, S, N> {
  __device__ __forceinline__ static void run(cx<R> *, const cx<R> *__restrict__) {}
};

// Runs a full transform; expects the data in s and a barrier before the call.
template <typename R, int N, int NTR, int NTH, int S>
__device__ __forceinline__ void fft_lds(cx<R> *s, const cx<R> *__restrict__ tw) {
  fft_passes<R, N, NTR, NTH, S, 1>::run(s, tw);
}

// ------------------------------------------------------------------------------------------------
// Non-negative double max via the ordering of IEEE bit patterns.
__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
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
