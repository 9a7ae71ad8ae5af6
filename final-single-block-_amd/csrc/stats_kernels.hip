// Statistics and spectra of the current state (stats.f90:1-97, spectra.f90:1-60 and the omega_x assembly
// of FOU3D.f90:307-338), accumulated on the device.
//
// Spectra are column-local (spectra.f90:29-60): sp(j, column) += |u|^2 (u1, u2, u3, p) and
// Re u1 Re u2_itp + Im u1 Im u2_itp (u2_itp = interp_v, FOU3D.f90:396-418).
//
// The plane statistics sum physical-space values over the (i, k) points of a plane.  They are computed
// from the spectral rows with the discrete sum rules of the reference's transforms, which are exact up
// to rounding, instead of transforming the planes:
//   * interior planes (u and v planes 1..nyu, in physical space after four_to_phys_u on the 3/2 grid of
//     Mx x Mz points): sum u = Mx Mz Re u(0,0); sum u v = Mx Mz Q(u, v) with
//       Q(u, v) = sum_{kx >= 1} 2 Re(u v*) + sum_{kx = 0} (Re(u_kz v_kz*) + Re(u_kz v_-kz)) / 2,
//     the kx = 0 term being FFTPACK's real synthesis, which ignores Im of the kx = 0 coefficient after
//     the z-transform (rft_buff.f:76-84); -kz is paired only when it is retained (kz = -Nz/2 has no
//     partner on the padded grid);
//   * the planes the reference leaves in spectral layout (u planes 0 and nyu+1, v plane 0; p planes,
//     which hold modes_to_planes_UVP's output): sums of the interleaved (Re, Im) entries;
//   * omega_x (v planes 0..nyu) on the unpadded Nx x Nz grid of four_to_phys_N: sum = Nx Nz Re w(0,0),
//     sum w^2 = Nx Nz Q'(w) with kx = 0 and kx = Nx/2 (the Nyquist of the unpadded x-transform) taking
//     the half-terms and kz = -Nz/2 paired with itself.
// Partial sums per chunk of columns go to a scratch array and are added in a fixed order, so the
// results are deterministic.  One kernel reads every field row once and does both: the spectra of the
// row's columns (read-modify-write) and the plane sums of its chunk.
#include "dns_kernels.h"

namespace dnsk {

constexpr int STAT_CHUNK = 1024;  // columns per partial sum
constexpr int STAT_NTH = 256;

__device__ __forceinline__ double re_dot(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }      // Re(a b*)
__device__ __forceinline__ double re_mul(double2 a, double2 b) { return a.x * b.x - a.y * b.y; }      // Re(a b)
__device__ __forceinline__ double sum_ri(double2 a) { return a.x + a.y; }

// interp_v (FOU3D.f90:396-418): v grid -> u grid row j (0..nyu+1) of one column
__device__ __forceinline__ double2 u2_on_u(const StatsArgs &a, const double2 *u2, size_t col, int j) {
  auto in = [&](int r) -> double2 {
    const double2 v = u2[(size_t)r * a.C + col], w = u2[(size_t)(r - 1) * a.C + col];
    const double d = a.yv[r] - a.yv[r - 1];
    return d2(((a.yu[r] - a.yv[r - 1]) * v.x + (a.yv[r] - a.yu[r]) * w.x) / d,
              ((a.yu[r] - a.yv[r - 1]) * v.y + (a.yv[r] - a.yu[r]) * w.y) / d);
  };
  if (j >= 1 && j <= a.nyu) return in(j);
  if (j == 0) {
    const double2 i1 = in(1), w = u2[col];
    return d2(a.gi1 * (w.x - i1.x) + i1.x, a.gi1 * (w.y - i1.y) + i1.y);
  }
  const double2 in_ = in(a.nyu), w = u2[(size_t)(a.nyv + 1) * a.C + col];
  return d2(a.gi2 * (w.x - in_.x) + in_.x, a.gi2 * (w.y - in_.y) + in_.y);
}

// interp_u (FOU3D.f90:373-392): u grid -> v grid row j (0..nyv+1) of one column
__device__ __forceinline__ double2 u_on_v(const StatsArgs &a, const double2 *u, size_t col, int j) {
  const double2 hi = u[(size_t)(j + 1) * a.C + col], lo = u[(size_t)j * a.C + col];
  const double d = a.yu[j + 1] - a.yu[j];
  return d2(((a.yv[j] - a.yu[j]) * hi.x + (a.yu[j + 1] - a.yv[j]) * lo.x) / d,
            ((a.yv[j] - a.yu[j]) * hi.y + (a.yu[j + 1] - a.yv[j]) * lo.y) / d);
}

// omega_x = -k1F_z u2 + (u3(j+1) - u3(j)) dthdyv(j) of one column, v row j (FOU3D.f90:315-336, 592-606)
__device__ __forceinline__ double2 omega_x(const StatsArgs &a, size_t col, int j) {
  const int k = (int)(col / a.Kx);
  const double kz = k < a.Nz / 2 ? a.bet * (double)k : -(a.bet * (double)(a.Nz - k));
  const double2 v = a.u2[(size_t)j * a.C + col];
  const double2 hi = a.u3[(size_t)(j + 1) * a.C + col], lo = a.u3[(size_t)j * a.C + col];
  const double2 dy = d2((hi.x - lo.x) * a.dthdyv[j], (hi.y - lo.y) * a.dthdyv[j]);
  // k1F_z u2 = (0 + i kz)(x + i y) = (0 x - kz y) + i (0 y + kz x); omega = -(k1F_z u2) + dy
  const double wr = 0.0 * v.x - kz * v.y, wi = 0.0 * v.y + kz * v.x;
  return d2(-wr + dy.x, -wi + dy.y);
}

// column of -kz at the same kx, or -1 (kz = -Nz/2: none on the padded grid, itself on the unpadded one)
__device__ __forceinline__ long partner(const StatsArgs &a, size_t col, bool padded) {
  const int i = (int)(col % a.Kx), k = (int)(col / a.Kx);
  if (k == a.Nz / 2) return padded ? -1 : (long)col;
  const int kk = k == 0 ? 0 : a.Nz - k;
  return (long)kk * a.Kx + i;
}

// SNAP: Pm, P2m read a separate copy of the pressure (dns_refresh_stats_pressure); otherwise they sum p itself
template <bool SNAP>
__global__ void __launch_bounds__(STAT_NTH) plane_stats_kernel(StatsArgs a) {
  __shared__ double red[STAT_NTH / 64][STAT_NQ];
  const int chunk = blockIdx.x, j = blockIdx.y, m = blockIdx.z;
  const MemberStrides ms = member_strides(a.C, a.nyu, a.nyv);
  StatsArgs b = a;
  b.u1 += m * ms.ru; b.u3 += m * ms.ru; b.u2 += m * ms.rv; b.p += m * ms.rp;
  if (SNAP) b.pm += m * ms.rp;
  const int nyu = a.nyu;
  const bool uin = j >= 1 && j <= nyu;          // u and v planes in physical space
  const size_t C = a.C, nu = (size_t)(nyu + 2) * C, nv = (size_t)(a.nyv + 2) * C;
  double *spU = a.sp + m * (3 * nu + nv + (size_t)nyu * C);
  double *spV = spU + nu, *spW = spV + nv, *spUV = spW + nu, *spP = spUV + nu;
  double s[STAT_NQ];
#pragma unroll
  for (int q = 0; q < STAT_NQ; q++) s[q] = 0.0;
  const size_t c0 = (size_t)chunk * STAT_CHUNK, c1 = min((size_t)a.C, c0 + STAT_CHUNK);
  for (size_t col = c0 + threadIdx.x; col < c1; col += STAT_NTH) {
    const int i = (int)(col % a.Kx);
    // ---- u planes: u1, u3, u2 interpolated
    {
      const double2 x1 = b.u1[(size_t)j * a.C + col], x3 = b.u3[(size_t)j * a.C + col];
      const double2 xv = u2_on_u(b, b.u2, col, j);
      const size_t o = (size_t)j * C + col;         // spectra.f90:29-60
      spU[o] = spU[o] + x1.x * x1.x + x1.y * x1.y;
      spW[o] = spW[o] + x3.x * x3.x + x3.y * x3.y;
      spUV[o] = spUV[o] + x1.x * xv.x + x1.y * xv.y;
      if (uin) {
        if (col == 0) { s[0] += x1.x; s[2] += x3.x; }
        if (i >= 1) {
          s[1] += 2.0 * re_dot(x1, x1); s[3] += 2.0 * re_dot(x3, x3);
          s[4] += 2.0 * re_dot(x1, xv); s[5] += 2.0 * re_dot(x1, x3);
        } else {
          const long pc = partner(b, col, true);
          double t1 = re_dot(x1, x1), t3 = re_dot(x3, x3), t4 = re_dot(x1, xv), t5 = re_dot(x1, x3);
          if (pc >= 0) {
            const double2 y1 = b.u1[(size_t)j * a.C + pc], y3 = b.u3[(size_t)j * a.C + pc];
            const double2 yv = u2_on_u(b, b.u2, (size_t)pc, j);
            t1 += re_mul(x1, y1); t3 += re_mul(x3, y3); t4 += re_mul(x1, yv); t5 += re_mul(x1, y3);
          }
          s[1] += 0.5 * t1; s[3] += 0.5 * t3; s[4] += 0.5 * t4; s[5] += 0.5 * t5;
        }
      } else {  // planes 0 and nyu+1: spectral layout
        s[0] += sum_ri(x1); s[1] += re_dot(x1, x1); s[2] += sum_ri(x3); s[3] += re_dot(x3, x3);
        s[4] += re_dot(x1, xv); s[5] += re_dot(x1, x3);
      }
    }
    // ---- v planes: u2, u3 interpolated (planes 0..nyu; plane nyu+1 is never filled)
    if (j <= nyu) {
      const double2 x2 = b.u2[(size_t)j * a.C + col];
      const double2 xw = u_on_v(b, b.u3, col, j);
      spV[(size_t)j * C + col] = spV[(size_t)j * C + col] + x2.x * x2.x + x2.y * x2.y;
      if (uin) {
        if (col == 0) s[6] += x2.x;
        if (i >= 1) {
          s[7] += 2.0 * re_dot(x2, x2); s[8] += 2.0 * re_dot(x2, xw);
        } else {
          const long pc = partner(b, col, true);
          double t7 = re_dot(x2, x2), t8 = re_dot(x2, xw);
          if (pc >= 0) {
            const double2 y2 = b.u2[(size_t)j * a.C + pc];
            const double2 yw = u_on_v(b, b.u3, (size_t)pc, j);
            t7 += re_mul(x2, y2); t8 += re_mul(x2, yw);
          }
          s[7] += 0.5 * t7; s[8] += 0.5 * t8;
        }
      } else {
        s[6] += sum_ri(x2); s[7] += re_dot(x2, x2); s[8] += re_dot(x2, xw);
      }
      // ---- omega_x on the unpadded grid
      const double2 w = omega_x(b, col, j);
      if (col == 0) s[9] += w.x;
      if (i >= 1 && i < a.Nx / 2) {
        s[10] += 2.0 * re_dot(w, w);
      } else {
        const double2 wp = omega_x(b, (size_t)partner(b, col, false), j);
        s[10] += 0.5 * (re_dot(w, w) + re_mul(w, wp));
      }
    }
    // ---- p planes 1..nyu: spectral layout; the spectrum of p, the plane sums of the ppPL pressure
    if (uin) {
      const double2 x = b.p[(size_t)(j - 1) * a.C + col];
      spP[(size_t)(j - 1) * C + col] = spP[(size_t)(j - 1) * C + col] + x.x * x.x + x.y * x.y;
      double2 xm = x;  // without a separate copy, the ppPL pressure is p itself (loaded once)
      if constexpr (SNAP) xm = b.pm[(size_t)(j - 1) * a.C + col];
      s[11] += sum_ri(xm);
      s[12] += re_dot(xm, xm);
    }
  }
  // block reduction (fixed order), one partial per (member, plane, chunk)
#pragma unroll
  for (int q = 0; q < STAT_NQ; q++) {
    double v = s[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < STAT_NQ) {
    double v = 0.0;
    for (int w = 0; w < STAT_NTH / 64; w++) v += red[w][threadIdx.x];
    a.partial[(((size_t)m * (nyu + 2) + j) * a.nchunk + chunk) * STAT_NQ + threadIdx.x] = v;
  }
}

// acc[m][q][j] += factor * sum over chunks (in order)
__global__ void stats_finish_kernel(StatsArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nj = a.nyu + 2;
  if (t >= a.nm * nj * STAT_NQ) return;
  const int q = t % STAT_NQ, j = (t / STAT_NQ) % nj, m = t / (STAT_NQ * nj);
  const double *pp = a.partial + ((size_t)m * nj + j) * a.nchunk * STAT_NQ + q;
  double v = 0.0;
  for (int c = 0; c < a.nchunk; c++) v += pp[(size_t)c * STAT_NQ];
  const bool uin = j >= 1 && j <= a.nyu;
  double f = 1.0;
  if (q <= 8 && uin) f = (double)a.Mx * (double)a.Mz;
  if (q == 9 || q == 10) f = (double)a.Nx * (double)a.Nz;
  a.acc[((size_t)m * STAT_NQ + q) * nj + j] += f * v;
}

int stats_nchunk(int C) { return (C + STAT_CHUNK - 1) / STAT_CHUNK; }

void launch_stats(const StatsArgs &a, hipStream_t st) {
  if (a.pm != a.p) plane_stats_kernel<true><<<dim3(a.nchunk, a.nyu + 2, a.nm), STAT_NTH, 0, st>>>(a);
  else plane_stats_kernel<false><<<dim3(a.nchunk, a.nyu + 2, a.nm), STAT_NTH, 0, st>>>(a);
  const int n = a.nm * (a.nyu + 2) * STAT_NQ;
  stats_finish_kernel<<<(n + 255) / 256, 256, 0, st>>>(a);
}

}  // namespace dnsk
