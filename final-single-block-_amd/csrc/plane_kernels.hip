// Plane stage of the nonlinear term (FOU3D.f90:51-371, ops_in_planes2 FOU3D.f90:1514-1652).
//
// Per y-plane the reference runs: zero-padded inverse FFT in z on the Kx retained x-modes (cft, pruned),
// inverse real FFT in x on all Mz rows (rft), pointwise products, forward real FFT in x, forward FFT in z
// on the Kx x-modes, the spectral derivatives der_x/der_z with their dealias mask, and keeps vv, uv, wv
// for the y-flux.  Here that work is three kernels over a batch of planes:
//   zinv_kernel : spectral plane (device layout [row][column]) -> I[plane][field][z][i]   (z-inverse)
//                 The y-interpolations interp_u/interp_v (FOU3D.f90:373-418) are fused into its loads.
//   xpass_kernel: I rows -> physical rows -> products -> P[plane][product][z][i]          (x-pass)
//                 Also the CFL maxima of dtc_calc (FOU3D.f90:613-651) at kRK = 1.
//   zfwd_kernel : P -> forward z FFT -> der_x + der_z (masked) and raw fluxes in device layout.
// Plane index w: w < nyu is u-plane j = w+1 (u1, u2_itp, u3 -> uu, uw, vv, ww); otherwise v-plane
// j = w-nyu+1 (u1_itp, u2, u3_itp -> uv, wv).  v-planes run to j = nyv+1 as the reference does on one
// rank (start.f90:1128-1180 gives the single rank v-planes 1..nyu).
#include "dns_kernels.h"

namespace dnsk {

// ------------------------------------------------------------------------------------------------
template <typename R, int MZ, int T, int NTH>
__global__ void __launch_bounds__(NTH) zinv_kernel(PlaneArgs a) {
  __shared__ cx<R> s[MZ * T];
  const int w = a.w0 + blockIdx.y;
  const int i0 = blockIdx.x * T;
  const int Kx = a.Kx, Nz = a.Nz, C = a.C;
  const bool vpl = w >= a.nyu;
  const int j = vpl ? w - a.nyu + 1 : w + 1;
  const cx<R> *tw = reinterpret_cast<const cx<R> *>(a.twz);
  cx<R> *I = reinterpret_cast<cx<R> *>(a.I) + (size_t)(w - a.w0) * 3 * MZ * Kx;
  // interpolation weights (interp_v for u-planes, interp_u for v-planes)
  double wa, wb, dd;
  if (!vpl) {
    wa = a.yu[j] - a.yv[j - 1];
    wb = a.yv[j] - a.yu[j];
    dd = a.yv[j] - a.yv[j - 1];
  } else {
    wa = a.yv[j] - a.yu[j];
    wb = a.yu[j + 1] - a.yv[j];
    dd = a.yu[j + 1] - a.yu[j];
  }
  for (int f = 0; f < 3; f++) {
    for (int idx = threadIdx.x; idx < MZ * T; idx += NTH) {
      const int z = idx / T, t = idx - z * T, i = i0 + t;
      double2 v = d2(0.0, 0.0);
      const bool kept = (z < Nz / 2) || (z >= MZ - Nz / 2);
      if (kept && i < Kx) {
        const int k = z < Nz / 2 ? z : z - (MZ - Nz);
        const size_t col = (size_t)k * Kx + i;
        if (!vpl) {
          if (f == 0) v = a.u1[(size_t)j * C + col];
          else if (f == 2) v = a.u3[(size_t)j * C + col];
          else {
            const double2 x1 = a.u2[(size_t)j * C + col], x0 = a.u2[(size_t)(j - 1) * C + col];
            v = d2((wa * x1.x + wb * x0.x) / dd, (wa * x1.y + wb * x0.y) / dd);
          }
        } else {
          if (f == 1) v = a.u2[(size_t)j * C + col];
          else {
            const double2 *u = f == 0 ? a.u1 : a.u3;
            const double2 x1 = u[(size_t)(j + 1) * C + col], x0 = u[(size_t)j * C + col];
            v = d2((wa * x1.x + wb * x0.x) / dd, (wa * x1.y + wb * x0.y) / dd);
          }
        }
      }
      s[idx] = from_d2<R>(v);
    }
    __syncthreads();
    fft_lds<R, MZ, T, NTH, +1>(s, tw);
    for (int idx = threadIdx.x; idx < MZ * T; idx += NTH) {
      const int z = idx / T, t = idx - z * T, i = i0 + t;
      if (i < Kx) I[((size_t)f * MZ + z) * Kx + i] = s[idx];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------------
// x-pass.  A real row x of length MX is packed as z_m = x_{2m} + i x_{2m+1} (H = MX/2 complex), so one
// length-H complex FFT does each real transform.  Inverse (rft_buff.f:76-84 semantics: Im c_0 and
// Im c_{MX/2} ignored, unnormalised Hermitian synthesis):
//   Z_n = (X_n + conj X_{H-n}) + i e^{+2 pi i n/MX} (X_n - conj X_{H-n}),  z = IFFT_H(Z)
// Forward (rft_buff.f:62-74: scaled 1/MX, Im c_0 = 0):
//   X_k = [ (Z_k + conj Z_{H-k}) - i e^{-2 pi i k/MX} (Z_k - conj Z_{H-k}) ] / (2 MX),  Z = FFT_H(z)
template <typename R, int MX, int RB, int NTH>
__global__ void __launch_bounds__(NTH) xpass_kernel(PlaneArgs a) {
  constexpr int H = MX / 2;
  constexpr int NA = RB * 3, NB = RB * 4;
  __shared__ cx<R> A[H * NA];
  __shared__ cx<R> B[H * NB];
  __shared__ double red[NTH / 64];
  const int w = a.w0 + blockIdx.y;
  const int z0 = blockIdx.x * RB;
  const int Kx = a.Kx, MZ = a.Mz;
  const bool vpl = w >= a.nyu;
  const int j = vpl ? w - a.nyu + 1 : w + 1;
  const cx<R> *twx = reinterpret_cast<const cx<R> *>(a.twx);  // e^{-2 pi i m/MX}, m < MX
  const cx<R> *twh = reinterpret_cast<const cx<R> *>(a.twh);  // e^{-2 pi i m/H}
  const cx<R> *I = reinterpret_cast<const cx<R> *>(a.I) + (size_t)(w - a.w0) * 3 * MZ * Kx;
  cx<R> *P = reinterpret_cast<cx<R> *>(a.P) + (size_t)(w - a.w0) * 4 * MZ * Kx;

  for (int idx = threadIdx.x; idx < H * NA; idx += NTH) {
    const int slot = idx / H, n = idx - slot * H;
    const int r = slot / 3, f = slot - r * 3;
    const cx<R> *row = I + ((size_t)f * MZ + (z0 + r)) * Kx;
    cx<R> xa = mk<R>(0, 0), xb = mk<R>(0, 0);
    if (n < Kx) xa = row[n];
    if (n == 0) xa.y = 0;
    if (n > 0 && H - n < Kx) xb = row[H - n];
    const cx<R> sum = cadd(xa, cconj(xb)), dif = csub(xa, cconj(xb));
    const cx<R> e = cconj(twx[n]);                        // e^{+2 pi i n/MX}
    A[n * NA + slot] = cadd(sum, cmul_si<1>(cmul(dif, e)));
  }
  __syncthreads();
  fft_lds<R, H, NA, NTH, +1>(A, twh);

  R m1 = 0, m3 = 0;
  const bool cfl = (a.kRK == 1) && !vpl && (j <= a.nyu - 1);
  for (int idx = threadIdx.x; idx < H * RB; idx += NTH) {
    const int r = idx / H, n = idx - r * H;
    const cx<R> zu = A[n * NA + r * 3 + 0], zv = A[n * NA + r * 3 + 1], zw = A[n * NA + r * 3 + 2];
    cx<R> *b = B + n * NB + r * 4;
    if (!vpl) {
      b[0] = mk<R>(zu.x * zu.x, zu.y * zu.y);  // uu
      b[1] = mk<R>(zu.x * zw.x, zu.y * zw.y);  // uw
      b[2] = mk<R>(zv.x * zv.x, zv.y * zv.y);  // vv
      b[3] = mk<R>(zw.x * zw.x, zw.y * zw.y);  // ww
      if (cfl) {
        m1 = fmax(m1, fmax(fabs(zu.x), fabs(zu.y)));
        m3 = fmax(m3, fmax(fabs(zw.x), fabs(zw.y)));
      }
    } else {
      b[0] = mk<R>(zu.x * zv.x, zu.y * zv.y);  // uv
      b[1] = mk<R>(zw.x * zv.x, zw.y * zv.y);  // wv
      b[2] = mk<R>(0, 0);
      b[3] = mk<R>(0, 0);
    }
  }
  if (cfl) {
    const double b1 = block_max<NTH>((double)m1, red);
    const double b3 = block_max<NTH>((double)m3, red);
    if (threadIdx.x == 0) {
      atomic_max_nonneg(&a.cfl[0], b1);
      atomic_max_nonneg(&a.cfl[1], b3);
    }
  }
  __syncthreads();
  fft_lds<R, H, NB, NTH, -1>(B, twh);

  const int np = vpl ? 2 : 4;
  const R sc = (R)(0.5 / (double)MX);
  for (int idx = threadIdx.x; idx < Kx * RB * np; idx += NTH) {
    const int q = idx / Kx, k = idx - q * Kx;
    const int r = q / np, p = q - r * np;
    const cx<R> zk = B[k * NB + r * 4 + p];
    const cx<R> zm = cconj(B[((H - k) % H) * NB + r * 4 + p]);
    const cx<R> sum = cadd(zk, zm), dif = csub(zk, zm);
    cx<R> x = csub(sum, cmul_si<1>(cmul(dif, twx[k])));
    x = cscale(x, sc);
    if (k == 0) x.y = 0;
    P[((size_t)p * MZ + (z0 + r)) * Kx + k] = x;
  }
}

// ------------------------------------------------------------------------------------------------
template <typename R, int MZ, int T, int NTH>
__global__ void __launch_bounds__(NTH) zfwd_kernel(PlaneArgs a) {
  __shared__ cx<R> s[MZ * T];
  constexpr int PERMAX = ((2 * MZ / 3) * T + NTH - 1) / NTH;
  const int w = a.w0 + blockIdx.y;
  const int i0 = blockIdx.x * T;
  const int Kx = a.Kx, Nz = a.Nz, C = a.C;
  const bool vpl = w >= a.nyu;
  const int j = vpl ? w - a.nyu + 1 : w + 1;
  const cx<R> *tw = reinterpret_cast<const cx<R> *>(a.twz);
  const cx<R> *P = reinterpret_cast<const cx<R> *>(a.P) + (size_t)(w - a.w0) * 4 * MZ * Kx;
  const double inv = 1.0 / (double)MZ;
  double2 acc0[PERMAX], acc1[PERMAX], acc2[PERMAX];
  const int np = vpl ? 2 : 4;
  for (int p = 0; p < np; p++) {
    for (int idx = threadIdx.x; idx < MZ * T; idx += NTH) {
      const int z = idx / T, t = idx - z * T, i = i0 + t;
      s[idx] = i < Kx ? P[((size_t)p * MZ + z) * Kx + i] : mk<R>(0, 0);
    }
    __syncthreads();
    fft_lds<R, MZ, T, NTH, -1>(s, tw);
#pragma unroll
    for (int q = 0; q < PERMAX; q++) {
      const int idx = threadIdx.x + q * NTH;
      if (idx < Nz * T) {
        const int k = idx / T, t = idx - k * T, i = i0 + t;
        const int zp = k < Nz / 2 ? k : k + MZ - Nz;
        double2 v = to_d2(s[zp * T + t]);
        v = d2(inv * v.x, inv * v.y);
        const double kx = a.alp * (double)i;
        const double kz = k < Nz / 2 ? a.bet * (double)k : -(a.bet * (double)(Nz - k));
        if (!vpl) {
          if (p == 0) acc0[q] = imul(kx, v);                    // der_x(uu)
          else if (p == 1) { acc0[q] = cadd(acc0[q], imul(kz, v)); acc1[q] = imul(kx, v); }  // der_z(uw), der_x(uw)
          else if (p == 2) acc2[q] = v;                         // vv (flux, unmasked)
          else acc1[q] = cadd(acc1[q], imul(kz, v));            // der_z(ww)
        } else {
          if (p == 0) { acc0[q] = imul(kx, v); acc1[q] = v; }   // der_x(uv), uv flux
          else { acc0[q] = cadd(acc0[q], imul(kz, v)); acc2[q] = v; }  // der_z(wv), wv flux
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < PERMAX; q++) {
    const int idx = threadIdx.x + q * NTH;
    if (idx < Nz * T) {
      const int k = idx / T, t = idx - k * T, i = i0 + t;
      if (i < Kx) {
        const size_t col = (size_t)k * Kx + i;
        const bool masked = (k == Nz / 2);  // kz = -Nz/2 row is inside der_x/der_z's zero band
        if (!vpl) {
          a.Nxz1[(size_t)(j - 1) * C + col] = masked ? d2(0.0, 0.0) : acc0[q];
          a.Nxz3[(size_t)(j - 1) * C + col] = masked ? d2(0.0, 0.0) : acc1[q];
          a.vvc[(size_t)j * C + col] = acc2[q];
        } else {
          if (j <= a.nyv) a.Nxz2[(size_t)(j - 1) * C + col] = masked ? d2(0.0, 0.0) : acc0[q];
          a.uvf[(size_t)j * C + col] = acc1[q];
          a.wvf[(size_t)j * C + col] = acc2[q];
        }
      }
    }
  }
}

// ghost plane 0 of u1/u3 is still spectral when dtc_calc reads it (FOU3D.f90:622-626 against the
// allocation at start.f90:549-551): its |Re|, |Im| join the CFL maxima.
__global__ void cfl_ghost_kernel(const double2 *u1, const double2 *u3, int C, double *cfl) {
  __shared__ double red[4];
  double m1 = 0.0, m3 = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
    const double2 a = u1[c], b = u3[c];
    m1 = fmax(m1, fmax(fabs(a.x), fabs(a.y)));
    m3 = fmax(m3, fmax(fabs(b.x), fabs(b.y)));
  }
  m1 = block_max<256>(m1, red);
  m3 = block_max<256>(m3, red);
  if (threadIdx.x == 0) {
    atomic_max_nonneg(&cfl[0], m1);
    atomic_max_nonneg(&cfl[1], m3);
  }
}

__global__ void zero_kernel(double *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------------------------------------------------
// reference-format plane FFTs, for dns_plane_fft_roundtrip: one plane per block, FP64.
template <int MX, int MZ, int NTH>
__global__ void __launch_bounds__(NTH) plane_roundtrip_kernel(const double *in, double *phys, double *back,
                                                               const double2 *twx, const double2 *twh,
                                                               const double2 *twz) {
  constexpr int H = MX / 2, IG = MX + 2, KX = (2 * MX / 3) / 2 + 1;
  constexpr int TZ = 8;
  __shared__ double2 s[(MZ * TZ > H * 8) ? MZ * TZ : H * 8];
  const double *pin = in + (size_t)blockIdx.x * IG * MZ;
  double *pph = phys + (size_t)blockIdx.x * IG * MZ;
  double *pbk = back + (size_t)blockIdx.x * IG * MZ;
  // copy in -> phys (whole plane), zero row MZ/2 (FOU3D.f90:693)
  for (int e = threadIdx.x; e < IG * MZ; e += NTH) pph[e] = (e / IG == MZ / 2) ? 0.0 : pin[e];
  __syncthreads();
  // z-inverse on the KX retained x-modes, TZ columns at a time
  for (int i0 = 0; i0 < KX; i0 += TZ) {
    for (int idx = threadIdx.x; idx < MZ * TZ; idx += NTH) {
      const int z = idx / TZ, t = idx % TZ, i = i0 + t;
      s[idx] = i < KX ? d2(pph[z * IG + 2 * i], pph[z * IG + 2 * i + 1]) : d2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, MZ, TZ, NTH, +1>(s, twz);
    for (int idx = threadIdx.x; idx < MZ * TZ; idx += NTH) {
      const int z = idx / TZ, t = idx % TZ, i = i0 + t;
      if (i < KX) { pph[z * IG + 2 * i] = s[idx].x; pph[z * IG + 2 * i + 1] = s[idx].y; }
    }
    __syncthreads();
  }
  // x-inverse on every row, 8 rows at a time (entries MX, MX+1 kept)
  for (int z0 = 0; z0 < MZ; z0 += 8) {
    for (int idx = threadIdx.x; idx < H * 8; idx += NTH) {
      const int r = idx / H, n = idx % H, z = z0 + r;
      double2 xa = d2(0.0, 0.0), xb = d2(0.0, 0.0);
      if (z < MZ) {
        xa = d2(pph[z * IG + 2 * n], n == 0 ? 0.0 : pph[z * IG + 2 * n + 1]);
        xb = n == 0 ? d2(pph[z * IG + MX], 0.0) : d2(pph[z * IG + 2 * (H - n)], pph[z * IG + 2 * (H - n) + 1]);
      }
      const double2 sum = cadd(xa, cconj(xb)), dif = csub(xa, cconj(xb));
      s[n * 8 + r] = cadd(sum, cmul_si<1>(cmul(dif, cconj(twx[n]))));
    }
    __syncthreads();
    fft_lds<double, H, 8, NTH, +1>(s, twh);
    for (int idx = threadIdx.x; idx < H * 8; idx += NTH) {
      const int r = idx / H, n = idx % H, z = z0 + r;
      if (z < MZ) { pph[z * IG + 2 * n] = s[n * 8 + r].x; pph[z * IG + 2 * n + 1] = s[n * 8 + r].y; }
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < IG * MZ; e += NTH) pbk[e] = pph[e];
  __syncthreads();
  // x-forward on every row
  for (int z0 = 0; z0 < MZ; z0 += 8) {
    for (int idx = threadIdx.x; idx < H * 8; idx += NTH) {
      const int r = idx / H, n = idx % H, z = z0 + r;
      s[n * 8 + r] = z < MZ ? d2(pbk[z * IG + 2 * n], pbk[z * IG + 2 * n + 1]) : d2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, H, 8, NTH, -1>(s, twh);
    for (int idx = threadIdx.x; idx < (H + 1) * 8; idx += NTH) {
      const int r = idx / (H + 1), k = idx % (H + 1), z = z0 + r;
      if (z < MZ) {
        const double2 zk = s[(k % H) * 8 + r], zm = cconj(s[((H - k) % H) * 8 + r]);
        const double2 sum = cadd(zk, zm), dif = csub(zk, zm);
        const double2 e = k < H ? twx[k] : d2(-1.0, 0.0);
        double2 x = csub(sum, cmul_si<1>(cmul(dif, e)));
        x = cscale(x, 0.5 / MX);
        if (k == 0 || k == H) x.y = 0.0;
        pbk[z * IG + 2 * k] = x.x;
        pbk[z * IG + 2 * k + 1] = x.y;
      }
    }
    __syncthreads();
  }
  // z-forward on the KX retained x-modes, scaled 1/MZ; zero row MZ/2 (FOU3D.f90:1670)
  for (int i0 = 0; i0 < KX; i0 += TZ) {
    for (int idx = threadIdx.x; idx < MZ * TZ; idx += NTH) {
      const int z = idx / TZ, t = idx % TZ, i = i0 + t;
      s[idx] = i < KX ? d2(pbk[z * IG + 2 * i], pbk[z * IG + 2 * i + 1]) : d2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<double, MZ, TZ, NTH, -1>(s, twz);
    for (int idx = threadIdx.x; idx < MZ * TZ; idx += NTH) {
      const int z = idx / TZ, t = idx % TZ, i = i0 + t;
      if (i < KX) {
        pbk[z * IG + 2 * i] = s[idx].x / MZ;
        pbk[z * IG + 2 * i + 1] = s[idx].y / MZ;
      }
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < IG; e += NTH) pbk[(MZ / 2) * IG + e] = 0.0;
}

// ------------------------------------------------------------------------------------------------
// launchers
template <int MX> constexpr int xpass_rows() { return MX >= 1152 ? 1 : (MX >= 576 ? 2 : 4); }

int xpass_rows_per_block(int Mx) { return Mx >= 1152 ? 1 : (Mx >= 576 ? 2 : 4); }

template <typename R, int MZ>
static void launch_zinv_t(const PlaneArgs &a, int nw, hipStream_t st) {
  constexpr int T = 8, NTH = 256;
  dim3 g((a.Kx + T - 1) / T, nw);
  zinv_kernel<R, MZ, T, NTH><<<g, NTH, 0, st>>>(a);
}
template <typename R, int MZ>
static void launch_zfwd_t(const PlaneArgs &a, int nw, hipStream_t st) {
  constexpr int T = 8, NTH = 256;
  dim3 g((a.Kx + T - 1) / T, nw);
  zfwd_kernel<R, MZ, T, NTH><<<g, NTH, 0, st>>>(a);
}
template <typename R, int MX>
static void launch_xpass_t(const PlaneArgs &a, int nw, hipStream_t st) {
  constexpr int RB = xpass_rows<MX>();
  constexpr int NTH = 256;
  dim3 g(a.Mz / RB, nw);
  xpass_kernel<R, MX, RB, NTH><<<g, NTH, 0, st>>>(a);
}

#define DNS_SIZES(X) X(12) X(24) X(48) X(96) X(144) X(192) X(288) X(384) X(576) X(1152)

bool plane_size_supported(int n) {
#define CASE(N) if (n == N) return true;
  DNS_SIZES(CASE)
#undef CASE
  return false;
}

template <typename R>
static int launch_plane_stage_t(const PlaneArgs &a, int nw, hipStream_t st) {
  bool ok = false;
#define CASEZ(N) if (a.Mz == N) { launch_zinv_t<R, N>(a, nw, st); ok = true; }
  DNS_SIZES(CASEZ)
#undef CASEZ
  if (!ok) return -1;
  ok = false;
#define CASEX(N) if (a.Mx == N) { launch_xpass_t<R, N>(a, nw, st); ok = true; }
  DNS_SIZES(CASEX)
#undef CASEX
  if (!ok) return -1;
  ok = false;
#define CASEZ2(N) if (a.Mz == N) { launch_zfwd_t<R, N>(a, nw, st); ok = true; }
  DNS_SIZES(CASEZ2)
#undef CASEZ2
  return ok ? 0 : -1;
}

int launch_plane_stage(const PlaneArgs &a, int nw, bool fp32, hipStream_t st) {
  return fp32 ? launch_plane_stage_t<float>(a, nw, st) : launch_plane_stage_t<double>(a, nw, st);
}

int launch_plane_stage_split(const PlaneArgs &a, int nw, bool fp32, hipStream_t st, int which) {
  bool ok = false;
  if (which == 0) {
#define CZ(N) if (a.Mz == N) { if (fp32) launch_zinv_t<float, N>(a, nw, st); else launch_zinv_t<double, N>(a, nw, st); ok = true; }
    DNS_SIZES(CZ)
#undef CZ
  } else if (which == 1) {
#define CXP(N) if (a.Mx == N) { if (fp32) launch_xpass_t<float, N>(a, nw, st); else launch_xpass_t<double, N>(a, nw, st); ok = true; }
    DNS_SIZES(CXP)
#undef CXP
  } else {
#define CZF(N) if (a.Mz == N) { if (fp32) launch_zfwd_t<float, N>(a, nw, st); else launch_zfwd_t<double, N>(a, nw, st); ok = true; }
    DNS_SIZES(CZF)
#undef CZF
  }
  return ok ? 0 : -1;
}

void launch_cfl_ghost(const double2 *u1, const double2 *u3, int C, double *cfl, hipStream_t st) {
  int nb = (C + 255) / 256;
  if (nb > 1024) nb = 1024;
  cfl_ghost_kernel<<<nb, 256, 0, st>>>(u1, u3, C, cfl);
}

void launch_zero(double *p, int n, hipStream_t st) { zero_kernel<<<(n + 255) / 256, 256, 0, st>>>(p, n); }

int launch_plane_roundtrip(int Mx, int Mz, const double *in, double *phys, double *back, int n,
                           const double2 *twx, const double2 *twh, const double2 *twz, hipStream_t st) {
#define RT(MXV, MZV) if (Mx == MXV && Mz == MZV) { plane_roundtrip_kernel<MXV, MZV, 256><<<n, 256, 0, st>>>(in, phys, back, twx, twh, twz); return 0; }
  RT(12, 12) RT(24, 24) RT(48, 48) RT(96, 96) RT(192, 192) RT(576, 576) RT(144, 144) RT(288, 288)
#undef RT
  return -1;
}

}  // namespace dnsk
