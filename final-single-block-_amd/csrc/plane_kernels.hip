// Plane stage of the nonlinear term: CFL helpers, the reference-format plane round trip, and the
// launchers of the plane kernels (plane_kernels.h) for every supported transform length.
#include "plane_launch.h"

namespace dnsk {

// ghost plane 0 of u1/u3 is still spectral when dtc_calc reads it (FOU3D.f90:622-626 against the
// allocation at start.f90:549-551): its |Re|, |Im| join the CFL maxima.
__global__ void cfl_ghost_kernel(const double2 *u1, const double2 *u3, int C, size_t ms, double *cfl) {
  __shared__ double red[4];
  u1 += blockIdx.y * ms;  // ensemble member blockIdx.y (member stride ms)
  u3 += blockIdx.y * ms;
  cfl += (size_t)blockIdx.y * 2 * CFL_SLOTS;
  double m1 = 0.0, m3 = 0.0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x) {
    const double2 a = u1[c], b = u3[c];
    m1 = fmax(m1, fmax(fabs(a.x), fabs(a.y)));
    m3 = fmax(m3, fmax(fabs(b.x), fabs(b.y)));
  }
  m1 = block_max<256>(m1, red);
  m3 = block_max<256>(m3, red);
  if (threadIdx.x == 0) {
    const int slot = blockIdx.x & (CFL_SLOTS - 1);
    atomic_max_nonneg(&cfl[slot], m1);
    atomic_max_nonneg(&cfl[CFL_SLOTS + slot], m3);
  }
}

// max over the slots -> out[0] = max|u1|, out[1] = max|u3| (one block of CFL_SLOTS threads)
__global__ void cfl_reduce_kernel(const double *slots, DnsState *st) {
  const int l = threadIdx.x;
  slots += (size_t)blockIdx.x * 2 * CFL_SLOTS;  // ensemble member blockIdx.x
  double *out = &st[blockIdx.x].cfl[0];
  double m1 = slots[l], m3 = slots[CFL_SLOTS + l];
  static_assert(CFL_SLOTS == 64, "one wavefront");
  for (int off = 32; off > 0; off >>= 1) {
    m1 = fmax(m1, __shfl_xor(m1, off, 64));
    m3 = fmax(m3, __shfl_xor(m3, off, 64));
  }
  if (l == 0) {
    out[0] = m1;
    out[1] = m3;
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
    fft_lds<double, Lay<MZ, TZ>, NTH, +1>(s, twz);
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
    fft_lds<double, Lay<H, 8>, NTH, +1>(s, twh);
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
    fft_lds<double, Lay<H, 8>, NTH, -1>(s, twh);
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
    fft_lds<double, Lay<MZ, TZ>, NTH, -1>(s, twz);
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

bool plane_size_supported(int n) {
#define CASE(N) if (n == N) return true;
  DNS_SIZES(CASE)
#undef CASE
  return false;
}

int xpass_rows_per_block(int Mx) { (void)Mx; return 1; }  // every Mz: tiles that do not divide it fall back

int plane_variants(int n) { return DNS_TUNED(n) ? 3 : 1; }

// which: 0 z-inverse, 1 x-pass u-planes, 2 x-pass v-planes, 3 z-forward u-planes, 4 z-forward v-planes.
// The batch is a.j0, a.nbj; nw = blockIdx.y extent (a.nbj + 1 field rows for the z-inverse, else a.nbj).
static int launch_which(const PlaneArgs &a, int nw, bool fp32, int which, int v, hipStream_t st) {
  switch (which) {
    case 0: return launch_zinv(a, nw, fp32, v, st);
    case 1: case 2:
      return fp32 ? launch_xpass_f32(a, nw, which == 2, v, st) : launch_xpass_f64(a, nw, which == 2, v, st);
    case 3: case 4: return launch_zfwd(a, nw, fp32, which == 4, v, st);
  }
  return -1;
}

// The x-pass's twiddle table twh holds H = Mx/2 entries e^{-2 pi i m/H} followed by H more, the register
// x-pass's lane twiddles twr (xreg_twiddles_kernel); this fills the second half.
int build_xpass_twiddles(void *twh, int Mx, bool fp32, hipStream_t st) {
#define CASE(N)                                                                                          \
  if (Mx == N) {                                                                                         \
    if (fp32)                                                                                            \
      xreg_twiddles_kernel<float, N / 2><<<(N / 2 + 255) / 256, 256, 0, st>>>(static_cast<cx<float> *>(twh)); \
    else                                                                                                 \
      xreg_twiddles_kernel<double, N / 2><<<(N / 2 + 255) / 256, 256, 0, st>>>(static_cast<cx<double> *>(twh)); \
    return 0;                                                                                            \
  }
  DNS_SIZES(CASE)
#undef CASE
  return -1;
}

int launch_plane_kernel(const PlaneArgs &a, int nw, bool fp32, int which, int variant, hipStream_t st) {
  if (a.Kx != a.Mx / 3 + 1 || a.Nz != 2 * a.Mz / 3) return -1;  // the kernels take both from Mx, Mz
  return launch_which(a, nw, fp32, which, variant, st);
}

void launch_cfl_ghost(const double2 *u1, const double2 *u3, int C, size_t ms, int nm, double *cfl, hipStream_t st) {
  int nb = (C + 255) / 256;
  if (nb > 1024) nb = 1024;
  cfl_ghost_kernel<<<dim3(nb, nm), 256, 0, st>>>(u1, u3, C, ms, cfl);
}

void launch_cfl_reduce(const double *slots, DnsState *out, int nm, hipStream_t st) {
  cfl_reduce_kernel<<<nm, CFL_SLOTS, 0, st>>>(slots, out);
}

// start of kRK = 1: CFL maxima, max|div| and the slots of every member to zero
__global__ void cfl_reset_kernel(DnsState *st, double *slots, int nm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nm) {
    st[i].cfl[0] = 0.0;
    st[i].cfl[1] = 0.0;
    st[i].divmax = 0.0;
  }
  if (i < nm * 2 * CFL_SLOTS) slots[i] = 0.0;
}
void launch_cfl_reset(DnsState *st, double *slots, int nm, hipStream_t s) {
  const int n = nm * 2 * CFL_SLOTS;
  cfl_reset_kernel<<<(n + 255) / 256, 256, 0, s>>>(st, slots, nm);
}

__global__ void divmax_reset_kernel(DnsState *st, int nm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nm) st[i].divmax = 0.0;
}
void launch_divmax_reset(DnsState *st, int nm, hipStream_t s) {
  divmax_reset_kernel<<<(nm + 255) / 256, 256, 0, s>>>(st, nm);
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
