// four_to_phys post-processing (four_to_phys.f90:60-81 and four_to_phys_N.f90: resize, cft, rft per plane):
// spectral planes of a field file to physical planes, FP64.
//
// A spectral plane is buffSR1(Nx+2, Nz): column i (0..Nx/2) as (Re, Im) pairs, z rows kz = 0..Nz/2-1
// then -Nz/2..-1.  The physical plane is buffSR2(Mx+2, Mz): u at x = 0..Mx-1 on each of the Mz rows,
// then two entries the x-transform leaves alone (the kx = Mx/2 coefficient of that row).
//
// Two kernels with the z-inverted columns in HBM between them (a 576 x 578 plane does not fit in LDS):
//   ftp_z_kernel: resize in z (rows kz >= 0 in place, kz < 0 at the top of the Mz rows; the kz = -Nz/2 row is
//     dropped when Mz > Nz and kept when Mz == Nz, as the two programs' resize routines do), then the
//     unnormalised inverse complex transform of length Mz of every retained column (cft, iopt = 1);
//   ftp_x_kernel: the unnormalised real inverse of length Mx of every row (rft, iopt = 1) as a complex
//     transform of length Mx/2 after the usual even/odd split: z_n = (c_n + conj c_{H-n}) +
//     i (c_n - conj c_{H-n}) e^{+2 pi i n/Mx}, c_0 and c_H taken real, u[2n] = Re z_n, u[2n+1] = Im z_n.
#include <mutex>

#include "dns_kernels.h"

namespace dnsk {

// Workspace of the allocation-free four_to_phys entry point: the z-inverted columns of a batch of planes
// and the twiddle tables, in a file-scope device array (each GPU holds its own copy until the process
// ends).  Host code never passes the array's name to a kernel or a copy: in host code it is a stand-in,
// not a device address.  Each device's address comes from hipGetSymbolAddress, taken once per device
// with that device current (four_to_phys_scratch).
static __device__ __align__(256) double2 scratch[FTP_SCRATCH_BYTES / sizeof(double2)];

int four_to_phys_scratch(void **ptr) {
  static std::mutex mu;
  static void *addr[64] = {};
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    addr[dev] = p;
  }
  *ptr = addr[dev];
  return 0;
}

// z-tile: T columns of one plane per workgroup (256-byte row segments where LDS allows), T * MZ complex
// in LDS; the 12.12.x plans of 576 and 1152 take 48 threads per column (one butterfly per thread and pass)
template <int MZ> struct FtpZ {
  static constexpr int T = MZ * 32 * 16 <= 150000 ? 32 : MZ * 16 * 16 <= 150000 ? 16 : 8;
  static constexpr int NTH = (MZ == 576 || MZ == 1152) ? 48 * T : 256;
};
// x-tile: RB rows per workgroup (8 rows and 256 threads measured best at 288 points)
template <int H> struct FtpX {
  static constexpr int RB = 8;
  static constexpr int NTH = 256;
};

template <int MZ, int T, int NTH, class PL>
__global__ void __launch_bounds__(NTH) ftp_z_kernel(const double *__restrict__ in, double2 *__restrict__ mid,
                                                    const double2 *__restrict__ twz, int Nx, int Nz, int drop,
                                                    int ip, int mp) {
  __shared__ double2 s[MZ * T];
  const int KX = Nx / 2 + 1, i0 = blockIdx.x * T, pl = blockIdx.y;
  const double *pin = in + (size_t)pl * ip * Nz;
  for (int idx = threadIdx.x; idx < MZ * T; idx += NTH) {
    const int z = idx / T, i = i0 + idx % T;
    const int r = z < Nz / 2 ? z : (z >= MZ - Nz / 2 + drop ? z + Nz - MZ : -1);
    double2 v = d2(0.0, 0.0);
    if (i < KX && r >= 0) v = reinterpret_cast<const double2 *>(pin + (size_t)r * ip)[i];  // ip is even
    s[idx] = v;
  }
  __syncthreads();
  fft_lds<double, Lay<MZ, T>, NTH, +1, PL>(s, twz);
  double2 *pm = mid + (size_t)pl * MZ * mp;
  for (int idx = threadIdx.x; idx < MZ * T; idx += NTH) {
    const int z = idx / T, i = i0 + idx % T;
    if (i < KX) pm[(size_t)z * mp + i] = s[idx];
  }
}

template <int H, int RB, int NTH, class PL>
__global__ void __launch_bounds__(NTH) ftp_x_kernel(const double2 *__restrict__ mid, double *__restrict__ out,
                                                    const double2 *__restrict__ twx, const double2 *__restrict__ twh,
                                                    int KX, int MZ, int mp, int op) {
  __shared__ double2 s[H * RB];
  const int z0 = blockIdx.x * RB, pl = blockIdx.y;
  const double2 *pm = mid + (size_t)pl * MZ * mp;
  double *po = out + (size_t)pl * MZ * op;
  for (int idx = threadIdx.x; idx < H * RB; idx += NTH) {
    const int r = idx / H, n = idx % H, z = z0 + r;
    double2 xa = d2(0.0, 0.0), xb = d2(0.0, 0.0);
    if (z < MZ) {
      const double2 *row = pm + (size_t)z * mp;
      const double2 cn = n < KX ? row[n] : d2(0.0, 0.0);
      xa = d2(cn.x, n == 0 ? 0.0 : cn.y);
      if (n == 0) {
        xb = d2(H < KX ? row[H].x : 0.0, 0.0);
      } else {
        xb = H - n < KX ? row[H - n] : d2(0.0, 0.0);
      }
    }
    const double2 sum = cadd(xa, cconj(xb)), dif = csub(xa, cconj(xb));
    s[n * RB + r] = cadd(sum, cmul_si<1>(cmul(dif, cconj(twx[n]))));
  }
  __syncthreads();
  fft_lds<double, Lay<H, RB>, NTH, +1, PL>(s, twh);
  for (int idx = threadIdx.x; idx < H * RB; idx += NTH) {
    const int r = idx / H, n = idx % H, z = z0 + r;
    if (z < MZ) {
      reinterpret_cast<double2 *>(po + (size_t)z * op)[n] = s[n * RB + r];   // op is even
    }
  }
  // entries Mx, Mx+1: the row's kx = Mx/2 coefficient, untouched by the x-transform
  if (threadIdx.x < RB) {
    const int z = z0 + threadIdx.x;
    if (z < MZ) {
      reinterpret_cast<double2 *>(po + (size_t)z * op)[H] = H < KX ? pm[(size_t)z * mp + H] : d2(0.0, 0.0);
    }
  }
}

// lengths 2^a 3^b of the transforms (z: Mz; x: Mx/2)
#define FTP_SIZES(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) \
  X(288) X(384) X(576) X(768) X(1152)

bool four_to_phys_supported(int Mx, int Mz) {
  bool zx = false, xx = false;
#define CASE(N) zx = zx || Mz == N; xx = xx || Mx / 2 == N;
  FTP_SIZES(CASE)
#undef CASE
  return zx && xx && Mx % 2 == 0;
}

int launch_four_to_phys(const FtpArgs &a, hipStream_t st) {
  const int KX = a.Nx / 2 + 1;
  const int drop = a.Mz > a.Nz ? 1 : 0;
  bool ok = false;
#define CASE(N)                                                                                        \
  if (a.Mz == N) {                                                                                     \
    using Z = FtpZ<N>;                                                                                 \
    ftp_z_kernel<N, Z::T, Z::NTH, typename ZPlanT<N>::type>                                          \
        <<<dim3((KX + Z::T - 1) / Z::T, a.n), Z::NTH, 0, st>>>(a.in, a.mid, a.twz, a.Nx, a.Nz, drop,   \
                                                               a.in_pitch, a.mid_pitch);               \
    ok = true;                                                                                         \
  }
  FTP_SIZES(CASE)
#undef CASE
  if (!ok) return -1;
  ok = false;
#define CASE(N)                                                                                       \
  if (a.Mx / 2 == N) {                                                                                \
    using X = FtpX<N>;                                                                                \
    ftp_x_kernel<N, X::RB, X::NTH, typename XPlanT<N>::type>                                          \
        <<<dim3((a.Mz + X::RB - 1) / X::RB, a.n), X::NTH, 0, st>>>(a.mid, a.out, a.twx, a.twh, KX, a.Mz, \
                                                                   a.mid_pitch, a.out_pitch);           \
    ok = true;                                                                                        \
  }
  FTP_SIZES(CASE)
#undef CASE
  return ok ? 0 : -1;
}

}  // namespace dnsk
