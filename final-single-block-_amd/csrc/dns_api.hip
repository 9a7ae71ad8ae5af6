// Host side of the C-ABI (include/dns_hip.h): set-up restated from start.f90 for one rank, device
// memory, and the launch sequences of the staged and fused RK3 substeps.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <vector>

#include "../../include/dns_hip.h"
#include "dns_kernels.h"

using dnsk::ColArgs;
using dnsk::DnsState;
using dnsk::PlaneArgs;

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) return fail(DNS_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

extern "C" const char *dns_last_error(void) { return g_err.c_str(); }

struct dns_handle {
  dns_params prm;
  int Mx, Mz, Nx, Nz, Kx, C, ny, nyu, nyv;
  bool fp32;
  int bj;  // plane index j per plane-stage group (u- and v-planes of each j)
  // variant per kernel: z-inverse, x-pass u/v, z-forward u/v, column sweeps up/down/final, one-pass column
  int var[DNS_NKERN] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool colpass = false;  // column stage: one-pass kernel (true) or the three sweeps
  double *stat_acc = nullptr, *stat_sp = nullptr, *stat_part = nullptr;  // statistics (allocated on first use)
  double2 *p_stat = nullptr;       // the pressure Pm, P2m sum after dns_refresh_stats_pressure (ppPL)
  int istat = 0;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  bool timing = false;
  int nm = 1;                      // ensemble members (same sizes, own Re, mpgx, state)
  std::vector<char> uploaded;      // per member
  std::vector<double> Re_m, mpgx_m;
  double dtv_den = 1.0;            // dtv = Re / dtv_den
  // host set-up (start.f90)
  std::vector<double> yu, yv, dthdyu, dthdyv, dyu2i, dyv2i;
  double dtv, gw1, gw2, gbc, gi1, gi2;
  double aRK[3], bRK[3], gRK[3], cRK[3], dRK[3];
  // device
  double2 *u1, *u2, *u3, *p, *psi, *div, *du1, *du2, *du3, *Nu1, *Nu2, *Nu3;
  double2 *Nxz1, *Nxz2, *Nxz3, *uvf, *vvc, *wvf;
  double *w3u, *w3v, *cP, *mscr;
  double *d_yu, *d_yv, *d_dthdyu, *d_dthdyv, *d_dyu2i, *d_dyv2i;
  void *I, *P, *twz, *twx, *twh;
  double2 *twz64, *twx64, *twh64;  // FP64 twiddles for the reference-format plane utilities
  double2 *stage;                  // layout conversion buffer
  double2 *stage2 = nullptr;       // FP64 copy of an FP32 array (mixed mode)
  DnsState *st;
  double *cfls = nullptr;  // CFL maxima spread over slots (plane stage at kRK = 1)
  std::vector<void *> allocs;
  hipGraphExec_t graph = nullptr;
  hipEvent_t ev[8];
  double tsum[7];
  int tcount;
};

// ------------------------------------------------------------------------------------------------
// set-up (restates start.f90 for one rank)
static double powi(double a, int b) {  // x**n with an integer exponent, by binary exponentiation
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

// ygrid : start.f90:665-806
static void host_grid(dns_handle *h) {
  const int nn = h->nyv + 1, ppp = h->prm.ppp;
  const double dyq = h->prm.dyq;
  const double qqq = nn / 2.0;
  const double bbb = (dyq - 1.0) /
      (qqq * (1.0 - powi(qqq, ppp - 1) + dyq * ((1.0 - qqq) * powi(qqq, ppp - 1) - powi(1.0 - qqq, ppp))));
  const double aaa = (1.0 - bbb * powi(qqq, ppp)) / qqq;
  const int ny = h->ny;
  h->yu.assign(ny + 1, 0.0);
  h->yv.assign(ny, 0.0);
  h->dthdyu.assign(ny + 1, 0.0);
  h->dthdyv.assign(ny, 0.0);
  h->dyu2i.assign(3 * (ny + 1), 0.0);
  h->dyv2i.assign(3 * ny, 0.0);
  std::vector<double> &yu = h->yu, &yv = h->yv;
  yu[0] = aaa * (-1.0 + .5 - qqq) + bbb * powi(-1.0 + .5 - qqq, ppp);
  for (int j = 0; j <= nn; j++) {
    yv[j] = aaa * (j - qqq) + bbb * powi(j - qqq, ppp);
    yu[j + 1] = aaa * (j + .5 - qqq) + bbb * powi(j + .5 - qqq, ppp);
  }
  h->dthdyu[0] = 1.0 / (aaa + ppp * bbb * powi(-1.0 + .5 - qqq, ppp - 1));
  for (int j = 0; j <= nn; j++) {
    h->dthdyv[j] = 1.0 / (aaa + ppp * bbb * powi(j - qqq, ppp - 1));
    h->dthdyu[j + 1] = 1.0 / (aaa + ppp * bbb * powi(j + .5 - qqq, ppp - 1));
  }
  auto U = [&](int c, int j) -> double & { return h->dyu2i[3 * j + c - 1]; };
  auto V = [&](int c, int j) -> double & { return h->dyv2i[3 * j + c - 1]; };
  auto sq = [](double x) { return x * x; };
  int j = 0;
  U(1, j) = 0.0;
  U(2, j) = -2.0 / sq(yu[j + 1] - yu[j]);
  U(3, j) = 1.0 / sq(yu[j + 1] - yu[j]);
  V(1, j) = 0.0;
  V(2, j) = -2.0 / sq(yv[j + 1] - yv[j]);
  V(3, j) = 1.0 / sq(yv[j + 1] - yv[j]);
  for (j = 1; j <= nn - 1; j++) {
    U(1, j) = 2.0 / ((yu[j] - yu[j - 1]) * (yu[j + 1] - yu[j - 1]));
    U(2, j) = -2.0 / ((yu[j + 1] - yu[j]) * (yu[j] - yu[j - 1]));
    U(3, j) = 2.0 / ((yu[j + 1] - yu[j]) * (yu[j + 1] - yu[j - 1]));
    V(1, j) = 2.0 / ((yv[j] - yv[j - 1]) * (yv[j + 1] - yv[j - 1]));
    V(2, j) = -2.0 / ((yv[j + 1] - yv[j]) * (yv[j] - yv[j - 1]));
    V(3, j) = 2.0 / ((yv[j + 1] - yv[j]) * (yv[j + 1] - yv[j - 1]));
  }
  j = nn;
  U(1, j) = 2.0 / ((yu[j] - yu[j - 1]) * (yu[j + 1] - yu[j - 1]));
  U(2, j) = -2.0 / ((yu[j + 1] - yu[j]) * (yu[j] - yu[j - 1]));
  U(3, j) = 2.0 / ((yu[j + 1] - yu[j]) * (yu[j + 1] - yu[j - 1]));
  V(1, j) = 1.0 / sq(yv[j] - yv[j - 1]);
  V(2, j) = -2.0 / sq(yv[j] - yv[j - 1]);
  V(3, j) = 0.0;
  j = nn + 1;
  U(1, j) = 1.0 / sq(yu[j] - yu[j - 1]);
  U(2, j) = -2.0 / sq(yu[j] - yu[j - 1]);
  U(3, j) = 0.0;
  // dtv (start.f90:476) with k2F_x(Nx/2) and k2F_z(Nz/2) of def_k (start.f90:838-862)
  const double k2x = -((h->prm.alp * (h->Nx / 2)) * (h->prm.alp * (h->Nx / 2)));
  const double k2z = -((h->prm.bet * (h->Nz / 2 - 1)) * (h->prm.bet * (h->Nz / 2 - 1)));
  h->dtv_den = -k2x - k2z + 4.0 / sq(yv[1] - yv[0]);
  h->dtv = h->prm.Re / h->dtv_den;
  // wall weights (start.f90:518-537)
  const int nyu = h->nyu, nyv = h->nyv;
  h->gw1 = -(yv[0] - yu[0]) / (yv[0] - yu[1]);
  h->gw2 = (yu[nyu + 1] - yv[nyv + 1]) / (yv[nyv + 1] - yu[nyu]);
  h->gbc = (yu[0] - yv[0]) / (yu[1] - yv[0]);
  h->gi1 = (yu[0] - yu[1]) / (yv[0] - yu[1]);
  h->gi2 = (yu[nyu + 1] - yu[nyu]) / (yv[nyv + 1] - yu[nyu]);
  // Runge-Kutta coefficients (Le & Moin 1991), start.f90:439-453
  const double a[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0}, g[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
  const double c[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0}, d[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};
  for (int q = 0; q < 3; q++) {
    h->aRK[q] = a[q]; h->bRK[q] = a[q]; h->gRK[q] = g[q]; h->cRK[q] = c[q]; h->dRK[q] = d[q];
  }
}

template <typename T>
static int dalloc(dns_handle *h, T **p, size_t n) {
  void *q = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc(&q, n * sizeof(T)));
  HIPCHK(hipMemset(q, 0, n * sizeof(T)));
  h->allocs.push_back(q);
  h->bytes += n * sizeof(T);
  *p = reinterpret_cast<T *>(q);
  return 0;
}

static bool fft_len_ok(int n) {
  int m = n;
  while (m % 2 == 0) m /= 2;
  while (m % 3 == 0) m /= 3;
  return n > 0 && m == 1;
}

template <typename R>
static std::vector<R> twiddles(int n) {  // e^{-2 pi i m/n} as (re, im) pairs
  std::vector<R> t(2 * (size_t)n);
  for (int m = 0; m < n; m++) {
    const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    t[2 * m] = (R)cosl(a);
    t[2 * m + 1] = (R)(-sinl(a));
  }
  return t;
}

static void destroy_all(dns_handle *h) {
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->graph) (void)hipGraphExecDestroy(h->graph);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->stream) {
    for (int q = 0; q < 8; q++) (void)hipEventDestroy(h->ev[q]);
    (void)hipStreamDestroy(h->stream);
  }
  delete h;
}

extern "C" int dns_create(const dns_params *prm, dns_handle **out) {
  if (!prm || !out) return fail(DNS_ERR_ARG, "dns_create: null argument");
  *out = nullptr;
  const int Mx = prm->Mx, Mz = prm->Mz, ny = prm->ny;
  if (Mx <= 0 || Mz <= 0 || Mx % 6 != 0 || Mz % 6 != 0)
    return fail(DNS_ERR_ARG, "dns_create: Mx=%d, Mz=%d must be positive multiples of 6 (Nx = 2Mx/3 even)", Mx, Mz);
  // FFTPACK factors 2, 3, 4 only; any other prime factor is fatal there (rft_buff.f:308-310)
  if (!fft_len_ok(Mx) || !fft_len_ok(Mz))
    return fail(DNS_ERR_ARG, "dns_create: Mx=%d, Mz=%d have a prime factor other than 2 and 3", Mx, Mz);
  if (!dnsk::plane_size_supported(Mx) || !dnsk::plane_size_supported(Mz))
    return fail(DNS_ERR_ARG, "dns_create: no kernel instantiated for Mx=%d, Mz=%d", Mx, Mz);
  if (Mz % dnsk::xpass_rows_per_block(Mx) != 0)
    return fail(DNS_ERR_ARG, "dns_create: Mz=%d must be a multiple of %d for Mx=%d", Mz, dnsk::xpass_rows_per_block(Mx), Mx);
  if (ny < 6) return fail(DNS_ERR_ARG, "dns_create: ny=%d too small", ny);
  if (prm->ppp < 1) return fail(DNS_ERR_ARG, "dns_create: pp=%d must be >= 1", prm->ppp);
  if (!(prm->Re > 0) || !(prm->CFL > 0) || !(prm->alp > 0) || !(prm->bet > 0))
    return fail(DNS_ERR_ARG, "dns_create: Re, CFL, alp, bet must be positive");
  if (prm->flag_ctpress != 0 && prm->flag_ctpress != 1)
    return fail(DNS_ERR_ARG, "dns_create: flag_ctpress=%d must be 0 or 1", prm->flag_ctpress);
  if (prm->precision != DNS_PREC_FP64 && prm->precision != DNS_PREC_MIXED)
    return fail(DNS_ERR_ARG, "dns_create: unknown precision %d", prm->precision);
  if (prm->column_path != 0 && prm->column_path != 1)
    return fail(DNS_ERR_ARG, "dns_create: column_path=%d must be 0 or 1", prm->column_path);

  dns_handle *h = new dns_handle();
  h->prm = *prm;
  h->Mx = Mx; h->Mz = Mz; h->ny = ny;
  h->Nx = 2 * Mx / 3; h->Nz = 2 * Mz / 3;
  h->Kx = h->Nx / 2 + 1; h->C = h->Kx * h->Nz;
  h->nyv = ny - 2; h->nyu = h->nyv + 1;
  h->fp32 = prm->precision == DNS_PREC_MIXED;
  h->nm = prm->members > 1 ? prm->members : 1;
  h->colpass = prm->column_path == 0 && dnsk::column_pass_supported(h->nyu);
  h->uploaded.assign(h->nm, 0);
  h->Re_m.assign(h->nm, prm->Re);
  h->mpgx_m.assign(h->nm, prm->mpgx);
  const int nplanes = 2 * h->nyu;
  h->bj = (prm->plane_batch > 0 && prm->plane_batch < nplanes) ? (prm->plane_batch + 1) / 2 : h->nyu;
  host_grid(h);

  auto chk = [&](int rc) { return rc; };
  int rc = 0;
  const size_t C = h->C, ru = (size_t)(h->nyu + 2) * h->nm, rv = (size_t)(h->nyv + 2) * h->nm,
               rp = (size_t)h->nyu * h->nm;  // rows of all members
  rc |= chk(dalloc(h, &h->u1, C * ru)); rc |= chk(dalloc(h, &h->u3, C * ru)); rc |= chk(dalloc(h, &h->u2, C * rv));
  rc |= chk(dalloc(h, &h->p, C * rp)); rc |= chk(dalloc(h, &h->psi, C * rp)); rc |= chk(dalloc(h, &h->div, C * rp));
  rc |= chk(dalloc(h, &h->du1, C * ru)); rc |= chk(dalloc(h, &h->du3, C * ru)); rc |= chk(dalloc(h, &h->du2, C * rv));
  rc |= chk(dalloc(h, &h->Nu1, C * rp)); rc |= chk(dalloc(h, &h->Nu3, C * rp)); rc |= chk(dalloc(h, &h->Nu2, C * rp));
  rc |= chk(dalloc(h, &h->Nxz1, C * rp)); rc |= chk(dalloc(h, &h->Nxz3, C * rp)); rc |= chk(dalloc(h, &h->Nxz2, C * rp));
  rc |= chk(dalloc(h, &h->uvf, C * rv)); rc |= chk(dalloc(h, &h->wvf, C * rv)); rc |= chk(dalloc(h, &h->vvc, C * ru));
  rc |= chk(dalloc(h, &h->w3u, C * ru)); rc |= chk(dalloc(h, &h->w3v, C * rv)); rc |= chk(dalloc(h, &h->cP, C * rp));
  rc |= chk(dalloc(h, &h->mscr, 2 * ru));
  rc |= chk(dalloc(h, &h->cfls, 2 * dnsk::CFL_SLOTS * (size_t)h->nm));
  rc |= chk(dalloc(h, &h->d_yu, (size_t)ny + 1)); rc |= chk(dalloc(h, &h->d_yv, (size_t)ny));
  rc |= chk(dalloc(h, &h->d_dthdyu, (size_t)ny + 1)); rc |= chk(dalloc(h, &h->d_dthdyv, (size_t)ny));
  rc |= chk(dalloc(h, &h->d_dyu2i, 3 * ((size_t)ny + 1))); rc |= chk(dalloc(h, &h->d_dyv2i, 3 * (size_t)ny));
  rc |= chk(dalloc(h, &h->stage, C * (h->nyu + 2)));  // one member's field
  if (h->fp32) rc |= chk(dalloc(h, &h->stage2, C * (h->nyu + 2)));
  rc |= chk(dalloc(h, &h->st, (size_t)h->nm));
  const size_t rsz = h->fp32 ? sizeof(float2) : sizeof(double2);
  const int Kp = h->fp32 ? dnsk::row_pitch<float>(h->Kx) : dnsk::row_pitch<double>(h->Kx);
  const size_t iel = dnsk::member_istride(h->bj, Mz, Kp) * h->nm, pel = dnsk::member_pstride(h->bj, Mz, Kp) * h->nm;
  char *Ib = nullptr, *Pb = nullptr, *tz = nullptr, *tx = nullptr, *th = nullptr;
  rc |= chk(dalloc(h, &Ib, iel * rsz)); rc |= chk(dalloc(h, &Pb, pel * rsz));
  rc |= chk(dalloc(h, &tz, (size_t)Mz * rsz)); rc |= chk(dalloc(h, &tx, (size_t)Mx * rsz));
  rc |= chk(dalloc(h, &th, (size_t)2 * (Mx / 2) * rsz));  // the table, then its pass-blocked copy
  rc |= chk(dalloc(h, &h->twz64, (size_t)Mz)); rc |= chk(dalloc(h, &h->twx64, (size_t)Mx));
  rc |= chk(dalloc(h, &h->twh64, (size_t)Mx / 2));
  if (rc) {
    std::string msg = g_err;
    destroy_all(h);
    return fail(DNS_ERR_HIP, "dns_create: device allocation failed: %s", msg.c_str());
  }
  h->I = Ib; h->P = Pb; h->twz = tz; h->twx = tx; h->twh = th;
  auto up = [&](void *dst, const void *src, size_t n) { return hipMemcpy(dst, src, n, hipMemcpyHostToDevice); };
  hipError_t e = hipSuccess;
  e = e ? e : up(h->d_yu, h->yu.data(), h->yu.size() * 8);
  e = e ? e : up(h->d_yv, h->yv.data(), h->yv.size() * 8);
  e = e ? e : up(h->d_dthdyu, h->dthdyu.data(), h->dthdyu.size() * 8);
  e = e ? e : up(h->d_dthdyv, h->dthdyv.data(), h->dthdyv.size() * 8);
  e = e ? e : up(h->d_dyu2i, h->dyu2i.data(), h->dyu2i.size() * 8);
  e = e ? e : up(h->d_dyv2i, h->dyv2i.data(), h->dyv2i.size() * 8);
  if (h->fp32) {
    auto a = twiddles<float>(Mz), b = twiddles<float>(Mx), c = twiddles<float>(Mx / 2);
    e = e ? e : up(tz, a.data(), a.size() * 4);
    e = e ? e : up(tx, b.data(), b.size() * 4);
    e = e ? e : up(th, c.data(), c.size() * 4);
  } else {
    auto a = twiddles<double>(Mz), b = twiddles<double>(Mx), c = twiddles<double>(Mx / 2);
    e = e ? e : up(tz, a.data(), a.size() * 8);
    e = e ? e : up(tx, b.data(), b.size() * 8);
    e = e ? e : up(th, c.data(), c.size() * 8);
  }
  if (!e && dnsk::build_xpass_twiddles(th, Mx, h->fp32, nullptr) != 0) {
    destroy_all(h);
    return fail(DNS_ERR_ARG, "dns_create: no x-pass plan for Mx = %d", Mx);
  }
  e = e ? e : hipDeviceSynchronize();
  e = e ? e : hipGetLastError();
  {
    auto a = twiddles<double>(Mz), b = twiddles<double>(Mx), c = twiddles<double>(Mx / 2);
    e = e ? e : up(h->twz64, a.data(), a.size() * 8);
    e = e ? e : up(h->twx64, b.data(), b.size() * 8);
    e = e ? e : up(h->twh64, c.data(), c.size() * 8);
  }
  for (int m = 0; m < h->nm; m++) {
    DnsState s0;
    memset(&s0, 0, sizeof s0);
    s0.mpgx = prm->mpgx;
    s0.Re = prm->Re;
    s0.dtv = h->dtv;
    e = e ? e : up(h->st + m, &s0, sizeof s0);
  }
  e = e ? e : hipStreamCreate(&h->stream);
  for (int q = 0; q < 8 && e == hipSuccess; q++) e = hipEventCreate(&h->ev[q]);
  if (e != hipSuccess) {
    destroy_all(h);
    return fail(DNS_ERR_HIP, "dns_create: %s", hipGetErrorString(e));
  }
  memset(h->tsum, 0, sizeof h->tsum);
  h->tcount = 0;
  *out = h;
  return DNS_OK;
}

extern "C" int dns_destroy(dns_handle *h) {
  if (!h) return DNS_OK;
  (void)hipStreamSynchronize(h->stream);
  destroy_all(h);
  return DNS_OK;
}

extern "C" int dns_ygrid(int ny, double dyq, int ppp, double *yu, double *yv, double *dthdyu, double *dthdyv) {
  if (ny < 6 || ppp < 1 || !yu || !yv || !dthdyu || !dthdyv) return fail(DNS_ERR_ARG, "dns_ygrid: bad argument");
  dns_handle h;
  h.prm.ny = ny;
  h.prm.dyq = dyq;
  h.prm.ppp = ppp;
  h.prm.alp = h.prm.bet = h.prm.Re = 1.0;
  h.ny = ny;
  h.nyv = ny - 2;
  h.nyu = ny - 1;
  h.Nx = h.Nz = 4;
  host_grid(&h);
  memcpy(yu, h.yu.data(), h.yu.size() * 8);
  memcpy(yv, h.yv.data(), h.yv.size() * 8);
  memcpy(dthdyu, h.dthdyu.data(), h.dthdyu.size() * 8);
  memcpy(dthdyv, h.dthdyv.data(), h.dthdyv.size() * 8);
  return DNS_OK;
}

extern "C" int dns_sizes(const dns_handle *h, long long *out) {
  if (!h || !out) return fail(DNS_ERR_ARG, "dns_sizes: null argument");
  const long long v[10] = {h->Nx, h->Nz, h->Kx, h->C, h->nyu, h->nyv, h->Mx, h->Mz, h->ny,
                           (long long)(h->bytes >> 20)};
  memcpy(out, v, sizeof v);
  return DNS_OK;
}

extern "C" int dns_setup_arrays(const dns_handle *h, double *yu, double *yv, double *dthdyu, double *dthdyv,
                                double *dyu2i, double *dyv2i, double *w) {
  if (!h) return fail(DNS_ERR_ARG, "dns_setup_arrays: null handle");
  if (yu) memcpy(yu, h->yu.data(), h->yu.size() * 8);
  if (yv) memcpy(yv, h->yv.data(), h->yv.size() * 8);
  if (dthdyu) memcpy(dthdyu, h->dthdyu.data(), h->dthdyu.size() * 8);
  if (dthdyv) memcpy(dthdyv, h->dthdyv.data(), h->dthdyv.size() * 8);
  if (dyu2i) memcpy(dyu2i, h->dyu2i.data(), h->dyu2i.size() * 8);
  if (dyv2i) memcpy(dyv2i, h->dyv2i.data(), h->dyv2i.size() * 8);
  if (w) {
    w[0] = h->dtv; w[1] = h->gw1; w[2] = h->gw2; w[3] = h->gbc; w[4] = h->gi1; w[5] = h->gi2;
  }
  return DNS_OK;
}

// ------------------------------------------------------------------------------------------------
static ColArgs col_args(dns_handle *h, int kRK) {
  ColArgs a;
  memset(&a, 0, sizeof a);
  a.Kx = h->Kx; a.Nx = h->Nx; a.Nz = h->Nz; a.C = h->C; a.nyu = h->nyu; a.nyv = h->nyv;
  a.kRK = kRK; a.flag_ctpress = h->prm.flag_ctpress;
  a.nm = h->nm;
  a.fixed = h->prm.fixed_dt > 0.0; a.fixed_dt = h->prm.fixed_dt;
  a.Re = h->prm.Re; a.alp = h->prm.alp; a.bet = h->prm.bet; a.CFL = h->prm.CFL; a.dtv = h->dtv;
  const int q = kRK - 1;
  a.aRK = h->aRK[q]; a.bRK = h->bRK[q]; a.gRK = h->gRK[q]; a.cRK = h->cRK[q]; a.dRK = h->dRK[q];
  a.gw1 = h->gw1; a.gw2 = h->gw2; a.gbc = h->gbc; a.gi1 = h->gi1; a.gi2 = h->gi2;
  a.dyp = h->yu[2] - h->yu[1];
  a.yu = h->d_yu; a.yv = h->d_yv; a.dthdyu = h->d_dthdyu; a.dthdyv = h->d_dthdyv;
  a.dyu2i = h->d_dyu2i; a.dyv2i = h->d_dyv2i;
  a.u1 = h->u1; a.u2 = h->u2; a.u3 = h->u3; a.p = h->p; a.psi = h->psi; a.div = h->div;
  a.du1 = h->du1; a.du2 = h->du2; a.du3 = h->du3;
  const int f = h->fp32 ? 1 : 0;  // nonlinear terms and fluxes are stored in the plane precision
  a.Nu1 = {h->Nu1, f}; a.Nu2 = {h->Nu2, f}; a.Nu3 = {h->Nu3, f};
  a.Nxz1 = {h->Nxz1, f}; a.Nxz2 = {h->Nxz2, f}; a.Nxz3 = {h->Nxz3, f};
  a.uvf = {h->uvf, f}; a.vvc = {h->vvc, f}; a.wvf = {h->wvf, f};
  a.w3u = h->w3u; a.w3v = h->w3v; a.cP = h->cP; a.mscr = h->mscr;
  a.st = h->st;
  return a;
}

static PlaneArgs plane_args(dns_handle *h, int kRK) {
  PlaneArgs a;
  memset(&a, 0, sizeof a);
  a.Mx = h->Mx; a.Mz = h->Mz; a.Nx = h->Nx; a.Nz = h->Nz; a.Kx = h->Kx; a.C = h->C;
  a.nyu = h->nyu; a.nyv = h->nyv; a.kRK = kRK;
  a.nm = h->nm; a.bjcap = h->bj;
  a.alp = h->prm.alp; a.bet = h->prm.bet;
  a.u1 = h->u1; a.u2 = h->u2; a.u3 = h->u3; a.yu = h->d_yu; a.yv = h->d_yv;
  a.I = h->I; a.P = h->P; a.twz = h->twz; a.twx = h->twx; a.twh = h->twh;
  a.Nxz1 = h->Nxz1; a.Nxz2 = h->Nxz2; a.Nxz3 = h->Nxz3; a.uvf = h->uvf; a.vvc = h->vvc; a.wvf = h->wvf;
  a.cfl = h->cfls;
  return a;
}

static int check_kRK(dns_handle *h, int kRK, const char *who) {
  if (!h) return fail(DNS_ERR_ARG, "%s: null handle", who);
  if (kRK < 1 || kRK > 3) return fail(DNS_ERR_ARG, "%s: kRK=%d not in 1..3", who, kRK);
  for (int m = 0; m < h->nm; m++)
    if (!h->uploaded[m]) return fail(DNS_ERR_STATE, "%s: no fields uploaded for member %d", who, m);
  return DNS_OK;
}

static void tmark(dns_handle *h, int q) {
  if (h->timing) (void)hipEventRecord(h->ev[q], h->stream);
}

// nonlinear plane stage + CFL maxima (FOU3D.f90:98-303 up to the column part)
static int plane_stage(dns_handle *h, int kRK) {
  if (kRK == 1) {
    dnsk::launch_cfl_reset(h->st, h->cfls, h->nm, h->stream);  // cfl[0], cfl[1], divmax, slots
    dnsk::launch_cfl_ghost(h->u1, h->u3, h->C, dnsk::member_strides(h->C, h->nyu, h->nyv).ru, h->nm, h->cfls,
                           h->stream);
  }
  PlaneArgs a = plane_args(h, kRK);
  const bool one_group = h->bj >= h->nyu;  // per-kernel event marks only with one launch per kernel
  tmark(h, 0);
  for (int j0 = 1; j0 <= h->nyu; j0 += h->bj) {
    const int nb = std::min(h->bj, h->nyu - j0 + 1);
    a.j0 = j0;
    a.nbj = nb;
    int rc = dnsk::launch_plane_kernel(a, nb + 1, h->fp32, 0, h->var[0], h->stream);
    if (one_group) tmark(h, 1);
    for (int k = 0; k < 2; k++) {  // x-pass, then z-forward; u-planes, then v-planes
      rc |= dnsk::launch_plane_kernel(a, nb, h->fp32, 1 + 2 * k, h->var[1 + 2 * k], h->stream);
      rc |= dnsk::launch_plane_kernel(a, nb, h->fp32, 2 + 2 * k, h->var[2 + 2 * k], h->stream);
      if (one_group) tmark(h, 2 + k);
    }
    if (rc) return fail(DNS_ERR_ARG, "plane stage: sizes %d x %d not instantiated", h->Mx, h->Mz);
  }
  if (!one_group) {
    tmark(h, 1);
    tmark(h, 2);
    tmark(h, 3);
  }
  if (kRK == 1) dnsk::launch_cfl_reduce(h->cfls, h->st, h->nm, h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

// ------------------------------------------------------------------------------------------------
extern "C" int dns_upload_member(dns_handle *h, int m, const double *u1, const double *u2, const double *u3,
                                 const double *p, double t, int iter, double QxT) {
  if (!h || !u1 || !u2 || !u3 || !p) return fail(DNS_ERR_ARG, "dns_upload_member: null argument");
  if (m < 0 || m >= h->nm) return fail(DNS_ERR_ARG, "dns_upload_member: member %d not in 0..%d", m, h->nm - 1);
  const size_t C = h->C;
  const dnsk::MemberStrides ms = dnsk::member_strides(h->C, h->nyu, h->nyv);
  struct F { const double *src; double2 *dst; int rows; } f[4] = {
      {u1, h->u1 + m * ms.ru, h->nyu + 2}, {u2, h->u2 + m * ms.rv, h->nyv + 2},
      {u3, h->u3 + m * ms.ru, h->nyu + 2}, {p, h->p + m * ms.rp, h->nyu}};
  for (auto &x : f) {
    HIPCHK(hipMemcpyAsync(h->stage, x.src, C * x.rows * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    dnsk::launch_ref_to_dev(h->stage, x.dst, (int)C, x.rows, h->stream);
  }
  // getini: Qx of mode (0,0) (flowrateIm, littleharsh.f90:473-492) and QxT (start.f90:1255-1263)
  double Qx = 0.0;
  for (int j = 1; j <= h->nyv + 1; j++) Qx = Qx + (h->yv[j] - h->yv[j - 1]) * u1[2 * j];
  DnsState s0;
  memset(&s0, 0, sizeof s0);
  s0.t = t;
  s0.iter = iter;
  s0.mpgx = h->mpgx_m[m];
  s0.Qx = Qx;
  s0.QxT = QxT > 0.0 ? QxT : Qx;
  s0.Re = h->Re_m[m];
  s0.dtv = h->Re_m[m] / h->dtv_den;
  HIPCHK(hipMemcpyAsync(h->st + m, &s0, sizeof s0, hipMemcpyHostToDevice, h->stream));
  // the carried nonlinear terms start at zero (their weight at kRK = 1 is cRK = 0)
  const size_t nb = ms.rp * (h->fp32 ? sizeof(float2) : sizeof(double2));
  HIPCHK(hipMemsetAsync(reinterpret_cast<char *>(h->Nu1) + m * nb, 0, nb, h->stream));
  HIPCHK(hipMemsetAsync(reinterpret_cast<char *>(h->Nu2) + m * nb, 0, nb, h->stream));
  HIPCHK(hipMemsetAsync(reinterpret_cast<char *>(h->Nu3) + m * nb, 0, nb, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->uploaded[m] = 1;
  return DNS_OK;
}

extern "C" int dns_upload_fields(dns_handle *h, const double *u1, const double *u2, const double *u3,
                                 const double *p, double t, int iter, double QxT) {
  if (!h) return fail(DNS_ERR_ARG, "dns_upload_fields: null argument");
  for (int m = 0; m < h->nm; m++) {
    const int rc = dns_upload_member(h, m, u1, u2, u3, p, t, iter, QxT);
    if (rc) return rc;
  }
  return DNS_OK;
}

extern "C" int dns_set_member_params(dns_handle *h, int m, double Re, double mpgx) {
  if (!h) return fail(DNS_ERR_ARG, "dns_set_member_params: null handle");
  if (m < 0 || m >= h->nm) return fail(DNS_ERR_ARG, "dns_set_member_params: member %d not in 0..%d", m, h->nm - 1);
  if (!(Re > 0)) return fail(DNS_ERR_ARG, "dns_set_member_params: Re must be positive");
  h->Re_m[m] = Re;
  h->mpgx_m[m] = mpgx;
  DnsState x;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(&x, h->st + m, sizeof x, hipMemcpyDeviceToHost));
  x.Re = Re;
  x.dtv = Re / h->dtv_den;
  x.mpgx = mpgx;
  HIPCHK(hipMemcpy(h->st + m, &x, sizeof x, hipMemcpyHostToDevice));
  return DNS_OK;
}

// nonlinear terms and fluxes are FP32 in mixed mode: widened into stage2 before the layout conversion
static const double2 *wide(dns_handle *h, const double2 *dev, int rows) {
  if (!h->fp32) return dev;
  dnsk::launch_widen(reinterpret_cast<const float2 *>(dev), h->stage2, (size_t)h->C * rows, h->stream);
  return h->stage2;
}

static int download_dev(dns_handle *h, const double2 *dev, double *dst, int rows) {
  const size_t C = h->C;
  dnsk::launch_dev_to_ref(dev, h->stage, (int)C, rows, h->stream);
  HIPCHK(hipMemcpyAsync(dst, h->stage, C * rows * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return DNS_OK;
}

static int upload_dev(dns_handle *h, double2 *dev, const double *src, int rows) {
  const size_t C = h->C;
  HIPCHK(hipMemcpyAsync(h->stage, src, C * rows * sizeof(double2), hipMemcpyHostToDevice, h->stream));
  dnsk::launch_ref_to_dev(h->stage, dev, (int)C, rows, h->stream);
  HIPCHK(hipStreamSynchronize(h->stream));
  return DNS_OK;
}

extern "C" int dns_download_member(dns_handle *h, int m, double *u1, double *u2, double *u3, double *p) {
  if (!h) return fail(DNS_ERR_ARG, "dns_download_member: null handle");
  if (m < 0 || m >= h->nm) return fail(DNS_ERR_ARG, "dns_download_member: member %d not in 0..%d", m, h->nm - 1);
  const dnsk::MemberStrides ms = dnsk::member_strides(h->C, h->nyu, h->nyv);
  int rc = 0;
  if (u1 && (rc = download_dev(h, h->u1 + m * ms.ru, u1, h->nyu + 2))) return rc;
  if (u2 && (rc = download_dev(h, h->u2 + m * ms.rv, u2, h->nyv + 2))) return rc;
  if (u3 && (rc = download_dev(h, h->u3 + m * ms.ru, u3, h->nyu + 2))) return rc;
  if (p && (rc = download_dev(h, h->p + m * ms.rp, p, h->nyu))) return rc;
  return DNS_OK;
}

extern "C" int dns_download_fields(dns_handle *h, double *u1, double *u2, double *u3, double *p) {
  return dns_download_member(h, 0, u1, u2, u3, p);
}

extern "C" int dns_download_nonlinear(dns_handle *h, double *Nu1, double *Nu2, double *Nu3) {
  if (!h) return fail(DNS_ERR_ARG, "dns_download_nonlinear: null handle");
  int rc = 0;
  if (Nu1 && (rc = download_dev(h, wide(h, h->Nu1, h->nyu), Nu1, h->nyu))) return rc;
  if (Nu2 && (rc = download_dev(h, wide(h, h->Nu2, h->nyv), Nu2, h->nyv))) return rc;
  if (Nu3 && (rc = download_dev(h, wide(h, h->Nu3, h->nyu), Nu3, h->nyu))) return rc;
  return DNS_OK;
}

extern "C" int dns_upload_nonlinear(dns_handle *h, const double *Nu1, const double *Nu2, const double *Nu3) {
  if (!h || !Nu1 || !Nu2 || !Nu3) return fail(DNS_ERR_ARG, "dns_upload_nonlinear: null argument");
  int rc = 0;
  double2 *dst[3] = {h->Nu1, h->Nu2, h->Nu3};
  const double *srcs[3] = {Nu1, Nu2, Nu3};
  const int rows[3] = {h->nyu, h->nyv, h->nyu};
  for (int k = 0; k < 3; k++) {
    if ((rc = upload_dev(h, h->fp32 ? h->stage2 : dst[k], srcs[k], rows[k]))) return rc;
    if (h->fp32) {
      dnsk::launch_narrow(h->stage2, reinterpret_cast<float2 *>(dst[k]), (size_t)h->C * rows[k], h->stream);
      HIPCHK(hipStreamSynchronize(h->stream));
    }
  }
  return DNS_OK;
}

extern "C" int dns_download_rhs(dns_handle *h, double *du1, double *du2, double *du3) {
  if (!h) return fail(DNS_ERR_ARG, "dns_download_rhs: null handle");
  const size_t C = h->C;
  int rc = 0;  // interior rows 1..nyu / 1..nyv start one row into the device arrays
  if (du1 && (rc = download_dev(h, h->du1 + C, du1, h->nyu))) return rc;
  if (du2 && (rc = download_dev(h, h->du2 + C, du2, h->nyv))) return rc;
  if (du3 && (rc = download_dev(h, h->du3 + C, du3, h->nyu))) return rc;
  return DNS_OK;
}

extern "C" int dns_download_flux(dns_handle *h, double *uv_f, double *vv_c, double *wv_f) {
  if (!h) return fail(DNS_ERR_ARG, "dns_download_flux: null handle");
  int rc = 0;
  if (uv_f && (rc = download_dev(h, wide(h, h->uvf, h->nyv + 2), uv_f, h->nyv + 2))) return rc;
  if (vv_c && (rc = download_dev(h, wide(h, h->vvc, h->nyu + 2), vv_c, h->nyu + 2))) return rc;
  if (wv_f && (rc = download_dev(h, wide(h, h->wvf, h->nyv + 2), wv_f, h->nyv + 2))) return rc;
  return DNS_OK;
}

extern "C" int dns_download_pressure_correction(dns_handle *h, double *psi, double *div) {
  if (!h) return fail(DNS_ERR_ARG, "dns_download_pressure_correction: null handle");
  int rc = 0;
  if (psi && (rc = download_dev(h, h->psi, psi, h->nyu))) return rc;
  if (div && (rc = download_dev(h, h->div, div, h->nyu))) return rc;
  return DNS_OK;
}

extern "C" int dns_get_member_scalars(dns_handle *h, int m, double *s) {
  if (!h || !s) return fail(DNS_ERR_ARG, "dns_get_member_scalars: null argument");
  if (m < 0 || m >= h->nm) return fail(DNS_ERR_ARG, "dns_get_member_scalars: member %d not in 0..%d", m, h->nm - 1);
  DnsState x;
  HIPCHK(hipMemcpyAsync(&x, h->st + m, sizeof x, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const double v[10] = {x.t, x.dt, x.dti, x.mpgx, x.Qx, x.dgx, x.dtc, x.QxT, x.iter, x.divmax};
  memcpy(s, v, sizeof v);
  if (x.nonfinite != 0.0) return fail(DNS_ERR_NONFINITE, "non-finite velocity reached the CFL maxima");
  return DNS_OK;
}

extern "C" int dns_get_scalars(dns_handle *h, double *s) { return dns_get_member_scalars(h, 0, s); }

// One history record of record_out (FOU3D.f90:2169-2200) for member m, from the device state and the
// (0,0) mode of u1 (column 0, rows 0..nyu+1):
//   rec = flag_ctpress, iter, t, dtv, dtc, dt, err, Qx, QxT, mpgx, dgx, Umax, Uslip
// err   = max|div| of the last v_corr (error, littleharsh.f90:370-409)
// Qx    = flowrateIm of u1(:,1) (littleharsh.f90:473-492), summed in the reference's order
// Umax  = maxvel (littleharsh.f90:550-570): max(0, Re u1(j,1)), j = 0..nyu+1
// Uslip = (u1(1,1)(-1 - yu(0)) + u1(0,1)(yu(1) + 1)) / (yu(1) - yu(0)), real part (FOU3D.f90:2171)
extern "C" int dns_history_record(dns_handle *h, int m, double *rec) {
  if (!h || !rec) return fail(DNS_ERR_ARG, "dns_history_record: null argument");
  if (m < 0 || m >= h->nm) return fail(DNS_ERR_ARG, "dns_history_record: member %d not in 0..%d", m, h->nm - 1);
  const dnsk::MemberStrides ms = dnsk::member_strides(h->C, h->nyu, h->nyv);
  const int nr = h->nyu + 2;
  std::vector<double2> col(nr);
  DnsState x;
  HIPCHK(hipMemcpy2DAsync(col.data(), sizeof(double2), h->u1 + m * ms.ru, (size_t)h->C * sizeof(double2),
                          sizeof(double2), nr, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(&x, h->st + m, sizeof x, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  double Q = 0.0;
  for (int j = 1; j <= h->nyv + 1; j++) Q = Q + (h->yv[j] - h->yv[j - 1]) * col[j].x;
  double Umax = 0.0;
  for (int j = 0; j <= h->nyu + 1; j++) Umax = std::max(Umax, col[j].x);
  const double Uslip = (col[1].x * (-1.0 - h->yu[0]) + col[0].x * (h->yu[1] + 1.0)) / (h->yu[1] - h->yu[0]);
  const double v[13] = {(double)h->prm.flag_ctpress, x.iter, x.t, x.dtv, x.dtc, x.dt, x.divmax, Q, x.QxT,
                        x.mpgx, x.dgx, Umax, Uslip};
  memcpy(rec, v, sizeof v);
  return DNS_OK;
}

extern "C" int dns_set_dt(dns_handle *h, double dt) {
  if (!h || !(dt > 0)) return fail(DNS_ERR_ARG, "dns_set_dt: bad argument");
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int m = 0; m < h->nm; m++) {
    double v[3];
    HIPCHK(hipMemcpy(v, h->st + m, 3 * sizeof(double), hipMemcpyDeviceToHost));
    v[1] = dt;
    v[2] = 1.0 / dt;
    HIPCHK(hipMemcpy(h->st + m, v, 3 * sizeof(double), hipMemcpyHostToDevice));
  }
  return DNS_OK;
}

// ------------------------------------------------------------------------------------------------
// staged entry points
extern "C" int dns_rhs0(dns_handle *h, int kRK) {
  int rc = check_kRK(h, kRK, "dns_rhs0");
  if (rc) return rc;
  dnsk::launch_rhs0(col_args(h, kRK), h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

extern "C" int dns_nonlinear(dns_handle *h, int kRK) {
  int rc = check_kRK(h, kRK, "dns_nonlinear");
  if (rc) return rc;
  if ((rc = plane_stage(h, kRK))) return rc;
  dnsk::launch_nl_col(col_args(h, kRK), h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

extern "C" int dns_solveU(dns_handle *h, int kRK) {
  int rc = check_kRK(h, kRK, "dns_solveU");
  if (rc) return rc;
  dnsk::launch_solveU(col_args(h, kRK), h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

extern "C" int dns_meanflow_ctU(dns_handle *h, int kRK) {
  int rc = check_kRK(h, kRK, "dns_meanflow_ctU");
  if (rc) return rc;
  dnsk::launch_meanflow(col_args(h, kRK), h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

extern "C" int dns_solveP(dns_handle *h, int kRK) {
  int rc = check_kRK(h, kRK, "dns_solveP");
  if (rc) return rc;
  dnsk::launch_solveP(col_args(h, kRK), h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

extern "C" int dns_vcorr(dns_handle *h, int kRK) {
  int rc = check_kRK(h, kRK, "dns_vcorr");
  if (rc) return rc;
  for (int m = 0; m < h->nm; m++) dnsk::launch_zero(&h->st[m].divmax, 1, h->stream);
  dnsk::launch_vcorr(col_args(h, kRK), h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

extern "C" int dns_meanflow_ctP(dns_handle *h) {
  if (!h) return fail(DNS_ERR_ARG, "dns_meanflow_ctP: null handle");
  dnsk::launch_qx(col_args(h, 1), h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

// ------------------------------------------------------------------------------------------------
// fused path
static int fused_substep(dns_handle *h, int kRK) {
  int rc = plane_stage(h, kRK);
  if (rc) return rc;
  ColArgs a = col_args(h, kRK);
  // max|div| of this substep's v_corr only, as error(div) sees it (littleharsh.f90:370-409); at kRK = 1
  // the plane stage's reset has cleared it
  if (kRK != 1) dnsk::launch_divmax_reset(h->st, h->nm, h->stream);
  if (h->colpass) {
    if (dnsk::launch_column_pass(a, h->var[8], h->stream))
      return fail(DNS_ERR_ARG, "column pass: %d rows not instantiated", h->nyu);
    // constant flow rate: meanflow_ctU on mode (0,0), which commutes with solveP and v_corr there
    if (h->prm.flag_ctpress == 0 && dnsk::launch_meanflow_post(a, h->stream))
      return fail(DNS_ERR_ARG, "column pass: %d rows exceed the mean-flow kernel", h->nyu);
    tmark(h, 4);
    tmark(h, 5);
    tmark(h, 6);
    HIPCHK(hipGetLastError());
    return DNS_OK;
  }
  dnsk::launch_fused_up(a, h->var[5], h->stream);
  tmark(h, 4);
  dnsk::launch_fused_down(a, h->var[6], h->stream);
  tmark(h, 5);
  if (h->prm.flag_ctpress == 0) dnsk::launch_meanflow(a, h->stream);
  dnsk::launch_fused_final(a, h->var[7], h->stream);
  tmark(h, 6);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

extern "C" int dns_rk3_substep(dns_handle *h, int kRK) {
  int rc = check_kRK(h, kRK, "dns_rk3_substep");
  if (rc) return rc;
  return fused_substep(h, kRK);
}

static void accumulate_timing(dns_handle *h) {
  float ms[7];
  (void)hipEventSynchronize(h->ev[6]);
  for (int q = 0; q < 6; q++) (void)hipEventElapsedTime(&ms[q], h->ev[q], h->ev[q + 1]);
  (void)hipEventElapsedTime(&ms[6], h->ev[0], h->ev[6]);
  for (int q = 0; q < 7; q++) h->tsum[q] += ms[q];
  h->tcount++;
}

extern "C" int dns_rk3_step(dns_handle *h, int nsteps) {
  int rc = check_kRK(h, 1, "dns_rk3_step");
  if (rc) return rc;
  for (int s = 0; s < nsteps; s++) {
    for (int kRK = 1; kRK <= 3; kRK++) {
      if ((rc = fused_substep(h, kRK))) return rc;
      if (h->timing) accumulate_timing(h);
    }
  }
  return DNS_OK;
}

extern "C" int dns_rk3_step_graph(dns_handle *h, int nsteps) {
  int rc = check_kRK(h, 1, "dns_rk3_step_graph");
  if (rc) return rc;
  if (!h->graph) {
    const bool tm = h->timing;
    h->timing = false;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int kRK = 1; kRK <= 3 && rc == 0; kRK++) rc = fused_substep(h, kRK);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    h->timing = tm;
    if (rc || e != hipSuccess) {  // the captured graph (if any) is not kept
      if (e == hipSuccess && g) (void)hipGraphDestroy(g);
      if (rc) return rc;
      return fail(DNS_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(DNS_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    h->graph = x;
  }
  for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(h->graph, h->stream));
  return DNS_OK;
}

// ------------------------------------------------------------------------------------------------
// Ensemble of several sizes: the handles (one per size, each with its same-size members) advance
// together with one grouped launch per kernel class per substep (ensemble.hip).
struct dns_ensemble {
  std::vector<dns_handle *> h;
  bool fp32 = false;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  hipEvent_t done = nullptr;
  bool timing = false;
  hipEvent_t ev[5] = {};
  double tsum[7] = {};
  int tcount = 0;
  // persistent launch (dns_ensemble_set_launch): one kernel per substep over a task queue
  int launch = 0;
  dnsk::PersistArgs *pargs = nullptr;  // device: the arguments of substeps kRK = 1, 2, 3
  dnsk::PTask *ptask = nullptr;        // device: the three queues one after another
  unsigned *pctr = nullptr;            // device: queue head, finished workgroups, timeout flag, counters
  int nwg = 0;                         // workgroups of a persistent launch
  int ntask[3] = {};
  // tile shapes of the grouped launches (dns_ensemble_set_tiles, dns_ensemble_autotune): variant of the
  // plane classes c = 0 z-inverse, 1/2 x-pass, 3/4 z-forward and of the column pass (var[5]); threads per
  // workgroup of the plane classes
  int var[6] = {1, 1, 1, 1, 1, 0};
  int wg[5] = {128, 128, 128, 128, 128};
};

static void ens_mark(dns_ensemble *e, int q) {
  if (e->timing) (void)hipEventRecord(e->ev[q], e->stream);
}

static int ens_substep(dns_ensemble *e, int kRK) {
  const int ng = (int)e->h.size();
  dnsk::EnsSmall S;
  memset(&S, 0, sizeof S);
  dnsk::EnsPlaneLaunch P;
  memset(&P, 0, sizeof P);
  dnsk::EnsColLaunch Cl;
  memset(&Cl, 0, sizeof Cl);
  int flags[dnsk::ENS_MAXG] = {0, 0, 0, 0};
  S.ng = P.ng = Cl.ng = ng;
  for (int g = 0; g < ng; g++) {
    dns_handle *h = e->h[g];
    S.st[g] = h->st;
    S.slots[g] = h->cfls;
    S.u1[g] = h->u1;
    S.u3[g] = h->u3;
    S.ms[g] = dnsk::member_strides(h->C, h->nyu, h->nyv).ru;
    S.nm[g] = h->nm;
    S.C[g] = h->C;
    P.a[g] = plane_args(h, kRK);
    P.a[g].j0 = 1;
    P.a[g].nbj = h->nyu;
    Cl.a[g] = col_args(h, kRK);
    flags[g] = h->prm.flag_ctpress == 0;
  }
  if (e->launch == 1) {  // everything of the substep in one launch
    ens_mark(e, 0);
    dnsk::launch_ens_persist(e->pargs + (kRK - 1), e->fp32, e->nwg, e->stream);
    for (int q = 1; q <= 4; q++) ens_mark(e, q);
    HIPCHK(hipGetLastError());
    return DNS_OK;
  }
  auto plane = [&](int c) { return dnsk::launch_ens_plane(P, e->fp32, c, e->wg[c], e->var[c], e->stream); };
  ens_mark(e, 0);
  dnsk::launch_ens_reset(S, kRK, e->stream);
  if (kRK == 1) dnsk::launch_ens_ghost(S, e->stream);
  int rc = plane(0);
  ens_mark(e, 1);
  rc |= plane(1);
  rc |= plane(2);
  ens_mark(e, 2);
  rc |= plane(3);
  rc |= plane(4);
  ens_mark(e, 3);
  if (kRK == 1) dnsk::launch_ens_cfl_reduce(S, e->stream);
  rc |= dnsk::launch_ens_column_pass(Cl, e->fp32, e->var[5], e->stream);
  dnsk::launch_ens_meanflow_post(Cl, flags, e->stream);
  ens_mark(e, 4);
  if (rc) return fail(DNS_ERR_ARG, "ensemble: a size is not instantiated for grouped launches");
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

static void ens_accumulate_timing(dns_ensemble *e) {
  float ms[5];
  (void)hipEventSynchronize(e->ev[4]);
  for (int q = 0; q < 4; q++) (void)hipEventElapsedTime(&ms[q], e->ev[q], e->ev[q + 1]);
  (void)hipEventElapsedTime(&ms[4], e->ev[0], e->ev[4]);
  e->tsum[0] += ms[0]; e->tsum[1] += ms[1]; e->tsum[2] += ms[2]; e->tsum[3] += ms[3]; e->tsum[6] += ms[4];
  e->tcount++;
}

static void ens_free_persist(dns_ensemble *e) {
  if (e->pargs) (void)hipFree(e->pargs);
  if (e->ptask) (void)hipFree(e->ptask);
  if (e->pctr) (void)hipFree(e->pctr);
  e->pargs = nullptr;
  e->ptask = nullptr;
  e->pctr = nullptr;
  e->launch = 0;
}

static void ens_destroy_all(dns_ensemble *e) {
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_free_persist(e);
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  if (e->done) (void)hipEventDestroy(e->done);
  for (int q = 0; q < 5; q++)
    if (e->ev[q]) (void)hipEventDestroy(e->ev[q]);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

extern "C" int dns_ensemble_create(dns_handle *const *hs, int n, dns_ensemble **out) {
  if (!hs || !out) return fail(DNS_ERR_ARG, "dns_ensemble_create: null argument");
  *out = nullptr;
  if (n < 1 || n > dnsk::ENS_MAXG) return fail(DNS_ERR_ARG, "dns_ensemble_create: %d groups (1..%d)", n, dnsk::ENS_MAXG);
  for (int g = 0; g < n; g++) {
    const dns_handle *h = hs[g];
    if (!h) return fail(DNS_ERR_ARG, "dns_ensemble_create: handle %d is null", g);
    for (int q = 0; q < g; q++)
      if (hs[q] == h) return fail(DNS_ERR_ARG, "dns_ensemble_create: handle %d appears twice", g);
    if (h->fp32 != hs[0]->fp32) return fail(DNS_ERR_ARG, "dns_ensemble_create: the groups' precisions differ");
    if (!dnsk::ens_plane_length_supported(h->Mx) || !dnsk::ens_plane_length_supported(h->Mz) ||
        !dnsk::ens_plane_rows_ok(h->Mx, h->Mz))
      return fail(DNS_ERR_ARG, "dns_ensemble_create: group %d has %d x %d physical points; grouped launches take "
                  "48, 96 or 192 in each direction", g, h->Mx, h->Mz);
    if (!dnsk::ens_column_rows_supported(h->nyu))
      return fail(DNS_ERR_ARG, "dns_ensemble_create: group %d has ny = %d; grouped launches take 9 <= ny <= 161", g,
                  h->ny);
    if (h->bj < h->nyu)
      return fail(DNS_ERR_ARG, "dns_ensemble_create: group %d splits its planes into batches (plane_batch)", g);
  }
  dns_ensemble *e = new dns_ensemble();
  e->h.assign(hs, hs + n);
  e->fp32 = hs[0]->fp32;
  hipError_t r = hipStreamCreate(&e->stream);
  if (r == hipSuccess) r = hipEventCreateWithFlags(&e->done, hipEventDisableTiming);
  for (int q = 0; q < 5 && r == hipSuccess; q++) r = hipEventCreate(&e->ev[q]);
  if (r != hipSuccess) {
    ens_destroy_all(e);
    return fail(DNS_ERR_HIP, "dns_ensemble_create: %s", hipGetErrorString(r));
  }
  *out = e;
  return DNS_OK;
}

extern "C" int dns_ensemble_destroy(dns_ensemble *e) {
  if (!e) return DNS_OK;
  ens_destroy_all(e);
  return DNS_OK;
}

// Work already queued on the handles' streams (uploads) comes first; the handles' later calls
// (downloads, their own steps) wait for the ensemble's steps.
static int ens_begin(dns_ensemble *e, const char *who) {
  if (!e) return fail(DNS_ERR_ARG, "%s: null ensemble", who);
  for (dns_handle *h : e->h) {
    for (int m = 0; m < h->nm; m++)
      if (!h->uploaded[m]) return fail(DNS_ERR_STATE, "%s: no fields uploaded for a member", who);
    HIPCHK(hipEventRecord(e->done, h->stream));
    HIPCHK(hipStreamWaitEvent(e->stream, e->done, 0));
  }
  return DNS_OK;
}
static int ens_end(dns_ensemble *e) {
  HIPCHK(hipEventRecord(e->done, e->stream));
  for (dns_handle *h : e->h) HIPCHK(hipStreamWaitEvent(h->stream, e->done, 0));
  return DNS_OK;
}

extern "C" int dns_ensemble_step(dns_ensemble *e, int nsteps) {
  int rc = ens_begin(e, "dns_ensemble_step");
  if (rc) return rc;
  for (int s = 0; s < nsteps; s++)
    for (int kRK = 1; kRK <= 3; kRK++) {
      if ((rc = ens_substep(e, kRK))) return rc;
      if (e->timing) ens_accumulate_timing(e);
    }
  return ens_end(e);
}

extern "C" int dns_ensemble_step_graph(dns_ensemble *e, int nsteps) {
  int rc = ens_begin(e, "dns_ensemble_step_graph");
  if (rc) return rc;
  if (!e->graph) {
    const bool tm = e->timing;
    e->timing = false;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (int kRK = 1; kRK <= 3 && rc == 0; kRK++) rc = ens_substep(e, kRK);
    hipError_t r = hipStreamEndCapture(e->stream, &g);
    e->timing = tm;
    if (rc || r != hipSuccess) {
      if (r == hipSuccess && g) (void)hipGraphDestroy(g);
      if (rc) return rc;
      return fail(DNS_ERR_HIP, "ensemble graph capture: %s", hipGetErrorString(r));
    }
    hipGraphExec_t x = nullptr;
    r = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (r != hipSuccess) return fail(DNS_ERR_HIP, "ensemble graph instantiate: %s", hipGetErrorString(r));
    e->graph = x;
  }
  for (int s = 0; s < nsteps; s++) HIPCHK(hipGraphLaunch(e->graph, e->stream));
  return ens_end(e);
}

extern "C" int dns_ensemble_sync(dns_ensemble *e) {
  if (!e) return fail(DNS_ERR_ARG, "dns_ensemble_sync: null ensemble");
  HIPCHK(hipStreamSynchronize(e->stream));
  if (e->launch == 1) {  // a dependency wait of the persistent launch that ran out (results not valid)
    unsigned err = 0;
    HIPCHK(hipMemcpy(&err, e->pctr + dnsk::PCTR_ERR, sizeof err, hipMemcpyDeviceToHost));
    if (err) {
      HIPCHK(hipMemset(e->pctr + dnsk::PCTR_ERR, 0, sizeof err));
      return fail(DNS_ERR_HIP, "dns_ensemble_sync: a task of the persistent launch waited too long for the "
                  "tasks it depends on; the fields of these steps are not valid");
    }
  }
  return DNS_OK;
}

// Tasks per queue entry by class (PC_RESET .. PC_MF): a task covers that many units (two plane blocks, or
// one column-pass workgroup) of one member.
static const int PERSIST_UNITS[dnsk::PC_N] = {1, 1, 16, 32, 32, 1, 16, 16, 8, 1};

extern "C" int dns_ensemble_set_launch(dns_ensemble *e, int mode, const int *units) {
  if (!e) return fail(DNS_ERR_ARG, "dns_ensemble_set_launch: null ensemble");
  if (mode != 0 && mode != 1) return fail(DNS_ERR_ARG, "dns_ensemble_set_launch: mode %d (0 or 1)", mode);
  HIPCHK(hipStreamSynchronize(e->stream));
  if (e->graph) {  // the captured step launches the kernels of the previous mode
    (void)hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
  }
  ens_free_persist(e);
  if (mode == 0) return DNS_OK;
  int un[dnsk::PC_N];
  for (int c = 0; c < dnsk::PC_N; c++) {
    un[c] = units && units[c] > 0 ? units[c] : PERSIST_UNITS[c];
    if (un[c] > 1 << 20) return fail(DNS_ERR_ARG, "dns_ensemble_set_launch: %d units per task", un[c]);
  }
  const int ng = (int)e->h.size();
  dnsk::PersistGroup q[dnsk::ENS_MAXG];
  int nm[dnsk::ENS_MAXG];
  for (int g = 0; g < ng; g++) {
    dns_handle *h = e->h[g];
    PlaneArgs a = plane_args(h, 1);
    a.j0 = 1;
    a.nbj = h->nyu;
    nm[g] = h->nm;
    if (dnsk::ens_persist_group(a, h->prm.flag_ctpress == 0, q[g]))
      return fail(DNS_ERR_ARG, "dns_ensemble_set_launch: group %d (%d x %d points, ny = %d) is not supported by the "
                  "persistent launch (each z-inverse chunk of planes must be whole: ny - 1 even)", g, h->Mx,
                  h->Mz, h->ny);
  }
  std::vector<dnsk::PTask> all, tk;
  int nctr = 0;
  dnsk::PersistArgs A[3];
  for (int k = 0; k < 3; k++) {
    if (dnsk::ens_persist_tasks(q, nm, ng, k + 1, un, tk, nctr))
      return fail(DNS_ERR_ARG, "dns_ensemble_set_launch: the task queue of substep %d does not order its tasks", k + 1);
    e->ntask[k] = (int)tk.size();
    memset(&A[k], 0, sizeof A[k]);
    A[k].ntask = (int)tk.size();
    A[k].nctr = nctr;
    A[k].kRK = k + 1;
    A[k].ng = ng;
    for (int g = 0; g < ng; g++) {
      dns_handle *h = e->h[g];
      A[k].pa[g] = plane_args(h, k + 1);
      A[k].pa[g].j0 = 1;
      A[k].pa[g].nbj = h->nyu;
      A[k].ca[g] = col_args(h, k + 1);
      A[k].q[g] = q[g];
    }
    all.insert(all.end(), tk.begin(), tk.end());
  }
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int per = dnsk::ens_persist_max_blocks(e->fp32);
  if (per < 1 || ncu < 1) return fail(DNS_ERR_HIP, "dns_ensemble_set_launch: the persistent kernel does not fit a CU");
  e->nwg = per * ncu;
  hipError_t r = hipMalloc((void **)&e->pargs, sizeof A);
  if (r == hipSuccess) r = hipMalloc((void **)&e->ptask, all.size() * sizeof(dnsk::PTask));
  if (r == hipSuccess) r = hipMalloc((void **)&e->pctr, (dnsk::PCTR0 + (size_t)nctr) * sizeof(unsigned));
  if (r == hipSuccess) r = hipMemcpy(e->ptask, all.data(), all.size() * sizeof(dnsk::PTask), hipMemcpyHostToDevice);
  if (r == hipSuccess) {
    size_t off = 0;
    for (int k = 0; k < 3; k++) {
      A[k].task = e->ptask + off;
      A[k].ctr = e->pctr;
      off += (size_t)e->ntask[k];
    }
    r = hipMemcpy(e->pargs, A, sizeof A, hipMemcpyHostToDevice);
  }
  if (r == hipSuccess) r = hipMemset(e->pctr, 0, (dnsk::PCTR0 + (size_t)nctr) * sizeof(unsigned));
  if (r != hipSuccess) {
    ens_free_persist(e);
    return fail(DNS_ERR_HIP, "dns_ensemble_set_launch: %s", hipGetErrorString(r));
  }
  e->launch = 1;
  return DNS_OK;
}

extern "C" int dns_ensemble_get_launch(dns_ensemble *e, int *mode, int *nwg, int *ntask) {
  if (!e || !mode) return fail(DNS_ERR_ARG, "dns_ensemble_get_launch: null argument");
  *mode = e->launch;
  if (nwg) *nwg = e->launch ? e->nwg : 0;
  if (ntask)
    for (int k = 0; k < 3; k++) ntask[k] = e->launch ? e->ntask[k] : 0;
  return DNS_OK;
}

extern "C" int dns_ensemble_set_timing(dns_ensemble *e, int on) {
  if (!e) return fail(DNS_ERR_ARG, "dns_ensemble_set_timing: null ensemble");
  e->timing = on != 0;
  memset(e->tsum, 0, sizeof e->tsum);
  e->tcount = 0;
  return DNS_OK;
}

extern "C" int dns_ensemble_get_timing(dns_ensemble *e, double *t, int *n) {
  if (!e || !t) return fail(DNS_ERR_ARG, "dns_ensemble_get_timing: null argument");
  for (int q = 0; q < 7; q++) t[q] = e->tcount ? e->tsum[q] / e->tcount : 0.0;
  if (n) *n = e->tcount;
  return DNS_OK;
}

// ------------------------------------------------------------------------------------------------
// Kernel variants.  Every variant of a kernel gives bit-identical results: the plane kernels' variants
// differ in tile shape (same butterflies, same order of operations per element), the column sweep's in
// how the three solves map to wavefronts (same arithmetic per column).
static int kern_len(const dns_handle *h, int c) { return (c == 1 || c == 2) ? h->Mx : h->Mz; }
static int kern_variants(const dns_handle *h, int c) {
  if (c < 5) return dnsk::plane_variants(kern_len(h, c));
  if (c == 8) return dnsk::column_pass_supported(h->nyu) ? dnsk::column_pass_variants(h->nyu) : 1;
  return 3;
}

// Replays of the executable graph may still be queued on the handle's stream: wait for them before the
// executable is destroyed.
static void drop_graph(dns_handle *h) {
  if (h->graph) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipGraphExecDestroy(h->graph);
  }
  h->graph = nullptr;
}

extern "C" int dns_get_plane_variants(dns_handle *h, int *var, int *count) {
  if (!h || !var) return fail(DNS_ERR_ARG, "dns_get_plane_variants: null argument");
  for (int c = 0; c < DNS_NKERN; c++) {
    var[c] = h->var[c];
    if (count) count[c] = kern_variants(h, c);
  }
  return DNS_OK;
}

extern "C" int dns_set_plane_variants(dns_handle *h, const int *var) {
  if (!h || !var) return fail(DNS_ERR_ARG, "dns_set_plane_variants: null argument");
  for (int c = 0; c < DNS_NKERN; c++)
    if (var[c] < 0 || var[c] >= kern_variants(h, c))
      return fail(DNS_ERR_ARG, "dns_set_plane_variants: kernel %d has no variant %d (it has %d)", c, var[c],
                  kern_variants(h, c));
  for (int c = 0; c < DNS_NKERN; c++) h->var[c] = var[c];
  drop_graph(h);
  return DNS_OK;
}

// Copies of what the column sweeps overwrite (u1, u3, u2, Nu1..3, p of every member, and the members'
// scalars), kept by the autotuners while they time and put back afterwards, on stream st.
struct TuneSave {
  double2 *buf = nullptr;
  std::vector<DnsState> st;
};
static void tune_arrays(dns_handle *h, double2 **src, size_t *len) {
  const size_t C = h->C;
  const size_t nu = C * (h->nyu + 2) * h->nm, nv = C * (h->nyv + 2) * h->nm, nn = C * h->nyu * h->nm;
  double2 *a[7] = {h->u1, h->u3, h->u2, h->Nu1, h->Nu2, h->Nu3, h->p};
  const size_t l[7] = {nu, nu, nv, nn, nn, nn, nn};
  for (int k = 0; k < 7; k++) {
    src[k] = a[k];
    len[k] = l[k];
  }
}
static int tune_save(dns_handle *h, hipStream_t st, TuneSave &sv) {
  double2 *src[7];
  size_t len[7], tot = 0;
  tune_arrays(h, src, len);
  for (size_t l : len) tot += l;
  sv.st.resize(h->nm);
  HIPCHK(hipMalloc(&sv.buf, tot * sizeof(double2)));
  HIPCHK(hipMemcpyAsync(sv.st.data(), h->st, h->nm * sizeof(DnsState), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double2 *q = sv.buf;
  for (int k = 0; k < 7; k++) {
    HIPCHK(hipMemcpyAsync(q, src[k], len[k] * sizeof(double2), hipMemcpyDeviceToDevice, st));
    q += len[k];
  }
  return DNS_OK;
}
static void tune_restore(dns_handle *h, hipStream_t st, TuneSave &sv) {
  if (!sv.buf) return;
  double2 *src[7];
  size_t len[7];
  tune_arrays(h, src, len);
  double2 *q = sv.buf;
  for (int k = 0; k < 7; k++) {
    (void)hipMemcpyAsync(src[k], q, len[k] * sizeof(double2), hipMemcpyDeviceToDevice, st);
    q += len[k];
  }
  (void)hipMemcpyAsync(h->st, sv.st.data(), h->nm * sizeof(DnsState), hipMemcpyHostToDevice, st);
  (void)hipStreamSynchronize(st);
  (void)hipFree(sv.buf);
  sv.buf = nullptr;
}

// Times every variant of every kernel that has more than one and keeps the fastest: plane kernels on
// the first plane batch with kRK = 2 arguments (no CFL maxima), the column sweep on the whole field.
// Rewrites scratch only (nonlinear terms, fluxes, eliminated rows), so it may be called before a step
// or between steps, not between the staged calls of one substep; the fields are saved and restored.
extern "C" int dns_autotune(dns_handle *h, int iters, int *chosen, double *ms) {
  if (!h) return fail(DNS_ERR_ARG, "dns_autotune: null handle");
  if (iters < 1 || iters > 1000) return fail(DNS_ERR_ARG, "dns_autotune: iters=%d not in 1..1000", iters);
  for (int m = 0; m < h->nm; m++)
    if (!h->uploaded[m]) return fail(DNS_ERR_STATE, "dns_autotune: no fields uploaded for member %d", m);
  PlaneArgs a = plane_args(h, 2);
  a.j0 = 1;
  a.nbj = std::min(h->bj, h->nyu);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // the column sweeps overwrite u1, u2, u3, p and Nu: keep copies and put them back afterwards
  TuneSave sv;
  int rc = tune_save(h, h->stream, sv);
  ColArgs ca = col_args(h, 2);
  for (int c = 0; c < DNS_NKERN && rc == DNS_OK; c++) {
    const int nw = c == 0 ? a.nbj + 1 : a.nbj;  // the first plane batch
    const int nv_ = kern_variants(h, c);
    int best = 0;
    double tbest = 1e300;
    for (int v = 0; v < 3; v++) {
      if (ms) ms[3 * c + v] = -1.0;
      // the sweeps are timed where they are in use, the one-pass shapes where that is
      if (nv_ < 2 || v >= nv_ || (c < 5 && nw <= 0) || (c >= 5 && c < 8 && h->colpass) || (c == 8 && !h->colpass))
        continue;
      auto launch = [&]() -> int {
        if (c < 5) return dnsk::launch_plane_kernel(a, nw, h->fp32, c, v, h->stream);
        if (c == 8) return dnsk::launch_column_pass(ca, v, h->stream);
        if (c == 5) dnsk::launch_fused_up(ca, v, h->stream);
        else if (c == 6) dnsk::launch_fused_down(ca, v, h->stream);
        else dnsk::launch_fused_final(ca, v, h->stream);
        return 0;
      };
      if (launch()) { rc = fail(DNS_ERR_ARG, "dns_autotune: size not instantiated"); break; }
      hipError_t e = hipEventRecord(e0, h->stream);
      for (int it = 0; it < iters && e == hipSuccess; it++) launch();
      if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e != hipSuccess) { rc = fail(DNS_ERR_HIP, "dns_autotune: %s", hipGetErrorString(e)); break; }
      const double tv = (double)t / iters;
      if (ms) ms[3 * c + v] = tv;
      // another variant replaces the first only when it is at least 2 % faster: closer timings are within
      // the run-to-run spread, and the one-pass column shapes timed on their own ranked within that spread
      // in the opposite order to their times inside the step
      if (v == 0 ? tv < tbest : tv < 0.98 * tbest) { tbest = tv; best = v; }
    }
    if ((c >= 5 && c < 8 && h->colpass) || (c == 8 && !h->colpass)) best = h->var[c];  // path not in use
    h->var[c] = best;
    if (chosen) chosen[c] = best;
  }
  tune_restore(h, h->stream, sv);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  drop_graph(h);
  if (rc == DNS_OK) HIPCHK(hipGetLastError());
  return rc;
}

// ---- tiles of the ensemble's grouped launches ------------------------------------------------------
static void ens_drop_graph(dns_ensemble *e) {
  if (e->graph) {
    (void)hipStreamSynchronize(e->stream);
    (void)hipGraphExecDestroy(e->graph);
  }
  e->graph = nullptr;
}
static PlaneArgs ens_plane_args(dns_handle *h, int kRK) {
  PlaneArgs a = plane_args(h, kRK);
  a.j0 = 1;
  a.nbj = h->nyu;
  return a;
}

// shape k = 0..4 of a plane class: (workgroup size, variant)
static const int ENS_SHAPE_WG[5] = {128, 128, 256, 256, 256}, ENS_SHAPE_V[5] = {0, 1, 0, 1, 2};

static bool ens_shape_ok(dns_ensemble *e, int c, int wg, int v) {
  for (dns_handle *h : e->h)
    if (!dnsk::ens_plane_variant_ok(ens_plane_args(h, 1), c, v, wg)) return false;
  return true;
}

extern "C" int dns_ensemble_set_tiles(dns_ensemble *e, const int *var, const int *wg) {
  if (!e || !var || !wg) return fail(DNS_ERR_ARG, "dns_ensemble_set_tiles: null argument");
  for (int c = 0; c < 5; c++)
    if (!ens_shape_ok(e, c, wg[c], var[c]))
      return fail(DNS_ERR_ARG, "dns_ensemble_set_tiles: class %d: variant %d in %d-thread workgroups is not "
                  "available for every group", c, var[c], wg[c]);
  if (var[5] < 0 || var[5] >= dnsk::ENS_NVAR)
    return fail(DNS_ERR_ARG, "dns_ensemble_set_tiles: column stage: no variant %d", var[5]);
  ens_drop_graph(e);
  for (int c = 0; c < 6; c++) e->var[c] = var[c];
  for (int c = 0; c < 5; c++) e->wg[c] = wg[c];
  return DNS_OK;
}

extern "C" int dns_ensemble_get_tiles(dns_ensemble *e, int *var, int *wg) {
  if (!e || !var || !wg) return fail(DNS_ERR_ARG, "dns_ensemble_get_tiles: null argument");
  for (int c = 0; c < 6; c++) var[c] = e->var[c];
  for (int c = 0; c < 5; c++) wg[c] = e->wg[c];
  return DNS_OK;
}

// Times every tile shape of every grouped class on the whole launch (all groups, all members) and keeps
// the fastest per class: five shapes per plane class (kRK = 2 arguments: no CFL maxima), three column-pass
// variants.  The fields the column pass overwrites are saved and put back, so the ensemble's state is
// unchanged.
extern "C" int dns_ensemble_autotune(dns_ensemble *e, int iters, int *var, int *wg, double *ms) {
  if (!e) return fail(DNS_ERR_ARG, "dns_ensemble_autotune: null ensemble");
  if (iters < 1 || iters > 1000) return fail(DNS_ERR_ARG, "dns_ensemble_autotune: iters=%d not in 1..1000", iters);
  int rc = ens_begin(e, "dns_ensemble_autotune");
  if (rc) return rc;
  const int ng = (int)e->h.size();
  if (ms)
    for (int i = 0; i < DNS_ENS_TUNE_MS; i++) ms[i] = -1.0;
  std::vector<TuneSave> sv(ng);
  for (int g = 0; g < ng && rc == DNS_OK; g++) rc = tune_save(e->h[g], e->stream, sv[g]);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == DNS_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(DNS_ERR_HIP, "dns_ensemble_autotune: event creation failed");
  // one warm-up launch, then iters timed ones
  auto timed = [&](auto launch, double &t) -> int {
    if (launch()) return fail(DNS_ERR_ARG, "dns_ensemble_autotune: a size is not instantiated");
    hipError_t r = hipEventRecord(e0, e->stream);
    for (int it = 0; it < iters && r == hipSuccess; it++) launch();
    if (r == hipSuccess) r = hipEventRecord(e1, e->stream);
    if (r == hipSuccess) r = hipEventSynchronize(e1);
    float f = 0.f;
    if (r == hipSuccess) r = hipEventElapsedTime(&f, e0, e1);
    if (r != hipSuccess) return fail(DNS_ERR_HIP, "dns_ensemble_autotune: %s", hipGetErrorString(r));
    t = (double)f / iters;
    return DNS_OK;
  };
  dnsk::EnsPlaneLaunch P;
  memset(&P, 0, sizeof P);
  dnsk::EnsColLaunch Cl;
  memset(&Cl, 0, sizeof Cl);
  P.ng = Cl.ng = ng;
  for (int g = 0; g < ng; g++) {
    P.a[g] = ens_plane_args(e->h[g], 2);
    Cl.a[g] = col_args(e->h[g], 2);
  }
  int nvar[6] = {1, 1, 1, 1, 1, 0}, nwg[5] = {128, 128, 128, 128, 128};
  for (int c = 0; c < 5 && rc == DNS_OK; c++) {
    double tb = 1e300;
    for (int k = 0; k < 5 && rc == DNS_OK; k++) {
      const int w = ENS_SHAPE_WG[k], v = ENS_SHAPE_V[k];
      if (!ens_shape_ok(e, c, w, v)) continue;
      double t = -1.0;
      rc = timed([&]() { return dnsk::launch_ens_plane(P, e->fp32, c, w, v, e->stream); }, t);
      if (ms) ms[c * 5 + k] = t;
      if (rc == DNS_OK && t < tb) { tb = t; nwg[c] = w; nvar[c] = v; }
    }
  }
  double tb = 1e300;
  for (int v = 0; v < dnsk::ENS_NVAR && rc == DNS_OK; v++) {
    double t = -1.0;
    rc = timed([&]() { return dnsk::launch_ens_column_pass(Cl, e->fp32, v, e->stream); }, t);
    if (ms) ms[25 + v] = t;
    if (rc == DNS_OK && t < tb) { tb = t; nvar[5] = v; }
  }
  for (int g = 0; g < ng; g++) tune_restore(e->h[g], e->stream, sv[g]);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc == DNS_OK) {
    ens_drop_graph(e);
    for (int c = 0; c < 6; c++) e->var[c] = nvar[c];
    for (int c = 0; c < 5; c++) e->wg[c] = nwg[c];
    if (var && wg) dns_ensemble_get_tiles(e, var, wg);
  }
  const int rc2 = ens_end(e);
  return rc ? rc : rc2;
}

// ---- statistics and spectra (stats.f90, spectra.f90) ---------------------------------------------
static size_t stats_sp_per_member(const dns_handle *h) {
  const size_t C = h->C, nu = (size_t)(h->nyu + 2) * C, nv = (size_t)(h->nyv + 2) * C;
  return 3 * nu + nv + (size_t)h->nyu * C;
}

extern "C" int dns_accumulate_stats(dns_handle *h) {
  if (!h) return fail(DNS_ERR_ARG, "dns_accumulate_stats: null handle");
  for (int m = 0; m < h->nm; m++)
    if (!h->uploaded[m]) return fail(DNS_ERR_STATE, "dns_accumulate_stats: no fields uploaded for member %d", m);
  const int nchunk = dnsk::stats_nchunk(h->C);
  if (!h->stat_acc) {
    int rc = 0;
    rc |= dalloc(h, &h->stat_acc, (size_t)h->nm * dnsk::STAT_NQ * (h->nyu + 2));
    rc |= dalloc(h, &h->stat_sp, (size_t)h->nm * stats_sp_per_member(h));
    rc |= dalloc(h, &h->stat_part, (size_t)h->nm * (h->nyu + 2) * nchunk * dnsk::STAT_NQ);
    if (rc) return fail(DNS_ERR_HIP, "dns_accumulate_stats: out of device memory");
  }
  dnsk::StatsArgs a;
  a.Kx = h->Kx; a.Nx = h->Nx; a.Nz = h->Nz; a.Mx = h->Mx; a.Mz = h->Mz; a.C = h->C;
  a.nyu = h->nyu; a.nyv = h->nyv; a.nm = h->nm; a.nchunk = nchunk;
  a.bet = h->prm.bet; a.gi1 = h->gi1; a.gi2 = h->gi2;
  a.yu = h->d_yu; a.yv = h->d_yv; a.dthdyv = h->d_dthdyv;
  a.u1 = h->u1; a.u2 = h->u2; a.u3 = h->u3; a.p = h->p; a.pm = h->p_stat ? h->p_stat : h->p;
  a.partial = h->stat_part; a.acc = h->stat_acc; a.sp = h->stat_sp;
  dnsk::launch_stats(a, h->stream);
  HIPCHK(hipGetLastError());
  h->istat++;
  return DNS_OK;
}

// ppPL = 0; modes_to_planes_UVP(ppPL, p) (FOU3D.f90:149-152): the plane sums Pm, P2m read this copy of
// every member's p from now on; the reference refreshes it at write steps and at the end of the run
// (littleharsh.f90:439-443), and reads it from the initial field file (start.f90:1369-1387).
extern "C" int dns_refresh_stats_pressure(dns_handle *h) {
  if (!h) return fail(DNS_ERR_ARG, "dns_refresh_stats_pressure: null handle");
  for (int m = 0; m < h->nm; m++)
    if (!h->uploaded[m]) return fail(DNS_ERR_STATE, "dns_refresh_stats_pressure: no fields uploaded for member %d", m);
  const size_t n = (size_t)h->nm * dnsk::member_strides(h->C, h->nyu, h->nyv).rp;
  if (!h->p_stat && dalloc(h, &h->p_stat, n)) return fail(DNS_ERR_HIP, "dns_refresh_stats_pressure: out of device memory");
  HIPCHK(hipMemcpyAsync(h->p_stat, h->p, n * sizeof(double2), hipMemcpyDeviceToDevice, h->stream));
  return DNS_OK;
}

extern "C" int dns_reset_stats(dns_handle *h) {
  if (!h) return fail(DNS_ERR_ARG, "dns_reset_stats: null handle");
  if (h->stat_acc) {
    HIPCHK(hipMemsetAsync(h->stat_acc, 0, (size_t)h->nm * dnsk::STAT_NQ * (h->nyu + 2) * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(h->stat_sp, 0, (size_t)h->nm * stats_sp_per_member(h) * sizeof(double), h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  h->istat = 0;
  return DNS_OK;
}

extern "C" int dns_get_stats(dns_handle *h, int m, double *stats, double *spectra, int *istat) {
  if (!h) return fail(DNS_ERR_ARG, "dns_get_stats: null handle");
  if (m < 0 || m >= h->nm) return fail(DNS_ERR_ARG, "dns_get_stats: member %d not in 0..%d", m, h->nm - 1);
  const size_t nj = h->nyu + 2;
  if (istat) *istat = h->istat;
  if (stats) {
    if (h->stat_acc) {
      HIPCHK(hipMemcpyAsync(stats, h->stat_acc + (size_t)m * dnsk::STAT_NQ * nj, dnsk::STAT_NQ * nj * sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
    } else {
      memset(stats, 0, dnsk::STAT_NQ * nj * sizeof(double));
    }
  }
  if (spectra) {
    const size_t per = stats_sp_per_member(h), C = h->C;
    std::vector<double> dev(per, 0.0);
    if (h->stat_sp) {
      HIPCHK(hipMemcpyAsync(dev.data(), h->stat_sp + (size_t)m * per, per * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
    }
    // device rows x columns -> the reference's sp(j, column) (j fastest), array by array
    const int rows[5] = {h->nyu + 2, h->nyv + 2, h->nyu + 2, h->nyu + 2, h->nyu};
    size_t off = 0;
    for (int k = 0; k < 5; k++) {
      const int nr = rows[k];
      for (size_t c = 0; c < C; c++)
        for (int j = 0; j < nr; j++) spectra[off + c * nr + j] = dev[off + (size_t)j * C + c];
      off += (size_t)nr * C;
    }
  }
  return DNS_OK;
}

// ---- four_to_phys post-processing (four_to_phys.f90, four_to_phys_N.f90) --------------------------
extern "C" int dns_four_to_phys(int Nx, int Nz, int Mx, int Mz, int nplanes, const double *in, double *out,
                                double *ms) {
  if (nplanes < 0 || (nplanes > 0 && (!in || !out))) return fail(DNS_ERR_ARG, "dns_four_to_phys: bad argument");
  if (Nx < 2 || Nz < 2 || Nx % 2 || Nz % 2 || Mx < Nx || Mz < Nz)
    return fail(DNS_ERR_ARG, "dns_four_to_phys: sizes %d x %d -> %d x %d (need even N <= M)", Nx, Nz, Mx, Mz);
  if (!dnsk::four_to_phys_supported(Mx, Mz))
    return fail(DNS_ERR_ARG, "dns_four_to_phys: %d x %d is not a supported transform size (Mz and Mx/2 in "
                "8..1152, 2^a 3^b)", Mx, Mz);
  if (ms) *ms = 0.0;
  if (nplanes == 0) return DNS_OK;
  // device rows padded to whole 128-byte lines (the host copies re-pitch the rows)
  const int ip = (Nx + 2 + 15) / 16 * 16, op = (Mx + 2 + 15) / 16 * 16, mp = (Nx / 2 + 1 + 7) / 8 * 8;
  const size_t nin = (size_t)ip * Nz, nout = (size_t)op * Mz, nmid = (size_t)mp * Mz;
  const size_t hin = (size_t)(Nx + 2) * Nz, hout = (size_t)(Mx + 2) * Mz;   // host plane sizes
  // planes per batch: at most ~4 GiB of device buffers
  const size_t per = (nin + nout) * sizeof(double) + nmid * sizeof(double2);
  const int nb = (int)std::max<size_t>(1, std::min<size_t>((size_t)nplanes, ((size_t)4 << 30) / per));
  double *din = nullptr, *dout = nullptr;
  double2 *dmid = nullptr, *tw = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = DNS_OK;
  auto hip_ok = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && rc == DNS_OK) rc = fail(DNS_ERR_HIP, "dns_four_to_phys: %s: %s", what, hipGetErrorString(e));
    return rc == DNS_OK;
  };
  const auto a = twiddles<double>(Mz), b = twiddles<double>(Mx), c = twiddles<double>(Mx / 2);
  if (hip_ok(hipMalloc(&din, nb * nin * sizeof(double)), "hipMalloc") &&
      hip_ok(hipMalloc(&dout, nb * nout * sizeof(double)), "hipMalloc") &&
      hip_ok(hipMalloc(&dmid, nb * nmid * sizeof(double2)), "hipMalloc") &&
      hip_ok(hipMalloc(&tw, (size_t)(Mz + Mx + Mx / 2) * sizeof(double2)), "hipMalloc") &&
      hip_ok(hipMemcpy(tw, a.data(), Mz * sizeof(double2), hipMemcpyHostToDevice), "hipMemcpy") &&
      hip_ok(hipMemcpy(tw + Mz, b.data(), Mx * sizeof(double2), hipMemcpyHostToDevice), "hipMemcpy") &&
      hip_ok(hipMemcpy(tw + Mz + Mx, c.data(), (Mx / 2) * sizeof(double2), hipMemcpyHostToDevice), "hipMemcpy") &&
      hip_ok(hipEventCreate(&e0), "hipEventCreate") && hip_ok(hipEventCreate(&e1), "hipEventCreate")) {
    double total = 0.0;
    for (int p0 = 0; p0 < nplanes && rc == DNS_OK; p0 += nb) {
      const int n = std::min(nb, nplanes - p0);
      if (!hip_ok(hipMemcpy2D(din, ip * sizeof(double), in + (size_t)p0 * hin, (Nx + 2) * sizeof(double),
                              (Nx + 2) * sizeof(double), (size_t)n * Nz, hipMemcpyHostToDevice), "upload"))
        break;
      dnsk::FtpArgs fa{Nx, Nz, Mx, Mz, n, ip, mp, op, din, dmid, dout, tw, tw + Mz, tw + Mz + Mx};
      (void)hipEventRecord(e0, 0);
      if (dnsk::launch_four_to_phys(fa, 0)) {
        rc = fail(DNS_ERR_ARG, "dns_four_to_phys: %d x %d not instantiated", Mx, Mz);
        break;
      }
      (void)hipEventRecord(e1, 0);
      if (!hip_ok(hipGetLastError(), "launch") ||
          !hip_ok(hipMemcpy2D(out + (size_t)p0 * hout, (Mx + 2) * sizeof(double), dout, op * sizeof(double),
                              (Mx + 2) * sizeof(double), (size_t)n * Mz, hipMemcpyDeviceToHost), "download"))
        break;
      float t = 0.0f;
      (void)hipEventElapsedTime(&t, e0, e1);
      total += t;
    }
    if (ms && rc == DNS_OK) *ms = total;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(din);
  (void)hipFree(dout);
  (void)hipFree(dmid);
  (void)hipFree(tw);
  return rc;
}

// The same per-plane work on device memory the caller owns, with no device allocation: the z-inverted
// columns and the twiddle tables live in the library's static workspace (dnsk::FTP_SCRATCH_BYTES per GPU,
// post_kernels.hip), so planes are converted in batches that fit it.  Same kernels and twiddles as
// dns_four_to_phys, so the same bits; only the row pitches differ (the caller's dense rows).
extern "C" int dns_four_to_phys_noalloc(int Nx, int Nz, int Mx, int Mz, int nplanes, const double *d_in,
                                        double *d_out, void *stream) {
  if (nplanes < 0 || (nplanes > 0 && (!d_in || !d_out)))
    return fail(DNS_ERR_ARG, "dns_four_to_phys_noalloc: bad argument");
  if (Nx < 2 || Nz < 2 || Nx % 2 || Nz % 2 || Mx < Nx || Mz < Nz)
    return fail(DNS_ERR_ARG, "dns_four_to_phys_noalloc: sizes %d x %d -> %d x %d (need even N <= M)", Nx, Nz, Mx,
                Mz);
  if (!dnsk::four_to_phys_supported(Mx, Mz))
    return fail(DNS_ERR_ARG, "dns_four_to_phys_noalloc: %d x %d is not a supported transform size (Mz and Mx/2 "
                "in 8..1152, 2^a 3^b)", Mx, Mz);
  if (((uintptr_t)d_in | (uintptr_t)d_out) % 16 != 0)
    return fail(DNS_ERR_ARG, "dns_four_to_phys_noalloc: d_in and d_out must be 16-byte aligned");
  if (nplanes == 0) return DNS_OK;
  void *base = nullptr;
  if (dnsk::four_to_phys_scratch(&base))
    return fail(DNS_ERR_HIP, "dns_four_to_phys_noalloc: no address for the workspace on this device");
  const int mp = (Nx / 2 + 1 + 7) / 8 * 8;                                  // complex per intermediate row
  const size_t ntw = (size_t)Mz + Mx + Mx / 2;                              // twiddles, complex
  const size_t midoff = (ntw * sizeof(double2) + 255) / 256 * 256;          // intermediate after them
  const size_t per = (size_t)Mz * mp * sizeof(double2);                     // intermediate of one plane
  if (midoff + per > dnsk::FTP_SCRATCH_BYTES)
    return fail(DNS_ERR_ARG, "dns_four_to_phys_noalloc: %d x %d -> %d x %d needs %zu bytes of workspace for one "
                "plane; the workspace holds %zu", Nx, Nz, Mx, Mz, midoff + per, dnsk::FTP_SCRATCH_BYTES);
  const int nb = (int)std::min<size_t>((size_t)nplanes, (dnsk::FTP_SCRATCH_BYTES - midoff) / per);
  hipStream_t st = static_cast<hipStream_t>(stream);
  double2 *tw = static_cast<double2 *>(base);
  double2 *mid = reinterpret_cast<double2 *>(static_cast<char *>(base) + midoff);
  const auto a = twiddles<double>(Mz), b = twiddles<double>(Mx), c = twiddles<double>(Mx / 2);
  HIPCHK(hipMemcpyAsync(tw, a.data(), Mz * sizeof(double2), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(tw + Mz, b.data(), Mx * sizeof(double2), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(tw + Mz + Mx, c.data(), (Mx / 2) * sizeof(double2), hipMemcpyHostToDevice, st));
  const size_t hin = (size_t)(Nx + 2) * Nz, hout = (size_t)(Mx + 2) * Mz;   // dense planes
  for (int p0 = 0; p0 < nplanes; p0 += nb) {
    const int n = std::min(nb, nplanes - p0);
    dnsk::FtpArgs fa{Nx, Nz, Mx, Mz, n, Nx + 2, mp, Mx + 2, d_in + (size_t)p0 * hin, mid,
                     d_out + (size_t)p0 * hout, tw, tw + Mz, tw + Mz + Mx};
    if (dnsk::launch_four_to_phys(fa, st))
      return fail(DNS_ERR_ARG, "dns_four_to_phys_noalloc: %d x %d not instantiated", Mx, Mz);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(st));  // the host twiddle tables are read before they go out of scope
  return DNS_OK;
}

extern "C" int dns_set_column_path(dns_handle *h, int path) {
  if (!h) return fail(DNS_ERR_ARG, "dns_set_column_path: null handle");
  if (path != 0 && path != 1) return fail(DNS_ERR_ARG, "dns_set_column_path: path=%d must be 0 or 1", path);
  h->prm.column_path = path;
  h->colpass = path == 0 && dnsk::column_pass_supported(h->nyu);
  drop_graph(h);
  return DNS_OK;
}

extern "C" int dns_get_column_path(dns_handle *h, int *path) {
  if (!h || !path) return fail(DNS_ERR_ARG, "dns_get_column_path: null argument");
  *path = h->colpass ? 0 : 1;
  return DNS_OK;
}

extern "C" int dns_sync(dns_handle *h) {
  if (!h) return fail(DNS_ERR_ARG, "dns_sync: null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  return DNS_OK;
}

extern "C" int dns_set_timing(dns_handle *h, int on) {
  if (!h) return fail(DNS_ERR_ARG, "dns_set_timing: null handle");
  h->timing = on != 0;
  memset(h->tsum, 0, sizeof h->tsum);
  h->tcount = 0;
  return DNS_OK;
}

extern "C" int dns_get_timing(dns_handle *h, double *t, int *n) {
  if (!h || !t) return fail(DNS_ERR_ARG, "dns_get_timing: null argument");
  for (int q = 0; q < 7; q++) t[q] = h->tcount ? h->tsum[q] / h->tcount : 0.0;
  if (n) *n = h->tcount;
  return DNS_OK;
}

// ------------------------------------------------------------------------------------------------
extern "C" int dns_plane_fft_roundtrip(dns_handle *h, const double *d_in, double *d_phys, double *d_back,
                                       int nplanes) {
  if (!h || !d_in || !d_phys || !d_back || nplanes < 0) return fail(DNS_ERR_ARG, "dns_plane_fft_roundtrip: bad argument");
  if (nplanes == 0) return DNS_OK;
  if (dnsk::launch_plane_roundtrip(h->Mx, h->Mz, d_in, d_phys, d_back, nplanes, h->twx64, h->twh64, h->twz64,
                                   h->stream))
    return fail(DNS_ERR_ARG, "dns_plane_fft_roundtrip: %d x %d not instantiated", h->Mx, h->Mz);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

extern "C" int dns_modes_to_planes(dns_handle *h, const double *d_cols, int nrows, double *d_planes) {
  if (!h || !d_cols || !d_planes || nrows <= 0) return fail(DNS_ERR_ARG, "dns_modes_to_planes: bad argument");
  dnsk::launch_modes_to_planes(reinterpret_cast<const double2 *>(d_cols), d_planes, h->C, nrows, h->Kx, h->Nz,
                               h->Mx, h->Mz, h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}

extern "C" int dns_planes_to_modes(dns_handle *h, const double *d_planes, int nrows, double *d_cols) {
  if (!h || !d_cols || !d_planes || nrows <= 0) return fail(DNS_ERR_ARG, "dns_planes_to_modes: bad argument");
  dnsk::launch_planes_to_modes(d_planes, reinterpret_cast<double2 *>(d_cols), h->C, nrows, h->Kx, h->Nz, h->Mx,
                               h->Mz, h->stream);
  HIPCHK(hipGetLastError());
  return DNS_OK;
}
