/* ORACLE TEST INFRASTRUCTURE -- never part of the product, never linked into it.
 *
 * FP64 CPU restatement of the reference's RK3 time step for one rank.  Each function cites the
 * reference lines it restates.  The FFTs are a plain Stockham mixed-radix (2/3/4) transform written
 * from the DFT definition; the conventions (normalisation, packing, pruning, zeroed rows) follow
 * rft_buff.f:24-86 and cft_buff.f:1-65 as used by FOU3D.f90.
 *
 * Pinning (what checks each part):
 *   - whole step, every stage: tests/test_oracle_ref.py against the reference's own compiled code
 *     (oracle/_ref/ref_harness) and the golden fixtures written from it;
 *   - FFT: DFT definition (numpy.fft), brute force on tiny sizes, round trip;
 *   - y grid: bit-exact against the y arrays stored in the reference's input fixtures;
 *   - tridiagonal solves: dense solves of the assembled matrices; divergence after the projection.
 */
#include "dns_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

/* ------------------------------------------------------------------------------------------------ */
/* Stockham mixed-radix FFT, radices 4, 2, 3.  sign = -1 forward (e^-), +1 backward (e^+); unnormalised. */
typedef struct {
  int n, nf, fac[40];
  cplx *tw[2]; /* tw[0]: e^{-2 pi i m/n}, tw[1]: e^{+2 pi i m/n} */
} fftplan;

static void plan_init(fftplan *pl, int n) {
  pl->n = n;
  pl->nf = 0;
  int m = n;
  while (m % 4 == 0) { pl->fac[pl->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { pl->fac[pl->nf++] = 2; m /= 2; }
  while (m % 3 == 0) { pl->fac[pl->nf++] = 3; m /= 3; }
  if (m != 1) { pl->nf = -1; } /* unsupported size: caller checks */
  for (int s = 0; s < 2; s++) {
    pl->tw[s] = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    double sg = s == 0 ? -1.0 : 1.0;
    for (int q = 0; q < n; q++) {
      double a = 2.0 * M_PI * (double)q / (double)n;
      pl->tw[s][q] = cos(a) + sg * sin(a) * I;
    }
  }
}

static void plan_free(fftplan *pl) {
  free(pl->tw[0]);
  free(pl->tw[1]);
}

static void fft_run(const fftplan *pl, cplx *x, cplx *work, int sign) {
  const int n = pl->n;
  const cplx *T = pl->tw[sign < 0 ? 0 : 1];
  const double sg = sign < 0 ? -1.0 : 1.0;
  const double s3 = 0.86602540378443864676;
  cplx *a = x, *b = work;
  int p = 1;
  for (int f = 0; f < pl->nf; f++) {
    const int R = pl->fac[f], m = n / R, tstep = n / (p * R);
    for (int i = 0; i < m; i++) {
      const int k = i % p;
      cplx u[4], v[4];
      for (int r = 0; r < R; r++) {
        u[r] = a[i + r * m];
        if (r && k) u[r] *= T[r * k * tstep];
      }
      if (R == 2) {
        v[0] = u[0] + u[1];
        v[1] = u[0] - u[1];
      } else if (R == 3) {
        cplx t1 = u[1] + u[2], t2 = u[0] - 0.5 * t1, t3 = (sg * s3 * I) * (u[1] - u[2]);
        v[0] = u[0] + t1;
        v[1] = t2 + t3;
        v[2] = t2 - t3;
      } else {
        cplx t0 = u[0] + u[2], t1 = u[0] - u[2], t2 = u[1] + u[3], t3 = (sg * I) * (u[1] - u[3]);
        v[0] = t0 + t2;
        v[1] = t1 + t3;
        v[2] = t0 - t2;
        v[3] = t1 - t3;
      }
      const int j0 = (i - k) * R + k;
      for (int s = 0; s < R; s++) b[j0 + s * p] = v[s];
    }
    cplx *t = a; a = b; b = t;
    p *= R;
  }
  if (a != x) memcpy(x, a, sizeof(cplx) * (size_t)n);
}

/* ------------------------------------------------------------------------------------------------ */
/* Plane operations on a (Mx+2) x Mz real buffer, exactly as FOU3D.f90 uses rft/cft.              */
typedef struct {
  int Mx, Mz, Kx;
  fftplan px, pz;
} planeops;

/* cft(u, ise=Mx+2, isa=2, m=Kx, iopt) : cft_buff.f:1-65; forward scales by 1/Mz */
static void cft_pruned(const planeops *po, double *pl, int sign, cplx *buf, cplx *work) {
  const int igal = po->Mx + 2, Mz = po->Mz;
  const double dn = 1.0 / Mz;
  for (int i = 0; i < po->Kx; i++) {
    for (int z = 0; z < Mz; z++) buf[z] = pl[z * igal + 2 * i] + pl[z * igal + 2 * i + 1] * I;
    fft_run(&po->pz, buf, work, sign);
    for (int z = 0; z < Mz; z++) {
      pl[z * igal + 2 * i] = sign < 0 ? dn * creal(buf[z]) : creal(buf[z]);
      pl[z * igal + 2 * i + 1] = sign < 0 ? dn * cimag(buf[z]) : cimag(buf[z]);
    }
  }
}

/* rft(u, isa=Mx+2, m=Mz, +1) : rft_buff.f:76-84 -- Hermitian synthesis; Im c0 and Im c_{n/2} ignored,
 * entries Mx, Mx+1 left as they were. */
static void rft_inverse_rows(const planeops *po, double *pl, cplx *buf, cplx *work) {
  const int igal = po->Mx + 2, n = po->Mx;
  for (int z = 0; z < po->Mz; z++) {
    double *row = pl + (size_t)z * igal;
    buf[0] = row[0];
    buf[n / 2] = row[n];
    for (int k = 1; k < n / 2; k++) {
      buf[k] = row[2 * k] + row[2 * k + 1] * I;
      buf[n - k] = row[2 * k] - row[2 * k + 1] * I;
    }
    fft_run(&po->px, buf, work, +1);
    for (int m = 0; m < n; m++) row[m] = creal(buf[m]);
  }
}

/* rft(u, isa=Mx+2, m=Mz, -1) : rft_buff.f:62-74 -- forward, scaled 1/n, Im c0 = Im c_{n/2} = 0 */
static void rft_forward_rows(const planeops *po, double *pl, cplx *buf, cplx *work) {
  const int igal = po->Mx + 2, n = po->Mx;
  const double dn = 1.0 / n;
  for (int z = 0; z < po->Mz; z++) {
    double *row = pl + (size_t)z * igal;
    for (int m = 0; m < n; m++) buf[m] = row[m];
    fft_run(&po->px, buf, work, -1);
    for (int k = 0; k <= n / 2; k++) {
      row[2 * k] = dn * creal(buf[k]);
      row[2 * k + 1] = dn * cimag(buf[k]);
    }
    row[1] = 0.0;
    row[n + 1] = 0.0;
  }
}

/* four_to_phys_du : FOU3D.f90:681-698 */
static void four_to_phys(const planeops *po, double *pl, cplx *buf, cplx *work) {
  const int igal = po->Mx + 2;
  memset(pl + (size_t)(po->Mz / 2) * igal, 0, sizeof(double) * igal);
  cft_pruned(po, pl, +1, buf, work);
  rft_inverse_rows(po, pl, buf, work);
}

/* phys_to_four_du : FOU3D.f90:1654-1672 */
static void phys_to_four(const planeops *po, double *pl, cplx *buf, cplx *work) {
  const int igal = po->Mx + 2;
  rft_forward_rows(po, pl, buf, work);
  cft_pruned(po, pl, -1, buf, work);
  memset(pl + (size_t)(po->Mz / 2) * igal, 0, sizeof(double) * igal);
}

static void planeops_init(planeops *po, int Mx, int Mz, int Nx) {
  po->Mx = Mx;
  po->Mz = Mz;
  po->Kx = Nx / 2 + 1;
  plan_init(&po->px, Mx);
  plan_init(&po->pz, Mz);
}

static void planeops_free(planeops *po) {
  plan_free(&po->px);
  plan_free(&po->pz);
}

/* ------------------------------------------------------------------------------------------------ */
struct orc {
  int Mx, Mz, Nx, Nz, Kx, C, ny, nyu, nyv, nthreads, ppp, flag_ctpress, iter;
  int nu_rows, nv_rows, np_rows; /* ny+1, ny, ny-1 */
  double Re, alp, bet, CFL, mpgx, dyq, t, dt, dti, dtc, dtv, Qx, QxT, dgx;
  double fixed_dt;  /* > 0: every step takes this dt instead of the CFL rule */
  double gw[2], gbc_u1, gbc_u3, gi[2];
  double aRK[3], bRK[3], gRK[3], cRK[3], dRK[3];
  double *yu, *yv, *dthdyu, *dthdyv, *dyu2i, *dyv2i; /* dy?2i[3*j + (c-1)] */
  double *kx, *kz, *k2x, *k2z;                       /* Im of k1F_x(i), k1F_z(k); k2F */
  double *DG;                                        /* DG[(col*np_rows + (j-1))*3 + (c-1)] */
  cplx *u1, *u2, *u3, *p, *psi, *div;
  cplx *du1, *du2, *du3, *Nu1, *Nu2, *Nu3;
  cplx *u1_itp, *u2_itp, *u3_itp, *uv_f, *vv_c, *wv_f;
  planeops po;
};

static double powi(double a, int b) { /* binary exponentiation, as compilers expand x**n */
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

#define D2U(o, c, j) ((o)->dyu2i[3 * (j) + (c)-1])
#define D2V(o, c, j) ((o)->dyv2i[3 * (j) + (c)-1])

/* ygrid : start.f90:665-806 */
static void setup_grid(orc_t *o) {
  const int nn = o->nyv + 1, ppp = o->ppp;
  const double dyq = o->dyq;
  const double qqq = nn / 2.0;
  const double bbb = (dyq - 1.0) /
      (qqq * (1.0 - powi(qqq, ppp - 1) + dyq * ((1.0 - qqq) * powi(qqq, ppp - 1) - powi(1.0 - qqq, ppp))));
  const double aaa = (1.0 - bbb * powi(qqq, ppp)) / qqq;
  double *yu = o->yu, *yv = o->yv;
  yu[0] = aaa * (-1.0 + .5 - qqq) + bbb * powi(-1.0 + .5 - qqq, ppp);
  for (int j = 0; j <= nn; j++) {
    yv[j] = aaa * (j - qqq) + bbb * powi(j - qqq, ppp);
    yu[j + 1] = aaa * (j + .5 - qqq) + bbb * powi(j + .5 - qqq, ppp);
  }
  o->dthdyu[0] = 1.0 / (aaa + ppp * bbb * powi(-1.0 + .5 - qqq, ppp - 1));
  for (int j = 0; j <= nn; j++) {
    o->dthdyv[j] = 1.0 / (aaa + ppp * bbb * powi(j - qqq, ppp - 1));
    o->dthdyu[j + 1] = 1.0 / (aaa + ppp * bbb * powi(j + .5 - qqq, ppp - 1));
  }
  int j = 0;
  D2U(o, 1, j) = 0.0;
  D2U(o, 2, j) = -2.0 / ((yu[j + 1] - yu[j]) * (yu[j + 1] - yu[j]));
  D2U(o, 3, j) = 1.0 / ((yu[j + 1] - yu[j]) * (yu[j + 1] - yu[j]));
  D2V(o, 1, j) = 0.0;
  D2V(o, 2, j) = -2.0 / ((yv[j + 1] - yv[j]) * (yv[j + 1] - yv[j]));
  D2V(o, 3, j) = 1.0 / ((yv[j + 1] - yv[j]) * (yv[j + 1] - yv[j]));
  for (j = 1; j <= nn - 1; j++) {
    D2U(o, 1, j) = 2.0 / ((yu[j] - yu[j - 1]) * (yu[j + 1] - yu[j - 1]));
    D2U(o, 2, j) = -2.0 / ((yu[j + 1] - yu[j]) * (yu[j] - yu[j - 1]));
    D2U(o, 3, j) = 2.0 / ((yu[j + 1] - yu[j]) * (yu[j + 1] - yu[j - 1]));
    D2V(o, 1, j) = 2.0 / ((yv[j] - yv[j - 1]) * (yv[j + 1] - yv[j - 1]));
    D2V(o, 2, j) = -2.0 / ((yv[j + 1] - yv[j]) * (yv[j] - yv[j - 1]));
    D2V(o, 3, j) = 2.0 / ((yv[j + 1] - yv[j]) * (yv[j + 1] - yv[j - 1]));
  }
  j = nn;
  D2U(o, 1, j) = 2.0 / ((yu[j] - yu[j - 1]) * (yu[j + 1] - yu[j - 1]));
  D2U(o, 2, j) = -2.0 / ((yu[j + 1] - yu[j]) * (yu[j] - yu[j - 1]));
  D2U(o, 3, j) = 2.0 / ((yu[j + 1] - yu[j]) * (yu[j + 1] - yu[j - 1]));
  D2V(o, 1, j) = 1.0 / ((yv[j] - yv[j - 1]) * (yv[j] - yv[j - 1]));
  D2V(o, 2, j) = -2.0 / ((yv[j] - yv[j - 1]) * (yv[j] - yv[j - 1]));
  D2V(o, 3, j) = 0.0;
  j = nn + 1;
  D2U(o, 1, j) = 1.0 / ((yu[j] - yu[j - 1]) * (yu[j] - yu[j - 1]));
  D2U(o, 2, j) = -2.0 / ((yu[j] - yu[j - 1]) * (yu[j] - yu[j - 1]));
  D2U(o, 3, j) = 0.0;
}

/* LU_buildP + LU_decP : tridLU_3D.f90:224-297, 366-405 (rows j = 1..nyu stored at j-1) */
static void setup_DG(orc_t *o) {
  const int n = o->np_rows, s = 1, e = o->nyu;
  double *Db = (double *)calloc((size_t)e + 3, sizeof(double)); /* index j - (s-1) */
  double *Gb = (double *)calloc((size_t)e + 3, sizeof(double));
  double *Dt = (double *)calloc((size_t)e + 3, sizeof(double));
  double *Gt = (double *)calloc((size_t)e + 3, sizeof(double));
#define IX(j) ((j) - (s - 1))
  for (int j = s; j <= e; j++) {
    Db[IX(j - 1)] = -1.0 * o->dthdyu[j];
    Gb[IX(j)] = -1.0 * o->dthdyv[j];
    Dt[IX(j)] = 1.0 * o->dthdyu[j];
    Gt[IX(j + 1)] = 1.0 * o->dthdyv[j];
  }
  for (int col = 0; col < o->C; col++) {
    const int i = col % o->Kx, k = col / o->Kx;
    const double k2x = o->k2x[i], k2z = o->k2z[k];
    double *a = o->DG + (size_t)col * n * 3;
#define A(c, j) a[((j)-s) * 3 + (c)-1]
    A(1, s) = 0.0;
    A(2, s) = (Dt[IX(s)] * Gb[IX(s)]) + k2x + k2z;
    A(3, s) = (Dt[IX(s)] * Gt[IX(s + 1)]);
    A(2, e) = (Db[IX(e - 1)] * Gt[IX(e)]) + k2x + k2z;
    A(1, e) = (Db[IX(e - 1)] * Gb[IX(e - 1)]);
    A(3, e) = 0.0;
    for (int j = s + 1; j <= e - 1; j++) {
      A(2, j) = (Db[IX(j - 1)] * Gt[IX(j)]) + (Dt[IX(j)] * Gb[IX(j)]) + k2x + k2z;
      A(1, j) = (Db[IX(j - 1)] * Gb[IX(j - 1)]);
      A(3, j) = (Dt[IX(j)] * Gt[IX(j + 1)]);
    }
    if (col == 0) {
      const double d = o->yu[2] - o->yu[1];
      A(1, s) = 0.0;
      A(2, s) = 1.0 / (d * d);
      A(3, s) = 0.0;
    }
    A(2, s) = 1.0 / A(2, s);
    for (int j = s + 1; j <= e; j++) {
      A(3, j - 1) = A(3, j - 1) * A(2, j - 1);
      A(2, j) = 1.0 / (A(2, j) - A(1, j) * A(3, j - 1));
    }
#undef A
  }
#undef IX
  free(Db); free(Gb); free(Dt); free(Gt);
}

orc_t *orc_create(int Mx, int Mz, int ny, double Re, double alp, double bet, double CFL, double mpgx,
                  double dyq, int ppp, int flag_ctpress) {
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->Mx = Mx; o->Mz = Mz; o->Nx = (2 * Mx) / 3; o->Nz = (2 * Mz) / 3;
  o->Kx = o->Nx / 2 + 1; o->C = o->Kx * o->Nz; o->ny = ny;
  o->nyv = ny - 2; o->nyu = o->nyv + 1;
  o->nu_rows = ny + 1; o->nv_rows = ny; o->np_rows = ny - 1;
  o->Re = Re; o->alp = alp; o->bet = bet; o->CFL = CFL; o->mpgx = mpgx; o->dyq = dyq; o->ppp = ppp;
  o->flag_ctpress = flag_ctpress;
#ifdef _OPENMP
  o->nthreads = omp_get_max_threads();
#else
  o->nthreads = 1;
#endif
  const double aR[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0}, gR[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
  const double cR[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0}, dR[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};
  for (int q = 0; q < 3; q++) { /* start.f90:439-453 */
    o->aRK[q] = aR[q]; o->bRK[q] = aR[q]; o->gRK[q] = gR[q]; o->cRK[q] = cR[q]; o->dRK[q] = dR[q];
  }
  o->yu = (double *)calloc((size_t)ny + 1, sizeof(double));
  o->yv = (double *)calloc((size_t)ny, sizeof(double));
  o->dthdyu = (double *)calloc((size_t)ny + 1, sizeof(double));
  o->dthdyv = (double *)calloc((size_t)ny, sizeof(double));
  o->dyu2i = (double *)calloc(3 * ((size_t)ny + 1), sizeof(double));
  o->dyv2i = (double *)calloc(3 * (size_t)ny, sizeof(double));
  setup_grid(o);
  /* def_k : start.f90:809-862 */
  o->kx = (double *)calloc(o->Kx, sizeof(double));
  o->k2x = (double *)calloc(o->Kx, sizeof(double));
  o->kz = (double *)calloc(o->Nz, sizeof(double));
  o->k2z = (double *)calloc(o->Nz, sizeof(double));
  for (int i = 0; i < o->Kx; i++) {
    o->kx[i] = alp * i;
    o->k2x[i] = -((alp * i) * (alp * i));
  }
  for (int k = 0; k < o->Nz; k++) {
    if (k < o->Nz / 2) {
      o->kz[k] = bet * k;
      o->k2z[k] = -((bet * k) * (bet * k));
    } else {
      o->kz[k] = -(bet * (o->Nz - k));
      o->k2z[k] = -((bet * (o->Nz - k)) * (bet * (o->Nz - k)));
    }
  }
  /* dtv : start.f90:476 */
  const double dyv0 = o->yv[1] - o->yv[0];
  o->dtv = Re / (-o->k2x[o->Nx / 2] - o->k2z[o->Nz / 2 - 1] + 4.0 / (dyv0 * dyv0)); /* k2F_z(Nz/2), 1-based */
  /* wall weights : start.f90:518-537 */
  const int nyu = o->nyu, nyv = o->nyv;
  o->gw[0] = -(o->yv[0] - o->yu[0]) / (o->yv[0] - o->yu[1]);
  o->gw[1] = (o->yu[nyu + 1] - o->yv[nyv + 1]) / (o->yv[nyv + 1] - o->yu[nyu]);
  o->gbc_u1 = (o->yu[0] - o->yv[0]) / (o->yu[1] - o->yv[0]);
  o->gbc_u3 = o->gbc_u1;
  o->gi[0] = (o->yu[0] - o->yu[1]) / (o->yv[0] - o->yu[1]);
  o->gi[1] = (o->yu[nyu + 1] - o->yu[nyu]) / (o->yv[nyv + 1] - o->yu[nyu]);
  o->DG = (double *)calloc((size_t)3 * o->np_rows * o->C, sizeof(double));
  setup_DG(o);
  const size_t C = o->C;
  o->u1 = (cplx *)calloc(C * o->nu_rows, sizeof(cplx));
  o->u3 = (cplx *)calloc(C * o->nu_rows, sizeof(cplx));
  o->u2 = (cplx *)calloc(C * o->nv_rows, sizeof(cplx));
  o->p = (cplx *)calloc(C * o->np_rows, sizeof(cplx));
  o->psi = (cplx *)calloc(C * o->np_rows, sizeof(cplx));
  o->div = (cplx *)calloc(C * o->np_rows, sizeof(cplx));
  o->du1 = (cplx *)calloc(C * o->nu_rows, sizeof(cplx));
  o->du3 = (cplx *)calloc(C * o->nu_rows, sizeof(cplx));
  o->du2 = (cplx *)calloc(C * o->nv_rows, sizeof(cplx));
  o->Nu1 = (cplx *)calloc(C * nyu, sizeof(cplx));
  o->Nu3 = (cplx *)calloc(C * nyu, sizeof(cplx));
  o->Nu2 = (cplx *)calloc(C * nyv, sizeof(cplx));
  o->u1_itp = (cplx *)calloc(C * o->nv_rows, sizeof(cplx));
  o->u3_itp = (cplx *)calloc(C * o->nv_rows, sizeof(cplx));
  o->u2_itp = (cplx *)calloc(C * o->nu_rows, sizeof(cplx));
  o->uv_f = (cplx *)calloc(C * o->nv_rows, sizeof(cplx));
  o->wv_f = (cplx *)calloc(C * o->nv_rows, sizeof(cplx));
  o->vv_c = (cplx *)calloc(C * o->nu_rows, sizeof(cplx));
  planeops_init(&o->po, Mx, Mz, o->Nx);
  if (o->po.px.nf < 0 || o->po.pz.nf < 0) {
    orc_destroy(o);
    return NULL;
  }
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  void *ptrs[] = {o->yu, o->yv, o->dthdyu, o->dthdyv, o->dyu2i, o->dyv2i, o->kx, o->kz, o->k2x, o->k2z,
                  o->DG, o->u1, o->u2, o->u3, o->p, o->psi, o->div, o->du1, o->du2, o->du3, o->Nu1,
                  o->Nu2, o->Nu3, o->u1_itp, o->u2_itp, o->u3_itp, o->uv_f, o->vv_c, o->wv_f};
  for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); q++) free(ptrs[q]);
  planeops_free(&o->po);
  free(o);
}

void orc_sizes(const orc_t *o, int *out) {
  out[0] = o->Nx; out[1] = o->Nz; out[2] = o->Kx; out[3] = o->C;
  out[4] = o->nyu; out[5] = o->nyv; out[6] = o->nthreads; out[7] = o->Mx;
}

void orc_set_threads(orc_t *o, int n) { o->nthreads = n > 0 ? n : 1; }

/* flowrateIm / flowrateRe : littleharsh.f90:473-513 (mixed u/v grid rule) */
static double flowrate_re_part(const orc_t *o, const cplx *ucol) {
  double Q = 0.0;
  for (int j = 1; j <= o->nyv + 1; j++) Q = Q + (o->yv[j] - o->yv[j - 1]) * creal(ucol[j]);
  return Q;
}

/* divergence : littleharsh.f90:242-278 */
static void divergence(orc_t *o) {
  const int nu = o->nu_rows, nv = o->nv_rows, npr = o->np_rows;
#pragma omp parallel for num_threads(o->nthreads)
  for (int col = 0; col < o->C; col++) {
    const int i = col % o->Kx, k = col / o->Kx;
    const cplx kx = o->kx[i] * I, kz = o->kz[k] * I;
    const cplx *u1 = o->u1 + (size_t)col * nu, *u3 = o->u3 + (size_t)col * nu, *u2 = o->u2 + (size_t)col * nv;
    cplx *dv = o->div + (size_t)col * npr;
    for (int j = 1; j <= o->nyu; j++)
      dv[j - 1] = kx * u1[j] + kz * u3[j] + (1.0 * o->dthdyu[j]) * (u2[j] - u2[j - 1]);
  }
}

void orc_set_state(orc_t *o, const double *u1, const double *u2, const double *u3, const double *p,
                   double t, int iter, double QxT) {
  const size_t C = o->C;
  memcpy(o->u1, u1, sizeof(cplx) * C * o->nu_rows);
  memcpy(o->u2, u2, sizeof(cplx) * C * o->nv_rows);
  memcpy(o->u3, u3, sizeof(cplx) * C * o->nu_rows);
  memcpy(o->p, p, sizeof(cplx) * C * o->np_rows);
  o->t = t;
  o->iter = iter;
  /* getini : start.f90:1255-1263 */
  o->Qx = flowrate_re_part(o, o->u1);
  o->QxT = QxT > 0.0 ? QxT : o->Qx;
  divergence(o);
}

void orc_get_state(const orc_t *o, double *u1, double *u2, double *u3, double *p) {
  const size_t C = o->C;
  if (u1) memcpy(u1, o->u1, sizeof(cplx) * C * o->nu_rows);
  if (u2) memcpy(u2, o->u2, sizeof(cplx) * C * o->nv_rows);
  if (u3) memcpy(u3, o->u3, sizeof(cplx) * C * o->nu_rows);
  if (p) memcpy(p, o->p, sizeof(cplx) * C * o->np_rows);
}

void orc_set_nu(orc_t *o, const double *Nu1, const double *Nu2, const double *Nu3) {
  const size_t C = o->C;
  memcpy(o->Nu1, Nu1, sizeof(cplx) * C * o->nyu);
  memcpy(o->Nu2, Nu2, sizeof(cplx) * C * o->nyv);
  memcpy(o->Nu3, Nu3, sizeof(cplx) * C * o->nyu);
}

void orc_get_nu(const orc_t *o, double *Nu1, double *Nu2, double *Nu3) {
  const size_t C = o->C;
  memcpy(Nu1, o->Nu1, sizeof(cplx) * C * o->nyu);
  memcpy(Nu2, o->Nu2, sizeof(cplx) * C * o->nyv);
  memcpy(Nu3, o->Nu3, sizeof(cplx) * C * o->nyu);
}

void orc_get_du(const orc_t *o, double *du1, double *du2, double *du3) {
  cplx *a = (cplx *)du1, *b = (cplx *)du2, *c = (cplx *)du3;
  for (int col = 0; col < o->C; col++) {
    for (int j = 1; j <= o->nyu; j++) {
      a[(size_t)col * o->nyu + j - 1] = o->du1[(size_t)col * o->nu_rows + j];
      c[(size_t)col * o->nyu + j - 1] = o->du3[(size_t)col * o->nu_rows + j];
    }
    for (int j = 1; j <= o->nyv; j++) b[(size_t)col * o->nyv + j - 1] = o->du2[(size_t)col * o->nv_rows + j];
  }
}

void orc_get_flux(const orc_t *o, double *uv_f, double *vv_c, double *wv_f) {
  const size_t C = o->C;
  memcpy(uv_f, o->uv_f, sizeof(cplx) * C * o->nv_rows);
  memcpy(vv_c, o->vv_c, sizeof(cplx) * C * o->nu_rows);
  memcpy(wv_f, o->wv_f, sizeof(cplx) * C * o->nv_rows);
}

void orc_get_psi(const orc_t *o, double *psi, double *div) {
  const size_t C = o->C;
  if (psi) memcpy(psi, o->psi, sizeof(cplx) * C * o->np_rows);
  if (div) memcpy(div, o->div, sizeof(cplx) * C * o->np_rows);
}

void orc_get_scalars(const orc_t *o, double *sc) {
  sc[0] = o->t; sc[1] = o->dt; sc[2] = o->dti; sc[3] = o->mpgx;
  sc[4] = o->Qx; sc[5] = o->dgx; sc[6] = o->dtc; sc[7] = o->QxT;
}

void orc_set_fixed_dt(orc_t *o, double dt) { o->fixed_dt = dt; }

void orc_set_dt(orc_t *o, double dt) {
  o->dt = dt;
  o->dti = 1.0 / dt;
}

void orc_get_setup(const orc_t *o, double *yu, double *yv, double *dthdyu, double *dthdyv, double *dyu2i,
                   double *dyv2i, double *DG, double *w) {
  const int ny = o->ny;
  if (yu) memcpy(yu, o->yu, sizeof(double) * (ny + 1));
  if (yv) memcpy(yv, o->yv, sizeof(double) * ny);
  if (dthdyu) memcpy(dthdyu, o->dthdyu, sizeof(double) * (ny + 1));
  if (dthdyv) memcpy(dthdyv, o->dthdyv, sizeof(double) * ny);
  if (dyu2i) memcpy(dyu2i, o->dyu2i, sizeof(double) * 3 * (ny + 1));
  if (dyv2i) memcpy(dyv2i, o->dyv2i, sizeof(double) * 3 * ny);
  if (DG) memcpy(DG, o->DG, sizeof(double) * 3 * (size_t)o->np_rows * o->C);
  if (w) {
    w[0] = o->dtv; w[1] = o->gw[0]; w[2] = o->gw[1]; w[3] = o->gbc_u1; w[4] = o->gi[0]; w[5] = o->gi[1];
  }
}

/* ------------------------------------------------------------------------------------------------ */
/* RHS0_u1/u2/u3 + laplacian_U/V : littleharsh.f90:281-368, 682-856 */
void orc_rhs0(orc_t *o, int kRK) {
  const int q = kRK - 1, nu = o->nu_rows, nv = o->nv_rows, npr = o->np_rows;
  const double C1 = o->aRK[q] / o->Re, C2 = -o->cRK[q];
#pragma omp parallel for num_threads(o->nthreads)
  for (int col = 0; col < o->C; col++) {
    const int i = col % o->Kx, k = col / o->Kx;
    const double k2 = o->k2x[i] + o->k2z[k];
    const cplx *u1 = o->u1 + (size_t)col * nu, *u3 = o->u3 + (size_t)col * nu, *u2 = o->u2 + (size_t)col * nv;
    const cplx *p = o->p + (size_t)col * npr;
    const cplx *N1 = o->Nu1 + (size_t)col * o->nyu, *N3 = o->Nu3 + (size_t)col * o->nyu;
    const cplx *N2 = o->Nu2 + (size_t)col * o->nyv;
    cplx *d1 = o->du1 + (size_t)col * nu, *d3 = o->du3 + (size_t)col * nu, *d2 = o->du2 + (size_t)col * nv;
    const cplx C3x = -(o->dRK[q] * (o->kx[i] * I)), C3z = -(o->dRK[q] * (o->kz[k] * I));
    for (int j = 1; j <= o->nyu; j++) {
      const double a = D2U(o, 1, j), b = k2 + D2U(o, 2, j), c = D2U(o, 3, j);
      const cplx L1 = a * u1[j - 1] + b * u1[j] + c * u1[j + 1];
      const cplx L3 = a * u3[j - 1] + b * u3[j] + c * u3[j + 1];
      d1[j] = C1 * L1 + C2 * N1[j - 1] + C3x * p[j - 1];
      d3[j] = C1 * L3 + C2 * N3[j - 1] + C3z * p[j - 1];
    }
    for (int j = 1; j <= o->nyv; j++) {
      const double a = D2V(o, 1, j), b = k2 + D2V(o, 2, j), c = D2V(o, 3, j);
      const cplx L2 = a * u2[j - 1] + b * u2[j] + c * u2[j + 1];
      const double C3 = -(o->dRK[q] * 1.0) * o->dthdyv[j];
      d2[j] = C1 * L2 + C2 * N2[j - 1] + C3 * (p[j] - p[j - 1]); /* p(j+1) - p(j), p stored at j-1 */
    }
  }
  /* mean pressure gradient on mode (0,0) : littleharsh.f90:727-733 */
  const double Cm = -(o->dRK[q] * o->mpgx);
  for (int j = 1; j <= o->nyu; j++) o->du1[j] = o->du1[j] + Cm;
}

/* interp_u / interp_v : FOU3D.f90:373-418 */
static void interp(orc_t *o) {
  const int nu = o->nu_rows, nv = o->nv_rows;
  const double *yu = o->yu, *yv = o->yv;
#pragma omp parallel for num_threads(o->nthreads)
  for (int col = 0; col < o->C; col++) {
    const cplx *u1 = o->u1 + (size_t)col * nu, *u3 = o->u3 + (size_t)col * nu, *u2 = o->u2 + (size_t)col * nv;
    cplx *a1 = o->u1_itp + (size_t)col * nv, *a3 = o->u3_itp + (size_t)col * nv, *a2 = o->u2_itp + (size_t)col * nu;
    for (int j = 0; j <= o->nyv + 1; j++) {
      const double w1 = yv[j] - yu[j], w0 = yu[j + 1] - yv[j], d = yu[j + 1] - yu[j];
      a1[j] = (w1 * u1[j + 1] + w0 * u1[j]) / d;
      a3[j] = (w1 * u3[j + 1] + w0 * u3[j]) / d;
    }
    for (int j = 1; j <= o->nyu; j++) {
      const double w1 = yu[j] - yv[j - 1], w0 = yv[j] - yu[j], d = yv[j] - yv[j - 1];
      a2[j] = (w1 * u2[j] + w0 * u2[j - 1]) / d;
    }
    a2[0] = o->gi[0] * (u2[0] - a2[1]) + a2[1];
    a2[o->nyu + 1] = o->gi[1] * (u2[o->nyv + 1] - a2[o->nyu]) + a2[o->nyu];
  }
}

/* modes_to_planes for one plane: spectral column values -> (Mx+2) x Mz buffer (FOU3D.f90:752-759) */
static void col_to_plane(const orc_t *o, const cplx *f, int nrows, int j, double *pl) {
  const int igal = o->Mx + 2;
  memset(pl, 0, sizeof(double) * (size_t)igal * o->Mz);
  for (int col = 0; col < o->C; col++) {
    const int i = col % o->Kx, k = col / o->Kx;
    const int kp = k < o->Nz / 2 ? k : k + o->Mz - o->Nz;
    const cplx v = f[(size_t)col * nrows + j];
    pl[kp * igal + 2 * i] = creal(v);
    pl[kp * igal + 2 * i + 1] = cimag(v);
  }
}

/* Plane stage of nonlinear: ops_in_planes2 (FOU3D.f90:1514-1652) with der_x / der_z (FOU3D.f90:420-523)
 * and the CFL maxima of dtc_calc (FOU3D.f90:613-651).  u-planes and v-planes 1..nyu (one rank). */
static void plane_stage(orc_t *o, int kRK, double *u1m_out, double *u3m_out) {
  const int igal = o->Mx + 2, Mz = o->Mz, Mx = o->Mx, nu = o->nu_rows, nv = o->nv_rows;
  const size_t plsz = (size_t)igal * Mz;
  const int nplanes = 2 * o->nyu;
  double u1m = 0.0, u3m = 0.0;
#pragma omp parallel num_threads(o->nthreads) reduction(max : u1m, u3m)
  {
    double *pa = (double *)malloc(sizeof(double) * plsz * 7);
    double *pb = pa + plsz, *pc = pb + plsz, *q1 = pc + plsz, *q2 = q1 + plsz, *q3 = q2 + plsz, *q4 = q3 + plsz;
    const int nbuf = Mx > Mz ? Mx : Mz;
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nbuf), *work = (cplx *)malloc(sizeof(cplx) * nbuf);
#pragma omp for schedule(dynamic)
    for (int w = 0; w < nplanes; w++) {
      const int vplane = w >= o->nyu, j = vplane ? w - o->nyu + 1 : w + 1;
      if (!vplane) {
        col_to_plane(o, o->u1, nu, j, pa);
        col_to_plane(o, o->u2_itp, nu, j, pb);
        col_to_plane(o, o->u3, nu, j, pc);
      } else {
        col_to_plane(o, o->u1_itp, nv, j, pa);
        col_to_plane(o, o->u2, nv, j, pb);
        col_to_plane(o, o->u3_itp, nv, j, pc);
      }
      four_to_phys(&o->po, pa, buf, work);
      four_to_phys(&o->po, pb, buf, work);
      four_to_phys(&o->po, pc, buf, work);
      if (!vplane && kRK == 1 && j <= o->nyu - 1) { /* dtc_calc sees u-planes 0..nyu-1 */
        for (size_t e = 0; e < plsz; e++) {
          u1m = fmax(u1m, fabs(pa[e]));
          u3m = fmax(u3m, fabs(pc[e]));
        }
      }
      for (int z = 0; z < Mz; z++)
        for (int x = 0; x < Mx; x++) {
          const size_t e = (size_t)z * igal + x;
          if (!vplane) {
            q1[e] = pa[e] * pa[e]; /* uu */
            q2[e] = pa[e] * pc[e]; /* uw */
            q3[e] = pb[e] * pb[e]; /* vv */
            q4[e] = pc[e] * pc[e]; /* ww */
          } else {
            q1[e] = pa[e] * pb[e]; /* uv */
            q2[e] = pc[e] * pb[e]; /* wv */
          }
        }
      const int nprod = vplane ? 2 : 4;
      double *qs[4] = {q1, q2, q3, q4};
      for (int s = 0; s < nprod; s++) phys_to_four(&o->po, qs[s], buf, work);
      /* der_x(a) + der_z(b) on the retained modes, kz = -Nz/2 row zeroed; raw flux values kept */
      for (int col = 0; col < o->C; col++) {
        const int i = col % o->Kx, k = col / o->Kx;
        const int kp = k < o->Nz / 2 ? k : k + o->Mz - o->Nz;
        const int masked = (k == o->Nz / 2);
        const size_t e = (size_t)kp * igal + 2 * i;
        const cplx kx = o->kx[i] * I, kz = o->kz[k] * I;
        if (!vplane) {
          const cplx uu = q1[e] + q1[e + 1] * I, uw = q2[e] + q2[e + 1] * I, ww = q4[e] + q4[e + 1] * I;
          const cplx vv = q3[e] + q3[e + 1] * I;
          o->Nu1[(size_t)col * o->nyu + j - 1] = masked ? 0.0 : kx * uu + kz * uw;
          o->Nu3[(size_t)col * o->nyu + j - 1] = masked ? 0.0 : kx * uw + kz * ww;
          o->vv_c[(size_t)col * nu + j] = vv;
        } else {
          const cplx uv = q1[e] + q1[e + 1] * I, wv = q2[e] + q2[e + 1] * I;
          if (j <= o->nyv) o->Nu2[(size_t)col * o->nyv + j - 1] = masked ? 0.0 : kx * uv + kz * wv;
          o->uv_f[(size_t)col * nv + j] = uv;
          o->wv_f[(size_t)col * nv + j] = wv;
        }
      }
    }
    free(pa); free(buf); free(work);
  }
  /* ghost plane 0 of u1PL/u3PL is still spectral when dtc_calc reads it */
  for (int col = 0; col < o->C; col++) {
    const cplx a = o->u1[(size_t)col * nu], b = o->u3[(size_t)col * nu];
    u1m = fmax(u1m, fmax(fabs(creal(a)), fabs(cimag(a))));
    u3m = fmax(u3m, fmax(fabs(creal(b)), fabs(cimag(b))));
  }
  *u1m_out = u1m;
  *u3m_out = u3m;
}

/* nonlinear : FOU3D.f90:51-371 */
void orc_nonlinear(orc_t *o, int kRK) {
  const int q = kRK - 1, nu = o->nu_rows, nv = o->nv_rows;
  interp(o);
  for (int col = 0; col < o->C; col++) { /* planes 0 of uv_fPL/wv_fPL are never computed: zero */
    o->uv_f[(size_t)col * nv] = 0.0;
    o->wv_f[(size_t)col * nv] = 0.0;
    o->vv_c[(size_t)col * nu] = 0.0;
    o->vv_c[(size_t)col * nu + o->nyu + 1] = 0.0;
  }
  double u1m, u3m;
  plane_stage(o, kRK, &u1m, &u3m);
  /* der_yu_h / der_yv_h and the sum : FOU3D.f90:268-281 */
#pragma omp parallel for num_threads(o->nthreads)
  for (int col = 0; col < o->C; col++) {
    cplx *N1 = o->Nu1 + (size_t)col * o->nyu, *N3 = o->Nu3 + (size_t)col * o->nyu, *N2 = o->Nu2 + (size_t)col * o->nyv;
    const cplx *uv = o->uv_f + (size_t)col * nv, *wv = o->wv_f + (size_t)col * nv, *vv = o->vv_c + (size_t)col * nu;
    for (int j = 1; j <= o->nyu; j++) {
      N1[j - 1] = N1[j - 1] + ((uv[j] - uv[j - 1]) * o->dthdyu[j]) * 1.0;
      N3[j - 1] = N3[j - 1] + ((wv[j] - wv[j - 1]) * o->dthdyu[j]) * 1.0;
    }
    for (int j = 1; j <= o->nyv; j++) N2[j - 1] = N2[j - 1] + ((vv[j + 1] - vv[j]) * o->dthdyv[j]) * 1.0;
  }
  /* CFL : FOU3D.f90:290-303, 642.  dthetavi is never assigned in the reference (declaration.f90:64),
   * so the v term is 1/(u2m*0) = +Inf and never limits dt. */
  if (kRK == 1) {
    const double d1 = 1.0 / (o->alp * (o->Nx / 2) * u1m), d3 = 1.0 / (o->bet * (o->Nz / 2) * u3m);
    const double dloc = fmin(d1, d3);
    o->dtc = dloc;
    /* fixed_dt > 0: the time step is given (a run at fixed dt; dtc stays the CFL diagnostic) */
    o->dt = o->fixed_dt > 0.0 ? o->fixed_dt : fmin(2.5 * o->dtv, o->CFL * o->dtc);
    o->dti = 1.0 / o->dt;
    o->t = o->t + o->dt;
  }
  /* du += -gRK * N : FOU3D.f90:359-368 */
  const double Cg = -o->gRK[q];
#pragma omp parallel for num_threads(o->nthreads)
  for (int col = 0; col < o->C; col++) {
    for (int j = 1; j <= o->nyu; j++) {
      o->du1[(size_t)col * nu + j] = o->du1[(size_t)col * nu + j] + Cg * o->Nu1[(size_t)col * o->nyu + j - 1];
      o->du3[(size_t)col * nu + j] = o->du3[(size_t)col * nu + j] + Cg * o->Nu3[(size_t)col * o->nyu + j - 1];
    }
    for (int j = 1; j <= o->nyv; j++)
      o->du2[(size_t)col * nv + j] = o->du2[(size_t)col * nv + j] + Cg * o->Nu2[(size_t)col * o->nyv + j - 1];
  }
}

/* LU_build + LU_dec + LUsolU : tridLU_3D.f90:38-94, 163-221, 327-364.  n = rows, s..e = 0..n-1 */
static void tri_solve(int n, const double *a1, const double *a2, const double *a3, cplx *u, const cplx *rhs,
                      double *w2, double *w3) {
  w2[0] = 1.0 / a2[0];
  w3[0] = a3[0];
  for (int j = 1; j < n; j++) {
    w3[j - 1] = w3[j - 1] * w2[j - 1];
    w3[j] = a3[j];
    w2[j] = 1.0 / (a2[j] - a1[j] * w3[j - 1]);
  }
  u[0] = rhs[0] * w2[0];
  for (int j = 1; j < n; j++) u[j] = (rhs[j] - a1[j] * u[j - 1]) * w2[j];
  for (int j = n - 2; j >= 0; j--) u[j] = u[j] - w3[j] * u[j + 1];
}

/* solveU : littleharsh.f90:863-894 */
void orc_solveU(orc_t *o, int kRK) {
  const int q = kRK - 1, nu = o->nu_rows, nv = o->nv_rows;
  const double beta = o->bRK[q] * o->dt / o->Re;
#pragma omp parallel num_threads(o->nthreads)
  {
    const int nmax = nu;
    double *a1 = (double *)malloc(sizeof(double) * nmax * 5);
    double *a2 = a1 + nmax, *a3 = a2 + nmax, *w2 = a3 + nmax, *w3 = w2 + nmax;
    cplx *rhs = (cplx *)malloc(sizeof(cplx) * nmax);
#pragma omp for
    for (int col = 0; col < o->C; col++) {
      const int i = col % o->Kx, k = col / o->Kx;
      const double k2x = o->k2x[i], k2z = o->k2z[k];
      /* u-grid matrix, shared by u1 and u3 */
      a1[0] = 0.0; a2[0] = 1.0; a3[0] = o->gw[0];
      for (int j = 1; j < nu - 1; j++) {
        a1[j] = -beta * D2U(o, 1, j);
        a2[j] = 1.0 - beta * (D2U(o, 2, j) + k2x + k2z);
        a3[j] = -beta * D2U(o, 3, j);
      }
      a1[nu - 1] = o->gw[1]; a2[nu - 1] = 1.0; a3[nu - 1] = 0.0;
      cplx *fields[2] = {o->u1 + (size_t)col * nu, o->u3 + (size_t)col * nu};
      cplx *dus[2] = {o->du1 + (size_t)col * nu, o->du3 + (size_t)col * nu};
      for (int f = 0; f < 2; f++) {
        rhs[0] = 0.0;
        for (int j = 1; j < nu - 1; j++) rhs[j] = fields[f][j] + o->dt * dus[f][j];
        rhs[nu - 1] = 0.0;
        tri_solve(nu, a1, a2, a3, fields[f], rhs, w2, w3);
      }
      /* v-grid */
      a1[0] = 0.0; a2[0] = 1.0; a3[0] = 0.0;
      for (int j = 1; j < nv - 1; j++) {
        a1[j] = -beta * D2V(o, 1, j);
        a2[j] = 1.0 - beta * (D2V(o, 2, j) + k2x + k2z);
        a3[j] = -beta * D2V(o, 3, j);
      }
      a1[nv - 1] = 0.0; a2[nv - 1] = 1.0; a3[nv - 1] = 0.0;
      cplx *u2 = o->u2 + (size_t)col * nv, *d2 = o->du2 + (size_t)col * nv;
      rhs[0] = 0.0;
      for (int j = 1; j < nv - 1; j++) rhs[j] = u2[j] + o->dt * d2[j];
      rhs[nv - 1] = 0.0;
      tri_solve(nv, a1, a2, a3, u2, rhs, w2, w3);
    }
    free(a1); free(rhs);
  }
}

/* meanflow_ctU -> flowrate_corr -> LUsol0 : littleharsh.f90:515-548, 591-607; tridLU_3D.f90:134-160, 299-325 */
void orc_meanflow_ctU(orc_t *o, int kRK) {
  const int q = kRK - 1, nu = o->nu_rows;
  const double beta = o->bRK[q] * o->dt / o->Re;
  double *a1 = (double *)malloc(sizeof(double) * nu * 5);
  double *a2 = a1 + nu, *a3 = a2 + nu, *w2 = a3 + nu, *w3 = w2 + nu;
  cplx *u11 = (cplx *)malloc(sizeof(cplx) * nu), *rhs = (cplx *)malloc(sizeof(cplx) * nu);
  a1[0] = 0.0; a2[0] = 1.0; a3[0] = o->gw[0];
  for (int j = 1; j < nu - 1; j++) {
    a1[j] = -beta * D2U(o, 1, j);
    a2[j] = 1.0 - beta * D2U(o, 2, j);
    a3[j] = -beta * D2U(o, 3, j);
  }
  a1[nu - 1] = o->gw[1]; a2[nu - 1] = 1.0; a3[nu - 1] = 0.0;
  const double bdt = -o->dRK[q] * o->dt;
  rhs[0] = 0.0;
  for (int j = 1; j <= o->nyv + 1; j++) rhs[j] = bdt;
  rhs[nu - 1] = 0.0;
  tri_solve(nu, a1, a2, a3, u11, rhs, w2, w3);
  /* flowrateRe's dummy is u(1:nn+2) while u11 is allocated 0:nn+2 (start.f90:261), so by sequence
   * association its u(j) is u11(j-1) (littleharsh.f90:505-511, 537) */
  double Qx1 = 0.0;
  for (int j = 1; j <= o->nyv + 1; j++) Qx1 = Qx1 + (o->yv[j] - o->yv[j - 1]) * creal(u11[j - 1]);
  const double Qx0 = flowrate_re_part(o, o->u1);
  const double g = (o->QxT - Qx0) / Qx1;
  o->mpgx = o->mpgx + g;
  o->dgx = g;
  for (int j = 0; j < nu; j++) o->u1[j] = o->u1[j] + g * creal(u11[j]);
  free(a1); free(u11); free(rhs);
}

/* solveP + LUsolP : littleharsh.f90:626-680; tridLU_3D.f90:98-132 */
void orc_solveP(orc_t *o, int kRK) {
  const int q = kRK - 1, npr = o->np_rows;
  const double dtirk = o->dti / (o->aRK[q] + o->bRK[q]);
  divergence(o);
#pragma omp parallel for num_threads(o->nthreads)
  for (int col = 0; col < o->C; col++) {
    cplx *x = o->psi + (size_t)col * npr;
    const cplx *dv = o->div + (size_t)col * npr;
    for (int j = 0; j < npr; j++) x[j] = dtirk * dv[j];
    if (col == 0) x[0] = 0.0;
    const double *a = o->DG + (size_t)col * npr * 3;
    x[0] = x[0] * a[1];
    for (int j = 1; j < npr; j++) x[j] = (x[j] - a[3 * j + 0] * x[j - 1]) * a[3 * j + 1];
    for (int j = npr - 2; j >= 0; j--) x[j] = x[j] - a[3 * j + 2] * x[j + 1];
    cplx *p = o->p + (size_t)col * npr;
    for (int j = 0; j < npr; j++) p[j] = p[j] + x[j];
  }
}

/* v_corr : littleharsh.f90:896-993 */
void orc_vcorr(orc_t *o, int kRK) {
  const int q = kRK - 1, nu = o->nu_rows, nv = o->nv_rows, npr = o->np_rows;
  const cplx dtrk = o->dt * (o->aRK[q] + o->bRK[q]);
  const cplx dtrk_u2 = dtrk * 1.0;
#pragma omp parallel for num_threads(o->nthreads)
  for (int col = 0; col < o->C; col++) {
    const int i = col % o->Kx, k = col / o->Kx;
    const cplx kx = dtrk * (o->kx[i] * I), kz = dtrk * (o->kz[k] * I);
    cplx *u1 = o->u1 + (size_t)col * nu, *u3 = o->u3 + (size_t)col * nu, *u2 = o->u2 + (size_t)col * nv;
    const cplx *ps = o->psi + (size_t)col * npr; /* psi(j) at j-1 */
    for (int j = 1; j <= o->nyu; j++) u1[j] = u1[j] - kx * ps[j - 1];
    u1[0] = o->gbc_u1 * u1[1];
    u1[nu - 1] = o->gbc_u1 * u1[nu - 2];
    for (int j = 1; j <= o->nyv; j++) u2[j] = u2[j] - (dtrk_u2 * o->dthdyv[j]) * (ps[j] - ps[j - 1]);
    u2[0] = 0.0;
    u2[nv - 1] = 0.0;
    for (int j = 1; j <= o->nyu; j++) u3[j] = u3[j] - kz * ps[j - 1];
    u3[0] = o->gbc_u3 * u3[1];
    u3[nu - 1] = o->gbc_u3 * u3[nu - 2];
  }
  divergence(o);
}

/* meanflow_ctP : littleharsh.f90:572-589 */
void orc_meanflow_ctP(orc_t *o) { o->Qx = flowrate_re_part(o, o->u1); }

/* main loop : littleharsh.f90:163-233 */
void orc_step(orc_t *o, int nsteps, double *hist) {
  for (int s = 0; s < nsteps; s++) {
    for (int kRK = 1; kRK <= 3; kRK++) {
      orc_rhs0(o, kRK);
      orc_nonlinear(o, kRK);
      orc_solveU(o, kRK);
      if (o->flag_ctpress == 0) orc_meanflow_ctU(o, kRK);
      orc_solveP(o, kRK);
      orc_vcorr(o, kRK);
    }
    if (o->flag_ctpress == 1) orc_meanflow_ctP(o);
    o->iter++;
    if (hist) {
      hist[4 * s + 0] = o->t;
      hist[4 * s + 1] = o->dt;
      hist[4 * s + 2] = o->Qx;
      hist[4 * s + 3] = o->mpgx;
    }
  }
}

void orc_plane_fft(int Mx, int Mz, int nplanes, const double *in, double *out) {
  planeops po;
  planeops_init(&po, Mx, Mz, (2 * Mx) / 3);
  const size_t plsz = (size_t)(Mx + 2) * Mz;
  const int nbuf = Mx > Mz ? Mx : Mz;
  cplx *buf = (cplx *)malloc(sizeof(cplx) * nbuf), *work = (cplx *)malloc(sizeof(cplx) * nbuf);
  for (int s = 0; s < nplanes; s++) {
    double *a = out + (size_t)s * plsz, *b = out + (size_t)(nplanes + s) * plsz;
    memcpy(a, in + (size_t)s * plsz, sizeof(double) * plsz);
    four_to_phys(&po, a, buf, work);
    memcpy(b, a, sizeof(double) * plsz);
    phys_to_four(&po, b, buf, work);
  }
  free(buf); free(work);
  planeops_free(&po);
}
