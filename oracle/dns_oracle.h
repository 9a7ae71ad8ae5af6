/* ORACLE TEST INFRASTRUCTURE -- never part of the product.
 *
 * FP64 CPU restatement of the reference's per-timestep hot path (littleharsh.f90, FOU3D.f90,
 * tridLU_3D.f90, rft_buff.f/cft_buff.f conventions, start.f90 set-up) for one rank.  Used only by
 * tests/, the smoke check and the CPU-baseline leg of bench.py, as the checker of the HIP path.
 *
 * Field layout follows the reference exactly: column fields are x(j, column) with j fastest,
 *   u1, u3 : j = 0 .. nyu+1      (ny+1 rows, ghost rows at both ends)
 *   u2     : j = 0 .. nyv+1      (ny rows, walls at both ends)
 *   p      : j = 1 .. nyu        (ny-1 rows)
 *   Nu1/Nu3: j = 1 .. nyu ; Nu2 : j = 1 .. nyv
 * column = k*Kx + i (k outer over the reference's z-mode order, i = 0..Nx/2 inner).
 * Complex values are interleaved (re, im) doubles.
 */
#ifndef DNS_ORACLE_H
#define DNS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

orc_t *orc_create(int Mx, int Mz, int ny, double Re, double alp, double bet, double CFL,
                  double mpgx, double dyq, int ppp, int flag_ctpress);
void orc_destroy(orc_t *o);
/* out[0..7] = Nx, Nz, Kx, C, nyu, nyv, nthreads, Mx */
void orc_sizes(const orc_t *o, int *out);
void orc_set_threads(orc_t *o, int n);

/* State in reference layout; QxT <= 0 keeps the reference's getini rule (QxT = flow rate of u1). */
void orc_set_state(orc_t *o, const double *u1, const double *u2, const double *u3, const double *p,
                   double t, int iter, double QxT);
void orc_get_state(const orc_t *o, double *u1, double *u2, double *u3, double *p);
void orc_set_nu(orc_t *o, const double *Nu1, const double *Nu2, const double *Nu3);
void orc_get_nu(const orc_t *o, double *Nu1, double *Nu2, double *Nu3);
void orc_get_du(const orc_t *o, double *du1, double *du2, double *du3);  /* interior rows only */
void orc_get_flux(const orc_t *o, double *uv_f, double *vv_c, double *wv_f); /* full v/u/v grids */
void orc_get_psi(const orc_t *o, double *psi, double *div);
/* sc[0..7] = t, dt, dti, mpgx, Qx, dgx, dtc, QxT */
void orc_get_scalars(const orc_t *o, double *sc);
void orc_set_dt(orc_t *o, double dt);
/* dt > 0: every RK3 step takes this dt at kRK = 1 instead of the CFL rule of FOU3D.f90:290-303 (a run at
 * fixed dt); t advances by dt, dtc is still the CFL diagnostic.  0 restores the CFL rule. */
void orc_set_fixed_dt(orc_t *o, double dt);

/* Set-up arrays: yu(ny+1) yv(ny) dthdyu(ny+1) dthdyv(ny) dyu2i(3*(ny+1)) dyv2i(3*ny), DG(3*(ny-1)*C),
 * w[0..5] = dtv, gw1, gw2, gbc, gi1, gi2 */
void orc_get_setup(const orc_t *o, double *yu, double *yv, double *dthdyu, double *dthdyv,
                   double *dyu2i, double *dyv2i, double *DG, double *w);

/* Stages of one RK3 substep, in the order of littleharsh.f90:169-220 */
void orc_rhs0(orc_t *o, int kRK);
void orc_nonlinear(orc_t *o, int kRK);
void orc_solveU(orc_t *o, int kRK);
void orc_meanflow_ctU(orc_t *o, int kRK);
void orc_solveP(orc_t *o, int kRK);
void orc_vcorr(orc_t *o, int kRK);
void orc_meanflow_ctP(orc_t *o);
/* nsteps full RK3 steps (littleharsh.f90:163-233); hist (may be NULL) gets 4 doubles per step:
 * t, dt, Qx, mpgx */
void orc_step(orc_t *o, int nsteps, double *hist);

/* four_to_phys_du then phys_to_four_du (FOU3D.f90:681-698, 1654-1672) on nplanes planes of
 * (Mx+2) x Mz doubles; out holds 2*nplanes planes (physical, then back to Fourier). */
void orc_plane_fft(int Mx, int Mz, int nplanes, const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif
