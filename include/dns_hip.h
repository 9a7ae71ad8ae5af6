/* dns_hip.h -- C-ABI of the MI355X implementation of the channel-flow DNS RK3 time step.
 *
 * The reference has no plugin layer: its hot path is a set of external Fortran subroutines with
 * implicit interfaces called from the main loop (littleharsh.f90:163-233), all state living in the
 * `declaration` module.  This ABI exports one entry point per replaced subroutine plus a fused step;
 * the Fortran interface module (final-single-block-_amd/fortran/dns_hip_mod.f90) binds them with
 * ISO_C_BINDING and wraps them under the reference's subroutine names.
 *
 * Conventions
 *  - Every function returns 0 on success and a non-zero status otherwise; dns_last_error() returns a
 *    message for the calling thread.  The Fortran wrappers turn a non-zero status into `error stop`,
 *    which keeps the reference's fatal `stop` semantics.
 *  - Fields stay resident in GPU memory between calls.  Host arrays cross the ABI only in
 *    dns_upload_fields / dns_download_fields, in the reference's column layout:
 *       complex(8) x(j, column), j fastest, column = (k-1)*(Nx/2+1) + i + 1   (start.f90:963-981)
 *       u1, u3: j = 0..nyu+1 ; u2: j = 0..nyv+1 ; p: j = 1..nyu        (littleharsh.f90:98-112)
 *    passed as interleaved (re, im) doubles.
 *  - One handle = one channel = one HIP stream.  Calls on one handle are not reentrant (the reference's
 *    FFT work arrays make it non-reentrant too); different handles are independent.
 */
#ifndef DNS_HIP_H
#define DNS_HIP_H
#ifdef __cplusplus
extern "C" {
#endif

#define DNS_OK 0
#define DNS_ERR_ARG 1     /* bad argument or unsupported size */
#define DNS_ERR_HIP 2     /* HIP runtime error */
#define DNS_ERR_STATE 3   /* call out of order (e.g. no fields uploaded) */
#define DNS_ERR_NONFINITE 4 /* a NaN/Inf reached the CFL maxima */

#define DNS_PREC_FP64 0   /* everything in FP64 */
#define DNS_PREC_MIXED 1  /* FFTs, plane intermediates and products in FP32; solves, RK3 update in FP64 */

typedef struct dns_handle dns_handle;

/* Run parameters: the live keys of the reference's input.in (start.f90:86-252). */
typedef struct dns_params {
  int Mx, Mz, ny;         /* Ngalx1, Ngalz1 (physical points, 3/2-rule), ny (input.in:15-19) */
  int ppp;                /* pp, polynomial exponent of the y grid (input.in:21) */
  int flag_ctpress;       /* flpress: 1 constant pressure gradient, 0 constant flow rate (input.in:36) */
  int precision;          /* DNS_PREC_FP64 or DNS_PREC_MIXED */
  double Re, alp, bet;    /* Re, alp, bet (input.in:10-14) */
  double CFL;             /* CFL (input.in:34) */
  double mpgx;            /* mpgx (input.in:38) */
  double dyq;             /* dyq (input.in:20) */
  double fixed_dt;        /* > 0: use this dt instead of the reference's CFL rule (FOU3D.f90:290-303) */
  int plane_batch;        /* planes per plane-stage launch group (u- and v-planes counted, rounded up to
                             an even number); 0 = all planes */
  int members;            /* ensemble members of these sizes (0 or 1: a single channel); each member has
                             its own fields, Re, mpgx, dt and state, and every launch covers all members */
  int column_path;        /* column stage: 0 one kernel per substep (held on chip; used when
                             8 <= ny-1 <= 512, else the sweeps; with flag_ctpress = 0 a one-workgroup
                             kernel applies meanflow_ctU after it), 1 three sweeps */
  int reserved[5];
} dns_params;

const char *dns_last_error(void);

/* start.f90:53-663 for one rank: grid, wavenumbers, RK3 coefficients, dtv, weights; device buffers.
 * Replaces `start` (start.f90:53); all set-up is host code, no kernel runs here. */
int dns_create(const dns_params *prm, dns_handle **out);
int dns_destroy(dns_handle *h);

/* Sizes: out[0..9] = Nx, Nz, Kx, ncol, nyu, nyv, Mx, Mz, ny, device bytes allocated (MiB) */
int dns_sizes(const dns_handle *h, long long *out);

/* Host-side set-up arrays (ygrid, start.f90:665-806), no GPU needed:
 * yu[ny+1], yv[ny], dthdyu[ny+1], dthdyv[ny], dyu2i[3*(ny+1)], dyv2i[3*ny] (3 coefficients per j),
 * w[0..5] = dtv, gridweighting(1:2), gridweighting_bc_u1, gridweighting_interp(1:2). Any may be NULL. */
int dns_setup_arrays(const dns_handle *h, double *yu, double *yv, double *dthdyu, double *dthdyv,
                     double *dyu2i, double *dyv2i, double *w);

/* ygrid (start.f90:665-806) without a handle or a GPU: yu[ny+1], yv[ny], dthdyu[ny+1], dthdyv[ny]. */
int dns_ygrid(int ny, double dyq, int ppp, double *yu, double *yv, double *dthdyu, double *dthdyv);

/* getini (start.f90:1221-1316) after the fields are in memory: uploads u1,u2,u3,p (reference column
 * layout, host memory), sets t and iter, Qx, and QxT (QxT <= 0: QxT = flow rate of u1, as getini).
 * The previous-substep nonlinear terms are cleared (they are multiplied by cRK(1) = 0 at kRK = 1).
 * dns_upload_fields loads the same fields into every ensemble member; dns_upload_member into member m. */
int dns_upload_fields(dns_handle *h, const double *u1, const double *u2, const double *u3, const double *p,
                      double t, int iter, double QxT);
int dns_upload_member(dns_handle *h, int m, const double *u1, const double *u2, const double *u3,
                      const double *p, double t, int iter, double QxT);
/* Copies the current fields back into host arrays in the reference layout (record_out's input):
 * member 0, or member m. */
int dns_download_fields(dns_handle *h, double *u1, double *u2, double *u3, double *p);
int dns_download_member(dns_handle *h, int m, double *u1, double *u2, double *u3, double *p);
/* Ensemble member m's Reynolds number (and with it dtv, start.f90:476) and mean pressure gradient;
 * the defaults are dns_params' Re and mpgx.  Takes effect at the next upload or step.  No counterpart
 * in the reference (one channel per run); it serves the ensemble form of the step. */
int dns_set_member_params(dns_handle *h, int m, double Re, double mpgx);
/* The following transfers act on ensemble member 0.
 * Nonlinear terms Nu1 (j=1..nyu), Nu2 (j=1..nyv), Nu3 (j=1..nyu) in the reference layout. */
int dns_download_nonlinear(dns_handle *h, double *Nu1, double *Nu2, double *Nu3);
int dns_upload_nonlinear(dns_handle *h, const double *Nu1, const double *Nu2, const double *Nu3);
/* Right-hand sides du1 (j=1..nyu), du2 (j=1..nyv), du3 (j=1..nyu) left by dns_rhs0 / dns_nonlinear. */
int dns_download_rhs(dns_handle *h, double *du1, double *du2, double *du3);
/* y-flux products uv_f (j=0..nyv+1), vv_c (j=0..nyu+1), wv_f (j=0..nyv+1) of the last dns_nonlinear. */
int dns_download_flux(dns_handle *h, double *uv_f, double *vv_c, double *wv_f);
/* psi and div (j = 1..nyu) of the last dns_solveP / dns_vcorr. */
int dns_download_pressure_correction(dns_handle *h, double *psi, double *div);

/* s[0..9] = t, dt, dti, mpgx, Qx, dgx, dtc, QxT, iter, max|div| (last dns_vcorr or fused step) of
 * member 0, or member m. */
int dns_get_scalars(dns_handle *h, double *s);
int dns_get_member_scalars(dns_handle *h, int m, double *s);
/* Sets dt (and dti) for a call sequence that starts at kRK > 1 (the reference sets dt at kRK = 1). */
int dns_set_dt(dns_handle *h, double dt);

/* ---- one entry point per replaced subroutine (staged path, same arithmetic order) ---- */
int dns_rhs0(dns_handle *h, int kRK);         /* RHS0_u1/u2/u3  littleharsh.f90:682-856 */
int dns_nonlinear(dns_handle *h, int kRK);    /* nonlinear      FOU3D.f90:51-371 (+ dtc_calc at kRK=1) */
int dns_solveU(dns_handle *h, int kRK);       /* solveU x3      littleharsh.f90:863-894, LUsolU tridLU_3D.f90:38 */
int dns_meanflow_ctU(dns_handle *h, int kRK); /* meanflow_ctU   littleharsh.f90:591-607 (flag_ctpress==0) */
int dns_solveP(dns_handle *h, int kRK);       /* solveP         littleharsh.f90:626-680, LUsolP tridLU_3D.f90:98 */
int dns_vcorr(dns_handle *h, int kRK);        /* v_corr         littleharsh.f90:896-993 */
int dns_meanflow_ctP(dns_handle *h);          /* meanflow_ctP   littleharsh.f90:572-589 (flag_ctpress==1) */

/* ---- fused path ---- */
/* One RK3 substep: RHS0 + nonlinear + solveU + meanflow_ctU + solveP + v_corr fused into the plane
 * stage (3 kernels) and the column stage (one kernel, or 3 sweeps: dns_set_column_path).
 * littleharsh.f90:165-222 */
int dns_rk3_substep(dns_handle *h, int kRK);
/* nsteps full RK3 steps (littleharsh.f90:163-233), including meanflow_ctP and iter += 1. */
int dns_rk3_step(dns_handle *h, int nsteps);
/* Records one RK3 step (3 substeps) into a hipGraph on the handle's stream and replays it nsteps times. */
int dns_rk3_step_graph(dns_handle *h, int nsteps);

/* ---- lower-level entry points, on device memory owned by the caller ---- */
/* four_to_phys_du then phys_to_four_du (FOU3D.f90:681-698, 1654-1672) on nplanes reference-format planes
 * of (Mx+2) x Mz doubles (x fastest), in place on d_planes: in = spectral, out[0..n) = physical,
 * and d_back (n planes) = back to spectral.  Planes use the handle's Mx, Mz, FP64. */
int dns_plane_fft_roundtrip(dns_handle *h, const double *d_in, double *d_phys, double *d_back, int nplanes);
/* modes_to_planes_UVP (FOU3D.f90:718-814) on device arrays: column field [ncol][nrows] (reference
 * layout, complex) -> nrows reference planes (Mx+2) x Mz with the 3/2 zero padding. */
int dns_modes_to_planes(dns_handle *h, const double *d_cols, int nrows, double *d_planes);
/* planes_to_modes_UVP (FOU3D.f90:1694-1824): the inverse gather. */
int dns_planes_to_modes(dns_handle *h, const double *d_planes, int nrows, double *d_cols);

/* One record of the history file hist_NNNNxNNNNxNNNN.dat that record_out appends (FOU3D.f90:2169-2200,
 * file opened at start.f90:273-284), for member m, at the current state (a step boundary):
 *   rec[13] = flag_ctpress, iter, t, dtv, dtc, dt, err, Qx, QxT, mpgx, dgx, Umax, Uslip
 * err = max|div| of the last v_corr (error, littleharsh.f90:370-409); Qx = flowrateIm of mode (0,0)
 * (littleharsh.f90:473-492); Umax = maxvel (littleharsh.f90:550-570); Uslip as FOU3D.f90:2171.
 * The file record holds QxT and dgx only when flag_ctpress = 0 (FOU3D.f90:2196-2200). */
int dns_history_record(dns_handle *h, int m, double *rec);

/* Statistics and spectra (stats.f90:1-97, spectra.f90:1-60, omega_x of FOU3D.f90:307-338), accumulated
 * on the device at the current state (the reference accumulates them at kRK = 1 of every nstat-th step,
 * from the fields at the start of that step: call dns_accumulate_stats before dns_rk3_step).
 * dns_get_stats: istat (number of accumulations), stats[13*(nyu+2)] = Um, U2m, Wm, W2m, UVm, UWm (u
 * planes 0..nyu+1), Vm, V2m, VWm, wxm, wx2m (v planes 0..nyu+1), Pm, P2m (p planes 1..nyu+1 at index j,
 * index 0 unused), each array nyu+2 long; spectra (may be NULL) = spU [ncol][nyu+2], spV [ncol][nyv+2],
 * spW, spUV [ncol][nyu+2], spP [ncol][nyu] in the reference's sp(j, column) layout.  Pm, P2m sum the
 * spectral p planes of ppPL (stats.f90:44-51): the current p until dns_refresh_stats_pressure is first
 * called, the copy it took from then on.  The physical-plane sums are formed with the exact discrete sum
 * rules of the reference's transforms (agreement with the reference to rounding, tested on its own
 * output). */
int dns_accumulate_stats(dns_handle *h);
int dns_get_stats(dns_handle *h, int m, double *stats, double *spectra, int *istat);
int dns_reset_stats(dns_handle *h);
/* ppPL = 0; modes_to_planes_UVP(ppPL, p) (FOU3D.f90:149-152; littleharsh.f90:439-443): copies every
 * member's current p into the pressure that Pm, P2m sum until the next call.  The reference fills ppPL from
 * the initial field file (start.f90:1369-1387) and refreshes it at each write step before that step's
 * statistics, so a statistics step that is not a write step sums the pressure of the last write step;
 * a driver reproduces that with one call after the upload and one at each write step before
 * dns_accumulate_stats. */
int dns_refresh_stats_pressure(dns_handle *h);

/* four_to_phys post-processing: the per-plane resize + cft + rft of the reference's post-processors
 * (four_to_phys.f90:60-81 and four_to_phys_N.f90:60-81, resize at four_to_phys.f90:318-337), for a batch of
 * planes on the GPU; needs no handle.  in: nplanes spectral planes buffSR1(Nx+2, Nz) as in a field file
 * (host memory); out: nplanes physical planes buffSR2(Mx+2, Mz).  Mx >= Nx, Mz >= Nz (Mz and Mx/2 in
 * 8..1152, of the form 2^a 3^b); the kz = -Nz/2 row is dropped when Mz > Nz (four_to_phys, Ngal = 3N/2)
 * and kept when Mz == Nz (four_to_phys_N, Ngal = N).  ms (may be NULL): device time of the transform
 * kernels in ms. */
int dns_four_to_phys(int Nx, int Nz, int Mx, int Mz, int nplanes, const double *in, double *out, double *ms);
/* dns_four_to_phys on device memory, with no device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  d_in: nplanes dense spectral planes (Nx+2) x Nz; d_out: nplanes dense physical
 * planes (Mx+2) x Mz; both device memory, 16-byte aligned; stream: a hipStream_t (NULL: the null stream);
 * returns after the work on that stream has finished.  The intermediate and the twiddle tables live in the
 * library's static workspace of 16 MiB per GPU (held until the process ends); planes go through it in
 * batches, and sizes whose single plane does not fit are rejected with DNS_ERR_ARG.  Every call on one GPU
 * shares that GPU's workspace: calls that overlap on one GPU (other threads, other streams) race.  Results
 * equal dns_four_to_phys bit for bit. */
int dns_four_to_phys_noalloc(int Nx, int Nz, int Mx, int Mz, int nplanes, const double *d_in, double *d_out,
                             void *stream);

/* Column stage of the fused path: path 0 = one kernel per substep (rows split into segments, tridiagonal
 * systems by the partition method, everything read and written once), 1 = three sweeps (thread per
 * column, reference elimination order).  Same equations as RHS0 + nonlinear (column part) + solveU +
 * solveP + v_corr (littleharsh.f90:163-233); results agree to rounding.  set: requested path (path 0
 * falls back to the sweeps where it does not apply); get: the path in use. */
int dns_set_column_path(dns_handle *h, int path);
int dns_get_column_path(dns_handle *h, int *path);

/* ---- ensemble of several sizes ----
 * No counterpart in the reference, which runs one channel per job; every member runs the reference's
 * single-channel RK3 loop (littleharsh.f90:163-233).  An ensemble joins 1..4 handles (one per size, each
 * with its same-size members, dns_params.members) so that all their members advance with one grouped launch
 * per kernel class per RK3 substep (z-inverse, x-pass u/v, z-forward u/v, column stage, and the small CFL
 * and reset kernels), on the ensemble's own stream.  Sizes: Mx and Mz each 48, 96 or 192, 9 <= ny <= 161,
 * one precision for all, no plane batches; others are rejected with DNS_ERR_ARG.  The handles stay owned
 * by the caller and must outlive the ensemble; their uploads, downloads and scalars work as before (work on
 * a handle's stream is ordered before and after the ensemble's steps).  A member's result equals the same
 * channel stepped alone, bit for bit. */
typedef struct dns_ensemble dns_ensemble;
int dns_ensemble_create(dns_handle *const *handles, int n, dns_ensemble **out);
int dns_ensemble_destroy(dns_ensemble *e);
/* nsteps RK3 steps of every member (iter += 1 per step, as dns_rk3_step) */
int dns_ensemble_step(dns_ensemble *e, int nsteps);
/* One RK3 step captured once in a hipGraph on the ensemble's stream and replayed nsteps times. */
int dns_ensemble_step_graph(dns_ensemble *e, int nsteps);
/* Waits for the ensemble's stream.  After persistent launches it also reports (DNS_ERR_HIP) a task that
 * waited too long for the tasks it depends on, in which case those steps' fields are not valid. */
int dns_ensemble_sync(dns_ensemble *e);
/* How a substep is launched.  mode 0 (the default): the grouped launches above.  mode 1: one persistent
 * launch per RK3 substep.  Its workgroups take tasks from a queue in device memory (a task is a run of
 * blocks of one kernel class for one member; the classes in the order of the grouped launches) and a task
 * starts once the tasks of its member that it depends on have finished (per-member completion counters),
 * so different members' classes overlap.  units[c] (c = 0..9: reset, ghost-plane CFL, z-inverse, x-pass u,
 * x-pass v, CFL reduction, z-forward u, z-forward v, column stage, mean-flow correction; NULL or entries
 * <= 0: defaults) sets how many units (two plane workgroups, or one column-stage workgroup) a task covers.
 * The same bodies run, so the results are the same bits as mode 0.  Needs ny - 1 even in every group
 * (DNS_ERR_ARG otherwise).  Drops a captured step graph.  No counterpart in the reference. */
int dns_ensemble_set_launch(dns_ensemble *e, int mode, const int *units);
/* mode in use; with mode 1 also the workgroups of a launch (nwg) and the tasks of substeps 1..3 (ntask[3]);
 * nwg and ntask may be NULL. */
int dns_ensemble_get_launch(dns_ensemble *e, int *mode, int *nwg, int *ntask);
/* Event timing of dns_ensemble_step as dns_set_timing/dns_get_timing: t[0] resets + z-inverse, t[1] x-pass,
 * t[2] z-forward, t[3] column stage (+ CFL reduction, mean-flow correction), t[6] whole substep (ms). */
int dns_ensemble_set_timing(dns_ensemble *e, int on);
int dns_ensemble_get_timing(dns_ensemble *e, double *t, int *n);
/* Tile shapes of the grouped launches (mode 0); no counterpart in the reference (its FFT and tridiagonal
 * work has no tiles).  var[c], c = 0 z-inverse, 1/2 x-pass u/v-planes, 3/4 z-forward u/v-planes: bodies of
 * 64 << var[c] threads (var 0..2), wg[c] threads per workgroup (128 or 256; a workgroup runs wg / (64 <<
 * var) bodies side by side); var[5]: column stage variant (0..2: a columns-per-workgroup count or form of
 * each group's row bucket with the same segments).  One shape per class for every group of the launch.
 * Defaults: plane variant 1 in 128-thread workgroups, column variant 0.  Every shape gives the same bits.
 * set rejects (DNS_ERR_ARG) a body wider than its workgroup, and several z-inverse bodies per workgroup
 * where a group's ny - 1 is odd; it drops a captured step graph. */
#define DNS_ENS_TUNE_MS 28
int dns_ensemble_set_tiles(dns_ensemble *e, const int *var, const int *wg);
int dns_ensemble_get_tiles(dns_ensemble *e, int *var, int *wg);
/* The ensemble's autotuner: times every shape of every grouped class on the whole launch (all groups and
 * members; plane classes with kRK = 2 arguments) and keeps the fastest per class; the fields the column
 * stage overwrites are saved and put back.  var[6], wg[5] (both may be NULL) receive the choice as
 * dns_ensemble_get_tiles; ms[DNS_ENS_TUNE_MS] (may be NULL) the time per launch (ms, -1 where not
 * available) of plane class c in shape k at ms[5c + k], k = 0..4 for (wg, var) = (128, 0), (128, 1),
 * (256, 0), (256, 1), (256, 2), and of column variant v at ms[25 + v]. */
int dns_ensemble_autotune(dns_ensemble *e, int iters, int *var, int *wg, double *ms);

/* Waits for all work on the handle's stream. */
int dns_sync(dns_handle *h);
/* Average device duration (ms) per substep of each kernel, from HIP events on the handle's stream,
 * over the dns_rk3_step calls made since dns_set_timing(h, 1): t[0] z-inverse, t[1] x-pass,
 * t[2] z-forward, t[3] column sweep up, t[4] column sweep down, t[5] final sweep (+ mean-flow
 * correction), t[6] whole substep; with the one-pass column stage t[3] holds it and t[4], t[5] are 0.  With plane_batch below the plane count t[0] holds the whole plane
 * stage and t[1], t[2] are 0.  n = number of substeps averaged. */
int dns_set_timing(dns_handle *h, int on);
int dns_get_timing(dns_handle *h, double *t, int *n);

/* Kernel variants, index c: 0 z-inverse, 1 x-pass of u-planes, 2 x-pass of v-planes, 3 z-forward of
 * u-planes, 4 z-forward of v-planes (tile shapes), 5 column sweep up (thread per column or one
 * wavefront per solve), 6 column sweep down, 7 final sweep, 8 one-pass column stage (columns per
 * workgroup at the bucket's segment count and length).  All variants of a kernel give
 * bit-identical results.  No counterpart in the reference (one code path per subroutine); they stand
 * in for the loop orders of four_to_phys_du / phys_to_four_du (FOU3D.f90:857-1231) and of the solves
 * (tridLU_3D.f90:38-221).
 * dns_get_plane_variants: var[DNS_NKERN] current choice, count[DNS_NKERN] (may be NULL) variants available.
 * dns_set_plane_variants: var[DNS_NKERN]; status 1 for a variant that does not exist. */
#define DNS_NKERN 9
int dns_get_plane_variants(dns_handle *h, int *var, int *count);
int dns_set_plane_variants(dns_handle *h, const int *var);
/* Autotuner: times each variant of each kernel that has several (iters launches; plane kernels on the
 * first plane batch) and keeps the fastest; another variant replaces the first only when it is at least
 * 2 % faster (closer timings are within the run-to-run spread).  chosen[DNS_NKERN] (may be NULL) receives the choice,
 * ms[3*DNS_NKERN] (may be NULL) the time per launch of variant v of kernel c at ms[3c+v] (-1 where not
 * timed).  Rewrites scratch only (the fields are saved and restored): call it before a step or between
 * steps.  Drops a captured graph (the next dns_rk3_step_graph captures the new choice). */
int dns_autotune(dns_handle *h, int iters, int *chosen, double *ms);

#ifdef __cplusplus
}
#endif
#endif
