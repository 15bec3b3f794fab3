/*
 * ifos2d_opt.h -- C ABI of the model-update side of an FWI iteration on the
 * device: reductions, PCG direction, step-length parabola, bounded model
 * update and L-BFGS.  All array arguments are device pointers (interior
 * [NY][NX] arrays, x fastest); scalars come back to the host.
 *
 * Sums run in FP64 as deterministic two-pass reductions (the reference sums
 * in FP32); the element-wise expressions are the reference's FP32 ones.
 * Errors: 0 on success, negative on failure, message in ifos_last_error().
 */
#ifndef IFOS2D_OPT_H
#define IFOS2D_OPT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ifos_opt ifos_opt;

/* Parameters of calc_mat_change_test (src/calc_mat_change_test.c:26 and the
 * globals it reads): the integers, then the floats, in this order. */
typedef struct ifos_update_params {
    int iter, INV_VP_ITER, INV_VS_ITER, INV_RHO_ITER, S, GRAD_METHOD, LBFGS_STEP_LENGTH, LBFGS_iter_start;
    float eps_scale, VPUPPERLIM, VPLOWERLIM, VSUPPERLIM, VSLOWERLIM, RHOUPPERLIM, RHOLOWERLIM, VP_VS_RATIO;
    float S_VP, S_VS, S_RHO;
} ifos_update_params;

/* Reduction scratch and a stream on one device. */
int  ifos_opt_create(int device, ifos_opt **out);
void ifos_opt_destroy(ifos_opt *o);

/* FP64 sum of a[i]*b[i] (replaces the FP32 sums of src/PCG.c:170-174 and
 * src/LBFGS.c:247-250, 279-282, 308-310). */
int ifos_opt_dot(ifos_opt *o, const float *a, const float *b, long n, double *out);
/* Largest a[i] (absolute = 0) or |a[i]| (absolute = 1), exact
 * (src/calc_mat_change_test.c:85-150). */
int ifos_opt_max(ifos_opt *o, const float *a, long n, int absolute, float *out);

/* replaces PCG() src/PCG.c:27 for one parameter (tapers, filters and the
 * jacobian files aside): dir = C*grad, conjugated with the Polak-Ribiere
 * beta (FP64 sums, clipped at 0) against p_last (iteration 2) or c_last
 * (later iterations) exactly when the reference conjugates.  p_last and
 * c_last hold the previous preconditioned gradient and direction and are
 * updated in place as the reference's _p.old / _c.old files are. */
int ifos_opt_pcg(ifos_opt *o, long n, const float *grad, float C, float *p_last, float *c_last, int iter,
                 int pcg_iter_start, float *dir, double *beta);

/* No counterpart in the reference.  ifos_opt_pcg for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernels and results, with no
 * ifos_opt context.  It runs on the current device's null stream and keeps
 * its FP64 reduction partials in a static device array of 1032 doubles
 * (8256 bytes) that each GPU holds until the process ends; it calls no
 * allocating routine.  n is limited to 2^31 - 1 cells (larger n: error -2).
 * Calls that overlap on one GPU share that array and race; serialise them. */
int ifos_opt_pcg_noalloc(long n, const float *grad, float C, float *p_last, float *c_last, int iter,
                         int pcg_iter_start, float *dir, double *beta);

/* replaces calc_opt_step() src/calc_opt_step.c:29 (host arrays: the
 * reference's L2t[1..3], epst1[1..3]): the parabola through the three
 * points, solved in FP64 on the device, with the same clamps. */
int ifos_opt_step_length(ifos_opt *o, const float *L2t, const float *epst, float *opteps);

/* replaces calc_mat_change_test() src/calc_mat_change_test.c:26,
 * PARAMETERIZATION 1: models vp, vs, rho minus the scaled directions,
 * bounded as there; vp0/vs0/rho0 are the starting models of the S == 1
 * limits.  eps[3] (host) receives EPSILON, EPSILON_u, EPSILON_rho. */
int ifos_opt_model_update(ifos_opt *o, long n, const ifos_update_params *u, const float *vp, const float *vs,
                          const float *rho, const float *g_vp, const float *g_vs, const float *g_rho,
                          const float *vp0, const float *vs0, const float *rho0, float *vp1, float *vs1,
                          float *rho1, float *eps);

/* L-BFGS memory: s_hist, y_hist [N_LBFGS][npar*ncell] (device), rho_hist
 * [N_LBFGS] (host, FP64).  Parameters in the reference's order vs, rho, vp.
 * replaces the s_LBFGS store of src/calc_mat_change_test.c:297-316:
 * s_w = (m1 - m0) / avg per parameter, w = iter mod N (N for 0). */
int ifos_opt_lbfgs_store_s(ifos_opt *o, long ncell, int npar, const float *const *m1, const float *const *m0,
                           const float *avg, float *s_hist, int N_LBFGS, int iter);
/* replaces lbfgs() src/LBFGS.c:28: grad[npar] (device, in/out) become the
 * L-BFGS direction; work holds 2*npar*ncell floats. */
int ifos_opt_lbfgs(ifos_opt *o, long ncell, int npar, float *const *grad, const float *avg, float *s_hist,
                   float *y_hist, double *rho_hist, float *work, int N_LBFGS, int iteration, int lbfgs_iter_start);

/* ---- gradient post-processing ------------------------------------------ */

/* Whole-gradient tapers of taper_grad (src/taper_grad.c:28), applied by
 * the reference inside PCG before the scaling (src/PCG.c:64-74):
 *   IFOS_TAPER_VERTICAL   sws 1: Gaussian flanks between columns GRADT1 and
 *                         GRADT4, zero at the sources and receivers;
 *   IFOS_TAPER_HORIZONTAL sws 2: zero above row GRADT1, Gaussian flank to
 *                         row GRADT2;
 *   IFOS_TAPER_SOURCES    sws 3: erf (SRTSHAPE 1) or log (2) taper of radius
 *                         SRTRADIUS around every source, zero in the
 *                         (2 FILTSIZE+1)^2 box around it.
 * The per-shot form of the source taper (taper_grad_shot sws 1) is
 * ifos_engine_set_shot_taper.  A taper read from a file (sws 4-6) is a
 * coefficient grid as ifos_model_read returns it. */
enum { IFOS_TAPER_VERTICAL = 1, IFOS_TAPER_HORIZONTAL = 2, IFOS_TAPER_SOURCES = 3 };
typedef struct ifos_taper {
    int kind;
    int GRADT1, GRADT2, GRADT3, GRADT4;
    float DH, SRTRADIUS;
    int SRTSHAPE, FILTSIZE;
    int nsrc;
    const int *srcpos;   /* [nsrc][2] grid indices iround(x/DH), iround(y/DH) */
    int ntr;
    const int *recpos;   /* [ntr][2] grid indices */
} ifos_taper;
/* Host: the taper's coefficients coef [NY][NX] and (zero != NULL) the
 * cells it sets to zero afterwards (zero[e] = 1), including the cells the
 * reference's box reaches through its scratch matrix's row wrap when a box
 * extends two or more columns left of the grid. */
int ifos_taper_coeff(const ifos_taper *t, int NX, int NY, float *coef, unsigned char *zero);
/* Device: grad[e] = zero[e] ? 0 : grad[e] * coef[e] for n cells; coef and
 * zero (may be NULL) are host arrays. */
int ifos_opt_taper(ifos_opt *o, long n, const float *coef, const unsigned char *zero, float *grad);
/* Device, in place: smooth() (src/smooth.c:24) with filter size filtsize
 * (FILT_SIZE_GRAD or FILT_SIZE; an even size grows by one away from zero;
 * 0 does nothing), edges extended by replication.  The value is what
 * median2d (src/median2D.c:3) returns: its sort covers the zero at t[0] and
 * the window without its last element, and it picks element k/2+1 of k.
 * *side receives the window side used.  Window sides up to 113. */
int ifos_opt_smooth(ifos_opt *o, int NX, int NY, int filtsize, float *grad, int *side);
/* Device, in place: spat_filt() (src/spat_filt.c:27, SPATFILTER, applied in
 * PCG after the tapers, src/PCG.c:101-105) for columns S+1..NX-S and rows
 * SPAT_FILT_1..NY-S, S = SPAT_FILT_SIZE (the caller reduces it by the
 * iteration when SPAT_FILT_ITER is set; S < 1 does nothing; S <= 64).  The
 * reference's weights are exp((-1/2) ...) with integer (-1/2), so the filter
 * is a mean over the 2S+1 cells of the row; same values. */
int ifos_opt_spat_filt(ifos_opt *o, int NX, int NY, int SPAT_FILT_SIZE, int SPAT_FILT_1, float *grad);

/* Source taper table helpers (host): table side n and entries m[n][n]
 * (returns n; fills m when cap >= n*n), the table cell that sits on the
 * source (1-based, ceil(n/2)), and one shot's normalisation lo, hi. */
int  ifos_srctaper_table(float DH, float SRTRADIUS, int SRTSHAPE, float *m, int cap);
int  ifos_srctaper_centre(int n);
void ifos_srctaper_norm(int NX, int NY, int n, const float *m, int is, int js, float *lo, float *hi);

#ifdef __cplusplus
}
#endif
#endif
