/*
 * ifos_oracle.h -- CPU restatement of IFOS2D's P-SV forward solver and
 * adjoint-state gradient.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; it is never linked into, loaded by or called from the product
 * library (ifos_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Every function names the reference file:line it restates.  The arithmetic
 * (operand order, float->double promotions of `2.0*` literals, division vs.
 * multiplication) follows the reference expression by expression, so that in
 * the default single-precision build the results are bit-identical to the
 * reference compiled with gcc -O3 -std=c99 on x86-64 (no FMA contraction).
 * Built with -DORC_FP64 every field is double ("the same scheme in FP64"),
 * which is the error yardstick for reduced-precision snapshot storage.
 *
 * Parity status of each part (how it is pinned) is listed in the header of
 * ifos_oracle.c.
 */
#ifndef IFOS_ORACLE_H
#define IFOS_ORACLE_H

#ifdef ORC_FP64
typedef double real;
#else
typedef float real;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Scalar parameters of one run; names follow the reference globals
 * (src/globvar.h:11-150). */
typedef struct orc_params {
    int NX, NY;            /* interior grid size (x fastest)              */
    int FDORDER;           /* 2-12 (10,12: see ifos_oracle.c header)      */
    int MAXRELERROR;       /* 0..4 row of the Holberg table               */
    int FW;                /* CPML width in cells                          */
    int FREE_SURF;         /* 1: stress-imaging free surface at j=1        */
    int BOUNDARY;          /* 1: no CPML on the left/right edges           */
    int L;                 /* relaxation mechanisms, 0 = elastic           */
    int VELOCITY;          /* 0: displacement misfit, 1: velocity          */
    int ADJOINT_TYPE;      /* 1: vx+vy, 2: vy, 3: vx                       */
    int PARAMETERIZATION;  /* 1: vp,vs,rho (the only one supported)        */
    int DTINV;             /* snapshot / correlation stride                */
    int NT;                /* number of time steps                         */
    int SOURCE_SHAPE;      /* 1 Ricker, 2 fumue, 4 sin^3, 5 d/dt Gauss, 6 spike, 8 int sin^3 */
    int SOURCE_TYPE;       /* default type for sources without column 8    */
    int TRKILL;            /* apply kill mask in misfit and residual       */
    float DH, DT;
    float VPPML, FPML, npower, k_max_PML;
    float F_REF;           /* reference frequency of the Q model           */
    float TAU;             /* not used when taus/taup are given            */
    float FL[8];           /* relaxation frequencies, FL[1..L] used        */
    int SNAP_PREC;         /* snapshot storage: 0 FP32 (the reference), 1 BF16,
                              2 FP16, each with a power-of-two max-abs scale per
                              field, slot and 64x8 tile (no reference counterpart) */
    int EPRECOND;          /* energy preconditioning: 0 off, 1 sqrt(Ws*Wr), 3
                              sqrt(Ws) times the receiver Green's-function term
                              (src/eprecond1.c)                              */
    int EPRECOND_PER_SHOT; /* 1: divide each shot's gradient by We/max(We)    */
    float EPSILON_WE;      /* regularisation of We, relative to max(We)      */
    int EPRECOND_FP64;     /* 1: Ws and Wr summed in FP64 from products of the
                              velocities formed in FP64, We from them in FP64
                              (no reference counterpart)                     */
} orc_params;

/* Sources: srcpos[l*8 + k] holds the reference's srcpos_loc[k+1][l+1]
 * (x grid index, y grid index, 0, tshift, fc, amp, azimuth, type). */
#define ORC_SRC_STRIDE 8

/* ---- setup (all outputs are plain arrays) ------------------------------ */
void orc_holbergcoeff(int FDORDER, int MAXRELERROR, float hc[7]);
/* 10 CPML profiles, each 2*FW long (index 0 == reference index 1), order:
 * K_x,a_x,b_x,K_x_half,a_x_half,b_x_half,K_y,a_y,b_y,K_y_half,a_y_half,b_y_half
 * (12 arrays). */
void orc_pml_pro(const orc_params *p, float *out12);
/* 16-bit snapshot codes: prec 1 BF16, 2 FP16, round to nearest even
 * (IEEE 754 roundTiesToEven, subnormals kept). */
unsigned short orc_f32_to_16(float x, int prec);
float orc_16_to_f32(unsigned short h, int prec);
/* Source wavelets: out[l*NT + (nt-1)] = signals[l+1][nt]. */
void orc_wavelet(const orc_params *p, int nsrc, const float *srcpos, float *out);

/* ---- whole-shot drivers ------------------------------------------------- */
/* Models are interior [NY][NX] row-major (x fastest).  taus/taup may be NULL
 * for L == 0.  recpos[2*itr] = x index, recpos[2*itr+1] = y index.
 * Seismograms are [ntr][NT].  final5 (optional) receives the interior of
 * vx,vy,sxx,syy,sxy after the last forward step, 5*[NY][NX]. */
int orc_forward(const orc_params *p, const float *vp, const float *vs,
                const float *rho, const float *taus, const float *taup,
                int nsrc, const float *srcpos, int ntr, const int *recpos,
                real *seis_vx, real *seis_vy, real *final5);

/* ---- single-operator probes (pinning the stencils and the free surface) -
 * Probe grids are padded, rows -nd+1..NY+nd and columns -nd+1..NX+nd with
 * nd = FDORDER/2+1, x fastest, (NY+2nd)*(NX+2nd) values each.
 * orc_probe_derivs: fields5 = vx, vy, sxx, syy, sxy; out8 = [8][NY][NX] the
 * unscaled staggered differences (grid spacing 1, no CPML) that the
 * velocity update forms from the stresses, sxx_x, sxy_x, sxy_y, syy_y
 * (src/update_v_PML.c), then those the stress update forms from the
 * velocities, vxx, vyx, vxy, vyy (src/update_s_elastic_PML.c).  Orders 2-12.
 * orc_probe_surface: one application of surface_elastic_PML (L = 0) or
 * surface_PML (L > 0) to fields8 = vx, vy, sxx, syy, sxy, ux, uy, uxy (in
 * and out), with zero CPML memory; prow = [L][NX], the memory variables
 * p_l of row 1 (in and out). */
int orc_probe_derivs(const orc_params *p, const real *fields5, real *out8);
int orc_probe_surface(const orc_params *p, const float *vp, const float *vs, const float *rho, const float *taus,
                      const float *taup, real *fields8, real *prow);

/* Forward + residual + adjoint + zero-lag correlation + chain rule for one
 * shot.  obs_* are observed traces [ntr][NT] exactly as delivered by
 * inseis (first sample already zeroed by the caller).  kill (optional,
 * ntr ints) marks killed traces.  Outputs: synthetic seismograms, the
 * per-shot gradients (Vp, Vs, rho) [NY][NX], and misfit scalars
 * misfit[0] = L2 from calc_res (FP32 running sum, test-shot semantics),
 * misfit[1] = energy (calc_energy), misfit[2] = L2 from calc_misfit,
 * misfit[3] = L2 summed in double, misfit[4] = energy summed in double. */
int orc_gradient(const orc_params *p, const float *vp, const float *vs,
                 const float *rho, const float *taus, const float *taup,
                 int nsrc, const float *srcpos, int ntr, const int *recpos,
                 const float *obs_vx, const float *obs_vy, const int *kill,
                 real *seis_vx, real *seis_vy,
                 real *grad_vp, real *grad_vs, real *grad_rho,
                 double *misfit);

/* orc_gradient, and with EPRECOND set the shot's energy weight We [NY][NX]
 * (after the regularisation of src/eprecond1.c:63-68, before the per-shot
 * normalisation) in We_out. */
int orc_gradient_we(const orc_params *p, const float *vp, const float *vs,
                    const float *rho, const float *taus, const float *taup,
                    int nsrc, const float *srcpos, int ntr, const int *recpos,
                    const float *obs_vx, const float *obs_vy, const int *kill,
                    real *seis_vx, real *seis_vy,
                    real *grad_vp, real *grad_vs, real *grad_rho,
                    double *misfit, real *We_out);

/* Residual only (calc_res, LNORM=2): diff[ntr][NT] time-reversed adjoint
 * source, returns the FP32 running L2 (all traces counted). */
double orc_calc_res(const orc_params *p, int ntr, const real *syn,
                    const float *obs, const int *kill, real *diff);

#ifdef __cplusplus
}
#endif
#endif
