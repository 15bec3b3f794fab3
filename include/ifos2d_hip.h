/*
 * ifos2d_hip.h -- C ABI of the MI355X engine for IFOS2D's P-SV forward
 * solver and adjoint-state gradient.
 *
 * Two layers share the same gfx950 kernels:
 *
 *  1. Drop-in entry points with the reference's names and argument lists
 *     (src/fd.h of IFOS2D 2.0.3).  They read the same global parameters by
 *     the same names (src/globvar.h) and operate on device-resident
 *     Numerical-Recipes arrays created by ifos_dev_matrix()/ifos_dev_f3tensor()
 *     / ifos_dev_vector(), whose row-pointer arrays have the shape of
 *     src/util.c:95,184,279 but point into device memory.  A driver that
 *     allocates its wavefield, material and CPML arrays with these, and its
 *     host-side arrays (sources, signals, seismogram sections) as before,
 *     links against this library instead of the CPU kernel files.
 *
 *  2. The native engine (ifos_engine_*): the whole shot -- forward run with
 *     on-device snapshots, residual, adjoint run with fused zero-lag
 *     correlation, chain rule -- runs on the device under captured HIP
 *     graphs.  Plain pointers and sizes only.
 *
 * Errors: every int-returning function returns 0 on success and a negative
 * code on failure; ifos_last_error() gives the message.  The drop-in void
 * functions report failures the way the reference does (src/util.c:26-36):
 * message on stderr, then exit(1).
 */
#ifndef IFOS2D_HIP_H
#define IFOS2D_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Run parameters.  Field names are the reference globals (src/globvar.h).  */
/* ------------------------------------------------------------------------ */
typedef struct ifos_config {
    int NX, NY;            /* interior grid, x fastest                         */
    int FDORDER;           /* 2..12, even                                      */
    int MAXRELERROR;       /* row of the Holberg table, 0..4                   */
    int FW;                /* CPML width in cells                              */
    int FREE_SURF;         /* free surface at j = 1                            */
    int BOUNDARY;          /* 1: no CPML on the left/right edges               */
    int L;                 /* relaxation mechanisms (0: elastic)               */
    int VELOCITY;          /* 0: displacement misfit / snapshots of stress     */
    int ADJOINT_TYPE;      /* 1: vx+vy residuals, 2: vy, 3: vx                 */
    int PARAMETERIZATION;  /* 1: vp, vs, rho                                   */
    int DTINV;             /* snapshot stride; NT % DTINV must be 0            */
    int NT;                /* time steps                                       */
    int SOURCE_SHAPE;      /* wavelet.c shapes 1,2,4,5,6,8                     */
    int SOURCE_TYPE;       /* default type of sources without column 8         */
    int TRKILL;            /* honour the kill mask in residual and misfit      */
    float DH, DT;
    float VPPML, FPML, npower, k_max_PML;
    float F_REF;           /* reference frequency of the Q model               */
    float TAU;
    float FL[8];           /* relaxation frequencies FL[1..L]                  */
} ifos_config;

/* Number of floats per source row in srcpos arrays: x index, y index, 0,
 * tshift, fc, amp, azimuth (deg), type -- the rows of srcpos_loc
 * (src/splitsrc.c:40-49) stored source-major. */
#define IFOS_SRC_STRIDE 8

const char *ifos_last_error(void);
int ifos_device_count(void);

/* Setup helpers (host), bit-identical to the reference routines.           */
/* replaces holbergcoeff()      src/holbergcoeff.c:37                       */
void ifos_holberg(int FDORDER, int MAXRELERROR, float hc[7]);
/* replaces PML_pro()           src/PML_pro.c:83; 12 profiles of 2*FW:
 * K_x a_x b_x K_x_half a_x_half b_x_half K_y a_y b_y K_y_half a_y_half b_y_half */
void ifos_pml_profiles(const ifos_config *c, float *out12);
/* replaces wavelet()           src/wavelet.c:29; out[l*NT + nt-1]           */
void ifos_wavelet(const ifos_config *c, int nsrc, const float *srcpos, float *out);

/* ------------------------------------------------------------------------ */
/* Native engine                                                            */
/* ------------------------------------------------------------------------ */
typedef struct ifos_engine ifos_engine;

/* Snapshot storage precision (forward wavefield kept for the correlation). */
enum { IFOS_SNAP_FP32 = 0 };

int  ifos_engine_create(const ifos_config *cfg, int device, ifos_engine **out);
void ifos_engine_destroy(ifos_engine *e);
/* Models: host arrays [NY][NX] (x fastest); taus/taup ignored when L == 0.
 * Performs matcopy/av_mue/av_rho/av_tau/prepare_update_s
 * (src/IFOS2D.c:906-943) and uploads the result. */
int  ifos_engine_set_model(ifos_engine *e, const float *vp, const float *vs, const float *rho,
                           const float *taus, const float *taup);
/* One shot: srcpos[nsrc][IFOS_SRC_STRIDE] (grid indices), recpos[ntr][2]
 * (x index, y index).  Source wavelets are generated like wavelet(). */
int  ifos_engine_set_shot(ifos_engine *e, int nsrc, const float *srcpos, int ntr, const int *recpos);
/* Forward run (src/IFOS2D.c:1651-1894); snapshots every DTINV steps if asked. */
int  ifos_engine_forward(ifos_engine *e, int store_snapshots);
/* Seismograms of the last forward run, [ntr][NT] each. */
int  ifos_engine_seismograms(ifos_engine *e, float *vx, float *vy);
/* Interior of vx, vy, sxx, syy, sxy after the last run, 5*[NY][NX]. */
int  ifos_engine_fields(ifos_engine *e, float *out5);
/* Residuals/adjoint sources from observed traces [ntr][NT] (first sample
 * zeroed as src/inseis.c:75 delivers them), optional kill mask (ntr ints).
 * obs_on_device != 0: obs pointers are device addresses.
 * misfit[0] = L2 (FP64 sum), misfit[1] = energy of the data (FP64 sum). */
int  ifos_engine_residual(ifos_engine *e, const float *obs_vx, const float *obs_vy, const int *kill,
                          int obs_on_device, double *misfit);
/* Adjoint run with the fused zero-lag correlation and the chain rule; the
 * per-shot gradient is added to the running sums (src/IFOS2D.c:2224-2724). */
int  ifos_engine_adjoint(ifos_engine *e);
/* forward(with snapshots) + residual + adjoint for the current shot */
int  ifos_engine_shot_gradient(ifos_engine *e, const float *obs_vx, const float *obs_vy, const int *kill,
                               int obs_on_device, double *misfit);
/* which = 0: gradient of the last shot, 1: sum over shots since the reset.
 * Arrays [NY][NX]: Vp, Vs, rho gradients. */
int  ifos_engine_gradient(ifos_engine *e, int which, float *g_vp, float *g_vs, float *g_rho);
int  ifos_engine_reset_gradient(ifos_engine *e);
/* Device memory helpers for callers without a GPU runtime of their own. */
void *ifos_dev_alloc(size_t bytes);
int  ifos_dev_free(void *p);
int  ifos_memcpy_h2d(void *dst, const void *src, size_t bytes);
int  ifos_memcpy_d2h(void *dst, const void *src, size_t bytes);
int  ifos_memcpy_d2d(void *dst, const void *src, size_t bytes);
/* Time one kernel of the engine on its own stream with HIP events:
 * kernel 0 = velocity update, 1 = stress update, 2 = post (surface/source/
 * receivers), 3 = correlation.  Runs reps launches on the live state, then
 * restores nothing: call between shots.  Returns the average in ms. */
int  ifos_engine_time_kernel(ifos_engine *e, int kernel, int reps, double *ms_avg);
/* Algorithmic HBM bytes per cell of one launch of that kernel. */
double ifos_engine_kernel_bytes_per_cell(ifos_engine *e, int kernel);
/* Synchronise and report the first NaN step seen by the stability probe (0: none). */
int  ifos_engine_nan_step(ifos_engine *e);

#ifdef __cplusplus
}
#endif
#endif
