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
    int BOUNDARY;          /* 1: periodic left/right edges, no CPML there      */
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
enum { IFOS_SNAP_FP32 = 0, IFOS_SNAP_BF16 = 1, IFOS_SNAP_FP16 = 2 };

int  ifos_engine_create(const ifos_config *cfg, int device, ifos_engine **out);
/* An engine with `slots` shot slots of one model: set_shots gives each slot
 * its own shot, and every forward / adjoint advances all of them in the same
 * kernel launches.  The reference loops over shots one at a time
 * (src/IFOS2D.c:1076-2724); per shot the results are identical to that
 * loop, and the gradient sum adds the shots in slot order as it does. */
int  ifos_engine_create_batch(const ifos_config *cfg, int device, int slots, ifos_engine **out);
void ifos_engine_destroy(ifos_engine *e);
/* Models: host arrays [NY][NX] (x fastest); taus/taup ignored when L == 0.
 * Performs matcopy/av_mue/av_rho/av_tau/prepare_update_s
 * (src/IFOS2D.c:906-943) and uploads the result. */
int  ifos_engine_set_model(ifos_engine *e, const float *vp, const float *vs, const float *rho,
                           const float *taus, const float *taup);
/* One shot: srcpos[nsrc][IFOS_SRC_STRIDE] (grid indices), recpos[ntr][2]
 * (x index, y index).  Source wavelets are generated like wavelet(). */
int  ifos_engine_set_shot(ifos_engine *e, int nsrc, const float *srcpos, int ntr, const int *recpos);
/* nshots shots (<= slots): shot k has nsrc[k] sources and ntr[k] receivers;
 * srcpos and recpos hold all shots' rows concatenated in shot order.  From
 * then on seismograms, observed traces, kill masks and adjoint sources are
 * [sum ntr][NT] in the same order.  set_shot is set_shots with one shot. */
int  ifos_engine_set_shots(ifos_engine *e, int nshots, const int *nsrc, const float *srcpos, const int *ntr,
                           const int *recpos);
/* The source signals of the current shots' sources, [nsrc][NT] in source
 * order, instead of the wavelets set_shots generated: the wavelet read
 * from a file (SOURCE_SHAPE 7, inseis_source_wavelet, src/IFOS2D.c:1527),
 * e.g. the wavelet convolved with an inverted source time function
 * (ifos2d_stf.h).  Used as given: the time filter is not applied to them,
 * as the reference skips it with INV_STF (src/IFOS2D.c:1556).  Kept until
 * the next set_shots. */
int  ifos_engine_set_wavelets(ifos_engine *e, const float *sig);
/* Per shot of the last residual: per_shot[2k] = L2, per_shot[2k+1] = energy (FP64). */
int  ifos_engine_shot_misfit(ifos_engine *e, double *per_shot);
/* Launch shape of the step kernels: spb = shots per workgroup (1, 2, 4; a
 * cell's materials are loaded once for all of them), grid = 0 for one
 * workgroup per (tile, shot group) or a multiple of 8 for a persistent grid
 * of that many workgroups. */
/* Time loop: mode 1 replays captured HIP graphs of a block of steps, 0
 * launches every kernel from the host, -1 (the default) chooses by size:
 * graphs below 2^19 cells x slots, where launching would be the bottleneck.
 * Results are the same bits either way.  The environment variable
 * IFOS_NO_GRAPH=1 (or 0) sets 0 (or 1) at creation.  get_graphs reports
 * what the next run will use. */
int  ifos_engine_set_graphs(ifos_engine *e, int mode);
int  ifos_engine_get_graphs(ifos_engine *e, int *on);
int  ifos_engine_set_launch(ifos_engine *e, int spb, int grid);
int  ifos_engine_get_launch(ifos_engine *e, int *spb, int *grid);
/* The persistent grids of the velocity and the stress step (set_launch sets
 * both; the autotuner picks each on its own). */
int  ifos_engine_get_launch_grids(ifos_engine *e, int *grid_vel, int *grid_str);
/* Order of the work items of a several-shot engine: 0 = every slot group of
 * a tile, then the next tile (a tile's materials are reused across shots);
 * 1 = every tile of a slot group, then the next group (a shot's
 * neighbouring tiles, whose halo rows they share, run close together). */
int  ifos_engine_set_work_order(ifos_engine *e, int slot_major);
int  ifos_engine_get_work_order(ifos_engine *e, int *slot_major);
/* Cells per thread along y in the velocity step: 1 (one cell per thread of
 * a 64x8 workgroup), 2 or 4 (adjacent rows per thread of a 64x4 or 64x2
 * workgroup, which share the y-stencil loads).  Results are the same bits
 * every way; the autotuner picks the fastest; an engine starts with 2 when it has
 * at least 2048 (tile, shot) work items, else with 1. */
int  ifos_engine_set_velocity_rows(ifos_engine *e, int rows);
int  ifos_engine_get_velocity_rows(ifos_engine *e, int *rows);
/* Velocity step with two rows per thread whose stencil values come from an
 * LDS copy of the tile and its halo (one row-contiguous read per value and
 * workgroup) instead of per-thread global loads; on = 1 also sets two rows
 * per thread, set_velocity_rows with 1 or 4 clears it.  Same bits.  The
 * autotuner does not pick it: on the 2000x800 grid it takes 28.7 us per
 * launch against 15.0 us for two rows from global memory. */
int  ifos_engine_set_velocity_lds(ifos_engine *e, int on);
int  ifos_engine_get_velocity_lds(ifos_engine *e, int *on);
/* One kernel per time step instead of the velocity and stress kernels: each
 * workgroup stages the old stresses of its 64x8 tile and a 2*FDORDER/2-cell
 * margin in LDS, forms the new velocities of the tile and of an
 * FDORDER/2-cell rim (what the neighbouring tiles form for those cells, the
 * same bits), then the tile's new stresses.  The wavefields and the
 * velocity-side CPML memory alternate between two sets from step to step,
 * because tiles read their neighbours' old values.  Same bits as the
 * two-kernel step; one shot per workgroup on this path. */
int  ifos_engine_set_fused(ifos_engine *e, int on);
int  ifos_engine_get_fused(ifos_engine *e, int *on);
/* Time every candidate launch shape (and, with several slots, both work
 * orders) on one forward step of the current model and shots (reps launches
 * of each kernel), the two-kernel step and the one-kernel step
 * (ifos_engine_set_fused), keep the fastest, zero the stepped wavefields;
 * reports the pick (grid: the stress step's or the one kernel's) and its
 * time per step. */
int  ifos_engine_autotune(ifos_engine *e, int reps, int *spb, int *grid, double *ms_per_step);
/* Forward run (src/IFOS2D.c:1651-1894); snapshots every DTINV steps if asked. */
int  ifos_engine_forward(ifos_engine *e, int store_snapshots);
/* Seismograms of the last forward run, [ntr][NT] each (all shots' traces). */
int  ifos_engine_seismograms(ifos_engine *e, float *vx, float *vy);
/* Interior of vx, vy, sxx, syy, sxy of slot 0 after the last run, 5*[NY][NX]. */
int  ifos_engine_fields(ifos_engine *e, float *out5);
/* Residuals/adjoint sources from observed traces [ntr][NT] (first sample
 * zeroed as src/inseis.c:75 delivers them), optional kill mask (ntr ints).
 * The L2 residual of src/calc_res.c (LNORM 2, the norm every workload here
 * uses); the other norms of LNORM 5-8 are not provided.
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
/* which = 0: gradient of the current shots (their sum in slot order; the
 * last shot's gradient for a one-slot engine), 1: sum over shots since the reset.
 * Arrays [NY][NX]: Vp, Vs, rho gradients. */
int  ifos_engine_gradient(ifos_engine *e, int which, float *g_vp, float *g_vs, float *g_rho);
int  ifos_engine_reset_gradient(ifos_engine *e);
/* The gradient of the shot in `slot` alone, from the last adjoint run (after
 * its energy preconditioning and tapers, before the sum): what a one-slot
 * engine's which = 0 gradient would be for that shot, bit for bit.  Kept
 * until the next forward run or set_shots.  A caller that sums shots of
 * several requests run in one launch adds these per request in shot order,
 * as the reference's loop adds shots (src/IFOS2D.c:2703-2713). */
int  ifos_engine_slot_gradient(ifos_engine *e, int slot, float *g_vp, float *g_vs, float *g_rho);
/* Energy preconditioning of the gradient (EPRECOND, EPSILON_WE,
 * EPRECOND_PER_SHOT of src/IFOS2D.c): mode 1 weights with
 * sqrt(Ws*Wr), mode 3 with sqrt(Ws) times the receiver Green's-function
 * term (src/eprecond1.c:8-69); Ws, Wr = sums over the time steps of
 * vx^2 + vy^2 of the forward and adjoint wavefields (src/eprecond.c),
 * accumulated inside the velocity step of gradient runs.  per_shot != 0
 * divides each shot's gradient by We/max|We| (src/IFOS2D.c:2627-2638);
 * otherwise 1/We is summed over shots and ifos_engine_apply_eprecond scales
 * the summed gradient by that sum over its maximum (src/IFOS2D.c:2737-2751).
 * mode 0 turns it off. */
int  ifos_engine_set_eprecond(ifos_engine *e, int mode, float epsilon_we, int per_shot);
/* on = 1: the energy sums Ws and Wr accumulate in FP64 (each step adds
 * vx*vx + vy*vy formed in FP64) and We is formed from them in FP64, then
 * rounded to FP32 (no reference counterpart: src/eprecond.c:8-21 sums in
 * FP32, as on = 0, the default, does bit for bit). */
int  ifos_engine_set_eprecond_fp64(ifos_engine *e, int on);
int  ifos_engine_apply_eprecond(ifos_engine *e);
/* which = 0: We of the first slot's shot of the last adjoint run; 1: the
 * sum of 1/We since the last apply.  [NY][NX]. */
int  ifos_engine_energy_weight(ifos_engine *e, int which, float *out);
/* Circular taper of each shot's gradient around its source
 * (SWS_TAPER_CIRCULAR_PER_SHOT, src/IFOS2D.c:2679-2689 calling
 * taper_grad_shot sws 1, src/taper_grad_shot.c:57-196): erf (SRTSHAPE 1) or
 * log (2) profile of radius SRTRADIUS metres, normalised over the grid, and
 * zero in the (2 FILTSIZE+1)^2 box around the source.  Applied after the
 * energy preconditioning and before the shot is added to the sum, to all
 * three gradients.  The taper centre is the shot's first source (the
 * reference's srcpos[.][ishot]).  on = 0 turns it off. */
int  ifos_engine_set_shot_taper(ifos_engine *e, int on, float SRTRADIUS, int SRTSHAPE, int FILTSIZE);
/* Per-shot taper grids read from files (SWS_TAPER_FILE_PER_SHOT,
 * taper_grad_shot sws 2-4, src/taper_grad_shot.c:199-288: the files
 * TAPER_FILE_NAME.shot<n>.vp/.vs/.rho as ifos_model_read returns them): the
 * shot in `slot` has its Vp, Vs and rho gradients multiplied by these
 * [NY][NX] grids, after the circular taper and before the shot sum, as
 * src/IFOS2D.c:2691-2698 orders them.  Slots never given grids keep their
 * gradients; set the grids again when the slot's shot changes.  All three
 * NULL turns it off. */
int  ifos_engine_set_shot_taper_file(ifos_engine *e, int slot, const float *vp, const float *vs, const float *rho);
/* Time-domain Butterworth filter of the multi-scale inversion (TIME_FILT,
 * timedomain_filt src/timedomain_filt.c:27 with the vendored libcseife):
 * lowpass of order ORDER at F_LOW_PASS Hz, then, when F_HIGH_PASS != 0, a
 * highpass of the same order at F_HIGH_PASS.  Applied, as the reference
 * does, to every shot's source wavelets (unless SOURCE_SHAPE is 6,
 * src/IFOS2D.c:1556-1561) and to the observed traces given to
 * ifos_engine_residual (src/IFOS2D.c:1985-2071; a device copy is filtered,
 * the caller's traces are left alone).  Double precision per trace as
 * there; bit-identical traces.  F_LOW_PASS = 0 turns it off; ORDER <= 12. */
int  ifos_engine_set_time_filter(ifos_engine *e, float F_LOW_PASS, int ORDER, float F_HIGH_PASS);
/* replaces timedomain_filt() src/timedomain_filt.c:27 on device traces
 * data [ntr][ns]: method 1 lowpass at fc (then a highpass at F_HIGH_PASS
 * when it is not 0, the global of that name), method 2 highpass at fc;
 * Butterworth of the given order (<= 12) in double per trace. */
int  ifos_time_filter(float *data, int ntr, int ns, float DT, float fc, int order, int method, float F_HIGH_PASS);
/* Storage of the forward snapshots (src/IFOS2D.c:1805-1871 keeps them FP32):
 * 0 = FP32 (default; gradients bit-identical to the reference), 1 = BF16,
 * 2 = FP16.  16-bit snapshots hold value / 2^e rounded to nearest-even, with
 * 2^e the smallest power of two above the field's largest magnitude in each
 * 64x8 tile of each snapshot slot; the correlation and gradient stay FP32.
 * Changing it frees the snapshot storage. */
int  ifos_engine_set_snapshot_precision(ifos_engine *e, int prec);
/* Device bytes of the snapshot storage of one shot at that precision
 * (0 for an invalid configuration). */
size_t ifos_snapshot_bytes(const ifos_config *c, int prec);
/* Device bytes an engine with `slots` slots holds for shots with ntr_total
 * receivers and nsrc_total sources in all, with snapshot storage at that
 * precision when snapshots != 0 (0 for an invalid configuration); the
 * admission budget of the request server (ifos2d_serve.h) counts this. */
/* Shots per launch for several-shot engines: the largest slot count whose
 * per-step working set (each slot's wavefields and memory variables, the
 * shared materials) fits in 160 MB of the 256 MB MALL; at least 1.  Larger
 * launches turn DRAM-bound: on the 1700x350, L=2 grid 4-5 shots per launch
 * gave 10.6 shot-gradients/s, 16 gave 8.7.  -2 for an invalid config. */
int ifos_engine_mall_slots(const ifos_config *c);
size_t ifos_engine_bytes(const ifos_config *c, int slots, int snap_prec, int snapshots, int ntr_total,
                         int nsrc_total);
/* Device memory helpers for callers without a GPU runtime of their own. */
void *ifos_dev_alloc(size_t bytes);
int  ifos_dev_free(void *p);
int  ifos_memcpy_h2d(void *dst, const void *src, size_t bytes);
int  ifos_memcpy_d2h(void *dst, const void *src, size_t bytes);
int  ifos_memcpy_d2d(void *dst, const void *src, size_t bytes);
/* Time one kernel of the engine on its own stream with HIP events:
 * kernel 0 = fused velocity step, 1 = fused stress step (forward step
 * without snapshot).  Runs reps launches on the live state, then restores
 * nothing: call between shots.  Returns the average in ms. */
int  ifos_engine_time_kernel(ifos_engine *e, int kernel, int reps, double *ms_avg);
/* Time reps device-to-device copies that read and write bytes/2 each, back
 * to back on the engine's stream (HIP events; no counterpart in the
 * reference): the streaming rate a kernel moving that many bytes per launch
 * can reach at that size.  Returns the average in ms. */
int  ifos_engine_time_copy(ifos_engine *e, size_t bytes, int reps, double *ms_avg);
/* Algorithmic HBM bytes per cell of one launch of that kernel. */
double ifos_engine_kernel_bytes_per_cell(ifos_engine *e, int kernel);
/* Synchronise and report the first NaN step seen by the stability probe (0: none). */
int  ifos_engine_nan_step(ifos_engine *e);

/* ------------------------------------------------------------------------ */
/* FP64 build of the same scheme (no reference counterpart; the reference   */
/* is FP32): the yardstick that reduced-precision results are measured      */
/* against.  Fields, materials, residuals, correlation and gradients are    */
/* double; the setup scalars (Holberg coefficients, CPML profiles, source   */
/* wavelets, GSLS factors) are the reference's FP32 values.  The forward    */
/* state is kept every segment_steps steps (a multiple of DTINV; 0 picks    */
/* DTINV*ceil(sqrt(NT/DTINV))) and the adjoint rebuilds one segment of      */
/* snapshots at a time, so memory is NT/C states plus C/DTINV snapshot      */
/* slots instead of NT/DTINV slots.  Energy preconditioning, tapers and     */
/* time filtering are not provided here.                                    */
/* ------------------------------------------------------------------------ */
typedef struct ifos_fp64 ifos_fp64;
int  ifos_fp64_create(const ifos_config *cfg, int device, int segment_steps, ifos_fp64 **out);
void ifos_fp64_destroy(ifos_fp64 *e);
/* Device bytes of such an engine (0 for an invalid configuration). */
size_t ifos_fp64_bytes(const ifos_config *cfg, int segment_steps, int ntr, int nsrc);
/* The segment length in use. */
int  ifos_fp64_segment_steps(ifos_fp64 *e);
/* Models as for ifos_engine_set_model (host [NY][NX] FP32). */
int  ifos_fp64_set_model(ifos_fp64 *e, const float *vp, const float *vs, const float *rho, const float *taus,
                         const float *taup);
/* One shot, arguments as for ifos_engine_set_shot. */
int  ifos_fp64_set_shot(ifos_fp64 *e, int nsrc, const float *srcpos, int ntr, const int *recpos);
/* Forward, residual (LNORM 2) against host observed traces [ntr][NT],
 * adjoint with the zero-lag correlation, chain rule; the shot's gradient is
 * added to the running sums.  misfit[0] = L2, misfit[1] = data energy. */
int  ifos_fp64_shot_gradient(ifos_fp64 *e, const float *obs_vx, const float *obs_vy, const int *kill,
                             double *misfit);
/* Seismograms of the last shot, [ntr][NT] each. */
int  ifos_fp64_seismograms(ifos_fp64 *e, double *vx, double *vy);
/* which = 0: the last shot's gradients, 1: the sums since the reset. */
int  ifos_fp64_gradient(ifos_fp64 *e, int which, double *g_vp, double *g_vs, double *g_rho);
int  ifos_fp64_reset_gradient(ifos_fp64 *e);

#ifdef __cplusplus
}
#endif
#endif
