// ifos_internal.h -- device-side argument blocks and kernel launchers shared
// by the native engine (ifos_engine.hip) and the reference-signature entry
// points (ifos_dropin.hip).
//
// Addressing convention: every grid field is passed as an "origin" pointer,
// i.e. the address element (j=0, i=0) would have, plus a row pitch in floats,
// so that field(j,i) = origin[j*pitch + i] for the full padded range
// j in [-nd+1, NY+nd], i in [-nd+1, NX+nd] (nd = FDORDER/2+1).  Both the
// Numerical-Recipes layout of the reference (src/util.c:184, pitch NX+2nd)
// and the engine's own aligned layout fit this form.
//
// CPML memory: x-strips psi(j,h) = p[(j-1)*px + (h-1)], j 1..NY, h 1..2FW;
// y-strips psi(h,i) = p[(h-1)*py + (i-1)], h 1..2FW, i 1..NX
// (shapes of src/IFOS2D.c:628-641).  Profiles are 0-based arrays of 2FW.
//
// Memory variables r,p,q and coefficients d,e,dip of the GSLS model are
// addressed as base[(j*pitch+i)*cs + (l-1)*ls]: the reference f3tensor
// [j][i][l] layout has cs=L, ls=1; the engine keeps planes, cs=1, ls=plane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// message for ifos_last_error() (defined in ifos_engine.hip)
extern "C" void ifos_set_error(const char *fmt, ...);

namespace ifos {

constexpr int kMaxL = 8;
constexpr int kMaxN = 6;  // FDORDER/2 up to 12/2

// Every buffer the kernels touch is device global memory.  Pointers that
// reach a kernel through memory (an argument block in device memory) are
// generic to the compiler, which would then emit flat loads that share a
// wait counter with the scalar loads of that block; these casts give global
// loads and stores instead.
#define IFOS_GAS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ IFOS_GAS T *gmem(T *p)
{
    return (IFOS_GAS T *)p;
}
template <class T>
__device__ __forceinline__ const IFOS_GAS T *gmem(const T *p)
{
    return (const IFOS_GAS T *)p;
}
// read-only data of a launch (tile lists, CPML profiles, flags) through the
// constant address space: the compiler may keep such loads scalar and
// schedule them freely, since nothing writes them while the kernel runs
#define IFOS_CAS __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const IFOS_CAS T *cmem(const T *p)
{
    return (const IFOS_CAS T *)p;
}

struct Cpml {
    const float *K_x, *a_x, *b_x, *K_xh, *a_xh, *b_xh;
    const float *K_y, *a_y, *b_y, *K_yh, *a_yh, *b_yh;
};

struct StepGeom {
    int NX, NY, pitch;
    int FW, FREE_SURF, BOUNDARY;
    int N;            // FDORDER/2
    float DT, DH;
    float hc[7];
    Cpml pml;
    int px, py;       // CPML strip pitches
    int K_one;        // every CPML K value is 1 (the divisions by K are identities)
    // tile order of the reference-signature update kernels (tile_of_block):
    // tnbx tiles of 64x4 cells per tile row; XCD x runs tiles [xlo[x], xlo[x+1])
    int tnbx;
    int xlo[9];
};

// update_v_PML (src/update_v_PML.c:29-596)
struct VArgs {
    StepGeom g;
    int sw, VELOCITY;
    const float *sxx, *syy, *sxy, *rip, *rjp;
    float *vx, *vy, *vxp1, *vyp1;   // vxp1/vyp1 may be null: not written
    float *psi_sxx_x, *psi_sxy_x, *psi_syy_y, *psi_sxy_y;
};

// update_s_elastic_PML (src/update_s_elastic_PML.c:28-569)
struct SElArgs {
    StepGeom g;
    const float *vx, *vy;
    float *sxx, *syy, *sxy;
    float *ux, *uy, *uxy;           // may be null: stress rates not needed
    // material: either precomputed f=rho*vs*vs, g=rho*(vp^2-2vs^2)
    // (bit-identical to the per-cell expression) or raw rho/vs/vp
    const float *f, *gl, *uipjp;
    const float *rho, *vs, *vp;     // used when f/gl are null
    float *psi_vxx, *psi_vyx, *psi_vyy, *psi_vxy;
};

// update_s_visc_PML (src/update_s_visc_PML.c:28-515)
struct SViscArgs {
    StepGeom g;
    int L;
    const float *vx, *vy;
    float *sxx, *syy, *sxy;
    float *ux, *uy, *uxy;
    const float *fipjp, *f, *gv;
    float *r, *p, *q;
    const float *d, *e, *dip;
    int cs, ls;
    float bip[kMaxL + 1], bjm[kMaxL + 1], cip[kMaxL + 1], cjm[kMaxL + 1];
    float *psi_vxx, *psi_vyx, *psi_vyy, *psi_vxy;
};

// point sources / receivers
struct PointArgs {
    int n;
    const int *ix, *iy;       // grid indices
    const int *type;          // SOURCE_TYPE per source (forward)
    const double *sin_azi, *cos_azi;  // sin/cos of the float azimuth in radians,
                                      // evaluated on the host with libm
    const float *sig, *sig1;  // [n][ld] time series, 1-based nt -> column nt-1
    int ld;
};

// free surface + explosive source + seismogram sampling, fused per step
struct PostArgs {
    StepGeom g;
    int L;
    int sw;                   // 0 forward, 1 adjoint
    int explosive;            // psource active (forward, SOURCE_TYPE==1)
    int NT;
    PointArgs src;            // explosive sources (forward)
    int ntr;
    const int *rx, *ry;       // receivers
    float *seis_vx, *seis_vy; // [ntr][NT]
    float *vx, *vy;
    float *sxx, *syy, *sxy, *ux, *uy, *uxy;
    // elastic surface material: f (=rho*vs^2) and gl at row 1, or raw
    const float *f, *gl;
    const float *rho, *vs, *vp, *taus, *taup;
    float *p, *q;             // visc memory variables (cs/ls below)
    int cs, ls;
    float sumu, sumpi;
    float eta[kMaxL + 1];
    float *psi_vxxs;
    const int *nt_dev;        // when non-null the step number is read here
    int nt;                   // otherwise this one is used
    int *nan_flag;            // set to 1 if vy(NY/2,NX/2) is NaN
};

// snapshot store (src/IFOS2D.c:1805-1871): forward step nt with
// (nt-1) % DTINV == 0 stores into slot (nt-1)/DTINV
struct SnapArgs {
    int NX, NY, pitch, DTINV;
    const float *a[5];
    float *base[5];           // [nsnap][NY][NX]
    const int *nt_dev;        // step number read here when non-null
    int nt;
};

// zero-lag correlation (src/IFOS2D.c:2381-2419): adjoint step nt pairs
// with forward step NT-nt+1; it runs when (NT-nt) % DTINV == 0 and reads
// snapshot slot (NT-nt)/DTINV
struct CorrArgs {
    int NX, NY, pitch, NT, DTINV;
    const float *vxp1, *vyp1, *sxx, *syy, *sxy;
    const float *F[5];        // rho_x, rho_y, x, y, u bases [nsnap][NY][NX]
    const int *nt_dev;
    int nt;
    const float *muss, *lamss, *vs;   // [NY][NX]
    float *wc_rho, *wc, *wc_u;        // [NY][NX]
};

// chain rule to Vp/Vs/rho and shot sum (src/IFOS2D.c:2471-2611, 2703-2713)
struct GradArgs {
    int n;                    // NX*NY
    float DT;
    const float *vp, *vs, *rho;       // [NY][NX]
    const float *wc, *wc_u, *wc_rho;  // slot 0; slot s adds s * wcs
    long wcs;
    int nslots;
    float *g_vp, *g_vs, *g_rho;       // gradient of the current shots (summed in slot order)
    float *s_vp, *s_vs, *s_rho;       // running sums over shots (may be null)
    // energy preconditioner: per-slot weights We (stride wes) and, per slot,
    // the largest |We|; per_shot: divide each shot's gradient by We/We_max,
    // else add 1/We to we_sum (null: no preconditioning)
    const float *we, *we_max;
    long wes;
    int per_shot;
    float *we_sum;
    // circular source taper of each shot (taper_grad_shot sws 1): table
    // tap_m [tap_n][tap_n] whose cell (tap_c, tap_c) sits on the source;
    // per slot the source cell tap_ij[2s..2s+1] and the normalisation
    // tap_lohi[2s..2s+1]; cells within tap_fs of the source become zero
    // (null tap_m: no taper)
    const float *tap_m;
    const int *tap_ij;
    const float *tap_lohi;
    int tap_n, tap_c, tap_fs, NX, NY;
    // per-shot taper grids read from files (taper_grad_shot sws 2-4): slot s
    // holds Vp, Vs and rho coefficients [3][n] at tf + s * tfs (null: none)
    const float *tf;
    long tfs;
    // each slot's own gradient after the preconditioning and tapers, before
    // the sum (null: not kept): Vp, Vs, rho at slot_g + s * slot_gs + {0, n, 2n}.
    // It may be the correlation buffer itself: a thread reads its cell's
    // correlations of a slot before it writes that cell's gradients there
    float *slot_g;
    long slot_gs;
};

// adjoint sources from seismograms (src/calc_res.c, LNORM 2)
struct ResArgs {
    int ntr, NT, VELOCITY;
    float DT;
    const float *syn, *obs;   // [ntr][NT]
    const int *kill;          // may be null
    float *diff;              // [ntr][NT], time reversed
    double *l2_trace, *en_trace;   // per-trace FP64 sums
};

// time-domain Butterworth filter of traces [ntr][ns] (stride between
// traces): a lowpass cascade, its result rounded to float, then a highpass
// cascade (either may be empty), per trace behind one leading zero sample
constexpr int kMaxFiltSections = 6;  // orders up to 12
struct FiltArgs {
    float *data;
    int ntr, ns;
    long stride;
    int nlow, nhigh;
    double c[2 * kMaxFiltSections][5];  // the lowpass sections, then the highpass ones
};

// launchers (ifos_kernels.hip)
void launch_time_filter(const FiltArgs &a, hipStream_t s);
void launch_update_v(const VArgs &a, hipStream_t s);
void launch_inject(const PointArgs &a, const StepGeom &g, float *vx, float *vy, const float *rip,
                   const float *rjp, int sw, int ADJOINT_TYPE, const int *nt_dev, int nt, int *nt_tick,
                   hipStream_t s);
void launch_update_s_elastic(const SElArgs &a, hipStream_t s);
void launch_update_s_visc(const SViscArgs &a, hipStream_t s);
void launch_post(const PostArgs &a, hipStream_t s);
void launch_snapshot(const SnapArgs &a, hipStream_t s);
void launch_correlate(const CorrArgs &a, hipStream_t s);
void launch_gradient(const GradArgs &a, hipStream_t s);
void launch_residual(const ResArgs &a, hipStream_t s);

// chain rule of one cell (src/IFOS2D.c:2471-2611, PARAMETERIZATION 1): the
// correlations wc (lambda), wc_u (mu) and wc_rho of a shot become the Vp, Vs
// and rho gradients; wl, wmu, wrs are the intermediate lambda', mu' and
// rho' (waveconv_lam, waveconv_mu, waveconv_rho_s)
__device__ __forceinline__ void chain_rule_cell(float DT, float rh, float u, float pv, float wc, float wcu, float wcr,
                                                float &svp, float &svs, float &srho, float &wl, float &wmu,
                                                float &wrs)
{
    const float muss = rh * u * u;
    const float lamss = (float)((double)(rh * pv * pv) - 2.0 * (double)muss);
    wl = -DT * wc;
    wl = (float)((1.0 / (4.0 * (double)(lamss + muss) * (double)(lamss + muss))) * (double)wl);
    svp = (float)(2.0 * (double)pv * (double)rh * (double)wl);
    wmu = -DT * wcu;
    svs = (float)(-4.0 * (double)rh * (double)u * (double)wl + 2.0 * (double)rh * (double)u * (double)wmu);
    wrs = -DT * wcr;
    srho = (float)(((double)(pv * pv) - (2.0 * (double)u * (double)u)) * (double)wl + (double)(u * u * wmu) +
                   (double)wrs);
}

}  // namespace ifos
