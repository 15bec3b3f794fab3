// ifos_engine.hip -- native whole-shot engine behind the C ABI
// (include/ifos2d_hip.h).  One process drives one GPU; all state of a shot
// stays in HBM: wavefields, CPML memory, materials, the forward snapshots
// (slot-major [k][j][i]), seismograms, adjoint sources and the gradient
// accumulators.  An engine holds S shot slots of one model (S = 1 for the
// one-shot-at-a-time path); every step kernel launch advances all active
// slots.  The time loop is recorded once per geometry as HIP graphs of
// U = k*DTINV steps and replayed NT/U times; kernels read the step number
// from a device counter so one graph serves every block of steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "host/ifos_setup.h"
#include "ifos2d_hip.h"
#include "ifos2d_opt.h"
#include "ifos_internal.h"
#include "ifos_step.h"

namespace ifos {
void set_step_control(const int *nt_dev, int DTINV, int NT);
void launch_inject_serial(const PointArgs &a, const StepGeom &g, float *vx, float *vy, const float *rip,
                          const float *rjp, int sw, int ADJOINT_TYPE, const int *nt_dev, int nt, int *nt_tick,
                          hipStream_t s);
}  // namespace ifos

using namespace ifos;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

extern "C" void ifos_set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

extern "C" const char *ifos_last_error(void) { return g_err.c_str(); }

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            ifos_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                                             \
        }                                                                                          \
    } while (0)

extern "C" int ifos_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ---------------------------------------------------------------------------
// engine state
// ---------------------------------------------------------------------------
// Engines of one process may run in several threads (the request server).
// Device allocation, release and graph capture of all engines are ordered
// by one lock, so that no capture on one engine's stream overlaps another
// thread's hipMalloc/hipFree.  Copies and fills go to each engine's own
// non-blocking stream, never to the legacy stream, for the same reason.
static std::mutex &device_lock()
{
    static std::mutex m;
    return m;
}
template <typename T>
static hipError_t dmalloc(T **p, size_t bytes)
{
    std::lock_guard<std::mutex> g(device_lock());
    return hipMalloc(p, bytes);
}
static hipError_t dfree(void *p)
{
    if (!p) return hipSuccess;
    std::lock_guard<std::mutex> g(device_lock());
    return hipFree(p);
}

struct ifos_engine {
    ifos_config c;
    int device = 0;
    hipStream_t stream = nullptr;
    int N = 0, H = 0, pitch = 0, rows = 0;
    long field = 0;       // floats per padded grid
    long off = 0;         // origin offset inside a padded grid
    long plane = 0;       // NX*NY
    int nsnap = 0;
    float hc[7];
    int slots = 1;        // shot slots held
    int nslots = 0;       // slots in use by the current shots

    // per-slot padded grids (origin pointers of slot 0; slot s adds s*ws)
    float *slab = nullptr;
    float *vx, *vy, *sxx, *syy, *sxy, *ux, *uy, *uxy, *vxp1, *vyp1;
    float *r = nullptr, *p = nullptr, *q = nullptr;
    float *vx2, *vy2, *sxx2, *syy2, *sxy2;  // second set of the one-kernel step
    long ws = 0;          // floats per slot: 15 wavefields + 3L memory variables
    bool fused = false;   // one kernel per time step (k_fused_step)
    int last_set = 0;     // wavefield set holding the state after the last run
    // shared material grids
    float *mslab = nullptr;
    float *rip, *rjp, *uipjp, *f, *g, *fipjp, *rho, *vs, *vp, *taus, *taup;
    float *d = nullptr, *e = nullptr, *dip = nullptr;
    float *mu_r, *pi_r, *muipjp_r, *tausipjp;  // relaxed moduli for the in-kernel GSLS coefficients

    // interior [NY][NX]
    float *interior = nullptr;
    float *muss, *lamss, *vs_i, *vp_i, *rho_i, *g_vp, *g_vs, *g_rho, *s_vp, *s_vs, *s_rho;
    float *wcbuf = nullptr;  // per slot: wc, wc_u, wc_rho (3 planes)

    // CPML
    float *pml = nullptr;   // 12 profiles
    float *psi = nullptr;   // per slot: 5 x-strips then 4 y-strips
    float *psi_sxx_x, *psi_sxy_x, *psi_vxx, *psi_vyx, *psi_vxxs, *psi_syy_y, *psi_sxy_y, *psi_vyy, *psi_vxy;
    float *psi2_sxx_x, *psi2_sxy_x, *psi2_syy_y, *psi2_sxy_y;  // second set of the one-kernel step
    long pss = 0;           // floats per slot

    // control
    int *ctl = nullptr;     // [0] step counter, [1] NaN step
    ifos_materials mat;
    bool model_set = false;

    // shots: points of all slots in global arrays, slot s owning
    // sources [src_off[s], src_off[s+1]) and traces [tr_off[s], tr_off[s+1])
    int nsrc = 0, ntr = 0, cap_src = 0, cap_tr = 0;
    std::vector<int> src_off, tr_off;
    int *explosive = nullptr;  // [slots] device
    int *src_i = nullptr;   // ix, iy, type  [3][cap_src]
    double *src_d = nullptr;  // sin, cos [2][cap_src]
    float *sig = nullptr;   // [cap_src][NT]
    int *rec_i = nullptr;   // rx, ry, kill [3][cap_tr]
    float *seis = nullptr;  // vx, vy, diff_vx, diff_vy, obs_vx, obs_vy: 6*[cap_tr][NT]
    double *trace_sums = nullptr;  // l2, en [2][cap_tr]
    std::vector<double> shot_misfit;  // l2, en per shot of the last residual

    // per-slot, per-tile point lists (sources, receivers) for the step kernels
    int nbx = 0, nby = 0;
    int *tiles = nullptr;
    size_t tiles_cap = 0;
    int off_src_start = 0, off_src_idx = 0, off_rec_start = 0, off_rec_idx = 0;
    int off_srcx_start = 0, off_srcx_idx = 0, off_recx_start = 0, off_recx_idx = 0;  // tile + rim lists
    LaunchShape shape;
    // energy preconditioner (src/eprecond.c, src/eprecond1.c): Ws and Wr per
    // slot, the 1/We sum over shots, the mode-3 geometric factor
    int ep_mode = 0, ep_per_shot = 0;
    float ep_eps = 0.0f;
    float *wbuf = nullptr;      // [slots][2][plane]: Ws then Wr (Ws becomes We after the adjoint)
    int ep_fp64 = 0;            // Ws and Wr summed in FP64 (wbuf64) instead of FP32
    double *wbuf64 = nullptr;   // [slots][2][plane] when ep_fp64
    float *we_sum = nullptr;    // [plane]
    double *ep_geom = nullptr;  // [plane], mode 3
    float *we_max = nullptr;    // [slots]
    float *red = nullptr;       // reduction partials
    // circular source taper of each shot's gradient (taper_grad_shot sws 1)
    int tap_on = 0, tap_n = 0, tap_fs = 0;
    float *tap_m = nullptr;         // [tap_n][tap_n]
    int *tap_ij = nullptr;          // [slots][2] source cell of each slot
    float *tap_lohi = nullptr;      // [slots][2] normalisation
    std::vector<float> tap_host;    // the table on the host
    std::vector<int> first_src;     // [slots][2] first source of each shot (0, 0: none)
    float *tfbuf = nullptr;         // [slots][3][plane] per-shot taper grids from files
    bool tf_on = false;
    int K_one = 0;                  // every CPML K value is 1
    // time-domain Butterworth filter of wavelets and observed traces (TIME_FILT)
    int filt_nlow = 0, filt_nhigh = 0;
    double filt_c[2 * kMaxFiltSections][5];
    std::vector<float> sig_host;    // unfiltered wavelets of the current shots
    bool sig_given = false;         // the wavelets were set by the caller (ifos_engine_set_wavelets)
    StepArgs *dargs = nullptr;  // device copies of the argument blocks: forward, forward with
                                // snapshots, adjoint, eager/timing

    // snapshots: per slot FP32 planes, or per-tile scales of all slots
    // followed by the 16-bit codes of all slots
    void *snap = nullptr;
    size_t snap_bytes = 0;
    int snap_prec = 0;        // 0 FP32, 1 BF16, 2 FP16
    bool snap_valid = false;  // the last forward stored snapshots for the current model and shot
    bool slot_grad_valid = false;  // wcbuf holds each slot's gradient of the last adjoint run

    // graphs
    hipGraphExec_t gexec[2] = {nullptr, nullptr};  // forward(with snapshots), adjoint
    int gsteps[2] = {0, 0};
    int gkey[2][6] = {{-1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1}};
    int graph_mode = -1;            // time loop: 1 captured graphs, 0 eager launches, -1 by size (use_graphs)
    int swizzle = 1;
};

#define ENG_TRY(expr)       \
    do {                    \
        int rc_ = (expr);   \
        if (rc_) return rc_; \
    } while (0)

static long pad_index(const ifos_engine *E, int j, int i) { return (long)j * E->pitch + i; }

static int validate(const ifos_config *c)
{
    if (c->NX < 1 || c->NY < 1) return ifos_set_error("NX, NY must be positive"), -2;
    if (c->FDORDER < 2 || c->FDORDER > 12 || (c->FDORDER & 1))
        return ifos_set_error("FDORDER must be one of 2,4,6,8,10,12 (got %d)", c->FDORDER), -2;
    if (c->MAXRELERROR < 0 || c->MAXRELERROR > 4) return ifos_set_error("MAXRELERROR must be 0..4"), -2;
    if (c->L < 0 || c->L > 5) return ifos_set_error("L must be 0..5 (got %d)", c->L), -2;
    if (c->PARAMETERIZATION != 1) return ifos_set_error("only PARAMETERIZATION=1 (vp, vs, rho) is supported"), -2;
    if (c->ADJOINT_TYPE < 1 || c->ADJOINT_TYPE > 3) return ifos_set_error("ADJOINT_TYPE must be 1, 2 or 3"), -2;
    if (c->DTINV < 1 || c->NT < 1 || c->NT % c->DTINV)
        return ifos_set_error("NT (%d) must be a positive multiple of DTINV (%d)", c->NT, c->DTINV), -2;
    if (c->FW < 0 || 2 * c->FW > c->NX || 2 * c->FW > c->NY)
        return ifos_set_error("FW (%d) must fit twice into NX and NY", c->FW), -2;
    if (!(c->DT > 0) || !(c->DH > 0)) return ifos_set_error("DT and DH must be positive"), -2;
    // BOUNDARY 1: periodic sides, no CPML on the left and right edges; the
    // step kernels copy the edge columns into the opposite halo as the
    // reference's exchanges do on one rank (src/exchange_v.c, src/exchange_s.c)
    if (c->BOUNDARY != 0 && c->BOUNDARY != 1) return ifos_set_error("BOUNDARY must be 0 or 1"), -2;
    if (c->BOUNDARY && c->NX < c->FDORDER) return ifos_set_error("BOUNDARY 1 needs NX >= FDORDER"), -2;
    return 0;
}

static int engine_create(const ifos_config *cfg, int device, int slots, ifos_engine **out)
{
    *out = nullptr;
    ENG_TRY(validate(cfg));
    if (slots < 1 || slots > 4096) return ifos_set_error("slots must be 1..4096 (got %d)", slots), -2;
    ifos_engine *E = new ifos_engine();
    E->c = *cfg;
    E->device = device;
    E->slots = slots;
    const char *ng = getenv("IFOS_NO_GRAPH");
    if (ng && ng[0] == '1') E->graph_mode = 0;
    if (ng && ng[0] == '0') E->graph_mode = 1;
    const char *ns = getenv("IFOS_NO_SWIZZLE");
    E->swizzle = (ns && ns[0] == '1') ? 0 : 1;
    auto fail = [&](int rc) {
        ifos_engine_destroy(E);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) {
        ifos_set_error("hipSetDevice(%d) failed", device);
        return fail(-1);
    }
    if (hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess) {
        ifos_set_error("stream creation failed");
        return fail(-1);
    }
    const ifos_config &c = E->c;
    E->N = c.FDORDER / 2;
    E->H = E->N + 1;
    const int A = 16;  // interior column 1 starts 64-byte aligned
    E->pitch = ((A + c.NX + E->H) + 63) / 64 * 64;
    E->rows = c.NY + 2 * E->H + 1;
    E->field = (long)E->pitch * E->rows;
    E->off = (long)E->H * E->pitch + (A - 1);
    E->plane = (long)c.NX * c.NY;
    E->nsnap = c.NT / c.DTINV;
    E->nbx = (c.NX + kTileX - 1) / kTileX;
    E->nby = (c.NY + kTileY - 1) / kTileY;
    // two rows per thread in the velocity step once there are enough work
    // items to fill the chip several times over (measured: faster on
    // 2000x800 and on 16 x 1700x350, slower on 400x200 and on one 1700x350
    // shot); the autotuner can change it
    E->shape.rows_v = (long)E->nbx * E->nby * slots >= 2048 ? 2 : 1;
    ifos_holberg(c.FDORDER, c.MAXRELERROR, E->hc);
    const int L = c.L;

    // per slot: 10 wavefields + 3L memory planes + the second set of
    // vx, vy, sxx, syy, sxy for the one-kernel step
    const int nwave = 15 + 3 * L;
    E->ws = (long)nwave * E->field;
    const size_t wbytes = sizeof(float) * (size_t)E->ws * slots;
    if (dmalloc(&E->slab, wbytes) != hipSuccess) {
        ifos_set_error("device allocation of %d x %d wavefield grids failed", slots, nwave);
        return fail(-3);
    }
    hipMemsetAsync(E->slab, 0, wbytes, E->stream);
    float *gp = E->slab + E->off;
    auto next = [&]() {
        float *r = gp;
        gp += E->field;
        return r;
    };
    E->vx = next(); E->vy = next(); E->sxx = next(); E->syy = next(); E->sxy = next();
    E->ux = next(); E->uy = next(); E->uxy = next(); E->vxp1 = next(); E->vyp1 = next();
    if (L) {
        E->r = next(); for (int l = 1; l < L; l++) next();
        E->p = next(); for (int l = 1; l < L; l++) next();
        E->q = next(); for (int l = 1; l < L; l++) next();
    }
    E->vx2 = next(); E->vy2 = next(); E->sxx2 = next(); E->syy2 = next(); E->sxy2 = next();
    // shared: 15 material grids + 3L coefficient planes
    const int nmat = 15 + 3 * L;
    if (dmalloc(&E->mslab, sizeof(float) * E->field * nmat) != hipSuccess) {
        ifos_set_error("device allocation of %d material grids failed", nmat);
        return fail(-3);
    }
    hipMemsetAsync(E->mslab, 0, sizeof(float) * E->field * nmat, E->stream);
    gp = E->mslab + E->off;
    E->rip = next(); E->rjp = next(); E->uipjp = next(); E->f = next(); E->g = next(); E->fipjp = next();
    E->rho = next(); E->vs = next(); E->vp = next(); E->taus = next(); E->taup = next();
    E->mu_r = next(); E->pi_r = next(); E->muipjp_r = next(); E->tausipjp = next();
    if (L) {
        E->d = next(); for (int l = 1; l < L; l++) next();
        E->e = next(); for (int l = 1; l < L; l++) next();
        E->dip = next(); for (int l = 1; l < L; l++) next();
    }

    // interior arrays: 11 shared + 3 correlation planes per slot
    const int nint = 11;
    if (dmalloc(&E->interior, sizeof(float) * E->plane * nint) != hipSuccess ||
        dmalloc(&E->wcbuf, sizeof(float) * E->plane * 3 * slots) != hipSuccess) {
        ifos_set_error("device allocation of interior arrays failed");
        return fail(-3);
    }
    hipMemsetAsync(E->interior, 0, sizeof(float) * E->plane * nint, E->stream);
    float **ia[nint] = {&E->muss, &E->lamss, &E->vs_i, &E->vp_i, &E->rho_i, &E->g_vp, &E->g_vs, &E->g_rho,
                        &E->s_vp, &E->s_vs, &E->s_rho};
    for (int k = 0; k < nint; k++) *ia[k] = E->interior + (long)k * E->plane;

    // CPML profiles and per-slot memory
    const int fw2 = c.FW > 0 ? 2 * c.FW : 1;
    std::vector<float> prof((size_t)12 * fw2, 0.0f);
    if (c.FW > 0) ifos_pml_profiles(&c, prof.data());
    // K_x, K_xh, K_y, K_yh are profiles 0, 3, 6 and 9
    E->K_one = 1;
    for (int q : {0, 3, 6, 9})
        for (int h = 0; h < fw2; h++)
            if (prof[(size_t)q * fw2 + h] != 1.0f) E->K_one = 0;
    if (dmalloc(&E->pml, sizeof(float) * prof.size()) != hipSuccess) return fail(-3);
    hipMemcpyAsync(E->pml, prof.data(), sizeof(float) * prof.size(), hipMemcpyHostToDevice, E->stream);
    const size_t xs = (size_t)c.NY * fw2, ys = (size_t)fw2 * c.NX;
    E->pss = (long)(7 * xs + 6 * ys);
    if (dmalloc(&E->psi, sizeof(float) * E->pss * slots) != hipSuccess) return fail(-3);
    hipMemsetAsync(E->psi, 0, sizeof(float) * E->pss * slots, E->stream);
    E->psi_sxx_x = E->psi; E->psi_sxy_x = E->psi + xs; E->psi_vxx = E->psi + 2 * xs;
    E->psi_vyx = E->psi + 3 * xs; E->psi_vxxs = E->psi + 4 * xs;
    float *yb = E->psi + 5 * xs;
    E->psi_syy_y = yb; E->psi_sxy_y = yb + ys; E->psi_vyy = yb + 2 * ys; E->psi_vxy = yb + 3 * ys;
    float *b2 = yb + 4 * ys;
    E->psi2_sxx_x = b2; E->psi2_sxy_x = b2 + xs; E->psi2_syy_y = b2 + 2 * xs; E->psi2_sxy_y = b2 + 2 * xs + ys;

    if (dmalloc(&E->ctl, sizeof(int) * 4) != hipSuccess) return fail(-3);
    hipMemsetAsync(E->ctl, 0, sizeof(int) * 4, E->stream);
    if (dmalloc(&E->explosive, sizeof(int) * slots) != hipSuccess) return fail(-3);
    hipMemsetAsync(E->explosive, 0, sizeof(int) * slots, E->stream);
    if (dmalloc(&E->dargs, sizeof(StepArgs) * 4) != hipSuccess) return fail(-3);
    if (hipStreamSynchronize(E->stream) != hipSuccess) {
        ifos_set_error("engine set-up failed on the device");
        return fail(-1);
    }
    *out = E;
    return 0;
}

extern "C" int ifos_engine_create(const ifos_config *cfg, int device, ifos_engine **out)
{
    return engine_create(cfg, device, 1, out);
}

extern "C" int ifos_engine_create_batch(const ifos_config *cfg, int device, int slots, ifos_engine **out)
{
    return engine_create(cfg, device, slots, out);
}

static void drop_graphs(ifos_engine *E)
{
    for (int k = 0; k < 2; k++) {
        if (E->gexec[k]) hipGraphExecDestroy(E->gexec[k]);
        E->gexec[k] = nullptr;
        for (int q = 0; q < 6; q++) E->gkey[k][q] = -1;
    }
}

extern "C" void ifos_engine_destroy(ifos_engine *E)
{
    if (!E) return;
    hipSetDevice(E->device);
    if (E->stream) hipStreamSynchronize(E->stream);
    drop_graphs(E);
    dfree(E->slab); dfree(E->mslab); dfree(E->interior); dfree(E->wcbuf); dfree(E->pml);
    dfree(E->psi); dfree(E->ctl); dfree(E->explosive); dfree(E->dargs);
    dfree(E->wbuf); dfree(E->wbuf64); dfree(E->we_sum); dfree(E->ep_geom); dfree(E->we_max); dfree(E->red);
    dfree(E->tap_m); dfree(E->tap_ij); dfree(E->tap_lohi); dfree(E->tfbuf);
    dfree(E->src_i); dfree(E->src_d); dfree(E->sig); dfree(E->rec_i); dfree(E->seis);
    dfree(E->trace_sums); dfree(E->snap); dfree(E->tiles);
    if (E->stream) hipStreamDestroy(E->stream);
    ifos_free_materials(&E->mat);
    delete E;
}

// copy an interior [NY][NX] host array into a padded device grid
static int upload_grid(ifos_engine *E, float *dst_origin, const float *host)
{
    float *dst = dst_origin + pad_index(E, 1, 1);
    HIP_TRY(hipMemcpy2DAsync(dst, sizeof(float) * E->pitch, host, sizeof(float) * E->c.NX, sizeof(float) * E->c.NX,
                             E->c.NY, hipMemcpyHostToDevice, E->stream));
    return 0;
}

extern "C" int ifos_engine_set_model(ifos_engine *E, const float *vp, const float *vs, const float *rho,
                                     const float *taus, const float *taup)
{
    HIP_TRY(hipSetDevice(E->device));
    const ifos_config &c = E->c;
    if (c.L && (!taus || !taup)) return ifos_set_error("viscoelastic model needs taus and taup"), -2;
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->snap_valid = false;
    ifos_free_materials(&E->mat);
    ifos_setup_materials(&c, vp, vs, rho, taus, taup, &E->mat);
    ifos_materials &m = E->mat;
    ENG_TRY(upload_grid(E, E->rip, m.rip));
    ENG_TRY(upload_grid(E, E->rjp, m.rjp));
    ENG_TRY(upload_grid(E, E->uipjp, m.uipjp));
    ENG_TRY(upload_grid(E, E->f, m.f));
    ENG_TRY(upload_grid(E, E->g, m.g));
    ENG_TRY(upload_grid(E, E->rho, rho));
    ENG_TRY(upload_grid(E, E->vs, vs));
    ENG_TRY(upload_grid(E, E->vp, vp));
    if (c.L) {
        ENG_TRY(upload_grid(E, E->fipjp, m.fipjp));
        ENG_TRY(upload_grid(E, E->taus, taus));
        ENG_TRY(upload_grid(E, E->taup, taup));
        ENG_TRY(upload_grid(E, E->mu_r, m.mu_r));
        ENG_TRY(upload_grid(E, E->pi_r, m.pi_r));
        ENG_TRY(upload_grid(E, E->muipjp_r, m.muipjp_r));
        ENG_TRY(upload_grid(E, E->tausipjp, m.tausipjp));
        for (int l = 0; l < c.L; l++) {
            ENG_TRY(upload_grid(E, E->d + l * E->field, m.d + (size_t)l * E->plane));
            ENG_TRY(upload_grid(E, E->e + l * E->field, m.e + (size_t)l * E->plane));
            ENG_TRY(upload_grid(E, E->dip + l * E->field, m.dip + (size_t)l * E->plane));
        }
    }
    const size_t pb = sizeof(float) * E->plane;
    HIP_TRY(hipMemcpyAsync(E->muss, m.muss, pb, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipMemcpyAsync(E->lamss, m.lamss, pb, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipMemcpyAsync(E->vs_i, vs, pb, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipMemcpyAsync(E->vp_i, vp, pb, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipMemcpyAsync(E->rho_i, rho, pb, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->model_set = true;
    return 0;
}

// the time-domain filter over traces [ntr][NT] in place, on the engine stream
static void filter_traces(ifos_engine *E, float *d, int ntr, bool apply)
{
    if (!apply || (!E->filt_nlow && !E->filt_nhigh) || ntr <= 0) return;
    FiltArgs f;
    f.data = d;
    f.ntr = ntr;
    f.ns = E->c.NT;
    f.stride = E->c.NT;
    f.nlow = E->filt_nlow;
    f.nhigh = E->filt_nhigh;
    memcpy(f.c, E->filt_c, sizeof f.c);
    launch_time_filter(f, E->stream);
}

// shots of the slots: slot s has nsrc[s] sources (rows of srcpos, all
// slots' rows concatenated) and ntr[s] receivers (rows of recpos, likewise)
extern "C" int ifos_engine_set_shots(ifos_engine *E, int nshots, const int *nsrc_of, const float *srcpos,
                                     const int *ntr_of, const int *recpos)
{
    HIP_TRY(hipSetDevice(E->device));
    const ifos_config &c = E->c;
    if (nshots < 1 || nshots > E->slots)
        return ifos_set_error("%d shots for an engine of %d slots", nshots, E->slots), -2;
    std::vector<int> so(nshots + 1, 0), to(nshots + 1, 0);
    for (int k = 0; k < nshots; k++) {
        if (nsrc_of[k] < 0 || ntr_of[k] < 0) return ifos_set_error("negative source or receiver count"), -2;
        so[k + 1] = so[k] + nsrc_of[k];
        to[k + 1] = to[k] + ntr_of[k];
    }
    const int nsrc = so[nshots], ntr = to[nshots];
    for (int l = 0; l < nsrc; l++) {
        int ix = (int)srcpos[l * IFOS_SRC_STRIDE], iy = (int)srcpos[l * IFOS_SRC_STRIDE + 1];
        if (ix < 1 || ix > c.NX || iy < 1 || iy > c.NY)
            return ifos_set_error("source %d at (%d,%d) outside the grid 1..%d x 1..%d", l + 1, ix, iy, c.NX, c.NY), -2;
    }
    for (int t = 0; t < ntr; t++) {
        int ix = recpos[2 * t], iy = recpos[2 * t + 1];
        if (ix < 1 || ix > c.NX || iy < 1 || iy > c.NY)
            return ifos_set_error("receiver %d at (%d,%d) outside the grid", t + 1, ix, iy), -2;
    }
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->snap_valid = false;
    E->slot_grad_valid = false;
    if (nsrc > E->cap_src) {
        dfree(E->src_i); dfree(E->src_d); dfree(E->sig);
        E->src_i = nullptr; E->src_d = nullptr; E->sig = nullptr;
        E->cap_src = nsrc;
        HIP_TRY(dmalloc(&E->src_i, sizeof(int) * 3 * nsrc));
        HIP_TRY(dmalloc(&E->src_d, sizeof(double) * 2 * nsrc));
        HIP_TRY(dmalloc(&E->sig, sizeof(float) * (size_t)nsrc * c.NT));
        drop_graphs(E);
    }
    if (ntr > E->cap_tr) {
        dfree(E->rec_i); dfree(E->seis); dfree(E->trace_sums);
        E->rec_i = nullptr; E->seis = nullptr; E->trace_sums = nullptr;
        E->cap_tr = ntr;
        HIP_TRY(dmalloc(&E->rec_i, sizeof(int) * 3 * ntr));
        HIP_TRY(dmalloc(&E->seis, sizeof(float) * 6 * (size_t)ntr * c.NT));
        HIP_TRY(dmalloc(&E->trace_sums, sizeof(double) * 2 * ntr));
        HIP_TRY(hipMemsetAsync(E->seis, 0, sizeof(float) * 6 * (size_t)ntr * c.NT, E->stream));
        drop_graphs(E);
    }
    if (nsrc != E->nsrc || ntr != E->ntr || nshots != E->nslots) drop_graphs(E);
    E->nsrc = nsrc;
    E->ntr = ntr;
    E->nslots = nshots;
    E->src_off = so;
    E->tr_off = to;
    E->shot_misfit.assign(2 * (size_t)nshots, 0.0);
    E->first_src.assign(2 * (size_t)nshots, 0);
    for (int k = 0; k < nshots; k++)
        if (nsrc_of[k]) {
            E->first_src[2 * k] = (int)srcpos[(size_t)so[k] * IFOS_SRC_STRIDE];
            E->first_src[2 * k + 1] = (int)srcpos[(size_t)so[k] * IFOS_SRC_STRIDE + 1];
        }
    // SOURCE_TYPE is rewritten by update_v_PML from the last source before
    // psource consults it (src/update_v_PML.c:558, src/IFOS2D.c:1758): per
    // shot, the last source's type decides whether explosive sources act
    std::vector<int> expl(nshots, 0);
    for (int k = 0; k < nshots; k++)
        expl[k] = nsrc_of[k] ? ((int)srcpos[(size_t)(so[k + 1] - 1) * IFOS_SRC_STRIDE + 7]) == 1 : c.SOURCE_TYPE == 1;
    HIP_TRY(hipMemcpyAsync(E->explosive, expl.data(), sizeof(int) * nshots, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    if (nsrc) {
        std::vector<int> si(3 * (size_t)E->cap_src, 0);
        std::vector<double> sd(2 * (size_t)E->cap_src, 0.0);
        for (int l = 0; l < nsrc; l++) {
            const float *sp = srcpos + (size_t)l * IFOS_SRC_STRIDE;
            si[l] = (int)sp[0];
            si[E->cap_src + l] = (int)sp[1];
            si[2 * E->cap_src + l] = (int)sp[7];
            float azi_rad = sp[6] * 3.141592653589793 / 180;  // src/update_v_PML.c:556
            sd[l] = sin(azi_rad);
            sd[E->cap_src + l] = cos(azi_rad);
        }
        std::vector<float> w((size_t)nsrc * c.NT);
        ifos_wavelet(&c, nsrc, srcpos, w.data());
        HIP_TRY(hipMemcpyAsync(E->src_i, si.data(), sizeof(int) * si.size(), hipMemcpyHostToDevice, E->stream));
        HIP_TRY(hipMemcpyAsync(E->src_d, sd.data(), sizeof(double) * sd.size(), hipMemcpyHostToDevice, E->stream));
        HIP_TRY(hipMemcpyAsync(E->sig, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, E->stream));
        filter_traces(E, E->sig, nsrc, c.SOURCE_SHAPE != 6);
        HIP_TRY(hipStreamSynchronize(E->stream));
        E->sig_host.swap(w);
        E->sig_given = false;
    }
    if (ntr) {
        std::vector<int> ri(3 * (size_t)E->cap_tr, 0);
        for (int t = 0; t < ntr; t++) {
            ri[t] = recpos[2 * t];
            ri[E->cap_tr + t] = recpos[2 * t + 1];
        }
        HIP_TRY(hipMemcpyAsync(E->rec_i, ri.data(), sizeof(int) * ri.size(), hipMemcpyHostToDevice, E->stream));
        HIP_TRY(hipStreamSynchronize(E->stream));
    }
    // bucket each slot's sources and receivers by 64x8 tile, keeping their order
    const int nblk = E->nbx * E->nby;
    auto bucket = [&](const std::vector<int> &off, auto x_of, auto y_of, std::vector<int> &start,
                      std::vector<int> &idx) {
        const int n = off[nshots];
        start.assign((size_t)nshots * (nblk + 1), 0);
        idx.assign(n > 0 ? n : 1, 0);
        int pos = 0;
        for (int k = 0; k < nshots; k++) {
            int *st = start.data() + (size_t)k * (nblk + 1);
            std::vector<int> tile(off[k + 1] - off[k]);
            for (int q = off[k]; q < off[k + 1]; q++) {
                tile[q - off[k]] = ((y_of(q) - 1) / kTileY) * E->nbx + (x_of(q) - 1) / kTileX;
                st[tile[q - off[k]] + 1]++;
            }
            st[0] = pos;
            for (int b = 0; b < nblk; b++) st[b + 1] += st[b];
            std::vector<int> fill(st, st + nblk);
            for (int q = off[k]; q < off[k + 1]; q++) idx[fill[tile[q - off[k]]]++] = q;
            pos = st[nblk];
        }
    };
    std::vector<int> s_start, s_idx, r_start, r_idx;
    bucket(so, [&](int q) { return (int)srcpos[(size_t)q * IFOS_SRC_STRIDE]; },
           [&](int q) { return (int)srcpos[(size_t)q * IFOS_SRC_STRIDE + 1]; }, s_start, s_idx);
    bucket(to, [&](int q) { return recpos[2 * q]; }, [&](int q) { return recpos[2 * q + 1]; }, r_start, r_idx);
    // the same per tile for the tile and its N-cell rim (k_fused_step): a
    // point is listed for every tile whose rim reaches it, in point order
    auto bucket_rim = [&](const std::vector<int> &off, auto x_of, auto y_of, std::vector<int> &start,
                          std::vector<int> &idx) {
        const int R = E->N;
        start.assign((size_t)nshots * (nblk + 1), 0);
        std::vector<std::vector<int>> per(nblk);
        idx.clear();
        for (int k = 0; k < nshots; k++) {
            for (auto &v : per) v.clear();
            for (int q = off[k]; q < off[k + 1]; q++) {
                const int x = x_of(q), y = y_of(q);
                const int bx0 = std::max(0, (x - 1 - R) / kTileX), bx1 = std::min(E->nbx - 1, (x - 1 + R) / kTileX);
                const int by0 = std::max(0, (y - 1 - R) / kTileY), by1 = std::min(E->nby - 1, (y - 1 + R) / kTileY);
                for (int by = by0; by <= by1; by++)
                    for (int bx = bx0; bx <= bx1; bx++) per[by * E->nbx + bx].push_back(q);
            }
            int *st = start.data() + (size_t)k * (nblk + 1);
            st[0] = (int)idx.size();
            for (int b = 0; b < nblk; b++) {
                idx.insert(idx.end(), per[b].begin(), per[b].end());
                st[b + 1] = (int)idx.size();
            }
        }
        if (idx.empty()) idx.push_back(0);
    };
    std::vector<int> sx_start, sx_idx, rx_start, rx_idx;
    bucket_rim(so, [&](int q) { return (int)srcpos[(size_t)q * IFOS_SRC_STRIDE]; },
               [&](int q) { return (int)srcpos[(size_t)q * IFOS_SRC_STRIDE + 1]; }, sx_start, sx_idx);
    bucket_rim(to, [&](int q) { return recpos[2 * q]; }, [&](int q) { return recpos[2 * q + 1]; }, rx_start, rx_idx);
    std::vector<int> blob;
    E->off_src_start = 0;
    blob.insert(blob.end(), s_start.begin(), s_start.end());
    E->off_src_idx = (int)blob.size();
    blob.insert(blob.end(), s_idx.begin(), s_idx.end());
    E->off_rec_start = (int)blob.size();
    blob.insert(blob.end(), r_start.begin(), r_start.end());
    E->off_rec_idx = (int)blob.size();
    blob.insert(blob.end(), r_idx.begin(), r_idx.end());
    E->off_srcx_start = (int)blob.size();
    blob.insert(blob.end(), sx_start.begin(), sx_start.end());
    E->off_srcx_idx = (int)blob.size();
    blob.insert(blob.end(), sx_idx.begin(), sx_idx.end());
    E->off_recx_start = (int)blob.size();
    blob.insert(blob.end(), rx_start.begin(), rx_start.end());
    E->off_recx_idx = (int)blob.size();
    blob.insert(blob.end(), rx_idx.begin(), rx_idx.end());
    if (blob.size() > E->tiles_cap) {
        dfree(E->tiles);
        E->tiles = nullptr;
        E->tiles_cap = blob.size() * 2;
        HIP_TRY(dmalloc(&E->tiles, sizeof(int) * E->tiles_cap));
        drop_graphs(E);
    }
    HIP_TRY(hipMemcpyAsync(E->tiles, blob.data(), sizeof(int) * blob.size(), hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

extern "C" int ifos_engine_set_wavelets(ifos_engine *E, const float *sig)
{
    HIP_TRY(hipSetDevice(E->device));
    if (!E->nsrc) return ifos_set_error("set_wavelets needs shots with sources"), -2;
    const size_t n = (size_t)E->nsrc * E->c.NT;
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->sig_host.assign(sig, sig + n);
    E->sig_given = true;
    HIP_TRY(hipMemcpyAsync(E->sig, E->sig_host.data(), sizeof(float) * n, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->snap_valid = false;
    return 0;
}

extern "C" int ifos_engine_set_shot(ifos_engine *E, int nsrc, const float *srcpos, int ntr, const int *recpos)
{
    return ifos_engine_set_shots(E, 1, &nsrc, srcpos, &ntr, recpos);
}

// ---------------------------------------------------------------------------
// argument blocks
// ---------------------------------------------------------------------------
static StepGeom geom(const ifos_engine *E)
{
    const ifos_config &c = E->c;
    StepGeom g;
    g.NX = c.NX; g.NY = c.NY; g.pitch = E->pitch;
    g.FW = c.FW; g.FREE_SURF = c.FREE_SURF; g.BOUNDARY = c.BOUNDARY;
    g.N = E->N; g.DT = c.DT; g.DH = c.DH;
    for (int k = 0; k < 7; k++) g.hc[k] = E->hc[k];
    const int fw2 = c.FW > 0 ? 2 * c.FW : 1;
    const float *P = E->pml;
    g.pml = {P, P + fw2, P + 2 * fw2, P + 3 * fw2, P + 4 * fw2, P + 5 * fw2,
             P + 6 * fw2, P + 7 * fw2, P + 8 * fw2, P + 9 * fw2, P + 10 * fw2, P + 11 * fw2};
    g.px = fw2;
    g.py = c.NX;
    g.K_one = E->K_one;
    return g;
}

static float *seis_buf(const ifos_engine *E, int k) { return E->seis + (size_t)k * E->cap_tr * E->c.NT; }

// argument block of the fused step kernels for this engine
static StepArgs step_args(ifos_engine *E, int sw, bool snapshots)
{
    const ifos_config &c = E->c;
    StepArgs a;
    memset(&a, 0, sizeof a);
    a.g = geom(E);
    a.L = c.L; a.VELOCITY = c.VELOCITY; a.ADJOINT_TYPE = c.ADJOINT_TYPE; a.NT = c.NT; a.DTINV = c.DTINV;
    a.sw = sw;
    a.plane = E->plane;
    a.vx = E->vx; a.vy = E->vy; a.sxx = E->sxx; a.syy = E->syy; a.sxy = E->sxy;
    a.vxp1 = E->vxp1; a.vyp1 = E->vyp1;
    a.r = E->r; a.p = E->p; a.q = E->q;
    a.ws = E->ws;
    a.ls = E->field;
    a.rip = E->rip; a.rjp = E->rjp;
    a.f = E->f; a.gl = E->g; a.uipjp = E->uipjp;
    a.mu = E->mu_r; a.pi = E->pi_r; a.muipjp = E->muipjp_r; a.taus = E->taus; a.taup = E->taup;
    a.tausipjp = E->tausipjp;
    for (int l = 0; l <= kMaxL; l++) {
        const bool on = l >= 1 && l <= c.L;
        a.eta[l] = on ? E->mat.eta[l] : 0.0f;
        a.bip[l] = on ? E->mat.bip[l] : 0.0f;
        a.bjm[l] = on ? E->mat.bjm[l] : 0.0f;
        a.cip[l] = on ? E->mat.cip[l] : 0.0f;
        a.cjm[l] = on ? E->mat.cjm[l] : 0.0f;
    }
    a.psi_sxx_x = E->psi_sxx_x; a.psi_sxy_x = E->psi_sxy_x; a.psi_syy_y = E->psi_syy_y; a.psi_sxy_y = E->psi_sxy_y;
    a.psi_vxx = E->psi_vxx; a.psi_vyx = E->psi_vyx; a.psi_vyy = E->psi_vyy; a.psi_vxy = E->psi_vxy;
    a.psi_vxxs = E->psi_vxxs;
    a.pss = E->pss;
    a.vx2 = E->vx2; a.vy2 = E->vy2; a.sxx2 = E->sxx2; a.syy2 = E->syy2; a.sxy2 = E->sxy2;
    a.psi2_sxx_x = E->psi2_sxx_x; a.psi2_sxy_x = E->psi2_sxy_x; a.psi2_syy_y = E->psi2_syy_y;
    a.psi2_sxy_y = E->psi2_sxy_y;
    TileList src;
    src.start = E->tiles + E->off_src_start; src.idx = E->tiles + E->off_src_idx;
    src.ix = E->src_i; src.iy = E->src_i + E->cap_src; src.type = E->src_i + 2 * E->cap_src;
    src.sin_azi = E->src_d; src.cos_azi = E->src_d + E->cap_src;
    src.sig = E->sig; src.sig1 = E->sig;
    TileList rec;
    rec.start = E->tiles + E->off_rec_start; rec.idx = E->tiles + E->off_rec_idx;
    rec.ix = E->rec_i; rec.iy = E->rec_i + E->cap_tr; rec.type = nullptr;
    rec.sin_azi = rec.cos_azi = nullptr;
    rec.sig = seis_buf(E, 2); rec.sig1 = seis_buf(E, 3);  // adjoint sources
    a.inj = sw == 0 ? src : rec;
    a.exp = src;
    a.rec = rec;
    TileList srcx = src, recx = rec;
    srcx.start = E->tiles + E->off_srcx_start; srcx.idx = E->tiles + E->off_srcx_idx;
    recx.start = E->tiles + E->off_recx_start; recx.idx = E->tiles + E->off_recx_idx;
    a.injx = sw == 0 ? srcx : recx;
    a.explosive = E->explosive;
    a.snap_prec = E->snap_prec;
    if (snapshots || sw == 1) {
        const size_t per = (size_t)E->nsnap * E->plane;
        const size_t ntile = (size_t)E->nbx * E->nby;
        a.sns = (long)(5 * per);
        a.scs = (long)(5 * E->nsnap * ntile);
        for (int k = 0; k < 5; k++) {
            if (E->snap_prec == 0) {
                a.snap[k] = (float *)E->snap + k * per;
            } else {
                // scales of all slots first (rounded up to 256 bytes), then the codes
                const size_t sbytes = (sizeof(float) * 5 * E->nsnap * ntile * E->slots + 255) / 256 * 256;
                a.sscale[k] = (float *)E->snap + k * E->nsnap * ntile;
                a.snap16[k] = (uint16_t *)((char *)E->snap + sbytes) + k * per;
            }
        }
    }
    a.muss = E->muss; a.lamss = E->lamss; a.vs_i = E->vs_i;
    a.wc = E->wcbuf; a.wc_u = E->wcbuf + E->plane; a.wc_rho = E->wcbuf + 2 * E->plane;
    a.wcs = 3 * E->plane;
    a.seis_vx = seis_buf(E, 0); a.seis_vy = seis_buf(E, 1);
    // energies for the preconditioner in gradient runs: Ws in the forward
    // run that stores snapshots, Wr in the adjoint (EPRECOND 1 only)
    if (E->ep_mode && ((sw == 0 && snapshots) || (sw == 1 && E->ep_mode == 1))) {
        if (E->ep_fp64) a.wacc64 = E->wbuf64 + (sw == 0 ? 0 : E->plane);
        else a.wacc = E->wbuf + (sw == 0 ? 0 : E->plane);
        a.wacs = 2 * E->plane;
    }
    a.base = E->ctl;
    a.nan_flag = E->ctl + 1;
    a.nbx = E->nbx;
    a.nby = E->nby;
    a.nslots = E->nslots;
    a.swizzle = E->swizzle;
    return a;
}

// copy an argument block into device slot idx (0 forward, 1 forward with
// snapshots, 2 adjoint, 3 eager runs and timing); kernels queued earlier on
// the stream have finished when it returns, so a slot can be rewritten
static int upload_args(ifos_engine *E, const StepArgs &a, int idx, const StepArgs **out)
{
    HIP_TRY(hipMemcpyAsync(E->dargs + idx, &a, sizeof a, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    *out = E->dargs + idx;
    return 0;
}

// Captured graphs pay where launching is the bottleneck: below about 2^19
// cell-slots a step's two kernels take less time than the host needs to
// issue them.  Above it eager launches keep up, and they ran 2 % faster on
// the 2000x800 grid (430 against 440 ms per shot gradient, each kernel a
// little shorter than inside a graph).
constexpr long kGraphCells = 1L << 19;

static bool use_graphs(const ifos_engine *E)
{
    if (E->graph_mode >= 0) return E->graph_mode == 1;
    return (long)E->c.NX * E->c.NY * E->nslots < kGraphCells;
}

// NT steps as graph units of `unit` steps: kernels take their position k in
// the unit and read the number of steps already done from a device counter,
// which a one-thread kernel advances at the end of each unit.
static int run_loop(ifos_engine *E, int sw, bool snapshots)
{
    const ifos_config &c = E->c;
    const int NT = c.NT;
    int unit = c.DTINV;
    while (unit < 32 && NT % (unit * 2) == 0) unit *= 2;
    const int kind = sw;
    // the one-kernel step takes one slot per work item
    LaunchShape ls = E->shape;
    if (E->fused) ls.spb = 1;
    const int key[6] = {E->nsrc, E->ntr,
                        (snapshots ? 1 : 0) + 2 * (int)(ls.cpml_weight * 4.0f) + 64 * ls.rows_v +
                            1024 * ls.lds_v + 4096 * (E->fused ? 1 : 0),
                        E->nslots,
                        ((ls.order * 8 + ls.spb) * 4096 + ls.grid_v) * 4096 + ls.grid, unit};
    StepArgs a = shaped_args(step_args(E, sw, snapshots), ls);
    const StepArgs *da = nullptr;
    if (use_graphs(E)) {
        ENG_TRY(upload_args(E, a, sw == 1 ? 2 : (snapshots ? 1 : 0), &da));
        if (!E->gexec[kind] || E->gsteps[kind] != unit || memcmp(E->gkey[kind], key, sizeof key)) {
            if (E->gexec[kind]) hipGraphExecDestroy(E->gexec[kind]);
            E->gexec[kind] = nullptr;
            hipGraph_t graph;
            std::lock_guard<std::mutex> capture_guard(device_lock());
            HIP_TRY(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < unit; k++) {
                if (E->fused) {
                    launch_fused_step(a, da, ls, k, E->stream);
                } else {
                    launch_vel_step(a, da, ls, k, E->stream);
                    launch_str_step(a, da, ls, k, E->stream);
                }
            }
            launch_tick(E->ctl, unit, E->stream);
            HIP_TRY(hipStreamEndCapture(E->stream, &graph));
            hipError_t ie = hipGraphInstantiate(&E->gexec[kind], graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                E->gexec[kind] = nullptr;
                ifos_set_error("graph instantiation failed: %s", hipGetErrorString(ie));
                return -1;
            }
            E->gsteps[kind] = unit;
            memcpy(E->gkey[kind], key, sizeof key);
        }
        for (int u = 0; u < NT / unit; u++) HIP_TRY(hipGraphLaunch(E->gexec[kind], E->stream));
    } else {
        a.base = nullptr;  // the step number is passed directly
        ENG_TRY(upload_args(E, a, 3, &da));
        for (int nt = 1; nt <= NT; nt++) {
            if (E->fused) {
                launch_fused_step(a, da, ls, nt, E->stream);
            } else {
                launch_vel_step(a, da, ls, nt, E->stream);
                launch_str_step(a, da, ls, nt, E->stream);
            }
        }
    }
    E->last_set = E->fused ? (NT & 1) : 0;
    HIP_TRY(hipGetLastError());
    return 0;
}

static int zero_state(ifos_engine *E)
{
    HIP_TRY(hipMemsetAsync(E->slab, 0, sizeof(float) * E->ws * E->nslots, E->stream));
    HIP_TRY(hipMemsetAsync(E->psi, 0, sizeof(float) * E->pss * E->nslots, E->stream));
    HIP_TRY(hipMemsetAsync(E->ctl, 0, sizeof(int) * 2, E->stream));
    return 0;
}

static size_t snapshot_bytes(const ifos_config &c, int prec)
{
    const size_t plane = (size_t)c.NX * c.NY, nsnap = (size_t)(c.NT / c.DTINV);
    const size_t ntile = (size_t)((c.NX + kTileX - 1) / kTileX) * ((c.NY + kTileY - 1) / kTileY);
    if (prec == 0) return sizeof(float) * 5 * nsnap * plane;
    return sizeof(uint16_t) * 5 * nsnap * plane + sizeof(float) * 5 * nsnap * ntile;
}

static int ensure_snapshots(ifos_engine *E)
{
    const size_t need = snapshot_bytes(E->c, E->snap_prec) * E->slots + 256;
    if (E->snap && E->snap_bytes >= need) return 0;
    dfree(E->snap);
    E->snap = nullptr;
    drop_graphs(E);
    if (dmalloc(&E->snap, need) != hipSuccess) {
        ifos_set_error("cannot allocate %.2f GB of snapshot storage", need / 1e9);
        return -3;
    }
    E->snap_bytes = need;
    return 0;
}

static int check_nan(ifos_engine *E)
{
    int h[2];
    HIP_TRY(hipMemcpyAsync(h, E->ctl, sizeof h, hipMemcpyDeviceToHost, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    if (h[1]) {
        ifos_set_error("Simulation is unstable ! (NaN at the probe cell in time step %d)", h[1]);
        return -4;
    }
    return 0;
}

extern "C" int ifos_engine_forward(ifos_engine *E, int store_snapshots)
{
    HIP_TRY(hipSetDevice(E->device));
    if (!E->model_set) return ifos_set_error("set_model must be called before forward"), -2;
    if (!E->tiles) return ifos_set_error("set_shot must be called before forward"), -2;
    if (store_snapshots) ENG_TRY(ensure_snapshots(E));
    E->snap_valid = false;
    E->slot_grad_valid = false;
    ENG_TRY(zero_state(E));
    // energies of this shot's wavefields start at zero (src/IFOS2D.c:1623-1631)
    if (E->ep_mode && store_snapshots)
        HIP_TRY(hipMemsetAsync(E->wbuf, 0, sizeof(float) * 2 * E->plane * E->nslots, E->stream));
    if (E->ep_mode && store_snapshots && E->ep_fp64)
        HIP_TRY(hipMemsetAsync(E->wbuf64, 0, sizeof(double) * 2 * E->plane * E->nslots, E->stream));
    ENG_TRY(run_loop(E, 0, store_snapshots != 0));
    ENG_TRY(check_nan(E));
    E->snap_valid = store_snapshots != 0;
    return 0;
}

extern "C" int ifos_engine_seismograms(ifos_engine *E, float *vx, float *vy)
{
    HIP_TRY(hipSetDevice(E->device));
    const size_t n = (size_t)E->ntr * E->c.NT;
    if (vx) HIP_TRY(hipMemcpyAsync(vx, seis_buf(E, 0), n * sizeof(float), hipMemcpyDeviceToHost, E->stream));
    if (vy) HIP_TRY(hipMemcpyAsync(vy, seis_buf(E, 1), n * sizeof(float), hipMemcpyDeviceToHost, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

extern "C" int ifos_engine_fields(ifos_engine *E, float *out5)
{
    HIP_TRY(hipSetDevice(E->device));
    float *set0[5] = {E->vx, E->vy, E->sxx, E->syy, E->sxy}, *set1[5] = {E->vx2, E->vy2, E->sxx2, E->syy2, E->sxy2};
    float **src = E->last_set ? set1 : set0;
    for (int k = 0; k < 5; k++)
        HIP_TRY(hipMemcpy2DAsync(out5 + (size_t)k * E->plane, sizeof(float) * E->c.NX, src[k] + pad_index(E, 1, 1),
                                 sizeof(float) * E->pitch, sizeof(float) * E->c.NX, E->c.NY, hipMemcpyDeviceToHost,
                                 E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

extern "C" int ifos_engine_residual(ifos_engine *E, const float *obs_vx, const float *obs_vy, const int *kill,
                                    int obs_on_device, double *misfit)
{
    HIP_TRY(hipSetDevice(E->device));
    const ifos_config &c = E->c;
    const int ntr = E->ntr;
    const size_t n = (size_t)ntr * c.NT;
    const float *ovx = obs_vx, *ovy = obs_vy;
    // an active component (ADJOINT_TYPE 1: both, 2: vy, 3: vx) needs its
    // observed traces; a missing one would leave the residual reading the
    // engine's copy from an earlier call
    if (ntr > 0 && !obs_vx && (c.ADJOINT_TYPE == 1 || c.ADJOINT_TYPE == 3))
        return ifos_set_error("observed vx traces missing (ADJOINT_TYPE %d uses them)", c.ADJOINT_TYPE), -2;
    if (ntr > 0 && !obs_vy && (c.ADJOINT_TYPE == 1 || c.ADJOINT_TYPE == 2))
        return ifos_set_error("observed vy traces missing (ADJOINT_TYPE %d uses them)", c.ADJOINT_TYPE), -2;
    if (!obs_on_device) {
        if (obs_vx) HIP_TRY(hipMemcpyAsync(seis_buf(E, 4), obs_vx, n * sizeof(float), hipMemcpyHostToDevice, E->stream));
        if (obs_vy) HIP_TRY(hipMemcpyAsync(seis_buf(E, 5), obs_vy, n * sizeof(float), hipMemcpyHostToDevice, E->stream));
        ovx = seis_buf(E, 4);
        ovy = seis_buf(E, 5);
    }
    if (E->filt_nlow || E->filt_nhigh) {
        // observed traces are filtered like the reference's sectionread
        // (src/IFOS2D.c:1985-2071), in the engine's copy
        if (obs_on_device) {
            if (obs_vx) HIP_TRY(hipMemcpyAsync(seis_buf(E, 4), obs_vx, n * sizeof(float), hipMemcpyDeviceToDevice, E->stream));
            if (obs_vy) HIP_TRY(hipMemcpyAsync(seis_buf(E, 5), obs_vy, n * sizeof(float), hipMemcpyDeviceToDevice, E->stream));
            ovx = seis_buf(E, 4);
            ovy = seis_buf(E, 5);
        }
        if (obs_vx) filter_traces(E, seis_buf(E, 4), ntr, true);
        if (obs_vy) filter_traces(E, seis_buf(E, 5), ntr, true);
        HIP_TRY(hipGetLastError());
    }
    const int *dkill = nullptr;
    if (c.TRKILL && kill) {
        HIP_TRY(hipMemcpyAsync(E->rec_i + 2 * E->cap_tr, kill, sizeof(int) * ntr, hipMemcpyHostToDevice, E->stream));
        dkill = E->rec_i + 2 * E->cap_tr;
    }
    double l2 = 0.0, en = 0.0;
    std::vector<double> h(2 * (size_t)ntr);
    std::fill(E->shot_misfit.begin(), E->shot_misfit.end(), 0.0);
    for (int comp = 0; comp < 2; comp++) {
        float *diff = seis_buf(E, 2 + comp);
        const bool active = comp == 0 ? (c.ADJOINT_TYPE == 1 || c.ADJOINT_TYPE == 3)
                                      : (c.ADJOINT_TYPE == 1 || c.ADJOINT_TYPE == 2);
        if (!active || ntr == 0) {
            if (n) HIP_TRY(hipMemsetAsync(diff, 0, n * sizeof(float), E->stream));
            continue;
        }
        ResArgs ra;
        ra.ntr = ntr; ra.NT = c.NT; ra.VELOCITY = c.VELOCITY; ra.DT = c.DT;
        ra.syn = seis_buf(E, comp); ra.obs = comp == 0 ? ovx : ovy; ra.kill = dkill; ra.diff = diff;
        ra.l2_trace = E->trace_sums; ra.en_trace = E->trace_sums + E->cap_tr;
        launch_residual(ra, E->stream);
        HIP_TRY(hipMemcpyAsync(h.data(), E->trace_sums, sizeof(double) * ntr, hipMemcpyDeviceToHost, E->stream));
        HIP_TRY(hipMemcpyAsync(h.data() + ntr, E->trace_sums + E->cap_tr, sizeof(double) * ntr,
                               hipMemcpyDeviceToHost, E->stream));
        HIP_TRY(hipStreamSynchronize(E->stream));
        for (int k = 0; k < E->nslots; k++)
            for (int t = E->tr_off[k]; t < E->tr_off[k + 1]; t++) {
                l2 += h[t];
                en += h[ntr + t];
                E->shot_misfit[2 * k] += h[t];
                E->shot_misfit[2 * k + 1] += h[ntr + t];
            }
    }
    if (misfit) {
        misfit[0] = l2;
        misfit[1] = en;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// energy preconditioner (src/eprecond1.c:8-69, src/IFOS2D.c:2615-2654, 2737-2751)
// ---------------------------------------------------------------------------
// We of each slot, written over its Ws plane: sqrt(Ws*Wr) (mode 1) or
// sqrt(Ws) times the geometric receiver term (mode 3)
__global__ void k_we_form(long n, int nslots, int mode, float *wbuf_, const double *geom_)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    for (int s = 0; s < nslots; s++) {
        auto W = gmem(wbuf_) + (long)s * 2 * n;
        if (mode == 1) W[e] = (float)sqrt((double)(W[e] * W[n + e]));
        else W[e] = (float)(sqrt((double)W[e]) * gmem(geom_)[e]);
    }
}

// We of every slot from FP64 sums: mode 1 sqrt(Ws*Wr), mode 3 sqrt(Ws) times
// the geometric receiver term, formed in FP64 and rounded to FP32 into wbuf
__global__ void k_we_form64(long n, int nslots, int mode, const double *w64_, float *wbuf_, const double *geom_)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    for (int s = 0; s < nslots; s++) {
        const auto W = gmem(w64_) + (long)s * 2 * n;
        auto out = gmem(wbuf_) + (long)s * 2 * n;
        if (mode == 1) out[e] = (float)sqrt(W[e] * W[n + e]);
        else out[e] = (float)(sqrt(W[e]) * gmem(geom_)[e]);
    }
}

// per-block maxima of |W| of one slot's We plane
__global__ __launch_bounds__(256) void k_absmax_part(const float *w_, long n, float *part_)
{
    __shared__ float lds[4];
    float m = 0.0f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(gmem(w_)[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (threadIdx.x % 64 == 0) lds[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) gmem(part_)[blockIdx.x] = fmaxf(fmaxf(lds[0], lds[1]), fmaxf(lds[2], lds[3]));
}

// We + epsilon*max(We) per slot (eps from the device array, one per slot)
__global__ void k_we_add(long n, int nslots, float *wbuf_, const float *eps_)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    for (int s = 0; s < nslots; s++) {
        auto W = gmem(wbuf_) + (long)s * 2 * n;
        W[e] = W[e] + gmem(eps_)[s];
    }
}

// summed gradient times We_sum / max(We_sum); We_sum back to zero
__global__ void k_apply_wesum(long n, float *we_sum_, float wmax, float *gvp_, float *gvs_, float *grho_)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    auto ws = gmem(we_sum_);
    const float w = ws[e] / wmax;
    gmem(gvp_)[e] = gmem(gvp_)[e] * w;
    gmem(gvs_)[e] = gmem(gvs_)[e] * w;
    gmem(grho_)[e] = gmem(grho_)[e] * w;
    ws[e] = 0.0f;
}

static constexpr int kMaxParts = 512;

static int plane_absmax(ifos_engine *E, const float *w, float *out)
{
    hipLaunchKernelGGL(k_absmax_part, dim3(kMaxParts), dim3(256), 0, E->stream, w, E->plane, E->red);
    HIP_TRY(hipGetLastError());
    std::vector<float> h(kMaxParts);
    HIP_TRY(hipMemcpyAsync(h.data(), E->red, sizeof(float) * kMaxParts, hipMemcpyDeviceToHost, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    float m = 0.0f;
    for (float v : h) m = fmaxf(m, v);
    *out = m;
    return 0;
}

static int energy_weights(ifos_engine *E, GradArgs &ga)
{
    const long n = E->plane;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (E->ep_fp64)
        hipLaunchKernelGGL(k_we_form64, dim3(nb), dim3(256), 0, E->stream, n, E->nslots, E->ep_mode, E->wbuf64,
                           E->wbuf, E->ep_geom);
    else
        hipLaunchKernelGGL(k_we_form, dim3(nb), dim3(256), 0, E->stream, n, E->nslots, E->ep_mode, E->wbuf,
                           E->ep_geom);
    HIP_TRY(hipGetLastError());
    std::vector<float> eps(E->nslots), wmax(E->nslots);
    for (int s = 0; s < E->nslots; s++) {
        float m;
        ENG_TRY(plane_absmax(E, E->wbuf + (long)s * 2 * n, &m));
        eps[s] = E->ep_eps * m;  // src/eprecond1.c:63-68
    }
    HIP_TRY(hipMemcpyAsync(E->we_max, eps.data(), sizeof(float) * E->nslots, hipMemcpyHostToDevice, E->stream));
    hipLaunchKernelGGL(k_we_add, dim3(nb), dim3(256), 0, E->stream, n, E->nslots, E->wbuf, E->we_max);
    HIP_TRY(hipGetLastError());
    if (E->ep_per_shot) {
        // global_maximum of the regularised weights (src/IFOS2D.c:2621)
        for (int s = 0; s < E->nslots; s++) ENG_TRY(plane_absmax(E, E->wbuf + (long)s * 2 * n, &wmax[s]));
        HIP_TRY(hipMemcpyAsync(E->we_max, wmax.data(), sizeof(float) * E->nslots, hipMemcpyHostToDevice, E->stream));
        HIP_TRY(hipStreamSynchronize(E->stream));
    }
    ga.we = E->wbuf;
    ga.wes = 2 * n;
    ga.we_max = E->we_max;
    ga.per_shot = E->ep_per_shot;
    ga.we_sum = E->we_sum;
    return 0;
}

extern "C" int ifos_engine_set_eprecond(ifos_engine *E, int mode, float epsilon_we, int per_shot)
{
    if (mode != 0 && mode != 1 && mode != 3) return ifos_set_error("EPRECOND must be 0, 1 or 3"), -2;
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->ep_eps = epsilon_we;
    E->ep_per_shot = per_shot ? 1 : 0;
    if (mode == E->ep_mode) return 0;
    drop_graphs(E);
    E->ep_mode = mode;
    if (!mode || E->wbuf) return 0;
    const size_t n = (size_t)E->plane;
    if (dmalloc(&E->wbuf, sizeof(float) * 2 * n * E->slots) != hipSuccess ||
        dmalloc(&E->we_sum, sizeof(float) * n) != hipSuccess ||
        dmalloc(&E->ep_geom, sizeof(double) * n) != hipSuccess ||
        dmalloc(&E->we_max, sizeof(float) * E->slots) != hipSuccess ||
        dmalloc(&E->red, sizeof(float) * kMaxParts) != hipSuccess) {
        E->ep_mode = 0;
        return ifos_set_error("device allocation for the energy preconditioner failed"), -3;
    }
    HIP_TRY(hipMemsetAsync(E->we_sum, 0, sizeof(float) * n, E->stream));
    // geometric receiver term of mode 3 (src/eprecond1.c:46-48), FP64 on the host
    const ifos_config &c = E->c;
    std::vector<double> g(n);
    const float xmin = c.DH, xmax = c.NX * c.DH;
    for (int j = 1; j <= c.NY; j++)
        for (int i = 1; i <= c.NX; i++) {
            const float x = i * c.DH, y = j * c.DH;
            g[(size_t)(j - 1) * c.NX + (i - 1)] = asinh((xmax - x) / y) - asinh((xmin - x) / y);
        }
    HIP_TRY(hipMemcpyAsync(E->ep_geom, g.data(), sizeof(double) * n, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

extern "C" int ifos_engine_set_eprecond_fp64(ifos_engine *E, int on)
{
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipStreamSynchronize(E->stream));
    if ((on != 0) != (E->ep_fp64 != 0)) drop_graphs(E);
    E->ep_fp64 = on != 0;
    if (E->ep_fp64 && !E->wbuf64) {
        if (dmalloc(&E->wbuf64, sizeof(double) * 2 * (size_t)E->plane * E->slots) != hipSuccess) {
            E->ep_fp64 = 0;
            return ifos_set_error("device allocation of the FP64 energy sums failed"), -3;
        }
    }
    return 0;
}

extern "C" int ifos_engine_apply_eprecond(ifos_engine *E)
{
    HIP_TRY(hipSetDevice(E->device));
    if (!E->ep_mode || E->ep_per_shot) return 0;
    float m;
    ENG_TRY(plane_absmax(E, E->we_sum, &m));  // global_maximum (src/IFOS2D.c:2740)
    const unsigned nb = (unsigned)((E->plane + 255) / 256);
    hipLaunchKernelGGL(k_apply_wesum, dim3(nb), dim3(256), 0, E->stream, E->plane, E->we_sum, m, E->s_vp, E->s_vs,
                       E->s_rho);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

extern "C" int ifos_engine_energy_weight(ifos_engine *E, int which, float *out)
{
    HIP_TRY(hipSetDevice(E->device));
    if (!E->ep_mode) return ifos_set_error("energy preconditioning is off"), -2;
    HIP_TRY(hipMemcpyAsync(out, which ? E->we_sum : E->wbuf, sizeof(float) * E->plane, hipMemcpyDeviceToHost,
                           E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

// per-slot source cells and normalisations of the circular taper
static int shot_taper_args(ifos_engine *E, GradArgs &ga)
{
    const ifos_config &c = E->c;
    std::vector<float> lohi(2 * (size_t)E->nslots);
    for (int k = 0; k < E->nslots; k++) {
        const int is = E->first_src[2 * k], js = E->first_src[2 * k + 1];
        if (!is) return ifos_set_error("the source taper needs a source in every shot (shot %d has none)", k + 1), -2;
        ifos_srctaper_norm(c.NX, c.NY, E->tap_n, E->tap_host.data(), is, js, &lohi[2 * k], &lohi[2 * k + 1]);
    }
    HIP_TRY(hipMemcpyAsync(E->tap_ij, E->first_src.data(), sizeof(int) * 2 * E->nslots, hipMemcpyHostToDevice,
                           E->stream));
    HIP_TRY(hipMemcpyAsync(E->tap_lohi, lohi.data(), sizeof(float) * 2 * E->nslots, hipMemcpyHostToDevice,
                           E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    ga.tap_m = E->tap_m;
    ga.tap_ij = E->tap_ij;
    ga.tap_lohi = E->tap_lohi;
    ga.tap_n = E->tap_n;
    ga.tap_c = ifos_srctaper_centre(E->tap_n);
    ga.tap_fs = E->tap_fs;
    ga.NX = c.NX;
    ga.NY = c.NY;
    return 0;
}

extern "C" int ifos_time_filter(float *data, int ntr, int ns, float DT, float fc, int order, int method,
                                float F_HIGH_PASS)
{
    if (ntr < 0 || ns < 0 || (ntr && !data)) return ifos_set_error("ifos_time_filter: bad arguments"), -2;
    if (method != 1 && method != 2) return ifos_set_error("ifos_time_filter: method must be 1 or 2"), -2;
    if (order < 1 || order > 2 * kMaxFiltSections)
        return ifos_set_error("filter order %d outside 1..%d", order, 2 * kMaxFiltSections), -2;
    FiltArgs f;
    memset(&f, 0, sizeof f);
    f.data = data;
    f.ntr = ntr;
    f.ns = ns;
    f.stride = ns;
    // timedomain_filt: method 1 lowpass at fc (+ highpass at F_HIGH_PASS when
    // set), method 2 highpass at fc
    if (method == 1) f.nlow = ifos_butterworth(0, DT, (double)fc, order, &f.c[0][0], kMaxFiltSections);
    const float fh = method == 2 ? fc : F_HIGH_PASS;
    if (method == 2 || F_HIGH_PASS != 0.0f)
        f.nhigh = ifos_butterworth(1, DT, (double)fh, order, &f.c[f.nlow][0], kMaxFiltSections);
    if (f.nlow < 0 || f.nhigh < 0) return ifos_set_error("filter design failed"), -2;
    thread_local hipStream_t st = nullptr;
    if (!st) HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    launch_time_filter(f, st);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}

extern "C" int ifos_engine_set_shot_taper_file(ifos_engine *E, int slot, const float *vp, const float *vs,
                                               const float *rho)
{
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipStreamSynchronize(E->stream));
    if (!vp && !vs && !rho) {
        E->tf_on = false;
        return 0;
    }
    if (!vp || !vs || !rho) return ifos_set_error("per-shot taper files need the Vp, Vs and rho grids"), -2;
    if (slot < 0 || slot >= E->slots) return ifos_set_error("slot %d outside 0..%d", slot, E->slots - 1), -2;
    const size_t pb = sizeof(float) * E->plane;
    if (!E->tfbuf) {
        HIP_TRY(dmalloc(&E->tfbuf, 3 * pb * E->slots));
        // slots without a file of their own keep their gradient (coefficient 1)
        std::vector<float> ones((size_t)3 * E->plane * E->slots, 1.0f);
        HIP_TRY(hipMemcpyAsync(E->tfbuf, ones.data(), 3 * pb * E->slots, hipMemcpyHostToDevice, E->stream));
        HIP_TRY(hipStreamSynchronize(E->stream));
    }
    float *d = E->tfbuf + (size_t)slot * 3 * E->plane;
    HIP_TRY(hipMemcpyAsync(d, vp, pb, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipMemcpyAsync(d + E->plane, vs, pb, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipMemcpyAsync(d + 2 * E->plane, rho, pb, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->tf_on = true;
    return 0;
}

extern "C" int ifos_engine_set_time_filter(ifos_engine *E, float F_LOW_PASS, int ORDER, float F_HIGH_PASS)
{
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipStreamSynchronize(E->stream));
    if (F_LOW_PASS != 0.0f && (ORDER < 1 || ORDER > 2 * kMaxFiltSections))
        return ifos_set_error("filter order %d outside 1..%d", ORDER, 2 * kMaxFiltSections), -2;
    E->filt_nlow = E->filt_nhigh = 0;
    if (F_LOW_PASS != 0.0f) {
        // timedomain_filt method 1: T0 = 1/fc, lowpass, then highpass at F_HIGH_PASS when set
        const int nl = ifos_butterworth(0, E->c.DT, (double)F_LOW_PASS, ORDER, &E->filt_c[0][0], kMaxFiltSections);
        const int nh = F_HIGH_PASS != 0.0f ? ifos_butterworth(1, E->c.DT, (double)F_HIGH_PASS, ORDER,
                                                              &E->filt_c[nl][0], kMaxFiltSections)
                                           : 0;
        if (nl < 0 || nh < 0) return ifos_set_error("filter design failed"), -2;
        E->filt_nlow = nl;
        E->filt_nhigh = nh;
    }
    // wavelets of shots already set: filter afresh from the unfiltered ones
    if (E->nsrc && !E->sig_host.empty() && !E->sig_given) {
        HIP_TRY(hipMemcpyAsync(E->sig, E->sig_host.data(), sizeof(float) * E->sig_host.size(), hipMemcpyHostToDevice,
                               E->stream));
        filter_traces(E, E->sig, E->nsrc, E->c.SOURCE_SHAPE != 6);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(E->stream));
    }
    return 0;
}

extern "C" int ifos_engine_set_shot_taper(ifos_engine *E, int on, float SRTRADIUS, int SRTSHAPE, int FILTSIZE)
{
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->tap_on = 0;
    if (!on) return 0;
    if (!(SRTRADIUS > 0.0f)) return ifos_set_error("SRTRADIUS must be positive"), -2;
    const int n = ifos_srctaper_table(E->c.DH, SRTRADIUS, SRTSHAPE, nullptr, 0);
    if (n > 4096) return ifos_set_error("SRTRADIUS %g gives a taper of %d x %d cells", SRTRADIUS, n, n), -2;
    E->tap_host.assign((size_t)n * n, 0.0f);
    ifos_srctaper_table(E->c.DH, SRTRADIUS, SRTSHAPE, E->tap_host.data(), n * n);
    if (n > E->tap_n) {
        dfree(E->tap_m);
        E->tap_m = nullptr;
        E->tap_n = 0;
        HIP_TRY(dmalloc(&E->tap_m, sizeof(float) * n * n));
    }
    if (!E->tap_ij) {
        HIP_TRY(dmalloc(&E->tap_ij, sizeof(int) * 2 * E->slots));
        HIP_TRY(dmalloc(&E->tap_lohi, sizeof(float) * 2 * E->slots));
    }
    HIP_TRY(hipMemcpyAsync(E->tap_m, E->tap_host.data(), sizeof(float) * n * n, hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    E->tap_n = n;
    E->tap_fs = FILTSIZE;
    E->tap_on = 1;
    return 0;
}

extern "C" int ifos_engine_adjoint(ifos_engine *E)
{
    HIP_TRY(hipSetDevice(E->device));
    if (!E->snap || !E->snap_valid)
        return ifos_set_error("adjoint needs a forward run with snapshots on the current model and shot"), -2;
    if (!E->tiles || !E->model_set) return ifos_set_error("adjoint needs set_model and set_shot"), -2;
    E->slot_grad_valid = false;
    ENG_TRY(zero_state(E));
    HIP_TRY(hipMemsetAsync(E->wcbuf, 0, sizeof(float) * 3 * E->plane * E->nslots, E->stream));
    ENG_TRY(run_loop(E, 1, false));
    GradArgs ga;
    memset(&ga, 0, sizeof ga);
    if (E->ep_mode) ENG_TRY(energy_weights(E, ga));
    if (E->tap_on) ENG_TRY(shot_taper_args(E, ga));
    if (E->tf_on) {
        ga.tf = E->tfbuf;
        ga.tfs = 3 * E->plane;
    }
    ga.n = (int)E->plane; ga.DT = E->c.DT;
    ga.vp = E->vp_i; ga.vs = E->vs_i; ga.rho = E->rho_i;
    ga.wc = E->wcbuf; ga.wc_u = E->wcbuf + E->plane; ga.wc_rho = E->wcbuf + 2 * E->plane;
    ga.wcs = 3 * E->plane;
    ga.nslots = E->nslots;
    ga.g_vp = E->g_vp; ga.g_vs = E->g_vs; ga.g_rho = E->g_rho;
    ga.s_vp = E->s_vp; ga.s_vs = E->s_vs; ga.s_rho = E->s_rho;
    // each slot's gradient replaces its correlations (no longer needed)
    ga.slot_g = E->wcbuf;
    ga.slot_gs = 3 * E->plane;
    launch_gradient(ga, E->stream);
    E->slot_grad_valid = true;
    return check_nan(E);
}

extern "C" int ifos_engine_shot_gradient(ifos_engine *E, const float *obs_vx, const float *obs_vy, const int *kill,
                                         int obs_on_device, double *misfit)
{
    ENG_TRY(ifos_engine_forward(E, 1));
    ENG_TRY(ifos_engine_residual(E, obs_vx, obs_vy, kill, obs_on_device, misfit));
    return ifos_engine_adjoint(E);
}

extern "C" int ifos_engine_gradient(ifos_engine *E, int which, float *g_vp, float *g_vs, float *g_rho)
{
    HIP_TRY(hipSetDevice(E->device));
    const size_t pb = sizeof(float) * E->plane;
    float *s0 = which ? E->s_vp : E->g_vp, *s1 = which ? E->s_vs : E->g_vs, *s2 = which ? E->s_rho : E->g_rho;
    if (g_vp) HIP_TRY(hipMemcpyAsync(g_vp, s0, pb, hipMemcpyDeviceToHost, E->stream));
    if (g_vs) HIP_TRY(hipMemcpyAsync(g_vs, s1, pb, hipMemcpyDeviceToHost, E->stream));
    if (g_rho) HIP_TRY(hipMemcpyAsync(g_rho, s2, pb, hipMemcpyDeviceToHost, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

extern "C" int ifos_engine_slot_gradient(ifos_engine *E, int slot, float *g_vp, float *g_vs, float *g_rho)
{
    HIP_TRY(hipSetDevice(E->device));
    if (slot < 0 || slot >= E->nslots) return ifos_set_error("slot %d: the engine has %d shots set", slot, E->nslots), -2;
    if (!E->slot_grad_valid) return ifos_set_error("no adjoint run since the last forward run"), -2;
    const size_t pb = sizeof(float) * E->plane;
    const float *b = E->wcbuf + (long)slot * 3 * E->plane;
    if (g_vp) HIP_TRY(hipMemcpyAsync(g_vp, b, pb, hipMemcpyDeviceToHost, E->stream));
    if (g_vs) HIP_TRY(hipMemcpyAsync(g_vs, b + E->plane, pb, hipMemcpyDeviceToHost, E->stream));
    if (g_rho) HIP_TRY(hipMemcpyAsync(g_rho, b + 2 * E->plane, pb, hipMemcpyDeviceToHost, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

extern "C" int ifos_engine_set_fused(ifos_engine *E, int on)
{
    HIP_TRY(hipSetDevice(E->device));
    if (on && E->c.BOUNDARY)
        return ifos_set_error("the one-kernel step has no periodic sides (BOUNDARY 1): its tiles form their "
                              "neighbours' velocities from stresses beyond the halo copies"), -2;
    HIP_TRY(hipStreamSynchronize(E->stream));
    if ((on != 0) != E->fused) drop_graphs(E);
    E->fused = on != 0;
    return 0;
}

extern "C" int ifos_engine_get_fused(ifos_engine *E, int *on)
{
    *on = E->fused ? 1 : 0;
    return 0;
}

extern "C" int ifos_engine_reset_gradient(ifos_engine *E)
{
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipMemsetAsync(E->s_vp, 0, sizeof(float) * 3 * E->plane, E->stream));  // s_vp, s_vs, s_rho
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

extern "C" int ifos_engine_set_snapshot_precision(ifos_engine *E, int prec)
{
    if (prec < 0 || prec > 2) return ifos_set_error("snapshot precision must be 0 (FP32), 1 (BF16) or 2 (FP16)"), -2;
    HIP_TRY(hipSetDevice(E->device));
    if (prec == E->snap_prec) return 0;
    HIP_TRY(hipStreamSynchronize(E->stream));
    drop_graphs(E);
    dfree(E->snap);
    E->snap = nullptr;
    E->snap_bytes = 0;
    E->snap_valid = false;
    E->snap_prec = prec;
    return 0;
}

// device bytes an engine of `slots` slots allocates (the sizes of
// engine_create and set_shots; snapshots when `snapshots` != 0)
extern "C" size_t ifos_engine_bytes(const ifos_config *c, int slots, int snap_prec, int snapshots, int ntr_total,
                                    int nsrc_total)
{
    if (validate(c) || slots < 1) return 0;
    const int N = c->FDORDER / 2, H = N + 1, A = 16;
    const size_t pitch = (size_t)((A + c->NX + H) + 63) / 64 * 64;
    const size_t field = pitch * (size_t)(c->NY + 2 * H + 1);
    const size_t plane = (size_t)c->NX * c->NY;
    const int fw2 = c->FW > 0 ? 2 * c->FW : 1;
    size_t b = sizeof(float) * field * (size_t)(15 + 3 * c->L) * slots;   // wavefields, both sets
    b += sizeof(float) * field * (size_t)(15 + 3 * c->L);                 // materials
    b += sizeof(float) * plane * (11 + 3 * (size_t)slots);                // interior + correlation
    b += sizeof(float) * (size_t)(7 * c->NY * fw2 + 6 * fw2 * c->NX) * slots;  // CPML memory
    b += sizeof(float) * (size_t)6 * ntr_total * c->NT;                   // seismograms etc.
    b += (sizeof(float) * c->NT + 3 * sizeof(int) + 2 * sizeof(double)) * (size_t)nsrc_total;
    b += (3 * sizeof(int) + 2 * sizeof(double)) * (size_t)ntr_total;
    if (snapshots) b += snapshot_bytes(*c, snap_prec) * slots + 256;
    return b;
}

// Shots per launch whose per-step working set stays in the MALL (the 256 MB
// Infinity Cache between the L2s and HBM): each slot touches its velocities,
// stresses and memory variables every step, all slots share the buoyancies
// and moduli.  A budget of 160 MB leaves room for the snapshot and trace
// streams.  On the 1700x350, L=2 grid this gives 5; launches of 4-5 shots ran
// at 10.6 shot-gradients/s there, 8 at 8.9 and 16 at 8.7 (DRAM-bound), 1 at
// 6.8 (the chip underfilled).
extern "C" int ifos_engine_mall_slots(const ifos_config *c)
{
    if (validate(c)) return -2;
    constexpr double kMallBudget = 160e6;
    const int N = c->FDORDER / 2, H = N + 1, A = 16;
    const double pitch = (double)(((A + c->NX + H) + 63) / 64 * 64);
    const double field = 4.0 * pitch * (double)(c->NY + 2 * H + 1);
    const double per_slot = field * (5 + 3 * c->L), shared = field * (c->L ? 8 : 5);
    const double n = (kMallBudget - shared) / per_slot;
    return n < 1.0 ? 1 : (n > 4096.0 ? 4096 : (int)n);
}

extern "C" size_t ifos_snapshot_bytes(const ifos_config *c, int prec)
{
    if (validate(c) || prec < 0 || prec > 2) return 0;
    return snapshot_bytes(*c, prec);
}

extern "C" int ifos_engine_shot_misfit(ifos_engine *E, double *per_shot)
{
    for (int k = 0; k < E->nslots; k++) {
        per_shot[2 * k] = E->shot_misfit[2 * k];
        per_shot[2 * k + 1] = E->shot_misfit[2 * k + 1];
    }
    return 0;
}

extern "C" int ifos_engine_set_launch(ifos_engine *E, int spb, int grid)
{
    if (spb != 1 && spb != 2 && spb != 4) return ifos_set_error("shots per workgroup must be 1, 2 or 4"), -2;
    if (grid < 0 || grid % 8) return ifos_set_error("persistent grid must be 0 or a positive multiple of 8"), -2;
    E->shape.spb = spb;
    E->shape.grid = grid;
    E->shape.grid_v = grid;
    return 0;
}

extern "C" int ifos_engine_get_launch_grids(ifos_engine *E, int *grid_vel, int *grid_str)
{
    *grid_vel = E->shape.grid_v;
    *grid_str = E->shape.grid;
    return 0;
}

extern "C" int ifos_engine_set_graphs(ifos_engine *E, int mode)
{
    if (mode < -1 || mode > 1) return ifos_set_error("graph mode must be -1, 0 or 1"), -2;
    E->graph_mode = mode;
    return 0;
}

extern "C" int ifos_engine_get_graphs(ifos_engine *E, int *on)
{
    *on = use_graphs(E) ? 1 : 0;
    return 0;
}

extern "C" int ifos_engine_get_launch(ifos_engine *E, int *spb, int *grid)
{
    *spb = E->shape.spb;
    *grid = E->shape.grid;
    return 0;
}

extern "C" int ifos_engine_set_velocity_rows(ifos_engine *E, int rows)
{
    if (rows != 1 && rows != 2 && rows != 4) return ifos_set_error("velocity rows per thread must be 1, 2 or 4"), -2;
    E->shape.rows_v = rows;
    if (rows != 2) E->shape.lds_v = 0;
    return 0;
}

extern "C" int ifos_engine_get_velocity_rows(ifos_engine *E, int *rows)
{
    *rows = E->shape.rows_v;
    return 0;
}

extern "C" int ifos_engine_set_velocity_lds(ifos_engine *E, int on)
{
    E->shape.lds_v = on ? 1 : 0;
    if (on) E->shape.rows_v = 2;
    return 0;
}

extern "C" int ifos_engine_get_velocity_lds(ifos_engine *E, int *on)
{
    *on = E->shape.lds_v;
    return 0;
}

extern "C" int ifos_engine_set_work_order(ifos_engine *E, int slot_major)
{
    E->shape.order = slot_major ? 1 : 0;
    return 0;
}

extern "C" int ifos_engine_get_work_order(ifos_engine *E, int *slot_major)
{
    *slot_major = E->shape.order;
    return 0;
}

extern "C" int ifos_engine_nan_step(ifos_engine *E)
{
    int h[2] = {0, 0};
    hipSetDevice(E->device);
    hipMemcpyAsync(h, E->ctl, sizeof h, hipMemcpyDeviceToHost, E->stream);
    hipStreamSynchronize(E->stream);
    return h[1];
}

// ---------------------------------------------------------------------------
// device memory helpers
// ---------------------------------------------------------------------------
extern "C" void *ifos_dev_alloc(size_t bytes)
{
    void *p = nullptr;
    if (dmalloc(&p, bytes) != hipSuccess) {
        ifos_set_error("dmalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}
extern "C" int ifos_dev_free(void *p)
{
    HIP_TRY(dfree(p));
    return 0;
}
// copies of the helpers run on a non-blocking stream of the calling thread
// (not the legacy stream, which would wait on a capturing engine stream)
static int copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    thread_local hipStream_t st = nullptr;
    if (!st) HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, kind, st));
    HIP_TRY(hipStreamSynchronize(st));
    return 0;
}
extern "C" int ifos_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    return copy_sync(dst, src, bytes, hipMemcpyHostToDevice);
}
extern "C" int ifos_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    return copy_sync(dst, src, bytes, hipMemcpyDeviceToHost);
}
extern "C" int ifos_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    return copy_sync(dst, src, bytes, hipMemcpyDeviceToDevice);
}

// ---------------------------------------------------------------------------
// kernel timing for the roofline figures
// ---------------------------------------------------------------------------

// device-to-device copies reading and writing bytes/2 each, reps of them back
// to back on the engine's stream between two events: the streaming rate a
// kernel moving those bytes could reach at that size, launch and drain
// included
extern "C" int ifos_engine_time_copy(ifos_engine *E, size_t bytes, int reps, double *ms_avg)
{
    HIP_TRY(hipSetDevice(E->device));
    if (reps < 1) reps = 1;
    const size_t half = std::max<size_t>(256, bytes / 2 / 256 * 256);
    void *src = nullptr, *dst = nullptr;
    if (dmalloc(&src, half) != hipSuccess || dmalloc(&dst, half) != hipSuccess) {
        dfree(src);
        return ifos_set_error("cannot allocate 2 x %zu bytes for the copy timing", half), -3;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = 0.0f;
    hipError_t e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e == hipSuccess) e = hipMemsetAsync(src, 0, half, E->stream);
    for (int k = 0; e == hipSuccess && k < 2; k++) e = hipMemcpyAsync(dst, src, half, hipMemcpyDeviceToDevice, E->stream);
    if (e == hipSuccess) e = hipEventRecord(t0, E->stream);
    for (int k = 0; e == hipSuccess && k < reps; k++)
        e = hipMemcpyAsync(dst, src, half, hipMemcpyDeviceToDevice, E->stream);
    if (e == hipSuccess) e = hipEventRecord(t1, E->stream);
    if (e == hipSuccess) e = hipEventSynchronize(t1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
    if (t0) hipEventDestroy(t0);
    if (t1) hipEventDestroy(t1);
    dfree(src);
    dfree(dst);
    if (e != hipSuccess) return ifos_set_error("copy timing: %s", hipGetErrorString(e)), -1;
    *ms_avg = ms / reps;
    return 0;
}
extern "C" double ifos_engine_kernel_bytes_per_cell(ifos_engine *E, int kernel)
{
    const int L = E->c.L;
    switch (kernel) {
    case 0: return 4.0 * (3 + 2 + 2 + 2);  // read sxx,syy,sxy,rip,rjp,vx,vy; write vx,vy
    case 1:
        if (L == 0) return 4.0 * (2 + 3 + 3 + 3);  // read vx,vy,sxx,syy,sxy,f,g,uipjp; write sxx,syy,sxy
        // read vx,vy, sxx,syy,sxy, 3L memory variables, 6 moduli/tau; write 3 stresses, 3L memory variables
        return 4.0 * (2 + 3 + 3 * L + 6 + 3 + 3 * L);
    case 2:
        // one-kernel step: read vx,vy,sxx,syy,sxy,rip,rjp and the stress
        // materials (3 elastic, 6 viscoelastic), 3L memory variables; write
        // vx,vy,sxx,syy,sxy and 3L memory variables
        if (L == 0) return 4.0 * (5 + 2 + 3 + 5);
        return 4.0 * (5 + 2 + 6 + 3 * L + 5 + 3 * L);
    default: return 0.0;
    }
}

// kernel 0: velocity step, 1: stress step (forward, no snapshot), on the
// live state with the step counter bypassed (nt = 2)
extern "C" int ifos_engine_time_kernel(ifos_engine *E, int kernel, int reps, double *ms_avg)
{
    HIP_TRY(hipSetDevice(E->device));
    if (!E->tiles || !E->model_set) return ifos_set_error("time_kernel needs set_model and set_shot"), -2;
    StepArgs a = shaped_args(step_args(E, 0, false), E->shape);
    a.base = nullptr;
    a.nan_flag = nullptr;
    const StepArgs *da = nullptr;
    ENG_TRY(upload_args(E, a, 3, &da));
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    LaunchShape ls1 = E->shape;
    ls1.spb = 1;
    if (kernel == 2) {
        a = shaped_args(step_args(E, 0, false), ls1);
        a.base = nullptr;
        a.nan_flag = nullptr;
        ENG_TRY(upload_args(E, a, 3, &da));
    }
    auto one = [&]() {
        if (kernel == 0) launch_vel_step(a, da, E->shape, 2, E->stream);
        else if (kernel == 1) launch_str_step(a, da, E->shape, 2, E->stream);
        else launch_fused_step(a, da, ls1, 2, E->stream);
    };
    one();
    HIP_TRY(hipEventRecord(t0, E->stream));
    for (int k = 0; k < reps; k++) one();
    HIP_TRY(hipEventRecord(t1, E->stream));
    HIP_TRY(hipEventSynchronize(t1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    *ms_avg = ms / reps;
    return 0;
}

// persistent-grid candidates of the tuner, workgroups per compute unit (0:
// one workgroup per work item).  Grids of 8-12 per CU beat one workgroup per
// tile by 4-10 % on 1-5 shots of 1.6 M and 0.6 M cells per launch.
static constexpr int kGridsPerCU[] = {0, 2, 3, 4, 6, 8, 10, 12};

// Launch-shape tuner: times one forward step (velocity + stress kernels, no
// snapshot) of the current model and shots for every candidate shape --
// 1, 2 or 4 shots per workgroup (up to the slots in use) times one
// workgroup per work item or a persistent grid of 2 to 12 workgroups per
// compute unit -- keeps the fastest, and zeroes the wavefields it stepped.
extern "C" int ifos_engine_autotune(ifos_engine *E, int reps, int *spb_out, int *grid_out, double *ms_out)
{
    HIP_TRY(hipSetDevice(E->device));
    if (!E->tiles || !E->model_set) return ifos_set_error("autotune needs set_model and set_shots"), -2;
    if (reps < 1) reps = 1;
    int ncu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, E->device));
    StepArgs a0 = step_args(E, 0, false);
    a0.base = nullptr;
    a0.nan_flag = nullptr;
    const StepArgs *da = nullptr;
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    double best = 1e30;
    LaunchShape pick = E->shape;
    // each kernel's persistent grid (and the velocity step's rows per
    // thread) is chosen on its own, timed in the alternation the time loop
    // runs (velocity, stress, velocity, ...) with events between the
    // kernels; shots per workgroup and the work order are common to both
    // (one argument block)
    constexpr int kMaxReps = 16;
    reps = std::min(reps, kMaxReps);
    hipEvent_t ev[2 * kMaxReps + 1];
    for (auto &x : ev) HIP_TRY(hipEventCreate(&x));
    for (float cw : {0.0f, 1.0f, 2.0f})
    for (int order = 0; order < (E->nslots > 1 ? 2 : 1); order++)
        for (int spb : {1, 2, 4}) {
            if (spb > 1 && spb > E->nslots) break;
            LaunchShape ls;
            ls.spb = spb;
            ls.order = order;
            ls.cpml_weight = cw;
            const StepArgs a = shaped_args(a0, ls);
            ENG_TRY(upload_args(E, a, 3, &da));
            float best_v = 1e30f, best_s = 1e30f;
            int grid_v = 0, grid_s = 0, rows_v = 1, lds_v = 0;
            for (int per_cu : kGridsPerCU)
            // rows per thread; the LDS-staged variant (lds_v) ran at half the
            // speed of two rows per thread on every workload measured, so it is
            // not a candidate here
            for (int vshape : {1, 2, 4}) {
                ls.grid = ls.grid_v = per_cu ? (per_cu * ncu + 7) / 8 * 8 : 0;
                ls.rows_v = vshape == 3 ? 2 : vshape;
                ls.lds_v = vshape == 3 ? 1 : 0;
                launch_vel_step(a, da, ls, 2, E->stream);
                launch_str_step(a, da, ls, 2, E->stream);
                HIP_TRY(hipEventRecord(ev[0], E->stream));
                for (int k = 0; k < reps; k++) {
                    launch_vel_step(a, da, ls, 2, E->stream);
                    HIP_TRY(hipEventRecord(ev[2 * k + 1], E->stream));
                    launch_str_step(a, da, ls, 2, E->stream);
                    HIP_TRY(hipEventRecord(ev[2 * k + 2], E->stream));
                }
                HIP_TRY(hipEventSynchronize(ev[2 * reps]));
                float mv = 0.0f, ms = 0.0f;
                for (int k = 0; k < reps; k++) {
                    float x = 0.0f, y = 0.0f;
                    HIP_TRY(hipEventElapsedTime(&x, ev[2 * k], ev[2 * k + 1]));
                    HIP_TRY(hipEventElapsedTime(&y, ev[2 * k + 1], ev[2 * k + 2]));
                    mv += x;
                    ms += y;
                }
                mv /= reps;
                ms /= reps;
                if (mv < best_v) best_v = mv, grid_v = ls.grid_v, rows_v = ls.rows_v, lds_v = ls.lds_v;
                if (ms < best_s) best_s = ms, grid_s = ls.grid;
            }
            if (best_v + best_s < best) {
                best = best_v + best_s;
                pick = ls;
                pick.grid_v = grid_v;
                pick.rows_v = rows_v;
                pick.lds_v = lds_v;
                pick.grid = grid_s;
            }
        }
    // the one-kernel step (one shot per workgroup; not with periodic sides):
    // every CPML cost weight, work order and grid, timed back to back
    bool fused = false;
    for (float cw : {0.0f, 1.0f, 2.0f})
        if (!E->c.BOUNDARY)
        for (int order = 0; order < (E->nslots > 1 ? 2 : 1); order++)
            for (int per_cu : kGridsPerCU) {
                LaunchShape ls = pick;
                ls.spb = 1;
                ls.order = order;
                ls.cpml_weight = cw;
                ls.grid = per_cu ? (per_cu * ncu + 7) / 8 * 8 : 0;
                const StepArgs a = shaped_args(a0, ls);
                ENG_TRY(upload_args(E, a, 3, &da));
                launch_fused_step(a, da, ls, 2, E->stream);
                HIP_TRY(hipEventRecord(ev[0], E->stream));
                for (int k = 0; k < reps; k++) launch_fused_step(a, da, ls, 2, E->stream);
                HIP_TRY(hipEventRecord(ev[1], E->stream));
                HIP_TRY(hipEventSynchronize(ev[1]));
                float mf = 0.0f;
                HIP_TRY(hipEventElapsedTime(&mf, ev[0], ev[1]));
                mf /= reps;
                if (mf < best) {
                    best = mf;
                    pick = ls;
                    fused = true;
                }
            }
    for (auto &x : ev) hipEventDestroy(x);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    E->shape = pick;
    E->fused = fused;
    drop_graphs(E);
    ENG_TRY(zero_state(E));
    HIP_TRY(hipStreamSynchronize(E->stream));
    if (spb_out) *spb_out = pick.spb;
    if (grid_out) *grid_out = pick.grid;
    if (ms_out) *ms_out = best;
    return 0;
}
