// ifos_engine.hip -- native whole-shot engine behind the C ABI
// (include/ifos2d_hip.h).  One process drives one GPU; all state of a shot
// stays in HBM: wavefields, CPML memory, materials, the forward snapshots
// (slot-major [k][j][i]), seismograms, adjoint sources and the gradient
// accumulators.  The time loop is recorded once per geometry as HIP graphs
// of U = k*DTINV steps and replayed NT/U times; kernels read the step number
// from a device counter so one graph serves every block of steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host/ifos_setup.h"
#include "ifos2d_hip.h"
#include "ifos_internal.h"

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

void ifos_set_error(const char *fmt, ...)
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

    // padded grids (origin pointers)
    float *slab = nullptr;
    float *vx, *vy, *sxx, *syy, *sxy, *ux, *uy, *uxy, *vxp1, *vyp1;
    float *rip, *rjp, *uipjp, *f, *g, *fipjp, *rho, *vs, *vp, *taus, *taup;
    float *r = nullptr, *p = nullptr, *q = nullptr, *d = nullptr, *e = nullptr, *dip = nullptr;
    int nwave = 0;  // grids zeroed at the start of a run (wavefields + memory variables)

    // interior [NY][NX]
    float *interior = nullptr;
    float *muss, *lamss, *vs_i, *vp_i, *rho_i, *wc, *wc_u, *wc_rho, *g_vp, *g_vs, *g_rho, *s_vp, *s_vs, *s_rho;

    // CPML
    float *pml = nullptr;   // 12 profiles
    float *psi = nullptr;   // 5 x-strips then 4 y-strips
    float *psi_sxx_x, *psi_sxy_x, *psi_vxx, *psi_vyx, *psi_vxxs, *psi_syy_y, *psi_sxy_y, *psi_vyy, *psi_vxy;
    size_t psi_bytes = 0;

    // control
    int *ctl = nullptr;     // [0] step counter, [1] NaN step
    ifos_materials mat;
    bool model_set = false;

    // shot
    int nsrc = 0, ntr = 0, cap_src = 0, cap_tr = 0;
    bool serial_adj = false;
    int explosive = 0;
    int *src_i = nullptr;   // ix, iy, type  [3][cap_src]
    double *src_d = nullptr;  // sin, cos [2][cap_src]
    float *sig = nullptr;   // [cap_src][NT]
    int *rec_i = nullptr;   // rx, ry, kill [3][cap_tr]
    float *seis = nullptr;  // vx, vy, diff_vx, diff_vy, obs_vx, obs_vy: 6*[cap_tr][NT]
    double *trace_sums = nullptr;  // l2, en [2][cap_tr]

    // snapshots
    float *snap = nullptr;
    size_t snap_bytes = 0;

    // graphs
    hipGraphExec_t gexec[2] = {nullptr, nullptr};  // forward(with snapshots), adjoint
    int gsteps[2] = {0, 0};
    int gkey[2][5] = {{-1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1}};
    bool use_graphs = true;
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
    return 0;
}

extern "C" int ifos_engine_create(const ifos_config *cfg, int device, ifos_engine **out)
{
    *out = nullptr;
    ENG_TRY(validate(cfg));
    ifos_engine *E = new ifos_engine();
    E->c = *cfg;
    E->device = device;
    const char *ng = getenv("IFOS_NO_GRAPH");
    E->use_graphs = !(ng && ng[0] == '1');
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
    ifos_holberg(c.FDORDER, c.MAXRELERROR, E->hc);

    // padded grids: 10 wavefields + 11 materials + 6*L memory planes
    const int L = c.L;
    const int ngrid = 10 + 11 + 6 * L;
    if (hipMalloc(&E->slab, sizeof(float) * E->field * ngrid) != hipSuccess) {
        ifos_set_error("device allocation of %ld grids failed", (long)ngrid);
        return fail(-1);
    }
    hipMemset(E->slab, 0, sizeof(float) * E->field * ngrid);
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
    E->nwave = 10 + 3 * L;
    E->rip = next(); E->rjp = next(); E->uipjp = next(); E->f = next(); E->g = next(); E->fipjp = next();
    E->rho = next(); E->vs = next(); E->vp = next(); E->taus = next(); E->taup = next();
    if (L) {
        E->d = next(); for (int l = 1; l < L; l++) next();
        E->e = next(); for (int l = 1; l < L; l++) next();
        E->dip = next(); for (int l = 1; l < L; l++) next();
    }

    // interior arrays
    const int nint = 14;
    if (hipMalloc(&E->interior, sizeof(float) * E->plane * nint) != hipSuccess) {
        ifos_set_error("device allocation of interior arrays failed");
        return fail(-1);
    }
    hipMemset(E->interior, 0, sizeof(float) * E->plane * nint);
    float **ia[nint] = {&E->muss, &E->lamss, &E->vs_i, &E->vp_i, &E->rho_i, &E->wc, &E->wc_u,
                        &E->wc_rho, &E->g_vp, &E->g_vs, &E->g_rho, &E->s_vp, &E->s_vs, &E->s_rho};
    for (int k = 0; k < nint; k++) *ia[k] = E->interior + (long)k * E->plane;

    // CPML profiles and memory
    const int fw2 = c.FW > 0 ? 2 * c.FW : 1;
    std::vector<float> prof((size_t)12 * fw2, 0.0f);
    if (c.FW > 0) ifos_pml_profiles(&c, prof.data());
    if (hipMalloc(&E->pml, sizeof(float) * prof.size()) != hipSuccess) return fail(-1);
    hipMemcpy(E->pml, prof.data(), sizeof(float) * prof.size(), hipMemcpyHostToDevice);
    const size_t xs = (size_t)c.NY * fw2, ys = (size_t)fw2 * c.NX;
    E->psi_bytes = sizeof(float) * (5 * xs + 4 * ys);
    if (hipMalloc(&E->psi, E->psi_bytes) != hipSuccess) return fail(-1);
    hipMemset(E->psi, 0, E->psi_bytes);
    E->psi_sxx_x = E->psi; E->psi_sxy_x = E->psi + xs; E->psi_vxx = E->psi + 2 * xs;
    E->psi_vyx = E->psi + 3 * xs; E->psi_vxxs = E->psi + 4 * xs;
    float *yb = E->psi + 5 * xs;
    E->psi_syy_y = yb; E->psi_sxy_y = yb + ys; E->psi_vyy = yb + 2 * ys; E->psi_vxy = yb + 3 * ys;

    if (hipMalloc(&E->ctl, sizeof(int) * 4) != hipSuccess) return fail(-1);
    hipMemset(E->ctl, 0, sizeof(int) * 4);
    *out = E;
    return 0;
}

static void drop_graphs(ifos_engine *E)
{
    for (int k = 0; k < 2; k++) {
        if (E->gexec[k]) hipGraphExecDestroy(E->gexec[k]);
        E->gexec[k] = nullptr;
        for (int q = 0; q < 5; q++) E->gkey[k][q] = -1;
    }
}

extern "C" void ifos_engine_destroy(ifos_engine *E)
{
    if (!E) return;
    hipSetDevice(E->device);
    if (E->stream) hipStreamSynchronize(E->stream);
    drop_graphs(E);
    hipFree(E->slab); hipFree(E->interior); hipFree(E->pml); hipFree(E->psi); hipFree(E->ctl);
    hipFree(E->src_i); hipFree(E->src_d); hipFree(E->sig); hipFree(E->rec_i); hipFree(E->seis);
    hipFree(E->trace_sums); hipFree(E->snap);
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

extern "C" int ifos_engine_set_shot(ifos_engine *E, int nsrc, const float *srcpos, int ntr, const int *recpos)
{
    HIP_TRY(hipSetDevice(E->device));
    const ifos_config &c = E->c;
    if (nsrc < 0 || ntr < 0) return ifos_set_error("negative source or receiver count"), -2;
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
    if (nsrc > E->cap_src) {
        hipFree(E->src_i); hipFree(E->src_d); hipFree(E->sig);
        E->cap_src = nsrc;
        HIP_TRY(hipMalloc(&E->src_i, sizeof(int) * 3 * nsrc));
        HIP_TRY(hipMalloc(&E->src_d, sizeof(double) * 2 * nsrc));
        HIP_TRY(hipMalloc(&E->sig, sizeof(float) * (size_t)nsrc * c.NT));
        drop_graphs(E);
    }
    if (ntr > E->cap_tr) {
        hipFree(E->rec_i); hipFree(E->seis); hipFree(E->trace_sums);
        E->cap_tr = ntr;
        HIP_TRY(hipMalloc(&E->rec_i, sizeof(int) * 3 * ntr));
        HIP_TRY(hipMalloc(&E->seis, sizeof(float) * 6 * (size_t)ntr * c.NT));
        HIP_TRY(hipMalloc(&E->trace_sums, sizeof(double) * 2 * ntr));
        HIP_TRY(hipMemset(E->seis, 0, sizeof(float) * 6 * (size_t)ntr * c.NT));
        drop_graphs(E);
    }
    if (nsrc != E->nsrc || ntr != E->ntr) drop_graphs(E);
    E->nsrc = nsrc;
    E->ntr = ntr;
    if (nsrc) {
        std::vector<int> si(3 * (size_t)E->cap_src, 0);
        std::vector<double> sd(2 * (size_t)E->cap_src, 0.0);
        for (int l = 0; l < nsrc; l++) {
            const float *s = srcpos + (size_t)l * IFOS_SRC_STRIDE;
            si[l] = (int)s[0];
            si[E->cap_src + l] = (int)s[1];
            si[2 * E->cap_src + l] = (int)s[7];
            float azi_rad = s[6] * 3.141592653589793 / 180;  // src/update_v_PML.c:556
            sd[l] = sin(azi_rad);
            sd[E->cap_src + l] = cos(azi_rad);
        }
        // SOURCE_TYPE is rewritten by update_v_PML from the last source before
        // psource consults it (src/update_v_PML.c:558, src/IFOS2D.c:1758)
        E->explosive = ((int)srcpos[(size_t)(nsrc - 1) * IFOS_SRC_STRIDE + 7]) == 1;
        std::vector<float> w((size_t)nsrc * c.NT);
        ifos_wavelet(&c, nsrc, srcpos, w.data());
        HIP_TRY(hipMemcpy(E->src_i, si.data(), sizeof(int) * si.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(E->src_d, sd.data(), sizeof(double) * sd.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(E->sig, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
    } else {
        E->explosive = c.SOURCE_TYPE == 1;
    }
    if (ntr) {
        std::vector<int> ri(3 * (size_t)E->cap_tr, 0);
        std::vector<long> seen;
        E->serial_adj = false;
        for (int t = 0; t < ntr; t++) {
            ri[t] = recpos[2 * t];
            ri[E->cap_tr + t] = recpos[2 * t + 1];
            seen.push_back((long)recpos[2 * t + 1] * (c.NX + 1) + recpos[2 * t]);
        }
        std::sort(seen.begin(), seen.end());
        for (size_t k = 1; k < seen.size(); k++)
            if (seen[k] == seen[k - 1]) E->serial_adj = true;
        HIP_TRY(hipMemcpy(E->rec_i, ri.data(), sizeof(int) * ri.size(), hipMemcpyHostToDevice));
    }
    return 0;
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
    return g;
}

static float *seis_buf(const ifos_engine *E, int k) { return E->seis + (size_t)k * E->cap_tr * E->c.NT; }

// one time step: sw = 0 forward, 1 adjoint; extra = snapshot (fwd) / correlation (adj)
static void enqueue_step(ifos_engine *E, int sw, bool extra, bool snapshots)
{
    const ifos_config &c = E->c;
    const StepGeom g = geom(E);
    int *ntd = E->ctl;
    set_step_control(ntd, c.DTINV, c.NT);
    hipStream_t s = E->stream;

    VArgs va;
    va.g = g; va.sw = sw; va.VELOCITY = c.VELOCITY;
    va.sxx = E->sxx; va.syy = E->syy; va.sxy = E->sxy; va.rip = E->rip; va.rjp = E->rjp;
    va.vx = E->vx; va.vy = E->vy;
    const bool need_p1 = sw == 1 || snapshots;
    va.vxp1 = need_p1 ? E->vxp1 : nullptr;
    va.vyp1 = need_p1 ? E->vyp1 : nullptr;
    va.psi_sxx_x = E->psi_sxx_x; va.psi_sxy_x = E->psi_sxy_x;
    va.psi_syy_y = E->psi_syy_y; va.psi_sxy_y = E->psi_sxy_y;
    launch_update_v(va, s);

    PointArgs pa;
    if (sw == 0) {
        pa.n = E->nsrc; pa.ix = E->src_i; pa.iy = E->src_i + E->cap_src; pa.type = E->src_i + 2 * E->cap_src;
        pa.sin_azi = E->src_d; pa.cos_azi = E->src_d + E->cap_src;
        pa.sig = E->sig; pa.sig1 = E->sig; pa.ld = c.NT;
        // point forces only; explosive sources are applied in the post kernel
    } else {
        pa.n = E->ntr; pa.ix = E->rec_i; pa.iy = E->rec_i + E->cap_tr; pa.type = nullptr;
        pa.sin_azi = pa.cos_azi = nullptr;
        pa.sig = seis_buf(E, 2); pa.sig1 = seis_buf(E, 3); pa.ld = c.NT;
    }
    if (sw == 1 && E->serial_adj)
        launch_inject_serial(pa, g, E->vx, E->vy, E->rip, E->rjp, sw, c.ADJOINT_TYPE, nullptr, 0, ntd, s);
    else
        launch_inject(pa, g, E->vx, E->vy, E->rip, E->rjp, sw, c.ADJOINT_TYPE, nullptr, 0, ntd, s);

    const bool rates = sw == 0 && snapshots && c.VELOCITY == 1;
    if (c.L == 0) {
        SElArgs sa;
        sa.g = g; sa.vx = E->vx; sa.vy = E->vy; sa.sxx = E->sxx; sa.syy = E->syy; sa.sxy = E->sxy;
        sa.ux = rates ? E->ux : nullptr; sa.uy = rates ? E->uy : nullptr; sa.uxy = rates ? E->uxy : nullptr;
        sa.f = E->f; sa.gl = E->g; sa.uipjp = E->uipjp; sa.rho = sa.vs = sa.vp = nullptr;
        sa.psi_vxx = E->psi_vxx; sa.psi_vyx = E->psi_vyx; sa.psi_vyy = E->psi_vyy; sa.psi_vxy = E->psi_vxy;
        launch_update_s_elastic(sa, s);
    } else {
        SViscArgs sa;
        sa.g = g; sa.L = c.L; sa.vx = E->vx; sa.vy = E->vy; sa.sxx = E->sxx; sa.syy = E->syy; sa.sxy = E->sxy;
        sa.ux = rates ? E->ux : nullptr; sa.uy = rates ? E->uy : nullptr; sa.uxy = rates ? E->uxy : nullptr;
        sa.fipjp = E->fipjp; sa.f = E->f; sa.gv = E->g;
        sa.r = E->r; sa.p = E->p; sa.q = E->q; sa.d = E->d; sa.e = E->e; sa.dip = E->dip;
        sa.cs = 1; sa.ls = (int)E->field;
        for (int l = 0; l <= kMaxL; l++) {
            sa.bip[l] = l <= c.L ? E->mat.bip[l] : 0.0f;
            sa.bjm[l] = l <= c.L ? E->mat.bjm[l] : 0.0f;
            sa.cip[l] = l <= c.L ? E->mat.cip[l] : 0.0f;
            sa.cjm[l] = l <= c.L ? E->mat.cjm[l] : 0.0f;
        }
        sa.psi_vxx = E->psi_vxx; sa.psi_vyx = E->psi_vyx; sa.psi_vyy = E->psi_vyy; sa.psi_vxy = E->psi_vxy;
        launch_update_s_visc(sa, s);
    }

    PostArgs po;
    po.g = g; po.L = c.L; po.sw = sw; po.explosive = (sw == 0) ? E->explosive : 0; po.NT = c.NT;
    po.src.n = E->nsrc; po.src.ix = E->src_i; po.src.iy = E->src_i + E->cap_src; po.src.type = nullptr;
    po.src.sin_azi = po.src.cos_azi = nullptr; po.src.sig = E->sig; po.src.sig1 = E->sig; po.src.ld = c.NT;
    po.ntr = sw == 0 ? E->ntr : 0;
    po.rx = E->rec_i; po.ry = E->rec_i + E->cap_tr;
    po.seis_vx = seis_buf(E, 0); po.seis_vy = seis_buf(E, 1);
    po.vx = E->vx; po.vy = E->vy; po.sxx = E->sxx; po.syy = E->syy; po.sxy = E->sxy;
    // stress rates are only consumed by VELOCITY=1 snapshots; their mirror
    // rows above the surface are never read
    po.ux = rates ? E->ux : nullptr; po.uy = rates ? E->uy : nullptr; po.uxy = nullptr;
    po.f = E->f; po.gl = E->g; po.rho = E->rho; po.vs = E->vs; po.vp = E->vp; po.taus = E->taus; po.taup = E->taup;
    po.p = E->p; po.q = E->q; po.cs = 1; po.ls = (int)E->field;
    po.sumu = E->mat.sumu; po.sumpi = E->mat.sumpi;
    for (int l = 0; l <= kMaxL; l++) po.eta[l] = l <= c.L ? E->mat.eta[l] : 0.0f;
    po.psi_vxxs = E->psi_vxxs;
    po.nt_dev = ntd; po.nt = 0; po.nan_flag = E->ctl + 1;
    if (c.L) { po.f = nullptr; po.gl = nullptr; }
    launch_post(po, s);

    if (extra && sw == 0 && snapshots) {
        SnapArgs na;
        na.NX = c.NX; na.NY = c.NY; na.pitch = E->pitch; na.DTINV = c.DTINV;
        const float *src[5] = {E->vxp1, E->vyp1, c.VELOCITY ? E->ux : E->sxx, c.VELOCITY ? E->uy : E->syy,
                               c.VELOCITY ? E->uxy : E->sxy};
        for (int k = 0; k < 5; k++) {
            na.a[k] = src[k];
            na.base[k] = E->snap + (size_t)k * E->nsnap * E->plane;
        }
        na.nt_dev = ntd; na.nt = 0;
        launch_snapshot(na, s);
    }
    if (extra && sw == 1) {
        CorrArgs ca;
        ca.NX = c.NX; ca.NY = c.NY; ca.pitch = E->pitch; ca.NT = c.NT; ca.DTINV = c.DTINV;
        ca.vxp1 = E->vxp1; ca.vyp1 = E->vyp1; ca.sxx = E->sxx; ca.syy = E->syy; ca.sxy = E->sxy;
        for (int k = 0; k < 5; k++) ca.F[k] = E->snap + (size_t)k * E->nsnap * E->plane;
        ca.nt_dev = ntd; ca.nt = 0;
        ca.muss = E->muss; ca.lamss = E->lamss; ca.vs = E->vs_i;
        ca.wc_rho = E->wc_rho; ca.wc = E->wc; ca.wc_u = E->wc_u;
        launch_correlate(ca, s);
    }
}

// Steps within a unit: forward units start at nt = 1 (mod DTINV) with the
// snapshot at the first step; adjoint units end with the correlation, since
// (NT - nt) % DTINV == 0 <=> nt % DTINV == 0 when DTINV divides NT.
static bool extra_at(const ifos_engine *E, int sw, int k) { return sw == 0 ? k % E->c.DTINV == 0 : (k + 1) % E->c.DTINV == 0; }

static int run_loop(ifos_engine *E, int sw, bool snapshots)
{
    const ifos_config &c = E->c;
    const int NT = c.NT;
    int unit = c.DTINV;
    while (unit < 32 && NT % (unit * 2) == 0) unit *= 2;
    const int kind = sw;  // graph slot
    const int key[5] = {E->nsrc, E->ntr, snapshots ? 1 : 0, E->serial_adj ? 1 : 0, E->explosive};
    if (E->use_graphs) {
        if (!E->gexec[kind] || E->gsteps[kind] != unit || memcmp(E->gkey[kind], key, sizeof key)) {
            if (E->gexec[kind]) hipGraphExecDestroy(E->gexec[kind]);
            E->gexec[kind] = nullptr;
            hipGraph_t graph;
            HIP_TRY(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < unit; k++) enqueue_step(E, sw, extra_at(E, sw, k), snapshots);
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
        for (int nt = 0; nt < NT; nt++) enqueue_step(E, sw, extra_at(E, sw, nt % c.DTINV), snapshots);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

static int zero_state(ifos_engine *E)
{
    HIP_TRY(hipMemsetAsync(E->slab, 0, sizeof(float) * E->field * E->nwave, E->stream));
    HIP_TRY(hipMemsetAsync(E->psi, 0, E->psi_bytes, E->stream));
    HIP_TRY(hipMemsetAsync(E->ctl, 0, sizeof(int) * 2, E->stream));
    return 0;
}

static int ensure_snapshots(ifos_engine *E)
{
    size_t need = sizeof(float) * 5 * (size_t)E->nsnap * E->plane;
    if (E->snap && E->snap_bytes >= need) return 0;
    hipFree(E->snap);
    E->snap = nullptr;
    drop_graphs(E);
    if (hipMalloc(&E->snap, need) != hipSuccess) {
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
    if (store_snapshots) ENG_TRY(ensure_snapshots(E));
    ENG_TRY(zero_state(E));
    ENG_TRY(run_loop(E, 0, store_snapshots != 0));
    return check_nan(E);
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
    float *src[5] = {E->vx, E->vy, E->sxx, E->syy, E->sxy};
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
    if (!obs_on_device) {
        if (obs_vx) HIP_TRY(hipMemcpyAsync(seis_buf(E, 4), obs_vx, n * sizeof(float), hipMemcpyHostToDevice, E->stream));
        if (obs_vy) HIP_TRY(hipMemcpyAsync(seis_buf(E, 5), obs_vy, n * sizeof(float), hipMemcpyHostToDevice, E->stream));
        ovx = seis_buf(E, 4);
        ovy = seis_buf(E, 5);
    }
    const int *dkill = nullptr;
    if (c.TRKILL && kill) {
        HIP_TRY(hipMemcpyAsync(E->rec_i + 2 * E->cap_tr, kill, sizeof(int) * ntr, hipMemcpyHostToDevice, E->stream));
        dkill = E->rec_i + 2 * E->cap_tr;
    }
    double l2 = 0.0, en = 0.0;
    std::vector<double> h(2 * (size_t)ntr);
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
        ra.l2f_trace = ra.enf_trace = nullptr;
        launch_residual(ra, E->stream);
        HIP_TRY(hipMemcpyAsync(h.data(), E->trace_sums, sizeof(double) * ntr, hipMemcpyDeviceToHost, E->stream));
        HIP_TRY(hipMemcpyAsync(h.data() + ntr, E->trace_sums + E->cap_tr, sizeof(double) * ntr,
                               hipMemcpyDeviceToHost, E->stream));
        HIP_TRY(hipStreamSynchronize(E->stream));
        for (int t = 0; t < ntr; t++) {
            l2 += h[t];
            en += h[ntr + t];
        }
    }
    if (misfit) {
        misfit[0] = l2;
        misfit[1] = en;
    }
    return 0;
}

extern "C" int ifos_engine_adjoint(ifos_engine *E)
{
    HIP_TRY(hipSetDevice(E->device));
    if (!E->snap) return ifos_set_error("adjoint needs a forward run with snapshots"), -2;
    ENG_TRY(zero_state(E));
    HIP_TRY(hipMemsetAsync(E->wc, 0, sizeof(float) * 3 * E->plane, E->stream));  // wc, wc_u, wc_rho
    ENG_TRY(run_loop(E, 1, false));
    GradArgs ga;
    ga.n = (int)E->plane; ga.DT = E->c.DT;
    ga.vp = E->vp_i; ga.vs = E->vs_i; ga.rho = E->rho_i;
    ga.wc = E->wc; ga.wc_u = E->wc_u; ga.wc_rho = E->wc_rho;
    ga.g_vp = E->g_vp; ga.g_vs = E->g_vs; ga.g_rho = E->g_rho;
    ga.s_vp = E->s_vp; ga.s_vs = E->s_vs; ga.s_rho = E->s_rho;
    launch_gradient(ga, E->stream);
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

extern "C" int ifos_engine_reset_gradient(ifos_engine *E)
{
    HIP_TRY(hipSetDevice(E->device));
    HIP_TRY(hipMemsetAsync(E->s_vp, 0, sizeof(float) * 3 * E->plane, E->stream));  // s_vp, s_vs, s_rho
    HIP_TRY(hipStreamSynchronize(E->stream));
    return 0;
}

extern "C" int ifos_engine_nan_step(ifos_engine *E)
{
    int h[2] = {0, 0};
    hipSetDevice(E->device);
    hipMemcpy(h, E->ctl, sizeof h, hipMemcpyDeviceToHost);
    return h[1];
}

// ---------------------------------------------------------------------------
// device memory helpers
// ---------------------------------------------------------------------------
extern "C" void *ifos_dev_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        ifos_set_error("hipMalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}
extern "C" int ifos_dev_free(void *p)
{
    HIP_TRY(hipFree(p));
    return 0;
}
extern "C" int ifos_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}
extern "C" int ifos_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
extern "C" int ifos_memcpy_d2d(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

// ---------------------------------------------------------------------------
// kernel timing for the roofline figures
// ---------------------------------------------------------------------------
extern "C" double ifos_engine_kernel_bytes_per_cell(ifos_engine *E, int kernel)
{
    const int L = E->c.L;
    switch (kernel) {
    case 0: return 4.0 * (5 + 2 + 2);  // read sxx,syy,sxy,rip,rjp,vx,vy; write vx,vy
    case 1:
        if (L == 0) return 4.0 * (2 + 3 + 3 + 3);  // read vx,vy,sxx,syy,sxy,f,g,uipjp; write sxx,syy,sxy
        return 4.0 * (2 + 3 + 3 + 6 * L + 3 + 3 * L);  // + r,p,q rw, d,e,dip, fipjp,f,g
    case 3: return 4.0 * (5 + 5 + 6 + 3);  // adjoint fields, snapshots, acc rw, moduli
    default: return 0.0;
    }
}

extern "C" int ifos_engine_time_kernel(ifos_engine *E, int kernel, int reps, double *ms_avg)
{
    HIP_TRY(hipSetDevice(E->device));
    const ifos_config &c = E->c;
    const StepGeom g = geom(E);
    set_step_control(nullptr, c.DTINV, c.NT);
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    auto one = [&]() {
        if (kernel == 0) {
            VArgs va;
            va.g = g; va.sw = 0; va.VELOCITY = c.VELOCITY;
            va.sxx = E->sxx; va.syy = E->syy; va.sxy = E->sxy; va.rip = E->rip; va.rjp = E->rjp;
            va.vx = E->vx; va.vy = E->vy; va.vxp1 = nullptr; va.vyp1 = nullptr;
            va.psi_sxx_x = E->psi_sxx_x; va.psi_sxy_x = E->psi_sxy_x;
            va.psi_syy_y = E->psi_syy_y; va.psi_sxy_y = E->psi_sxy_y;
            launch_update_v(va, E->stream);
        } else if (kernel == 1) {
            if (c.L == 0) {
                SElArgs sa;
                sa.g = g; sa.vx = E->vx; sa.vy = E->vy; sa.sxx = E->sxx; sa.syy = E->syy; sa.sxy = E->sxy;
                sa.ux = sa.uy = sa.uxy = nullptr;
                sa.f = E->f; sa.gl = E->g; sa.uipjp = E->uipjp; sa.rho = sa.vs = sa.vp = nullptr;
                sa.psi_vxx = E->psi_vxx; sa.psi_vyx = E->psi_vyx; sa.psi_vyy = E->psi_vyy; sa.psi_vxy = E->psi_vxy;
                launch_update_s_elastic(sa, E->stream);
            } else {
                SViscArgs sa;
                sa.g = g; sa.L = c.L; sa.vx = E->vx; sa.vy = E->vy; sa.sxx = E->sxx; sa.syy = E->syy; sa.sxy = E->sxy;
                sa.ux = sa.uy = sa.uxy = nullptr;
                sa.fipjp = E->fipjp; sa.f = E->f; sa.gv = E->g;
                sa.r = E->r; sa.p = E->p; sa.q = E->q; sa.d = E->d; sa.e = E->e; sa.dip = E->dip;
                sa.cs = 1; sa.ls = (int)E->field;
                for (int l = 0; l <= kMaxL; l++) {
                    sa.bip[l] = l <= c.L ? E->mat.bip[l] : 0.0f;
                    sa.bjm[l] = l <= c.L ? E->mat.bjm[l] : 0.0f;
                    sa.cip[l] = l <= c.L ? E->mat.cip[l] : 0.0f;
                    sa.cjm[l] = l <= c.L ? E->mat.cjm[l] : 0.0f;
                }
                sa.psi_vxx = E->psi_vxx; sa.psi_vyx = E->psi_vyx; sa.psi_vyy = E->psi_vyy; sa.psi_vxy = E->psi_vxy;
                launch_update_s_visc(sa, E->stream);
            }
        }
    };
    one();  // warm
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
