// ifos_fp64.hip -- the P-SV scheme in FP64 on the device, with checkpointed
// recomputation of the forward wavefield: the error yardstick for runs with
// 16-bit snapshots (no reference counterpart; the reference is FP32 only).
//
// Every field, material and accumulator is double; the setup scalars (Holberg
// coefficients, CPML profiles, source wavelets, the GSLS factors eta, b, c)
// are the FP32 values the reference computes, promoted where they are used.
// Each expression keeps the reference's operand order (src/update_v_PML.c:29,
// src/update_s_elastic_PML.c:28, src/update_s_visc_PML.c:28,
// src/surface_elastic_PML.c:26, src/surface_PML.c:25, src/psource.c:26,
// src/calc_res.c:136, src/IFOS2D.c:2381-2419 and :2471-2611), and the library
// is built with -ffp-contract=off, so results are those of the same scheme
// evaluated in double on the CPU.
//
// Memory: the reference keeps every DTINV-th forward wavefield
// (src/IFOS2D.c:544-598), 5 fields x NT/DTINV slots; in double that is
// 640 GB for a 4000x1000 grid with NT 8000 and DTINV 2.  Here the forward run
// keeps its whole state (velocities, stresses, memory variables, CPML
// memory) at the start of every segment of C steps.  The adjoint run, which
// needs the forward wavefields in reverse time order, rebuilds one segment at
// a time from its checkpoint, storing only that segment's C/DTINV snapshots.
// The forward run is thus done twice; the checkpoints take NT/C states and
// the snapshots C/DTINV slots, with C near sqrt(NT*DTINV) by default.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ifos2d_hip.h"
#include "host/ifos_setup.h"
#include "ifos_internal.h"

namespace {

#define F64_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            ifos_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                                             \
        }                                                                                          \
    } while (0)

constexpr int kBX = 64, kBY = 4;

// Wavefield state of one run (forward or adjoint).  The arrays are views
// into one contiguous block, so a checkpoint is one device copy.
struct State {
    double *vx, *vy, *sxx, *syy, *sxy;          // padded grids (origin pointers)
    double *r, *p, *q;                          // L padded planes each (origin of plane 0)
    double *psi_sxx_x, *psi_sxy_x, *psi_vxx, *psi_vyx, *psi_vxxs;  // [NY][2FW]
    double *psi_syy_y, *psi_sxy_y, *psi_vyy, *psi_vxy;             // [2FW][NX]
};

// Everything a kernel reads besides the state: geometry, scalars, materials.
struct Geo {
    int NX, NY, N, FW, L, P;       // P: pitch of a padded grid
    long plane;                    // doubles per padded grid
    int FREE_SURF, BOUNDARY, VELOCITY, ADJOINT_TYPE;
    float DT, DH;
    float hc[7];
    float eta[9], bip[9], bjm[9], cip[9], cjm[9];
    float sumu, sumpi;
    const float *pml;              // 12 profiles of 2FW (ifos_pml_profiles order)
    // materials, padded grids of double (origin pointers)
    const double *rho, *vp, *vs, *taus, *taup, *rip, *rjp, *uipjp, *fipjp, *f, *g;
    const double *dip, *d, *e;     // L planes each
    double *ux, *uy, *uxy, *vxp1, *vyp1;
};

__device__ __forceinline__ long at(const Geo &g, int j, int i) { return (long)j * g.P + i; }

enum { PK_X = 0, PA_X, PB_X, PK_XH, PA_XH, PB_XH, PK_Y, PA_Y, PB_Y, PK_YH, PA_YH, PB_YH };
// profile k at 1-based strip index h
__device__ __forceinline__ float prof(const Geo &g, int k, int h) { return g.pml[k * 2 * g.FW + h - 1]; }

// ---------------------------------------------------------------------------
// update_v_PML (src/update_v_PML.c:29-596): one thread per interior cell
template <int N>
__global__ __launch_bounds__(kBX *kBY) void k64_vel(Geo g, State s, int sw)
{
    const int i = blockIdx.x * kBX + threadIdx.x + 1, j = blockIdx.y * kBY + threadIdx.y + 1;
    if (i > g.NX || j > g.NY) return;
    const long c = at(g, j, i);
    const int P = g.P;
    const double *sxx = s.sxx + c, *syy = s.syy + c, *sxy = s.sxy + c;
    double sxx_x = g.hc[1] * (sxx[1] - sxx[0]);
    double sxy_x = g.hc[1] * (sxy[0] - sxy[-1]);
    double sxy_y = g.hc[1] * (sxy[0] - sxy[-P]);
    double syy_y = g.hc[1] * (syy[P] - syy[0]);
#pragma unroll
    for (int m = 2; m <= N; m++) {
        sxx_x = sxx_x + g.hc[m] * (sxx[m] - sxx[1 - m]);
        sxy_x = sxy_x + g.hc[m] * (sxy[m - 1] - sxy[-m]);
        sxy_y = sxy_y + g.hc[m] * (sxy[(m - 1) * P] - sxy[-m * P]);
        syy_y = syy_y + g.hc[m] * (syy[m * P] - syy[(1 - m) * P]);
    }
    const int FW = g.FW, NX = g.NX, NY = g.NY;
    if (!g.BOUNDARY && i <= FW) {
        const int h = i;
        double &p1 = s.psi_sxx_x[(long)(j - 1) * 2 * FW + h - 1], &p2 = s.psi_sxy_x[(long)(j - 1) * 2 * FW + h - 1];
        p1 = prof(g, PB_XH, h) * p1 + prof(g, PA_XH, h) * sxx_x;
        sxx_x = sxx_x / prof(g, PK_XH, h) + p1;
        p2 = prof(g, PB_X, h) * p2 + prof(g, PA_X, h) * sxy_x;
        sxy_x = sxy_x / prof(g, PK_X, h) + p2;
    }
    if (!g.BOUNDARY && i >= NX - FW + 1) {
        const int h = i - NX + 2 * FW;
        double &p1 = s.psi_sxx_x[(long)(j - 1) * 2 * FW + h - 1], &p2 = s.psi_sxy_x[(long)(j - 1) * 2 * FW + h - 1];
        p1 = prof(g, PB_XH, h) * p1 + prof(g, PA_XH, h) * sxx_x;
        sxx_x = sxx_x / prof(g, PK_XH, h) + p1;
        p2 = prof(g, PB_X, h) * p2 + prof(g, PA_X, h) * sxy_x;
        sxy_x = sxy_x / prof(g, PK_X, h) + p2;
    }
    if (!g.FREE_SURF && j <= FW) {
        const int h = j;
        double &p1 = s.psi_syy_y[(long)(h - 1) * NX + i - 1], &p2 = s.psi_sxy_y[(long)(h - 1) * NX + i - 1];
        p1 = prof(g, PB_YH, h) * p1 + prof(g, PA_YH, h) * syy_y;
        syy_y = syy_y / prof(g, PK_YH, h) + p1;
        p2 = prof(g, PB_Y, h) * p2 + prof(g, PA_Y, h) * sxy_y;
        sxy_y = sxy_y / prof(g, PK_Y, h) + p2;
    }
    if (j >= NY - FW + 1) {
        const int h = j - NY + 2 * FW;
        double &p1 = s.psi_syy_y[(long)(h - 1) * NX + i - 1], &p2 = s.psi_sxy_y[(long)(h - 1) * NX + i - 1];
        p1 = prof(g, PB_YH, h) * p1 + prof(g, PA_YH, h) * syy_y;
        syy_y = syy_y / prof(g, PK_YH, h) + p1;
        p2 = prof(g, PB_Y, h) * p2 + prof(g, PA_Y, h) * sxy_y;
        sxy_y = sxy_y / prof(g, PK_Y, h) + p2;
    }
    const double rip = g.rip[c], rjp = g.rjp[c];
    const float DT = g.DT, DH = g.DH;
    if (sw == 0) {
        g.vxp1[c] = rip * (sxx_x + sxy_y) / DH;
        g.vyp1[c] = rjp * (sxy_x + syy_y) / DH;
    } else if (g.VELOCITY == 0) {
        g.vxp1[c] += s.vx[c] * DT;
        g.vyp1[c] += s.vy[c] * DT;
    } else {
        g.vxp1[c] = s.vx[c];
        g.vyp1[c] = s.vy[c];
    }
    s.vx[c] += DT * rip * (sxx_x + sxy_y) / DH;
    s.vy[c] += DT * rjp * (sxy_x + syy_y) / DH;
}

// Point forces (forward) and adjoint sources after the velocity update
// (src/update_v_PML.c:553-578).  serial: one thread applies the points in
// order (points may share a cell); otherwise one thread per point (all
// points at distinct cells).
struct Points {
    int n;
    const int *ix, *iy, *type;
    const double *sin_azi, *cos_azi;
    const float *fsig;    // forward wavelets [n][NT]
    const double *res0, *res1;  // adjoint sources [n][NT]
    int NT;
};

__device__ __forceinline__ void apply_point(const Geo &g, const State &s, const Points &p, int l, int nt, int sw)
{
    const int i = p.ix[l], j = p.iy[l];
    const long c = at(g, j, i);
    const float DT = g.DT, DH = g.DH;
    if (sw == 0) {
        const double s0 = p.fsig[(long)l * p.NT + nt - 1], s1 = s0;
        const int type = p.type[l];
        if (type == 2) s.vx[c] += (DT * g.rip[c] * s1) / (DH * DH);
        if (type == 3) s.vy[c] += (DT * g.rjp[c] * s1) / (DH * DH);
        if (type == 4) {
            s.vx[c] += (DT * g.rip[c] * p.sin_azi[l] * s0) / (DH * DH);
            s.vy[c] += (DT * g.rjp[c] * p.cos_azi[l] * s0) / (DH * DH);
        }
    } else {
        const double s0 = p.res0[(long)l * p.NT + nt - 1], s1 = p.res1[(long)l * p.NT + nt - 1];
        if (g.ADJOINT_TYPE == 1) {
            s.vx[c] += s0;
            s.vy[c] += s1;
        }
        if (g.ADJOINT_TYPE == 2) s.vy[c] += s1;
        if (g.ADJOINT_TYPE == 3) s.vx[c] += s0;
    }
}

__global__ void k64_points(Geo g, State s, Points p, int nt, int sw, int serial)
{
    if (serial) {
        if (blockIdx.x || threadIdx.x) return;
        for (int l = 0; l < p.n; l++) apply_point(g, s, p, l, nt, sw);
        return;
    }
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < p.n) apply_point(g, s, p, l, nt, sw);
}

// ---------------------------------------------------------------------------
// stress update (src/update_s_elastic_PML.c:28-569 for L == 0,
// src/update_s_visc_PML.c:28-515 otherwise) with the stress-side CPML
// (:166-219)
template <int N>
__global__ __launch_bounds__(kBX *kBY) void k64_str(Geo g, State s)
{
    const int i = blockIdx.x * kBX + threadIdx.x + 1, j = blockIdx.y * kBY + threadIdx.y + 1;
    if (i > g.NX || j > g.NY) return;
    const long c = at(g, j, i);
    const int P = g.P, L = g.L;
    const float DT = g.DT, DH = g.DH;
    const float scale = L ? (float)(1.0 / DH) : DT / DH;
    const double *vx = s.vx + c, *vy = s.vy + c;
    double vxx = g.hc[1] * (vx[0] - vx[-1]);
    double vyx = g.hc[1] * (vy[1] - vy[0]);
    double vxy = g.hc[1] * (vx[P] - vx[0]);
    double vyy = g.hc[1] * (vy[0] - vy[-P]);
#pragma unroll
    for (int m = 2; m <= N; m++) {
        vxx = vxx + g.hc[m] * (vx[m - 1] - vx[-m]);
        vyx = vyx + g.hc[m] * (vy[m] - vy[1 - m]);
        vxy = vxy + g.hc[m] * (vx[m * P] - vx[(1 - m) * P]);
        vyy = vyy + g.hc[m] * (vy[(m - 1) * P] - vy[-m * P]);
    }
    vxx = vxx * scale;
    vyx = vyx * scale;
    vxy = vxy * scale;
    vyy = vyy * scale;
    const int FW = g.FW, NX = g.NX, NY = g.NY;
    if (!g.BOUNDARY && i <= FW) {
        const int h = i;
        double &p1 = s.psi_vxx[(long)(j - 1) * 2 * FW + h - 1], &p2 = s.psi_vyx[(long)(j - 1) * 2 * FW + h - 1];
        p1 = prof(g, PB_X, h) * p1 + prof(g, PA_X, h) * vxx;
        vxx = vxx / prof(g, PK_X, h) + p1;
        p2 = prof(g, PB_XH, h) * p2 + prof(g, PA_XH, h) * vyx;
        vyx = vyx / prof(g, PK_XH, h) + p2;
    }
    if (!g.BOUNDARY && i >= NX - FW + 1) {
        const int h = i - NX + 2 * FW;
        double &p1 = s.psi_vxx[(long)(j - 1) * 2 * FW + h - 1], &p2 = s.psi_vyx[(long)(j - 1) * 2 * FW + h - 1];
        p1 = prof(g, PB_X, h) * p1 + prof(g, PA_X, h) * vxx;
        vxx = vxx / prof(g, PK_X, h) + p1;
        p2 = prof(g, PB_XH, h) * p2 + prof(g, PA_XH, h) * vyx;
        vyx = vyx / prof(g, PK_XH, h) + p2;
    }
    if (!g.FREE_SURF && j <= FW) {
        const int h = j;
        double &p1 = s.psi_vyy[(long)(h - 1) * NX + i - 1], &p2 = s.psi_vxy[(long)(h - 1) * NX + i - 1];
        p1 = prof(g, PB_Y, h) * p1 + prof(g, PA_Y, h) * vyy;
        p2 = prof(g, PB_YH, h) * p2 + prof(g, PA_YH, h) * vxy;
        vyy = vyy / prof(g, PK_Y, h) + p1;
        vxy = vxy / prof(g, PK_YH, h) + p2;
    }
    if (j >= NY - FW + 1) {
        const int h = j - NY + 2 * FW;
        double &p1 = s.psi_vyy[(long)(h - 1) * NX + i - 1], &p2 = s.psi_vxy[(long)(h - 1) * NX + i - 1];
        p1 = prof(g, PB_Y, h) * p1 + prof(g, PA_Y, h) * vyy;
        vyy = vyy / prof(g, PK_Y, h) + p1;
        p2 = prof(g, PB_YH, h) * p2 + prof(g, PA_YH, h) * vxy;
        vxy = vxy / prof(g, PK_YH, h) + p2;
    }
    if (L == 0) {
        const double fipjp = g.uipjp[c];
        const double f = g.rho[c] * g.vs[c] * g.vs[c];
        const double gg = g.rho[c] * ((g.vp[c] * g.vp[c]) - 2 * g.vs[c] * g.vs[c]);
        s.sxy[c] += fipjp * (vyx + vxy);
        s.sxx[c] += gg * (vxx + vyy) + (2.0 * f * vxx);
        s.syy[c] += gg * (vxx + vyy) + (2.0 * f * vyy);
        g.uxy[c] = (fipjp / DT) * (vyx + vxy);
        g.ux[c] = (gg / DT) * (vxx + vyy) + ((2.0 * f / DT) * vxx);
        g.uy[c] = (gg / DT) * (vxx + vyy) + ((2.0 * f / DT) * vyy);
        return;
    }
    const float dthalbe = DT / 2.0;
    const long pl = g.plane;
    double sumr = 0.0, sump = 0.0, sumq = 0.0;
    for (int l = 0; l < L; l++) {
        sumr += s.r[c + l * pl];
        sump += s.p[c + l * pl];
        sumq += s.q[c + l * pl];
    }
    const double fipjp = g.fipjp[c], f = g.f[c], gg = g.g[c];
    double sxy = s.sxy[c], sxx = s.sxx[c], syy = s.syy[c];
    sxy += (fipjp * (vxy + vyx)) + (dthalbe * sumr);
    sxx += (gg * (vxx + vyy)) - (2.0 * f * vyy) + (dthalbe * sump);
    syy += (gg * (vxx + vyy)) - (2.0 * f * vxx) + (dthalbe * sumq);
    double uxy = ((fipjp / DT) * (vxy + vyx)) + (0.5 * sumr);
    double ux = ((gg / DT) * (vxx + vyy)) - ((2.0 * f / DT) * vyy) + (0.5 * sump);
    double uy = ((gg / DT) * (vxx + vyy)) - ((2.0 * f / DT) * vxx) + (0.5 * sumq);
    sumr = sump = sumq = 0.0;
    for (int l = 0; l < L; l++) {
        const long o = c + l * pl;
        const double rr = g.bip[l + 1] * (s.r[o] * g.cip[l + 1] - (g.dip[o] * (vxy + vyx)));
        const double pp = g.bjm[l + 1] * (s.p[o] * g.cjm[l + 1] - (g.e[o] * (vxx + vyy)) + (2.0 * g.d[o] * vyy));
        const double qq = g.bjm[l + 1] * (s.q[o] * g.cjm[l + 1] - (g.e[o] * (vxx + vyy)) + (2.0 * g.d[o] * vxx));
        s.r[o] = rr;
        s.p[o] = pp;
        s.q[o] = qq;
        sumr += rr;
        sump += pp;
        sumq += qq;
    }
    s.sxy[c] = sxy + (dthalbe * sumr);
    s.sxx[c] = sxx + (dthalbe * sump);
    s.syy[c] = syy + (dthalbe * sumq);
    g.uxy[c] = uxy + (0.5 * sumr);
    g.ux[c] = ux + (0.5 * sump);
    g.uy[c] = uy + (0.5 * sumq);
}

// explosive sources after the stress update (src/psource.c:26-75), forward
// only, in source order by one thread; amp is FP32 as there
__global__ void k64_psource(Geo g, State s, Points p, int nt)
{
    if (blockIdx.x || threadIdx.x) return;
    const int NT = p.NT;
    const float DH = g.DH;
    for (int l = 0; l < p.n; l++) {
        const float *sg = p.fsig + (long)l * NT - 1;  // 1-based
        float amp = 0.0f;
        if (nt == 1) amp = sg[nt + 1] / (2.0 * DH * DH);
        if ((nt > 1) && (nt < NT)) amp = (sg[nt + 1] - sg[nt - 1]) / (2.0 * DH * DH);
        if (nt == NT) amp = -sg[nt - 1] / (2.0 * DH * DH);
        const long c = at(g, p.iy[l], p.ix[l]);
        s.sxx[c] += 0.5 * amp;
        s.syy[c] += 0.5 * amp;
    }
}

// free surface at j = 1, one thread per column (src/surface_elastic_PML.c:
// 26-124, src/surface_PML.c:25-177)
template <int N>
__global__ void k64_surface(Geo g, State s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > g.NX) return;
    const int j = 1, P = g.P, L = g.L, FW = g.FW, NX = g.NX;
    const long c = at(g, j, i), pl = g.plane;
    const float DT = g.DT;
    const float dh24 = 1.0 / g.DH;
    if (L)
        for (int l = 0; l < L; l++) s.q[c + l * pl] = 0.0;
    s.syy[c] = 0.0;
    g.uy[c] = 0.0;
    double vxx = 0.0, vyy = 0.0;
    for (int m = 1; m <= N; m++) {
        s.syy[c - m * P] = -s.syy[c + m * P];
        s.sxy[c - m * P] = -s.sxy[c + (m - 1) * P];
        g.uy[c - m * P] = -g.uy[c + m * P];
        g.uxy[c - m * P] = -g.uxy[c + (m - 1) * P];
        vxx += g.hc[m] * (s.vx[c + m - 1] - s.vx[c - m]);
        vyy += g.hc[m] * (s.vy[c + (m - 1) * P] - s.vy[c - m * P]);
    }
    vxx *= dh24;
    vyy *= dh24;
    if (!g.BOUNDARY && i <= FW) {
        const int h = i;
        double &p1 = s.psi_vxxs[(long)(j - 1) * 2 * FW + h - 1];
        p1 = prof(g, PB_X, h) * p1 + prof(g, PA_X, h) * vxx;
        vxx = vxx / prof(g, PK_X, h) + p1;
    }
    if (!g.BOUNDARY && i >= NX - FW + 1) {
        const int h = i - NX + 2 * FW;
        double &p1 = s.psi_vxxs[(long)(j - 1) * 2 * FW + h - 1];
        p1 = prof(g, PB_X, h) * p1 + prof(g, PA_X, h) * vxx;
        vxx = vxx / prof(g, PK_X, h) + p1;
    }
    if (L == 0) {
        const double fjm = g.rho[c] * g.vs[c] * g.vs[c] * 2.0;
        const double gg = g.rho[c] * ((g.vp[c] * g.vp[c]) - 2 * g.vs[c] * g.vs[c]);
        s.sxx[c] += -DT * ((gg * gg) / (gg + fjm) * vxx + gg * vyy);
        g.ux[c] = ((2.0 * gg * fjm + fjm * fjm) / (gg + fjm)) * vxx;
        return;
    }
    const float dthalbe = DT / 2.0;
    const double mu = (g.vs[c] * g.vs[c] * g.rho[c]) / (1.0 + g.sumu * g.taus[c]);
    const double pim = (g.vp[c] * g.vp[c] * g.rho[c]) / (1.0 + g.sumpi * g.taup[c]);
    double sump = 0.0;
    for (int l = 0; l < L; l++) sump += s.p[c + l * pl];
    const double fjm = mu * 2.0 * (1.0 + L * g.taus[c]);
    const double gg = pim * (1.0 + L * g.taup[c]);
    s.sxx[c] += -(DT * (gg - fjm) * (gg - fjm) * vxx / gg) - (DT * (gg - fjm) * vyy) - (dthalbe * sump);
    g.ux[c] += -((gg - fjm) * (gg - fjm) * vxx / gg) - ((gg - fjm) * vyy) - (0.5 * sump);
    sump = 0.0;
    for (int l = 0; l < L; l++) {
        const double bjm = g.eta[l + 1] / (1.0 + (g.eta[l + 1] * 0.5));
        const double djm = 2.0 * mu * g.taus[c];
        const double e = pim * g.taup[c];
        s.p[c + l * pl] += bjm * (((djm - e) * ((fjm / gg) - 1.0) * vxx) - ((djm - e) * vyy));
        sump += s.p[c + l * pl];
    }
    s.sxx[c] += (dthalbe * sump);
    g.ux[c] += (0.5 * sump);
}

// seismo_ssg (src/seismo_ssg.c:26-55): vx, vy at the receivers
__global__ void k64_seismo(Geo g, State s, int ntr, const int *rx, const int *ry, int NT, int nt, double *svx,
                           double *svy)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ntr) return;
    const long c = at(g, ry[r], rx[r]);
    svx[(long)r * NT + nt - 1] = s.vx[c];
    svy[(long)r * NT + nt - 1] = s.vy[c];
}

// snapshot of one forward step (src/IFOS2D.c:1805-1871) into slot k of the
// segment buffer, 5 interior planes [k][f][NY][NX]
__global__ __launch_bounds__(kBX *kBY) void k64_snap(Geo g, State s, double *snap, long k)
{
    const int i = blockIdx.x * kBX + threadIdx.x + 1, j = blockIdx.y * kBY + threadIdx.y + 1;
    if (i > g.NX || j > g.NY) return;
    const long c = at(g, j, i), pn = (long)g.NX * g.NY, e = (long)(j - 1) * g.NX + (i - 1);
    double *o = snap + k * 5 * pn + e;
    o[0] = g.vxp1[c];
    o[pn] = g.vyp1[c];
    o[2 * pn] = g.VELOCITY ? g.ux[c] : s.sxx[c];
    o[3 * pn] = g.VELOCITY ? g.uy[c] : s.syy[c];
    o[4 * pn] = g.VELOCITY ? g.uxy[c] : s.sxy[c];
}

// zero-lag correlation of the adjoint state with snapshot slot k
// (src/IFOS2D.c:2381-2419; the mu term parenthesised as coded at :2398-2404)
__global__ __launch_bounds__(kBX *kBY) void k64_corr(Geo g, State s, const double *snap, long k, double *wc,
                                                     double *wc_u, double *wc_rho)
{
    const int i = blockIdx.x * kBX + threadIdx.x + 1, j = blockIdx.y * kBY + threadIdx.y + 1;
    if (i > g.NX || j > g.NY) return;
    const long c = at(g, j, i), pn = (long)g.NX * g.NY, e = (long)(j - 1) * g.NX + (i - 1);
    const double *F = snap + k * 5 * pn + e;
    const double Frx = F[0], Fry = F[pn], Fx = F[2 * pn], Fy = F[3 * pn], Fu = F[4 * pn];
    const double sxx = s.sxx[c], syy = s.syy[c], sxy = s.sxy[c];
    wc_rho[e] += (g.vxp1[c] * Frx) + (g.vyp1[c] * Fry);
    wc[e] += (Fx + Fy) * (sxx + syy);
    const double muss = g.rho[c] * g.vs[c] * g.vs[c];
    const double lamss = g.rho[c] * g.vp[c] * g.vp[c] - 2.0 * muss;
    if (g.vs[c] > 0.0)
        wc_u[e] += ((1.0 / (muss * muss)) * (Fu * sxy)) +
                   (.25 * (Fx + Fy * (sxx + syy)) / ((lamss + muss) * (lamss + muss))) +
                   (.25 * (Fx - Fy * (sxx - syy)) / (muss * muss));
}

// chain rule to Vp, Vs, rho (src/IFOS2D.c:2471-2611, PARAMETERIZATION 1)
// and the shot sum (:2703-2713)
__global__ __launch_bounds__(kBX *kBY) void k64_chain(Geo g, const double *wc, const double *wc_u,
                                                      const double *wc_rho, double *gvp, double *gvs, double *grho,
                                                      double *svp, double *svs, double *srho)
{
    const int i = blockIdx.x * kBX + threadIdx.x + 1, j = blockIdx.y * kBY + threadIdx.y + 1;
    if (i > g.NX || j > g.NY) return;
    const long c = at(g, j, i), e = (long)(j - 1) * g.NX + (i - 1);
    const float DT = g.DT;
    const double rh = g.rho[c], u = g.vs[c], pv = g.vp[c];
    double wl = -DT * wc[e];
    const double muss = rh * u * u;
    const double lamss = rh * pv * pv - 2.0 * muss;
    wl = (1.0 / (4.0 * (lamss + muss) * (lamss + muss))) * wl;
    const double a = 2.0 * pv * rh * wl;
    const double wmu = -DT * wc_u[e];
    const double b = -4.0 * rh * u * wl + 2.0 * rh * u * wmu;
    const double wrs = -DT * wc_rho[e];
    const double r = ((pv * pv) - (2.0 * u * u)) * wl + u * u * wmu + wrs;
    gvp[e] = a;
    gvs[e] = b;
    grho[e] = r;
    svp[e] += a;
    svs[e] += b;
    srho[e] += r;
}

// residual of one trace (src/calc_res.c:136-282, LNORM 2): integrated
// (VELOCITY 0) or plain traces, time-reversed difference, and the trace's
// misfit and data energy
__global__ void k64_residual(int ntr, int NT, int VELOCITY, float DT, const double *syn, const float *obs,
                             const int *kill, double *diff, double *l2, double *en)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ntr) return;
    const double *sy = syn + (long)r * NT;
    const float *ob = obs + (long)r * NT;
    double *d = diff + (long)r * NT;
    double a = 0.0, b = 0.0, L2 = 0.0, E = 0.0;
    const bool dead = kill && kill[r] == 1;
    for (int t = 0; t < NT; t++) {
        double is, id;
        if (VELOCITY == 0) {
            a += sy[t];
            b += (double)ob[t];
            is = a * DT;
            id = b * DT;
        } else {
            is = sy[t];
            id = ob[t];
        }
        const double df = dead ? 0.0 : is - id;
        d[NT - 1 - t] = df;
        if (!dead) {
            L2 += df * df;
            E += id * id;
        }
    }
    l2[r] = L2;
    en[r] = E;
}

inline dim3 grid2(int NX, int NY) { return dim3((NX + kBX - 1) / kBX, (NY + kBY - 1) / kBY); }

}  // namespace

// ---------------------------------------------------------------------------
struct ifos_fp64 {
    ifos_config cfg{};
    int device = 0, N = 0, nd = 0, P = 0, rows = 0;
    long plane = 0, npts = 0;  // doubles per padded grid; interior cells
    hipStream_t st = nullptr;
    Geo g{};
    // device memory
    double *mat = nullptr;      // materials block
    double *work = nullptr;     // ux uy uxy vxp1 vyp1, adjoint vxp1 vyp1
    float *pml = nullptr;
    size_t state_doubles = 0;
    double *fwd = nullptr, *adj = nullptr;  // two state blocks
    State sf{}, sa{};
    double *ckpt = nullptr;     // [ncp][state]
    int C = 0, ncp = 0;
    double *snap = nullptr;     // [C/DTINV][5][NY][NX]
    double *acc = nullptr;      // wc, wc_u, wc_rho, g (3), sum (3): 9 interior planes
    // shot
    int nsrc = 0, ntr = 0;
    int *pts_i = nullptr;       // device: src ix, iy, type (3*nsrc), rec x, y (2*ntr)
    double *pts_d = nullptr;    // device: sin, cos (2*nsrc)
    float *fsig = nullptr;      // device [nsrc][NT]
    double *seis = nullptr;     // device [2][ntr][NT]
    double *res = nullptr;      // device [2][ntr][NT]
    float *obs = nullptr;       // device [2][ntr][NT]
    int *kill = nullptr;        // device [ntr]
    double *red = nullptr;      // device [4][ntr]
    bool explosive = false, src_serial = true, rec_serial = false;
    size_t cap_pts = 0;
};

static void f64_free(ifos_fp64 *e)
{
    void *ptrs[] = {e->mat, e->work, e->pml, e->fwd, e->adj, e->ckpt, e->snap, e->acc, e->pts_i, e->pts_d,
                    e->fsig, e->seis, e->res, e->obs, e->kill, e->red};
    for (void *p : ptrs)
        if (p) hipFree(p);
}

static size_t state_doubles(const ifos_config *c, long plane)
{
    const long fw2 = c->FW > 0 ? 2L * c->FW : 1;
    return (size_t)plane * (5 + 3 * c->L) + (size_t)fw2 * c->NY * 5 + (size_t)fw2 * c->NX * 4;
}

static void state_views(ifos_fp64 *e, double *base, State &s)
{
    const ifos_config &c = e->cfg;
    const long pl = e->plane, off = (long)(e->nd - 1) * e->P + (e->nd - 1);  // (j=0,i=0) in a padded grid
    const long fw2 = c.FW > 0 ? 2L * c.FW : 1;
    double *p = base;
    double **grids[5] = {&s.vx, &s.vy, &s.sxx, &s.syy, &s.sxy};
    for (auto g : grids) { *g = p + off; p += pl; }
    s.r = p + off; p += pl * c.L;
    s.p = p + off; p += pl * c.L;
    s.q = p + off; p += pl * c.L;
    double **xs[5] = {&s.psi_sxx_x, &s.psi_sxy_x, &s.psi_vxx, &s.psi_vyx, &s.psi_vxxs};
    for (auto x : xs) { *x = p; p += fw2 * c.NY; }
    double **ys[4] = {&s.psi_syy_y, &s.psi_sxy_y, &s.psi_vyy, &s.psi_vxy};
    for (auto y : ys) { *y = p; p += fw2 * c.NX; }
}

static int default_segment(const ifos_config *c)
{
    // C = DTINV * ceil(sqrt(NT / DTINV)): about sqrt(NT*DTINV) steps
    const int slots = c->NT / c->DTINV;
    int k = (int)std::ceil(std::sqrt((double)slots));
    if (k < 1) k = 1;
    return c->DTINV * k;
}

extern "C" size_t ifos_fp64_bytes(const ifos_config *c, int segment_steps, int ntr, int nsrc)
{
    if (!c || c->NX < 1 || c->NY < 1 || c->DTINV < 1 || c->NT < 1 || c->NT % c->DTINV) return 0;
    const int nd = c->FDORDER / 2 + 1;
    const long plane = (long)(c->NY + 2 * nd + 1) * (c->NX + 2 * nd);
    const int C = segment_steps > 0 ? segment_steps : default_segment(c);
    if (C % c->DTINV) return 0;
    const long ncp = (c->NT + C - 1) / C;
    const size_t st = state_doubles(c, plane);
    const long cells = (long)c->NX * c->NY;
    size_t d = (size_t)plane * (11 + 3 * c->L) + (size_t)plane * 7 + st * (2 + ncp) + (size_t)(C / c->DTINV) * 5 * cells +
               (size_t)9 * cells + (size_t)ntr * c->NT * 4 + 2 * (size_t)nsrc + 4 * (size_t)ntr;
    return d * sizeof(double) + (size_t)ntr * c->NT * 2 * sizeof(float) + (size_t)nsrc * c->NT * sizeof(float) +
           sizeof(float) * 12 * (c->FW > 0 ? 2 * c->FW : 1) + sizeof(int) * (3 * (size_t)nsrc + 3 * (size_t)ntr);
}

extern "C" int ifos_fp64_create(const ifos_config *cfg, int device, int segment_steps, ifos_fp64 **out)
{
    *out = nullptr;
    if (!cfg || cfg->NX < 1 || cfg->NY < 1 || cfg->FDORDER < 2 || cfg->FDORDER > 12 || (cfg->FDORDER & 1))
        return ifos_set_error("fp64: invalid grid or FDORDER"), -2;
    if (cfg->DTINV < 1 || cfg->NT < 1 || cfg->NT % cfg->DTINV)
        return ifos_set_error("fp64: NT must be a positive multiple of DTINV"), -2;
    if (cfg->L < 0 || cfg->L > 7) return ifos_set_error("fp64: L must be 0..7"), -2;
    if (cfg->PARAMETERIZATION != 1) return ifos_set_error("fp64: PARAMETERIZATION 1 only"), -2;
    if (cfg->BOUNDARY != 0) return ifos_set_error("fp64: BOUNDARY 0 only (no periodic halo copies)"), -2;
    if (cfg->FW < 0 || 2 * cfg->FW > cfg->NX || (!cfg->FREE_SURF && 2 * cfg->FW > cfg->NY) || cfg->FW > cfg->NY)
        return ifos_set_error("fp64: CPML wider than the grid"), -2;
    const int C = segment_steps > 0 ? segment_steps : default_segment(cfg);
    if (C % cfg->DTINV) return ifos_set_error("fp64: segment of %d steps is not a multiple of DTINV", C), -2;
    ifos_fp64 *e = new ifos_fp64();
    e->cfg = *cfg;
    e->device = device;
    e->N = cfg->FDORDER / 2;
    e->nd = e->N + 1;
    e->P = cfg->NX + 2 * e->nd;
    e->rows = cfg->NY + 2 * e->nd + 1;
    e->plane = (long)e->rows * e->P;
    e->npts = (long)cfg->NX * cfg->NY;
    e->C = C;
    e->ncp = (cfg->NT + C - 1) / C;
    e->state_doubles = state_doubles(cfg, e->plane);
    auto fail = [&](const char *what) {
        ifos_set_error("fp64: %s", what);
        f64_free(e);
        if (e->st) hipStreamDestroy(e->st);
        delete e;
        return -1;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess)
        return fail("device setup failed");
    const int L = cfg->L;
    const size_t fw2 = cfg->FW > 0 ? 2 * cfg->FW : 1;
    if (hipMalloc(&e->mat, sizeof(double) * e->plane * (11 + 3 * L)) != hipSuccess ||
        hipMalloc(&e->work, sizeof(double) * e->plane * 7) != hipSuccess ||
        hipMalloc(&e->pml, sizeof(float) * 12 * fw2) != hipSuccess ||
        hipMalloc(&e->fwd, sizeof(double) * e->state_doubles) != hipSuccess ||
        hipMalloc(&e->adj, sizeof(double) * e->state_doubles) != hipSuccess ||
        hipMalloc(&e->ckpt, sizeof(double) * e->state_doubles * e->ncp) != hipSuccess ||
        hipMalloc(&e->snap, sizeof(double) * (size_t)(C / cfg->DTINV) * 5 * e->npts) != hipSuccess ||
        hipMalloc(&e->acc, sizeof(double) * 9 * e->npts) != hipSuccess)
        return fail("out of device memory");
    if (hipMemsetAsync(e->mat, 0, sizeof(double) * e->plane * (11 + 3 * L), e->st) != hipSuccess ||
        hipMemsetAsync(e->work, 0, sizeof(double) * e->plane * 7, e->st) != hipSuccess ||
        hipMemsetAsync(e->acc, 0, sizeof(double) * 9 * e->npts, e->st) != hipSuccess)
        return fail("clearing device memory failed");
    state_views(e, e->fwd, e->sf);
    state_views(e, e->adj, e->sa);
    // geometry and the FP32 setup scalars
    Geo &g = e->g;
    g.NX = cfg->NX; g.NY = cfg->NY; g.N = e->N; g.FW = cfg->FW; g.L = L; g.P = e->P; g.plane = e->plane;
    g.FREE_SURF = cfg->FREE_SURF; g.BOUNDARY = cfg->BOUNDARY; g.VELOCITY = cfg->VELOCITY;
    g.ADJOINT_TYPE = cfg->ADJOINT_TYPE; g.DT = cfg->DT; g.DH = cfg->DH;
    ifos_holberg(cfg->FDORDER, cfg->MAXRELERROR, g.hc);
    std::vector<float> prof(12 * fw2, 0.0f);
    if (cfg->FW > 0) ifos_pml_profiles(cfg, prof.data());
    if (hipMemcpy(e->pml, prof.data(), sizeof(float) * prof.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail("profile upload failed");
    g.pml = e->pml;
    const long off = (long)(e->nd - 1) * e->P + (e->nd - 1);
    double *m = e->mat;
    const double **mats[11] = {&g.rho, &g.vp, &g.vs, &g.taus, &g.taup, &g.rip, &g.rjp, &g.uipjp, &g.fipjp, &g.f, &g.g};
    for (int k = 0; k < 11; k++) *mats[k] = m + (long)k * e->plane + off;
    g.dip = m + 11L * e->plane + off;
    g.d = m + (11L + L) * e->plane + off;
    g.e = m + (11L + 2 * L) * e->plane + off;
    double **wk[5] = {&g.ux, &g.uy, &g.uxy, &g.vxp1, &g.vyp1};
    for (int k = 0; k < 5; k++) *wk[k] = e->work + (long)k * e->plane + off;
    *out = e;
    return 0;
}

extern "C" void ifos_fp64_destroy(ifos_fp64 *e)
{
    if (!e) return;
    hipSetDevice(e->device);
    if (e->st) hipStreamSynchronize(e->st);
    f64_free(e);
    if (e->st) hipStreamDestroy(e->st);
    delete e;
}

// Materials in double from the FP32 models (src/av_rho.c:21-35,
// src/av_mue.c:45-62, src/av_tau.c:21-31, src/prepare_update_s.c:21-108
// with eta from src/readmod.c:47-51), the periodic one-cell halo of a
// single-rank matcopy (src/matcopy_elastic.c:48-140)
extern "C" int ifos_fp64_set_model(ifos_fp64 *e, const float *vp, const float *vs, const float *rho,
                                   const float *taus, const float *taup)
{
    F64_TRY(hipSetDevice(e->device));
    const ifos_config &c = e->cfg;
    const int NX = c.NX, NY = c.NY, L = c.L, P = e->P;
    if (L && (!taus || !taup)) return ifos_set_error("fp64: viscoelastic model needs taus and taup"), -2;
    const long pl = e->plane, off = (long)(e->nd - 1) * P + (e->nd - 1);
    std::vector<double> h((size_t)pl * (11 + 3 * L), 0.0);
    auto G = [&](int k, int j, int i) -> double & { return h[(size_t)k * pl + off + (long)j * P + i]; };
    enum { RHO = 0, VP, VS, TS, TP, RIP, RJP, UIPJP, FIPJP, FF, GG };
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            const size_t k = (size_t)(j - 1) * NX + (i - 1);
            G(RHO, j, i) = rho[k];
            G(VP, j, i) = vp[k];
            G(VS, j, i) = vs[k];
            if (L) {
                G(TS, j, i) = taus[k];
                G(TP, j, i) = taup[k];
            }
        }
    for (int k = RHO; k <= TP; k++) {
        if (!L && (k == TS || k == TP)) continue;
        for (int i = 0; i <= NX + 1; i++) {
            G(k, NY + 1, i) = G(k, 1, i);
            G(k, 0, i) = G(k, NY, i);
        }
        for (int j = 0; j <= NY + 1; j++) {
            G(k, j, NX + 1) = G(k, j, 1);
            G(k, j, 0) = G(k, j, NX);
        }
    }
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            G(RJP, j, i) = 1.0 / (0.5 * (G(RHO, j, i) + G(RHO, j + 1, i)));
            G(RIP, j, i) = 1.0 / (0.5 * (G(RHO, j, i) + G(RHO, j, i + 1)));
            const double u1 = G(RHO, j, i) * G(VS, j, i) * G(VS, j, i);
            const double u2 = G(RHO, j, i + 1) * G(VS, j, i + 1) * G(VS, j, i + 1);
            const double u3 = G(RHO, j + 1, i) * G(VS, j + 1, i) * G(VS, j + 1, i);
            const double u4 = G(RHO, j + 1, i + 1) * G(VS, j + 1, i + 1) * G(VS, j + 1, i + 1);
            G(UIPJP, j, i) = 4.0 / ((1.0 / u1) + (1.0 / u2) + (1.0 / u3) + (1.0 / u4));
            if ((u1 == 0.0) || (u2 == 0.0) || (u3 == 0.0) || (u4 == 0.0)) G(UIPJP, j, i) = 0.0;
        }
    Geo &g = e->g;
    if (L) {
        float pts[9];
        const double PI_D = 3.141592653589793;
        for (int l = 1; l <= L; l++) {
            pts[l] = 1.0 / (2.0 * PI_D * c.FL[l]);
            g.eta[l] = c.DT / pts[l];
        }
        const float ws = 2.0 * PI_D * c.F_REF;
        float sumu = 0.0, sumpi = 0.0;
        for (int l = 1; l <= L; l++) {
            sumu = sumu + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l]));
            sumpi = sumpi + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l]));
        }
        g.sumu = sumu;
        g.sumpi = sumpi;
        for (int l = 1; l <= L; l++) {
            g.bip[l] = 1.0 / (1.0 + (g.eta[l] * 0.5));
            g.bjm[l] = 1.0 / (1.0 + (g.eta[l] * 0.5));
            g.cip[l] = 1.0 - (g.eta[l] * 0.5);
            g.cjm[l] = 1.0 - (g.eta[l] * 0.5);
        }
        auto T = [&](int base, int l, int j, int i) -> double & {
            return h[(size_t)(11 + base * L + (l - 1)) * pl + off + (long)j * P + i];
        };
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                const double tsipjp =
                    0.25 * (G(TS, j, i) + G(TS, j, i + 1) + G(TS, j + 1, i) + G(TS, j + 1, i + 1));
                const double mu = (G(VS, j, i) * G(VS, j, i) * G(RHO, j, i)) / (1.0 + sumu * G(TS, j, i));
                const double pim = (G(VP, j, i) * G(VP, j, i) * G(RHO, j, i)) / (1.0 + sumpi * G(TP, j, i));
                const double muipjp = G(UIPJP, j, i) / (1.0 + sumu * tsipjp);
                G(FIPJP, j, i) = muipjp * c.DT * (1.0 + L * tsipjp);
                G(FF, j, i) = mu * c.DT * (1.0 + L * G(TS, j, i));
                G(GG, j, i) = pim * c.DT * (1.0 + L * G(TP, j, i));
                for (int l = 1; l <= L; l++) {
                    T(0, l, j, i) = muipjp * g.eta[l] * tsipjp;
                    T(1, l, j, i) = mu * g.eta[l] * G(TS, j, i);
                    T(2, l, j, i) = pim * g.eta[l] * G(TP, j, i);
                }
            }
    }
    F64_TRY(hipMemcpy(e->mat, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int ifos_fp64_set_shot(ifos_fp64 *e, int nsrc, const float *srcpos, int ntr, const int *recpos)
{
    F64_TRY(hipSetDevice(e->device));
    const ifos_config &c = e->cfg;
    const int NT = c.NT;
    if (nsrc < 0 || ntr < 0) return ifos_set_error("fp64: negative point count"), -2;
    for (int l = 0; l < nsrc; l++) {
        const int x = (int)srcpos[l * IFOS_SRC_STRIDE], y = (int)srcpos[l * IFOS_SRC_STRIDE + 1];
        if (x < 1 || x > c.NX || y < 1 || y > c.NY) return ifos_set_error("fp64: source %d outside the grid", l), -2;
    }
    for (int r = 0; r < ntr; r++)
        if (recpos[2 * r] < 1 || recpos[2 * r] > c.NX || recpos[2 * r + 1] < 1 || recpos[2 * r + 1] > c.NY)
            return ifos_set_error("fp64: receiver %d outside the grid", r), -2;
    void *ptrs[] = {e->pts_i, e->pts_d, e->fsig, e->seis, e->res, e->obs, e->kill, e->red};
    for (void *p : ptrs)
        if (p) hipFree(p);
    e->pts_i = nullptr; e->pts_d = nullptr; e->fsig = nullptr; e->seis = nullptr;
    e->res = nullptr; e->obs = nullptr; e->kill = nullptr; e->red = nullptr;
    e->nsrc = nsrc;
    e->ntr = ntr;
    const size_t ns = nsrc > 0 ? nsrc : 1, nr = ntr > 0 ? ntr : 1;
    F64_TRY(hipMalloc(&e->pts_i, sizeof(int) * (3 * ns + 2 * nr)));
    F64_TRY(hipMalloc(&e->pts_d, sizeof(double) * 2 * ns));
    F64_TRY(hipMalloc(&e->fsig, sizeof(float) * ns * NT));
    F64_TRY(hipMalloc(&e->seis, sizeof(double) * 2 * nr * NT));
    F64_TRY(hipMalloc(&e->res, sizeof(double) * 2 * nr * NT));
    F64_TRY(hipMalloc(&e->obs, sizeof(float) * 2 * nr * NT));
    F64_TRY(hipMalloc(&e->kill, sizeof(int) * nr));
    F64_TRY(hipMalloc(&e->red, sizeof(double) * 4 * nr));
    std::vector<int> hi(3 * ns + 2 * nr, 0);
    std::vector<double> hd(2 * ns, 0.0);
    const double PI_D = 3.141592653589793;
    for (int l = 0; l < nsrc; l++) {
        const float *s = srcpos + l * IFOS_SRC_STRIDE;
        hi[l] = (int)s[0];
        hi[ns + l] = (int)s[1];
        hi[2 * ns + l] = (int)s[7];
        const float azi_rad = s[6] * PI_D / 180;
        hd[l] = sin(azi_rad);
        hd[ns + l] = cos(azi_rad);
    }
    for (int r = 0; r < ntr; r++) {
        hi[3 * ns + r] = recpos[2 * r];
        hi[3 * ns + nr + r] = recpos[2 * r + 1];
    }
    // SOURCE_TYPE is rewritten from the last source by update_v_PML
    // (src/update_v_PML.c:558) before psource reads it
    e->explosive = c.SOURCE_TYPE == 1;
    if (nsrc > 0) e->explosive = hi[2 * ns + nsrc - 1] == 1;
    std::vector<long> cells;
    for (int l = 0; l < nsrc; l++) cells.push_back((long)hi[ns + l] * (c.NX + 1) + hi[l]);
    std::sort(cells.begin(), cells.end());
    e->src_serial = std::adjacent_find(cells.begin(), cells.end()) != cells.end();
    cells.clear();
    for (int r = 0; r < ntr; r++) cells.push_back((long)recpos[2 * r + 1] * (c.NX + 1) + recpos[2 * r]);
    std::sort(cells.begin(), cells.end());
    e->rec_serial = std::adjacent_find(cells.begin(), cells.end()) != cells.end();
    std::vector<float> sig(ns * NT, 0.0f);
    if (nsrc > 0) ifos_wavelet(&c, nsrc, srcpos, sig.data());
    F64_TRY(hipMemcpy(e->pts_i, hi.data(), sizeof(int) * hi.size(), hipMemcpyHostToDevice));
    F64_TRY(hipMemcpy(e->pts_d, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice));
    F64_TRY(hipMemcpy(e->fsig, sig.data(), sizeof(float) * sig.size(), hipMemcpyHostToDevice));
    return 0;
}

namespace {

template <int N>
void launch_step(ifos_fp64 *e, const Geo &g, const State &s, int nt, int sw)
{
    const ifos_config &c = e->cfg;
    const dim3 grid = grid2(c.NX, c.NY), blk(kBX, kBY);
    const size_t ns = e->nsrc > 0 ? e->nsrc : 1, nr = e->ntr > 0 ? e->ntr : 1;
    hipLaunchKernelGGL(k64_vel<N>, grid, blk, 0, e->st, g, s, sw);
    Points p{};
    p.NT = c.NT;
    if (sw == 0) {
        p.n = e->nsrc;
        p.ix = e->pts_i;
        p.iy = e->pts_i + ns;
        p.type = e->pts_i + 2 * ns;
        p.sin_azi = e->pts_d;
        p.cos_azi = e->pts_d + ns;
        p.fsig = e->fsig;
    } else {
        p.n = e->ntr;
        p.ix = e->pts_i + 3 * ns;
        p.iy = e->pts_i + 3 * ns + nr;
        p.res0 = e->res;
        p.res1 = e->res + nr * c.NT;
    }
    if (p.n > 0) {
        const int serial = sw == 0 ? e->src_serial : e->rec_serial;
        hipLaunchKernelGGL(k64_points, dim3(serial ? 1 : (p.n + 255) / 256), dim3(serial ? 1 : 256), 0, e->st, g, s,
                           p, nt, sw, serial);
    }
    hipLaunchKernelGGL(k64_str<N>, grid, blk, 0, e->st, g, s);
    if (sw == 0 && e->explosive && e->nsrc > 0) hipLaunchKernelGGL(k64_psource, dim3(1), dim3(1), 0, e->st, g, s, p, nt);
    if (c.FREE_SURF) hipLaunchKernelGGL(k64_surface<N>, dim3((c.NX + 255) / 256), dim3(256), 0, e->st, g, s);
}

void step(ifos_fp64 *e, const Geo &g, const State &s, int nt, int sw)
{
    switch (e->N) {
    case 1: launch_step<1>(e, g, s, nt, sw); break;
    case 2: launch_step<2>(e, g, s, nt, sw); break;
    case 3: launch_step<3>(e, g, s, nt, sw); break;
    case 4: launch_step<4>(e, g, s, nt, sw); break;
    case 5: launch_step<5>(e, g, s, nt, sw); break;
    default: launch_step<6>(e, g, s, nt, sw); break;
    }
}

}  // namespace

// forward steps [nt0, nt1] on the forward state; with record != 0 the
// seismograms are sampled and the state before every segment start is kept;
// with snaps != 0 the snapshot steps are stored in the segment buffer,
// slot (nt - seg0 - 1) / DTINV
static int run_fwd(ifos_fp64 *e, int nt0, int nt1, bool record, bool snaps, int seg0)
{
    const ifos_config &c = e->cfg;
    const dim3 grid = grid2(c.NX, c.NY), blk(kBX, kBY);
    const size_t nr = e->ntr > 0 ? e->ntr : 1, ns = e->nsrc > 0 ? e->nsrc : 1;
    for (int nt = nt0; nt <= nt1; nt++) {
        if (record && (nt - 1) % e->C == 0)
            F64_TRY(hipMemcpyAsync(e->ckpt + (size_t)((nt - 1) / e->C) * e->state_doubles, e->fwd,
                                   sizeof(double) * e->state_doubles, hipMemcpyDeviceToDevice, e->st));
        step(e, e->g, e->sf, nt, 0);
        if (record && e->ntr > 0)
            hipLaunchKernelGGL(k64_seismo, dim3((e->ntr + 255) / 256), dim3(256), 0, e->st, e->g, e->sf, e->ntr,
                               e->pts_i + 3 * ns, e->pts_i + 3 * ns + nr, c.NT, nt, e->seis, e->seis + nr * c.NT);
        if (snaps && (nt - 1) % c.DTINV == 0)
            hipLaunchKernelGGL(k64_snap, grid, blk, 0, e->st, e->g, e->sf, e->snap, (long)((nt - seg0 - 1) / c.DTINV));
    }
    F64_TRY(hipGetLastError());
    return 0;
}

extern "C" int ifos_fp64_shot_gradient(ifos_fp64 *e, const float *obs_vx, const float *obs_vy, const int *kill,
                                       double *misfit)
{
    F64_TRY(hipSetDevice(e->device));
    const ifos_config &c = e->cfg;
    const int NT = c.NT, NX = c.NX, NY = c.NY;
    const size_t nr = e->ntr > 0 ? e->ntr : 1;
    const long pn = e->npts;
    // forward run with checkpoints and seismograms
    F64_TRY(hipMemsetAsync(e->fwd, 0, sizeof(double) * e->state_doubles, e->st));
    F64_TRY(hipMemsetAsync(e->work, 0, sizeof(double) * e->plane * 5, e->st));
    if (run_fwd(e, 1, NT, true, false, 0)) return -1;
    // residuals (src/IFOS2D.c:1965-2060; LNORM 2), misfit and energy in FP64
    std::vector<int> hk(nr, 0);
    if (kill && c.TRKILL)
        for (int r = 0; r < e->ntr; r++) hk[r] = kill[r];
    F64_TRY(hipMemcpyAsync(e->kill, hk.data(), sizeof(int) * nr, hipMemcpyHostToDevice, e->st));
    F64_TRY(hipMemsetAsync(e->res, 0, sizeof(double) * 2 * nr * NT, e->st));
    double l2 = 0.0, en = 0.0;
    const float *obs[2] = {obs_vx, obs_vy};
    const bool use[2] = {c.ADJOINT_TYPE == 1 || c.ADJOINT_TYPE == 3, c.ADJOINT_TYPE == 1 || c.ADJOINT_TYPE == 2};
    std::vector<double> red(2 * nr);
    for (int k = 0; k < 2; k++) {
        if (!use[k] || e->ntr == 0) continue;
        if (!obs[k]) return ifos_set_error("fp64: observed traces of an active component missing"), -2;
        F64_TRY(hipMemcpyAsync(e->obs + k * nr * NT, obs[k], sizeof(float) * e->ntr * NT, hipMemcpyHostToDevice, e->st));
        hipLaunchKernelGGL(k64_residual, dim3((e->ntr + 63) / 64), dim3(64), 0, e->st, e->ntr, NT, c.VELOCITY, c.DT,
                           e->seis + k * nr * NT, e->obs + k * nr * NT, e->kill, e->res + k * nr * NT, e->red,
                           e->red + nr);
        F64_TRY(hipMemcpyAsync(red.data(), e->red, sizeof(double) * 2 * nr, hipMemcpyDeviceToHost, e->st));
        F64_TRY(hipStreamSynchronize(e->st));
        for (int r = 0; r < e->ntr; r++) {
            l2 += red[r];
            en += red[nr + r];
        }
    }
    if (misfit) {
        misfit[0] = l2;
        misfit[1] = en;
    }
    // adjoint run; forward segments rebuilt from their checkpoints in reverse
    F64_TRY(hipMemsetAsync(e->adj, 0, sizeof(double) * e->state_doubles, e->st));
    F64_TRY(hipMemsetAsync(e->acc, 0, sizeof(double) * 3 * pn, e->st));
    double *wc = e->acc, *wc_u = e->acc + pn, *wc_rho = e->acc + 2 * pn;
    // the adjoint run keeps its own vxp1/vyp1 (its integrated velocity with
    // VELOCITY 0 is state), since the forward rebuild rewrites the forward's
    // every step; ux/uy/uxy are rewritten by every step and only read for
    // forward snapshots, so the runs share them
    const long off = (long)(e->nd - 1) * e->P + (e->nd - 1);
    Geo ga = e->g;
    ga.vxp1 = e->work + 5 * e->plane + off;
    ga.vyp1 = e->work + 6 * e->plane + off;
    F64_TRY(hipMemsetAsync(e->work + 5 * e->plane, 0, sizeof(double) * 2 * e->plane, e->st));
    const dim3 grid = grid2(NX, NY), blk(kBX, kBY);
    int loaded = -1;
    int rc = 0;
    for (int nt = 1; nt <= NT && !rc; nt++) {
        step(e, ga, e->sa, nt, 1);
        if ((NT - nt) % c.DTINV == 0) {
            const int nf = NT - nt + 1;  // forward step paired with this adjoint step
            const int q = (nf - 1) / e->C;
            if (q != loaded) {
                // rebuild segment q: restore its checkpoint, run its steps storing snapshots
                F64_TRY(hipMemcpyAsync(e->fwd, e->ckpt + (size_t)q * e->state_doubles,
                                       sizeof(double) * e->state_doubles, hipMemcpyDeviceToDevice, e->st));
                const int a = q * e->C + 1, b = std::min(NT, (q + 1) * e->C);
                rc = run_fwd(e, a, b, false, true, q * e->C);
                loaded = q;
            }
            const long k = (nf - q * e->C - 1) / c.DTINV;
            hipLaunchKernelGGL(k64_corr, grid, blk, 0, e->st, ga, e->sa, e->snap, k, wc, wc_u, wc_rho);
        }
    }
    if (!rc) {
        hipLaunchKernelGGL(k64_chain, grid, blk, 0, e->st, e->g, wc, wc_u, wc_rho, e->acc + 3 * pn, e->acc + 4 * pn,
                           e->acc + 5 * pn, e->acc + 6 * pn, e->acc + 7 * pn, e->acc + 8 * pn);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(e->st) != hipSuccess) {
            ifos_set_error("fp64: adjoint run failed");
            rc = -1;
        }
    }
    hipStreamSynchronize(e->st);
    return rc;
}

extern "C" int ifos_fp64_seismograms(ifos_fp64 *e, double *vx, double *vy)
{
    F64_TRY(hipSetDevice(e->device));
    const size_t nr = e->ntr > 0 ? e->ntr : 1, n = (size_t)e->ntr * e->cfg.NT;
    if (vx) F64_TRY(hipMemcpy(vx, e->seis, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (vy) F64_TRY(hipMemcpy(vy, e->seis + nr * e->cfg.NT, sizeof(double) * n, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int ifos_fp64_gradient(ifos_fp64 *e, int which, double *g_vp, double *g_vs, double *g_rho)
{
    F64_TRY(hipSetDevice(e->device));
    const long pn = e->npts;
    const double *base = e->acc + (which ? 6 : 3) * pn;
    double *out[3] = {g_vp, g_vs, g_rho};
    for (int k = 0; k < 3; k++)
        if (out[k]) F64_TRY(hipMemcpy(out[k], base + k * pn, sizeof(double) * pn, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int ifos_fp64_reset_gradient(ifos_fp64 *e)
{
    F64_TRY(hipSetDevice(e->device));
    F64_TRY(hipMemset(e->acc + 6 * e->npts, 0, sizeof(double) * 3 * e->npts));
    return 0;
}

extern "C" int ifos_fp64_segment_steps(ifos_fp64 *e) { return e->C; }
