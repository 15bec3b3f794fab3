// ifos_step.hip -- fused time-step kernels of the native engine.
//
// One time step is two kernels:
//   k_vel_step   velocity update (src/update_v_PML.c) + the point forces or
//                adjoint residual injection that follow it (:553-578) + the
//                acceleration snapshot (forward) or the adjoint vxp1 update
//   k_str_step   stress update (src/update_s_*_PML.c) + explosive source
//                (src/psource.c) + free surface (src/surface_*_PML.c) +
//                snapshot store (src/IFOS2D.c:1805-1871) or zero-lag
//                correlation (:2381-2419) + receiver sampling
//                (src/seismo_ssg.c) + stability probe
// Each thread owns one cell of a 64x8 tile; point sources and receivers are
// bucketed per tile on the host (CSR lists in source order), so a thread only
// scans the few points of its own tile.  The free surface needs the new
// stresses of rows 2..N+1 of its column: tiles are 8 rows high and N <= 6, so
// the top tile row holds them; they are exchanged through LDS after a barrier.
// The arithmetic is the per-cell code of ifos_cellmath.h, bit-identical to
// the reference.
#include "ifos_cellmath.h"
#include "ifos_step.h"

namespace ifos {

__device__ __forceinline__ int step_number(const StepArgs &a, int k) { return a.base ? *a.base + k + 1 : k; }

// Workgroups are dealt round-robin to the 8 XCDs (workgroup h runs on XCD
// h % 8), and each XCD has its own L2.  With swizzle on, XCD x gets one
// contiguous run of tiles in row-major order, so the x- and y-halos a tile
// shares with its neighbours are mostly fetched into the same L2 once.
__device__ __forceinline__ int tile_of(const StepArgs &a, int h)
{
    const int nblk = a.nbx * a.nby;
    if (!a.swizzle) return h;
    const int per = nblk / 8, rem = nblk % 8, x = h % 8, loc = h / 8;
    return x < rem ? x * (per + 1) + loc : rem * (per + 1) + (x - rem) * per + loc;
}

// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(512) void k_vel_step(StepArgs a, int k)
{
    const StepGeom &g = a.g;
    const int blk = tile_of(a, blockIdx.x);
    const int bx = blk % a.nbx, by = blk / a.nbx;
    const int i = bx * kTileX + threadIdx.x + 1;
    const int j = by * kTileY + threadIdx.y + 1;
    if (i > g.NX || j > g.NY) return;
    const int nt = step_number(a, k);
    const int P = g.pitch;
    const long c = (long)j * P + i;
    float sxx_x, sxy_x, sxy_y, syy_y;
    stress_derivs<N>(g.hc, a.sxx + c, a.syy + c, a.sxy + c, P, sxx_x, sxy_x, sxy_y, syy_y);
    cpml_velocity(g, j, i, a.psi_sxx_x, a.psi_sxy_x, a.psi_syy_y, a.psi_sxy_y, sxx_x, sxy_x, sxy_y, syy_y);
    const float rip = a.rip[c], rjp = a.rjp[c];
    const float vx0 = a.vx[c], vy0 = a.vy[c];
    const long e = (long)(j - 1) * g.NX + (i - 1);
    if (a.sw == 0) {
        if (a.snap[0] && (nt - 1) % a.DTINV == 0) {
            const long o = (long)((nt - 1) / a.DTINV) * a.plane + e;
            a.snap[0][o] = vel_rate(rip, sxx_x + sxy_y, g.DH);
            a.snap[1][o] = vel_rate(rjp, sxy_x + syy_y, g.DH);
        }
    } else if (a.VELOCITY == 0) {
        a.vxp1[c] += vx0 * g.DT;
        a.vyp1[c] += vy0 * g.DT;
    } else if ((a.NT - nt) % a.DTINV == 0) {
        a.vxp1[c] = vx0;
        a.vyp1[c] = vy0;
    }
    float vx = vel_new(vx0, g.DT, rip, sxx_x + sxy_y, g.DH);
    float vy = vel_new(vy0, g.DT, rjp, sxy_x + syy_y, g.DH);

    // point forces (forward) / residual injection (adjoint) of this tile, in order
    const int b0 = a.inj.start[blk], b1 = a.inj.start[blk + 1];
    for (int q = b0; q < b1; q++) {
        const int l = a.inj.idx[q];
        if (a.inj.ix[l] != i || a.inj.iy[l] != j) continue;
        const float s0 = a.inj.sig[(long)l * a.NT + nt - 1];
        const float s1 = a.inj.sig1[(long)l * a.NT + nt - 1];
        if (a.sw == 0) {
            const int type = a.inj.type[l];
            if (type == 2) vx += (g.DT * rip * s1) / (g.DH * g.DH);
            if (type == 3) vy += (g.DT * rjp * s1) / (g.DH * g.DH);
            if (type == 4) {
                vx += (g.DT * rip * a.inj.sin_azi[l] * s0) / (g.DH * g.DH);
                vy += (g.DT * rjp * a.inj.cos_azi[l] * s0) / (g.DH * g.DH);
            }
        } else {
            if (a.ADJOINT_TYPE == 1) { vx += s0; vy += s1; }
            if (a.ADJOINT_TYPE == 2) vy += s1;
            if (a.ADJOINT_TYPE == 3) vx += s0;
        }
    }
    a.vx[c] = vx;
    a.vy[c] = vy;
}

// ---------------------------------------------------------------------------
template <int N, int L>
__global__ __launch_bounds__(512) void k_str_step(StepArgs a, int k)
{
    __shared__ float s_syy[kTileY][kTileX];
    __shared__ float s_sxy[kTileY][kTileX];
    const StepGeom &g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int blk = tile_of(a, blockIdx.x);
    const int bx = blk % a.nbx, by = blk / a.nbx;
    const int i = bx * kTileX + tx + 1;
    const int j = by * kTileY + ty + 1;
    const bool live = i <= g.NX && j <= g.NY;
    const bool top = by == 0 && g.FREE_SURF;
    const int nt = step_number(a, k);
    const int P = g.pitch;
    const long c = (long)j * P + i;
    const long e = (long)(j - 1) * g.NX + (i - 1);
    const bool snap_step = a.sw == 0 && a.snap[0] && (nt - 1) % a.DTINV == 0;
    const bool corr_step = a.sw == 1 && (a.NT - nt) % a.DTINV == 0;
    const bool rates = snap_step && a.VELOCITY == 1;

    float sxx = 0.0f, syy = 0.0f, sxy = 0.0f, ux = 0.0f, uy = 0.0f, uxy = 0.0f;
    float rr[L > 0 ? L : 1], pp[L > 0 ? L : 1], qq[L > 0 ? L : 1];
    ViscCell vc;
    float fe = 0.0f, ge = 0.0f;
    if (live) {
        float vxx, vyx, vxy, vyy;
        const float scale = L ? (float)(1.0 / (double)g.DH) : g.DT / g.DH;
        velocity_derivs<N>(g.hc, a.vx + c, a.vy + c, P, scale, vxx, vyx, vxy, vyy);
        cpml_stress(g, j, i, a.psi_vxx, a.psi_vyx, a.psi_vyy, a.psi_vxy, vxx, vyx, vxy, vyy);
        sxx = a.sxx[c];
        syy = a.syy[c];
        sxy = a.sxy[c];
        if (L == 0) {
            fe = a.f[c];
            ge = a.gl[c];
            const float fipjp = a.uipjp[c];
            if (rates) rates_elastic(fe, ge, fipjp, g.DT, vxx, vyx, vxy, vyy, ux, uy, uxy);
            stress_elastic(fe, ge, fipjp, vxx, vyx, vxy, vyy, sxx, syy, sxy);
        } else {
            visc_coeffs(L, g.DT, a.mu[c], a.pi[c], a.muipjp[c], a.taus[c], a.taup[c], a.tausipjp[c], vc);
#pragma unroll
            for (int l = 0; l < L; l++) {
                rr[l] = a.r[c + (long)l * a.ls];
                pp[l] = a.p[c + (long)l * a.ls];
                qq[l] = a.q[c + (long)l * a.ls];
            }
            stress_visc<(L > 0 ? L : 1)>(vc, a.eta, a.bip, a.bjm, a.cip, a.cjm, g.DT, vxx, vyx, vxy, vyy, sxx, syy,
                                       sxy, rr, pp, qq, rates, ux, uy, uxy);
        }
        // explosive sources of this tile, applied after the stress update
        // and before the free surface (src/IFOS2D.c:1758-1773)
        if (a.sw == 0 && a.explosive) {
            for (int q = a.exp.start[blk]; q < a.exp.start[blk + 1]; q++) {
                const int l = a.exp.idx[q];
                if (a.exp.ix[l] != i || a.exp.iy[l] != j) continue;
                const float amp = explosive_amp(a.exp.sig + (long)l * a.NT - 1, nt, a.NT, g.DH);
                sxx = add_half(sxx, amp);
                syy = add_half(syy, amp);
            }
        }
    }
    if (top) {
        // the free surface of column i reads the new syy of rows 2..N+1 and
        // sxy of rows 1..N; they come from the other threads of this tile
        s_syy[ty][tx] = live ? syy : 0.0f;
        s_sxy[ty][tx] = live ? sxy : 0.0f;
        __syncthreads();
        if (live && j == 1) {
            syy = 0.0f;
            if (rates) uy = 0.0f;
            if (L) {
#pragma unroll
                for (int l = 0; l < L; l++) qq[l] = 0.0f;
            }
#pragma unroll
            for (int m = 1; m <= N; m++) {
                a.syy[c - (long)m * P] = -s_syy[m][tx];
                a.sxy[c - (long)m * P] = -s_sxy[m - 1][tx];
            }
            float vxx, vyy;
            surface_derivs<N>(g, i, a.vx + c, a.vy + c, P, a.psi_vxxs, vxx, vyy);
            if (L == 0) {
                surface_elastic_update((float)((double)fe * 2.0), ge, g.DT, vxx, vyy, sxx, rates, ux);
            } else {
                surface_visc_update<(L > 0 ? L : 1)>(vc.mu, vc.pi, vc.taus, vc.taup, a.eta, g.DT, vxx, vyy, sxx, pp,
                                                     rates, ux);
            }
        }
    }
    if (!live) return;
    a.sxx[c] = sxx;
    a.syy[c] = syy;
    a.sxy[c] = sxy;
    if (L) {
#pragma unroll
        for (int l = 0; l < L; l++) {
            a.r[c + (long)l * a.ls] = rr[l];
            a.p[c + (long)l * a.ls] = pp[l];
            a.q[c + (long)l * a.ls] = qq[l];
        }
    }
    if (snap_step) {
        const long o = (long)((nt - 1) / a.DTINV) * a.plane + e;
        a.snap[2][o] = a.VELOCITY ? ux : sxx;
        a.snap[3][o] = a.VELOCITY ? uy : syy;
        a.snap[4][o] = a.VELOCITY ? uxy : sxy;
    }
    if (corr_step) {
        const long o = (long)((a.NT - nt) / a.DTINV) * a.plane + e;
        float wr = a.wc_rho[e], w = a.wc[e], wu = a.wc_u[e];
        correlate_cell(a.vxp1[c], a.vyp1[c], sxx, syy, sxy, a.snap[0][o], a.snap[1][o], a.snap[2][o], a.snap[3][o],
                       a.snap[4][o], a.muss[e], a.lamss[e], a.vs_i[e] > 0.0f, wr, w, wu);
        a.wc_rho[e] = wr;
        a.wc[e] = w;
        a.wc_u[e] = wu;
    }
    if (a.sw == 0) {
        for (int q = a.rec.start[blk]; q < a.rec.start[blk + 1]; q++) {
            const int t = a.rec.idx[q];
            if (a.rec.ix[t] != i || a.rec.iy[t] != j) continue;
            a.seis_vx[(long)t * a.NT + nt - 1] = a.vx[c];
            a.seis_vy[(long)t * a.NT + nt - 1] = a.vy[c];
        }
    }
    if (a.nan_flag && j == g.NY / 2 && i == g.NX / 2) {
        const float v = a.vy[c];
        if (isnan(v) && *a.nan_flag == 0) *a.nan_flag = nt;
    }
}

__global__ void k_tick(int *base, int by)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *base += by;
}

// ---------------------------------------------------------------------------
void launch_vel_step(const StepArgs &a, int k, hipStream_t s)
{
    dim3 b(kTileX, kTileY), gr(a.nbx * a.nby);
    switch (a.g.N) {
    case 1: hipLaunchKernelGGL(k_vel_step<1>, gr, b, 0, s, a, k); break;
    case 2: hipLaunchKernelGGL(k_vel_step<2>, gr, b, 0, s, a, k); break;
    case 3: hipLaunchKernelGGL(k_vel_step<3>, gr, b, 0, s, a, k); break;
    case 4: hipLaunchKernelGGL(k_vel_step<4>, gr, b, 0, s, a, k); break;
    case 5: hipLaunchKernelGGL(k_vel_step<5>, gr, b, 0, s, a, k); break;
    case 6: hipLaunchKernelGGL(k_vel_step<6>, gr, b, 0, s, a, k); break;
    default: break;
    }
}

template <int N>
static void str_n(const StepArgs &a, int k, hipStream_t s)
{
    dim3 b(kTileX, kTileY), gr(a.nbx * a.nby);
    switch (a.L) {
    case 0: hipLaunchKernelGGL((k_str_step<N, 0>), gr, b, 0, s, a, k); break;
    case 1: hipLaunchKernelGGL((k_str_step<N, 1>), gr, b, 0, s, a, k); break;
    case 2: hipLaunchKernelGGL((k_str_step<N, 2>), gr, b, 0, s, a, k); break;
    case 3: hipLaunchKernelGGL((k_str_step<N, 3>), gr, b, 0, s, a, k); break;
    case 4: hipLaunchKernelGGL((k_str_step<N, 4>), gr, b, 0, s, a, k); break;
    case 5: hipLaunchKernelGGL((k_str_step<N, 5>), gr, b, 0, s, a, k); break;
    default: break;
    }
}

void launch_str_step(const StepArgs &a, int k, hipStream_t s)
{
    switch (a.g.N) {
    case 1: str_n<1>(a, k, s); break;
    case 2: str_n<2>(a, k, s); break;
    case 3: str_n<3>(a, k, s); break;
    case 4: str_n<4>(a, k, s); break;
    case 5: str_n<5>(a, k, s); break;
    case 6: str_n<6>(a, k, s); break;
    default: break;
    }
}

void launch_tick(int *base, int by, hipStream_t s) { hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s, base, by); }

}  // namespace ifos
