// ifos_kernels.hip -- gfx950 kernels of the P-SV hot path.
//
// Arithmetic contract: every kernel evaluates the reference expression in the
// same operand order, with the same float->double promotions of `2.0*` and
// `0.5*` literals and true divisions where the reference divides.  The
// library is compiled with -ffp-contract=off (no FMA contraction, like gcc
// -std=c99 on x86-64) and with IEEE denormals, so FP32 results are
// bit-identical to the reference's CPU run.
//
// Step counter: kernels that need the time-step number read it from a
// device int (nt_dev) so that a captured graph of a block of time steps can
// be replayed.  The single-block inject kernel advances the counter; kernels
// that run before it within a step (update_v) see the previous value.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ifos_cellmath.h"
#include "ifos_internal.h"

namespace ifos {

#define IFOS_KERNEL __global__ __launch_bounds__(256)

__device__ __forceinline__ int step_of(const int *nt_dev, int nt) { return nt_dev ? *nt_dev : nt; }

// Tile of this workgroup for the update kernels: workgroups are dealt
// round-robin to the 8 XCDs (workgroup h runs on XCD h % 8) and each XCD has
// its own L2, so XCD x works through one contiguous run [xlo[x], xlo[x+1])
// of the row-major 64x4 tiles and the halo rows neighbouring tiles share are
// fetched into one L2.  False for the workgroups past the end of a run.
__device__ __forceinline__ bool tile_of_block(const StepGeom &g, int &i, int &j)
{
    const int h = blockIdx.x, x = h & 7;
    const int t = g.xlo[x] + (h >> 3);
    if (t >= g.xlo[x + 1]) return false;
    const int by = t / g.tnbx, bx = t - by * g.tnbx;
    i = bx * blockDim.x + threadIdx.x + 1;
    j = by * blockDim.y + threadIdx.y + 1;
    return true;
}

// ---------------------------------------------------------------------------
// update_v_PML  (src/update_v_PML.c:29-596; orders 2-8 of the reference and
// the same form for orders 10/12)
// ---------------------------------------------------------------------------
template <int N>
IFOS_KERNEL void k_update_v(VArgs a, const int *nt_dev, int DTINV, int NT)
{
    const StepGeom &g = a.g;
    int i, j;
    if (!tile_of_block(g, i, j) || i > g.NX || j > g.NY) return;
    const int P = g.pitch;
    const long c = (long)j * P + i;
    float sxx_x, sxy_x, sxy_y, syy_y;
    stress_derivs<N>(g.hc, a.sxx + c, a.syy + c, a.sxy + c, P, sxx_x, sxy_x, sxy_y, syy_y);
    cpml_velocity(g, j, i, a.psi_sxx_x, a.psi_sxy_x, a.psi_syy_y, a.psi_sxy_y, sxx_x, sxy_x, sxy_y, syy_y);

    const float rip = a.rip[c], rjp = a.rjp[c];
    const float vx = a.vx[c], vy = a.vy[c];
    if (a.vxp1) {
        // this kernel runs before the step counter advances: nt = *nt_dev + 1
        const int nt = nt_dev ? *nt_dev + 1 : 0;
        if (a.sw == 0) {
            if (!nt_dev || (nt - 1) % DTINV == 0) {
                a.vxp1[c] = rip * (sxx_x + sxy_y) / g.DH;
                a.vyp1[c] = rjp * (sxy_x + syy_y) / g.DH;
            }
        } else if (a.VELOCITY == 0) {
            a.vxp1[c] += vx * g.DT;
            a.vyp1[c] += vy * g.DT;
        } else if (!nt_dev || (NT - nt) % DTINV == 0) {
            a.vxp1[c] = vx;
            a.vyp1[c] = vy;
        }
    }
    a.vx[c] = vx + g.DT * rip * (sxx_x + sxy_y) / g.DH;
    a.vy[c] = vy + g.DT * rjp * (sxy_x + syy_y) / g.DH;
}

// ---------------------------------------------------------------------------
// point forces after the velocity update (src/update_v_PML.c:553-578) and
// the step counter.  One block.  Forward sources are applied serially in
// source order (several may share a cell); adjoint sources sit on distinct
// receiver cells and are applied in parallel unless serial is requested.
// ---------------------------------------------------------------------------
IFOS_KERNEL void k_inject(PointArgs a, StepGeom g, float *vx, float *vy, const float *rip, const float *rjp,
                          int sw, int ADJOINT_TYPE, const int *nt_dev, int nt_arg, int *nt_tick, int serial)
{
    __shared__ int s_nt;
    if (threadIdx.x == 0) s_nt = nt_tick ? (*nt_tick + 1) : step_of(nt_dev, nt_arg);
    __syncthreads();
    const int nt = s_nt;
    const int P = g.pitch;
    if (sw == 0) {
        if (threadIdx.x == 0) {
            for (int l = 0; l < a.n; l++) {
                const long c = (long)a.iy[l] * P + a.ix[l];
                const float s0 = a.sig[(long)l * a.ld + nt - 1];
                const float s1 = a.sig1[(long)l * a.ld + nt - 1];
                const int type = a.type[l];
                if (type == 2) vx[c] += (g.DT * rip[c] * s1) / (g.DH * g.DH);
                if (type == 3) vy[c] += (g.DT * rjp[c] * s1) / (g.DH * g.DH);
                if (type == 4) {
                    vx[c] += (g.DT * rip[c] * a.sin_azi[l] * s0) / (g.DH * g.DH);
                    vy[c] += (g.DT * rjp[c] * a.cos_azi[l] * s0) / (g.DH * g.DH);
                }
            }
        }
    } else if (serial) {
        if (threadIdx.x == 0) {
            for (int l = 0; l < a.n; l++) {
                const long c = (long)a.iy[l] * P + a.ix[l];
                const float s0 = a.sig[(long)l * a.ld + nt - 1];
                const float s1 = a.sig1[(long)l * a.ld + nt - 1];
                if (ADJOINT_TYPE == 1) { vx[c] += s0; vy[c] += s1; }
                if (ADJOINT_TYPE == 2) vy[c] += s1;
                if (ADJOINT_TYPE == 3) vx[c] += s0;
            }
        }
    } else {
        for (int l = threadIdx.x; l < a.n; l += blockDim.x) {
            const long c = (long)a.iy[l] * P + a.ix[l];
            const float s0 = a.sig[(long)l * a.ld + nt - 1];
            const float s1 = a.sig1[(long)l * a.ld + nt - 1];
            if (ADJOINT_TYPE == 1) { vx[c] += s0; vy[c] += s1; }
            if (ADJOINT_TYPE == 2) vy[c] += s1;
            if (ADJOINT_TYPE == 3) vx[c] += s0;
        }
    }
    __syncthreads();
    if (nt_tick && threadIdx.x == 0) *nt_tick = nt;
}

// ---------------------------------------------------------------------------
// update_s_elastic_PML  (src/update_s_elastic_PML.c:28-569)
// ---------------------------------------------------------------------------
template <int N>
IFOS_KERNEL void k_update_s_elastic(SElArgs a, const int *nt_dev, int DTINV)
{
    const StepGeom &g = a.g;
    int i, j;
    if (!tile_of_block(g, i, j) || i > g.NX || j > g.NY) return;
    const long c = (long)j * g.pitch + i;
    const float dhi = g.DT / g.DH;
    float vxx, vyx, vxy, vyy;
    velocity_derivs<N>(g.hc, a.vx + c, a.vy + c, g.pitch, dhi, vxx, vyx, vxy, vyy);
    cpml_stress(g, j, i, a.psi_vxx, a.psi_vyx, a.psi_vyy, a.psi_vxy, vxx, vyx, vxy, vyy);

    const float fipjp = a.uipjp[c];
    float f, gl;
    if (a.f) {
        f = a.f[c];
        gl = a.gl[c];
    } else {
        const float rho = a.rho[c], u = a.vs[c], pi = a.vp[c];
        f = rho * u * u;
        gl = rho * ((pi * pi) - 2 * u * u);
    }
    const float sxx = a.sxx[c], syy = a.syy[c], sxy = a.sxy[c];
    a.sxy[c] = sxy + fipjp * (vyx + vxy);
    a.sxx[c] = (float)((double)sxx + ((double)(gl * (vxx + vyy)) + (2.0 * (double)f * (double)vxx)));
    a.syy[c] = (float)((double)syy + ((double)(gl * (vxx + vyy)) + (2.0 * (double)f * (double)vyy)));
    if (a.ux && (!nt_dev || (*nt_dev - 1) % DTINV == 0)) {
        a.uxy[c] = (fipjp / g.DT) * (vyx + vxy);
        const double f2 = (2.0 * (double)f) / (double)g.DT;
        a.ux[c] = (float)((double)((gl / g.DT) * (vxx + vyy)) + (f2 * (double)vxx));
        a.uy[c] = (float)((double)((gl / g.DT) * (vxx + vyy)) + (f2 * (double)vyy));
    }
}

// ---------------------------------------------------------------------------
// update_s_visc_PML  (src/update_s_visc_PML.c:28-515), generalized
// standard linear solid with L relaxation mechanisms
// ---------------------------------------------------------------------------
template <int N, int L>
IFOS_KERNEL void k_update_s_visc(SViscArgs a, const int *nt_dev, int DTINV)
{
    const StepGeom &g = a.g;
    int i, j;
    if (!tile_of_block(g, i, j) || i > g.NX || j > g.NY) return;
    const long c = (long)j * g.pitch + i;
    const float dhi = 1.0 / (double)g.DH;
    const float dthalbe = g.DT / 2.0;
    float vxx, vyx, vxy, vyy;
    velocity_derivs<N>(g.hc, a.vx + c, a.vy + c, g.pitch, dhi, vxx, vyx, vxy, vyy);
    cpml_stress(g, j, i, a.psi_vxx, a.psi_vyx, a.psi_vyy, a.psi_vxy, vxx, vyx, vxy, vyy);

    const long mc = c * a.cs;
    float rl[L], pl[L], ql[L];
    float sumr = 0.0f, sump = 0.0f, sumq = 0.0f;
#pragma unroll
    for (int l = 0; l < L; l++) {
        rl[l] = a.r[mc + (long)l * a.ls];
        pl[l] = a.p[mc + (long)l * a.ls];
        ql[l] = a.q[mc + (long)l * a.ls];
        sumr += rl[l];
        sump += pl[l];
        sumq += ql[l];
    }
    const float fipjp = a.fipjp[c], f = a.f[c], gv = a.gv[c];
    float sxy = a.sxy[c] + ((fipjp * (vxy + vyx)) + (dthalbe * sumr));
    float sxx = (float)((double)a.sxx[c] +
                        (((double)(gv * (vxx + vyy)) - (2.0 * (double)f * (double)vyy)) + (double)(dthalbe * sump)));
    float syy = (float)((double)a.syy[c] +
                        (((double)(gv * (vxx + vyy)) - (2.0 * (double)f * (double)vxx)) + (double)(dthalbe * sumq)));
    const bool rates = a.ux && (!nt_dev || (*nt_dev - 1) % DTINV == 0);
    float uxy = 0.0f, ux = 0.0f, uy = 0.0f;
    if (rates) {
        const double f2 = (2.0 * (double)f) / (double)g.DT;
        uxy = (float)((double)((fipjp / g.DT) * (vxy + vyx)) + (0.5 * (double)sumr));
        ux = (float)(((double)((gv / g.DT) * (vxx + vyy)) - (f2 * (double)vyy)) + (0.5 * (double)sump));
        uy = (float)(((double)((gv / g.DT) * (vxx + vyy)) - (f2 * (double)vxx)) + (0.5 * (double)sumq));
    }
    sumr = sump = sumq = 0.0f;
#pragma unroll
    for (int l = 0; l < L; l++) {
        const long o = mc + (long)l * a.ls;
        const float dip = a.dip[o], d = a.d[o], e = a.e[o];
        const float r = a.bip[l + 1] * (rl[l] * a.cip[l + 1] - (dip * (vxy + vyx)));
        const float p = (float)((double)a.bjm[l + 1] *
                                ((double)(pl[l] * a.cjm[l + 1] - (e * (vxx + vyy))) + (2.0 * (double)d * (double)vyy)));
        const float q = (float)((double)a.bjm[l + 1] *
                                ((double)(ql[l] * a.cjm[l + 1] - (e * (vxx + vyy))) + (2.0 * (double)d * (double)vxx)));
        a.r[o] = r;
        a.p[o] = p;
        a.q[o] = q;
        sumr += r;
        sump += p;
        sumq += q;
    }
    a.sxy[c] = sxy + (dthalbe * sumr);
    a.sxx[c] = sxx + (dthalbe * sump);
    a.syy[c] = syy + (dthalbe * sumq);
    if (rates) {
        a.uxy[c] = (float)((double)uxy + (0.5 * (double)sumr));
        a.ux[c] = (float)((double)ux + (0.5 * (double)sump));
        a.uy[c] = (float)((double)uy + (0.5 * (double)sumq));
    }
}

// ---------------------------------------------------------------------------
// per-step post kernel: explosive source (src/psource.c:26-75), free surface
// (src/surface_elastic_PML.c:26-124, src/surface_PML.c:25-177), receiver
// sampling (src/seismo_ssg.c:26-55) and the stability probe
// (src/IFOS2D.c:1660-1665).  Threads [0,NX) own one column each: the
// explosive sources located in that column are applied first, in source
// order, then the free surface of that column, which only reads and writes
// that column (plus vx of the surface row, untouched here).  Threads
// [NX, NX+ntr) sample the receivers.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void surface_column(const PostArgs &a, int i, float &vxx_out, float &vyy_out)
{
    const StepGeom &g = a.g;
    const int P = g.pitch;
    const long c1 = (long)1 * P + i;
    // Every value read here is loaded before the first store: none of them
    // lies in the rows written (the mirrored rows 1-N..0, row 1 of syy and
    // uy), and no other thread of the launch writes them, so the loads need
    // not wait on the stores one round trip at a time.
    float syy_s[N], sxy_s[N], uy_s[N], uxy_s[N], vxp[N], vxm[N], vyp[N], vym[N];
#pragma unroll
    for (int m = 1; m <= N; m++) {
        syy_s[m - 1] = a.syy[c1 + (long)m * P];
        sxy_s[m - 1] = a.sxy[c1 + (long)(m - 1) * P];
        uy_s[m - 1] = a.uy ? a.uy[c1 + (long)m * P] : 0.0f;
        uxy_s[m - 1] = a.uxy ? a.uxy[c1 + (long)(m - 1) * P] : 0.0f;
        vxp[m - 1] = a.vx[c1 + m - 1];
        vxm[m - 1] = a.vx[c1 - m];
        vyp[m - 1] = a.vy[c1 + (long)(m - 1) * P];
        vym[m - 1] = a.vy[c1 - (long)m * P];
    }
    const Cpml &pm = g.pml;
    const bool left = !g.BOUNDARY && i <= g.FW, right = !g.BOUNDARY && i >= g.NX - g.FW + 1;
    const int hl = i - 1, hr = i - g.NX + 2 * g.FW - 1;
    const float psl = left ? a.psi_vxxs[hl] : 0.0f, psr = right ? a.psi_vxxs[hr] : 0.0f;
    a.syy[c1] = 0.0f;
    if (a.uy) a.uy[c1] = 0.0f;
    float vxx = 0.0f, vyy = 0.0f;
#pragma unroll
    for (int m = 1; m <= N; m++) {
        a.syy[c1 - (long)m * P] = -syy_s[m - 1];
        a.sxy[c1 - (long)m * P] = -sxy_s[m - 1];
        if (a.uy) a.uy[c1 - (long)m * P] = -uy_s[m - 1];
        if (a.uxy) a.uxy[c1 - (long)m * P] = -uxy_s[m - 1];
        vxx += g.hc[m] * (vxp[m - 1] - vxm[m - 1]);
        vyy += g.hc[m] * (vyp[m - 1] - vym[m - 1]);
    }
    const float dh24 = 1.0 / (double)g.DH;
    vxx *= dh24;
    vyy *= dh24;
    if (left) {
        const float ps = pm.b_x[hl] * psl + pm.a_x[hl] * vxx;
        a.psi_vxxs[hl] = ps;
        vxx = vxx / pm.K_x[hl] + ps;
    }
    if (right) {
        // the left strip's update above wrote another entry (hl != hr)
        const float ps = pm.b_x[hr] * psr + pm.a_x[hr] * vxx;
        a.psi_vxxs[hr] = ps;
        vxx = vxx / pm.K_x[hr] + ps;
    }
    vxx_out = vxx;
    vyy_out = vyy;
}

template <int N>
IFOS_KERNEL void k_post(PostArgs a)
{
    const StepGeom &g = a.g;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = step_of(a.nt_dev, a.nt);
    const int P = g.pitch;
    if (t < g.NX) {
        const int i = t + 1;
        if (a.explosive) {
            for (int l = 0; l < a.src.n; l++) {
                if (a.src.ix[l] != i) continue;
                const long c = (long)a.src.iy[l] * P + i;
                const float *s = a.src.sig + (long)l * a.src.ld - 1;  // 1-based
                float amp = 0.0f;
                const double den = 2.0 * (double)g.DH * (double)g.DH;
                if (nt == 1 && a.NT > 1) amp = (float)((double)s[nt + 1] / den);
                if ((nt > 1) && (nt < a.NT)) amp = (float)((double)(s[nt + 1] - s[nt - 1]) / den);
                if (nt == a.NT) amp = (float)((double)(-s[nt - 1]) / den);
                a.sxx[c] = (float)((double)a.sxx[c] + 0.5 * (double)amp);
                a.syy[c] = (float)((double)a.syy[c] + 0.5 * (double)amp);
            }
        }
        if (g.FREE_SURF) {
            const long c1 = (long)P + i;
            const float DT = g.DT;
            // the column's own values are loaded with the surface stencil's,
            // before surface_column's stores (which write none of them)
            const float sxx1 = a.sxx[c1];
            const float ux1 = a.ux ? a.ux[c1] : 0.0f;
            const long mc = c1 * a.cs;
            float m0, m1, m2 = 0.0f, m3 = 0.0f, m4 = 0.0f;
            if (a.L == 0) {
                if (a.f) {
                    m0 = a.f[c1];
                    m1 = a.gl[c1];
                } else {
                    m0 = a.rho[c1];
                    m1 = a.vs[c1];
                    m2 = a.vp[c1];
                }
            } else {
                m0 = a.vs[c1];
                m1 = a.vp[c1];
                m2 = a.rho[c1];
                m3 = a.taus[c1];
                m4 = a.taup[c1];
            }
            float pv[kMaxL];
#pragma unroll
            for (int l = 0; l < kMaxL; l++) pv[l] = l < a.L ? a.p[mc + (long)l * a.ls] : 0.0f;
            float vxx, vyy;
            surface_column<N>(a, i, vxx, vyy);
            if (a.L == 0) {
                float fjm, gl;
                if (a.f) {
                    fjm = (float)((double)m0 * 2.0);
                    gl = m1;
                } else {
                    const float rho = m0, u = m1, pi = m2;
                    fjm = (float)((double)(rho * u * u) * 2.0);
                    gl = rho * ((pi * pi) - 2 * u * u);
                }
                a.sxx[c1] = sxx1 + (-DT * ((gl * gl) / (gl + fjm) * vxx + gl * vyy));
                if (a.ux)
                    a.ux[c1] = (float)(((2.0 * (double)gl * (double)fjm + (double)(fjm * fjm)) / (double)(gl + fjm)) *
                                       (double)vxx);
            } else {
                const int L = a.L;
                for (int l = 0; l < L; l++) a.q[mc + (long)l * a.ls] = 0.0f;
                const float u = m0, pi = m1, rho = m2;
                const float taus = m3, taup = m4;
                const float mu = (float)((double)(u * u * rho) / (1.0 + (double)(a.sumu * taus)));
                const float pim = (float)((double)(pi * pi * rho) / (1.0 + (double)(a.sumpi * taup)));
                float sump = 0.0f;
#pragma unroll
                for (int l = 0; l < kMaxL; l++)
                    if (l < L) sump += pv[l];
                const float fjm = (float)((double)mu * 2.0 * (1.0 + (double)(L * taus)));
                const float gv = (float)((double)pim * (1.0 + (double)(L * taup)));
                const float dthalbe = DT / 2.0;
                const float sxx2 = sxx1 + ((-(DT * (gv - fjm) * (gv - fjm) * vxx / gv) - (DT * (gv - fjm) * vyy)) -
                                           (dthalbe * sump));
                const float ux2 = (float)((double)ux1 +
                                          ((double)(-((gv - fjm) * (gv - fjm) * vxx / gv) - ((gv - fjm) * vyy)) -
                                           (0.5 * (double)sump)));
                sump = 0.0f;
#pragma unroll
                for (int l = 1; l <= kMaxL; l++) {
                    if (l > L) continue;
                    const float bjm = (float)((double)a.eta[l] / (1.0 + ((double)a.eta[l] * 0.5)));
                    const float djm = (float)(2.0 * (double)mu * (double)taus);
                    const float e = pim * taup;
                    const float pn = (float)((double)pv[l - 1] +
                                             (double)bjm * ((((double)(djm - e) * ((double)(fjm / gv) - 1.0)) *
                                                             (double)vxx) -
                                                            (double)((djm - e) * vyy)));
                    a.p[mc + (long)(l - 1) * a.ls] = pn;
                    sump += pn;
                }
                a.sxx[c1] = sxx2 + (dthalbe * sump);
                if (a.ux) a.ux[c1] = (float)((double)ux2 + (0.5 * (double)sump));
            }
        }
    } else if (t < g.NX + a.ntr) {
        const int itr = t - g.NX;
        const long c = (long)a.ry[itr] * P + a.rx[itr];
        a.seis_vx[(long)itr * a.NT + nt - 1] = a.vx[c];
        a.seis_vy[(long)itr * a.NT + nt - 1] = a.vy[c];
    } else if (t == g.NX + a.ntr && a.nan_flag) {
        const float v = a.vy[(long)(g.NY / 2) * P + g.NX / 2];
        if (isnan(v) && *a.nan_flag == 0) *a.nan_flag = nt;
    }
}

// ---------------------------------------------------------------------------
// snapshot store  (src/IFOS2D.c:1805-1871), slot-major [k][j][i]
// ---------------------------------------------------------------------------
IFOS_KERNEL void k_snapshot(SnapArgs a)
{
    const int nt = step_of(a.nt_dev, a.nt);
    if ((nt - 1) % a.DTINV != 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > a.NX || j > a.NY) return;
    const long plane = (long)a.NX * a.NY;
    const long k = (long)((nt - 1) / a.DTINV);
    const long o = k * plane + (long)(j - 1) * a.NX + (i - 1);
    const long c = (long)j * a.pitch + i;
#pragma unroll
    for (int f = 0; f < 5; f++) a.base[f][o] = a.a[f][c];
}

// ---------------------------------------------------------------------------
// zero-lag correlation of the adjoint with the stored forward wavefield
// (src/IFOS2D.c:2381-2419; the parenthesisation of the mu term is the coded
// one of lines 2398-2404)
// ---------------------------------------------------------------------------
IFOS_KERNEL void k_correlate(CorrArgs a)
{
    const int nt = step_of(a.nt_dev, a.nt);
    if ((a.NT - nt) % a.DTINV != 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > a.NX || j > a.NY) return;
    const long plane = (long)a.NX * a.NY;
    const long e = (long)(j - 1) * a.NX + (i - 1);
    const long o = (long)((a.NT - nt) / a.DTINV) * plane + e;
    const long c = (long)j * a.pitch + i;
    const float Frx = a.F[0][o], Fry = a.F[1][o], Fx = a.F[2][o], Fy = a.F[3][o], Fu = a.F[4][o];
    const float sxx = a.sxx[c], syy = a.syy[c], sxy = a.sxy[c];
    a.wc_rho[e] = a.wc_rho[e] + ((a.vxp1[c] * Frx) + (a.vyp1[c] * Fry));
    a.wc[e] = a.wc[e] + ((Fx + Fy) * (sxx + syy));
    if (a.vs[e] > 0.0f) {
        const float muss = a.muss[e], lamss = a.lamss[e];
        const double t1 = (1.0 / (double)(muss * muss)) * (double)(Fu * sxy);
        const double t2 = (0.25 * (double)(Fx + Fy * (sxx + syy))) / (double)((lamss + muss) * (lamss + muss));
        const double t3 = (0.25 * (double)(Fx - Fy * (sxx - syy))) / (double)(muss * muss);
        a.wc_u[e] = (float)((double)a.wc_u[e] + ((t1 + t2) + t3));
    }
}

// ---------------------------------------------------------------------------
// chain rule from lambda/mu/rho' to Vp/Vs/rho (src/IFOS2D.c:2471-2611,
// PARAMETERIZATION 1) and the sum over shots (src/IFOS2D.c:2703-2713)
// ---------------------------------------------------------------------------
// Is interior cell (i, j) zeroed by the (2F+1)^2 box around source (is, js)?
// The reference writes the box into a scratch matrix(0,NY+1,0,NX+1) with
// contiguous rows, for columns i' with (i'-1)/NX == 0 and rows 1..NY
// (src/taper_grad_shot.c:168-181): columns 1..NX hit their own cell, columns
// i' <= -2 hit column NX+2+i' of the row above.
__device__ __forceinline__ bool in_box(int i, int j, int is, int js, int F, int NX, int NY)
{
    const int j0 = max(js - F, 1), j1 = min(js + F, NY);
    if (j >= j0 && j <= j1 && i >= is - F && i <= is + F) return true;
    const int iw = i - NX - 2;  // the box column that wraps onto (i, j)
    return j + 1 >= j0 && j + 1 <= j1 && iw >= is - F && iw <= is + F && iw >= 2 - NX;
}

IFOS_KERNEL void k_gradient(GradArgs a)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.n) return;
    const float DT = a.DT;
    const float rh = a.rho[e], u = a.vs[e], pv = a.vp[e];
    float gvp = 0.0f, gvs = 0.0f, grho = 0.0f;
    float tvp = 0.0f, tvs = 0.0f, trho = 0.0f;
    const int ci = a.tap_m ? (int)(e % a.NX) + 1 : 0, cj = a.tap_m ? (int)(e / a.NX) + 1 : 0;
    if (a.s_vp) {
        tvp = a.s_vp[e];
        tvs = a.s_vs[e];
        trho = a.s_rho[e];
    }
    // one shot after the other, in slot order, as the reference's shot loop
    // adds them (src/IFOS2D.c:2703-2713)
    for (int s = 0; s < a.nslots; s++) {
        const long o = (long)s * a.wcs + e;
        float svp, svs, srho, wl, wmu, wrs;
        chain_rule_cell(DT, rh, u, pv, a.wc[o], a.wc_u[o], a.wc_rho[o], svp, svs, srho, wl, wmu, wrs);
        // energy preconditioning of the shot (src/IFOS2D.c:2627-2654)
        if (a.we) {
            const float We = a.we[(long)s * a.wes + e];
            if (a.per_shot) {
                const float w = We / a.we_max[s];
                svp = svp / (w);
                svs = svs / (w);
                srho = srho / (w);
            } else {
                a.we_sum[e] = (float)((double)a.we_sum[e] + 1.0 / (double)We);
            }
        }
        // circular taper around the shot's source, then the zeroed box
        // (src/IFOS2D.c:2679-2689, src/taper_grad_shot.c:146-187)
        if (a.tap_m) {
            const int is = a.tap_ij[2 * s], js = a.tap_ij[2 * s + 1];
            const int tx = ci - is + a.tap_c, ty = cj - js + a.tap_c;
            const float v = (tx >= 1 && tx <= a.tap_n && ty >= 1 && ty <= a.tap_n)
                                ? a.tap_m[(ty - 1) * a.tap_n + (tx - 1)] : 1.0f;
            const float t = (v - a.tap_lohi[2 * s]) / a.tap_lohi[2 * s + 1];
            svp *= t;
            svs *= t;
            srho *= t;
            if (in_box(ci, cj, is, js, a.tap_fs, a.NX, a.NY)) svp = svs = srho = 0.0f;
        }
        // taper grids of this shot read from files (src/IFOS2D.c:2691-2698)
        if (a.tf) {
            const float *t = a.tf + (long)s * a.tfs;
            svp *= t[e];
            svs *= t[a.n + e];
            srho *= t[2L * a.n + e];
        }
        if (a.slot_g) {
            float *sg = a.slot_g + (long)s * a.slot_gs + e;
            sg[0] = svp;
            sg[a.n] = svs;
            sg[2L * a.n] = srho;
        }
        gvp = s ? gvp + svp : svp;
        gvs = s ? gvs + svs : svs;
        grho = s ? grho + srho : srho;
        tvp = tvp + svp;
        tvs = tvs + svs;
        trho = trho + srho;
    }
    a.g_vp[e] = gvp;
    a.g_vs[e] = gvs;
    a.g_rho[e] = grho;
    if (a.s_vp) {
        a.s_vp[e] = tvp;
        a.s_vs[e] = tvs;
        a.s_rho[e] = trho;
    }
}

// ---------------------------------------------------------------------------
// residual / adjoint source (src/calc_res.c:136-282, LNORM 2) with the
// misfit and data energy (src/calc_misfit.c:193-194, src/calc_energy.c:157)
// summed per trace in FP64.
//
// One wavefront per trace, four per workgroup.  Samples are read and the
// time-reversed adjoint source written 64 at a time (one contiguous 256-byte
// row per load or store); each lane sums its own samples' squares in FP64
// and a shuffle tree combines the 64 partial sums.  With VELOCITY 0 the
// traces are integrated first by a sequential FP32 running sum as in the
// reference: the sum is carried through the lanes of each 64-sample chunk
// in sample order (a readlane per sample, the sum uniform across the wave),
// so every partial sum has the reference's bits.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lane_value(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

IFOS_KERNEL void k_residual(ResArgs a)
{
    const int lane = threadIdx.x & 63;
    const int itr = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (itr >= a.ntr) return;  // the whole wavefront leaves together
    const float *syn = a.syn + (long)itr * a.NT;
    const float *obs = a.obs + (long)itr * a.NT;
    float *diff = a.diff + (long)itr * a.NT;
    const bool killed = a.kill && a.kill[itr] == 1;
    double l2 = 0.0, en = 0.0;
    float is = 0.0f, id = 0.0f;  // running integrals (VELOCITY 0), equal in every lane
    for (int k0 = 0; k0 < a.NT; k0 += 64) {
        const int k = k0 + lane;
        const bool in = k < a.NT;
        if (killed) {
            if (in) diff[a.NT - 1 - k] = 0.0f;
            continue;
        }
        const float s = in ? syn[k] : 0.0f;
        const float o = in ? obs[k] : 0.0f;
        float ys = s, yd = o;
        if (a.VELOCITY == 0) {
            const int m = min(64, a.NT - k0);
            for (int l = 0; l < m; l++) {
                is += lane_value(s, l);
                id += lane_value(o, l);
                if (lane == l) {
                    ys = is * a.DT;
                    yd = id * a.DT;
                }
            }
        }
        if (in) {
            const float d = ys - yd;
            diff[a.NT - 1 - k] = d;
            l2 += (double)d * (double)d;
            en += (double)yd * (double)yd;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        l2 += __shfl_xor(l2, off, 64);
        en += __shfl_xor(en, off, 64);
    }
    if (lane == 0) {
        a.l2_trace[itr] = l2;
        a.en_trace[itr] = en;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 grid2d(int NX, int NY, dim3 b) { return dim3((NX + b.x - 1) / b.x, (NY + b.y - 1) / b.y); }

// XCD runs of the 64x4 tiles of a grid (tile_of_block) with equal cost, a
// tile that reaches into an absorbing strip counting double; returns the
// number of workgroups to launch.  The last few geometries are kept, since
// the update kernels run with the same one every step.
static int tile_runs(StepGeom &g)
{
    constexpr int TW = 64, TH = 4;
    struct Entry {
        int NX, NY, FW, FREE_SURF, BOUNDARY, grid;
        int xlo[9];
    };
    static thread_local Entry cache[4];
    static thread_local int next = 0;
    g.tnbx = (g.NX + TW - 1) / TW;
    for (const Entry &e : cache)
        if (e.grid && e.NX == g.NX && e.NY == g.NY && e.FW == g.FW && e.FREE_SURF == g.FREE_SURF &&
            e.BOUNDARY == g.BOUNDARY) {
            memcpy(g.xlo, e.xlo, sizeof g.xlo);
            return e.grid;
        }
    const int nby = (g.NY + TH - 1) / TH, nt = g.tnbx * nby;
    std::vector<long> cum(nt + 1, 0);
    for (int t = 0; t < nt; t++) {
        const int by = t / g.tnbx, bx = t - by * g.tnbx;
        const int i0 = bx * TW + 1, i1 = std::min(i0 + TW - 1, g.NX);
        const int j0 = by * TH + 1, j1 = std::min(j0 + TH - 1, g.NY);
        const bool strip = (!g.BOUNDARY && (i0 <= g.FW || i1 >= g.NX - g.FW + 1)) ||
                           (!g.FREE_SURF && j0 <= g.FW) || j1 >= g.NY - g.FW + 1;
        cum[t + 1] = cum[t] + (strip ? 2 : 1);
    }
    int widest = 0;
    g.xlo[0] = 0;
    for (int x = 1; x <= 8; x++) {
        const long want = cum[nt] * x / 8;
        int t = g.xlo[x - 1];
        while (t < nt && cum[t] < want) t++;
        g.xlo[x] = x == 8 ? nt : t;
        widest = std::max(widest, g.xlo[x] - g.xlo[x - 1]);
    }
    Entry &e = cache[next];
    next = (next + 1) % 4;
    e = Entry{g.NX, g.NY, g.FW, g.FREE_SURF, g.BOUNDARY, 8 * widest, {}};
    memcpy(e.xlo, g.xlo, sizeof g.xlo);
    return e.grid;
}

#define IFOS_DISPATCH_N(N, CALL) \
    switch (N) {                 \
    case 1: CALL(1); break;      \
    case 2: CALL(2); break;      \
    case 3: CALL(3); break;      \
    case 4: CALL(4); break;      \
    case 5: CALL(5); break;      \
    case 6: CALL(6); break;      \
    default: break;              \
    }

struct StepCtl {
    const int *nt_dev;
    int DTINV, NT;
};
static thread_local StepCtl g_ctl = {nullptr, 1, 1};

void set_step_control(const int *nt_dev, int DTINV, int NT)
{
    g_ctl.nt_dev = nt_dev;
    g_ctl.DTINV = DTINV;
    g_ctl.NT = NT;
}

void launch_update_v(const VArgs &a0, hipStream_t s)
{
    VArgs a = a0;
    dim3 b(64, 4), gr(tile_runs(a.g));
#define L_(n) hipLaunchKernelGGL(k_update_v<n>, gr, b, 0, s, a, g_ctl.nt_dev, g_ctl.DTINV, g_ctl.NT)
    IFOS_DISPATCH_N(a.g.N, L_)
#undef L_
}

void launch_inject(const PointArgs &a, const StepGeom &g, float *vx, float *vy, const float *rip,
                   const float *rjp, int sw, int ADJOINT_TYPE, const int *nt_dev, int nt, int *nt_tick,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_inject, dim3(1), dim3(256), 0, s, a, g, vx, vy, rip, rjp, sw, ADJOINT_TYPE, nt_dev, nt,
                       nt_tick, 0);
}

void launch_inject_serial(const PointArgs &a, const StepGeom &g, float *vx, float *vy, const float *rip,
                          const float *rjp, int sw, int ADJOINT_TYPE, const int *nt_dev, int nt, int *nt_tick,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_inject, dim3(1), dim3(256), 0, s, a, g, vx, vy, rip, rjp, sw, ADJOINT_TYPE, nt_dev, nt,
                       nt_tick, 1);
}

void launch_update_s_elastic(const SElArgs &a0, hipStream_t s)
{
    SElArgs a = a0;
    dim3 b(64, 4), gr(tile_runs(a.g));
#define L_(n) hipLaunchKernelGGL(k_update_s_elastic<n>, gr, b, 0, s, a, g_ctl.nt_dev, g_ctl.DTINV)
    IFOS_DISPATCH_N(a.g.N, L_)
#undef L_
}

template <int N>
static void launch_visc_n(const SViscArgs &a0, hipStream_t s)
{
    SViscArgs a = a0;
    dim3 b(64, 4), gr(tile_runs(a.g));
    switch (a.L) {
    case 1: hipLaunchKernelGGL((k_update_s_visc<N, 1>), gr, b, 0, s, a, g_ctl.nt_dev, g_ctl.DTINV); break;
    case 2: hipLaunchKernelGGL((k_update_s_visc<N, 2>), gr, b, 0, s, a, g_ctl.nt_dev, g_ctl.DTINV); break;
    case 3: hipLaunchKernelGGL((k_update_s_visc<N, 3>), gr, b, 0, s, a, g_ctl.nt_dev, g_ctl.DTINV); break;
    case 4: hipLaunchKernelGGL((k_update_s_visc<N, 4>), gr, b, 0, s, a, g_ctl.nt_dev, g_ctl.DTINV); break;
    case 5: hipLaunchKernelGGL((k_update_s_visc<N, 5>), gr, b, 0, s, a, g_ctl.nt_dev, g_ctl.DTINV); break;
    default: break;
    }
}

void launch_update_s_visc(const SViscArgs &a, hipStream_t s)
{
#define L_(n) launch_visc_n<n>(a, s)
    IFOS_DISPATCH_N(a.g.N, L_)
#undef L_
}

void launch_post(const PostArgs &a, hipStream_t s)
{
    const int n = a.g.NX + a.ntr + 1;
    dim3 b(256), gr((n + 255) / 256);
#define L_(nn) hipLaunchKernelGGL(k_post<nn>, gr, b, 0, s, a)
    IFOS_DISPATCH_N(a.g.N, L_)
#undef L_
}

void launch_snapshot(const SnapArgs &a, hipStream_t s)
{
    dim3 b(64, 4), gr = grid2d(a.NX, a.NY, b);
    hipLaunchKernelGGL(k_snapshot, gr, b, 0, s, a);
}

void launch_correlate(const CorrArgs &a, hipStream_t s)
{
    dim3 b(64, 4), gr = grid2d(a.NX, a.NY, b);
    hipLaunchKernelGGL(k_correlate, gr, b, 0, s, a);
}

void launch_gradient(const GradArgs &a, hipStream_t s)
{
    dim3 b(256), gr((unsigned)((a.n + 255) / 256));
    hipLaunchKernelGGL(k_gradient, gr, b, 0, s, a);
}

// Time-domain Butterworth filter (timedomain_filt, src/timedomain_filt.c:27,
// libcseife seife_butterworth / seife_rekfl), one thread per trace.  The
// reference runs each section over the whole trace in double; a cascade of
// causal sections gives every section the same inputs when it is run sample
// by sample instead, so the state of all sections stays in registers.  The
// reference filters ns+1 samples whose first is 0 (its vector's element 0),
// and stores the lowpass result as float before the highpass pass.
template <int NS>
__device__ __forceinline__ double cascade(const double (*c)[5], double (*st)[4], int n, double x)
{
#pragma unroll
    for (int k = 0; k < NS; k++) {
        if (k >= n) break;
        const double y = c[k][0] * x + c[k][1] * st[k][0] + c[k][2] * st[k][1] + c[k][3] * st[k][2] +
                         c[k][4] * st[k][3];
        st[k][1] = st[k][0];
        st[k][0] = x;
        st[k][3] = st[k][2];
        st[k][2] = y;
        x = y;
    }
    return x;
}

IFOS_KERNEL void k_time_filter(FiltArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntr) return;
    float *tr = a.data + (long)t * a.stride;
    double lo[kMaxFiltSections][4], hi[kMaxFiltSections][4];
    for (int k = 0; k < kMaxFiltSections; k++)
        for (int q = 0; q < 4; q++) lo[k][q] = hi[k][q] = 0.0;
    // the leading zero sample of both passes
    if (a.nlow) cascade<kMaxFiltSections>(a.c, lo, a.nlow, 0.0);
    if (a.nhigh) cascade<kMaxFiltSections>(a.c + a.nlow, hi, a.nhigh, 0.0);
    for (int j = 0; j < a.ns; j++) {
        float v = tr[j];
        if (a.nlow) v = (float)cascade<kMaxFiltSections>(a.c, lo, a.nlow, (double)v);
        if (a.nhigh) v = (float)cascade<kMaxFiltSections>(a.c + a.nlow, hi, a.nhigh, (double)v);
        tr[j] = v;
    }
}

void launch_time_filter(const FiltArgs &a, hipStream_t s)
{
    if (a.ntr <= 0 || (!a.nlow && !a.nhigh)) return;
    dim3 b(64), gr((a.ntr + 63) / 64);
    hipLaunchKernelGGL(k_time_filter, gr, b, 0, s, a);
}

void launch_residual(const ResArgs &a, hipStream_t s)
{
    if (a.ntr <= 0) return;
    dim3 b(256), gr((a.ntr + 3) / 4);
    hipLaunchKernelGGL(k_residual, gr, b, 0, s, a);
}

}  // namespace ifos
