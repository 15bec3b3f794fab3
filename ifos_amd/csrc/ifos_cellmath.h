// ifos_cellmath.h -- per-cell arithmetic of the P-SV scheme, shared by the
// reference-signature kernels and the fused step kernels.
//
// Every function evaluates the reference's expression with the same operand
// order and the same float->double promotions; the library is compiled with
// -ffp-contract=off, so results are bit-identical to the reference's CPU run.
#pragma once
#include "ifos_internal.h"

namespace ifos {

// staggered derivatives of the stresses at a vx/vy node
// (src/update_v_PML.c:358-381 for order 8; the other orders extend the sum)
template <int N>
__device__ __forceinline__ void stress_derivs(const float *hc, const float *sxx_, const float *syy_, const float *sxy_,
                                              int P, float &sxx_x, float &sxy_x, float &sxy_y, float &syy_y)
{
    const auto sxx = gmem(sxx_), syy = gmem(syy_), sxy = gmem(sxy_);
    sxx_x = hc[1] * (sxx[1] - sxx[0]);
    sxy_x = hc[1] * (sxy[0] - sxy[-1]);
    sxy_y = hc[1] * (sxy[0] - sxy[-P]);
    syy_y = hc[1] * (syy[P] - syy[0]);
#pragma unroll
    for (int m = 2; m <= N; m++) {
        sxx_x = sxx_x + hc[m] * (sxx[m] - sxx[1 - m]);
        sxy_x = sxy_x + hc[m] * (sxy[m - 1] - sxy[-m]);
        sxy_y = sxy_y + hc[m] * (sxy[(m - 1) * P] - sxy[-m * P]);
        syy_y = syy_y + hc[m] * (syy[m * P] - syy[(1 - m) * P]);
    }
}

// stress_derivs for R vertically adjacent cells, rows j + r (r = 0 .. R-1;
// row j at the given pointers): the y derivatives of all R rows read one
// column of 2N+R-1 values of sxy (rows j-N .. j+N+R-2) and of syy (rows
// j-N+1 .. j+N+R-1) instead of 2N per row; each row's sums keep
// stress_derivs' order.  lim is the last row offset from j inside the
// padded grid: for R > 2 the reads of rows past the model's last row, which
// only feed cells that are not stored, are clamped to it (R <= 2 never reads
// past it)
template <int N, int R>
__device__ __forceinline__ void stress_derivs_rows(const float *hc, const float *sxx_, const float *syy_,
                                                   const float *sxy_, int P, int lim, float (&sxx_x)[R],
                                                   float (&sxy_x)[R], float (&sxy_y)[R], float (&syy_y)[R])
{
    const auto sxx = gmem(sxx_), syy = gmem(syy_), sxy = gmem(sxy_);
    constexpr int K = 2 * N + R - 1;
    auto row = [&](int o) { return R > 2 ? min(o, lim) : o; };
    float cy[K], cs[K];  // cy[k]: sxy row k-N; cs[k]: syy row k-N+1
#pragma unroll
    for (int k = 0; k < K; k++) {
        cy[k] = sxy[row(k - N) * P];
        cs[k] = syy[row(k - N + 1) * P];
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const auto sx = sxx + row(r) * P, sy = sxy + row(r) * P;
        sxx_x[r] = hc[1] * (sx[1] - sx[0]);
        sxy_x[r] = hc[1] * (cy[N + r] - sy[-1]);
        sxy_y[r] = hc[1] * (cy[N + r] - cy[N + r - 1]);
        syy_y[r] = hc[1] * (cs[N + r] - cs[N + r - 1]);
#pragma unroll
        for (int m = 2; m <= N; m++) {
            sxx_x[r] = sxx_x[r] + hc[m] * (sx[m] - sx[1 - m]);
            sxy_x[r] = sxy_x[r] + hc[m] * (sy[m - 1] - sy[-m]);
            sxy_y[r] = sxy_y[r] + hc[m] * (cy[N + r + m - 1] - cy[N + r - m]);
            syy_y[r] = syy_y[r] + hc[m] * (cs[N + r + m - 1] - cs[N + r - m]);
        }
    }
}

// x / K[h] of the CPML update.  With every profile value K = 1 (k_max_PML = 1,
// the usual case) the division is skipped: x / 1 is exactly x, so the
// result is the reference's bit for bit and the correctly rounded division
// sequence is saved in every absorbing cell.
__device__ __forceinline__ float divK(float x, const float *K, int h, int K_one)
{
    if (K_one) return x;
    return x / cmem(K)[h];
}

// velocity-side CPML (src/update_v_PML.c:383-434)
__device__ __forceinline__ void cpml_velocity(const StepGeom &g, int j, int i, float *psi_sxx_x, float *psi_sxy_x,
                                              float *psi_syy_y, float *psi_sxy_y, float &sxx_x, float &sxy_x,
                                              float &sxy_y, float &syy_y)
{
    const Cpml &pm = g.pml;
    if (!g.BOUNDARY && i <= g.FW) {
        const int h = i - 1;
        auto p1 = gmem(psi_sxx_x) + (long)(j - 1) * g.px + h;
        auto p2 = gmem(psi_sxy_x) + (long)(j - 1) * g.px + h;
        *p1 = cmem(pm.b_xh)[h] * *p1 + cmem(pm.a_xh)[h] * sxx_x;
        sxx_x = divK(sxx_x, pm.K_xh, h, g.K_one) + *p1;
        *p2 = cmem(pm.b_x)[h] * *p2 + cmem(pm.a_x)[h] * sxy_x;
        sxy_x = divK(sxy_x, pm.K_x, h, g.K_one) + *p2;
    }
    if (!g.BOUNDARY && i >= g.NX - g.FW + 1) {
        const int h = i - g.NX + 2 * g.FW - 1;
        auto p1 = gmem(psi_sxx_x) + (long)(j - 1) * g.px + h;
        auto p2 = gmem(psi_sxy_x) + (long)(j - 1) * g.px + h;
        *p1 = cmem(pm.b_xh)[h] * *p1 + cmem(pm.a_xh)[h] * sxx_x;
        sxx_x = divK(sxx_x, pm.K_xh, h, g.K_one) + *p1;
        *p2 = cmem(pm.b_x)[h] * *p2 + cmem(pm.a_x)[h] * sxy_x;
        sxy_x = divK(sxy_x, pm.K_x, h, g.K_one) + *p2;
    }
    if (!g.FREE_SURF && j <= g.FW) {
        const int h = j - 1;
        auto p1 = gmem(psi_syy_y) + (long)h * g.py + (i - 1);
        auto p2 = gmem(psi_sxy_y) + (long)h * g.py + (i - 1);
        *p1 = cmem(pm.b_yh)[h] * *p1 + cmem(pm.a_yh)[h] * syy_y;
        syy_y = divK(syy_y, pm.K_yh, h, g.K_one) + *p1;
        *p2 = cmem(pm.b_y)[h] * *p2 + cmem(pm.a_y)[h] * sxy_y;
        sxy_y = divK(sxy_y, pm.K_y, h, g.K_one) + *p2;
    }
    if (j >= g.NY - g.FW + 1) {
        const int h = j - g.NY + 2 * g.FW - 1;
        auto p1 = gmem(psi_syy_y) + (long)h * g.py + (i - 1);
        auto p2 = gmem(psi_sxy_y) + (long)h * g.py + (i - 1);
        *p1 = cmem(pm.b_yh)[h] * *p1 + cmem(pm.a_yh)[h] * syy_y;
        syy_y = divK(syy_y, pm.K_yh, h, g.K_one) + *p1;
        *p2 = cmem(pm.b_y)[h] * *p2 + cmem(pm.a_y)[h] * sxy_y;
        sxy_y = divK(sxy_y, pm.K_y, h, g.K_one) + *p2;
    }
}

// staggered derivatives of the particle velocities at a stress node
// (src/update_s_elastic_PML.c:340-360 for order 8), times `scale`
template <int N>
__device__ __forceinline__ void velocity_derivs(const float *hc, const float *vx_, const float *vy_, int P, float scale,
                                                float &vxx, float &vyx, float &vxy, float &vyy)
{
    const auto vx = gmem(vx_), vy = gmem(vy_);
    vxx = hc[1] * (vx[0] - vx[-1]);
    vyx = hc[1] * (vy[1] - vy[0]);
    vxy = hc[1] * (vx[P] - vx[0]);
    vyy = hc[1] * (vy[0] - vy[-P]);
#pragma unroll
    for (int m = 2; m <= N; m++) {
        vxx = vxx + hc[m] * (vx[m - 1] - vx[-m]);
        vyx = vyx + hc[m] * (vy[m] - vy[1 - m]);
        vxy = vxy + hc[m] * (vx[m * P] - vx[(1 - m) * P]);
        vyy = vyy + hc[m] * (vy[(m - 1) * P] - vy[-m * P]);
    }
    vxx = vxx * scale;
    vyx = vyx * scale;
    vxy = vxy * scale;
    vyy = vyy * scale;
}

// stress-side CPML (src/update_s_elastic_PML.c:362-415)
__device__ __forceinline__ void cpml_stress(const StepGeom &g, int j, int i, float *psi_vxx, float *psi_vyx,
                                            float *psi_vyy, float *psi_vxy, float &vxx, float &vyx, float &vxy,
                                            float &vyy)
{
    const Cpml &pm = g.pml;
    if (!g.BOUNDARY && i <= g.FW) {
        const int h = i - 1;
        auto p1 = gmem(psi_vxx) + (long)(j - 1) * g.px + h;
        auto p2 = gmem(psi_vyx) + (long)(j - 1) * g.px + h;
        *p1 = cmem(pm.b_x)[h] * *p1 + cmem(pm.a_x)[h] * vxx;
        vxx = divK(vxx, pm.K_x, h, g.K_one) + *p1;
        *p2 = cmem(pm.b_xh)[h] * *p2 + cmem(pm.a_xh)[h] * vyx;
        vyx = divK(vyx, pm.K_xh, h, g.K_one) + *p2;
    }
    if (!g.BOUNDARY && i >= g.NX - g.FW + 1) {
        const int h = i - g.NX + 2 * g.FW - 1;
        auto p1 = gmem(psi_vxx) + (long)(j - 1) * g.px + h;
        auto p2 = gmem(psi_vyx) + (long)(j - 1) * g.px + h;
        *p1 = cmem(pm.b_x)[h] * *p1 + cmem(pm.a_x)[h] * vxx;
        vxx = divK(vxx, pm.K_x, h, g.K_one) + *p1;
        *p2 = cmem(pm.b_xh)[h] * *p2 + cmem(pm.a_xh)[h] * vyx;
        vyx = divK(vyx, pm.K_xh, h, g.K_one) + *p2;
    }
    if (!g.FREE_SURF && j <= g.FW) {
        const int h = j - 1;
        auto p1 = gmem(psi_vyy) + (long)h * g.py + (i - 1);
        auto p2 = gmem(psi_vxy) + (long)h * g.py + (i - 1);
        *p1 = cmem(pm.b_y)[h] * *p1 + cmem(pm.a_y)[h] * vyy;
        *p2 = cmem(pm.b_yh)[h] * *p2 + cmem(pm.a_yh)[h] * vxy;
        vyy = divK(vyy, pm.K_y, h, g.K_one) + *p1;
        vxy = divK(vxy, pm.K_yh, h, g.K_one) + *p2;
    }
    if (j >= g.NY - g.FW + 1) {
        const int h = j - g.NY + 2 * g.FW - 1;
        auto p1 = gmem(psi_vyy) + (long)h * g.py + (i - 1);
        auto p2 = gmem(psi_vxy) + (long)h * g.py + (i - 1);
        *p1 = cmem(pm.b_y)[h] * *p1 + cmem(pm.a_y)[h] * vyy;
        vyy = divK(vyy, pm.K_y, h, g.K_one) + *p1;
        *p2 = cmem(pm.b_yh)[h] * *p2 + cmem(pm.a_yh)[h] * vxy;
        vxy = divK(vxy, pm.K_yh, h, g.K_one) + *p2;
    }
}

// velocity increment (src/update_v_PML.c:448-460)
__device__ __forceinline__ float vel_new(float v, float DT, float r, float s, float DH) { return v + DT * r * s / DH; }
// acceleration kept for the density correlation (sw == 0)
__device__ __forceinline__ float vel_rate(float r, float s, float DH) { return r * s / DH; }

// elastic stress update (src/update_s_elastic_PML.c:232-234): returns the
// new sxx, syy, sxy
__device__ __forceinline__ void stress_elastic(float f, float gl, float fipjp, float vxx, float vyx, float vxy,
                                               float vyy, float &sxx, float &syy, float &sxy)
{
    sxy = sxy + fipjp * (vyx + vxy);
    sxx = (float)((double)sxx + ((double)(gl * (vxx + vyy)) + (2.0 * (double)f * (double)vxx)));
    syy = (float)((double)syy + ((double)(gl * (vxx + vyy)) + (2.0 * (double)f * (double)vyy)));
}

// elastic stress rates (src/update_s_elastic_PML.c:236-238)
__device__ __forceinline__ void rates_elastic(float f, float gl, float fipjp, float DT, float vxx, float vyx,
                                              float vxy, float vyy, float &ux, float &uy, float &uxy)
{
    uxy = (fipjp / DT) * (vyx + vxy);
    const double f2 = (2.0 * (double)f) / (double)DT;
    ux = (float)((double)((gl / DT) * (vxx + vyy)) + (f2 * (double)vxx));
    uy = (float)((double)((gl / DT) * (vxx + vyy)) + (f2 * (double)vyy));
}

// GSLS relaxation-coefficient set of one cell (src/prepare_update_s.c:66-78)
// rebuilt from the relaxed moduli and the tau values, bit-identical to the
// reference's stored f, g, fipjp and d, e, dip
struct ViscCell {
    float f, g, fipjp;
    float mu, pi, muipjp, taus, taup, tausipjp;
};

__device__ __forceinline__ void visc_coeffs(int L, float DT, float mu, float pi, float muipjp, float taus, float taup,
                                            float tausipjp, ViscCell &c)
{
    c.mu = mu;
    c.pi = pi;
    c.muipjp = muipjp;
    c.taus = taus;
    c.taup = taup;
    c.tausipjp = tausipjp;
    c.fipjp = (float)((double)(muipjp * DT) * (1.0 + (double)(L * tausipjp)));
    c.f = (float)((double)(mu * DT) * (1.0 + (double)(L * taus)));
    c.g = (float)((double)(pi * DT) * (1.0 + (double)(L * taup)));
}

// viscoelastic stress update with L memory variables
// (src/update_s_visc_PML.c:486-514).  r/p/q hold the old memory variables on
// entry and the new ones on exit.  d/e/dip per mechanism are formed as
// (modulus*eta_l)*tau like src/prepare_update_s.c:75-77.
template <int L>
__device__ __forceinline__ void stress_visc(const ViscCell &c, const float *eta, const float *bip, const float *bjm,
                                            const float *cip, const float *cjm, float DT, float vxx, float vyx,
                                            float vxy, float vyy, float &sxx, float &syy, float &sxy, float *r,
                                            float *p, float *q, bool rates, float &ux, float &uy, float &uxy)
{
    const float dthalbe = DT / 2.0;
    float sumr = 0.0f, sump = 0.0f, sumq = 0.0f;
#pragma unroll
    for (int l = 0; l < L; l++) {
        sumr += r[l];
        sump += p[l];
        sumq += q[l];
    }
    sxy = sxy + ((c.fipjp * (vxy + vyx)) + (dthalbe * sumr));
    sxx = (float)((double)sxx + (((double)(c.g * (vxx + vyy)) - (2.0 * (double)c.f * (double)vyy)) +
                                 (double)(dthalbe * sump)));
    syy = (float)((double)syy + (((double)(c.g * (vxx + vyy)) - (2.0 * (double)c.f * (double)vxx)) +
                                 (double)(dthalbe * sumq)));
    if (rates) {
        const double f2 = (2.0 * (double)c.f) / (double)DT;
        uxy = (float)((double)((c.fipjp / DT) * (vxy + vyx)) + (0.5 * (double)sumr));
        ux = (float)(((double)((c.g / DT) * (vxx + vyy)) - (f2 * (double)vyy)) + (0.5 * (double)sump));
        uy = (float)(((double)((c.g / DT) * (vxx + vyy)) - (f2 * (double)vxx)) + (0.5 * (double)sumq));
    }
    sumr = sump = sumq = 0.0f;
#pragma unroll
    for (int l = 0; l < L; l++) {
        const float dip = c.muipjp * eta[l + 1] * c.tausipjp;
        const float d = c.mu * eta[l + 1] * c.taus;
        const float e = c.pi * eta[l + 1] * c.taup;
        r[l] = bip[l + 1] * (r[l] * cip[l + 1] - (dip * (vxy + vyx)));
        p[l] = (float)((double)bjm[l + 1] *
                       ((double)(p[l] * cjm[l + 1] - (e * (vxx + vyy))) + (2.0 * (double)d * (double)vyy)));
        q[l] = (float)((double)bjm[l + 1] *
                       ((double)(q[l] * cjm[l + 1] - (e * (vxx + vyy))) + (2.0 * (double)d * (double)vxx)));
        sumr += r[l];
        sump += p[l];
        sumq += q[l];
    }
    sxy = sxy + (dthalbe * sumr);
    sxx = sxx + (dthalbe * sump);
    syy = syy + (dthalbe * sumq);
    if (rates) {
        uxy = (float)((double)uxy + (0.5 * (double)sumr));
        ux = (float)((double)ux + (0.5 * (double)sump));
        uy = (float)((double)uy + (0.5 * (double)sumq));
    }
}

// explosive source amplitude at step nt (src/psource.c:46-49)
__device__ __forceinline__ float explosive_amp(const float *s_ /* 1-based */, int nt, int NT, float DH)
{
    const auto s = gmem(s_);
    float amp = 0.0f;
    const double den = 2.0 * (double)DH * (double)DH;
    if (nt == 1 && NT > 1) amp = (float)((double)s[nt + 1] / den);
    if ((nt > 1) && (nt < NT)) amp = (float)((double)(s[nt + 1] - s[nt - 1]) / den);
    if (nt == NT) amp = (float)((double)(-s[nt - 1]) / den);
    return amp;
}
__device__ __forceinline__ float add_half(float v, float amp) { return (float)((double)v + 0.5 * (double)amp); }

// horizontal derivatives on the free-surface row and the surface CPML
// (src/surface_elastic_PML.c:66-91); vx, vy point at (1,i)
template <int N>
__device__ __forceinline__ void surface_derivs(const StepGeom &g, int i, const float *vx_, const float *vy_, int P,
                                               float *psi_vxxs, float &vxx, float &vyy)
{
    const auto vx = gmem(vx_), vy = gmem(vy_);
    vxx = 0.0f;
    vyy = 0.0f;
#pragma unroll
    for (int m = 1; m <= N; m++) {
        vxx += g.hc[m] * (vx[m - 1] - vx[-m]);
        vyy += g.hc[m] * (vy[(m - 1) * P] - vy[-m * P]);
    }
    const float dh24 = 1.0 / (double)g.DH;
    vxx *= dh24;
    vyy *= dh24;
    const Cpml &pm = g.pml;
    if (!g.BOUNDARY && i <= g.FW) {
        const int h = i - 1;
        auto ps = gmem(psi_vxxs) + h;
        *ps = cmem(pm.b_x)[h] * *ps + cmem(pm.a_x)[h] * vxx;
        vxx = divK(vxx, pm.K_x, h, g.K_one) + *ps;
    }
    if (!g.BOUNDARY && i >= g.NX - g.FW + 1) {
        const int h = i - g.NX + 2 * g.FW - 1;
        auto ps = gmem(psi_vxxs) + h;
        *ps = cmem(pm.b_x)[h] * *ps + cmem(pm.a_x)[h] * vxx;
        vxx = divK(vxx, pm.K_x, h, g.K_one) + *ps;
    }
}

// elastic free-surface update of sxx and ux at j=1 (src/surface_elastic_PML.c:95-105)
__device__ __forceinline__ void surface_elastic_update(float fjm, float gl, float DT, float vxx, float vyy, float &sxx,
                                                       bool want_ux, float &ux)
{
    sxx = sxx + (-DT * ((gl * gl) / (gl + fjm) * vxx + gl * vyy));
    if (want_ux) ux = (float)(((2.0 * (double)gl * (double)fjm + (double)(fjm * fjm)) / (double)(gl + fjm)) * (double)vxx);
}

// viscoelastic free-surface update of sxx, ux and the p memory variables at
// j=1 (src/surface_PML.c:120-160); q must already be zeroed by the caller
template <int L>
__device__ __forceinline__ void surface_visc_update(float mu, float pim, float taus, float taup, const float *eta,
                                                    float DT, float vxx, float vyy, float &sxx, float *p, bool want_ux,
                                                    float &ux)
{
    float sump = 0.0f;
#pragma unroll
    for (int l = 0; l < L; l++) sump += p[l];
    const float fjm = (float)((double)mu * 2.0 * (1.0 + (double)(L * taus)));
    const float gv = (float)((double)pim * (1.0 + (double)(L * taup)));
    const float dthalbe = DT / 2.0;
    sxx = sxx + ((-(DT * (gv - fjm) * (gv - fjm) * vxx / gv) - (DT * (gv - fjm) * vyy)) - (dthalbe * sump));
    if (want_ux)
        ux = (float)((double)ux +
                     ((double)(-((gv - fjm) * (gv - fjm) * vxx / gv) - ((gv - fjm) * vyy)) - (0.5 * (double)sump)));
    sump = 0.0f;
#pragma unroll
    for (int l = 1; l <= L; l++) {
        const float bjm = (float)((double)eta[l] / (1.0 + ((double)eta[l] * 0.5)));
        const float djm = (float)(2.0 * (double)mu * (double)taus);
        const float e = pim * taup;
        float &pp = p[l - 1];
        pp = (float)((double)pp +
                     (double)bjm * ((((double)(djm - e) * ((double)(fjm / gv) - 1.0)) * (double)vxx) -
                                    (double)((djm - e) * vyy)));
        sump += pp;
    }
    sxx = sxx + (dthalbe * sump);
    if (want_ux) ux = (float)((double)ux + (0.5 * (double)sump));
}

// zero-lag correlation of one cell (src/IFOS2D.c:2386-2405)
__device__ __forceinline__ void correlate_cell(float vxp1, float vyp1, float sxx, float syy, float sxy, float Frx,
                                               float Fry, float Fx, float Fy, float Fu, float muss, float lamss,
                                               bool shear, float &wc_rho, float &wc, float &wc_u)
{
    wc_rho = wc_rho + ((vxp1 * Frx) + (vyp1 * Fry));
    wc = wc + ((Fx + Fy) * (sxx + syy));
    if (shear) {
        const double t1 = (1.0 / (double)(muss * muss)) * (double)(Fu * sxy);
        const double t2 = (0.25 * (double)(Fx + Fy * (sxx + syy))) / (double)((lamss + muss) * (lamss + muss));
        const double t3 = (0.25 * (double)(Fx - Fy * (sxx - syy))) / (double)(muss * muss);
        wc_u = (float)((double)wc_u + ((t1 + t2) + t3));
    }
}

}  // namespace ifos
