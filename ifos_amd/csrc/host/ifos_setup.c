/*
 * ifos_setup.c -- host-side setup of the P-SV hot path: finite-difference
 * coefficients, CPML profiles, source wavelets, material averaging and the
 * GSLS coefficients.  These run once per model / shot and stay on the host
 * as in the reference; their float results must match it bit for bit, so
 * every expression keeps the reference's operand order and promotions, and
 * the file is compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ifos2d_hip.h"
#include "ifos_setup.h"

#define IFOS_PI (3.141592653589793)

/* ---- holbergcoeff (src/holbergcoeff.c:37-108) --------------------------- */
/* Row 0: Taylor coefficients (exact rationals); rows 1-4: Holberg (1987)
 * coefficients for a maximum relative group-velocity error of 0.1, 0.5, 1 and
 * 3 %.  Column 0 is the points-per-wavelength figure used by the CFL and
 * dispersion check. */
static const float kHolberg[5][6][7] = {
    {{23.0, 1.0, 0, 0, 0, 0, 0},
     {8.0, 9.0 / 8.0, -1.0 / 24.0, 0, 0, 0, 0},
     {6.0, 75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0, 0, 0, 0},
     {5.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0, 0, 0},
     {5.0, 19845.0 / 16384.0, -735.0 / 8192.0, 567.0 / 40960.0, -405.0 / 229376.0, 35.0 / 294912.0, 0},
     {4.0, 160083.0 / 131072.0, -12705.0 / 131072.0, 22869.0 / 1310720.0, -5445.0 / 1835008.0, 847.0 / 2359296.0,
      -63.0 / 2883584.0}},
    {{49.7, 1.0010, 0, 0, 0, 0, 0},
     {8.32, 1.1382, -0.046414, 0, 0, 0, 0},
     {4.77, 1.1965, -0.078804, 0.0081781, 0, 0, 0},
     {3.69, 1.2257, -0.099537, 0.018063, -0.0026274, 0, 0},
     {3.19, 1.2415, -0.11231, 0.026191, -0.0064682, 0.001191, 0},
     {2.91, 1.2508, -0.12034, 0.032131, -0.010142, 0.0029857, -0.00066667}},
    {{22.2, 1.0050, 0, 0, 0, 0, 0},
     {5.65, 1.1534, -0.052806, 0, 0, 0, 0},
     {3.74, 1.2111, -0.088313, 0.011768, 0, 0, 0},
     {3.11, 1.2367, -0.10815, 0.023113, -0.0046905, 0, 0},
     {2.80, 1.2496, -0.11921, 0.031130, -0.0093272, 0.0025161, 0},
     {2.62, 1.2568, -0.12573, 0.036423, -0.013132, 0.0047484, -0.0015979}},
    {{15.8, 1.0100, 0, 0, 0, 0, 0},
     {4.80, 1.1640, -0.057991, 0, 0, 0, 0},
     {3.39, 1.2192, -0.094070, 0.014608, 0, 0, 0},
     {2.90, 1.2422, -0.11269, 0.026140, -0.0064054, 0, 0},
     {2.65, 1.2534, -0.12257, 0.033755, -0.011081, 0.0036784, 0},
     {2.51, 1.2596, -0.12825, 0.038550, -0.014763, 0.0058619, -0.0024538}},
    {{9.16, 1.0300, 0, 0, 0, 0, 0},
     {3.47, 1.1876, -0.072518, 0, 0, 0, 0},
     {2.91, 1.2341, -0.10569, 0.022589, 0, 0, 0},
     {2.61, 1.2516, -0.12085, 0.032236, -0.011459, 0, 0},
     {2.45, 1.2596, -0.12829, 0.038533, -0.014681, 0.0072580, 0},
     {2.36, 1.2640, -0.13239, 0.042217, -0.017803, 0.0081959, -0.0051848}},
};

void ifos_holberg(int FDORDER, int MAXRELERROR, float hc[7])
{
    int row = FDORDER / 2 - 1;
    if (row < 0) row = 0;
    if (row > 5) row = 5;
    if (MAXRELERROR < 0) MAXRELERROR = 0;
    if (MAXRELERROR > 4) MAXRELERROR = 4;
    for (int k = 0; k < 7; k++) hc[k] = kHolberg[MAXRELERROR][row][k];
}

/* ---- PML_pro (src/PML_pro.c:83-330) ------------------------------------- */
/* Komatitsch & Martin (2007) convolutional PML.  The reference compares |d|
 * through the int abs() of stdlib.h, so a stays 0 where |d| < 1; kept.  The
 * y half-step profile of the bottom strip is built from d0_x
 * (src/PML_pro.c:299); kept (d0_x == d0_y for square cells). */
typedef struct {
    float *d, *K, *al, *a, *b;
} prof_t;

static int iabs_trunc(float v)
{
    int k = (int)v;
    return k < 0 ? -k : k;
}

static void pml_fill(const ifos_config *c, int lower, int N, float d0, float d0h, prof_t full, prof_t half,
                     int clamp)
{
    const float DH = c->DH, DT = c->DT;
    const float alpha_max = 2.0 * IFOS_PI * (c->FPML / 2.0);
    float thick = c->FW * DH;
    float origin = lower ? thick : (N - 1) * DH - thick;
    int i0 = lower ? 1 : N - c->FW + 1, i1 = lower ? c->FW : N;
    for (int i = i0; i <= i1; i++) {
        int h = lower ? i : i - N + 2 * c->FW;
        full.K[h] = 1.0;
        half.K[h] = 1.0;
        float xval = DH * (i - 1);
        float absc = lower ? origin - xval : xval - origin;
        if (absc >= 0.0) {
            float an = absc / thick;
            full.d[h] = d0 * pow(an, c->npower);
            full.K[h] = 1.0 + (c->k_max_PML - 1.0) * pow(an, c->npower);
            full.al[h] = alpha_max * (1.0 - an);
        }
        absc = lower ? origin - (xval + DH / 2.0) : xval + DH / 2.0 - origin;
        if (absc >= 0.0) {
            float an = absc / thick;
            half.d[h] = d0h * pow(an, c->npower);
            half.K[h] = 1.0 + (c->k_max_PML - 1.0) * pow(an, c->npower);
            half.al[h] = alpha_max * (1.0 - an);
        }
        if (clamp) {
            if (full.al[h] < 0.0) full.al[h] = 0.0;
            if (half.al[h] < 0.0) half.al[h] = 0.0;
        }
        full.b[h] = exp(-(full.d[h] / full.K[h] + full.al[h]) * DT);
        half.b[h] = exp(-(half.d[h] / half.K[h] + half.al[h]) * DT);
        if (iabs_trunc(full.d[h]) > 1.0e-6)
            full.a[h] = full.d[h] * (full.b[h] - 1.0) / (full.K[h] * (full.d[h] + full.K[h] * full.al[h]));
        if (iabs_trunc(half.d[h]) > 1.0e-6)
            half.a[h] = half.d[h] * (half.b[h] - 1.0) / (half.K[h] * (half.d[h] + half.K[h] * half.al[h]));
    }
}

void ifos_pml_profiles(const ifos_config *c, float *out12)
{
    int n = 2 * c->FW;
    if (n <= 0) return;
    float *w = (float *)calloc((size_t)20 * (n + 1), sizeof(float));
    prof_t P[4];
    for (int k = 0; k < 4; k++) {
        float *b = w + (size_t)k * 5 * (n + 1);
        P[k].d = b;
        P[k].K = b + (n + 1);
        P[k].al = b + 2 * (n + 1);
        P[k].a = b + 3 * (n + 1);
        P[k].b = b + 4 * (n + 1);
    }
    float thick = c->FW * c->DH;
    float Rcoef = 0.001;
    float d0_x = -(c->npower + 1) * c->VPPML * log(Rcoef) / (2.0 * thick);
    float d0_y = -(c->npower + 1) * c->VPPML * log(Rcoef) / (2.0 * thick);
    /* x: full = P[0], half = P[1]; y: full = P[2], half = P[3] */
    pml_fill(c, 1, c->NX, d0_x, d0_x, P[0], P[1], 1);
    pml_fill(c, 0, c->NX, d0_x, d0_x, P[0], P[1], 1);
    pml_fill(c, 1, c->NY, d0_y, d0_y, P[2], P[3], 0);
    pml_fill(c, 0, c->NY, d0_y, d0_x, P[2], P[3], 0);
    const float *src[12] = {P[0].K, P[0].a, P[0].b, P[1].K, P[1].a, P[1].b,
                            P[2].K, P[2].a, P[2].b, P[3].K, P[3].a, P[3].b};
    for (int k = 0; k < 12; k++) memcpy(out12 + (size_t)k * n, src[k] + 1, sizeof(float) * (size_t)n);
    free(w);
}

/* ---- wavelet (src/wavelet.c:29-125, P-SV shapes computed in closed form) - */
#define IROUND(x) ((int)(floor)((x) + 0.5))
void ifos_wavelet(const ifos_config *c, int nsrc, const float *srcpos, float *out)
{
    for (int nt = 1; nt <= c->NT; nt++) {
        float t = (float)nt * c->DT;
        for (int k = 0; k < nsrc; k++) {
            const float *s = srcpos + (size_t)k * IFOS_SRC_STRIDE;
            float tshift = s[3], fc = s[4], a = s[5];
            float ts = 1.0 / fc, amp = 0.0f, tau, ag;
            switch (c->SOURCE_SHAPE) {
            case 1: /* Ricker */
                tau = IFOS_PI * (t - 1.5 * ts - tshift) / (ts);
                amp = (((1.0 - 2.0 * tau * tau) * exp(-tau * tau)));
                break;
            case 2: /* Fuchs-Mueller */
                if ((t < tshift) || (t > (tshift + ts))) amp = 0.0;
                else amp = ((sin(2.0 * IFOS_PI * (t - tshift) * fc) - 0.5 * sin(4.0 * IFOS_PI * (t - tshift) * fc)));
                break;
            case 4: /* sin^3 */
                if ((t < tshift) || (t > (tshift + ts))) amp = 0.0;
                else amp = pow(sin(IFOS_PI * (t - tshift) / ts), 3.0);
                break;
            case 5: /* first derivative of a Gaussian */
                ts = 1.2 / fc;
                ag = IFOS_PI * IFOS_PI * fc * fc;
                amp = -2.0 * ag * (t - ts) * exp(-ag * (t - ts) * (t - ts));
                break;
            case 6: /* band-limited spike */
                amp = 0.0;
                if (nt == 1 + IROUND(tshift / c->DT)) amp = 1.0;
                break;
            case 8: /* integral of sin^3 */
                if (t < tshift) amp = 0.0;
                if ((t >= tshift) && (t <= (tshift + ts)))
                    amp = (ts / (0.75 * IFOS_PI)) * (0.5 - 0.75 * cos(IFOS_PI * (t - tshift) / ts) +
                                                     0.25 * pow(cos(IFOS_PI * (t - tshift) / ts), 3.0));
                if (t > (tshift + ts)) amp = ts / (0.75 * IFOS_PI);
                break;
            default:
                amp = 0.0;
            }
            out[(size_t)k * c->NT + (nt - 1)] = amp * a;
        }
    }
}

/* ---- materials ----------------------------------------------------------- */
/* Padded host scratch with a one-cell halo: a(j,i), j,i in 0..N+1. */
#define PAD(a, j, i) (a)[(size_t)(j) * (NX + 2) + (size_t)(i)]

/* Periodic one-cell halo of a single-rank run: matcopy_elastic / matcopy
 * send row 1 to row NY+1, row NY to row 0, then column 1 to column NX+1 and
 * column NX to column 0 over all rows (src/matcopy_elastic.c:48-140). */
static void halo_wrap(float *a, int NX, int NY)
{
    for (int i = 0; i <= NX + 1; i++) {
        PAD(a, NY + 1, i) = PAD(a, 1, i);
        PAD(a, 0, i) = PAD(a, NY, i);
    }
    for (int j = 0; j <= NY + 1; j++) {
        PAD(a, j, NX + 1) = PAD(a, j, 1);
        PAD(a, j, 0) = PAD(a, j, NX);
    }
}

static float *pad_copy(const float *src, int NX, int NY)
{
    float *a = (float *)calloc((size_t)(NX + 2) * (NY + 2), sizeof(float));
    for (int j = 1; j <= NY; j++) memcpy(&PAD(a, j, 1), src + (size_t)(j - 1) * NX, sizeof(float) * NX);
    halo_wrap(a, NX, NY);
    return a;
}

int ifos_setup_materials(const ifos_config *c, const float *vp, const float *vs, const float *rho,
                         const float *taus, const float *taup, ifos_materials *m)
{
    const int NX = c->NX, NY = c->NY, L = c->L;
    const size_t n = (size_t)NX * NY;
    memset(m, 0, sizeof(*m));
    float *R = pad_copy(rho, NX, NY), *U = pad_copy(vs, NX, NY), *V = pad_copy(vp, NX, NY);
    float *TS = L ? pad_copy(taus, NX, NY) : NULL, *TP = L ? pad_copy(taup, NX, NY) : NULL;
    m->rip = (float *)malloc(n * sizeof(float));
    m->rjp = (float *)malloc(n * sizeof(float));
    m->uipjp = (float *)malloc(n * sizeof(float));
    m->muss = (float *)malloc(n * sizeof(float));
    m->lamss = (float *)malloc(n * sizeof(float));
    m->f = (float *)malloc(n * sizeof(float));
    m->g = (float *)malloc(n * sizeof(float));
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            size_t k = (size_t)(j - 1) * NX + (i - 1);
            /* av_rho (src/av_rho.c:27-31) */
            m->rjp[k] = 1.0 / (0.5 * (PAD(R, j, i) + PAD(R, j + 1, i)));
            m->rip[k] = 1.0 / (0.5 * (PAD(R, j, i) + PAD(R, j, i + 1)));
            /* av_mue, PARAMETERIZATION 1 (src/av_mue.c:45-62) */
            float u1 = PAD(R, j, i) * PAD(U, j, i) * PAD(U, j, i);
            float u2 = PAD(R, j, i + 1) * PAD(U, j, i + 1) * PAD(U, j, i + 1);
            float u3 = PAD(R, j + 1, i) * PAD(U, j + 1, i) * PAD(U, j + 1, i);
            float u4 = PAD(R, j + 1, i + 1) * PAD(U, j + 1, i + 1) * PAD(U, j + 1, i + 1);
            m->uipjp[k] = 4.0 / ((1.0 / u1) + (1.0 / u2) + (1.0 / u3) + (1.0 / u4));
            if ((u1 == 0.0) || (u2 == 0.0) || (u3 == 0.0) || (u4 == 0.0)) m->uipjp[k] = 0.0;
            /* moduli used by the correlation and chain rule (src/IFOS2D.c:2395-2396) */
            float muss = PAD(R, j, i) * PAD(U, j, i) * PAD(U, j, i);
            float lamss = PAD(R, j, i) * PAD(V, j, i) * PAD(V, j, i) - 2.0 * muss;
            m->muss[k] = muss;
            m->lamss[k] = lamss;
            /* elastic moduli of update_s_elastic_PML (src/update_s_elastic_PML.c:228-230) */
            float rr = PAD(R, j, i), uu = PAD(U, j, i), pp = PAD(V, j, i);
            m->f[k] = rr * uu * uu;
            m->g[k] = rr * ((pp * pp) - 2 * uu * uu);
        }
    if (L) {
        /* eta (src/readmod.c:47-51) and prepare_update_s (src/prepare_update_s.c:21-108) */
        float pts[IFOS_MAXL + 1];
        for (int l = 1; l <= L; l++) {
            pts[l] = 1.0 / (2.0 * IFOS_PI * c->FL[l]);
            m->eta[l] = c->DT / pts[l];
        }
        float ws = 2.0 * IFOS_PI * c->F_REF;
        float sumu = 0.0, sumpi = 0.0;
        for (int l = 1; l <= L; l++) {
            sumu = sumu + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l]));
            sumpi = sumpi + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l]));
        }
        m->sumu = sumu;
        m->sumpi = sumpi;
        for (int l = 1; l <= L; l++) {
            m->bip[l] = 1.0 / (1.0 + (m->eta[l] * 0.5));
            m->bjm[l] = 1.0 / (1.0 + (m->eta[l] * 0.5));
            m->cip[l] = 1.0 - (m->eta[l] * 0.5);
            m->cjm[l] = 1.0 - (m->eta[l] * 0.5);
        }
        m->fipjp = (float *)malloc(n * sizeof(float));
        m->d = (float *)malloc(n * L * sizeof(float));
        m->e = (float *)malloc(n * L * sizeof(float));
        m->dip = (float *)malloc(n * L * sizeof(float));
        m->mu_r = (float *)malloc(n * sizeof(float));
        m->pi_r = (float *)malloc(n * sizeof(float));
        m->muipjp_r = (float *)malloc(n * sizeof(float));
        m->tausipjp = (float *)malloc(n * sizeof(float));
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                size_t k = (size_t)(j - 1) * NX + (i - 1);
                /* av_tau (src/av_tau.c:25-27) */
                float tsipjp = 0.25 * (PAD(TS, j, i) + PAD(TS, j, i + 1) + PAD(TS, j + 1, i) + PAD(TS, j + 1, i + 1));
                float u = PAD(U, j, i), pv = PAD(V, j, i), rh = PAD(R, j, i);
                float ts = PAD(TS, j, i), tp = PAD(TP, j, i);
                float mu = (u * u * rh) / (1.0 + sumu * ts);
                float pi = (pv * pv * rh) / (1.0 + sumpi * tp);
                float muipjp = m->uipjp[k] / (1.0 + sumu * tsipjp);
                m->mu_r[k] = mu;
                m->pi_r[k] = pi;
                m->muipjp_r[k] = muipjp;
                m->tausipjp[k] = tsipjp;
                m->fipjp[k] = muipjp * c->DT * (1.0 + L * tsipjp);
                m->f[k] = mu * c->DT * (1.0 + L * ts);
                m->g[k] = pi * c->DT * (1.0 + L * tp);
                for (int l = 1; l <= L; l++) {
                    m->dip[(size_t)(l - 1) * n + k] = muipjp * m->eta[l] * tsipjp;
                    m->d[(size_t)(l - 1) * n + k] = mu * m->eta[l] * ts;
                    m->e[(size_t)(l - 1) * n + k] = pi * m->eta[l] * tp;
                }
            }
    }
    free(R); free(U); free(V); free(TS); free(TP);
    return 0;
}

void ifos_free_materials(ifos_materials *m)
{
    free(m->rip); free(m->rjp); free(m->uipjp); free(m->muss); free(m->lamss);
    free(m->f); free(m->g); free(m->fipjp); free(m->d); free(m->e); free(m->dip);
    free(m->mu_r); free(m->pi_r); free(m->muipjp_r); free(m->tausipjp);
    memset(m, 0, sizeof(*m));
}
