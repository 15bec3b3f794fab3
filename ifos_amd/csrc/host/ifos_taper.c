/*
 * ifos_taper.c -- host-side taper coefficients of the gradient
 * post-processing: the circular source taper table (taper_grad_shot sws 1,
 * src/taper_grad_shot.c:57-122, and taper_grad sws 3) and the whole-gradient
 * coefficient grids of taper_grad (src/taper_grad.c:28).  They are built
 * once per shot set on the host, like the CPML profiles; the device applies
 * them (ifos_opt_taper, and the per-shot taper fused into the gradient
 * kernel).  The float results equal the reference's bit for bit: each
 * expression keeps its operand types (float products, double divisions,
 * glibc erf/log/exp), and the file is compiled with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ifos2d_opt.h"

void ifos_set_error(const char *fmt, ...);

/* Side of the source taper table: ceil(2 R / DH), at least 5. */
static int srctaper_side(float DH, float R)
{
    const int side = (int)ceil(2.0 * R / DH);
    return side > 5 ? side : 5;
}

int ifos_srctaper_table(float DH, float SRTRADIUS, int SRTSHAPE, float *m, int cap)
{
    const int n = srctaper_side(DH, SRTRADIUS);
    if (!m || cap < n * n) return n;
    const float maxrad = (float)sqrt(2.0 * SRTRADIUS * SRTRADIUS);
    const double half = ((float)n) / 2.0;
    float lo = 0.0f;
    for (int r = 0; r < n; r++) {
        const float y = (float)(((float)(r + 1) - half - 0.5) * DH);
        for (int q = 0; q < n; q++) {
            const float x = (float)(((float)(q + 1) - half - 0.5) * DH);
            const float rad = (float)sqrt((double)(x * x + y * y));
            float v;
            switch (SRTSHAPE) {
            case 1: v = (float)erf((double)(1.0f * rad / maxrad)); break;
            case 2: v = rad > 0 ? (float)log((double)rad) : 0.0f; break;
            default: v = 0.0f;
            }
            m[r * n + q] = v;
            if ((r == 0 && q == 0) || v < lo) lo = v;
        }
    }
    for (int k = 0; k < n * n; k++) m[k] = m[k] - lo;
    /* normalised by the largest edge-centre value, then clipped at 1 */
    const int h = n / 2 - 1;  /* 0-based index of the reference's n/2 */
    const float e[4] = {m[h], m[h * n], m[h * n + n - 1], m[(n - 1) * n + h]};
    float top = e[0];
    for (int k = 1; k < 4; k++) top = e[k] > top ? e[k] : top;
    for (int k = 0; k < n * n; k++) {
        const float v = m[k] / top;
        m[k] = v > 1.0f ? 1.0f : v;
    }
    return n;
}

int ifos_srctaper_centre(int n)
{
    return (int)ceil((float)n / 2);
}

/* Normalisation of one shot's taper: msum = table entries placed around the
 * source (in the grid) and 1 elsewhere; lo = min msum, hi = max(msum - lo). */
void ifos_srctaper_norm(int NX, int NY, int n, const float *m, int is, int js, float *lo, float *hi)
{
    const int c = ifos_srctaper_centre(n);
    long covered = 0;
    float vmin = 1.0f, vmax = -INFINITY;
    int any = 0;
    for (int r = 0; r < n; r++) {
        const int y = js + (r + 1) - c;
        if (y < 1 || y > NY) continue;
        for (int q = 0; q < n; q++) {
            const int x = is + (q + 1) - c;
            if (x < 1 || x > NX) continue;
            const float v = m[r * n + q];
            if (!any || v < vmin) vmin = v;
            if (!any || v > vmax) vmax = v;
            any = 1;
            covered++;
        }
    }
    if (covered < (long)NX * NY) {
        if (!any || 1.0f < vmin) vmin = 1.0f;
        if (!any || 1.0f > vmax) vmax = 1.0f;
    }
    *lo = vmin;
    *hi = vmax - vmin;
}

/* Marks box cells of the reference's (NY+2) x (NX+2) scratch matrix whose
 * rows are contiguous: column i <= -2 of row j is column NX+2+i of row j-1. */
static void mark_cell(int NX, int NY, unsigned char *zero, int i, int j)
{
    const long p = (long)j * (NX + 2) + i;
    if (p < 0) return;
    const long row = p / (NX + 2), col = p % (NX + 2);
    if (row >= 1 && row <= NY && col >= 1 && col <= NX) zero[(row - 1) * NX + (col - 1)] = 1;
}

static float gauss_flank(int k, int len)
{
    const float a = 3;
    const double q = (double)(a * k) / (len / 2.0);
    return (float)exp(-(1.0 / 2.0) * q * q);
}

int ifos_taper_coeff(const ifos_taper *t, int NX, int NY, float *coef, unsigned char *zero)
{
    if (!t || NX < 1 || NY < 1 || !coef) return ifos_set_error("ifos_taper_coeff: bad arguments"), -2;
    const long plane = (long)NX * NY;
    if (zero) memset(zero, 0, (size_t)plane);
    if (t->kind == IFOS_TAPER_VERTICAL) {
        const int len1 = t->GRADT2 - t->GRADT1, len2 = t->GRADT4 - t->GRADT3, w = t->GRADT4 - t->GRADT1;
        for (int i = 1; i <= NX; i++) {
            float v = 0.0f;
            if (i > t->GRADT1 && i < t->GRADT4) {
                const int k = i - t->GRADT1;  /* window index 1 .. w-1 */
                v = 1.0f;
                if (k <= len1) v = gauss_flank(k - len1, len1);
                if (k >= w - len2) v = gauss_flank(k - (w - len2), len2);
            }
            for (int j = 0; j < NY; j++) coef[(long)j * NX + (i - 1)] = v;
        }
        if (zero) {
            for (int s = 0; s < t->nsrc; s++) {
                const int i = t->srcpos[2 * s], j = t->srcpos[2 * s + 1];
                if ((i - 1) / NX == 0 && (j - 1) / NY == 0) mark_cell(NX, NY, zero, i, j);
            }
            for (int r = 0; r < t->ntr; r++) {
                const int i = t->recpos[2 * r], j = t->recpos[2 * r + 1];
                if ((i - 1) / NX == 0 && (j - 1) / NY == 0) mark_cell(NX, NY, zero, i, j);
            }
        }
        return 0;
    }
    if (t->kind == IFOS_TAPER_HORIZONTAL) {
        const int len = t->GRADT2 - t->GRADT1;
        for (int j = 1; j <= NY; j++) {
            float v = 0.0f;
            if (j >= t->GRADT1 && j <= t->GRADT2) {
                const int k = j - t->GRADT1 + 1;
                v = k <= len ? gauss_flank(k - len, len) : 1.0f;
            }
            if (j > t->GRADT2) v = 1.0f;
            for (int i = 0; i < NX; i++) coef[(long)(j - 1) * NX + i] = v;
        }
        return 0;
    }
    if (t->kind == IFOS_TAPER_SOURCES) {
        const int n = srctaper_side(t->DH, t->SRTRADIUS);
        float *m = (float *)malloc(sizeof(float) * n * n);
        if (!m) return ifos_set_error("ifos_taper_coeff: out of memory"), -1;
        ifos_srctaper_table(t->DH, t->SRTRADIUS, t->SRTSHAPE, m, n * n);
        const int c = ifos_srctaper_centre(n);
        for (long e = 0; e < plane; e++) coef[e] = 1.0f;
        for (int s = 0; s < t->nsrc; s++) {
            const int is = t->srcpos[2 * s], js = t->srcpos[2 * s + 1];
            for (int r = 0; r < n; r++) {
                const int y = js + (r + 1) - c;
                if (y < 1 || y > NY) continue;
                for (int q = 0; q < n; q++) {
                    const int x = is + (q + 1) - c;
                    if (x < 1 || x > NX) continue;
                    float *d = &coef[(long)(y - 1) * NX + (x - 1)];
                    if (*d > m[r * n + q]) *d = m[r * n + q];
                }
            }
        }
        free(m);
        float lo = coef[0], hi;
        for (long e = 1; e < plane; e++) lo = coef[e] < lo ? coef[e] : lo;
        for (long e = 0; e < plane; e++) coef[e] = coef[e] - lo;
        hi = coef[0];
        for (long e = 1; e < plane; e++) hi = coef[e] > hi ? coef[e] : hi;
        for (long e = 0; e < plane; e++) coef[e] = coef[e] / hi;
        if (zero) {
            const int F = t->FILTSIZE;
            for (int s = 0; s < t->nsrc; s++) {
                const int is = t->srcpos[2 * s], js = t->srcpos[2 * s + 1];
                for (int i = is - F; i <= is + F; i++)
                    for (int j = js - F; j <= js + F; j++)
                        if ((i - 1) / NX == 0 && (j - 1) / NY == 0 && j > 0) mark_cell(NX, NY, zero, i, j);
            }
        }
        return 0;
    }
    return ifos_set_error("ifos_taper_coeff: unknown taper kind %d", t->kind), -2;
}
