/*
 * post_oracle.c -- CPU restatement of the reference's gradient
 * post-processing.  TEST INFRASTRUCTURE ONLY: it is the checker for the
 * device tapers and the median smoothing, never part of the product.
 *
 *   orc_taper_sources     taper_grad_shot sws 1 src/taper_grad_shot.c:57-196,
 *                         taper_grad sws 3 src/taper_grad.c:252-398
 *   orc_taper_vertical    taper_grad sws 1 src/taper_grad.c:60-172
 *   orc_taper_horizontal  taper_grad sws 2 src/taper_grad.c:179-246
 *   orc_spat_filt         spat_filt() src/spat_filt.c:27-150 (SPATFILTER)
 *   orc_median_smooth     smooth() src/smooth.c:24-160 (filter 1 and 2 differ
 *                         only in which global holds the size), median2d()
 *                         src/median2D.c:3-20 with the reference's off-by-one
 *                         sort (see orc_median_smooth)
 *
 * FP32 with the reference's double promotions (compile with
 * -ffp-contract=off).  Pinned against the reference's own functions,
 * compiled from source into oracle/_ref (tests/test_post.py), and against
 * closed forms (a window's median by brute force, the erf taper's values).
 *
 * The zeroed box around a source is written, as in the reference, into a
 * (NY+2) x (NX+2) scratch matrix whose rows are contiguous: a box column
 * i <= -2 lands in column NX+2+i of the row above, which is then copied
 * back (src/taper_grad_shot.c:168-179 has no lower bound on i beyond
 * (i-1)/NX == 0).  The restatement keeps that, so sources closer than
 * FILTSIZE-1 cells to the left edge zero the same cells as the reference.
 */
#include "post_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* --- circular source taper ------------------------------------------------ */

/* side of the local taper table: ceil(2 R / DH), at least 5
 * (src/taper_grad_shot.c:63-65) */
static int srctaper_side(float DH, float R)
{
    int n = (int)ceil(2.0 * R / DH);
    return n < 5 ? 5 : n;
}

/* local table m [n][n] (row = y), src/taper_grad_shot.c:80-122 */
static void srctaper_table(int n, float DH, float R, int shape, float *m)
{
    const float a = 1.0f;
    const float maxrad = (float)sqrt(2.0 * R * R);
    for (int jy = 1; jy <= n; jy++)
        for (int ix = 1; ix <= n; ix++) {
            float x = (float)(((float)ix - ((float)n) / 2.0 - 0.5) * DH);
            float y = (float)(((float)jy - ((float)n) / 2.0 - 0.5) * DH);
            float rad = (float)sqrt((double)(x * x + y * y));
            float v = 0.0f;  /* matrix() zero-fills; other shapes leave 0 */
            if (shape == 1) v = (float)erf((double)(a * rad / maxrad));
            else if (shape == 2) v = rad > 0 ? (float)log((double)rad) : 0.0f;
            m[(jy - 1) * n + (ix - 1)] = v;
        }
    /* shift to a zero minimum (minimum_m, src/util.c:56) */
    float lo = m[0];
    for (int k = 1; k < n * n; k++)
        if (m[k] < lo) lo = m[k];
    for (int k = 0; k < n * n; k++) m[k] -= lo;
    /* scale by the largest of the four edge-centre values, clip at 1 */
    const int h = n / 2;
    float edge[4] = {m[(1 - 1) * n + (h - 1)], m[(h - 1) * n + (1 - 1)], m[(h - 1) * n + (n - 1)],
                     m[(n - 1) * n + (h - 1)]};
    float hi = edge[0];
    for (int k = 1; k < 4; k++)
        if (edge[k] > hi) hi = edge[k];
    for (int k = 0; k < n * n; k++) {
        m[k] /= hi;
        if (m[k] > 1.0f) m[k] = 1.0f;
    }
}

/* scratch matrix(0,NY+1,0,NX+1) of the reference, rows contiguous */
#define PADIDX(j, i) ((long)(j) * (NX + 2) + (i))

void orc_taper_sources(int NX, int NY, float DH, float SRTRADIUS, int SRTSHAPE, int FILTSIZE, int nshots,
                       const int *isrc, const int *jsrc, float *g, float *coef)
{
    const int n = srctaper_side(DH, SRTRADIUS);
    float *m = malloc(sizeof(float) * n * n);
    srctaper_table(n, DH, SRTRADIUS, SRTSHAPE, m);
    const int c = (int)ceil((float)n / 2);
    const long plane = (long)NX * NY;
    float *msum = malloc(sizeof(float) * plane);
    float *mm = malloc(sizeof(float) * plane);
    for (long e = 0; e < plane; e++) msum[e] = 1.0f;
    for (int s = 0; s < nshots; s++) {
        for (long e = 0; e < plane; e++) mm[e] = 1.0f;
        for (int iy = 1; iy <= n; iy++)
            for (int ix = 1; ix <= n; ix++) {
                int xx = isrc[s] + ix - c, yy = jsrc[s] + iy - c;
                if (xx < 1 || xx > NX || yy < 1 || yy > NY) continue;
                mm[(long)(yy - 1) * NX + (xx - 1)] = m[(iy - 1) * n + (ix - 1)];
            }
        for (long e = 0; e < plane; e++)
            if (msum[e] > mm[e]) msum[e] = mm[e];
    }
    float lo = msum[0];
    for (long e = 1; e < plane; e++)
        if (msum[e] < lo) lo = msum[e];
    for (long e = 0; e < plane; e++) msum[e] -= lo;
    float hi = msum[0];
    for (long e = 1; e < plane; e++)
        if (msum[e] > hi) hi = msum[e];
    float *tmp = calloc((size_t)(NY + 2) * (NX + 2), sizeof(float));
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            const long e = (long)(j - 1) * NX + (i - 1);
            const float t = msum[e] / hi;
            if (coef) coef[e] = t;
            g[e] *= t;
            tmp[PADIDX(j, i)] = g[e];
        }
    for (int s = 0; s < nshots; s++)
        for (int i = isrc[s] - FILTSIZE; i <= isrc[s] + FILTSIZE; i++)
            for (int j = jsrc[s] - FILTSIZE; j <= jsrc[s] + FILTSIZE; j++)
                if ((i - 1) / NX == 0 && (j - 1) / NY == 0 && j > 0) tmp[PADIDX(j, i)] = 0.0f;
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) g[(long)(j - 1) * NX + (i - 1)] = tmp[PADIDX(j, i)];
    free(tmp);
    free(mm);
    free(msum);
    free(m);
}

/* --- vertical and horizontal tapers -------------------------------------- */

/* Gaussian flank value exp(-(1/2) (a k / (len/2))^2), a = 3, with the
 * reference's float/double mix (src/taper_grad.c:86-98) */
static float flank(int k, int len)
{
    const float a = 3;
    double q = (double)(a * k) / (len / 2.0);
    return (float)exp(-(1.0 / 2.0) * q * q);
}

void orc_taper_vertical(int NX, int NY, int GRADT1, int GRADT2, int GRADT3, int GRADT4, int nshots,
                        const int *isrc, const int *jsrc, int ntr, const int *recpos, float *g, float *coef)
{
    const int t1 = GRADT2 - GRADT1, t2 = GRADT4 - GRADT3, w = GRADT4 - GRADT1;
    float *win = malloc(sizeof(float) * (w > 0 ? w + 1 : 1));
    for (int k = 1; k <= w; k++) {
        win[k] = 1.0f;
        if (k <= t1) win[k] = flank(k - t1, t1);
        if (k >= w - t2) win[k] = flank(k - (w - t2), t2);
    }
    float *tmp = calloc((size_t)(NY + 2) * (NX + 2), sizeof(float));
    for (int j = 1; j <= NY; j++) {
        int h = 1;  /* columns strictly between GRADT1 and GRADT4 take win[1], win[2], ... */
        for (int i = 1; i <= NX; i++) {
            float t = 0.0f;
            if (i > GRADT1 && i < GRADT4) t = win[h++];
            const long e = (long)(j - 1) * NX + (i - 1);
            if (coef) coef[e] = t;
            g[e] *= t;
            tmp[PADIDX(j, i)] = g[e];
        }
    }
    const long cap = (long)(NY + 2) * (NX + 2);
    for (int s = 0; s < nshots; s++) {
        int i = isrc[s], j = jsrc[s];
        long p = PADIDX(j, i);
        if ((i - 1) / NX == 0 && (j - 1) / NY == 0 && p >= 0 && p < cap) tmp[p] = 0.0f;
    }
    for (int r = 0; r < ntr; r++) {
        int i = recpos[2 * r], j = recpos[2 * r + 1];
        long p = PADIDX(j, i);
        if ((i - 1) / NX == 0 && (j - 1) / NY == 0 && p >= 0 && p < cap) tmp[p] = 0.0f;
    }
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) g[(long)(j - 1) * NX + (i - 1)] = tmp[PADIDX(j, i)];
    free(tmp);
    free(win);
}

void orc_taper_horizontal(int NX, int NY, int GRADT1, int GRADT2, float *g, float *coef)
{
    const int t1 = GRADT2 - GRADT1, w = GRADT2 - GRADT1 + 1;
    float *win = malloc(sizeof(float) * (w > 0 ? w + 1 : 1));
    for (int k = 1; k <= w; k++) win[k] = k <= t1 ? flank(k - t1, t1) : 1.0f;
    int h = 1;  /* rows GRADT1..GRADT2 take win[1], win[2], ...; deeper rows 1 */
    for (int j = 1; j <= NY; j++) {
        float t = 0.0f;
        if (j >= GRADT1 && j <= GRADT2) t = win[h];
        if (j > GRADT2) t = 1.0f;
        for (int i = 1; i <= NX; i++) {
            const long e = (long)(j - 1) * NX + (i - 1);
            if (coef) coef[e] = t;
            g[e] *= t;
        }
        if (j >= GRADT1) h++;
    }
    free(win);
}

/* --- median smoothing ----------------------------------------------------- */

static int cmp_float(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

int orc_median_smooth(int NX, int NY, int filtsize, const float *in, float *out)
{
    const long plane = (long)NX * NY;
    if (filtsize == 0) {
        memcpy(out, in, sizeof(float) * plane);
        return 0;
    }
    if (!(filtsize % 2)) filtsize += filtsize > 0 ? 1 : -1;
    const int side = abs(filtsize), hfs = side / 2, k = side * side;
    float *win = malloc(sizeof(float) * k);
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            /* window in smooth()'s order: x offset outer, y offset inner
             * (filterpart[ii][jj], src/smooth.c:116-121) */
            int q = 0;
            for (int di = -hfs; di <= hfs; di++) {
                int ii = i + di < 1 ? 1 : (i + di > NX ? NX : i + di);
                for (int dj = -hfs; dj <= hfs; dj++) {
                    int jj = j + dj < 1 ? 1 : (j + dj > NY ? NY : j + dj);
                    win[q++] = in[(long)(jj - 1) * NX + (ii - 1)];
                }
            }
            float med;
            if (k == 1) {
                med = win[0];
            } else {
                /* median2d copies the window to t[1..k] of a zero-filled
                 * vector and calls quicksort(t, 1, k), which sorts the k
                 * elements t[0..k-1] (src/util.c:437-446): the zero at t[0]
                 * and the window without its last element; it then returns
                 * t[k/2+1] of that order.  Kept as the reference computes it. */
                memmove(win + 1, win, sizeof(float) * (k - 1));
                win[0] = 0.0f;
                qsort(win, k, sizeof(float), cmp_float);
                med = win[k / 2 + 1];
            }
            out[(long)(j - 1) * NX + (i - 1)] = med;
        }
    free(win);
    return side;
}

/* --- spatial filter --------------------------------------------------------- */

/* spat_filt (src/spat_filt.c:27-150) on the whole gradient: for columns
 * S+1..NX-S and rows J1..NY-S, a weighted mean over columns i-S..i+S of the
 * same row.  The weight is (1/(sqrt(2 pi) S)) * exp((-1/2) * ...) where
 * (-1/2) is integer division, so every weight is exp(0) = 1 times the
 * prefactor: a boxcar mean, accumulated in float with double products as
 * there.  Other cells keep their values.  S = SPAT_FILT_SIZE (already
 * reduced by the iteration when SPAT_FILT_ITER is set), J1 = SPAT_FILT_1. */
void orc_spat_filt(int NX, int NY, int S, int J1, const float *in, float *out)
{
    const long plane = (long)NX * NY;
    memcpy(out, in, sizeof(float) * plane);
    if (S < 1) return;
    const double PI = 3.141592653589793;
    for (int i = 1 + S; i <= NX - S; i++)
        for (int j = J1; j <= NY - S; j++) {
            float acc = 0.0f, norm = 0.0f;
            for (int i1 = i - S; i1 <= i + S; i1++) {
                const double d = sqrt((double)((i - i1) * (i - i1) + (i - i1) * (i - i1)));
                const double wgt = (1 / (sqrt(2 * PI) * S)) * exp((-1 / 2) * (d * d) / (S * S));
                acc = (float)((double)acc + (double)in[(long)(j - 1) * NX + (i1 - 1)] * wgt);
                norm = (float)((double)norm + wgt);
            }
            out[(long)(j - 1) * NX + (i - 1)] = acc / norm;
        }
}
