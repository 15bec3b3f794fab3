/*
 * filt_oracle.c -- CPU restatement of the reference's time-domain
 * Butterworth filter.  TEST INFRASTRUCTURE ONLY: the checker of the device
 * filter, never part of the product.
 *
 *   orc_time_filter   timedomain_filt() src/timedomain_filt.c:27-93 with
 *                     seife_lpb / seife_hpb, seife_butterworth, seife_rfk,
 *                     seife_rekfl and seife_first of the vendored
 *                     contrib/libcseife (cseife.c:58-68, 156-197,
 *                     cseife_rfk.c, cseife_rekfl.c:53-89)
 *
 * Double precision as there; compile with -ffp-contract=off.  Pinned
 * against the reference's own timedomain_filt and libcseife compiled from
 * source into oracle/_ref (tests/test_filter.py), and against closed forms
 * (DC gain of a lowpass section, a highpass removing a constant).
 */
#include <math.h>
#include <stdlib.h>

/* one recursive section y[j] = f0 x[j] + f1 x[j-1] + f2 x[j-2] + g1 y[j-1]
 * + g2 y[j-2], zero initial state, in place */
static void section(double *x, int n, const double c[5])
{
    double x1 = 0.0, x2 = 0.0, y1 = 0.0, y2 = 0.0;
    for (int j = 0; j < n; j++) {
        const double xn = x[j];
        const double y = c[0] * xn + c[1] * x1 + c[2] * x2 + c[3] * y1 + c[4] * y2;
        x[j] = y;
        x2 = x1;
        x1 = xn;
        y2 = y1;
        y1 = y;
    }
}

/* coefficients {f0, f1, f2, g1, g2} of a first-order (h < 0) or
 * second-order (damping h) low- or highpass with period t0 in samples */
static void coeffs(int highpass, int second, double t0, double h, double c[5])
{
    const double eps = 8.0 * atan(1.0) / t0;
    if (!second) {
        c[3] = (2.0 - eps) / (2.0 + eps);
        c[4] = 0.0;
        c[2] = 0.0;
        if (highpass) {
            c[0] = 2.0 / (2.0 + eps);
            c[1] = -(c[0]);
        } else {
            c[0] = eps / (2.0 + eps);
            c[1] = c[0];
        }
        return;
    }
    const double e2 = eps * eps;
    const double a = 1.0 - eps * h + e2 / 4.0;
    const double b = -2.0 + e2 / 2.0;
    const double cc = 1.0 + eps * h + e2 / 4.0;
    c[3] = -b / cc;
    c[4] = -a / cc;
    if (highpass) {
        c[0] = 1.0 / cc;
        c[1] = -(c[0]) - c[0];
        c[2] = c[0];
    } else {
        c[0] = e2 / 4.0 / cc;
        c[1] = c[0] + c[0];
        c[2] = c[0];
    }
}

/* seife_butterworth on x[0..n-1]: an odd order starts with a first-order
 * section, then order/2 second-order sections with damping
 * sin(pi/2 (2j-1)/order); a highpass first subtracts x[0] */
static void butterworth(double *x, int n, double dt, int highpass, double t0, int order)
{
    if (highpass) {
        const double x0 = x[0];
        for (int i = 0; i < n; i++) x[i] -= x0;
    }
    double c[5];
    const int half = order / 2;
    if (order > 2 * half) {
        coeffs(highpass, 0, t0 / dt, 0.0, c);
        section(x, n, c);
    }
    const double pih = 2.0 * atan(1.0);
    for (int j = 1; j <= half; j++) {
        const double h = sin(pih * (2 * j - 1) / order);
        coeffs(highpass, 1, t0 / dt, h, c);
        section(x, n, c);
    }
}

/* timedomain_filt(data, fc, order, ntr, ns, method) on data [ntr][ns]:
 * method 1 lowpass at fc (then a highpass at fhp when fhp != 0), method 2
 * highpass at fc.  Each trace is filtered in double behind a leading zero
 * sample (the reference passes ns+1 samples of a vector whose element 0 is
 * 0) and stored back as float after each of the two passes. */
void orc_time_filter(float *data, int ntr, int ns, float DT, float fc, int order, int method, float fhp)
{
    double *x = malloc(sizeof(double) * (ns + 1));
    const double t0 = 1.0 / (double)fc;
    double t0_hp = 0.0;
    if (fhp != 0.0f) t0_hp = 1.0 / (double)fhp;
    if (method == 2) t0_hp = 1.0 / (double)fc;
    for (int pass = 0; pass < 2; pass++) {
        const int highpass = pass == 1;
        if (!highpass && method != 1) continue;
        if (highpass && !(method == 2 || fhp != 0.0f)) continue;
        for (int t = 0; t < ntr; t++) {
            float *tr = data + (long)t * ns;
            x[0] = 0.0;
            for (int j = 0; j < ns; j++) x[j + 1] = (double)tr[j];
            butterworth(x, ns + 1, DT, highpass, highpass ? t0_hp : t0, order);
            for (int j = 0; j < ns; j++) tr[j] = (float)x[j + 1];
        }
    }
    free(x);
}
