/*
 * ifos_filter.c -- coefficients of the time-domain Butterworth filter of
 * the multi-scale inversion (TIME_FILT, src/timedomain_filt.c:27), as the
 * reference's vendored libcseife designs it: an odd order starts with a
 * first-order section, then order/2 second-order sections with damping
 * sin(pi/2 (2j-1)/order) (contrib/libcseife/cseife.c:164-195), each section
 * from the recursive-filter design of cseife_rfk.c.  Evaluated on the host
 * in double with glibc atan/sin and the reference's operand order, so the
 * device filter reproduces the reference's trace values bit for bit.
 */
#include <math.h>

#include "ifos_setup.h"

/* {f0, f1, f2, g1, g2} of one section; period t0 in samples */
static void design(int highpass, int second, double t0, double h, double *c)
{
    const double eps = 8.0 * atan(1.0) / t0;
    if (!second) {
        const double den = 2.0 + eps;
        c[3] = (2.0 - eps) / den;
        c[2] = c[4] = 0.0;
        c[0] = highpass ? 2.0 / den : eps / den;
        c[1] = highpass ? -(c[0]) : c[0];
        return;
    }
    const double eq = eps * eps;
    const double lo = 1.0 - eps * h + eq / 4.0;
    const double mid = -2.0 + eq / 2.0;
    const double hi = 1.0 + eps * h + eq / 4.0;
    c[3] = -mid / hi;
    c[4] = -lo / hi;
    if (highpass) {
        c[0] = 1.0 / hi;
        c[1] = -(c[0]) - c[0];
    } else {
        c[0] = eq / 4.0 / hi;
        c[1] = c[0] + c[0];
    }
    c[2] = c[0];
}

int ifos_butterworth(int highpass, double dt, double fc, int order, double *coef, int max_sections)
{
    if (order < 1 || !(fc > 0.0)) return 0;
    const int half = order / 2, first = order > 2 * half;
    if (half + first > max_sections) return -1;
    const double t0 = (1.0 / fc) / dt;
    int n = 0;
    if (first) design(highpass, 0, t0, 0.0, coef + 5 * n++);
    const double pih = 2.0 * atan(1.0);
    for (int j = 1; j <= half; j++) design(highpass, 1, t0, sin(pih * (2 * j - 1) / order), coef + 5 * n++);
    return n;
}
