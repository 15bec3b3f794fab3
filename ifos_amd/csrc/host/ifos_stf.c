/* ifos_stf.c -- source-time-function inversion (include/ifos2d_stf.h).
 *
 * The fdlsq engine follows contrib/stfinv/stfinvfdleastsquares.cc:113-161
 * and contrib/stfinv/stfinvfourier.cc (padding :222-248, time shift
 * :540-553, impulse-response taper :577-660, output scaling :456-474);
 * the weights follow contrib/stfinv/stfinvbase.cc:103-117, 359-364.  The
 * Fourier transforms are this file's own: an iterative radix-2 transform,
 * and Bluestein's chirp-z algorithm (a length-n DFT as a convolution of
 * power-of-two length) for every other n, all in double precision.  The
 * scalings are those of contrib/fourier/fftwaffar.cc:263-283: transforms
 * are unscaled, the impulse response is multiplied by 1/(n dt).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ifos2d_stf.h"

extern void ifos_set_error(const char *fmt, ...);

#define PI_D 3.141592653589793

/* ---- transforms ------------------------------------------------------- */

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* iterative radix-2 DFT of length n (a power of two), in place */
static void fft_pow2(double *re, double *im, int n, int sign)
{
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = sign * 2.0 * PI_D / len;
        const int half = len / 2;
        for (int k = 0; k < half; k++) {
            /* twiddle from the angle directly, not by repeated products */
            const double wr = cos(ang * k), wi = sin(ang * k);
            for (int i = k; i < n; i += len) {
                const int j = i + half;
                const double xr = re[j] * wr - im[j] * wi;
                const double xi = re[j] * wi + im[j] * wr;
                re[j] = re[i] - xr;
                im[j] = im[i] - xi;
                re[i] += xr;
                im[i] += xi;
            }
        }
    }
}

int ifos_dft(double *re, double *im, int n, int sign)
{
    if (n < 1 || (sign != 1 && sign != -1)) return ifos_set_error("ifos_dft: n must be >= 1, sign +1 or -1"), -2;
    if (is_pow2(n)) {
        fft_pow2(re, im, n, sign);
        return 0;
    }
    /* Bluestein: t k = (t^2 + k^2 - (k - t)^2) / 2, so
     * X_k = c_k sum_t (x_t c_t) conj(c_{k-t}) with c_m = exp(sign i pi m^2 / n) */
    int M = 1;
    while (M < 2 * n - 1) M <<= 1;
    double *cr = malloc(sizeof(double) * n), *ci = malloc(sizeof(double) * n);
    double *ar = calloc(M, sizeof(double)), *ai = calloc(M, sizeof(double));
    double *br = calloc(M, sizeof(double)), *bi = calloc(M, sizeof(double));
    if (!cr || !ci || !ar || !ai || !br || !bi) {
        free(cr); free(ci); free(ar); free(ai); free(br); free(bi);
        return ifos_set_error("ifos_dft: out of memory"), -3;
    }
    for (int m = 0; m < n; m++) {
        const long long q = ((long long)m * m) % (2LL * n);  /* exp is periodic in m^2 with period 2n */
        const double ang = sign * PI_D * (double)q / n;
        cr[m] = cos(ang);
        ci[m] = sin(ang);
    }
    for (int t = 0; t < n; t++) {
        ar[t] = re[t] * cr[t] - im[t] * ci[t];
        ai[t] = re[t] * ci[t] + im[t] * cr[t];
    }
    br[0] = cr[0];
    bi[0] = -ci[0];
    for (int m = 1; m < n; m++) {
        br[m] = br[M - m] = cr[m];
        bi[m] = bi[M - m] = -ci[m];
    }
    fft_pow2(ar, ai, M, -1);
    fft_pow2(br, bi, M, -1);
    for (int k = 0; k < M; k++) {
        const double xr = ar[k] * br[k] - ai[k] * bi[k];
        const double xi = ar[k] * bi[k] + ai[k] * br[k];
        ar[k] = xr;
        ai[k] = xi;
    }
    fft_pow2(ar, ai, M, 1);
    for (int k = 0; k < n; k++) {
        const double yr = ar[k] / M, yi = ai[k] / M;
        re[k] = yr * cr[k] - yi * ci[k];
        im[k] = yr * ci[k] + yi * cr[k];
    }
    free(cr); free(ci); free(ar); free(ai); free(br); free(bi);
    return 0;
}

/* ---- parameters ------------------------------------------------------- */

int ifos_stf_parse(const char *para, ifos_stf_params *p)
{
    memset(p, 0, sizeof *p);
    p->waterlevel = 1e-3;
    p->fpad = 1.5;
    if (!para) return ifos_set_error("STF: empty engine string"), -2;
    char buf[1024];
    strncpy(buf, para, sizeof buf - 1);
    buf[sizeof buf - 1] = 0;
    char *rest = buf, *id = rest, *colon = strchr(rest, ':');
    if (colon) {
        *colon = 0;
        rest = colon + 1;
    } else {
        rest = rest + strlen(rest);
    }
    if (strcmp(id, "fdlsq") != 0 && strcmp(id, "fbd") != 0)
        return ifos_set_error("STF: engine \"%s\" is not provided (only fdlsq)", id), -2;
    while (*rest) {
        char *item = rest, *next = strchr(rest, ':');
        if (next) {
            *next = 0;
            rest = next + 1;
        } else {
            rest = rest + strlen(rest);
        }
        char *val = strchr(item, '=');
        if (val) *val++ = 0;
        else val = item + strlen(item);
        if (!strcmp(item, "waterlevel")) p->waterlevel = atof(val);
        else if (!strcmp(item, "fpad")) p->fpad = atof(val);
        else if (!strcmp(item, "fpow2")) p->fpow2 = !strcmp(val, "true");
        else if (!strcmp(item, "fdiv")) p->fdiv = *val ? atoi(val) : 100;
        else if (!strcmp(item, "tshift")) {
            p->use_tshift = 1;
            p->tshift = atof(val);
        } else if (!strcmp(item, "exp")) {
            p->use_exp = 1;
            p->exp = atof(val);
        } else if (!strcmp(item, "irtap")) {
            p->use_irtap = 1;
            char *s = val;
            for (int k = 0; k < 4; k++) {
                p->irtap[k] = strtod(s, &s);
                while (*s == ',' || *s == ';' || *s == ' ') s++;
            }
        } else if (!strcmp(item, "DEBUG") || !strcmp(item, "verbose") || !*item) {
            /* diagnostics of the library, no effect on the result */
        } else {
            return ifos_set_error("STF: unknown option \"%s\"", item), -2;
        }
    }
    if (!(p->waterlevel > 0)) return ifos_set_error("STF: waterlevel must be larger than 0"), -2;
    if (!(p->fpad >= 1)) return ifos_set_error("STF: fpad must be at least 1"), -2;
    if (p->fdiv < 0) return ifos_set_error("STF: fdiv must be larger than 0"), -2;
    if (p->use_irtap && !(p->irtap[0] <= p->irtap[1] && p->irtap[1] <= p->irtap[2] && p->irtap[2] <= p->irtap[3]))
        return ifos_set_error("STF: irtap times not in increasing order"), -2;
    return 0;
}

/* padded transform length (contrib/stfinv/stfinvfourier.cc:222-248) */
static int fft_length(const ifos_stf_params *p, int ns)
{
    unsigned int n = (unsigned int)(p->fpad * ns);
    if (n < (unsigned int)ns) n = (unsigned int)ns;
    if (p->fpow2) {
        unsigned int m = 2;
        while (m < n) m *= 2;
        n = m;
    } else if (p->fdiv > 0) {
        const unsigned int rest = n % (unsigned int)p->fdiv;
        n -= rest;
        n += (unsigned int)p->fdiv;
    }
    return (int)n;
}

/* real series of n values (hermitian spectrum spec[0..n/2]) from the
 * coefficients, unscaled, as a complex-to-real transform: the imaginary
 * parts of the zero and (n even) Nyquist coefficients are not used */
static int c2r(const double *sr, const double *si, int n, double *out, double *wr, double *wi)
{
    for (int k = 0; k <= n / 2; k++) {
        wr[k] = sr[k];
        wi[k] = si[k];
    }
    wi[0] = 0.0;
    if (n % 2 == 0) wi[n / 2] = 0.0;
    for (int k = n / 2 + 1; k < n; k++) {
        wr[k] = sr[n - k];
        wi[k] = -si[n - k];
    }
    if (ifos_dft(wr, wi, n, 1)) return -1;
    for (int t = 0; t < n; t++) out[t] = wr[t];
    return 0;
}

int ifos_stf_fdlsq(const ifos_stf_params *p, int ntr, int ns, float dt, const float *obs, const float *syn,
                   const double *offset, float *stf, float *conv)
{
    if (ntr < 1 || ns < 1 || !(dt > 0)) return ifos_set_error("STF: need ntr >= 1, ns >= 1, dt > 0"), -2;
    const int n = fft_length(p, ns), nf = n / 2 + 1;
    double *w = malloc(sizeof(double) * ntr);
    double *Dr = calloc((size_t)ntr * nf, sizeof(double)), *Di = calloc((size_t)ntr * nf, sizeof(double));
    double *Sr = calloc((size_t)ntr * nf, sizeof(double)), *Si = calloc((size_t)ntr * nf, sizeof(double));
    double *xr = malloc(sizeof(double) * n), *xi = malloc(sizeof(double) * n);
    double *Hr = calloc(nf, sizeof(double)), *Hi = calloc(nf, sizeof(double)), *h = malloc(sizeof(double) * n);
    int rc = 0;
    if (!w || !Dr || !Di || !Sr || !Si || !xr || !xi || !Hr || !Hi || !h) {
        ifos_set_error("STF: out of memory");
        rc = -3;
        goto done;
    }
    /* weights offset^exp (stfinvbase.cc:103-117) */
    for (int k = 0; k < ntr; k++) w[k] = p->use_exp ? pow(offset ? offset[k] : 0.0, p->exp) : 1.0;
    /* spectra of the zero-padded traces (forward transform, sign -1) and
     * the waterlevel: the given fraction of the weighted synthetic energy */
    double energy = 0.0;
    for (int k = 0; k < ntr; k++) {
        for (int pass = 0; pass < 2; pass++) {
            const float *src = (pass ? syn : obs) + (size_t)k * ns;
            for (int t = 0; t < n; t++) {
                xr[t] = t < ns ? (double)src[t] : 0.0;
                xi[t] = 0.0;
            }
            if (pass) {
                double e = 0.0;
                for (int t = 0; t < ns; t++) e += (double)src[t] * (double)src[t];
                energy += w[k] * w[k] * e;
            }
            if (ifos_dft(xr, xi, n, -1)) {
                rc = -1;
                goto done;
            }
            double *dr = (pass ? Sr : Dr) + (size_t)k * nf, *di = (pass ? Si : Di) + (size_t)k * nf;
            for (int f = 0; f < nf; f++) {
                dr[f] = xr[f];
                di[f] = xi[f];
            }
        }
    }
    const double waterlevel = energy * p->waterlevel;
    /* least squares per frequency (stfinvfdleastsquares.cc:141-155):
     * H = sum w^2 conj(S) D / (waterlevel + sum w^2 conj(S) S) */
    for (int f = 0; f < nf; f++) {
        double nr = 0.0, ni = 0.0, den = waterlevel;
        for (int k = 0; k < ntr; k++) {
            const size_t o = (size_t)k * nf + f;
            const double ww = w[k] * w[k];
            nr += ww * (Sr[o] * Dr[o] + Si[o] * Di[o]);
            ni += ww * (Sr[o] * Di[o] - Si[o] * Dr[o]);
            den += ww * (Sr[o] * Sr[o] + Si[o] * Si[o]);
        }
        Hr[f] = nr / den;
        Hi[f] = ni / den;
    }
    /* taper of the impulse response, periodic in n (stfinvfourier.cc:577-660) */
    if (p->use_irtap) {
        if (c2r(Hr, Hi, n, h, xr, xi)) {
            rc = -1;
            goto done;
        }
        const double T = (double)dt * n;
        if (!((p->irtap[3] - p->irtap[0]) < T)) {
            ifos_set_error("STF: taper width is larger than the padded time series");
            rc = -2;
            goto done;
        }
        const double tt3 = p->irtap[2], tt4 = p->irtap[3], tt1 = p->irtap[0] + T, tt2 = p->irtap[1] + T;
        const int l3 = (int)floor(tt3 / dt), l2 = (int)ceil(tt2 / dt);
        for (int t = 0; t < n; t++) h[t] *= 1.0 / (n * (double)dt);
        for (int l = l3; l <= l2; ++l) {
            const double t = l * (double)dt;
            int rl = l % n;
            if (rl < 0) rl += n;
            double factor = 1.0;
            if ((t >= tt3) && (t <= tt4) && (tt4 > tt3)) factor = 0.5 + 0.5 * cos(PI_D * (t - tt3) / (tt4 - tt3));
            else if ((t > tt4) && (t < tt1)) factor = 0.0;
            else if ((t >= tt1) && (t <= tt2) && (tt2 > tt1)) factor = 0.5 - 0.5 * cos(PI_D * (t - tt1) / (tt2 - tt1));
            h[rl] = h[rl] * factor;
        }
        for (int t = 0; t < n; t++) {
            xr[t] = h[t];
            xi[t] = 0.0;
        }
        if (ifos_dft(xr, xi, n, -1)) {
            rc = -1;
            goto done;
        }
        for (int f = 0; f < nf; f++) {
            Hr[f] = xr[f] * dt;
            Hi[f] = xi[f] * dt;
        }
    }
    /* the synthetics convolved with the filter, before any time shift
     * (stfinvfourier.cc:287-295, 489-507); scaled by dt/(n dt) = 1/n */
    if (conv) {
        double *cr = malloc(sizeof(double) * nf), *ci = malloc(sizeof(double) * nf);
        if (!cr || !ci) {
            free(cr);
            free(ci);
            ifos_set_error("STF: out of memory");
        rc = -3;
            goto done;
        }
        for (int k = 0; k < ntr && !rc; k++) {
            for (int f = 0; f < nf; f++) {
                const size_t o = (size_t)k * nf + f;
                cr[f] = Sr[o] * Hr[f] - Si[o] * Hi[f];
                ci[f] = Sr[o] * Hi[f] + Si[o] * Hr[f];
            }
            if (c2r(cr, ci, n, h, xr, xi)) rc = -1;
            else
                for (int t = 0; t < ns; t++) conv[(size_t)k * ns + t] = (float)(h[t] * (double)dt * (1.0 / (n * (double)dt)));
        }
        free(cr);
        free(ci);
        if (rc) goto done;
    }
    /* delay of the filter by tshift seconds: H(f) exp(-2 pi i f tshift),
     * f = j / (n dt) (stfinvfourier.cc:540-553) */
    if (p->use_tshift)
        for (int f = 0; f < nf; f++) {
            const double ang = -2.0 * PI_D * p->tshift * ((double)f / ((double)dt * n));
            const double c = cos(ang), s = sin(ang);
            const double r = Hr[f] * c - Hi[f] * s, i = Hr[f] * s + Hi[f] * c;
            Hr[f] = r;
            Hi[f] = i;
        }
    if (c2r(Hr, Hi, n, h, xr, xi)) {
        rc = -1;
        goto done;
    }
    for (int t = 0; t < ns; t++) stf[t] = (float)(h[t] * (1.0 / (n * (double)dt)));
done:
    free(w); free(Dr); free(Di); free(Sr); free(Si); free(xr); free(xi); free(Hr); free(Hi); free(h);
    return rc;
}

/* ---- convolution and taper of stf() ----------------------------------- */

int ifos_conv_fd(const float *a, const float *b, float *out, int ns)
{
    if (ns < 1) return ifos_set_error("conv_fd: ns must be >= 1"), -2;
    /* the padded length of src/conv_FD.c:41, 2^(ceil(log2 ns) + 2) >= 2 ns:
     * the cyclic convolution of the padded series is the linear one */
    const int npad = (int)(pow(2.0, ceil(log((double)ns) / log(2.0)) + 2.0));
    double *ar = calloc(npad, sizeof(double)), *ai = calloc(npad, sizeof(double));
    double *br = calloc(npad, sizeof(double)), *bi = calloc(npad, sizeof(double));
    if (!ar || !ai || !br || !bi) {
        free(ar); free(ai); free(br); free(bi);
        return ifos_set_error("conv_fd: out of memory"), -3;
    }
    for (int j = 0; j < ns; j++) {
        ar[j] = a[j];
        br[j] = b[j];
    }
    int rc = ifos_dft(ar, ai, npad, 1) || ifos_dft(br, bi, npad, 1);
    if (!rc) {
        for (int i = 0; i < npad; i++) {
            const double r = (ar[i] * br[i]) - (ai[i] * bi[i]);
            const double im = (ar[i] * bi[i]) + (ai[i] * br[i]);
            ar[i] = r;
            ai[i] = im;
        }
        rc = ifos_dft(ar, ai, npad, -1);
    }
    if (!rc)
        for (int j = 0; j < ns; j++) out[j] = (float)((1.0 / (double)npad) * ar[j]);
    free(ar); free(ai); free(br); free(bi);
    return rc ? -1 : 0;
}

void ifos_stf_taper(float *trace, int ns, float fc, float DT)
{
    /* src/taper.c:40-80: window of taperlength = ceil(2/fc/DT) samples
     * rising as exp(-a^2 (taperlength - i)^2), a from a 99.9 % damping,
     * placed backwards ending at sample 2*taperlength, then continued
     * from there to the end (1-based indices as there) */
    const int taperlength = (int)(ceil(2.0 / fc / DT));
    const int taperduration = 2 * taperlength;
    float *window = malloc(sizeof(float) * (ns + 1)), *amp1 = malloc(sizeof(float) * (ns + 1));
    if (!window || !amp1) {
        free(window);
        free(amp1);
        return;
    }
    const float damping = 99.9;
    const float amp = 1.0 - damping / 100.0;
    const float a = sqrt(-log(amp) / ((taperlength - 1) * (taperlength - 1)));
    for (int i = 1; i <= ns; i++) {
        window[i] = 1.0;
        amp1[i] = 0.0;
    }
    for (int i = 1; i <= taperlength && i <= ns; i++) amp1[i] = exp(-(a * a * (taperlength - i) * (taperlength - i)));
    int h = 1;
    for (int i = taperduration; i >= (taperduration - taperlength + 3); i--) {
        if (i >= 1 && i <= ns && h <= ns) window[i] = amp1[h];
        h++;
    }
    for (int i = taperduration; i <= ns; i++)
        if (i >= 1) window[i] = amp1[i];
    for (int j = 1; j <= ns; j++) trace[j - 1] *= window[j];
    free(window);
    free(amp1);
}
