/*
 * ifos2d_stf.h -- source-time-function inversion of the multi-scale
 * inversion (INV_STF, src/stf.c:24), host C.
 *
 * The reference hands observed and synthetic vertical-component traces of a
 * shot to the stfinv library (contrib/stfinv, engine string PARA such as
 * "fdlsq:tshift=0.0"), which returns a correction filter; stf() convolves
 * the source wavelet with it (conv_FD, src/conv_FD.c:25, FFTW), optionally
 * tapers the result (taper, src/taper.c:24) and writes it as the wavelet of
 * the later forward runs (SOURCE_SHAPE 7).  FFTW is not part of this build:
 * the transforms here are the library's own (radix-2, and Bluestein's
 * algorithm for other lengths), in double precision.
 *
 * Errors: 0 on success, negative on failure, message in ifos_last_error().
 */
#ifndef IFOS2D_STF_H
#define IFOS2D_STF_H

#ifdef __cplusplus
extern "C" {
#endif

/* Options of the Fourier-domain least-squares engine "fdlsq"
 * (contrib/stfinv/stfinvfdleastsquares.cc, stfinvfourier.cc,
 * stfinvbase.cc), as parsed from "fdlsq:key=value:key=value". */
typedef struct ifos_stf_params {
    double waterlevel;   /* waterlevel=, relative to the weighted synthetic energy (default 1e-3) */
    double fpad;         /* fpad=, padding factor >= 1 (default 1.5)                          */
    int fpow2;           /* fpow2=true: next power of two                                     */
    int fdiv;            /* fdiv=: next multiple of this above the padded length (0: unset)   */
    int use_tshift;      /* tshift= given: delay of the correction filter in seconds          */
    double tshift;
    int use_exp;         /* exp= given: trace weights offset^exp                              */
    double exp;
    int use_irtap;       /* irtap=t1,t2,t3,t4 given: cosine taper of the filter's impulse     */
    double irtap[4];     /* response, periodic in the padded length (seconds)                 */
} ifos_stf_params;

/* Parse an engine string as stfinv does (the engine ID before the first
 * ':', then ':'-separated key=value pairs).  Only "fdlsq" (and its older
 * name "fbd") are provided; other engine IDs are an error. */
int ifos_stf_parse(const char *para, ifos_stf_params *out);

/* replaces initstfinvengine() + runstfinvengine() for the fdlsq engine
 * (src/stf.c:216-219).  obs, syn: [ntr][ns] recorded and synthetic traces;
 * offset[ntr]: source-receiver offsets (any unit; only used with exp);
 * stf[ns] receives the correction filter's impulse response (scaled by
 * 1/(n dt) as the library returns it), conv[ntr][ns] (may be NULL) the
 * synthetics convolved with it. */
int ifos_stf_fdlsq(const ifos_stf_params *p, int ntr, int ns, float dt, const float *obs, const float *syn,
                   const double *offset, float *stf, float *conv);

/* replaces conv_FD() src/conv_FD.c:25: out[0..ns-1] = the first ns samples
 * of the linear convolution of a and b (0-based arrays of ns samples). */
int ifos_conv_fd(const float *a, const float *b, float *out, int ns);

/* replaces taper() src/taper.c:24 (TAPER_STF): the Cerjan-type window of
 * the reference applied in place to one trace of ns samples (0-based). */
void ifos_stf_taper(float *trace, int ns, float fc, float DT);

/* Complex discrete Fourier transform of length n >= 1 in place,
 * X[k] = sum_t x[t] exp(sign * 2 pi i t k / n), sign = -1 or +1, unscaled. */
int ifos_dft(double *re, double *im, int n, int sign);

#ifdef __cplusplus
}
#endif
#endif
