/* post_oracle.h -- CPU restatement of the reference's gradient
 * post-processing (TEST INFRASTRUCTURE ONLY): the tapers of
 * src/taper_grad_shot.c and src/taper_grad.c and the median smoothing of
 * src/smooth.c.  Arrays are interior [NY][NX], x fastest; positions are
 * 1-based grid indices (the reference's iround(x/DH), iround(y/DH)). */
#ifndef POST_ORACLE_H
#define POST_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Circular taper around the sources isrc/jsrc[0..nshots-1] and zeroing of
 * the (2*FILTSIZE+1)^2 box around each: taper_grad_shot sws 1
 * (src/taper_grad_shot.c:57-196) with one shot, taper_grad sws 3
 * (src/taper_grad.c:252-398) with all shots.  g is tapered in place; coef
 * (may be NULL) receives the taper coefficients before the boxes. */
void orc_taper_sources(int NX, int NY, float DH, float SRTRADIUS, int SRTSHAPE, int FILTSIZE, int nshots,
                       const int *isrc, const int *jsrc, float *g, float *coef);
/* Vertical taper, taper_grad sws 1 (src/taper_grad.c:60-172): Gaussian
 * flanks between columns GRADT1..GRADT4, zero at the sources and at the
 * receivers recpos [ntr][2]. */
void orc_taper_vertical(int NX, int NY, int GRADT1, int GRADT2, int GRADT3, int GRADT4, int nshots,
                        const int *isrc, const int *jsrc, int ntr, const int *recpos, float *g, float *coef);
/* Horizontal taper, taper_grad sws 2 (src/taper_grad.c:179-246). */
void orc_taper_horizontal(int NX, int NY, int GRADT1, int GRADT2, float *g, float *coef);
/* "Median" filter of smooth() (src/smooth.c:24-160, median2d
 * src/median2D.c:3): window filtsize x filtsize (an even size grows by one
 * away from zero), edges extended by replication.  As the reference
 * computes it, the value is element k/2+1 (0-based) of the sorted set
 * {0, window without its (+hfs,+hfs) corner}, k = window size (identity for
 * k = 1).  Returns the window size used (0: nothing done, out = in). */
int orc_median_smooth(int NX, int NY, int filtsize, const float *in, float *out);
/* spat_filt() (src/spat_filt.c:27): the boxcar mean over columns i-S..i+S
 * its integer (-1/2) makes of the Gaussian, for columns S+1..NX-S and rows
 * J1..NY-S (S = SPAT_FILT_SIZE, J1 = SPAT_FILT_1); other cells unchanged. */
void orc_spat_filt(int NX, int NY, int S, int J1, const float *in, float *out);

#ifdef __cplusplus
}
#endif
#endif
