/* ifos_setup.h -- host-side material preparation shared by the engine and
 * the drop-in layer. */
#ifndef IFOS_SETUP_H
#define IFOS_SETUP_H

#include "ifos2d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IFOS_MAXL 8

/* Interior [NY][NX] arrays (x fastest); d, e, dip are L planes. */
typedef struct ifos_materials {
    float *rip, *rjp, *uipjp;   /* av_rho, av_mue                          */
    float *f, *g;               /* elastic: rho*vs^2, rho*(vp^2-2vs^2);
                                   viscoelastic: prepare_update_s f, g      */
    float *fipjp;               /* viscoelastic                            */
    float *d, *e, *dip;         /* viscoelastic, [L][NY][NX]               */
    float *muss, *lamss;        /* moduli of the correlation / chain rule  */
    float *mu_r, *pi_r, *muipjp_r, *tausipjp;  /* viscoelastic: relaxed moduli
                                   and averaged tau of prepare_update_s      */
    float eta[IFOS_MAXL + 1], bip[IFOS_MAXL + 1], bjm[IFOS_MAXL + 1];
    float cip[IFOS_MAXL + 1], cjm[IFOS_MAXL + 1];
    float sumu, sumpi;
} ifos_materials;

int ifos_setup_materials(const ifos_config *c, const float *vp, const float *vs, const float *rho,
                         const float *taus, const float *taup, ifos_materials *m);
void ifos_free_materials(ifos_materials *m);

/* Butterworth low- or highpass of the given order and corner fc (Hz) for
 * sample interval dt as cascaded sections {f0, f1, f2, g1, g2} (coef
 * [max_sections][5]); returns the section count, 0 for no filter, -1 if
 * max_sections is too small. */
int ifos_butterworth(int highpass, double dt, double fc, int order, double *coef, int max_sections);

#ifdef __cplusplus
}
#endif
#endif
