/* opt_oracle.h -- CPU restatement of the reference's optimizer routines
 * (TEST INFRASTRUCTURE ONLY).  Arrays are interior [NY][NX], x fastest. */
#ifndef OPT_ORACLE_H
#define OPT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters of calc_mat_change_test (src/calc_mat_change_test.c:26):
 * the integers, then the floats, in this order. */
typedef struct orc_update_params {
    int iter, INV_VP_ITER, INV_VS_ITER, INV_RHO_ITER, S, GRAD_METHOD, LBFGS_STEP_LENGTH, LBFGS_iter_start;
    float eps_scale, VPUPPERLIM, VPLOWERLIM, VSUPPERLIM, VSLOWERLIM, RHOUPPERLIM, RHOLOWERLIM, VP_VS_RATIO;
    float S_VP, S_VS, S_RHO;
} orc_update_params;

/* PCG for one parameter (src/PCG.c:27): gradp = C*grad; the conjugate
 * direction with the Polak-Ribiere beta (returned) when the iteration
 * allows it.  p_last / c_last: the previous gradp / direction. */
float orc_pcg(int NX, int NY, const float *grad, float C, const float *p_last, const float *c_last, int iter,
              int pcg_iter_start, float *gradp, float *dir);
/* calc_opt_step (src/calc_opt_step.c:29): L2t, epst = the reference's
 * L2t[1..3], epst1[1..3]. */
float orc_opt_step(const float *L2t3, const float *epst3);
/* calc_mat_change_test, PARAMETERIZATION 1: new models and the three
 * step lengths EPSILON, EPSILON_u, EPSILON_rho. */
void orc_mat_change(int NX, int NY, const orc_update_params *u, const float *vp, const float *vs, const float *rho,
                    const float *g_vp, const float *g_vs, const float *g_rho, const float *vp0, const float *vs0,
                    const float *rho0, float *vp1, float *vs1, float *rho1, float *eps);
/* lbfgs_core (src/LBFGS.c:226): s, y [N_LBFGS][n]; rho, alpha [N_LBFGS];
 * q [n] (modified), r [n] (result). */
void orc_lbfgs_core(int iteration, int N_LBFGS, int n, const float *s, const float *y, float *rho, float *q,
                    float *alpha, float *r);

#ifdef __cplusplus
}
#endif
#endif
