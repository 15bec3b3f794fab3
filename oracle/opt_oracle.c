/*
 * opt_oracle.c -- CPU restatement of the reference's optimizer routines for
 * the P-SV gradient path (TEST INFRASTRUCTURE ONLY: used by tests/ to check
 * the gfx950 optimizer kernels).
 *
 *   orc_pcg          PCG() for one parameter, src/PCG.c:27-292, without
 *                    tapers, filters and the disk files of the previous
 *                    iteration (they are passed in instead)
 *   orc_opt_step     calc_opt_step() src/calc_opt_step.c:29-127 with
 *                    solvelin() src/solvelin.c:30-83 (Gauss, n = 3)
 *   orc_mat_change   calc_mat_change_test() src/calc_mat_change_test.c:26,
 *                    PARAMETERIZATION 1, elastic (not ACOUSTIC), the model
 *                    update of one test or final step
 *   orc_lbfgs_core   lbfgs_core() src/LBFGS.c:226-322
 *
 * All arithmetic is FP32 in the reference's order: sums run over x in the
 * outer loop and y in the inner loop where the reference does so.  Pinned
 * against the reference's own compiled files (oracle/_ref): calc_opt_step,
 * lbfgs_core, calc_mat_change_test, and PCG itself, driven through the
 * state files it writes and reads over several iterations
 * (tests/test_oracle_pins.py), bit for bit; and by closed forms
 * (tests/test_optimizer.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "opt_oracle.h"

float orc_pcg(int NX, int NY, const float *grad, float C, const float *p_last, const float *c_last, int iter,
              int pcg_iter_start, float *gradp, float *dir)
{
    int use1 = 1, use2 = 1;
    if ((iter - pcg_iter_start) < 2) {
        use2 = 0;
        if ((iter - pcg_iter_start) < 1) use1 = 0;
    }
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) {
            const size_t e = (size_t)j * NX + i;
            dir[e] = C * grad[e];
            gradp[e] = dir[e];
        }
    float beta = 0.0f;
    if ((iter > 1) && use1) {
        float betaz = 0.0f, betan = 0.0f;
        for (int i = 0; i < NX; i++)
            for (int j = 0; j < NY; j++) {
                const size_t e = (size_t)j * NX + i;
                betaz += (gradp[e]) * ((gradp[e]) - (p_last[e]));
                betan += (p_last[e]) * (p_last[e]);
            }
        beta = 0.0f;
        if (betan != 0.0f) beta = betaz / betan;
        if (beta < 0.0) beta = 0.0;
        for (int i = 0; i < NX; i++)
            for (int j = 0; j < NY; j++) {
                const size_t e = (size_t)j * NX + i;
                if (iter == 2) dir[e] = gradp[e] + p_last[e] * beta;
                if ((iter > 2) && use2) dir[e] = gradp[e] + c_last[e] * beta;
            }
    }
    return beta;
}

/* Gauss elimination without pivoting, FP32 (src/solvelin.c:51-71) */
static void solve3(float A[4][4], float b[4], float x[4])
{
    const int rows = 3, columns = 3;
    float a, c;
    for (int k = 1; k <= rows - 1; k++)
        for (int n = k; n <= rows - 1; n++) {
            a = A[n + 1][k] / A[k][k];
            for (int m = 1; m <= columns; m++) A[n + 1][m] = A[n + 1][m] - a * A[k][m];
            b[n + 1] = b[n + 1] - a * b[k];
        }
    for (int k = rows; k >= 1; k--) {
        c = b[k];
        for (int m = columns; m >= k + 1; m--) c = c - A[k][m] * x[m];
        x[k] = c / A[k][k];
    }
}

float orc_opt_step(const float *L2t3, const float *epst3)
{
    float A[4][4], b[4], x[4] = {0, 0, 0, 0};
    float L2t[4], epst[4];
    for (int i = 1; i <= 3; i++) {
        L2t[i] = L2t3[i - 1];
        epst[i] = epst3[i - 1];
    }
    for (int i = 1; i <= 3; i++) {
        A[i][3] = (epst[i] * epst[i]);
        A[i][2] = (epst[i]);
        A[i][1] = (1.0);
        b[i] = (L2t[i]);
    }
    solve3(A, b, x);
    float opteps = -x[2] / (2.0 * x[3]);
    if (((2.0 * x[3]) < 0.0) && (L2t[2] > L2t[3])) opteps = epst[3];
    if (opteps > epst[3]) opteps = epst[3];
    if (opteps < 0.0) opteps = epst[1];
    return opteps;
}

static void grid_max(int NX, int NY, const float *m, const float *g, float *mmax, float *gmax)
{
    float a = 0.0f, b = 0.0f;
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            const size_t e = (size_t)(j - 1) * NX + (i - 1);
            if (m[e] > a) a = m[e];
            if ((i * j == 1) || (b == 0.0)) {
                b = fabs(g[e]);
            } else {
                if (fabs(g[e]) > b) b = fabs(g[e]);
            }
        }
    *mmax = a;
    *gmax = b;
}

void orc_mat_change(int NX, int NY, const orc_update_params *u, const float *vp, const float *vs, const float *rho,
                    const float *g_vp, const float *g_vs, const float *g_rho, const float *vp0, const float *vs0,
                    const float *rho0, float *vp1, float *vs1, float *rho1, float *eps)
{
    float pimax, gradmax, umax, gradmax_u, rhomax, gradmax_rho;
    grid_max(NX, NY, vp, g_vp, &pimax, &gradmax);
    grid_max(NX, NY, vs, g_vs, &umax, &gradmax_u);
    grid_max(NX, NY, rho, g_rho, &rhomax, &gradmax_rho);
    const int nonorm = u->GRAD_METHOD == 2 && (u->iter > u->LBFGS_iter_start) && u->LBFGS_STEP_LENGTH;
    float EPSILON = u->eps_scale * (pimax / gradmax);
    if (u->iter < u->INV_VP_ITER) EPSILON = 0.0;
    if (nonorm) {
        EPSILON = u->eps_scale;
        if (u->iter < u->INV_VP_ITER) EPSILON = 0.0;
    }
    float EPSILON_u = u->eps_scale * (umax / gradmax_u);
    if (u->iter < u->INV_VS_ITER) EPSILON_u = 0.0;
    if (nonorm) {
        EPSILON_u = u->eps_scale;
        if (u->iter < u->INV_VS_ITER) EPSILON_u = 0.0;
    }
    float EPSILON_rho = u->eps_scale * (rhomax / gradmax_rho);
    if (u->iter < u->INV_RHO_ITER) EPSILON_rho = 0.0;
    if (nonorm) {
        EPSILON_rho = u->eps_scale;
        if (u->iter < u->INV_RHO_ITER) EPSILON_rho = 0.0;
    }
    for (int i = 0; i < NX; i++)
        for (int j = 0; j < NY; j++) {
            const size_t e = (size_t)j * NX + i;
            float pinp1 = vp[e] - EPSILON * g_vp[e];
            float unp1 = vs[e] - EPSILON_u * g_vs[e];
            float rhonp1 = rho[e] - EPSILON_rho * g_rho[e];
            if (pinp1 < u->VPLOWERLIM) pinp1 = u->VPLOWERLIM;
            if (pinp1 > u->VPUPPERLIM) pinp1 = u->VPUPPERLIM;
            if ((unp1 < u->VSLOWERLIM) && (unp1 > 1e-6)) unp1 = u->VSLOWERLIM;
            if (unp1 > u->VSUPPERLIM) unp1 = vs[e];
            if (u->VP_VS_RATIO > 1) {
                if ((pinp1 / unp1) < u->VP_VS_RATIO) pinp1 = unp1 * u->VP_VS_RATIO;
            }
            if (u->S == 1) {
                float pro_vs = ((vs0[e] - unp1) / vs0[e]) * 100.0;
                if (((fabs(pro_vs)) > u->S_VS) && (pro_vs > 0.0)) unp1 = vs0[e] - vs0[e] * u->S_VS / 100.0;
                if (((fabs(pro_vs)) > u->S_VS) && (pro_vs < 0.0)) unp1 = vs0[e] + vs0[e] * u->S_VS / 100.0;
                float pro_vp = ((vp0[e] - pinp1) / vp0[e]) * 100.0;
                if (((fabs(pro_vp)) > u->S_VP) && (pro_vp > 0.0)) pinp1 = vp0[e] - vp0[e] * u->S_VP / 100.0;
                if (((fabs(pro_vp)) > u->S_VP) && (pro_vp < 0.0)) pinp1 = vp0[e] + vp0[e] * u->S_VP / 100.0;
                float pro_rho = ((rho0[e] - rhonp1) / rho0[e]) * 100.0;
                if (((fabs(pro_rho)) > u->S_RHO) && (pro_rho > 0.0)) rhonp1 = rho0[e] - rho0[e] * u->S_RHO / 100.0;
                if (((fabs(pro_rho)) > u->S_RHO) && (pro_rho < 0.0)) rhonp1 = rho0[e] + rho0[e] * u->S_RHO / 100.0;
            }
            if ((rhonp1 < u->RHOLOWERLIM) && (u->INV_RHO_ITER < u->iter)) rhonp1 = rho[e];
            if ((rhonp1 > u->RHOUPPERLIM) && (u->INV_RHO_ITER < u->iter)) rhonp1 = rho[e];
            if (pinp1 < 0.0) pinp1 = vp[e];
            if (unp1 < 0.0) unp1 = vs[e];
            if (rhonp1 < 0.0) rhonp1 = rho[e];
            vp1[e] = pinp1;
            vs1[e] = unp1;
            rho1[e] = rhonp1;
        }
    eps[0] = EPSILON;
    eps[1] = EPSILON_u;
    eps[2] = EPSILON_rho;
}

void orc_lbfgs_core(int iteration, int N_LBFGS, int n, const float *s, const float *y, float *rho, float *q,
                    float *alpha, float *r)
{
    int m = iteration - N_LBFGS;
    if (m < 1) m = 1;
    int w = (iteration - 1) % N_LBFGS;
    if (w == 0) w = N_LBFGS;
    float dum1 = 0.0, dum2 = 0.0;
    const float *sw = s + (size_t)(w - 1) * n, *yw = y + (size_t)(w - 1) * n;
    for (int l = 0; l < n; l++) {
        dum1 += yw[l] * sw[l];
        dum2 += yw[l] * yw[l];
    }
    rho[w - 1] = 1 / dum1;
    const float h0 = dum1 / dum2;
    for (int v = iteration - 1; v >= m; v--) {
        w = v % N_LBFGS;
        if (w == 0) w = N_LBFGS;
        const float *sv = s + (size_t)(w - 1) * n, *yv = y + (size_t)(w - 1) * n;
        float a = 0.0;
        for (int l = 0; l < n; l++) a += rho[w - 1] * sv[l] * q[l];
        alpha[w - 1] = a;
        for (int l = 0; l < n; l++) q[l] = q[l] - alpha[w - 1] * yv[l];
    }
    for (int l = 0; l < n; l++) r[l] = h0 * q[l];
    for (int v = m; v <= iteration - 1; v++) {
        w = v % N_LBFGS;
        if (w == 0) w = N_LBFGS;
        const float *sv = s + (size_t)(w - 1) * n, *yv = y + (size_t)(w - 1) * n;
        float beta = 0.0;
        for (int l = 0; l < n; l++) beta += rho[w - 1] * yv[l] * r[l];
        for (int l = 0; l < n; l++) r[l] = r[l] + sv[l] * (alpha[w - 1] - beta);
    }
}
