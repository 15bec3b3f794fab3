/*
 * ref_harness.c -- drives the reference's own hot-path C files, compiled
 * from /root/reference/src by oracle/Makefile into oracle/_ref/.
 *
 * TEST INFRASTRUCTURE ONLY.  Used to pin the oracle restatement
 * (ifos_oracle.c) and to write the golden fixtures under tests/golden/.
 *
 * The reference kernels (update_v_PML, update_s_elastic_PML,
 * update_s_visc_PML, psource, surface_elastic_PML, surface_PML, seismo_ssg,
 * holbergcoeff, PML_pro, matcopy*, av_*, prepare_update_s, wavelet,
 * zero_fdveps*, exchange_v, exchange_s) are called unmodified, with the globals they read defined by
 * the reference's own globvar.h.  The per-step call order restates
 * src/IFOS2D.c:1651-1803 (forward) and :2224-2419 (adjoint); the snapshot
 * store, zero-lag correlation and chain rule are inline code in the
 * reference's main() and are restated here the same way as in the oracle.
 * Residual computation (calc_res) needs FFTW at link time in the reference,
 * which is absent, so the adjoint sources are passed in by the caller.
 *
 * Single MPI rank (MPICH from the image): INDEX[] all point to rank 0 and
 * POS = (0,0), as initproc() sets them for NPROCX = NPROCY = 1.
 */
#include "fd.h"
#include "globvar.h"

static int harness_mpi_ready = 0;
static char *harness_bsend_buf = NULL;

static void harness_setup(const int *ip, const float *fp)
{
    if (!harness_mpi_ready) {
        int inited = 0;
        MPI_Initialized(&inited);
        if (!inited) MPI_Init(NULL, NULL);
        int sz = 1 << 24;
        harness_bsend_buf = (char *)malloc((size_t)sz);
        MPI_Buffer_attach(harness_bsend_buf, sz);
        harness_mpi_ready = 1;
    }
    /* ip / fp mirror orc_params field order (oracle/ifos_oracle.h) */
    NX = ip[0]; NY = ip[1]; FDORDER = ip[2]; MAXRELERROR = ip[3]; FW = ip[4];
    FREE_SURF = ip[5]; BOUNDARY = ip[6]; L = ip[7]; VELOCITY = ip[8];
    ADJOINT_TYPE = ip[9]; PARAMETERIZATION = ip[10]; NT = ip[12];
    SOURCE_SHAPE = ip[13]; SOURCE_TYPE = ip[14];
    DH = fp[0]; DT = fp[1]; VPPML = fp[2]; FPML = fp[3]; npower = fp[4];
    k_max_PML = fp[5]; F_REF = fp[6]; TAU = fp[7];
    NXG = NX; NYG = NY; IENDX = NX; IENDY = NY;
    NPROCX = NPROCY = 1; NPROC = 1; MYID = 0;
    POS[1] = POS[2] = 0;
    INDEX[1] = INDEX[2] = INDEX[3] = INDEX[4] = 0;
    WAVETYPE = 1; ACOUSTIC = 0; VERBOSE = 0; SEISMO = 1; NDT = 1;
    IDX = IDY = 1;
    FP = stderr;
}

float ref_hc[7];

void ref_holbergcoeff(const int *ip, const float *fp, float *hc_out)
{
    harness_setup(ip, fp);
    float *hc = holbergcoeff();
    for (int k = 0; k < 7; k++) hc_out[k] = hc[k];
    free_vector(hc, 0, 6);
}

typedef struct {
    float *d_x, *K_x, *al_x, *a_x, *b_x, *d_xh, *K_xh, *al_xh, *a_xh, *b_xh;
    float *d_y, *K_y, *al_y, *a_y, *b_y, *d_yh, *K_yh, *al_yh, *a_yh, *b_yh;
} pml_t;

static void pml_make(pml_t *P)
{
    int n = 2 * FW;
    float **v[20] = {&P->d_x, &P->K_x, &P->al_x, &P->a_x, &P->b_x, &P->d_xh, &P->K_xh, &P->al_xh, &P->a_xh, &P->b_xh,
                     &P->d_y, &P->K_y, &P->al_y, &P->a_y, &P->b_y, &P->d_yh, &P->K_yh, &P->al_yh, &P->a_yh, &P->b_yh};
    for (int k = 0; k < 20; k++) *v[k] = vector(1, n > 0 ? n : 1);
    if (FW > 0)
        PML_pro(P->d_x, P->K_x, P->al_x, P->a_x, P->b_x, P->d_xh, P->K_xh, P->al_xh, P->a_xh, P->b_xh,
                P->d_y, P->K_y, P->al_y, P->a_y, P->b_y, P->d_yh, P->K_yh, P->al_yh, P->a_yh, P->b_yh);
}

static void pml_free(pml_t *P)
{
    int n = 2 * FW;
    float *v[20] = {P->d_x, P->K_x, P->al_x, P->a_x, P->b_x, P->d_xh, P->K_xh, P->al_xh, P->a_xh, P->b_xh,
                    P->d_y, P->K_y, P->al_y, P->a_y, P->b_y, P->d_yh, P->K_yh, P->al_yh, P->a_yh, P->b_yh};
    for (int k = 0; k < 20; k++) free_vector(v[k], 1, n > 0 ? n : 1);
}

void ref_pml_pro(const int *ip, const float *fp, float *out12)
{
    harness_setup(ip, fp);
    pml_t P;
    pml_make(&P);
    int n = 2 * FW;
    float *src[12] = {P.K_x, P.a_x, P.b_x, P.K_xh, P.a_xh, P.b_xh, P.K_y, P.a_y, P.b_y, P.K_yh, P.a_yh, P.b_yh};
    for (int k = 0; k < 12; k++)
        for (int i = 1; i <= n; i++) out12[k * n + i - 1] = src[k][i];
    pml_free(&P);
}

/* wavelet(): srcpos rows as in the reference srcpos_loc[1..8][1..nsrc] */
void ref_wavelet(const int *ip, const float *fp, int nsrc, const float *srcpos, float *out)
{
    harness_setup(ip, fp);
    float **sp = matrix(1, 8, 1, nsrc);
    for (int l = 1; l <= nsrc; l++)
        for (int k = 1; k <= 8; k++) sp[k][l] = srcpos[(l - 1) * 8 + k - 1];
    float **sig = wavelet(sp, nsrc, 1, 0, 0);
    for (int l = 1; l <= nsrc; l++)
        for (int nt = 1; nt <= NT; nt++) out[(size_t)(l - 1) * NT + nt - 1] = sig[l][nt];
    free_matrix(sp, 1, 8, 1, nsrc);
}

/* ------------------------------------------------------------------------ */
typedef struct {
    int nd, nsnap;
    float *hc;
    pml_t P;
    float **prho, **ppi, **pu, **ptaus, **ptaup, *peta, *etaip, *etajm;
    float **prip, **prjp, **puipjp, **ptausipjp, **fipjp, **f, **g;
    float *bip, *bjm, *cip, *cjm;
    float ***dip, ***d, ***e, ***pr, ***pp, ***pq;
    float **pvx, **pvy, **psxx, **psyy, **psxy, **ux, **uy, **uxy, **uyx, **pvxp1, **pvyp1;
    float **pvz, **psxz, **psyz, **uyz, **uxz, **pvxm1, **pvym1, **uttx, **utty;
    float ***pt, ***po;
    float **psi_sxx_x, **psi_sxy_x, **psi_vxx, **psi_vyx, **psi_vxxs, **psi_sxz_x, **psi_vzx;
    float **psi_syy_y, **psi_sxy_y, **psi_vyy, **psi_vxy, **psi_syz_y, **psi_vzy;
    /* halo buffers of exchange_v / exchange_s (src/IFOS2D.c:649-652) */
    float **buf_lr, **buf_rl, **buf_tb, **buf_bt;
} run_t;

static void run_init(run_t *R, const float *vp, const float *vs, const float *rho, const float *taus,
                     const float *taup, const float *FLv)
{
    int nd = FDORDER / 2 + 1;
    R->nd = nd;
    R->hc = holbergcoeff();
    pml_make(&R->P);
#define GRID matrix(-nd + 1, NY + nd, -nd + 1, NX + nd)
#define TENS(n) f3tensor(-nd + 1, NY + nd, -nd + 1, NX + nd, 1, n)
    R->prho = GRID; R->ppi = GRID; R->pu = GRID;
    R->prip = GRID; R->prjp = GRID; R->puipjp = GRID;
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            size_t k = (size_t)(j - 1) * NX + (i - 1);
            R->prho[j][i] = rho[k]; R->ppi[j][i] = vp[k]; R->pu[j][i] = vs[k];
        }
    int nl = L > 0 ? L : 1;
    if (L) {
        FL = vector(1, L);
        for (int l = 1; l <= L; l++) FL[l] = FLv[l];
        R->ptaus = GRID; R->ptaup = GRID; R->ptausipjp = GRID;
        R->fipjp = GRID; R->f = GRID; R->g = GRID;
        R->peta = vector(1, L); R->etaip = vector(1, L); R->etajm = vector(1, L);
        R->bip = vector(1, L); R->bjm = vector(1, L); R->cip = vector(1, L); R->cjm = vector(1, L);
        R->dip = TENS(L); R->d = TENS(L); R->e = TENS(L);
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                size_t k = (size_t)(j - 1) * NX + (i - 1);
                R->ptaus[j][i] = taus[k]; R->ptaup[j][i] = taup[k];
            }
        /* readmod: eta = DT / tau_sigma (src/readmod.c:47-51) */
        for (int l = 1; l <= L; l++) {
            float pts = 1.0 / (2.0 * PI * FL[l]);
            R->peta[l] = DT / pts;
        }
        matcopy(R->prho, R->ppi, R->pu, R->ptaus, R->ptaup);
    } else {
        matcopy_elastic(R->prho, R->ppi, R->pu);
    }
    av_mue(R->pu, R->puipjp, R->prho);
    av_rho(R->prho, R->prip, R->prjp);
    if (L) {
        av_tau(R->ptaus, R->ptausipjp);
        prepare_update_s(R->etajm, R->etaip, R->peta, R->fipjp, R->pu, R->puipjp, R->ppi, R->prho, R->ptaus,
                         R->ptaup, R->ptausipjp, R->f, R->g, R->bip, R->bjm, R->cip, R->cjm, R->dip, R->d, R->e);
    }
    R->pvx = GRID; R->pvy = GRID; R->psxx = GRID; R->psyy = GRID; R->psxy = GRID;
    R->ux = GRID; R->uy = GRID; R->uxy = GRID; R->uyx = GRID; R->pvxp1 = GRID; R->pvyp1 = GRID;
    R->pvz = GRID; R->psxz = GRID; R->psyz = GRID; R->uyz = GRID; R->uxz = GRID;
    R->pvxm1 = GRID; R->pvym1 = GRID; R->uttx = GRID; R->utty = GRID;
    R->pr = TENS(nl); R->pp = TENS(nl); R->pq = TENS(nl); R->pt = TENS(nl); R->po = TENS(nl);
    int fw2 = FW > 0 ? 2 * FW : 1;
    R->psi_sxx_x = matrix(1, NY, 1, fw2); R->psi_sxy_x = matrix(1, NY, 1, fw2);
    R->psi_vxx = matrix(1, NY, 1, fw2); R->psi_vyx = matrix(1, NY, 1, fw2);
    R->psi_vxxs = matrix(1, NY, 1, fw2); R->psi_sxz_x = matrix(1, NY, 1, fw2); R->psi_vzx = matrix(1, NY, 1, fw2);
    R->psi_syy_y = matrix(1, fw2, 1, NX); R->psi_sxy_y = matrix(1, fw2, 1, NX);
    R->psi_vyy = matrix(1, fw2, 1, NX); R->psi_vxy = matrix(1, fw2, 1, NX);
    R->psi_syz_y = matrix(1, fw2, 1, NX); R->psi_vzy = matrix(1, fw2, 1, NX);
    const int fdo3 = 2 * nd; /* WAVETYPE 1 (src/IFOS2D.c:311-325) */
    R->buf_lr = matrix(1, NY, 1, fdo3); R->buf_rl = matrix(1, NY, 1, fdo3);
    R->buf_tb = matrix(1, NX, 1, fdo3); R->buf_bt = matrix(1, NX, 1, fdo3);
#undef GRID
#undef TENS
}

static void run_zero(run_t *R)
{
    int nd = R->nd;
    if (L)
        zero_fdveps_visc(-nd + 1, NY + nd, -nd + 1, NX + nd, R->pvx, R->pvy, R->pvz, R->psxx, R->psyy, R->psxy,
                         R->psxz, R->psyz, R->ux, R->uy, R->uxy, R->pvxp1, R->pvyp1, R->psi_sxx_x, R->psi_sxy_x,
                         R->psi_sxz_x, R->psi_vxx, R->psi_vyx, R->psi_vzx, R->psi_syy_y, R->psi_sxy_y,
                         R->psi_syz_y, R->psi_vyy, R->psi_vxy, R->psi_vzy, R->psi_vxxs, R->pr, R->pp, R->pq,
                         R->pt, R->po);
    else
        zero_fdveps(-nd + 1, NY + nd, -nd + 1, NX + nd, R->pvx, R->pvy, R->pvz, R->psxx, R->psyy, R->psxy,
                    R->psxz, R->psyz, R->ux, R->uy, R->uxy, R->pvxp1, R->pvyp1, R->psi_sxx_x, R->psi_sxy_x,
                    R->psi_sxz_x, R->psi_vxx, R->psi_vyx, R->psi_vzx, R->psi_syy_y, R->psi_sxy_y, R->psi_syz_y,
                    R->psi_vyy, R->psi_vxy, R->psi_vzy, R->psi_vxxs);
}

/* one time step, driver order of src/IFOS2D.c:1682-1800 */
static void run_step(run_t *R, int nt, int sw, float **srcpos_loc, float **sig, float **sig1, int nsrc)
{
    pml_t *P = &R->P;
    update_v_PML(1, NX, 1, NY, nt, R->pvx, R->pvxp1, R->pvxm1, R->pvy, R->pvyp1, R->pvym1, R->uttx, R->utty,
                 R->psxx, R->psyy, R->psxy, R->prip, R->prjp, srcpos_loc, sig, sig1, nsrc, NULL, R->hc, 0, sw,
                 P->K_x, P->a_x, P->b_x, P->K_xh, P->a_xh, P->b_xh, P->K_y, P->a_y, P->b_y, P->K_yh, P->a_yh,
                 P->b_yh, R->psi_sxx_x, R->psi_syy_y, R->psi_sxy_y, R->psi_sxy_x);
    /* on one rank the exchanges fill only the periodic halo columns of
     * BOUNDARY 1; with BOUNDARY 0 they send to the rank itself and unpack
     * nothing */
    exchange_v(R->pvx, R->pvy, R->pvz, R->buf_lr, R->buf_rl, R->buf_tb, R->buf_bt, NULL, NULL, 1);
    if (L)
        update_s_visc_PML(1, NX, 1, NY, R->pvx, R->pvy, R->ux, R->uy, R->uxy, R->uyx, R->psxx, R->psyy, R->psxy,
                          R->ppi, R->pu, R->puipjp, R->prho, R->hc, 0, R->pr, R->pp, R->pq, R->fipjp, R->f, R->g,
                          R->bip, R->bjm, R->cip, R->cjm, R->d, R->e, R->dip, P->K_x, P->a_x, P->b_x, P->K_xh,
                          P->a_xh, P->b_xh, P->K_y, P->a_y, P->b_y, P->K_yh, P->a_yh, P->b_yh, R->psi_vxx,
                          R->psi_vyy, R->psi_vxy, R->psi_vyx);
    else
        update_s_elastic_PML(1, NX, 1, NY, R->pvx, R->pvy, R->ux, R->uy, R->uxy, R->uyx, R->psxx, R->psyy,
                             R->psxy, R->ppi, R->pu, R->puipjp, NULL, R->prho, R->hc, 0, P->K_x, P->a_x, P->b_x,
                             P->K_xh, P->a_xh, P->b_xh, P->K_y, P->a_y, P->b_y, P->K_yh, P->a_yh, P->b_yh,
                             R->psi_vxx, R->psi_vyy, R->psi_vxy, R->psi_vyx);
    if (sw == 0 && SOURCE_TYPE == 1) psource(nt, R->psxx, R->psyy, NULL, srcpos_loc, sig, nsrc, 0);
    if (FREE_SURF) {
        if (L)
            surface_PML(1, R->pvx, R->pvy, R->psxx, R->psyy, R->psxy, R->psyz, R->pp, R->pq, R->ppi, R->pu,
                        R->prho, R->ptaup, R->ptaus, R->etajm, R->peta, R->hc, P->K_x, P->a_x, P->b_x, R->psi_vxxs,
                        R->ux, R->uy, R->uxy, R->uyz, R->psxz, R->uxz);
        else
            surface_elastic_PML(1, R->pvx, R->pvy, R->psxx, R->psyy, R->psxy, R->psyz, R->ppi, R->pu, R->prho,
                                R->hc, P->K_x, P->a_x, P->b_x, R->psi_vxxs, R->ux, R->uy, R->uxy, R->uyz, R->psxz,
                                R->uxz);
    }
    exchange_s(R->psxx, R->psyy, R->psxy, R->psxz, R->psyz, R->buf_lr, R->buf_rl, R->buf_tb, R->buf_bt, NULL, NULL,
               1);
}

static void copy_interior(float **a, float *out)
{
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) out[(size_t)(j - 1) * NX + (i - 1)] = a[j][i];
}

/* forward run: seismograms [ntr][NT], optional final wavefields, optional
 * snapshots (5 fields x nsnap planes of [NY][NX]) */
/* energy accumulators of the preconditioner (set by ref_gradient_we) */
static float **harness_Ws = NULL, **harness_Wr = NULL;

static void run_forward(run_t *R, int nsrc, const float *srcpos, int ntr, const int *recpos, float *seis_vx,
                        float *seis_vy, float **snap, int *help, int dtinv)
{
    float **sp = matrix(1, 8, 1, nsrc);
    for (int l = 1; l <= nsrc; l++)
        for (int k = 1; k <= 8; k++) sp[k][l] = srcpos[(l - 1) * 8 + k - 1];
    float **signals = wavelet(sp, nsrc, 1, 0, 0);
    int **rp = imatrix(1, 3, 1, ntr > 0 ? ntr : 1);
    for (int t = 1; t <= ntr; t++) { rp[1][t] = recpos[2 * (t - 1)]; rp[2][t] = recpos[2 * (t - 1) + 1]; rp[3][t] = t; }
    float **svx = matrix(1, ntr > 0 ? ntr : 1, 1, NT), **svy = matrix(1, ntr > 0 ? ntr : 1, 1, NT);
    run_zero(R);
    int hin = 1, hin1 = 1;
    for (int nt = 1; nt <= NT; nt++) {
        run_step(R, nt, 0, sp, signals, signals, nsrc);
        seismo_ssg(nt, ntr, rp, svx, svy, NULL, NULL, NULL, NULL, R->pvx, R->pvy, NULL, R->psxx, R->psyy, NULL,
                   R->ppi, R->pu, R->hc);
        if (nt == hin1) {
            if (snap) {
                size_t off = (size_t)(hin - 1) * NX * NY;
                copy_interior(R->pvxp1, snap[0] + off);
                copy_interior(R->pvyp1, snap[1] + off);
                copy_interior(VELOCITY ? R->ux : R->psxx, snap[2] + off);
                copy_interior(VELOCITY ? R->uy : R->psyy, snap[3] + off);
                copy_interior(VELOCITY ? R->uxy : R->psxy, snap[4] + off);
            }
            hin++;
            hin1 += dtinv;
            help[nt] = 1;
        }
        if (harness_Ws) eprecond(harness_Ws, R->pvx, R->pvy);  /* src/IFOS2D.c:1872-1874 */
    }
    for (int t = 1; t <= ntr; t++)
        for (int nt = 1; nt <= NT; nt++) {
            seis_vx[(size_t)(t - 1) * NT + nt - 1] = svx[t][nt];
            seis_vy[(size_t)(t - 1) * NT + nt - 1] = svy[t][nt];
        }
    free_matrix(sp, 1, 8, 1, nsrc);
}

int ref_forward(const int *ip, const float *fp, const float *FLv, const float *vp, const float *vs, const float *rho,
                const float *taus, const float *taup, int nsrc, const float *srcpos, int ntr, const int *recpos,
                float *seis_vx, float *seis_vy, float *final5)
{
    harness_setup(ip, fp);
    run_t R;
    run_init(&R, vp, vs, rho, taus, taup, FLv);
    int *help = (int *)calloc((size_t)NT + 2, sizeof(int));
    run_forward(&R, nsrc, srcpos, ntr, recpos, seis_vx, seis_vy, NULL, help, ip[11] > 0 ? ip[11] : 1);
    if (final5) {
        size_t pl = (size_t)NX * NY;
        copy_interior(R.pvx, final5);
        copy_interior(R.pvy, final5 + pl);
        copy_interior(R.psxx, final5 + 2 * pl);
        copy_interior(R.psyy, final5 + 3 * pl);
        copy_interior(R.psxy, final5 + 4 * pl);
    }
    free(help);
    return 0;
}

/* forward with snapshots, then the adjoint run driven by the given
 * time-reversed residuals diff_vx/diff_vy [ntr][NT], correlation and chain
 * rule.  Outputs per-shot gradients [NY][NX]. */
int ref_gradient_we(const int *ip, const float *fp, const float *FLv, const float *vp, const float *vs,
                    const float *rho, const float *taus, const float *taup, int nsrc, const float *srcpos, int ntr,
                    const int *recpos, const float *diff_vx, const float *diff_vy, float *seis_vx, float *seis_vy,
                    float *grad_vp, float *grad_vs, float *grad_rho, int ep_mode, float epsilon_we, int per_shot,
                    float *We_out);

int ref_gradient(const int *ip, const float *fp, const float *FLv, const float *vp, const float *vs,
                 const float *rho, const float *taus, const float *taup, int nsrc, const float *srcpos, int ntr,
                 const int *recpos, const float *diff_vx, const float *diff_vy, float *seis_vx, float *seis_vy,
                 float *grad_vp, float *grad_vs, float *grad_rho)
{
    return ref_gradient_we(ip, fp, FLv, vp, vs, rho, taus, taup, nsrc, srcpos, ntr, recpos, diff_vx, diff_vy,
                           seis_vx, seis_vy, grad_vp, grad_vs, grad_rho, 0, 0.0f, 0, NULL);
}

/* with energy preconditioning: the reference's eprecond() accumulates the
 * forward (and for EPRECOND 1 the adjoint) energies every time step,
 * eprecond1() forms We, and the per-shot division of src/IFOS2D.c:2627-2638
 * follows; We_out receives We after eprecond1. */
int ref_gradient_we(const int *ip, const float *fp, const float *FLv, const float *vp, const float *vs,
                    const float *rho, const float *taus, const float *taup, int nsrc, const float *srcpos, int ntr,
                    const int *recpos, const float *diff_vx, const float *diff_vy, float *seis_vx, float *seis_vy,
                    float *grad_vp, float *grad_vs, float *grad_rho, int ep_mode, float epsilon_we, int per_shot,
                    float *We_out)
{
    harness_setup(ip, fp);
    EPRECOND = ep_mode;
    const int nd = FDORDER / 2 + 1;
    float **We = NULL;
    if (ep_mode == 1 || ep_mode == 3) {
        harness_Ws = matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
        harness_Wr = matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
        We = matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    }
    int DTINV = ip[11];
    run_t R;
    run_init(&R, vp, vs, rho, taus, taup, FLv);
    int nsnap = NT / DTINV + (NT % DTINV ? 1 : 0);
    size_t pl = (size_t)NX * NY;
    float *snap[5];
    for (int k = 0; k < 5; k++) snap[k] = (float *)calloc((size_t)nsnap * pl, sizeof(float));
    int *help = (int *)calloc((size_t)NT + 2, sizeof(int));
    run_forward(&R, nsrc, srcpos, ntr, recpos, seis_vx, seis_vy, snap, help, DTINV);

    float **srcpos_loc_back = matrix(1, 6, 1, ntr);
    for (int t = 1; t <= ntr; t++) {
        srcpos_loc_back[1][t] = recpos[2 * (t - 1)];
        srcpos_loc_back[2][t] = recpos[2 * (t - 1) + 1];
    }
    float **dvx = matrix(1, ntr, 1, NT), **dvy = matrix(1, ntr, 1, NT);
    for (int t = 1; t <= ntr; t++)
        for (int nt = 1; nt <= NT; nt++) {
            dvx[t][nt] = diff_vx[(size_t)(t - 1) * NT + nt - 1];
            dvy[t][nt] = diff_vy[(size_t)(t - 1) * NT + nt - 1];
        }
    float *wc = (float *)calloc(pl, sizeof(float)), *wcu = (float *)calloc(pl, sizeof(float));
    float *wcr = (float *)calloc(pl, sizeof(float));
    run_zero(&R);
    int hin = 1;
    for (int nt = 1; nt <= NT; nt++) {
        run_step(&R, nt, 1, srcpos_loc_back, dvx, dvy, ntr);
        if (help[NT - nt + 1] == 1) {
            size_t off = (size_t)(nsnap - hin) * pl;
            for (int j = 1; j <= NY; j++)
                for (int i = 1; i <= NX; i++) {
                    size_t c = (size_t)(j - 1) * NX + (i - 1);
                    float Frx = snap[0][off + c], Fry = snap[1][off + c];
                    float Fx = snap[2][off + c], Fy = snap[3][off + c], Fu = snap[4][off + c];
                    float sxx = R.psxx[j][i], syy = R.psyy[j][i], sxy = R.psxy[j][i];
                    float muss, lamss;
                    wcr[c] += (R.pvxp1[j][i] * Frx) + (R.pvyp1[j][i] * Fry);
                    wc[c] += (Fx + Fy) * (sxx + syy);
                    muss = R.prho[j][i] * R.pu[j][i] * R.pu[j][i];
                    lamss = R.prho[j][i] * R.ppi[j][i] * R.ppi[j][i] - 2.0 * muss;
                    if (R.pu[j][i] > 0.0)
                        wcu[c] += ((1.0 / (muss * muss)) * (Fu * sxy)) +
                                  (.25 * (Fx + Fy * (sxx + syy)) / ((lamss + muss) * (lamss + muss))) +
                                  (.25 * (Fx - Fy * (sxx - syy)) / (muss * muss));
                }
            hin++;
        }
        if (ep_mode == 1) eprecond(harness_Wr, R.pvx, R.pvy);  /* src/IFOS2D.c:2420-2425 */
    }
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            size_t c = (size_t)(j - 1) * NX + (i - 1);
            float rh = R.prho[j][i], u = R.pu[j][i], pv = R.ppi[j][i];
            float wl = -DT * wc[c], muss = rh * u * u, lamss = rh * pv * pv - 2.0 * muss;
            wl = (1.0 / (4.0 * (lamss + muss) * (lamss + muss))) * wl;
            grad_vp[c] = 2.0 * pv * rh * wl;
            float wmu = -DT * wcu[c];
            grad_vs[c] = -4.0 * rh * u * wl + 2.0 * rh * u * wmu;
            float wrs = -DT * wcr[c];
            grad_rho[c] = ((pv * pv) - (2.0 * u * u)) * wl + u * u * wmu + wrs;
        }
    if (We) {
        eprecond1(We, harness_Ws, harness_Wr, epsilon_we);
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++)
                if (We_out) We_out[(size_t)(j - 1) * NX + (i - 1)] = We[j][i];
        if (per_shot) {
            const float We_max = global_maximum(We);
            for (int j = 1; j <= NY; j++)
                for (int i = 1; i <= NX; i++) {
                    size_t c = (size_t)(j - 1) * NX + (i - 1);
                    We[j][i] = We[j][i] / We_max;
                    grad_vp[c] = grad_vp[c] / (We[j][i]);
                    grad_vs[c] = grad_vs[c] / (We[j][i]);
                    grad_rho[c] = grad_rho[c] / (We[j][i]);
                }
        }
        free_matrix(harness_Ws, -nd + 1, NY + nd, -nd + 1, NX + nd);
        free_matrix(harness_Wr, -nd + 1, NY + nd, -nd + 1, NX + nd);
        free_matrix(We, -nd + 1, NY + nd, -nd + 1, NX + nd);
        harness_Ws = harness_Wr = NULL;
    }
    free(wc); free(wcu); free(wcr); free(help);
    for (int k = 0; k < 5; k++) free(snap[k]);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Optimizer routines of the reference, called unmodified: calc_opt_step
 * (+ solvelin), lbfgs_core and calc_mat_change_test in its test-step mode
 * (itest = 1: computes the updated models, writes no files).
 * ------------------------------------------------------------------------- */
static FILE *harness_null = NULL;

static void harness_setup_opt(int nx, int ny)
{
    int ip[16] = {0};
    float fp[8] = {0};
    ip[0] = nx; ip[1] = ny; ip[2] = 2; ip[10] = 1; ip[12] = 1;
    harness_setup(ip, fp);
    if (!harness_null) harness_null = fopen("/dev/null", "w");
    FP = harness_null;
    if (!FL) FL = vector(1, 8);
}

/* L2t[0..2], epst[0..2] are the reference's L2t[1..3], epst1[1..3] */
float ref_calc_opt_step(const float *L2t3, const float *epst3)
{
    harness_setup_opt(1, 1);
    float *L2t = vector(1, 4), *epst = vector(1, 3);
    for (int k = 1; k <= 3; k++) {
        L2t[k] = L2t3[k - 1];
        epst[k] = epst3[k - 1];
    }
    L2t[4] = 0.0f;
    float e = calc_opt_step(L2t, NULL, NULL, epst, 1, 1.0f);
    free_vector(L2t, 1, 4);
    free_vector(epst, 1, 3);
    return e;
}

/* s, y: [N_LBFGS][n] (row w-1 = the reference's vector w), rho [N_LBFGS],
 * q, r: [n]; alpha [N_LBFGS].  q is modified as in the reference. */
void ref_lbfgs_core(int iteration, int N_LBFGS, int n, float *s, float *y, float *rho, float *q, float *alpha,
                    float *r)
{
    harness_setup_opt(1, 1);
    float **S = matrix(1, N_LBFGS, 1, n), **Y = matrix(1, N_LBFGS, 1, n);
    float *R = vector(1, N_LBFGS), *Q = vector(1, n), *A = vector(1, N_LBFGS), *Rr = vector(1, n);
    for (int w = 1; w <= N_LBFGS; w++) {
        R[w] = rho[w - 1];
        A[w] = alpha[w - 1];
        for (int l = 1; l <= n; l++) {
            S[w][l] = s[(size_t)(w - 1) * n + l - 1];
            Y[w][l] = y[(size_t)(w - 1) * n + l - 1];
        }
    }
    for (int l = 1; l <= n; l++) Q[l] = q[l - 1];
    lbfgs_core(iteration, N_LBFGS, n, S, Y, R, Q, A, Rr);
    for (int w = 1; w <= N_LBFGS; w++) {
        rho[w - 1] = R[w];
        alpha[w - 1] = A[w];
    }
    for (int l = 1; l <= n; l++) {
        q[l - 1] = Q[l];
        r[l - 1] = Rr[l];
    }
    free_matrix(S, 1, N_LBFGS, 1, n);
    free_matrix(Y, 1, N_LBFGS, 1, n);
    free_vector(R, 1, N_LBFGS);
    free_vector(Q, 1, n);
    free_vector(A, 1, N_LBFGS);
    free_vector(Rr, 1, n);
}

/* Test-step model update (PARAMETERIZATION 1, itest = 1).  Arrays are
 * interior [NY][NX]: models vp, vs, rho, directions g_vp, g_vs, g_rho,
 * starting models (for the S == 1 limits); outputs vp1, vs1, rho1 and
 * eps[3] = EPSILON, EPSILON_u, EPSILON_rho.
 * ip: iter, INV_VP_ITER, INV_VS_ITER, INV_RHO_ITER, S, GRAD_METHOD,
 *     LBFGS_STEP_LENGTH, LBFGS_iter_start
 * fp: eps_scale, VPUPPERLIM, VPLOWERLIM, VSUPPERLIM, VSLOWERLIM, RHOUPPERLIM,
 *     RHOLOWERLIM, VP_VS_RATIO, S_VP, S_VS, S_RHO */
void ref_mat_change(int nx, int ny, const int *ip, const float *fp, const float *vp, const float *vs,
                    const float *rho, const float *g_vp, const float *g_vs, const float *g_rho, const float *vp0,
                    const float *vs0, const float *rho0, float *vp1, float *vs1, float *rho1, float *eps)
{
    harness_setup_opt(nx, ny);
    INV_VP_ITER = ip[1]; INV_VS_ITER = ip[2]; INV_RHO_ITER = ip[3]; S = ip[4]; GRAD_METHOD = ip[5];
    LBFGS_STEP_LENGTH = ip[6];
    VPUPPERLIM = fp[1]; VPLOWERLIM = fp[2]; VSUPPERLIM = fp[3]; VSLOWERLIM = fp[4];
    RHOUPPERLIM = fp[5]; RHOLOWERLIM = fp[6]; VP_VS_RATIO = fp[7]; S_VP = fp[8]; S_VS = fp[9]; S_RHO = fp[10];
    float **a[12];
    const float *src[12] = {g_vp, g_rho, g_vs, rho, NULL, vp, NULL, vs, NULL, vs0, vp0, rho0};
    for (int k = 0; k < 12; k++) {
        a[k] = matrix(1, ny, 1, nx);
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) a[k][j][i] = src[k] ? src[k][(size_t)(j - 1) * nx + i - 1] : 0.0f;
    }
    /* (waveconv, waveconv_rho, waveconv_u, rho, rhonp1, pi, pinp1, u, unp1, ...)
     * as src/IFOS2D.c:3239 passes them */
    calc_mat_change_test(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], ip[0], 1, 0, fp[0], 1, 1, a[9],
                         a[10], a[11], 1, NULL, 5, 3, 1.0f, 1.0f, 1.0f, ip[7]);
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            const size_t e = (size_t)(j - 1) * nx + i - 1;
            vp1[e] = a[6][j][i];
            vs1[e] = a[8][j][i];
            rho1[e] = a[4][j][i];
        }
    eps[0] = EPSILON;
    eps[1] = EPSILON_u;
    eps[2] = EPSILON_rho;
    for (int k = 0; k < 12; k++) free_matrix(a[k], 1, ny, 1, nx);
}

/* ---------------------------------------------------------------------------
 * Gradient post-processing of the reference, called unmodified:
 * taper_grad_shot (src/taper_grad_shot.c:29), taper_grad
 * (src/taper_grad.c:28) and smooth (src/smooth.c:24).  Gradients are
 * interior [NY][NX]; source positions in metres (srcpos[1..2][n]),
 * receivers as grid indices [ntr][2].  taper_grad writes
 * taper_coeff_*.bin into the working directory, as the reference does.
 * ------------------------------------------------------------------------- */
static float **harness_grid_in(int nx, int ny, const float *g)
{
    const int nd = 3;
    float **a = matrix(-nd + 1, ny + nd, -nd + 1, nx + nd);
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) a[j][i] = g[(size_t)(j - 1) * nx + i - 1];
    return a;
}

static void harness_grid_out(int nx, int ny, float **a, float *g)
{
    const int nd = 3;
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) g[(size_t)(j - 1) * nx + i - 1] = a[j][i];
    free_matrix(a, -nd + 1, ny + nd, -nd + 1, nx + nd);
}

static float **harness_srcpos(int nshots, const float *xs, const float *ys)
{
    float **s = matrix(1, 8, 1, nshots);
    for (int n = 1; n <= nshots; n++) {
        s[1][n] = xs[n - 1];
        s[2][n] = ys[n - 1];
    }
    return s;
}

/* taper_grad_shot(..., ishot, sws) for one gradient; sws 1 (circular) */
void ref_taper_grad_shot(int nx, int ny, float dh, float srtradius, int srtshape, int filtsize, int nshots,
                         const float *xs, const float *ys, int ishot, float *g)
{
    harness_setup_opt(nx, ny);
    DH = dh; SRTRADIUS = srtradius; SRTSHAPE = srtshape; FILTSIZE = filtsize;
    USE_WORKFLOW = 0;
    float **w = harness_grid_in(nx, ny, g);
    /* taper_coeff is matrix(1,NY,1,NX) in src/IFOS2D.c:645; a wider one
     * here keeps the box writes left of column 1 inside the allocation */
    float **tc = matrix(-2, ny + 3, -2 * nx, nx + 3);
    float **sp = harness_srcpos(nshots, xs, ys);
    int **rp = imatrix(1, 3, 1, 1);
    taper_grad_shot(w, tc, sp, nshots, rp, 1, ishot, 1);
    harness_grid_out(nx, ny, w, g);
    free_matrix(tc, -2, ny + 3, -2 * nx, nx + 3);
    free_matrix(sp, 1, 8, 1, nshots);
    free_imatrix(rp, 1, 3, 1, 1);
}

/* taper_grad(..., sws) for sws 1 (vertical), 2 (horizontal), 3 (sources) */
void ref_taper_grad(int nx, int ny, float dh, const int *gradt, float srtradius, int srtshape, int filtsize,
                    int nshots, const float *xs, const float *ys, int ntr, const int *recpos, int sws, float *g)
{
    harness_setup_opt(nx, ny);
    DH = dh; SRTRADIUS = srtradius; SRTSHAPE = srtshape; FILTSIZE = filtsize;
    GRADT1 = gradt[0]; GRADT2 = gradt[1]; GRADT3 = gradt[2]; GRADT4 = gradt[3];
    USE_WORKFLOW = 0;
    float **w = harness_grid_in(nx, ny, g);
    float **tc = matrix(-2, ny + 3, -2 * nx, nx + 3);
    float **sp = harness_srcpos(nshots, xs, ys);
    int **rp = imatrix(1, 3, 1, ntr > 0 ? ntr : 1);
    for (int n = 1; n <= ntr; n++) {
        rp[1][n] = recpos[2 * (n - 1)];
        rp[2][n] = recpos[2 * (n - 1) + 1];
    }
    taper_grad(w, tc, sp, nshots, rp, ntr, sws);
    harness_grid_out(nx, ny, w, g);
    free_matrix(tc, -2, ny + 3, -2 * nx, nx + 3);
    free_matrix(sp, 1, 8, 1, nshots);
    free_imatrix(rp, 1, 3, 1, ntr > 0 ? ntr : 1);
}

/* smooth(mat, 1, 1, ...) with FILT_SIZE_GRAD = filtsize; returns the
 * FILT_SIZE_GRAD the call leaves behind (an even size grows by one) */
int ref_smooth(int nx, int ny, int filtsize, float *g)
{
    harness_setup_opt(nx, ny);
    FILT_SIZE_GRAD = filtsize;
    GRAD_FILT_WAVELENGTH = 0;
    TIME_FILT = 0;
    float **w = harness_grid_in(nx, ny, g);
    smooth(w, 1, 1, 1.0f, 1.0f);
    harness_grid_out(nx, ny, w, g);
    return FILT_SIZE_GRAD;
}

/* create_trkill_table (src/create_trkill_table.c:21) on a cleared table:
 * recpos [ntr][2] grid indices, srcpos [nsrc][2] metres, out [ntr][nsrc] */
void ref_trkill_offsets(int ntr, const int *recpos, int nsrc, const float *srcpos, float dh, float lower, float upper,
                        int *out)
{
    harness_setup_opt(1, 1);
    DH = dh;
    int **kt = imatrix(1, ntr, 1, nsrc);
    int **rp = imatrix(1, 3, 1, ntr);
    float **sp = matrix(1, 8, 1, nsrc);
    for (int r = 1; r <= ntr; r++) {
        rp[1][r] = recpos[2 * (r - 1)];
        rp[2][r] = recpos[2 * (r - 1) + 1];
        for (int s = 1; s <= nsrc; s++) kt[r][s] = 0;
    }
    for (int s = 1; s <= nsrc; s++) {
        sp[1][s] = srcpos[2 * (s - 1)];
        sp[2][s] = srcpos[2 * (s - 1) + 1];
    }
    create_trkill_table(kt, ntr, rp, nsrc, sp, 1, lower, upper);
    for (int r = 1; r <= ntr; r++)
        for (int s = 1; s <= nsrc; s++) out[(r - 1) * nsrc + (s - 1)] = kt[r][s];
    free_imatrix(kt, 1, ntr, 1, nsrc);
    free_imatrix(rp, 1, 3, 1, ntr);
    free_matrix(sp, 1, 8, 1, nsrc);
}

/* timedomain_filt (src/timedomain_filt.c:27) on data [ntr][ns]: method 1
 * lowpass at fc (plus a highpass at fhp when fhp != 0), 2 highpass at fc */
void ref_time_filter(float *data, int ntr, int ns, float dt, float fc, int order, int method, float fhp)
{
    harness_setup_opt(1, 1);
    DT = dt;
    F_HIGH_PASS = fhp;
    NT = ns;
    float **d = matrix(1, ntr, 1, ns);
    for (int t = 1; t <= ntr; t++)
        for (int j = 1; j <= ns; j++) d[t][j] = data[(size_t)(t - 1) * ns + j - 1];
    timedomain_filt(d, fc, order, ntr, ns, method);
    for (int t = 1; t <= ntr; t++)
        for (int j = 1; j <= ns; j++) data[(size_t)(t - 1) * ns + j - 1] = d[t][j];
    free_matrix(d, 1, ntr, 1, ns);
}

/* spat_filt (src/spat_filt.c:27) with sws 1 on gradient g [NY][NX]: the
 * routine reads the gradient from <prefix>_g.old as PCG writes it
 * (src/PCG.c:77-85, x outer) and writes the filtered one back there */
void ref_spat_filt(int nx, int ny, int size, int first_row, const char *prefix, float *g)
{
    harness_setup_opt(nx, ny);
    SPAT_FILT_SIZE = size;
    SPAT_FILT_1 = first_row;
    SPAT_FILT_ITER = 0;
    snprintf(JACOBIAN, STRING_SIZE, "%s", prefix);
    char path[STRING_SIZE + 16];
    snprintf(path, sizeof path, "%s_g.old", prefix);
    FILE *f = fopen(path, "wb");
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++) fwrite(&g[(size_t)(j - 1) * nx + i - 1], sizeof(float), 1, f);
    fclose(f);
    float **w = harness_grid_in(nx, ny, g);
    spat_filt(w, 1, 1);
    harness_grid_out(nx, ny, w, g);
}

/* ---------------------------------------------------------------------------
 * Single-operator probes: the reference's own update_v_PML,
 * update_s_elastic_PML, surface_elastic_PML and surface_PML applied once to
 * given fields.  Probe grids are padded (rows -nd+1..NY+nd, columns
 * -nd+1..NX+nd, x fastest), as orc_probe_* in ifos_oracle.h.
 * ------------------------------------------------------------------------- */
static void probe_in(float **a, const float *src, int nd)
{
    const int w = NX + 2 * nd;
    for (int j = -nd + 1; j <= NY + nd; j++)
        for (int i = -nd + 1; i <= NX + nd; i++) a[j][i] = src[(size_t)(j + nd - 1) * w + (size_t)(i + nd - 1)];
}
static void probe_out(float **a, float *dst, int nd)
{
    const int w = NX + 2 * nd;
    for (int j = -nd + 1; j <= NY + nd; j++)
        for (int i = -nd + 1; i <= NX + nd; i++) dst[(size_t)(j + nd - 1) * w + (size_t)(i + nd - 1)] = a[j][i];
}
static void probe_zero(float **a, int nd)
{
    for (int j = -nd + 1; j <= NY + nd; j++)
        for (int i = -nd + 1; i <= NX + nd; i++) a[j][i] = 0.0f;
}

/* The staggered difference sums of the reference kernels at any FDORDER,
 * read off single updates.  DT = DH = 1 and unit materials make every update
 * return one sum, unscaled and unrounded beyond the sum itself:
 *  - update_v_PML, vx = vy = 0, rho = 1 (rip = rjp = 1): with sxy = 0 it
 *    leaves vx = sxx_x, vy = syy_y; with sxx = syy = 0, vx = sxy_y and
 *    vy = sxy_x (the zero terms add exact zeros, also in the single sums of
 *    orders 10/12, src/update_v_PML.c:466-547, whose DT*DT/DH factor is 1);
 *  - update_s_elastic_PML with vs = 0, vp = rho = 1 (f = 0, g = 1): with
 *    vy = 0 it leaves ux = vxx, with vx = 0 uy = vyy;
 *  - the same with vs = vp = rho = 1 (uipjp = 1): with vx = 0 it leaves
 *    sxy = vyx, with vy = 0 sxy = vxy.
 * BOUNDARY = 1 and FREE_SURF = 1 keep the CPML off every cell but the bottom
 * strip j > NY-FW (orders 2-8 apply it there; orders 10/12 have none), which
 * the caller leaves out.  out8 as orc_probe_derivs. */
void ref_probe_derivs(const int *ip, const float *fp, const float *fields5, float *out8)
{
    harness_setup(ip, fp);
    DT = 1.0f; DH = 1.0f; BOUNDARY = 1; FREE_SURF = 1; L = 0;
    const int nd = FDORDER / 2 + 1;
    const size_t pl = (size_t)(NX + 2 * nd) * (size_t)(NY + 2 * nd), ipl = (size_t)NX * NY;
    float *ones = (float *)malloc(sizeof(float) * ipl), *zeros = (float *)calloc(ipl, sizeof(float));
    for (size_t k = 0; k < ipl; k++) ones[k] = 1.0f;
    const float *VX = fields5, *VY = fields5 + pl, *SXX = fields5 + 2 * pl, *SYY = fields5 + 3 * pl,
                *SXY = fields5 + 4 * pl;
    for (int pass = 0; pass < 2; pass++) {
        /* pass 0: vs = 0 (ux, uy read the normal derivatives); pass 1: vs = 1 */
        run_t R;
        run_init(&R, ones, pass ? ones : zeros, ones, NULL, NULL, NULL);
        pml_t *P = &R.P;
        for (int k = 0; k < 6; k++) {
            run_zero(&R);
            if (pass == 0 && k == 0) { probe_in(R.psxx, SXX, nd); probe_in(R.psyy, SYY, nd); }
            else if (pass == 0 && k == 1) { probe_in(R.psxy, SXY, nd); }
            else if (pass == 0 && k == 2) { probe_in(R.pvx, VX, nd); }
            else if (pass == 0 && k == 3) { probe_in(R.pvy, VY, nd); }
            else if (pass == 1 && k == 4) { probe_in(R.pvy, VY, nd); }
            else if (pass == 1 && k == 5) { probe_in(R.pvx, VX, nd); }
            else continue;
            if (k <= 1)
                update_v_PML(1, NX, 1, NY, 1, R.pvx, R.pvxp1, R.pvxm1, R.pvy, R.pvyp1, R.pvym1, R.uttx, R.utty,
                             R.psxx, R.psyy, R.psxy, R.prip, R.prjp, NULL, NULL, NULL, 0, NULL, R.hc, 0, 0,
                             P->K_x, P->a_x, P->b_x, P->K_xh, P->a_xh, P->b_xh, P->K_y, P->a_y, P->b_y, P->K_yh,
                             P->a_yh, P->b_yh, R.psi_sxx_x, R.psi_syy_y, R.psi_sxy_y, R.psi_sxy_x);
            else
                update_s_elastic_PML(1, NX, 1, NY, R.pvx, R.pvy, R.ux, R.uy, R.uxy, R.uyx, R.psxx, R.psyy,
                                     R.psxy, R.ppi, R.pu, R.puipjp, NULL, R.prho, R.hc, 0, P->K_x, P->a_x, P->b_x,
                                     P->K_xh, P->a_xh, P->b_xh, P->K_y, P->a_y, P->b_y, P->K_yh, P->a_yh, P->b_yh,
                                     R.psi_vxx, R.psi_vyy, R.psi_vxy, R.psi_vyx);
            /* out8 order: sxx_x, sxy_x, sxy_y, syy_y, vxx, vyx, vxy, vyy */
            float **src0 = NULL, **src1 = NULL;
            int o0 = -1, o1 = -1;
            if (k == 0) { src0 = R.pvx; o0 = 0; src1 = R.pvy; o1 = 3; }
            if (k == 1) { src0 = R.pvx; o0 = 2; src1 = R.pvy; o1 = 1; }
            if (k == 2) { src0 = R.ux; o0 = 4; }
            if (k == 3) { src0 = R.uy; o0 = 7; }
            if (k == 4) { src0 = R.psxy; o0 = 5; }
            if (k == 5) { src0 = R.psxy; o0 = 6; }
            copy_interior(src0, out8 + (size_t)o0 * ipl);
            if (src1) copy_interior(src1, out8 + (size_t)o1 * ipl);
        }
    }
    free(ones);
    free(zeros);
}

/* One surface_elastic_PML (L = 0) or surface_PML (L > 0) call of the
 * reference on the given fields (fields8 = vx, vy, sxx, syy, sxy, ux, uy,
 * uxy, padded, in and out) with zero CPML memory; prow [L][NX] holds the
 * memory variables p_l of row 1 (in and out). */
void ref_probe_surface(const int *ip, const float *fp, const float *FLv, const float *vp, const float *vs,
                       const float *rho, const float *taus, const float *taup, float *fields8, float *prow)
{
    harness_setup(ip, fp);
    run_t R;
    run_init(&R, vp, vs, rho, taus, taup, FLv);
    run_zero(&R);
    const int nd = R.nd;
    const size_t pl = (size_t)(NX + 2 * nd) * (size_t)(NY + 2 * nd);
    float **f[8] = {R.pvx, R.pvy, R.psxx, R.psyy, R.psxy, R.ux, R.uy, R.uxy};
    for (int k = 0; k < 8; k++) probe_in(f[k], fields8 + (size_t)k * pl, nd);
    for (int l = 1; l <= L; l++)
        for (int i = 1; i <= NX; i++) R.pp[1][i][l] = prow[(size_t)(l - 1) * NX + (i - 1)];
    pml_t *P = &R.P;
    if (L)
        surface_PML(1, R.pvx, R.pvy, R.psxx, R.psyy, R.psxy, R.psyz, R.pp, R.pq, R.ppi, R.pu, R.prho, R.ptaup,
                    R.ptaus, R.etajm, R.peta, R.hc, P->K_x, P->a_x, P->b_x, R.psi_vxxs, R.ux, R.uy, R.uxy, R.uyz,
                    R.psxz, R.uxz);
    else
        surface_elastic_PML(1, R.pvx, R.pvy, R.psxx, R.psyy, R.psxy, R.psyz, R.ppi, R.pu, R.prho, R.hc, P->K_x,
                            P->a_x, P->b_x, R.psi_vxxs, R.ux, R.uy, R.uxy, R.uyz, R.psxz, R.uxz);
    for (int k = 0; k < 8; k++) probe_out(f[k], fields8 + (size_t)k * pl, nd);
    for (int l = 1; l <= L; l++)
        for (int i = 1; i <= NX; i++) prow[(size_t)(l - 1) * NX + (i - 1)] = R.pp[1][i][l];
}

/* ---------------------------------------------------------------------------
 * PCG() (src/PCG.c:27) called unmodified for iterations 1..niter on the
 * gradients grads [niter][3][NY][NX] (Vp, Vs, rho), with every taper and
 * filter off.  Its state between iterations goes through the files it writes
 * and reads (<prefix>_p.old.0.0, _c.old.0.0, ... for each parameter).
 * dirs [niter][3][NY][NX] receives waveconv after each call (the search
 * direction), gradps the scaled gradients gradp.
 * ------------------------------------------------------------------------- */
void ref_pcg(int nx, int ny, int niter, int pcg_iter_start, const float *C3, const char *prefix,
             const float *grads, float *dirs, float *gradps)
{
    harness_setup_opt(nx, ny);
    IDX = IDY = 1;
    SPATFILTER = 0; GRAD_FILTER = 0; FORWARD_ONLY = 0; ACOUSTIC = 0; VERBOSE = 0;
    SWS_TAPER_GRAD_VERT = SWS_TAPER_GRAD_HOR = SWS_TAPER_GRAD_SOURCES = SWS_TAPER_FILE = 0;
    snprintf(JACOBIAN, STRING_SIZE, "%s", prefix);
    const size_t pl = (size_t)nx * ny;
    float **w[3], **gp[3];
    for (int k = 0; k < 3; k++) {
        w[k] = matrix(1, ny, 1, nx);
        gp[k] = matrix(1, ny, 1, nx);
    }
    float **tc = matrix(1, ny, 1, nx);
    float **sp = matrix(1, 8, 1, 1);
    int **rp = imatrix(1, 3, 1, 1);
    for (int it = 1; it <= niter; it++) {
        for (int k = 0; k < 3; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++)
                    w[k][j][i] = grads[((size_t)(it - 1) * 3 + k) * pl + (size_t)(j - 1) * nx + (i - 1)];
        PCG(w[0], tc, 1, sp, rp, 1, it, C3[0], gp[0], -1, w[1], C3[1], gp[1], w[2], C3[2], gp[2], 0.0f, 0.0f,
            pcg_iter_start);
        for (int k = 0; k < 3; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    size_t e = ((size_t)(it - 1) * 3 + k) * pl + (size_t)(j - 1) * nx + (i - 1);
                    dirs[e] = w[k][j][i];
                    gradps[e] = gp[k][j][i];
                }
    }
    for (int k = 0; k < 3; k++) {
        free_matrix(w[k], 1, ny, 1, nx);
        free_matrix(gp[k], 1, ny, 1, nx);
    }
    free_matrix(tc, 1, ny, 1, nx);
    free_matrix(sp, 1, 8, 1, 1);
    free_imatrix(rp, 1, 3, 1, 1);
}

/* ---------------------------------------------------------------------------
 * File readers of the reference, called unmodified: sources() (SRCREC 1),
 * splitsrc() on one rank, receiver() (READREC 1 and 0), inseis().
 * ------------------------------------------------------------------------- */
int ref_sources(const char *path, int source_type, int max_src, float *out, int *nsrc_out, float *fc_max)
{
    harness_setup_opt(1, 1);
    SRCREC = 1;
    SOURCE_TYPE = source_type;
    PLANE_WAVE_DEPTH = 0.0f;
    RUN_MULTIPLE_SHOTS = 1;
    snprintf(SOURCE_FILE, STRING_SIZE, "%s", path);
    int n = 0;
    float **sp = sources(&n);
    if (n > max_src) return -1;
    for (int l = 1; l <= n; l++)
        for (int k = 1; k <= 8; k++) out[(size_t)(l - 1) * 8 + k - 1] = sp[k][l];
    *nsrc_out = n;
    *fc_max = 1.0f / TS;
    return 0;
}

int ref_splitsrc(int nx, int ny, float dh, int nsrc, const float *srcpos_m, float *out, int *nloc)
{
    harness_setup_opt(nx, ny);
    DH = dh;
    IENDX = nx;
    IENDY = ny;
    float **sp = matrix(1, 8, 1, nsrc);
    for (int l = 1; l <= nsrc; l++)
        for (int k = 1; k <= 8; k++) sp[k][l] = srcpos_m[(size_t)(l - 1) * 8 + k - 1];
    int n = 0;
    float **loc = splitsrc(sp, &n, nsrc);
    for (int l = 1; l <= n; l++)
        for (int k = 1; k <= 8; k++) out[(size_t)(l - 1) * 8 + k - 1] = loc[k][l];
    *nloc = n;
    return 0;
}

int ref_receivers(int readrec, const char *path, float dh, float rx, float ry, float x1, float y1, float x2, float y2,
                  int ngeoph, int max_tr, int *out, int *ntr_out)
{
    harness_setup_opt(1, 1);
    DH = dh;
    READREC = readrec;
    REFREC[1] = rx;
    REFREC[2] = ry;
    REFREC[3] = 0.0f;
    XREC1 = x1; YREC1 = y1; XREC2 = x2; YREC2 = y2;
    NGEOPH = ngeoph;
    VERBOSE = 0;
    snprintf(REC_FILE, STRING_SIZE, "%s", path ? path : "");
    int n = 0;
    int **rp = receiver(&n, NULL, 0);
    if (n > max_tr) return -1;
    for (int t = 1; t <= n; t++) {
        out[2 * (t - 1)] = rp[1][t];
        out[2 * (t - 1) + 1] = rp[2][t];
    }
    *ntr_out = n;
    return 0;
}

/* inseis(..., sws 1) on <prefix>_vx.su.shot<shot>: out [ntr][ns] */
void ref_inseis(const char *prefix, int shot, int ntr, int ns, float *out)
{
    harness_setup_opt(1, 1);
    snprintf(DATA_DIR, STRING_SIZE, "%s", prefix);
    /* one spare column: the reference writes section[t][ns+1] */
    float **sec = matrix(1, ntr, 1, ns + 1);
    inseis(NULL, shot, sec, ntr, ns, 1, 0);
    for (int t = 1; t <= ntr; t++)
        for (int k = 1; k <= ns; k++) out[(size_t)(t - 1) * ns + k - 1] = sec[t][k];
    free_matrix(sec, 1, ntr, 1, ns + 1);
}

/* taper (src/taper.c:24, TAPER_STF of stf()) on one trace of ns samples,
 * 0-based; the reference writes its window up to sample 2*ceil(2/fc/DT),
 * so the trace is copied into a vector that long when that is beyond ns */
void ref_stf_taper(float *trace, int ns, float fc, float dt)
{
    harness_setup_opt(1, 1);
    DT = dt;
    MYID = 0;
    const int tl = (int)(ceil(2.0 / fc / DT));
    const int n = ns > 2 * tl ? ns : 2 * tl;
    float *v = vector(1, n);
    for (int j = 1; j <= n; j++) v[j] = j <= ns ? trace[j - 1] : 0.0f;
    taper(v, ns, fc);
    for (int j = 1; j <= ns; j++) trace[j - 1] = v[j];
    free_vector(v, 1, n);
}
