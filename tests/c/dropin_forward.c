/*
 * dropin_forward.c -- a C client of the drop-in layer (include/ifos2d_dropin.h)
 * making the calls of the reference's forward loop, src/IFOS2D.c:1651-1803:
 * per time step update_v_PML, update_s_elastic_PML, psource, surface_elastic_PML
 * and seismo_ssg, on device arrays allocated with the reference's bounds
 * (src/IFOS2D.c:371-641).  Host setup uses the library's holbergcoeff /
 * PML_pro / wavelet / av_* equivalents.  Elastic (L = 0), explosive source,
 * SEISMO 1.
 *
 * Built as a shared library (tests run it in-process through
 * dropin_forward_run) and, with DROPIN_MAIN defined, as a program:
 * dropin_forward IN OUT
 *   IN : int NX NY NT FDORDER FW FREE_SURF VELOCITY nsrc ntr;
 *        float DT DH VPPML FPML npower k_max_PML;
 *        float srcpos[nsrc][8]; int recpos[ntr][2];
 *        float vp[NY][NX] vs[NY][NX] rho[NY][NX]
 *   OUT: float seis_vx[ntr][NT] seis_vy[ntr][NT] final[5][NY][NX]
 *        (vx, vy, sxx, syy, sxy)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ifos2d_dropin.h"
#include "ifos2d_hip.h"
#include "host/ifos_setup.h"

/* the reference's globals (src/globvar.h), defined by the library */
extern float DT, DH;
extern int NX, NY, NT, FDORDER, FW, FREE_SURF, BOUNDARY, L, PARAMETERIZATION, VELOCITY, ADJOINT_TYPE, SOURCE_TYPE,
    SEISMO;

#define FAIL(m)                                          \
    do {                                                 \
        fprintf(stderr, "dropin_forward: %s\n", (m));    \
        return 1;                                        \
    } while (0)

/* host NR matrix (src/util.c:184): rows of one block, offset by nrl / ncl */
static float **host_matrix(int nrl, int nrh, int ncl, int nch)
{
    const long nr = nrh - nrl + 1, nc = nch - ncl + 1;
    float *b = calloc((size_t)(nr * nc), sizeof(float));
    float **rows = malloc(sizeof(float *) * nr);
    for (long r = 0; r < nr; r++) rows[r] = b + r * nc - ncl;
    return rows - nrl;
}

static int **host_imatrix(int nrl, int nrh, int ncl, int nch)
{
    const long nr = nrh - nrl + 1, nc = nch - ncl + 1;
    int *b = calloc((size_t)(nr * nc), sizeof(int));
    int **rows = malloc(sizeof(int *) * nr);
    for (long r = 0; r < nr; r++) rows[r] = b + r * nc - ncl;
    return rows - nrl;
}

/* upload an interior [NY][NX] array into a device matrix of the wavefield bounds */
static int upload(float **dev, const float *interior, int nd)
{
    const int nx = NX + 2 * nd, ny = NY + 2 * nd;
    float *h = calloc((size_t)nx * ny, sizeof(float));
    for (int j = 0; j < NY; j++) memcpy(h + (size_t)(j + nd) * nx + nd, interior + (size_t)j * NX, sizeof(float) * NX);
    const int rc = ifos_dev_copy_array(dev, h, 0);
    free(h);
    return rc;
}

static int download(float **dev, float *interior, int nd)
{
    const int nx = NX + 2 * nd, ny = NY + 2 * nd;
    float *h = calloc((size_t)nx * ny, sizeof(float));
    const int rc = ifos_dev_copy_array(dev, h, 1);
    for (int j = 0; j < NY; j++) memcpy(interior + (size_t)j * NX, h + (size_t)(j + nd) * nx + nd, sizeof(float) * NX);
    free(h);
    return rc;
}

int dropin_forward_run(const char *in_path, const char *out_path)
{
    FILE *in = fopen(in_path, "rb");
    if (!in) FAIL("cannot open input");
    int hdr[9];
    float fl[6];
    if (fread(hdr, sizeof(int), 9, in) != 9 || fread(fl, sizeof(float), 6, in) != 6) FAIL("short input");
    ifos_config c;
    memset(&c, 0, sizeof c);
    c.NX = hdr[0]; c.NY = hdr[1]; c.NT = hdr[2]; c.FDORDER = hdr[3]; c.FW = hdr[4]; c.FREE_SURF = hdr[5];
    c.VELOCITY = hdr[6];
    const int nsrc = hdr[7], ntr = hdr[8];
    c.DT = fl[0]; c.DH = fl[1]; c.VPPML = fl[2]; c.FPML = fl[3]; c.npower = fl[4]; c.k_max_PML = fl[5];
    c.MAXRELERROR = 1; c.PARAMETERIZATION = 1; c.DTINV = 1; c.ADJOINT_TYPE = 1; c.SOURCE_SHAPE = 1;
    c.SOURCE_TYPE = 1;
    float *src = malloc(sizeof(float) * 8 * nsrc);
    int *rec = malloc(sizeof(int) * 2 * ntr);
    const size_t n = (size_t)c.NX * c.NY;
    float *vp = malloc(sizeof(float) * n), *vs = malloc(sizeof(float) * n), *rho = malloc(sizeof(float) * n);
    if (fread(src, sizeof(float), 8 * (size_t)nsrc, in) != 8 * (size_t)nsrc ||
        fread(rec, sizeof(int), 2 * (size_t)ntr, in) != 2 * (size_t)ntr || fread(vp, sizeof(float), n, in) != n ||
        fread(vs, sizeof(float), n, in) != n || fread(rho, sizeof(float), n, in) != n)
        FAIL("short input");
    fclose(in);

    /* globals the kernels read (src/globvar.h) */
    NX = c.NX; NY = c.NY; NT = c.NT; FDORDER = c.FDORDER; FW = c.FW; FREE_SURF = c.FREE_SURF; BOUNDARY = 0;
    L = 0; PARAMETERIZATION = 1; VELOCITY = c.VELOCITY; ADJOINT_TYPE = 1; SOURCE_TYPE = 1; SEISMO = 1;
    DT = c.DT; DH = c.DH;
    const int nd = FDORDER / 2 + 1, fw2 = 2 * FW;

    /* device arrays with the reference's bounds (src/IFOS2D.c:371-641) */
    float **pvx = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **pvy = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **psxx = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **psyy = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **psxy = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **ux = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **uy = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **uxy = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **uyx = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **pvxp1 = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **pvyp1 = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **pvxm1 = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **pvym1 = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **ppi = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **pu = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **prho = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **prip = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **prjp = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **puipjp = ifos_dev_matrix(-nd + 1, NY + nd, -nd + 1, NX + nd);
    float **psi_sxx_x = ifos_dev_matrix(1, NY, 1, fw2), **psi_sxy_x = ifos_dev_matrix(1, NY, 1, fw2);
    float **psi_vxx = ifos_dev_matrix(1, NY, 1, fw2), **psi_vyx = ifos_dev_matrix(1, NY, 1, fw2);
    float **psi_vxxs = ifos_dev_matrix(1, NY, 1, fw2);
    float **psi_syy_y = ifos_dev_matrix(1, fw2, 1, NX), **psi_sxy_y = ifos_dev_matrix(1, fw2, 1, NX);
    float **psi_vyy = ifos_dev_matrix(1, fw2, 1, NX), **psi_vxy = ifos_dev_matrix(1, fw2, 1, NX);

    /* host setup: averaged materials (av_rho / av_mue on the periodic halo of
     * matcopy_elastic), Holberg coefficients, CPML profiles, wavelet */
    ifos_materials m;
    if (ifos_setup_materials(&c, vp, vs, rho, NULL, NULL, &m)) FAIL("materials");
    if (upload(ppi, vp, nd) || upload(pu, vs, nd) || upload(prho, rho, nd) || upload(prip, m.rip, nd) ||
        upload(prjp, m.rjp, nd) || upload(puipjp, m.uipjp, nd))
        FAIL("upload");
    ifos_free_materials(&m);
    float hc[7];
    ifos_holberg(FDORDER, 1, hc);
    float *prof = malloc(sizeof(float) * 12 * fw2);
    ifos_pml_profiles(&c, prof);
    float *K[12];
    for (int k = 0; k < 12; k++) K[k] = prof + (size_t)k * fw2 - 1; /* 1-based vectors */
    float *w = malloc(sizeof(float) * (size_t)nsrc * NT);
    ifos_wavelet(&c, nsrc, src, w);
    float **srcpos_loc = host_matrix(1, 8, 1, nsrc);
    float **signals = host_matrix(1, nsrc, 1, NT);
    for (int l = 1; l <= nsrc; l++) {
        for (int k = 1; k <= 8; k++) srcpos_loc[k][l] = src[(l - 1) * 8 + k - 1];
        for (int t = 1; t <= NT; t++) signals[l][t] = w[(size_t)(l - 1) * NT + t - 1];
    }
    int **recpos = host_imatrix(1, 3, 1, ntr);
    for (int t = 1; t <= ntr; t++) {
        recpos[1][t] = rec[2 * (t - 1)];
        recpos[2][t] = rec[2 * (t - 1) + 1];
        recpos[3][t] = t;
    }
    float **sectionvx = host_matrix(1, ntr, 1, NT), **sectionvy = host_matrix(1, ntr, 1, NT);

    /* the forward loop of src/IFOS2D.c:1651-1803 */
    zero_fdveps(-nd + 1, NY + nd, -nd + 1, NX + nd, pvx, pvy, NULL, psxx, psyy, psxy, NULL, NULL, pvxm1, pvym1, uxy,
                pvxp1, pvyp1, psi_sxx_x, psi_sxy_x, NULL, psi_vxx, psi_vyx, NULL, psi_syy_y, psi_sxy_y, NULL, psi_vyy,
                psi_vxy, NULL, psi_vxxs);
    for (int nt = 1; nt <= NT; nt++) {
        update_v_PML(1, NX, 1, NY, nt, pvx, pvxp1, pvxm1, pvy, pvyp1, pvym1, NULL, NULL, psxx, psyy, psxy, prip, prjp,
                     srcpos_loc, signals, signals, nsrc, NULL, hc, 0, 0, K[0], K[1], K[2], K[3], K[4], K[5], K[6],
                     K[7], K[8], K[9], K[10], K[11], psi_sxx_x, psi_syy_y, psi_sxy_y, psi_sxy_x);
        update_s_elastic_PML(1, NX, 1, NY, pvx, pvy, ux, uy, uxy, uyx, psxx, psyy, psxy, ppi, pu, puipjp, NULL, prho,
                             hc, 0, K[0], K[1], K[2], K[3], K[4], K[5], K[6], K[7], K[8], K[9], K[10], K[11], psi_vxx,
                             psi_vyy, psi_vxy, psi_vyx);
        if (SOURCE_TYPE == 1) psource(nt, psxx, psyy, NULL, srcpos_loc, signals, nsrc, 0);
        if (FREE_SURF)
            surface_elastic_PML(1, pvx, pvy, psxx, psyy, psxy, NULL, ppi, pu, prho, hc, K[0], K[1], K[2], psi_vxxs,
                                ux, uy, uxy, NULL, NULL, NULL);
        seismo_ssg(nt, ntr, recpos, sectionvx, sectionvy, NULL, NULL, NULL, NULL, pvx, pvy, NULL, psxx, psyy, NULL,
                   ppi, pu, hc);
    }
    if (ifos_dev_sync()) FAIL("sync");

    FILE *out = fopen(out_path, "wb");
    if (!out) FAIL("cannot open output");
    for (int t = 1; t <= ntr; t++) fwrite(&sectionvx[t][1], sizeof(float), NT, out);
    for (int t = 1; t <= ntr; t++) fwrite(&sectionvy[t][1], sizeof(float), NT, out);
    float *fin = malloc(sizeof(float) * n);
    float **fields[5] = {pvx, pvy, psxx, psyy, psxy};
    for (int k = 0; k < 5; k++) {
        if (download(fields[k], fin, nd)) FAIL("download");
        fwrite(fin, sizeof(float), n, out);
    }
    fclose(out);
    void *all[] = {pvx, pvy, psxx, psyy, psxy, ux, uy, uxy, uyx, pvxp1, pvyp1, pvxm1, pvym1, ppi, pu, prho, prip, prjp,
                   puipjp, psi_sxx_x, psi_sxy_x, psi_vxx, psi_vyx, psi_vxxs, psi_syy_y, psi_sxy_y, psi_vyy, psi_vxy};
    for (size_t k = 0; k < sizeof all / sizeof all[0]; k++) ifos_dev_free_array(all[k]);
    free(fin);
    printf("dropin_forward: %d x %d, %d steps, %d receivers done\n", NX, NY, NT, ntr);
    return 0;
}

#ifdef DROPIN_MAIN
int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: dropin_forward IN OUT\n");
        return 2;
    }
    return dropin_forward_run(argv[1], argv[2]);
}
#endif
