/*
 * engine_gradient.c -- a C client of the native engine (include/ifos2d_hip.h)
 * and of the source-time-function calls (include/ifos2d_stf.h), making the
 * calls that replace the shot loop of src/IFOS2D.c:1076-2724 and the STF
 * step of src/stf.c:24.
 *
 * engine_gradient_run(IN, OUT): one engine, the model set once, then per
 * shot ifos_engine_set_shot + ifos_engine_shot_gradient (the misfits added
 * up as the reference adds L2 and energy over shots), and the gradient summed
 * over the shots (src/IFOS2D.c:2703-2713).
 *   IN : ifos_config (its C layout); int nshots;
 *        per shot: int nsrc, ntr; float srcpos[nsrc][IFOS_SRC_STRIDE];
 *        int recpos[ntr][2]; float obs_vx[ntr][NT], obs_vy[ntr][NT];
 *        float vp[NY][NX], vs[NY][NX], rho[NY][NX] (+ taus, taup when L > 0)
 *   OUT: double misfit[2]; float g_vp[NY][NX], g_vs[NY][NX], g_rho[NY][NX]
 *
 * engine_stf_run(IN, PARA, OUT): for the first shot of IN, the synthetics
 * with the generated wavelet, the fdlsq correction filter from the observed
 * and synthetic vertical traces, the wavelet convolved with it (conv_FD),
 * that wavelet given to the engine (SOURCE_SHAPE 7), and the forward run
 * with it.
 *   OUT: float stf[NT], new_wavelet[NT], seis_vy[ntr][NT]
 *
 * Built as a shared library that the tests call in-process.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ifos2d_hip.h"
#include "ifos2d_stf.h"

typedef struct shot {
    int nsrc, ntr;
    float *srcpos, *obs_vx, *obs_vy;
    int *recpos;
} shot;

typedef struct input {
    ifos_config cfg;
    int nshots;
    shot *shots;
    float *m[5];  /* vp vs rho taus taup */
} input;

static int read_all(FILE *f, void *p, size_t n)
{
    return fread(p, 1, n, f) == n ? 0 : -1;
}

static void free_input(input *in)
{
    if (in->shots)
        for (int k = 0; k < in->nshots; k++) {
            free(in->shots[k].srcpos);
            free(in->shots[k].recpos);
            free(in->shots[k].obs_vx);
            free(in->shots[k].obs_vy);
        }
    free(in->shots);
    for (int k = 0; k < 5; k++) free(in->m[k]);
    memset(in, 0, sizeof *in);
}

static int read_input(const char *path, input *in)
{
    memset(in, 0, sizeof *in);
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int rc = read_all(f, &in->cfg, sizeof in->cfg) || read_all(f, &in->nshots, sizeof(int));
    if (!rc && (in->nshots < 1 || in->nshots > 1024)) rc = -1;
    const size_t NT = (size_t)in->cfg.NT, n = (size_t)in->cfg.NX * in->cfg.NY;
    if (!rc) in->shots = calloc((size_t)in->nshots, sizeof(shot));
    for (int k = 0; !rc && k < in->nshots; k++) {
        shot *s = &in->shots[k];
        rc = read_all(f, &s->nsrc, sizeof(int)) || read_all(f, &s->ntr, sizeof(int));
        if (rc || s->nsrc < 0 || s->ntr < 1) {
            rc = -1;
            break;
        }
        s->srcpos = malloc(sizeof(float) * IFOS_SRC_STRIDE * (size_t)(s->nsrc ? s->nsrc : 1));
        s->recpos = malloc(sizeof(int) * 2 * (size_t)s->ntr);
        s->obs_vx = malloc(sizeof(float) * NT * (size_t)s->ntr);
        s->obs_vy = malloc(sizeof(float) * NT * (size_t)s->ntr);
        rc = read_all(f, s->srcpos, sizeof(float) * IFOS_SRC_STRIDE * (size_t)s->nsrc) ||
             read_all(f, s->recpos, sizeof(int) * 2 * (size_t)s->ntr) ||
             read_all(f, s->obs_vx, sizeof(float) * NT * (size_t)s->ntr) ||
             read_all(f, s->obs_vy, sizeof(float) * NT * (size_t)s->ntr);
    }
    const int nm = in->cfg.L > 0 ? 5 : 3;
    for (int k = 0; !rc && k < nm; k++) {
        in->m[k] = malloc(sizeof(float) * n);
        rc = read_all(f, in->m[k], sizeof(float) * n);
    }
    fclose(f);
    if (rc) free_input(in);
    return rc;
}

#define TRY(call)                                                                 \
    do {                                                                          \
        if ((call) != 0) {                                                        \
            fprintf(stderr, "engine_gradient: %s: %s\n", #call, ifos_last_error()); \
            rc = 1;                                                               \
            goto done;                                                            \
        }                                                                         \
    } while (0)

int engine_gradient_run(const char *in_path, const char *out_path)
{
    input in;
    if (read_input(in_path, &in)) return 2;
    const size_t n = (size_t)in.cfg.NX * in.cfg.NY;
    int rc = 0;
    ifos_engine *eng = NULL;
    float *g = malloc(sizeof(float) * 3 * n);
    double L2 = 0.0, energy = 0.0;
    TRY(ifos_engine_create(&in.cfg, 0, &eng));
    TRY(ifos_engine_set_model(eng, in.m[0], in.m[1], in.m[2], in.m[3], in.m[4]));
    TRY(ifos_engine_reset_gradient(eng));
    for (int k = 0; k < in.nshots; k++) {
        const shot *s = &in.shots[k];
        double misfit[2];
        TRY(ifos_engine_set_shot(eng, s->nsrc, s->srcpos, s->ntr, s->recpos));
        TRY(ifos_engine_shot_gradient(eng, s->obs_vx, s->obs_vy, NULL, 0, misfit));
        L2 += misfit[0];
        energy += misfit[1];
    }
    TRY(ifos_engine_gradient(eng, 1, g, g + n, g + 2 * n));
    {
        FILE *f = fopen(out_path, "wb");
        if (!f) {
            rc = 3;
            goto done;
        }
        const double mis[2] = {L2, energy};
        if (fwrite(mis, sizeof(double), 2, f) != 2 || fwrite(g, sizeof(float), 3 * n, f) != 3 * n) rc = 3;
        fclose(f);
    }
done:
    if (eng) ifos_engine_destroy(eng);
    free(g);
    free_input(&in);
    return rc;
}

int engine_stf_run(const char *in_path, const char *para, const char *out_path)
{
    input in;
    if (read_input(in_path, &in)) return 2;
    const int NT = in.cfg.NT;
    const shot *s = &in.shots[0];
    int rc = 0;
    ifos_engine *eng = NULL;
    ifos_stf_params p;
    float *wav = malloc(sizeof(float) * (size_t)NT * (size_t)(s->nsrc ? s->nsrc : 1));
    float *syn_vx = malloc(sizeof(float) * (size_t)NT * s->ntr), *syn_vy = malloc(sizeof(float) * (size_t)NT * s->ntr);
    float *stf = malloc(sizeof(float) * NT), *neww = malloc(sizeof(float) * NT);
    if (s->nsrc != 1) {
        fprintf(stderr, "engine_stf_run: one source per shot\n");
        rc = 4;
        goto done;
    }
    TRY(ifos_stf_parse(para, &p));
    TRY(ifos_engine_create(&in.cfg, 0, &eng));
    TRY(ifos_engine_set_model(eng, in.m[0], in.m[1], in.m[2], in.m[3], in.m[4]));
    TRY(ifos_engine_set_shot(eng, s->nsrc, s->srcpos, s->ntr, s->recpos));
    TRY(ifos_engine_forward(eng, 0));
    TRY(ifos_engine_seismograms(eng, syn_vx, syn_vy));
    /* initstfinvengine + runstfinvengine (src/stf.c:216-219), then conv_FD */
    TRY(ifos_stf_fdlsq(&p, s->ntr, NT, in.cfg.DT, s->obs_vy, syn_vy, NULL, stf, NULL));
    ifos_wavelet(&in.cfg, s->nsrc, s->srcpos, wav);
    TRY(ifos_conv_fd(wav, stf, neww, NT));
    TRY(ifos_engine_set_wavelets(eng, neww));
    TRY(ifos_engine_forward(eng, 0));
    TRY(ifos_engine_seismograms(eng, syn_vx, syn_vy));
    {
        FILE *f = fopen(out_path, "wb");
        if (!f) {
            rc = 3;
            goto done;
        }
        if (fwrite(stf, sizeof(float), NT, f) != (size_t)NT || fwrite(neww, sizeof(float), NT, f) != (size_t)NT ||
            fwrite(syn_vy, sizeof(float), (size_t)NT * s->ntr, f) != (size_t)NT * s->ntr)
            rc = 3;
        fclose(f);
    }
done:
    if (eng) ifos_engine_destroy(eng);
    free(wav);
    free(syn_vx);
    free(syn_vy);
    free(stf);
    free(neww);
    free_input(&in);
    return rc;
}
