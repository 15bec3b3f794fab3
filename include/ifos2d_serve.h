/*
 * ifos2d_serve.h -- request server over the native engine (no counterpart
 * in the reference, which runs one inversion per process): forward-modelling
 * and gradient requests of different grids and shot counts share one GPU.
 *
 *  - Admission: a request runs when the device bytes of its engine
 *    (ifos_engine_bytes: wavefields, materials, CPML memory, traces and, for
 *    gradients, snapshots) fit in the HBM budget next to every engine alive.
 *  - Residency: finished engines stay allocated and are reused by later
 *    requests of the same configuration, snapshot precision and slot count
 *    (and keep their model when the model id matches); idle engines are
 *    evicted, least recently used first, when a new request needs the room.
 *  - Continuous batching: work is admitted as launches of up to max_slots
 *    shots on one engine, and of at most max(2, ifos_engine_mall_slots(cfg))
 *    shots for a grid whose wider launches would not stay in the MALL.  A
 *    launch takes the oldest waiting request's next shots and fills its free
 *    slots with the next shots of later requests of the same configuration,
 *    kind, snapshot precision and model id (!= 0), so requests that arrive
 *    while others run join the next launch of their kind.  Worker threads each drive their own engine and stream, so
 *    launches of different requests overlap on the device.
 *  - Order: first come first served, except that requests behind one that
 *    does not fit may start while it waits; after 8 such overtakes the
 *    waiting request is served before any other.
 *
 * Results are bit-identical to running the request alone on one engine:
 * the kernels are deterministic, each shot's gradient is taken from its own
 * slot (ifos_engine_slot_gradient), and a request's shots are added to its
 * sum in shot order in FP32, as a one-slot engine adds them.
 * Errors: 0 on success, negative on failure, message in ifos_last_error().
 */
#ifndef IFOS2D_SERVE_H
#define IFOS2D_SERVE_H

#include <stddef.h>

#include "ifos2d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { IFOS_REQ_FORWARD = 0, IFOS_REQ_GRADIENT = 1 };

/* One request.  Arrays are host memory and are copied at submission.
 * nsrc[k], ntr[k]: sources and receivers of shot k; srcpos, recpos: all
 * shots' rows concatenated (as ifos_engine_set_shots).  Gradient requests
 * give observed traces [sum ntr][NT] and optionally a kill mask. */
typedef struct ifos_request {
    ifos_config cfg;
    int kind;                 /* IFOS_REQ_FORWARD or IFOS_REQ_GRADIENT      */
    int snap_prec;            /* gradient requests: 0 FP32, 1 BF16, 2 FP16  */
    int nshots;
    const int *nsrc;
    const float *srcpos;
    const int *ntr;
    const int *recpos;
    const float *vp, *vs, *rho, *taus, *taup;   /* [NY][NX]                  */
    long long model_id;       /* != 0: same id = same model, may stay resident */
    const float *obs_vx, *obs_vy;
    const int *kill;
} ifos_request;

/* Outputs of a finished request; pointers may be NULL to skip a field. */
typedef struct ifos_result {
    float *seis_vx, *seis_vy;             /* forward: [sum ntr][NT]           */
    float *g_vp, *g_vs, *g_rho;           /* gradient: [NY][NX], shot sum     */
    double *misfit;                       /* gradient: [nshots][2] L2, energy */
    double times[3];                      /* submitted, started, finished (s) */
    int slots;                            /* most of its shots in one launch  */
    int reused;                           /* 1: it started on a resident engine */
    int shared;                           /* its launches that also carried
                                             shots of other requests         */
} ifos_result;

typedef struct ifos_server_stats {
    long long submitted, finished, failed;
    long long engines_created, engines_reused, engines_evicted;
    size_t budget, bytes_in_use, peak_bytes;
    long long launches, shared_launches;  /* engine launches; those carrying
                                             shots of several requests      */
} ifos_server_stats;

typedef struct ifos_server ifos_server;

/* workers: concurrent requests; max_slots: shots per launch per request. */
int  ifos_server_create(int device, size_t hbm_budget, int workers, int max_slots, ifos_server **out);
/* Finishes the admitted requests, drops the queued ones, frees everything. */
void ifos_server_destroy(ifos_server *s);
/* Queue a request; *id identifies it.  Fails at once when the request can
 * never fit in the budget. */
int  ifos_server_submit(ifos_server *s, const ifos_request *r, long long *id);
/* Wait up to timeout_s seconds (< 0: forever): 0 finished, 1 still
 * running, -5 the request failed (message in ifos_last_error()). */
int  ifos_server_wait(ifos_server *s, long long id, double timeout_s);
/* Copy a finished request's outputs, then forget it. */
int  ifos_server_result(ifos_server *s, long long id, ifos_result *out);
int  ifos_server_stats_get(ifos_server *s, ifos_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif
