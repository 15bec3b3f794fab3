// ifos_step.h -- argument block of the fused time-step kernels.
#pragma once
#include "ifos_internal.h"

namespace ifos {

constexpr int kTileX = 64;  // one wavefront across x
constexpr int kTileY = 8;   // >= FDORDER/2 + 1 for the free surface

// points of one kind bucketed by 64x8 tile: the entries of tile b are
// idx[start[b] .. start[b+1]) (indices into ix/iy/...), in the caller's order
struct TileList {
    const int *start, *idx;
    const int *ix, *iy, *type;
    const double *sin_azi, *cos_azi;
    const float *sig, *sig1;   // [n][NT]
};

struct StepArgs {
    StepGeom g;
    int L, VELOCITY, ADJOINT_TYPE, NT, DTINV, sw, explosive;
    long plane;                // NX*NY
    float *vx, *vy, *sxx, *syy, *sxy;
    float *vxp1, *vyp1;        // adjoint only
    const float *rip, *rjp;
    const float *f, *gl, *uipjp;                                 // elastic
    const float *mu, *pi, *muipjp, *taus, *taup, *tausipjp;      // viscoelastic
    float *r, *p, *q;
    long ls;                   // plane stride of the memory variables
    float eta[kMaxL + 1], bip[kMaxL + 1], bjm[kMaxL + 1], cip[kMaxL + 1], cjm[kMaxL + 1];
    float *psi_sxx_x, *psi_sxy_x, *psi_syy_y, *psi_sxy_y;
    float *psi_vxx, *psi_vyx, *psi_vyy, *psi_vxy, *psi_vxxs;
    TileList inj, exp, rec;
    float *snap[5];            // rho_x, rho_y, x, y, u : [nsnap][NY][NX]; null: no snapshots
    const float *muss, *lamss, *vs_i;
    float *wc_rho, *wc, *wc_u;
    float *seis_vx, *seis_vy;
    const int *base;           // steps completed before the current graph unit
    int *nan_flag;
    int nbx, nby;              // tiles in x and y
    int swizzle;               // 1: XCD-contiguous tile order (see tile_of)
};

void launch_vel_step(const StepArgs &a, int k, hipStream_t s);
void launch_str_step(const StepArgs &a, int k, hipStream_t s);
void launch_tick(int *base, int by, hipStream_t s);

}  // namespace ifos
