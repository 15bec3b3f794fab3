// ifos_step.h -- argument block of the fused time-step kernels.
//
// The kernels advance S shot slots of one model at once.  Every array that
// belongs to a shot is laid out slot-major with a fixed stride, and the
// pointers below address slot 0; slot s adds s * (its stride).  Sources and
// receivers of all slots share global point arrays (seismograms and adjoint
// sources are [trace][NT] over all slots' traces), reached through per-slot,
// per-tile CSR lists.
#pragma once
#include "ifos_internal.h"

namespace ifos {

constexpr int kTileX = 64;  // one wavefront across x
constexpr int kTileY = 8;   // >= FDORDER/2 + 1 for the free surface

// points of one kind bucketed by slot and 64x8 tile: the entries of tile b of
// slot s are idx[start[s*(ntile+1) + b] .. start[s*(ntile+1) + b + 1])
// (indices into ix/iy/...), in the caller's order
struct TileList {
    const int *start, *idx;
    const int *ix, *iy, *type;
    const double *sin_azi, *cos_azi;
    const float *sig, *sig1;   // [n][NT]
};

struct StepArgs {
    StepGeom g;
    int L, VELOCITY, ADJOINT_TYPE, NT, DTINV, sw;
    long plane;                // NX*NY
    // per-slot padded grids (stride ws floats between slots)
    float *vx, *vy, *sxx, *syy, *sxy;
    float *vxp1, *vyp1;        // adjoint only
    float *r, *p, *q;
    long ws;
    long ls;                   // plane stride of the memory variables
    // materials, shared by all slots
    const float *rip, *rjp;
    const float *f, *gl, *uipjp;                                 // elastic
    const float *mu, *pi, *muipjp, *taus, *taup, *tausipjp;      // viscoelastic
    float eta[kMaxL + 1], bip[kMaxL + 1], bjm[kMaxL + 1], cip[kMaxL + 1], cjm[kMaxL + 1];
    // per-slot CPML memory (stride pss floats)
    float *psi_sxx_x, *psi_sxy_x, *psi_syy_y, *psi_sxy_y;
    float *psi_vxx, *psi_vyx, *psi_vyy, *psi_vxy, *psi_vxxs;
    long pss;
    // second wavefield set of the one-kernel step (k_fused_step): step nt
    // reads set (nt - 1) & 1 and writes set nt & 1 (set 0 is the one above)
    float *vx2, *vy2, *sxx2, *syy2, *sxy2;
    float *psi2_sxx_x, *psi2_sxy_x, *psi2_syy_y, *psi2_sxy_y;
    TileList inj, exp, rec;
    // the injection points (sources forward, receivers adjoint) of each tile
    // and of the N-cell rim around it, for k_fused_step
    TileList injx;
    const int *explosive;      // [slot]: 1 when the slot's sources act as explosive (forward)
    // forward snapshots rho_x, rho_y, x, y, u, each [nsnap][NY][NX] per slot:
    // FP32 in snap (snap_prec 0), or 16-bit codes in snap16 (1: BF16,
    // 2: FP16) of the value divided by a power-of-two scale, one scale per
    // field, slot and tile in sscale[field][slot][tile]; all null: none.
    // Slot strides: sns elements (snap / snap16), scs floats (sscale).
    float *snap[5];
    uint16_t *snap16[5];
    float *sscale[5];
    int snap_prec;
    long sns, scs;
    // correlation
    const float *muss, *lamss, *vs_i;
    float *wc_rho, *wc, *wc_u;  // per slot, stride wcs floats
    long wcs;
    float *seis_vx, *seis_vy;   // [trace][NT], traces of all slots
    float *wacc;                // energy accumulator of the preconditioner (per slot, stride wacs), or null
    double *wacc64;             // the same in FP64 (products formed in FP64), used instead when set
    long wacs;
    const int *base;            // steps completed before the current graph unit
    int *nan_flag;
    int nbx, nby;               // tiles in x and y
    int nslots;                 // active shot slots
    int swizzle;                // 1: XCD-contiguous work order (see work_range)
    int ngroups;                // slot groups per tile (set by the launcher)
    int slot_major;             // work order: 0 all slot groups of a tile, then the next tile;
                                // 1 all tiles of a slot group, then the next group
    // division by nbx, by ngroups and by the tile count as multiply-high and
    // shift (set by the launcher): the quotients of every work item without
    // integer division
    uint32_t nbx_mul, nbx_sh, ng_mul, ng_sh, nt_mul, nt_sh;
    // with swizzle: the work items of XCD x are [xlo[x], xlo[x+1]), runs of
    // equal estimated cost (set by the launcher)
    int xlo[9];
};

// launch shape of the step kernels: spb slots per workgroup (1, 2 or 4; the
// materials of a cell are loaded once for all of them), when grid > 0 a
// persistent grid of that many workgroups (a multiple of 8) looping over the
// work items (grid == 0 launches one workgroup per work item), and the order
// of the work items
struct LaunchShape {
    int spb = 1;
    int grid = 0;    // stress step
    int grid_v = 0;  // velocity step
    int order = 0;   // StepArgs::slot_major
    int rows_v = 1;  // velocity step: cells per thread along y (1, 2 or 4: a 64x(8/rows) block)
    int lds_v = 0;   // velocity step: two rows per thread with the tile's stresses staged in LDS
    float cpml_weight = 1.0f;  // extra cost of a tile in an absorbing strip, for the XCD split
};

// a: host copy of the argument block (grid sizes), da: the same block in
// device memory, which the kernels read (kept off the kernel-argument
// registers, whose uniform values would not all fit)
void launch_vel_step(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s);
void launch_str_step(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s);
void launch_tick(int *base, int by, hipStream_t s);
// one kernel per time step (velocity and stress), spb 1; ls.grid as for the stress step
void launch_fused_step(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s);
// the argument block as the kernels of launch shape ls must see it (slot
// groups and the division constants filled in): upload this one
StepArgs shaped_args(const StepArgs &a, const LaunchShape &ls);

}  // namespace ifos
