// ifos_step.hip -- fused time-step kernels of the native engine.
//
// One time step is two kernels:
//   k_vel_step   velocity update (src/update_v_PML.c) + the point forces or
//                adjoint residual injection that follow it (:553-578) + the
//                acceleration snapshot (forward) or the adjoint vxp1 update
//   k_str_step   stress update (src/update_s_*_PML.c) + explosive source
//                (src/psource.c) + free surface (src/surface_*_PML.c) +
//                snapshot store (src/IFOS2D.c:1805-1871) or zero-lag
//                correlation (:2381-2419) + receiver sampling
//                (src/seismo_ssg.c) + stability probe
// Each launch advances every active shot slot (one shot of the same model per
// slot).  Work items are (tile, group of spb slots) pairs; a thread owns one
// cell of a 64x8 tile and loops over the slots of its group, loading the
// cell's materials once.  Point sources and receivers are bucketed per slot
// and tile on the host (CSR lists in source order), so a thread only scans
// the few points of its own tile.  The free surface needs the new
// stresses of rows 2..N+1 of its column: tiles are 8 rows high and N <= 6, so
// the top tile row holds them; they are exchanged through LDS after a barrier.
// The arithmetic is the per-cell code of ifos_cellmath.h, bit-identical to
// the reference.
#include <algorithm>
#include <vector>

#include "ifos_cellmath.h"
#include "ifos_step.h"

namespace ifos {

__device__ __forceinline__ int step_number(const StepArgs &a, int k) { return a.base ? *gmem(a.base) + k + 1 : k; }

// Per-launch step flags, decided on the host: kSnapStep for a forward step
// that stores a snapshot, (nt-1) % DTINV == 0 (src/IFOS2D.c:1805); kCorrStep
// for an adjoint step that correlates, (NT-nt) % DTINV == 0
// (src/IFOS2D.c:2381).  Inside a graph unit (a.base set) step nt is
// base + k + 1 with base a multiple of the unit, and the unit a multiple of
// DTINV dividing NT, so both tests depend on k alone.
// The other flags restate fields of the argument block that every wave
// would otherwise load before it can branch: kForward (sw == 0), kAdjInt
// (adjoint with VELOCITY == 0: the integrated adjoint velocity,
// src/update_v_PML.c:140-143) and kEnergy (the preconditioner's energy sums).
constexpr int kSnapStep = 1, kCorrStep = 2, kForward = 4, kAdjInt = 8, kEnergy = 16;

static int step_flags(const StepArgs &a, int k)
{
    const int D = a.DTINV > 0 ? a.DTINV : 1;
    int f = 0;
    if (a.sw == 0 && (a.snap[0] || a.snap16[0]) && (a.base ? k % D == 0 : (k - 1) % D == 0)) f |= kSnapStep;
    if (a.sw == 1 && (a.base ? (k + 1) % D == 0 : (a.NT - k) % D == 0)) f |= kCorrStep;
    if (a.sw == 0) f |= kForward;
    if (a.sw == 1 && a.VELOCITY == 0) f |= kAdjInt;
    if (a.wacc || a.wacc64) f |= kEnergy;
    return f;
}

// a step whose kernels need no more than the update, the sources and the
// receivers: most steps of a run (every step but each DTINV-th of the
// forward run with snapshots and of the adjoint run with VELOCITY 1)
static bool lean_step(int flags) { return (flags & ~kForward) == 0; }

// Workgroups are dealt round-robin to the 8 XCDs (workgroup h runs on XCD
// h % 8), and each XCD has its own L2.  With swizzle on, XCD x works through
// one contiguous run [xlo[x], xlo[x+1]) of the work items in row-major tile
// order, so the halos neighbouring tiles share, and the materials the slot
// groups of a tile share, are mostly fetched into one L2 once.  The runs have
// equal estimated cost rather than equal length (shaped_args).

// n / d for 0 <= n < 2^31 from (mul, sh) of make_div: t = mulhi(n, mul),
// q = (((n - t) >> 1) + t) >> sh; sh = 0 with mul = 0 marks d = 1
__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint32_t mul, uint32_t sh)
{
    if (mul == 0) return n;
    const uint32_t t = __umulhi(n, mul);
    return (((n - t) >> 1) + t) >> sh;
}

// The work-item mapping of a launch, read from the argument block into
// locals once per workgroup, so that these loads issue together instead of
// one by one where they are first used.
struct ItemMap {
    int nbx, ntiles, nslots, ngroups, slot_major, swizzle;
    uint32_t nbx_mul, nbx_sh, ng_mul, ng_sh, nt_mul, nt_sh;
    int lo, hi;  // this workgroup's XCD run of items
};

__device__ __forceinline__ ItemMap item_map(const StepArgs &a)
{
    ItemMap m;
    m.nbx = a.nbx;
    m.ntiles = a.nbx * a.nby;
    m.nslots = a.nslots;
    m.ngroups = a.ngroups;
    m.slot_major = a.slot_major;
    m.swizzle = a.swizzle;
    m.nbx_mul = a.nbx_mul;
    m.nbx_sh = a.nbx_sh;
    m.ng_mul = a.ng_mul;
    m.ng_sh = a.ng_sh;
    m.nt_mul = a.nt_mul;
    m.nt_sh = a.nt_sh;
    const int x = blockIdx.x % 8;
    m.lo = a.xlo[x];
    m.hi = a.xlo[x + 1];
    return m;
}

// tile and first slot of work item it
__device__ __forceinline__ void item_of(const ItemMap &m, int it, int spb, int &blk, int &s0)
{
    if (m.slot_major) {
        const int grp = (int)fdiv(it, m.nt_mul, m.nt_sh);
        blk = it - grp * m.ntiles;
        s0 = grp * spb;
    } else {
        blk = (int)fdiv(it, m.ng_mul, m.ng_sh);
        s0 = (it - blk * m.ngroups) * spb;
    }
}

// items [first, end) with the given stride for this workgroup: one item per
// workgroup, or a persistent grid (gridDim.x a multiple of 8) looping over
// the items of its XCD's run
__device__ __forceinline__ void work_range(const ItemMap &m, int nitems, bool persistent, int &first, int &end,
                                           int &stride)
{
    const int h = blockIdx.x;
    int lo = 0, hi = nitems;
    if (m.swizzle) {
        lo = m.lo;
        hi = m.hi;
    }
    if (!persistent) {
        first = m.swizzle ? lo + h / 8 : h;
        end = m.swizzle ? (first < hi ? first + 1 : first) : first + 1;
        stride = 1;
    } else if (m.swizzle) {
        first = lo + h / 8;
        end = hi;
        stride = gridDim.x / 8;
    } else {
        first = h;
        end = nitems;
        stride = gridDim.x;
    }
}

// Memory variables r_l, p_l, q_l of a slot are interleaved per 64 cells of
// the padded grid: component k (r_l: l, p_l: L + l, q_l: 2L + l) of cell c
// sits at ((c - 1) / 64 * 3L + k) * 64 + (c - 1) % 64 from the slot's block
// of 3L fields.  The pitch is a multiple of 64 and a tile's cells start at
// column 1, so the 64 cells of a wavefront's row are one chunk and all 3L
// components of it one contiguous run: two streams (read, write) instead of
// 3L each, which DRAM serves faster (a copy with 32 separate arrays ran at
// 5.2 TB/s, the same bytes interleaved at 5.7 TB/s).  Cells are interior
// (c >= pitch + 1), so the index stays inside the block.
__device__ __forceinline__ long mv_index(long c, int k, int L3)
{
    const long cc = c - 1;
    return ((cc >> 6) * L3 + k) * 64 + (cc & 63);
}

// ---------------------------------------------------------------------------
// 16-bit snapshots.  A tile's values of one field are divided by 2^e, the
// smallest power of two above their largest magnitude (exact in binary
// floating point), and rounded to nearest-even BF16 or FP16.
__device__ __forceinline__ bool has_snap(const StepArgs &a) { return a.snap[0] || a.snap16[0]; }

__device__ __forceinline__ int scale_exp(float m)
{
    if (!(m > 0.0f) || isinf(m)) return 0;
    return min(max(ilogbf(m) + 1, -120), 120);
}

__device__ __forceinline__ uint16_t enc16(float x, float inv_scale, int prec)
{
    const float y = x * inv_scale;
    if (prec == 1) {
        uint32_t u = __float_as_uint(y);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    }
    return __builtin_bit_cast(uint16_t, (_Float16)y);
}

__device__ __forceinline__ float dec16(uint16_t h, float scale, int prec)
{
    if (prec == 1) return __uint_as_float((uint32_t)h << 16) * scale;
    return (float)__builtin_bit_cast(_Float16, h) * scale;
}

// largest |v| over the block of NW waves (blockDim.y == NW), for NF values at
// once; every thread of the block must call it (it holds two barriers)
template <int NF, int NW = kTileY>
__device__ __forceinline__ void block_absmax(float (&v)[NF], float (*lds)[kTileY])
{
#pragma unroll
    for (int f = 0; f < NF; f++) {
        float m = fabsf(v[f]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if (threadIdx.x == 0) lds[f][threadIdx.y] = m;
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; f++) {
        float m = lds[f][0];
#pragma unroll
        for (int w = 1; w < NW; w++) m = fmaxf(m, lds[f][w]);
        v[f] = m;
    }
    __syncthreads();
}

// store NF snapshot fields (first field index f0) of one cell of slot s for
// snapshot slot k; with 16-bit storage every thread of the block must call it
template <int NF>
__device__ __forceinline__ void snap_store(const StepArgs &a, int s, int f0, long k, long e, int blk, bool live,
                                           const float (&val)[NF], float (*lds)[kTileY])
{
    const long o = (long)s * a.sns + k * a.plane + e;
    if (a.snap_prec == 0) {
        if (live)
#pragma unroll
            for (int f = 0; f < NF; f++) gmem(a.snap[f0 + f])[o] = val[f];
        return;
    }
    float m[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) m[f] = live ? val[f] : 0.0f;
    block_absmax<NF>(m, lds);
    const long ntile = (long)a.nbx * a.nby;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const int ex = scale_exp(m[f]);
        if (live) gmem(a.snap16[f0 + f])[o] = enc16(val[f], ldexpf(1.0f, -ex), a.snap_prec);
        if (threadIdx.x == 0 && threadIdx.y == 0) gmem(a.sscale[f0 + f])[(long)s * a.scs + k * ntile + blk] = ldexpf(1.0f, ex);
    }
}

// snap_store for the R cells (rows j .. j+R-1, interior indices e + r*NX) of
// a thread in a 64x(8/R) block of R-row threads; the 16-bit scale is the
// same 64x8 tile maximum as in snap_store
template <int NF, int R>
__device__ __forceinline__ void snap_store_rows(const StepArgs &a, int s, int f0, long k, long e, int blk,
                                                const bool (&live)[R], const float (&val)[R][NF],
                                                float (*lds)[kTileY])
{
    const long o = (long)s * a.sns + k * a.plane + e;
    const int NX = a.g.NX;
    if (a.snap_prec == 0) {
#pragma unroll
        for (int r = 0; r < R; r++)
            if (live[r])
#pragma unroll
                for (int f = 0; f < NF; f++) gmem(a.snap[f0 + f])[o + r * NX] = val[r][f];
        return;
    }
    float m[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        m[f] = live[0] ? fabsf(val[0][f]) : 0.0f;
#pragma unroll
        for (int r = 1; r < R; r++) m[f] = fmaxf(m[f], live[r] ? fabsf(val[r][f]) : 0.0f);
    }
    block_absmax<NF, kTileY / R>(m, lds);
    const long ntile = (long)a.nbx * a.nby;
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const int ex = scale_exp(m[f]);
        const float inv = ldexpf(1.0f, -ex);
#pragma unroll
        for (int r = 0; r < R; r++)
            if (live[r]) gmem(a.snap16[f0 + f])[o + r * NX] = enc16(val[r][f], inv, a.snap_prec);
        if (threadIdx.x == 0 && threadIdx.y == 0) gmem(a.sscale[f0 + f])[(long)s * a.scs + k * ntile + blk] = ldexpf(1.0f, ex);
    }
}

__device__ __forceinline__ float snap_load(const StepArgs &a, int s, int f, long k, long e, int blk)
{
    const long o = (long)s * a.sns + k * a.plane + e;
    if (a.snap_prec == 0) return gmem(a.snap[f])[o];
    const float sc = gmem(a.sscale[f])[(long)s * a.scs + k * (long)a.nbx * a.nby + blk];
    return dec16(gmem(a.snap16[f])[o], sc, a.snap_prec);
}

// ---------------------------------------------------------------------------
// Periodic sides (BOUNDARY 1).  On one rank the reference's exchanges after
// each half step copy the opposite edge into the halo columns
// (src/exchange_v.c:143-228, src/exchange_s.c:145-229): after the velocity
// step vy to columns NX+1..NX+N and 2-N..0, vx to NX+1..NX+N-1 and 1-N..0;
// after the stress step (sources and free surface included) sxx to
// NX+1..NX+N and 2-N..0, sxy to NX+1..NX+N-1 and 1-N..0.  The cell that owns
// the final value also stores it into its halo copy.  No kernel of the same
// launch reads those halo columns: the velocity step reads stress halos, the
// stress step velocity halos.  The padded rows hold at least 16 columns left
// of column 1 and N+1 right of column NX, so the copies stay in their row.
template <int N, typename Ptr>
__device__ __forceinline__ void wrap_pair(const StepGeom &g, Ptr a, Ptr b, long c, int i, float av, float bv)
{
    // a: the field with N copies on the right and N-1 on the left (vy, sxx),
    // b: N-1 on the right and N on the left (vx, sxy)
    if (i <= N) {
        a[c + g.NX] = av;
        if (i <= N - 1) b[c + g.NX] = bv;
    }
    if (i >= g.NX - N + 1) {
        b[c - g.NX] = bv;
        if (i >= g.NX - N + 2) a[c - g.NX] = av;
    }
}

// velocity step of one cell of slot s
template <int N>
__device__ __forceinline__ void vel_cell(const StepArgs &a, int s, int blk, int i, int j, bool live, long c, long e,
                                         int nt, int flags, float rip, float rjp, float (*s_red)[kTileY])
{
    const StepGeom &g = a.g;
    const int P = g.pitch;
    const long wo = (long)s * a.ws, po = (long)s * a.pss;
    const auto vxs = gmem(a.vx) + wo, vys = gmem(a.vy) + wo;
    float sxx_x = 0.0f, sxy_x = 0.0f, sxy_y = 0.0f, syy_y = 0.0f, vx0 = 0.0f, vy0 = 0.0f;
    if (live) {
        stress_derivs<N>(g.hc, a.sxx + wo + c, a.syy + wo + c, a.sxy + wo + c, P, sxx_x, sxy_x, sxy_y, syy_y);
        cpml_velocity(g, j, i, a.psi_sxx_x + po, a.psi_sxy_x + po, a.psi_syy_y + po, a.psi_sxy_y + po, sxx_x, sxy_x,
                      sxy_y, syy_y);
        vx0 = vxs[c];
        vy0 = vys[c];
    }
    // bounds of this tile's point-force / residual list, loaded behind the
    // stencil loads so that their latency overlaps them
    const int so = __builtin_amdgcn_readfirstlane(s * (a.nbx * a.nby + 1) + blk);  // uniform
    const int q0 = cmem(a.inj.start)[so], q1 = cmem(a.inj.start)[so + 1];
    if (flags & kForward) {
        if (flags & kSnapStep) {
            const float val[2] = {vel_rate(rip, sxx_x + sxy_y, g.DH), vel_rate(rjp, sxy_x + syy_y, g.DH)};
            snap_store<2>(a, s, 0, (nt - 1) / a.DTINV, e, blk, live, val, s_red);
        }
    } else if (!live) {
    } else if (flags & kAdjInt) {
        gmem(a.vxp1)[wo + c] += vx0 * g.DT;
        gmem(a.vyp1)[wo + c] += vy0 * g.DT;
    } else if (flags & kCorrStep) {
        gmem(a.vxp1)[wo + c] = vx0;
        gmem(a.vyp1)[wo + c] = vy0;
    }
    if (!live) return;
    float vx = vel_new(vx0, g.DT, rip, sxx_x + sxy_y, g.DH);
    float vy = vel_new(vy0, g.DT, rjp, sxy_x + syy_y, g.DH);

    // point forces (forward) / residual injection (adjoint) of this tile, in order
    // the point index is the same for the whole wave: the list, position and
    // signal reads are scalar loads, the cell match is per lane
    for (int q = q0; q < q1; q++) {
        const int l = cmem(a.inj.idx)[q];
        const int px = cmem(a.inj.ix)[l], py = cmem(a.inj.iy)[l];
        const float s0 = cmem(a.inj.sig)[(long)l * a.NT + nt - 1];
        const float s1 = cmem(a.inj.sig1)[(long)l * a.NT + nt - 1];
        if (px != i || py != j) continue;
        if (a.sw == 0) {
            const int type = cmem(a.inj.type)[l];
            if (type == 2) vx += (g.DT * rip * s1) / (g.DH * g.DH);
            if (type == 3) vy += (g.DT * rjp * s1) / (g.DH * g.DH);
            if (type == 4) {
                vx += (g.DT * rip * cmem(a.inj.sin_azi)[l] * s0) / (g.DH * g.DH);
                vy += (g.DT * rjp * cmem(a.inj.cos_azi)[l] * s0) / (g.DH * g.DH);
            }
        } else {
            if (a.ADJOINT_TYPE == 1) { vx += s0; vy += s1; }
            if (a.ADJOINT_TYPE == 2) vy += s1;
            if (a.ADJOINT_TYPE == 3) vx += s0;
        }
    }
    vxs[c] = vx;
    vys[c] = vy;
    if (g.BOUNDARY) wrap_pair<N>(g, vys, vxs, c, i, vy, vx);
    // energy of the new particle velocities (eprecond, src/eprecond.c:8-21)
    if (flags & kEnergy) {
        if (a.wacc64) {
            const auto W = gmem(a.wacc64) + (long)s * a.wacs;
            W[e] += ((double)vx * (double)vx) + ((double)vy * (double)vy);
        } else {
            const auto W = gmem(a.wacc) + (long)s * a.wacs;
            W[e] += ((vx * vx) + (vy * vy));
        }
    }
}

// velocity step of the R vertically adjacent cells (rows j .. j+R-1) of an
// R-row thread, slot s; the same arithmetic as vel_cell for each cell.  All
// loads of the R cells issue before any store, so their memory latencies
// overlap; live[r] implies live[r-1].
// Derivs(sxx_x, sxy_x, sxy_y, syy_y) fills the stress derivatives of the R
// cells (called only when live[0]): from global memory (vel_cell_rows) or
// from the workgroup's LDS copy of the tile (k_vel_step_lds).
template <int N, int R, typename Derivs>
__device__ __forceinline__ void vel_rows_with(const StepArgs &a, int s, int blk, int i, int j, const bool (&live)[R],
                                              long c, long e, int nt, int flags, const float (&rip)[R],
                                              const float (&rjp)[R], float (*s_red)[kTileY], Derivs derivs)
{
    const StepGeom &g = a.g;
    const int P = g.pitch, NX = g.NX;
    const long wo = (long)s * a.ws, po = (long)s * a.pss;
    const auto vxs = gmem(a.vx) + wo, vys = gmem(a.vy) + wo;
    float sxx_x[R], sxy_x[R], sxy_y[R], syy_y[R], vx0[R], vy0[R];
#pragma unroll
    for (int r = 0; r < R; r++) sxx_x[r] = sxy_x[r] = sxy_y[r] = syy_y[r] = vx0[r] = vy0[r] = 0.0f;
    if (live[0]) {
        derivs(sxx_x, sxy_x, sxy_y, syy_y);
#pragma unroll
        for (int r = 0; r < R; r++)
            if (live[r]) {
                vx0[r] = vxs[c + r * P];
                vy0[r] = vys[c + r * P];
            }
#pragma unroll
        for (int r = 0; r < R; r++)
            if (live[r])
                cpml_velocity(g, j + r, i, a.psi_sxx_x + po, a.psi_sxy_x + po, a.psi_syy_y + po, a.psi_sxy_y + po,
                              sxx_x[r], sxy_x[r], sxy_y[r], syy_y[r]);
    }
    const int so = __builtin_amdgcn_readfirstlane(s * (a.nbx * a.nby + 1) + blk);  // uniform
    const int q0 = cmem(a.inj.start)[so], q1 = cmem(a.inj.start)[so + 1];
    if (flags & kForward) {
        if (flags & kSnapStep) {
            float val[R][2];
#pragma unroll
            for (int r = 0; r < R; r++) {
                val[r][0] = vel_rate(rip[r], sxx_x[r] + sxy_y[r], g.DH);
                val[r][1] = vel_rate(rjp[r], sxy_x[r] + syy_y[r], g.DH);
            }
            snap_store_rows<2, R>(a, s, 0, (nt - 1) / a.DTINV, e, blk, live, val, s_red);
        }
    } else if (flags & kAdjInt) {
#pragma unroll
        for (int r = 0; r < R; r++)
            if (live[r]) {
                gmem(a.vxp1)[wo + c + r * P] += vx0[r] * g.DT;
                gmem(a.vyp1)[wo + c + r * P] += vy0[r] * g.DT;
            }
    } else if (flags & kCorrStep) {
#pragma unroll
        for (int r = 0; r < R; r++)
            if (live[r]) {
                gmem(a.vxp1)[wo + c + r * P] = vx0[r];
                gmem(a.vyp1)[wo + c + r * P] = vy0[r];
            }
    }
    if (!live[0]) return;
    float vx[R], vy[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        vx[r] = vel_new(vx0[r], g.DT, rip[r], sxx_x[r] + sxy_y[r], g.DH);
        vy[r] = vel_new(vy0[r], g.DT, rjp[r], sxy_x[r] + syy_y[r], g.DH);
    }
    // point forces / residual injection of this tile, in list order; a point
    // lands on at most one of the R cells
    for (int q = q0; q < q1; q++) {
        const int l = cmem(a.inj.idx)[q];
        const int px = cmem(a.inj.ix)[l], py = cmem(a.inj.iy)[l];
        const float s0 = cmem(a.inj.sig)[(long)l * a.NT + nt - 1];
        const float s1 = cmem(a.inj.sig1)[(long)l * a.NT + nt - 1];
        if (px != i || py < j || py >= j + R) continue;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (py != j + r) continue;
            if (a.sw == 0) {
                const int type = cmem(a.inj.type)[l];
                if (type == 2) vx[r] += (g.DT * rip[r] * s1) / (g.DH * g.DH);
                if (type == 3) vy[r] += (g.DT * rjp[r] * s1) / (g.DH * g.DH);
                if (type == 4) {
                    vx[r] += (g.DT * rip[r] * cmem(a.inj.sin_azi)[l] * s0) / (g.DH * g.DH);
                    vy[r] += (g.DT * rjp[r] * cmem(a.inj.cos_azi)[l] * s0) / (g.DH * g.DH);
                }
            } else {
                if (a.ADJOINT_TYPE == 1) { vx[r] += s0; vy[r] += s1; }
                if (a.ADJOINT_TYPE == 2) vy[r] += s1;
                if (a.ADJOINT_TYPE == 3) vx[r] += s0;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
        if (live[r]) {
            vxs[c + r * P] = vx[r];
            vys[c + r * P] = vy[r];
            if (g.BOUNDARY) wrap_pair<N>(g, vys, vxs, c + r * P, i, vy[r], vx[r]);
        }
    if (flags & kEnergy) {
        if (a.wacc64) {
            const auto W = gmem(a.wacc64) + (long)s * a.wacs;
#pragma unroll
            for (int r = 0; r < R; r++)
                if (live[r])
                    W[e + r * NX] += ((double)vx[r] * (double)vx[r]) + ((double)vy[r] * (double)vy[r]);
        } else {
            const auto W = gmem(a.wacc) + (long)s * a.wacs;
#pragma unroll
            for (int r = 0; r < R; r++)
                if (live[r]) W[e + r * NX] += ((vx[r] * vx[r]) + (vy[r] * vy[r]));
        }
    }
}

template <int N, int R>
__device__ __forceinline__ void vel_cell_rows(const StepArgs &a, int s, int blk, int i, int j, const bool (&live)[R],
                                              long c, long e, int nt, int flags, const float (&rip)[R],
                                              const float (&rjp)[R], float (*s_red)[kTileY])
{
    const StepGeom &g = a.g;
    const long wo = (long)s * a.ws;
    vel_rows_with<N, R>(a, s, blk, i, j, live, c, e, nt, flags, rip, rjp, s_red,
                        [&](float(&sxx_x)[R], float(&sxy_x)[R], float(&sxy_y)[R], float(&syy_y)[R]) {
                            // the padded grid holds rows 1-H .. NY+H with H = N+1; with
                            // j <= NY, R = 2 reads up to row NY+N+1 and larger R clamp to it
                            stress_derivs_rows<N, R>(g.hc, a.sxx + wo + c, a.syy + wo + c, a.sxy + wo + c, g.pitch,
                                                     g.NY + N + 1 - j, sxx_x, sxy_x, sxy_y, syy_y);
                        });
}

// RY = 1: a 64x8 block, one cell per thread.  RY = 2 or 4: a 64x(8/RY)
// block, each thread owns rows RY*ty .. RY*ty+RY-1 of the 64x8 tile
// (vel_cell_rows).
template <int N, int SPB, int RY, bool LEAN>
__global__ __launch_bounds__(512) void k_vel_step(const StepArgs *__restrict__ pa, int k, int persistent, int flags_)
{
    // LEAN: a step with no snapshot, correlation, adjoint integration or
    // energy sum (lean_step), compiled without that code and its registers
    const int flags = LEAN ? (flags_ & kForward) : flags_;
    const StepArgs &a = *pa;
    __shared__ float s_red[2][kTileY];
    const StepGeom &g = a.g;
    const ItemMap m = item_map(a);
    const int nt = step_number(a, k);
    int first, end, stride;
    work_range(m, m.ntiles * m.ngroups, persistent != 0, first, end, stride);
    for (int it = first; it < end; it += stride) {
        int blk, s0;
        item_of(m, it, SPB, blk, s0);
        // uniform across the workgroup: keep them in scalar registers
        blk = __builtin_amdgcn_readfirstlane(blk);
        s0 = __builtin_amdgcn_readfirstlane(s0);
        const int by = (int)fdiv(blk, m.nbx_mul, m.nbx_sh), bx = blk - by * m.nbx;
        const int i = bx * kTileX + threadIdx.x + 1;
        const int j = by * kTileY + RY * threadIdx.y + 1;
        const long c = (long)j * g.pitch + i;
        const long e = (long)(j - 1) * g.NX + (i - 1);
        if constexpr (RY == 1) {
            const bool live = i <= g.NX && j <= g.NY;
            const float rip = live ? gmem(a.rip)[c] : 0.0f, rjp = live ? gmem(a.rjp)[c] : 0.0f;
#pragma unroll
            for (int u = 0; u < SPB; u++)
                if (s0 + u < m.nslots) vel_cell<N>(a, s0 + u, blk, i, j, live, c, e, nt, flags, rip, rjp, s_red);
        } else {
            bool live[RY];
            float rip[RY], rjp[RY];
#pragma unroll
            for (int r = 0; r < RY; r++) {
                live[r] = i <= g.NX && j + r <= g.NY;
                // unconditional loads, then a select: live lanes read their
                // own cell, the others cell (1, 1), which every grid has (no
                // per-lane branch, so the pointers are loaded once)
                const long ci = live[r] ? c + r * g.pitch : (long)g.pitch + 1;
                const float vr = gmem(a.rip)[ci], vj = gmem(a.rjp)[ci];
                rip[r] = live[r] ? vr : 0.0f;
                rjp[r] = live[r] ? vj : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < SPB; u++)
                if (s0 + u < m.nslots)
                    vel_cell_rows<N, RY>(a, s0 + u, blk, i, j, live, c, e, nt, flags, rip, rjp, s_red);
        }
    }
}

// Velocity step with the tile's stresses staged in LDS: a 64x4 workgroup of
// two-row threads first copies, for its 64x8 tile, sxx (rows j0..j0+7,
// columns i0-N..i0+63+N), sxy (rows j0-N..j0+7+N, same columns) and syy (rows
// j0-N..j0+7+N, columns i0..i0+63) from HBM into LDS with row-contiguous
// loads, then every stencil value comes from LDS.  Coordinates outside the
// padded grid are clamped to its edge; they only feed cells outside the
// model, which are not stored.  Same arithmetic as vel_cell.
template <int N>
struct VelTile {
    static constexpr int W = kTileX + 2 * N;  // LDS column x <-> grid column i0 - N + x
    static constexpr int H = kTileY + 2 * N;  // LDS row y <-> grid row j0 - N + y (sxy, syy)
    float sxx[kTileY][W];                     // LDS row y <-> grid row j0 + y
    float sxy[H][W];
    float syy[H][kTileX];                     // LDS column x <-> grid column i0 + x
};

template <int N>
__device__ __forceinline__ void stage_tile(VelTile<N> &t, const float *sxx_, const float *sxy_, const float *syy_,
                                           int P, int i0, int j0, int NX, int NY)
{
    constexpr int W = VelTile<N>::W, H = VelTile<N>::H;
    const auto sxx = gmem(sxx_), sxy = gmem(sxy_), syy = gmem(syy_);
    const int lo_i = -N, hi_i = NX + N + 1, lo_j = -N, hi_j = NY + N + 1;  // inside the padded grid
    const int tid = threadIdx.y * kTileX + threadIdx.x, nth = kTileX * (kTileY / 2);
    for (int q = tid; q < kTileY * W; q += nth) {
        const int y = q / W, x = q - y * W;
        const int ii = min(max(i0 - N + x, lo_i), hi_i), jj = min(j0 + y, hi_j);
        t.sxx[y][x] = sxx[(long)jj * P + ii];
    }
    for (int q = tid; q < H * W; q += nth) {
        const int y = q / W, x = q - y * W;
        const int ii = min(max(i0 - N + x, lo_i), hi_i), jj = min(max(j0 - N + y, lo_j), hi_j);
        t.sxy[y][x] = sxy[(long)jj * P + ii];
    }
    for (int q = tid; q < H * kTileX; q += nth) {
        const int y = q / kTileX, x = q - y * kTileX;
        const int ii = min(i0 + x, hi_i), jj = min(max(j0 - N + y, lo_j), hi_j);
        t.syy[y][x] = syy[(long)jj * P + ii];
    }
}

template <int N, int SPB>
__global__ __launch_bounds__(256) void k_vel_step_lds(const StepArgs *__restrict__ pa, int k, int persistent,
                                                      int flags)
{
    constexpr int R = 2;
    const StepArgs &a = *pa;
    __shared__ float s_red[2][kTileY];
    __shared__ VelTile<N> t;
    const StepGeom &g = a.g;
    const ItemMap m = item_map(a);
    const int nt = step_number(a, k);
    int first, end, stride;
    work_range(m, m.ntiles * m.ngroups, persistent != 0, first, end, stride);
    for (int it = first; it < end; it += stride) {
        int blk, s0;
        item_of(m, it, SPB, blk, s0);
        blk = __builtin_amdgcn_readfirstlane(blk);
        s0 = __builtin_amdgcn_readfirstlane(s0);
        const int by = (int)fdiv(blk, m.nbx_mul, m.nbx_sh), bx = blk - by * m.nbx;
        const int i0 = bx * kTileX + 1, j0 = by * kTileY + 1;
        const int tx = threadIdx.x, jr = R * threadIdx.y;  // this thread's rows: j0 + jr, j0 + jr + 1
        const int i = i0 + tx, j = j0 + jr;
        const long c = (long)j * g.pitch + i;
        const long e = (long)(j - 1) * g.NX + (i - 1);
        bool live[R];
        float rip[R], rjp[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            live[r] = i <= g.NX && j + r <= g.NY;
            rip[r] = live[r] ? gmem(a.rip)[c + r * g.pitch] : 0.0f;
            rjp[r] = live[r] ? gmem(a.rjp)[c + r * g.pitch] : 0.0f;
        }
#pragma unroll 1
        for (int u = 0; u < SPB; u++) {
            const int sl = s0 + u;
            if (sl >= m.nslots) break;  // uniform: the whole workgroup leaves together
            const long wo = (long)sl * a.ws;
            __syncthreads();  // the previous item's or slot's reads of t are done
            stage_tile<N>(t, a.sxx + wo, a.sxy + wo, a.syy + wo, g.pitch, i0, j0, g.NX, g.NY);
            __syncthreads();
            vel_rows_with<N, R>(a, sl, blk, i, j, live, c, e, nt, flags, rip, rjp, s_red,
                                [&](float(&sxx_x)[R], float(&sxy_x)[R], float(&sxy_y)[R], float(&syy_y)[R]) {
                                    const float *hc = g.hc;
#pragma unroll
                                    for (int r = 0; r < R; r++) {
                                        const float *sx = &t.sxx[jr + r][tx + N];
                                        const float *sy = &t.sxy[jr + r + N][tx + N];
                                        const float *cy = &t.sxy[jr + r + N][tx + N];
                                        const float *cs = &t.syy[jr + r + N][tx];
                                        constexpr int Wd = VelTile<N>::W;
                                        sxx_x[r] = hc[1] * (sx[1] - sx[0]);
                                        sxy_x[r] = hc[1] * (sy[0] - sy[-1]);
                                        sxy_y[r] = hc[1] * (cy[0] - cy[-Wd]);
                                        syy_y[r] = hc[1] * (cs[kTileX] - cs[0]);
#pragma unroll
                                        for (int mm = 2; mm <= N; mm++) {
                                            sxx_x[r] = sxx_x[r] + hc[mm] * (sx[mm] - sx[1 - mm]);
                                            sxy_x[r] = sxy_x[r] + hc[mm] * (sy[mm - 1] - sy[-mm]);
                                            sxy_y[r] = sxy_y[r] + hc[mm] * (cy[(mm - 1) * Wd] - cy[-mm * Wd]);
                                            syy_y[r] = syy_y[r] + hc[mm] * (cs[mm * kTileX] - cs[(1 - mm) * kTileX]);
                                        }
                                    }
                                });
        }
    }
}

// ---------------------------------------------------------------------------
// per-cell material values of the stress step, loaded once per work item
template <int L>
struct StressMat {
    float fe, ge, fipjp;  // elastic
    ViscCell vc;          // viscoelastic
};

// stress step of one cell of slot s
template <int N, int L>
__device__ __forceinline__ void str_cell(const StepArgs &a, int s, int blk, int i, int j, bool live, bool top, long c,
                                         long e, int nt, int flags, const StressMat<L> &M, float (*s_syy)[kTileX],
                                         float (*s_sxy)[kTileX], float (*s_red)[kTileY])
{
    const StepGeom &g = a.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int P = g.pitch;
    const long wo = (long)s * a.ws, po = (long)s * a.pss;
    const auto vxs = gmem(a.vx) + wo, vys = gmem(a.vy) + wo;
    const bool snap_step = (flags & kSnapStep) != 0;
    const bool corr_step = (flags & kCorrStep) != 0;
    const bool rates = snap_step && a.VELOCITY == 1;
    // bounds of this tile's explosive-source and receiver lists, loaded
    // before the stencil work so that their latency is hidden behind it
    int expl = 0, ex0 = 0, ex1 = 0, rc0 = 0, rc1 = 0;
    if (flags & kForward) {
        const int so = __builtin_amdgcn_readfirstlane(s * (a.nbx * a.nby + 1) + blk);  // uniform
        expl = cmem(a.explosive)[s];
        ex0 = cmem(a.exp.start)[so];
        ex1 = cmem(a.exp.start)[so + 1];
        rc0 = cmem(a.rec.start)[so];
        rc1 = cmem(a.rec.start)[so + 1];
    }

    float sxx = 0.0f, syy = 0.0f, sxy = 0.0f, ux = 0.0f, uy = 0.0f, uxy = 0.0f;
    float rr[L > 0 ? L : 1], pp[L > 0 ? L : 1], qq[L > 0 ? L : 1];
    if (live) {
        float vxx, vyx, vxy, vyy;
        const float scale = L ? (float)(1.0 / (double)g.DH) : g.DT / g.DH;
        velocity_derivs<N>(g.hc, a.vx + wo + c, a.vy + wo + c, P, scale, vxx, vyx, vxy, vyy);
        cpml_stress(g, j, i, a.psi_vxx + po, a.psi_vyx + po, a.psi_vyy + po, a.psi_vxy + po, vxx, vyx, vxy, vyy);
        sxx = gmem(a.sxx)[wo + c];
        syy = gmem(a.syy)[wo + c];
        sxy = gmem(a.sxy)[wo + c];
        if (L == 0) {
            if (rates) rates_elastic(M.fe, M.ge, M.fipjp, g.DT, vxx, vyx, vxy, vyy, ux, uy, uxy);
            stress_elastic(M.fe, M.ge, M.fipjp, vxx, vyx, vxy, vyy, sxx, syy, sxy);
        } else {
#pragma unroll
            for (int l = 0; l < L; l++) {
                rr[l] = gmem(a.r)[wo + mv_index(c, l, 3 * L)];
                pp[l] = gmem(a.r)[wo + mv_index(c, L + l, 3 * L)];
                qq[l] = gmem(a.r)[wo + mv_index(c, 2 * L + l, 3 * L)];
            }
            stress_visc<(L > 0 ? L : 1)>(M.vc, a.eta, a.bip, a.bjm, a.cip, a.cjm, g.DT, vxx, vyx, vxy, vyy, sxx, syy,
                                       sxy, rr, pp, qq, rates, ux, uy, uxy);
        }
        // explosive sources of this tile, applied after the stress update
        // and before the free surface (src/IFOS2D.c:1758-1773)
        if (expl) {
            for (int q = ex0; q < ex1; q++) {
                const int l = cmem(a.exp.idx)[q];
                if (cmem(a.exp.ix)[l] != i || cmem(a.exp.iy)[l] != j) continue;
                const float amp = explosive_amp(a.exp.sig + (long)l * a.NT - 1, nt, a.NT, g.DH);
                sxx = add_half(sxx, amp);
                syy = add_half(syy, amp);
            }
        }
    }
    if (top) {
        // the free surface of column i reads the new syy of rows 2..N+1 and
        // sxy of rows 1..N; they come from the other threads of this tile
        s_syy[ty][tx] = live ? syy : 0.0f;
        s_sxy[ty][tx] = live ? sxy : 0.0f;
        __syncthreads();
        if (live && j == 1) {
            syy = 0.0f;
            if (rates) uy = 0.0f;
            if (L) {
#pragma unroll
                for (int l = 0; l < L; l++) qq[l] = 0.0f;
            }
#pragma unroll
            for (int m = 1; m <= N; m++) {
                gmem(a.syy)[wo + c - (long)m * P] = -s_syy[m][tx];
                gmem(a.sxy)[wo + c - (long)m * P] = -s_sxy[m - 1][tx];
            }
            float vxx, vyy;
            surface_derivs<N>(g, i, a.vx + wo + c, a.vy + wo + c, P, a.psi_vxxs + po, vxx, vyy);
            if (L == 0) {
                surface_elastic_update((float)((double)M.fe * 2.0), M.ge, g.DT, vxx, vyy, sxx, rates, ux);
            } else {
                surface_visc_update<(L > 0 ? L : 1)>(M.vc.mu, M.vc.pi, M.vc.taus, M.vc.taup, a.eta, g.DT, vxx, vyy,
                                                     sxx, pp, rates, ux);
            }
        }
        __syncthreads();  // the exchange rows are reused by the next slot
    }
    if (snap_step) {
        const float val[3] = {a.VELOCITY ? ux : sxx, a.VELOCITY ? uy : syy, a.VELOCITY ? uxy : sxy};
        snap_store<3>(a, s, 2, (nt - 1) / a.DTINV, e, blk, live, val, s_red);
    }
    if (!live) return;
    gmem(a.sxx)[wo + c] = sxx;
    gmem(a.syy)[wo + c] = syy;
    gmem(a.sxy)[wo + c] = sxy;
    if (g.BOUNDARY) wrap_pair<N>(g, gmem(a.sxx) + wo, gmem(a.sxy) + wo, c, i, sxx, sxy);
    if (L) {
#pragma unroll
        for (int l = 0; l < L; l++) {
            gmem(a.r)[wo + mv_index(c, l, 3 * L)] = rr[l];
            gmem(a.r)[wo + mv_index(c, L + l, 3 * L)] = pp[l];
            gmem(a.r)[wo + mv_index(c, 2 * L + l, 3 * L)] = qq[l];
        }
    }
    if (corr_step) {
        const long ks = (a.NT - nt) / a.DTINV, wco = (long)s * a.wcs;
        float wr = gmem(a.wc_rho)[wco + e], w = gmem(a.wc)[wco + e], wu = gmem(a.wc_u)[wco + e];
        correlate_cell(gmem(a.vxp1)[wo + c], gmem(a.vyp1)[wo + c], sxx, syy, sxy, snap_load(a, s, 0, ks, e, blk),
                       snap_load(a, s, 1, ks, e, blk), snap_load(a, s, 2, ks, e, blk),
                       snap_load(a, s, 3, ks, e, blk), snap_load(a, s, 4, ks, e, blk), gmem(a.muss)[e],
                       gmem(a.lamss)[e], gmem(a.vs_i)[e] > 0.0f, wr, w, wu);
        gmem(a.wc_rho)[wco + e] = wr;
        gmem(a.wc)[wco + e] = w;
        gmem(a.wc_u)[wco + e] = wu;
    }
    if (flags & kForward) {
        for (int q = rc0; q < rc1; q++) {
            const int t = cmem(a.rec.idx)[q];
            if (cmem(a.rec.ix)[t] != i || cmem(a.rec.iy)[t] != j) continue;
            gmem(a.seis_vx)[(long)t * a.NT + nt - 1] = vxs[c];
            gmem(a.seis_vy)[(long)t * a.NT + nt - 1] = vys[c];
        }
    }
    if (a.nan_flag && j == g.NY / 2 && i == g.NX / 2) {
        const float v = vys[c];
        if (isnan(v) && *gmem(a.nan_flag) == 0) *gmem(a.nan_flag) = nt;
    }
}

template <int N, int L, int SPB, bool LEAN>
__global__ __launch_bounds__(512) void k_str_step(const StepArgs *__restrict__ pa, int k, int persistent, int flags_)
{
    const int flags = LEAN ? (flags_ & kForward) : flags_;  // as in k_vel_step
    const StepArgs &a = *pa;
    __shared__ float s_syy[kTileY][kTileX];
    __shared__ float s_sxy[kTileY][kTileX];
    __shared__ float s_red[3][kTileY];
    const StepGeom &g = a.g;
    const ItemMap m = item_map(a);
    const int nt = step_number(a, k);
    int first, end, stride;
    work_range(m, m.ntiles * m.ngroups, persistent != 0, first, end, stride);
    for (int it = first; it < end; it += stride) {
        int blk, s0;
        item_of(m, it, SPB, blk, s0);
        // uniform across the workgroup: keep them in scalar registers
        blk = __builtin_amdgcn_readfirstlane(blk);
        s0 = __builtin_amdgcn_readfirstlane(s0);
        const int by = (int)fdiv(blk, m.nbx_mul, m.nbx_sh), bx = blk - by * m.nbx;
        const int i = bx * kTileX + threadIdx.x + 1;
        const int j = by * kTileY + threadIdx.y + 1;
        const bool live = i <= g.NX && j <= g.NY;
        const bool top = by == 0 && g.FREE_SURF;
        const long c = (long)j * g.pitch + i;
        const long e = (long)(j - 1) * g.NX + (i - 1);
        StressMat<L> M;
        M.fe = M.ge = M.fipjp = 0.0f;
        M.vc = ViscCell{};
        if (live) {
            if (L == 0) {
                M.fe = gmem(a.f)[c];
                M.ge = gmem(a.gl)[c];
                M.fipjp = gmem(a.uipjp)[c];
            } else {
                visc_coeffs(L, g.DT, gmem(a.mu)[c], gmem(a.pi)[c], gmem(a.muipjp)[c], gmem(a.taus)[c], gmem(a.taup)[c], gmem(a.tausipjp)[c], M.vc);
            }
        }
#pragma unroll
        for (int u = 0; u < SPB; u++)
            if (s0 + u < m.nslots)
                str_cell<N, L>(a, s0 + u, blk, i, j, live, top, c, e, nt, flags, M, s_syy, s_sxy, s_red);
    }
}

__global__ void k_tick(int *base, int by)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *base += by;
}

// ---------------------------------------------------------------------------
// One time step in one kernel.  A workgroup owns a 64x8 tile.  It first
// stages the old stresses it needs in LDS, computes the new velocities of the
// tile and of an N-cell rim around it (the rim's values are what the
// neighbouring tiles compute for their own cells, from the same inputs and
// the same arithmetic, so they are the same bits), then updates the tile's
// stresses from those velocities.  Wavefields and the velocity-side CPML
// memory are read from one buffer and written to the other (the buffers swap
// every step), because a tile reads its neighbours' old values while they
// write their new ones.  Per cell and step this moves one read and one write
// of vx, vy, sxx, syy, sxy instead of the two kernels' two reads of the
// stresses and of the velocities.
template <int N>
struct FusedLds {
    static constexpr int VX = kTileX + 2 * N, VY = kTileY + 2 * N;  // velocity region
    static constexpr int AX = kTileX + 4 * N - 1, SY = kTileY + 4 * N - 1;
    float sxx[VY][AX];   // rows j0+1-N .., columns i0+2-2N ..
    float syy[SY][VX];   // rows j0+2-2N .., columns i0+1-N ..
    float sxy[SY][AX];   // rows j0+1-2N .., columns i0+1-2N ..
    float vx[VY][VX];    // new velocities, rows j0+1-N .., columns i0+1-N ..
    float vy[VY][VX];
    float acc[2][kTileY][kTileX];  // accelerations of the tile's cells (forward snapshot steps)
};

// velocity-side CPML (the arithmetic of cpml_velocity) reading the memory of
// the previous step from the *_r arrays and, when `own`, writing the new
// memory to the *_w arrays
__device__ __forceinline__ void cpml_velocity_rw(const StepGeom &g, int j, int i, bool own, const float *rxx,
                                                 const float *rxy, const float *ryy, const float *ryx, float *wxx,
                                                 float *wxy, float *wyy, float *wyx, float &sxx_x, float &sxy_x,
                                                 float &sxy_y, float &syy_y)
{
    const Cpml &pm = g.pml;
    auto side_x = [&](int h) {
        const long o = (long)(j - 1) * g.px + h;
        const float p1 = cmem(pm.b_xh)[h] * gmem(rxx)[o] + cmem(pm.a_xh)[h] * sxx_x;
        sxx_x = divK(sxx_x, pm.K_xh, h, g.K_one) + p1;
        const float p2 = cmem(pm.b_x)[h] * gmem(rxy)[o] + cmem(pm.a_x)[h] * sxy_x;
        sxy_x = divK(sxy_x, pm.K_x, h, g.K_one) + p2;
        if (own) {
            gmem(wxx)[o] = p1;
            gmem(wxy)[o] = p2;
        }
    };
    auto side_y = [&](int h) {
        const long o = (long)h * g.py + (i - 1);
        const float p1 = cmem(pm.b_yh)[h] * gmem(ryy)[o] + cmem(pm.a_yh)[h] * syy_y;
        syy_y = divK(syy_y, pm.K_yh, h, g.K_one) + p1;
        const float p2 = cmem(pm.b_y)[h] * gmem(ryx)[o] + cmem(pm.a_y)[h] * sxy_y;
        sxy_y = divK(sxy_y, pm.K_y, h, g.K_one) + p2;
        if (own) {
            gmem(wyy)[o] = p1;
            gmem(wyx)[o] = p2;
        }
    };
    if (!g.BOUNDARY && i <= g.FW) side_x(i - 1);
    if (!g.BOUNDARY && i >= g.NX - g.FW + 1) side_x(i - g.NX + 2 * g.FW - 1);
    if (!g.FREE_SURF && j <= g.FW) side_y(j - 1);
    if (j >= g.NY - g.FW + 1) side_y(j - g.NY + 2 * g.FW - 1);
}

template <int N, int L>
__global__ __launch_bounds__(512) void k_fused_step(const StepArgs *__restrict__ pa, int k, int persistent,
                                                     int flags)
{
    using T = FusedLds<N>;
    constexpr int VX = T::VX, VY = T::VY, AX = T::AX, SY = T::SY;
    const StepArgs &a = *pa;
    __shared__ T st;
    __shared__ float s_syy[kTileY][kTileX];
    __shared__ float s_sxy[kTileY][kTileX];
    __shared__ float s_red[3][kTileY];
    const StepGeom &g = a.g;
    const ItemMap m = item_map(a);
    const int nt = step_number(a, k);
    const int tid = threadIdx.y * kTileX + threadIdx.x, tx = threadIdx.x, ty = threadIdx.y;
    const int P = g.pitch, NX = g.NX, NY = g.NY, H = N + 1;
    // buffers: step nt reads set (nt - 1) & 1 and writes set nt & 1
    const bool odd = (nt & 1) != 0;
    const float *rvx = odd ? a.vx : a.vx2, *rvy = odd ? a.vy : a.vy2;
    const float *rsxx = odd ? a.sxx : a.sxx2, *rsyy = odd ? a.syy : a.syy2, *rsxy = odd ? a.sxy : a.sxy2;
    float *wvx = odd ? a.vx2 : a.vx, *wvy = odd ? a.vy2 : a.vy;
    float *wsxx = odd ? a.sxx2 : a.sxx, *wsyy = odd ? a.syy2 : a.syy, *wsxy = odd ? a.sxy2 : a.sxy;
    const float *rp_xx = odd ? a.psi_sxx_x : a.psi2_sxx_x, *rp_xy = odd ? a.psi_sxy_x : a.psi2_sxy_x;
    const float *rp_yy = odd ? a.psi_syy_y : a.psi2_syy_y, *rp_yx = odd ? a.psi_sxy_y : a.psi2_sxy_y;
    float *wp_xx = odd ? a.psi2_sxx_x : a.psi_sxx_x, *wp_xy = odd ? a.psi2_sxy_x : a.psi_sxy_x;
    float *wp_yy = odd ? a.psi2_syy_y : a.psi_syy_y, *wp_yx = odd ? a.psi2_sxy_y : a.psi_sxy_y;
    const bool snap_step = (flags & kSnapStep) != 0, corr_step = (flags & kCorrStep) != 0;
    const bool fwd = (flags & kForward) != 0;
    const bool rates = snap_step && a.VELOCITY == 1;
    int first, end, stride;
    work_range(m, m.ntiles * m.ngroups, persistent != 0, first, end, stride);
    for (int it = first; it < end; it += stride) {
        int blk, s;
        item_of(m, it, 1, blk, s);
        blk = __builtin_amdgcn_readfirstlane(blk);
        s = __builtin_amdgcn_readfirstlane(s);
        const int by = (int)fdiv(blk, m.nbx_mul, m.nbx_sh), bx = blk - by * m.nbx;
        const int i0 = bx * kTileX, j0 = by * kTileY;
        const long wo = (long)s * a.ws, po = (long)s * a.pss;
        // ---- stage the old stresses (zero outside the padded grid, where
        // they only feed velocities of cells outside the model)
        auto stage = [&](const float *src, float *dst, int rows, int cols, int jr, int ic) {
            const auto sg = gmem(src) + wo;
            for (int q = tid; q < rows * cols; q += kTileX * kTileY) {
                const int r = q / cols, cc = q - r * cols;
                const int gj = jr + r, gi = ic + cc;
                const bool in = gj >= 1 - H && gj <= NY + H && gi >= -15 && gi <= P - 16;
                dst[q] = in ? sg[(long)gj * P + gi] : 0.0f;
            }
        };
        __syncthreads();  // the previous item's readers are done with the LDS
        stage(rsxx, &st.sxx[0][0], VY, AX, j0 + 1 - N, i0 + 2 - 2 * N);
        stage(rsyy, &st.syy[0][0], SY, VX, j0 + 2 - 2 * N, i0 + 1 - N);
        stage(rsxy, &st.sxy[0][0], SY, AX, j0 + 1 - 2 * N, i0 + 1 - 2 * N);
        __syncthreads();
        // ---- velocities of the tile and its rim
        const int so = __builtin_amdgcn_readfirstlane(s * (a.nbx * a.nby + 1) + blk);  // uniform
        const int q0 = cmem(a.injx.start)[so], q1 = cmem(a.injx.start)[so + 1];
        for (int q = tid; q < VX * VY; q += kTileX * kTileY) {
            const int ry = q / VX, rx = q - ry * VX;
            const int gj = j0 + 1 - N + ry, gi = i0 + 1 - N + rx;
            float vx = 0.0f, vy = 0.0f;
            if (gj >= 1 && gj <= NY && gi >= 1 && gi <= NX) {
                const bool own = rx >= N && rx < N + kTileX && ry >= N && ry < N + kTileY;
                const float *X = &st.sxx[ry][rx + N - 1];
                const float *Y = &st.syy[ry + N - 1][rx];
                const float *S = &st.sxy[ry + N][rx + N];
                float sxx_x = g.hc[1] * (X[1] - X[0]);
                float sxy_x = g.hc[1] * (S[0] - S[-1]);
                float sxy_y = g.hc[1] * (S[0] - S[-AX]);
                float syy_y = g.hc[1] * (Y[VX] - Y[0]);
#pragma unroll
                for (int mm = 2; mm <= N; mm++) {
                    sxx_x = sxx_x + g.hc[mm] * (X[mm] - X[1 - mm]);
                    sxy_x = sxy_x + g.hc[mm] * (S[mm - 1] - S[-mm]);
                    sxy_y = sxy_y + g.hc[mm] * (S[(mm - 1) * AX] - S[-mm * AX]);
                    syy_y = syy_y + g.hc[mm] * (Y[mm * VX] - Y[(1 - mm) * VX]);
                }
                cpml_velocity_rw(g, gj, gi, own, rp_xx + po, rp_xy + po, rp_yy + po, rp_yx + po, wp_xx + po,
                                 wp_xy + po, wp_yy + po, wp_yx + po, sxx_x, sxy_x, sxy_y, syy_y);
                const long c = (long)gj * P + gi;
                const float vx0 = gmem(rvx)[wo + c], vy0 = gmem(rvy)[wo + c];
                const float rip = gmem(a.rip)[c], rjp = gmem(a.rjp)[c];
                if (own) {
                    if (fwd) {
                        if (snap_step) {
                            st.acc[0][ry - N][rx - N] = vel_rate(rip, sxx_x + sxy_y, g.DH);
                            st.acc[1][ry - N][rx - N] = vel_rate(rjp, sxy_x + syy_y, g.DH);
                        }
                    } else if (flags & kAdjInt) {
                        gmem(a.vxp1)[wo + c] += vx0 * g.DT;
                        gmem(a.vyp1)[wo + c] += vy0 * g.DT;
                    } else if (corr_step) {
                        gmem(a.vxp1)[wo + c] = vx0;
                        gmem(a.vyp1)[wo + c] = vy0;
                    }
                }
                vx = vel_new(vx0, g.DT, rip, sxx_x + sxy_y, g.DH);
                vy = vel_new(vy0, g.DT, rjp, sxy_x + syy_y, g.DH);
                // point forces / residual injection at this cell, in list order
                for (int qq = q0; qq < q1; qq++) {
                    const int l = cmem(a.injx.idx)[qq];
                    if (cmem(a.injx.ix)[l] != gi || cmem(a.injx.iy)[l] != gj) continue;
                    const float s0 = cmem(a.injx.sig)[(long)l * a.NT + nt - 1];
                    const float s1 = cmem(a.injx.sig1)[(long)l * a.NT + nt - 1];
                    if (a.sw == 0) {
                        const int type = cmem(a.injx.type)[l];
                        if (type == 2) vx += (g.DT * rip * s1) / (g.DH * g.DH);
                        if (type == 3) vy += (g.DT * rjp * s1) / (g.DH * g.DH);
                        if (type == 4) {
                            vx += (g.DT * rip * cmem(a.injx.sin_azi)[l] * s0) / (g.DH * g.DH);
                            vy += (g.DT * rjp * cmem(a.injx.cos_azi)[l] * s0) / (g.DH * g.DH);
                        }
                    } else {
                        if (a.ADJOINT_TYPE == 1) { vx += s0; vy += s1; }
                        if (a.ADJOINT_TYPE == 2) vy += s1;
                        if (a.ADJOINT_TYPE == 3) vx += s0;
                    }
                }
                if (own) {
                    gmem(wvx)[wo + c] = vx;
                    gmem(wvy)[wo + c] = vy;
                }
            }
            st.vx[ry][rx] = vx;
            st.vy[ry][rx] = vy;
        }
        __syncthreads();
        // ---- the tile's cells: one per thread
        const int i = i0 + 1 + tx, j = j0 + 1 + ty;
        const bool live = i <= NX && j <= NY;
        const bool top = by == 0 && g.FREE_SURF;
        const long c = (long)j * P + i, e = (long)(j - 1) * NX + (i - 1);
        const float nvx = st.vx[ty + N][tx + N], nvy = st.vy[ty + N][tx + N];
        if (fwd && snap_step) {
            const float val[2] = {st.acc[0][ty][tx], st.acc[1][ty][tx]};
            snap_store<2>(a, s, 0, (nt - 1) / a.DTINV, e, blk, live, val, s_red);
        }
        if ((flags & kEnergy) && live) {
            if (a.wacc64) {
                const auto W = gmem(a.wacc64) + (long)s * a.wacs;
                W[e] += ((double)nvx * (double)nvx) + ((double)nvy * (double)nvy);
            } else {
                const auto W = gmem(a.wacc) + (long)s * a.wacs;
                W[e] += ((nvx * nvx) + (nvy * nvy));
            }
        }
        int expl = 0, ex0 = 0, ex1 = 0, rc0 = 0, rc1 = 0;
        if (fwd) {
            expl = cmem(a.explosive)[s];
            ex0 = cmem(a.exp.start)[so];
            ex1 = cmem(a.exp.start)[so + 1];
            rc0 = cmem(a.rec.start)[so];
            rc1 = cmem(a.rec.start)[so + 1];
        }
        float sxx = 0.0f, syy = 0.0f, sxy = 0.0f, ux = 0.0f, uy = 0.0f, uxy = 0.0f;
        float rr[L > 0 ? L : 1], pp[L > 0 ? L : 1], qq[L > 0 ? L : 1];
        StressMat<L> M;
        M.fe = M.ge = M.fipjp = 0.0f;
        M.vc = ViscCell{};
        const float *V = &st.vx[ty + N][tx + N], *Wv = &st.vy[ty + N][tx + N];
        if (live) {
            if (L == 0) {
                M.fe = gmem(a.f)[c];
                M.ge = gmem(a.gl)[c];
                M.fipjp = gmem(a.uipjp)[c];
            } else {
                visc_coeffs(L, g.DT, gmem(a.mu)[c], gmem(a.pi)[c], gmem(a.muipjp)[c], gmem(a.taus)[c],
                            gmem(a.taup)[c], gmem(a.tausipjp)[c], M.vc);
            }
            const float scale = L ? (float)(1.0 / (double)g.DH) : g.DT / g.DH;
            float vxx = g.hc[1] * (V[0] - V[-1]);
            float vyx = g.hc[1] * (Wv[1] - Wv[0]);
            float vxy = g.hc[1] * (V[VX] - V[0]);
            float vyy = g.hc[1] * (Wv[0] - Wv[-VX]);
#pragma unroll
            for (int mm = 2; mm <= N; mm++) {
                vxx = vxx + g.hc[mm] * (V[mm - 1] - V[-mm]);
                vyx = vyx + g.hc[mm] * (Wv[mm] - Wv[1 - mm]);
                vxy = vxy + g.hc[mm] * (V[mm * VX] - V[(1 - mm) * VX]);
                vyy = vyy + g.hc[mm] * (Wv[(mm - 1) * VX] - Wv[-mm * VX]);
            }
            vxx = vxx * scale;
            vyx = vyx * scale;
            vxy = vxy * scale;
            vyy = vyy * scale;
            cpml_stress(g, j, i, a.psi_vxx + po, a.psi_vyx + po, a.psi_vyy + po, a.psi_vxy + po, vxx, vyx, vxy, vyy);
            sxx = st.sxx[ty + N][tx + 2 * N - 1];
            syy = st.syy[ty + 2 * N - 1][tx + N];
            sxy = st.sxy[ty + 2 * N][tx + 2 * N];
            if (L == 0) {
                if (rates) rates_elastic(M.fe, M.ge, M.fipjp, g.DT, vxx, vyx, vxy, vyy, ux, uy, uxy);
                stress_elastic(M.fe, M.ge, M.fipjp, vxx, vyx, vxy, vyy, sxx, syy, sxy);
            } else {
#pragma unroll
                for (int l = 0; l < L; l++) {
                    rr[l] = gmem(a.r)[wo + mv_index(c, l, 3 * L)];
                    pp[l] = gmem(a.r)[wo + mv_index(c, L + l, 3 * L)];
                    qq[l] = gmem(a.r)[wo + mv_index(c, 2 * L + l, 3 * L)];
                }
                stress_visc<(L > 0 ? L : 1)>(M.vc, a.eta, a.bip, a.bjm, a.cip, a.cjm, g.DT, vxx, vyx, vxy, vyy, sxx,
                                           syy, sxy, rr, pp, qq, rates, ux, uy, uxy);
            }
            if (expl) {
                for (int q = ex0; q < ex1; q++) {
                    const int l = cmem(a.exp.idx)[q];
                    if (cmem(a.exp.ix)[l] != i || cmem(a.exp.iy)[l] != j) continue;
                    const float amp = explosive_amp(a.exp.sig + (long)l * a.NT - 1, nt, a.NT, g.DH);
                    sxx = add_half(sxx, amp);
                    syy = add_half(syy, amp);
                }
            }
        }
        if (top) {
            s_syy[ty][tx] = live ? syy : 0.0f;
            s_sxy[ty][tx] = live ? sxy : 0.0f;
            __syncthreads();
            if (live && j == 1) {
                syy = 0.0f;
                if (rates) uy = 0.0f;
                if (L) {
#pragma unroll
                    for (int l = 0; l < L; l++) qq[l] = 0.0f;
                }
#pragma unroll
                for (int mm = 1; mm <= N; mm++) {
                    gmem(wsyy)[wo + c - (long)mm * P] = -s_syy[mm][tx];
                    gmem(wsxy)[wo + c - (long)mm * P] = -s_sxy[mm - 1][tx];
                }
                // surface_derivs on the new velocities of row 1 (from LDS)
                float vxx = 0.0f, vyy = 0.0f;
#pragma unroll
                for (int mm = 1; mm <= N; mm++) {
                    vxx += g.hc[mm] * (V[mm - 1] - V[-mm]);
                    vyy += g.hc[mm] * (Wv[(mm - 1) * VX] - Wv[-mm * VX]);
                }
                const float dh24 = 1.0 / (double)g.DH;
                vxx *= dh24;
                vyy *= dh24;
                const Cpml &pm = g.pml;
                auto surf = [&](int h) {
                    auto ps = gmem(a.psi_vxxs + po) + h;
                    *ps = cmem(pm.b_x)[h] * *ps + cmem(pm.a_x)[h] * vxx;
                    vxx = divK(vxx, pm.K_x, h, g.K_one) + *ps;
                };
                if (!g.BOUNDARY && i <= g.FW) surf(i - 1);
                if (!g.BOUNDARY && i >= g.NX - g.FW + 1) surf(i - g.NX + 2 * g.FW - 1);
                if (L == 0) {
                    surface_elastic_update((float)((double)M.fe * 2.0), M.ge, g.DT, vxx, vyy, sxx, rates, ux);
                } else {
                    surface_visc_update<(L > 0 ? L : 1)>(M.vc.mu, M.vc.pi, M.vc.taus, M.vc.taup, a.eta, g.DT, vxx,
                                                         vyy, sxx, pp, rates, ux);
                }
            }
        }
        if (snap_step) {
            const float val[3] = {a.VELOCITY ? ux : sxx, a.VELOCITY ? uy : syy, a.VELOCITY ? uxy : sxy};
            snap_store<3>(a, s, 2, (nt - 1) / a.DTINV, e, blk, live, val, s_red);
        }
        if (!live) continue;
        gmem(wsxx)[wo + c] = sxx;
        gmem(wsyy)[wo + c] = syy;
        gmem(wsxy)[wo + c] = sxy;
        if (L) {
#pragma unroll
            for (int l = 0; l < L; l++) {
                gmem(a.r)[wo + mv_index(c, l, 3 * L)] = rr[l];
                gmem(a.r)[wo + mv_index(c, L + l, 3 * L)] = pp[l];
                gmem(a.r)[wo + mv_index(c, 2 * L + l, 3 * L)] = qq[l];
            }
        }
        if (corr_step) {
            const long ks = (a.NT - nt) / a.DTINV, wco = (long)s * a.wcs;
            float wr = gmem(a.wc_rho)[wco + e], w = gmem(a.wc)[wco + e], wu = gmem(a.wc_u)[wco + e];
            correlate_cell(gmem(a.vxp1)[wo + c], gmem(a.vyp1)[wo + c], sxx, syy, sxy,
                           snap_load(a, s, 0, ks, e, blk), snap_load(a, s, 1, ks, e, blk),
                           snap_load(a, s, 2, ks, e, blk), snap_load(a, s, 3, ks, e, blk),
                           snap_load(a, s, 4, ks, e, blk), gmem(a.muss)[e], gmem(a.lamss)[e],
                           gmem(a.vs_i)[e] > 0.0f, wr, w, wu);
            gmem(a.wc_rho)[wco + e] = wr;
            gmem(a.wc)[wco + e] = w;
            gmem(a.wc_u)[wco + e] = wu;
        }
        if (fwd) {
            for (int q = rc0; q < rc1; q++) {
                const int t = cmem(a.rec.idx)[q];
                if (cmem(a.rec.ix)[t] != i || cmem(a.rec.iy)[t] != j) continue;
                gmem(a.seis_vx)[(long)t * a.NT + nt - 1] = nvx;
                gmem(a.seis_vy)[(long)t * a.NT + nt - 1] = nvy;
            }
        }
        if (a.nan_flag && j == NY / 2 && i == NX / 2) {
            if (isnan(nvy) && *gmem(a.nan_flag) == 0) *gmem(a.nan_flag) = nt;
        }
    }
}

// ---------------------------------------------------------------------------
// multiplier and shift of fdiv for divisor d >= 1 (the round-up method for
// unsigned division by an invariant: l = ceil(log2 d),
// mul = floor(2^32 (2^l - d) / d) + 1, shift l - 1)
static void make_div(uint32_t d, uint32_t &mul, uint32_t &sh)
{
    if (d <= 1) {
        mul = 0;
        sh = 0;
        return;
    }
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    mul = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    sh = l - 1;
}

// the launcher's copy of the argument block with the per-shape fields set
StepArgs shaped_args(const StepArgs &a0, const LaunchShape &ls)
{
    StepArgs a = a0;
    a.ngroups = (a.nslots + ls.spb - 1) / ls.spb;
    a.slot_major = ls.order ? 1 : 0;
    make_div((uint32_t)a.nbx, a.nbx_mul, a.nbx_sh);
    make_div((uint32_t)a.ngroups, a.ng_mul, a.ng_sh);
    make_div((uint32_t)(a.nbx * a.nby), a.nt_mul, a.nt_sh);
    // XCD runs of equal estimated cost: a tile in an absorbing strip costs
    // 1 + cpml_weight (its cells update the CPML memory and divide by K);
    // the tiles of the bottom strip would otherwise all fall to the last XCD
    const StepGeom &g = a.g;
    const int ntiles = a.nbx * a.nby, nitems = ntiles * a.ngroups;
    auto in_cpml = [&](int blk) {
        const int by = blk / a.nbx, bx = blk - by * a.nbx;
        const int i0 = bx * kTileX + 1, i1 = i0 + kTileX - 1, j0 = by * kTileY + 1, j1 = j0 + kTileY - 1;
        return (!g.BOUNDARY && (i0 <= g.FW || i1 >= g.NX - g.FW + 1)) || (!g.FREE_SURF && j0 <= g.FW) ||
               j1 >= g.NY - g.FW + 1;
    };
    std::vector<double> cum(nitems + 1, 0.0);
    for (int it = 0; it < nitems; it++) {
        const int blk = a.slot_major ? it % ntiles : it / a.ngroups;
        cum[it + 1] = cum[it] + 1.0 + (in_cpml(blk) ? (double)ls.cpml_weight : 0.0);
    }
    a.xlo[0] = 0;
    for (int x = 1; x < 8; x++) {
        const double target = cum[nitems] * x / 8.0;
        a.xlo[x] = (int)(std::lower_bound(cum.begin(), cum.end(), target) - cum.begin());
        a.xlo[x] = std::max(a.xlo[x], a.xlo[x - 1]);
    }
    a.xlo[8] = nitems;
    return a;
}

static int grid_of(const StepArgs &a, const LaunchShape &ls, int grid)
{
    const int nitems = a.nbx * a.nby * ((a.nslots + ls.spb - 1) / ls.spb);
    if (grid > 0) return std::min(grid, (nitems + 7) / 8 * 8);
    if (!a.swizzle) return nitems;
    int widest = 0;  // one workgroup per item of the largest XCD run
    for (int x = 0; x < 8; x++) widest = std::max(widest, a.xlo[x + 1] - a.xlo[x]);
    return 8 * widest;
}

template <int N, int SPB, bool LEAN>
static void vel_launch_rows(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, int flags, int gr,
                            hipStream_t s)
{
    const int pers = ls.grid_v > 0 ? 1 : 0;
    if (ls.rows_v == 4)
        hipLaunchKernelGGL((k_vel_step<N, SPB, 4, LEAN>), dim3(gr), dim3(kTileX, kTileY / 4), 0, s, da, k, pers,
                           flags);
    else if (ls.rows_v == 2)
        hipLaunchKernelGGL((k_vel_step<N, SPB, 2, LEAN>), dim3(gr), dim3(kTileX, kTileY / 2), 0, s, da, k, pers,
                           flags);
    else
        hipLaunchKernelGGL((k_vel_step<N, SPB, 1, LEAN>), dim3(gr), dim3(kTileX, kTileY), 0, s, da, k, pers, flags);
}

template <int N, int SPB>
static void vel_launch(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    const int gr = grid_of(a, ls, ls.grid_v);
    const int flags = step_flags(a, k);
    if (ls.lds_v)
        hipLaunchKernelGGL((k_vel_step_lds<N, SPB>), dim3(gr), dim3(kTileX, kTileY / 2), 0, s, da, k,
                           ls.grid_v > 0 ? 1 : 0, flags);
    else if (lean_step(flags))
        vel_launch_rows<N, SPB, true>(a, da, ls, k, flags, gr, s);
    else
        vel_launch_rows<N, SPB, false>(a, da, ls, k, flags, gr, s);
}

template <int N>
static void vel_n(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    switch (ls.spb) {
    case 1: vel_launch<N, 1>(a, da, ls, k, s); break;
    case 2: vel_launch<N, 2>(a, da, ls, k, s); break;
    case 4: vel_launch<N, 4>(a, da, ls, k, s); break;
    default: break;
    }
}

void launch_vel_step(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    switch (a.g.N) {
    case 1: vel_n<1>(a, da, ls, k, s); break;
    case 2: vel_n<2>(a, da, ls, k, s); break;
    case 3: vel_n<3>(a, da, ls, k, s); break;
    case 4: vel_n<4>(a, da, ls, k, s); break;
    case 5: vel_n<5>(a, da, ls, k, s); break;
    case 6: vel_n<6>(a, da, ls, k, s); break;
    default: break;
    }
}

template <int N, int L, int SPB>
static void str_launch(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    const int gr = grid_of(a, ls, ls.grid);
    const int flags = step_flags(a, k);
    if (lean_step(flags))
        hipLaunchKernelGGL((k_str_step<N, L, SPB, true>), dim3(gr), dim3(kTileX, kTileY), 0, s, da, k,
                           ls.grid > 0 ? 1 : 0, flags);
    else
        hipLaunchKernelGGL((k_str_step<N, L, SPB, false>), dim3(gr), dim3(kTileX, kTileY), 0, s, da, k,
                           ls.grid > 0 ? 1 : 0, flags);
}

template <int N, int L>
static void str_nl(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    switch (ls.spb) {
    case 1: str_launch<N, L, 1>(a, da, ls, k, s); break;
    case 2: str_launch<N, L, 2>(a, da, ls, k, s); break;
    case 4: str_launch<N, L, 4>(a, da, ls, k, s); break;
    default: break;
    }
}

template <int N>
static void str_n(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    switch (a.L) {
    case 0: str_nl<N, 0>(a, da, ls, k, s); break;
    case 1: str_nl<N, 1>(a, da, ls, k, s); break;
    case 2: str_nl<N, 2>(a, da, ls, k, s); break;
    case 3: str_nl<N, 3>(a, da, ls, k, s); break;
    case 4: str_nl<N, 4>(a, da, ls, k, s); break;
    case 5: str_nl<N, 5>(a, da, ls, k, s); break;
    default: break;
    }
}

void launch_str_step(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    switch (a.g.N) {
    case 1: str_n<1>(a, da, ls, k, s); break;
    case 2: str_n<2>(a, da, ls, k, s); break;
    case 3: str_n<3>(a, da, ls, k, s); break;
    case 4: str_n<4>(a, da, ls, k, s); break;
    case 5: str_n<5>(a, da, ls, k, s); break;
    case 6: str_n<6>(a, da, ls, k, s); break;
    default: break;
    }
}

template <int N, int L>
static void fused_launch(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    const int gr = grid_of(a, ls, ls.grid);
    hipLaunchKernelGGL((k_fused_step<N, L>), dim3(gr), dim3(kTileX, kTileY), 0, s, da, k, ls.grid > 0 ? 1 : 0,
                       step_flags(a, k));
}

template <int N>
static void fused_n(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    switch (a.L) {
    case 0: fused_launch<N, 0>(a, da, ls, k, s); break;
    case 1: fused_launch<N, 1>(a, da, ls, k, s); break;
    case 2: fused_launch<N, 2>(a, da, ls, k, s); break;
    case 3: fused_launch<N, 3>(a, da, ls, k, s); break;
    case 4: fused_launch<N, 4>(a, da, ls, k, s); break;
    case 5: fused_launch<N, 5>(a, da, ls, k, s); break;
    default: break;
    }
}

void launch_fused_step(const StepArgs &a, const StepArgs *da, const LaunchShape &ls, int k, hipStream_t s)
{
    switch (a.g.N) {
    case 1: fused_n<1>(a, da, ls, k, s); break;
    case 2: fused_n<2>(a, da, ls, k, s); break;
    case 3: fused_n<3>(a, da, ls, k, s); break;
    case 4: fused_n<4>(a, da, ls, k, s); break;
    case 5: fused_n<5>(a, da, ls, k, s); break;
    case 6: fused_n<6>(a, da, ls, k, s); break;
    default: break;
    }
}

void launch_tick(int *base, int by, hipStream_t s) { hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, s, base, by); }

}  // namespace ifos
