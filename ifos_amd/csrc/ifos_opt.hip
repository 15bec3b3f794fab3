// ifos_opt.hip -- model-update side of an FWI iteration on the device:
// the reductions, the PCG direction, the step-length parabola, the bounded
// model update and L-BFGS, behind the C ABI of include/ifos2d_opt.h.
//
// Reductions (dot products, the Polak-Ribiere and L-BFGS sums) run in FP64
// as a deterministic two-pass tree: fixed per-block partial sums, then one
// block adds the partials in a fixed order, so a result does not depend on
// scheduling.  The reference sums in FP32 (src/PCG.c:170-174,
// src/LBFGS.c:247-250); everything else keeps its FP32 expressions, so with
// equal scalars the element-wise results are the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "ifos2d_opt.h"
#include "ifos_internal.h"

using namespace ifos;

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;  // partials per reduction

#define OPT_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            ifos_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return -1;                                                                             \
        }                                                                                          \
    } while (0)

__device__ __forceinline__ double block_sum(double v, double *lds)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kRedThreads / 64; w++) s += lds[w];
    __syncthreads();
    return s;
}

// partial[b] = sum over the grid-stride share of block b of a[i]*b[i]
// (mode 0), or a[i]*(a[i] - b[i]) (mode 1, the Polak-Ribiere numerator)
__global__ __launch_bounds__(kRedThreads) void k_dot_partial(const float *a_, const float *b_, long n, int mode,
                                                             double *partial_)
{
    __shared__ double lds[kRedThreads / 64];
    const auto a = gmem(a_), b = gmem(b_);
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double x = a[i], y = b[i];
        s += mode == 0 ? x * y : x * (x - y);
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) gmem(partial_)[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void k_sum_partials(const double *partial_, int np, double *out_)
{
    __shared__ double lds[kRedThreads / 64];
    const auto partial = gmem(partial_);
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) s += partial[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) *gmem(out_) = s;
}

// per-block maxima of a (mode 0) or |a| (mode 1); max is exact in any order
__global__ __launch_bounds__(kRedThreads) void k_max_partial(const float *a_, long n, int absolute, float *partial_)
{
    __shared__ float lds[kRedThreads / 64];
    const auto a = gmem(a_);
    float m = absolute ? 0.0f : -INFINITY;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        m = fmaxf(m, absolute ? fabsf(a[i]) : a[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (threadIdx.x % 64 == 0) lds[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kRedThreads / 64; w++) m = fmaxf(m, lds[w]);
        gmem(partial_)[blockIdx.x] = m;
    }
}

// PCG direction (src/PCG.c:113-119,194-212): gp = C*grad;
// dir = gp + last*beta when the iteration conjugates; p_last <- gp;
// c_last <- dir when beta was formed
__global__ void k_pcg_scale(const float *grad_, float C, float *dir_, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gmem(dir_)[i] = C * gmem(grad_)[i];
}

__global__ void k_pcg_direction(float *dir_, float *p_last_, float *c_last_, long n, float beta, int form,
                                int store_c)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const auto dir = gmem(dir_), p_last = gmem(p_last_), c_last = gmem(c_last_);
    const float gp = dir[i];
    float d = gp;
    if (form == 1) d = gp + p_last[i] * beta;
    if (form == 2) d = gp + c_last[i] * beta;
    p_last[i] = gp;
    if (store_c) c_last[i] = d;
    dir[i] = d;
}

// parabola L2(e) = x1 + x2 e + x3 e^2 through three points, FP64, with the
// clamps of src/calc_opt_step.c:95-112
__global__ void k_step_length(const float *in_, float *out_)
{
    if (threadIdx.x || blockIdx.x) return;
    const auto in = gmem(in_);
    double A[4][4], b[4], x[4] = {0, 0, 0, 0};
    float L2t[4], epst[4];
    for (int i = 1; i <= 3; i++) {
        L2t[i] = in[i - 1];
        epst[i] = in[3 + i - 1];
        A[i][1] = 1.0;
        A[i][2] = epst[i];
        A[i][3] = (double)epst[i] * (double)epst[i];
        b[i] = L2t[i];
    }
    // Gauss elimination in the order of src/solvelin.c:51-71
    for (int k = 1; k <= 2; k++)
        for (int nn = k; nn <= 2; nn++) {
            const double a = A[nn + 1][k] / A[k][k];
            for (int m = 1; m <= 3; m++) A[nn + 1][m] -= a * A[k][m];
            b[nn + 1] -= a * b[k];
        }
    for (int k = 3; k >= 1; k--) {
        double c = b[k];
        for (int m = 3; m >= k + 1; m--) c -= A[k][m] * x[m];
        x[k] = c / A[k][k];
    }
    float opteps = (float)(-x[2] / (2.0 * x[3]));
    if (((2.0 * x[3]) < 0.0) && (L2t[2] > L2t[3])) opteps = epst[3];
    if (opteps > epst[3]) opteps = epst[3];
    if (opteps < 0.0f) opteps = epst[1];
    *gmem(out_) = opteps;
}

// bounded model update of one cell, src/calc_mat_change_test.c:225-320
// (PARAMETERIZATION 1, elastic)
__global__ void k_model_update(long n, ifos_update_params u, float e_vp, float e_vs, float e_rho, const float *vp_,
                               const float *vs_, const float *rho_, const float *gvp_, const float *gvs_,
                               const float *grho_, const float *vp0_, const float *vs0_, const float *rho0_,
                               float *vp1_, float *vs1_, float *rho1_)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const auto vp = gmem(vp_), vs = gmem(vs_), rho = gmem(rho_), vp0 = gmem(vp0_), vs0 = gmem(vs0_),
               rho0 = gmem(rho0_);
    float pinp1 = vp[e] - e_vp * gmem(gvp_)[e];
    float unp1 = vs[e] - e_vs * gmem(gvs_)[e];
    float rhonp1 = rho[e] - e_rho * gmem(grho_)[e];
    if (pinp1 < u.VPLOWERLIM) pinp1 = u.VPLOWERLIM;
    if (pinp1 > u.VPUPPERLIM) pinp1 = u.VPUPPERLIM;
    if ((unp1 < u.VSLOWERLIM) && (unp1 > 1e-6)) unp1 = u.VSLOWERLIM;
    if (unp1 > u.VSUPPERLIM) unp1 = vs[e];
    if (u.VP_VS_RATIO > 1) {
        if ((pinp1 / unp1) < u.VP_VS_RATIO) pinp1 = unp1 * u.VP_VS_RATIO;
    }
    if (u.S == 1) {
        float pro_vs = ((vs0[e] - unp1) / vs0[e]) * 100.0;
        if (((fabs(pro_vs)) > u.S_VS) && (pro_vs > 0.0)) unp1 = vs0[e] - vs0[e] * u.S_VS / 100.0;
        if (((fabs(pro_vs)) > u.S_VS) && (pro_vs < 0.0)) unp1 = vs0[e] + vs0[e] * u.S_VS / 100.0;
        float pro_vp = ((vp0[e] - pinp1) / vp0[e]) * 100.0;
        if (((fabs(pro_vp)) > u.S_VP) && (pro_vp > 0.0)) pinp1 = vp0[e] - vp0[e] * u.S_VP / 100.0;
        if (((fabs(pro_vp)) > u.S_VP) && (pro_vp < 0.0)) pinp1 = vp0[e] + vp0[e] * u.S_VP / 100.0;
        float pro_rho = ((rho0[e] - rhonp1) / rho0[e]) * 100.0;
        if (((fabs(pro_rho)) > u.S_RHO) && (pro_rho > 0.0)) rhonp1 = rho0[e] - rho0[e] * u.S_RHO / 100.0;
        if (((fabs(pro_rho)) > u.S_RHO) && (pro_rho < 0.0)) rhonp1 = rho0[e] + rho0[e] * u.S_RHO / 100.0;
    }
    if ((rhonp1 < u.RHOLOWERLIM) && (u.INV_RHO_ITER < u.iter)) rhonp1 = rho[e];
    if ((rhonp1 > u.RHOUPPERLIM) && (u.INV_RHO_ITER < u.iter)) rhonp1 = rho[e];
    if (pinp1 < 0.0) pinp1 = vp[e];
    if (unp1 < 0.0) unp1 = vs[e];
    if (rhonp1 < 0.0) rhonp1 = rho[e];
    gmem(vp1_)[e] = pinp1;
    gmem(vs1_)[e] = unp1;
    gmem(rho1_)[e] = rhonp1;
}

// y = a*x + y style updates of the L-BFGS recursion
__global__ void k_axpy(float *y_, const float *x_, float a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gmem(y_)[i] = gmem(y_)[i] + gmem(x_)[i] * a;
}

__global__ void k_scale_into(float *y_, const float *x_, float a, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gmem(y_)[i] = a * gmem(x_)[i];
}

// s = (m1 - m0) / avg (src/calc_mat_change_test.c:299-316)
__global__ void k_model_diff(float *s_, const float *m1_, const float *m0_, float avg, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gmem(s_)[i] = (gmem(m1_)[i] - gmem(m0_)[i]) / avg;
}

inline unsigned blocks_for(long n, int t) { return (unsigned)((n + t - 1) / t); }

#define ENGINE_OK(expr)      \
    do {                     \
        int rc_ = (expr);    \
        if (rc_) return rc_; \
    } while (0)

}  // namespace

struct ifos_opt {
    int device = 0;
    hipStream_t stream = nullptr;
    double *partial = nullptr;  // [kRedBlocks] then one result slot
    float *fpart = nullptr;     // [kRedBlocks] then scratch for the step length
    float *work = nullptr;      // gradient post-processing: taper coefficients / smoothing input
    unsigned char *mask = nullptr;
    size_t work_cap = 0, mask_cap = 0;
};

extern "C" int ifos_opt_create(int device, ifos_opt **out)
{
    *out = nullptr;
    ifos_opt *o = new ifos_opt();
    o->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&o->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&o->partial, sizeof(double) * (kRedBlocks + 8)) != hipSuccess ||
        hipMalloc(&o->fpart, sizeof(float) * (kRedBlocks + 16)) != hipSuccess) {
        ifos_set_error("optimizer context: device setup failed");
        ifos_opt_destroy(o);
        return -1;
    }
    *out = o;
    return 0;
}

extern "C" void ifos_opt_destroy(ifos_opt *o)
{
    if (!o) return;
    hipSetDevice(o->device);
    if (o->stream) hipStreamSynchronize(o->stream);
    hipFree(o->partial);
    hipFree(o->fpart);
    hipFree(o->work);
    hipFree(o->mask);
    if (o->stream) hipStreamDestroy(o->stream);
    delete o;
}

// FP64 sum of a*b (mode 0) or a*(a-b) (mode 1) on stream st, with the
// kRedBlocks partials and the result slot in partial[0 .. kRedBlocks]
static int dot_on(hipStream_t st, double *partial, const float *a, const float *b, long n, int mode, double *out)
{
    if (n < 0) return ifos_set_error("negative length"), -2;
    hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, st, a, b, n, mode, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, st, partial, kRedBlocks, partial + kRedBlocks);
    OPT_TRY(hipGetLastError());
    OPT_TRY(hipMemcpyAsync(out, partial + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st));
    OPT_TRY(hipStreamSynchronize(st));
    return 0;
}

static int dot_mode(ifos_opt *o, const float *a, const float *b, long n, int mode, double *out)
{
    OPT_TRY(hipSetDevice(o->device));
    return dot_on(o->stream, o->partial, a, b, n, mode, out);
}

extern "C" int ifos_opt_dot(ifos_opt *o, const float *a, const float *b, long n, double *out)
{
    return dot_mode(o, a, b, n, 0, out);
}

extern "C" int ifos_opt_max(ifos_opt *o, const float *a, long n, int absolute, float *out)
{
    OPT_TRY(hipSetDevice(o->device));
    if (n < 1) return ifos_set_error("empty array"), -2;
    hipLaunchKernelGGL(k_max_partial, dim3(kRedBlocks), dim3(kRedThreads), 0, o->stream, a, n, absolute, o->fpart);
    OPT_TRY(hipGetLastError());
    std::vector<float> h(kRedBlocks);
    OPT_TRY(hipMemcpyAsync(h.data(), o->fpart, sizeof(float) * kRedBlocks, hipMemcpyDeviceToHost, o->stream));
    OPT_TRY(hipStreamSynchronize(o->stream));
    float m = h[0];
    for (int b = 1; b < kRedBlocks; b++) m = fmaxf(m, h[b]);
    *out = m;
    return 0;
}

// PCG direction on stream st with reduction scratch partial[kRedBlocks + 1]
static int pcg_on(hipStream_t st, double *partial, long n, const float *grad, float C, float *p_last, float *c_last,
                  int iter, int pcg_iter_start, float *dir, double *beta_out)
{
    int use1 = 1, use2 = 1;
    if ((iter - pcg_iter_start) < 2) {
        use2 = 0;
        if ((iter - pcg_iter_start) < 1) use1 = 0;
    }
    hipLaunchKernelGGL(k_pcg_scale, dim3(blocks_for(n, 256)), dim3(256), 0, st, grad, C, dir, n);
    OPT_TRY(hipGetLastError());
    const bool conj = (iter > 1) && use1;
    float beta = 0.0f;
    double b64 = 0.0;
    if (conj) {
        double num = 0.0, den = 0.0;
        ENGINE_OK(dot_on(st, partial, dir, p_last, n, 1, &num));
        ENGINE_OK(dot_on(st, partial, p_last, p_last, n, 0, &den));
        if (den != 0.0) b64 = num / den;
        if (b64 < 0.0) b64 = 0.0;
        beta = (float)b64;
    }
    const int form = !conj ? 0 : (iter == 2 ? 1 : ((iter > 2 && use2) ? 2 : 0));
    hipLaunchKernelGGL(k_pcg_direction, dim3(blocks_for(n, 256)), dim3(256), 0, st, dir, p_last, c_last, n, beta,
                       form, conj ? 1 : 0);
    OPT_TRY(hipGetLastError());
    OPT_TRY(hipStreamSynchronize(st));
    if (beta_out) *beta_out = b64;
    return 0;
}

extern "C" int ifos_opt_pcg(ifos_opt *o, long n, const float *grad, float C, float *p_last, float *c_last, int iter,
                            int pcg_iter_start, float *dir, double *beta_out)
{
    OPT_TRY(hipSetDevice(o->device));
    return pcg_on(o->stream, o->partial, n, grad, C, p_last, c_last, iter, pcg_iter_start, dir, beta_out);
}

// ---------------------------------------------------------------------------
// ifos_opt_pcg_noalloc: the PCG direction with its reduction scratch in a
// file-scope device array instead of an ifos_opt context, for callers that
// cannot allocate device memory.  Each GPU holds its own copy of the array
// for the life of the process; host code never passes the array's name
// (a host-side stand-in, not a device address) to a kernel, only the device
// address hipGetSymbolAddress reports for the current device.  Every entry
// of the scratch is written before it is read (each block writes its
// partial, then the sum writes the result slot), so nothing is cleared.
constexpr int kNoAllocScratch = kRedBlocks + 8;  // doubles: partials + result slot
static __device__ __align__(256) double scratch[kNoAllocScratch];

constexpr int kNoAllocMaxDevices = 64;
constexpr long kNoAllocMaxCells = 2147483647L;  // element-wise grids stay below 2^23 blocks
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kNoAllocMaxDevices];

static int noalloc_scratch(double **out)
{
    int dev = -1;
    OPT_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return ifos_set_error("device %d beyond the %d supported", dev, kNoAllocMaxDevices), -2;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double *p = nullptr;
        OPT_TRY(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        if (!p) return ifos_set_error("no device address for the reduction scratch on device %d", dev), -1;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" int ifos_opt_pcg_noalloc(long n, const float *grad, float C, float *p_last, float *c_last, int iter,
                                    int pcg_iter_start, float *dir, double *beta_out)
{
    if (n < 0 || n > kNoAllocMaxCells)
        return ifos_set_error("ifos_opt_pcg_noalloc: %ld cells, the limit is %ld", n, kNoAllocMaxCells), -2;
    double *partial = nullptr;
    ENGINE_OK(noalloc_scratch(&partial));
    return pcg_on(nullptr, partial, n, grad, C, p_last, c_last, iter, pcg_iter_start, dir, beta_out);
}

extern "C" int ifos_opt_step_length(ifos_opt *o, const float *L2t, const float *epst, float *opteps)
{
    OPT_TRY(hipSetDevice(o->device));
    float h[6] = {L2t[0], L2t[1], L2t[2], epst[0], epst[1], epst[2]};
    OPT_TRY(hipMemcpyAsync(o->fpart + kRedBlocks, h, sizeof h, hipMemcpyHostToDevice, o->stream));
    hipLaunchKernelGGL(k_step_length, dim3(1), dim3(64), 0, o->stream, o->fpart + kRedBlocks, o->fpart + kRedBlocks + 8);
    OPT_TRY(hipGetLastError());
    OPT_TRY(hipMemcpyAsync(opteps, o->fpart + kRedBlocks + 8, sizeof(float), hipMemcpyDeviceToHost, o->stream));
    OPT_TRY(hipStreamSynchronize(o->stream));
    return 0;
}

extern "C" int ifos_opt_model_update(ifos_opt *o, long n, const ifos_update_params *u, const float *vp,
                                     const float *vs, const float *rho, const float *g_vp, const float *g_vs,
                                     const float *g_rho, const float *vp0, const float *vs0, const float *rho0,
                                     float *vp1, float *vs1, float *rho1, float *eps)
{
    OPT_TRY(hipSetDevice(o->device));
    // maxima of the models (from 0) and of the direction magnitudes
    float mm[3], gm[3];
    const float *m[3] = {vp, vs, rho}, *g[3] = {g_vp, g_vs, g_rho};
    for (int k = 0; k < 3; k++) {
        ENGINE_OK(ifos_opt_max(o, m[k], n, 0, &mm[k]));
        ENGINE_OK(ifos_opt_max(o, g[k], n, 1, &gm[k]));
        if (mm[k] < 0.0f) mm[k] = 0.0f;
    }
    // step lengths, src/calc_mat_change_test.c:166-211 (FP32 as there)
    const bool nonorm = u->GRAD_METHOD == 2 && (u->iter > u->LBFGS_iter_start) && u->LBFGS_STEP_LENGTH;
    const int inv[3] = {u->INV_VP_ITER, u->INV_VS_ITER, u->INV_RHO_ITER};
    float e[3];
    for (int k = 0; k < 3; k++) {
        e[k] = u->eps_scale * (mm[k] / gm[k]);
        if (u->iter < inv[k]) e[k] = 0.0f;
        if (nonorm) {
            e[k] = u->eps_scale;
            if (u->iter < inv[k]) e[k] = 0.0f;
        }
        eps[k] = e[k];
    }
    hipLaunchKernelGGL(k_model_update, dim3(blocks_for(n, 256)), dim3(256), 0, o->stream, n, *u, e[0], e[1], e[2], vp,
                       vs, rho, g_vp, g_vs, g_rho, vp0, vs0, rho0, vp1, vs1, rho1);
    OPT_TRY(hipGetLastError());
    OPT_TRY(hipStreamSynchronize(o->stream));
    return 0;
}

extern "C" int ifos_opt_lbfgs_store_s(ifos_opt *o, long ncell, int npar, const float *const *m1,
                                      const float *const *m0, const float *avg, float *s_hist, int N_LBFGS, int iter)
{
    OPT_TRY(hipSetDevice(o->device));
    int w = iter % N_LBFGS;
    if (w == 0) w = N_LBFGS;
    const long nvec = ncell * npar;
    for (int k = 0; k < npar; k++)
        hipLaunchKernelGGL(k_model_diff, dim3(blocks_for(ncell, 256)), dim3(256), 0, o->stream,
                           s_hist + (long)(w - 1) * nvec + k * ncell, m1[k], m0[k], avg[k], ncell);
    OPT_TRY(hipGetLastError());
    OPT_TRY(hipStreamSynchronize(o->stream));
    return 0;
}

extern "C" int ifos_opt_lbfgs(ifos_opt *o, long ncell, int npar, float *const *grad, const float *avg, float *s_hist,
                              float *y_hist, double *rho_hist, float *work, int N_LBFGS, int iteration,
                              int lbfgs_iter_start)
{
    OPT_TRY(hipSetDevice(o->device));
    const long nvec = ncell * npar;
    auto vec = [&](float *base, int w) { return base + (long)(w - 1) * nvec; };
    const unsigned nb = blocks_for(ncell, 256);
    if (iteration == lbfgs_iter_start) {
        // src/LBFGS.c:84-100: y_w = -grad * avg
        int w = iteration % N_LBFGS;
        if (w == 0) w = N_LBFGS;
        for (int k = 0; k < npar; k++)
            hipLaunchKernelGGL(k_scale_into, dim3(nb), dim3(256), 0, o->stream, vec(y_hist, w) + k * ncell, grad[k],
                               -avg[k], ncell);
        OPT_TRY(hipGetLastError());
        OPT_TRY(hipStreamSynchronize(o->stream));
        return 0;
    }
    if (iteration < lbfgs_iter_start) return 0;
    // y_w += grad*avg (completes grad(i) - grad(i-1)); q = grad*avg (src/LBFGS.c:122-156)
    int w = (iteration - 1) % N_LBFGS;
    if (w == 0) w = N_LBFGS;
    float *q = work, *r = work + nvec;
    for (int k = 0; k < npar; k++) {
        hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(256), 0, o->stream, vec(y_hist, w) + k * ncell, grad[k], avg[k],
                           ncell);
        hipLaunchKernelGGL(k_scale_into, dim3(nb), dim3(256), 0, o->stream, q + k * ncell, grad[k], avg[k], ncell);
    }
    OPT_TRY(hipGetLastError());
    // two-loop recursion (src/LBFGS.c:226-322) with FP64 dot products
    int m = iteration - N_LBFGS;
    if (m < 1) m = 1;
    double ys = 0.0, yy = 0.0;
    ENGINE_OK(ifos_opt_dot(o, vec(y_hist, w), vec(s_hist, w), nvec, &ys));
    ENGINE_OK(ifos_opt_dot(o, vec(y_hist, w), vec(y_hist, w), nvec, &yy));
    rho_hist[w - 1] = 1.0 / ys;
    const double h0 = ys / yy;
    std::vector<double> alpha(N_LBFGS + 1, 0.0);
    const unsigned nbv = blocks_for(nvec, 256);
    for (int v = iteration - 1; v >= m; v--) {
        int wv = v % N_LBFGS;
        if (wv == 0) wv = N_LBFGS;
        double sq = 0.0;
        ENGINE_OK(ifos_opt_dot(o, vec(s_hist, wv), q, nvec, &sq));
        alpha[wv] = rho_hist[wv - 1] * sq;
        hipLaunchKernelGGL(k_axpy, dim3(nbv), dim3(256), 0, o->stream, q, vec(y_hist, wv), (float)-alpha[wv], nvec);
    }
    hipLaunchKernelGGL(k_scale_into, dim3(nbv), dim3(256), 0, o->stream, r, q, (float)h0, nvec);
    for (int v = m; v <= iteration - 1; v++) {
        int wv = v % N_LBFGS;
        if (wv == 0) wv = N_LBFGS;
        double yr = 0.0;
        ENGINE_OK(ifos_opt_dot(o, vec(y_hist, wv), r, nvec, &yr));
        const double beta = rho_hist[wv - 1] * yr;
        hipLaunchKernelGGL(k_axpy, dim3(nbv), dim3(256), 0, o->stream, r, vec(s_hist, wv), (float)(alpha[wv] - beta),
                           nvec);
    }
    OPT_TRY(hipGetLastError());
    // y for the next iteration and the denormalised direction (src/LBFGS.c:166-200)
    w = iteration % N_LBFGS;
    if (w == 0) w = N_LBFGS;
    for (int k = 0; k < npar; k++) {
        hipLaunchKernelGGL(k_scale_into, dim3(nb), dim3(256), 0, o->stream, vec(y_hist, w) + k * ncell, grad[k],
                           -avg[k], ncell);
        hipLaunchKernelGGL(k_scale_into, dim3(nb), dim3(256), 0, o->stream, grad[k], r + k * ncell, avg[k], ncell);
    }
    OPT_TRY(hipGetLastError());
    OPT_TRY(hipStreamSynchronize(o->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Gradient post-processing: tapers (coefficients built on the host,
// ifos_taper_coeff) and the smoothing of smooth() (src/smooth.c:24).
// ---------------------------------------------------------------------------
namespace {

__global__ __launch_bounds__(256) void k_taper(float *g, const float *coef, const unsigned char *zero, long n)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float v = g[e] * coef[e];
    g[e] = (zero && zero[e]) ? 0.0f : v;
}

// order-preserving map of a float's bits to an unsigned key
__device__ __forceinline__ unsigned int fkey(float v)
{
    const unsigned int u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float funkey(unsigned int k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

constexpr int kSmTile = 16;        // outputs per workgroup side
constexpr int kSmMaxHalf = 56;     // (16 + 2*56)^2 keys = 64 KiB of LDS

// One workgroup: a 16 x 16 block of outputs.  The (16+2h)^2 window union
// is loaded once into LDS as keys, with the grid's edge values repeated
// outside (src/smooth.c:87-104).  Each thread then selects element
// rank = k/2 + 1 of its window with the (+h,+h) corner taken as +0, the
// set median2d sorts (src/median2D.c:9-16, quicksort over t[0..k-1]), by a
// 32-step bitwise search: the largest key x with #{key < x} <= rank.
__global__ __launch_bounds__(256) void k_smooth(const float *in, float *out, int NX, int NY, int h)
{
    extern __shared__ unsigned int keys[];
    const int W = kSmTile + 2 * h;
    const int x0 = blockIdx.x * kSmTile, y0 = blockIdx.y * kSmTile;
    for (int q = threadIdx.x; q < W * W; q += blockDim.x) {
        const int ty = q / W, tx = q - ty * W;
        const int x = min(max(x0 + tx - h, 0), NX - 1), y = min(max(y0 + ty - h, 0), NY - 1);
        keys[q] = fkey(in[(long)y * NX + x]);
    }
    __syncthreads();
    const int lx = threadIdx.x % kSmTile, ly = threadIdx.x / kSmTile;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= NX || y >= NY) return;
    const int side = 2 * h + 1, rank = side * side / 2 + 1;
    const unsigned int corner = keys[(ly + 2 * h) * W + lx + 2 * h], zero = fkey(0.0f);
    unsigned int prefix = 0;
    for (int bit = 31; bit >= 0; bit--) {
        const unsigned int cand = prefix | (1u << bit);
        int cnt = (int)(zero < cand) - (int)(corner < cand);
        for (int r = 0; r < side; r++) {
            const unsigned int *row = keys + (ly + r) * W + lx;
            for (int c = 0; c < side; c++) cnt += row[c] < cand;
        }
        if (cnt <= rank) prefix = cand;
    }
    out[(long)y * NX + x] = funkey(prefix);
}

// spat_filt (src/spat_filt.c:27): per output cell of the window a weighted
// mean along x, accumulated in float with double products in the
// reference's order, acc += g * pref * e[d], norm += pref * e[d]
constexpr int kSpatMax = 64;
struct SpatArgs {
    const float *in;
    float *out;
    int NX, NY, S, J1;
    double pref;
    double e[2 * kSpatMax + 1];
};

__global__ __launch_bounds__(256) void k_spat_filt(SpatArgs a)
{
    const int i = a.S + 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = a.J1 + (int)blockIdx.y;
    if (i > a.NX - a.S || j > a.NY - a.S) return;
    const float *row = a.in + (long)(j - 1) * a.NX;
    float acc = 0.0f, norm = 0.0f;
    for (int i1 = i - a.S; i1 <= i + a.S; i1++) {
        const double w = a.e[i1 - i + a.S];
        acc = (float)((double)acc + ((double)row[i1 - 1] * a.pref) * w);
        norm = (float)((double)norm + a.pref * w);
    }
    a.out[(long)(j - 1) * a.NX + (i - 1)] = acc / norm;
}

int ensure_work(ifos_opt *o, size_t floats, size_t bytes)
{
    if (floats > o->work_cap) {
        hipFree(o->work);
        o->work = nullptr;
        o->work_cap = 0;
        OPT_TRY(hipMalloc(&o->work, sizeof(float) * floats));
        o->work_cap = floats;
    }
    if (bytes > o->mask_cap) {
        hipFree(o->mask);
        o->mask = nullptr;
        o->mask_cap = 0;
        OPT_TRY(hipMalloc(&o->mask, bytes));
        o->mask_cap = bytes;
    }
    return 0;
}

}  // namespace

extern "C" int ifos_opt_taper(ifos_opt *o, long n, const float *coef, const unsigned char *zero, float *grad)
{
    OPT_TRY(hipSetDevice(o->device));
    if (n < 0 || !coef || !grad) return ifos_set_error("ifos_opt_taper: bad arguments"), -2;
    if (n == 0) return 0;
    if (ensure_work(o, (size_t)n, zero ? (size_t)n : 0)) return -1;
    OPT_TRY(hipMemcpyAsync(o->work, coef, sizeof(float) * n, hipMemcpyHostToDevice, o->stream));
    if (zero) OPT_TRY(hipMemcpyAsync(o->mask, zero, (size_t)n, hipMemcpyHostToDevice, o->stream));
    hipLaunchKernelGGL(k_taper, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, o->stream, grad, o->work,
                       zero ? o->mask : (const unsigned char *)nullptr, n);
    OPT_TRY(hipGetLastError());
    OPT_TRY(hipStreamSynchronize(o->stream));
    return 0;
}

extern "C" int ifos_opt_smooth(ifos_opt *o, int NX, int NY, int filtsize, float *grad, int *side)
{
    OPT_TRY(hipSetDevice(o->device));
    if (NX < 1 || NY < 1 || !grad) return ifos_set_error("ifos_opt_smooth: bad arguments"), -2;
    if (side) *side = 0;
    if (filtsize == 0) return 0;
    if (!(filtsize % 2)) filtsize += filtsize > 0 ? 1 : -1;
    const int s = filtsize < 0 ? -filtsize : filtsize, h = s / 2;
    if (h > kSmMaxHalf)
        return ifos_set_error("ifos_opt_smooth: window side %d exceeds %d", s, 2 * kSmMaxHalf + 1), -2;
    if (side) *side = s;
    if (h == 0) return 0;  // a 1 x 1 window returns each value itself
    const size_t n = (size_t)NX * NY;
    if (ensure_work(o, n, 0)) return -1;
    OPT_TRY(hipMemcpyAsync(o->work, grad, sizeof(float) * n, hipMemcpyDeviceToDevice, o->stream));
    const int W = kSmTile + 2 * h;
    dim3 grid((NX + kSmTile - 1) / kSmTile, (NY + kSmTile - 1) / kSmTile);
    hipLaunchKernelGGL(k_smooth, grid, dim3(kSmTile * kSmTile), sizeof(unsigned int) * W * W, o->stream, o->work,
                       grad, NX, NY, h);
    OPT_TRY(hipGetLastError());
    OPT_TRY(hipStreamSynchronize(o->stream));
    return 0;
}

extern "C" int ifos_opt_spat_filt(ifos_opt *o, int NX, int NY, int SPAT_FILT_SIZE, int SPAT_FILT_1, float *grad)
{
    OPT_TRY(hipSetDevice(o->device));
    if (NX < 1 || NY < 1 || !grad) return ifos_set_error("ifos_opt_spat_filt: bad arguments"), -2;
    const int S = SPAT_FILT_SIZE;
    if (S < 1) return 0;
    if (S > kSpatMax) return ifos_set_error("ifos_opt_spat_filt: SPAT_FILT_SIZE %d exceeds %d", S, kSpatMax), -2;
    const int ni = NX - 2 * S, nj = NY - S - SPAT_FILT_1 + 1;
    if (ni < 1 || nj < 1) return 0;
    const size_t n = (size_t)NX * NY;
    if (ensure_work(o, n, 0)) return -1;
    OPT_TRY(hipMemcpyAsync(o->work, grad, sizeof(float) * n, hipMemcpyDeviceToDevice, o->stream));
    SpatArgs a;
    a.in = o->work;
    a.out = grad;
    a.NX = NX;
    a.NY = NY;
    a.S = S;
    a.J1 = SPAT_FILT_1;
    // the weight of offset d, with the reference's integer (-1/2) exponent
    // factor (src/spat_filt.c:97-98), evaluated on the host
    const double PI = 3.141592653589793;
    a.pref = 1 / (sqrt(2 * PI) * S);
    for (int d = -S; d <= S; d++) {
        const double r = sqrt((double)(d * d + d * d));
        a.e[d + S] = exp((-1 / 2) * (r * r) / (S * S));
    }
    dim3 b(256), gr((ni + 255) / 256, nj);
    hipLaunchKernelGGL(k_spat_filt, gr, b, 0, o->stream, a);
    OPT_TRY(hipGetLastError());
    OPT_TRY(hipStreamSynchronize(o->stream));
    return 0;
}
