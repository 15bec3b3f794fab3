// ifos_dropin.hip -- entry points with the reference's names and argument
// lists (src/fd.h of IFOS2D 2.0.3), running the gfx950 kernels on
// device-resident Numerical-Recipes arrays.
//
// Device NR arrays: ifos_dev_matrix(nrl,nrh,ncl,nch) returns a host
// row-pointer array of the shape util.c's matrix() builds
// (src/util.c:184-206) whose rows point into one device allocation;
// ifos_dev_f3tensor and ifos_dev_vector likewise (src/util.c:279, :95).
// A registry keyed by the returned pointer recovers base and bounds, so the
// kernels see plain origin pointers and pitches.  Host-side arrays the
// reference keeps on the host (srcpos_loc, signals, seismogram sections,
// the CPML profile vectors, hc) may stay host arrays: they are staged per
// call.  Work is queued on one library stream in call order; ifos_dev_sync()
// waits for it (required before the host reads device arrays back).
//
// Parameters are read from the reference's global variables by their names
// (src/globvar.h).  This library defines them, so it also runs standalone; a
// driver that defines them itself (IFOS2D.c includes globvar.h) supplies its
// own definitions, which take precedence at dynamic link time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "host/ifos_setup.h"
#include "ifos2d_dropin.h"
#include "ifos2d_hip.h"
#include "ifos_cellmath.h"
#include "ifos_internal.h"

using namespace ifos;

namespace ifos {
void set_step_control(const int *nt_dev, int DTINV, int NT);
void launch_inject_serial(const PointArgs &a, const StepGeom &g, float *vx, float *vy, const float *rip,
                          const float *rjp, int sw, int ADJOINT_TYPE, const int *nt_dev, int nt, int *nt_tick,
                          hipStream_t s);
void launch_inject(const PointArgs &a, const StepGeom &g, float *vx, float *vy, const float *rip, const float *rjp,
                   int sw, int ADJOINT_TYPE, const int *nt_dev, int nt, int *nt_tick, hipStream_t s);
}
extern "C" void ifos_set_error(const char *fmt, ...);

// ---------------------------------------------------------------------------
// reference globals read by the kernels (src/globvar.h)
// ---------------------------------------------------------------------------
extern "C" {
float DT, DH, F_REF, *FL;
int NX, NY, NT, FDORDER, FW, FREE_SURF, BOUNDARY, L, PARAMETERIZATION, VELOCITY, ADJOINT_TYPE, SOURCE_TYPE;
int NPROCX = 1, NPROCY = 1, MYID, POS[3], WAVETYPE = 1, ACOUSTIC, SEISMO = 1;
}

// ---------------------------------------------------------------------------
// registry of device NR arrays
// ---------------------------------------------------------------------------
namespace {

struct DevArr {
    void *dev;        // device allocation
    void *rows;       // host row-pointer block (as allocated)
    void *rows2;      // second level for f3tensor
    int kind;         // 1 vector, 2 matrix, 3 f3tensor
    int nrl, nrh, ncl, nch, ndl, ndh;
    size_t bytes;
};

std::mutex g_mu;
std::unordered_map<const void *, DevArr> g_reg;
hipStream_t g_stream = nullptr;

hipStream_t stream()
{
    if (!g_stream) hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
    return g_stream;
}

[[noreturn]] void fatal(const char *what)
{
    // src/util.c:26-36: message, then abort of the run
    fprintf(stderr, "\n Error: %s\n", what);
    exit(1);
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
        fatal(buf);
    }
}

const DevArr *lookup(const void *p)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_reg.find(p);
    return it == g_reg.end() ? nullptr : &it->second;
}

// origin pointer and pitch of a registered device matrix
float *origin_of(float **m, int *pitch, const char *name)
{
    const DevArr *a = lookup(m);
    if (!a || a->kind != 2) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s is not a device matrix created by ifos_dev_matrix()", name);
        fatal(buf);
    }
    const int p = a->nch - a->ncl + 1;
    *pitch = p;
    return (float *)a->dev - (long)a->nrl * p - a->ncl;
}

// base of element [1][1] of a device CPML strip matrix
float *strip_of(float **m, int *pitch, const char *name)
{
    int p;
    float *o = origin_of(m, &p, name);
    *pitch = p;
    return o + p + 1;
}

// A time loop's source series on the device.  update_v_PML and psource read
// one to three samples per source and step (signals[l][nt], or nt-1..nt+1 for
// the explosive amplitude); staging them cost a copy per step.  The whole
// block signals[1..nsrc][1..NT] is uploaded when first seen and again
// whenever a sample a step reads differs, bit for bit, from the copy last
// uploaded (compared on the host every step), so each step reads the host's
// values of that moment, as the reference does.
struct SeriesCache {
    float **key = nullptr;
    int nsrc = 0, ns = 0;
    std::vector<float> mirror;  // sample t of source l at l * ns + t - 1
    float *dev = nullptr;
    size_t cap = 0;
};

// device block [nsrc][NT] (source-major) of host signals, checked at samples t0..t1
const float *series_dev(SeriesCache &c, float **signals, int nsrc, int t0, int t1)
{
    const int T = NT;
    t0 = std::max(t0, 1);
    t1 = std::min(t1, T);
    bool fresh = c.key != signals || c.nsrc != nsrc || c.ns != T;
    for (int l = 1; l <= nsrc && !fresh; l++)
        fresh = memcmp(&signals[l][t0], &c.mirror[(size_t)(l - 1) * T + t0 - 1], sizeof(float) * (t1 - t0 + 1)) != 0;
    if (fresh) {
        // kernels queued earlier may still read the old block
        hip_check(hipStreamSynchronize(stream()), "series sync");
        const size_t n = (size_t)nsrc * T;
        c.mirror.resize(n);
        for (int l = 1; l <= nsrc; l++) memcpy(&c.mirror[(size_t)(l - 1) * T], &signals[l][1], sizeof(float) * T);
        if (n > c.cap) {
            if (c.dev) hip_check(hipFree(c.dev), "hipFree");
            hip_check(hipMalloc(&c.dev, sizeof(float) * n), "hipMalloc(series)");
            c.cap = n;
        }
        hip_check(hipMemcpyAsync(c.dev, c.mirror.data(), sizeof(float) * n, hipMemcpyHostToDevice, stream()),
                  "series upload");
        c.key = signals;
        c.nsrc = nsrc;
        c.ns = T;
    }
    return c.dev;
}

// small per-loop tables (positions, types, azimuths) on the device,
// uploaded again only when their bytes change
struct BlobCache {
    std::vector<char> host;
    char *dev = nullptr;
    size_t cap = 0;
};

const char *blob_dev(BlobCache &c, const void *data, size_t n)
{
    if (c.host.size() == n && !memcmp(c.host.data(), data, n)) return c.dev;
    hip_check(hipStreamSynchronize(stream()), "table sync");
    c.host.assign((const char *)data, (const char *)data + n);
    if (n > c.cap) {
        if (c.dev) hip_check(hipFree(c.dev), "hipFree");
        hip_check(hipMalloc(&c.dev, n), "hipMalloc(table)");
        c.cap = n;
    }
    hip_check(hipMemcpyAsync(c.dev, c.host.data(), n, hipMemcpyHostToDevice, stream()), "table upload");
    return c.dev;
}

SeriesCache g_vsig[2], g_psig;
BlobCache g_vsrc, g_psrc;

// device copy of a 1-based host or device vector of n entries, element 1 first
struct VecCache {
    std::vector<float> host;
    float *dev = nullptr;
};
std::unordered_map<const void *, VecCache> g_vcache;

const float *vec1_dev(const float *v, int n)
{
    if (n <= 0) n = 1;
    const DevArr *a = lookup(v);
    if (a && a->kind == 1) return (const float *)a->dev + (1 - a->nrl);
    std::lock_guard<std::mutex> lk(g_mu);
    VecCache &c = g_vcache[v];
    if (c.host.size() != (size_t)n || memcmp(c.host.data(), v + 1, sizeof(float) * n)) {
        c.host.assign(v + 1, v + 1 + n);
        if (c.dev) hip_check(hipFree(c.dev), "hipFree");
        hip_check(hipMalloc(&c.dev, sizeof(float) * n), "hipMalloc");
        hip_check(hipMemcpyAsync(c.dev, c.host.data(), sizeof(float) * n, hipMemcpyHostToDevice, stream()),
                  "profile upload");
    }
    return c.dev;
}

// per-call staging of small host data
struct Stage {
    void *dev = nullptr;
    size_t cap = 0;
    void *get(size_t n)
    {
        if (n > cap) {
            // earlier calls' kernels may still read the old buffer
            hip_check(hipStreamSynchronize(stream()), "stage sync");
            if (dev) hip_check(hipFree(dev), "hipFree");
            cap = n * 2 + 256;
            hip_check(hipMalloc(&dev, cap), "hipMalloc(stage)");
        }
        return dev;
    }
};
Stage g_stage[4];

// Pinned host staging of the small per-call uploads (source positions and
// samples): a ring of slots, each reused only after the copy that last read
// it has completed, so a call returns without waiting for the device.
struct PinnedRing {
    static constexpr int kSlots = 16;
    char *host[kSlots] = {};
    size_t cap[kSlots] = {};
    hipEvent_t ev[kSlots] = {};
    int next = 0;
    char *get(size_t n, int &slot)
    {
        slot = next;
        next = (next + 1) % kSlots;
        if (ev[slot])
            hip_check(hipEventSynchronize(ev[slot]), "staging wait");
        else
            hip_check(hipEventCreateWithFlags(&ev[slot], hipEventDisableTiming), "staging event");
        if (cap[slot] < n) {
            if (host[slot]) hip_check(hipHostFree(host[slot]), "hipHostFree");
            cap[slot] = n * 2 + 256;
            hip_check(hipHostMalloc((void **)&host[slot], cap[slot]), "hipHostMalloc(staging)");
        }
        return host[slot];
    }
    // after the copy from slot has been queued
    void queued(int slot) { hip_check(hipEventRecord(ev[slot], stream()), "staging record"); }
};
PinnedRing g_ring;

// Seismogram sections: seismo_ssg gathers each step's receiver samples into
// a device buffer [2][ntr][NT] per pair of host sections and returns at
// once; ifos_dev_sync() copies the samples gathered since the last flush
// into the host sections (the reference reads them after its time loop,
// src/IFOS2D.c:1903-1961).
struct SectionBuf {
    float **s0 = nullptr, **s1 = nullptr;  // host sections (s1 null for SEISMO 2)
    int ntr = 0, ns = 0, lo = 0, hi = -1;   // samples lo..hi (1-based) not yet flushed
    float *dev = nullptr;
    int *dpos = nullptr;
    std::vector<int> pos;                   // receiver cells of the positions in dpos
};
std::unordered_map<const void *, SectionBuf> g_sections;

__global__ void k_gather_col(int n, const int *ix, const int *iy, int pitch, const float *a, const float *b,
                             int mode, float *out, int ld, int col)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long c = (long)iy[t] * pitch + ix[t];
    out[(long)t * ld + col] = mode == 2 ? -a[c] - b[c] : a[c];
    if (mode == 1) out[(long)(n + t) * ld + col] = b[c];
}

// copy the unflushed samples of every section pair to the host (after a
// stream synchronisation)
hipError_t flush_sections()
{
    for (auto &kv : g_sections) {
        SectionBuf &sb = kv.second;
        if (sb.hi < sb.lo) continue;
        const int n = sb.s1 ? 2 * sb.ntr : sb.ntr;
        std::vector<float> h((size_t)n * sb.ns);
        hipError_t e = hipMemcpy(h.data(), sb.dev, sizeof(float) * h.size(), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return e;
        for (int t = 1; t <= sb.ntr; t++)
            for (int k = sb.lo; k <= sb.hi; k++) {
                sb.s0[t][k] = h[(size_t)(t - 1) * sb.ns + k - 1];
                if (sb.s1) sb.s1[t][k] = h[(size_t)(sb.ntr + t - 1) * sb.ns + k - 1];
            }
        sb.lo = 0;
        sb.hi = -1;
    }
    return hipSuccess;
}

StepGeom make_geom(int nx1, int nx2, int ny1, int ny2, int pitch, const float *hc, const float *K_x,
                   const float *a_x, const float *b_x, const float *K_xh, const float *a_xh, const float *b_xh,
                   const float *K_y, const float *a_y, const float *b_y, const float *K_yh, const float *a_yh,
                   const float *b_yh)
{
    if (nx1 != 1 || ny1 != 1) fatal("the device kernels update the whole grid: nx1 and ny1 must be 1");
    if (NPROCX != 1 || NPROCY != 1) fatal("one GPU holds the whole domain: NPROCX = NPROCY = 1 required");
    if (BOUNDARY != 0 && BOUNDARY != 1) fatal("BOUNDARY must be 0 or 1");
    StepGeom g;
    g.NX = nx2; g.NY = ny2; g.pitch = pitch;
    g.FW = FW; g.FREE_SURF = FREE_SURF; g.BOUNDARY = BOUNDARY;
    g.N = FDORDER / 2; g.DT = DT; g.DH = DH;
    for (int k = 0; k < 7; k++) g.hc[k] = hc[k];
    const int n = 2 * FW;
    const float *z = nullptr;
    g.pml = {K_x ? vec1_dev(K_x, n) : z, a_x ? vec1_dev(a_x, n) : z, b_x ? vec1_dev(b_x, n) : z,
             K_xh ? vec1_dev(K_xh, n) : z, a_xh ? vec1_dev(a_xh, n) : z, b_xh ? vec1_dev(b_xh, n) : z,
             K_y ? vec1_dev(K_y, n) : z, a_y ? vec1_dev(a_y, n) : z, b_y ? vec1_dev(b_y, n) : z,
             K_yh ? vec1_dev(K_yh, n) : z, a_yh ? vec1_dev(a_yh, n) : z, b_yh ? vec1_dev(b_yh, n) : z};
    g.px = n > 0 ? n : 1;
    g.py = nx2;
    g.K_one = 0;  // the drop-in layer always divides
    return g;
}

// point-wise additions in order: field[c_l] = (float)((double)field[c_l] + w*v_l)
__global__ void k_point_add(int n, const int *ix, const int *iy, const double *amp, int pitch, float *f1, float *f2)
{
    if (threadIdx.x != 0) return;
    for (int l = 0; l < n; l++) {
        const long c = (long)iy[l] * pitch + ix[l];
        f1[c] = (float)((double)f1[c] + 0.5 * amp[l]);
        if (f2) f2[c] = (float)((double)f2[c] + 0.5 * amp[l]);
    }
}

// psource (src/psource.c:26-56) for every source in order, from the device
// copy of the series (sig: sample t of source l at l * ns + t - 1): the
// amplitude is the reference's, formed in double from the float samples
// (explosive_amp for the forward run, the sample itself for the adjoint),
// then added half to sxx and half to syy
__global__ void k_psource_series(int n, const int *ix, const int *iy, const float *sig, int ns, int nt, float DH,
                                 int sw, int pitch, float *f1, float *f2)
{
    if (threadIdx.x != 0) return;
    for (int l = 0; l < n; l++) {
        const float *s = sig + (long)l * ns - 1;  // 1-based
        const float a = sw == 0 ? explosive_amp(s, nt, ns, DH) : s[nt];
        const long c = (long)iy[l] * pitch + ix[l];
        f1[c] = (float)((double)f1[c] + 0.5 * (double)a);
        if (f2) f2[c] = (float)((double)f2[c] + 0.5 * (double)a);
    }
}

// one-rank exchanges with BOUNDARY 1 (src/exchange_v.c:143-228,
// src/exchange_s.c:145-229): each row's halo columns take the opposite
// edge; a gets N columns on the right (NX+1..NX+N from 1..N) and N-1 on the
// left (2-N..0 from NX-N+2..NX), b N-1 on the right and N on the left.  One
// thread per row; the columns read and the columns written are disjoint.
__global__ void k_exchange_x(float *a, float *b, int pitch, int NX, int NY, int N)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j > NY) return;
    float *ra = a + (long)j * pitch, *rb = b + (long)j * pitch;
    for (int l = 1; l <= N; l++) ra[NX + l] = ra[l];
    for (int l = 1; l <= N - 1; l++) rb[NX + l] = rb[l];
    for (int l = 1; l <= N - 1; l++) ra[1 - l] = ra[NX - l + 1];
    for (int l = 1; l <= N; l++) rb[1 - l] = rb[NX - l + 1];
}

void exchange_x(float **a, float **b, const char *na, const char *nb)
{
    const int N = FDORDER / 2;
    if (NX < FDORDER) fatal("BOUNDARY 1 needs NX >= FDORDER");
    int Pa, Pb;
    float *oa = origin_of(a, &Pa, na), *ob = origin_of(b, &Pb, nb);
    for (float **m : {a, b}) {
        const DevArr *d = lookup(m);
        if (d->nrl > 1 || d->nrh < NY || d->ncl > 1 - N || d->nch < NX + N)
            fatal("exchange: the wavefield matrices need rows 1..NY and columns 1-FDORDER/2..NX+FDORDER/2");
    }
    if (Pa != Pb) fatal("exchange: the two wavefield matrices must have the same column range");
    hipLaunchKernelGGL(k_exchange_x, dim3((NY + 255) / 256), dim3(256), 0, stream(), oa, ob, Pa, NX, NY, N);
    hip_check(hipGetLastError(), "exchange launch");
}

__global__ void k_zero_rect(float *origin, int pitch, int j0, int j1, int i0, int i1)
{
    const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = j0 + blockIdx.y * blockDim.y + threadIdx.y;
    if (i > i1 || j > j1) return;
    origin[(long)j * pitch + i] = 0.0f;
}

void zero_dev(void *p)
{
    if (!p) return;
    const DevArr *a = lookup(p);
    if (!a) fatal("zero_fdveps: array is not a device array created by ifos_dev_*()");
    hip_check(hipMemsetAsync(a->dev, 0, a->bytes, stream()), "zero");
}

}  // namespace

// ---------------------------------------------------------------------------
// device NR arrays
// ---------------------------------------------------------------------------
extern "C" float **ifos_dev_matrix(int nrl, int nrh, int ncl, int nch)
{
    const long nr = nrh - nrl + 1, nc = nch - ncl + 1;
    float *dev = nullptr;
    hip_check(hipMalloc(&dev, sizeof(float) * nr * nc), "ifos_dev_matrix");
    hip_check(hipMemset(dev, 0, sizeof(float) * nr * nc), "ifos_dev_matrix zero");
    float **rows = (float **)malloc(sizeof(float *) * nr);
    for (long r = 0; r < nr; r++) rows[r] = dev + r * nc - ncl;
    float **m = rows - nrl;
    std::lock_guard<std::mutex> lk(g_mu);
    g_reg[m] = DevArr{dev, rows, nullptr, 2, nrl, nrh, ncl, nch, 0, 0, sizeof(float) * nr * nc};
    return m;
}

extern "C" float ***ifos_dev_f3tensor(int nrl, int nrh, int ncl, int nch, int ndl, int ndh)
{
    const long nr = nrh - nrl + 1, nc = nch - ncl + 1, nd = ndh - ndl + 1;
    float *dev = nullptr;
    hip_check(hipMalloc(&dev, sizeof(float) * nr * nc * nd), "ifos_dev_f3tensor");
    hip_check(hipMemset(dev, 0, sizeof(float) * nr * nc * nd), "ifos_dev_f3tensor zero");
    float **cols = (float **)malloc(sizeof(float *) * nr * nc);
    float ***rows = (float ***)malloc(sizeof(float **) * nr);
    for (long r = 0; r < nr; r++) {
        for (long c = 0; c < nc; c++) cols[r * nc + c] = dev + (r * nc + c) * nd - ndl;
        rows[r] = cols + r * nc - ncl;
    }
    float ***t = rows - nrl;
    std::lock_guard<std::mutex> lk(g_mu);
    g_reg[t] = DevArr{dev, rows, cols, 3, nrl, nrh, ncl, nch, ndl, ndh, sizeof(float) * nr * nc * nd};
    return t;
}

extern "C" float *ifos_dev_vector(int nl, int nh)
{
    const long n = nh - nl + 1;
    float *dev = nullptr;
    hip_check(hipMalloc(&dev, sizeof(float) * n), "ifos_dev_vector");
    hip_check(hipMemset(dev, 0, sizeof(float) * n), "ifos_dev_vector zero");
    float *v = dev - nl;
    std::lock_guard<std::mutex> lk(g_mu);
    g_reg[v] = DevArr{dev, nullptr, nullptr, 1, nl, nh, 0, 0, 0, 0, sizeof(float) * n};
    return v;
}

extern "C" void ifos_dev_free_array(void *p)
{
    DevArr a;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_reg.find(p);
        if (it == g_reg.end()) return;
        a = it->second;
        g_reg.erase(it);
    }
    hipStreamSynchronize(stream());
    hipFree(a.dev);
    free(a.rows);
    free(a.rows2);
}

// copy between a host NR array (util.c layout, contiguous) and a device one
// of the same bounds; dir 0 = host to device, 1 = device to host
extern "C" int ifos_dev_copy_array(void *dev_arr, const void *host_first_elem, int dir)
{
    const DevArr *a = lookup(dev_arr);
    if (!a) return ifos_set_error("not a device array"), -2;
    hipError_t e = hipSuccess;
    if (a->kind != 3) {
        e = dir == 0 ? hipMemcpyAsync(a->dev, host_first_elem, a->bytes, hipMemcpyHostToDevice, stream())
                     : hipMemcpyAsync((void *)host_first_elem, a->dev, a->bytes, hipMemcpyDeviceToHost, stream());
        if (e == hipSuccess) e = hipStreamSynchronize(stream());
    } else {
        // the host tensor is util.c's contiguous [j][i][l] block; the device
        // holds it plane by plane: transpose one plane at a time
        const long plane = (long)(a->nrh - a->nrl + 1) * (a->nch - a->ncl + 1);
        const int nd = a->ndh - a->ndl + 1;
        std::vector<float> buf(plane);
        float *h = (float *)host_first_elem;
        for (int l = 0; l < nd && e == hipSuccess; l++) {
            float *d = (float *)a->dev + (long)l * plane;
            if (dir == 0) {
                for (long c = 0; c < plane; c++) buf[c] = h[c * nd + l];
                e = hipMemcpyAsync(d, buf.data(), sizeof(float) * plane, hipMemcpyHostToDevice, stream());
                if (e == hipSuccess) e = hipStreamSynchronize(stream());
            } else {
                e = hipMemcpyAsync(buf.data(), d, sizeof(float) * plane, hipMemcpyDeviceToHost, stream());
                if (e == hipSuccess) e = hipStreamSynchronize(stream());
                for (long c = 0; c < plane; c++) h[c * nd + l] = buf[c];
            }
        }
    }
    if (e != hipSuccess) return ifos_set_error("copy failed: %s", hipGetErrorString(e)), -1;
    return 0;
}

extern "C" int ifos_dev_sync(void)
{
    hipError_t e = hipStreamSynchronize(stream());
    if (e == hipSuccess) e = flush_sections();
    if (e != hipSuccess) return ifos_set_error("%s", hipGetErrorString(e)), -1;
    return 0;
}

// one element of a device matrix (the NaN probe of src/IFOS2D.c:1660)
extern "C" float ifos_dev_read(float **m, int j, int i)
{
    int p;
    float *o = origin_of(m, &p, "ifos_dev_read");
    float v = 0.0f;
    hip_check(hipMemcpyAsync(&v, o + (long)j * p + i, sizeof v, hipMemcpyDeviceToHost, stream()), "read");
    hip_check(hipStreamSynchronize(stream()), "read sync");
    return v;
}

// ---------------------------------------------------------------------------
// update_v_PML  (src/update_v_PML.c:29, src/fd.h:384-389)
// ---------------------------------------------------------------------------
extern "C" void update_v_PML(int nx1, int nx2, int ny1, int ny2, int nt, float **vx, float **vxp1, float **vxm1,
                             float **vy, float **vyp1, float **vym1, float **uttx, float **utty, float **sxx,
                             float **syy, float **sxy, float **rip, float **rjp, float **srcpos_loc,
                             float **signals, float **signals1, int nsrc, float **absorb_coeff, float *hc,
                             int infoout, int sw, float *K_x, float *a_x, float *b_x, float *K_x_half,
                             float *a_x_half, float *b_x_half, float *K_y, float *a_y, float *b_y, float *K_y_half,
                             float *a_y_half, float *b_y_half, float **psi_sxx_x, float **psi_syy_y,
                             float **psi_sxy_y, float **psi_sxy_x)
{
    (void)vxm1; (void)vym1; (void)uttx; (void)utty; (void)absorb_coeff; (void)infoout;
    int P, p2;
    VArgs a;
    a.vx = origin_of(vx, &P, "vx");
    a.vy = origin_of(vy, &p2, "vy");
    a.sxx = origin_of(sxx, &p2, "sxx");
    a.syy = origin_of(syy, &p2, "syy");
    a.sxy = origin_of(sxy, &p2, "sxy");
    a.rip = origin_of(rip, &p2, "rip");
    a.rjp = origin_of(rjp, &p2, "rjp");
    a.vxp1 = origin_of(vxp1, &p2, "vxp1");
    a.vyp1 = origin_of(vyp1, &p2, "vyp1");
    a.g = make_geom(nx1, nx2, ny1, ny2, P, hc, K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y, K_y_half,
                    a_y_half, b_y_half);
    int px = 1, py = 1;
    a.psi_sxx_x = strip_of(psi_sxx_x, &px, "psi_sxx_x");
    a.psi_sxy_x = strip_of(psi_sxy_x, &px, "psi_sxy_x");
    a.psi_syy_y = strip_of(psi_syy_y, &py, "psi_syy_y");
    a.psi_sxy_y = strip_of(psi_sxy_y, &py, "psi_sxy_y");
    a.g.px = px;
    a.g.py = py;
    a.sw = sw;
    a.VELOCITY = VELOCITY;
    set_step_control(nullptr, 1, NT);
    launch_update_v(a, stream());

    if (nsrc <= 0) return;
    // point sources after the loop (src/update_v_PML.c:553-578): positions,
    // types and azimuths from the host arrays, this step's samples from the
    // device copies of the source series
    std::vector<int> pos(3 * nsrc);
    std::vector<double> sc(2 * nsrc);
    int any = 0;
    for (int l = 1; l <= nsrc; l++) {
        pos[l - 1] = (int)srcpos_loc[1][l];
        pos[nsrc + l - 1] = (int)srcpos_loc[2][l];
        if (sw == 0) {
            float azi_rad = srcpos_loc[7][l] * 3.141592653589793 / 180;
            SOURCE_TYPE = (int)srcpos_loc[8][l];
            pos[2 * nsrc + l - 1] = SOURCE_TYPE;
            sc[l - 1] = sin(azi_rad);
            sc[nsrc + l - 1] = cos(azi_rad);
            if (SOURCE_TYPE >= 2 && SOURCE_TYPE <= 4) any = 1;
        } else {
            any = 1;
        }
    }
    if (!any) return;
    if (lookup(signals) || lookup(signals1)) fatal("update_v_PML: signals are host arrays in this interface");
    if (nt < 1 || nt > NT) fatal("update_v_PML: time step outside 1..NT");
    const float *sig0 = series_dev(g_vsig[0], signals, nsrc, nt, nt);
    const float *sig1 = signals1 == signals ? sig0 : series_dev(g_vsig[1], signals1, nsrc, nt, nt);
    const size_t o1 = sizeof(int) * pos.size(), nb = o1 + sizeof(double) * sc.size();
    std::vector<char> tab(nb);
    memcpy(tab.data(), pos.data(), o1);
    memcpy(tab.data() + o1, sc.data(), sizeof(double) * sc.size());
    const char *st = blob_dev(g_vsrc, tab.data(), nb);
    PointArgs pa;
    pa.n = nsrc;
    pa.ix = (const int *)st;
    pa.iy = (const int *)st + nsrc;
    pa.type = (const int *)st + 2 * nsrc;
    pa.sin_azi = (const double *)(st + o1);
    pa.cos_azi = (const double *)(st + o1) + nsrc;
    pa.sig = sig0;
    pa.sig1 = sig1;
    pa.ld = NT;  // sample nt of source l at l * NT + nt - 1
    // in source order, as the reference loop: serially where positions may
    // repeat; adjoint sources at distinct cells are added in parallel (each
    // cell then receives one addition, so the order does not matter)
    bool distinct = sw == 1;
    if (distinct) {
        std::unordered_set<long> seen;
        for (int l = 0; l < nsrc && distinct; l++) distinct = seen.insert((long)pos[nsrc + l] * P + pos[l]).second;
    }
    if (distinct)
        launch_inject(pa, a.g, a.vx, a.vy, a.rip, a.rjp, sw, ADJOINT_TYPE, nullptr, nt, nullptr, stream());
    else
        launch_inject_serial(pa, a.g, a.vx, a.vy, a.rip, a.rjp, sw, ADJOINT_TYPE, nullptr, nt, nullptr, stream());
}

// ---------------------------------------------------------------------------
// update_s_elastic_PML  (src/update_s_elastic_PML.c:28, src/fd.h:352-357)
// ---------------------------------------------------------------------------
extern "C" void update_s_elastic_PML(int nx1, int nx2, int ny1, int ny2, float **vx, float **vy, float **ux,
                                     float **uy, float **uxy, float **uyx, float **sxx, float **syy, float **sxy,
                                     float **pi, float **u, float **uipjp, float **absorb_coeff, float **rho,
                                     float *hc, int infoout, float *K_x, float *a_x, float *b_x, float *K_x_half,
                                     float *a_x_half, float *b_x_half, float *K_y, float *a_y, float *b_y,
                                     float *K_y_half, float *a_y_half, float *b_y_half, float **psi_vxx,
                                     float **psi_vyy, float **psi_vxy, float **psi_vyx)
{
    (void)uyx; (void)absorb_coeff; (void)infoout;
    if (PARAMETERIZATION != 1) fatal("update_s_elastic_PML: only PARAMETERIZATION=1 is supported on the GPU");
    int P, p2;
    SElArgs a;
    a.vx = origin_of(vx, &P, "vx");
    a.vy = origin_of(vy, &p2, "vy");
    a.sxx = origin_of(sxx, &p2, "sxx");
    a.syy = origin_of(syy, &p2, "syy");
    a.sxy = origin_of(sxy, &p2, "sxy");
    a.ux = origin_of(ux, &p2, "ux");
    a.uy = origin_of(uy, &p2, "uy");
    a.uxy = origin_of(uxy, &p2, "uxy");
    a.uipjp = origin_of(uipjp, &p2, "uipjp");
    a.rho = origin_of(rho, &p2, "rho");
    a.vs = origin_of(u, &p2, "u");
    a.vp = origin_of(pi, &p2, "pi");
    a.f = nullptr;
    a.gl = nullptr;
    a.g = make_geom(nx1, nx2, ny1, ny2, P, hc, K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y, K_y_half,
                    a_y_half, b_y_half);
    int px = 1, py = 1;
    a.psi_vxx = strip_of(psi_vxx, &px, "psi_vxx");
    a.psi_vyx = strip_of(psi_vyx, &px, "psi_vyx");
    a.psi_vyy = strip_of(psi_vyy, &py, "psi_vyy");
    a.psi_vxy = strip_of(psi_vxy, &py, "psi_vxy");
    a.g.px = px;
    a.g.py = py;
    set_step_control(nullptr, 1, NT);
    launch_update_s_elastic(a, stream());
}

// Device tensors are stored plane by plane (third index slowest): element
// (j, i, l) at origin + j * pitch + i + (l - 1) * plane.  Returns the depth.
static int tensor_origin(float ***t, float **origin, long *plane, int pitch, const char *name)
{
    const DevArr *a = lookup(t);
    if (!a || a->kind != 3) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s is not a device tensor created by ifos_dev_f3tensor()", name);
        fatal(buf);
    }
    const int nc = a->nch - a->ncl + 1, nd = a->ndh - a->ndl + 1;
    const long pl = (long)(a->nrh - a->nrl + 1) * nc;
    if (nc != pitch) fatal("memory-variable tensors must share the wavefield grid bounds");
    *origin = (float *)a->dev - ((long)a->nrl * nc + a->ncl) + (long)(1 - a->ndl) * pl;
    *plane = pl;
    return nd;
}

// ---------------------------------------------------------------------------
// update_s_visc_PML  (src/update_s_visc_PML.c:28, src/fd.h:363-369)
// ---------------------------------------------------------------------------
extern "C" void update_s_visc_PML(int nx1, int nx2, int ny1, int ny2, float **vx, float **vy, float **ux,
                                  float **uy, float **uxy, float **uyx, float **sxx, float **syy, float **sxy,
                                  float **pi, float **u, float **uipjp, float **rho, float *hc, int infoout,
                                  float ***r, float ***p, float ***q, float **fipjp, float **f, float **g,
                                  float *bip, float *bjm, float *cip, float *cjm, float ***d, float ***e,
                                  float ***dip, float *K_x, float *a_x, float *b_x, float *K_x_half,
                                  float *a_x_half, float *b_x_half, float *K_y, float *a_y, float *b_y,
                                  float *K_y_half, float *a_y_half, float *b_y_half, float **psi_vxx,
                                  float **psi_vyy, float **psi_vxy, float **psi_vyx)
{
    (void)uyx; (void)pi; (void)u; (void)uipjp; (void)rho; (void)infoout;
    if (L < 1 || L > 5) fatal("update_s_visc_PML: L must be 1..5 on the GPU");
    int P, p2;
    SViscArgs a;
    a.L = L;
    a.vx = origin_of(vx, &P, "vx");
    a.vy = origin_of(vy, &p2, "vy");
    a.sxx = origin_of(sxx, &p2, "sxx");
    a.syy = origin_of(syy, &p2, "syy");
    a.sxy = origin_of(sxy, &p2, "sxy");
    a.ux = origin_of(ux, &p2, "ux");
    a.uy = origin_of(uy, &p2, "uy");
    a.uxy = origin_of(uxy, &p2, "uxy");
    a.fipjp = origin_of(fipjp, &p2, "fipjp");
    a.f = origin_of(f, &p2, "f");
    a.gv = origin_of(g, &p2, "g");
    float *o;
    long pl = 0;
    float ***ts[6] = {r, p, q, d, e, dip};
    const char *tn[6] = {"r", "p", "q", "d", "e", "dip"};
    float *org[6];
    for (int k = 0; k < 6; k++) {
        const int nd = tensor_origin(ts[k], &org[k], &pl, P, tn[k]);
        if (nd < L) fatal("update_s_visc_PML: memory-variable tensors need L planes");
    }
    a.r = org[0]; a.p = org[1]; a.q = org[2]; a.d = org[3]; a.e = org[4]; a.dip = org[5];
    a.cs = 1;
    a.ls = (int)pl;
    for (int l = 0; l <= kMaxL; l++) {
        a.bip[l] = (l >= 1 && l <= L) ? bip[l] : 0.0f;
        a.bjm[l] = (l >= 1 && l <= L) ? bjm[l] : 0.0f;
        a.cip[l] = (l >= 1 && l <= L) ? cip[l] : 0.0f;
        a.cjm[l] = (l >= 1 && l <= L) ? cjm[l] : 0.0f;
    }
    a.g = make_geom(nx1, nx2, ny1, ny2, P, hc, K_x, a_x, b_x, K_x_half, a_x_half, b_x_half, K_y, a_y, b_y, K_y_half,
                    a_y_half, b_y_half);
    int px = 1, py = 1;
    a.psi_vxx = strip_of(psi_vxx, &px, "psi_vxx");
    a.psi_vyx = strip_of(psi_vyx, &px, "psi_vyx");
    a.psi_vyy = strip_of(psi_vyy, &py, "psi_vyy");
    a.psi_vxy = strip_of(psi_vxy, &py, "psi_vxy");
    a.g.px = px;
    a.g.py = py;
    set_step_control(nullptr, 1, NT);
    launch_update_s_visc(a, stream());
}

// ---------------------------------------------------------------------------
// psource  (src/psource.c:26, src/fd.h:206-207): the amplitude is formed on
// the host exactly as the reference does, the addition runs on the device in
// source order
// ---------------------------------------------------------------------------
extern "C" void psource(int nt, float **sxx, float **syy, float **sp, float **srcpos_loc, float **signals, int nsrc,
                        int sw)
{
    (void)sp;
    if (nsrc <= 0) return;
    if (nt < 1 || nt > NT) fatal("psource: time step outside 1..NT");
    if (lookup(signals)) fatal("psource: signals are host arrays in this interface");
    int P, p2;
    float *oxx = origin_of(sxx, &P, "sxx"), *oyy = origin_of(syy, &p2, "syy");
    std::vector<int> pos(2 * nsrc);
    for (int l = 1; l <= nsrc; l++) {
        pos[l - 1] = (int)srcpos_loc[1][l];
        pos[nsrc + l - 1] = (int)srcpos_loc[2][l];
    }
    // the amplitude reads samples nt-1..nt+1 (sw 0) or nt (sw 1)
    const float *sig = series_dev(g_psig, signals, nsrc, sw == 0 ? nt - 1 : nt, sw == 0 ? nt + 1 : nt);
    const int *dpos = (const int *)blob_dev(g_psrc, pos.data(), sizeof(int) * pos.size());
    hipLaunchKernelGGL(k_psource_series, dim3(1), dim3(64), 0, stream(), nsrc, dpos, dpos + nsrc, sig, NT, nt, DH, sw,
                       P, oxx, oyy);
}

// ---------------------------------------------------------------------------
// surface_elastic_PML (src/surface_elastic_PML.c:26, src/fd.h:286) and
// surface_PML (src/surface_PML.c:25, src/fd.h:288)
// ---------------------------------------------------------------------------
static void surface_common(int ndepth, PostArgs &po, float **vx, float **vy, float **sxx, float **syy, float **sxy,
                           float **ux, float **uy, float **uxy, float *hc, float *K_x, float *a_x, float *b_x,
                           float **psi_vxx)
{
    if (ndepth != 1) fatal("surface: the free surface must be at ndepth = 1");
    int P, p2;
    po.vx = origin_of(vx, &P, "vx");
    po.vy = origin_of(vy, &p2, "vy");
    po.sxx = origin_of(sxx, &p2, "sxx");
    po.syy = origin_of(syy, &p2, "syy");
    po.sxy = origin_of(sxy, &p2, "sxy");
    po.ux = ux ? origin_of(ux, &p2, "ux") : nullptr;
    po.uy = uy ? origin_of(uy, &p2, "uy") : nullptr;
    po.uxy = uxy ? origin_of(uxy, &p2, "uxy") : nullptr;
    po.g = make_geom(1, NX, 1, NY, P, hc, K_x, a_x, b_x, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr, nullptr, nullptr);
    po.g.FREE_SURF = 1;
    int px = 1;
    po.psi_vxxs = strip_of(psi_vxx, &px, "psi_vxx");
    po.g.px = px;
    po.sw = 0;
    po.explosive = 0;
    po.NT = NT;
    po.src.n = 0;
    po.ntr = 0;
    po.nt_dev = nullptr;
    po.nt = 1;
    po.nan_flag = nullptr;
}

extern "C" void surface_elastic_PML(int ndepth, float **vx, float **vy, float **sxx, float **syy, float **sxy,
                                    float **syz, float **pi, float **u, float **rho, float *hc, float *K_x, float *a_x,
                                    float *b_x, float **psi_vxx, float **ux, float **uy, float **uxy, float **uyz,
                                    float **sxz, float **uxz)
{
    (void)syz; (void)uyz; (void)sxz; (void)uxz;
    if (PARAMETERIZATION != 1) fatal("surface_elastic_PML: only PARAMETERIZATION=1 is supported on the GPU");
    PostArgs po;
    memset(&po, 0, sizeof po);
    surface_common(ndepth, po, vx, vy, sxx, syy, sxy, ux, uy, uxy, hc, K_x, a_x, b_x, psi_vxx);
    int p2;
    po.L = 0;
    po.f = nullptr;
    po.gl = nullptr;
    po.rho = origin_of(rho, &p2, "rho");
    po.vs = origin_of(u, &p2, "u");
    po.vp = origin_of(pi, &p2, "pi");
    launch_post(po, stream());
}

extern "C" void surface_PML(int ndepth, float **vx, float **vy, float **sxx, float **syy, float **sxy, float **syz,
                            float ***p, float ***q, float **ppi, float **pu, float **prho, float **ptaup,
                            float **ptaus, float *etajm, float *peta, float *hc, float *K_x, float *a_x, float *b_x,
                            float **psi_vxx, float **ux, float **uy, float **uxy, float **uyz, float **sxz,
                            float **uxz)
{
    (void)syz; (void)uyz; (void)sxz; (void)uxz;
    if (PARAMETERIZATION != 1) fatal("surface_PML: only PARAMETERIZATION=1 is supported on the GPU");
    PostArgs po;
    memset(&po, 0, sizeof po);
    surface_common(ndepth, po, vx, vy, sxx, syy, sxy, ux, uy, uxy, hc, K_x, a_x, b_x, psi_vxx);
    int p2;
    po.L = L;
    po.rho = origin_of(prho, &p2, "prho");
    po.vs = origin_of(pu, &p2, "pu");
    po.vp = origin_of(ppi, &p2, "ppi");
    po.taus = origin_of(ptaus, &p2, "ptaus");
    po.taup = origin_of(ptaup, &p2, "ptaup");
    float *o;
    long pl = 0;
    if (tensor_origin(p, &o, &pl, po.g.pitch, "p") < L) fatal("surface_PML: p needs L planes");
    po.p = o;
    if (tensor_origin(q, &o, &pl, po.g.pitch, "q") < L) fatal("surface_PML: q needs L planes");
    po.q = o;
    po.cs = 1;
    po.ls = (int)pl;
    // the relaxed-modulus sums of src/surface_PML.c:46-60
    float pts[kMaxL + 1];
    for (int l = 1; l <= L; l++) pts[l] = 1.0 / (2.0 * 3.141592653589793 * FL[l]);
    float ws = 2.0 * 3.141592653589793 * F_REF;
    float sumu = 0.0, sumpi = 0.0;
    for (int l = 1; l <= L; l++) {
        sumu = sumu + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l]));
        sumpi = sumpi + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l]));
    }
    po.sumu = sumu;
    po.sumpi = sumpi;
    for (int l = 0; l <= kMaxL; l++) po.eta[l] = (l >= 1 && l <= L) ? peta[l] : 0.0f;
    for (int l = 1; l <= L; l++) etajm[l] = peta[l];
    launch_post(po, stream());
}

// ---------------------------------------------------------------------------
// seismo_ssg  (src/seismo_ssg.c:26, src/fd.h:262-265), SEISMO 1 and 2
// ---------------------------------------------------------------------------
extern "C" void seismo_ssg(int lsamp, int ntr, int **recpos, float **sectionvx, float **sectionvy,
                           float **sectionvz, float **sectionp, float **sectioncurl, float **sectiondiv, float **vx,
                           float **vy, float **vz, float **sxx, float **syy, float **sp, float **pi, float **u,
                           float *hc)
{
    (void)sectionvz; (void)sectioncurl; (void)sectiondiv; (void)vz; (void)sp; (void)pi; (void)u; (void)hc;
    if (ntr <= 0) return;
    if (SEISMO != 1 && SEISMO != 2) fatal("seismo_ssg: SEISMO must be 1 or 2 on the GPU");
    int P, p2;
    const float *a, *b;
    if (SEISMO == 1) {
        a = origin_of(vx, &P, "vx");
        b = origin_of(vy, &p2, "vy");
    } else {
        a = origin_of(sxx, &P, "sxx");
        b = origin_of(syy, &p2, "syy");
    }
    if (lsamp < 1 || lsamp > NT) fatal("seismo_ssg: sample index outside 1..NT");
    std::vector<int> pos(2 * ntr);
    for (int t = 1; t <= ntr; t++) {
        pos[t - 1] = recpos[1][t];
        pos[ntr + t - 1] = recpos[2][t];
    }
    float **s0 = SEISMO == 1 ? sectionvx : sectionp, **s1 = SEISMO == 1 ? sectionvy : nullptr;
    SectionBuf &sb = g_sections[s0];
    if (sb.ntr != ntr || sb.ns != NT || sb.s1 != s1) {
        // a new section pair, or one reused with another shape: flush what it
        // holds and size its buffer for this shape
        hip_check(hipStreamSynchronize(stream()), "section sync");
        hip_check(flush_sections(), "section flush");
        if (sb.dev) hip_check(hipFree(sb.dev), "hipFree");
        if (sb.dpos) hip_check(hipFree(sb.dpos), "hipFree");
        sb = SectionBuf();
        sb.s0 = s0;
        sb.s1 = s1;
        sb.ntr = ntr;
        sb.ns = NT;
        hip_check(hipMalloc(&sb.dev, sizeof(float) * 2 * (size_t)ntr * NT), "hipMalloc(section)");
        hip_check(hipMalloc(&sb.dpos, sizeof(int) * 2 * ntr), "hipMalloc(section positions)");
    }
    if (sb.pos != pos) {
        int slot;
        char *blob = g_ring.get(sizeof(int) * pos.size(), slot);
        memcpy(blob, pos.data(), sizeof(int) * pos.size());
        hip_check(hipMemcpyAsync(sb.dpos, blob, sizeof(int) * pos.size(), hipMemcpyHostToDevice, stream()),
                  "recpos");
        g_ring.queued(slot);
        sb.pos = pos;
    }
    hipLaunchKernelGGL(k_gather_col, dim3((ntr + 63) / 64), dim3(64), 0, stream(), ntr, sb.dpos, sb.dpos + ntr, P,
                       a, b, SEISMO == 1 ? 1 : 2, sb.dev, NT, lsamp - 1);
    sb.lo = sb.hi < sb.lo ? lsamp : std::min(sb.lo, lsamp);
    sb.hi = std::max(sb.hi, lsamp);
}

// ---------------------------------------------------------------------------
// exchange_v / exchange_s (src/exchange_v.c:28, src/exchange_s.c:28,
// src/fd.h:116-124) on the one rank that holds the whole grid: with
// BOUNDARY 1 the rank is its own left and right neighbour and the halo
// columns take the opposite edge (vy and sxx: N right, N-1 left; vx and sxy:
// N-1 right, N left); with BOUNDARY 0 nothing changes, as in the reference,
// where a single rank sends to itself and unpacks nothing.  The buffers and
// requests are the driver's MPI ones and are not used; WAVETYPE 1 only.
// ---------------------------------------------------------------------------
extern "C" void exchange_v(float **vx, float **vy, float **vz, float **bufferlef_to_rig, float **bufferrig_to_lef,
                           float **buffertop_to_bot, float **bufferbot_to_top, void *req_send, void *req_rec,
                           int wavetyp_start)
{
    (void)vz; (void)bufferlef_to_rig; (void)bufferrig_to_lef; (void)buffertop_to_bot; (void)bufferbot_to_top;
    (void)req_send; (void)req_rec;
    if (wavetyp_start != 1) fatal("exchange_v: WAVETYPE 1 (P-SV) only");
    if (BOUNDARY) exchange_x(vy, vx, "vy", "vx");
}

extern "C" void exchange_s(float **sxx, float **syy, float **sxy, float **sxz, float **syz, float **bufferlef_to_rig,
                           float **bufferrig_to_lef, float **buffertop_to_bot, float **bufferbot_to_top,
                           void *req_send, void *req_rec, int wavetyp_start)
{
    (void)syy; (void)sxz; (void)syz; (void)bufferlef_to_rig; (void)bufferrig_to_lef; (void)buffertop_to_bot;
    (void)bufferbot_to_top; (void)req_send; (void)req_rec;
    if (wavetyp_start != 1) fatal("exchange_s: WAVETYPE 1 (P-SV) only");
    if (BOUNDARY) exchange_x(sxx, sxy, "sxx", "sxy");
}

// ---------------------------------------------------------------------------
// zero_fdveps / zero_fdveps_visc (src/zero_fdveps.c:28, src/fd.h:428-433):
// every device array passed is cleared
// ---------------------------------------------------------------------------
extern "C" void zero_fdveps(int ny1, int ny2, int nx1, int nx2, float **vx, float **vy, float **vz, float **sxx,
                            float **syy, float **sxy, float **sxz, float **syz, float **vxm1, float **vym1,
                            float **uxy, float **vxp1, float **vyp1, float **psi_sxx_x, float **psi_sxy_x,
                            float **psi_sxz_x, float **psi_vxx, float **psi_vyx, float **psi_vzx, float **psi_syy_y,
                            float **psi_sxy_y, float **psi_syz_y, float **psi_vyy, float **psi_vxy, float **psi_vzy,
                            float **psi_vxxs)
{
    (void)ny1; (void)ny2; (void)nx1; (void)nx2;
    void *all[] = {vx, vy, sxx, syy, sxy, vxm1, vym1, uxy, vxp1, vyp1, psi_sxx_x, psi_sxy_x, psi_vxx,
                   psi_vyx, psi_syy_y, psi_sxy_y, psi_vyy, psi_vxy, psi_vxxs};
    for (void *p : all) zero_dev(p);
    void *opt[] = {vz, sxz, syz, psi_sxz_x, psi_vzx, psi_syz_y, psi_vzy};
    for (void *p : opt)
        if (p && lookup(p)) zero_dev(p);
}

extern "C" void zero_fdveps_visc(int ny1, int ny2, int nx1, int nx2, float **vx, float **vy, float **vz,
                                 float **sxx, float **syy, float **sxy, float **sxz, float **syz, float **vxm1,
                                 float **vym1, float **uxy, float **vxp1, float **vyp1, float **psi_sxx_x,
                                 float **psi_sxy_x, float **psi_sxz_x, float **psi_vxx, float **psi_vyx,
                                 float **psi_vzx, float **psi_syy_y, float **psi_sxy_y, float **psi_syz_y,
                                 float **psi_vyy, float **psi_vxy, float **psi_vzy, float **psi_vxxs, float ***pr,
                                 float ***pp, float ***pq, float ***pt, float ***po)
{
    zero_fdveps(ny1, ny2, nx1, nx2, vx, vy, vz, sxx, syy, sxy, sxz, syz, vxm1, vym1, uxy, vxp1, vyp1, psi_sxx_x,
                psi_sxy_x, psi_sxz_x, psi_vxx, psi_vyx, psi_vzx, psi_syy_y, psi_sxy_y, psi_syz_y, psi_vyy, psi_vxy,
                psi_vzy, psi_vxxs);
    zero_dev(pr);
    zero_dev(pp);
    zero_dev(pq);
    if (pt && lookup(pt)) zero_dev(pt);
    if (po && lookup(po)) zero_dev(po);
}

// ---------------------------------------------------------------------------
// Snapshot store, zero-lag correlation and chain rule.  The reference runs
// them inline in main (src/IFOS2D.c:1805-1871, 2381-2419, 2471-2611) on the
// arrays main holds; these entry points run the same loops on device arrays
// (WAVETYPE 1, not ACOUSTIC, IDXI = IDYI = IDX = IDY = 1, PARAMETERIZATION 1).
// Snapshot tensors keep the reference's f3tensor layout, time fastest.
// ---------------------------------------------------------------------------
namespace {

struct Mat {
    float *o;  // origin: element [j][i] at o + j * pitch + i
};

// the kernels below touch rows 1..NY and columns 1..NX of every array
void check_covers(const void *arr, const char *name)
{
    const DevArr *d = lookup(arr);
    if (!d || d->nrl > 1 || d->nrh < NY || d->ncl > 1 || d->nch < NX) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s must be a device array covering rows 1..NY and columns 1..NX", name);
        fatal(buf);
    }
}

// origin of a device matrix whose row length must equal P (P <= 0: set it)
Mat mat_of(float **m, int &P, const char *name)
{
    int p;
    float *o = origin_of(m, &p, name);
    check_covers(m, name);
    if (P <= 0) P = p;
    if (p != P) {
        char buf[256];
        snprintf(buf, sizeof buf, "%s must have the wavefield grid's column bounds", name);
        fatal(buf);
    }
    return Mat{o};
}

struct SnapStoreArgs {
    int NX, NY, P, k;  // k: 0-based time slot
    long plane;
    const float *src[5];
    float *dst[5];     // tensor origins: element (j, i, l) at dst + j * P + i + (l - 1) * plane
};

__global__ void k_dropin_snapshot(SnapStoreArgs a)
{
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y;
    if (i > a.NX) return;
    const long c = (long)j * a.P + i, t = c + (long)a.k * a.plane;
#pragma unroll
    for (int f = 0; f < 5; f++) a.dst[f][t] = a.src[f][c];
}

struct CorrStoreArgs {
    int NX, NY, P, k;
    long plane;
    const float *vxp1, *vyp1, *sxx, *syy, *sxy, *rho, *vp, *vs;
    const float *F[5];     // forward_prop_rho_x, _rho_y, _x, _y, _u
    float *wc, *wc_u, *wc_rho;
};

__global__ void k_dropin_correlate(CorrStoreArgs a)
{
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y;
    if (i > a.NX) return;
    const long c = (long)j * a.P + i, t = c + (long)a.k * a.plane;
    const float rh = a.rho[c], u = a.vs[c], pv = a.vp[c];
    // muss, lamss as main forms them (src/IFOS2D.c:2395-2396)
    const float muss = rh * u * u;
    const float lamss = (float)((double)(rh * pv * pv) - 2.0 * (double)muss);
    float wr = a.wc_rho[c], wl = a.wc[c], wu = a.wc_u[c];
    correlate_cell(a.vxp1[c], a.vyp1[c], a.sxx[c], a.syy[c], a.sxy[c], a.F[0][t], a.F[1][t], a.F[2][t], a.F[3][t],
                   a.F[4][t], muss, lamss, u > 0.0f, wr, wl, wu);
    a.wc_rho[c] = wr;
    a.wc[c] = wl;
    a.wc_u[c] = wu;
}

struct AssembleStoreArgs {
    int NX, NY, P;
    float DT;
    const float *rho, *vp, *vs;
    float *wc, *wc_u, *wc_rho;
    float *lam, *mu, *rho_s;  // may be null
};

__global__ void k_dropin_assemble(AssembleStoreArgs a)
{
    const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y;
    if (i > a.NX) return;
    const long c = (long)j * a.P + i;
    float svp, svs, srho, wl, wmu, wrs;
    chain_rule_cell(a.DT, a.rho[c], a.vs[c], a.vp[c], a.wc[c], a.wc_u[c], a.wc_rho[c], svp, svs, srho, wl, wmu, wrs);
    a.wc[c] = svp;
    a.wc_u[c] = svs;
    a.wc_rho[c] = srho;
    if (a.lam) a.lam[c] = wl;
    if (a.mu) a.mu[c] = wmu;
    if (a.rho_s) a.rho_s[c] = wrs;
}

void check_loop_domain(const char *who)
{
    if (NPROCX != 1 || NPROCY != 1) fatal("one GPU holds the whole domain: NPROCX = NPROCY = 1 required");
    if (NX < 1 || NY < 1) {
        char buf[128];
        snprintf(buf, sizeof buf, "%s: NX and NY must be set", who);
        fatal(buf);
    }
}

}  // namespace

extern "C" void ifos_snapshot_store(int hin, float **pvxp1, float **pvyp1, float **psxx, float **psyy, float **psxy,
                                    float **ux, float **uy, float **uxy, float ***forward_prop_rho_x,
                                    float ***forward_prop_rho_y, float ***forward_prop_x, float ***forward_prop_y,
                                    float ***forward_prop_u)
{
    check_loop_domain("ifos_snapshot_store");
    int P = 0;
    SnapStoreArgs a;
    a.src[0] = mat_of(pvxp1, P, "pvxp1").o;
    a.src[1] = mat_of(pvyp1, P, "pvyp1").o;
    if (VELOCITY == 0) {
        a.src[2] = mat_of(psxx, P, "psxx").o;
        a.src[3] = mat_of(psyy, P, "psyy").o;
        a.src[4] = mat_of(psxy, P, "psxy").o;
    } else {
        a.src[2] = mat_of(ux, P, "ux").o;
        a.src[3] = mat_of(uy, P, "uy").o;
        a.src[4] = mat_of(uxy, P, "uxy").o;
    }
    float ***t[5] = {forward_prop_rho_x, forward_prop_rho_y, forward_prop_x, forward_prop_y, forward_prop_u};
    const char *tn[5] = {"forward_prop_rho_x", "forward_prop_rho_y", "forward_prop_x", "forward_prop_y",
                         "forward_prop_u"};
    for (int f = 0; f < 5; f++) {
        const int nd = tensor_origin(t[f], &a.dst[f], &a.plane, P, tn[f]);
        check_covers(t[f], tn[f]);
        if (hin < 1 || hin > nd) fatal("ifos_snapshot_store: slot outside the snapshot tensors");
    }
    a.NX = NX; a.NY = NY; a.P = P; a.k = hin - 1;
    hipLaunchKernelGGL(k_dropin_snapshot, dim3((NX + 255) / 256, NY), dim3(256), 0, stream(), a);
}

extern "C" void ifos_correlate(int k, float **pvxp1, float **pvyp1, float **psxx, float **psyy, float **psxy,
                               float **prho, float **ppi, float **pu, float ***forward_prop_rho_x,
                               float ***forward_prop_rho_y, float ***forward_prop_x, float ***forward_prop_y,
                               float ***forward_prop_u, float **waveconv_shot, float **waveconv_u_shot,
                               float **waveconv_rho_shot)
{
    check_loop_domain("ifos_correlate");
    int P = 0;
    CorrStoreArgs a;
    a.vxp1 = mat_of(pvxp1, P, "pvxp1").o;
    a.vyp1 = mat_of(pvyp1, P, "pvyp1").o;
    a.sxx = mat_of(psxx, P, "psxx").o;
    a.syy = mat_of(psyy, P, "psyy").o;
    a.sxy = mat_of(psxy, P, "psxy").o;
    a.rho = mat_of(prho, P, "prho").o;
    a.vp = mat_of(ppi, P, "ppi").o;
    a.vs = mat_of(pu, P, "pu").o;
    a.wc = mat_of(waveconv_shot, P, "waveconv_shot").o;
    a.wc_u = mat_of(waveconv_u_shot, P, "waveconv_u_shot").o;
    a.wc_rho = mat_of(waveconv_rho_shot, P, "waveconv_rho_shot").o;
    float ***t[5] = {forward_prop_rho_x, forward_prop_rho_y, forward_prop_x, forward_prop_y, forward_prop_u};
    const char *tn[5] = {"forward_prop_rho_x", "forward_prop_rho_y", "forward_prop_x", "forward_prop_y",
                         "forward_prop_u"};
    for (int f = 0; f < 5; f++) {
        float *o;
        const int nd = tensor_origin(t[f], &o, &a.plane, P, tn[f]);
        check_covers(t[f], tn[f]);
        if (k < 1 || k > nd) fatal("ifos_correlate: slot outside the snapshot tensors");
        a.F[f] = o;
    }
    a.NX = NX; a.NY = NY; a.P = P; a.k = k - 1;
    hipLaunchKernelGGL(k_dropin_correlate, dim3((NX + 255) / 256, NY), dim3(256), 0, stream(), a);
}

extern "C" void ifos_gradient_assemble(float **prho, float **ppi, float **pu, float **waveconv_shot,
                                       float **waveconv_u_shot, float **waveconv_rho_shot, float **waveconv_lam,
                                       float **waveconv_mu, float **waveconv_rho_s)
{
    check_loop_domain("ifos_gradient_assemble");
    if (PARAMETERIZATION != 1) fatal("ifos_gradient_assemble: PARAMETERIZATION 1 only");
    int P = 0;
    AssembleStoreArgs a;
    a.rho = mat_of(prho, P, "prho").o;
    a.vp = mat_of(ppi, P, "ppi").o;
    a.vs = mat_of(pu, P, "pu").o;
    a.wc = mat_of(waveconv_shot, P, "waveconv_shot").o;
    a.wc_u = mat_of(waveconv_u_shot, P, "waveconv_u_shot").o;
    a.wc_rho = mat_of(waveconv_rho_shot, P, "waveconv_rho_shot").o;
    a.lam = waveconv_lam ? mat_of(waveconv_lam, P, "waveconv_lam").o : nullptr;
    a.mu = waveconv_mu ? mat_of(waveconv_mu, P, "waveconv_mu").o : nullptr;
    a.rho_s = waveconv_rho_s ? mat_of(waveconv_rho_s, P, "waveconv_rho_s").o : nullptr;
    a.NX = NX; a.NY = NY; a.P = P; a.DT = DT;
    hipLaunchKernelGGL(k_dropin_assemble, dim3((NX + 255) / 256, NY), dim3(256), 0, stream(), a);
}
