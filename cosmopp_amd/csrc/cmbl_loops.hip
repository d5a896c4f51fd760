// Device-resident optimiser and sampler loops over a map's likelihood.
//
//   cmbl_lbfgs  LBFGS_General with the More-Thuente line search
//               (include/lbfgs_general.hpp:144-405, include/line_search.hpp:
//               12-272): every vector in device memory, each phase's vector
//               work in fused passes, captured in HIP graphs (one per
//               direction computation, one per line-search trial); the host
//               keeps only the scalar control flow of the line search.
//   cmbl_hmc    HMCGeneral (include/hmc_general.hpp:97-246): one HIP graph per
//               trajectory (kick / drift passes around the evaluations), the
//               reference's momentum draws (BasicLBFGSFunc::whitenoise,
//               source/lbfgs.cpp:401-416) made on host threads trajectories
//               ahead and uploaded asynchronously.
//
// Every dot product of these loops runs the plain dot product's summation
// order (kDotParts blocks of 256 threads with grid-stride FMA accumulation,
// block sums, then the final block's wave sums added in order), and every
// elementwise update the BLAS-1 entry point's operation (fma(c, x, y), c * x,
// x / y), so the loops produce the same iterates, bit for bit, as the
// reference's LBFGS_General / HMCGeneral running on cmbl::DeviceLargeVector.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <future>
#include <map>
#include <random>
#include <string>
#include <vector>

// HIP contracts a * b + c into fma by default, across statements; the scalar
// steps here (alpha - rho * dot, the line search) must round each operation
// as the reference's host code does, and the element updates already spell
// out their fma, so contraction is off for this file.
#pragma clang fp contract(off)

#include "cmbl_device.hpp"
#include "cmbl_internal.hpp"
#include "cmbl_linesearch.hpp"

namespace cmbl {

constexpr int kPassThreads = 256;
constexpr int kMaxPassDots = 3;

// Block sums of a fused pass, kDotParts per dot product.
struct PassSync {
    double* part;  // [kMaxPassDots * kDotParts]
};

// The plain dot product's final stage (1024 threads, thread t holding
// 0 + part[t], wave sums, thread 0 adding the 16 wave sums in order) replayed
// by the 256 threads of one block.  Result valid in thread 0.
__device__ double final_replay(const double* part, double* sh16)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int v = w; v < kDotParts / 64; v += kPassThreads / 64) {
        double a = 0.0;
        a += part[64 * v + lane];
        a = wave_sum(a);
        if (lane == 0) sh16[v] = a;
    }
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kDotParts / 64; ++i) t += sh16[i];
    __syncthreads();
    return t;
}

// One pass over n elements: Op::fetch loads element i's inputs, Op::apply
// updates element i and accumulates its dot-product terms
// (a[d] = fma(x, y, a[d])), and each block stores its kDots block sums.
// Launched with exactly kDotParts blocks.  Each thread fetches four of its
// elements before applying them in their grid-stride order, so more loads
// are in flight while the accumulation order stays that of the plain dot
// product.
// The element loop of a pass from element i on (the thread's elements are
// i, i + stride, ...), accumulating into a; then the block sums.
template <class Op>
__device__ __forceinline__ void pass_loop(Op& op, long i, double* a)
{
    const long stride = (long)gridDim.x * kPassThreads;
    for (; i + 3 * stride < op.n; i += 4 * stride) {
        typename Op::In v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = op.fetch(i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u) op.apply(i + u * stride, v[u], a);
    }
    for (; i < op.n; i += stride) op.apply(i, op.fetch(i), a);
}
template <class Op>
__device__ __forceinline__ void pass_store(const double* a, const PassSync& ps, double* sh)
{
    constexpr int ND = Op::kDots;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const double t = block_sum<kPassThreads>(a[d], sh);
        if (threadIdx.x == 0) ps.part[d * kDotParts + blockIdx.x] = t;
    }
}

template <class Op>
__global__ __launch_bounds__(kPassThreads) void k_pass(Op op, PassSync ps)
{
    constexpr int ND = Op::kDots;
    __shared__ double sh[kPassThreads / 64];
    op.load();
    double a[ND > 0 ? ND : 1];
#pragma unroll
    for (int d = 0; d < (ND > 0 ? ND : 1); ++d) a[d] = 0.0;
    pass_loop(op, blockIdx.x * (long)kPassThreads + threadIdx.x, a);
    if (ND > 0) pass_store<Op>(a, ps, sh);
}

// A pass chained to its predecessor (a one-product pass of the two-loop
// recursion): every block replays the predecessor's dot product from its
// block sums `prev`, exactly as the predecessor's final stage would, while
// its first elements' loads are in flight; Op::scalar turns the product into
// the scalar this pass needs, the first block also storing what that final
// stage stored.  The final stage's launch, and its wait between two passes,
// are saved; `prev` and ps.part are different buffers.
template <class Op>
__global__ __launch_bounds__(kPassThreads) void k_pass_chained(Op op, PassSync ps, const double* __restrict__ prev)
{
    constexpr int ND = Op::kDots;
    __shared__ double sh[kPassThreads / 64];
    __shared__ double sh16[kDotParts / 64];
    __shared__ double rprev;
    op.load_pointers();
    const long stride = (long)gridDim.x * kPassThreads;
    long i = blockIdx.x * (long)kPassThreads + threadIdx.x;
    const bool full = i + 3 * stride < op.n;
    typename Op::In v[4];
    if (full) {
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = op.fetch(i + u * stride);
    }
    const double r = final_replay(prev, sh16);
    if (threadIdx.x == 0) rprev = r;
    __syncthreads();
    op.scalar(rprev, blockIdx.x == 0 && threadIdx.x == 0);
    double a[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) a[d] = 0.0;
    if (full) {
#pragma unroll
        for (int u = 0; u < 4; ++u) op.apply(i + u * stride, v[u], a);
        i += 4 * stride;
    }
    pass_loop(op, i, a);
    pass_store<Op>(a, ps, sh);
}

// The pass's final stage, one block: the kDots results from the block sums,
// then Op::finish.  (Completing the sums inside the pass in its last block
// would need a device-scope release by every block, i.e. an L2 write-back per
// block on a chip whose XCDs have separate L2s; the kernel boundary gives
// the same ordering once.)
template <class Op>
__global__ __launch_bounds__(kPassThreads) void k_pass_final(Op op, PassSync ps)
{
    constexpr int ND = Op::kDots;
    __shared__ double sh16[kDotParts / 64];
    op.load();
    double r[ND > 0 ? ND : 1];
    for (int d = 0; d < ND; ++d) r[d] = final_replay(ps.part + d * kDotParts, sh16);
    if (threadIdx.x == 0) op.finish(r);
}

template <class Op>
static int run_pass(const Op& op, const PassSync& ps, hipStream_t s)
{
    hipLaunchKernelGGL(k_pass<Op>, dim3(kDotParts), dim3(kPassThreads), 0, s, op, ps);
    if (Op::kDots > 0) hipLaunchKernelGGL(k_pass_final<Op>, dim3(1), dim3(kPassThreads), 0, s, op, ps);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

// The pieces of a chain of passes: the first pass without its final stage,
// the chained passes, and the last pass's final stage.
template <class Op>
static int run_pass_open(const Op& op, const PassSync& ps, const double* prev, hipStream_t s)
{
    if (prev)
        hipLaunchKernelGGL(k_pass_chained<Op>, dim3(kDotParts), dim3(kPassThreads), 0, s, op, ps, prev);
    else
        hipLaunchKernelGGL(k_pass<Op>, dim3(kDotParts), dim3(kPassThreads), 0, s, op, ps);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}
template <class Op>
static int run_pass_close(const Op& op, const PassSync& ps, hipStream_t s)
{
    hipLaunchKernelGGL(k_pass_final<Op>, dim3(1), dim3(kPassThreads), 0, s, op, ps);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

// ====================================================================== L-BFGS
// Scalars the host reads after a pass (host-mapped memory).
struct LbHost {
    double stp;  // trial step, written by the host before a trial
    double f;    // -2lnL at the last trial point
    double gz;   // g . z at the last trial point (the line-search slope is -gz)
    double gg;   // g . g at the last trial point (gradient norm squared)
    double zz;   // z . z of the new direction
    double zg;   // z . g of the new direction
    double ys;   // s . y of the newest history pair
    double yy;   // |y|^2 of the newest pair, formed as norm() * norm()
};

// Device state of one L-BFGS context.  History index i (0 = newest) lives in
// slot slot[i] of S and Y; rho, alpha and coef are kept per slot.
struct LbDev {
    long n;
    int m;
    double *x, *x0, *g, *g0, *q, *z;
    double *S, *Y;
    int* slot;
    double *rho, *alpha, *coef, *h0k;
    const double* f;  // the evaluation's -2lnL
    LbHost* h;        // device view of the host-mapped scalars
};

// Start: x0 = x, g0 = g (xPrev, gPrev of setStarting), g . g, H0k = 1.
struct OpStart {
    static constexpr int kDots = 1;
    LbDev d;
    long n;
    struct In {
        double x, g;
    };
    __device__ void load() {}
    __device__ In fetch(long i) const { return In{d.x[i], d.g[i]}; }
    __device__ void apply(long i, const In& v, double* a)
    {
        d.x0[i] = 1.0 * v.x;
        d.g0[i] = 1.0 * v.g;
        a[0] = fma(v.g, v.g, a[0]);
    }
    __device__ void finish(const double* r)
    {
        d.h->gg = r[0];
        d.h->f = *d.f;
        *d.h0k = 1.0;
    }
};

// History update after a line search: s = x - xPrev, y = g - gPrev into the
// oldest slot (which becomes index 0), xPrev = x, gPrev = g; s . y, y . y and
// s . g (the first product of the next two-loop recursion, q = g).
struct OpHistory {
    static constexpr int kDots = 3;
    LbDev d;
    long n;
    double *sn, *yn;
    __device__ void load()
    {
        const int k = d.slot[d.m - 1];
        sn = d.S + (long)k * d.n;
        yn = d.Y + (long)k * d.n;
    }
    struct In {
        double x, x0, g, g0;
    };
    __device__ In fetch(long i) const { return In{d.x[i], d.x0[i], d.g[i], d.g0[i]}; }
    __device__ void apply(long i, const In& v, double* a)
    {
        const double s = v.x - v.x0;
        const double y = v.g - v.g0;
        sn[i] = s;
        yn[i] = y;
        d.x0[i] = 1.0 * v.x;
        d.g0[i] = 1.0 * v.g;
        a[0] = fma(s, y, a[0]);
        a[1] = fma(y, y, a[1]);
        a[2] = fma(s, v.g, a[2]);
    }
    __device__ void finish(const double* r)
    {
        const int k = d.slot[d.m - 1];
        for (int i = d.m - 1; i > 0; --i) d.slot[i] = d.slot[i - 1];
        d.slot[0] = k;
        const double ys = r[0];
        double yy = sqrt(r[1]);
        yy = yy * yy;
        d.h->ys = ys;
        d.h->yy = yy;
        d.rho[k] = 1 / ys;
        *d.h0k = ys / yy;
        d.alpha[k] = d.rho[k] * r[2];
    }
};

// First loop of the two-loop recursion, step i of mu: q = q_in - alpha_i y_i
// (q_in = g for i = 0); then either alpha_{i+1} = rho_{i+1} s_{i+1} . q, or,
// at i = mu - 1, z = H0k q and coef_i = alpha_i - rho_i y_i . z.
struct OpLoop1 {
    static constexpr int kDots = 1;
    LbDev d;
    long n;
    int i, mu;
    const double *qin, *yi, *sn;
    double c, h;
    int ki, kn;
    __device__ void load()
    {
        ki = d.slot[i];
        kn = i + 1 < mu ? d.slot[i + 1] : -1;
        qin = i == 0 ? d.g : d.q;
        yi = d.Y + (long)ki * d.n;
        sn = kn >= 0 ? d.S + (long)kn * d.n : nullptr;
        c = -d.alpha[ki];
        h = *d.h0k;
    }
    // Chained to step i - 1 (i >= 1), whose final stage stores
    // alpha_i = rho_i s_i . q (its product r).
    __device__ void load_pointers()
    {
        ki = d.slot[i];
        kn = i + 1 < mu ? d.slot[i + 1] : -1;
        qin = d.q;
        yi = d.Y + (long)ki * d.n;
        sn = kn >= 0 ? d.S + (long)kn * d.n : nullptr;
        h = *d.h0k;
    }
    __device__ void scalar(double r, bool writer)
    {
        const double a = d.rho[ki] * r;
        if (writer) d.alpha[ki] = a;
        c = -a;
    }
    struct In {
        double y, q, s;
    };
    __device__ In fetch(long e) const { return In{yi[e], qin[e], kn >= 0 ? sn[e] : 0.0}; }
    __device__ void apply(long e, const In& v, double* a)
    {
        const double qv = fma(c, v.y, v.q);
        if (kn >= 0) {
            d.q[e] = qv;
            a[0] = fma(v.s, qv, a[0]);
        } else {
            const double zv = h * qv;
            d.z[e] = zv;
            a[0] = fma(v.y, zv, a[0]);
        }
    }
    __device__ void finish(const double* r)
    {
        if (kn >= 0) {
            d.alpha[kn] = d.rho[kn] * r[0];
        } else {
            const double beta = d.rho[ki] * r[0];
            d.coef[ki] = d.alpha[ki] - beta;
        }
    }
};

// Second loop, step i (mu - 1 down to 0): z += coef_i s_i; then either
// coef_{i-1} = alpha_{i-1} - rho_{i-1} y_{i-1} . z, or, at i = 0, z . z and
// z . g for the host.
struct OpLoop2 {
    static constexpr int kDots = 2;
    LbDev d;
    long n;
    int i;
    const double *si, *yp;
    double c;
    int ki, kp;
    __device__ void load()
    {
        ki = d.slot[i];
        kp = i > 0 ? d.slot[i - 1] : -1;
        si = d.S + (long)ki * d.n;
        yp = kp >= 0 ? d.Y + (long)kp * d.n : nullptr;
        c = d.coef[ki];
    }
    // Chained to the first loop's last step or to step i + 1 of this loop,
    // whose final stage stores coef_i = alpha_i - rho_i r (r: its product
    // y_i . z).
    __device__ void load_pointers()
    {
        ki = d.slot[i];
        kp = i > 0 ? d.slot[i - 1] : -1;
        si = d.S + (long)ki * d.n;
        yp = kp >= 0 ? d.Y + (long)kp * d.n : nullptr;
    }
    __device__ void scalar(double r, bool writer)
    {
        const double beta = d.rho[ki] * r;
        const double cc = d.alpha[ki] - beta;
        if (writer) d.coef[ki] = cc;
        c = cc;
    }
    struct In {
        double s, z, o;  // o: y_{i-1}, or g at i = 0
    };
    __device__ In fetch(long e) const { return In{si[e], d.z[e], kp >= 0 ? yp[e] : d.g[e]}; }
    __device__ void apply(long e, const In& v, double* a)
    {
        const double zv = fma(c, v.s, v.z);
        d.z[e] = zv;
        if (kp >= 0) {
            a[0] = fma(v.o, zv, a[0]);
        } else {
            a[0] = fma(zv, zv, a[0]);
            a[1] = fma(zv, v.o, a[1]);
        }
    }
    __device__ void finish(const double* r)
    {
        if (kp >= 0) {
            const double beta = d.rho[kp] * r[0];
            d.coef[kp] = d.alpha[kp] - beta;
        } else {
            d.h->zz = r[0];
            d.h->zg = r[1];
        }
    }
};

// No history yet: z = H0k g (q = g), z . z, z . g.
struct OpDir0 {
    static constexpr int kDots = 2;
    LbDev d;
    long n;
    double h;
    __device__ void load() { h = *d.h0k; }
    struct In {
        double g;
    };
    __device__ In fetch(long e) const { return In{d.g[e]}; }
    __device__ void apply(long e, const In& v, double* a)
    {
        const double zv = h * v.g;
        d.z[e] = zv;
        a[0] = fma(zv, zv, a[0]);
        a[1] = fma(zv, v.g, a[1]);
    }
    __device__ void finish(const double* r)
    {
        d.h->zz = r[0];
        d.h->zg = r[1];
    }
};

// Steepest-descent fallback: z = g.
struct OpZeqG {
    static constexpr int kDots = 0;
    LbDev d;
    long n;
    __device__ void load() {}
    struct In {
        double g;
    };
    __device__ In fetch(long e) const { return In{d.g[e]}; }
    __device__ void apply(long e, const In& v, double*) { d.z[e] = 1.0 * v.g; }
    __device__ void finish(const double*) {}
};

// Trial point of the line search along s = -z: x = x0 + stp s.
struct OpTrial {
    static constexpr int kDots = 0;
    LbDev d;
    long n;
    double stp;
    __device__ void load() { stp = d.h->stp; }
    struct In {
        double z, x0;
    };
    __device__ In fetch(long e) const { return In{d.z[e], d.x0[e]}; }
    __device__ void apply(long e, const In& v, double*) { d.x[e] = fma(stp, -v.z, v.x0); }
    __device__ void finish(const double*) {}
};

// After the trial's evaluation: g . z (slope -g . z), g . g, and -2lnL.
struct OpTrialDots {
    static constexpr int kDots = 2;
    LbDev d;
    long n;
    __device__ void load() {}
    struct In {
        double g, z;
    };
    __device__ In fetch(long e) const { return In{d.g[e], d.z[e]}; }
    __device__ void apply(long, const In& v, double* a)
    {
        a[0] = fma(v.g, v.z, a[0]);
        a[1] = fma(v.g, v.g, a[1]);
    }
    __device__ void finish(const double* r)
    {
        d.h->gz = r[0];
        d.h->gg = r[1];
        d.h->f = *d.f;
    }
};

// A HIP graph per launch sequence, captured on first use.
struct GraphCache {
    std::map<long, hipGraphExec_t> exec;
    ~GraphCache() { clear(); }
    void clear()
    {
        for (auto& kv : exec) (void)hipGraphExecDestroy(kv.second);
        exec.clear();
    }
    // Run `launch(s)` on stream s: captured into a graph under `key` the first
    // time, replayed afterwards.  With graphs off, launches directly.
    template <class F>
    int run(bool graphs, long key, hipStream_t s, F launch)
    {
        if (!graphs) return launch(s);
        auto it = exec.find(key);
        if (it == exec.end()) {
            CMBL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = launch(s);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &graph);
            if (rc != CMBL_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            hipGraphExec_t ge = nullptr;
            const hipError_t ei = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate");
            it = exec.emplace(key, ge).first;
        }
        CMBL_HIP(hipGraphLaunch(it->second, s));
        return CMBL_OK;
    }
};

// Device buffers of a loop context, freed together.
struct LoopAllocs {
    std::vector<void*> dev;
    std::vector<void*> host;
    ~LoopAllocs()
    {
        for (void* p : dev) (void)hipFree(p);
        for (void* p : host) (void)hipHostFree(p);
    }
    template <class T>
    int dalloc(T** p, size_t count)
    {
        void* q = nullptr;
        CMBL_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
        dev.push_back(q);
        *p = static_cast<T*>(q);
        return CMBL_OK;
    }
    template <class T>
    int halloc(T** p, T** dp, size_t count)
    {
        void* q = nullptr;
        CMBL_HIP(hipHostMalloc(&q, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocMapped));
        host.push_back(q);
        std::memset(q, 0, std::max<size_t>(count, 1) * sizeof(T));
        *p = static_cast<T*>(q);
        void* d = nullptr;
        CMBL_HIP(hipHostGetDevicePointer(&d, q, 0));
        *dp = static_cast<T*>(d);
        return CMBL_OK;
    }
};

}  // namespace cmbl

using namespace cmbl;

struct cmbl_lbfgs {
    cmbl_map* map = nullptr;
    long n = 0;
    int m = 0;
    bool graphs = true;
    hipStream_t stream = nullptr;
    LoopAllocs mem;
    LbDev d{};
    LbHost* h = nullptr;  // host view
    PassSync ps{};
    double* d_f = nullptr;
    GraphCache cache;
    // reference state (include/lbfgs_general.hpp:128-155)
    double val = 0, gradNorm = 0, gg = 0;
    int iter = 0;
    bool pending = false;  // history update of the last line search not yet applied
    long evaluations = 0;
    ~cmbl_lbfgs()
    {
        cache.clear();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct cmbl_hmc;

namespace cmbl {

static int lb_eval(cmbl_lbfgs* lb, hipStream_t s)
{
    return eval_device(lb->map, lb->d.x, lb->d_f, lb->d.g, s, false);
}

// Vector work before the trial evaluations of an iteration: the pending
// history update, then the two-loop recursion with mu stored pairs.
static int lb_direction(cmbl_lbfgs* lb, bool pending, int mu, hipStream_t s)
{
    const LbDev& d = lb->d;
    if (pending) CMBL_TRY(run_pass(OpHistory{d, d.n, nullptr, nullptr}, lb->ps, s));
    if (mu == 0) return run_pass(OpDir0{d, d.n, 0.0}, lb->ps, s);
    // the 2 mu passes form one chain: each pass replays its predecessor's
    // product (block sums alternating between two buffers), and only the
    // last one has a final stage of its own
    const double* prev = nullptr;
    int buf = 0;
    for (int i = 0; i < mu; ++i) {
        OpLoop1 op{};
        op.d = d;
        op.n = d.n;
        op.i = i;
        op.mu = mu;
        const PassSync ps{lb->ps.part + (long)buf * kMaxPassDots * kDotParts};
        CMBL_TRY(run_pass_open(op, ps, prev, s));
        prev = ps.part;
        buf ^= 1;
    }
    for (int i = mu - 1; i >= 0; --i) {
        OpLoop2 op{};
        op.d = d;
        op.n = d.n;
        op.i = i;
        const PassSync ps{lb->ps.part + (long)buf * kMaxPassDots * kDotParts};
        CMBL_TRY(run_pass_open(op, ps, prev, s));
        prev = ps.part;
        buf ^= 1;
        if (i == 0) CMBL_TRY(run_pass_close(op, ps, s));
    }
    return CMBL_OK;
}

static int lb_trial(cmbl_lbfgs* lb, hipStream_t s)
{
    const LbDev& d = lb->d;
    CMBL_TRY(run_pass(OpTrial{d, d.n, 0.0}, lb->ps, s));
    CMBL_TRY(lb_eval(lb, s));
    return run_pass(OpTrialDots{d, d.n}, lb->ps, s);
}

static int lb_sync(cmbl_lbfgs* lb)
{
    CMBL_HIP(hipStreamSynchronize(lb->stream));
    return CMBL_OK;
}

// setStarting (include/lbfgs_general.hpp:167-200): evaluate at x, xPrev = x,
// gPrev = g, H0k = 1, no history.
static int lb_start(cmbl_lbfgs* lb)
{
    hipStream_t s = lb->stream;
    CMBL_TRY(lb_eval(lb, s));
    CMBL_TRY(run_pass(OpStart{lb->d, lb->d.n}, lb->ps, s));
    CMBL_TRY(lb_sync(lb));
    ++lb->evaluations;
    lb->val = lb->h->f;
    lb->gg = lb->h->gg;
    lb->gradNorm = std::sqrt(lb->gg);
    lb->iter = 0;
    lb->pending = false;
    if (!std::isfinite(lb->val)) return fail(CMBL_ERR_NONFINITE, "-2lnL at the starting point is not finite");
    return CMBL_OK;
}

}  // namespace cmbl

extern "C" {

int cmbl_lbfgs_create(cmbl_lbfgs** out, cmbl_map* map, const double* d_x0, int m, int flags)
{
    if (!out || !map || !d_x0) return fail(CMBL_ERR_ARG, "null pointer");
    *out = nullptr;
    if (m < 1) return fail(CMBL_ERR_ARG, "history length m must be >= 1");
    if (!map->has_data) return fail(CMBL_ERR_STATE, "cmbl_map_set_data has not been called");
    cmbl_lbfgs* lb = new cmbl_lbfgs();
    auto bail = [&](int rc) {
        const std::string msg = cmbl_last_error();
        delete lb;
        set_error(msg);
        return rc;
    };
    lb->map = map;
    lb->n = map->nparam;
    lb->m = m;
    lb->graphs = (flags & CMBL_LOOP_EAGER) == 0;
    int rc;
    if (hipStreamCreateWithFlags(&lb->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(CMBL_ERR_HIP, "hipStreamCreate failed"));
    LbDev& d = lb->d;
    d.n = lb->n;
    d.m = m;
    const size_t n = (size_t)lb->n;
    if ((rc = lb->mem.dalloc(&d.x, n)) || (rc = lb->mem.dalloc(&d.x0, n)) || (rc = lb->mem.dalloc(&d.g, n)) ||
        (rc = lb->mem.dalloc(&d.g0, n)) || (rc = lb->mem.dalloc(&d.q, n)) || (rc = lb->mem.dalloc(&d.z, n)) ||
        (rc = lb->mem.dalloc(&d.S, n * m)) || (rc = lb->mem.dalloc(&d.Y, n * m)) ||
        (rc = lb->mem.dalloc(&d.slot, m)) || (rc = lb->mem.dalloc(&d.rho, m)) ||
        (rc = lb->mem.dalloc(&d.alpha, m)) || (rc = lb->mem.dalloc(&d.coef, m)) ||
        (rc = lb->mem.dalloc(&d.h0k, 1)) || (rc = lb->mem.dalloc(&lb->d_f, 1)) ||
        (rc = lb->mem.dalloc(&lb->ps.part, (size_t)2 * kMaxPassDots * kDotParts)))
        return bail(rc);
    if ((rc = lb->mem.halloc(&lb->h, &d.h, 1))) return bail(rc);
    d.f = lb->d_f;
    std::vector<int> slots(m);
    for (int i = 0; i < m; ++i) slots[i] = i;
    std::vector<double> zeros(m, 0.0);
    // the caller's x0 may still be in flight on its streams; the copy goes on
    // this context's stream (a device-to-device hipMemcpy does not wait for
    // completion, and this stream does not order with the null stream)
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(d.slot, slots.data(), sizeof(int) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.rho, zeros.data(), sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.alpha, zeros.data(), sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.coef, zeros.data(), sizeof(double) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpyAsync(d.x, d_x0, sizeof(double) * n, hipMemcpyDeviceToDevice, lb->stream) != hipSuccess)
        return bail(fail(CMBL_ERR_HIP, "L-BFGS set-up copies failed"));
    if ((rc = lb_start(lb))) return bail(rc);
    *out = lb;
    return CMBL_OK;
}

void cmbl_lbfgs_destroy(cmbl_lbfgs* lb)
{
    if (!lb) return;
    (void)hipStreamSynchronize(lb->stream);
    delete lb;
}

int cmbl_lbfgs_set_starting(cmbl_lbfgs* lb, const double* d_x0)
{
    if (!lb || !d_x0) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipStreamSynchronize(lb->stream));
    CMBL_HIP(hipMemcpyAsync(lb->d.x, d_x0, sizeof(double) * lb->n, hipMemcpyDeviceToDevice, lb->stream));
    return lb_start(lb);
}

int cmbl_lbfgs_minimize(cmbl_lbfgs* lb, double epsilon, double gnorm_tol, int max_iter, double* d_res, double* f_out,
                        int* iterations, int* evaluations, cmbl_lbfgs_callback cb, void* user)
{
    if (!lb) return fail(CMBL_ERR_ARG, "null context");
    if (!(epsilon > 0)) return fail(CMBL_ERR_ARG, "epsilon must be > 0");
    if (!(gnorm_tol >= 0)) return fail(CMBL_ERR_ARG, "gnorm_tol must be >= 0");
    hipStream_t s = lb->stream;
    LbHost* h = lb->h;
    int thisIter = 0, functionEval = 0, convergedIters = 0;
    if (cb) {
        CMBL_TRY(lb_sync(lb));
        cb(user, thisIter, lb->val, lb->gradNorm, lb->d.x, lb->d.g, lb->d.z);
    }
    while (true) {
        // the two-loop recursion (include/lbfgs_general.hpp:222-238), after the
        // previous iteration's history update (:349-382) when that is still
        // pending (without a callback it shares the direction's launch)
        const int mu = std::min(lb->m, lb->iter);
        const bool pending = lb->pending;
        const long key = (pending ? 1L << 20 : 0L) | mu;
        CMBL_TRY(lb->cache.run(lb->graphs, key, s, [&](hipStream_t st) { return lb_direction(lb, pending, mu, st); }));
        CMBL_TRY(lb_sync(lb));
        if (pending) {
            lb->pending = false;
            if (h->ys == 0 || h->yy == 0) break;  // the reference: "LBFGS has not moved"
            if (thisIter > max_iter) break;
        }
        const double zNorm = std::sqrt(h->zz);
        const double zg = h->zg / zNorm;
        bool usingCG = false;
        // slope along s = -z: g . s = -(g . z), bit for bit
        double dginit = -h->zg;
        if (zg / lb->gradNorm < 0.01) {
            // steepest descent at this step (:241-252)
            CMBL_TRY(run_pass(OpZeqG{lb->d, lb->d.n}, lb->ps, s));
            usingCG = true;
            dginit = -lb->gg;
        }
        const double oldVal = lb->val;
        const double rate = (lb->iter == 0 || usingCG) ? 1.0 / lb->gradNorm : 1.0;
        if (!(dginit < 0)) return fail(CMBL_ERR_STATE, "the search direction is not a descent direction");
        // More-Thuente line search (:255-275), one graph launch per trial
        MoreThuente mt;
        double stp = mt.begin(lb->val, dginit, rate);
        while (true) {
            h->stp = stp;
            CMBL_TRY(lb->cache.run(lb->graphs, 1L << 21, s, [&](hipStream_t st) { return lb_trial(lb, st); }));
            CMBL_TRY(lb_sync(lb));
            ++lb->evaluations;
            if (!std::isfinite(h->f)) {
                (void)cmbl_map_check_finite(lb->map, s);
                return fail(CMBL_ERR_NONFINITE, "-2lnL at a line-search trial point is not finite");
            }
            if (mt.feed(h->f, -h->gz, &stp) != 0) break;
        }
        functionEval += mt.nfev;
        lb->val = h->f;
        lb->gg = h->gg;
        lb->gradNorm = std::sqrt(lb->gg);
        ++lb->iter;
        ++thisIter;
        // convergence (:320-347)
        const double deltaVal = std::fabs(lb->val - oldVal);
        const double valMax = std::max(std::fabs(lb->val), std::fabs(oldVal));
        const double ratio = deltaVal / std::max(valMax, 1.0);
        const int minIter = 10;
        if (ratio < epsilon && lb->iter >= minIter) {
            if (++convergedIters >= 3) break;
        } else {
            convergedIters = 0;
        }
        if (lb->gradNorm < gnorm_tol) break;
        if (!cb) {
            lb->pending = true;
            continue;
        }
        // with a callback the history update runs on its own, so that the
        // callback sees x, g and the direction z of the finished iteration
        CMBL_TRY(run_pass(OpHistory{lb->d, lb->d.n, nullptr, nullptr}, lb->ps, s));
        CMBL_TRY(lb_sync(lb));
        if (h->ys == 0 || h->yy == 0) break;
        cb(user, thisIter, lb->val, lb->gradNorm, lb->d.x, lb->d.g, lb->d.z);
        if (thisIter > max_iter) break;
    }
    if (d_res) CMBL_HIP(hipMemcpyAsync(d_res, lb->d.x, sizeof(double) * lb->n, hipMemcpyDeviceToDevice, s));
    CMBL_TRY(lb_sync(lb));
    if (f_out) *f_out = lb->val;
    if (iterations) *iterations = thisIter;
    if (evaluations) *evaluations = functionEval;
    return CMBL_OK;
}

int cmbl_lbfgs_gradient(cmbl_lbfgs* lb, double* d_g)
{
    if (!lb || !d_g) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipMemcpyAsync(d_g, lb->d.g, sizeof(double) * lb->n, hipMemcpyDeviceToDevice, lb->stream));
    CMBL_HIP(hipStreamSynchronize(lb->stream));
    return CMBL_OK;
}

int cmbl_lbfgs_state(const cmbl_lbfgs* lb, double* f, double* gnorm, int* iter, long* evaluations)
{
    if (!lb) return fail(CMBL_ERR_ARG, "null context");
    if (f) *f = lb->val;
    if (gnorm) *gnorm = lb->gradNorm;
    if (iter) *iter = lb->iter;
    if (evaluations) *evaluations = lb->evaluations;
    return CMBL_OK;
}

}  // extern "C"

// ====================================================================== HMC
namespace cmbl {

// Scalars of one trajectory (host-mapped).
struct HmcHost {
    double tau;   // step size (written by the host)
    double c;     // kick coefficient -tau / 2 / 2 (the derivative is of -2lnL)
    double f;     // -2lnL at the end of the trajectory
    double oldP;  // p . M^-1 p at its start
    double newP;  // p . M^-1 p at its end
};

struct HmcDev {
    long n;
    double *x, *p, *d, *prev, *dprev, *mass, *msq;
    double* noise[2];  // momentum draws of even / odd trajectories
    const double* f;
    HmcHost* h;
};

// Start of a trajectory (include/hmc_general.hpp:169-173, 233-246): momenta
// p = noise * sqrt(M), prev = x (and the gradient there), p . M^-1 p.
struct OpHmcInit {
    static constexpr int kDots = 1;
    HmcDev d;
    long n;
    const double* noise;
    __device__ void load() {}
    struct In {
        double noise, msq, x, d, mass;
    };
    __device__ In fetch(long e) const { return In{noise[e], d.msq[e], d.x[e], d.d[e], d.mass[e]}; }
    __device__ void apply(long e, const In& v, double* a)
    {
        const double pv = v.noise * v.msq;
        d.p[e] = pv;
        d.prev[e] = 1.0 * v.x;
        d.dprev[e] = 1.0 * v.d;
        const double t = (1.0 * pv) / v.mass;
        a[0] = fma(t, pv, a[0]);
    }
    __device__ void finish(const double* r) { d.h->oldP = r[0]; }
};

// Leapfrog (:177-189): the closing half kick of the previous step when
// `twice`, the opening half kick p += c d, and the drift x += tau p / M.
struct OpHmcKickDrift {
    static constexpr int kDots = 0;
    HmcDev d;
    long n;
    int twice;
    double c, tau;
    __device__ void load()
    {
        c = d.h->c;
        tau = d.h->tau;
    }
    struct In {
        double d, p, mass, x;
    };
    __device__ In fetch(long e) const { return In{d.d[e], d.p[e], d.mass[e], d.x[e]}; }
    __device__ void apply(long e, const In& v, double*)
    {
        double pv = v.p;
        if (twice) pv = fma(c, v.d, pv);
        pv = fma(c, v.d, pv);
        d.p[e] = pv;
        const double t = (1.0 * pv) / v.mass;
        d.x[e] = fma(tau, t, v.x);
    }
    __device__ void finish(const double*) {}
};

// End of a trajectory: the last half kick, p . M^-1 p, and -2lnL.
struct OpHmcFinal {
    static constexpr int kDots = 1;
    HmcDev d;
    long n;
    double c;
    __device__ void load() { c = d.h->c; }
    struct In {
        double d, p, mass;
    };
    __device__ In fetch(long e) const { return In{d.d[e], d.p[e], d.mass[e]}; }
    __device__ void apply(long e, const In& v, double* a)
    {
        const double pv = fma(c, v.d, v.p);
        d.p[e] = pv;
        const double t = (1.0 * pv) / v.mass;
        a[0] = fma(t, pv, a[0]);
    }
    __device__ void finish(const double* r)
    {
        d.h->newP = r[0];
        d.h->f = *d.f;
    }
};

// Value at the starting point: -2lnL (and |d|^2, unused).
struct OpHmcStart {
    static constexpr int kDots = 1;
    HmcDev d;
    long n;
    __device__ void load() {}
    struct In {
        double d;
    };
    __device__ In fetch(long e) const { return In{d.d[e]}; }
    __device__ void apply(long, const In& v, double* a) { a[0] = fma(v.d, v.d, a[0]); }
    __device__ void finish(const double*) { d.h->f = *d.f; }
};

// Rejected trajectory: back to the previous point and its gradient.
struct OpHmcReject {
    static constexpr int kDots = 0;
    HmcDev d;
    long n;
    __device__ void load() {}
    struct In {
        double prev, dprev;
    };
    __device__ In fetch(long e) const { return In{d.prev[e], d.dprev[e]}; }
    __device__ void apply(long e, const In& v, double*)
    {
        d.x[e] = 1.0 * v.prev;
        d.d[e] = 1.0 * v.dprev;
    }
    __device__ void finish(const double*) {}
};

}  // namespace cmbl

struct cmbl_hmc {
    cmbl_map* map = nullptr;
    long n = 0;
    double tauMax = 0;
    int nMax = 0;
    int flags = 0;
    bool graphs = true;
    hipStream_t stream = nullptr;
    LoopAllocs mem;
    HmcDev d{};
    HmcHost* h = nullptr;
    PassSync ps{};
    double* d_f = nullptr;
    GraphCache cache;
    std::mt19937 ugen;                               // UniformRealGenerator(seed, 0, 1)
    std::uniform_real_distribution<double> udist{0.0, 1.0};
    int gaussSeed = 0;
    double like = 0;
    bool started = false;
    long trajectories = 0, accepted = 0, evaluations = 0, gradCalls = 0;
    // momentum draws: pinned host buffers filled by worker threads kSlots
    // trajectories ahead, uploaded on a copy stream into the device buffer of
    // the trajectory's parity while the previous trajectory runs.  The
    // workers keep drawing ahead between runs (the seeds of later
    // trajectories are known), so a run does not start by waiting for draws.
    static constexpr int kSlots = 12;
    std::vector<double*> slot;
    std::vector<hipEvent_t> slotEv;  // upload from the slot complete
    std::vector<std::future<int>> fut;  // worker filling each slot
    std::vector<long> slotSeed;         // seed drawn (or being drawn) into each slot
    hipStream_t copy = nullptr;
    hipEvent_t upDone[2] = {nullptr, nullptr}, used[2] = {nullptr, nullptr};
    ~cmbl_hmc()
    {
        for (auto& f : fut)
            if (f.valid()) f.wait();
        cache.clear();
        if (stream) (void)hipStreamSynchronize(stream);
        if (copy) (void)hipStreamSynchronize(copy);
        for (hipEvent_t e : slotEv) (void)hipEventDestroy(e);
        for (int i = 0; i < 2; ++i) {
            if (upDone[i]) (void)hipEventDestroy(upDone[i]);
            if (used[i]) (void)hipEventDestroy(used[i]);
        }
        if (copy) (void)hipStreamDestroy(copy);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace cmbl {

static int hmc_eval(cmbl_hmc* hm, hipStream_t s) { return eval_device(hm->map, hm->d.x, hm->d_f, hm->d.d, s, false); }

// One trajectory of nsteps leapfrog steps (momenta already in d.noise).
static int hmc_trajectory(cmbl_hmc* hm, int nsteps, int parity, hipStream_t s)
{
    const HmcDev& d = hm->d;
    CMBL_TRY(run_pass(OpHmcInit{d, d.n, d.noise[parity]}, hm->ps, s));
    for (int j = 0; j < nsteps; ++j) {
        OpHmcKickDrift kd{};
        kd.d = d;
        kd.n = d.n;
        kd.twice = j > 0;
        CMBL_TRY(run_pass(kd, hm->ps, s));
        CMBL_TRY(hmc_eval(hm, s));
    }
    return run_pass(OpHmcFinal{d, d.n, 0.0}, hm->ps, s);
}

}  // namespace cmbl

extern "C" {

int cmbl_hmc_create(cmbl_hmc** out, cmbl_map* map, const double* d_start, const double* d_masses, double tau_max,
                    int n_max, int seed, int flags)
{
    if (!out || !map || !d_start || !d_masses) return fail(CMBL_ERR_ARG, "null pointer");
    *out = nullptr;
    if (!(tau_max > 0) || n_max < 1) return fail(CMBL_ERR_ARG, "tau_max must be > 0 and n_max >= 1");
    if (seed == 0) return fail(CMBL_ERR_ARG, "seed 0 means time(0) in the reference; pass a nonzero seed");
    if (!map->has_data) return fail(CMBL_ERR_STATE, "cmbl_map_set_data has not been called");
    cmbl_hmc* hm = new cmbl_hmc();
    auto bail = [&](int rc) {
        const std::string msg = cmbl_last_error();
        delete hm;
        set_error(msg);
        return rc;
    };
    hm->map = map;
    hm->n = map->nparam;
    hm->tauMax = tau_max;
    hm->nMax = n_max;
    hm->flags = flags;
    hm->graphs = (flags & CMBL_LOOP_EAGER) == 0;
    hm->ugen.seed(seed);
    hm->gaussSeed = seed + 1;
    if (hipStreamCreateWithFlags(&hm->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&hm->copy, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(CMBL_ERR_HIP, "hipStreamCreate failed"));
    for (int i = 0; i < 2; ++i)
        if (hipEventCreateWithFlags(&hm->upDone[i], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&hm->used[i], hipEventDisableTiming) != hipSuccess)
            return bail(fail(CMBL_ERR_HIP, "hipEventCreate failed"));
    HmcDev& d = hm->d;
    d.n = hm->n;
    const size_t n = (size_t)hm->n;
    int rc;
    if ((rc = hm->mem.dalloc(&d.x, n)) || (rc = hm->mem.dalloc(&d.p, n)) || (rc = hm->mem.dalloc(&d.d, n)) ||
        (rc = hm->mem.dalloc(&d.prev, n)) || (rc = hm->mem.dalloc(&d.dprev, n)) ||
        (rc = hm->mem.dalloc(&d.mass, n)) || (rc = hm->mem.dalloc(&d.msq, n)) ||
        (rc = hm->mem.dalloc(&d.noise[0], n)) || (rc = hm->mem.dalloc(&d.noise[1], n)) ||
        (rc = hm->mem.dalloc(&hm->d_f, 1)) || (rc = hm->mem.dalloc(&hm->ps.part, (size_t)kMaxPassDots * kDotParts)))
        return bail(rc);
    if ((rc = hm->mem.halloc(&hm->h, &d.h, 1))) return bail(rc);
    d.f = hm->d_f;
    for (int i = 0; i < cmbl_hmc::kSlots; ++i) {
        double* hp = nullptr;
        double* dp = nullptr;
        if ((rc = hm->mem.halloc(&hp, &dp, n))) return bail(rc);
        hm->slot.push_back(hp);
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return bail(fail(CMBL_ERR_HIP, "hipEventCreate failed"));
        hm->slotEv.push_back(e);
    }
    hm->fut.resize(cmbl_hmc::kSlots);
    hm->slotSeed.assign(cmbl_hmc::kSlots, LONG_MIN);
    // x = start, M = masses, sqrt(M) as the reference forms it (pow(M, 0.5))
    // copies on this context's stream, ordered before the square root (a
    // device-to-device hipMemcpy does not wait for completion, and this
    // stream does not order with the null stream)
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpyAsync(d.x, d_start, sizeof(double) * n, hipMemcpyDeviceToDevice, hm->stream) != hipSuccess ||
        hipMemcpyAsync(d.mass, d_masses, sizeof(double) * n, hipMemcpyDeviceToDevice, hm->stream) != hipSuccess ||
        hipMemcpyAsync(d.msq, d_masses, sizeof(double) * n, hipMemcpyDeviceToDevice, hm->stream) != hipSuccess)
        return bail(fail(CMBL_ERR_HIP, "HMC set-up copies failed"));
    if ((rc = launch_vec(kVecPow, d.msq, nullptr, 0.5, hm->n, hm->stream))) return bail(rc);
    if (hipStreamSynchronize(hm->stream) != hipSuccess) return bail(fail(CMBL_ERR_HIP, "HMC set-up failed"));
    *out = hm;
    return CMBL_OK;
}

void cmbl_hmc_destroy(cmbl_hmc* hm) { delete hm; }

int cmbl_hmc_run(cmbl_hmc* hm, int max_iters, cmbl_hmc_callback cb, void* user)
{
    if (!hm) return fail(CMBL_ERR_ARG, "null context");
    if (max_iters < 1) return fail(CMBL_ERR_ARG, "max_iters must be >= 1");
    hipStream_t s = hm->stream, cs = hm->copy;
    HmcHost* h = hm->h;
    const HmcDev& d = hm->d;
    const long n = hm->n;
    const size_t bytes = sizeof(double) * (size_t)n;
    // f_->set(x); currentLike = f_->value() (include/hmc_general.hpp:139-144)
    CMBL_TRY(hmc_eval(hm, s));
    CMBL_TRY(run_pass(OpHmcStart{d, d.n}, hm->ps, s));
    CMBL_HIP(hipStreamSynchronize(s));
    ++hm->evaluations;
    hm->like = h->f;
    if (!std::isfinite(hm->like)) return fail(CMBL_ERR_NONFINITE, "-2lnL at the starting point is not finite");
    if (cb) cb(user, d.x, hm->like);

    // momenta of trajectory k: whitenoise(gaussSeed + k), drawn on worker
    // threads up to kSlots trajectories ahead into pinned slots (a slot is
    // refilled once its upload has completed), uploaded on the copy stream
    // into noise[k % 2] once trajectory k - 2 has finished with it
    const int W = cmbl_hmc::kSlots;
    const long seed0 = hm->gaussSeed;
    auto slot_of = [W](long seed) { return (int)(((seed % W) + W) % W); };
    auto produce = [hm, n](long seed, int sl) -> int {
        if (hipEventSynchronize(hm->slotEv[sl]) != hipSuccess) return CMBL_ERR_HIP;
        return cmbl_whitenoise((int)seed, 1.0, n, hm->slot[sl]);
    };
    auto schedule = [&](long seed) {
        const int sl = slot_of(seed);
        if (hm->fut[sl].valid()) (void)hm->fut[sl].get();  // an unused earlier draw in that slot
        hm->slotSeed[sl] = seed;
        hm->fut[sl] = std::async(std::launch::async, produce, seed, sl);
    };
    auto upload = [&](int k) -> int {
        const long seed = seed0 + k;
        const int sl = slot_of(seed), par = k & 1;
        if (hm->slotSeed[sl] != seed || !hm->fut[sl].valid()) return fail(CMBL_ERR_STATE, "momentum draw missing");
        const int rc = hm->fut[sl].get();
        hm->slotSeed[sl] = LONG_MIN;
        if (rc != CMBL_OK) return fail(rc, "momentum draw failed");
        CMBL_HIP(hipStreamWaitEvent(cs, hm->used[par], 0));
        CMBL_HIP(hipMemcpyAsync(d.noise[par], hm->slot[sl], bytes, hipMemcpyHostToDevice, cs));
        CMBL_HIP(hipEventRecord(hm->upDone[par], cs));
        CMBL_HIP(hipEventRecord(hm->slotEv[sl], cs));
        schedule(seed + W);  // also past this run: the next run starts with its draws made
        return CMBL_OK;
    };
    // draws left by the previous run are kept; the rest of the window starts now
    for (int k = 0; k < W; ++k)
        if (hm->slotSeed[slot_of(seed0 + k)] != seed0 + k) schedule(seed0 + k);
    int rc = upload(0);
    for (int iter = 0; iter < max_iters && rc == CMBL_OK; ++iter) {
        double tau;
        int steps;
        if (hm->flags & CMBL_HMC_FIXED) {
            tau = hm->tauMax;
            steps = hm->nMax;
        } else {
            // the master's draws (:158-164)
            tau = hm->udist(hm->ugen) * hm->tauMax;
            steps = (int)std::ceil(hm->nMax * hm->udist(hm->ugen));
        }
        if (!(tau > 0 && tau <= hm->tauMax) || !(steps > 0 && steps <= hm->nMax)) {
            rc = fail(CMBL_ERR_STATE, "step size or step count out of range");
            break;
        }
        ++hm->gradCalls;  // f_->derivative(d_): the gradient at x is already held
        ++hm->gaussSeed;
        const int par = iter & 1;
        h->tau = tau;
        h->c = -tau / 2 / 2;
        if (hipStreamWaitEvent(s, hm->upDone[par], 0) != hipSuccess) {
            rc = fail(CMBL_ERR_HIP, "waiting for the momentum upload failed");
            break;
        }
        const long key = (long)steps * 2 + par;
        if ((rc = hm->cache.run(hm->graphs, key, s,
                                [&](hipStream_t st) { return hmc_trajectory(hm, steps, par, st); })))
            break;
        if (hipEventRecord(hm->used[par], s) != hipSuccess) {
            rc = fail(CMBL_ERR_HIP, "event record failed");
            break;
        }
        // next trajectory's momenta go up while this one runs
        if (iter + 1 < max_iters && (rc = upload(iter + 1)) != CMBL_OK) break;
        if (hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(CMBL_ERR_HIP, "trajectory failed");
            break;
        }
        hm->evaluations += steps;
        hm->gradCalls += steps;
        const double newLike = h->f;
        // Metropolis test (:196-206)
        const double deltaLike = newLike + h->newP - hm->like - h->oldP;
        const double pacc = std::exp(-deltaLike / 2);
        const double q = hm->udist(hm->ugen);
        const bool accept = q <= pacc;
        ++hm->trajectories;
        if (accept) {
            ++hm->accepted;
            hm->like = newLike;
        } else {
            if ((rc = run_pass(OpHmcReject{d, d.n}, hm->ps, s))) break;
            if (hipStreamSynchronize(s) != hipSuccess) {
                rc = fail(CMBL_ERR_HIP, "reject copy failed");
                break;
            }
        }
        if (cb) cb(user, d.x, hm->like);
    }
    // uploads finish before the call returns; the workers drawing ahead keep
    // going (the destructor waits for them)
    if (hipStreamSynchronize(cs) != hipSuccess && rc == CMBL_OK) rc = fail(CMBL_ERR_HIP, "momentum uploads failed");
    return rc;
}

int cmbl_hmc_state(const cmbl_hmc* hm, double* like, long* trajectories, long* accepted, long* evaluations)
{
    if (!hm) return fail(CMBL_ERR_ARG, "null context");
    if (like) *like = hm->like;
    if (trajectories) *trajectories = hm->trajectories;
    if (accepted) *accepted = hm->accepted;
    if (evaluations) *evaluations = hm->evaluations;
    return CMBL_OK;
}

int cmbl_hmc_last_trajectory(const cmbl_hmc* hm, double* like_end, double* kinetic_start, double* kinetic_end)
{
    if (!hm) return fail(CMBL_ERR_ARG, "null context");
    if (like_end) *like_end = hm->h->f;
    if (kinetic_start) *kinetic_start = hm->h->oldP;
    if (kinetic_end) *kinetic_end = hm->h->newP;
    return CMBL_OK;
}

int cmbl_hmc_position(cmbl_hmc* hm, double* d_out)
{
    if (!hm || !d_out) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipMemcpyAsync(d_out, hm->d.x, sizeof(double) * hm->n, hipMemcpyDeviceToDevice, hm->stream));
    CMBL_HIP(hipStreamSynchronize(hm->stream));
    return CMBL_OK;
}

}  // extern "C"

// ====================================================================== Gibbs CG
// Math::ConjugateGradient::solve (include/conjugate_gradient.hpp:45-102) on
// a Gibbs system with every vector and scalar on the device: each iteration
// is the operator, then three fused passes (p . Ap; the x and r updates with
// r . r, the preconditioner and r . z; the new direction), with alpha, beta
// and the stopping test formed by the passes' final kernels.  Blocks of
// kCgBlock iterations are captured in one HIP graph and replayed; the host
// reads the iteration count and the stop flag after each block.  Once the
// stop flag is set (|r|^2 <= precision^2, the iteration limit, or a
// breakdown), the rest of the block leaves x, r, z and p alone, so the
// result and the count are those of a loop that checks after every
// iteration.  The passes run the plain dot product's order and the BLAS-1
// operations of the host-driven loop, so the iterates are the same, bit for
// bit.
namespace cmbl {

constexpr int kCgBlock = 8;

struct CgScalars {
    double rz, alpha, beta, eps, prec2;
    int done, error, iters, max_iter;
};

struct CgView {
    double *x, *r, *z, *p, *ap;
    const double* dl;
    CgScalars* st;    // device copy
    CgScalars* host;  // host-mapped mirror, written after each iteration
};

__device__ __forceinline__ int packed_l_of(long p)
{
    long l = (long)sqrt((double)p);
    while (l * l > p) --l;
    while ((l + 1) * (l + 1) <= p) ++l;
    return (int)l;
}

// x = 0, r = b - A 0 = b, z = M^-1 r, p = z, rz = r . z; sets the limits.
struct OpCgStart {
    static constexpr int kDots = 1;
    CgView c;
    long n;
    const double* b;
    double prec2;
    int max_iter;
    struct In {
        double b;
    };
    __device__ void load() {}
    __device__ In fetch(long e) const { return In{b[e]}; }
    __device__ void apply(long e, const In& v, double* a)
    {
        const double r = 1.0 * v.b;
        const double z = r / c.dl[packed_l_of(e)];
        c.x[e] = 0.0;
        c.r[e] = r;
        c.z[e] = z;
        c.p[e] = 1.0 * z;
        a[0] = fma(r, z, a[0]);
    }
    __device__ void finish(const double* r)
    {
        CgScalars& st = *c.st;
        st.rz = r[0];
        st.alpha = st.beta = st.eps = 0.0;
        st.prec2 = prec2;
        st.done = st.error = st.iters = 0;
        st.max_iter = max_iter;
        *c.host = st;
    }
};

// p . Ap, alpha = rz / p . Ap (a zero p . Ap stops the solve as a breakdown).
struct OpCgPap {
    static constexpr int kDots = 1;
    CgView c;
    long n;
    struct In {
        double p, ap;
    };
    __device__ void load() {}
    __device__ In fetch(long e) const { return In{c.p[e], c.ap[e]}; }
    __device__ void apply(long, const In& v, double* a) { a[0] = fma(v.p, v.ap, a[0]); }
    __device__ void finish(const double* r)
    {
        CgScalars& st = *c.st;
        if (st.done) return;
        if (r[0] == 0.0) {
            st.error = 1;
            st.done = 1;
            *c.host = st;
            return;
        }
        st.alpha = st.rz / r[0];
    }
};

// x += alpha p, r -= alpha Ap, r . r, z = M^-1 r, r . z; then beta and the
// stopping test.
struct OpCgUpdate {
    static constexpr int kDots = 2;
    CgView c;
    long n;
    double alpha;
    int skip;
    struct In {
        double x, p, r, ap;
    };
    __device__ void load()
    {
        skip = c.st->done;
        alpha = c.st->alpha;
    }
    __device__ In fetch(long e) const
    {
        if (skip) return In{0.0, 0.0, 0.0, 0.0};
        return In{c.x[e], c.p[e], c.r[e], c.ap[e]};
    }
    __device__ void apply(long e, const In& v, double* a)
    {
        if (skip) return;
        const double x = fma(alpha, v.p, v.x);
        const double r = fma(-alpha, v.ap, v.r);
        a[0] = fma(r, r, a[0]);
        const double z = r / c.dl[packed_l_of(e)];
        a[1] = fma(r, z, a[1]);
        c.x[e] = x;
        c.r[e] = r;
        c.z[e] = z;
    }
    __device__ void finish(const double* r)
    {
        CgScalars& st = *c.st;
        if (st.done) return;
        st.eps = r[0];
        st.iters += 1;
        if (!isfinite(r[0])) {
            st.error = 2;
            st.done = 1;
        } else if (st.rz == 0.0) {
            st.error = 3;
            st.done = 1;
        } else {
            st.beta = r[1] / st.rz;
            st.rz = r[1];
            if (!(r[0] > st.prec2) || st.iters >= st.max_iter) st.done = 1;
        }
        *c.host = st;
    }
};

// p = z + beta p (skipped once the solve has stopped: p is not used again).
struct OpCgDir {
    static constexpr int kDots = 0;
    CgView c;
    long n;
    double beta;
    int skip;
    struct In {
        double z, p;
    };
    __device__ void load()
    {
        skip = c.st->done;
        beta = c.st->beta;
    }
    __device__ In fetch(long e) const
    {
        if (skip) return In{0.0, 0.0};
        return In{c.z[e], c.p[e]};
    }
    __device__ void apply(long e, const In& v, double*)
    {
        if (skip) return;
        c.p[e] = fma(1.0, v.z, beta * v.p);
    }
    __device__ void finish(const double*) {}
};

struct GibbsCg {
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    LoopAllocs mem;
    PassSync ps{};
    CgView v{};
    CgScalars* h = nullptr;  // host view of v.host
    GraphCache cache;
    ~GibbsCg()
    {
        cache.clear();
        if (stream) (void)hipStreamSynchronize(stream);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

void gibbs_cg_free(GibbsCg* cg) { delete cg; }

static int gibbs_cg_make(cmbl_gibbs* gb)
{
    if (gb->cg) return CMBL_OK;
    GibbsCg* cg = new GibbsCg();
    auto bail = [&](int rc) {
        const std::string msg = cmbl_last_error();
        delete cg;
        set_error(msg);
        return rc;
    };
    if (hipStreamCreateWithFlags(&cg->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&cg->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&cg->ev_out, hipEventDisableTiming) != hipSuccess)
        return bail(fail(CMBL_ERR_HIP, "Gibbs CG stream set-up failed"));
    const size_t n = (size_t)gb->n;
    int rc;
    CgView& v = cg->v;
    if ((rc = cg->mem.dalloc(&v.x, n)) || (rc = cg->mem.dalloc(&v.st, 1)) ||
        (rc = cg->mem.dalloc(&cg->ps.part, (size_t)kMaxPassDots * kDotParts)) ||
        (rc = cg->mem.halloc(&cg->h, &v.host, 1)))
        return bail(rc);
    v.r = gb->d_r;
    v.z = gb->d_z;
    v.p = gb->d_p;
    v.ap = gb->d_ap;
    v.dl = gb->d_dl;
    gb->cg = cg;
    return CMBL_OK;
}

static int gibbs_cg_block(cmbl_gibbs* gb, hipStream_t s)
{
    const CgView& v = gb->cg->v;
    const long n = gb->n;
    for (int i = 0; i < kCgBlock; ++i) {
        CMBL_TRY(cmbl_gibbs_apply_device(gb, v.p, v.ap, s));
        CMBL_TRY(run_pass(OpCgPap{v, n}, gb->cg->ps, s));
        CMBL_TRY(run_pass(OpCgUpdate{v, n, 0.0, 0}, gb->cg->ps, s));
        CMBL_TRY(run_pass(OpCgDir{v, n, 0.0, 0}, gb->cg->ps, s));
    }
    return CMBL_OK;
}

}  // namespace cmbl

extern "C" {

int cmbl_gibbs_solve_device(cmbl_gibbs* gb, const double* b, double* x, double precision, int max_iter,
                            int* iterations, void* stream)
{
    if (!gb || !gb->mp) return fail(CMBL_ERR_ARG, "gibbs context is null");
    if (!b || !x) return fail(CMBL_ERR_ARG, "null pointer");
    if (!(precision > 0.0)) return fail(CMBL_ERR_ARG, "precision must be > 0");
    if (max_iter < 1) return fail(CMBL_ERR_ARG, "max_iter must be >= 1");
    if (iterations) *iterations = 0;
    CMBL_TRY(gibbs_cg_make(gb));
    GibbsCg* cg = gb->cg;
    hipStream_t caller = (hipStream_t)stream;
    hipStream_t s = cg->stream;
    CMBL_HIP(hipEventRecord(cg->ev_in, caller));
    CMBL_HIP(hipStreamWaitEvent(s, cg->ev_in, 0));
    OpCgStart st0{cg->v, gb->n, b, precision * precision, max_iter};
    CMBL_TRY(run_pass(st0, cg->ps, s));
    for (;;) {
        CMBL_TRY(cg->cache.run(true, kCgBlock, s, [&](hipStream_t q) { return gibbs_cg_block(gb, q); }));
        CMBL_HIP(hipStreamSynchronize(s));
        if (__atomic_load_n(&cg->h->done, __ATOMIC_ACQUIRE)) break;
    }
    const CgScalars r = *cg->h;
    CMBL_HIP(hipMemcpyAsync(x, cg->v.x, sizeof(double) * gb->n, hipMemcpyDeviceToDevice, s));
    CMBL_HIP(hipEventRecord(cg->ev_out, s));
    CMBL_HIP(hipStreamWaitEvent(caller, cg->ev_out, 0));
    CMBL_HIP(hipStreamSynchronize(s));
    if (iterations) *iterations = r.iters;
    if (r.error == 1) return fail(CMBL_ERR_STATE, "CG breakdown: p.Ap = 0");
    if (r.error == 2) return fail(CMBL_ERR_NONFINITE, "CG residual is not finite");
    if (r.error == 3) return fail(CMBL_ERR_STATE, "CG breakdown: r.z = 0");
    if (r.eps > r.prec2)
        return fail(CMBL_ERR_NOCONV,
                    "CG reached " + std::to_string(max_iter) + " iterations, |r|^2 = " + std::to_string(r.eps));
    return CMBL_OK;
}

}  // extern "C"

// ====================================================================== line search
struct cmbl_line_search {
    cmbl::MoreThuente mt;
    bool started = false;
};

extern "C" {

int cmbl_line_search_create(cmbl_line_search** out, double ftol, double gtol, double xtol, double stpmin,
                            double stpmax, int maxfev)
{
    if (!out) return fail(CMBL_ERR_ARG, "out is null");
    *out = nullptr;
    if (!(ftol >= 0) || !(gtol >= 0) || !(xtol >= 0) || !(stpmin >= 0) || !(stpmax >= stpmin) || maxfev < 1)
        return fail(CMBL_ERR_ARG, "line search parameters out of range");
    cmbl_line_search* ls = new cmbl_line_search();
    ls->mt.ftol = ftol;
    ls->mt.gtol = gtol;
    ls->mt.xtol = xtol;
    ls->mt.stpmin = stpmin;
    ls->mt.stpmax = stpmax;
    ls->mt.maxfev = maxfev;
    *out = ls;
    return CMBL_OK;
}

void cmbl_line_search_destroy(cmbl_line_search* ls) { delete ls; }

int cmbl_line_search_begin(cmbl_line_search* ls, double f0, double dg0, double stp0, double* stp)
{
    if (!ls || !stp) return fail(CMBL_ERR_ARG, "null pointer");
    if (!(stp0 > 0)) return fail(CMBL_ERR_ARG, "the initial step must be > 0");
    if (!(dg0 < 0)) return fail(CMBL_ERR_ARG, "the initial slope must be < 0 (a descent direction)");
    *stp = ls->mt.begin(f0, dg0, stp0);
    ls->started = true;
    return CMBL_OK;
}

int cmbl_line_search_feed(cmbl_line_search* ls, double f, double dg, double* stp, int* info)
{
    if (!ls || !stp || !info) return fail(CMBL_ERR_ARG, "null pointer");
    if (!ls->started) return fail(CMBL_ERR_STATE, "cmbl_line_search_begin has not been called");
    *info = ls->mt.feed(f, dg, stp);
    if (*info != 0) {
        *stp = ls->mt.stp;
        ls->started = false;
    }
    return CMBL_OK;
}

}  // extern "C"
