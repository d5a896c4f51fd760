// HIP kernels of the map-level likelihood for gfx950 (CDNA4, wave64, FP64).
//
// Stages of one likelihood + gradient evaluation (all FP64):
//   k_pre     packed a_lm (almToVector layout, source/cmb_gibbs.cpp:320-357)
//             -> m-major B_l N_lm a_lm, and the prior partial sums
//   k_synth   Legendre synthesis (the a_lm -> ring Fourier part of HEALPix
//             alm2map): per (m, ring pair) parity sums E_m, O_m
//   k_ring    per ring pair: phase fold + inverse ring DFT -> map s,
//             r = d - s, w r, chi^2 partial (source/cmb_gibbs.cpp:381-387),
//             forward ring DFT + unfold -> Psi_m (map2alm's ring part with
//             ring weight Npix/4pi, source/cmb_gibbs.cpp:390-391)
//   k_adj     adjoint Legendre: u_lm = sum_rings lambda_lm Psi_m
//   k_post    gradient assembly into the packed layout
//   k_final   deterministic -2lnL reduction
#include <atomic>
#include <cstdlib>

#include "cmbl_device.hpp"
#include "cmbl_internal.hpp"

namespace cmbl {

// ================================================================== helpers

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b)  // a * conj(b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// e^{i pi q / n} for integer q (argument reduced exactly modulo 2n)
__device__ __forceinline__ double2 cis_pi_frac(long q, long n)
{
    const long two_n = 2 * n;
    long r = q % two_n;
    if (r < 0) r += two_n;
    double s, c;
    sincospi((double)r / (double)n, &s, &c);
    return make_double2(c, s);
}

// Ordering point for LDS data exchanged inside one wavefront.  The Legendre
// kernels run one wavefront per workgroup and the LDS executes the operations
// of a wavefront in issue order, so a store is visible to every later load of
// the same wavefront: the compiler only must not move memory operations across
// this point.  Unlike __syncthreads() it costs no lgkmcnt(0) wait.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Global-address-space views of device memory: loads and stores through them
// compile to global_* instructions (counted by vmcnt only) rather than flat_*
// ones, whose lgkmcnt share makes every LDS wait also wait for memory, and
// which the compiler must order after earlier LDS and global stores.
typedef __attribute__((address_space(1))) double gdouble;
typedef const __attribute__((address_space(1))) double gcdouble;
typedef double __attribute__((ext_vector_type(2))) dvec2;

// Views of double2 arrays in global memory (16-byte loads and stores).
struct GCView2 {
    const __attribute__((address_space(1))) dvec2* p;
    __device__ __forceinline__ double2 operator[](long i) const
    {
        const dvec2 v = p[i];
        return make_double2(v.x, v.y);
    }
    __device__ __forceinline__ GCView2 operator+(long k) const { return GCView2{p + k}; }
};
struct GView2 {
    __attribute__((address_space(1))) dvec2* p;
    struct Ref {
        __attribute__((address_space(1))) dvec2* q;
        __device__ __forceinline__ void operator=(double2 a) const
        {
            dvec2 v;
            v.x = a.x;
            v.y = a.y;
            *q = v;
        }
        __device__ __forceinline__ operator double2() const
        {
            const dvec2 v = *q;
            return make_double2(v.x, v.y);
        }
    };
    __device__ __forceinline__ Ref operator[](long i) const { return Ref{p + i}; }
    __device__ __forceinline__ GView2 operator+(long k) const { return GView2{p + k}; }
};

__device__ __forceinline__ gcdouble* gview(const double* p) { return (gcdouble*)p; }
__device__ __forceinline__ gdouble* gview(double* p) { return (gdouble*)p; }
__device__ __forceinline__ GCView2 gview(const double2* p)
{
    return GCView2{(const __attribute__((address_space(1))) dvec2*)p};
}
__device__ __forceinline__ GView2 gview(double2* p) { return GView2{(__attribute__((address_space(1))) dvec2*)p}; }

// ================================================================== k_pre / k_post
// The packed parameter vector (almToVector: l-major, index l^2 + 2m - 1 + c)
// and the m-major device arrays (alm, Nrm, upart: index moff(L, m) + l - m)
// are transposes of each other over (l, m).  Both kernels work on kLM x kLM
// tiles of the triangle m <= l <= L, one tile per block, and transpose through
// LDS so that each side is read or written in its own contiguous order.
constexpr int kLM = 32;

__device__ __forceinline__ void tile_of(int b, int& tl, int& tm)
{
    int t = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while (t * (t + 1) / 2 > b) --t;
    while ((t + 1) * (t + 2) / 2 <= b) ++t;
    tl = t;
    tm = b - t * (t + 1) / 2;
}

// Packed index of (l, m, c) (c = 0 real, 1 imaginary), -1 for the absent Im a_l0.
__device__ __forceinline__ long packed_index(int l, int m, int c)
{
    const long li = (long)l * l;
    if (m == 0) return c == 0 ? li : -1;
    return li + 2 * m - 1 + c;
}

// alm = B_l N_lm a_lm (m-major) from the packed x, and the prior partial sum
// of the tile: sum invcl_l (|a_l0|^2 + 2 sum_m>0 |a_lm|^2).
__device__ __forceinline__ void pre_tile(const MapDesc& md, const double* __restrict__ x, int tile)
{
    __shared__ double T[kLM][2 * kLM + 1];
    __shared__ double sh[4];
    const int L = md.L;
    int tl, tm;
    tile_of(tile, tl, tm);
    const int l0 = tl * kLM, m0 = tm * kLM;
    // all loads of a thread are issued before their first use (entries
    // outside the triangle read entry 0 and are dropped)
    constexpr int kX = kLM * 2 * kLM / 256, kE = kLM * kLM / 256;
    double xv[kX];
#pragma unroll
    for (int e = 0; e < kX; ++e) {
        const int i = threadIdx.x + 256 * e;
        const int lr = i / (2 * kLM), c = i % (2 * kLM);
        const int l = l0 + lr, m = m0 + (c >> 1);
        const long pi = (l <= L && m <= l) ? packed_index(l, m, c & 1) : -1;
        xv[e] = gview(x)[pi >= 0 ? pi : 0];
        xv[e] = pi >= 0 ? xv[e] : 0.0;
    }
    double sc[kE], ic[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const int i = threadIdx.x + 256 * e;
        const int l = l0 + i % kLM, m = m0 + i / kLM;
        const bool ok = l <= L && m <= l;
        sc[e] = gview(md.beam)[ok ? l : 0] * gview(md.Nrm)[ok ? moff(L, m) + (l - m) : 0];
        ic[e] = gview(md.invcl)[ok ? l : 0];
    }
#pragma unroll
    for (int e = 0; e < kX; ++e) {
        const int i = threadIdx.x + 256 * e;
        T[i / (2 * kLM)][i % (2 * kLM)] = xv[e];
    }
    __syncthreads();
    double pr = 0.0;
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const int i = threadIdx.x + 256 * e;
        const int mr = i / kLM, lr = i % kLM;
        const int l = l0 + lr, m = m0 + mr;
        if (l <= L && m <= l) {
            const double re = T[lr][2 * mr], im = T[lr][2 * mr + 1];
            const long k = moff(L, m) + (l - m);
            md.alm[k] = make_double2(sc[e] * re, sc[e] * im);
            pr += ic[e] * ((m == 0) ? re * re : 2.0 * (re * re + im * im));
        }
    }
    const double t = block_sum<256>(pr, sh);
    if (threadIdx.x == 0) md.prior[tile] = t;
}

__global__ __launch_bounds__(256) void k_pre(const MapDesc* __restrict__ md_, const double* __restrict__ x)
{
    pre_tile(*md_, x, blockIdx.x);
}

// Map of block b in a batched launch (wave-uniform, at most kBatchGroup steps).
__device__ __forceinline__ int batch_map(const BatchArgs& a, int b)
{
    int i = 0;
    while (i + 1 < a.n && a.tile0[i + 1] <= b) ++i;
    return i;
}

__global__ __launch_bounds__(256) void k_pre_b(const MapDesc* __restrict__ mds, const BatchArgs a)
{
    const int i = batch_map(a, blockIdx.x);
    pre_tile(mds[i], a.x[i], blockIdx.x - a.tile0[i]);
}

// u = N_lm B_l sum_rb upart (fixed order) -> the packed output of `mode`
// (PostMode): the gradient forms kPostGrad / kPostLike, or the Gibbs forms in
// the HEALPix complex-a_lm convention (no factor 2 for m >= 1).  kPostAddAlm
// scales by c instead of B_l and adds to g.
__device__ __forceinline__ void post_tile(const MapDesc& md, const double* __restrict__ x, double* __restrict__ g,
                                          int mode, double c, int tile)
{
    __shared__ double T[kLM][2 * kLM + 1];
    const int L = md.L;
    int tl, tm;
    tile_of(tile, tl, tm);
    const int l0 = tl * kLM, m0 = tm * kLM;
    constexpr int kE = kLM * kLM / 256;  // elements per thread
    // every load of the thread's elements is issued before the sums (entries
    // outside the triangle read entry 0 and are dropped)
    long k[kE];
    bool ok[kE];
    double sc[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const int i = threadIdx.x + 256 * e;
        const int l = l0 + i % kLM, m = m0 + i / kLM;
        ok[e] = l <= L && m <= l;
        k[e] = ok[e] ? moff(L, m) + (l - m) : 0;
        const double s = (mode == kPostAddAlm || mode == kPostQuad) ? c : gview(md.beam)[ok[e] ? l : 0];
        sc[e] = gview(md.Nrm)[k[e]] * s;
    }
    const GCView2 u = gview(reinterpret_cast<const double2*>(md.upart));
    const long stride = md.nalm;  // double2 per ring block
    double ur[kE], ui[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) ur[e] = ui[e] = 0.0;
    // eight partials of two elements in flight per thread, added in
    // ring-block order
#pragma unroll
    for (int e0 = 0; e0 < kE; e0 += 2) {
        int rb = 0;
        for (; rb + 8 <= md.nrb5; rb += 8) {
            double2 v[2][8];
#pragma unroll
            for (int e = 0; e < 2; ++e)
#pragma unroll
                for (int j = 0; j < 8; ++j) v[e][j] = u[k[e0 + e] + (rb + j) * stride];
#pragma unroll
            for (int e = 0; e < 2; ++e)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    ur[e0 + e] += v[e][j].x;
                    ui[e0 + e] += v[e][j].y;
                }
        }
        for (; rb < md.nrb5; ++rb) {
            double2 v[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) v[e] = u[k[e0 + e] + rb * stride];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                ur[e0 + e] += v[e].x;
                ui[e0 + e] += v[e].y;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const int i = threadIdx.x + 256 * e;
        const int mr = i / kLM, lr = i % kLM;
        T[lr][2 * mr] = ok[e] ? ur[e] * sc[e] : 0.0;
        T[lr][2 * mr + 1] = ok[e] ? ui[e] * sc[e] : 0.0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLM * 2 * kLM; i += 256) {
        const int lr = i / (2 * kLM), c2 = i % (2 * kLM);
        const int l = l0 + lr, m = m0 + (c2 >> 1);
        const long pi = (l <= L && m <= l) ? packed_index(l, m, c2 & 1) : -1;
        if (pi < 0) continue;
        const double u = T[lr][c2];
        switch (mode) {
        case kPostGrad: g[pi] = (m == 0) ? u : 2.0 * u; break;
        case kPostLike:
            g[pi] = (m == 0) ? -2.0 * u + 2.0 * md.invcl[l] * x[pi] : -4.0 * u + 4.0 * md.invcl[l] * x[pi];
            break;
        case kPostGibbs: g[pi] = x[pi] + u; break;
        case kPostAlm:
        case kPostQuad: g[pi] = u; break;
        default: g[pi] += u; break;  // kPostAddAlm
        }
    }
}

__global__ __launch_bounds__(256) void k_post(const MapDesc* __restrict__ md_, const double* __restrict__ x,
                                              double* __restrict__ g, int mode, double c)
{
    post_tile(*md_, x, g, mode, c, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_post_b(const MapDesc* __restrict__ mds, const BatchArgs a, int mode)
{
    const int i = batch_map(a, blockIdx.x);
    post_tile(mds[i], a.x[i], a.g[i], mode, 0.0, blockIdx.x - a.tile0[i]);
}

// y_p = x_p / d_l over the packed layout (l = floor(sqrt(p))): the Gibbs
// preconditioner (source/cmb_gibbs.cpp:407-423 divides every a_lm by d_l).
// y may be x (in place).
__global__ __launch_bounds__(256) void k_ldiv(double* y, const double* x, const double* __restrict__ dl, int L)
{
    const long n = (long)(L + 1) * (L + 1);
    for (long p = blockIdx.x * 256L + threadIdx.x; p < n; p += (long)gridDim.x * 256L) {
        long l = (long)sqrt((double)p);
        while (l * l > p) --l;
        while ((l + 1) * (l + 1) <= p) ++l;
        y[p] = x[p] / dl[l];
    }
}

// ================================================================== spin-2 synthesis
// alm2map_pol (HEALPix; called at source/likelihood.cpp:585 and
// source/c_matrix_generator.cpp:577-578): T = sum a^T Y, and with
// W_lm = (2Y_lm + -2Y_lm)/2 = F1_lm(theta) e^{i m phi},
// X_lm = (2Y_lm - -2Y_lm)/2 = F2_lm(theta) e^{i m phi},
//   Q = -sum (a^E W + i a^B X),   U = -sum (a^B W - i a^E X).
// With kappa_l = 2 / sqrt((l-1) l (l+1) (l+2)), q_lm = sqrt((2l+1)/(2l-1) (l^2 - m^2)),
// c = cos(theta), s2 = sin^2(theta) and lambda_lm = Y_lm(theta, 0):
//   F1_lm = kappa_l [-((l - m^2)/s2 + l(l-1)/2) lambda_lm + c q_lm lambda_{l-1,m} / s2]
//   F2_lm = kappa_l (m/s2) [-(l-1) c lambda_lm + q_lm lambda_{l-1,m}].
// Per field the sums over l need four scalar Legendre sums of one recursion:
//   Sa = sum kappa a lambda_l,  Sb = sum l kappa a lambda_l,  Sc = sum l^2 kappa a lambda_l,
//   D3 = sum kappa q a lambda_{l-1}  (coefficients shifted by one l),
// and with D1 = Sb - m^2 Sa, D2 = (Sc - Sb)/2, D4 = Sb - Sa:
//   sum_l a F1 = -D1/s2 - D2 + c D3/s2,   sum_l a F2 = m (D3 - c D4)/s2.
// k_synth_pol runs the recursion once per (m, ring block) for the T sum and
// the eight field sums and forms Q_m, U_m per ring in its epilogue.

// Record of (l, m) (m-major, kPolRec complex): N a^T, N kappa_l a^E,
// N kappa_{l+1} q_{l+1,m} a^E_{l+1,m}, and the same two for B.
__global__ __launch_bounds__(256) void k_pre_pol(const MapDesc* __restrict__ md_, const double* __restrict__ t,
                                                 const double* __restrict__ e, const double* __restrict__ b,
                                                 double2* __restrict__ rec)
{
    __shared__ double T[3][kLM + 1][2 * kLM + 1];
    const MapDesc& md = *md_;
    const int L = md.L;
    int tl, tm;
    tile_of(blockIdx.x, tl, tm);
    const int l0 = tl * kLM, m0 = tm * kLM;
    // rows l0 .. l0 + kLM: the extra row feeds the shifted coefficients
    for (int i = threadIdx.x; i < (kLM + 1) * 2 * kLM; i += 256) {
        const int lr = i / (2 * kLM), c = i % (2 * kLM);
        const int l = l0 + lr, m = m0 + (c >> 1);
        const long pi = (l <= L && m <= l) ? packed_index(l, m, c & 1) : -1;
        T[0][lr][c] = pi >= 0 ? t[pi] : 0.0;
        T[1][lr][c] = pi >= 0 ? e[pi] : 0.0;
        T[2][lr][c] = pi >= 0 ? b[pi] : 0.0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLM * kLM; i += 256) {
        const int mr = i / kLM, lr = i % kLM;
        const int l = l0 + lr, m = m0 + mr;
        if (l > L || m > l) continue;
        const long k = moff(L, m) + (l - m);
        const double nr = md.Nrm[k];
        const double fl = l, m2 = (double)m * m;
        const double ck = l >= 2 ? nr * 2.0 / sqrt((fl - 1.0) * fl * (fl + 1.0) * (fl + 2.0)) : 0.0;
        const double l1 = fl + 1.0;
        const double dk = (l + 1 <= L && l + 1 >= 2)
                              ? nr * 2.0 / sqrt((l1 - 1.0) * l1 * (l1 + 1.0) * (l1 + 2.0)) *
                                    sqrt((2.0 * l1 + 1.0) / (2.0 * l1 - 1.0) * (l1 * l1 - m2))
                              : 0.0;
        double2* o = rec + k * kPolRec;
        o[0] = make_double2(nr * T[0][lr][2 * mr], nr * T[0][lr][2 * mr + 1]);
        o[1] = make_double2(ck * T[1][lr][2 * mr], ck * T[1][lr][2 * mr + 1]);
        o[2] = make_double2(dk * T[1][lr + 1][2 * mr], dk * T[1][lr + 1][2 * mr + 1]);
        o[3] = make_double2(ck * T[2][lr][2 * mr], ck * T[2][lr][2 * mr + 1]);
        o[4] = make_double2(dk * T[2][lr + 1][2 * mr], dk * T[2][lr + 1][2 * mr + 1]);
    }
}

// ================================================================== MASTER coupling
// Mode-coupling matrix of a mask with pseudo spectrum W (source/master.cpp:180-215):
//   M[l1][l2] = (2 l2 + 1) / (4 pi) sum_{l3 <= L} (2 l3 + 1) W_l3 (l1 l2 l3; 0 0 0)^2.
// Only squares of the 3j symbols enter, and with g = (l1 + l2 + l3) / 2,
// a = g - l1, b = g - l2, c = g - l3 their closed form gives two rational
// recursions (no square roots):
//   (l1 l2 l1-l2)^2:  1 / (2 l1 + 1) at l2 = 0, times (2c+1) b / ((c+1)(2b-1)) per l2 step;
//   l3 -> l3 + 2:     times (2a+1)(2b+1) c (g+1) / ((a+1)(b+1)(2c-1)(2g+3)).
// k_w3j_start tabulates the first, one thread per l1 (s0[l1 (L+1) + l2], l2 <= l1).
__global__ __launch_bounds__(256) void k_w3j_start(int L, double* __restrict__ s0)
{
    const int l1 = blockIdx.x * 256 + threadIdx.x;
    if (l1 > L) return;
    double s = 1.0 / (2.0 * l1 + 1.0);
    double* row = s0 + (long)l1 * (L + 1);
    row[0] = s;
    for (int k = 0; k < l1; ++k) {
        const double b = l1 - k, c = k;
        s *= (2.0 * c + 1.0) * b / ((c + 1.0) * (2.0 * b - 1.0));
        row[k + 1] = s;
    }
}

// One thread per pair l2 <= l1 (consecutive threads: consecutive l2, so the
// l3 loops of a wavefront have similar lengths); writes M[l1][l2] and M[l2][l1].
__global__ __launch_bounds__(256) void k_coupling(int L, const double* __restrict__ W, const double* __restrict__ s0,
                                                  double* __restrict__ M)
{
    const long np = (long)(L + 1) * (L + 2) / 2;
    for (long p = blockIdx.x * 256L + threadIdx.x; p < np; p += (long)gridDim.x * 256L) {
        long l1 = (long)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
        while (l1 * (l1 + 1) / 2 > p) --l1;
        while ((l1 + 1) * (l1 + 2) / 2 <= p) ++l1;
        const int i1 = (int)l1, i2 = (int)(p - l1 * (l1 + 1) / 2);
        double s = s0[l1 * (L + 1) + i2];
        double acc = 0.0;
        const int top = min(i1 + i2, L);
        for (int l3 = i1 - i2; l3 <= top; l3 += 2) {
            acc = fma((2.0 * l3 + 1.0) * W[l3], s, acc);
            const int g = (i1 + i2 + l3) >> 1;
            const double a = g - i1, b = g - i2, c = g - l3;
            s *= (2.0 * a + 1.0) * (2.0 * b + 1.0) * c * (g + 1.0) /
                 ((a + 1.0) * (b + 1.0) * (2.0 * c - 1.0) * (2.0 * g + 3.0));
        }
        const double inv4pi = 0.25 / 3.141592653589793;
        M[l1 * (L + 1) + i2] = (2.0 * i2 + 1.0) * inv4pi * acc;
        M[(long)i2 * (L + 1) + l1] = (2.0 * i1 + 1.0) * inv4pi * acc;
    }
}

// NEST <-> RING reordering of a map: out[i] = in[pixel of i in the other
// scheme].  Writes are coalesced, reads gathered.
__global__ __launch_bounds__(256) void k_reorder(const double* __restrict__ in, double* __restrict__ out, long nside,
                                                 int to_ring)
{
    const long npix = 12 * nside * nside;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < npix; i += (long)gridDim.x * 256L)
        out[i] = in[to_ring ? ring_to_nest(nside, i) : nest_to_ring(nside, i)];
}

// ================================================================== Legendre stages
//
// One wavefront (= one workgroup) per work item: degree m of one ring block
// of one map.  Lane `lane` handles ring pairs p_k = rb*64*K + k*64 + lane,
// k < K.  The recurrence mu_l = A_lm x mu_{l-1} - mu_{l-2} runs in registers,
// starting from the state precomputed at set-up for the item's first live l.
// The wave-uniform coefficients (A_lm and, for the synthesis, B_l N_lm a_lm)
// are staged through a per-wave LDS double buffer, kChunk l-steps at a time:
// every lane issues one vector load of the next chunk before the current
// chunk is processed, so the loads overlap kChunk steps of FP64 work.
// Lanes whose scale count is > 0 are negligible (|lambda| < 2^-100): the wave
// runs a "scaled" path with rescale checks until every lane is live, then a
// check-free path of 4 FP64 VALU operations per (l, ring pair).

template <int K>
__device__ __forceinline__ bool all_live(const int (&sc)[K])
{
    int ok = 1;
#pragma unroll
    for (int k = 0; k < K; ++k) ok &= (sc[k] == 0);
    return __all(ok);
}
template <int K>
__device__ __forceinline__ bool any_live(const int (&sc)[K])
{
    int ok = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) ok |= (sc[k] == 0);
    return __any(ok);
}

template <int K>
__device__ __forceinline__ bool rescale(double (&mu1)[K], double (&mu2)[K], int (&sc)[K])
{
    int need = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) need |= (sc[k] > 0) & (fabs(mu1[k]) > kScaleHi);
    if (!__any(need)) return false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (sc[k] > 0 && fabs(mu1[k]) > kScaleHi) {
            mu1[k] *= kScaleDn;
            mu2[k] *= kScaleDn;
            sc[k] -= 1;
        }
    }
    return true;
}

// Phase state shared by both Legendre kernels.
template <int K>
struct LegState {
    double x[K];
    double mu1[K], mu2[K];  // mu_l, mu_{l-1}
    double gate[K];         // 1 for live lanes, 0 otherwise (scaled path only)
    int sc[K];
    bool live, all;
    __device__ __forceinline__ void refresh()
    {
        live = any_live<K>(sc);
        all = all_live<K>(sc);
#pragma unroll
        for (int k = 0; k < K; ++k) gate[k] = sc[k] == 0 ? 1.0 : 0.0;
    }
    __device__ __forceinline__ void advance(double Al)
    {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double nu = fma(Al * x[k], mu1[k], -mu2[k]);
            mu2[k] = mu1[k];
            mu1[k] = nu;
        }
    }
};

// -------------------------------------------------------------- synthesis
template <int K>
struct SynAcc {
    double er[K], ei[K], orr[K], oi[K];
};

template <int K, bool ODD>
__device__ __forceinline__ void syn_live(LegState<K>& st, SynAcc<K>& ac, double Al, double ar, double ai)
{
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double nu = fma(Al * st.x[k], st.mu1[k], -st.mu2[k]);
        st.mu2[k] = st.mu1[k];
        st.mu1[k] = nu;
        if (ODD) {
            ac.orr[k] = fma(nu, ar, ac.orr[k]);
            ac.oi[k] = fma(nu, ai, ac.oi[k]);
        } else {
            ac.er[k] = fma(nu, ar, ac.er[k]);
            ac.ei[k] = fma(nu, ai, ac.ei[k]);
        }
    }
}

// Scaled lanes (count > 0) accumulate like live ones: their scaled mu
// (|mu| < 2^945 between rescale checks) makes meaningless sums, which are
// cleared when the lane turns live (its count reaching 0 at a rescale check)
// and never stored while it is scaled, so the result equals gating each term
// by the lane's state at the last check, without the multiplies.

// The item's first l (the recurrence state is already at it).
template <int K, bool ODD>
__device__ __forceinline__ void syn_first(const LegState<K>& st, SynAcc<K>& ac, double ar, double ai)
{
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (ODD) {
            ac.orr[k] = fma(st.mu1[k], ar, ac.orr[k]);
            ac.oi[k] = fma(st.mu1[k], ai, ac.oi[k]);
        } else {
            ac.er[k] = fma(st.mu1[k], ar, ac.er[k]);
            ac.ei[k] = fma(st.mu1[k], ai, ac.ei[k]);
        }
    }
}

// Rescale check of the synthesis' scaled path: lanes above 2^900 drop by
// 2^-1000; a lane whose count reaches 0 clears its sums.
template <int K>
__device__ __forceinline__ void syn_rescale(LegState<K>& st, SynAcc<K>& ac)
{
    int need = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) need |= (st.sc[k] > 0) & (fabs(st.mu1[k]) > kScaleHi);
    if (!__any(need)) return;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const bool r = st.sc[k] > 0 && fabs(st.mu1[k]) > kScaleHi;
        st.mu1[k] = r ? st.mu1[k] * kScaleDn : st.mu1[k];
        st.mu2[k] = r ? st.mu2[k] * kScaleDn : st.mu2[k];
        st.sc[k] -= r ? 1 : 0;
        const bool join = r && st.sc[k] == 0;
        ac.er[k] = join ? 0.0 : ac.er[k];
        ac.ei[k] = join ? 0.0 : ac.ei[k];
        ac.orr[k] = join ? 0.0 : ac.orr[k];
        ac.oi[k] = join ? 0.0 : ac.oi[k];
    }
    st.all = all_live<K>(st.sc);
}

// Four live steps with coefficients c (A), r/i (B_l N_lm a_lm), parity
// pattern ODD0, !ODD0, ODD0, !ODD0.
template <int K, bool ODD0>
__device__ __forceinline__ void syn_live4(LegState<K>& st, SynAcc<K>& ac, const double (&c)[4], const double (&r)[4],
                                          const double (&i)[4])
{
    syn_live<K, ODD0>(st, ac, c[0], r[0], i[0]);
    syn_live<K, !ODD0>(st, ac, c[1], r[1], i[1]);
    syn_live<K, ODD0>(st, ac, c[2], r[2], i[2]);
    syn_live<K, !ODD0>(st, ac, c[3], r[3], i[3]);
}

template <int K, bool ODD0>
__device__ __forceinline__ void syn_load4(int j, const double* cA, const double2* ca, double (&c)[4], double (&r)[4],
                                          double (&i)[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        c[q] = cA[j + q];
        r[q] = ca[j + q].x;
        i[q] = ca[j + q].y;
    }
}

// Groups of four live steps.  Two register sets alternate: each iteration
// does eight steps (compute X while Y loads, compute Y while the next X
// loads), so loads overlap arithmetic without register copies.  The speculative load of the
// next X may read up to kChunkPad entries past the chunk; the LDS rows are
// padded for it and the values are not used.
template <int K, bool ODD0>
__device__ __forceinline__ void syn_live_run(LegState<K>& st, SynAcc<K>& ac, int& l, int cend, int cb,
                                             const double* cA, const double2* ca)
{
    if (l + 3 > cend) return;
    double xc[4], xr[4], xi[4], yc[4], yr[4], yi[4];
    syn_load4<K, ODD0>(l - cb, cA, ca, xc, xr, xi);
    for (; l + 7 <= cend; l += 8) {
        // each set's loads are issued after the first step that uses the
        // other set, so the wait before that step covers only loads issued
        // half an iteration earlier
        syn_live<K, ODD0>(st, ac, xc[0], xr[0], xi[0]);
        __builtin_amdgcn_sched_barrier(0);
        syn_load4<K, ODD0>(l + 4 - cb, cA, ca, yc, yr, yi);
        syn_live<K, !ODD0>(st, ac, xc[1], xr[1], xi[1]);
        syn_live<K, ODD0>(st, ac, xc[2], xr[2], xi[2]);
        syn_live<K, !ODD0>(st, ac, xc[3], xr[3], xi[3]);
        syn_live<K, ODD0>(st, ac, yc[0], yr[0], yi[0]);
        __builtin_amdgcn_sched_barrier(0);
        syn_load4<K, ODD0>(l + 8 - cb, cA, ca, xc, xr, xi);
        syn_live<K, !ODD0>(st, ac, yc[1], yr[1], yi[1]);
        syn_live<K, ODD0>(st, ac, yc[2], yr[2], yi[2]);
        syn_live<K, !ODD0>(st, ac, yc[3], yr[3], yi[3]);
    }
    if (l + 3 <= cend) {
        syn_live4<K, ODD0>(st, ac, xc, xr, xi);
        l += 4;
    }
}

// Rows of eight steps (parity pattern ODD0, !ODD0, ...) while some lane is
// scaled, with one rescale check per row.  Eight steps multiply |mu| by at
// most prod (A_l + 1) < 2^45 (A_{m+1} = sqrt(2m + 2) is the largest factor),
// so from 2^900 nothing overflows between checks; a lane whose scale count
// reaches 0 inside a row joins at the row's end, dropping terms below 2^-55
// of O(1).
template <int K, bool ODD0>
__device__ __forceinline__ void syn_scaled_rows(LegState<K>& st, SynAcc<K>& ac, int& l, int cend, int cb,
                                                const double* cA, const double2* ca)
{
    for (; !st.all && l + 7 <= cend; l += 8) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double2 al = ca[l + q - cb];
            if (ODD0 ^ (q & 1))
                syn_live<K, true>(st, ac, cA[l + q - cb], al.x, al.y);
            else
                syn_live<K, false>(st, ac, cA[l + q - cb], al.x, al.y);
        }
        syn_rescale<K>(st, ac);
    }
}

// Steps l = l .. cend of one staged chunk (coefficients cA[l - cb], ca[l - cb]).
// Scaled path: rows while eight steps remain, then single steps with a
// rescale check after every second step.
template <int K>
__device__ __forceinline__ void syn_chunk(LegState<K>& st, SynAcc<K>& ac, int& l, int cend, int cb, int m,
                                          const double* cA, const double2* ca)
{
    if (!st.all && l + 7 <= cend) {
        if ((l - m) & 1)
            syn_scaled_rows<K, true>(st, ac, l, cend, cb, cA, ca);
        else
            syn_scaled_rows<K, false>(st, ac, l, cend, cb, cA, ca);
    }
    while (!st.all && l <= cend) {
        for (int rep = 0; rep < 2 && l <= cend; ++rep, ++l) {
            const double2 al = ca[l - cb];
            if ((l - m) & 1)
                syn_live<K, true>(st, ac, cA[l - cb], al.x, al.y);
            else
                syn_live<K, false>(st, ac, cA[l - cb], al.x, al.y);
        }
        syn_rescale<K>(st, ac);
    }
    if (l > cend) return;
    // all lanes live: parity pattern fixed by (l - m)
    if ((l - m) & 1)
        syn_live_run<K, true>(st, ac, l, cend, cb, cA, ca);
    else
        syn_live_run<K, false>(st, ac, l, cend, cb, cA, ca);
    for (; l <= cend; ++l) {
        const int j = l - cb;
        if ((l - m) & 1)
            syn_live<K, true>(st, ac, cA[j], ca[j].x, ca[j].y);
        else
            syn_live<K, false>(st, ac, cA[j], ca[j].x, ca[j].y);
    }
}

template <int K>
__global__ __launch_bounds__(64) void k_synth(const WorkItem* __restrict__ items, const MapDesc* __restrict__ maps)
{
    constexpr int CH = kSynChunk, NC = kSynChunk / 64;
    __shared__ double sA[2][CH + kChunkPad];
    __shared__ double2 sa[2][CH + kChunkPad];
    const WorkItem it = items[blockIdx.x];
    if (it.mcnt == 0) return;
    const MapDesc& md = maps[it.map];
    const int L = md.L;
    const int npad = md.npair_pad;
    const int lane = threadIdx.x;
    const int m = it.m0;
    const int l0 = it.lfirst;
    const long base = moff(L, m) - m;
    gcdouble* __restrict__ A = (gcdouble*)(md.A + base);
    gcdouble* __restrict__ a = (gcdouble*)(md.alm + base);  // (re, im) interleaved

    LegState<K> st;
    SynAcc<K> ac;
    int p[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        p[k] = it.rb * 64 * K + k * 64 + lane;
        const long q = (long)m * npad + p[k];
        st.x[k] = md.ring_x[p[k]];
        st.mu1[k] = md.s1_cur[q];
        st.mu2[k] = md.s1_prev[q];
        st.sc[k] = md.s1_k[q];
        ac.er[k] = ac.ei[k] = ac.orr[k] = ac.oi[k] = 0.0;
    }
    st.all = all_live<K>(st.sc);
    {
        const double ar = a[2 * l0], ai = a[2 * l0 + 1];
        if ((l0 - m) & 1)
            syn_first<K, true>(st, ac, ar, ai);
        else
            syn_first<K, false>(st, ac, ar, ai);
    }
    // staged chunks of CH steps from l0+1 to L, NC coefficients per lane
    int cb = l0 + 1;
    double tA[NC];
    double2 ta[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int j = cb + 64 * c + lane;
        tA[c] = 0.0;
        ta[c] = make_double2(0.0, 0.0);
        if (j <= L) {
            tA[c] = A[j];
            ta[c] = make_double2(a[2 * j], a[2 * j + 1]);
        }
        sA[0][64 * c + lane] = tA[c];
        sa[0][64 * c + lane] = ta[c];
    }
    wave_sync();
    int buf = 0;
    int l = cb;
    while (cb <= L) {
        const int nb = cb + CH;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int j = nb + 64 * c + lane;
            if (j <= L) {
                tA[c] = A[j];
                ta[c] = make_double2(a[2 * j], a[2 * j + 1]);
            }
        }
        const int cend = min(L, nb - 1);
        syn_chunk<K>(st, ac, l, cend, cb, m, sA[buf], sa[buf]);
        if (nb <= L) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                sA[buf ^ 1][64 * c + lane] = tA[c];
                sa[buf ^ 1][64 * c + lane] = ta[c];
            }
        }
        wave_sync();
        buf ^= 1;
        cb = nb;
    }
    // lanes still scaled at L store zeros
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (p[k] < md.npair) {
            const bool z = st.sc[k] != 0;
            double2* dst = reinterpret_cast<double2*>(md.P + ((long)p[k] * (L + 1) + m) * 4);
            dst[0] = make_double2(z ? 0.0 : ac.er[k], z ? 0.0 : ac.ei[k]);
            dst[1] = make_double2(z ? 0.0 : ac.orr[k], z ? 0.0 : ac.oi[k]);
        }
    }
}

// -------------------------------------------------------------- shared-recurrence synthesis
// G maps of one geometry (same nside, lmax and ring pairs per lane, so the
// same ring tables, recurrence coefficients, work items and start states)
// in one wave: the recurrence runs once per (l, ring pair) and feeds the G
// maps' parity sums, 2K + 2KG FP64 instructions per l-step against 4KG for
// the maps one by one.  Item.map indexes the group table (rows of four map
// indices, the first G used);
// the first map supplies the geometry.  Each map's sums are formed by the
// same operations in the same order as in k_synth, so its P is the same.
// Compile-time loops over the maps of a group: every index is a constant
// when the accumulators are laid out, so they stay in registers (a runtime
// loop, even one that is unrolled later, left them in scratch).
template <class F, int... I>
__device__ __forceinline__ void for_each_map(std::integer_sequence<int, I...>, F&& f)
{
    (f(std::integral_constant<int, I>{}), ...);
}

template <int K, int G, bool ODD>
__device__ __forceinline__ void syng_step(LegState<K>& st, SynAcc<K> (&ac)[G], double Al, const double2 (&av)[G])
{
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double nu = fma(Al * st.x[k], st.mu1[k], -st.mu2[k]);
        st.mu2[k] = st.mu1[k];
        st.mu1[k] = nu;
        for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
            constexpr int g = decltype(gi)::value;
            if (ODD) {
                ac[g].orr[k] = fma(nu, av[g].x, ac[g].orr[k]);
                ac[g].oi[k] = fma(nu, av[g].y, ac[g].oi[k]);
            } else {
                ac[g].er[k] = fma(nu, av[g].x, ac[g].er[k]);
                ac[g].ei[k] = fma(nu, av[g].y, ac[g].ei[k]);
            }
        });
    }
}

template <int K, int G, bool ODD>
__device__ __forceinline__ void syng_stepj(LegState<K>& st, SynAcc<K> (&ac)[G], int j, const double* cA,
                                           const double2 (*ca)[kSynChunk / 2 + kChunkPad])
{
    double2 av[G];
    for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
        constexpr int g = decltype(gi)::value;
        av[g] = ca[g][j];
    });
    syng_step<K, G, ODD>(st, ac, cA[j], av);
}

// Rescale check: as syn_rescale, a joining lane clears every map's sums.
template <int K, int G>
__device__ __forceinline__ void syng_rescale(LegState<K>& st, SynAcc<K> (&ac)[G])
{
    int need = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) need |= (st.sc[k] > 0) & (fabs(st.mu1[k]) > kScaleHi);
    if (!__any(need)) return;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const bool r = st.sc[k] > 0 && fabs(st.mu1[k]) > kScaleHi;
        st.mu1[k] = r ? st.mu1[k] * kScaleDn : st.mu1[k];
        st.mu2[k] = r ? st.mu2[k] * kScaleDn : st.mu2[k];
        st.sc[k] -= r ? 1 : 0;
        const bool join = r && st.sc[k] == 0;
        for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
            constexpr int g = decltype(gi)::value;
            ac[g].er[k] = join ? 0.0 : ac[g].er[k];
            ac[g].ei[k] = join ? 0.0 : ac[g].ei[k];
            ac[g].orr[k] = join ? 0.0 : ac[g].orr[k];
            ac[g].oi[k] = join ? 0.0 : ac[g].oi[k];
        });
    }
    st.all = all_live<K>(st.sc);
}

// One step of either parity, the parity by value (both parities' sums take an
// fma, the other one's with zero coefficients: fma(nu, 0, s) = s for s != -0).
template <int K, int G>
__device__ __forceinline__ void syng_step_any(LegState<K>& st, SynAcc<K> (&ac)[G], bool odd, int j, const double* cA,
                                              const double2 (*ca)[kSynChunk / 2 + kChunkPad])
{
    const double Al = cA[j];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double nu = fma(Al * st.x[k], st.mu1[k], -st.mu2[k]);
        st.mu2[k] = st.mu1[k];
        st.mu1[k] = nu;
    }
    for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
        constexpr int g = decltype(gi)::value;
        const double2 a = ca[g][j];
        const double er = odd ? 0.0 : a.x, ei = odd ? 0.0 : a.y, orr = odd ? a.x : 0.0, oi = odd ? a.y : 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ac[g].er[k] = fma(st.mu1[k], er, ac[g].er[k]);
            ac[g].ei[k] = fma(st.mu1[k], ei, ac[g].ei[k]);
            ac[g].orr[k] = fma(st.mu1[k], orr, ac[g].orr[k]);
            ac[g].oi[k] = fma(st.mu1[k], oi, ac[g].oi[k]);
        }
    });
}

// Steps l..cend of one staged chunk.  Every run of steps starts at an even
// l - m (an odd first step is taken on its own, parity by value), so all
// paired steps are (even, odd) in one code path: two paths that differ only
// in which parity's sums a step updates get merged by the compiler into a
// choice of sums by address, which keeps them in scratch.  While some lane
// is scaled: rows of eight steps with a rescale check after each, and pairs
// with a check after each near the chunk's end (a lane joining inside a row
// drops terms below 2^-55 of O(1), as in k_synth); then live pairs.
template <int K, int G>
__device__ __forceinline__ void syng_chunk(LegState<K>& st, SynAcc<K> (&ac)[G], int& l, int cend, int cb, int m,
                                           const double* cA, const double2 (*ca)[kSynChunk / 2 + kChunkPad])
{
    while (l <= cend) {
        if ((l - m) & 1) {
            syng_step_any<K, G>(st, ac, true, l - cb, cA, ca);
            ++l;
            if (!st.all) syng_rescale<K, G>(st, ac);
            continue;
        }
        if (!st.all) {
            const int n = (l + 7 <= cend) ? 8 : 2;
#pragma unroll 1
            for (int q = 0; q < n && l + 1 <= cend; q += 2, l += 2) {
                syng_stepj<K, G, false>(st, ac, l - cb, cA, ca);
                syng_stepj<K, G, true>(st, ac, l + 1 - cb, cA, ca);
            }
            if (l == cend && n == 2) {
                syng_stepj<K, G, false>(st, ac, l - cb, cA, ca);
                ++l;
            }
            syng_rescale<K, G>(st, ac);
            continue;
        }
#pragma unroll 1
        for (; l + 1 <= cend; l += 2) {
            syng_stepj<K, G, false>(st, ac, l - cb, cA, ca);
            syng_stepj<K, G, true>(st, ac, l + 1 - cb, cA, ca);
        }
        if (l <= cend) {
            syng_stepj<K, G, false>(st, ac, l - cb, cA, ca);
            ++l;
        }
    }
}

template <int K, int G>
__global__ __launch_bounds__(64) void k_synth_g(const WorkItem* __restrict__ items, const MapDesc* __restrict__ maps,
                                                const int* __restrict__ gmaps)
{
    constexpr int CH = kSynChunk / 2;  // 64 steps: the G coefficient rows share the LDS budget
    __shared__ double sA[2][CH + kChunkPad];
    __shared__ double2 sa[2][G][CH + kChunkPad];
    const WorkItem it = items[blockIdx.x];
    if (it.mcnt == 0) return;
    const int* gm = gmaps + (long)it.map * 4;  // group table rows of four entries
    const MapDesc& md = maps[gm[0]];
    const int L = md.L;
    const int npad = md.npair_pad;
    const int lane = threadIdx.x;
    const int m = it.m0;
    const int l0 = it.lfirst;
    const long base = moff(L, m) - m;
    gcdouble* __restrict__ A = (gcdouble*)(md.A + base);
    gcdouble* a[G];
    for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
        constexpr int g = decltype(gi)::value;
        a[g] = (gcdouble*)(maps[gm[g]].alm + base);
    });

    LegState<K> st;
    SynAcc<K> ac[G];
    int p[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        p[k] = it.rb * 64 * K + k * 64 + lane;
        const long q = (long)m * npad + p[k];
        st.x[k] = md.ring_x[p[k]];
        st.mu1[k] = md.s1_cur[q];
        st.mu2[k] = md.s1_prev[q];
        st.sc[k] = md.s1_k[q];
        for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
            constexpr int g = decltype(gi)::value;
            ac[g].er[k] = ac[g].ei[k] = ac[g].orr[k] = ac[g].oi[k] = 0.0;
        });
    }
    st.all = all_live<K>(st.sc);
    // the item's first l, its parity by value (a branch here would be merged
    // into a choice of accumulator by address, which keeps the sums in
    // scratch); fma(mu, 0, +0) = +0 leaves the other parity's sums as they are
    const bool odd0 = ((l0 - m) & 1) != 0;
    for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
        constexpr int g = decltype(gi)::value;
        const double ar = a[g][2 * l0], ai = a[g][2 * l0 + 1];
        const double er = odd0 ? 0.0 : ar, ei = odd0 ? 0.0 : ai, orr = odd0 ? ar : 0.0, oi = odd0 ? ai : 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            ac[g].er[k] = fma(st.mu1[k], er, ac[g].er[k]);
            ac[g].ei[k] = fma(st.mu1[k], ei, ac[g].ei[k]);
            ac[g].orr[k] = fma(st.mu1[k], orr, ac[g].orr[k]);
            ac[g].oi[k] = fma(st.mu1[k], oi, ac[g].oi[k]);
        }
    });
    int cb = l0 + 1;
    double tA = 0.0;
    double2 ta[G];
    {
        const int j = cb + lane;
        tA = j <= L ? A[j] : 0.0;
        for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
            constexpr int g = decltype(gi)::value;
            ta[g] = j <= L ? make_double2(a[g][2 * j], a[g][2 * j + 1]) : make_double2(0.0, 0.0);
            sa[0][g][lane] = ta[g];
        });
        sA[0][lane] = tA;
    }
    wave_sync();
    int buf = 0;
    int l = cb;
    while (cb <= L) {
        const int nb = cb + CH;
        const int j = nb + lane;
        if (j <= L) {
            tA = A[j];
            for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
                constexpr int g = decltype(gi)::value;
                ta[g] = make_double2(a[g][2 * j], a[g][2 * j + 1]);
            });
        }
        syng_chunk<K, G>(st, ac, l, min(L, nb - 1), cb, m, sA[buf], sa[buf]);
        if (nb <= L) {
            sA[buf ^ 1][lane] = tA;
            for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
                constexpr int g = decltype(gi)::value;
                sa[buf ^ 1][g][lane] = ta[g];
            });
        }
        wave_sync();
        buf ^= 1;
        cb = nb;
    }
    // lanes still scaled at L store zeros
    for_each_map(std::make_integer_sequence<int, G>{}, [&](auto gi) {
        constexpr int g = decltype(gi)::value;
        const MapDesc& mg = maps[gm[g]];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (p[k] < md.npair) {
                const bool z = st.sc[k] != 0;
                double2* dst = reinterpret_cast<double2*>(mg.P + ((long)p[k] * (L + 1) + m) * 4);
                dst[0] = make_double2(z ? 0.0 : ac[g].er[k], z ? 0.0 : ac[g].ei[k]);
                dst[1] = make_double2(z ? 0.0 : ac[g].orr[k], z ? 0.0 : ac[g].oi[k]);
            }
        }
    });
}

// -------------------------------------------------------------- fused spin-2 synthesis
// Accumulators of one ring pair, per slot (0: steps at even offsets from the
// item's first l, 1: odd offsets), 18 doubles: T (re, im), then per field f
// (E, B) at 2 + 8 f: Sa, Sb, Sc, D3.
template <int K>
struct PolAcc {
    double v[2][K][18];
};

template <int K, int SLOT>
__device__ __forceinline__ void pol_accum(PolAcc<K>& ac, int k, double nu, double fl, double fl2,
                                          const double2 (&c)[kPolRec])
{
    double* a = ac.v[SLOT][k];
    a[0] = fma(nu, c[0].x, a[0]);
    a[1] = fma(nu, c[0].y, a[1]);
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        double* q = a + 2 + 8 * f;
        const double ur = nu * c[1 + 2 * f].x, ui = nu * c[1 + 2 * f].y;
        q[0] += ur;
        q[1] += ui;
        q[2] = fma(fl, ur, q[2]);
        q[3] = fma(fl, ui, q[3]);
        q[4] = fma(fl2, ur, q[4]);
        q[5] = fma(fl2, ui, q[5]);
        q[6] = fma(nu, c[2 + 2 * f].x, q[6]);
        q[7] = fma(nu, c[2 + 2 * f].y, q[7]);
    }
}

template <int K, int SLOT>
__device__ __forceinline__ void pol_live(LegState<K>& st, PolAcc<K>& ac, double Al, int l,
                                         const double2 (&c)[kPolRec])
{
    const double fl = l, fl2 = fl * fl;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double nu = fma(Al * st.x[k], st.mu1[k], -st.mu2[k]);
        st.mu2[k] = st.mu1[k];
        st.mu1[k] = nu;
        pol_accum<K, SLOT>(ac, k, nu, fl, fl2, c);
    }
}

// Gated step: negligible lanes contribute gate 0 (no branch, so the
// accumulators stay in place across steps).
template <int K, int SLOT>
__device__ __forceinline__ void pol_gated(const LegState<K>& st, PolAcc<K>& ac, int l, const double2 (&c)[kPolRec])
{
    const double fl = l, fl2 = fl * fl;
#pragma unroll
    for (int k = 0; k < K; ++k) pol_accum<K, SLOT>(ac, k, st.gate[k] * st.mu1[k], fl, fl2, c);
}

template <int K>
__device__ __forceinline__ void pol_load(const double2 (*cr)[kPolRec], int j, double2 (&c)[kPolRec])
{
#pragma unroll
    for (int i = 0; i < kPolRec; ++i) c[i] = cr[j][i];
}

// Steps l .. cend of one staged chunk in pairs: the accumulators are indexed
// by the step's position relative to the item's first l (slot 0 for even
// offsets), so no parity decision sits between accumulator updates.  Chunks
// start at odd offsets and hold an even number of steps except the last,
// whose extra step reads the zero coefficients staged beyond lmax.  Scaled
// pairs carry a rescale check each, then live pairs.
template <int K>
__device__ __forceinline__ void pol_chunk(LegState<K>& st, PolAcc<K>& ac, int& l, int cend, int cb,
                                          const double* cA, const double2 (*cr)[kPolRec])
{
    double2 c[kPolRec];
    while (!st.all && l <= cend) {
        pol_load<K>(cr, l - cb, c);
        st.advance(cA[l - cb]);
        pol_gated<K, 1>(st, ac, l, c);
        pol_load<K>(cr, l + 1 - cb, c);
        st.advance(cA[l + 1 - cb]);
        pol_gated<K, 0>(st, ac, l + 1, c);
        l += 2;
        if (rescale<K>(st.mu1, st.mu2, st.sc)) st.refresh();
    }
    for (; l <= cend; l += 2) {
        pol_load<K>(cr, l - cb, c);
        pol_live<K, 1>(st, ac, cA[l - cb], l, c);
        pol_load<K>(cr, l + 1 - cb, c);
        pol_live<K, 0>(st, ac, cA[l + 1 - cb], l + 1, c);
    }
}

// One (m, ring block) item of the synthesis work list: T, Q, U parity sums of
// its ring pairs into pt, pq, pu ([pair][m] (E.re, E.im, O.re, O.im), the
// form the ring stage reads; Q and U as ((north + south)/2, (north - south)/2)).
template <int K>
__global__ __launch_bounds__(64) void k_synth_pol(const WorkItem* __restrict__ items, const MapDesc* __restrict__ maps,
                                                  const double2* __restrict__ rec, double* __restrict__ pt,
                                                  double* __restrict__ pq, double* __restrict__ pu)
{
    __shared__ double sA[2][kChunk + kChunkPad];
    __shared__ double2 sr[2][kChunk + kChunkPad][kPolRec];
    const WorkItem it = items[blockIdx.x];
    if (it.mcnt == 0) return;
    const MapDesc& md = maps[it.map];
    const int L = md.L;
    const int npad = md.npair_pad;
    const int lane = threadIdx.x;
    const int m = it.m0;
    const int l0 = it.lfirst;
    const long base = moff(L, m) - m;
    gcdouble* __restrict__ A = (gcdouble*)(md.A + base);
    gcdouble* __restrict__ R = (gcdouble*)(rec + base * kPolRec);  // 2 kPolRec doubles per l

    LegState<K> st;
    PolAcc<K> ac;
    int p[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        p[k] = it.rb * 64 * K + k * 64 + lane;
        const long q = (long)m * npad + p[k];
        st.x[k] = md.ring_x[p[k]];
        st.mu1[k] = md.s1_cur[q];
        st.mu2[k] = md.s1_prev[q];
        st.sc[k] = md.s1_k[q];
#pragma unroll
        for (int i = 0; i < 18; ++i) ac.v[0][k][i] = ac.v[1][k][i] = 0.0;
    }
    st.refresh();
    {
        double2 c[kPolRec];
#pragma unroll
        for (int i = 0; i < kPolRec; ++i)
            c[i] = make_double2(R[(long)l0 * 2 * kPolRec + 2 * i], R[(long)l0 * 2 * kPolRec + 2 * i + 1]);
        pol_gated<K, 0>(st, ac, l0, c);
    }
    // staged chunks of steps l0+1 .. L
    int cb = l0 + 1;
    double tA = 0.0;
    double2 tc[kPolRec];
#pragma unroll
    for (int i = 0; i < kPolRec; ++i) tc[i] = make_double2(0.0, 0.0);
    if (cb + lane <= L) {
        tA = A[cb + lane];
#pragma unroll
        for (int i = 0; i < kPolRec; ++i)
            tc[i] = make_double2(R[(long)(cb + lane) * 2 * kPolRec + 2 * i], R[(long)(cb + lane) * 2 * kPolRec + 2 * i + 1]);
    }
    sA[0][lane] = tA;
#pragma unroll
    for (int i = 0; i < kPolRec; ++i) sr[0][lane][i] = tc[i];
    wave_sync();
    int buf = 0;
    int l = cb;
    while (cb <= L) {
        const int nb = cb + kChunk;
        // entries beyond lmax stage zeros: the last chunk's extra step reads one
        tA = 0.0;
#pragma unroll
        for (int i = 0; i < kPolRec; ++i) tc[i] = make_double2(0.0, 0.0);
        if (nb + lane <= L) {
            tA = A[nb + lane];
#pragma unroll
            for (int i = 0; i < kPolRec; ++i)
                tc[i] = make_double2(R[(long)(nb + lane) * 2 * kPolRec + 2 * i],
                                     R[(long)(nb + lane) * 2 * kPolRec + 2 * i + 1]);
        }
        pol_chunk<K>(st, ac, l, min(L, nb - 1), cb, sA[buf], sr[buf]);
        if (nb <= L) {
            sA[buf ^ 1][lane] = tA;
#pragma unroll
            for (int i = 0; i < kPolRec; ++i) sr[buf ^ 1][lane][i] = tc[i];
        }
        wave_sync();
        buf ^= 1;
        cb = nb;
    }
    // epilogue: per ring pair, north (E + O at c) and south (E - O at -c)
    const double fm = m, fm2 = fm * fm;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (p[k] >= md.npair) continue;
        const long o = ((long)p[k] * (L + 1) + m) * 4;
        const bool swap = ((l0 - m) & 1) != 0;  // slot 0 holds the parity of l0 - m
        double e[18], d[18];
#pragma unroll
        for (int i = 0; i < 18; ++i) {
            e[i] = swap ? ac.v[1][k][i] : ac.v[0][k][i];
            d[i] = swap ? ac.v[0][k][i] : ac.v[1][k][i];
        }
        *reinterpret_cast<double4*>(pt + o) = make_double4(e[0], e[1], d[0], d[1]);
        const double c = st.x[k], is2 = 1.0 / md.ring_s2[p[k]];
        double W[2][2][2], X[2][2][2];  // [field][north/south][re/im]
#pragma unroll
        for (int f = 0; f < 2; ++f) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const double sg = h ? -1.0 : 1.0;  // south: E - O, c -> -c
                const double ch = sg * c;
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int b = 2 + 8 * f + r;
                    const double Sa = e[b] + sg * d[b], Sb = e[b + 2] + sg * d[b + 2];
                    const double Sc = e[b + 4] + sg * d[b + 4], D3 = e[b + 6] + sg * d[b + 6];
                    const double D1 = Sb - fm2 * Sa, D2 = 0.5 * (Sc - Sb), D4 = Sb - Sa;
                    W[f][h][r] = -is2 * D1 - D2 + ch * is2 * D3;
                    X[f][h][r] = fm * is2 * (D3 - ch * D4);
                }
            }
        }
        // Q = -(W_E + i X_B), U = -(W_B - i X_E)
        double Q[2][2], U[2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            Q[h][0] = -(W[0][h][0] - X[1][h][1]);
            Q[h][1] = -(W[0][h][1] + X[1][h][0]);
            U[h][0] = -(W[1][h][0] + X[0][h][1]);
            U[h][1] = -(W[1][h][1] - X[0][h][0]);
        }
        *reinterpret_cast<double4*>(pq + o) = make_double4(0.5 * (Q[0][0] + Q[1][0]), 0.5 * (Q[0][1] + Q[1][1]),
                                                           0.5 * (Q[0][0] - Q[1][0]), 0.5 * (Q[0][1] - Q[1][1]));
        *reinterpret_cast<double4*>(pu + o) = make_double4(0.5 * (U[0][0] + U[1][0]), 0.5 * (U[0][1] + U[1][1]),
                                                           0.5 * (U[0][0] - U[1][0]), 0.5 * (U[0][1] - U[1][1]));
    }
}

// -------------------------------------------------------------- adjoint
// Per l-step each lane forms sum_k mu_l(p_k) Psi^{parity}(p_k); every
// kRedChunk steps the wave reduces the 64 lanes (register folds and an LDS
// tile) and writes the ring-block partial u_lm (re, im) with one coalesced
// store.  Steps outside whole rows (the item's last partial row) go through a
// per-lane LDS buffer reduced when it holds kRedChunk steps or at the end.

// A lane's ring-stage inputs Psi_m (even and odd parts) of its K ring pairs.
// Lanes whose scale count is > 0 hold zeros, so that the same four-operation
// step serves the scaled and the live path: their scaled mu (|mu| < 2^945
// between rescale checks) times 0 contributes exactly 0.  A lane that turns
// live loads its inputs at the rescale check where its count reaches 0.
template <int K>
struct AdjIn {
    double per[K], pei[K], por[K], poi[K];
};

// Where the inputs of lane `lane`'s ring pair pbase + 64 k + lane live:
// Q + p * row + 4 m (row = 4 (L + 1) doubles per ring pair); pairs past npair
// (padding) have none.
struct AdjSrc {
    const double* Q;
    long row;
    int m, pbase, npair, lane;
};

template <int K>
__device__ __forceinline__ void adj_load_in(AdjIn<K>& in, int k, const AdjSrc& s)
{
    const int p = s.pbase + k * 64 + s.lane;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    if (p < s.npair) {
        gcdouble* src = (gcdouble*)(s.Q + (long)p * s.row + 4 * (long)s.m);
        v0 = src[0];
        v1 = src[1];
        v2 = src[2];
        v3 = src[3];
    }
    // one unconditional store per field: stores in both arms of a branch
    // would be merged through a pointer and keep the inputs in memory
    in.per[k] = v0;
    in.pei[k] = v1;
    in.por[k] = v2;
    in.poi[k] = v3;
}

// Rescale check of the adjoint's scaled path: lanes above 2^900 drop by
// 2^-1000; a lane whose count reaches 0 loads its inputs.
template <int K>
__device__ __forceinline__ void adj_rescale(LegState<K>& st, AdjIn<K>& in, const AdjSrc& s)
{
    int need = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) need |= (st.sc[k] > 0) & (fabs(st.mu1[k]) > kScaleHi);
    if (!__any(need)) return;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (st.sc[k] > 0 && fabs(st.mu1[k]) > kScaleHi) {
            st.mu1[k] *= kScaleDn;
            st.mu2[k] *= kScaleDn;
            st.sc[k] -= 1;
            if (st.sc[k] == 0) adj_load_in<K>(in, k, s);
        }
    }
    st.all = all_live<K>(st.sc);
}

// Psi (re, im) of parity o of ring pair k.  All four inputs are read and the
// values selected: selecting array elements, or merging the even and odd
// variants of a step, would keep the inputs in addressable memory.
template <int K>
__device__ __forceinline__ void adj_psi(const AdjIn<K>& in, int k, bool o, double& pr, double& pi)
{
    const double er = in.per[k], ei = in.pei[k], orr = in.por[k], oi = in.poi[k];
    pr = o ? orr : er;
    pi = o ? oi : ei;
}


// Lanes 0..31 get a summed over lanes (j, j + 32), lanes 32..63 get b summed
// likewise (v_permlane32_swap exchanges lanes 32..63 of its first operand
// with lanes 0..31 of its second).
__device__ __forceinline__ double swap32_add(double a, double b)
{
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b),
                                                     false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b),
                                                     false, false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

// Rows of 16 lanes: rows 0 and 2 get a summed over row pairs (0,1) and (2,3),
// rows 1 and 3 get b summed likewise (v_permlane16_swap exchanges the odd rows
// of its first operand with the even rows of its second).
__device__ __forceinline__ double swap16_add(double a, double b)
{
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b),
                                                     false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b),
                                                     false, false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

// s plus s rotated by N lanes within each row of 16 (DPP row_ror:N); for
// N = 8 that is s plus s of lane j ^ 8.
template <int N>
__device__ __forceinline__ double ror_add(double s)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(s), 0x120 + N, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(s), 0x120 + N, 0xF, 0xF, false);
    return s + __hiloint2double(hi, lo);
}

// the row tile (records 33 q + u, q < 8, u < 32) fits in the lane rows
static_assert(2 * (33 * 7 + 32) <= 64 * kRedStride, "row tile exceeds the reduction buffer");

struct AdjRed {
    double* red;   // [64 * kRedStride]
    double* mine;  // this lane's row
    double* dst;   // output row, dst[2l], dst[2l+1] (global memory)
    int cnt, lbase, lane;
    __device__ __forceinline__ void flush()
    {
        wave_sync();
        const int c = lane & 15;
        const int q = lane >> 4;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) s += red[(16 * q + i) * kRedStride + c];
        s += __shfl_xor(s, 16);
        s += __shfl_xor(s, 32);
        if (lane < 2 * cnt) ((gdouble*)dst)[2 * lbase + lane] = s;
        wave_sync();
        lbase += cnt;
        cnt = 0;
    }
    __device__ __forceinline__ void push(double sr, double si)
    {
        mine[2 * cnt] = sr;
        mine[2 * cnt + 1] = si;
        if (++cnt == kRedChunk) flush();
    }
    // Rows of kRedChunk = 8 steps.  Steps 2p and 2p + 1 are folded as soon as
    // both exist, once over lane pairs (j, j + 32) with permlane32 swaps
    // (fold_pair): z[2p + c] then holds step 2p in lanes 0..31 and step 2p + 1
    // in lanes 32..63 (c = 0 real, 1 imaginary).  tile_store puts lane L's four
    // (re, im) records as records (q, L & 31), q = 2p + (L >> 5), at index
    // 33 q + (L & 31); tile_load gives lane j the records (j & 7, 4 (j >> 3)
    // + r), r < 4 (the skew of one record per q puts the eight lanes of a read
    // on distinct banks); finish_row sums them and completes the sum over the
    // eight lanes that share j & 7 (distances 32, 16, 8): the real part ends in
    // lanes 0..31, the imaginary part in lanes 32..63.  A permlane swap costs
    // about two FP64 issue slots, so the tile carries the second fold: 20 swaps
    // per row instead of 28, for 4 KB of LDS traffic each way.
    __device__ __forceinline__ static void fold_pair(double are, double aim, double bre, double bim, double& z0,
                                                     double& z1)
    {
        z0 = swap32_add(are, bre);
        z1 = swap32_add(aim, bim);
    }
    __device__ __forceinline__ void tile_store(const double (&z)[8])
    {
        double2* t = reinterpret_cast<double2*>(red) + 33 * (lane >> 5) + (lane & 31);
#pragma unroll
        for (int p = 0; p < 4; ++p) t[66 * p] = make_double2(z[2 * p], z[2 * p + 1]);
        wave_sync();
    }
    __device__ __forceinline__ void tile_load(double2 (&t4)[4]) const
    {
        const double2* t = reinterpret_cast<const double2*>(red) + 33 * (lane & 7) + 4 * (lane >> 3);
#pragma unroll
        for (int r = 0; r < 4; ++r) t4[r] = t[r];
    }
    __device__ __forceinline__ void finish_row(const double2 (&t4)[4])
    {
        double re = (t4[0].x + t4[1].x) + (t4[2].x + t4[3].x);
        double im = (t4[0].y + t4[1].y) + (t4[2].y + t4[3].y);
        double s = swap32_add(re, im);
        s = swap16_add(s, s);
        s = ror_add<8>(s);
        // four lanes hold each value and store it (no branch, so the store
        // stays in the loop's block)
        gdouble* g = (gdouble*)dst;
        g[2 * lbase + 2 * (lane & 7) + (lane >> 5)] = s;
        lbase += kRedChunk;
    }
    // One l on its own (the item's first): the whole-wave sum of (sr, si),
    // the real part in lanes 0..31 and the imaginary part in lanes 32..63, all
    // in registers, stored at lbase.  Keeping it out of the push buffer lets
    // the staged chunks (64 steps from the next l) fall on row boundaries.
    __device__ __forceinline__ void store_one(double sr, double si)
    {
        double s = swap32_add(sr, si);
        s = swap16_add(s, s);
        s = ror_add<8>(s);
        s = ror_add<4>(s);
        s = ror_add<2>(s);
        s = ror_add<1>(s);
        gdouble* g = (gdouble*)dst;
        g[2 * lbase + (lane >> 5)] = s;
        lbase += 1;
    }
};

template <int K, bool ODD>
__device__ __forceinline__ void adj_live(LegState<K>& st, const AdjIn<K>& in, AdjRed& rd, double Al)
{
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double nu = fma(Al * st.x[k], st.mu1[k], -st.mu2[k]);
        st.mu2[k] = st.mu1[k];
        st.mu1[k] = nu;
        double pr, pi;
        adj_psi<K>(in, k, ODD, pr, pi);
        sr = fma(nu, pr, sr);
        si = fma(nu, pi, si);
    }
    rd.push(sr, si);
}

// The item's first l (the recurrence state is already at it).
template <int K, bool ODD>
__device__ __forceinline__ void adj_first(const LegState<K>& st, const AdjIn<K>& in, AdjRed& rd)
{
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double pr, pi;
        adj_psi<K>(in, k, ODD, pr, pi);
        sr = fma(st.mu1[k], pr, sr);
        si = fma(st.mu1[k], pi, si);
    }
    rd.store_one(sr, si);
}

// One live step returning its (real, imaginary) lane sums; o = parity of l - m.
template <int K>
__device__ __forceinline__ void adj_step_v(LegState<K>& st, const AdjIn<K>& in, bool o, double Al, double& sr,
                                           double& si)
{
    sr = 0.0;
    si = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double nu = fma(Al * st.x[k], st.mu1[k], -st.mu2[k]);
        st.mu2[k] = st.mu1[k];
        st.mu1[k] = nu;
        double pr, pi;
        adj_psi<K>(in, k, o, pr, pi);
        sr = fma(nu, pr, sr);
        si = fma(nu, pi, si);
    }
}

template <int K>
__device__ __forceinline__ void adj_step(LegState<K>& st, const AdjIn<K>& in, AdjRed& rd, int l, int m, double Al)
{
    if ((l - m) & 1)
        adj_live<K, true>(st, in, rd, Al);
    else
        adj_live<K, false>(st, in, rd, Al);
}

// Steps l..cend of one staged chunk while some lane is still scaled: single
// steps up to a row boundary (rescale check after each), then rows of
// kRedChunk = 8 steps reduced like the live rows with one rescale check per
// row, then (at the item's end) the rest one step at a time.  It returns once
// every lane is live; the caller finishes the chunk with live rows.  Eight steps multiply |mu| by at most prod (A_l + 1) < 2^45
// (A_{m+1} = sqrt(2m + 2) is the largest factor), so from 2^900 no value
// overflows between checks; a lane whose scale count reaches 0 inside a row
// joins at the row's end, dropping terms below 2^-55 of O(1).
template <int K, bool ODD0>
__device__ __forceinline__ void adj_scaled_rows(LegState<K>& st, AdjIn<K>& in, AdjRed& rd, int& l, int cend, int cb,
                                                const double* cA, const AdjSrc& s)
{
    for (; !st.all && l + 7 <= cend; l += 8) {
        double z[8];
        double2 t4[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            double are, aim, bre, bim;
            adj_step_v<K>(st, in, ODD0, cA[l + 2 * p - cb], are, aim);
            adj_step_v<K>(st, in, !ODD0, cA[l + 2 * p + 1 - cb], bre, bim);
            AdjRed::fold_pair(are, aim, bre, bim, z[2 * p], z[2 * p + 1]);
        }
        rd.tile_store(z);
        rd.tile_load(t4);
        rd.finish_row(t4);
        wave_sync();
        adj_rescale<K>(st, in, s);
    }
}

template <int K>
__device__ __forceinline__ void adj_chunk_scaled(LegState<K>& st, AdjIn<K>& in, AdjRed& rd, int& l, int cend,
                                                 int cb, int m, const double* cA, const AdjSrc& s)
{
    while (!st.all && rd.cnt != 0 && l <= cend) {
        adj_step<K>(st, in, rd, l, m, cA[l - cb]);
        ++l;
        adj_rescale<K>(st, in, s);
    }
    if (!st.all && l + 7 <= cend) {
        if ((l - m) & 1)
            adj_scaled_rows<K, true>(st, in, rd, l, cend, cb, cA, s);
        else
            adj_scaled_rows<K, false>(st, in, rd, l, cend, cb, cA, s);
    }
    while (!st.all && l <= cend) {
        adj_step<K>(st, in, rd, l, m, cA[l - cb]);
        ++l;
        adj_rescale<K>(st, in, s);
    }
}

// Steps l..cend of one staged chunk with every lane live: single steps up to
// the next reduction row boundary, then rows of kRedChunk = 8 steps with two
// 4-coefficient register sets alternating (the speculative load of the next
// set may read up to 4 entries past the chunk, into the row padding).  A
// row's tile is stored at its end and read back after the next row's first
// step, and the sum finishes halfway through that row, so the LDS round trip
// overlaps FP64 work.
template <int K>
__device__ __forceinline__ void adj_chunk_live(LegState<K>& st, const AdjIn<K>& in, AdjRed& rd, int& l, int cend,
                                               int cb, int m, const double* cA)
{
    for (; rd.cnt != 0 && l <= cend; ++l) adj_step<K>(st, in, rd, l, m, cA[l - cb]);
    if (l + 7 <= cend) {
        const bool odd = ((l - m) & 1) != 0;
        double x[4], y[4];
        double2 t4[4];
        bool pend = false;  // the tile holds a row not yet summed
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = cA[l - cb + q];
        for (; l + 7 <= cend; l += 8) {
            double z[8], are, aim, bre, bim;
            // each set's loads are issued after the first step that uses the
            // other set, so the wait before that step covers only LDS traffic
            // issued half an iteration earlier; the previous row's tile read
            // rides along with the Y loads
            adj_step_v<K>(st, in, odd, x[0], are, aim);
            __builtin_amdgcn_sched_barrier(0);
            if (pend) rd.tile_load(t4);
#pragma unroll
            for (int q = 0; q < 4; ++q) y[q] = cA[l - cb + 4 + q];
            adj_step_v<K>(st, in, !odd, x[1], bre, bim);
            AdjRed::fold_pair(are, aim, bre, bim, z[0], z[1]);
            adj_step_v<K>(st, in, odd, x[2], are, aim);
            adj_step_v<K>(st, in, !odd, x[3], bre, bim);
            AdjRed::fold_pair(are, aim, bre, bim, z[2], z[3]);
            if (pend) rd.finish_row(t4);
            adj_step_v<K>(st, in, odd, y[0], are, aim);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 4; ++q) x[q] = cA[l - cb + 8 + q];
            adj_step_v<K>(st, in, !odd, y[1], bre, bim);
            AdjRed::fold_pair(are, aim, bre, bim, z[4], z[5]);
            adj_step_v<K>(st, in, odd, y[2], are, aim);
            adj_step_v<K>(st, in, !odd, y[3], bre, bim);
            AdjRed::fold_pair(are, aim, bre, bim, z[6], z[7]);
            rd.tile_store(z);
            pend = true;
        }
        rd.tile_load(t4);
        rd.finish_row(t4);
        wave_sync();
    }
    for (; l <= cend; ++l) adj_step<K>(st, in, rd, l, m, cA[l - cb]);
}

// K = 4 (the adjoint's ring pairs per lane from nside 512 up) runs three
// waves per SIMD: it needs 166 VGPRs (168 fit three waves); a four-wave bound
// (128) spills around the step loops and measured slower.
template <int K>
__global__ __launch_bounds__(64, K == 4 ? 3 : 1) void k_adj(const WorkItem* __restrict__ items, const MapDesc* __restrict__ maps)
{
    __shared__ __align__(16) double red[64 * kRedStride];
    __shared__ double sA[2][kAdjChunk + kChunkPad];
    const WorkItem it = items[blockIdx.x];
    if (it.mcnt == 0) return;
    const MapDesc& md = maps[it.map];
    const int L = md.L;
    const int lend = it.lend;  // last l of this segment
    const int lane = threadIdx.x;
    const int m = it.m0;
    const int l0 = it.lfirst;
    const long base = moff(L, m) - m;
    gcdouble* __restrict__ A = (gcdouble*)(md.A + base);

    LegState<K> st;
    AdjIn<K> in;
    AdjSrc src;
    src.Q = md.Q;
    src.row = 4 * (long)(L + 1);
    src.m = m;
    src.pbase = it.rb * 64 * K;
    src.npair = md.npair;
    src.lane = lane;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = src.pbase + k * 64 + lane;
        const long q = (long)it.sidx * 64 * K + k * 64 + lane;  // the segment's start state
        st.x[k] = md.ring_x[p];
        st.mu1[k] = md.s5_cur[q];
        st.mu2[k] = md.s5_prev[q];
        st.sc[k] = md.s5_k[q];
        adj_load_in<K>(in, k, src);
    }
    // lanes still scaled start from zero inputs
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const bool z = st.sc[k] != 0;
        in.per[k] = z ? 0.0 : in.per[k];
        in.pei[k] = z ? 0.0 : in.pei[k];
        in.por[k] = z ? 0.0 : in.por[k];
        in.poi[k] = z ? 0.0 : in.poi[k];
    }
    st.all = all_live<K>(st.sc);
    AdjRed rd;
    rd.red = red;
    rd.mine = red + lane * kRedStride;
    rd.dst = md.upart + 2 * (long)it.rb * md.nalm + 2 * base;
    rd.cnt = 0;
    rd.lbase = l0;
    rd.lane = lane;
    if ((l0 - m) & 1)
        adj_first<K, true>(st, in, rd);
    else
        adj_first<K, false>(st, in, rd);

    // staged chunks of kAdjChunk = 128 steps, two coefficients per lane (each
    // chunk ends with one unpipelined row reduction, so longer chunks drain
    // the pipeline less often)
    int cb = l0 + 1;
    double tA0 = (cb + lane <= lend) ? A[cb + lane] : 0.0;
    double tA1 = (cb + 64 + lane <= lend) ? A[cb + 64 + lane] : 0.0;
    sA[0][lane] = tA0;
    sA[0][64 + lane] = tA1;
    wave_sync();
    int buf = 0;
    int l = cb;
    // two loops so that the scaled-path state is dead in the second
    while (cb <= lend && !st.all) {
        const int nb = cb + kAdjChunk;
        if (nb + lane <= lend) tA0 = A[nb + lane];
        if (nb + 64 + lane <= lend) tA1 = A[nb + 64 + lane];
        adj_chunk_scaled<K>(st, in, rd, l, min(lend, nb - 1), cb, m, sA[buf], src);
        // every lane turned live inside the chunk: its rest runs as live rows
        if (st.all && l <= min(lend, nb - 1)) adj_chunk_live<K>(st, in, rd, l, min(lend, nb - 1), cb, m, sA[buf]);
        if (nb <= lend) {
            sA[buf ^ 1][lane] = tA0;
            sA[buf ^ 1][64 + lane] = tA1;
        }
        wave_sync();
        buf ^= 1;
        cb = nb;
    }
    while (cb <= lend) {
        const int nb = cb + kAdjChunk;
        if (nb + lane <= lend) tA0 = A[nb + lane];
        if (nb + 64 + lane <= lend) tA1 = A[nb + 64 + lane];
        adj_chunk_live<K>(st, in, rd, l, min(lend, nb - 1), cb, m, sA[buf]);
        if (nb <= lend) {
            sA[buf ^ 1][lane] = tA0;
            sA[buf ^ 1][64 + lane] = tA1;
        }
        wave_sync();
        buf ^= 1;
        cb = nb;
    }
    if (rd.cnt > 0) rd.flush();
}

// -------------------------------------------------------------- shared adjoint on the matrix cores
//
// Maps of one geometry share lambda_lm(theta).  For a group of up to kAdjG = 8
// such maps (a smaller group leaves columns empty) the adjoint of one
// (m, block of 64 ring pairs) is a matrix product,
//   U[l][c] = sum_p Lambda[l][p] Psi[p][c],  c = (map, re / im),
// the ring-pair sum of 16 l-rows by 16 columns taking one FP64 matrix
// instruction per 4 ring pairs (v_mfma_f64_16x16x4f64): the sum over ring
// pairs happens inside the instruction (no cross-lane reduction) and the
// recurrence runs once for the 8 maps.  Lane layout of that instruction:
// A[i][k] comes from lane i + 16 k, B[k][j] from lane j + 16 k, and D[i][j]
// lands in lane j + 16 (i % 4), register i / 4.
// Each lane runs the recurrence of one ring pair and writes lambda_l (times a
// gate that is 0 while the lane is scaled) into an LDS tile of kMmaSteps
// l-steps; the 16 rows of one parity (l - m even pairs with Psi_E, odd with
// Psi_O) are the A operands, and the Psi of the block's 64 ring pairs stay in
// registers as B operands for the whole item.  A workgroup is kMmaWaves
// waves on consecutive 64-pair blocks of one m, each starting at its own
// first live l on a common grid of tiles; their products are added in LDS
// in wave order, so one partial per kMmaWaves blocks goes to memory.
constexpr int kMmaSteps = 32;
constexpr int kMmaRow = 65;  // doubles per tile row: the 16 rows of one parity start 4 banks apart

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Rescale check of a scaled lane (every 8 steps), without branches.
__device__ __forceinline__ void mma_rescale(double& mu1, double& mu2, int& sc, double& g, bool& all)
{
    const bool r = sc > 0 && fabs(mu1) > kScaleHi;
    mu1 = r ? mu1 * kScaleDn : mu1;
    mu2 = r ? mu2 * kScaleDn : mu2;
    sc -= r ? 1 : 0;
    g = sc == 0 ? 1.0 : 0.0;
    all = __all(sc == 0);
}

// kMmaSteps recurrence steps of one ring pair into tile column `lane`: row s
// holds lambda at l = c0 + s.  SCALED: times the gate, rescale checks every 8
// steps.  Coefficients are read from LDS eight steps ahead of their use.
template <bool SCALED, int RS = kMmaRow>
__device__ __forceinline__ void mma_fill(double* __restrict__ T, const double* __restrict__ cA, int lane, double x,
                                         double& mu1, double& mu2, int& sc, double& g, bool& all)
{
    double c[8], cn[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) c[u] = cA[u];
#pragma unroll
    for (int s8 = 0; s8 < kMmaSteps; s8 += 8) {
        if (s8 + 8 < kMmaSteps) {
#pragma unroll
            for (int u = 0; u < 8; ++u) cn[u] = cA[s8 + 8 + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double nu = fma(c[u] * x, mu1, -mu2);
            mu2 = mu1;
            mu1 = nu;
            T[(s8 + u) * RS + lane] = SCALED ? mu1 * g : mu1;
        }
        if (SCALED) mma_rescale(mu1, mu2, sc, g, all);
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] = cn[u];
    }
}

// The tile holding a wave's first live l at row r0: rows below it are zero,
// row r0 is the start state itself, later rows advance.  (Rescale checks
// before r0 find nothing: start states are at most 2^900.)
template <int RS = kMmaRow>
__device__ __forceinline__ void mma_fill_first(double* __restrict__ T, const double* __restrict__ cA, int lane,
                                               int r0, double x, double& mu1, double& mu2, int& sc, double& g,
                                               bool& all)
{
#pragma unroll
    for (int s = 0; s < kMmaSteps; ++s) {
        const bool adv = s > r0;
        const double nu = fma(cA[s] * x, mu1, -mu2);
        mu2 = adv ? mu1 : mu2;
        mu1 = adv ? nu : mu1;
        T[s * RS + lane] = s >= r0 ? mu1 * g : 0.0;
        if ((s & 7) == 7) mma_rescale(mu1, mu2, sc, g, all);
    }
}

__global__ __launch_bounds__(64 * kMmaWaves, 2) void k_adj_mma(const WorkItem* __restrict__ items,
                                                               const MmaSub* __restrict__ subs,
                                                               const MapDesc* __restrict__ maps,
                                                               const int* __restrict__ groups)
{
    __shared__ double T[kMmaWaves][kMmaSteps * kMmaRow];
    __shared__ double sA[2][kMmaSteps];
    const WorkItem it = items[blockIdx.x];
    if (it.mcnt == 0) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int* __restrict__ gm = groups + kAdjG * it.map;
    const MapDesc& md = maps[gm[0]];
    const int L = md.L;
    const int lend = it.lend;
    const int m = it.m0;
    const int l0 = it.lfirst;  // the earliest first live l of the workgroup's blocks
    const long base = moff(L, m) - m;
    gcdouble* __restrict__ A = (gcdouble*)(md.A + base);
    const int lfw = subs[it.sidx].lf[w];  // this wave's first live l, -1: no live item
    const int rb = it.rb * kMmaWaves + w;  // this wave's 64-pair block
    double* __restrict__ Tw = T[w];

    // the recurrence of ring pair rb * 64 + lane, from the wave's start state
    // (a wave without a live item may lie past the ring tables: npair_pad
    // need not be a multiple of kMmaWaves * 64)
    double x = 0.0, mu1 = 0.0, mu2 = 0.0;
    int sc = 1;
    if (lfw >= 0) {
        x = md.ring_x[rb * 64 + lane];
        const long q = (long)subs[it.sidx].sidx[w] * 64 + lane;
        mu1 = md.s5_cur[q];
        mu2 = md.s5_prev[q];
        sc = md.s5_k[q];
    }
    double g = sc == 0 ? 1.0 : 0.0;
    bool all = __all(sc == 0);

    // B operands: B[k][j] = Psi of ring pair rb * 64 + 4 b + k, column j =
    // (map j / 2 of the group, re / im j % 2); a group's empty columns
    // (entry -1) hold zeros and are not stored
    const int j = lane & 15, kq = lane >> 4;
    const int mapj = gm[j >> 1];
    const MapDesc& mj = maps[mapj >= 0 ? mapj : gm[0]];
    double bE[16], bO[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) bE[b] = bO[b] = 0.0;
    if (lfw >= 0 && mapj >= 0) {
        const long row = 4 * (long)(L + 1);
        // padding pairs (past npair) read the last ring pair's row and are
        // zeroed afterwards, so that all 32 loads are issued together
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const int pp = min(rb * 64 + 4 * b + kq, md.npair - 1);
            gcdouble* src = (gcdouble*)(mj.Q + (long)pp * row + 4 * (long)m + (j & 1));
            bE[b] = src[0];
            bO[b] = src[2];
        }
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const bool pad = rb * 64 + 4 * b + kq >= md.npair;
            bE[b] = pad ? 0.0 : bE[b];
            bO[b] = pad ? 0.0 : bO[b];
        }
    }
    // this workgroup's partial (kMmaWaves blocks), written by wave 0
    gdouble* __restrict__ dst = (gdouble*)(mj.upart + 2 * ((long)it.rb * md.nalm + base) + (j & 1));

    // coefficients of the current tile in sA[buf], of the next one in tA
    const int t = threadIdx.x;
    double tA = 0.0;
    if (t < kMmaSteps) {
        sA[0][t] = (l0 + t <= lend) ? A[l0 + t] : 0.0;
        tA = (l0 + kMmaSteps + t <= lend) ? A[l0 + kMmaSteps + t] : 0.0;
    }
    __syncthreads();
    int buf = 0;
    for (int c0 = l0; c0 <= lend; c0 += kMmaSteps) {
        const int n2 = c0 + 2 * kMmaSteps;
        // tile rows past lend hold finite values that no stored output uses
        const bool active = lfw >= 0 && lfw < c0 + kMmaSteps;
        if (active) {
            if (lfw >= c0)
                mma_fill_first(Tw, sA[buf], lane, lfw - c0, x, mu1, mu2, sc, g, all);
            else if (all)
                mma_fill<false>(Tw, sA[buf], lane, x, mu1, mu2, sc, g, all);
            else
                mma_fill<true>(Tw, sA[buf], lane, x, mu1, mu2, sc, g, all);
        }
        wave_sync();
        const int sE = (c0 - m) & 1;  // first tile row with l - m even
        dbl4 dE = {0.0, 0.0, 0.0, 0.0}, dO = {0.0, 0.0, 0.0, 0.0};
        if (active) {
            // the ring-pair sums on the matrix cores, 16 l-rows of each parity
            const double* tE = Tw + (sE + 2 * j) * kMmaRow + kq;
            const double* tO = Tw + ((sE ^ 1) + 2 * j) * kMmaRow + kq;
            double aE[16], aO[16];
#pragma unroll
            for (int b = 0; b < 16; ++b) {
                aE[b] = tE[4 * b];
                aO[b] = tO[4 * b];
            }
#pragma unroll
            for (int b = 0; b < 16; ++b) {
                dE = __builtin_amdgcn_mfma_f64_16x16x4f64(aE[b], bE[b], dE, 0, 0, 0);
                dO = __builtin_amdgcn_mfma_f64_16x16x4f64(aO[b], bO[b], dO, 0, 0, 0);
            }
        }
        // waves 1.. hand their sums to wave 0 through their own tile (its
        // A operands are in registers; LDS runs a wave's operations in order)
        if (w > 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Tw[r * 64 + lane] = dE[r];
                Tw[(4 + r) * 64 + lane] = dO[r];
            }
        }
        __syncthreads();
        if (w == 0) {
#pragma unroll
            for (int v = 1; v < kMmaWaves; ++v) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    dE[r] += T[v][r * 64 + lane];
                    dO[r] += T[v][(4 + r) * 64 + lane];
                }
            }
            // row i = kq + 4 r of each parity: l = c0 + s0 + 2 i
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int lE = c0 + sE + 2 * (kq + 4 * r);
                const int lO = c0 + (sE ^ 1) + 2 * (kq + 4 * r);
                if (mapj >= 0 && lE <= lend) dst[2 * (long)lE] = dE[r];
                if (mapj >= 0 && lO <= lend) dst[2 * (long)lO] = dO[r];
            }
        }
        if (t < kMmaSteps) {
            sA[buf ^ 1][t] = tA;
            tA = (n2 + t <= lend) ? A[n2 + t] : 0.0;
        }
        __syncthreads();
        buf ^= 1;
    }
}

// -------------------------------------------------------------- shared synthesis on the matrix cores
//
// The synthesis of a group of up to kAdjG = 8 maps of one geometry (a smaller
// group leaves columns empty) as matrix products: for a block of 64 ring
// pairs, P[p][c] = sum_l Lambda[l][p] a_l[c] per parity, c = (map, re / im),
// accumulated over the item's l range in registers.  Each lane runs the
// recurrence of one ring pair (once for the group) into an LDS tile of
// kMmaSteps l-steps as in k_adj_mma; an instruction takes 16 ring pairs by 4
// l-steps of one parity (A, read across the tile's rows) times those 4
// l-steps' a_lm of the 8 maps (B, from global memory, a tile ahead).  A tile
// is 2 parities x 4 groups of 4 l-steps x 4 blocks of 16 pairs = 32
// instructions; the item's sums (E, O per pair and map) are stored once at
// its end.  No sums cross waves, so a workgroup is one wave; an item of K
// ring pairs per lane runs as K workgroups, all from the item's first l.
constexpr int kMmaRowS = 72;  // tile row stride: rows two apart start 32 banks apart

template <int K>
__global__ __launch_bounds__(64, 2) void k_synth_mma(const WorkItem* __restrict__ items,
                                                     const MapDesc* __restrict__ maps, const int* __restrict__ groups)
{
    __shared__ double T[kMmaSteps * kMmaRowS];
    __shared__ double sA[2][kMmaSteps];
    const WorkItem it = items[blockIdx.x / K];
    if (it.mcnt == 0) return;
    const int w = blockIdx.x % K;
    const int* __restrict__ gm = groups + kAdjG * it.map;
    const MapDesc& md = maps[gm[0]];
    const int L = md.L;
    const int m = it.m0;
    const int l0 = it.lfirst;
    const int lane = threadIdx.x;
    const long base = moff(L, m) - m;
    gcdouble* __restrict__ A = (gcdouble*)(md.A + base);
    const int pbase = it.rb * 64 * K + 64 * w;  // this wave's 64 ring pairs

    // the recurrence of ring pair pbase + lane, from the item's start state
    const long q = (long)m * md.npair_pad + pbase + lane;
    const double x = md.ring_x[pbase + lane];
    double mu1 = md.s1_cur[q], mu2 = md.s1_prev[q];
    int sc = md.s1_k[q];
    double g = sc == 0 ? 1.0 : 0.0;
    bool all = __all(sc == 0);

    // B operands: B[k][j] = a_l of column j = (map j / 2, re / im j % 2) at
    // the k-th l-step of a group; empty columns hold zeros and are not stored
    const int j = lane & 15, kq = lane >> 4;
    const int mapj = gm[j >> 1];
    gcdouble* __restrict__ aj = (gcdouble*)(maps[mapj >= 0 ? mapj : gm[0]].alm + base) + (j & 1);
    auto load_b = [&](int c0, double(&bE)[4], double(&bO)[4]) {
        const int sE = (c0 - m) & 1;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int lE = c0 + sE + 2 * (4 * h + kq), lO = c0 + (sE ^ 1) + 2 * (4 * h + kq);
            const double e = aj[2 * (long)min(lE, L)], o = aj[2 * (long)min(lO, L)];
            bE[h] = (mapj >= 0 && lE <= L) ? e : 0.0;
            bO[h] = (mapj >= 0 && lO <= L) ? o : 0.0;
        }
    };
    double bE[4], bO[4], nE[4], nO[4];
    load_b(l0, bE, bO);

    dbl4 dE[4], dO[4];
#pragma unroll
    for (int pb = 0; pb < 4; ++pb) dE[pb] = dO[pb] = dbl4{0.0, 0.0, 0.0, 0.0};

    // coefficients of the current tile in sA[buf], of the next one in tA
    double tA = 0.0;
    if (lane < kMmaSteps) {
        sA[0][lane] = (l0 + lane <= L) ? A[l0 + lane] : 0.0;
        tA = (l0 + kMmaSteps + lane <= L) ? A[l0 + kMmaSteps + lane] : 0.0;
    }
    wave_sync();
    int buf = 0;
    for (int c0 = l0; c0 <= L; c0 += kMmaSteps) {
        const int n2 = c0 + 2 * kMmaSteps;
        if (c0 + kMmaSteps <= L) load_b(c0 + kMmaSteps, nE, nO);
        // tile rows past L hold finite values whose B operands are zero
        if (c0 == l0)
            mma_fill_first<kMmaRowS>(T, sA[buf], lane, 0, x, mu1, mu2, sc, g, all);
        else if (all)
            mma_fill<false, kMmaRowS>(T, sA[buf], lane, x, mu1, mu2, sc, g, all);
        else
            mma_fill<true, kMmaRowS>(T, sA[buf], lane, x, mu1, mu2, sc, g, all);
        wave_sync();
        const int sE = (c0 - m) & 1;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const double* tE = T + (sE + 2 * (4 * h + kq)) * kMmaRowS + j;
            const double* tO = T + ((sE ^ 1) + 2 * (4 * h + kq)) * kMmaRowS + j;
#pragma unroll
            for (int pb = 0; pb < 4; ++pb) {
                dE[pb] = __builtin_amdgcn_mfma_f64_16x16x4f64(tE[16 * pb], bE[h], dE[pb], 0, 0, 0);
                dO[pb] = __builtin_amdgcn_mfma_f64_16x16x4f64(tO[16 * pb], bO[h], dO[pb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            bE[h] = nE[h];
            bO[h] = nO[h];
        }
        if (lane < kMmaSteps) {
            sA[buf ^ 1][lane] = tA;
            tA = (n2 + lane <= L) ? A[n2 + lane] : 0.0;
        }
        wave_sync();
        buf ^= 1;
    }
    // D[i][j] lands in lane j + 16 (i % 4), register i / 4: ring pair
    // pbase + 16 pb + kq + 4 r
    if (mapj >= 0) {
        gdouble* __restrict__ P = (gdouble*)maps[mapj].P;
#pragma unroll
        for (int pb = 0; pb < 4; ++pb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int pair = pbase + 16 * pb + kq + 4 * r;
                if (pair < md.npair) {
                    const long o = ((long)pair * (L + 1) + m) * 4 + (j & 1);
                    P[o] = dE[pb][r];
                    P[o + 2] = dO[pb][r];
                }
            }
    }
}

// -------------------------------------------------------------- set-up passes
// Both run the scaled recurrence from the seed lambda_mm.
// k_scan: first l at which each ring pair becomes live (L+1 if never).
template <int K>
__global__ __launch_bounds__(64) void k_scan(const WorkItem* __restrict__ items, const MapDesc* __restrict__ maps,
                                             int* __restrict__ lstart)
{
    const WorkItem it = items[blockIdx.x];
    const MapDesc& md = maps[it.map];
    const int L = md.L;
    const int npad = md.npair_pad;
    const int lane = threadIdx.x;
    for (int mi = 0; mi < it.mcnt; ++mi) {
        const int m = it.m0 + mi;
        const double* __restrict__ A = md.A + moff(L, m) - m;
        LegState<K> st;
        int p[K], ls[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            p[k] = it.rb * 64 * K + k * 64 + lane;
            st.x[k] = md.ring_x[p[k]];
            st.mu1[k] = md.seed_v[(long)m * npad + p[k]];
            st.sc[k] = md.seed_k[(long)m * npad + p[k]];
            st.mu2[k] = 0.0;
            ls[k] = st.sc[k] == 0 ? m : L + 1;
        }
        bool all = all_live<K>(st.sc);
        int l = m;
        while (!all && l < L) {
            ++l;
            st.advance(A[l]);
            if (rescale<K>(st.mu1, st.mu2, st.sc)) {
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (st.sc[k] == 0 && ls[k] > L) ls[k] = l + 1;
                all = all_live<K>(st.sc);
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) lstart[(long)m * npad + p[k]] = ls[k];
    }
}

// k_capture: the state (mu_{l-1}, mu_l, scale count) at l = item.lfirst, in
// slot sidx (adjoint segments) or at (m, pair).
template <int K>
__global__ __launch_bounds__(64) void k_capture(const WorkItem* __restrict__ items, const MapDesc* __restrict__ maps,
                                                double* __restrict__ prev, double* __restrict__ cur,
                                                int* __restrict__ kout)
{
    const WorkItem it = items[blockIdx.x];
    if (it.mcnt == 0) return;
    const MapDesc& md = maps[it.map];
    const int L = md.L;
    const int npad = md.npair_pad;
    const int lane = threadIdx.x;
    const int m = it.m0;
    const double* __restrict__ A = md.A + moff(L, m) - m;
    LegState<K> st;
    int p[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        p[k] = it.rb * 64 * K + k * 64 + lane;
        st.x[k] = md.ring_x[p[k]];
        st.mu1[k] = md.seed_v[(long)m * npad + p[k]];
        st.sc[k] = md.seed_k[(long)m * npad + p[k]];
        st.mu2[k] = 0.0;
    }
    for (int l = m + 1; l <= it.lfirst; ++l) {
        st.advance(A[l]);
        rescale<K>(st.mu1, st.mu2, st.sc);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const long q = it.sidx >= 0 ? (long)it.sidx * 64 * K + k * 64 + lane : (long)m * npad + p[k];
        prev[q] = st.mu2[k];
        cur[q] = st.mu1[k];
        kout[q] = st.sc[k];
    }
}

// ================================================================== ring stage
//
// One workgroup per ring pair.  The north and south rings of a pair share
// nph = n and phi0, so they travel as the real and imaginary parts of one
// complex sequence of length n held in LDS:
//   synthesis: Z_k = G_N,k + i G_S,k  ->  inverse DFT  ->  s_N + i s_S
//   adjoint:   f_N + i f_S  ->  forward DFT  ->  Y, split into H_N, H_S.
// Power-of-two n uses the FFT directly; other n (the polar rings, n = 4i)
// use Bluestein's chirp-z with a power-of-two length M >= 2n-1.  All
// transforms take and return natural order.  The phase of the first pixel,
// phi0 = pi/n, enters per Fourier bin: e^{i pi m/n} = e^{i pi k/n} (-1)^j for
// m = k + j n.
//
// LDS layout: element e of a sequence lives at zp(e) = e + e/16, one spare
// slot per 16 values, so the stride-16 accesses of the FFT passes fall on
// distinct banks.

__device__ __forceinline__ int zp(int e) { return e + (e >> 4); }

__device__ __forceinline__ double2 tw_get(const double2* __restrict__ tw, int t, int sign)
{
    double2 w = gview(tw)[t];
    if (sign > 0) w.y = -w.y;
    return w;
}

// x e^{SIGN 2 pi i j / 16}, j = 0..7 (cos, sin of pi j / 8 as literals).
template <int SIGN>
__device__ __forceinline__ double2 tw16(double2 x, int j)
{
    constexpr double C1 = 0.9238795325112867, S1 = 0.3826834323650898, R2 = 0.7071067811865476;
    double c, sn;
    switch (j) {
    case 0: return x;
    case 4: return SIGN > 0 ? make_double2(-x.y, x.x) : make_double2(x.y, -x.x);
    case 1: c = C1; sn = S1; break;
    case 2: c = R2; sn = R2; break;
    case 3: c = S1; sn = C1; break;
    case 5: c = -S1; sn = C1; break;
    case 6: c = -R2; sn = R2; break;
    default: c = -C1; sn = S1; break;
    }
    sn *= SIGN;
    return make_double2(x.x * c - x.y * sn, x.x * sn + x.y * c);
}

// In-register R-point DFT (R <= 16), radix-2 decimation in frequency:
// natural order in, u[bitrev(q)] = sum_r u_r e^{SIGN 2 pi i r q / R} out.
template <int R, int SIGN>
__device__ __forceinline__ void dft_reg(double2 (&u)[R])
{
#pragma unroll
    for (int s = R / 2; s >= 1; s /= 2) {
#pragma unroll
        for (int blk = 0; blk < R; blk += 2 * s) {
#pragma unroll
            for (int a = 0; a < s; ++a) {
                const double2 x0 = u[blk + a], x1 = u[blk + a + s];
                u[blk + a] = cadd(x0, x1);
                u[blk + a + s] = tw16<SIGN>(csub(x0, x1), a * (8 / s));
            }
        }
    }
}

template <int LR>
__device__ __forceinline__ constexpr int brev_c(int q)
{
    int r = 0;
    for (int b = 0; b < LR; ++b) r |= ((q >> b) & 1) << (LR - 1 - b);
    return r;
}

// One Stockham pass of radix R over M points with p = product of the radices
// already applied: butterfly i (k = i mod p) reads x[i + r M/R], twiddles by
// e^{SIGN 2 pi i r k/(p R)}, transforms, and writes y[(i - k) R + k + q p].
// Every thread loads all its butterflies before the first barrier and stores
// after it, so the pass works in place.  The twiddle loads (from the table in
// global memory, mostly L2 hits) are issued first, so their latency runs
// alongside the LDS loads and the barrier instead of after them.
constexpr int kLogMaxFFT = 13;
static_assert((1 << kLogMaxFFT) == kMaxFFT, "twiddle table length");

// Pointwise factors of a Bluestein transform fused into the first and last
// pass of its two FFTs: kFuseChirpIn multiplies each loaded input by the
// chirp (inputs at n and beyond are zero and not read), kFuseFilt each output
// by the chirp filter, kFuseChirpOut each output by 1/M and the chirp (outputs
// at n and beyond are not stored).  OS is the sign of the ring transform: for
// OS > 0 the chirps are conjugated and the filter is read reversed and
// conjugated.
enum { kFuseNone = 0, kFuseChirpIn = 1, kFuseFilt = 2, kFuseChirpOut = 3 };
struct BsFuse {
    const double2* chirp;
    const double2* filt;
    int n;
    double inv;  // 1 / M
};

// The tables hold half of each sequence: the chirp of even length n has
// c_{n-j} = c_j ((n - j)^2 = j^2 mod 2n), and the filter, the FFT of a
// sequence even in j (b_{M-j} = b_j), has F_{M-r} = F_r; entries 0..n/2 and
// 0..M/2 are stored.
template <int OS>
__device__ __forceinline__ double2 bs_chirp(BsFuse f, int j)
{
    double2 c = gview(f.chirp)[min(j, f.n - j)];
    if (OS > 0) c.y = -c.y;
    return c;
}

template <int OS>
__device__ __forceinline__ double2 bs_filt(BsFuse f, int r, int M)
{
    double2 g = gview(f.filt)[min(r, M - r)];
    if (OS > 0) g.y = -g.y;  // F_{M-r} conjugated, F_{M-r} = F_r
    return g;
}

template <int R, int LR, int SIGN, int MAXM, int PRE = kFuseNone, int POST = kFuseNone, int OS = -1>
__device__ __forceinline__ void stockham_pass(double2* Z, int M, int p, int lp, const double2* __restrict__ tw,
                                              BsFuse fz = BsFuse{})
{
    constexpr int NT = ring_threads(MAXM);
    constexpr int B = (MAXM / R + NT - 1) / NT;  // butterflies per thread, at most
    const int tid = threadIdx.x;
    const int nb = M / R;
    double2 u[B][R];
    double2 w1v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int i = tid + b * NT;
        const int k = i & (p - 1);
        w1v[b] = make_double2(1.0, 0.0);
        if (i < nb && k) w1v[b] = tw_get(tw, k << (kLogMaxFFT - lp - LR), SIGN);  // k kMaxFFT / (p R)
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int i = tid + b * NT;
        if (i < nb) {
            if (PRE == kFuseChirpIn) {
                // chirps first (global loads in flight), then the LDS loads
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int e = i + r * nb;
                    u[b][r] = e < fz.n ? bs_chirp<OS>(fz, e) : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int e = i + r * nb;
                    if (e < fz.n) u[b][r] = cmul(Z[zp(e)], u[b][r]);
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) u[b][r] = Z[zp(i + r * nb)];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int i = tid + b * NT;
        if (i < nb) {
            const int k = i & (p - 1);
            if (k) {
                const double2 w1 = w1v[b];
                double2 w = w1;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    u[b][r] = cmul(u[b][r], w);
                    if (r + 1 < R) w = cmul(w, w1);
                }
            }
            dft_reg<R, SIGN>(u[b]);
            const int j = (i - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int e = j + q * p;
                double2 v = u[b][brev_c<LR>(q)];
                if (POST == kFuseFilt) {
                    v = cmul(v, bs_filt<OS>(fz, e, M));
                } else if (POST == kFuseChirpOut) {
                    if (e >= fz.n) continue;
                    v = cmul(make_double2(v.x * fz.inv, v.y * fz.inv), bs_chirp<OS>(fz, e));
                }
                Z[zp(e)] = v;
            }
        }
    }
    __syncthreads();
}

// stockham_pass over two transforms of length M = 2^logM side by side in
// LDS (transform t at Z[t M ..)): butterfly g belongs to transform g / (M/R).
template <int R, int LR, int SIGN, int MAXM>
__device__ __forceinline__ void stockham_pass_two(double2* Z, int logM, int p, int lp, const double2* __restrict__ tw)
{
    constexpr int NT = ring_threads(MAXM);
    constexpr int B = (MAXM / R + NT - 1) / NT;  // butterflies per thread, at most (2 M <= MAXM)
    const int tid = threadIdx.x;
    const int lnb = logM - LR;  // log2 of the butterflies per transform
    const int nb = 1 << lnb;
    double2 u[B][R];
    double2 w1v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int g = tid + b * NT;
        const int i = g & (nb - 1);
        const int k = i & (p - 1);
        w1v[b] = make_double2(1.0, 0.0);
        if (g < 2 * nb && k) w1v[b] = tw_get(tw, k << (kLogMaxFFT - lp - LR), SIGN);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int g = tid + b * NT;
        if (g < 2 * nb) {
            const int base = (g >> lnb) << logM;
            const int i = g & (nb - 1);
#pragma unroll
            for (int r = 0; r < R; ++r) u[b][r] = Z[zp(base + i + r * nb)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int g = tid + b * NT;
        if (g < 2 * nb) {
            const int base = (g >> lnb) << logM;
            const int i = g & (nb - 1);
            const int k = i & (p - 1);
            if (k) {
                const double2 w1 = w1v[b];
                double2 w = w1;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    u[b][r] = cmul(u[b][r], w);
                    if (r + 1 < R) w = cmul(w, w1);
                }
            }
            dft_reg<R, SIGN>(u[b]);
            const int j = (i - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) Z[zp(base + j + q * p)] = u[b][brev_c<LR>(q)];
        }
    }
    __syncthreads();
}

// fft() of two transforms of length M = 2^logM side by side (2 M <= MAXM).
template <int SIGN, int MAXM>
__device__ void fft_two(double2* Z, int logM, const double2* __restrict__ tw)
{
    int p = 1, lp = 0, rem = logM;
    for (; rem >= 3; rem -= 3, p <<= 3, lp += 3) stockham_pass_two<8, 3, SIGN, MAXM>(Z, logM, p, lp, tw);
    if (rem == 2)
        stockham_pass_two<4, 2, SIGN, MAXM>(Z, logM, p, lp, tw);
    else if (rem == 1)
        stockham_pass_two<2, 1, SIGN, MAXM>(Z, logM, p, lp, tw);
}

// Unnormalised DFT of Z[0..M) in place, natural order in and out:
// Z_q <- sum_r Z_r e^{SIGN 2 pi i r q / M}, M = 2^logM <= MAXM <= kMaxFFT, M >= 2.
// Radix-8 passes, then one radix-4/2 pass for the remaining factor.  MAXM
// bounds the butterflies per thread and so the registers.
template <int SIGN, int MAXM>
__device__ void fft(double2* Z, int M, int logM, const double2* __restrict__ tw)
{
    // radix 8: one butterfly per thread at M = 8 x threads, within the 128
    // VGPRs that four waves per SIMD allow (radix 16 spills there)
    constexpr int LR = 3;
    int p = 1, lp = 0, rem = logM;
    for (; rem >= LR; rem -= LR, p <<= LR, lp += LR) stockham_pass<1 << LR, LR, SIGN, MAXM>(Z, M, p, lp, tw);
    if (rem == 2)
        stockham_pass<4, 2, SIGN, MAXM>(Z, M, p, lp, tw);
    else if (rem == 1)
        stockham_pass<2, 1, SIGN, MAXM>(Z, M, p, lp, tw);
}

// The FFT of fft() with pointwise factors fused into its first (PRE) and last
// (POST) pass; M >= 8.
template <int SIGN, int MAXM, int PRE, int POST, int OS>
__device__ void fft_fused(double2* Z, int M, int logM, const double2* __restrict__ tw, BsFuse fz)
{
    if (logM == 3) {
        stockham_pass<8, 3, SIGN, MAXM, PRE, POST, OS>(Z, M, 1, 0, tw, fz);
        return;
    }
    stockham_pass<8, 3, SIGN, MAXM, PRE, kFuseNone, OS>(Z, M, 1, 0, tw, fz);
    int p = 8, lp = 3, rem = logM - 3;
    for (; rem > 3; rem -= 3, p <<= 3, lp += 3) stockham_pass<8, 3, SIGN, MAXM>(Z, M, p, lp, tw);
    if (rem == 3)
        stockham_pass<8, 3, SIGN, MAXM, kFuseNone, POST, OS>(Z, M, p, lp, tw, fz);
    else if (rem == 2)
        stockham_pass<4, 2, SIGN, MAXM, kFuseNone, POST, OS>(Z, M, p, lp, tw, fz);
    else
        stockham_pass<2, 1, SIGN, MAXM, kFuseNone, POST, OS>(Z, M, p, lp, tw, fz);
}

// Bluestein DFT of z[0..n) (natural order; M >= 2n, so M >= 8): result in
// natural order in z[0..n).  chirp[j] = e^{-i pi j^2/n}, j <= n/2; filt =
// FFT_M of e^{+i pi j^2/n} (wrapped), entries 0..M/2.  For SIGN > 0 the chirps
// are conjugated and the filter is read reversed and conjugated.  The chirp
// products, the zero padding, the filter product and the final chirp and 1/M
// ride on the first and last passes of the two FFTs.
template <int SIGN, int MAXM>
__device__ void bluestein(double2* z, int n, int M, int logM, const double2* __restrict__ chirp,
                          const double2* __restrict__ filt, const double2* __restrict__ tw)
{
    const BsFuse fz{chirp, filt, n, ldexp(1.0, -logM)};  // 1 / M, M = 2^logM
    fft_fused<-1, MAXM, kFuseChirpIn, kFuseFilt, SIGN>(z, M, logM, tw, fz);
    fft_fused<+1, MAXM, kFuseNone, kFuseChirpOut, SIGN>(z, M, logM, tw, fz);
}

// Length-n DFT in LDS (natural order): the FFT for power-of-two n, Bluestein otherwise.
template <int SIGN, int MAXM>
__device__ __forceinline__ void ring_dft(double2* Z, int n, int M, int logM, bool pow2, const double2* chirp,
                                         const double2* filt, const double2* __restrict__ tw)
{
    if (pow2)
        fft<SIGN, MAXM>(Z, n, logM, tw);
    else
        bluestein<SIGN, MAXM>(Z, n, M, logM, chirp, filt, tw);
}

// Alias sums of Fourier bin k over the parity rows (E, O) of a ring pair:
// sum_j (-1)^{j shift} F_{k+jn} + sum_{j>=1} (-1)^{j shift} conj(F_{jn-k}),
// for j = r, r+R, r+2R, ... (one thread group of R).
__device__ __forceinline__ void fold_aliases(const double2* __restrict__ row, int L, int n, int k, int r, int R,
                                             bool shifted, double2& gN, double2& gS)
{
    gN = make_double2(0.0, 0.0);
    gS = make_double2(0.0, 0.0);
    for (int j = r; k + (long)j * n <= L; j += R) {
        const int mm = k + j * n;
        const double2 e = row[2 * mm], o = row[2 * mm + 1];
        const double sg = (shifted && (j & 1)) ? -1.0 : 1.0;
        gN.x += sg * (e.x + o.x);
        gN.y += sg * (e.y + o.y);
        gS.x += sg * (e.x - o.x);
        gS.y += sg * (e.y - o.y);
    }
    for (int j = 1 + r; (long)j * n - k <= L; j += R) {
        const int mm = j * n - k;
        const double2 e = row[2 * mm], o = row[2 * mm + 1];
        const double sg = (shifted && (j & 1)) ? -1.0 : 1.0;
        gN.x += sg * (e.x + o.x);
        gN.y -= sg * (e.y + o.y);
        gS.x += sg * (e.x - o.x);
        gS.y -= sg * (e.y - o.y);
    }
}

// e^{i pi (k0 + t D) / n} for t = 0, 1, 2, ...: the first two factors from
// sincospi, each later term by one complex multiply (a thread's bins and
// samples step by a fixed D; the rounding grows by about an ulp per step, over
// at most 16 steps).  A phase by sincospi costs a 64-bit remainder, an FP64
// division and the polynomial, about a hundred instructions.
struct PhaseSeq {
    double2 cur, step;
    __device__ __forceinline__ void init(long k0, long D, long n)
    {
        cur = cis_pi_frac(k0, n);
        step = cis_pi_frac(D, n);
    }
    __device__ __forceinline__ double2 next()
    {
        const double2 c = cur;
        cur = cmul(cur, step);
        return c;
    }
};

// Bin of the packed pair spectrum (natural order): phase ph = e^{i pi k/n}
// (bin k) for shifted rings, no south ring at the equator.
__device__ __forceinline__ void fold_store(double2* Z, int slot, double2 ph, bool shifted, bool no_south,
                                           double2 gN, double2 gS)
{
    if (no_south) gS = make_double2(0.0, 0.0);
    if (shifted) {
        gN = cmul(gN, ph);
        gS = cmul(gS, ph);
    }
    Z[zp(slot)] = make_double2(gN.x - gS.y, gN.y + gS.x);
}

// Unfold bin k and its mirror n - k of Y into Q (the aliases m = k + j n,
// j = r, r + R, ...): Psi_X,m = e^{-i pi k/n} (-1)^j H_X,k; ph = e^{i pi k/n}
// (used for shifted rings).
__device__ __forceinline__ void unfold_bin(double2* qrow, int L, int n, int k, int r, int R, bool shifted,
                                           double2 ph, double2 y, double2 z2)
{
    const double2 yc = make_double2(z2.x, -z2.y);
    double2 hN = make_double2(0.5 * (y.x + yc.x), 0.5 * (y.y + yc.y));
    double2 hS = make_double2(0.5 * (y.y - yc.y), -0.5 * (y.x - yc.x));
    if (shifted) {
        ph.y = -ph.y;
        hN = cmul(hN, ph);
        hS = cmul(hS, ph);
    }
    const double2 pe = cadd(hN, hS), po = csub(hN, hS);
    auto q = gview(qrow);
    for (int j = r; k + (long)j * n <= L; j += R) {
        const int mm = k + j * n;
        const double sg = (shifted && (j & 1)) ? -1.0 : 1.0;
        q[2 * mm] = make_double2(sg * pe.x, sg * pe.y);
        q[2 * mm + 1] = make_double2(sg * po.x, sg * po.y);
    }
}

// Y_k of a split ring from its two half transforms: SE (global scratch) and
// Z (LDS), Y_k = E_{k mod h} + w O_{k mod h}, w = e^{-2 pi i k/n}.
__device__ __forceinline__ double2 split_bin(const double2* SE, const double2* Z, int k, int h, double2 w)
{
    const int kk = k % h;
    return cadd(gview(SE)[kk], cmul(Z[zp(kk)], w));
}

// A ring pair whose length-n Bluestein transform would not fit in LDS (n = 4i
// with 2n - 1 > kMaxFFT): each direction splits into two length-h = n/2 DFTs
// over the even and odd samples, each done by Bluestein in LDS, the first kept
// in global scratch while the second runs (decimation in time):
//   synthesis  s_j = E_j + w^j O_j,  s_{j+h} = E_j - w^j O_j,  w = e^{2 pi i/n},
//              E, O = inverse DFT_h of the even / odd bins;
//   analysis   Y_k = FE_{k mod h} + w^{-k} FO_{k mod h},
//              FE, FO = DFT_h of the even / odd samples.
// Scratch per ring: SE[h] (a half transform) and SF[n] (weighted residual).
template <int MODE, int MAXM>
__device__ void ring_split(const MapDesc& md, int p, int n, int M, int logM, bool shifted, long oN, long oS,
                           double* __restrict__ map_out, const double* __restrict__ map_in,
                           const double2* __restrict__ tw, double2* Z, double* sh)
{
    const int L = md.L;
    constexpr int NT = ring_threads(MAXM);
    const int tid = threadIdx.x;
    const int h = n >> 1;
    const double2* filt = md.bs_filt + md.ring_bsofs[p];
    const double2* chirp = md.bs_chirp + md.ring_chofs[p];
    double2* SE = md.ring_scr + md.ring_scrofs[p];
    double2* SF = SE + h;

    if (MODE != kRingAna) {
        const double2* row = reinterpret_cast<const double2*>(md.P + (long)p * (L + 1) * 4);
        for (int r = 0; r < 2; ++r) {
            PhaseSeq ps;
            if (shifted) ps.init(2 * tid + r, 2 * NT, n);
            for (int kk = tid; kk < h; kk += NT) {
                const int k = 2 * kk + r;
                double2 gN, gS;
                fold_aliases(row, L, n, k, 0, 1, shifted, gN, gS);
                fold_store(Z, kk, shifted ? ps.next() : make_double2(1.0, 0.0), shifted, oS < 0, gN, gS);
            }
            __syncthreads();
            bluestein<+1, MAXM>(Z, h, M, logM, chirp, filt, tw);
            if (r == 0) {
                for (int j = tid; j < h; j += NT) SE[j] = Z[zp(j)];
                __syncthreads();
            }
        }
        double chi = 0.0;
        PhaseSeq tw2;  // e^{2 pi i j / n}
        tw2.init(2 * tid, 2 * NT, n);
        for (int j = tid; j < h; j += NT) {
            const double2 e = SE[j];
            const double2 o = cmul(Z[zp(j)], tw2.next());
            const double2 sv[2] = {cadd(e, o), csub(e, o)};
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const long t = j + half * h;
                const double2 v = sv[half];
                if (MODE == kRingSyn) {
                    map_out[oN + t] = v.x;
                    if (oS >= 0) map_out[oS + t] = v.y;
                } else if (MODE == kRingWeight) {
                    SF[t] = make_double2(md.w[oN + t] * v.x, oS >= 0 ? md.w[oS + t] * v.y : 0.0);
                } else {
                    const double rN = md.d[oN + t] - v.x;
                    const double fN = md.w[oN + t] * rN;
                    chi = fma(fN, rN, chi);
                    double fS = 0.0;
                    if (oS >= 0) {
                        const double rS = md.d[oS + t] - v.y;
                        fS = md.w[oS + t] * rS;
                        chi = fma(fS, rS, chi);
                    }
                    SF[t] = make_double2(fN, fS);
                }
            }
        }
        if (MODE == kRingSyn) return;
        if (MODE == kRingLike) {
            const double t = block_sum<NT>(chi, sh);
            if (tid == 0) md.chi[p] = t;
        } else {
            __syncthreads();
        }
    }

    for (int r = 0; r < 2; ++r) {
        for (int j = tid; j < h; j += NT) {
            const long t = 2 * j + r;
            Z[zp(j)] = (MODE == kRingAna) ? make_double2(map_in[oN + t], oS >= 0 ? map_in[oS + t] : 0.0) : SF[t];
        }
        __syncthreads();
        bluestein<-1, MAXM>(Z, h, M, logM, chirp, filt, tw);
        if (r == 0) {
            for (int j = tid; j < h; j += NT) SE[j] = Z[zp(j)];
            __syncthreads();
        }
    }

    double2* qrow = reinterpret_cast<double2*>(md.Q + (long)p * (L + 1) * 4);
    PhaseSeq wk, pk;  // e^{2 pi i k/n} and e^{i pi k/n}
    wk.init(2 * tid, 2 * NT, n);
    if (shifted) pk.init(tid, NT, n);
    for (int k = tid; k < n && k <= L; k += NT) {
        const int nk = (k == 0) ? 0 : n - k;
        const double2 a = wk.next();  // e^{-2 pi i (n-k)/n} = a, e^{-2 pi i k/n} = conj(a)
        const double2 ph = shifted ? pk.next() : make_double2(1.0, 0.0);
        unfold_bin(qrow, L, n, k, 0, 1, shifted, ph, split_bin(md.ring_scr + md.ring_scrofs[p], Z, k, h,
                                                                 make_double2(a.x, -a.y)),
                   split_bin(md.ring_scr + md.ring_scrofs[p], Z, nk, h, a));
    }
}

// Two power-of-two rings of length n = MAXM / 2 in one workgroup (each longer
// than its map's lmax and unsplit; the pair of a RingItem with pair2 >= 0):
// ring r's spectrum and samples live at Z[r n ..), a thread holds CNT / 2
// samples of each, and every step is the single ring's (one-alias fold,
// transform, residual or weights, transform, unfold) with the two transforms
// side by side.  Rings of half the launch's length would otherwise leave half
// of their workgroup's threads idle in every FFT pass.
constexpr int clog2(int v) { return v <= 1 ? 0 : 1 + clog2(v / 2); }

template <int MODE, int MAXM>
__device__ void ring_two(const RingItem& it, const MapDesc* __restrict__ maps, double* __restrict__ map_out,
                         const double* __restrict__ map_in, const double2* __restrict__ tw, double2* Z, double* sh)
{
    constexpr int NT = ring_threads(MAXM);
    constexpr int n = MAXM / 2;
    constexpr int logn = clog2(n);
    constexpr int CNT = MAXM / NT;  // samples per thread, both rings
    constexpr int CH = CNT / 2;     // per ring
    static_assert(CNT % 2 == 0 && (1 << logn) == n, "two rings per workgroup");
    const int tid = threadIdx.x;
    const MapDesc* mdr[2] = {&maps[it.map], &maps[it.map2]};
    const int pr[2] = {it.pair, it.pair2};
    bool shf[2];
    long oN[2], oS[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        shf[r] = mdr[r]->ring_shift[pr[r]] != 0;
        oN[r] = mdr[r]->ring_ofsN[pr[r]];
        oS[r] = mdr[r]->ring_ofsS[pr[r]];
    }

    if (MODE != kRingAna) {
        // fold, n > L: at most the alias k and the mirror n - k per bin
        double2 e[CNT], o[CNT];
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int r = t / CH;
            const int L = mdr[r]->L;
            auto rw = gview(reinterpret_cast<const double2*>(mdr[r]->P + (long)pr[r] * (L + 1) * 4));
            const int k = tid + (t - r * CH) * NT;
            const int m = k <= L ? k : n - k;
            if (k <= L || (k >= 1 && n - k <= L)) {
                e[t] = rw[2 * m];
                o[t] = rw[2 * m + 1];
            }
        }
        PhaseSeq ps;
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int r = t / CH;
            const int L = mdr[r]->L;
            auto rw = gview(reinterpret_cast<const double2*>(mdr[r]->P + (long)pr[r] * (L + 1) * 4));
            const int k = tid + (t - r * CH) * NT;
            if (t % CH == 0 && shf[r]) ps.init(tid, NT, n);
            const double sg = shf[r] ? -1.0 : 1.0;  // (-1)^{j shift} at j = 1
            const double2 ph = shf[r] ? ps.next() : make_double2(1.0, 0.0);
            double2 gN = make_double2(0.0, 0.0), gS = make_double2(0.0, 0.0);
            const bool mirror = k >= 1 && n - k <= L;
            if (k <= L) {
                gN.x += 1.0 * (e[t].x + o[t].x);
                gN.y += 1.0 * (e[t].y + o[t].y);
                gS.x += 1.0 * (e[t].x - o[t].x);
                gS.y += 1.0 * (e[t].y - o[t].y);
                if (mirror) {
                    e[t] = rw[2 * (n - k)];
                    o[t] = rw[2 * (n - k) + 1];
                }
            }
            if (mirror) {
                gN.x += sg * (e[t].x + o[t].x);
                gN.y -= sg * (e[t].y + o[t].y);
                gS.x += sg * (e[t].x - o[t].x);
                gS.y -= sg * (e[t].y - o[t].y);
            }
            fold_store(Z, r * n + k, ph, shf[r], oS[r] < 0, gN, gS);
        }
        __syncthreads();
        fft_two<+1, MAXM>(Z, logn, tw);
    }

    if (MODE == kRingSyn) {
        gdouble* out = gview(map_out);
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int r = t / CH;
            const int j = tid + (t - r * CH) * NT;
            const double2 v = Z[zp(r * n + j)];
            out[oN[r] + j] = v.x;
            if (oS[r] >= 0) out[oS[r] + j] = v.y;
        }
        return;
    }

    if (MODE == kRingLike) {
        double dN[CNT], wN[CNT], dS[CNT], wS[CNT];
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int r = t / CH;
            gcdouble* dd = gview(mdr[r]->d);
            gcdouble* ww = gview(mdr[r]->w);
            const int j = tid + (t - r * CH) * NT;
            dN[t] = dd[oN[r] + j];
            wN[t] = ww[oN[r] + j];
            if (oS[r] >= 0) {
                dS[t] = dd[oS[r] + j];
                wS[t] = ww[oS[r] + j];
            }
        }
        double chi[2] = {0.0, 0.0};
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int r = t / CH;
            const int j = tid + (t - r * CH) * NT;
            const double2 v = Z[zp(r * n + j)];
            const double rN = dN[t] - v.x;
            const double fN = wN[t] * rN;
            chi[r] = fma(fN, rN, chi[r]);
            double fS = 0.0;
            if (oS[r] >= 0) {
                const double rS = dS[t] - v.y;
                fS = wS[t] * rS;
                chi[r] = fma(fS, rS, chi[r]);
            }
            Z[zp(r * n + j)] = make_double2(fN, fS);
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const double c = block_sum<NT>(chi[r], sh);
            if (tid == 0) mdr[r]->chi[pr[r]] = c;
        }
    } else if (MODE == kRingWeight) {
        double wN[CNT], wS[CNT];
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int r = t / CH;
            gcdouble* ww = gview(mdr[r]->w);
            const int j = tid + (t - r * CH) * NT;
            wN[t] = ww[oN[r] + j];
            wS[t] = oS[r] >= 0 ? ww[oS[r] + j] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int r = t / CH;
            const int j = tid + (t - r * CH) * NT;
            const double2 v = Z[zp(r * n + j)];
            Z[zp(r * n + j)] = make_double2(wN[t] * v.x, oS[r] >= 0 ? wS[t] * v.y : 0.0);
        }
        __syncthreads();
    } else {  // kRingAna
        gcdouble* in = gview(map_in);
        double2 mv[CNT];
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int r = t / CH;
            const int j = tid + (t - r * CH) * NT;
            mv[t] = make_double2(in[oN[r] + j], oS[r] >= 0 ? in[oS[r] + j] : 0.0);
        }
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int r = t / CH;
            const int j = tid + (t - r * CH) * NT;
            Z[zp(r * n + j)] = mv[t];
        }
        __syncthreads();
    }
    fft_two<-1, MAXM>(Z, logn, tw);

    // unfold, one alias per bin (n > L)
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int L = mdr[r]->L;
        double2* qrow = reinterpret_cast<double2*>(mdr[r]->Q + (long)pr[r] * (L + 1) * 4);
        PhaseSeq ps;
        if (shf[r]) ps.init(tid, NT, n);
        for (int k = tid; k < n; k += NT) {
            const double2 ph = shf[r] ? ps.next() : make_double2(1.0, 0.0);
            if (k > L) continue;
            const int nk = (k == 0) ? 0 : n - k;
            unfold_bin(qrow, L, n, k, 0, 1, shf[r], ph, Z[zp(r * n + k)], Z[zp(r * n + nk)]);
        }
    }
}

// Ring kernels up to M = 4096 are held to four waves per SIMD (128 VGPRs):
// two workgroups of 512 threads, or four of 256 up to M = 2048, per CU.
template <int MODE, int MAXM>
__global__ __launch_bounds__(ring_threads(MAXM), MAXM > 4096 ? 1 : 4) void k_ring(const RingItem* __restrict__ items,
                                                       const MapDesc* __restrict__ maps,
                                                       double* __restrict__ map_out,
                                                       const double* __restrict__ map_in,
                                                       const double2* __restrict__ tw)
{
    constexpr int NT = ring_threads(MAXM);
    extern __shared__ double2 Z[];
    __shared__ double sh[NT / 64];
    const RingItem it = items[blockIdx.x];
    if constexpr (MAXM <= 4096) {
        if (it.pair2 >= 0) {
            ring_two<MODE, MAXM>(it, maps, map_out, map_in, tw, Z, sh);
            return;
        }
    }
    const MapDesc& md = maps[it.map];
    const int L = md.L;
    const int p = it.pair;
    const int n = md.ring_nph[p];
    const int M = md.ring_M[p];
    int logM = 0;
    while ((1 << logM) < M) ++logM;
    const bool pow2 = (M == n);
    const bool shifted = md.ring_shift[p] != 0;
    const long oN = md.ring_ofsN[p];
    const long oS = md.ring_ofsS[p];
    if (md.ring_scrofs[p] >= 0) {
        ring_split<MODE, MAXM>(md, p, n, M, logM, shifted, oN, oS, map_out, map_in, tw, Z, sh);
        return;
    }
    const double2* filt = pow2 ? nullptr : md.bs_filt + md.ring_bsofs[p];
    const double2* chirp = pow2 ? nullptr : md.bs_chirp + md.ring_chofs[p];
    const int tid = threadIdx.x;
    constexpr int CNT = (MAXM + NT - 1) / NT;  // ring samples per thread, at most
    gdouble* out = gview(map_out);
    gcdouble* in = gview(map_in);
    gcdouble* dd = gview(md.d);
    gcdouble* ww = gview(md.w);
    // for n < NT the aliases of each bin are split over R thread
    // groups (reads stay coalesced over k), partial sums added in fixed order
    const int R = n < NT ? NT / n : 1;

    if (MODE != kRingAna) {
        const double2* row = reinterpret_cast<const double2*>(md.P + (long)p * (L + 1) * 4);
        // fold: Z_k = G_N,k + i G_S,k
        if (R == 1 && n > L) {
            // at most the alias k (k <= L) and the mirror n - k (n - k <= L)
            // per bin, and both only where n - L <= k <= L: one load per bin
            // issued for every bin of the thread before the sums, the rare
            // second one where the sums need it
            auto rw = gview(row);
            double2 e[CNT], o[CNT];
#pragma unroll
            for (int t = 0; t < CNT; ++t) {
                const int k = tid + t * NT;
                const int m = k <= L ? k : n - k;
                if (k < n && (k <= L || (k >= 1 && n - k <= L))) {
                    e[t] = rw[2 * m];
                    o[t] = rw[2 * m + 1];
                }
            }
            const double sg = shifted ? -1.0 : 1.0;  // (-1)^{j shift} at j = 1
            PhaseSeq ps;
            if (shifted) ps.init(tid, NT, n);
#pragma unroll
            for (int t = 0; t < CNT; ++t) {
                const int k = tid + t * NT;
                if (k >= n) continue;
                const double2 ph = shifted ? ps.next() : make_double2(1.0, 0.0);
                double2 gN = make_double2(0.0, 0.0), gS = make_double2(0.0, 0.0);
                const bool mirror = k >= 1 && n - k <= L;
                if (k <= L) {
                    gN.x += 1.0 * (e[t].x + o[t].x);
                    gN.y += 1.0 * (e[t].y + o[t].y);
                    gS.x += 1.0 * (e[t].x - o[t].x);
                    gS.y += 1.0 * (e[t].y - o[t].y);
                    if (mirror) {
                        e[t] = rw[2 * (n - k)];
                        o[t] = rw[2 * (n - k) + 1];
                    }
                }
                if (mirror) {
                    gN.x += sg * (e[t].x + o[t].x);
                    gN.y -= sg * (e[t].y + o[t].y);
                    gS.x += sg * (e[t].x - o[t].x);
                    gS.y -= sg * (e[t].y - o[t].y);
                }
                fold_store(Z, k, ph, shifted, oS < 0, gN, gS);
            }
        } else if (R == 1) {
            PhaseSeq ps;
            if (shifted) ps.init(tid, NT, n);
            for (int k = tid; k < n; k += NT) {
                double2 gN, gS;
                fold_aliases(row, L, n, k, 0, 1, shifted, gN, gS);
                fold_store(Z, k, shifted ? ps.next() : make_double2(1.0, 0.0), shifted, oS < 0, gN, gS);
            }
        } else {
            double2 gN = make_double2(0.0, 0.0), gS = make_double2(0.0, 0.0);
            if (tid < R * n) {
                fold_aliases(row, L, n, tid % n, tid / n, R, shifted, gN, gS);
                Z[zp(2 * tid)] = gN;
                Z[zp(2 * tid + 1)] = gS;
            }
            __syncthreads();
            if (tid < n) {
                gN = make_double2(0.0, 0.0);
                gS = make_double2(0.0, 0.0);
                for (int r = 0; r < R; ++r) {
                    gN = cadd(gN, Z[zp(2 * (r * n + tid))]);
                    gS = cadd(gS, Z[zp(2 * (r * n + tid) + 1)]);
                }
            }
            __syncthreads();
            if (tid < n)
                fold_store(Z, tid, shifted ? cis_pi_frac(tid, n) : make_double2(1.0, 0.0), shifted, oS < 0, gN, gS);
        }
        __syncthreads();
        ring_dft<+1, MAXM>(Z, n, M, logM, pow2, chirp, filt, tw);
    }

    if (MODE == kRingSyn) {
        for (int j = tid; j < n; j += NT) {
            const double2 v = Z[zp(j)];
            out[oN + j] = v.x;
            if (oS >= 0) out[oS + j] = v.y;
        }
        return;
    }

    // per-pixel inputs of the thread's samples, all loaded before use
    if (MODE == kRingLike) {
        double dN[CNT], wN[CNT], dS[CNT], wS[CNT];
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int j = tid + t * NT;
            if (j < n) {
                dN[t] = dd[oN + j];
                wN[t] = ww[oN + j];
                if (oS >= 0) {
                    dS[t] = dd[oS + j];
                    wS[t] = ww[oS + j];
                }
            }
        }
        double chi = 0.0;
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int j = tid + t * NT;
            if (j >= n) continue;
            const double2 v = Z[zp(j)];
            const double rN = dN[t] - v.x;
            const double fN = wN[t] * rN;
            chi = fma(fN, rN, chi);
            double fS = 0.0;
            if (oS >= 0) {
                const double rS = dS[t] - v.y;
                fS = wS[t] * rS;
                chi = fma(fS, rS, chi);
            }
            Z[zp(j)] = make_double2(fN, fS);
        }
        const double t = block_sum<NT>(chi, sh);
        if (tid == 0) md.chi[p] = t;
    } else if (MODE == kRingWeight) {
        double wN[CNT], wS[CNT];
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int j = tid + t * NT;
            if (j < n) {
                wN[t] = ww[oN + j];
                wS[t] = oS >= 0 ? ww[oS + j] : 0.0;
            }
        }
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int j = tid + t * NT;
            if (j >= n) continue;
            const double2 v = Z[zp(j)];
            Z[zp(j)] = make_double2(wN[t] * v.x, oS >= 0 ? wS[t] * v.y : 0.0);
        }
        __syncthreads();
    } else {  // kRingAna: load the map to analyse
        double2 mv[CNT];
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int j = tid + t * NT;
            if (j < n) mv[t] = make_double2(in[oN + j], oS >= 0 ? in[oS + j] : 0.0);
        }
#pragma unroll
        for (int t = 0; t < CNT; ++t) {
            const int j = tid + t * NT;
            if (j < n) Z[zp(j)] = mv[t];
        }
        __syncthreads();
    }
    ring_dft<-1, MAXM>(Z, n, M, logM, pow2, chirp, filt, tw);

    // unfold, with the aliases of bin k split over R thread groups as in the fold
    double2* qrow = reinterpret_cast<double2*>(md.Q + (long)p * (L + 1) * 4);
    if (R == 1) {
        PhaseSeq ps;
        if (shifted) ps.init(tid, NT, n);
        for (int k = tid; k < n; k += NT) {
            const double2 ph = shifted ? ps.next() : make_double2(1.0, 0.0);
            if (k > L) continue;
            const int nk = (k == 0) ? 0 : n - k;
            unfold_bin(qrow, L, n, k, 0, 1, shifted, ph, Z[zp(k)], Z[zp(nk)]);
        }
    } else if (tid < R * n) {  // R n <= NT: one (bin, group) per thread
        const int k = tid % n, r = tid / n;
        if (k <= L) {
            const int nk = (k == 0) ? 0 : n - k;
            unfold_bin(qrow, L, n, k, r, R, shifted, shifted ? cis_pi_frac(k, n) : make_double2(1.0, 0.0), Z[zp(k)],
                       Z[zp(nk)]);
        }
    }
}

// Bluestein set-up for one ring pair of Bluestein length n (nph, or nph/2 for
// a split ring; even): chirp[j] = e^{-i pi j^2/n} for j <= n/2 and filt =
// FFT_M of b_j = e^{+i pi j^2/n} for |j| < n wrapped mod M, entries 0..M/2 (the
// rest follow by symmetry).
__global__ __launch_bounds__(ring_threads(kMaxFFT)) void k_bluestein_setup(const int* __restrict__ pairs,
                                                                  const MapDesc* __restrict__ md_,
                                                                  double2* __restrict__ filt_out,
                                                                  double2* __restrict__ chirp_out,
                                                                  const double2* __restrict__ tw)
{
    extern __shared__ double2 Z[];
    const MapDesc& md = *md_;
    const int p = pairs[blockIdx.x];
    const int n = md.ring_scrofs[p] >= 0 ? md.ring_nph[p] / 2 : md.ring_nph[p];  // Bluestein length
    const int M = md.ring_M[p];
    int logM = 0;
    while ((1 << logM) < M) ++logM;
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
        int jj = -1;
        if (j < n)
            jj = j;
        else if (j > M - n)
            jj = M - j;
        Z[zp(j)] = jj >= 0 ? cis_pi_frac(((long)jj * jj) % (2L * n), n) : make_double2(0.0, 0.0);
    }
    double2* ch = chirp_out + md.ring_chofs[p];
    for (int j = threadIdx.x; j <= n / 2; j += blockDim.x) {
        double2 c = cis_pi_frac(((long)j * j) % (2L * n), n);
        c.y = -c.y;
        ch[j] = c;
    }
    __syncthreads();
    fft<-1, kMaxFFT>(Z, M, logM, tw);
    double2* dst = filt_out + md.ring_bsofs[p];
    for (int j = threadIdx.x; j <= M / 2; j += blockDim.x) dst[j] = Z[zp(j)];
}

// ================================================================== k_final
// One block: -2lnL = sum_pairs chi + sum_blocks prior, fixed order.
__device__ __forceinline__ void final_map(const MapDesc& md, double* __restrict__ f)
{
    __shared__ double sh[16];
    double a = 0.0;
    for (int i = threadIdx.x; i < md.npair; i += 1024) a += md.chi[i];
    for (int i = threadIdx.x; i < md.nprior; i += 1024) a += md.prior[i];
    const double t = block_sum<1024>(a, sh);
    if (threadIdx.x == 0) {
        *f = t;
        if (!isfinite(t)) *md.nonfinite = 1;
    }
}

__global__ __launch_bounds__(1024) void k_final(const MapDesc* __restrict__ md_, double* __restrict__ f)
{
    final_map(*md_, f);
}

// one block per map of the group
__global__ __launch_bounds__(1024) void k_final_b(const MapDesc* __restrict__ mds, const BatchArgs a)
{
    final_map(mds[blockIdx.x], a.f + blockIdx.x);
}

// ================================================================== BLAS-1

// y and x may be the same vector (in-place scaling y = c y, y += c y).
__global__ void k_vec(int op, double* y, const double* x, double c, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        switch (op) {
        case kVecCopy: y[i] = c * x[i]; break;
        case kVecSet: y[i] = c; break;
        case kVecAxpy: y[i] = fma(c, x[i], y[i]); break;
        case kVecMul: y[i] *= x[i]; break;
        case kVecDiv: y[i] /= x[i]; break;
        case kVecPow: y[i] = pow(y[i], c); break;
        }
    }
}

__global__ __launch_bounds__(256) void k_dot_part(const double* __restrict__ x, const double* __restrict__ y, long n,
                                                  double* __restrict__ part)
{
    __shared__ double sh[4];
    double a = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) a = fma(x[i], y[i], a);
    const double t = block_sum<256>(a, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(1024) void k_dot_final(const double* __restrict__ part, int np, double* __restrict__ out)
{
    __shared__ double sh[16];
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += 1024) a += part[i];
    const double t = block_sum<1024>(a, sh);
    if (threadIdx.x == 0) *out = t;
}

// ================================================================== launchers

int lm_tiles(int L)
{
    const int nt = (L + kLM) / kLM;
    return nt * (nt + 1) / 2;
}

int launch_pre(const MapDesc& md, const MapDesc* d_md, const double* d_x, hipStream_t s)
{
    hipLaunchKernelGGL(k_pre, dim3(lm_tiles(md.L)), dim3(256), 0, s, d_md, d_x);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

template <int K>
static int synth_k(const WorkItem* items, int n, const MapDesc* d_maps, hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    hipLaunchKernelGGL(k_synth<K>, dim3(n), dim3(64), 0, s, items, d_maps);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}
template <int K>
static int adj_k(const WorkItem* items, int n, const MapDesc* d_maps, hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    hipLaunchKernelGGL(k_adj<K>, dim3(n), dim3(64), 0, s, items, d_maps);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}
int launch_adj_mma(const WorkItem* items, const MmaSub* subs, int n, const MapDesc* d_maps, const int* d_groups,
                   hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    hipLaunchKernelGGL(k_adj_mma, dim3(n), dim3(64 * kMmaWaves), 0, s, items, subs, d_maps, d_groups);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}
int launch_synth_mma(int K, const WorkItem* items, int n, const MapDesc* d_maps, const int* d_groups, hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    if (K == 1)
        hipLaunchKernelGGL(k_synth_mma<1>, dim3(n), dim3(64), 0, s, items, d_maps, d_groups);
    else if (K == 2)
        hipLaunchKernelGGL(k_synth_mma<2>, dim3(2 * n), dim3(64), 0, s, items, d_maps, d_groups);
    else
        return fail(CMBL_ERR_ARG, "unsupported rings-per-lane K for the matrix-core synthesis");
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}
template <int K>
static int scan_k(const WorkItem* items, int n, const MapDesc* d_maps, int* d_lstart, hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    hipLaunchKernelGGL(k_scan<K>, dim3(n), dim3(64), 0, s, items, d_maps, d_lstart);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

template <int K, int G>
static int synth_g(const WorkItem* items, int n, const MapDesc* d_maps, const int* d_groups, hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    hipLaunchKernelGGL((k_synth_g<K, G>), dim3(n), dim3(64), 0, s, items, d_maps, d_groups);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_synth_group(int K, int G, const WorkItem* items, int n, const MapDesc* d_maps, const int* d_groups,
                       hipStream_t s)
{
    if (G == 4) {
        switch (K) {
        case 1: return synth_g<1, 4>(items, n, d_maps, d_groups, s);
        case 2: return synth_g<2, 4>(items, n, d_maps, d_groups, s);
        }
    } else if (G == 2) {
        switch (K) {
        case 1: return synth_g<1, 2>(items, n, d_maps, d_groups, s);
        case 2: return synth_g<2, 2>(items, n, d_maps, d_groups, s);
        case 4: return synth_g<4, 2>(items, n, d_maps, d_groups, s);
        }
    }
    return fail(CMBL_ERR_ARG, "unsupported shared-synthesis shape");
}

int launch_synth(int K, const WorkItem* items, int n, const MapDesc* d_maps, hipStream_t s)
{
    switch (K) {
    case 1: return synth_k<1>(items, n, d_maps, s);
    case 2: return synth_k<2>(items, n, d_maps, s);
    case 4: return synth_k<4>(items, n, d_maps, s);
    case 8: return synth_k<8>(items, n, d_maps, s);
    }
    return fail(CMBL_ERR_ARG, "unsupported rings-per-lane K");
}
int launch_adj(int K, const WorkItem* items, int n, const MapDesc* d_maps, hipStream_t s)
{
    switch (K) {
    case 1: return adj_k<1>(items, n, d_maps, s);
    case 2: return adj_k<2>(items, n, d_maps, s);
    case 4: return adj_k<4>(items, n, d_maps, s);
    case 8: return adj_k<8>(items, n, d_maps, s);
    }
    return fail(CMBL_ERR_ARG, "unsupported rings-per-lane K");
}
template <int K>
static int capture_k(const WorkItem* items, int n, const MapDesc* d_maps, double* prev, double* cur, int* kk,
                     hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    hipLaunchKernelGGL(k_capture<K>, dim3(n), dim3(64), 0, s, items, d_maps, prev, cur, kk);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}
int launch_capture(int K, const WorkItem* items, int n, const MapDesc* d_maps, double* prev, double* cur, int* kk,
                   hipStream_t s)
{
    switch (K) {
    case 1: return capture_k<1>(items, n, d_maps, prev, cur, kk, s);
    case 2: return capture_k<2>(items, n, d_maps, prev, cur, kk, s);
    case 4: return capture_k<4>(items, n, d_maps, prev, cur, kk, s);
    case 8: return capture_k<8>(items, n, d_maps, prev, cur, kk, s);
    }
    return fail(CMBL_ERR_ARG, "unsupported rings-per-lane K");
}

int launch_scan(int K, const WorkItem* items, int n, const MapDesc* d_maps, int* d_lstart, hipStream_t s)
{
    switch (K) {
    case 1: return scan_k<1>(items, n, d_maps, d_lstart, s);
    case 2: return scan_k<2>(items, n, d_maps, d_lstart, s);
    case 4: return scan_k<4>(items, n, d_maps, d_lstart, s);
    case 8: return scan_k<8>(items, n, d_maps, d_lstart, s);
    }
    return fail(CMBL_ERR_ARG, "unsupported rings-per-lane K");
}

// Raise a kernel's dynamic-LDS limit to kMaxLds once per device (function
// attributes belong to a device; one bit per device id in `done`).
static int allow_dynamic_lds(const void* fn, std::atomic<unsigned long long>& done)
{
    int dev = 0;
    CMBL_HIP(hipGetDevice(&dev));
    const unsigned long long bit = 1ull << (dev & 63);
    if (done.load(std::memory_order_acquire) & bit) return CMBL_OK;
    CMBL_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds));
    done.fetch_or(bit, std::memory_order_acq_rel);
    return CMBL_OK;
}

template <int MODE, int MAXM>
static int ring_mm(const RingItem* items, int n, const MapDesc* d_maps, int lds, double* out, const double* in,
                   const double2* tw, hipStream_t s)
{
    static std::atomic<unsigned long long> lds_set{0};
    CMBL_TRY(allow_dynamic_lds((const void*)k_ring<MODE, MAXM>, lds_set));
    hipLaunchKernelGGL((k_ring<MODE, MAXM>), dim3(n), dim3(ring_threads(MAXM)), lds, s, items, d_maps, out, in, tw);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

// One instantiation per transform-length class (the classes' LDS sizes):
// the bound on the transform length bounds the registers of the FFT passes.
template <int MODE>
static int ring_m(const RingItem* items, int n, const MapDesc* d_maps, int lds, double* out, const double* in,
                  const double2* tw, hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    if (lds > kMaxLds) return fail(CMBL_ERR_ARG, "ring transform longer than the LDS FFT");
    if (lds > lds_bytes(4096)) return ring_mm<MODE, 8192>(items, n, d_maps, lds, out, in, tw, s);
    if (lds > lds_bytes(2048)) return ring_mm<MODE, 4096>(items, n, d_maps, lds, out, in, tw, s);
    return ring_mm<MODE, 2048>(items, n, d_maps, lds, out, in, tw, s);
}

int launch_ring(int mode, const RingItem* items, int n, const MapDesc* d_maps, int lds, double* out,
                const double* in, const double2* tw, hipStream_t s)
{
    switch (mode) {
    case kRingLike: return ring_m<kRingLike>(items, n, d_maps, lds, out, in, tw, s);
    case kRingSyn: return ring_m<kRingSyn>(items, n, d_maps, lds, out, in, tw, s);
    case kRingAna: return ring_m<kRingAna>(items, n, d_maps, lds, out, in, tw, s);
    case kRingWeight: return ring_m<kRingWeight>(items, n, d_maps, lds, out, in, tw, s);
    }
    return fail(CMBL_ERR_ARG, "bad ring mode");
}

int launch_bluestein_setup(const MapDesc* d_md, const int* d_pairs, int npairs_bs, double2* d_out,
                           double2* d_chirp, int lds, const double2* tw, hipStream_t s)
{
    if (npairs_bs <= 0) return CMBL_OK;
    static std::atomic<unsigned long long> lds_set{0};
    CMBL_TRY(allow_dynamic_lds((const void*)k_bluestein_setup, lds_set));
    hipLaunchKernelGGL(k_bluestein_setup, dim3(npairs_bs), dim3(ring_threads(kMaxFFT)), lds, s, d_pairs, d_md, d_out, d_chirp,
                       tw);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_post(const MapDesc& md, const MapDesc* d_md, const double* d_x, double* d_g, int mode, hipStream_t s,
                double c)
{
    hipLaunchKernelGGL(k_post, dim3(lm_tiles(md.L)), dim3(256), 0, s, d_md, d_x, d_g, mode, c);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_coupling(const double* w, int L, double* s0, double* M, hipStream_t s)
{
    hipLaunchKernelGGL(k_w3j_start, dim3((L + 256) / 256), dim3(256), 0, s, L, s0);
    CMBL_HIP(hipGetLastError());
    const long np = (long)(L + 1) * (L + 2) / 2;
    const int blocks = (int)std::min<long>((np + 255) / 256, 16384);
    hipLaunchKernelGGL(k_coupling, dim3(blocks), dim3(256), 0, s, L, w, s0, M);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_pre_pol(const MapDesc& md, const MapDesc* d_md, const double* t, const double* e, const double* b,
                   double2* d_rec, hipStream_t s)
{
    hipLaunchKernelGGL(k_pre_pol, dim3(lm_tiles(md.L)), dim3(256), 0, s, d_md, t, e, b, d_rec);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

template <int K>
static int synth_pol_k(const WorkItem* items, int n, const MapDesc* d_md, const double2* rec, double* pt, double* pq,
                       double* pu, hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    hipLaunchKernelGGL(k_synth_pol<K>, dim3(n), dim3(64), 0, s, items, d_md, rec, pt, pq, pu);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_synth_pol(int K, const WorkItem* items, int n, const MapDesc* d_md, const double2* d_rec, double* d_pt,
                     double* d_pq, double* d_pu, hipStream_t s)
{
    switch (K) {
    case 1: return synth_pol_k<1>(items, n, d_md, d_rec, d_pt, d_pq, d_pu, s);
    case 2: return synth_pol_k<2>(items, n, d_md, d_rec, d_pt, d_pq, d_pu, s);
    case 4: return synth_pol_k<4>(items, n, d_md, d_rec, d_pt, d_pq, d_pu, s);
    case 8: return synth_pol_k<8>(items, n, d_md, d_rec, d_pt, d_pq, d_pu, s);
    }
    return fail(CMBL_ERR_ARG, "unsupported rings-per-lane K");
}

int launch_reorder(const double* in, double* out, long nside, int to_ring, hipStream_t s)
{
    const long npix = 12 * nside * nside;
    const int blocks = (int)std::min<long>((npix + 255) / 256, 8192);
    hipLaunchKernelGGL(k_reorder, dim3(blocks), dim3(256), 0, s, in, out, nside, to_ring);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_ldiv(double* y, const double* x, const double* dl, int L, hipStream_t s)
{
    const long n = (long)(L + 1) * (L + 1);
    const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_ldiv, dim3(blocks), dim3(256), 0, s, y, x, dl, L);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_pre_batch(const MapDesc* d_mds, const BatchArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_pre_b, dim3(a.tile0[a.n]), dim3(256), 0, s, d_mds, a);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_post_batch(const MapDesc* d_mds, const BatchArgs& a, int mode, hipStream_t s)
{
    hipLaunchKernelGGL(k_post_b, dim3(a.tile0[a.n]), dim3(256), 0, s, d_mds, a, mode);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_final_batch(const MapDesc* d_mds, const BatchArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_final_b, dim3(a.n), dim3(1024), 0, s, d_mds, a);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_final(const MapDesc* d_md, double* d_f, hipStream_t s)
{
    hipLaunchKernelGGL(k_final, dim3(1), dim3(1024), 0, s, d_md, d_f);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_vec(int op, double* y, const double* x, double c, long n, hipStream_t s)
{
    if (n <= 0) return CMBL_OK;
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_vec, dim3((unsigned)blocks), dim3(256), 0, s, op, y, x, c, n);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

int launch_dot(const double* x, const double* y, long n, double* d_out, double* d_work, hipStream_t s)
{
    const int np = kDotParts;
    hipLaunchKernelGGL(k_dot_part, dim3(np), dim3(256), 0, s, x, y, n, d_work);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(1024), 0, s, d_work, np, d_out);
    CMBL_HIP(hipGetLastError());
    return CMBL_OK;
}

}  // namespace cmbl
