// Internal definitions shared by the HIP kernels (cmbl_kernels.hip) and the
// host orchestration (cmbl_map.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "cmbl.h"

namespace cmbl {

// ------------------------------------------------------------------ layout
//
// a_lm tables are m-major: entry (l, m) lives at moff(L, m) + (l - m).
// Ring pairs: pair p = 0 .. npair-1 holds north ring p and its mirror south
// ring nring-1-p (the equator, p = 2*nside-1, has no mirror).  Pair tables
// are padded to npair_pad (a multiple of 64 * max K) with zero seeds.
// Stage buffers, four doubles per (pair, m), pair-major:
//   P = (E, O): the even / odd parity sums of the synthesis (zeroed once at
//       set-up; entries of (m, ring block) items that can never be live are
//       never written and stay zero);
//   Q = (PsiE, PsiO) = (Psi_N + Psi_S, Psi_N - Psi_S): ring transforms of the
//       weighted residual, input of the adjoint Legendre stage.

__host__ __device__ inline long moff(int L, int m) { return (long)m * (L + 1) - (long)m * (m - 1) / 2; }

// ---------------------------------------------------------------- HEALPix NEST <-> RING
// The 12 base pixels (faces) of the HEALPix sphere (Gorski et al. 2005): face
// f spans rings around jrll[f] * nside and starts at longitude index
// jpll[f] (in units of pi/4).  Inside a face, NEST numbers pixel (ix, iy) by
// interleaving the bits of ix (even positions) and iy (odd positions).
__host__ __device__ inline int face_ring(int f) { return 2 + f / 4; }                            // jrll
__host__ __device__ inline int face_phi(int f) { return f < 4 || f >= 8 ? 2 * (f & 3) + 1 : 2 * (f & 3); }  // jpll

// Morton bit spreading: bit b of x moves to bit 2b (x < 2^32).
__host__ __device__ inline unsigned long spread_bits(unsigned long x)
{
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFUL;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFUL;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0FUL;
    x = (x | (x << 2)) & 0x3333333333333333UL;
    x = (x | (x << 1)) & 0x5555555555555555UL;
    return x;
}
__host__ __device__ inline long interleave_bits(long x, long y)
{
    return (long)(spread_bits((unsigned long)x) | (spread_bits((unsigned long)y) << 1));
}
// Inverse: the even-position bits of v, packed.
__host__ __device__ inline long even_bits(long v)
{
    unsigned long x = (unsigned long)v & 0x5555555555555555UL;
    x = (x | (x >> 1)) & 0x3333333333333333UL;
    x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0FUL;
    x = (x | (x >> 4)) & 0x00FF00FF00FF00FFUL;
    x = (x | (x >> 8)) & 0x0000FFFF0000FFFFUL;
    x = (x | (x >> 16)) & 0x00000000FFFFFFFFUL;
    return (long)x;
}
__host__ __device__ inline long isqrt_hd(long v)
{
    long r = (long)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// NEST index -> RING index.  Pixel (ix, iy) of face f lies on ring
// jr = jrll nside - ix - iy - 1 (counted from the north pole, 1 .. 4 nside - 1)
// at position jp = (jpll nr + ix - iy + 1 + shift) / 2 along that ring of nr
// pixels per quarter.
__host__ __device__ inline long nest_to_ring(long nside, long ipnest)
{
    const long fpix = nside * nside, npix = 12 * fpix, ncap = 2 * nside * (nside - 1);
    const int f = (int)(ipnest / fpix);
    const long in = ipnest - f * fpix;
    const long ix = even_bits(in), iy = even_bits(in >> 1);
    const long jr = face_ring(f) * nside - ix - iy - 1;
    long nr, first, shift;
    if (jr < nside) {
        nr = jr;
        first = 2 * nr * (nr - 1);
        shift = 0;
    } else if (jr > 3 * nside) {
        nr = 4 * nside - jr;
        first = npix - 2 * nr * (nr + 1);
        shift = 0;
    } else {
        nr = nside;
        first = ncap + (jr - nside) * 4 * nside;
        shift = (jr - nside) & 1;
    }
    long jp = (face_phi(f) * nr + ix - iy + 1 + shift) / 2;
    if (jp > 4 * nr) jp -= 4 * nr;
    if (jp < 1) jp += 4 * nr;
    return first + jp - 1;
}

// RING index -> NEST index: ring number and position give the face and, by
// inverting the two relations above, (ix, iy).
__host__ __device__ inline long ring_to_nest(long nside, long ipring)
{
    const long fpix = nside * nside, npix = 12 * fpix, ncap = 2 * nside * (nside - 1);
    long jr, jp, nr, shift;
    int f;
    if (ipring < ncap) {
        jr = (1 + isqrt_hd(1 + 2 * ipring)) >> 1;
        jp = ipring - 2 * jr * (jr - 1) + 1;
        nr = jr;
        shift = 0;
        f = (int)((jp - 1) / nr);
    } else if (ipring < npix - ncap) {
        const long ip = ipring - ncap;
        jr = ip / (4 * nside) + nside;
        jp = ip % (4 * nside) + 1;
        nr = nside;
        shift = (jr - nside) & 1;
        // the two diagonals through the pixel name the face
        const long ire = jr - nside + 1, irm = 2 * nside + 2 - ire;
        const long ifm = (jp - ire / 2 + nside - 1) / nside, ifp = (jp - irm / 2 + nside - 1) / nside;
        f = (int)(ifp == ifm ? (ifp | 4) : (ifp < ifm ? ifp : ifm + 8));
    } else {
        const long ip = npix - ipring;
        const long r = (1 + isqrt_hd(2 * ip - 1)) >> 1;
        jp = 4 * r + 1 - (ip - 2 * r * (r - 1));
        nr = r;
        shift = 0;
        jr = 4 * nside - r;
        f = (int)(8 + (jp - 1) / nr);
    }
    const long irt = jr - face_ring(f) * nside + 1;
    long ipt = 2 * jp - face_phi(f) * nr - shift - 1;
    if (ipt >= 2 * nside) ipt -= 8 * nside;
    const long ix = (ipt - irt) >> 1, iy = (-ipt - irt) >> 1;
    return f * fpix + interleave_bits(ix, iy);
}

// Legendre scaling: true value = v * 2^(-1000 k).  A lane whose scale count k
// is > 0 has |true| <= 2^-100 and contributes nothing; it is rescaled when
// |v| exceeds 2^900.
constexpr double kScaleHi = 0x1p+900;
constexpr double kScaleDn = 0x1p-1000;
constexpr double kScaleUp = 0x1p+1000;
constexpr double kScaleLo = 0x1p-100;

constexpr int kWave = 64;
constexpr int kRingThreads = 512;  // threads of a ring workgroup up to M = 4096
// Threads of a ring workgroup for transforms up to maxm: 256 up to 2048, so
// that four such workgroups share a CU where two would leave most threads
// idle; 1024 at 8192, whose LDS (139 KB) allows one workgroup per CU, so that
// the CU still holds sixteen waves.
constexpr int ring_threads(int maxm) { return maxm <= 2048 ? 256 : maxm <= 4096 ? kRingThreads : 1024; }
constexpr int kMaxFFT = 8192;          // largest LDS FFT (complex doubles)
// LDS bytes of a sequence of n complex doubles in the ring stage's padded
// layout (one spare slot per 16 values).
constexpr int lds_bytes(int n) { return (n + n / 16) * 16; }
constexpr int kMaxLds = lds_bytes(kMaxFFT);  // 136 KiB
constexpr int kRedChunk = 8;           // l-steps per cross-lane reduction in the adjoint
constexpr int kRedStride = 17;         // doubles per lane row in the reduction buffer
constexpr int kChunk = 64;             // l-steps per LDS-staged coefficient chunk
constexpr int kChunkPad = 8;           // padding for speculative reads past a chunk
constexpr int kAdjChunk = 2 * kChunk;  // l-steps per staged chunk in the adjoint
constexpr int kSynChunk = 2 * kChunk;  // l-steps per staged chunk in the synthesis

struct MapDesc {
    int nside, L, npair, npair_pad, nring;
    int nrb5;                 // ring blocks of the adjoint Legendre stage
    long npix, nalm;
    // ring-pair tables [npair_pad]
    const double* ring_x;     // cos(theta) of the north ring
    const double* ring_s2;    // sin^2(theta), from 1 - z without cancellation
    const int* ring_nph;      // pixels per ring
    const int* ring_shift;    // 1 if phi0 = pi/nph, 0 if phi0 = 0
    const long* ring_ofsN;    // first pixel of the north ring
    const long* ring_ofsS;    // first pixel of the south ring, -1 for the equator
    const int* ring_M;        // FFT length used (nph if a power of two, else Bluestein length)
    const long* ring_bsofs;   // offset of the Bluestein filter in bs_filt (-1 if none)
    const long* ring_chofs;   // offset of the Bluestein chirp in bs_chirp (-1 if none)
    const double2* bs_filt;   // FFT_M of e^{+i pi j^2/b}, entries 0..M/2, per non-power-of-two ring
    const double2* bs_chirp;  // e^{-i pi j^2/b}, j <= b/2, per non-power-of-two ring
    // Rings whose Bluestein length for nph would exceed kMaxFFT are split into
    // two transforms of b = nph/2 (even and odd samples); the others use b = nph.
    const long* ring_scrofs;  // offset of the split ring's scratch in ring_scr (-1 if not split)
    double2* ring_scr;        // per split ring: 3 nph/2 entries (one half-transform, nph samples)
    // Legendre tables
    const double* seed_v;     // [(L+1) * npair_pad] scaled lambda_mm
    const int* seed_k;        // [(L+1) * npair_pad] scale count
    // recurrence state at each work item's first live l, per stage layout:
    // (mu_{l-1}, mu_l) scaled, and the scale count; [(L+1) * npair_pad]
    const double* s1_prev;
    const double* s1_cur;
    const int* s1_k;
    const double* s5_prev;
    const double* s5_cur;
    const int* s5_k;
    const double* A;          // [nalm] recurrence coefficient of mu_l = A x mu_{l-1} - mu_{l-2}
    const double* Nrm;        // [nalm] lambda_lm = Nrm * mu_lm
    const double* beam;       // [L+1]
    const double* invcl;      // [L+1] 1/C_l or 0
    // data
    const double* d;          // [npix]
    const double* w;          // [npix]
    // per-evaluation workspace
    double2* alm;             // [nalm] B_l N_lm a_lm, m-major
    double* P;                // [npair * (L+1) * 4] synthesis parity sums
    double* Q;                // [npair * (L+1) * 4] adjoint ring transforms
    double* upart;            // [nrb5 * nalm * 2]
    double* chi;              // [npair] per-pair chi^2
    double* prior;            // [nprior_blocks]
    int nprior;
    int* nonfinite;           // set to 1 by the -2lnL reduction when the value is NaN/Inf
                              // (host-mapped memory, cleared by the host)
};

// One Legendre work item: degree m of one ring block of one map, starting at
// the first l at which any of its ring pairs is live.  mcnt = 0 marks padding.
struct WorkItem {
    int map;
    int m0;
    int mcnt;
    int rb;
    int lfirst;
    int lend;  // last l of the item (the adjoint's segments end before L)
    int sidx;  // start-state slot of an adjoint segment (-1: indexed by m)
    int pad_;
};

// Ring items of one launch: consecutive entries of the ring list that share
// an LDS size (the largest transform length of the class).
struct RingClass {
    int first, count, lds;
};

struct ItemCost {
    WorkItem it;
    double cost;
};

// A ring pair of a map; with pair2 >= 0 also a second ring pair (of map2) of
// the same power-of-two length, half the launch's largest, in the same
// workgroup.
struct RingItem {
    int map;
    int pair;
    int map2 = -1;
    int pair2 = -1;
};

// kRingLike: s -> f = w (d - s) and chi partials -> analysis;  kRingSyn: s -> map;
// kRingAna: map -> analysis;  kRingWeight: s -> f = w s -> analysis (Gibbs operator).
enum RingMode { kRingLike = 0, kRingSyn = 1, kRingAna = 2, kRingWeight = 3 };

// k_post output modes (u = B_l N_lm sum of the ring-block partials, complex):
//   kPostGrad    g = J^T v  (real-parameter derivative: 2u for m >= 1)
//   kPostLike    g = -2 J^T (w r) + prior gradient
//   kPostGibbs   y = x + u                 (I + D Y^T N^-1 Y D, HEALPix complex convention)
//   kPostAlm     y = u                     (packed D Y^T v, complex convention)
//   kPostAddAlm  y += c u / B_l            (packed c Y^T v without the l-scaling)
//   kPostQuad    y = c u / B_l             (quadrature analysis, c = 4 pi / Npix)
enum PostMode { kPostGrad = 0, kPostLike = 1, kPostGibbs = 2, kPostAlm = 3, kPostAddAlm = 4, kPostQuad = 5 };

// One likelihood + gradient evaluation of a map on stream s (device x, f, g);
// marks = false records no stage-timing events (for HIP graph capture).
int eval_device(::cmbl_map* mp, const double* x, double* f, double* g, hipStream_t s, bool marks = true);

// Thread-local error string.
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define CMBL_TRY(expr)                \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != CMBL_OK) return rc_; \
    } while (0)

#define CMBL_HIP(call)                                  \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return ::cmbl::hip_fail(e_, #call); \
    } while (0)

// ------------------------------------------------------------------ launchers
// (defined in cmbl_kernels.hip)
int launch_pre(const MapDesc& md, const MapDesc* d_md, const double* d_x, hipStream_t s);
int launch_synth(int K, const WorkItem* items, int nitems, const MapDesc* d_maps, hipStream_t s);
// Shared-recurrence synthesis of groups of G maps of one geometry (G = 2, 4).
int launch_synth_group(int K, int G, const WorkItem* items, int nitems, const MapDesc* d_maps, const int* d_groups,
                       hipStream_t s);
int launch_adj(int K, const WorkItem* items, int nitems, const MapDesc* d_maps, hipStream_t s);
// Shared adjoint of groups of kAdjG maps of one geometry on the FP64 matrix
// cores: one ring pair per lane, a workgroup of kMmaWaves waves on
// consecutive 64-pair blocks of one m; item.sidx indexes `subs` (each wave's
// first live l and start-state slot); item.rb counts blocks of kMmaWaves * 64
// pairs, whose partials the maps' descriptors address (nrb5 of them).
constexpr int kAdjG = 8;
constexpr int kAdjMinG = 3;  // fewest maps of a group (the rest of its columns empty, entry -1)
constexpr int kSynMinG = 6;  // fewest maps of a matrix-core synthesis group (fewer share k_synth_g)
constexpr int kMmaWaves = 4;
struct MmaSub {
    int lf[kMmaWaves];    // first live l of the wave's block, -1 if it has none
    int sidx[kMmaWaves];  // its start-state slot
};
// Shared synthesis of groups of up to kAdjG maps of one geometry (ring
// pairs per lane K = 1 or 2) on the FP64 matrix cores; d_groups: kAdjG map
// indices per group (-1: empty).
int launch_synth_mma(int K, const WorkItem* items, int nitems, const MapDesc* d_maps, const int* d_groups,
                     hipStream_t s);
int launch_adj_mma(const WorkItem* items, const MmaSub* subs, int nitems, const MapDesc* d_maps, const int* d_groups,
                   hipStream_t s);
int lm_tiles(int L);  // blocks of k_pre / k_post (prior partial sums)
int launch_ldiv(double* y, const double* x, const double* dl, int L, hipStream_t s);
// spin-2 synthesis: kPolRec complex coefficients per (l, m); parity sums of
// T, Q and U in three buffers laid out as MapDesc::P
constexpr int kPolRec = 5;
int launch_pre_pol(const MapDesc& md, const MapDesc* d_md, const double* t, const double* e, const double* b,
                   double2* d_rec, hipStream_t s);
int launch_synth_pol(int K, const WorkItem* items, int nitems, const MapDesc* d_md, const double2* d_rec,
                     double* d_pt, double* d_pq, double* d_pu, hipStream_t s);
int launch_reorder(const double* in, double* out, long nside, int to_ring, hipStream_t s);
int launch_coupling(const double* w, int L, double* s0, double* M, hipStream_t s);
int launch_ring(int mode, const RingItem* items, int nitems, const MapDesc* d_maps, int lds_bytes,
                double* d_map_out, const double* d_map_in, const double2* d_tw, hipStream_t s);
int launch_post(const MapDesc& md, const MapDesc* d_md, const double* d_x, double* d_g, int mode,
                hipStream_t s, double c = 0.0);
int launch_final(const MapDesc* d_md, double* d_f, hipStream_t s);

// One launch of the a_lm preparation, the gradient assembly or the -2lnL
// reduction over up to kBatchGroup maps of a batch (kernel argument, by
// value): map i of the group owns blocks tile0[i] .. tile0[i+1]-1.
constexpr int kBatchGroup = 64;
struct BatchArgs {
    const double* x[kBatchGroup];
    double* g[kBatchGroup];
    double* f;  // f[i]: -2lnL of map i of the group
    int tile0[kBatchGroup + 1];
    int n;
};
int launch_pre_batch(const MapDesc* d_mds, const BatchArgs& a, hipStream_t s);
int launch_post_batch(const MapDesc* d_mds, const BatchArgs& a, int mode, hipStream_t s);
int launch_final_batch(const MapDesc* d_mds, const BatchArgs& a, hipStream_t s);
int launch_bluestein_setup(const MapDesc* d_md, const int* d_pairs, int npairs_bs, double2* d_out,
                           double2* d_chirp, int lds_bytes, const double2* d_tw, hipStream_t s);
int launch_scan(int K, const WorkItem* items, int nitems, const MapDesc* d_maps, int* d_lstart,
                hipStream_t s);
int launch_capture(int K, const WorkItem* items, int nitems, const MapDesc* d_maps, double* d_prev, double* d_cur,
                   int* d_k, hipStream_t s);

// BLAS-1.  The dot product runs kDotParts blocks of 256 threads (grid-stride
// FMA accumulation, block sums) and one final block over the kDotParts
// partial sums; the fused loop kernels replicate this order exactly.
constexpr int kDotParts = 1024;
enum { kVecCopy = 0, kVecSet, kVecAxpy, kVecMul, kVecDiv, kVecPow };
int launch_vec(int op, double* y, const double* x, double c, long n, hipStream_t s);
int launch_dot(const double* x, const double* y, long n, double* d_out, double* d_work, hipStream_t s);

// Device-resident CG state of a Gibbs system (defined with the loops).
struct GibbsCg;
void gibbs_cg_free(GibbsCg* cg);

}  // namespace cmbl

// Opaque handles of the C ABI.
struct cmbl_map {
    int nside = 0, L = 0, npair = 0, npair_pad = 0, nring = 0;
    int K1 = 1, K5 = 1, nrb1 = 0, nrb5 = 0;
    cmbl_tuning tuning{1, 1, 4096};
    long npix = 0, nalm = 0, nparam = 0;
    int device = 0;
    int lds_ring = 0;
    bool has_beam = false, has_cl = false, has_data = false;
    long generation = 0;            // incremented by every cmbl_map_set_*
    int* h_nonfinite = nullptr;     // host-mapped flag written by the -2lnL reduction
    std::vector<void*> allocs;
    cmbl::MapDesc desc{};
    cmbl::MapDesc* d_desc = nullptr;
    cmbl::WorkItem* d_items1 = nullptr;
    cmbl::WorkItem* d_items5 = nullptr;
    cmbl::RingItem* d_ring_items = nullptr;
    int nitems1 = 0, nitems5 = 0, nring_items = 0;
    std::vector<cmbl::RingClass> ring_classes;
    std::vector<int> ring_M;  // transform length per pair
    std::vector<int> ring_two;  // n for power-of-two unsplit pairs with n > lmax (two may share a workgroup), else 0
    std::vector<cmbl::WorkItem> items1, items5;   // launch order (XCD-interleaved)
    std::vector<cmbl::ItemCost> cost1, cost5;      // live items with cost estimates
    std::vector<cmbl::RingItem> ring_items;
    double2* d_tw = nullptr;
    double* d_f = nullptr;
    double* d_x = nullptr;  // staging for the host entry points
    double* d_g = nullptr;
    double* d_map = nullptr;
    // spin-2 synthesis (allocated at its first call): coefficient records,
    // parity sums of T, Q, U and one descriptor per sum for the ring stage
    cmbl::MapDesc* d_desc_pol = nullptr;  // [3]
    double2* d_pol_rec = nullptr;         // [kPolRec * nalm]
    double* d_pol_P = nullptr;            // [3 * npair * (lmax + 1) * 4]
    double* d_pol_in = nullptr;   // [3 * nparam] host-call staging
    double* d_pol_out = nullptr;  // [3 * npix]
    // mutable views of the workspace / data owned by this map
    double* d_beam = nullptr;
    double* d_invcl = nullptr;
    double* d_data = nullptr;
    double* d_w = nullptr;
    // stage timing: a pool of event sets recorded without host synchronisation
    bool timing = false;
    std::vector<hipEvent_t> ev;  // kEvSets * (CMBL_NSTAGES + 1)
    int ev_next = 0, ev_pending = 0;
    double stage_ms[CMBL_NSTAGES] = {0, 0, 0, 0, 0, 0};
    long timed_evals = 0;
    static constexpr int kEvSets = 64;
};

// Gibbs constrained-realisation system: its own map whose "beam" is
// D_l = B_l sqrt(C_l) and whose weights are mask/sigma^2.
struct cmbl_gibbs {
    cmbl_map* mp = nullptr;
    int L = 0;
    long n = 0, npix = 0, good = 0;
    double sigma = 0.0;
    std::vector<double> beam;  // B_l
    double* d_dl = nullptr;  // [L+1] preconditioner divisors 1 + C_l B_l^2 / N_l
    double *d_r = nullptr, *d_z = nullptr, *d_p = nullptr, *d_ap = nullptr;  // [n] CG vectors
    double *d_b = nullptr, *d_x = nullptr;                                   // [n] host-call staging
    double *d_m0 = nullptr, *d_m1 = nullptr, *d_m2 = nullptr;               // [npix] maps
    cmbl::GibbsCg* cg = nullptr;  // device-resident CG, made at the first solve
    std::vector<void*> allocs;
};

struct cmbl_batch {
    std::vector<cmbl_map*> maps;
    cmbl::MapDesc* d_descs = nullptr;
    // work lists grouped by K (index = log2 K)
    cmbl::WorkItem* d_items1[4] = {nullptr, nullptr, nullptr, nullptr};
    cmbl::WorkItem* d_items5[4] = {nullptr, nullptr, nullptr, nullptr};
    int n1[4] = {0, 0, 0, 0}, n5[4] = {0, 0, 0, 0};
    // shared-recurrence synthesis: groups of G maps of one geometry, their
    // items by (K, G) (index log2 K, G = 2 or 4) and the group table
    struct SynGroupList {
        int K = 0, G = 0, n = 0;
        cmbl::WorkItem* d_items = nullptr;
    };
    std::vector<SynGroupList> syn_groups;
    int* d_gmaps = nullptr;
    int ngroups = 0;
    // shared adjoint on the matrix cores: groups of kAdjG maps of one
    // geometry with one ring pair per lane, their items and group table
    // matrix-core synthesis: groups of up to kAdjG maps of one geometry, items by K - 1
    cmbl::WorkItem* d_items_smma[2] = {nullptr, nullptr};
    int n_smma[2] = {0, 0};
    int* d_smaps = nullptr;
    cmbl::WorkItem* d_items_mma = nullptr;
    cmbl::MmaSub* d_subs_mma = nullptr;
    int n_mma = 0;
    int* d_amaps = nullptr;
    cmbl::RingItem* d_ring_items = nullptr;
    int nring_items = 0;
    int lds_ring = 0;
    std::vector<cmbl::RingClass> ring_classes;
    double2* d_tw = nullptr;
    std::vector<void*> allocs;
};
