// Host side of the C ABI (include/cmbl.h): map contexts, set-up tables,
// evaluation orchestration, batches, inputs and device-vector BLAS-1.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <random>
#include <tuple>
#include <string>
#include <vector>

#include "cmbl_internal.hpp"

namespace cmbl {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}
int hip_fail(hipError_t e, const char* what)
{
    g_err = std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ") in " + what;
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? CMBL_ERR_ALLOC : CMBL_ERR_HIP;
}

static const long double kPiL = 3.141592653589793238462643383279502884L;

// ------------------------------------------------------------------ geometry
// HEALPix RING scheme (Gorski et al. 2005); pair p = north ring p+1.
struct RingGeom {
    double z, sint, s2;
    int nph;
    int shift;  // phi0 = pi/nph
    long ofs;
};

static RingGeom ring_geom(int nside, long i /* 1-based ring */)
{
    const long N = nside;
    const long npix = 12 * N * N;
    const long in = (i <= 2 * N) ? i : 4 * N - i;
    RingGeom g{};
    if (in < N) {
        const double t = (double)(in * in) / (3.0 * (double)N * (double)N);
        g.z = 1.0 - t;
        g.s2 = t * (2.0 - t);
        g.sint = std::sqrt(g.s2);
        g.nph = (int)(4 * in);
        g.shift = 1;
        g.ofs = 2 * in * (in - 1);
    } else {
        g.z = (double)(2 * N - in) * 2.0 / (3.0 * (double)N);
        g.s2 = (1.0 - g.z) * (1.0 + g.z);
        g.sint = std::sqrt(g.s2);
        g.nph = (int)(4 * N);
        g.shift = ((in - N) % 2 == 0) ? 1 : 0;
        g.ofs = 2 * N * (N - 1) + (in - N) * 4 * N;
    }
    if (i > 2 * N) {
        g.z = -g.z;
        g.ofs = npix - g.ofs - g.nph;
    }
    return g;
}

static long isqrt_l(long v)
{
    long r = (long)std::sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

static void pix2ang(int nside, long pix, double* theta, double* phi)
{
    const long N = nside;
    const long npix = 12 * N * N;
    const long ncap = 2 * N * (N - 1);
    const double halfpi = 0.5 * 3.141592653589793;
    if (pix < ncap) {
        const long iring = (1 + isqrt_l(1 + 2 * pix)) >> 1;
        const long iphi = (pix + 1) - 2 * iring * (iring - 1);
        const double z = 1.0 - (double)(iring * iring) / (3.0 * (double)N * (double)N);
        *theta = std::acos(z);
        *phi = ((double)iphi - 0.5) * halfpi / (double)iring;
    } else if (pix < npix - ncap) {
        const long ip = pix - ncap;
        const long iring = ip / (4 * N) + N;
        const long iphi = ip % (4 * N) + 1;
        const double fodd = ((iring + N) & 1) ? 1.0 : 0.5;
        const double z = (double)(2 * N - iring) * 2.0 / (3.0 * (double)N);
        *theta = std::acos(z);
        *phi = ((double)iphi - fodd) * halfpi / (double)N;
    } else {
        const long ip = npix - pix;
        const long iring = (1 + isqrt_l(2 * ip - 1)) >> 1;
        const long iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
        const double z = -1.0 + (double)(iring * iring) / (3.0 * (double)N * (double)N);
        *theta = std::acos(z);
        *phi = ((double)iphi - 0.5) * halfpi / (double)iring;
    }
}

static bool is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }

// ------------------------------------------------------------------ allocation helpers
template <typename T>
static int dalloc(cmbl_map* mp, T** p, size_t count)
{
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    CMBL_HIP(hipMalloc(&q, bytes));
    mp->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return CMBL_OK;
}
template <typename T>
static int dupload(cmbl_map* mp, T** p, const std::vector<T>& h)
{
    int rc = dalloc(mp, p, h.size());
    if (rc) return rc;
    CMBL_HIP(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return CMBL_OK;
}



// Work lists for a Legendre stage with K rings per lane.  One item per live
// (m, ring block); lfirst is the first l at which any of its pairs is live.
// The cost estimate counts VALU work per lane: 3 per scaled step, 4 per live
// step, K rings.
static const int kXcds = 8;

static void build_items(const cmbl_map* mp, int K, const std::vector<int>& lstart, int map_index,
                        std::vector<ItemCost>* out)
{
    const int L = mp->L;
    const int npad = mp->npair_pad;
    const int nrb = npad / (64 * K);
    for (int rb = 0; rb < nrb; ++rb) {
        for (int m = 0; m <= L; ++m) {
            int lmin = L + 1, lmax = -1;
            for (int j = 0; j < 64 * K; ++j) {
                const int v = lstart[(size_t)m * npad + rb * 64 * K + j];
                if (v <= L) {
                    lmin = std::min(lmin, v);
                    lmax = std::max(lmax, v);
                }
            }
            if (lmin > L) continue;
            const double cost = K * (3.0 * (lmax - lmin) + 4.0 * (L - lmax + 1)) + 64.0 * K;
            out->push_back(ItemCost{WorkItem{map_index, m, 1, rb, lmin, L, -1, 0}, cost});
        }
    }
}

// Adjoint items cut along l into segments of at most S steps, each starting
// from its own captured recurrence state (slot sidx): the adjoint's outputs
// are per l, so segments of one (m, ring block) write disjoint parts of its
// row, and the long low-m items no longer set the launch's tail.
static int adj_segment_len(int L) { return std::max(256, ((L / 4 + 127) / 128) * 128); }
// Segments only where the adjoint has few items: below this count a launch
// is a few waves per SIMD (three fit at once) and its longest items set its
// length; at nside 512 / lmax 1024 (4,100 items) segments took the adjoint
// from 0.153 to 0.122 ms, at nside 1024 / lmax 2048 (16,000) they changed
// nothing and at nside 2048 (65,000) cost up to 4%.
static const size_t kSegmentBelow = 12288;

static void segment_items(std::vector<ItemCost>* v, int L, int S, int K, std::vector<ItemCost>* whole)
{
    std::vector<ItemCost> out;
    whole->clear();
    for (const ItemCost& c : *v) {
        whole->push_back(c);
        whole->back().it.sidx = (int)out.size();  // the first segment's start state is the item's
        const double len = (double)(L - c.it.lfirst + 1);
        for (int l0 = c.it.lfirst; l0 <= L; l0 += S) {
            ItemCost sg = c;
            sg.it.lfirst = l0;
            sg.it.lend = std::min(L, l0 + S - 1);
            sg.it.sidx = (int)out.size();
            sg.cost = (c.cost - 64.0 * K) * (double)(sg.it.lend - l0 + 1) / len + 64.0 * K;
            out.push_back(sg);
        }
    }
    *v = out;
}

// XCD-aware order.  Workgroup i runs on XCD i % 8, so the list is built as
// eight interleaved per-XCD lists.  Items of four consecutive m of one map
// (which share the 128-byte lines of the pair-major stage buffers) go to the
// same XCD; groups are spread over the XCDs by total cost; each XCD list is
// ordered by cost, largest first; short lists are padded with empty items.
static std::vector<WorkItem> order_items(std::vector<ItemCost>& v)
{
    std::vector<WorkItem> out;
    if (v.empty()) return out;
    std::vector<std::pair<long, double>> groups;  // (key, cost)
    std::vector<long> keys(v.size());
    {
        std::vector<std::pair<long, size_t>> kv;
        for (size_t i = 0; i < v.size(); ++i) {
            keys[i] = ((long)v[i].it.map << 32) | (long)(v[i].it.m0 >> 2);
            kv.push_back({keys[i], i});
        }
        std::sort(kv.begin(), kv.end());
        for (size_t i = 0; i < kv.size(); ++i) {
            if (groups.empty() || groups.back().first != kv[i].first) groups.push_back({kv[i].first, 0.0});
            groups.back().second += v[kv[i].second].cost;
        }
    }
    std::sort(groups.begin(), groups.end(),
              [](const std::pair<long, double>& a, const std::pair<long, double>& b) { return a.second > b.second; });
    std::vector<double> load(kXcds, 0.0);
    std::vector<std::pair<long, int>> assign;
    for (const auto& g : groups) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += g.second;
        assign.push_back({g.first, x});
    }
    std::sort(assign.begin(), assign.end());
    std::vector<std::vector<ItemCost>> lists(kXcds);
    for (size_t i = 0; i < v.size(); ++i) {
        const auto itx = std::lower_bound(assign.begin(), assign.end(), std::make_pair(keys[i], -1));
        lists[itx->second].push_back(v[i]);
    }
    size_t len = 0;
    for (auto& l : lists) {
        std::stable_sort(l.begin(), l.end(), [](const ItemCost& a, const ItemCost& b) { return a.cost > b.cost; });
        len = std::max(len, l.size());
    }
    out.resize(len * kXcds, WorkItem{0, 0, 0, 0, 0, 0, -1, 0});
    for (int x = 0; x < kXcds; ++x)
        for (size_t j = 0; j < lists[x].size(); ++j) out[j * kXcds + x] = lists[x][j].it;
    return out;
}

// Ring items ordered by transform length, largest first, and cut into at
// most two launches by LDS size: M > 4096 (139 KB, one workgroup per CU) and
// the rest.  Up to M = 4096 the ring kernels are held by their registers to
// two workgroups per CU whatever their LDS, so one launch carries every ring
// up to 4096 and the small ones fill its tail instead of running as launches
// of a few hundred workgroups.  In that launch, power-of-two rings of half its
// largest length (Two(item) == n, n > lmax, unsplit) go two to a workgroup:
// alone they would leave half of its threads idle in every FFT pass.  The
// pairs lead their launch (each weighs about one ring of the full length).
template <typename MOf, typename TwoOf>
static std::vector<RingClass> ring_classes(std::vector<RingItem>* items, MOf Mof, TwoOf Two)
{
    std::stable_sort(items->begin(), items->end(),
                     [&](const RingItem& a, const RingItem& b) { return Mof(a) > Mof(b); });
    auto cls = [](int M) { return M > 4096 ? 0 : 1; };
    std::vector<RingItem> out_items;
    std::vector<RingClass> out;
    for (size_t i = 0; i < items->size();) {
        const int Mfirst = Mof((*items)[i]);
        size_t j = i;
        while (j < items->size() && cls(Mof((*items)[j])) == cls(Mfirst)) ++j;
        // the launch's kernel template: 2048 or 4096 for the class up to 4096
        const int maxm = Mfirst > 4096 ? 0 : (Mfirst > 2048 ? 4096 : 2048);
        const int half = (maxm > 0 && Mfirst == maxm) ? maxm / 2 : 0;
        const size_t start = out_items.size();
        std::vector<RingItem> rest;
        int pending = -1;
        for (size_t k = i; k < j; ++k) {
            const RingItem& r = (*items)[k];
            if (half > 0 && Two(r) == half) {
                if (pending < 0) {
                    pending = (int)k;
                } else {
                    RingItem b = (*items)[pending];
                    b.map2 = r.map;
                    b.pair2 = r.pair;
                    out_items.push_back(b);
                    pending = -1;
                }
            } else {
                rest.push_back(r);
            }
        }
        if (pending >= 0) rest.insert(rest.begin(), (*items)[pending]);
        out_items.insert(out_items.end(), rest.begin(), rest.end());
        // at least 2 * kRingThreads entries: the fold's partial sums live in LDS
        out.push_back(RingClass{(int)start, (int)(out_items.size() - start), lds_bytes(std::max(Mfirst, 2 * kRingThreads))});
        i = j;
    }
    *items = out_items;
    return out;
}

static int run_ring(int mode, const RingItem* d_items, const std::vector<RingClass>& classes, const MapDesc* d_maps,
                    double* out, const double* in, const double2* tw, hipStream_t s)
{
    for (const RingClass& c : classes)
        CMBL_TRY(launch_ring(mode, d_items + c.first, c.count, d_maps, c.lds, out, in, tw, s));
    return CMBL_OK;
}

static int choose_k(int npair, int kmax)
{
    int k = 1;
    while (k < kmax && npair >= 64 * k * 2 * 4) k *= 2;  // keep >= 4 ring blocks per m
    return k;
}

// ------------------------------------------------------------------ launch-shape tuning
// Per geometry: ring pairs per lane of the Legendre synthesis (k_synth) and
// adjoint (k_adj), and the length above which Bluestein rings are split in
// two half transforms.  Chosen by cmbl_tune (timed candidates) and kept in a
// table keyed by (device architecture, nside, lmax); cmbl_map_create takes a
// geometry's entry, else the rule of choose_k.  The table file next to the
// library, tuning_<arch>.txt, is read at the first map creation (or the file
// named by CMBL_TUNING_FILE).
struct TuneEntry {
    std::string arch;
    int nside, lmax;
    cmbl_tuning t;
};
static std::mutex g_tune_mu;
static std::vector<TuneEntry> g_tune;
static bool g_tune_loaded = false;

static bool tuning_valid(const cmbl_tuning& t)
{
    auto k_ok = [](int k) { return k == 1 || k == 2 || k == 4 || k == 8; };
    return k_ok(t.k_synth) && k_ok(t.k_adj) && (t.split_m == 2048 || t.split_m == 4096 || t.split_m == 8192);
}

static std::string device_arch()
{
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    std::string a = prop.gcnArchName;
    return a.substr(0, a.find(':'));
}

static int tuning_read(const std::string& path, bool must_exist)
{
    std::ifstream in(path);
    if (!in) return must_exist ? fail(CMBL_ERR_ARG, "cannot open tuning file " + path) : CMBL_OK;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ls(line);
        TuneEntry e;
        if (!(ls >> e.arch >> e.nside >> e.lmax >> e.t.k_synth >> e.t.k_adj >> e.t.split_m) || !tuning_valid(e.t))
            return fail(CMBL_ERR_ARG, "bad line in tuning file " + path + ": " + line);
        bool replaced = false;
        for (TuneEntry& o : g_tune)
            if (o.arch == e.arch && o.nside == e.nside && o.lmax == e.lmax) {
                o.t = e.t;
                replaced = true;
            }
        if (!replaced) g_tune.push_back(e);
    }
    return CMBL_OK;
}

static void tuning_load_default(const std::string& arch)
{
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    if (const char* f = std::getenv("CMBL_TUNING_FILE")) {
        (void)tuning_read(f, false);
        return;
    }
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&tuning_load_default), &info) && info.dli_fname) {
        std::string dir = info.dli_fname;
        const size_t slash = dir.rfind('/');
        dir = slash == std::string::npos ? "." : dir.substr(0, slash);
        (void)tuning_read(dir + "/tuning_" + arch + ".txt", false);
    }
}

// The tuning a new map of this geometry gets: the table's entry, else the rule.
static cmbl_tuning default_tuning(int nside, int lmax)
{
    const std::string arch = device_arch();
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        tuning_load_default(arch);
        for (const TuneEntry& e : g_tune)
            if (e.arch == arch && e.nside == nside && e.lmax == lmax) return e.t;
    }
    // synthesis: no cross-lane reduction, so small ring blocks (less mixed
    // live/negligible work per wave); adjoint: four rings per lane amortise
    // the per-l lane reduction
    return cmbl_tuning{choose_k(2 * nside, 2), choose_k(2 * nside, 4), 4096};
}

static int map_create(cmbl_map* mp, int nside, int lmax, const cmbl_tuning& tune)
{
    const int N = nside;
    const int L = lmax;
    mp->nside = N;
    mp->L = L;
    mp->npix = 12L * N * N;
    mp->nring = 4 * N - 1;
    mp->npair = 2 * N;
    mp->nalm = (long)(L + 1) * (L + 2) / 2;
    mp->nparam = (long)(L + 1) * (L + 1);
    mp->tuning = tune;
    CMBL_HIP(hipGetDevice(&mp->device));
    mp->K1 = tune.k_synth;
    mp->K5 = tune.k_adj;
    const int kmax = std::max(mp->K1, mp->K5);
    mp->npair_pad = ((mp->npair + 64 * kmax - 1) / (64 * kmax)) * (64 * kmax);
    mp->nrb1 = mp->npair_pad / (64 * mp->K1);
    mp->nrb5 = mp->npair_pad / (64 * mp->K5);
    const int npad = mp->npair_pad;

    // Bluestein rings longer than split_limit are split into two half-length
    // transforms: above 4096 a ring's LDS (139 KB at 8192) allows one
    // workgroup per CU, while two halves of 4096 run two per CU
    const int split_limit = tune.split_m;

    // ---- ring-pair tables
    std::vector<double> rx(npad, 0.0), rs(npad, 0.0), rs2(npad, 1.0);
    std::vector<int> nph(npad, 0), shift(npad, 0), Mfft(npad, 0);
    std::vector<long> ofsN(npad, 0), ofsS(npad, -1), bsofs(npad, -1), chofs(npad, -1), scrofs(npad, -1);
    std::vector<int> bs_pairs;
    long bs_total = 0, ch_total = 0, scr_total = 0;
    int maxM = 1;
    for (int p = 0; p < mp->npair; ++p) {
        const RingGeom g = ring_geom(N, p + 1);
        rx[p] = g.z;
        rs[p] = g.sint;
        rs2[p] = g.s2;
        nph[p] = g.nph;
        shift[p] = g.shift;
        ofsN[p] = g.ofs;
        ofsS[p] = (p + 1 == 2 * N) ? -1 : ring_geom(N, 4L * N - (p + 1)).ofs;
        if (is_pow2(g.nph)) {
            Mfft[p] = g.nph;
        } else {
            // Bluestein length b (nph, or nph/2 for a split ring; nph = 4i is even)
            int b = g.nph, M = 1;
            while (M < 2 * b - 1) M <<= 1;
            if (M > split_limit) {
                b = g.nph / 2;
                M = 1;
                while (M < 2 * b - 1) M <<= 1;
                scrofs[p] = scr_total;
                scr_total += 3L * b;
            }
            Mfft[p] = M;
            // half tables (the chirp and the filter are symmetric)
            bsofs[p] = bs_total;
            bs_total += M / 2 + 1;
            chofs[p] = ch_total;
            ch_total += b / 2 + 1;
            bs_pairs.push_back(p);
        }
        maxM = std::max(maxM, Mfft[p]);
    }
    if (maxM > kMaxFFT)
        return fail(CMBL_ERR_ARG, "ring transform length " + std::to_string(maxM) + " exceeds the LDS limit " +
                                      std::to_string(kMaxFFT) + " (nside > 2048 is not supported)");
    mp->lds_ring = lds_bytes(std::max(maxM, 2 * kRingThreads));

    // ---- recurrence tables (long double on the host, rounded once)
    std::vector<double> A(mp->nalm, 0.0), Nrm(mp->nalm, 1.0);
    for (int m = 0; m <= L; ++m) {
        const long b = moff(L, m) - m;
        long double Nm2 = 1.0L, Nm1 = 1.0L;  // N_{l-2}, N_{l-1}
        long double alpha_prev = 0.0L;
        for (int l = m; l <= L; ++l) {
            if (l == m) {
                Nrm[b + l] = 1.0;
                continue;
            }
            const long double ll = l, mm = m;
            const long double alpha = std::sqrt((4.0L * ll * ll - 1.0L) / (ll * ll - mm * mm));
            if (l == m + 1) {
                A[b + l] = (double)alpha;
                Nrm[b + l] = 1.0;
                Nm2 = 1.0L;
                Nm1 = 1.0L;
            } else {
                const long double Nl = alpha / alpha_prev * Nm2;
                A[b + l] = (double)(alpha * Nm1 / Nl);
                Nrm[b + l] = (double)Nl;
                Nm2 = Nm1;
                Nm1 = Nl;
            }
            alpha_prev = alpha;
        }
    }

    // ---- seeds lambda_mm(theta_p) in scaled form, m-major [m][p]
    std::vector<double> seed_v((size_t)(L + 1) * npad, 0.0);
    std::vector<int> seed_k((size_t)(L + 1) * npad, 0);
    const long double up = ldexpl(1.0L, 1000), lo = ldexpl(1.0L, -100);
    for (int p = 0; p < mp->npair; ++p) {
        long double r = 1.0L / std::sqrt(4.0L * kPiL);
        int k = 0;
        const long double s = rs[p];
        for (int m = 0; m <= L; ++m) {
            if (m > 0) {
                r *= -std::sqrt((2.0L * m - 1.0L) / (2.0L * m)) * s;
                while (r != 0.0L && fabsl(r) <= lo) {
                    r *= up;
                    ++k;
                }
            }
            seed_v[(size_t)m * npad + p] = (double)(r * std::sqrt(2.0L * m + 1.0L));
            seed_k[(size_t)m * npad + p] = k;
        }
    }

    // ---- twiddles e^{-2 pi i k / kMaxFFT}
    std::vector<double2> tw(kMaxFFT / 2);
    for (int k = 0; k < kMaxFFT / 2; ++k) {
        const long double a = 2.0L * kPiL * k / kMaxFFT;
        tw[k] = make_double2((double)cosl(a), (double)-sinl(a));
    }

    // ---- device uploads
    MapDesc& d = mp->desc;
    std::memset(&d, 0, sizeof(d));
    d.nside = N;
    d.L = L;
    d.npair = mp->npair;
    d.npair_pad = npad;
    d.nring = mp->nring;
    d.nrb5 = mp->nrb5;
    d.npix = mp->npix;
    d.nalm = mp->nalm;
    double *p_rx, *p_A, *p_N, *p_sv;
    int *p_nph, *p_sh, *p_M, *p_sk;
    long *p_oN, *p_oS, *p_bo;
    CMBL_TRY(dupload(mp, &p_rx, rx));
    double* p_rs2;
    CMBL_TRY(dupload(mp, &p_rs2, rs2));
    d.ring_s2 = p_rs2;
    CMBL_TRY(dupload(mp, &p_nph, nph));
    CMBL_TRY(dupload(mp, &p_sh, shift));
    CMBL_TRY(dupload(mp, &p_oN, ofsN));
    CMBL_TRY(dupload(mp, &p_oS, ofsS));
    CMBL_TRY(dupload(mp, &p_M, Mfft));
    CMBL_TRY(dupload(mp, &p_bo, bsofs));
    long* p_co;
    CMBL_TRY(dupload(mp, &p_co, chofs));
    d.ring_chofs = p_co;
    long* p_so;
    CMBL_TRY(dupload(mp, &p_so, scrofs));
    d.ring_scrofs = p_so;
    CMBL_TRY(dalloc(mp, &d.ring_scr, (size_t)std::max<long>(scr_total, 1)));
    CMBL_TRY(dupload(mp, &p_A, A));
    CMBL_TRY(dupload(mp, &p_N, Nrm));
    CMBL_TRY(dupload(mp, &p_sv, seed_v));
    CMBL_TRY(dupload(mp, &p_sk, seed_k));
    CMBL_TRY(dupload(mp, &mp->d_tw, tw));
    d.ring_x = p_rx;
    d.ring_nph = p_nph;
    d.ring_shift = p_sh;
    d.ring_ofsN = p_oN;
    d.ring_ofsS = p_oS;
    d.ring_M = p_M;
    d.ring_bsofs = p_bo;
    d.A = p_A;
    d.Nrm = p_N;
    d.seed_v = p_sv;
    d.seed_k = p_sk;
    double2* p_bs;
    CMBL_TRY(dalloc(mp, &p_bs, (size_t)std::max<long>(bs_total, 1)));
    double2* p_ch;
    CMBL_TRY(dalloc(mp, &p_ch, (size_t)std::max<long>(ch_total, 1)));
    d.bs_chirp = p_ch;
    d.bs_filt = p_bs;
    // beam = 1, no prior until set
    CMBL_TRY(dupload(mp, &mp->d_beam, std::vector<double>(L + 1, 1.0)));
    CMBL_TRY(dupload(mp, &mp->d_invcl, std::vector<double>(L + 1, 0.0)));
    d.beam = mp->d_beam;
    d.invcl = mp->d_invcl;
    CMBL_TRY(dalloc(mp, &mp->d_data, mp->npix));
    CMBL_TRY(dalloc(mp, &mp->d_w, mp->npix));
    CMBL_HIP(hipMemset(mp->d_data, 0, sizeof(double) * mp->npix));
    CMBL_HIP(hipMemset(mp->d_w, 0, sizeof(double) * mp->npix));
    d.d = mp->d_data;
    d.w = mp->d_w;
    // workspace
    CMBL_TRY(dalloc(mp, &d.alm, mp->nalm));
    CMBL_TRY(dalloc(mp, &d.P, (size_t)mp->npair * (L + 1) * 4));
    CMBL_HIP(hipMemset(d.P, 0, sizeof(double) * (size_t)mp->npair * (L + 1) * 4));
    CMBL_TRY(dalloc(mp, &d.Q, (size_t)mp->npair * (L + 1) * 4));
    CMBL_TRY(dalloc(mp, &d.upart, (size_t)mp->nrb5 * mp->nalm * 2));
    CMBL_HIP(hipMemset(d.upart, 0, sizeof(double) * (size_t)mp->nrb5 * mp->nalm * 2));
    CMBL_TRY(dalloc(mp, &d.chi, mp->npair));
    d.nprior = lm_tiles(L);
    CMBL_TRY(dalloc(mp, &d.prior, d.nprior));
    CMBL_TRY(dalloc(mp, &mp->d_f, 1));
    CMBL_TRY(dalloc(mp, &mp->d_x, mp->nparam));
    CMBL_TRY(dalloc(mp, &mp->d_g, mp->nparam));
    CMBL_TRY(dalloc(mp, &mp->d_map, mp->npix));
    CMBL_TRY(dalloc(mp, &mp->d_desc, 1));
    // non-finite flag of the -2lnL reduction, in host-mapped memory so that the
    // host reads it without a copy
    CMBL_HIP(hipHostMalloc((void**)&mp->h_nonfinite, sizeof(int), hipHostMallocMapped));
    *mp->h_nonfinite = 0;
    CMBL_HIP(hipHostGetDevicePointer((void**)&d.nonfinite, mp->h_nonfinite, 0));
    CMBL_HIP(hipMemcpy(mp->d_desc, &d, sizeof(d), hipMemcpyHostToDevice));

    // ---- Bluestein filters
    if (!bs_pairs.empty()) {
        int* d_pairs;
        CMBL_TRY(dupload(mp, &d_pairs, bs_pairs));
        CMBL_TRY(launch_bluestein_setup(mp->d_desc, d_pairs, (int)bs_pairs.size(), p_bs, p_ch, mp->lds_ring, mp->d_tw,
                                        nullptr));
        CMBL_HIP(hipDeviceSynchronize());
    }

    // ---- scan: first live l per (m, pair); work lists; start states
    std::vector<WorkItem> all;
    for (int rb = 0; rb < mp->nrb1; ++rb)
        for (int m = 0; m <= L; m += 4) all.push_back(WorkItem{0, m, std::min(4, L + 1 - m), rb, m, L, -1, 0});
    WorkItem* d_all;
    int* d_ls;
    CMBL_TRY(dupload(mp, &d_all, all));
    CMBL_TRY(dalloc(mp, &d_ls, (size_t)(L + 1) * npad));
    CMBL_TRY(launch_scan(mp->K1, d_all, (int)all.size(), mp->d_desc, d_ls, nullptr));
    std::vector<int> lstart((size_t)(L + 1) * npad);
    CMBL_HIP(hipMemcpy(lstart.data(), d_ls, sizeof(int) * lstart.size(), hipMemcpyDeviceToHost));
    // padding pairs never constrain liveness
    for (int m = 0; m <= L; ++m)
        for (int p = mp->npair; p < npad; ++p) lstart[(size_t)m * npad + p] = L + 1;

    std::vector<ItemCost> c1, c5;
    build_items(mp, mp->K1, lstart, 0, &c1);
    build_items(mp, mp->K5, lstart, 0, &c5);
    std::vector<ItemCost> c5whole;  // batches take whole items: they have items enough
    if (c5.size() < kSegmentBelow) {
        segment_items(&c5, L, adj_segment_len(L), mp->K5, &c5whole);
    } else {
        for (size_t i = 0; i < c5.size(); ++i) c5[i].it.sidx = (int)i;  // whole items, one slot each
        c5whole = c5;
    }
    mp->cost1 = c1;
    mp->cost5 = c5whole;
    mp->items1 = order_items(c1);
    mp->items5 = order_items(c5);
    mp->nitems1 = (int)mp->items1.size();
    mp->nitems5 = (int)mp->items5.size();
    CMBL_TRY(dupload(mp, &mp->d_items1, mp->items1));
    CMBL_TRY(dupload(mp, &mp->d_items5, mp->items5));
    {
        const size_t nst = (size_t)(L + 1) * npad;
        const size_t nst5 = std::max<size_t>(c5.size(), 1) * 64 * mp->K5;  // one slot per segment
        double *p1, *c1p, *p5, *c5p;
        int *k1, *k5;
        CMBL_TRY(dalloc(mp, &p1, nst));
        CMBL_TRY(dalloc(mp, &c1p, nst));
        CMBL_TRY(dalloc(mp, &k1, nst));
        CMBL_TRY(dalloc(mp, &p5, nst5));
        CMBL_TRY(dalloc(mp, &c5p, nst5));
        CMBL_TRY(dalloc(mp, &k5, nst5));
        CMBL_TRY(launch_capture(mp->K1, mp->d_items1, mp->nitems1, mp->d_desc, p1, c1p, k1, nullptr));
        CMBL_TRY(launch_capture(mp->K5, mp->d_items5, mp->nitems5, mp->d_desc, p5, c5p, k5, nullptr));
        d.s1_prev = p1;
        d.s1_cur = c1p;
        d.s1_k = k1;
        d.s5_prev = p5;
        d.s5_cur = c5p;
        d.s5_k = k5;
        CMBL_HIP(hipMemcpy(mp->d_desc, &d, sizeof(d), hipMemcpyHostToDevice));
    }
    // ring items: largest transform first
    for (int p = 0; p < mp->npair; ++p) mp->ring_items.push_back(RingItem{0, p});
    mp->ring_M.assign(Mfft.begin(), Mfft.begin() + mp->npair);
    mp->ring_two.assign(mp->npair, 0);
    for (int p = 0; p < mp->npair; ++p)
        if (Mfft[p] == nph[p] && nph[p] > L && scrofs[p] < 0) mp->ring_two[p] = nph[p];
    mp->ring_classes = ring_classes(
        &mp->ring_items, [&](const RingItem& r) { return Mfft[r.pair]; },
        [&](const RingItem& r) { return mp->ring_two[r.pair]; });
    mp->nring_items = (int)mp->ring_items.size();
    CMBL_TRY(dupload(mp, &mp->d_ring_items, mp->ring_items));
    CMBL_HIP(hipDeviceSynchronize());
    return CMBL_OK;
}

static void map_free(cmbl_map* mp)
{
    for (hipEvent_t e : mp->ev) (void)hipEventDestroy(e);
    mp->ev.clear();
    for (void* p : mp->allocs) (void)hipFree(p);
    mp->allocs.clear();
    if (mp->h_nonfinite) (void)hipHostFree(mp->h_nonfinite);
    mp->h_nonfinite = nullptr;
}

// Report (and clear) a non-finite -2lnL raised by an evaluation of this map
// that has completed.  Does not wait for evaluations still in flight.
static int take_nonfinite(cmbl_map* mp)
{
    if (mp->h_nonfinite && __atomic_load_n(mp->h_nonfinite, __ATOMIC_ACQUIRE) != 0) {
        __atomic_store_n(mp->h_nonfinite, 0, __ATOMIC_RELEASE);
        return fail(CMBL_ERR_NONFINITE, "an earlier evaluation of this map produced a non-finite -2lnL "
                                        "(NaN/Inf in a_lm or data)");
    }
    return CMBL_OK;
}

static int check_map(const cmbl_map* mp)
{
    if (!mp) return fail(CMBL_ERR_ARG, "null map");
    return CMBL_OK;
}

// Stage timing: accumulate the oldest recorded event sets (blocks on them).
static int drain_timing(cmbl_map* mp, int keep)
{
    const int S = CMBL_NSTAGES + 1;
    while (mp->ev_pending > keep) {
        const int set = (mp->ev_next - mp->ev_pending + cmbl_map::kEvSets) % cmbl_map::kEvSets;
        hipEvent_t* e = &mp->ev[(size_t)set * S];
        CMBL_HIP(hipEventSynchronize(e[CMBL_NSTAGES]));
        for (int i = 0; i < CMBL_NSTAGES; ++i) {
            float ms = 0.0f;
            CMBL_HIP(hipEventElapsedTime(&ms, e[i], e[i + 1]));
            mp->stage_ms[i] += ms;
        }
        ++mp->timed_evals;
        --mp->ev_pending;
    }
    return CMBL_OK;
}

struct StageMarks {
    cmbl_map* mp;
    hipStream_t s;
    bool on;
    hipEvent_t* e = nullptr;
    int begin()
    {
        if (!on || !mp->timing) return CMBL_OK;
        CMBL_TRY(drain_timing(mp, cmbl_map::kEvSets - 1));
        e = &mp->ev[(size_t)mp->ev_next * (CMBL_NSTAGES + 1)];
        mp->ev_next = (mp->ev_next + 1) % cmbl_map::kEvSets;
        ++mp->ev_pending;
        return mark(0);
    }
    int mark(int i)
    {
        if (e) CMBL_HIP(hipEventRecord(e[i], s));
        return CMBL_OK;
    }
};

// One evaluation on stream s with device x; f and g are device pointers.
// With marks, stage-timing events are recorded when the map's timing is on.
int eval_device(cmbl_map* mp, const double* x, double* f, double* g, hipStream_t s, bool marks)
{
    const MapDesc& d = mp->desc;
    StageMarks t{mp, s, marks};
    CMBL_TRY(t.begin());
    CMBL_TRY(launch_pre(d, mp->d_desc, x, s));
    CMBL_TRY(t.mark(1));
    CMBL_TRY(launch_synth(mp->K1, mp->d_items1, mp->nitems1, mp->d_desc, s));
    CMBL_TRY(t.mark(2));
    CMBL_TRY(run_ring(kRingLike, mp->d_ring_items, mp->ring_classes, mp->d_desc, nullptr, nullptr, mp->d_tw, s));
    CMBL_TRY(t.mark(3));
    CMBL_TRY(launch_adj(mp->K5, mp->d_items5, mp->nitems5, mp->d_desc, s));
    CMBL_TRY(t.mark(4));
    CMBL_TRY(launch_post(d, mp->d_desc, x, g, kPostLike, s));
    CMBL_TRY(t.mark(5));
    CMBL_TRY(launch_final(mp->d_desc, f, s));
    CMBL_TRY(t.mark(6));
    return CMBL_OK;
}

}  // namespace cmbl

using namespace cmbl;

// ====================================================================== C ABI
extern "C" {

const char* cmbl_last_error(void) { return g_err.c_str(); }
const char* cmbl_version(void) { return "cosmopp_amd 0.1 (gfx950)"; }

static int map_create_checked(cmbl_map** out, int nside, int lmax, const cmbl_tuning* given)
{
    if (!out) return fail(CMBL_ERR_ARG, "out is null");
    *out = nullptr;
    if (nside < 1 || !is_pow2(nside)) return fail(CMBL_ERR_ARG, "nside must be a positive power of two");
    if (lmax < 2 || lmax > 4 * nside) return fail(CMBL_ERR_ARG, "lmax must be in [2, 4*nside]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CMBL_ERR_HIP, "no HIP device available");
    cmbl_tuning t;
    if (given) {
        if (!tuning_valid(*given))
            return fail(CMBL_ERR_ARG, "tuning: ring pairs per lane must be 1, 2, 4 or 8 and the split length "
                                      "2048, 4096 or 8192");
        t = *given;
    } else {
        t = default_tuning(nside, lmax);
        // test overrides: CMBL_K1 / CMBL_K5 in {1, 2, 4, 8}, CMBL_SPLIT_M in {2048, 4096, 8192}
        cmbl_tuning o = t;
        if (const char* e = std::getenv("CMBL_K1")) o.k_synth = std::atoi(e);
        if (const char* e = std::getenv("CMBL_K5")) o.k_adj = std::atoi(e);
        if (const char* e = std::getenv("CMBL_SPLIT_M")) o.split_m = std::atoi(e);
        if (tuning_valid(o)) t = o;
    }
    cmbl_map* mp = new cmbl_map();
    const int rc = map_create(mp, nside, lmax, t);
    if (rc != CMBL_OK) {
        const std::string msg = g_err;
        map_free(mp);
        delete mp;
        g_err = msg;
        return rc;
    }
    *out = mp;
    return CMBL_OK;
}

int cmbl_map_create(cmbl_map** out, int nside, int lmax) { return map_create_checked(out, nside, lmax, nullptr); }

int cmbl_map_create_tuned(cmbl_map** out, int nside, int lmax, const cmbl_tuning* tuning)
{
    if (!tuning) return fail(CMBL_ERR_ARG, "tuning is null");
    return map_create_checked(out, nside, lmax, tuning);
}

int cmbl_map_get_tuning(const cmbl_map* mp, cmbl_tuning* t)
{
    if (!mp || !t) return fail(CMBL_ERR_ARG, "null pointer");
    *t = mp->tuning;
    return CMBL_OK;
}

int cmbl_tuning_load(const char* path)
{
    if (!path) return fail(CMBL_ERR_ARG, "path is null");
    std::lock_guard<std::mutex> lock(g_tune_mu);
    return tuning_read(path, true);
}

int cmbl_tuning_save(const char* path)
{
    if (!path) return fail(CMBL_ERR_ARG, "path is null");
    std::lock_guard<std::mutex> lock(g_tune_mu);
    std::ofstream out(path);
    if (!out) return fail(CMBL_ERR_ARG, std::string("cannot write ") + path);
    out << "# arch nside lmax k_synth k_adj split_m  (cmbl_tune: fastest of timed candidates)\n";
    for (const TuneEntry& e : g_tune)
        out << e.arch << " " << e.nside << " " << e.lmax << " " << e.t.k_synth << " " << e.t.k_adj << " "
            << e.t.split_m << "\n";
    return out ? CMBL_OK : fail(CMBL_ERR_ARG, std::string("write failed: ") + path);
}

// Average ms of the synthesis, ring and adjoint stages of one evaluation of
// a map created with tuning t (the stages' work does not depend on the data
// or the a_lm values, so zero data, unit weights and x = 1 are used).
static int time_candidate(int nside, int lmax, const cmbl_tuning& t, int reps, double* synth, double* ring,
                          double* adj)
{
    cmbl_map* mp = nullptr;
    CMBL_TRY(map_create_checked(&mp, nside, lmax, &t));
    int rc = CMBL_OK;
    double* x = nullptr;
    double ms[CMBL_NSTAGES] = {0, 0, 0, 0, 0, 0};
    long nev = 0;
    if (hipMalloc(&x, sizeof(double) * mp->nparam) != hipSuccess) rc = fail(CMBL_ERR_ALLOC, "tuning x");
    if (rc == CMBL_OK) rc = launch_vec(kVecSet, x, nullptr, 1.0, mp->nparam, nullptr);
    if (rc == CMBL_OK) rc = launch_vec(kVecSet, mp->d_w, nullptr, 1.0, mp->npix, nullptr);
    mp->has_data = true;
    for (int i = 0; i < 2 && rc == CMBL_OK; ++i) rc = eval_device(mp, x, mp->d_f, mp->d_g, nullptr);
    if (rc == CMBL_OK) rc = cmbl_map_set_timing(mp, 1);
    for (int i = 0; i < reps && rc == CMBL_OK; ++i) rc = eval_device(mp, x, mp->d_f, mp->d_g, nullptr);
    if (rc == CMBL_OK) rc = cmbl_map_stage_times(mp, ms, &nev);
    if (rc == CMBL_OK && nev > 0) {
        *synth = ms[1] / nev;
        *ring = ms[2] / nev;
        *adj = ms[3] / nev;
    }
    const std::string msg = g_err;
    if (x) (void)hipFree(x);
    cmbl_map_destroy(mp);
    g_err = msg;
    return rc;
}

// The first candidate, in preference order, within 2% of the fastest: timing
// noise of a percent does not flip the choice between runs.
static int pick(const std::vector<int>& order, const double* t)
{
    double best = 1e300;
    for (int i : order)
        if (t[i] >= 0) best = std::min(best, t[i]);
    for (int i : order)
        if (t[i] >= 0 && t[i] <= 1.02 * best) return i;
    return order[0];
}

int cmbl_tune(int nside, int lmax, int reps, cmbl_tuning* best, double* times)
{
    if (!best) return fail(CMBL_ERR_ARG, "best is null");
    if (reps < 1) return fail(CMBL_ERR_ARG, "reps must be >= 1");
    if (nside < 1 || !is_pow2(nside) || lmax < 2 || lmax > 4 * nside)
        return fail(CMBL_ERR_ARG, "nside must be a power of two and lmax in [2, 4*nside]");
    const cmbl_tuning rule{choose_k(2 * nside, 2), choose_k(2 * nside, 4), 4096};
    double t[11];
    for (double& v : t) v = -1.0;
    // ring pairs per lane: candidates 1, 2, 4, 8 for both Legendre stages
    const int ks[4] = {1, 2, 4, 8};
    for (int c = 0; c < 4; ++c) {
        double ring = 0;
        CMBL_TRY(time_candidate(nside, lmax, cmbl_tuning{ks[c], ks[c], rule.split_m}, reps, &t[c], &ring, &t[4 + c]));
        t[9] = ring;
    }
    auto order_from = [&](int k) {
        std::vector<int> o{k == 1 ? 0 : k == 2 ? 1 : k == 4 ? 2 : 3};
        for (int c = 0; c < 4; ++c)
            if (c != o[0]) o.push_back(c);
        return o;
    };
    cmbl_tuning choice{ks[pick(order_from(rule.k_synth), t)], 0, rule.split_m};
    std::vector<int> oa = order_from(rule.k_adj);
    for (int& i : oa) i += 4;
    choice.k_adj = ks[pick(oa, t) - 4];
    // split length: only rings of nph > 1024 (nside >= 512) are Bluestein
    // transforms longer than 2048
    if (nside >= 512) {
        const int splits[3] = {2048, 4096, 8192};
        for (int c = 0; c < 3; ++c) {
            double sy = 0, ad = 0;
            CMBL_TRY(time_candidate(nside, lmax, cmbl_tuning{choice.k_synth, choice.k_adj, splits[c]}, reps, &sy,
                                    &t[8 + c], &ad));
        }
        choice.split_m = splits[pick({9, 8, 10}, t) - 8];
    }
    if (times)
        for (int i = 0; i < 11; ++i) times[i] = t[i];
    *best = choice;
    return cmbl_tuning_set(nside, lmax, &choice);
}

int cmbl_tuning_set(int nside, int lmax, const cmbl_tuning* t)
{
    if (!t || !tuning_valid(*t)) return fail(CMBL_ERR_ARG, "invalid tuning");
    const std::string arch = device_arch();
    std::lock_guard<std::mutex> lock(g_tune_mu);
    tuning_load_default(arch);
    for (TuneEntry& e : g_tune)
        if (e.arch == arch && e.nside == nside && e.lmax == lmax) {
            e.t = *t;
            return CMBL_OK;
        }
    g_tune.push_back(TuneEntry{arch, nside, lmax, *t});
    return CMBL_OK;
}

void cmbl_map_destroy(cmbl_map* mp)
{
    if (!mp) return;
    (void)hipDeviceSynchronize();
    map_free(mp);
    delete mp;
}

int cmbl_map_nside(const cmbl_map* mp) { return mp ? mp->nside : -1; }
int cmbl_map_lmax(const cmbl_map* mp) { return mp ? mp->L : -1; }
long cmbl_map_npix(const cmbl_map* mp) { return mp ? mp->npix : -1; }
long cmbl_map_nparams(const cmbl_map* mp) { return mp ? mp->nparam : -1; }

int cmbl_map_set_timing(cmbl_map* mp, int on)
{
    CMBL_TRY(check_map(mp));
    CMBL_TRY(drain_timing(mp, 0));
    if (on && mp->ev.empty()) {
        mp->ev.resize((size_t)cmbl_map::kEvSets * (CMBL_NSTAGES + 1));
        for (hipEvent_t& e : mp->ev) CMBL_HIP(hipEventCreate(&e));
    }
    mp->timing = on != 0;
    for (double& v : mp->stage_ms) v = 0.0;
    mp->timed_evals = 0;
    return CMBL_OK;
}

int cmbl_map_stage_times(cmbl_map* mp, double* ms, long* nevals)
{
    CMBL_TRY(check_map(mp));
    if (!ms) return fail(CMBL_ERR_ARG, "ms is null");
    CMBL_TRY(drain_timing(mp, 0));
    for (int i = 0; i < CMBL_NSTAGES; ++i) ms[i] = mp->stage_ms[i];
    if (nevals) *nevals = mp->timed_evals;
    return CMBL_OK;
}

int cmbl_map_set_beam(cmbl_map* mp, const double* beam)
{
    CMBL_TRY(check_map(mp));
    if (!beam) return fail(CMBL_ERR_ARG, "beam is null");
    CMBL_HIP(hipMemcpy(mp->d_beam, beam, sizeof(double) * (mp->L + 1), hipMemcpyHostToDevice));
    mp->has_beam = true;
    ++mp->generation;
    return CMBL_OK;
}

int cmbl_map_set_cl(cmbl_map* mp, const double* cl)
{
    CMBL_TRY(check_map(mp));
    if (!cl) return fail(CMBL_ERR_ARG, "cl is null");
    std::vector<double> inv(mp->L + 1, 0.0);
    for (int l = 0; l <= mp->L; ++l) inv[l] = cl[l] > 0 ? 1.0 / cl[l] : 0.0;
    CMBL_HIP(hipMemcpy(mp->d_invcl, inv.data(), sizeof(double) * (mp->L + 1), hipMemcpyHostToDevice));
    mp->has_cl = true;
    ++mp->generation;
    return CMBL_OK;
}

int cmbl_map_set_data(cmbl_map* mp, const double* d, const double* w)
{
    CMBL_TRY(check_map(mp));
    if (!d || !w) return fail(CMBL_ERR_ARG, "data or weight is null");
    CMBL_HIP(hipMemcpy(mp->d_data, d, sizeof(double) * mp->npix, hipMemcpyHostToDevice));
    CMBL_HIP(hipMemcpy(mp->d_w, w, sizeof(double) * mp->npix, hipMemcpyHostToDevice));
    mp->has_data = true;
    ++mp->generation;
    return CMBL_OK;
}

int cmbl_map_set_data_device(cmbl_map* mp, const double* d, const double* w, void* stream)
{
    CMBL_TRY(check_map(mp));
    if (!d || !w) return fail(CMBL_ERR_ARG, "data or weight is null");
    hipStream_t s = (hipStream_t)stream;
    CMBL_HIP(hipMemcpyAsync(mp->d_data, d, sizeof(double) * mp->npix, hipMemcpyDeviceToDevice, s));
    CMBL_HIP(hipMemcpyAsync(mp->d_w, w, sizeof(double) * mp->npix, hipMemcpyDeviceToDevice, s));
    mp->has_data = true;
    ++mp->generation;
    return CMBL_OK;
}

int cmbl_value_and_grad_device(cmbl_map* mp, const double* x, double* f, double* g, void* stream)
{
    CMBL_TRY(check_map(mp));
    if (!x || !f) return fail(CMBL_ERR_ARG, "x or f is null");
    if (!mp->has_data) return fail(CMBL_ERR_STATE, "cmbl_map_set_data has not been called");
    CMBL_TRY(take_nonfinite(mp));
    hipStream_t s = (hipStream_t)stream;
    return eval_device(mp, x, f, g ? g : mp->d_g, s);
}

int cmbl_map_check_finite(cmbl_map* mp, void* stream)
{
    CMBL_TRY(check_map(mp));
    CMBL_HIP(hipStreamSynchronize((hipStream_t)stream));
    return take_nonfinite(mp);
}

long cmbl_map_generation(const cmbl_map* mp) { return mp ? mp->generation : -1; }

int cmbl_value_and_grad(cmbl_map* mp, const double* x, double* f, double* g)
{
    CMBL_TRY(check_map(mp));
    if (!x || !f) return fail(CMBL_ERR_ARG, "x or f is null");
    if (!mp->has_data) return fail(CMBL_ERR_STATE, "cmbl_map_set_data has not been called");
    CMBL_TRY(take_nonfinite(mp));
    CMBL_HIP(hipMemcpy(mp->d_x, x, sizeof(double) * mp->nparam, hipMemcpyHostToDevice));
    CMBL_TRY(eval_device(mp, mp->d_x, mp->d_f, mp->d_g, nullptr));
    CMBL_HIP(hipMemcpy(f, mp->d_f, sizeof(double), hipMemcpyDeviceToHost));
    if (g) CMBL_HIP(hipMemcpy(g, mp->d_g, sizeof(double) * mp->nparam, hipMemcpyDeviceToHost));
    if (!std::isfinite(*f)) {
        (void)take_nonfinite(mp);
        return fail(CMBL_ERR_NONFINITE, "-2lnL is not finite (NaN/Inf in a_lm or data)");
    }
    return CMBL_OK;
}

int cmbl_forward_device(cmbl_map* mp, const double* x, double* map, void* stream)
{
    CMBL_TRY(check_map(mp));
    if (!x || !map) return fail(CMBL_ERR_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    CMBL_TRY(launch_pre(mp->desc, mp->d_desc, x, s));
    CMBL_TRY(launch_synth(mp->K1, mp->d_items1, mp->nitems1, mp->d_desc, s));
    CMBL_TRY(run_ring(kRingSyn, mp->d_ring_items, mp->ring_classes, mp->d_desc, map, nullptr, mp->d_tw, s));
    return CMBL_OK;
}

int cmbl_forward(cmbl_map* mp, const double* x, double* map)
{
    CMBL_TRY(check_map(mp));
    if (!x || !map) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipMemcpy(mp->d_x, x, sizeof(double) * mp->nparam, hipMemcpyHostToDevice));
    CMBL_TRY(cmbl_forward_device(mp, mp->d_x, mp->d_map, nullptr));
    CMBL_HIP(hipMemcpy(map, mp->d_map, sizeof(double) * mp->npix, hipMemcpyDeviceToHost));
    return CMBL_OK;
}

int cmbl_backward_device(cmbl_map* mp, const double* v, double* g, void* stream)
{
    CMBL_TRY(check_map(mp));
    if (!v || !g) return fail(CMBL_ERR_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    CMBL_TRY(run_ring(kRingAna, mp->d_ring_items, mp->ring_classes, mp->d_desc, nullptr, v, mp->d_tw, s));
    CMBL_TRY(launch_adj(mp->K5, mp->d_items5, mp->nitems5, mp->d_desc, s));
    CMBL_TRY(launch_post(mp->desc, mp->d_desc, nullptr, g, kPostGrad, s));
    return CMBL_OK;
}

int cmbl_backward(cmbl_map* mp, const double* v, double* g)
{
    CMBL_TRY(check_map(mp));
    if (!v || !g) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipMemcpy(mp->d_map, v, sizeof(double) * mp->npix, hipMemcpyHostToDevice));
    CMBL_TRY(cmbl_backward_device(mp, mp->d_map, mp->d_g, nullptr));
    CMBL_HIP(hipMemcpy(g, mp->d_g, sizeof(double) * mp->nparam, hipMemcpyDeviceToHost));
    return CMBL_OK;
}

// ---------------------------------------------------------------- quadrature analysis
int cmbl_map2alm_quadrature_device(cmbl_map* mp, const double* map, double* alm, void* stream)
{
    CMBL_TRY(check_map(mp));
    if (!map || !alm) return fail(CMBL_ERR_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    CMBL_TRY(run_ring(kRingAna, mp->d_ring_items, mp->ring_classes, mp->d_desc, nullptr, map, mp->d_tw, s));
    CMBL_TRY(launch_adj(mp->K5, mp->d_items5, mp->nitems5, mp->d_desc, s));
    CMBL_TRY(launch_post(mp->desc, mp->d_desc, nullptr, alm, kPostQuad, s, 4.0 * 3.141592653589793 / (double)mp->npix));
    return CMBL_OK;
}

int cmbl_map2alm_quadrature(cmbl_map* mp, const double* map, double* alm)
{
    CMBL_TRY(check_map(mp));
    if (!map || !alm) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipMemcpy(mp->d_map, map, sizeof(double) * mp->npix, hipMemcpyHostToDevice));
    CMBL_TRY(cmbl_map2alm_quadrature_device(mp, mp->d_map, mp->d_g, nullptr));
    CMBL_HIP(hipMemcpy(alm, mp->d_g, sizeof(double) * mp->nparam, hipMemcpyDeviceToHost));
    return CMBL_OK;
}

// ---------------------------------------------------------------- spin-2 synthesis
// Workspace of cmbl_alm2map_pol, allocated at its first call: the
// coefficient records and the T, Q, U parity-sum buffers (zeroed once: entries
// of (m, ring block) items with no live l are never written), with one
// descriptor per buffer that differs from the map's only in P.
static int pol_setup(cmbl_map* mp)
{
    if (mp->d_desc_pol) return CMBL_OK;
    const size_t pstride = (size_t)mp->npair * (mp->L + 1) * 4;
    CMBL_TRY(dalloc(mp, &mp->d_pol_rec, (size_t)kPolRec * mp->nalm));
    CMBL_TRY(dalloc(mp, &mp->d_pol_P, 3 * pstride));
    CMBL_HIP(hipMemset(mp->d_pol_P, 0, sizeof(double) * 3 * pstride));
    std::vector<MapDesc> v(3, mp->desc);
    for (int i = 0; i < 3; ++i) v[i].P = mp->d_pol_P + (size_t)i * pstride;
    MapDesc* d;
    CMBL_TRY(dupload(mp, &d, v));
    mp->d_desc_pol = d;
    return CMBL_OK;
}

int cmbl_alm2map_pol_device(cmbl_map* mp, const double* t, const double* e, const double* b, double* map_t,
                            double* map_q, double* map_u, void* stream)
{
    CMBL_TRY(check_map(mp));
    if (!t || !e || !b || !map_t || !map_q || !map_u) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_TRY(pol_setup(mp));
    hipStream_t s = (hipStream_t)stream;
    MapDesc* dm = mp->d_desc_pol;
    const size_t pstride = (size_t)mp->npair * (mp->L + 1) * 4;
    double* P = mp->d_pol_P;
    CMBL_TRY(launch_pre_pol(mp->desc, mp->d_desc, t, e, b, mp->d_pol_rec, s));
    CMBL_TRY(launch_synth_pol(mp->K1, mp->d_items1, mp->nitems1, mp->d_desc, mp->d_pol_rec, P, P + pstride,
                              P + 2 * pstride, s));
    CMBL_TRY(run_ring(kRingSyn, mp->d_ring_items, mp->ring_classes, dm + 0, map_t, nullptr, mp->d_tw, s));
    CMBL_TRY(run_ring(kRingSyn, mp->d_ring_items, mp->ring_classes, dm + 1, map_q, nullptr, mp->d_tw, s));
    CMBL_TRY(run_ring(kRingSyn, mp->d_ring_items, mp->ring_classes, dm + 2, map_u, nullptr, mp->d_tw, s));
    return CMBL_OK;
}

int cmbl_alm2map_pol(cmbl_map* mp, const double* t, const double* e, const double* b, double* map_t, double* map_q,
                     double* map_u)
{
    CMBL_TRY(check_map(mp));
    if (!t || !e || !b || !map_t || !map_q || !map_u) return fail(CMBL_ERR_ARG, "null pointer");
    if (!mp->d_pol_in) {
        CMBL_TRY(dalloc(mp, &mp->d_pol_in, 3 * (size_t)mp->nparam));
        CMBL_TRY(dalloc(mp, &mp->d_pol_out, 3 * (size_t)mp->npix));
    }
    double* in = mp->d_pol_in;
    double* out = mp->d_pol_out;
    const size_t n = mp->nparam, np = mp->npix;
    CMBL_HIP(hipMemcpy(in, t, sizeof(double) * n, hipMemcpyHostToDevice));
    CMBL_HIP(hipMemcpy(in + n, e, sizeof(double) * n, hipMemcpyHostToDevice));
    CMBL_HIP(hipMemcpy(in + 2 * n, b, sizeof(double) * n, hipMemcpyHostToDevice));
    CMBL_TRY(cmbl_alm2map_pol_device(mp, in, in + n, in + 2 * n, out, out + np, out + 2 * np, nullptr));
    CMBL_HIP(hipMemcpy(map_t, out, sizeof(double) * np, hipMemcpyDeviceToHost));
    CMBL_HIP(hipMemcpy(map_q, out + np, sizeof(double) * np, hipMemcpyDeviceToHost));
    CMBL_HIP(hipMemcpy(map_u, out + 2 * np, sizeof(double) * np, hipMemcpyDeviceToHost));
    return CMBL_OK;
}

// ---------------------------------------------------------------- Gibbs CG
void cmbl_gibbs_destroy(cmbl_gibbs* gb)
{
    if (!gb) return;
    (void)hipDeviceSynchronize();
    gibbs_cg_free(gb->cg);
    gb->cg = nullptr;
    for (void* p : gb->allocs) (void)hipFree(p);
    cmbl_map_destroy(gb->mp);
    delete gb;
}

int cmbl_gibbs_create(cmbl_gibbs** out, int nside, int lmax, const double* cl, const double* beam,
                      const double* mask, double pixel_noise)
{
    if (!out || !cl || !beam || !mask) return fail(CMBL_ERR_ARG, "null pointer");
    *out = nullptr;
    if (!(pixel_noise > 0.0) || !std::isfinite(pixel_noise)) return fail(CMBL_ERR_ARG, "pixel noise must be > 0");
    for (int l = 0; l <= lmax && lmax >= 0; ++l)
        if (!(cl[l] >= 0.0) || !std::isfinite(cl[l]) || !std::isfinite(beam[l]))
            return fail(CMBL_ERR_ARG, "C_l must be finite and >= 0, B_l finite");
    cmbl_gibbs* gb = new cmbl_gibbs();
    auto bail = [&](int rc) {
        const std::string msg = g_err;
        cmbl_gibbs_destroy(gb);
        g_err = msg;
        return rc;
    };
    int rc = cmbl_map_create(&gb->mp, nside, lmax);
    if (rc) return bail(rc);
    gb->L = lmax;
    gb->n = gb->mp->nparam;
    gb->npix = gb->mp->npix;
    gb->sigma = pixel_noise;
    gb->beam.assign(beam, beam + lmax + 1);
    // D_l = B_l sqrt(C_l) (source/cmb_gibbs.cpp:370-375); w = mask / sigma^2 (:381-387)
    std::vector<double> dl(lmax + 1), w(gb->npix), zero(gb->npix, 0.0);
    for (int l = 0; l <= lmax; ++l) dl[l] = beam[l] * std::sqrt(cl[l]);
    for (long i = 0; i < gb->npix; ++i) {
        const bool good = mask[i] > 0.5;
        w[i] = good ? 1.0 / (pixel_noise * pixel_noise) : 0.0;
        gb->good += good;
    }
    if (gb->good == 0) {
        g_err = "the mask has no pixel > 0.5";
        return bail(CMBL_ERR_ARG);
    }
    if ((rc = cmbl_map_set_beam(gb->mp, dl.data())) || (rc = cmbl_map_set_data(gb->mp, zero.data(), w.data())))
        return bail(rc);
    // preconditioner divisors (source/cmb_gibbs.cpp:412-418)
    const double nl = pixel_noise * pixel_noise * 4 * 3.141592653589793 / (double)gb->good;
    for (int l = 0; l <= lmax; ++l) dl[l] = 1.0 + cl[l] * beam[l] * beam[l] / nl;
    auto alloc = [&](double** p, long count) {
        void* q = nullptr;
        if (hipMalloc(&q, sizeof(double) * (size_t)std::max<long>(count, 1)) != hipSuccess) {
            g_err = "hipMalloc failed for the Gibbs workspace";
            return CMBL_ERR_ALLOC;
        }
        gb->allocs.push_back(q);
        *p = static_cast<double*>(q);
        return CMBL_OK;
    };
    for (double** v : {&gb->d_r, &gb->d_z, &gb->d_p, &gb->d_ap, &gb->d_b, &gb->d_x})
        if ((rc = alloc(v, gb->n))) return bail(rc);
    for (double** v : {&gb->d_m0, &gb->d_m1, &gb->d_m2})
        if ((rc = alloc(v, gb->npix))) return bail(rc);
    if ((rc = alloc(&gb->d_dl, lmax + 1))) return bail(rc);
    if (hipMemcpy(gb->d_dl, dl.data(), sizeof(double) * (lmax + 1), hipMemcpyHostToDevice) != hipSuccess) {
        g_err = "upload of the preconditioner failed";
        return bail(CMBL_ERR_HIP);
    }
    *out = gb;
    return CMBL_OK;
}

static int check_gibbs(const cmbl_gibbs* gb)
{
    if (!gb || !gb->mp) return fail(CMBL_ERR_ARG, "gibbs context is null");
    return CMBL_OK;
}

int cmbl_gibbs_set_cl(cmbl_gibbs* gb, const double* cl)
{
    CMBL_TRY(check_gibbs(gb));
    if (!cl) return fail(CMBL_ERR_ARG, "cl is null");
    const int L = gb->L;
    for (int l = 0; l <= L; ++l)
        if (!(cl[l] >= 0.0) || !std::isfinite(cl[l])) return fail(CMBL_ERR_ARG, "C_l must be finite and >= 0");
    // D_l = B_l sqrt(C_l) and the preconditioner divisors, as at creation
    std::vector<double> dl(L + 1);
    for (int l = 0; l <= L; ++l) dl[l] = gb->beam[l] * std::sqrt(cl[l]);
    CMBL_TRY(cmbl_map_set_beam(gb->mp, dl.data()));
    const double nl = gb->sigma * gb->sigma * 4 * 3.141592653589793 / (double)gb->good;
    for (int l = 0; l <= L; ++l) dl[l] = 1.0 + cl[l] * gb->beam[l] * gb->beam[l] / nl;
    CMBL_HIP(hipMemcpy(gb->d_dl, dl.data(), sizeof(double) * (L + 1), hipMemcpyHostToDevice));
    return CMBL_OK;
}

int cmbl_gibbs_apply_device(cmbl_gibbs* gb, const double* x, double* y, void* stream)
{
    CMBL_TRY(check_gibbs(gb));
    if (!x || !y) return fail(CMBL_ERR_ARG, "null pointer");
    cmbl_map* mp = gb->mp;
    hipStream_t s = (hipStream_t)stream;
    // D x -> Y -> w s -> Y^T -> D -> + x, one pass per stage
    CMBL_TRY(launch_pre(mp->desc, mp->d_desc, x, s));
    CMBL_TRY(launch_synth(mp->K1, mp->d_items1, mp->nitems1, mp->d_desc, s));
    CMBL_TRY(run_ring(kRingWeight, mp->d_ring_items, mp->ring_classes, mp->d_desc, nullptr, nullptr, mp->d_tw, s));
    CMBL_TRY(launch_adj(mp->K5, mp->d_items5, mp->nitems5, mp->d_desc, s));
    CMBL_TRY(launch_post(mp->desc, mp->d_desc, x, y, kPostGibbs, s));
    return CMBL_OK;
}

int cmbl_gibbs_precondition_device(cmbl_gibbs* gb, const double* x, double* y, void* stream)
{
    CMBL_TRY(check_gibbs(gb));
    if (!x || !y) return fail(CMBL_ERR_ARG, "null pointer");
    return launch_ldiv(y, x, gb->d_dl, gb->L, (hipStream_t)stream);
}

int cmbl_gibbs_rhs_device(cmbl_gibbs* gb, const double* d, const double* omega0, const double* omega1, double* b,
                          void* stream)
{
    CMBL_TRY(check_gibbs(gb));
    if (!d || !omega0 || !omega1 || !b) return fail(CMBL_ERR_ARG, "null pointer");
    cmbl_map* mp = gb->mp;
    hipStream_t s = (hipStream_t)stream;
    // w (d + sigma omega1) = N^-1 d + N^-1/2 omega1 on the good pixels
    CMBL_TRY(cmbl_vec_copy(gb->d_m0, d, 1.0, gb->npix, s));
    CMBL_TRY(cmbl_vec_axpy(gb->d_m0, omega1, gb->sigma, gb->npix, s));
    CMBL_TRY(cmbl_vec_mul(gb->d_m0, mp->d_w, gb->npix, s));
    CMBL_TRY(run_ring(kRingAna, mp->d_ring_items, mp->ring_classes, mp->d_desc, nullptr, gb->d_m0, mp->d_tw, s));
    CMBL_TRY(launch_adj(mp->K5, mp->d_items5, mp->nitems5, mp->d_desc, s));
    CMBL_TRY(launch_post(mp->desc, mp->d_desc, nullptr, b, kPostAlm, s));
    // + sqrt(4pi/Npix) Y^T omega0 (map2alm with ring weight sqrt(Npix/4pi), :479)
    CMBL_TRY(run_ring(kRingAna, mp->d_ring_items, mp->ring_classes, mp->d_desc, nullptr, omega0, mp->d_tw, s));
    CMBL_TRY(launch_adj(mp->K5, mp->d_items5, mp->nitems5, mp->d_desc, s));
    CMBL_TRY(launch_post(mp->desc, mp->d_desc, nullptr, b, kPostAddAlm, s,
                         std::sqrt(4.0 * 3.141592653589793 / (double)gb->npix)));
    return CMBL_OK;
}

int cmbl_gibbs_apply(cmbl_gibbs* gb, const double* x, double* y)
{
    CMBL_TRY(check_gibbs(gb));
    if (!x || !y) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipMemcpy(gb->d_x, x, sizeof(double) * gb->n, hipMemcpyHostToDevice));
    CMBL_TRY(cmbl_gibbs_apply_device(gb, gb->d_x, gb->d_b, nullptr));
    CMBL_HIP(hipMemcpy(y, gb->d_b, sizeof(double) * gb->n, hipMemcpyDeviceToHost));
    return CMBL_OK;
}

int cmbl_gibbs_precondition(cmbl_gibbs* gb, const double* x, double* y)
{
    CMBL_TRY(check_gibbs(gb));
    if (!x || !y) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipMemcpy(gb->d_x, x, sizeof(double) * gb->n, hipMemcpyHostToDevice));
    CMBL_TRY(cmbl_gibbs_precondition_device(gb, gb->d_x, gb->d_b, nullptr));
    CMBL_HIP(hipMemcpy(y, gb->d_b, sizeof(double) * gb->n, hipMemcpyDeviceToHost));
    return CMBL_OK;
}

int cmbl_gibbs_rhs(cmbl_gibbs* gb, const double* d, const double* omega0, const double* omega1, double* b)
{
    CMBL_TRY(check_gibbs(gb));
    if (!d || !omega0 || !omega1 || !b) return fail(CMBL_ERR_ARG, "null pointer");
    const size_t bytes = sizeof(double) * gb->npix;
    CMBL_HIP(hipMemcpy(gb->d_m1, d, bytes, hipMemcpyHostToDevice));
    CMBL_HIP(hipMemcpy(gb->d_m2, omega1, bytes, hipMemcpyHostToDevice));
    CMBL_HIP(hipMemcpy(gb->mp->d_map, omega0, bytes, hipMemcpyHostToDevice));
    CMBL_TRY(cmbl_gibbs_rhs_device(gb, gb->d_m1, gb->mp->d_map, gb->d_m2, gb->d_b, nullptr));
    CMBL_HIP(hipMemcpy(b, gb->d_b, sizeof(double) * gb->n, hipMemcpyDeviceToHost));
    return CMBL_OK;
}

int cmbl_gibbs_solve(cmbl_gibbs* gb, const double* b, double* x, double precision, int max_iter, int* iterations)
{
    CMBL_TRY(check_gibbs(gb));
    if (!b || !x) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipMemcpy(gb->d_b, b, sizeof(double) * gb->n, hipMemcpyHostToDevice));
    const int rc = cmbl_gibbs_solve_device(gb, gb->d_b, gb->d_x, precision, max_iter, iterations, nullptr);
    if (rc != CMBL_OK && rc != CMBL_ERR_NOCONV) return rc;
    const std::string msg = g_err;
    CMBL_HIP(hipMemcpy(x, gb->d_x, sizeof(double) * gb->n, hipMemcpyDeviceToHost));
    g_err = msg;
    return rc;
}

// ---------------------------------------------------------------- batch
int cmbl_batch_create(cmbl_batch** out, cmbl_map* const* maps, int nmaps)
{
    if (!out || !maps || nmaps < 1) return fail(CMBL_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    cmbl_batch* b = new cmbl_batch();
    auto cleanup = [&](int rc) {
        for (void* p : b->allocs) (void)hipFree(p);
        delete b;
        return rc;
    };
    std::vector<MapDesc> descs;
    std::vector<ItemCost> w1[4], w5[4];
    std::vector<RingItem> rings;
    // maps of one geometry (nside, lmax, synthesis ring pairs per lane) share
    // the synthesis recurrence in groups of four (then two)
    // (groups of up to kAdjG such maps, at least kSynMinG, run the synthesis
    // on the matrix cores first; CMBL_BATCH_MMA=0 turns that off)
    const char* mma_env = getenv("CMBL_BATCH_MMA");
    const bool mma_on = !(mma_env && mma_env[0] == '0');
    std::vector<int> group_of(nmaps, -1);
    std::vector<int> gtable;                   // G map indices per group
    std::vector<std::pair<int, int>> gshape;   // (K, G) per group
    std::vector<int> sgroup_of(nmaps, -1);
    std::vector<int> stable;                   // kAdjG map indices per matrix-core synthesis group (-1: empty)
    std::vector<int> sK;                       // its ring pairs per lane
    {
        std::map<std::tuple<int, int, int>, std::vector<int>> byg;
        for (int i = 0; i < nmaps; ++i)
            if (maps[i]) byg[std::make_tuple(maps[i]->nside, maps[i]->L, maps[i]->K1)].push_back(i);
        for (auto& kv : byg) {
            const std::vector<int>& v = kv.second;
            const int K = std::get<2>(kv.first);
            size_t at = 0;
            while (mma_on && (K == 1 || K == 2) && v.size() - at >= (size_t)kSynMinG) {
                const size_t n = std::min<size_t>(kAdjG, v.size() - at);
                const int sid = (int)sK.size();
                sK.push_back(K);
                for (size_t j = 0; j < (size_t)kAdjG; ++j) {
                    stable.push_back(j < n ? v[at + j] : -1);
                    if (j < n) sgroup_of[v[at + j]] = sid;
                }
                at += n;
            }
            for (int G : {4, 2}) {
                const bool ok = (G == 4 && (K == 1 || K == 2)) || (G == 2 && K <= 4);
                for (; ok && at + G <= v.size(); at += G) {
                    const int gid = (int)gshape.size();
                    gshape.push_back({K, G});
                    for (int j = 0; j < G; ++j) {
                        gtable.push_back(v[at + j]);
                        group_of[v[at + j]] = gid;
                    }
                    gtable.resize(gtable.size() + 4 - G, -1);  // table rows of 4
                }
            }
        }
    }
    // maps of one geometry whose adjoint runs one ring pair per lane share it
    // on the matrix cores in groups of kAdjG (CMBL_BATCH_MMA=0 turns this off)
    std::vector<int> agroup_of(nmaps, -1);
    std::vector<int> atable;  // kAdjG map indices per group
    {
        std::map<std::pair<int, int>, std::vector<int>> bya;
        for (int i = 0; mma_on && i < nmaps; ++i)
            if (maps[i] && maps[i]->K5 == 1) bya[std::make_pair(maps[i]->nside, maps[i]->L)].push_back(i);
        // full groups, then the rest as one group with empty columns when it
        // holds at least kAdjMinG maps (below that the one-pair-per-lane
        // adjoint of each map is cheaper than the products' fixed cost)
        for (auto& kv : bya)
            for (size_t at = 0; at < kv.second.size(); at += kAdjG) {
                const size_t n = std::min<size_t>(kAdjG, kv.second.size() - at);
                if (n < (size_t)kAdjMinG) break;
                const int gid = (int)(atable.size() / kAdjG);
                for (size_t j = 0; j < kAdjG; ++j) {
                    atable.push_back(j < n ? kv.second[at + j] : -1);
                    if (j < n) agroup_of[kv.second[at + j]] = gid;
                }
            }
    }
    std::vector<ItemCost> ws[2];  // items of the matrix-core synthesis by K - 1, map = its group
    std::vector<ItemCost> wa;  // items of the shared adjoint, map = group, sidx = entry of subs
    std::vector<MmaSub> subs;
    std::map<std::pair<int, int>, std::vector<ItemCost>> wg;  // (K, G) -> items
    for (int i = 0; i < nmaps; ++i) {
        cmbl_map* mp = maps[i];
        if (!mp) return cleanup(fail(CMBL_ERR_ARG, "null map in batch"));
        for (int j = 0; j < i; ++j)
            if (maps[j] == mp)
                return cleanup(fail(CMBL_ERR_ARG, "map " + std::to_string(i) + " appears twice in the batch (the "
                                                  "entries would share one workspace)"));
        b->maps.push_back(mp);
        descs.push_back(mp->desc);
        // maps of one nside and split threshold have the same Bluestein
        // chirps and filters (same rings, lengths and values): the batch's
        // ring stage reads one copy, which stays in the caches
        for (int j = 0; j < i; ++j)
            if (maps[j]->nside == mp->nside && maps[j]->tuning.split_m == mp->tuning.split_m) {
                descs.back().bs_filt = descs[j].bs_filt;
                descs.back().bs_chirp = descs[j].bs_chirp;
                descs.back().ring_bsofs = descs[j].ring_bsofs;
                descs.back().ring_chofs = descs[j].ring_chofs;
                break;
            }
        if (agroup_of[i] >= 0) {
            // the shared adjoint writes one partial per kMmaWaves ring blocks
            // (the last group may be partial), into the batch's own buffer
            // (the map's is laid out per block)
            const int nrb = (mp->npair_pad + 64 * kMmaWaves - 1) / (64 * kMmaWaves);
            const size_t n = (size_t)nrb * mp->nalm * 2;
            void* pb = nullptr;
            if (hipMalloc(&pb, sizeof(double) * n) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
            b->allocs.push_back(pb);
            if (hipMemset(pb, 0, sizeof(double) * n) != hipSuccess) return cleanup(fail(CMBL_ERR_HIP, "batch"));
            descs.back().upart = (double*)pb;
            descs.back().nrb5 = nrb;
        }
        int k1 = 0, k5 = 0;
        while ((1 << k1) < mp->K1) ++k1;
        while ((1 << k5) < mp->K5) ++k5;
        const int gid = group_of[i];
        const int sid = sgroup_of[i];
        if (sid >= 0) {
            if (stable[(size_t)kAdjG * sid] == i)  // the group's first map brings the group's items
                for (ItemCost ic : mp->cost1) {
                    ic.it.map = sid;
                    ws[mp->K1 - 1].push_back(ic);
                }
        } else if (gid < 0) {
            for (ItemCost ic : mp->cost1) {
                ic.it.map = i;
                w1[k1].push_back(ic);
            }
        } else if (gtable[4 * gid] == i) {  // the group's first map brings the group's items
            for (ItemCost ic : mp->cost1) {
                ic.it.map = gid;
                ic.cost *= 0.5 + 0.5 * gshape[gid].second;  // recurrence once, sums G times
                wg[gshape[gid]].push_back(ic);
            }
        }
        const int agid = agroup_of[i];
        if (agid < 0) {
            for (ItemCost ic : mp->cost5) {
                ic.it.map = i;
                w5[k5].push_back(ic);
            }
        } else if (atable[(size_t)kAdjG * agid] == i) {
            // the group's first map brings its items, merged per (m, kMmaWaves
            // consecutive ring blocks)
            std::map<std::pair<int, int>, size_t> at;
            for (const ItemCost& ic : mp->cost5) {
                const std::pair<int, int> key(ic.it.m0, ic.it.rb / kMmaWaves);
                auto f = at.find(key);
                if (f == at.end()) {
                    f = at.emplace(key, wa.size()).first;
                    ItemCost g = ic;
                    g.it.map = agid;
                    g.it.rb = key.second;
                    g.it.lend = mp->L;
                    g.it.sidx = (int)subs.size();
                    g.cost = 0.0;
                    wa.push_back(g);
                    MmaSub sb;
                    for (int w = 0; w < kMmaWaves; ++w) sb.lf[w] = sb.sidx[w] = -1;
                    subs.push_back(sb);
                }
                ItemCost& g = wa[f->second];
                MmaSub& sb = subs[g.it.sidx];
                sb.lf[ic.it.rb % kMmaWaves] = ic.it.lfirst;
                sb.sidx[ic.it.rb % kMmaWaves] = ic.it.sidx;
                g.it.lfirst = std::min(g.it.lfirst, ic.it.lfirst);
                g.cost += ic.cost;
            }
        }
        // the map's own two-ring workgroups are cut again for the batch's launches
        for (const RingItem& r : mp->ring_items) {
            rings.push_back(RingItem{i, r.pair});
            if (r.pair2 >= 0) rings.push_back(RingItem{i, r.pair2});
        }
        b->lds_ring = std::max(b->lds_ring, mp->lds_ring);
    }
    b->d_tw = maps[0]->d_tw;
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(MapDesc) * descs.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
    b->allocs.push_back(q);
    b->d_descs = (MapDesc*)q;
    if (hipMemcpy(b->d_descs, descs.data(), sizeof(MapDesc) * descs.size(), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
    for (int k = 0; k < 4; ++k) {
        for (int stage = 0; stage < 2; ++stage) {
            std::vector<ItemCost>& wc = stage ? w5[k] : w1[k];
            if (wc.empty()) continue;
            const std::vector<WorkItem> w = order_items(wc);
            if (hipMalloc(&q, sizeof(WorkItem) * w.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
            b->allocs.push_back(q);
            if (hipMemcpy(q, w.data(), sizeof(WorkItem) * w.size(), hipMemcpyHostToDevice) != hipSuccess)
                return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
            if (stage) {
                b->d_items5[k] = (WorkItem*)q;
                b->n5[k] = (int)w.size();
            } else {
                b->d_items1[k] = (WorkItem*)q;
                b->n1[k] = (int)w.size();
            }
        }
    }
    for (auto& kv : wg) {
        const std::vector<WorkItem> w = order_items(kv.second);
        if (hipMalloc(&q, sizeof(WorkItem) * w.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
        b->allocs.push_back(q);
        if (hipMemcpy(q, w.data(), sizeof(WorkItem) * w.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
        cmbl_batch::SynGroupList gl;
        gl.K = kv.first.first;
        gl.G = kv.first.second;
        gl.n = (int)w.size();
        gl.d_items = (WorkItem*)q;
        b->syn_groups.push_back(gl);
    }
    for (int k = 0; k < 2; ++k) {
        if (ws[k].empty()) continue;
        const std::vector<WorkItem> w = order_items(ws[k]);
        if (hipMalloc(&q, sizeof(WorkItem) * w.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
        b->allocs.push_back(q);
        if (hipMemcpy(q, w.data(), sizeof(WorkItem) * w.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
        b->d_items_smma[k] = (WorkItem*)q;
        b->n_smma[k] = (int)w.size();
    }
    if (!stable.empty()) {
        if (hipMalloc(&q, sizeof(int) * stable.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
        b->allocs.push_back(q);
        if (hipMemcpy(q, stable.data(), sizeof(int) * stable.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
        b->d_smaps = (int*)q;
    }
    if (!wa.empty()) {
        const std::vector<WorkItem> w = order_items(wa);
        if (hipMalloc(&q, sizeof(WorkItem) * w.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
        b->allocs.push_back(q);
        if (hipMemcpy(q, w.data(), sizeof(WorkItem) * w.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
        b->d_items_mma = (WorkItem*)q;
        b->n_mma = (int)w.size();
        if (hipMalloc(&q, sizeof(MmaSub) * subs.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
        b->allocs.push_back(q);
        if (hipMemcpy(q, subs.data(), sizeof(MmaSub) * subs.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
        b->d_subs_mma = (MmaSub*)q;
        if (hipMalloc(&q, sizeof(int) * atable.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
        b->allocs.push_back(q);
        if (hipMemcpy(q, atable.data(), sizeof(int) * atable.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
        b->d_amaps = (int*)q;
    }
    if (!gtable.empty()) {
        // rows of four entries; a G = 2 group reads the first two of its row
        if (hipMalloc(&q, sizeof(int) * gtable.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
        b->allocs.push_back(q);
        if (hipMemcpy(q, gtable.data(), sizeof(int) * gtable.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
        b->d_gmaps = (int*)q;
        b->ngroups = (int)gshape.size();
    }
    b->ring_classes = ring_classes(
        &rings, [&](const RingItem& r) { return maps[r.map]->ring_M[r.pair]; },
        [&](const RingItem& r) { return maps[r.map]->ring_two[r.pair]; });
    if (hipMalloc(&q, sizeof(RingItem) * rings.size()) != hipSuccess) return cleanup(fail(CMBL_ERR_ALLOC, "batch"));
    b->allocs.push_back(q);
    b->d_ring_items = (RingItem*)q;
    b->nring_items = (int)rings.size();
    if (hipMemcpy(q, rings.data(), sizeof(RingItem) * rings.size(), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(CMBL_ERR_HIP, "batch upload"));
    *out = b;
    return CMBL_OK;
}

int cmbl_batch_check_finite(cmbl_batch* b, void* stream, int* first_map)
{
    if (!b) return fail(CMBL_ERR_ARG, "null batch");
    if (first_map) *first_map = -1;
    CMBL_HIP(hipStreamSynchronize((hipStream_t)stream));
    int bad = -1;
    for (int i = 0; i < (int)b->maps.size(); ++i)
        if (take_nonfinite(b->maps[i]) != CMBL_OK && bad < 0) bad = i;
    if (bad < 0) return CMBL_OK;
    if (first_map) *first_map = bad;
    return fail(CMBL_ERR_NONFINITE, "map " + std::to_string(bad) + " of the batch produced a non-finite -2lnL");
}

void cmbl_batch_destroy(cmbl_batch* b)
{
    if (!b) return;
    (void)hipDeviceSynchronize();
    for (void* p : b->allocs) (void)hipFree(p);
    delete b;
}

int cmbl_batch_value_and_grad_device(cmbl_batch* b, const double* const* x, double* f, double* const* g,
                                     void* stream)
{
    if (!b || !x || !f || !g) return fail(CMBL_ERR_ARG, "null pointer");
    hipStream_t s = (hipStream_t)stream;
    const int n = (int)b->maps.size();
    for (int i = 0; i < n; ++i) {
        if (!b->maps[i]->has_data) return fail(CMBL_ERR_STATE, "map without data in batch");
        if (!x[i] || !g[i]) return fail(CMBL_ERR_ARG, "null pointer");
    }
    for (int i = 0; i < n; ++i) {
        const int rc = take_nonfinite(b->maps[i]);
        if (rc != CMBL_OK) return fail(rc, "map " + std::to_string(i) + " of the batch: " + g_err);
    }
    // groups of up to kBatchGroup maps share one launch of k_pre, k_post, k_final
    std::vector<BatchArgs> groups;
    for (int g0 = 0; g0 < n; g0 += kBatchGroup) {
        BatchArgs a{};
        a.n = std::min(kBatchGroup, n - g0);
        a.f = f + g0;
        a.tile0[0] = 0;
        for (int j = 0; j < a.n; ++j) {
            a.x[j] = x[g0 + j];
            a.g[j] = g[g0 + j];
            a.tile0[j + 1] = a.tile0[j] + lm_tiles(b->maps[g0 + j]->L);
        }
        groups.push_back(a);
    }
    for (size_t q = 0; q < groups.size(); ++q)
        CMBL_TRY(launch_pre_batch(b->d_descs + q * kBatchGroup, groups[q], s));
    for (int k = 0; k < 4; ++k) CMBL_TRY(launch_synth(1 << k, b->d_items1[k], b->n1[k], b->d_descs, s));
    for (const cmbl_batch::SynGroupList& gl : b->syn_groups)
        CMBL_TRY(launch_synth_group(gl.K, gl.G, gl.d_items, gl.n, b->d_descs, b->d_gmaps, s));
    for (int k = 0; k < 2; ++k) CMBL_TRY(launch_synth_mma(k + 1, b->d_items_smma[k], b->n_smma[k], b->d_descs, b->d_smaps, s));
    CMBL_TRY(run_ring(kRingLike, b->d_ring_items, b->ring_classes, b->d_descs, nullptr, nullptr, b->d_tw, s));
    CMBL_TRY(launch_adj_mma(b->d_items_mma, b->d_subs_mma, b->n_mma, b->d_descs, b->d_amaps, s));
    for (int k = 0; k < 4; ++k) CMBL_TRY(launch_adj(1 << k, b->d_items5[k], b->n5[k], b->d_descs, s));
    for (size_t q = 0; q < groups.size(); ++q) {
        CMBL_TRY(launch_post_batch(b->d_descs + q * kBatchGroup, groups[q], kPostLike, s));
        CMBL_TRY(launch_final_batch(b->d_descs + q * kBatchGroup, groups[q], s));
    }
    return CMBL_OK;
}

// ---------------------------------------------------------------- inputs
int cmbl_simulate_alm(const double* cl, int lmax, long seed, double* x)
{
    if (!cl || !x || lmax < 0) return fail(CMBL_ERR_ARG, "bad simulate_alm arguments");
    if (seed == 0) return fail(CMBL_ERR_ARG, "seed 0 means time(0) in the reference; pass a nonzero seed");
    std::mt19937 gen((int)seed);
    std::normal_distribution<> dist(0, 1);
    for (int l = 0; l <= lmax; ++l) {
        x[(long)l * l] = dist(gen) * std::sqrt(cl[l]);
        for (int m = 1; m <= l; ++m) {
            const double re = dist(gen) * std::sqrt(cl[l] / 2);
            const double im = dist(gen) * std::sqrt(cl[l] / 2);
            x[(long)l * l + 2 * m - 1] = re;
            x[(long)l * l + 2 * m] = im;
        }
    }
    return CMBL_OK;
}

int cmbl_simulate_white_noise(int nside, double sigma, long seed, double* map)
{
    if (!map || nside < 1) return fail(CMBL_ERR_ARG, "bad white-noise arguments");
    if (seed == 0) return fail(CMBL_ERR_ARG, "seed 0 means time(0) in the reference; pass a nonzero seed");
    std::mt19937 gen((int)seed);
    std::normal_distribution<> dist(0, sigma);
    const long npix = 12L * nside * nside;
    for (long i = 0; i < npix; ++i) map[i] = dist(gen);
    return CMBL_OK;
}

int cmbl_beam(int lmax, double fwhm, double* beam)
{
    if (!beam || lmax < 0) return fail(CMBL_ERR_ARG, "bad beam arguments");
    if (fwhm < 0) return fail(CMBL_ERR_ARG, "invalid fwhm");
    for (int l = 0; l <= lmax; ++l) {
        if (fwhm == 0) {
            beam[l] = 1.0;
            continue;
        }
        const double sigma = std::sqrt(8 * std::log(2.0)) / (fwhm * 3.141592653589793 / 180);
        beam[l] = std::exp(-l * (l + 1) / (2 * sigma * sigma));
    }
    return CMBL_OK;
}

int cmbl_pix2ang_ring(int nside, long pix, double* theta, double* phi)
{
    if (nside < 1 || pix < 0 || pix >= 12L * nside * nside || !theta || !phi)
        return fail(CMBL_ERR_ARG, "bad pix2ang arguments");
    pix2ang(nside, pix, theta, phi);
    return CMBL_OK;
}

int cmbl_mask_latitude(int nside, double cut, double* mask)
{
    if (!mask || nside < 1) return fail(CMBL_ERR_ARG, "bad mask arguments");
    const long npix = 12L * nside * nside;
    const double halfpi = 3.141592653589793 / 2;
    for (long p = 0; p < npix; ++p) {
        double th, ph;
        pix2ang(nside, p, &th, &ph);
        if (th < halfpi + cut && th > halfpi - cut) mask[p] = 0;
    }
    return CMBL_OK;
}

int cmbl_mask_regions(int nside, int nreg, const double* theta, const double* phi, const double* angle,
                      double* mask)
{
    if (!mask || nside < 1 || nreg < 0 || (nreg > 0 && (!theta || !phi || !angle)))
        return fail(CMBL_ERR_ARG, "bad mask arguments");
    const int nring = 4 * nside - 1;
    // ring theta table (pixel centres of a ring share theta)
    std::vector<double> rth(nring);
    std::vector<long> rofs(nring);
    std::vector<int> rn(nring);
    for (int r = 0; r < nring; ++r) {
        const RingGeom g = ring_geom(nside, r + 1);
        double th, ph;
        pix2ang(nside, g.ofs, &th, &ph);
        rth[r] = th;
        rofs[r] = g.ofs;
        rn[r] = g.nph;
    }
    for (int j = 0; j < nreg; ++j) {
        const double x = std::sin(theta[j]) * std::cos(phi[j]), y = std::sin(theta[j]) * std::sin(phi[j]),
                     z = std::cos(theta[j]);
        const double ca = std::cos(angle[j]);
        const double margin = 1e-9;
        for (int r = 0; r < nring; ++r) {
            if (std::fabs(rth[r] - theta[j]) > angle[j] + margin) continue;
            for (int i = 0; i < rn[r]; ++i) {
                const long p = rofs[r] + i;
                double th, ph;
                pix2ang(nside, p, &th, &ph);
                const double x1 = std::sin(th) * std::cos(ph), y1 = std::sin(th) * std::sin(ph), z1 = std::cos(th);
                const double dot = x1 * x + y1 * y + z1 * z;
                if (dot > ca) mask[p] = 0;
            }
        }
    }
    return CMBL_OK;
}

int cmbl_weights_from_mask(long npix, const double* mask, double sigma, double* w)
{
    if (!mask || !w || npix < 0 || !(sigma > 0)) return fail(CMBL_ERR_ARG, "bad weight arguments");
    for (long i = 0; i < npix; ++i) w[i] = mask[i] > 0.5 ? 1.0 / (sigma * sigma) : 0.0;
    return CMBL_OK;
}

int cmbl_whitenoise(int seed, double amplitude, long n, double* out)
{
    if (!out || n < 0) return fail(CMBL_ERR_ARG, "bad whitenoise arguments");
    // one process: the master seed list has a single entry
    std::mt19937 g1(seed);
    std::uniform_real_distribution<double> u(0, 1);
    const int s1 = static_cast<int>(std::ceil(1000000 * u(g1)));
    std::mt19937 g2(s1);
    std::normal_distribution<> gauss(0, 1);
    for (long i = 0; i < n; ++i) out[i] = amplitude * gauss(g2);
    return CMBL_OK;
}

// ---------------------------------------------------------------- device memory
int cmbl_device_alloc(void** ptr, unsigned long bytes)
{
    if (!ptr) return fail(CMBL_ERR_ARG, "ptr is null");
    *ptr = nullptr;
    CMBL_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return CMBL_OK;
}
int cmbl_device_free(void* ptr)
{
    if (ptr) CMBL_HIP(hipFree(ptr));
    return CMBL_OK;
}
int cmbl_copy_to_device(void* dst, const void* src, unsigned long bytes)
{
    if (bytes && (!dst || !src)) return fail(CMBL_ERR_ARG, "null pointer");
    if (bytes) CMBL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return CMBL_OK;
}
int cmbl_copy_to_host(void* dst, const void* src, unsigned long bytes)
{
    if (bytes && (!dst || !src)) return fail(CMBL_ERR_ARG, "null pointer");
    if (bytes) CMBL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return CMBL_OK;
}
int cmbl_copy_device(void* dst, const void* src, unsigned long bytes, void* stream)
{
    if (bytes && (!dst || !src)) return fail(CMBL_ERR_ARG, "null pointer");
    if (bytes) CMBL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return CMBL_OK;
}
int cmbl_stream_synchronize(void* stream)
{
    CMBL_HIP(hipStreamSynchronize((hipStream_t)stream));
    return CMBL_OK;
}

// ---------------------------------------------------------------- device vectors
int cmbl_vec_copy(double* y, const double* x, double c, long n, void* s)
{
    return launch_vec(0, y, x, c, n, (hipStream_t)s);
}
int cmbl_vec_set(double* y, double v, long n, void* s) { return launch_vec(1, y, nullptr, v, n, (hipStream_t)s); }
int cmbl_vec_axpy(double* y, const double* x, double c, long n, void* s)
{
    return launch_vec(2, y, x, c, n, (hipStream_t)s);
}
int cmbl_vec_mul(double* y, const double* x, long n, void* s) { return launch_vec(3, y, x, 0, n, (hipStream_t)s); }
int cmbl_vec_div(double* y, const double* x, long n, void* s) { return launch_vec(4, y, x, 0, n, (hipStream_t)s); }
int cmbl_vec_pow(double* y, double p, long n, void* s) { return launch_vec(5, y, nullptr, p, n, (hipStream_t)s); }

// Dot-product workspaces (kDotParts partial sums + 1 result), one per
// (device, stream): calls queued on one stream run in order and may share
// one, calls on different streams may not.
struct DotWorkspace {
    int device;
    hipStream_t stream;
    double* p;
};
static std::mutex g_dot_mu;
static std::vector<DotWorkspace> g_dot_ws;

static int dot_work(hipStream_t s, double** out)
{
    int dev = 0;
    CMBL_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_dot_mu);
    for (const DotWorkspace& w : g_dot_ws)
        if (w.device == dev && w.stream == s) {
            *out = w.p;
            return CMBL_OK;
        }
    double* p = nullptr;
    CMBL_HIP(hipMalloc(&p, sizeof(double) * (kDotParts + 1)));
    g_dot_ws.push_back(DotWorkspace{dev, s, p});
    *out = p;
    return CMBL_OK;
}

}  // extern "C"

namespace cmbl {
// Scratch of cmbl_vec_dot_device_static: the dot product's kDotParts partial
// sums (8 KiB), one copy on every GPU the library runs on, held until the
// process ends.  Host code never uses the array's name as an address: its
// device address is looked up once per device with hipGetSymbolAddress.
static __device__ __align__(256) double scratch[kDotParts];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

static int static_scratch(double** out)
{
    int dev = 0;
    CMBL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(CMBL_ERR_ARG, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        CMBL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CMBL_ERR_HIP, "hipGetSymbolAddress returned no address for the dot scratch");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return CMBL_OK;
}
}  // namespace cmbl

extern "C" {

int cmbl_vec_dot_device_static(const double* x, const double* y, long n, double* d_out, void* s)
{
    if (!x || !y || !d_out || n < 0) return fail(CMBL_ERR_ARG, "bad dot arguments");
    double* w = nullptr;
    CMBL_TRY(static_scratch(&w));
    return launch_dot(x, y, n, d_out, w, (hipStream_t)s);
}

int cmbl_vec_release_workspaces(void)
{
    std::lock_guard<std::mutex> lock(g_dot_mu);
    int cur = 0;
    CMBL_HIP(hipGetDevice(&cur));
    int rc = CMBL_OK;
    for (const DotWorkspace& w : g_dot_ws) {
        if (hipSetDevice(w.device) != hipSuccess || hipDeviceSynchronize() != hipSuccess || hipFree(w.p) != hipSuccess)
            rc = fail(CMBL_ERR_HIP, "freeing a dot-product workspace failed");
    }
    g_dot_ws.clear();
    (void)hipSetDevice(cur);
    return rc;
}

int cmbl_vec_dot_device(const double* x, const double* y, long n, double* d_out, void* s)
{
    if (!x || !y || !d_out || n < 0) return fail(CMBL_ERR_ARG, "bad dot arguments");
    double* w = nullptr;
    CMBL_TRY(dot_work((hipStream_t)s, &w));
    return launch_dot(x, y, n, d_out, w, (hipStream_t)s);
}

int cmbl_reorder_device(int nside, const double* in, double* out, int to_ring, void* stream)
{
    if (nside < 1 || (nside & (nside - 1))) return fail(CMBL_ERR_ARG, "nside must be a power of two");
    if (!in || !out || in == out) return fail(CMBL_ERR_ARG, "null or aliased buffers");
    return launch_reorder(in, out, nside, to_ring, (hipStream_t)stream);
}

int cmbl_reorder(int nside, const double* in, double* out, int to_ring)
{
    if (nside < 1 || (nside & (nside - 1))) return fail(CMBL_ERR_ARG, "nside must be a power of two");
    if (!in || !out) return fail(CMBL_ERR_ARG, "null pointer");
    const size_t bytes = sizeof(double) * 12 * (size_t)nside * nside;
    double *d_in = nullptr, *d_out = nullptr;
    CMBL_HIP(hipMalloc(&d_in, bytes));
    if (hipMalloc(&d_out, bytes) != hipSuccess) {
        (void)hipFree(d_in);
        return fail(CMBL_ERR_ALLOC, "hipMalloc failed for the reorder buffers");
    }
    int rc = CMBL_OK;
    if (hipMemcpy(d_in, in, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(CMBL_ERR_HIP, "copy to device failed");
    if (rc == CMBL_OK) rc = launch_reorder(d_in, d_out, nside, to_ring, nullptr);
    if (rc == CMBL_OK && hipMemcpy(out, d_out, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(CMBL_ERR_HIP, "copy to host failed");
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return rc;
}

int cmbl_vec_dot(const double* x, const double* y, long n, double* out, void* s)
{
    if (!x || !y || !out || n < 0) return fail(CMBL_ERR_ARG, "bad dot arguments");
    double* w = nullptr;
    CMBL_TRY(dot_work((hipStream_t)s, &w));
    CMBL_TRY(launch_dot(x, y, n, w + kDotParts, w, (hipStream_t)s));
    CMBL_HIP(hipMemcpyAsync(out, w + kDotParts, sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)s));
    CMBL_HIP(hipStreamSynchronize((hipStream_t)s));
    return CMBL_OK;
}

}  // extern "C"
