/*
 * CPU FP64 restatement of the Cosmo++ map-level CMB likelihood path.
 * TEST INFRASTRUCTURE ONLY (parity oracle and timed CPU baseline); the
 * product never links, loads or calls this file.
 *
 * Functions, the reference they follow, and what pins each one
 * (tests/test_oracle_*.py):
 *   orc_ring_info / orc_pix2ang_ring  HEALPix RING geometry (Gorski et al. 2005;
 *       used at source/utils.cpp:227,243, source/cmb_gibbs.cpp:136).
 *       Pinned: ring table vs per-pixel pix2ang formula, sum nph = Npix,
 *       equal-area invariant, north/south symmetry.
 *   orc_lambda  normalised lambda_lm = Y_lm(theta,0) with Condon-Shortley
 *       phase, as include/spherical_harmonics.hpp:29-78.  Pinned: the
 *       reference KATs source/test_spherical_harmonics.cpp:28-76 and
 *       source/test_legendre.cpp:27-107 (via lambda <-> P_lm), brute force
 *       against the reference SphericalHarmonics (oracle/_ref), the
 *       addition theorem sum_m |Y_lm|^2 = (2l+1)/4pi, sectoral closed form.
 *   orc_fft  complex DFT (radix-2 + Bluestein).  Pinned: naive DFT, Parseval.
 *   orc_synthesis  HEALPix alm2map (called at source/cmb_gibbs.cpp:255,378).
 *       Pinned: brute force sum_lm a_lm Y_lm(theta_p, phi_p) with Y from the
 *       reference SphericalHarmonics (golden fixture), closed-form monopole /
 *       dipole maps (source/cmb_gibbs.cpp:133-143).
 *   orc_adjoint  map2alm with ring weight Npix/4pi = exact transpose
 *       (source/cmb_gibbs.cpp:390-391).  Pinned: adjoint identity
 *       <Y a, f> = <a, Y^T f> and brute force.
 *   orc_neg2lnL  -2lnL and gradient (composition of source/cmb_gibbs.cpp:
 *       381-387 weighting, :541-549 sigma_l with m>0 doubled, :180-181
 *       C_0=C_1=0).  Pinned: exact quadratic => central finite differences,
 *       w = 0 (prior only) and noiseless-data (data term 0) closed cases.
 *   orc_alm2map_pol  HEALPix alm2map_pol (called at source/likelihood.cpp:585,
 *       source/c_matrix_generator.cpp:577-578).  Pinned: brute force against
 *       spin-weighted harmonics from the explicit sum of Goldberg et al. 1967
 *       (itself pinned by s = 0 -> Y_lm and by the eth-operator definition
 *       2Y = sqrt((l-2)!/(l+2)!) eth eth Y), the closed form of the l = 2,
 *       m = 0 E mode, E <-> B rotation.  Parity with HEALPix's own Q/U
 *       sign convention rests on its documentation (Q + iU = sum a_2 2Y,
 *       a^E = -(a_2 + a_-2)/2, a^B = i (a_2 - a_-2)/2): unpinned by any
 *       reference file or run.
 *   orc_beam  Utils::beamFunction (source/utils.cpp:54-64).  Closed form.
 *   orc_mask_latitude / orc_mask_regions  source/test_like_high.cpp:123-130,
 *       Utils::maskRegions (source/utils.cpp:235-253).  Pinned: brute-force
 *       pixel counts and spherical-cap area.
 *   orc_gauss  Math::GaussianGenerator(seed, 0, 1) draws (include/random.hpp:
 *       32-52: std::mt19937 + std::normal_distribution<double>), restated as
 *       MT19937 (Matsumoto & Nishimura 1998) feeding 53-bit uniforms made of
 *       two 32-bit words, then the Marsaglia polar method returning the
 *       second variate of each pair first.  Pinned: the reference's own draws
 *       for fixed seeds (tests/golden/rng.npz) and, where oracle/_ref builds,
 *       the reference generator run live.
 * Bit-level parity with libsharp's summation order is not reproducible here
 * (HEALPix is absent); agreement is asserted to 1e-10 relative.
 */
#include "cmbl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793238462643383279502884

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* Gaussian draws                                                      */
/* ------------------------------------------------------------------ */

typedef struct {
    unsigned int mt[624];
    int idx;
    int have_saved;
    double saved;
} orc_rng;

static void rng_seed(orc_rng* g, unsigned int seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (unsigned int)i;
    g->idx = 624;
    g->have_saved = 0;
}

static unsigned int rng_u32(orc_rng* g)
{
    if (g->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            const unsigned int y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->idx = 0;
    }
    unsigned int y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* Uniform in [0, 1) from two 32-bit words, low word first. */
static double rng_unit(orc_rng* g)
{
    const double lo = (double)rng_u32(g);
    const double hi = (double)rng_u32(g);
    double u = (lo + hi * 4294967296.0) / 18446744073709551616.0;
    if (u >= 1.0) u = nextafter(1.0, 0.0);
    return u;
}

static double rng_normal(orc_rng* g)
{
    if (g->have_saved) {
        g->have_saved = 0;
        return g->saved;
    }
    double x, y, r2;
    do {
        x = 2.0 * rng_unit(g) - 1.0;
        y = 2.0 * rng_unit(g) - 1.0;
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double f = sqrt(-2.0 * log(r2) / r2);
    g->saved = x * f;
    g->have_saved = 1;
    return y * f;
}

void orc_gauss(long seed, long n, double* out)
{
    orc_rng g;
    rng_seed(&g, (unsigned int)(int)seed);
    for (long i = 0; i < n; ++i) out[i] = rng_normal(&g);
}

/* ------------------------------------------------------------------ */
/* Geometry                                                            */
/* ------------------------------------------------------------------ */

int orc_nring(int nside) { return 4 * nside - 1; }

void orc_ring_info(int nside, int ring, orc_ring* out)
{
    const long N = nside;
    const long npix = 12 * N * N;
    const long i = ring + 1; /* 1-based ring number, north to south */
    const long in = (i <= 2 * N) ? i : 4 * N - i; /* northern mirror ring */
    double z, sint, phi0;
    long nph, ofs;
    if (in < N) {
        /* polar cap: 1 - z = in^2 / (3 N^2) */
        const double t = (double)(in * in) / (3.0 * (double)N * (double)N);
        z = 1.0 - t;
        sint = sqrt(t * (2.0 - t));
        nph = 4 * in;
        phi0 = ORC_PI / (double)nph;
        ofs = 2 * in * (in - 1);
    } else {
        z = (double)(2 * N - in) * 2.0 / (3.0 * (double)N);
        sint = sqrt((1.0 - z) * (1.0 + z));
        nph = 4 * N;
        phi0 = ((in - N) % 2 == 0) ? ORC_PI / (double)nph : 0.0;
        ofs = 2 * N * (N - 1) + (in - N) * 4 * N;
    }
    if (i > 2 * N) {
        z = -z;
        ofs = npix - ofs - nph;
    }
    out->z = z;
    out->sint = sint;
    out->phi0 = phi0;
    out->nph = nph;
    out->ofs = ofs;
}

static long isqrt_l(long v)
{
    long r = (long)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

void orc_pix2ang_ring(int nside, long pix, double* theta, double* phi)
{
    const long N = nside;
    const long npix = 12 * N * N;
    const long ncap = 2 * N * (N - 1);
    const double halfpi = 0.5 * ORC_PI;
    if (pix < ncap) {
        const long iring = (1 + isqrt_l(1 + 2 * pix)) >> 1;
        const long iphi = (pix + 1) - 2 * iring * (iring - 1);
        const double z = 1.0 - (double)(iring * iring) / (3.0 * (double)N * (double)N);
        *theta = acos(z);
        *phi = ((double)iphi - 0.5) * halfpi / (double)iring;
    } else if (pix < npix - ncap) {
        const long ip = pix - ncap;
        const long iring = ip / (4 * N) + N;
        const long iphi = ip % (4 * N) + 1;
        const double fodd = ((iring + N) & 1) ? 1.0 : 0.5;
        const double z = (double)(2 * N - iring) * 2.0 / (3.0 * (double)N);
        *theta = acos(z);
        *phi = ((double)iphi - fodd) * halfpi / (double)N;
    } else {
        const long ip = npix - pix;
        const long iring = (1 + isqrt_l(2 * ip - 1)) >> 1;
        const long iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
        const double z = -1.0 + (double)(iring * iring) / (3.0 * (double)N * (double)N);
        *theta = acos(z);
        *phi = ((double)iphi - 0.5) * halfpi / (double)iring;
    }
}

/* ------------------------------------------------------------------ */
/* Normalised associated Legendre recursion with exponent scaling      */
/* ------------------------------------------------------------------ */
/* true value = v * 2^(-ORC_SCL * e).  While e > 0 the true value is below
 * 2^(ORC_HI - ORC_SCL) = 2^-200 and is treated as 0. */
#define ORC_SCL 600
#define ORC_HI 400
#define ORC_LO (-400)

typedef struct {
    double v1, v2; /* lambda_{l}, lambda_{l-1} (scaled) */
    int e;
    int l;
    int m;
    double x;
} orc_lam;

static void lam_init(orc_lam* st, int m, double x, double sint)
{
    double v = 1.0 / sqrt(4.0 * ORC_PI);
    int e = 0;
    const double up = ldexp(1.0, ORC_SCL);
    const double lo = ldexp(1.0, ORC_LO);
    for (int k = 1; k <= m; ++k) {
        v *= -sqrt((2.0 * k - 1.0) / (2.0 * k)) * sint;
        if (fabs(v) < lo && v != 0.0) {
            v *= up;
            ++e;
        }
    }
    v *= sqrt(2.0 * m + 1.0);
    st->v1 = v;
    st->v2 = 0.0;
    st->e = e;
    st->l = m;
    st->m = m;
    st->x = x;
}

static inline double lam_value(const orc_lam* st) { return st->e == 0 ? st->v1 : 0.0; }

static inline void lam_step(orc_lam* st)
{
    const int l = st->l + 1;
    const int m = st->m;
    double v;
    if (l == m + 1) {
        v = sqrt(2.0 * m + 3.0) * st->x * st->v1;
    } else {
        const double a = sqrt((4.0 * l * (double)l - 1.0) / ((double)l * l - (double)m * m));
        const double lm1 = l - 1.0;
        const double binv = sqrt((lm1 * lm1 - (double)m * m) / (4.0 * lm1 * lm1 - 1.0));
        v = a * (st->x * st->v1 - binv * st->v2);
    }
    st->v2 = st->v1;
    st->v1 = v;
    st->l = l;
    if (st->e > 0) {
        const double av = fabs(v);
        if (av > ldexp(1.0, ORC_HI)) {
            const double dn = ldexp(1.0, -ORC_SCL);
            st->v1 *= dn;
            st->v2 *= dn;
            --st->e;
        } else if (av < ldexp(1.0, ORC_LO) && fabs(st->v2) < ldexp(1.0, ORC_LO) && av != 0.0) {
            const double up = ldexp(1.0, ORC_SCL);
            st->v1 *= up;
            st->v2 *= up;
            ++st->e;
        }
    }
}

void orc_lambda(int L, int m, double x, double sint, double* out)
{
    orc_lam st;
    lam_init(&st, m, x, sint);
    out[0] = lam_value(&st);
    for (int l = m + 1; l <= L; ++l) {
        lam_step(&st);
        out[l - m] = lam_value(&st);
    }
}

/* ------------------------------------------------------------------ */
/* FFT                                                                 */
/* ------------------------------------------------------------------ */

void orc_dft(int n, const double* re, const double* im, double* ore, double* oim, int sign)
{
    for (int k = 0; k < n; ++k) {
        long double sr = 0, si = 0;
        for (int j = 0; j < n; ++j) {
            const long r = ((long)j * k) % n;
            const double ang = sign * 2.0 * ORC_PI * (double)r / (double)n;
            const double c = cos(ang), s = sin(ang);
            sr += (long double)re[j] * c - (long double)im[j] * s;
            si += (long double)re[j] * s + (long double)im[j] * c;
        }
        ore[k] = (double)sr;
        oim[k] = (double)si;
    }
}

static void fft_pow2(int n, double* re, double* im, int sign)
{
    if (n <= 1) return;
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    double* twr = (double*)malloc(sizeof(double) * (n / 2));
    double* twi = (double*)malloc(sizeof(double) * (n / 2));
    for (int k = 0; k < n / 2; ++k) {
        const double ang = 2.0 * ORC_PI * (double)k / (double)n;
        twr[k] = cos(ang);
        twi[k] = sign * sin(ang);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        const int stride = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; ++k) {
                const double wr = twr[k * stride], wi = twi[k * stride];
                const int a = i + k, b = i + k + half;
                const double vr = re[b] * wr - im[b] * wi;
                const double vi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - vr;
                im[b] = im[a] - vi;
                re[a] += vr;
                im[a] += vi;
            }
        }
    }
    free(twr);
    free(twi);
}

void orc_fft(int n, double* re, double* im, int sign)
{
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        fft_pow2(n, re, im, sign);
        return;
    }
    /* Bluestein: jk = (j^2 + k^2 - (k-j)^2) / 2 */
    int M = 1;
    while (M < 2 * n - 1) M <<= 1;
    double* cr = (double*)malloc(sizeof(double) * n);
    double* ci = (double*)malloc(sizeof(double) * n);
    for (int j = 0; j < n; ++j) {
        const long r = ((long)j * j) % (2L * n);
        const double ang = sign * ORC_PI * (double)r / (double)n;
        cr[j] = cos(ang);
        ci[j] = sin(ang);
    }
    double* ar = (double*)calloc(M, sizeof(double));
    double* ai = (double*)calloc(M, sizeof(double));
    double* br = (double*)calloc(M, sizeof(double));
    double* bi = (double*)calloc(M, sizeof(double));
    for (int j = 0; j < n; ++j) {
        ar[j] = re[j] * cr[j] - im[j] * ci[j];
        ai[j] = re[j] * ci[j] + im[j] * cr[j];
    }
    br[0] = cr[0];
    bi[0] = -ci[0];
    for (int j = 1; j < n; ++j) {
        br[j] = br[M - j] = cr[j];
        bi[j] = bi[M - j] = -ci[j];
    }
    fft_pow2(M, ar, ai, -1);
    fft_pow2(M, br, bi, -1);
    for (int k = 0; k < M; ++k) {
        const double tr = ar[k] * br[k] - ai[k] * bi[k];
        const double ti = ar[k] * bi[k] + ai[k] * br[k];
        ar[k] = tr;
        ai[k] = ti;
    }
    fft_pow2(M, ar, ai, +1);
    const double inv = 1.0 / (double)M;
    for (int k = 0; k < n; ++k) {
        const double tr = ar[k] * inv, ti = ai[k] * inv;
        re[k] = tr * cr[k] - ti * ci[k];
        im[k] = tr * ci[k] + ti * cr[k];
    }
    free(cr); free(ci); free(ar); free(ai); free(br); free(bi);
}

/* e^{i m phi0} with the argument reduced exactly (phi0 is 0 or pi/nph) */
static void phase(long m, const orc_ring* r, double* c, double* s)
{
    if (r->phi0 == 0.0) {
        *c = 1.0;
        *s = 0.0;
        return;
    }
    const long two_n = 2 * r->nph;
    long q = m % two_n;
    if (q < 0) q += two_n;
    const double ang = ORC_PI * (double)q / (double)r->nph;
    *c = cos(ang);
    *s = sin(ang);
}

/* ------------------------------------------------------------------ */
/* Synthesis and its transpose                                         */
/* ------------------------------------------------------------------ */

/* m-major complex coefficients B_l a_lm: index off(m) + (l - m) */
static long moff(int L, int m) { return (long)m * (L + 1) - (long)m * (m - 1) / 2; }

/* Real values of ring r from its coefficients F_m, m = 0..L:
 * f(phi_j) = F_0 + 2 Re sum_{m>=1} F_m e^{i m phi_j} (the real-field sum over
 * +-m), by folding m onto the ring's nph bins and one DFT. */
static void ring_from_F(int nside, int L, int r, const double* Fr, const double* Fi, double* out)
{
    orc_ring rg;
    orc_ring_info(nside, r, &rg);
    const long n = rg.nph;
    double* gr = (double*)calloc(n, sizeof(double));
    double* gi = (double*)calloc(n, sizeof(double));
    for (long m = 0; m <= L; ++m) {
        double c, s;
        phase(m, &rg, &c, &s);
        const double fr = Fr[m], fi = Fi[m];
        const double pr = fr * c - fi * s, pi = fr * s + fi * c;
        const long b = m % n;
        gr[b] += pr;
        gi[b] += pi;
        if (m > 0) {
            const long b2 = ((-m) % n + n) % n;
            gr[b2] += pr;
            gi[b2] -= pi;
        }
    }
    orc_fft((int)n, gr, gi, +1);
    for (long j = 0; j < n; ++j) out[j] = gr[j];
    free(gr);
    free(gi);
}

/* Every ring of the map from F[ring][m]. */
static void rings_from_F(int nside, int L, const double* Fr, const double* Fi, double* map)
{
    const int nring = orc_nring(nside);
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nring; ++r) {
        orc_ring rg;
        orc_ring_info(nside, r, &rg);
        ring_from_F(nside, L, r, Fr + (size_t)r * (L + 1), Fi + (size_t)r * (L + 1), map + rg.ofs);
    }
}

/* m-major complex B_l a_lm from the packed real vector. */
static void pack_to_mmajor(int L, const double* x, const double* beam, double* ar, double* ai)
{
    for (int m = 0; m <= L; ++m)
        for (int l = m; l <= L; ++l) {
            const double b = beam ? beam[l] : 1.0;
            const long k = moff(L, m) + (l - m);
            if (m == 0) {
                ar[k] = b * x[(long)l * l];
                ai[k] = 0.0;
            } else {
                ar[k] = b * x[(long)l * l + 2 * m - 1];
                ai[k] = b * x[(long)l * l + 2 * m];
            }
        }
}

/* Ring coefficients F_m of ring pair p (north ring p, south ring nring-1-p):
 * F_N = E + O, F_S = E - O from the even / odd (l - m) parity sums of
 * lambda_lm(theta_p) B_l a_lm.  Fs may be NULL (the equator has no mirror). */
static void pair_F(int nside, int L, int p, const double* ar, const double* ai, double* Fnr, double* Fni,
                   double* Fsr, double* Fsi)
{
    orc_ring rg;
    orc_ring_info(nside, p, &rg);
    for (int m = 0; m <= L; ++m) {
        orc_lam st;
        lam_init(&st, m, rg.z, rg.sint);
        double er = 0, ei = 0, orr = 0, oi = 0;
        const long base = moff(L, m) - m;
        for (int l = m; l <= L; ++l) {
            if (l > m) lam_step(&st);
            const double lam = lam_value(&st);
            if (lam != 0.0) {
                if (((l - m) & 1) == 0) {
                    er += lam * ar[base + l];
                    ei += lam * ai[base + l];
                } else {
                    orr += lam * ar[base + l];
                    oi += lam * ai[base + l];
                }
            }
        }
        Fnr[m] = er + orr;
        Fni[m] = ei + oi;
        if (Fsr) {
            Fsr[m] = er - orr;
            Fsi[m] = ei - oi;
        }
    }
}

void orc_synthesis(int nside, int L, const double* x, const double* beam, double* map)
{
    const int nring = orc_nring(nside);
    const int npair = 2 * nside;
    const long nalm = (long)(L + 1) * (L + 2) / 2;
    double* ar = (double*)malloc(sizeof(double) * nalm);
    double* ai = (double*)malloc(sizeof(double) * nalm);
    pack_to_mmajor(L, x, beam, ar, ai);
    /* F[ring][m] for both rings of each pair */
    double* Fr = (double*)malloc(sizeof(double) * (size_t)nring * (L + 1));
    double* Fi = (double*)malloc(sizeof(double) * (size_t)nring * (L + 1));

#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < npair; ++p) {
        const int rn = p, rs = nring - 1 - p; /* rs == rn for the equator */
        const size_t on = (size_t)rn * (L + 1), os = (size_t)rs * (L + 1);
        pair_F(nside, L, p, ar, ai, Fr + on, Fi + on, rs != rn ? Fr + os : NULL, rs != rn ? Fi + os : NULL);
    }

    rings_from_F(nside, L, Fr, Fi, map);
    free(Fr); free(Fi); free(ar); free(ai);
}

void orc_synthesis_rings(int nside, int L, const double* x, const double* beam, int nsel, const int* rings,
                         double* out)
{
    const int nring = orc_nring(nside);
    const long nalm = (long)(L + 1) * (L + 2) / 2;
    double* ar = (double*)malloc(sizeof(double) * nalm);
    double* ai = (double*)malloc(sizeof(double) * nalm);
    pack_to_mmajor(L, x, beam, ar, ai);
    long* ofs = (long*)malloc(sizeof(long) * (nsel + 1));
    ofs[0] = 0;
    for (int i = 0; i < nsel; ++i) {
        orc_ring rg;
        orc_ring_info(nside, rings[i], &rg);
        ofs[i + 1] = ofs[i] + rg.nph;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nsel; ++i) {
        const int r = rings[i];
        const int north = r < 2 * nside;
        const int p = north ? r : nring - 1 - r;
        double* F = (double*)malloc(sizeof(double) * 4 * (L + 1));
        double *Fnr = F, *Fni = F + (L + 1), *Fsr = F + 2 * (L + 1), *Fsi = F + 3 * (L + 1);
        const int has_south = p != nring - 1 - p;
        pair_F(nside, L, p, ar, ai, Fnr, Fni, has_south ? Fsr : NULL, has_south ? Fsi : NULL);
        if (north)
            ring_from_F(nside, L, r, Fnr, Fni, out + ofs[i]);
        else
            ring_from_F(nside, L, r, Fsr, Fsi, out + ofs[i]);
        free(F);
    }
    free(ofs); free(ar); free(ai);
}

/* HEALPix alm2map_pol (external; called at source/likelihood.cpp:585 and
 * source/c_matrix_generator.cpp:577-578): T by orc_synthesis, and
 *   Q_m = -sum_l (a^E F1_lm + i a^B F2_lm),  U_m = -sum_l (a^B F1_lm - i a^E F2_lm)
 * per ring, F1 and F2 evaluated for each l from lambda_lm and lambda_{l-1,m}
 * by the relations of the spin-2 harmonics to the scalar ones
 * (W = (2Y + -2Y)/2 = F1 e^{i m phi}, X = (2Y - -2Y)/2 = F2 e^{i m phi}):
 *   F1 = kappa_l [-((l - m^2)/s^2 + l(l-1)/2) lambda_l + (c/s^2) q_lm lambda_{l-1}]
 *   F2 = kappa_l (m/s^2) [-(l-1) c lambda_l + q_lm lambda_{l-1}]
 * kappa_l = 2/sqrt((l-1) l (l+1) (l+2)), q_lm = sqrt((2l+1)/(2l-1) (l^2 - m^2)).
 * The south ring of a pair takes c -> -c and lambda_l -> (-1)^{l+m} lambda_l. */
void orc_alm2map_pol(int nside, int L, const double* t, const double* e, const double* b, double* mt,
                     double* mq, double* mu)
{
    orc_synthesis(nside, L, t, NULL, mt);
    const int nring = orc_nring(nside);
    const int npair = 2 * nside;
    const long nalm = (long)(L + 1) * (L + 2) / 2;
    double* er = (double*)calloc(nalm, sizeof(double));
    double* ei = (double*)calloc(nalm, sizeof(double));
    double* br = (double*)calloc(nalm, sizeof(double));
    double* bi = (double*)calloc(nalm, sizeof(double));
    for (int m = 0; m <= L; ++m)
        for (int l = m; l <= L; ++l) {
            const long k = moff(L, m) + (l - m);
            const long i = (long)l * l + (m == 0 ? 0 : 2 * m - 1);
            er[k] = e[i];
            br[k] = b[i];
            ei[k] = m == 0 ? 0.0 : e[i + 1];
            bi[k] = m == 0 ? 0.0 : b[i + 1];
        }
    const size_t nf = (size_t)nring * (L + 1);
    double* Qr = (double*)calloc(nf, sizeof(double));
    double* Qi = (double*)calloc(nf, sizeof(double));
    double* Ur = (double*)calloc(nf, sizeof(double));
    double* Ui = (double*)calloc(nf, sizeof(double));

#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < npair; ++p) {
        orc_ring rg;
        orc_ring_info(nside, p, &rg);
        const int rn = p, rs = nring - 1 - p;
        const double s2 = rg.sint * rg.sint;
        for (int m = 0; m <= L; ++m) {
            orc_lam st;
            lam_init(&st, m, rg.z, rg.sint);
            double acc[2][4] = {{0}}; /* [north/south][Q.re, Q.im, U.re, U.im] */
            const long base = moff(L, m) - m;
            const double fm = m;
            for (int l = m; l <= L; ++l) {
                if (l > m) lam_step(&st);
                if (l < 2) continue;
                const double lam = lam_value(&st);
                const double lamp = (l > m && st.e == 0) ? st.v2 : 0.0;
                if (lam == 0.0 && lamp == 0.0) continue;
                const double fl = l;
                const double kap = 2.0 / sqrt((fl - 1.0) * fl * (fl + 1.0) * (fl + 2.0));
                const double q = sqrt((2.0 * fl + 1.0) / (2.0 * fl - 1.0) * (fl * fl - fm * fm));
                const double ar = er[base + l], ai = ei[base + l], cr = br[base + l], ci = bi[base + l];
                for (int h = 0; h < 2; ++h) {
                    /* south ring: (-1)^{l+m} on lambda_l, the opposite sign on lambda_{l-1} */
                    const double sl = (h == 1 && ((l + m) & 1)) ? -1.0 : 1.0;
                    const double c = h == 0 ? rg.z : -rg.z;
                    const double ll = sl * lam, lpp = (h == 0 ? 1.0 : -sl) * lamp;
                    const double f1 = kap * (-((fl - fm * fm) / s2 + 0.5 * fl * (fl - 1.0)) * ll + c / s2 * q * lpp);
                    const double f2 = kap * (fm / s2) * (-(fl - 1.0) * c * ll + q * lpp);
                    acc[h][0] -= ar * f1 - ci * f2;
                    acc[h][1] -= ai * f1 + cr * f2;
                    acc[h][2] -= cr * f1 + ai * f2;
                    acc[h][3] -= ci * f1 - ar * f2;
                }
            }
            Qr[(size_t)rn * (L + 1) + m] = acc[0][0];
            Qi[(size_t)rn * (L + 1) + m] = acc[0][1];
            Ur[(size_t)rn * (L + 1) + m] = acc[0][2];
            Ui[(size_t)rn * (L + 1) + m] = acc[0][3];
            if (rs != rn) {
                Qr[(size_t)rs * (L + 1) + m] = acc[1][0];
                Qi[(size_t)rs * (L + 1) + m] = acc[1][1];
                Ur[(size_t)rs * (L + 1) + m] = acc[1][2];
                Ui[(size_t)rs * (L + 1) + m] = acc[1][3];
            }
        }
    }
    rings_from_F(nside, L, Qr, Qi, mq);
    rings_from_F(nside, L, Ur, Ui, mu);
    free(Qr); free(Qi); free(Ur); free(Ui);
    free(er); free(ei); free(br); free(bi);
}

/* u_lm = sum_p f_p lambda_lm(theta_p) e^{-i m phi_p}, m-major complex */
static void orc_adjoint_u(int nside, int L, const double* f, double* ur, double* ui)
{
    const int nring = orc_nring(nside);
    const int npair = 2 * nside;
    double* Pr = (double*)malloc(sizeof(double) * (size_t)nring * (L + 1));
    double* Pi = (double*)malloc(sizeof(double) * (size_t)nring * (L + 1));

#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nring; ++r) {
        orc_ring rg;
        orc_ring_info(nside, r, &rg);
        const long n = rg.nph;
        double* hr = (double*)malloc(sizeof(double) * n);
        double* hi = (double*)calloc(n, sizeof(double));
        for (long j = 0; j < n; ++j) hr[j] = f[rg.ofs + j];
        orc_fft((int)n, hr, hi, -1);
        for (long m = 0; m <= L; ++m) {
            double c, s;
            phase(m, &rg, &c, &s);
            const long b = m % n;
            /* e^{-i m phi0} H_b */
            Pr[(size_t)r * (L + 1) + m] = hr[b] * c + hi[b] * s;
            Pi[(size_t)r * (L + 1) + m] = hi[b] * c - hr[b] * s;
        }
        free(hr);
        free(hi);
    }

#pragma omp parallel for schedule(dynamic, 1)
    for (int m = 0; m <= L; ++m) {
        const long base = moff(L, m) - m;
        for (int l = m; l <= L; ++l) {
            ur[base + l] = 0.0;
            ui[base + l] = 0.0;
        }
        for (int p = 0; p < npair; ++p) {
            orc_ring rg;
            orc_ring_info(nside, p, &rg);
            const int rn = p, rs = nring - 1 - p;
            double er = Pr[(size_t)rn * (L + 1) + m], ei = Pi[(size_t)rn * (L + 1) + m];
            double orr = er, oi = ei;
            if (rs != rn) {
                const double sr = Pr[(size_t)rs * (L + 1) + m], si = Pi[(size_t)rs * (L + 1) + m];
                er += sr; ei += si;
                orr -= sr; oi -= si;
            }
            orc_lam st;
            lam_init(&st, m, rg.z, rg.sint);
            for (int l = m; l <= L; ++l) {
                if (l > m) lam_step(&st);
                const double lam = lam_value(&st);
                if (lam != 0.0) {
                    if (((l - m) & 1) == 0) {
                        ur[base + l] += lam * er;
                        ui[base + l] += lam * ei;
                    } else {
                        ur[base + l] += lam * orr;
                        ui[base + l] += lam * oi;
                    }
                }
            }
        }
    }
    free(Pr);
    free(Pi);
}

void orc_adjoint(int nside, int L, const double* map, const double* beam, double* out)
{
    const long nalm = (long)(L + 1) * (L + 2) / 2;
    double* ur = (double*)malloc(sizeof(double) * nalm);
    double* ui = (double*)malloc(sizeof(double) * nalm);
    orc_adjoint_u(nside, L, map, ur, ui);
    for (int l = 0; l <= L; ++l) {
        const double b = beam ? beam[l] : 1.0;
        out[(long)l * l] = b * ur[moff(L, 0) + l];
        for (int m = 1; m <= l; ++m) {
            const long k = moff(L, m) + (l - m);
            out[(long)l * l + 2 * m - 1] = 2.0 * b * ur[k];
            out[(long)l * l + 2 * m] = 2.0 * b * ui[k];
        }
    }
    free(ur);
    free(ui);
}

double orc_neg2lnL(int nside, int L, const double* x, const double* beam, const double* cl,
                   const double* d, const double* w, double* g)
{
    const long npix = 12L * nside * nside;
    const long n = (long)(L + 1) * (L + 1);
    double* s = (double*)malloc(sizeof(double) * npix);
    orc_synthesis(nside, L, x, beam, s);
    long double chi2 = 0;
    for (long p = 0; p < npix; ++p) {
        const double r = d[p] - s[p];
        const double f = w[p] * r;
        chi2 += (long double)f * r;
        s[p] = f; /* reuse as the weighted residual */
    }
    long double prior = 0;
    for (int l = 2; l <= L; ++l) {
        if (!(cl[l] > 0)) continue;
        long double t = (long double)x[(long)l * l] * x[(long)l * l];
        for (int m = 1; m <= l; ++m) {
            const double re = x[(long)l * l + 2 * m - 1], im = x[(long)l * l + 2 * m];
            t += 2.0L * ((long double)re * re + (long double)im * im);
        }
        prior += t / cl[l];
    }
    if (g) {
        orc_adjoint(nside, L, s, beam, g); /* J^T (w r) */
        for (long k = 0; k < n; ++k) g[k] *= -2.0;
        for (int l = 2; l <= L; ++l) {
            if (!(cl[l] > 0)) continue;
            g[(long)l * l] += 2.0 * x[(long)l * l] / cl[l];
            for (int m = 1; m <= l; ++m) {
                g[(long)l * l + 2 * m - 1] += 4.0 * x[(long)l * l + 2 * m - 1] / cl[l];
                g[(long)l * l + 2 * m] += 4.0 * x[(long)l * l + 2 * m] / cl[l];
            }
        }
    }
    free(s);
    return (double)(chi2 + prior);
}

/* ------------------------------------------------------------------ */
/* Inputs                                                              */
/* ------------------------------------------------------------------ */

void orc_beam(int L, double fwhm, double* beam)
{
    for (int l = 0; l <= L; ++l) {
        if (fwhm == 0) {
            beam[l] = 1.0;
            continue;
        }
        const double sigma = sqrt(8 * log(2.0)) / (fwhm * 3.141592653589793 / 180);
        beam[l] = exp(-l * (l + 1) / (2 * sigma * sigma));
    }
}

void orc_mask_latitude(int nside, double cut, double* mask)
{
    const long npix = 12L * nside * nside;
    const double halfpi = 3.141592653589793 / 2;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < npix; ++p) {
        double th, ph;
        orc_pix2ang_ring(nside, p, &th, &ph);
        if (th < halfpi + cut && th > halfpi - cut) mask[p] = 0;
    }
}

void orc_mask_regions(int nside, int nreg, const double* theta, const double* phi,
                      const double* angle, double* mask)
{
    const long npix = 12L * nside * nside;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < npix; ++p) {
        double th, ph;
        orc_pix2ang_ring(nside, p, &th, &ph);
        const double x1 = sin(th) * cos(ph), y1 = sin(th) * sin(ph), z1 = cos(th);
        for (int j = 0; j < nreg; ++j) {
            const double x = sin(theta[j]) * cos(phi[j]), y = sin(theta[j]) * sin(phi[j]),
                         z = cos(theta[j]);
            const double dot = x1 * x + y1 * y + z1 * z;
            if (dot > cos(angle[j])) mask[p] = 0;
        }
    }
}
