// Thin C entry points over the reference's own header-only numerics, so the
// tests can pin the oracle against the reference code itself.
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile (target `ref`) from the
// headers where they lie under /root/reference/include, into oracle/_ref/.
// Nothing here is part of the product.
//
// Reference code exercised:
//   Math::SphericalHarmonics::calculate   include/spherical_harmonics.hpp:29-37
//   Math::Legendre / AssociatedLegendre   include/legendre.hpp:26-103
//   Math::GaussianGenerator               include/random.hpp:32-52
//   Math::UniformRealGenerator            include/random.hpp:10-30
//   Math::moreThuenteSearch               include/line_search.hpp:12-272
#include <cmath>
#include <vector>

#include <spherical_harmonics.hpp>
#include <legendre.hpp>
#include <random.hpp>
#include <wigner_3j.hpp>
#include <line_search.hpp>

extern "C" {

void ref_sh(int l, int m, double theta, double phi, double* re, double* im)
{
    Math::SphericalHarmonics sh;
    const ComplexDouble v = sh.calculate(l, m, theta, phi);
    *re = v.real();
    *im = v.imag();
}

double ref_legendre(int l, double x)
{
    Math::Legendre p;
    return p.calculate(l, x);
}

double ref_assoc_legendre(int l, int m, double x)
{
    Math::AssociatedLegendre p;
    return p.calculate(l, m, x);
}

void ref_gauss_draws(int seed, double mean, double sigma, long n, double* out)
{
    Math::GaussianGenerator g(seed, mean, sigma);
    for (long i = 0; i < n; ++i) out[i] = g.generate();
}

void ref_uniform_draws(int seed, double lo, double hi, long n, double* out)
{
    Math::UniformRealGenerator g(seed, lo, hi);
    for (long i = 0; i < n; ++i) out[i] = g.generate();
}

// lambda_lm(theta) = Re Y_lm(theta, 0) for 0 <= m <= l <= L, index l(l+1)/2 + m
void ref_lambda_table(int L, double theta, double* out)
{
    Math::SphericalHarmonics sh;
    for (int l = 0; l <= L; ++l)
        for (int m = 0; m <= l; ++m)
            out[(long)l * (l + 1) / 2 + m] = sh.calculate(l, m, theta, 0.0).real();
}

// Brute-force real-map synthesis at arbitrary points:
//   s = sum_l [ a_l0 Y_l0 + 2 sum_{m>=1} Re(a_lm Y_lm) ],  a from the packed
// almToVector layout, each Y_lm from the reference SphericalHarmonics.
// Points must be grouped by theta (consecutive equal theta reuse the table).
void ref_brute_synthesis(int L, const double* x_packed, const double* beam, long npts,
                         const double* theta, const double* phi, double* out)
{
    std::vector<double> lam((size_t)(L + 1) * (L + 2) / 2);
    double last = -1.0;
    for (long p = 0; p < npts; ++p) {
        if (theta[p] != last) {
            ref_lambda_table(L, theta[p], lam.data());
            last = theta[p];
        }
        long double s = 0;
        for (int m = 0; m <= L; ++m) {
            const double c = std::cos(m * phi[p]), sn = std::sin(m * phi[p]);
            for (int l = m; l <= L; ++l) {
                const double y = lam[(size_t)l * (l + 1) / 2 + m];
                const double b = beam ? beam[l] : 1.0;
                if (m == 0) {
                    s += (long double)b * x_packed[(long)l * l] * y;
                } else {
                    const double re = x_packed[(long)l * l + 2 * m - 1];
                    const double im = x_packed[(long)l * l + 2 * m];
                    s += 2.0L * b * y * (re * c - im * sn);
                }
            }
        }
        out[p] = (double)s;
    }
}

// (l1 l2 l3; 0 0 0)^2 from the reference's Math::Wigner3JZeroM for every
// l1, l2, l3 <= L, into out[(l1 (L+1) + l2) (L+1) + l3] (zero where the
// symbol vanishes).
struct W3jCollect {
    int L;
    double* out;
    void process(int l1, int l2, int l3, double val)
    {
        if (l1 > L || l2 > L || l3 > L) return;
        out[((long)l1 * (L + 1) + l2) * (L + 1) + l3] = val * val;
    }
};

void ref_w3j2_table(int L, double* out)
{
    W3jCollect c{L, out};
    Math::Wigner3JZeroM<W3jCollect> w(c);
    w.calculate(L * 3 / 2 + 1);
}

// One-dimensional test functions of More and Thuente (ACM TOMS 20, 1994,
// section 5) and their derivatives, for the line-search comparison.
static void phi(int fn, double a, double* f, double* g)
{
    switch (fn) {
    case 1: {
        const double b = 2.0;
        *f = -a / (a * a + b);
        *g = (a * a - b) / ((a * a + b) * (a * a + b));
        break;
    }
    case 2: {
        const double b = 0.004, t = a + b;
        *f = std::pow(t, 5) - 2.0 * std::pow(t, 4);
        *g = 5.0 * std::pow(t, 4) - 8.0 * std::pow(t, 3);
        break;
    }
    case 3: {
        const double b = 0.01, l = 39.0, pi = 3.141592653589793;
        double f0, g0;
        if (a <= 1 - b) {
            f0 = 1 - a;
            g0 = -1;
        } else if (a >= 1 + b) {
            f0 = a - 1;
            g0 = 1;
        } else {
            f0 = (a - 1) * (a - 1) / (2 * b) + b / 2;
            g0 = (a - 1) / b;
        }
        *f = f0 + 2 * (1 - b) / (l * pi) * std::sin(l * pi * a / 2);
        *g = g0 + (1 - b) * std::cos(l * pi * a / 2);
        break;
    }
    default: {  // 4, 5, 6
        const double b1 = fn == 5 ? 0.01 : 0.001, b2 = fn == 6 ? 0.01 : 0.001;
        const double g1 = std::sqrt(1 + b1 * b1) - b1, g2 = std::sqrt(1 + b2 * b2) - b2;
        const double r1 = std::sqrt((1 - a) * (1 - a) + b2 * b2), r2 = std::sqrt(a * a + b1 * b1);
        *f = g1 * r1 + g2 * r2;
        *g = g1 * (a - 1) / r1 + g2 * a / r2;
        break;
    }
    }
}

void ref_phi(int fn, double a, double* f, double* g) { phi(fn, a, f, g); }

struct V1 {
    double v = 0;
    void copy(const V1& o, double c = 1.) { v = c * o.v; }
    double dotProduct(const V1& o) const { return v * o.v; }
    void add(const V1& o, double c = 1.) { v += c * o.v; }
};

struct Phi1 {
    int fn;
    double x;
    double* trials;
    int n, cap;
    Phi1(int f, double* t, int c) : fn(f), x(0.0), trials(t), n(0), cap(c) {}
    void set(const V1& p)
    {
        x = p.v;
        if (n < cap) trials[n] = x;
        ++n;
    }
    double value()
    {
        double f, g;
        phi(fn, x, &f, &g);
        return f;
    }
    void derivative(V1* res)
    {
        double f, g;
        phi(fn, x, &f, &g);
        res->v = g;
    }
};

// The reference's moreThuenteSearch from a = 0 along s = 1: trial steps into
// trials (at most cap), their count into *ntrials, the final step into *stp;
// returns info.
int ref_mt_trace(int fn, double stp0, double ftol, double gtol, double xtol, double stpmin, double stpmax, int maxfev,
                 double* trials, int cap, int* ntrials, double* stp)
{
    Phi1 func(fn, trials, cap);
    V1 x0, g0, s, x, g;
    s.v = 1.0;
    double f0, d0;
    phi(fn, 0.0, &f0, &d0);
    g0.v = d0;
    double f = f0, st = stp0;
    int nfev = 0;
    const int info = Math::moreThuenteSearch(&func, x0, f, g0, s, st, ftol, gtol, xtol, stpmin, stpmax, maxfev, &x, &g,
                                             nfev);
    *ntrials = func.n;
    *stp = st;
    return info;
}

}  // extern "C"
