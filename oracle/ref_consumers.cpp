// Drop-in check: the reference's own optimisers and samplers, compiled from
// /root/reference/source (lbfgs.cpp, hmc.cpp, cosmo_mpi.cpp, macros.cpp) and
// the header-only LBFGS_General / HMCGeneral (include/lbfgs_general.hpp,
// include/hmc_general.hpp), driving the MI355X likelihood through
// include/cmbl_likelihood.hpp.  TEST INFRASTRUCTURE ONLY: built by
// oracle/Makefile (target `consumers`) into oracle/_ref/.
//
// Usage: ref_consumers <mode> <nside> <lmax> [iters] [fwhm_deg] [sigma] [vectors]
// (defaults 200, 5, 10, 1; vectors = 0 leaves the result vectors out of the
// output, for timing runs at large sizes).
//   lbfgs          Math::LBFGS on cmbl::MapLikelihood (host vectors)
//   lbfgs_device   Math::LBFGS_General on cmbl::DeviceLargeVector / DeviceMapFunction
//   hmc            Math::HMC on cmbl::MapLikelihood (LikelihoodWithDerivs)
//   hmc_device     Math::HMCGeneral on device vectors
//   cg             Math::ConjugateGradient on cmbl::GibbsSystem (the Gibbs
//                  constrained-realisation system), against the device solve
//   lbfgs_compare  Math::LBFGS_General on device vectors and the library's
//                  device-resident cmbl::DeviceLBFGS from the same start:
//                  per-iteration (f, gradient norm) and final x of both
//   hmc_compare    Math::HMCGeneral on device vectors and cmbl::DeviceHMC
//                  (same masses, tauMax, nMax, seed): the chain's -2lnL per
//                  step and final point of both
// Prints one JSON object on stdout.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include <function.hpp>
#include <likelihood_function.hpp>
#include <lbfgs.hpp>
#include <lbfgs_general.hpp>
#include <hmc.hpp>
#include <hmc_general.hpp>
#include <conjugate_gradient.hpp>

#include <cmbl_likelihood.hpp>

namespace {

struct Inputs {
    std::vector<double> cl, beam, d, w, x0;
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool g_vectors = true;

Inputs make_inputs(cmbl::MapContext& ctx, int nside, int lmax, double fwhm_deg, double sigma)
{
    Inputs in;
    const long npix = 12L * nside * nside;
    in.cl.assign(lmax + 1, 0.0);
    for (int l = 2; l <= lmax; ++l) in.cl[l] = 1e3 / (l * (l + 1.0));
    in.beam.resize(lmax + 1);
    cmbl::raise_on_error(cmbl_beam(lmax, fwhm_deg, in.beam.data()));
    std::vector<double> a((size_t)(lmax + 1) * (lmax + 1));
    cmbl::raise_on_error(cmbl_simulate_alm(in.cl.data(), lmax, 501, a.data()));
    ctx.setBeam(in.beam);
    ctx.setCl(in.cl);
    std::vector<double> s(npix), noise(npix), mask(npix, 1.0);
    cmbl::raise_on_error(cmbl_forward(ctx.get(), a.data(), s.data()));
    cmbl::raise_on_error(cmbl_simulate_white_noise(nside, sigma, 502, noise.data()));
    in.d.resize(npix);
    for (long i = 0; i < npix; ++i) in.d[i] = s[i] + noise[i];
    cmbl::raise_on_error(cmbl_mask_latitude(nside, 20.0 * M_PI / 180.0, mask.data()));
    in.w.resize(npix);
    cmbl::raise_on_error(cmbl_weights_from_mask(npix, mask.data(), sigma, in.w.data()));
    ctx.setData(in.d, in.w);
    in.x0.assign(a.size(), 0.0);
    return in;
}

void print_vec(const char* name, const std::vector<double>& v)
{
    if (!g_vectors) {
        std::printf("\"%s\": []", name);
        return;
    }
    std::printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? "," : "", v[i]);
    std::printf("]");
}

struct LbTrace {
    std::vector<double> f, gn;
    bool full = false;
    long n = 0;
    std::vector<double> xs, gs, zs;  // per iteration, when full
    void keep(std::vector<double>& dst, const double* d)
    {
        std::vector<double> h(n);
        cmbl::raise_on_error(cmbl_copy_to_host(h.data(), d, sizeof(double) * n));
        dst.insert(dst.end(), h.begin(), h.end());
    }
    void operator()(int, double fv, double g, const cmbl::DeviceLargeVector& x, const cmbl::DeviceLargeVector& gr,
                    const cmbl::DeviceLargeVector& z)
    {
        f.push_back(fv);
        gn.push_back(g);
        if (full) {
            keep(xs, x.data());
            keep(gs, gr.data());
            keep(zs, z.data());
        }
    }
    void operator()(int, double fv, double g, const double* x, const double* gr, const double* z)
    {
        f.push_back(fv);
        gn.push_back(g);
        if (full) {
            keep(xs, x);
            keep(gs, gr);
            keep(zs, z);
        }
    }
};

struct ChainTrace {
    long n = 0;
    std::vector<double> like, last;
    void operator()(const cmbl::DeviceLargeVector& x, double l)
    {
        like.push_back(l);
        x.download(&last);
    }
    void operator()(const double* x, double l)
    {
        like.push_back(l);
        last.resize(n);
        cmbl::raise_on_error(cmbl_copy_to_host(last.data(), x, sizeof(double) * n));
    }
};

std::vector<double> hmc_masses(const Inputs& in, int lmax)
{
    std::vector<double> m((size_t)(lmax + 1) * (lmax + 1));
    // mass ~ inverse prior variance keeps the leapfrog stable across l
    for (int l = 0; l <= lmax; ++l)
        for (int k = 0; k < 2 * l + 1; ++k) m[(size_t)l * l + k] = l >= 2 ? 2.0 / in.cl[l] : 1.0;
    return m;
}

struct ChainStats {
    long n = 0;
    double like_sum = 0;
    std::vector<double> mean;
    void operator()(const cmbl::DeviceLargeVector& x, double like)
    {
        std::vector<double> h;
        x.download(&h);
        if (mean.empty()) mean.assign(h.size(), 0.0);
        for (size_t i = 0; i < h.size(); ++i) mean[i] += h[i];
        like_sum += like;
        ++n;
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <lbfgs|lbfgs_device|hmc|hmc_device|cg> <nside> <lmax> [iters]\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const int nside = std::atoi(argv[2]);
    const int lmax = std::atoi(argv[3]);
    const int iters = argc > 4 ? std::atoi(argv[4]) : 200;
    const double fwhm = argc > 5 ? std::atof(argv[5]) : 5.0;
    const double sigma = argc > 6 ? std::atof(argv[6]) : 10.0;
    g_vectors = argc > 7 ? std::atoi(argv[7]) != 0 : true;
    try {
        cmbl::MapContext ctx(nside, lmax);
        Inputs in = make_inputs(ctx, nside, lmax, fwhm, sigma);
        const long n = ctx.nParams();
        if (mode == "lbfgs") {
            cmbl::MapLikelihood like(ctx);
            Math::LBFGS lbfgs((int)n, like, like, in.x0, 10);
            std::vector<double> res;
            const double f = lbfgs.minimize(&res, 1e-14, 1e-8, iters);
            std::vector<double> g;
            like.evaluate(res, &g);
            double gn = 0;
            for (double v : g) gn += v * v;
            std::printf("{\"mode\": \"lbfgs\", \"f\": %.17g, \"gnorm\": %.17g, \"evals\": %ld, ", f, std::sqrt(gn),
                        like.evaluations());
            print_vec("x", res);
            std::printf("}\n");
        } else if (mode == "lbfgs_device") {
            cmbl::DeviceLargeVectorFactory factory(n);
            cmbl::DeviceMapFunction func(ctx);
            cmbl::DeviceLargeVector start(n);
            start.upload(in.x0);
            Math::LBFGS_General<cmbl::DeviceLargeVector, cmbl::DeviceLargeVectorFactory, cmbl::DeviceMapFunction> lbfgs(
                &factory, &func, start, 10);
            cmbl::DeviceLargeVector res(n);
            const double t0 = now_s();
            const double f = lbfgs.minimize(&res, 1e-14, 1e-8, iters);
            const double t1 = now_s();
            cmbl::DeviceLargeVector g(n);
            lbfgs.getGradient(&g);
            std::vector<double> h;
            res.download(&h);
            std::printf("{\"mode\": \"lbfgs_device\", \"f\": %.17g, \"gnorm\": %.17g, \"evals\": %ld, "
                        "\"seconds\": %.6f, ", f, g.norm(), func.evaluations(), t1 - t0);
            print_vec("x", h);
            std::printf("}\n");
        } else if (mode == "hmc") {
            cmbl::MapLikelihood like(ctx);
            const std::string root = "hmc_chain_tmp";
            Math::HMC hmc((int)n, like, root, 0.05, 10, 777);
            for (long i = 0; i < n; ++i) hmc.setParam((int)i, "a" + std::to_string(i), 1.0, in.x0[i]);
            hmc.run(iters);
            std::ifstream chain(root + ".txt");
            long lines = 0;
            std::string line;
            while (std::getline(chain, line)) ++lines;
            std::remove((root + ".txt").c_str());
            std::printf("{\"mode\": \"hmc\", \"chain_lines\": %ld, \"evals\": %ld}\n", lines, like.evaluations());
        } else if (mode == "hmc_device") {
            cmbl::DeviceLargeVectorFactory factory(n);
            cmbl::DeviceMapFunction func(ctx);
            cmbl::DeviceLargeVector start(n), masses(n);
            start.upload(in.x0);
            std::vector<double> m(n);
            // mass ~ inverse prior variance keeps the leapfrog stable across l
            for (int l = 0; l <= lmax; ++l)
                for (int k = 0; k < 2 * l + 1; ++k) m[(size_t)l * l + k] = l >= 2 ? 2.0 / in.cl[l] : 1.0;
            masses.upload(m);
            Math::HMCGeneral<cmbl::DeviceLargeVector, cmbl::DeviceLargeVectorFactory, cmbl::DeviceMapFunction> hmc(
                &factory, &func, start, masses, 0.5, 10, 778);
            ChainStats cb;
            const double t0 = now_s();
            hmc.run(iters, &cb);
            const double t1 = now_s();
            for (double& v : cb.mean) v /= (double)cb.n;
            std::printf("{\"mode\": \"hmc_device\", \"samples\": %ld, \"evals\": %ld, \"mean_like\": %.17g, "
                        "\"seconds\": %.6f, ", cb.n, func.evaluations(), cb.like_sum / cb.n, t1 - t0);
            print_vec("mean", cb.mean);
            std::printf("}\n");
        } else if (mode == "lbfgs_compare" || mode == "lbfgs_compare_full") {
            // the reference's LBFGS_General on device vectors ...
            cmbl::DeviceLargeVectorFactory factory(n);
            cmbl::DeviceMapFunction func(ctx);
            cmbl::DeviceLargeVector start(n);
            start.upload(in.x0);
            Math::LBFGS_General<cmbl::DeviceLargeVector, cmbl::DeviceLargeVectorFactory, cmbl::DeviceMapFunction> ref(
                &factory, &func, start, 10);
            cmbl::DeviceLargeVector res(n), res2(n);
            LbTrace tr, tr2;
            tr.full = tr2.full = mode == "lbfgs_compare_full";
            tr.n = tr2.n = n;
            const double t0 = now_s();
            const double f = ref.minimize(&res, 1e-14, 1e-8, iters, &tr);
            const double t1 = now_s();
            // ... and the library's device-resident loop from the same start
            cmbl::DeviceLBFGS dev(ctx, start, 10);
            const double f2 = dev.minimize(&res2, 1e-14, 1e-8, iters, &tr2);
            const double t2 = now_s();
            std::vector<double> h, h2;
            res.download(&h);
            res2.download(&h2);
            std::printf("{\"mode\": \"lbfgs_compare\", \"f_ref\": %.17g, \"f_dev\": %.17g, \"evals_ref\": %ld, "
                        "\"evals_dev\": %d, \"seconds_ref\": %.6f, \"seconds_dev\": %.6f, ",
                        f, f2, func.evaluations(), dev.evaluations(), t1 - t0, t2 - t1);
            print_vec("trace_f_ref", tr.f);
            std::printf(", ");
            print_vec("trace_f_dev", tr2.f);
            std::printf(", ");
            print_vec("trace_gn_ref", tr.gn);
            std::printf(", ");
            print_vec("trace_gn_dev", tr2.gn);
            std::printf(", ");
            print_vec("x_ref", h);
            std::printf(", ");
            print_vec("x_dev", h2);
            if (tr.full) {
                const char* names[6] = {"xs_ref", "xs_dev", "gs_ref", "gs_dev", "zs_ref", "zs_dev"};
                const std::vector<double>* v[6] = {&tr.xs, &tr2.xs, &tr.gs, &tr2.gs, &tr.zs, &tr2.zs};
                for (int k = 0; k < 6; ++k) {
                    std::printf(", ");
                    print_vec(names[k], *v[k]);
                }
            }
            std::printf("}\n");
        } else if (mode == "hmc_compare") {
            cmbl::DeviceLargeVectorFactory factory(n);
            cmbl::DeviceMapFunction func(ctx);
            cmbl::DeviceLargeVector start(n), masses(n);
            start.upload(in.x0);
            masses.upload(hmc_masses(in, lmax));
            ChainTrace tr, tr2;
            tr.n = tr2.n = n;
            double t0, t1, t2;
            std::vector<double> h2;
            {
                Math::HMCGeneral<cmbl::DeviceLargeVector, cmbl::DeviceLargeVectorFactory, cmbl::DeviceMapFunction>
                    ref(&factory, &func, start, masses, 0.5, 10, 778);
                t0 = now_s();
                ref.run(iters, &tr);
                t1 = now_s();
            }
            cmbl::DeviceHMC dev(ctx, start, masses, 0.5, 10, 778);
            dev.run(iters, &tr2);
            t2 = now_s();
            cmbl::DeviceLargeVector xd(n);
            dev.position(&xd);
            xd.download(&h2);
            std::printf("{\"mode\": \"hmc_compare\", \"accepted_dev\": %ld, \"seconds_ref\": %.6f, "
                        "\"seconds_dev\": %.6f, ", dev.accepted(), t1 - t0, t2 - t1);
            print_vec("like_ref", tr.like);
            std::printf(", ");
            print_vec("like_dev", tr2.like);
            std::printf(", ");
            print_vec("x_ref", tr.last);
            std::printf(", ");
            print_vec("x_dev", h2);
            std::printf("}\n");
        } else if (mode == "cg") {
            // Gibbs system on the same data: mask 20 deg cut, sigma 10, white
            // noise maps omega0, omega1 of unit variance (seeds 601, 602)
            const long npix = 12L * nside * nside;
            std::vector<double> mask(npix, 1.0), w0(npix), w1(npix), b, xcg, xdev;
            cmbl::raise_on_error(cmbl_mask_latitude(nside, 20.0 * M_PI / 180.0, mask.data()));
            cmbl::raise_on_error(cmbl_simulate_white_noise(nside, 1.0, 601, w0.data()));
            cmbl::raise_on_error(cmbl_simulate_white_noise(nside, 1.0, 602, w1.data()));
            cmbl::GibbsSystem sys(nside, lmax, in.cl, in.beam, mask, sigma);
            sys.rhs(in.d, w0, w1, b);
            double bn = 0;
            for (double v : b) bn += v * v;
            // absolute precision of the reference's solve(); 1e-10 |b| relative
            const double precision = 1e-10 * std::sqrt(bn);
            Math::ConjugateGradient<cmbl::GibbsSystem> cg((int)b.size(), &sys, b);
            int it_cg = 0;
            const double t0 = now_s();
            xcg = cg.solve(precision, &it_cg);
            const double t1 = now_s();
            const int it_dev = sys.solve(b, xdev, precision, 100000);
            const double t2 = now_s();
            std::printf("{\"mode\": \"cg\", \"iter_cg\": %d, \"iter_dev\": %d, \"precision\": %.17g, "
                        "\"seconds_cg\": %.6f, \"seconds_dev\": %.6f, ", it_cg, it_dev, precision, t1 - t0, t2 - t1);
            print_vec("b", b);
            std::printf(", ");
            print_vec("x_cg", xcg);
            std::printf(", ");
            print_vec("x_dev", xdev);
            std::printf("}\n");
        } else {
            std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
            return 2;
        }
    } catch (std::exception& e) {
        std::fprintf(stderr, "EXCEPTION: %s\n", e.what());
        return 1;
    }
    return 0;
}
