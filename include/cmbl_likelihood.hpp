// cmbl_likelihood.hpp - C++ host classes over the C ABI (cmbl.h) that plug the
// MI355X likelihood into the Cosmo++ optimisers and samplers.
//
// Include the Cosmo++ headers (function.hpp, likelihood_function.hpp) or
// cmbl_function.hpp before this header; both provide the same interfaces.
//
//   cmbl::MapLikelihood      Math::RealFunctionMultiDim + RealFunctionMultiToMulti
//                            (include/function.hpp:138,168) and
//                            Math::LikelihoodWithDerivs (include/likelihood_function.hpp:39,46):
//                            usable by Math::LBFGS (include/lbfgs.hpp:126-172),
//                            Math::CG, Math::HMC (source/hmc.cpp:31-163).
//   cmbl::DeviceLargeVector  the LargeVector concept of LBFGS_General / HMCGeneral
//                            (include/lbfgs_general.hpp:14-31, include/hmc_general.hpp:14-37)
//                            on device memory.
//   cmbl::DeviceLargeVectorFactory, cmbl::DeviceMapFunction
//                            the Factory and Function concepts
//                            (include/lbfgs_general.hpp:33-53, include/hmc_general.hpp:39-61).
//   cmbl::DeviceLBFGS, cmbl::DeviceHMC
//                            LBFGS_General (include/lbfgs_general.hpp:58-405) and
//                            HMCGeneral (include/hmc_general.hpp:64-246) with the
//                            whole loop on the GPU (cmbl_lbfgs / cmbl_hmc).
//
// Values are -2lnL and derivatives of -2lnL, as in the reference.  Failures
// throw cmbl::Error (a std::exception, like StandardException).
#ifndef CMBL_LIKELIHOOD_HPP
#define CMBL_LIKELIHOOD_HPP

#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <utility>
#include <vector>

#include "cmbl.h"

namespace cmbl
{

class Error : public std::exception
{
public:
    Error(int code, const std::string& msg) : code_(code), msg_("cmbl error " + std::to_string(code) + ": " + msg) {}
    const char* what() const noexcept override { return msg_.c_str(); }
    int code() const { return code_; }

private:
    int code_;
    std::string msg_;
};

inline void raise_on_error(int rc)
{
    if (rc != CMBL_OK) throw Error(rc, cmbl_last_error());
}

/// One HEALPix map with its likelihood tables on the current GPU.
class MapContext
{
public:
    MapContext(int nside, int lmax) { raise_on_error(cmbl_map_create(&m_, nside, lmax)); }
    ~MapContext() { cmbl_map_destroy(m_); }
    MapContext(const MapContext&) = delete;
    MapContext& operator=(const MapContext&) = delete;

    void setBeam(const std::vector<double>& b) { need(b.size(), lmax() + 1); raise_on_error(cmbl_map_set_beam(m_, b.data())); }
    void setCl(const std::vector<double>& c) { need(c.size(), lmax() + 1); raise_on_error(cmbl_map_set_cl(m_, c.data())); }
    void setData(const std::vector<double>& d, const std::vector<double>& w)
    {
        need(d.size(), nPix());
        need(w.size(), nPix());
        raise_on_error(cmbl_map_set_data(m_, d.data(), w.data()));
    }
    int nside() const { return cmbl_map_nside(m_); }
    int lmax() const { return cmbl_map_lmax(m_); }
    long nPix() const { return cmbl_map_npix(m_); }
    long nParams() const { return cmbl_map_nparams(m_); }
    cmbl_map* get() { return m_; }

private:
    static void need(size_t have, long want)
    {
        if ((long)have != want) throw Error(CMBL_ERR_ARG, "size " + std::to_string(have) + " != " + std::to_string(want));
    }
    cmbl_map* m_ = nullptr;
};

/// -2lnL(a_lm) and its gradient behind the reference's function interfaces.
/// The last point, value and gradient are cached by contents and by the
/// map's generation (any setBeam / setCl / setData invalidates them), so that
/// value + derivative at one point (LBFGS, line searches) and the n
/// per-component calls of HMC::calculateDerivs cost one device evaluation.
class MapLikelihood : public Math::RealFunctionMultiDim,
                      public Math::RealFunctionMultiToMulti,
                      public Math::LikelihoodWithDerivs
{
public:
    explicit MapLikelihood(MapContext& ctx) : ctx_(ctx), n_(ctx.nParams()) {}

    /// RealFunctionMultiDim::evaluate
    double evaluate(const std::vector<double>& x) const override
    {
        update(x.data(), (long)x.size());
        return f_;
    }
    /// RealFunctionMultiToMulti::evaluate (res is resized)
    void evaluate(const std::vector<double>& x, std::vector<double>* res) const override
    {
        update(x.data(), (long)x.size());
        res->resize(n_);
        std::memcpy(res->data(), g_.data(), sizeof(double) * n_);
    }
    /// LikelihoodWithDerivs::calculate
    double calculate(double* params, int nParams) override
    {
        update(params, nParams);
        return f_;
    }
    /// LikelihoodWithDerivs::calculateDeriv
    double calculateDeriv(double* params, int nParams, int i) override
    {
        if (i < 0 || i >= nParams) throw Error(CMBL_ERR_ARG, "invalid index " + std::to_string(i));
        update(params, nParams);
        return g_[i];
    }
    long evaluations() const { return evals_; }

private:
    void update(const double* x, long n) const
    {
        if (n != n_) throw Error(CMBL_ERR_ARG, "expected " + std::to_string(n_) + " parameters");
        const long gen = cmbl_map_generation(ctx_.get());
        if (valid_ && gen == gen_ && std::memcmp(x, x_.data(), sizeof(double) * n) == 0) return;
        x_.assign(x, x + n);
        g_.resize(n);
        valid_ = false;
        raise_on_error(cmbl_value_and_grad(ctx_.get(), x_.data(), &f_, g_.data()));
        valid_ = true;
        gen_ = gen;
        ++evals_;
    }
    MapContext& ctx_;
    long n_;
    mutable std::vector<double> x_, g_;
    mutable double f_ = 0;
    mutable bool valid_ = false;
    mutable long gen_ = -1;
    mutable long evals_ = 0;
};

/// LargeVector on device memory: BLAS-1 on the GPU, scalars returned to the host.
class DeviceLargeVector
{
public:
    explicit DeviceLargeVector(long n, void* stream = nullptr) : n_(n), stream_(stream)
    {
        raise_on_error(cmbl_device_alloc(reinterpret_cast<void**>(&d_), sizeof(double) * (n > 0 ? n : 1)));
        setToZero();
    }
    ~DeviceLargeVector() { cmbl_device_free(d_); }
    DeviceLargeVector(const DeviceLargeVector&) = delete;
    DeviceLargeVector& operator=(const DeviceLargeVector&) = delete;

    void copy(const DeviceLargeVector& o, double c = 1.) { same(o); raise_on_error(cmbl_vec_copy(d_, o.d_, c, n_, stream_)); }
    void setToZero() { raise_on_error(cmbl_vec_set(d_, 0.0, n_, stream_)); }
    double norm() const { return std::sqrt(dotProduct(*this)); }
    double dotProduct(const DeviceLargeVector& o) const
    {
        same(o);
        double r = 0;
        raise_on_error(cmbl_vec_dot(d_, o.d_, n_, &r, stream_));
        return r;
    }
    void add(const DeviceLargeVector& o, double c = 1.) { same(o); raise_on_error(cmbl_vec_axpy(d_, o.d_, c, n_, stream_)); }
    void multiply(const DeviceLargeVector& o) { same(o); raise_on_error(cmbl_vec_mul(d_, o.d_, n_, stream_)); }
    void divide(const DeviceLargeVector& o) { same(o); raise_on_error(cmbl_vec_div(d_, o.d_, n_, stream_)); }
    void pow(double p) { raise_on_error(cmbl_vec_pow(d_, p, n_, stream_)); }
    void swap(DeviceLargeVector& o)
    {
        same(o);
        std::swap(d_, o.d_);
    }

    long size() const { return n_; }
    double* data() { return d_; }
    const double* data() const { return d_; }
    void* stream() const { return stream_; }
    void upload(const std::vector<double>& h)
    {
        if ((long)h.size() != n_) throw Error(CMBL_ERR_ARG, "size mismatch");
        raise_on_error(cmbl_stream_synchronize(stream_));
        raise_on_error(cmbl_copy_to_device(d_, h.data(), sizeof(double) * n_));
    }
    void download(std::vector<double>* h) const
    {
        h->resize(n_);
        raise_on_error(cmbl_stream_synchronize(stream_));
        raise_on_error(cmbl_copy_to_host(h->data(), d_, sizeof(double) * n_));
    }

private:
    void same(const DeviceLargeVector& o) const
    {
        if (o.n_ != n_) throw Error(CMBL_ERR_ARG, "vector sizes differ");
    }
    long n_;
    void* stream_;
    double* d_ = nullptr;
};

class DeviceLargeVectorFactory
{
public:
    explicit DeviceLargeVectorFactory(long n, void* stream = nullptr) : n_(n), stream_(stream) {}
    DeviceLargeVector* giveMeOne() { return new DeviceLargeVector(n_, stream_); }

private:
    long n_;
    void* stream_;
};

/// The Function concept (set / value / derivative / whitenoise) on device vectors.
class DeviceMapFunction
{
public:
    explicit DeviceMapFunction(MapContext& ctx, void* stream = nullptr)
        : ctx_(ctx), stream_(stream), x_(ctx.nParams(), stream), g_(ctx.nParams(), stream)
    {
        raise_on_error(cmbl_device_alloc(reinterpret_cast<void**>(&d_f_), sizeof(double)));
    }
    ~DeviceMapFunction() { cmbl_device_free(d_f_); }

    void set(const DeviceLargeVector& x)
    {
        x_.copy(x);
        fresh_ = false;
    }
    double value()
    {
        eval();
        return f_;
    }
    void derivative(DeviceLargeVector* res)
    {
        eval();
        res->copy(g_);
    }
    /// BasicLBFGSFunc::whitenoise (source/lbfgs.cpp:401-416), one process.
    void whitenoise(int seed, DeviceLargeVector* x, double amplitude)
    {
        std::vector<double> h(x->size());
        raise_on_error(cmbl_whitenoise(seed, amplitude, x->size(), h.data()));
        x->upload(h);
    }
    long evaluations() const { return evals_; }

private:
    void eval()
    {
        const long gen = cmbl_map_generation(ctx_.get());
        if (fresh_ && gen == gen_) return;
        raise_on_error(cmbl_value_and_grad_device(ctx_.get(), x_.data(), d_f_, g_.data(), stream_));
        raise_on_error(cmbl_stream_synchronize(stream_));
        raise_on_error(cmbl_copy_to_host(&f_, d_f_, sizeof(double)));
        if (!std::isfinite(f_)) {
            (void)cmbl_map_check_finite(ctx_.get(), stream_);  // clears the map's flag
            throw Error(CMBL_ERR_NONFINITE, "-2lnL is not finite");
        }
        fresh_ = true;
        gen_ = gen;
        ++evals_;
    }
    MapContext& ctx_;
    void* stream_;
    DeviceLargeVector x_, g_;
    double* d_f_ = nullptr;
    double f_ = 0;
    bool fresh_ = false;
    long gen_ = -1;
    long evals_ = 0;
};

/// LBFGS_General with the More-Thuente line search, run on the GPU: the
/// vectors never leave device memory, the vector work of each phase is
/// captured in HIP graphs, and the iterates equal those of the reference's
/// LBFGS_General on DeviceLargeVector / DeviceMapFunction bit for bit.
class DeviceLBFGS
{
public:
    DeviceLBFGS(MapContext& ctx, const DeviceLargeVector& starting, int m = 10, bool graphs = true)
    {
        raise_on_error(cmbl_stream_synchronize(starting.stream()));
        raise_on_error(cmbl_lbfgs_create(&lb_, ctx.get(), starting.data(), m, graphs ? 0 : CMBL_LOOP_EAGER));
    }
    ~DeviceLBFGS() { cmbl_lbfgs_destroy(lb_); }
    DeviceLBFGS(const DeviceLBFGS&) = delete;
    DeviceLBFGS& operator=(const DeviceLBFGS&) = delete;

    void setStarting(const DeviceLargeVector& starting)
    {
        raise_on_error(cmbl_stream_synchronize(starting.stream()));
        raise_on_error(cmbl_lbfgs_set_starting(lb_, starting.data()));
    }
    double minimize(DeviceLargeVector* res, double epsilon = 1e-3, double gNormTol = 1e-5, int maxIter = 1000000)
    {
        double f = 0;
        raise_on_error(cmbl_lbfgs_minimize(lb_, epsilon, gNormTol, maxIter, res ? res->data() : nullptr, &f,
                                           &iterations_, &evaluations_, nullptr, nullptr));
        return f;
    }
    /// callback(iter, f, gradNorm, d_x, d_g, d_z) after every iteration.
    template <class CallBack>
    double minimize(DeviceLargeVector* res, double epsilon, double gNormTol, int maxIter, CallBack* callback)
    {
        double f = 0;
        raise_on_error(cmbl_lbfgs_minimize(lb_, epsilon, gNormTol, maxIter, res ? res->data() : nullptr, &f,
                                           &iterations_, &evaluations_, callback ? &call<CallBack> : nullptr,
                                           callback));
        return f;
    }
    void getGradient(DeviceLargeVector* g) { raise_on_error(cmbl_lbfgs_gradient(lb_, g->data())); }
    int iterations() const { return iterations_; }
    int evaluations() const { return evaluations_; }

private:
    template <class CallBack>
    static void call(void* user, int iter, double f, double gn, const double* x, const double* g, const double* z)
    {
        (*static_cast<CallBack*>(user))(iter, f, gn, x, g, z);
    }
    cmbl_lbfgs* lb_ = nullptr;
    int iterations_ = 0, evaluations_ = 0;
};

/// HMCGeneral run on the GPU: one HIP graph per trajectory, the reference's
/// momentum draws made ahead on host threads.  With fixedSteps every
/// trajectory takes step tauMax and nMax steps.
class DeviceHMC
{
public:
    DeviceHMC(MapContext& ctx, const DeviceLargeVector& starting, const DeviceLargeVector& masses, double tauMax,
              int nMax, int seed, bool fixedSteps = false, bool graphs = true)
    {
        raise_on_error(cmbl_stream_synchronize(starting.stream()));
        raise_on_error(cmbl_stream_synchronize(masses.stream()));
        raise_on_error(cmbl_hmc_create(&hm_, ctx.get(), starting.data(), masses.data(), tauMax, nMax, seed,
                                       (fixedSteps ? CMBL_HMC_FIXED : 0) | (graphs ? 0 : CMBL_LOOP_EAGER)));
    }
    ~DeviceHMC() { cmbl_hmc_destroy(hm_); }
    DeviceHMC(const DeviceHMC&) = delete;
    DeviceHMC& operator=(const DeviceHMC&) = delete;

    /// callback(d_x, like) at the start and after every trajectory.
    template <class CallBack>
    void run(int maxIters, CallBack* callback)
    {
        raise_on_error(cmbl_hmc_run(hm_, maxIters, callback ? &call<CallBack> : nullptr, callback));
    }
    void run(int maxIters) { raise_on_error(cmbl_hmc_run(hm_, maxIters, nullptr, nullptr)); }
    long accepted() const
    {
        long a = 0;
        raise_on_error(cmbl_hmc_state(hm_, nullptr, nullptr, &a, nullptr));
        return a;
    }
    long evaluations() const
    {
        long e = 0;
        raise_on_error(cmbl_hmc_state(hm_, nullptr, nullptr, nullptr, &e));
        return e;
    }
    void position(DeviceLargeVector* x) { raise_on_error(cmbl_hmc_position(hm_, x->data())); }

private:
    template <class CallBack>
    static void call(void* user, const double* x, double like)
    {
        (*static_cast<CallBack*>(user))(x, like);
    }
    cmbl_hmc* hm_ = nullptr;
};

/// The Gibbs constrained-realisation system of CMBGibbsSampler::generateSignal
/// (source/cmb_gibbs.cpp:437-530) on the GPU.  multiplyByMatrix and
/// preconditioner make it a CGTreats for the reference's
/// Math::ConjugateGradient (include/conjugate_gradient.hpp:13-19), the way
/// CmbGibbsCGTreats is (source/cmb_gibbs.cpp:303-434); solve() runs the same
/// CG on the device without moving vectors to the host.
class GibbsSystem
{
public:
    GibbsSystem(int nside, int lmax, const std::vector<double>& cl, const std::vector<double>& beam,
                const std::vector<double>& mask, double pixelNoise)
        : n_((long)(lmax + 1) * (lmax + 1)), npix_(12L * nside * nside)
    {
        if ((long)cl.size() != lmax + 1 || (long)beam.size() != lmax + 1 || (long)mask.size() != npix_)
            throw Error(CMBL_ERR_ARG, "GibbsSystem: cl, beam or mask has the wrong size");
        raise_on_error(cmbl_gibbs_create(&g_, nside, lmax, cl.data(), beam.data(), mask.data(), pixelNoise));
    }
    ~GibbsSystem() { cmbl_gibbs_destroy(g_); }
    GibbsSystem(const GibbsSystem&) = delete;
    GibbsSystem& operator=(const GibbsSystem&) = delete;

    long size() const { return n_; }
    cmbl_gibbs* get() { return g_; }

    void multiplyByMatrix(const std::vector<double>& original, std::vector<double>& result) const
    {
        need(original.size());
        result.resize(n_);
        raise_on_error(cmbl_gibbs_apply(g_, original.data(), result.data()));
    }
    void preconditioner(const std::vector<double>& original, std::vector<double>& result) const
    {
        need(original.size());
        result.resize(n_);
        raise_on_error(cmbl_gibbs_precondition(g_, original.data(), result.data()));
    }
    /// b of generateSignal from the data map (monopole and dipole removed)
    /// and the two white-noise maps.
    void rhs(const std::vector<double>& d, const std::vector<double>& omega0, const std::vector<double>& omega1,
             std::vector<double>& b) const
    {
        if ((long)d.size() != npix_ || (long)omega0.size() != npix_ || (long)omega1.size() != npix_)
            throw Error(CMBL_ERR_ARG, "GibbsSystem::rhs: maps must have npix values");
        b.resize(n_);
        raise_on_error(cmbl_gibbs_rhs(g_, d.data(), omega0.data(), omega1.data(), b.data()));
    }
    /// Device CG from x = 0; returns the iteration count.
    int solve(const std::vector<double>& b, std::vector<double>& x, double precision, int maxIter) const
    {
        need(b.size());
        x.resize(n_);
        int it = 0;
        raise_on_error(cmbl_gibbs_solve(g_, b.data(), x.data(), precision, maxIter, &it));
        return it;
    }

private:
    void need(size_t have) const
    {
        if ((long)have != n_) throw Error(CMBL_ERR_ARG, "GibbsSystem: vector size " + std::to_string(have));
    }
    long n_, npix_;
    cmbl_gibbs* g_ = nullptr;
};

}  // namespace cmbl

#endif
