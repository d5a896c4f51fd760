// The Gibbs sampler step of CMBGibbsSampler (source/cmb_gibbs.cpp:63-549)
// with its maps and a_lm vectors on the GPU:
//   construction  map_ = map + noise, s = map2alm(map, weight 1) / B_l,
//                 sigma_l of s (:63-162);
//   generateW     synthesis of s B, the weighted residual with four noise
//                 draws per pixel projected on the monopole / dipole maps,
//                 and the 4x4 conjugate-gradient solve for their amplitudes
//                 (:240-301);
//   generateSignal residual without monopole / dipole, the right-hand side and
//                 the constrained-realisation CG of cmbl_gibbs, then
//                 s = x sqrt(C_l) and sigma_l (:437-549);
//   generateCl    C_l = sigma_l / rho with rho a sum of 2l - 1 squared draws
//                 (:178-193).
// The draws come from one host generator (mt19937 + normal distribution, as
// GaussianGenerator, include/random.hpp:32-52) in the reference's order:
// 4 per pixel in generateW, 2 per pixel in generateSignal, then 2l - 1 per l.
// The beam is the caller's B_l or else the Gaussian of the FWHM; the
// reference multiplies the Gaussian by the HEALPix pixel window that
// Utils::readPixelWindowFunction reads from HEALPix's data files
// (source/utils.cpp:67-130), which a caller passes in through beam.
#include <cmath>
#include <cstring>
#include <ctime>
#include <random>
#include <string>
#include <vector>

#include "cmbl_device.hpp"
#include "cmbl_internal.hpp"

namespace cmbl {

constexpr int kSamplerBlocks = 512;

// generateW's pixel pass (:259-287): v = (map_ - s) / sigma^2 + sum_k omega_k
// / sigma on pixels with mask > 0.5 (0 elsewhere), and per block the four
// partial sums of Y_k v.
__global__ __launch_bounds__(256) void k_gibbs_w(long npix, const double* __restrict__ map,
                                                 const double* __restrict__ s, const double* __restrict__ mask,
                                                 const double* __restrict__ omega, const double* __restrict__ Y,
                                                 double sigma, double* __restrict__ part)
{
    __shared__ double sh[4];
    double b[4] = {0.0, 0.0, 0.0, 0.0};
    for (long i = blockIdx.x * 256L + threadIdx.x; i < npix; i += (long)gridDim.x * 256) {
        double v = map[i] - s[i];
        double o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = omega[4 * i + k];
        if (mask[i] > 0.5) {
            v /= (sigma * sigma);
#pragma unroll
            for (int k = 0; k < 4; ++k) o[k] /= sigma;
        } else {
            v = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) o[k] = 0.0;
        }
        v += (o[0] + o[1] + o[2] + o[3]);
#pragma unroll
        for (int k = 0; k < 4; ++k) b[k] = fma(Y[k * npix + i], v, b[k]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double t = block_sum<256>(b[k], sh);
        if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = t;
    }
}

// Four sums of nb block partials each, in block order.
__global__ __launch_bounds__(256) void k_gibbs_w_final(const double* __restrict__ part, int nb, double* __restrict__ b)
{
    __shared__ double sh[4];
    for (int k = 0; k < 4; ++k) {
        double a = 0.0;
        for (int i = threadIdx.x; i < nb; i += 256) a += part[k * nb + i];
        const double t = block_sum<256>(a, sh);
        if (threadIdx.x == 0) b[k] = t;
    }
}

// generateSignal's data map (:446): map_ - a00 Y00 - a1m1 Y1m1 - a10 Y10 - a11 Y11.
__global__ __launch_bounds__(256) void k_gibbs_d(long npix, const double* __restrict__ map,
                                                 const double* __restrict__ Y, double a0, double a1, double a2,
                                                 double a3, double* __restrict__ out)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < npix; i += (long)gridDim.x * 256) {
        double v = map[i] - a0 * Y[i];
        v = v - a1 * Y[npix + i];
        v = v - a2 * Y[2 * npix + i];
        v = v - a3 * Y[3 * npix + i];
        out[i] = v;
    }
}

// y_p = x_p f_l over the packed layout (l = floor(sqrt(p))).
__global__ __launch_bounds__(256) void k_lscale(double* __restrict__ y, const double* __restrict__ x,
                                                const double* __restrict__ f, int L)
{
    const long n = (long)(L + 1) * (L + 1);
    for (long p = blockIdx.x * 256L + threadIdx.x; p < n; p += (long)gridDim.x * 256L) {
        long l = (long)sqrt((double)p);
        while (l * l > p) --l;
        while ((l + 1) * (l + 1) <= p) ++l;
        y[p] = x[p] * f[l];
    }
}

// calculateSigmaL (:541-549): |s_l0|^2 + sum_m 2 |s_lm|^2, one thread per l.
__global__ __launch_bounds__(256) void k_sigma_l(const double* __restrict__ s, int L, double* __restrict__ out)
{
    const int l = blockIdx.x * 256 + threadIdx.x;
    if (l > L) return;
    const long b = (long)l * l;
    double v = s[b] * s[b];
    for (int m = 1; m <= l; ++m) {
        const double re = s[b + 2 * m - 1], im = s[b + 2 * m];
        v += 2 * (re * re + im * im);
    }
    out[l] = v;
}

}  // namespace cmbl

using namespace cmbl;

struct cmbl_gibbs_sampler {
    int nside = 0, L = 0;
    long npix = 0, n = 0;
    double sigma = 0;
    cmbl_gibbs* gb = nullptr;  // constrained-realisation system (beam D = B sqrt(C))
    cmbl_map* syn = nullptr;   // synthesis with beam B (generateW)
    std::vector<double> beam, cl, sigmaL;
    double G[16] = {};         // sum_i Y_j Y_k mask_i / sigma^2 (the 4x4 operator of generateW)
    double a[4] = {0, 0, 0, 0};
    std::mt19937 gen;
    std::normal_distribution<double> dist{0.0, 1.0};
    // device
    double *d_map = nullptr, *d_mask = nullptr, *d_Y = nullptr, *d_omega = nullptr, *d_s = nullptr, *d_x = nullptr,
           *d_b = nullptr, *d_dm = nullptr, *d_w0 = nullptr, *d_w1 = nullptr, *d_sqcl = nullptr, *d_beam = nullptr,
           *d_sig = nullptr, *d_part = nullptr, *d_b4 = nullptr;
    std::vector<void*> allocs;
    std::vector<double> h_draws;
    ~cmbl_gibbs_sampler()
    {
        (void)hipDeviceSynchronize();
        for (void* p : allocs) (void)hipFree(p);
        cmbl_gibbs_destroy(gb);
        cmbl_map_destroy(syn);
    }
};

namespace cmbl {

static int gs_alloc(cmbl_gibbs_sampler* gs, double** p, size_t count)
{
    void* q = nullptr;
    CMBL_HIP(hipMalloc(&q, sizeof(double) * std::max<size_t>(count, 1)));
    gs->allocs.push_back(q);
    *p = static_cast<double*>(q);
    return CMBL_OK;
}

static int gs_sigma_l(cmbl_gibbs_sampler* gs)
{
    hipLaunchKernelGGL(k_sigma_l, dim3((gs->L + 256) / 256), dim3(256), 0, 0, gs->d_s, gs->L, gs->d_sig);
    CMBL_HIP(hipGetLastError());
    gs->sigmaL.resize(gs->L + 1);
    CMBL_HIP(hipMemcpy(gs->sigmaL.data(), gs->d_sig, sizeof(double) * (gs->L + 1), hipMemcpyDeviceToHost));
    return CMBL_OK;
}

static int gs_upload_sqrt_cl(cmbl_gibbs_sampler* gs)
{
    std::vector<double> q(gs->L + 1);
    for (int l = 0; l <= gs->L; ++l) q[l] = std::sqrt(gs->cl[l]);
    CMBL_HIP(hipMemcpy(gs->d_sqcl, q.data(), sizeof(double) * (gs->L + 1), hipMemcpyHostToDevice));
    return CMBL_OK;
}

// Math::ConjugateGradient::solve (include/conjugate_gradient.hpp:45-102) for
// the 4x4 system of generateW with the identity preconditioner.
static void cg4(const double (&G)[16], const double (&b)[4], double precision, int max_iter, double (&x)[4])
{
    auto mul = [&](const double (&v)[4], double (&out)[4]) {
        for (int k = 0; k < 4; ++k) {
            out[k] = 0;
            for (int j = 0; j < 4; ++j) out[k] += G[4 * k + j] * v[j];
        }
    };
    double r[4], z[4], p[4], ap[4];
    for (int k = 0; k < 4; ++k) x[k] = 0;
    mul(x, r);
    for (int k = 0; k < 4; ++k) r[k] = b[k] - r[k];
    for (int k = 0; k < 4; ++k) z[k] = p[k] = r[k];
    double eps;
    int it = 0;
    do {
        double rz = 0, pap = 0;
        for (int k = 0; k < 4; ++k) rz += r[k] * z[k];
        mul(p, ap);
        for (int k = 0; k < 4; ++k) pap += p[k] * ap[k];
        if (pap == 0 || rz == 0) return;
        eps = 0;
        const double alpha = rz / pap;
        for (int k = 0; k < 4; ++k) {
            x[k] += alpha * p[k];
            r[k] -= alpha * ap[k];
            eps += r[k] * r[k];
        }
        for (int k = 0; k < 4; ++k) z[k] = r[k];
        double rzNew = 0;
        for (int k = 0; k < 4; ++k) rzNew += r[k] * z[k];
        const double beta = rzNew / rz;
        for (int k = 0; k < 4; ++k) p[k] = z[k] + beta * p[k];
    } while (eps > precision * precision && ++it < max_iter);
}

}  // namespace cmbl

extern "C" {

void cmbl_gibbs_sampler_destroy(cmbl_gibbs_sampler* gs) { delete gs; }

int cmbl_gibbs_sampler_create(cmbl_gibbs_sampler** out, int nside, int lmax, const double* map,
                              const double* noise_map, const double* mask, double pixel_noise, double fwhm_deg,
                              const double* beam, const double* starting_cl, long seed)
{
    if (!out || !map || !noise_map || !mask || !starting_cl) return fail(CMBL_ERR_ARG, "null pointer");
    *out = nullptr;
    if (lmax < 2) return fail(CMBL_ERR_ARG, "lmax must be >= 2");
    if (!(pixel_noise > 0.0) || !std::isfinite(pixel_noise)) return fail(CMBL_ERR_ARG, "pixel noise must be > 0");
    if (seed == 0) seed = (long)std::time(nullptr);  // as the reference's default seed (:172-173)
    cmbl_gibbs_sampler* gs = new cmbl_gibbs_sampler();
    auto bail = [&](int rc) {
        const std::string msg = cmbl_last_error();
        delete gs;
        set_error(msg);
        return rc;
    };
    gs->nside = nside;
    gs->L = lmax;
    gs->npix = 12L * nside * nside;
    gs->n = (long)(lmax + 1) * (lmax + 1);
    gs->sigma = pixel_noise;
    gs->gen.seed((unsigned)(int)seed);
    gs->beam.resize(lmax + 1);
    gs->cl.assign(starting_cl, starting_cl + lmax + 1);
    int rc;
    if (beam)
        gs->beam.assign(beam, beam + lmax + 1);
    else if ((rc = cmbl_beam(lmax, fwhm_deg, gs->beam.data())))
        return bail(rc);
    for (int l = 0; l <= lmax; ++l)
        if (!std::isfinite(gs->beam[l]) || gs->beam[l] == 0.0)
            return bail(fail(CMBL_ERR_ARG, "B_l must be finite and nonzero (s = a_lm / B_l)"));
    if ((rc = cmbl_gibbs_create(&gs->gb, nside, lmax, gs->cl.data(), gs->beam.data(), mask, pixel_noise)))
        return bail(rc);
    if ((rc = cmbl_map_create(&gs->syn, nside, lmax)) || (rc = cmbl_map_set_beam(gs->syn, gs->beam.data())))
        return bail(rc);
    const long np = gs->npix, n = gs->n;
    for (auto pr : {std::make_pair(&gs->d_map, np), std::make_pair(&gs->d_mask, np), std::make_pair(&gs->d_Y, 4 * np),
                    std::make_pair(&gs->d_omega, 4 * np), std::make_pair(&gs->d_s, n), std::make_pair(&gs->d_x, n),
                    std::make_pair(&gs->d_b, n), std::make_pair(&gs->d_dm, np), std::make_pair(&gs->d_w0, np),
                    std::make_pair(&gs->d_w1, np), std::make_pair(&gs->d_sqcl, (long)lmax + 1),
                    std::make_pair(&gs->d_beam, (long)lmax + 1), std::make_pair(&gs->d_sig, (long)lmax + 1),
                    std::make_pair(&gs->d_part, 4L * kSamplerBlocks), std::make_pair(&gs->d_b4, 4L)})
        if ((rc = gs_alloc(gs, pr.first, (size_t)pr.second))) return bail(rc);
    // map_ = map + noise (:92-93); monopole and dipole maps (:133-143); their
    // weighted Gram matrix, the 4x4 operator of generateW (:202-230)
    std::vector<double> h(np), Y(4 * np);
    for (long i = 0; i < np; ++i) h[i] = map[i] + noise_map[i];
    const double pi = 3.141592653589793;
    const double y0 = 1.0 / std::sqrt(4.0 * pi), f = std::sqrt(3.0 / (4.0 * pi));
    for (long i = 0; i < np; ++i) {
        double th, ph;
        if ((rc = cmbl_pix2ang_ring(nside, i, &th, &ph))) return bail(rc);
        Y[i] = y0;
        Y[np + i] = f * std::sin(th) * std::sin(ph);
        Y[2 * np + i] = f * std::cos(th);
        Y[3 * np + i] = f * std::sin(th) * std::cos(ph);
    }
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) {
            double g = 0;
            for (long i = 0; i < np; ++i)
                if (mask[i] > 0.5) g += Y[k * np + i] * (Y[j * np + i] / (pixel_noise * pixel_noise));
            gs->G[4 * k + j] = g;
        }
    if (hipMemcpy(gs->d_map, h.data(), sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(gs->d_mask, mask, sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(gs->d_Y, Y.data(), sizeof(double) * 4 * np, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(gs->d_beam, gs->beam.data(), sizeof(double) * (lmax + 1), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(gs->d_dm, map, sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(CMBL_ERR_HIP, "sampler uploads failed"));
    // s = map2alm(map, weight 1) / B_l (:145-159), sigma_l
    if ((rc = cmbl_map2alm_quadrature_device(gs->gb->mp, gs->d_dm, gs->d_s, nullptr)) ||
        (rc = launch_ldiv(gs->d_s, gs->d_s, gs->d_beam, lmax, nullptr)) || (rc = gs_sigma_l(gs)))
        return bail(rc);
    *out = gs;
    return CMBL_OK;
}

int cmbl_gibbs_sampler_step(cmbl_gibbs_sampler* gs, double precision, int max_iter, int* iterations)
{
    if (!gs) return fail(CMBL_ERR_ARG, "null sampler");
    if (!(precision > 0) || max_iter < 1) return fail(CMBL_ERR_ARG, "precision must be > 0 and max_iter >= 1");
    const long np = gs->npix;
    const int L = gs->L;
    // ---- generateW: s B -> map, residual, four draws per pixel, projections
    CMBL_TRY(cmbl_forward_device(gs->syn, gs->d_s, gs->d_dm, nullptr));
    gs->h_draws.resize(4 * np);
    for (long i = 0; i < 4 * np; ++i) gs->h_draws[i] = gs->dist(gs->gen);
    CMBL_HIP(hipMemcpy(gs->d_omega, gs->h_draws.data(), sizeof(double) * 4 * np, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_gibbs_w, dim3(kSamplerBlocks), dim3(256), 0, 0, np, gs->d_map, gs->d_dm, gs->d_mask,
                       gs->d_omega, gs->d_Y, gs->sigma, gs->d_part);
    hipLaunchKernelGGL(k_gibbs_w_final, dim3(1), dim3(256), 0, 0, gs->d_part, kSamplerBlocks, gs->d_b4);
    CMBL_HIP(hipGetLastError());
    double b4[4];
    CMBL_HIP(hipMemcpy(b4, gs->d_b4, sizeof(b4), hipMemcpyDeviceToHost));
    cg4(gs->G, b4, 1e-20, 1000, gs->a);
    // ---- generateSignal: data without monopole / dipole, two draws per pixel
    hipLaunchKernelGGL(k_gibbs_d, dim3(kSamplerBlocks), dim3(256), 0, 0, np, gs->d_map, gs->d_Y, gs->a[0], gs->a[1],
                       gs->a[2], gs->a[3], gs->d_dm);
    CMBL_HIP(hipGetLastError());
    gs->h_draws.resize(2 * np);
    for (long i = 0; i < np; ++i) {
        gs->h_draws[i] = gs->dist(gs->gen);       // omega0
        gs->h_draws[np + i] = gs->dist(gs->gen);  // omega1
    }
    CMBL_HIP(hipMemcpy(gs->d_w0, gs->h_draws.data(), sizeof(double) * np, hipMemcpyHostToDevice));
    CMBL_HIP(hipMemcpy(gs->d_w1, gs->h_draws.data() + np, sizeof(double) * np, hipMemcpyHostToDevice));
    CMBL_TRY(cmbl_gibbs_rhs_device(gs->gb, gs->d_dm, gs->d_w0, gs->d_w1, gs->d_b, nullptr));
    int it = 0;
    const int rc = cmbl_gibbs_solve_device(gs->gb, gs->d_b, gs->d_x, precision, max_iter, &it, nullptr);
    if (iterations) *iterations = it;
    if (rc != CMBL_OK && rc != CMBL_ERR_NOCONV) return rc;
    // fix for sqrt(C_l) (:519-527), sigma_l
    CMBL_TRY(gs_upload_sqrt_cl(gs));
    hipLaunchKernelGGL(k_lscale, dim3((unsigned)std::min<long>((gs->n + 255) / 256, 4096)), dim3(256), 0, 0, gs->d_s,
                       gs->d_x, gs->d_sqcl, L);
    CMBL_HIP(hipGetLastError());
    CMBL_TRY(gs_sigma_l(gs));
    // ---- generateCl (:178-193)
    gs->cl[0] = 0;
    gs->cl[1] = 0;
    for (int l = 2; l <= L; ++l) {
        double rho = 0;
        for (int i = 0; i < 2 * l - 1; ++i) {
            const double r = gs->dist(gs->gen);
            rho += r * r;
        }
        gs->cl[l] = gs->sigmaL[l] / rho;
    }
    CMBL_TRY(cmbl_gibbs_set_cl(gs->gb, gs->cl.data()));
    return rc;
}

int cmbl_gibbs_sampler_cl(const cmbl_gibbs_sampler* gs, double* cl)
{
    if (!gs || !cl) return fail(CMBL_ERR_ARG, "null pointer");
    std::memcpy(cl, gs->cl.data(), sizeof(double) * (gs->L + 1));
    return CMBL_OK;
}

int cmbl_gibbs_sampler_sigma_l(const cmbl_gibbs_sampler* gs, double* sigma_l)
{
    if (!gs || !sigma_l) return fail(CMBL_ERR_ARG, "null pointer");
    std::memcpy(sigma_l, gs->sigmaL.data(), sizeof(double) * (gs->L + 1));
    return CMBL_OK;
}

int cmbl_gibbs_sampler_signal(const cmbl_gibbs_sampler* gs, double* s_packed)
{
    if (!gs || !s_packed) return fail(CMBL_ERR_ARG, "null pointer");
    CMBL_HIP(hipMemcpy(s_packed, gs->d_s, sizeof(double) * gs->n, hipMemcpyDeviceToHost));
    return CMBL_OK;
}

int cmbl_gibbs_sampler_dipole(const cmbl_gibbs_sampler* gs, double* a4)
{
    if (!gs || !a4) return fail(CMBL_ERR_ARG, "null pointer");
    for (int k = 0; k < 4; ++k) a4[k] = gs->a[k];
    return CMBL_OK;
}

}  // extern "C"
