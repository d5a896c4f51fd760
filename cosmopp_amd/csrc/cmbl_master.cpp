// MASTER pseudo-C_l estimator on the GPU path (Master, source/master.cpp):
// the mask's pseudo spectrum by quadrature analysis, the mode-coupling matrix
// (k_coupling), the K matrix with the beam and optional binning, and the
// binned / unbinned spectrum of a masked map.  K is factorised once (LU with
// partial pivoting, host) and reused for every map.
#include <cmath>
#include <string>
#include <vector>

#include "cmbl_internal.hpp"

namespace cmbl {
namespace {

double lfactor(int l) { return l == 0 ? 1.0 / (2 * 3.141592653589793) : l * (l + 1.0) / (2 * 3.141592653589793); }

// ps(alm, l) of source/master.cpp:62-73 on the packed layout: (|a_l0|^2 + 2 sum |a_lm|^2) / (2l + 1).
void pseudo_from_packed(const double* a, int L, double* cl)
{
    for (int l = 0; l <= L; ++l) {
        const long b = (long)l * l;
        double r = a[b] * a[b];
        for (int m = 1; m <= l; ++m) r += 2 * (a[b + 2 * m - 1] * a[b + 2 * m - 1] + a[b + 2 * m] * a[b + 2 * m]);
        cl[l] = r / (2 * l + 1);
    }
}

// In-place LU with partial pivoting of the n x n row-major matrix; false if singular.
bool lu_factor(std::vector<double>& A, std::vector<int>& piv, int n)
{
    piv.resize(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)p * n + k])) p = i;
        piv[k] = p;
        if (A[(size_t)p * n + k] == 0.0) return false;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
        const double inv = 1.0 / A[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            double* ri = &A[(size_t)i * n];
            const double* rk = &A[(size_t)k * n];
            const double f = ri[k] * inv;
            ri[k] = f;
            for (int j = k + 1; j < n; ++j) ri[j] -= f * rk[j];
        }
    }
    return true;
}

void lu_solve(const std::vector<double>& A, const std::vector<int>& piv, int n, double* x)
{
    for (int k = 0; k < n; ++k) std::swap(x[k], x[piv[k]]);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) x[i] -= A[(size_t)i * n + j] * x[j];
    for (int i = n - 1; i >= 0; --i) {
        for (int j = i + 1; j < n; ++j) x[i] -= A[(size_t)i * n + j] * x[j];
        x[i] /= A[(size_t)i * n + i];
    }
}

int coupling_host(const double* w, int L, std::vector<double>* M)
{
    const size_t n = (size_t)(L + 1);
    double *d_w = nullptr, *d_s0 = nullptr, *d_M = nullptr;
    M->assign(n * n, 0.0);
    int rc = CMBL_OK;
    if (hipMalloc(&d_w, sizeof(double) * n) != hipSuccess || hipMalloc(&d_s0, sizeof(double) * n * n) != hipSuccess ||
        hipMalloc(&d_M, sizeof(double) * n * n) != hipSuccess)
        rc = fail(CMBL_ERR_ALLOC, "hipMalloc failed for the coupling matrix");
    if (rc == CMBL_OK && hipMemcpy(d_w, w, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(CMBL_ERR_HIP, "copy of W failed");
    if (rc == CMBL_OK) rc = launch_coupling(d_w, L, d_s0, d_M, nullptr);
    if (rc == CMBL_OK && hipMemcpy(M->data(), d_M, sizeof(double) * n * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(CMBL_ERR_HIP, "copy of the coupling matrix failed");
    if (d_w) (void)hipFree(d_w);
    if (d_s0) (void)hipFree(d_s0);
    if (d_M) (void)hipFree(d_M);
    return rc;
}

}  // namespace
}  // namespace cmbl

struct cmbl_master {
    cmbl_map* mp = nullptr;
    int L = 0, nb = 0;
    long npix = 0, n = 0;
    std::vector<double> mask, beam, W, M, K;
    std::vector<int> bins, piv;
};

using namespace cmbl;

extern "C" {

int cmbl_pseudo_cl(const double* alm, int lmax, double* cl)
{
    if (!alm || !cl || lmax < 0) return fail(CMBL_ERR_ARG, "bad arguments");
    pseudo_from_packed(alm, lmax, cl);
    return CMBL_OK;
}

int cmbl_coupling_matrix(const double* w, int lmax, double* M)
{
    if (!w || !M || lmax < 0) return fail(CMBL_ERR_ARG, "bad arguments");
    std::vector<double> h;
    CMBL_TRY(coupling_host(w, lmax, &h));
    std::copy(h.begin(), h.end(), M);
    return CMBL_OK;
}

int cmbl_coupling_matrix_device(const double* d_w, int lmax, double* d_work, double* d_M, void* stream)
{
    if (!d_w || !d_work || !d_M || lmax < 0) return fail(CMBL_ERR_ARG, "bad arguments");
    return launch_coupling(d_w, lmax, d_work, d_M, (hipStream_t)stream);
}

void cmbl_master_destroy(cmbl_master* m)
{
    if (!m) return;
    cmbl_map_destroy(m->mp);
    delete m;
}

int cmbl_master_create(cmbl_master** out, int nside, int lmax, const double* mask, const double* beam, const int* bins,
                       int nedges)
{
    if (!out || !mask || !beam) return fail(CMBL_ERR_ARG, "null pointer");
    *out = nullptr;
    int L = lmax;
    if (nedges > 0) {
        if (!bins || nedges < 2) return fail(CMBL_ERR_ARG, "bins need at least two edges");
        for (int i = 0; i < nedges; ++i)
            if (bins[i] < 0 || (i && bins[i] <= bins[i - 1])) return fail(CMBL_ERR_ARG, "bin edges must increase");
        L = bins[nedges - 1] - 1;  // source/master.cpp:91
        if (lmax > 0 && lmax != L) return fail(CMBL_ERR_ARG, "lmax must be 0 or the last bin edge - 1");
    }
    if (L < 2) return fail(CMBL_ERR_ARG, "lmax must be >= 2");
    cmbl_master* m = new cmbl_master();
    auto bail = [&](int rc) {
        const std::string msg = cmbl_last_error();
        cmbl_master_destroy(m);
        set_error(msg);
        return rc;
    };
    int rc = cmbl_map_create(&m->mp, nside, L);
    if (rc) return bail(rc);
    m->L = L;
    m->npix = 12L * nside * nside;
    m->n = (long)(L + 1) * (L + 1);
    m->mask.assign(mask, mask + m->npix);
    m->beam.assign(beam, beam + L + 1);
    if (nedges > 0) m->bins.assign(bins, bins + nedges);
    m->nb = nedges > 0 ? nedges - 1 : L + 1;
    // mask pseudo spectrum: map2alm with weight 1 (source/master.cpp:112-125)
    std::vector<double> alm(m->n);
    if ((rc = cmbl_map2alm_quadrature(m->mp, m->mask.data(), alm.data()))) return bail(rc);
    m->W.resize(L + 1);
    pseudo_from_packed(alm.data(), L, m->W.data());
    if ((rc = coupling_host(m->W.data(), L, &m->M))) return bail(rc);
    // K (source/master.cpp:353-402)
    const int nb = m->nb;
    m->K.assign((size_t)nb * nb, 0.0);
    if (nedges > 0) {
        for (int b = 0; b < nb; ++b)
            for (int b1 = 0; b1 < nb; ++b1) {
                double k = 0.0;
                for (int l = m->bins[b]; l < m->bins[b + 1]; ++l) {
                    const double p = lfactor(l) / (m->bins[b + 1] - m->bins[b]);
                    for (int l1 = m->bins[b1]; l1 < m->bins[b1 + 1]; ++l1)
                        k += p * m->M[(size_t)l * (L + 1) + l1] * m->beam[l1] * m->beam[l1] / lfactor(l1);
                }
                m->K[(size_t)b * nb + b1] = k;
            }
    } else {
        for (int l = 0; l <= L; ++l)
            for (int l1 = 0; l1 <= L; ++l1)
                m->K[(size_t)l * nb + l1] =
                    m->M[(size_t)l * (L + 1) + l1] * m->beam[l1] * m->beam[l1] * lfactor(l) / lfactor(l1);
    }
    if (!lu_factor(m->K, m->piv, nb)) {
        set_error("the K matrix is singular (mask or beam leave a mode unconstrained)");
        return bail(CMBL_ERR_STATE);
    }
    *out = m;
    return CMBL_OK;
}

int cmbl_master_nbins(const cmbl_master* m) { return m ? m->nb : -1; }

int cmbl_master_mask_spectrum(const cmbl_master* m, double* w)
{
    if (!m || !w) return fail(CMBL_ERR_ARG, "null pointer");
    std::copy(m->W.begin(), m->W.end(), w);
    return CMBL_OK;
}

int cmbl_master_coupling(const cmbl_master* m, double* M)
{
    if (!m || !M) return fail(CMBL_ERR_ARG, "null pointer");
    std::copy(m->M.begin(), m->M.end(), M);
    return CMBL_OK;
}

int cmbl_master_calculate(cmbl_master* m, const double* map, double* ps, double* pseudo)
{
    if (!m || !map || !ps) return fail(CMBL_ERR_ARG, "null pointer");
    const int L = m->L, nb = m->nb;
    // masked map -> map2alm with weight 1 -> pseudo spectrum (source/master.cpp:132-153)
    std::vector<double> masked(m->npix), alm(m->n), c(L + 1);
    for (long i = 0; i < m->npix; ++i) masked[i] = map[i] * m->mask[i];
    CMBL_TRY(cmbl_map2alm_quadrature(m->mp, masked.data(), alm.data()));
    pseudo_from_packed(alm.data(), L, c.data());
    if (pseudo) std::copy(c.begin(), c.end(), pseudo);
    // K^-1 applied to the binned / l-weighted pseudo spectrum (source/master.cpp:410-436)
    std::vector<double> v(nb, 0.0);
    if (!m->bins.empty()) {
        for (int b = 0; b < nb; ++b)
            for (int l = m->bins[b]; l < m->bins[b + 1]; ++l)
                v[b] += lfactor(l) / (m->bins[b + 1] - m->bins[b]) * c[l];
    } else {
        for (int l = 0; l <= L; ++l) v[l] = c[l] * lfactor(l);
    }
    lu_solve(m->K, m->piv, nb, v.data());
    std::copy(v.begin(), v.end(), ps);
    return CMBL_OK;
}

}  // extern "C"
