"""cosmopp_amd - MI355X-native map-level CMB likelihood (HIP, gfx950, FP64).

Python binding of the C ABI in include/cmbl.h (library
cosmopp_amd/libcosmopp_amd.so).  The compute path is the HIP library only:
if the library is missing or no GPU is usable, every entry point raises.

Reference interface mirrored (aslanyan/cosmopp 3.1):
  MapLikelihood.evaluate(x)            Math::RealFunctionMultiDim::evaluate   include/function.hpp:138
  MapLikelihood.evaluate_grad(x, res)  Math::RealFunctionMultiToMulti::evaluate include/function.hpp:168
  MapLikelihood.calculate / calculate_deriv
                                       Math::LikelihoodWithDerivs            include/likelihood_function.hpp:39,46
  simulate_alm / simulate_white_noise  Simulate::simulateAlm / simulateWhiteNoise source/simulate.cpp:162-203
  beam                                 Utils::beamFunction                   source/utils.cpp:54-64
  mask_regions                         Utils::maskRegions                    source/utils.cpp:235-253
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcosmopp_amd.so")

CMBL_OK = 0
CMBL_ERR_ARG = 1
CMBL_ERR_HIP = 2
CMBL_ERR_ALLOC = 3
CMBL_ERR_STATE = 4
CMBL_ERR_NONFINITE = 5
CMBL_ERR_NOCONV = 6


class CmblError(RuntimeError):
    """Raised for a non-zero status of the C ABI (StandardException analogue)."""

    def __init__(self, code, msg):
        super().__init__("cmbl error %d: %s" % (code, msg))
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_lib = None

# name -> (argtypes, restype)
_SIGS = {
    "cmbl_last_error": ([], ctypes.c_char_p),
    "cmbl_version": ([], ctypes.c_char_p),
    "cmbl_map_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "cmbl_map_destroy": ([_vp], None),
    "cmbl_map_create_tuned": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "cmbl_map_get_tuning": ([_vp, _vp], ctypes.c_int),
    "cmbl_tune": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _dp], ctypes.c_int),
    "cmbl_tuning_set": ([ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "cmbl_tuning_load": ([ctypes.c_char_p], ctypes.c_int),
    "cmbl_tuning_save": ([ctypes.c_char_p], ctypes.c_int),
    "cmbl_map_nside": ([_vp], ctypes.c_int),
    "cmbl_map_lmax": ([_vp], ctypes.c_int),
    "cmbl_map_npix": ([_vp], ctypes.c_long),
    "cmbl_map_nparams": ([_vp], ctypes.c_long),
    "cmbl_map_set_beam": ([_vp, _vp], ctypes.c_int),
    "cmbl_map_set_cl": ([_vp, _vp], ctypes.c_int),
    "cmbl_map_set_data": ([_vp, _vp, _vp], ctypes.c_int),
    "cmbl_map_set_data_device": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_map_generation": ([_vp], ctypes.c_long),
    "cmbl_map_check_finite": ([_vp, _vp], ctypes.c_int),
    "cmbl_batch_check_finite": ([_vp, _vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cmbl_vec_release_workspaces": ([], ctypes.c_int),
    "cmbl_vec_dot_device_static": ([_vp, _vp, ctypes.c_long, _vp, _vp], ctypes.c_int),
    "cmbl_lbfgs_create": ([ctypes.POINTER(_vp), _vp, _vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "cmbl_lbfgs_destroy": ([_vp], None),
    "cmbl_lbfgs_set_starting": ([_vp, _vp], ctypes.c_int),
    "cmbl_lbfgs_minimize": ([_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _vp, _dp,
                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _vp, _vp], ctypes.c_int),
    "cmbl_lbfgs_gradient": ([_vp, _vp], ctypes.c_int),
    "cmbl_lbfgs_state": ([_vp, _dp, _dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
    "cmbl_hmc_create": ([ctypes.POINTER(_vp), _vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                         ctypes.c_int], ctypes.c_int),
    "cmbl_hmc_destroy": ([_vp], None),
    "cmbl_hmc_run": ([_vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "cmbl_hmc_state": ([_vp, _dp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long),
                        ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
    "cmbl_hmc_position": ([_vp, _vp], ctypes.c_int),
    "cmbl_hmc_last_trajectory": ([_vp, _dp, _dp, _dp], ctypes.c_int),
    "cmbl_line_search_create": ([ctypes.POINTER(_vp), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int], ctypes.c_int),
    "cmbl_line_search_destroy": ([_vp], None),
    "cmbl_line_search_begin": ([_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp], ctypes.c_int),
    "cmbl_line_search_feed": ([_vp, ctypes.c_double, ctypes.c_double, _dp, ctypes.POINTER(ctypes.c_int)],
                              ctypes.c_int),
    "cmbl_value_and_grad": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_value_and_grad_device": ([_vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_forward": ([_vp, _vp, _vp], ctypes.c_int),
    "cmbl_forward_device": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_backward": ([_vp, _vp, _vp], ctypes.c_int),
    "cmbl_backward_device": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_map_set_timing": ([_vp, ctypes.c_int], ctypes.c_int),
    "cmbl_map_stage_times": ([_vp, _dp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
    "cmbl_batch_create": ([ctypes.POINTER(_vp), _vp, ctypes.c_int], ctypes.c_int),
    "cmbl_batch_destroy": ([_vp], None),
    "cmbl_batch_value_and_grad_device": ([_vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_gibbs_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_double],
                          ctypes.c_int),
    "cmbl_gibbs_destroy": ([_vp], None),
    "cmbl_gibbs_apply": ([_vp, _vp, _vp], ctypes.c_int),
    "cmbl_gibbs_apply_device": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_gibbs_precondition": ([_vp, _vp, _vp], ctypes.c_int),
    "cmbl_gibbs_precondition_device": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_gibbs_rhs": ([_vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_gibbs_rhs_device": ([_vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_gibbs_solve": ([_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cmbl_gibbs_solve_device": ([_vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _vp],
                                ctypes.c_int),
    "cmbl_gibbs_set_cl": ([_vp, _vp], ctypes.c_int),
    "cmbl_gibbs_sampler_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_double,
                                   ctypes.c_double, _vp, _vp, ctypes.c_long], ctypes.c_int),
    "cmbl_gibbs_sampler_destroy": ([_vp], None),
    "cmbl_gibbs_sampler_step": ([_vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cmbl_gibbs_sampler_cl": ([_vp, _vp], ctypes.c_int),
    "cmbl_gibbs_sampler_sigma_l": ([_vp, _vp], ctypes.c_int),
    "cmbl_gibbs_sampler_signal": ([_vp, _vp], ctypes.c_int),
    "cmbl_gibbs_sampler_dipole": ([_vp, _vp], ctypes.c_int),
    "cmbl_simulate_alm": ([_vp, ctypes.c_int, ctypes.c_long, _vp], ctypes.c_int),
    "cmbl_simulate_white_noise": ([ctypes.c_int, ctypes.c_double, ctypes.c_long, _vp], ctypes.c_int),
    "cmbl_beam": ([ctypes.c_int, ctypes.c_double, _vp], ctypes.c_int),
    "cmbl_mask_latitude": ([ctypes.c_int, ctypes.c_double, _vp], ctypes.c_int),
    "cmbl_mask_regions": ([ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_weights_from_mask": ([ctypes.c_long, _vp, ctypes.c_double, _vp], ctypes.c_int),
    "cmbl_pix2ang_ring": ([ctypes.c_int, ctypes.c_long, _dp, _dp], ctypes.c_int),
    "cmbl_whitenoise": ([ctypes.c_int, ctypes.c_double, ctypes.c_long, _vp], ctypes.c_int),
    "cmbl_map2alm_quadrature": ([_vp, _vp, _vp], ctypes.c_int),
    "cmbl_map2alm_quadrature_device": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_alm2map_pol": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_alm2map_pol_device": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_pseudo_cl": ([_vp, ctypes.c_int, _vp], ctypes.c_int),
    "cmbl_coupling_matrix": ([_vp, ctypes.c_int, _vp], ctypes.c_int),
    "cmbl_coupling_matrix_device": ([_vp, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_master_create": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int],
                           ctypes.c_int),
    "cmbl_master_destroy": ([_vp], None),
    "cmbl_master_nbins": ([_vp], ctypes.c_int),
    "cmbl_master_mask_spectrum": ([_vp, _vp], ctypes.c_int),
    "cmbl_master_coupling": ([_vp, _vp], ctypes.c_int),
    "cmbl_master_calculate": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "cmbl_fits_map_info": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "cmbl_fits_read_map": ([ctypes.c_char_p, ctypes.c_int, _vp, ctypes.c_long], ctypes.c_int),
    "cmbl_fits_write_map": ([ctypes.c_char_p, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "cmbl_nest2ring": ([ctypes.c_int, ctypes.c_long, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
    "cmbl_ring2nest": ([ctypes.c_int, ctypes.c_long, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
    "cmbl_reorder": ([ctypes.c_int, _vp, _vp, ctypes.c_int], ctypes.c_int),
    "cmbl_reorder_device": ([ctypes.c_int, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "cmbl_device_alloc": ([ctypes.POINTER(_vp), ctypes.c_ulong], ctypes.c_int),
    "cmbl_device_free": ([_vp], ctypes.c_int),
    "cmbl_copy_to_device": ([_vp, _vp, ctypes.c_ulong], ctypes.c_int),
    "cmbl_copy_to_host": ([_vp, _vp, ctypes.c_ulong], ctypes.c_int),
    "cmbl_copy_device": ([_vp, _vp, ctypes.c_ulong, _vp], ctypes.c_int),
    "cmbl_stream_synchronize": ([_vp], ctypes.c_int),
    "cmbl_vec_copy": ([_vp, _vp, ctypes.c_double, ctypes.c_long, _vp], ctypes.c_int),
    "cmbl_vec_set": ([_vp, ctypes.c_double, ctypes.c_long, _vp], ctypes.c_int),
    "cmbl_vec_axpy": ([_vp, _vp, ctypes.c_double, ctypes.c_long, _vp], ctypes.c_int),
    "cmbl_vec_mul": ([_vp, _vp, ctypes.c_long, _vp], ctypes.c_int),
    "cmbl_vec_div": ([_vp, _vp, ctypes.c_long, _vp], ctypes.c_int),
    "cmbl_vec_pow": ([_vp, ctypes.c_double, ctypes.c_long, _vp], ctypes.c_int),
    "cmbl_vec_dot": ([_vp, _vp, ctypes.c_long, _dp, _vp], ctypes.c_int),
    "cmbl_vec_dot_device": ([_vp, _vp, ctypes.c_long, _vp, _vp], ctypes.c_int),
}


def declared_symbols():
    """Every entry point bound here (the declarations of include/cmbl.h)."""
    return sorted(_SIGS)


def lib():
    """The loaded HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("cosmopp_amd: %s is missing; run `python -m cosmopp_amd.build` "
                               "(there is no CPU fallback)" % LIB_PATH)
        # A torch wheel for ROCm bundles its own HIP and HSA runtimes.  Its HSA
        # runtime has the soname this library needs (libhsa-runtime64.so.1), so
        # loading torch first makes both share one HSA runtime; loaded the other
        # way round, torch's HIP opens a second HSA runtime that finds no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = res
        _lib = l
    return _lib


def _check(rc):
    if rc != CMBL_OK:
        raise CmblError(rc, lib().cmbl_last_error().decode())


def _arr(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


def _p(a):
    return a.ctypes.data_as(_vp)


# ------------------------------------------------------------------ inputs (host)

def simulate_alm(cl, lmax, seed):
    """Packed a_lm drawn as Simulate::simulateAlm (mt19937 + std::normal_distribution)."""
    cl = _arr(cl)
    out = np.zeros((lmax + 1) ** 2)
    _check(lib().cmbl_simulate_alm(_p(cl), lmax, seed, _p(out)))
    return out


def simulate_white_noise(nside, sigma, seed):
    out = np.zeros(12 * nside * nside)
    _check(lib().cmbl_simulate_white_noise(nside, sigma, seed, _p(out)))
    return out


def beam(lmax, fwhm_deg):
    out = np.zeros(lmax + 1)
    _check(lib().cmbl_beam(lmax, fwhm_deg, _p(out)))
    return out


def mask_latitude(nside, cut_rad, mask=None):
    mask = np.ones(12 * nside * nside) if mask is None else _arr(mask).copy()
    _check(lib().cmbl_mask_latitude(nside, cut_rad, _p(mask)))
    return mask


def mask_regions(nside, theta, phi, angle, mask=None):
    mask = np.ones(12 * nside * nside) if mask is None else _arr(mask).copy()
    th, ph, an = _arr(theta), _arr(phi), _arr(angle)
    _check(lib().cmbl_mask_regions(nside, th.size, _p(th), _p(ph), _p(an), _p(mask)))
    return mask


def weights_from_mask(mask, sigma):
    mask = _arr(mask)
    w = np.zeros_like(mask)
    _check(lib().cmbl_weights_from_mask(mask.size, _p(mask), sigma, _p(w)))
    return w


def pix2ang_ring(nside, pix):
    t, p = ctypes.c_double(), ctypes.c_double()
    _check(lib().cmbl_pix2ang_ring(nside, pix, ctypes.byref(t), ctypes.byref(p)))
    return t.value, p.value


def whitenoise(seed, amplitude, n):
    """BasicLBFGSFunc::whitenoise for one process (host array)."""
    out = np.zeros(n)
    _check(lib().cmbl_whitenoise(seed, amplitude, n, _p(out)))
    return out


def cl_shape(lmax, amp=1e3):
    """C_l = amp / (l (l+1)) for l >= 2, C_0 = C_1 = 0."""
    cl = np.zeros(lmax + 1)
    l = np.arange(2, lmax + 1)
    cl[2:] = amp / (l * (l + 1.0))
    return cl


# ------------------------------------------------------------------ map context

# ------------------------------------------------------------------ map files
def read_map(path, column=1):
    """(map, nside, nested) from a HEALPix FITS file, in the file's ordering
    (read_Healpix_map_from_fits)."""
    nside, nested, npix, ncol = ctypes.c_int(), ctypes.c_int(), ctypes.c_long(), ctypes.c_int()
    b = os.fsencode(path)
    _check(lib().cmbl_fits_map_info(b, ctypes.byref(nside), ctypes.byref(nested), ctypes.byref(npix),
                                    ctypes.byref(ncol)))
    out = np.zeros(npix.value)
    _check(lib().cmbl_fits_read_map(b, column, _p(out), npix.value))
    return out, nside.value, bool(nested.value)


def write_map(path, m, nside, nested=False, overwrite=True):
    """Write a HEALPix FITS map (write_Healpix_map_to_fits), float64."""
    _check(lib().cmbl_fits_write_map(os.fsencode(path), _p(_arr(m, 12 * nside * nside)), nside, int(nested),
                                     int(overwrite)))


def nest2ring(nside, ip):
    out = ctypes.c_long()
    _check(lib().cmbl_nest2ring(nside, int(ip), ctypes.byref(out)))
    return out.value


def ring2nest(nside, ip):
    out = ctypes.c_long()
    _check(lib().cmbl_ring2nest(nside, int(ip), ctypes.byref(out)))
    return out.value


def reorder(m, nside, to_ring):
    """NESTED -> RING (to_ring) or RING -> NESTED on the GPU."""
    m = _arr(m, 12 * nside * nside)
    out = np.zeros_like(m)
    _check(lib().cmbl_reorder(nside, _p(m), _p(out), int(bool(to_ring))))
    return out


class Tuning(ctypes.Structure):
    """Launch shapes of a map (cmbl_tuning): ring pairs per lane of the
    Legendre synthesis and adjoint, Bluestein split length."""
    _fields_ = [("k_synth", ctypes.c_int), ("k_adj", ctypes.c_int), ("split_m", ctypes.c_int)]

    def as_tuple(self):
        return (self.k_synth, self.k_adj, self.split_m)


def tune(nside, lmax, reps=10):
    """Autotuner (cmbl_tune): returns (Tuning, stage times in ms:
    synthesis at k = 1, 2, 4, 8; adjoint at k = 1, 2, 4, 8; ring at split 2048, 4096, 8192)."""
    t = Tuning()
    times = (ctypes.c_double * 11)()
    _check(lib().cmbl_tune(int(nside), int(lmax), int(reps), ctypes.byref(t), times))
    return t, list(times)


def tuning_load(path):
    _check(lib().cmbl_tuning_load(os.fsencode(path)))


def tuning_save(path):
    _check(lib().cmbl_tuning_save(os.fsencode(path)))


class Map:
    """One HEALPix map with its likelihood tables on the current GPU (cmbl_map).
    tuning: (k_synth, k_adj, split_m) to override the tuning table."""

    def __init__(self, nside, lmax, tuning=None):
        h = _vp()
        if tuning is None:
            _check(lib().cmbl_map_create(ctypes.byref(h), nside, lmax))
        else:
            t = Tuning(*tuning)
            _check(lib().cmbl_map_create_tuned(ctypes.byref(h), nside, lmax, ctypes.byref(t)))
        self._h = h
        self.nside = nside
        self.lmax = lmax
        self.npix = 12 * nside * nside
        self.nparams = (lmax + 1) ** 2

    @property
    def handle(self):
        return self._h

    @property
    def tuning(self):
        t = Tuning()
        _check(lib().cmbl_map_get_tuning(self._h, ctypes.byref(t)))
        return t.as_tuple()

    def close(self):
        if self._h:
            lib().cmbl_map_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_beam(self, b):
        b = _arr(b, self.lmax + 1)
        _check(lib().cmbl_map_set_beam(self._h, _p(b)))

    def set_cl(self, cl):
        cl = _arr(cl, self.lmax + 1)
        _check(lib().cmbl_map_set_cl(self._h, _p(cl)))

    def set_data(self, d, w):
        d, w = _arr(d, self.npix), _arr(w, self.npix)
        _check(lib().cmbl_map_set_data(self._h, _p(d), _p(w)))

    def set_data_device(self, d_ptr, w_ptr, stream=0):
        _check(lib().cmbl_map_set_data_device(self._h, d_ptr, w_ptr, stream))

    def value_and_grad(self, x, grad=True):
        x = _arr(x, self.nparams)
        f = np.zeros(1)
        g = np.zeros(self.nparams) if grad else None
        _check(lib().cmbl_value_and_grad(self._h, _p(x), _p(f), None if g is None else _p(g)))
        return (f[0], g) if grad else f[0]

    def value_and_grad_device(self, x_ptr, f_ptr, g_ptr, stream=0):
        _check(lib().cmbl_value_and_grad_device(self._h, x_ptr, f_ptr, g_ptr, stream))

    def check_finite(self, stream=0):
        """Wait for stream; raise CmblError(CMBL_ERR_NONFINITE) if an evaluation
        since the last check produced a non-finite -2lnL."""
        _check(lib().cmbl_map_check_finite(self._h, stream))

    @property
    def generation(self):
        """Number of set_* calls so far (a cached value is stale once it changes)."""
        return lib().cmbl_map_generation(self._h)

    def forward(self, x):
        x = _arr(x, self.nparams)
        out = np.zeros(self.npix)
        _check(lib().cmbl_forward(self._h, _p(x), _p(out)))
        return out

    def forward_device(self, x_ptr, map_ptr, stream=0):
        _check(lib().cmbl_forward_device(self._h, x_ptr, map_ptr, stream))

    def map2alm_quadrature(self, m):
        """HEALPix map2alm with ring weight 1: packed complex a_lm, no beam."""
        m = _arr(m, self.npix)
        out = np.zeros(self.nparams)
        _check(lib().cmbl_map2alm_quadrature(self._h, _p(m), _p(out)))
        return out

    def alm2map_pol(self, t, e, b):
        """HEALPix alm2map_pol: packed T, E, B a_lm (no beam) -> RING maps (T, Q, U)."""
        t, e, b = (_arr(v, self.nparams) for v in (t, e, b))
        out = [np.zeros(self.npix) for _ in range(3)]
        _check(lib().cmbl_alm2map_pol(self._h, _p(t), _p(e), _p(b), *(_p(o) for o in out)))
        return tuple(out)

    def alm2map_pol_device(self, t_ptr, e_ptr, b_ptr, mt_ptr, mq_ptr, mu_ptr, stream=0):
        _check(lib().cmbl_alm2map_pol_device(self._h, t_ptr, e_ptr, b_ptr, mt_ptr, mq_ptr, mu_ptr, stream))

    def backward(self, v):
        v = _arr(v, self.npix)
        out = np.zeros(self.nparams)
        _check(lib().cmbl_backward(self._h, _p(v), _p(out)))
        return out

    def backward_device(self, v_ptr, g_ptr, stream=0):
        _check(lib().cmbl_backward_device(self._h, v_ptr, g_ptr, stream))

    STAGES = ("pre", "synthesis", "ring", "adjoint", "post", "final")

    def set_timing(self, on=True):
        """Enable (and reset) or disable HIP-event timing of the evaluation stages."""
        _check(lib().cmbl_map_set_timing(self._h, 1 if on else 0))

    def stage_times(self):
        """(dict stage -> accumulated ms, number of timed evaluations)."""
        ms = (ctypes.c_double * 6)()
        n = ctypes.c_long()
        _check(lib().cmbl_map_stage_times(self._h, ms, ctypes.byref(n)))
        return dict(zip(self.STAGES, list(ms))), n.value


class Batch:
    """Many maps evaluated with one launch per stage (cmbl_batch)."""

    def __init__(self, maps):
        self.maps = list(maps)
        arr = (_vp * len(self.maps))(*[m.handle for m in self.maps])
        h = _vp()
        _check(lib().cmbl_batch_create(ctypes.byref(h), ctypes.cast(arr, _vp), len(self.maps)))
        self._h = h

    def close(self):
        if self._h:
            lib().cmbl_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def value_and_grad_device(self, x_ptrs, f_ptr, g_ptrs, stream=0):
        n = len(self.maps)
        xs = (_vp * n)(*x_ptrs)
        gs = (_vp * n)(*g_ptrs)
        _check(lib().cmbl_batch_value_and_grad_device(self._h, ctypes.cast(xs, _vp), f_ptr,
                                                      ctypes.cast(gs, _vp), stream))

    def check_finite(self, stream=0):
        """Wait for stream; raise CmblError(CMBL_ERR_NONFINITE) naming the first
        map whose evaluation produced a non-finite -2lnL."""
        first = ctypes.c_int(-1)
        _check(lib().cmbl_batch_check_finite(self._h, stream, ctypes.byref(first)))


def pseudo_cl(alm, lmax):
    """ps(alm, l) (source/master.cpp:62-73) of a packed complex a_lm vector."""
    out = np.zeros(lmax + 1)
    _check(lib().cmbl_pseudo_cl(_p(_arr(alm, (lmax + 1) ** 2)), lmax, _p(out)))
    return out


def coupling_matrix(w, lmax):
    """MASTER mode-coupling matrix from the mask pseudo spectrum w (GPU)."""
    M = np.zeros((lmax + 1, lmax + 1))
    _check(lib().cmbl_coupling_matrix(_p(_arr(w, lmax + 1)), lmax, _p(M)))
    return M


class Master:
    """MASTER estimator (class Master, source/master.cpp): mask and beam fixed,
    optional bin edges; calculate(map) -> (spectrum per bin or per l, pseudo spectrum)."""

    def __init__(self, nside, mask, beam, bins=None, lmax=0):
        h = _vp()
        edges = None if bins is None else np.ascontiguousarray(bins, dtype=np.int32)
        L = int(edges[-1]) - 1 if edges is not None else lmax
        self.lmax = L
        _check(lib().cmbl_master_create(ctypes.byref(h), nside, 0 if edges is not None else lmax,
                                        _p(_arr(mask, 12 * nside * nside)), _p(_arr(beam)[:L + 1].copy()),
                                        None if edges is None else edges.ctypes.data_as(_vp),
                                        0 if edges is None else len(edges)))
        self._h = h
        self.nbins = lib().cmbl_master_nbins(h)
        self.npix = 12 * nside * nside

    def close(self):
        if self._h:
            lib().cmbl_master_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mask_spectrum(self):
        w = np.zeros(self.lmax + 1)
        _check(lib().cmbl_master_mask_spectrum(self._h, _p(w)))
        return w

    def coupling(self):
        M = np.zeros((self.lmax + 1, self.lmax + 1))
        _check(lib().cmbl_master_coupling(self._h, _p(M)))
        return M

    def calculate(self, m):
        ps = np.zeros(self.nbins)
        pseudo = np.zeros(self.lmax + 1)
        _check(lib().cmbl_master_calculate(self._h, _p(_arr(m, self.npix)), _p(ps), _p(pseudo)))
        return ps, pseudo


class Gibbs:
    """The Gibbs constrained-realisation system A x = b of
    CMBGibbsSampler::generateSignal (source/cmb_gibbs.cpp:437-530), with the
    CGTreats interface of Math::ConjugateGradient
    (include/conjugate_gradient.hpp:13-19): multiply_by_matrix / preconditioner,
    plus the right-hand side and the device CG solve."""

    def __init__(self, nside, lmax, cl, beam, mask, pixel_noise):
        self.nside, self.lmax = nside, lmax
        self.n = (lmax + 1) ** 2
        self.npix = 12 * nside * nside
        h = _vp()
        _check(lib().cmbl_gibbs_create(ctypes.byref(h), nside, lmax, _p(_arr(cl, lmax + 1)),
                                       _p(_arr(beam, lmax + 1)), _p(_arr(mask, self.npix)), float(pixel_noise)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().cmbl_gibbs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def multiply_by_matrix(self, x):
        x = _arr(x, self.n)
        y = np.zeros(self.n)
        _check(lib().cmbl_gibbs_apply(self._h, _p(x), _p(y)))
        return y

    def preconditioner(self, x):
        x = _arr(x, self.n)
        y = np.zeros(self.n)
        _check(lib().cmbl_gibbs_precondition(self._h, _p(x), _p(y)))
        return y

    def rhs(self, d, omega0, omega1):
        b = np.zeros(self.n)
        _check(lib().cmbl_gibbs_rhs(self._h, _p(_arr(d, self.npix)), _p(_arr(omega0, self.npix)),
                                    _p(_arr(omega1, self.npix)), _p(b)))
        return b

    def solve(self, b, precision=1e-7, max_iter=10000):
        """(x, iterations); raises CmblError(CMBL_ERR_NOCONV) past max_iter."""
        b = _arr(b, self.n)
        x = np.zeros(self.n)
        it = ctypes.c_int(0)
        _check(lib().cmbl_gibbs_solve(self._h, _p(b), _p(x), float(precision), int(max_iter), ctypes.byref(it)))
        return x, it.value


    def set_cl(self, cl):
        """Replace C_l (D_l = B_l sqrt(C_l) and the preconditioner follow)."""
        _check(lib().cmbl_gibbs_set_cl(self._h, _p(_arr(cl, self.lmax + 1))))


class GibbsSampler:
    """CMBGibbsSampler (include/cmb_gibbs.hpp:12-100, source/cmb_gibbs.cpp:63-680)
    with its maps, a_lm and solves on the GPU.

    Arguments in the order of the reference constructor (:63): map, noise_map
    and mask are RING-ordered arrays (reorder NESTED ones with
    :func:`reorder`); fwhm in degrees; seed 0 takes the time.  beam, if
    given, replaces the Gaussian B_l (the reference multiplies it by the
    HEALPix pixel window, read from HEALPix's data files)."""

    def __init__(self, map, noise_map, mask, pixel_noise, lmax, fwhm, starting_cl, seed=0, beam=None):
        map = np.ascontiguousarray(map, dtype=np.float64)
        npix = map.size
        nside = int(round(math.sqrt(npix / 12)))
        if 12 * nside * nside != npix:
            raise CmblError(CMBL_ERR_ARG, "map size %d is not 12 nside^2" % npix)
        if len(starting_cl) < lmax + 1:
            raise CmblError(CMBL_ERR_ARG, "starting_cl needs lmax + 1 values")
        self.nside, self.lmax, self.npix = nside, lmax, npix
        h = _vp()
        b = None if beam is None else _arr(beam, lmax + 1)
        _check(lib().cmbl_gibbs_sampler_create(ctypes.byref(h), nside, lmax, _p(map), _p(_arr(noise_map, npix)),
                                               _p(_arr(mask, npix)), float(pixel_noise), float(fwhm),
                                               None if b is None else _p(b),
                                               _p(_arr(np.asarray(starting_cl)[:lmax + 1], lmax + 1)), int(seed)))
        self._h = h
        self.iterations = 0

    def close(self):
        if self._h:
            lib().cmbl_gibbs_sampler_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, precision=1e-20, max_iter=100000):
        """CMBGibbsSampler::step (source/cmb_gibbs.cpp:533-538); the signal
        CG's iteration count goes to self.iterations."""
        it = ctypes.c_int(0)
        _check(lib().cmbl_gibbs_sampler_step(self._h, float(precision), int(max_iter), ctypes.byref(it)))
        self.iterations = it.value

    def _get(self, fn, n):
        out = np.zeros(n)
        _check(fn(self._h, _p(out)))
        return out

    def get_cl(self):
        return self._get(lib().cmbl_gibbs_sampler_cl, self.lmax + 1)

    def get_sigma_l(self):
        return self._get(lib().cmbl_gibbs_sampler_sigma_l, self.lmax + 1)

    def signal(self):
        """Current s, packed a_lm ((lmax+1)^2)."""
        return self._get(lib().cmbl_gibbs_sampler_signal, (self.lmax + 1) ** 2)

    def dipole(self):
        """a00, a1-1, a10, a11 of the last generateW."""
        return self._get(lib().cmbl_gibbs_sampler_dipole, 4)

    getCl = get_cl
    getSigmaL = get_sigma_l

    def generate_chain(self, n_samples, burn_in=10):
        """generateChain (source/cmb_gibbs.cpp:572-596): sigma_l after each
        post-burn-in step."""
        if burn_in < 0 or n_samples < 1:
            raise CmblError(CMBL_ERR_ARG, "invalid burn_in %d / n_samples %d" % (burn_in, n_samples))
        for _ in range(burn_in):
            self.step()
        chain = []
        for _ in range(n_samples):
            self.step()
            chain.append(self.get_sigma_l())
        return chain

    @staticmethod
    def calculate_likelihood(cl, sigma_l_or_chain, lmax):
        """-2lnL of C_l given one sigma_l (source/cmb_gibbs.cpp:551-570), or
        the Blackwell-Rao average over a chain (:598-624)."""
        if lmax < 2 or len(cl) < lmax + 1:
            raise CmblError(CMBL_ERR_ARG, "lmax must be >= 2 and cl must hold lmax + 1 values")
        s = np.asarray(sigma_l_or_chain, dtype=np.float64)
        if s.ndim == 1:
            if s.size < lmax + 1:
                raise CmblError(CMBL_ERR_ARG, "sigma_l must hold lmax + 1 values")
            res = 0.0
            for l in range(2, lmax + 1):
                if not cl[l] > 0:
                    raise CmblError(CMBL_ERR_ARG, "C_%d must be > 0" % l)
                sig = s[l] / float(2 * l + 1)
                res -= (2 * l - 1) * math.log(sig)
                res += (2 * l + 1) * math.log(cl[l])
                res += (2 * l + 1) * sig / cl[l]
            return res
        if s.shape[0] == 0:
            raise CmblError(CMBL_ERR_ARG, "the chain is empty")
        like = sorted(GibbsSampler.calculate_likelihood(cl, row, lmax) for row in s)
        acc = 0.0
        for v in like:
            if v - like[0] < 20:
                acc += math.exp(-(v - like[0]) / 2.0)
        acc /= len(like)
        return like[0] - 2 * math.log(acc)

    @staticmethod
    def write_chain(chain, path):
        """writeChainIntoFile (source/cmb_gibbs.cpp:626-648): int32 size,
        int32 lmax, then size x (lmax+1) doubles, native byte order."""
        c = np.asarray(chain, dtype=np.float64)
        if c.ndim != 2 or c.shape[0] == 0 or c.shape[1] == 0:
            raise CmblError(CMBL_ERR_ARG, "the chain is empty")
        with open(path, "wb") as f:
            f.write(np.array([c.shape[0], c.shape[1] - 1], dtype=np.int32).tobytes())
            f.write(np.ascontiguousarray(c).tobytes())

    @staticmethod
    def read_chain(path):
        """readChainFromFile (source/cmb_gibbs.cpp:650-680)."""
        try:
            with open(path, "rb") as f:
                raw = f.read()
        except OSError as e:
            raise CmblError(CMBL_ERR_ARG, "Cannot read the input file %s." % path) from e
        if len(raw) < 8:
            raise CmblError(CMBL_ERR_ARG, "The input file %s is too short." % path)
        size, lmax = (int(v) for v in np.frombuffer(raw[:8], dtype=np.int32))
        if size <= 0:
            raise CmblError(CMBL_ERR_ARG, "The input file %s contains an invalid chain. The size of the chain is %d, "
                            "it needs to be positive." % (path, size))
        if lmax < 2:
            raise CmblError(CMBL_ERR_ARG, "The input file %s contains an invalid chain. l_max = %d, it needs to be "
                            ">= 2." % (path, lmax))
        body = np.frombuffer(raw[8:], dtype=np.float64)
        if body.size < size * (lmax + 1):
            raise CmblError(CMBL_ERR_ARG, "The input file %s holds fewer samples than its header says." % path)
        return [body[i * (lmax + 1):(i + 1) * (lmax + 1)].copy() for i in range(size)]


CMBL_LOOP_EAGER = 1
CMBL_HMC_FIXED = 2

_LBFGS_CB = ctypes.CFUNCTYPE(None, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp)
_HMC_CB = ctypes.CFUNCTYPE(None, _vp, _vp, ctypes.c_double)


class LBFGS:
    """Device-resident LBFGS_General with the More-Thuente line search
    (include/lbfgs_general.hpp:58-405) over a map's -2lnL (cmbl_lbfgs).

    x0_ptr: device pointer to the starting point (nparams doubles).  The
    constructor evaluates there (setStarting).  graphs=False launches the
    same kernels without HIP graphs.
    """

    def __init__(self, mp, x0_ptr, m=10, graphs=True):
        self.map = mp
        h = _vp()
        _check(lib().cmbl_lbfgs_create(ctypes.byref(h), mp.handle, x0_ptr, int(m), 0 if graphs else CMBL_LOOP_EAGER))
        self._h = h

    def close(self):
        if self._h:
            lib().cmbl_lbfgs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_starting(self, x0_ptr):
        _check(lib().cmbl_lbfgs_set_starting(self._h, x0_ptr))

    def minimize(self, res_ptr=None, epsilon=1e-3, gnorm_tol=1e-5, max_iter=1000000, callback=None):
        """LBFGS_General::minimize: returns (-2lnL, iterations, evaluations).
        callback(iter, f, gnorm, x_ptr, g_ptr, z_ptr) is called per iteration."""
        f = ctypes.c_double()
        it, ev = ctypes.c_int(), ctypes.c_int()
        cb = None
        if callback is not None:
            cb = _LBFGS_CB(lambda user, i, fv, gn, x, g, z: callback(i, fv, gn, x, g, z))
        _check(lib().cmbl_lbfgs_minimize(self._h, float(epsilon), float(gnorm_tol), int(max_iter), res_ptr,
                                         ctypes.byref(f), ctypes.byref(it), ctypes.byref(ev),
                                         ctypes.cast(cb, _vp) if cb else None, None))
        return f.value, it.value, ev.value

    def gradient(self, g_ptr):
        _check(lib().cmbl_lbfgs_gradient(self._h, g_ptr))

    def state(self):
        f, gn = ctypes.c_double(), ctypes.c_double()
        it, ev = ctypes.c_int(), ctypes.c_long()
        _check(lib().cmbl_lbfgs_state(self._h, ctypes.byref(f), ctypes.byref(gn), ctypes.byref(it), ctypes.byref(ev)))
        return dict(f=f.value, gnorm=gn.value, iter=it.value, evaluations=ev.value)


class HMC:
    """Device-resident HMCGeneral (include/hmc_general.hpp:64-246) over a
    map's -2lnL (cmbl_hmc).  start_ptr, masses_ptr: device pointers.
    fixed=True uses tau_max and n_max on every trajectory."""

    def __init__(self, mp, start_ptr, masses_ptr, tau_max, n_max, seed, fixed=False, graphs=True):
        self.map = mp
        flags = (CMBL_HMC_FIXED if fixed else 0) | (0 if graphs else CMBL_LOOP_EAGER)
        h = _vp()
        _check(lib().cmbl_hmc_create(ctypes.byref(h), mp.handle, start_ptr, masses_ptr, float(tau_max), int(n_max),
                                     int(seed), flags))
        self._h = h

    def close(self):
        if self._h:
            lib().cmbl_hmc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, max_iters, callback=None):
        """HMCGeneral::run: callback(x_ptr, like) at the start and after every trajectory."""
        cb = None
        if callback is not None:
            cb = _HMC_CB(lambda user, x, like: callback(x, like))
        _check(lib().cmbl_hmc_run(self._h, int(max_iters), ctypes.cast(cb, _vp) if cb else None, None))

    def state(self):
        like = ctypes.c_double()
        t, a, e = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
        _check(lib().cmbl_hmc_state(self._h, ctypes.byref(like), ctypes.byref(t), ctypes.byref(a), ctypes.byref(e)))
        return dict(like=like.value, trajectories=t.value, accepted=a.value, evaluations=e.value)

    def position(self, out_ptr):
        _check(lib().cmbl_hmc_position(self._h, out_ptr))

    def last_trajectory(self):
        """(-2lnL at the end point, kinetic term at start, kinetic term at end)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().cmbl_hmc_last_trajectory(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value


class LineSearch:
    """The More-Thuente line search as a state machine (cmbl_line_search):
    begin(f0, dg0, stp0) -> first step; feed(f, dg) -> (info, step)."""

    def __init__(self, ftol=0.1, gtol=0.5, xtol=1e-15, stpmin=1e-15, stpmax=1e15, maxfev=100):
        h = _vp()
        _check(lib().cmbl_line_search_create(ctypes.byref(h), ftol, gtol, xtol, stpmin, stpmax, maxfev))
        self._h = h

    def close(self):
        if self._h:
            lib().cmbl_line_search_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def begin(self, f0, dg0, stp0):
        stp = ctypes.c_double()
        _check(lib().cmbl_line_search_begin(self._h, f0, dg0, stp0, ctypes.byref(stp)))
        return stp.value

    def feed(self, f, dg):
        stp = ctypes.c_double()
        info = ctypes.c_int()
        _check(lib().cmbl_line_search_feed(self._h, f, dg, ctypes.byref(stp), ctypes.byref(info)))
        return info.value, stp.value


class MapLikelihood:
    """-2lnL(a_lm) with gradient behind the reference's function interfaces.

    evaluate(x) -> -2lnL                      (RealFunctionMultiDim::evaluate)
    evaluate_grad(x, res) -> fills res        (RealFunctionMultiToMulti::evaluate; res is resized)
    calculate(params) / calculate_deriv(params, i)  (LikelihoodWithDerivs)
    The last (x, value, gradient) is cached by contents and by the map's
    generation (set_beam / set_cl / set_data invalidate it), so value +
    derivative at one point cost one device evaluation.
    """

    def __init__(self, mp):
        self.map = mp
        self._x = None
        self._f = None
        self._g = None
        self._gen = None
        self.evaluations = 0

    def _eval(self, x):
        x = _arr(x, self.map.nparams)
        gen = self.map.generation
        if self._x is None or self._gen != gen or not np.array_equal(self._x, x):
            self._f, self._g = self.map.value_and_grad(x)
            self._x = x.copy()
            self._gen = gen
            self.evaluations += 1
        return self._f, self._g

    def evaluate(self, x):
        return self._eval(x)[0]

    def evaluate_grad(self, x, res):
        g = self._eval(x)[1]
        res[:] = g
        return res

    def calculate(self, params):
        return self.evaluate(params)

    def calculate_deriv(self, params, i):
        if not 0 <= i < self.map.nparams:
            raise IndexError("invalid parameter index %d" % i)
        return self._eval(params)[1][i]
