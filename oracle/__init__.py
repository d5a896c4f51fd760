"""ctypes binding of the CPU oracle (oracle/cmbl_oracle.c) and, when built,
of the reference shim (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY.  Importable from tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg; the product package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class _Ring(ctypes.Structure):
    _fields_ = [("z", ctypes.c_double), ("sint", ctypes.c_double), ("phi0", ctypes.c_double),
                ("nph", ctypes.c_long), ("ofs", ctypes.c_long)]


def build(reference="/root/reference"):
    """Compile the oracle (and the reference shim when the reference is present)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    if os.path.isdir(os.path.join(reference, "include")):
        subprocess.run(["make", "-s", "-C", _HERE, "-B", "ref", "REFERENCE=" + reference], check=True)
        lib_path = os.path.join(os.path.dirname(_HERE), "cosmopp_amd", "libcosmopp_amd.so")
        if os.path.exists(lib_path):
            subprocess.run(["make", "-s", "-C", _HERE, "-B", "consumers", "REFERENCE=" + reference], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    lib.orc_nring.argtypes = [ctypes.c_int]
    lib.orc_nring.restype = ctypes.c_int
    lib.orc_ring_info.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Ring)]
    lib.orc_pix2ang_ring.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]
    lib.orc_lambda.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
    lib.orc_fft.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
    lib.orc_dft.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
    lib.orc_synthesis.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_void_p, _dp]
    lib.orc_synthesis_rings.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_void_p, ctypes.c_int,
                                        np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), _dp]
    lib.orc_alm2map_pol.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_adjoint.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_void_p, _dp]
    lib.orc_neg2lnL.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_void_p]
    lib.orc_neg2lnL.restype = ctypes.c_double
    lib.orc_beam.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
    lib.orc_mask_latitude.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
    lib.orc_mask_regions.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
    lib.orc_gauss.argtypes = [ctypes.c_long, ctypes.c_long, _dp]
    lib.orc_num_threads.restype = ctypes.c_int
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def nring(nside):
    return lib().orc_nring(nside)


def ring_info(nside, ring):
    r = _Ring()
    lib().orc_ring_info(nside, ring, ctypes.byref(r))
    return dict(z=r.z, sint=r.sint, phi0=r.phi0, nph=r.nph, ofs=r.ofs)


def pix2ang_ring(nside, pix):
    t, p = ctypes.c_double(), ctypes.c_double()
    lib().orc_pix2ang_ring(nside, pix, ctypes.byref(t), ctypes.byref(p))
    return t.value, p.value


def lam(L, m, x, sint):
    out = np.zeros(L - m + 1)
    lib().orc_lambda(L, m, x, sint, out)
    return out


def fft(z, sign):
    re = np.ascontiguousarray(np.real(z), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(z), dtype=np.float64).copy()
    lib().orc_fft(len(re), re, im, sign)
    return re + 1j * im


def dft(z, sign):
    re = np.ascontiguousarray(np.real(z), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(z), dtype=np.float64)
    ore, oim = np.zeros_like(re), np.zeros_like(im)
    lib().orc_dft(len(re), re, im, ore, oim, sign)
    return ore + 1j * oim


def synthesis(nside, L, x, beam=None):
    out = np.zeros(12 * nside * nside)
    x = np.ascontiguousarray(x, dtype=np.float64)
    b = None if beam is None else np.ascontiguousarray(beam, dtype=np.float64)
    lib().orc_synthesis(nside, L, x, _ptr(b), out)
    return out


def synthesis_rings(nside, L, x, rings, beam=None):
    """Synthesis on the listed rings only: the pixels of each ring in turn."""
    rings = np.ascontiguousarray(rings, dtype=np.int32)
    total = sum(ring_info(nside, int(r))["nph"] for r in rings)
    out = np.zeros(total)
    x = np.ascontiguousarray(x, dtype=np.float64)
    b = None if beam is None else np.ascontiguousarray(beam, dtype=np.float64)
    lib().orc_synthesis_rings(nside, L, x, _ptr(b), len(rings), rings, out)
    return out


def alm2map_pol(nside, L, t, e, b):
    """HEALPix alm2map_pol: packed T, E, B -> RING maps (T, Q, U), no beam."""
    npix = 12 * nside * nside
    mt, mq, mu = np.zeros(npix), np.zeros(npix), np.zeros(npix)
    t, e, b = (np.ascontiguousarray(v, dtype=np.float64) for v in (t, e, b))
    lib().orc_alm2map_pol(nside, L, t, e, b, mt, mq, mu)
    return mt, mq, mu


def adjoint(nside, L, m, beam=None):
    out = np.zeros((L + 1) ** 2)
    m = np.ascontiguousarray(m, dtype=np.float64)
    b = None if beam is None else np.ascontiguousarray(beam, dtype=np.float64)
    lib().orc_adjoint(nside, L, m, _ptr(b), out)
    return out


def neg2lnL(nside, L, x, beam, cl, d, w, grad=True):
    x = np.ascontiguousarray(x, dtype=np.float64)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (beam, cl, d, w)]
    g = np.zeros((L + 1) ** 2) if grad else None
    f = lib().orc_neg2lnL(nside, L, x, *args, _ptr(g))
    return (f, g) if grad else f


def beam(L, fwhm_deg):
    out = np.zeros(L + 1)
    lib().orc_beam(L, fwhm_deg, out)
    return out


def mask_latitude(nside, cut_rad, mask=None):
    mask = np.ones(12 * nside * nside) if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
    lib().orc_mask_latitude(nside, cut_rad, mask)
    return mask


def mask_regions(nside, theta, phi, angle, mask=None):
    mask = np.ones(12 * nside * nside) if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
    th, ph, an = (np.ascontiguousarray(a, dtype=np.float64) for a in (theta, phi, angle))
    lib().orc_mask_regions(nside, len(th), th, ph, an, mask)
    return mask


def num_threads():
    return lib().orc_num_threads()


def set_threads(n):
    lib().orc_set_threads(int(n))


# ---------------------------------------------------------------- Gibbs CG system
# CmbGibbsCGTreats (source/cmb_gibbs.cpp:303-434) over the oracle's synthesis
# and adjoint.  HEALPix map2alm returns complex a_lm (u = Y^T_c f); in the
# packed real layout that is the real-parameter transpose (adjoint(), which has
# the factor 2 for m >= 1) halved for m >= 1.

def m_half(L):
    """1 for the packed entries with m = 0, 1/2 for m >= 1."""
    p = np.arange((L + 1) ** 2)
    l = np.floor(np.sqrt(p)).astype(np.int64)
    l -= (l * l > p)
    l += ((l + 1) * (l + 1) <= p)
    return np.where(p == l * l, 1.0, 0.5)


def packed_l(L):
    p = np.arange((L + 1) ** 2)
    l = np.floor(np.sqrt(p)).astype(np.int64)
    l -= (l * l > p)
    l += ((l + 1) * (l + 1) <= p)
    return l


def gibbs_weights(mask, sigma):
    """N^-1 = 1/sigma^2 where mask > 0.5 (source/cmb_gibbs.cpp:381-387)."""
    mask = np.asarray(mask, dtype=np.float64)
    return np.where(mask > 0.5, 1.0 / (sigma * sigma), 0.0)


def gibbs_apply(nside, L, x, cl, beam, mask, sigma):
    """multiplyByMatrix (source/cmb_gibbs.cpp:359-405): x + D Y^T_c N^-1 Y D x, D = B sqrt(C)."""
    D = np.asarray(beam, dtype=np.float64) * np.sqrt(np.asarray(cl, dtype=np.float64))
    s = synthesis(nside, L, x, D)
    u = adjoint(nside, L, gibbs_weights(mask, sigma) * s, D) * m_half(L)
    return np.asarray(x, dtype=np.float64) + u


def gibbs_precondition(L, x, cl, beam, mask, sigma):
    """preconditioner (source/cmb_gibbs.cpp:407-423): x_lm / (1 + C_l B_l^2 / N_l)."""
    good = int(np.count_nonzero(np.asarray(mask) > 0.5))
    nl = sigma * sigma * 4 * np.pi / good
    cl = np.asarray(cl, dtype=np.float64)
    b = np.asarray(beam, dtype=np.float64)
    d = 1.0 + cl * b * b / nl
    return np.asarray(x, dtype=np.float64) / d[packed_l(L)]


def gibbs_rhs(nside, L, d, omega0, omega1, cl, beam, mask, sigma):
    """generateSignal's b (source/cmb_gibbs.cpp:444-487):
    D Y^T_c (N^-1 d + N^-1/2 omega1) + sqrt(4 pi / Npix) Y^T_c omega0."""
    good = np.asarray(mask) > 0.5
    m1 = np.where(good, np.asarray(d) / (sigma * sigma), 0.0)
    o1 = np.where(good, np.asarray(omega1) / sigma, 0.0)
    D = np.asarray(beam, dtype=np.float64) * np.sqrt(np.asarray(cl, dtype=np.float64))
    half = m_half(L)
    npix = 12 * nside * nside
    b = adjoint(nside, L, m1 + o1, D) * half
    b += np.sqrt(4 * np.pi / npix) * adjoint(nside, L, np.asarray(omega0, dtype=np.float64)) * half
    return b


# ---------------------------------------------------------------- MASTER (source/master.cpp)

def quadrature(nside, L, m):
    """HEALPix map2alm with ring weight 1: (4 pi / Npix) Y^T_c m, packed complex a_lm."""
    return 4 * np.pi / (12 * nside * nside) * adjoint(nside, L, m) * m_half(L)


def pseudo_cl(alm, L):
    """ps(alm, l) (source/master.cpp:62-73): (|a_l0|^2 + 2 sum_m |a_lm|^2) / (2l + 1)."""
    a = np.asarray(alm, dtype=np.float64)
    out = np.zeros(L + 1)
    for l in range(L + 1):
        seg = a[l * l:(l + 1) * (l + 1)]
        out[l] = (seg[0] ** 2 + 2 * np.sum(seg[1:] ** 2)) / (2 * l + 1)
    return out


def w3j2_table(L):
    """(l1 l2 l3; 0 0 0)^2, dense (L+1)^3, from the closed form
    (2g-2l1)!(2g-2l2)!(2g-2l3)!/(2g+1)! [g!/((g-l1)!(g-l2)!(g-l3)!)]^2, 2g = l1+l2+l3 even."""
    from math import lgamma
    l = np.arange(L + 1)
    l1, l2, l3 = np.meshgrid(l, l, l, indexing="ij")
    J = l1 + l2 + l3
    ok = (J % 2 == 0) & (l3 >= np.abs(l1 - l2)) & (l3 <= l1 + l2)
    g = J // 2
    lg = np.vectorize(lgamma, otypes=[np.float64])
    out = np.zeros(l1.shape)
    gg, a, b, c = g[ok], (g - l1)[ok], (g - l2)[ok], (g - l3)[ok]
    val = (lg(2 * a + 1) + lg(2 * b + 1) + lg(2 * c + 1) - lg(2 * gg + 2)
           + 2 * (lg(gg + 1) - lg(a + 1) - lg(b + 1) - lg(c + 1)))
    out[ok] = np.exp(val)
    return out


def coupling(w, L, table=None):
    """M[l1][l2] = (2 l2 + 1)/(4 pi) sum_{l3 <= L} (2 l3 + 1) W_l3 (l1 l2 l3; 000)^2
    (Wigner3jTreats::process, source/master.cpp:190-201)."""
    t = w3j2_table(L) if table is None else table
    l = np.arange(L + 1)
    s = np.einsum("abc,c->ab", t, (2 * l + 1) * np.asarray(w, dtype=np.float64)[:L + 1])
    return s * (2 * l[None, :] + 1) / (4 * np.pi)


def lfactor(l):
    l = np.asarray(l, dtype=np.float64)
    return np.where(l == 0, 1.0 / (2 * np.pi), l * (l + 1) / (2 * np.pi))


def master(nside, mask, beam, m, bins=None, lmax=None):
    """Master::construct + calculate (source/master.cpp:76-153, 353-436):
    returns (spectrum per bin or per l, pseudo spectrum, W, M)."""
    L = bins[-1] - 1 if bins is not None else lmax
    W = pseudo_cl(quadrature(nside, L, mask), L)
    M = coupling(W, L)
    b2 = np.asarray(beam, dtype=np.float64)[:L + 1] ** 2
    lf = lfactor(np.arange(L + 1))
    c = pseudo_cl(quadrature(nside, L, np.asarray(m) * np.asarray(mask)), L)
    if bins is None:
        K = M * b2[None, :] * lf[:, None] / lf[None, :]
        v = c * lf
    else:
        nb = len(bins) - 1
        P = np.zeros((nb, L + 1))
        Q = np.zeros((L + 1, nb))
        for b in range(nb):
            ls = np.arange(bins[b], bins[b + 1])
            P[b, ls] = lf[ls] / (bins[b + 1] - bins[b])
            Q[ls, b] = 1.0 / lf[ls]
        K = P @ (M * b2[None, :]) @ Q
        v = P @ c
    return np.linalg.solve(K, v), c, W, M


def cg_solve(apply, precond, b, precision=1e-7, max_iter=None):
    """Math::ConjugateGradient::solve (include/conjugate_gradient.hpp:45-102):
    x0 = 0, r = b - A x0, z = M^-1 r, p = z; loop until |r|^2 <= precision^2.
    Returns (x, iterations)."""
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    r = b - apply(x)
    z = precond(r)
    p = z.copy()
    it = 0
    while True:
        rz = float(np.dot(r, z))
        ap = apply(p)
        pap = float(np.dot(p, ap))
        if pap == 0:
            raise ZeroDivisionError("p.Ap = 0")
        alpha = rz / pap
        x += alpha * p
        r -= alpha * ap
        eps = float(np.dot(r, r))
        z = precond(r)
        rz_new = float(np.dot(r, z))
        beta = rz_new / rz
        p = z + beta * p
        it += 1
        if eps <= precision * precision or (max_iter is not None and it >= max_iter):
            return x, it


# ---------------------------------------------------------------- Gibbs sampler (source/cmb_gibbs.cpp)

def gauss(seed, n):
    """n draws of Math::GaussianGenerator(seed, 0, 1) (include/random.hpp:32-52)."""
    out = np.zeros(int(n))
    lib().orc_gauss(int(seed), int(n), out)
    return out


class _Draws:
    """One GaussianGenerator's stream, consumed in order."""

    def __init__(self, seed):
        self.seed, self.buf, self.pos = int(seed), np.zeros(0), 0

    def take(self, n):
        if self.pos + n > self.buf.size:
            self.buf = gauss(self.seed, max(2 * self.buf.size, self.pos + n, 1 << 16))
        out = self.buf[self.pos:self.pos + n]
        self.pos += n
        return out


def sigma_l(s, L):
    """calculateSigmaL (source/cmb_gibbs.cpp:541-549): |s_l0|^2 + 2 sum_m |s_lm|^2."""
    s = np.asarray(s, dtype=np.float64)
    return np.array([s[l * l] ** 2 + 2 * np.sum(s[l * l + 1:(l + 1) * (l + 1)] ** 2) for l in range(L + 1)])


class GibbsSampler:
    """CMBGibbsSampler (source/cmb_gibbs.cpp:63-549) on the oracle's transforms.

    map, noise and mask RING-ordered; beam B_l (the reference's pixel window
    times the Gaussian, source/cmb_gibbs.cpp:116-117, is the caller's)."""

    def __init__(self, nside, L, m, noise, mask, sigma, beam, starting_cl, seed):
        self.nside, self.L, self.sigma = nside, L, float(sigma)
        self.npix = 12 * nside * nside
        m = np.asarray(m, dtype=np.float64)
        self.data = m + np.asarray(noise, dtype=np.float64)          # :92-93
        self.mask = np.asarray(mask, dtype=np.float64)
        self.good = self.mask > 0.5
        self.beam = np.asarray(beam, dtype=np.float64)[:L + 1].copy()
        self.cl = np.asarray(starting_cl, dtype=np.float64)[:L + 1].copy()
        th = np.zeros(self.npix)
        ph = np.zeros(self.npix)
        for p in range(self.npix):
            th[p], ph[p] = pix2ang_ring(nside, p)
        f = np.sqrt(3.0 / (4.0 * np.pi))                               # :133-143
        self.Y = np.stack([np.full(self.npix, 1.0 / np.sqrt(4.0 * np.pi)), f * np.sin(th) * np.sin(ph),
                           f * np.cos(th), f * np.sin(th) * np.cos(ph)])
        self.s = quadrature(nside, L, m) / self.beam[packed_l(L)]     # :145-159
        self.sigma_l = sigma_l(self.s, L)
        self.a = np.zeros(4)
        self.draws = _Draws(seed)
        self.iterations = 0

    def _w_operator(self, u):
        """CmbGibbsMDTreats::multiplyByMatrix (source/cmb_gibbs.cpp:202-230)."""
        mm = u[0] * self.Y[0] + u[1] * self.Y[1] + u[2] * self.Y[2] + u[3] * self.Y[3]
        mm = np.where(self.good, mm / (self.sigma * self.sigma), 0.0)
        return self.Y @ mm

    def generate_w(self):
        """source/cmb_gibbs.cpp:240-301."""
        v = self.data - synthesis(self.nside, self.L, self.s, self.beam)
        o = self.draws.take(4 * self.npix).reshape(self.npix, 4)
        v = np.where(self.good, v / (self.sigma * self.sigma), 0.0)
        o = np.where(self.good[:, None], o / self.sigma, 0.0)
        v = v + (((o[:, 0] + o[:, 1]) + o[:, 2]) + o[:, 3])
        b = self.Y @ v
        self.a, _ = cg_solve(self._w_operator, lambda r: r.copy(), b, precision=1e-20)

    def generate_signal(self, precision=1e-20, max_iter=None):
        """source/cmb_gibbs.cpp:437-531."""
        o = self.draws.take(2 * self.npix).reshape(self.npix, 2)
        d = self.data - self.a[0] * self.Y[0] - self.a[1] * self.Y[1] - self.a[2] * self.Y[2] - self.a[3] * self.Y[3]
        args = (self.cl, self.beam, self.mask, self.sigma)
        b = gibbs_rhs(self.nside, self.L, d, o[:, 0], o[:, 1], *args)
        x, self.iterations = cg_solve(lambda v: gibbs_apply(self.nside, self.L, v, *args),
                                      lambda v: gibbs_precondition(self.L, v, *args), b, precision, max_iter)
        self.s = x * np.sqrt(self.cl)[packed_l(self.L)]
        self.sigma_l = sigma_l(self.s, self.L)

    def generate_cl(self):
        """source/cmb_gibbs.cpp:178-193."""
        self.cl[0] = self.cl[1] = 0.0
        for l in range(2, self.L + 1):
            r = self.draws.take(2 * l - 1)
            rho = 0.0
            for v in r:
                rho += v * v
            self.cl[l] = self.sigma_l[l] / rho

    def step(self, precision=1e-20, max_iter=None):
        """source/cmb_gibbs.cpp:533-538."""
        self.generate_w()
        self.generate_signal(precision, max_iter)
        self.generate_cl()


# ---------------------------------------------------------------- reference shim
_ref = None


def ref_available():
    return os.path.exists(_REF)


def ref():
    """The reference's own header-only numerics (only in the build container)."""
    global _ref
    if _ref is None:
        r = ctypes.CDLL(_REF)
        r.ref_sh.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        r.ref_legendre.argtypes = [ctypes.c_int, ctypes.c_double]
        r.ref_legendre.restype = ctypes.c_double
        r.ref_assoc_legendre.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        r.ref_assoc_legendre.restype = ctypes.c_double
        r.ref_gauss_draws.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_long, _dp]
        r.ref_uniform_draws.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_long, _dp]
        r.ref_lambda_table.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        r.ref_w3j2_table.argtypes = [ctypes.c_int, _dp]
        r.ref_phi.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double)]
        r.ref_mt_trace.argtypes = [ctypes.c_int] + [ctypes.c_double] * 6 + [ctypes.c_int, _dp, ctypes.c_int,
                                                                            ctypes.POINTER(ctypes.c_int),
                                                                            ctypes.POINTER(ctypes.c_double)]
        r.ref_mt_trace.restype = ctypes.c_int
        r.ref_brute_synthesis.argtypes = [ctypes.c_int, _dp, ctypes.c_void_p, ctypes.c_long, _dp, _dp, _dp]
        _ref = r
    return _ref


def ref_sh(l, m, theta, phi):
    re, im = ctypes.c_double(), ctypes.c_double()
    ref().ref_sh(l, m, theta, phi, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def ref_lambda_table(L, theta):
    out = np.zeros((L + 1) * (L + 2) // 2)
    ref().ref_lambda_table(L, theta, out)
    return out


def ref_brute_synthesis(L, x, beam, theta, phi):
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.zeros(len(theta))
    b = None if beam is None else np.ascontiguousarray(beam, dtype=np.float64)
    ref().ref_brute_synthesis(L, np.ascontiguousarray(x, dtype=np.float64), _ptr(b), len(theta),
                              theta, phi, out)
    return out


def ref_phi(fn, a):
    """More-Thuente test function fn (1..6) and its derivative at a."""
    f, g = ctypes.c_double(), ctypes.c_double()
    ref().ref_phi(fn, a, ctypes.byref(f), ctypes.byref(g))
    return f.value, g.value


def ref_mt_trace(fn, stp0, ftol=0.1, gtol=0.5, xtol=1e-15, stpmin=1e-15, stpmax=1e15, maxfev=100):
    """The reference's moreThuenteSearch on test function fn from 0 along +1:
    (info, trial steps, final step)."""
    trials = np.zeros(maxfev + 1)
    n, stp = ctypes.c_int(), ctypes.c_double()
    info = ref().ref_mt_trace(fn, stp0, ftol, gtol, xtol, stpmin, stpmax, maxfev, trials, trials.size,
                              ctypes.byref(n), ctypes.byref(stp))
    return info, list(trials[:n.value]), stp.value


def ref_w3j2_table(L):
    """(l1 l2 l3; 0 0 0)^2 for l's <= L from the reference's Wigner3JZeroM, dense (L+1)^3."""
    out = np.zeros((L + 1) ** 3)
    ref().ref_w3j2_table(L, out)
    return out.reshape(L + 1, L + 1, L + 1)


def ref_gauss(seed, mean, sigma, n):
    out = np.zeros(n)
    ref().ref_gauss_draws(seed, mean, sigma, n, out)
    return out


def ref_uniform(seed, lo, hi, n):
    out = np.zeros(n)
    ref().ref_uniform_draws(seed, lo, hi, n, out)
    return out
