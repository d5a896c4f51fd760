"""CPU: the oracle's spin-2 synthesis (HEALPix alm2map_pol, called at
source/likelihood.cpp:585 and source/c_matrix_generator.cpp:577-578).

The reference holds no spin-weighted code or polarisation fixture, so the
oracle is pinned by mathematics:
- spin-weighted harmonics from the explicit sum of Goldberg et al. (1967),
  which must reduce to Y_lm (Condon-Shortley phase, the oracle's lambda_lm,
  itself pinned by the reference's KATs) at s = 0, and must equal
  sqrt((l-2)!/(l+2)!) eth eth Y_lm and sqrt((l-2)!/(l+2)!) ethbar ethbar Y_lm
  (finite differences) at s = +-2;
- brute force Q + iU = sum a_2 2Y with a_2 = -(a^E + i a^B) over all m of
  the real field (HEALPix's definitions), at every pixel centre;
- closed form of the l = 2, m = 0 E mode; the E -> B rotation
  (Q, U)[B = a] = (-U, Q)[E = a]; l < 2 of E and B ignored.
What no file of the reference pins is HEALPix's own Q/U sign convention; the
definitions above follow its documentation.
"""
import math

import numpy as np
import pytest

import oracle


def goldberg(s, l, m, th, ph):
    """sY_lm(theta, phi), explicit sum of Goldberg et al. 1967."""
    if l < abs(s) or l < abs(m):
        return 0.0
    f = math.factorial
    pre = (-1) ** m * math.sqrt(f(l + m) * f(l - m) * (2 * l + 1) / (4 * math.pi * f(l + s) * f(l - s)))
    cot = 1.0 / math.tan(th / 2)
    terms = []
    for r in range(0, l - s + 1):
        k = r + s - m
        if 0 <= k <= l + s:
            terms.append(math.comb(l - s, r) * math.comb(l + s, k) * (-1) ** (l - r - s) * cot ** (2 * r + s - m))
    return pre * math.sin(th / 2) ** (2 * l) * math.fsum(terms) * complex(math.cos(m * ph), math.sin(m * ph))


@pytest.mark.parametrize("l,m,th", [(0, 0, 0.3), (3, 1, 0.7), (5, 4, 2.1), (4, 0, 1.3), (7, 6, 2.9), (6, 2, 0.05)])
def test_goldberg_spin0_is_ylm(l, m, th):
    lam = oracle.lam(l, m, math.cos(th), math.sin(th))[l - m]
    ph = 0.37
    y = goldberg(0, l, m, th, ph)
    assert abs(y - lam * complex(math.cos(m * ph), math.sin(m * ph))) < 1e-12 * max(1.0, abs(lam))


def _eth(f, s, th, ph, sign, h=1e-4):
    """eth (sign +1) or ethbar (sign -1) of a spin-s function f(theta, phi):
    eth = -(sin)^s (d_th + i/sin d_ph) (sin)^-s, ethbar = -(sin)^-s (d_th - i/sin d_ph) (sin)^s."""
    p = -s if sign > 0 else s
    g = lambda t: math.sin(t) ** p * f(t, ph)
    dth = (g(th + h) - g(th - h)) / (2 * h)
    dph = (f(th, ph + h) - f(th, ph - h)) / (2 * h) * math.sin(th) ** p
    return -(math.sin(th) ** (-p)) * (dth + sign * 1j / math.sin(th) * dph)


@pytest.mark.parametrize("l,m,th,ph", [(2, 0, 0.9, 0.2), (3, 1, 0.7, 0.4), (4, -2, 1.9, 1.1), (5, 3, 2.4, 2.0)])
def test_goldberg_spin2_from_eth(l, m, th, ph):
    c = math.sqrt(math.factorial(l - 2) / math.factorial(l + 2))
    y0 = lambda t, p: goldberg(0, l, m, t, p)
    y1 = lambda t, p: _eth(y0, 0, t, p, +1)
    yb1 = lambda t, p: _eth(y0, 0, t, p, -1)
    ep = c * _eth(y1, 1, th, ph, +1)
    em = c * _eth(yb1, -1, th, ph, -1)  # (-1)^s with s = 2
    assert abs(ep - goldberg(2, l, m, th, ph)) < 2e-6
    assert abs(em - goldberg(-2, l, m, th, ph)) < 2e-6


def packed_to_complex(x, L):
    """Packed almToVector layout -> dict (l, m) -> complex a_lm, m >= 0."""
    a = {}
    for l in range(L + 1):
        a[(l, 0)] = complex(x[l * l], 0.0)
        for m in range(1, l + 1):
            a[(l, m)] = complex(x[l * l + 2 * m - 1], x[l * l + 2 * m])
    return a


def brute_pol(nside, L, e, b):
    """Q + iU = sum_{l >= 2, |m| <= l} a_2,lm 2Y_lm at each pixel centre,
    a_2 = -(a^E + i a^B), a_{l,-m} = (-1)^m conj(a_lm) for E and B."""
    E, B = packed_to_complex(e, L), packed_to_complex(b, L)
    npix = 12 * nside * nside
    q, u = np.zeros(npix), np.zeros(npix)
    for p in range(npix):
        th, ph = oracle.pix2ang_ring(nside, p)
        z = 0j
        for l in range(2, L + 1):
            for m in range(-l, l + 1):
                if m >= 0:
                    ae, ab = E[(l, m)], B[(l, m)]
                else:
                    sg = (-1) ** (-m)
                    ae, ab = sg * E[(l, -m)].conjugate(), sg * B[(l, -m)].conjugate()
                z += -(ae + 1j * ab) * goldberg(2, l, m, th, ph)
        q[p], u[p] = z.real, z.imag
    return q, u


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nside,L", [(2, 5), (4, 8)])
def test_pol_brute_force(nside, L):
    rng = np.random.default_rng(nside * 100 + L)
    n = (L + 1) ** 2
    t, e, b = rng.standard_normal((3, n))
    mt, mq, mu = oracle.alm2map_pol(nside, L, t, e, b)
    q, u = brute_pol(nside, L, e, b)
    assert rel(mq, q) < 1e-10
    assert rel(mu, u) < 1e-10
    assert np.array_equal(mt, oracle.synthesis(nside, L, t))


def test_pol_e20_closed_form():
    """a^E_20 = 1: Q = -sqrt(15/(32 pi)) sin^2(theta), U = 0."""
    nside, L = 8, 6
    e = np.zeros((L + 1) ** 2)
    e[4] = 1.0
    z = np.zeros_like(e)
    _, q, u = oracle.alm2map_pol(nside, L, z, e, z)
    th = np.array([oracle.pix2ang_ring(nside, p)[0] for p in range(12 * nside * nside)])
    assert np.abs(q + math.sqrt(15.0 / (32.0 * math.pi)) * np.sin(th) ** 2).max() < 1e-14
    assert np.abs(u).max() < 1e-14


def test_pol_rotation_and_low_l():
    nside, L = 8, 16
    rng = np.random.default_rng(3)
    a = rng.standard_normal((L + 1) ** 2)
    z = np.zeros_like(a)
    _, qe, ue = oracle.alm2map_pol(nside, L, z, a, z)
    _, qb, ub = oracle.alm2map_pol(nside, L, z, z, a)
    assert np.allclose(qb, -ue, rtol=0, atol=1e-13 * np.abs(ue).max())
    assert np.allclose(ub, qe, rtol=0, atol=1e-13 * np.abs(qe).max())
    low = np.zeros_like(a)
    low[:4] = rng.standard_normal(4)  # l = 0, 1
    _, q, u = oracle.alm2map_pol(nside, L, z, low, low)
    assert not q.any() and not u.any()
