"""Pin the oracle's map-level synthesis at lmax 2048 and 4096 by a closed form.

A beamed point source at n0, a_lm = B_l Y*_lm(n0), synthesises by the
addition theorem to

    s(n) = sum_l B_l (2l+1)/(4 pi) P_l(n . n0),

a sum over l alone.  The a_lm come from an independent NumPy evaluation of
the normalised Legendre functions (sectoral seed from log-gamma, the
three-term recursion in l), the right-hand side from the Legendre
polynomial recursion (l+1) P_{l+1} = (2l+1) mu P_l - l P_{l-1} written for
t = 1 - mu and D_l = P_l - P_{l-1}:
    (l+1) D_{l+1} = l D_l - (2l+1) t P_l,   P_{l+1} = P_l + D_{l+1},
with t = |n - n0|^2 / 2 from the chord.  Near the source ds/dmu is ~1e11,
so forming mu itself in double precision would cost 1e-10 of the peak;
t keeps its relative precision.  A wrong
phase convention, a dropped (-1)^m, a conjugation or a transposed (l, m)
index moves or distorts the source and fails the comparison; the pixel
centres come from pix2ang_ring, independent of the oracle's ring tables.
Rings: the ones nearest the source, the polar caps, the equator and a
southern ring, at nside 1024 / lmax 2048 and nside 2048 / lmax 4096 (the
synthesis is restricted to the selected rings so the check runs in seconds).
"""
import math

import numpy as np
import pytest

import oracle


def lambda_table_numpy(L, theta):
    """lambda_lm(theta) = Y_lm(theta, 0), Condon-Shortley phase, m-major list of arrays."""
    x, s = math.cos(theta), math.sin(theta)
    m = np.arange(L + 1, dtype=np.float64)
    # lambda_mm = (-1)^m sqrt((2m+1)/(4 pi) (2m)! / (4^m m!^2)) sin^m theta
    lg = np.array([math.lgamma(2 * k + 1) - k * math.log(4.0) - 2 * math.lgamma(k + 1) for k in range(L + 1)])
    logv = 0.5 * (np.log((2 * m + 1) / (4 * math.pi)) + lg) + m * math.log(s)
    seed = np.where(m % 2 == 1, -1.0, 1.0) * np.exp(logv)
    out = np.zeros((L + 1, L + 1))  # [l][m]
    out[np.arange(L + 1), np.arange(L + 1)] = seed
    mm = np.arange(L)
    out[mm + 1, mm] = x * np.sqrt(2 * mm + 3.0) * seed[:-1]
    for l in range(2, L + 1):
        mv = np.arange(0, l - 1, dtype=np.float64)
        a = np.sqrt((4.0 * l * l - 1) / (l * l - mv * mv))
        b = np.sqrt(((l - 1.0) ** 2 - mv * mv) / (4.0 * (l - 1.0) ** 2 - 1))
        out[l, : l - 1] = a * (x * out[l - 1, : l - 1] - b * out[l - 2, : l - 1])
    return out


def point_source_alm(L, theta0, phi0, beam):
    lam = lambda_table_numpy(L, theta0)
    x = np.zeros((L + 1) ** 2)
    for l in range(L + 1):
        x[l * l] = beam[l] * lam[l, 0]
        m = np.arange(1, l + 1)
        # conj(Y_lm(n0)) = lambda_lm(theta0) e^{-i m phi0}
        x[l * l + 2 * m - 1] = beam[l] * lam[l, m] * np.cos(m * phi0)
        x[l * l + 2 * m] = -beam[l] * lam[l, m] * np.sin(m * phi0)
    return x


def legendre_sum(L, beam, t):
    """sum_l B_l (2l+1)/(4 pi) P_l(1 - t)."""
    p = 1.0 - t
    d = -t
    acc = beam[0] / (4 * math.pi) + beam[1] * 3.0 * p / (4 * math.pi)
    for l in range(1, L):
        d = (l * d - (2 * l + 1) * t * p) / (l + 1)
        p = p + d
        acc += beam[l + 1] * (2 * l + 3) * p / (4 * math.pi)
    return acc


@pytest.mark.parametrize("nside,lmax,fwhm_arcmin", [(1024, 2048, 10.0), (2048, 4096, 5.0)])
def test_point_source_addition_theorem(nside, lmax, fwhm_arcmin):
    theta0, phi0 = 1.2, 0.7
    beam = oracle.beam(lmax, fwhm_arcmin / 60.0)
    x = point_source_alm(lmax, theta0, phi0, beam)
    nring = oracle.nring(nside)
    # ring of the source and its neighbours, polar caps, equator, a southern ring
    zs = np.array([oracle.ring_info(nside, r)["z"] for r in range(nring)])
    r0 = int(np.argmin(np.abs(zs - math.cos(theta0))))
    rings = sorted({0, 1, r0 - 1, r0, r0 + 1, nside, 2 * nside - 1, 3 * nside, nring - 1})
    s = oracle.synthesis_rings(nside, lmax, x, rings)
    pix = np.concatenate([oracle.ring_info(nside, r)["ofs"] + np.arange(oracle.ring_info(nside, r)["nph"])
                          for r in rings])
    ang = np.array([oracle.pix2ang_ring(nside, int(p)) for p in pix])
    th, ph = ang[:, 0], ang[:, 1]
    # t = 1 - n . n0 = |n - n0|^2 / 2
    dx = np.sin(th) * np.cos(ph) - math.sin(theta0) * math.cos(phi0)
    dy = np.sin(th) * np.sin(ph) - math.sin(theta0) * math.sin(phi0)
    dz = np.cos(th) - math.cos(theta0)
    ref = legendre_sum(lmax, beam, 0.5 * (dx * dx + dy * dy + dz * dz))
    peak = np.max(np.abs(ref))
    # the source's peak lies on the selected rings (the map is not trivially small there)
    assert peak > 0.5 * np.sum(beam * (2 * np.arange(lmax + 1) + 1)) / (4 * math.pi)
    err = np.max(np.abs(s - ref)) / peak
    print("nside %d lmax %d: max |oracle - closed form| / peak = %.3e" % (nside, lmax, err))
    assert err < 1e-12
    # the check separates a plausible mistake: conjugated a_lm (the source mirrored in phi)
    s_bad = oracle.synthesis_rings(nside, lmax, point_source_alm(lmax, theta0, -phi0, beam), rings)
    assert np.max(np.abs(s_bad - ref)) / peak > 0.1
