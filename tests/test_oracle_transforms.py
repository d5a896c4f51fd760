"""Pin the oracle's geometry, ring DFTs, synthesis, adjoint and likelihood.

- ring tables vs the per-pixel pix2ang_ring formula (Gorski et al. 2005),
  pixel count, equal-area and north/south symmetry;
- orc_fft vs a naive DFT for every ring length class (powers of two,
  4*i with odd / prime i) and Parseval;
- synthesis vs brute force sum_lm a_lm Y_lm(theta_p, phi_p) with Y from the
  reference SphericalHarmonics (tests/golden/brute_*.npz);
- closed forms: monopole and dipole maps (source/cmb_gibbs.cpp:133-143);
- adjoint identity <Y a, f> = <a, Y^T f>, including aliased rings (lmax > 2 nside);
- likelihood: exact-quadratic central finite differences, w = 0 (prior only),
  noiseless data (data term 0).
"""
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("nside", [1, 2, 4, 8, 32])
def test_ring_tables_match_pix2ang(nside):
    npix = 12 * nside * nside
    seen = np.zeros(npix, dtype=int)
    for r in range(oracle.nring(nside)):
        ri = oracle.ring_info(nside, r)
        for j in range(ri["nph"]):
            p = ri["ofs"] + j
            seen[p] += 1
            th, ph = oracle.pix2ang_ring(nside, p)
            assert abs(math.cos(th) - ri["z"]) < 1e-14
            assert abs(math.sin(th) - ri["sint"]) < 1e-12
            assert abs(ph - (ri["phi0"] + 2 * math.pi * j / ri["nph"])) < 1e-12
    assert np.all(seen == 1)


def test_ring_symmetry_and_area():
    nside = 16
    nr = oracle.nring(nside)
    for r in range(nr):
        a, b = oracle.ring_info(nside, r), oracle.ring_info(nside, nr - 1 - r)
        assert a["nph"] == b["nph"] and abs(a["z"] + b["z"]) < 1e-15 and a["phi0"] == b["phi0"]
    # equal area: the pixels north of ring r's upper edge... check cumulative counts vs cap area
    for i in range(1, nside):
        ri = oracle.ring_info(nside, i - 1)
        # pixels in rings 1..i: 2 i (i+1); ring i centre at 1 - z = i^2/(3 N^2)
        assert abs((1 - ri["z"]) - i * i / (3.0 * nside * nside)) < 1e-15


@pytest.mark.parametrize("n", [1, 2, 4, 8, 64, 4096, 12, 20, 28, 36, 44, 4 * 97, 4 * 255, 4 * 1021])
def test_fft_vs_dft(n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        a = oracle.fft(z, sign)
        if n <= 1024:
            b = oracle.dft(z, sign)
        else:
            b = np.fft.fft(z) if sign < 0 else np.fft.ifft(z) * n
        assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b) * math.log2(max(n, 2)), (n, sign)
    # Parseval
    a = oracle.fft(z, -1)
    assert abs(np.vdot(a, a).real - n * np.vdot(z, z).real) <= 1e-12 * n * np.vdot(z, z).real


@pytest.mark.parametrize("name", ["brute_n8_l16.npz", "brute_n16_l40.npz"])
def test_synthesis_vs_reference_brute_force(name):
    z = np.load(os.path.join(GOLDEN, name))
    nside, lmax = int(z["nside"]), int(z["lmax"])
    s = oracle.synthesis(nside, lmax, z["x"], z["beam"])
    ref = z["map"]
    assert np.linalg.norm(s - ref) <= 1e-13 * np.linalg.norm(ref)


def _packed_single(lmax, l, m, re, im=0.0):
    x = np.zeros((lmax + 1) ** 2)
    if m == 0:
        x[l * l] = re
    else:
        x[l * l + 2 * m - 1] = re
        x[l * l + 2 * m] = im
    return x


def test_monopole_dipole_closed_forms():
    """Y_00, Y_10 and Y_11 maps as the closed forms of source/cmb_gibbs.cpp:133-143."""
    nside, lmax = 8, 12
    npix = 12 * nside * nside
    th = np.array([oracle.pix2ang_ring(nside, p)[0] for p in range(npix)])
    ph = np.array([oracle.pix2ang_ring(nside, p)[1] for p in range(npix)])
    f = math.sqrt(3.0 / (4.0 * math.pi))
    s = oracle.synthesis(nside, lmax, _packed_single(lmax, 0, 0, 1.0))
    np.testing.assert_allclose(s, 1.0 / math.sqrt(4 * math.pi), rtol=1e-14)
    s = oracle.synthesis(nside, lmax, _packed_single(lmax, 1, 0, 1.0))
    np.testing.assert_allclose(s, f * np.cos(th), atol=1e-14)
    # real map of a_11 = (a, b): 2 Re(a_11 Y_11), Y_11 = -f/sqrt2 sin(th) e^{i phi}
    a, b = 0.7, -0.3
    s = oracle.synthesis(nside, lmax, _packed_single(lmax, 1, 1, a, b))
    expect = 2 * (-f / math.sqrt(2)) * np.sin(th) * (a * np.cos(ph) - b * np.sin(ph))
    np.testing.assert_allclose(s, expect, atol=1e-14)


@pytest.mark.parametrize("nside,lmax", [(1, 2), (2, 8), (4, 16), (8, 12), (8, 32), (16, 64)])
def test_adjoint_identity(nside, lmax):
    rng = np.random.default_rng(nside * 100 + lmax)
    x = rng.standard_normal((lmax + 1) ** 2)
    for l in range(lmax + 1):
        x[l * l] = x[l * l]  # Im a_l0 is not a parameter: nothing to zero in the packed layout
    f = rng.standard_normal(12 * nside * nside)
    b = oracle.beam(lmax, 3.0)
    lhs = np.dot(oracle.synthesis(nside, lmax, x, b), f)
    rhs = np.dot(x, oracle.adjoint(nside, lmax, f, b))
    assert abs(lhs - rhs) <= 1e-13 * (abs(lhs) + np.linalg.norm(x) * np.linalg.norm(f))


def _inputs(nside, lmax, seed=5, sigma=10.0):
    rng = np.random.default_rng(seed)
    cl = np.zeros(lmax + 1)
    l = np.arange(2, lmax + 1)
    cl[2:] = 1e3 / (l * (l + 1.0))
    x_true = np.zeros((lmax + 1) ** 2)
    for ll in range(lmax + 1):
        x_true[ll * ll:(ll + 1) ** 2] = rng.standard_normal(2 * ll + 1) * math.sqrt(cl[ll] / 2)
    b = oracle.beam(lmax, 5.0)
    d = oracle.synthesis(nside, lmax, x_true, b) + sigma * rng.standard_normal(12 * nside * nside)
    mask = oracle.mask_latitude(nside, math.radians(20))
    w = np.where(mask > 0.5, 1.0 / sigma ** 2, 0.0)
    x = x_true + 0.1 * rng.standard_normal(x_true.size) * np.sqrt(np.repeat(cl, 2 * np.arange(lmax + 1) + 1) + 1e-3)
    return x, b, cl, d, w, x_true


def test_gradient_finite_differences():
    nside, lmax = 8, 24
    x, b, cl, d, w, _ = _inputs(nside, lmax)
    f, g = oracle.neg2lnL(nside, lmax, x, b, cl, d, w)
    rng = np.random.default_rng(9)
    for _ in range(5):
        v = rng.standard_normal(x.size)
        h = 1e-3 * np.linalg.norm(x) / np.linalg.norm(v)
        fp = oracle.neg2lnL(nside, lmax, x + h * v, b, cl, d, w, grad=False)
        fm = oracle.neg2lnL(nside, lmax, x - h * v, b, cl, d, w, grad=False)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - np.dot(g, v)) <= 1e-8 * abs(np.dot(np.abs(g), np.abs(v)))


def test_prior_only_and_noiseless_cases():
    nside, lmax = 8, 16
    x, b, cl, d, w, x_true = _inputs(nside, lmax)
    # w = 0: -2lnL = prior exactly
    f0, g0 = oracle.neg2lnL(nside, lmax, x, b, cl, d, np.zeros_like(w))
    prior = 0.0
    gp = np.zeros_like(x)
    for l in range(2, lmax + 1):
        prior += (x[l * l] ** 2 + 2 * np.sum(x[l * l + 1:(l + 1) ** 2] ** 2)) / cl[l]
        gp[l * l] = 2 * x[l * l] / cl[l]
        gp[l * l + 1:(l + 1) ** 2] = 4 * x[l * l + 1:(l + 1) ** 2] / cl[l]
    assert abs(f0 - prior) <= 1e-12 * prior
    np.testing.assert_allclose(g0, gp, rtol=1e-13, atol=1e-300)
    # noiseless data at the truth: data term and its gradient vanish
    d0 = oracle.synthesis(nside, lmax, x_true, b)
    f1, g1 = oracle.neg2lnL(nside, lmax, x_true, b, cl, d0, w)
    prior_t = sum((x_true[l * l] ** 2 + 2 * np.sum(x_true[l * l + 1:(l + 1) ** 2] ** 2)) / cl[l]
                  for l in range(2, lmax + 1))
    assert abs(f1 - prior_t) <= 1e-10 * prior_t


def test_mask_regions_counts():
    """A single cap of radius a holds ~ (1 - cos a)/2 of the pixels; maskRegions predicate."""
    nside = 32
    a = math.radians(10.0)
    m = oracle.mask_regions(nside, [1.0], [2.0], [a])
    frac = 1 - m.mean()
    assert abs(frac - (1 - math.cos(a)) / 2) < 0.01
    # brute force of the predicate on a few pixels
    x = np.array([math.sin(1.0) * math.cos(2.0), math.sin(1.0) * math.sin(2.0), math.cos(1.0)])
    for p in range(0, 12 * nside * nside, 97):
        th, ph = oracle.pix2ang_ring(nside, p)
        y = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
        assert (m[p] == 0) == (y[0] * x[0] + y[1] * x[1] + y[2] * x[2] > math.cos(a))
