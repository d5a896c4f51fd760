"""GPU: quadrature analysis (map2alm with ring weight 1), the MASTER
mode-coupling matrix (k_coupling) and the MASTER estimator (class Master,
source/master.cpp) through the C ABI, against the oracle.

The estimator runs on the reference's own test mask (tests/golden/mask1.fits,
NESTED, reordered to RING on the GPU), unbinned and binned.
"""
import math
import os
import time

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MASK = os.path.join(ROOT, "tests", "golden", "mask1.fits")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("nside,lmax", [(8, 16), (32, 64), (64, 200)])
def test_quadrature_analysis_matches_oracle(nside, lmax):
    print("case quadrature nside=%d lmax=%d" % (nside, lmax), flush=True)
    rng = np.random.default_rng(nside + lmax)
    m = rng.standard_normal(12 * nside * nside)
    mp = cm.Map(nside, lmax)
    a = mp.map2alm_quadrature(m)
    assert rel(a, oracle.quadrature(nside, lmax, m)) < 1e-10
    assert rel(cm.pseudo_cl(a, lmax), oracle.pseudo_cl(a, lmax)) < 1e-14
    mp.close()


@pytest.mark.parametrize("lmax", [24, 64, 128])
def test_coupling_matrix_matches_oracle(lmax):
    print("case coupling lmax=%d" % lmax, flush=True)
    rng = np.random.default_rng(lmax)
    w = np.abs(rng.standard_normal(lmax + 1)) / (1.0 + np.arange(lmax + 1)) ** 2
    assert rel(cm.coupling_matrix(w, lmax), oracle.coupling(w, lmax)) < 1e-12
    full = np.zeros(lmax + 1)
    full[0] = 4 * np.pi
    assert np.allclose(cm.coupling_matrix(full, lmax), np.eye(lmax + 1), atol=1e-13)


def test_coupling_matrix_lmax2048_rows_and_time():
    """Full size: rows against the oracle's closed form at sampled l1, and the time."""
    lmax = 2048
    print("case coupling lmax=%d" % lmax, flush=True)
    rng = np.random.default_rng(7)
    w = np.abs(rng.standard_normal(lmax + 1)) / (1.0 + np.arange(lmax + 1)) ** 2
    cm.coupling_matrix(w[:33], 32)  # warm up the library
    t0 = time.perf_counter()
    M = cm.coupling_matrix(w, lmax)
    print("coupling lmax 2048: %.1f ms (incl. copies)" % ((time.perf_counter() - t0) * 1e3), flush=True)
    from math import lgamma
    l = np.arange(lmax + 1)
    for l1 in (0, 1, 17, 500, 1999, 2048):
        row = np.zeros(lmax + 1)
        for l2 in range(lmax + 1):
            l3 = np.arange(abs(l1 - l2), min(l1 + l2, lmax) + 1)
            l3 = l3[(l1 + l2 + l3) % 2 == 0]
            if l3.size == 0:
                continue
            g = (l1 + l2 + l3) // 2
            lg = np.vectorize(lgamma)
            v = np.exp(lg(2 * (g - l1) + 1) + lg(2 * (g - l2) + 1) + lg(2 * (g - l3) + 1) - lg(2 * g + 2)
                       + 2 * (lg(g + 1) - lg(g - l1 + 1) - lg(g - l2 + 1) - lg(g - l3 + 1)))
            row[l2] = (2 * l2 + 1) / (4 * np.pi) * np.sum((2 * l3 + 1) * w[l3] * v)
        assert rel(M[l1], row) < 1e-9  # the lgamma closed form is the less accurate side


def ring_mask():
    m, nside, nested = cm.read_map(MASK)
    return cm.reorder(m, nside, to_ring=True), nside


@pytest.mark.parametrize("binned", [False, True])
def test_master_on_reference_mask_matches_oracle(binned):
    print("case master binned=%s" % binned, flush=True)
    mask, nside = ring_mask()
    lmax = 64
    bins = [2, 6, 10, 16, 24, 32, 40, 48, 56, 65] if binned else None
    beam = cm.beam(lmax, 2.0)
    cl = cm.cl_shape(lmax)
    x = cm.simulate_alm(cl, lmax, 33)
    mp = cm.Map(nside, lmax)
    mp.set_beam(beam)
    sky = mp.forward(x) + cm.simulate_white_noise(nside, 0.01, 34)
    mp.close()
    ms = cm.Master(nside, mask, beam, bins=bins, lmax=0 if binned else lmax)
    ps, pseudo = ms.calculate(sky)
    ops, opseudo, W, M = oracle.master(nside, mask, beam, sky, bins=bins, lmax=lmax)
    assert rel(ms.mask_spectrum(), W) < 1e-10
    assert rel(ms.coupling(), M) < 1e-10
    assert rel(pseudo, opseudo) < 1e-10
    assert rel(ps, ops) < 1e-8
    ms.close()
