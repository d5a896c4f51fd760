"""GPU: the Gibbs constrained-realisation system (CmbGibbsCGTreats,
source/cmb_gibbs.cpp:303-434) and its CG solve (include/conjugate_gradient.hpp:45-102)
through the C ABI, against the oracle.

Bar: operator, right-hand side within 1e-10 norm-wise relative; preconditioner
to rounding; CG solution within 1e-8 of the oracle's CG (same algorithm, sums
in a different order) with |A x - b| at the requested precision.
"""
import math

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def problem(nside, lmax, seed, sigma=5.0, cut_deg=20.0, fwhm=3.0):
    rng = np.random.default_rng(seed)
    cl = cm.cl_shape(lmax)
    beam = cm.beam(lmax, fwhm)
    mask = cm.mask_latitude(nside, math.radians(cut_deg))
    return rng, cl, beam, mask, sigma


@pytest.mark.parametrize("nside,lmax", [(4, 16), (8, 16), (16, 40), (32, 64), (64, 100)])
def test_gibbs_apply_precondition_rhs_match_oracle(nside, lmax):
    print("case gibbs nside=%d lmax=%d" % (nside, lmax), flush=True)
    rng, cl, beam, mask, sigma = problem(nside, lmax, nside + lmax)
    n, npix = (lmax + 1) ** 2, 12 * nside * nside
    g = cm.Gibbs(nside, lmax, cl, beam, mask, sigma)
    x = rng.standard_normal(n)
    assert rel(g.multiply_by_matrix(x), oracle.gibbs_apply(nside, lmax, x, cl, beam, mask, sigma)) < 1e-10
    assert rel(g.preconditioner(x), oracle.gibbs_precondition(lmax, x, cl, beam, mask, sigma)) < 1e-15
    d, w0, w1 = (rng.standard_normal(npix) for _ in range(3))
    assert rel(g.rhs(d, w0, w1), oracle.gibbs_rhs(nside, lmax, d, w0, w1, cl, beam, mask, sigma)) < 1e-10
    g.close()


def test_gibbs_split_rings_nside2048():
    """The Gibbs ring pass (f = w s) through the split polar-ring transforms."""
    nside, lmax = 2048, 64
    print("case gibbs split nside=%d lmax=%d" % (nside, lmax), flush=True)
    rng, cl, beam, mask, sigma = problem(nside, lmax, 2048, fwhm=20.0)
    g = cm.Gibbs(nside, lmax, cl, beam, mask, sigma)
    x = rng.standard_normal((lmax + 1) ** 2)
    assert rel(g.multiply_by_matrix(x), oracle.gibbs_apply(nside, lmax, x, cl, beam, mask, sigma)) < 1e-10
    g.close()


def test_gibbs_paired_rings_nside512():
    """The Gibbs ring pass (f = w s) with the equatorial rings (2048 pixels,
    half the launch's 4096) two to a workgroup."""
    nside, lmax = 512, 256
    print("case gibbs paired nside=%d lmax=%d" % (nside, lmax), flush=True)
    rng, cl, beam, mask, sigma = problem(nside, lmax, 512, fwhm=30.0)
    g = cm.Gibbs(nside, lmax, cl, beam, mask, sigma)
    x = rng.standard_normal((lmax + 1) ** 2)
    assert rel(g.multiply_by_matrix(x), oracle.gibbs_apply(nside, lmax, x, cl, beam, mask, sigma)) < 1e-10
    g.close()


@pytest.mark.parametrize("nside,lmax", [(8, 16), (32, 64)])
def test_gibbs_solve_matches_oracle_cg(nside, lmax):
    print("case gibbs solve nside=%d lmax=%d" % (nside, lmax), flush=True)
    rng, cl, beam, mask, sigma = problem(nside, lmax, 7 * nside + lmax)
    npix = 12 * nside * nside
    g = cm.Gibbs(nside, lmax, cl, beam, mask, sigma)
    b = g.rhs(rng.standard_normal(npix) * 30.0, rng.standard_normal(npix), rng.standard_normal(npix))
    prec = 1e-9 * np.linalg.norm(b)
    x, it = g.solve(b, precision=prec)
    A = lambda v: oracle.gibbs_apply(nside, lmax, v, cl, beam, mask, sigma)
    M = lambda v: oracle.gibbs_precondition(lmax, v, cl, beam, mask, sigma)
    xo, ito = oracle.cg_solve(A, M, b, precision=prec)
    assert abs(it - ito) <= 2
    assert rel(x, xo) < 1e-8
    assert np.linalg.norm(A(x) - b) <= 10 * prec
    g.close()


def test_gibbs_iteration_limit_and_argument_errors():
    nside, lmax = 8, 16
    rng, cl, beam, mask, sigma = problem(nside, lmax, 1)
    g = cm.Gibbs(nside, lmax, cl, beam, mask, sigma)
    npix = 12 * nside * nside
    b = g.rhs(rng.standard_normal(npix), rng.standard_normal(npix), rng.standard_normal(npix))
    with pytest.raises(cm.CmblError) as e:
        g.solve(b, precision=1e-30, max_iter=2)
    assert e.value.code == cm.CMBL_ERR_NOCONV
    g.close()
    with pytest.raises(cm.CmblError) as e:
        cm.Gibbs(nside, lmax, cl, beam, np.zeros(npix), sigma)  # no good pixel
    assert e.value.code == cm.CMBL_ERR_ARG
    bad = cl.copy()
    bad[5] = -1.0
    with pytest.raises(cm.CmblError) as e:
        cm.Gibbs(nside, lmax, bad, beam, mask, sigma)
    assert e.value.code == cm.CMBL_ERR_ARG
