"""GPU: every ring-pairs-per-lane variant of the Legendre kernels against the
oracle.  The CMBL_K1 / CMBL_K5 overrides (read when a map is created) pick
K = 1, 2, 4 or 8 ring pairs per lane for the synthesis and the adjoint,
whatever the default choice for the map would be.  At lmax
= 3 nside - 1 most (m, ring block) items start with scaled lanes that join
at row ends, so the scaled path (zero adjoint inputs loaded when a lane
turns live, synthesis sums cleared at that point) carries much of the work.
"""
import math

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("nside,lmax,k1,k5", [(64, 191, 1, 1), (64, 191, 2, 2), (64, 191, 4, 4), (64, 191, 8, 8),
                                              (256, 767, 2, 4), (256, 767, 4, 8)])
def test_k_variants_match_oracle(nside, lmax, k1, k5, monkeypatch):
    print("case K variants nside=%d lmax=%d K1=%d K5=%d" % (nside, lmax, k1, k5), flush=True)
    monkeypatch.setenv("CMBL_K1", str(k1))
    monkeypatch.setenv("CMBL_K5", str(k5))
    rng = np.random.default_rng(1000 * k1 + 10 * k5 + nside)
    n = (lmax + 1) ** 2
    # B_lmax = 0.1: every l carries weight in the comparison
    b = cm.beam(lmax, math.degrees(math.sqrt(8.0 * math.log(2.0)) * math.sqrt(2.0 * math.log(10.0) / (lmax * (lmax + 1.0)))))
    mp = cm.Map(nside, lmax)
    mp.set_beam(b)
    x = rng.standard_normal(n)
    assert rel(mp.forward(x), oracle.synthesis(nside, lmax, x, b)) < TOL
    v = rng.standard_normal(12 * nside * nside)
    assert rel(mp.backward(v), oracle.adjoint(nside, lmax, v, b)) < TOL
    cl = cm.cl_shape(lmax)
    d = rng.standard_normal(12 * nside * nside)
    w = cm.weights_from_mask(cm.mask_latitude(nside, math.radians(20.0)), 10.0)
    mp.set_cl(cl)
    mp.set_data(d, w)
    f, g = mp.value_and_grad(x)
    fo, go = oracle.neg2lnL(nside, lmax, x, b, cl, d, w)
    assert abs(f - fo) <= TOL * abs(fo)
    assert rel(g, go) < TOL
    mp.close()
