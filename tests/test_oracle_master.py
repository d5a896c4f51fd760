"""CPU: the oracle's MASTER pieces (source/master.cpp).

- (l1 l2 l3; 0 0 0)^2 from the closed form against the reference's own
  Math::Wigner3JZeroM (include/wigner_3j.hpp) for every l <= 24
  (tests/golden/w3j.npz, from tests/golden/make_golden.py), its orthogonality
  sum_l3 (2 l3 + 1) (...)^2 = 1 and (l 0 l)^2 = 1/(2l+1);
- the coupling matrix is the identity for a full-sky mask (W_l = 4 pi delta_l0),
  the quadrature analysis of a unit map gives that W;
- the pseudo spectrum of a packed a_lm follows ps() of the reference.
"""
import os

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def test_w3j_matches_reference_recursion():
    g = np.load(os.path.join(GOLDEN, "w3j.npz"))
    L = int(g["lmax"][0])
    ref = g["w3j2"]
    o = oracle.w3j2_table(L)
    assert np.array_equal(ref != 0, o != 0)
    nz = ref != 0
    assert np.max(np.abs(o[nz] - ref[nz]) / ref[nz]) < 1e-12


def test_w3j_orthogonality_and_closed_cases():
    L = 60
    t = oracle.w3j2_table(L)
    l = np.arange(L + 1)
    s = np.einsum("abc,c->ab", t, 2 * l + 1)
    for l1 in range(L + 1):
        for l2 in range(L + 1 - l1):  # every l3 <= l1 + l2 is inside the table
            assert abs(s[l1, l2] - 1.0) < 1e-12
    assert np.allclose(t[l, 0, l], 1.0 / (2 * l + 1), rtol=1e-14)


def test_full_sky_coupling_is_identity():
    L = 40
    w = np.zeros(L + 1)
    w[0] = 4 * np.pi
    assert np.allclose(oracle.coupling(w, L), np.eye(L + 1), atol=1e-13)
    nside = 16
    wq = oracle.pseudo_cl(oracle.quadrature(nside, L, np.ones(12 * nside * nside)), L)
    assert abs(wq[0] - 4 * np.pi) < 1e-12 * 4 * np.pi
    assert np.max(np.abs(wq[1:])) < 1e-6 * 4 * np.pi  # HEALPix quadrature of a constant


def test_pseudo_cl_definition():
    L = 5
    rng = np.random.default_rng(1)
    a = rng.standard_normal((L + 1) ** 2)
    ps = oracle.pseudo_cl(a, L)
    for l in range(L + 1):
        tot = a[l * l] ** 2 + 2 * sum(a[l * l + 2 * m - 1] ** 2 + a[l * l + 2 * m] ** 2 for m in range(1, l + 1))
        assert abs(ps[l] - tot / (2 * l + 1)) < 1e-14 * max(tot, 1)
