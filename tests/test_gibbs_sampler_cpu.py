"""CPU: the oracle's restatement of CMBGibbsSampler (source/cmb_gibbs.cpp:63-549)
and the host side of the sampler's interface.

What pins the oracle here:
- its draws equal the reference's GaussianGenerator bit for bit
  (tests/golden/rng.npz, and the live reference generator where oracle/_ref
  builds);
- the noise terms of generateSignal's right-hand side have the covariance
  that makes x = A^-1 b a draw of the conditional signal: with M = 1 for
  m = 0 and 2 for m > 0 (the packed complex convention), A M^-1 is symmetric
  and Cov(b) = A M^-1 up to HEALPix's quadrature error, so the mean of
  b^T M A^-1 b over draws is the dimension (a dropped sqrt(4pi/Npix), a 1/sigma
  for 1/sigma^2 or a lost m > 0 factor move it by far more than the bound);
- the sampled chain recovers the simulated sky's spectrum, the acceptance
  criterion of the reference's own test (source/test_cmb_gibbs.cpp:206-229).
"""
import os

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle
from cases import gibbs_acceptance, gibbs_case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_gauss_draws_equal_reference_fixture():
    z = np.load(os.path.join(GOLDEN, "rng.npz"))
    assert np.array_equal(oracle.gauss(int(z["gauss_seed"][0]), z["gauss"].size), z["gauss"])
    # normal_distribution(0, 20) returns 20 z + 0 of the same stream
    assert np.array_equal(20.0 * oracle.gauss(int(z["gauss_sigma_seed"][0]), z["gauss_sigma"].size),
                          z["gauss_sigma"])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref (reference generator) not built here")
@pytest.mark.parametrize("seed", [1, 21, 123456789, -5, 2 ** 31 - 1])
def test_gauss_draws_equal_live_reference(seed):
    assert np.array_equal(oracle.gauss(seed, 5000), oracle.ref_gauss(seed, 0.0, 1.0, 5000))


def test_rhs_noise_covariance_is_the_operator():
    nside, L, sigma = 8, 6, 2.0
    npix, n = 12 * nside * nside, (L + 1) ** 2
    cl = np.full(L + 1, 4 * np.pi * sigma ** 2 / npix) * (1 + 0.5 * np.sin(np.arange(L + 1)))
    beam = np.exp(-0.01 * np.arange(L + 1) ** 2)
    mask = np.ones(npix)
    A = np.stack([oracle.gibbs_apply(nside, L, e, cl, beam, mask, sigma) for e in np.eye(n)], axis=1)
    Mv = np.where(np.arange(n) == oracle.packed_l(L) ** 2, 1.0, 2.0)
    AM = A / Mv[None, :]
    assert np.abs(AM - AM.T).max() < 1e-12 * np.abs(AM).max()
    W = Mv[:, None] * np.linalg.inv(A)
    rng = np.random.default_rng(5)
    K = 3000
    q = np.empty(K)
    zero = np.zeros(npix)
    for k in range(K):
        b = oracle.gibbs_rhs(nside, L, zero, rng.standard_normal(npix), rng.standard_normal(npix), cl, beam, mask,
                             sigma)
        q[k] = b @ W @ b
    # standard error of the mean is ~0.18 here; quadrature error ~0.3
    assert abs(q.mean() - n) < 1.2, q.mean()


def test_oracle_sampler_recovers_the_realisation():
    c = gibbs_case(0)
    gs = oracle.GibbsSampler(c["nside"], c["lmax"], c["map"], c["noise"], c["mask"], c["sigma"], c["beam"], c["cl0"],
                             c["seed"])
    chain = []
    for k in range(110):
        gs.step()
        assert gs.cl[0] == 0.0 and gs.cl[1] == 0.0
        assert np.all(gs.cl[2:] > 0)
        if k >= 10:
            chain.append(gs.sigma_l.copy())
    assert gibbs_acceptance(c, chain) == []


def test_chain_likelihood_minimum_and_blackwell_rao():
    L = 12
    rng = np.random.default_rng(3)
    sig = rng.uniform(1.0, 5.0, L + 1) * (2 * np.arange(L + 1) + 1)
    best = sig / (2 * np.arange(L + 1) + 1)  # d/dC_l of (2l+1)(ln C + sigma/C) vanishes at C = sigma
    f0 = cm.GibbsSampler.calculate_likelihood(best, sig, L)
    for l in range(2, L + 1):
        for f in (0.97, 1.03):
            c = best.copy()
            c[l] *= f
            assert cm.GibbsSampler.calculate_likelihood(c, sig, L) > f0
    # a chain of identical samples averages to the single-sample value
    assert cm.GibbsSampler.calculate_likelihood(best, [sig] * 7, L) == pytest.approx(f0, rel=1e-14)
    # two samples far apart in likelihood: only the best one counts (cut at 20)
    far = sig * 50
    v = cm.GibbsSampler.calculate_likelihood(best, [sig, far], L)
    assert v == pytest.approx(f0 - 2 * np.log(0.5), rel=1e-12)
    with pytest.raises(cm.CmblError):
        cm.GibbsSampler.calculate_likelihood(np.zeros(L + 1), sig, L)
    with pytest.raises(cm.CmblError):
        cm.GibbsSampler.calculate_likelihood(best, sig, 1)


def test_chain_file_round_trip(tmp_path):
    chain = [np.arange(6, dtype=np.float64) + k for k in range(4)]
    p = str(tmp_path / "chain.bin")
    cm.GibbsSampler.write_chain(chain, p)
    raw = open(p, "rb").read()
    assert np.frombuffer(raw[:8], dtype=np.int32).tolist() == [4, 5]
    assert len(raw) == 8 + 4 * 6 * 8
    back = cm.GibbsSampler.read_chain(p)
    assert len(back) == 4 and all(np.array_equal(a, b) for a, b in zip(chain, back))
    bad = str(tmp_path / "bad.bin")
    with open(bad, "wb") as f:
        f.write(np.array([3, 1], dtype=np.int32).tobytes())
    with pytest.raises(cm.CmblError, match="l_max = 1"):
        cm.GibbsSampler.read_chain(bad)
    with pytest.raises(cm.CmblError):
        cm.GibbsSampler.read_chain(str(tmp_path / "missing.bin"))


def test_sampler_rejects_bad_arguments_before_the_gpu():
    npix = 12 * 16 * 16
    z, one = np.zeros(npix), np.ones(npix)
    with pytest.raises(cm.CmblError, match="lmax must be >= 2"):
        cm.GibbsSampler(z, z, one, 1.0, 1, 10.0, np.ones(2))
    with pytest.raises(cm.CmblError, match="pixel noise"):
        cm.GibbsSampler(z, z, one, 0.0, 8, 10.0, np.ones(9))
    with pytest.raises(cm.CmblError):
        cm.GibbsSampler(np.zeros(100), np.zeros(100), np.ones(100), 1.0, 8, 10.0, np.ones(9))
