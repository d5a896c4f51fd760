"""GPU: CMBGibbsSampler (source/cmb_gibbs.cpp:63-549) on the device against the
oracle's restatement, at the sizes of the reference's own sampler test
(source/test_cmb_gibbs.cpp:64-99: nside 16 / Gibbs lmax 30, nside 32 / lmax 50,
and nside 32 / lmax 50 on the reference's mask1.fits).

Both sides consume one GaussianGenerator stream in the reference's order, so
after each step the monopole / dipole amplitudes, the signal s, sigma_l and
C_l agree to round-off (1e-10 relative); the chain then passes the reference
test's acceptance criterion (source/test_cmb_gibbs.cpp:206-229).
"""
import os

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle
from cases import gibbs_acceptance, gibbs_case

pytestmark = pytest.mark.gpu

MASK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mask1.fits")


def _ring_mask1():
    m, nside, nested = cm.read_map(MASK)
    return cm.reorder(m, nside, to_ring=True) if nested else m


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def _samplers(i, explicit_beam=True):
    c = gibbs_case(i, ring_mask=_ring_mask1() if i == 4 else None)
    dev = cm.GibbsSampler(c["map"], c["noise"], c["mask"], c["sigma"], c["lmax"], c["fwhm"], c["cl0"],
                          seed=c["seed"], beam=c["beam"] if explicit_beam else None)
    orc = oracle.GibbsSampler(c["nside"], c["lmax"], c["map"], c["noise"], c["mask"], c["sigma"], c["beam"],
                              c["cl0"], c["seed"])
    return c, dev, orc


@pytest.mark.parametrize("i,steps", [(0, 3), (1, 2), (4, 2)])
def test_sampler_steps_match_oracle(i, steps):
    print("case gibbs sampler subtest=%d steps=%d" % (i, steps), flush=True)
    c, dev, orc = _samplers(i, explicit_beam=(i != 0))
    assert _rel(dev.signal(), orc.s) < 1e-10
    assert _rel(dev.get_sigma_l(), orc.sigma_l) < 1e-10
    for k in range(steps):
        dev.step()
        orc.step()
        print("  step %d: CG iterations device %d oracle %d" % (k, dev.iterations, orc.iterations), flush=True)
        assert _rel(dev.dipole(), orc.a) < 1e-10
        assert _rel(dev.signal(), orc.s) < 1e-10
        assert _rel(dev.get_sigma_l(), orc.sigma_l) < 1e-10
        cl = dev.get_cl()
        assert cl[0] == 0.0 and cl[1] == 0.0
        assert np.max(np.abs(cl[2:] - orc.cl[2:]) / orc.cl[2:]) < 1e-10
        assert abs(dev.iterations - orc.iterations) <= max(2, orc.iterations // 50)


def test_sampler_is_deterministic_for_a_seed():
    print("case gibbs sampler determinism", flush=True)
    c = gibbs_case(0)
    runs = []
    for _ in range(2):
        gs = cm.GibbsSampler(c["map"], c["noise"], c["mask"], c["sigma"], c["lmax"], c["fwhm"], c["cl0"],
                             seed=c["seed"])
        out = []
        for _ in range(3):
            gs.step()
            out.append((gs.get_cl(), gs.signal(), gs.dipole()))
        gs.close()
        runs.append(out)
    for a, b in zip(*runs):
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("i,samples", [(0, 300), (1, 100)])
def test_device_chain_recovers_the_realisation(i, samples):
    print("case gibbs chain subtest=%d samples=%d" % (i, samples), flush=True)
    c = gibbs_case(i)
    gs = cm.GibbsSampler(c["map"], c["noise"], c["mask"], c["sigma"], c["lmax"], c["fwhm"], c["cl0"], seed=c["seed"])
    chain = gs.generate_chain(samples, burn_in=10)
    assert len(chain) == samples
    assert gibbs_acceptance(c, chain) == []
    # the Blackwell-Rao likelihood is lower at the realisation's spectrum than
    # at a spectrum twice as large
    orig = np.maximum(c["cl_original"], 1e-30)
    assert cm.GibbsSampler.calculate_likelihood(orig, chain, c["lmax"]) < \
        cm.GibbsSampler.calculate_likelihood(2 * orig, chain, c["lmax"])
