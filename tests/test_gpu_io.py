"""GPU: NEST <-> RING reordering of whole maps (cmbl_reorder), against the
host index conversions (pinned on the CPU in tests/test_io_cpu.py)."""
import os

import numpy as np
import pytest

import cosmopp_amd as cm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MASK = os.path.join(ROOT, "tests", "golden", "mask1.fits")


def test_reorder_reference_mask():
    print("case reorder mask1", flush=True)
    m, nside, nested = cm.read_map(MASK)
    assert nested
    perm = np.array([cm.ring2nest(nside, r) for r in range(m.size)])
    ring = cm.reorder(m, nside, to_ring=True)
    assert np.array_equal(ring, m[perm])
    assert np.array_equal(cm.reorder(ring, nside, to_ring=False), m)


@pytest.mark.parametrize("nside", [1, 2, 64, 1024])
def test_reorder_round_trip_and_samples(nside):
    print("case reorder nside=%d" % nside, flush=True)
    rng = np.random.default_rng(nside)
    npix = 12 * nside * nside
    m = rng.standard_normal(npix)
    ring = cm.reorder(m, nside, to_ring=True)
    assert np.array_equal(cm.reorder(ring, nside, to_ring=False), m)
    for r in rng.integers(0, npix, 500):
        assert ring[r] == m[cm.ring2nest(nside, int(r))]
