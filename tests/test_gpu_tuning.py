"""GPU: the launch-shape autotuner and the tuning table.

- cmbl_tune returns a valid choice, registers it, and a map created
  afterwards for that geometry uses it;
- explicitly tuned maps (every ring-pairs-per-lane choice, every split
  length where splits occur) match the oracle to 1e-10;
- the table survives a save / load round trip.
"""
import math
import os
import tempfile

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def test_tune_registers_choice():
    print("case tune nside=32 lmax=64", flush=True)
    t, times = cm.tune(32, 64, reps=3)
    assert t.k_synth in (1, 2, 4, 8) and t.k_adj in (1, 2, 4, 8) and t.split_m == 4096
    assert all(v > 0 for v in times[:8]) and times[8] < 0 and times[10] < 0
    mp = cm.Map(32, 64)
    assert mp.tuning == t.as_tuple()
    mp.close()


@pytest.mark.parametrize("tuning", [(1, 8, 4096), (8, 1, 4096), (2, 4, 2048), (4, 2, 8192)])
def test_tuned_maps_match_oracle(tuning):
    nside, lmax = 1024, 48
    print("case tuned map %s nside=%d lmax=%d" % (tuning, nside, lmax), flush=True)
    rng = np.random.default_rng(sum(tuning))
    b = cm.beam(lmax, 2.0)
    mp = cm.Map(nside, lmax, tuning=tuning)
    assert mp.tuning == tuning
    mp.set_beam(b)
    x = rng.standard_normal((lmax + 1) ** 2)
    assert rel(mp.forward(x), oracle.synthesis(nside, lmax, x, b)) < 1e-10
    v = rng.standard_normal(12 * nside * nside)
    assert rel(mp.backward(v), oracle.adjoint(nside, lmax, v, b)) < 1e-10
    mp.close()


def test_tuning_table_round_trip():
    t = cm.Tuning(2, 8, 2048)
    cm._check(cm.lib().cmbl_tuning_set(16, 24, cm.ctypes.byref(t)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tuning.txt")
        cm.tuning_save(path)
        text = open(path).read()
        assert " 16 24 2 8 2048" in text
        cm._check(cm.lib().cmbl_tuning_set(16, 24, cm.ctypes.byref(cm.Tuning(1, 1, 4096))))
        cm.tuning_load(path)
    mp = cm.Map(16, 24)
    assert mp.tuning == (2, 8, 2048)
    mp.close()
    with pytest.raises(cm.CmblError):
        cm.Map(16, 24, tuning=(3, 1, 4096))
