"""CPU-side checks of the C ABI library (no GPU compute is launched here).

- the library loads and exports every entry point include/cmbl.h declares;
- argument errors come back as status codes with a message;
- the host input routines match the reference: simulateAlm / simulateWhiteNoise
  draw order on the reference's own RNG draws (tests/golden/rng.npz, from
  include/random.hpp), beamFunction closed form, masks and pix2ang vs the oracle.
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "cmbl.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cmbl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = cm.lib()
    declared = header_symbols()
    assert len(declared) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", cm.LIB_PATH], stdout=subprocess.PIPE, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        getattr(lib, s)
    assert sorted(cm.declared_symbols()) == declared


_RUNTIMES = r"""
import sys
sys.path.insert(0, %r)
import cosmopp_amd as cm
cm.lib()
try:
    import torch
except ImportError:
    pass
names = set()
for line in open("/proc/self/maps"):
    path = line.split()[-1]
    for key in ("libamdhip64", "libhsa-runtime64"):
        if key in path:
            names.add((key, path))
for key in ("libamdhip64", "libhsa-runtime64"):
    print(key, len([p for k, p in names if k == key]))
"""


def test_one_hip_runtime_when_torch_comes_second():
    """Loading the library and then torch maps one HIP and one HSA runtime
    (two HSA runtimes in a process leave the second without a GPU)."""
    out = subprocess.run(["python3", "-c", _RUNTIMES % ROOT], stdout=subprocess.PIPE, text=True,
                         check=True).stdout.split()
    counts = dict(zip(out[0::2], map(int, out[1::2])))
    assert counts == {"libamdhip64": 1, "libhsa-runtime64": 1}, counts


def test_argument_errors_are_status_codes():
    with pytest.raises(cm.CmblError) as e:
        cm.Map(3, 8)  # nside not a power of two
    assert e.value.code == cm.CMBL_ERR_ARG
    with pytest.raises(cm.CmblError) as e:
        cm.Map(4, 17)  # lmax > 4 nside
    assert e.value.code == cm.CMBL_ERR_ARG
    with pytest.raises(cm.CmblError):
        cm.simulate_alm(np.ones(5), 4, 0)  # seed 0 would be time(0) in the reference
    with pytest.raises(cm.CmblError):
        cm.weights_from_mask(np.ones(4), 0.0)


def test_simulate_alm_matches_reference_draws():
    z = np.load(os.path.join(GOLDEN, "rng.npz"))
    seed = int(z["gauss_seed"][0])
    g = z["gauss"]
    lmax = 40
    cl = np.linspace(0.5, 3.0, lmax + 1)
    x = cm.simulate_alm(cl, lmax, seed)
    k = 0
    for l in range(lmax + 1):
        assert x[l * l] == g[k] * math.sqrt(cl[l])
        k += 1
        for m in range(1, l + 1):
            assert x[l * l + 2 * m - 1] == g[k] * math.sqrt(cl[l] / 2)
            assert x[l * l + 2 * m] == g[k + 1] * math.sqrt(cl[l] / 2)
            k += 2


def test_white_noise_matches_reference_draws():
    z = np.load(os.path.join(GOLDEN, "rng.npz"))
    seed = int(z["gauss_sigma_seed"][0])
    n = cm.simulate_white_noise(4, 20.0, seed)  # 192 pixels
    assert np.array_equal(n, z["gauss_sigma"][: n.size])


def test_beam_closed_form():
    b = cm.beam(100, 2.0)
    sigma = math.sqrt(8 * math.log(2.0)) / (2.0 * math.pi / 180)
    for l in (0, 1, 10, 100):
        assert b[l] == math.exp(-l * (l + 1) / (2 * sigma * sigma))
    assert np.all(cm.beam(10, 0.0) == 1.0)
    np.testing.assert_array_equal(b, oracle.beam(100, 2.0))


def test_masks_match_oracle():
    nside = 32
    m1 = cm.mask_latitude(nside, math.radians(20))
    np.testing.assert_array_equal(m1, oracle.mask_latitude(nside, math.radians(20)))
    z = np.load(os.path.join(GOLDEN, "rng.npz"))
    th = z["uni"][:25]
    rng = np.random.default_rng(3)
    ph = rng.uniform(0, 2 * math.pi, 25)
    an = rng.uniform(math.pi / 60, math.pi / 40, 25)
    m2 = cm.mask_regions(nside, th, ph, an, m1)
    np.testing.assert_array_equal(m2, oracle.mask_regions(nside, th, ph, an, m1))
    w = cm.weights_from_mask(m2, 30.0)
    np.testing.assert_array_equal(w, np.where(m2 > 0.5, 1.0 / 900.0, 0.0))


def test_pix2ang_matches_oracle():
    for nside in (1, 4, 64):
        for p in range(0, 12 * nside * nside, max(1, nside)):
            assert cm.pix2ang_ring(nside, p) == oracle.pix2ang_ring(nside, p)
