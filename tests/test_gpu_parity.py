"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar: -2lnL within 1e-10 relative, maps / gradients within 1e-10 norm-wise
relative (||d|| / ||ref||), gradient vs central finite differences within
1e-6 relative.  Sizes are chosen so the oracle finishes in seconds.
"""
import math

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle
from cases import CONFIG1, CONFIG2, CONFIG4, Case

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


# (nside, lmax): regular, lmax = 4 nside (every ring aliased), lmax < nside, tiny maps
GEOMS = [(1, 2), (1, 4), (2, 8), (4, 5), (8, 32), (16, 32), (32, 20), (64, 128), (128, 300)]


@pytest.mark.parametrize("nside,lmax", GEOMS)
def test_forward_matches_oracle(nside, lmax):
    print("case forward nside=%d lmax=%d" % (nside, lmax), flush=True)
    rng = np.random.default_rng(nside + 7 * lmax)
    x = rng.standard_normal((lmax + 1) ** 2)
    b = cm.beam(lmax, 3.0)
    mp = cm.Map(nside, lmax)
    mp.set_beam(b)
    s = mp.forward(x)
    ref = oracle.synthesis(nside, lmax, x, b)
    assert rel(s, ref) < TOL
    mp.close()


@pytest.mark.parametrize("nside,lmax", GEOMS)
def test_backward_matches_oracle(nside, lmax):
    print("case backward nside=%d lmax=%d" % (nside, lmax), flush=True)
    rng = np.random.default_rng(3 * nside + lmax)
    v = rng.standard_normal(12 * nside * nside)
    b = cm.beam(lmax, 3.0)
    mp = cm.Map(nside, lmax)
    mp.set_beam(b)
    g = mp.backward(v)
    ref = oracle.adjoint(nside, lmax, v, b)
    assert rel(g, ref) < TOL
    # adjoint identity on the device pair
    x = rng.standard_normal((lmax + 1) ** 2)
    lhs = np.dot(mp.forward(x), v)
    rhs = np.dot(x, g)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(v) * math.sqrt(x.size)
    mp.close()


@pytest.mark.parametrize("nside,lmax", [(2048, 96), (1024, 64), (256, 1024), (512, 256)])
def test_large_geometries_match_oracle(nside, lmax):
    """Polar rings whose Bluestein length would exceed 4096 (nph = 4i > 2048,
    nside >= 1024) take the split transform (two Bluestein halves of nph/2, one
    held in global scratch; at nside 2048 in a launch of their own), the others
    the unsplit paths in one launch; nside 256 / lmax 1024 aliases every ring
    in the 256-thread small-ring kernels; at nside 512 the equatorial rings
    (2048 pixels, half the launch's 4096) run two to a workgroup.  Forward,
    backward and -2lnL + gradient against the oracle."""
    print("case large geometry nside=%d lmax=%d" % (nside, lmax), flush=True)
    rng = np.random.default_rng(nside + lmax)
    n = (lmax + 1) ** 2
    # B_lmax = 0.1: every l carries weight in the comparison
    b = cm.beam(lmax, math.degrees(math.sqrt(8.0 * math.log(2.0)) * math.sqrt(2.0 * math.log(10.0) / (lmax * (lmax + 1.0)))))
    assert 0.09 < b[lmax] < 0.11
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


@pytest.mark.parametrize("nside,lmax,cut", [(16, 32, 20.0), (64, 128, 20.0), (32, 128, 10.0), (128, 256, 30.0)])
def test_value_and_grad_matches_oracle(nside, lmax, cut):
    print("case like nside=%d lmax=%d" % (nside, lmax), flush=True)
    c = Case(nside, lmax, 2.0, 10.0, cut, seed=nside + lmax)
    mp = c.make_map()
    f, g = mp.value_and_grad(c.x)
    fo, go = oracle.neg2lnL(nside, lmax, c.x, c.beam, c.cl, c.d, c.w)
    assert abs(f - fo) / abs(fo) < TOL
    assert rel(g, go) < TOL
    # deterministic: bitwise identical on repeat
    f2, g2 = mp.value_and_grad(c.x)
    assert f2 == f and np.array_equal(g2, g)
    mp.close()


def test_config1_parity():
    """Plain method, CPU-runnable configuration: nside 64, lmax 128."""
    print("case config1", flush=True)
    c = Case(**CONFIG1, seed=1)
    mp = c.make_map()
    f, g = mp.value_and_grad(c.x)
    fo, go = oracle.neg2lnL(c.nside, c.lmax, c.x, c.beam, c.cl, c.d, c.w)
    assert abs(f - fo) / abs(fo) < TOL
    assert rel(g, go) < TOL
    mp.close()


@pytest.fixture(scope="module")
def config2():
    print("case config2 setup", flush=True)
    c = Case(**CONFIG2, seed=2)
    mp = c.make_map()
    yield c, mp
    mp.close()


def test_config2_parity(config2):
    """nside 512, lmax 1024, 200 point-source holes."""
    c, mp = config2
    print("case config2 parity", flush=True)
    f, g = mp.value_and_grad(c.x)
    fo, go = oracle.neg2lnL(c.nside, c.lmax, c.x, c.beam, c.cl, c.d, c.w)
    assert abs(f - fo) / abs(fo) < TOL
    assert rel(g, go) < TOL
    s = mp.forward(c.x)
    assert rel(s, oracle.synthesis(c.nside, c.lmax, c.x, c.beam)) < TOL


def test_config2_finite_differences(config2):
    """Gradient vs central differences (step 1e-6 relative: h = 1e-6 ||x|| / ||v||)
    on 1000 seeded random directions v ~ N(0, I).

    -2lnL is exactly quadratic, so the difference quotient equals <g, v> up to
    the rounding of f (|f| ~ 3e6, one ulp 4.7e-10, so ~2e-3 absolute in the
    quotient).  Two error measures per direction:
      rel      |fd - <g,v>| / |<g,v>|
      floored  |fd - <g,v>| / max(|<g,v>|, ||g|| ||v|| / sqrt(n)),
    the floor being the typical size of a random projection: for the few
    directions nearly orthogonal to g, |<g,v>| is a thousandth of that and
    the rounding of f alone exceeds 1e-6 of it.  Asserted: floored <= 1e-6 on
    every direction, rel <= 1e-6 on the median and the 90th percentile.
    """
    c, mp = config2
    ndir = 1000
    print("case config2 fd ndir=%d" % ndir, flush=True)
    f, g = mp.value_and_grad(c.x)
    rng = np.random.default_rng(77)
    rel_e, floored = [], []
    gn = np.linalg.norm(g)
    for i in range(ndir):
        v = rng.standard_normal(c.x.size)
        h = 1e-6 * np.linalg.norm(c.x) / np.linalg.norm(v)
        fp = mp.value_and_grad(c.x + h * v, grad=False)
        fm = mp.value_and_grad(c.x - h * v, grad=False)
        fd = (fp - fm) / (2 * h)
        gv = np.dot(g, v)
        rel_e.append(abs(fd - gv) / abs(gv))
        floored.append(abs(fd - gv) / max(abs(gv), gn * np.linalg.norm(v) / math.sqrt(v.size)))
        if (i + 1) % 250 == 0:
            print("config2 fd %d directions" % (i + 1), flush=True)
    rel_e, floored = np.array(rel_e), np.array(floored)
    print("config2 fd (1000 directions): rel |fd-<g,v>|/|<g,v>| median %.3e p90 %.3e p99 %.3e max %.3e; "
          "floored max %.3e median %.3e"
          % (np.median(rel_e), np.percentile(rel_e, 90), np.percentile(rel_e, 99), rel_e.max(),
             floored.max(), np.median(floored)), flush=True)
    assert floored.max() < 1e-6
    assert np.median(rel_e) < 1e-6 and np.percentile(rel_e, 90) < 1e-6


@pytest.fixture(scope="module")
def config4():
    print("case config4 setup", flush=True)
    c = Case(**CONFIG4, seed=4)
    mp = c.make_map()
    yield c, mp
    mp.close()


def test_config4_parity(config4):
    """nside 1024, lmax 2048, FWHM 10 arcmin, sigma 20 uK, 20 degree cut:
    -2lnL and gradient against the oracle at the headline workload's size."""
    c, mp = config4
    print("case config4 value_and_grad", flush=True)
    f, g = mp.value_and_grad(c.x)
    fo, go = oracle.neg2lnL(c.nside, c.lmax, c.x, c.beam, c.cl, c.d, c.w)
    print("config4: -2lnL rel %.3e, gradient rel %.3e" % (abs(f - fo) / abs(fo), rel(g, go)), flush=True)
    assert abs(f - fo) / abs(fo) < TOL
    assert rel(g, go) < TOL


def test_config4_forward_backward(config4):
    """The two halves of the pair at config 4: forward map and J^T v."""
    c, mp = config4
    print("case config4 forward/backward", flush=True)
    s = mp.forward(c.x)
    so = oracle.synthesis(c.nside, c.lmax, c.x, c.beam)
    assert rel(s, so) < TOL
    v = c.w * (c.d - so)
    gb = mp.backward(v)
    go = oracle.adjoint(c.nside, c.lmax, v, c.beam)
    print("config4: forward rel %.3e, backward rel %.3e" % (rel(s, so), rel(gb, go)), flush=True)
    assert rel(gb, go) < TOL


def test_nonfinite_input_is_reported():
    c = Case(8, 16, 2.0, 10.0, 20.0, seed=3)
    mp = c.make_map()
    x = c.x.copy()
    x[5] = np.nan
    with pytest.raises(cm.CmblError) as e:
        mp.value_and_grad(x)
    assert e.value.code == cm.CMBL_ERR_NONFINITE
    # the context stays usable
    f, _ = mp.value_and_grad(c.x)
    assert np.isfinite(f)
    mp.close()


def test_state_errors():
    mp = cm.Map(4, 8)
    with pytest.raises(cm.CmblError) as e:
        mp.value_and_grad(np.zeros(81))
    assert e.value.code == cm.CMBL_ERR_STATE
    mp.close()
