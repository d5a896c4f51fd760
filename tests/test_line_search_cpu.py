"""The device-resident L-BFGS's line search (cmbl_line_search, a state
machine on the host) against the reference's own moreThuenteSearch
(include/line_search.hpp:12-272, compiled from the reference header into
oracle/_ref/libref.so) on the one-dimensional test functions of More and
Thuente (ACM TOMS 20, 1994, section 5): the same trial steps, bit for bit,
the same final step and the same result code, for several initial steps and
both the reference's L-BFGS parameters and the paper's.  Skipped where the
reference tree is absent (the shim is built from it)."""
import pytest

import cosmopp_amd as cm
import oracle

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference shim not built")

PARAMS = [dict(ftol=0.1, gtol=0.5, xtol=1e-15, stpmin=1e-15, stpmax=1e15, maxfev=100),
          dict(ftol=1e-3, gtol=0.1, xtol=1e-10, stpmin=1e-20, stpmax=1e20, maxfev=60),
          dict(ftol=1e-3, gtol=1e-3, xtol=1e-12, stpmin=1e-20, stpmax=1e20, maxfev=60)]


def mine(fn, stp0, p):
    ls = cm.LineSearch(**p)
    f0, g0 = oracle.ref_phi(fn, 0.0)
    stp = ls.begin(f0, g0, stp0)
    trials = []
    while True:
        trials.append(stp)
        f, g = oracle.ref_phi(fn, stp)
        info, stp = ls.feed(f, g)
        if info != 0:
            ls.close()
            return info, trials, stp


@needs_ref
@pytest.mark.parametrize("fn", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("stp0", [1e-3, 1e-1, 1.0, 10.0, 1e3])
@pytest.mark.parametrize("pi", [0, 1, 2])
def test_line_search_matches_reference(fn, stp0, pi):
    p = PARAMS[pi]
    ref = oracle.ref_mt_trace(fn, stp0, **p)
    got = mine(fn, stp0, p)
    assert got[0] == ref[0], (got, ref)
    assert got[1] == ref[1]
    assert got[2] == ref[2]


def test_line_search_rejects_bad_input():
    ls = cm.LineSearch()
    with pytest.raises(cm.CmblError):
        ls.begin(1.0, 0.5, 1.0)  # ascent direction
    with pytest.raises(cm.CmblError):
        ls.feed(1.0, -1.0)  # not started
    ls.close()
