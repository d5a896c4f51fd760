"""GPU: the reference's own optimisers, samplers and CG solver drive the
MI355X likelihood and Gibbs system.

oracle/_ref/ref_consumers is compiled from the reference sources
(source/lbfgs.cpp, source/hmc.cpp, include/lbfgs_general.hpp,
include/hmc_general.hpp, include/conjugate_gradient.hpp, ...) plus
include/cmbl_likelihood.hpp.  The L-BFGS results are checked against the
exact minimiser of the quadratic -2lnL, built column by column from the
oracle's gradients; the reference's ConjugateGradient on cmbl::GibbsSystem
against the device CG and the oracle's.
"""
import json
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

import cosmopp_amd as cm
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "ref_consumers")

needs_bin = pytest.mark.skipif(not os.path.exists(BIN), reason="reference consumers not built (needs the reference tree)")

NSIDE, LMAX = 8, 16


def run(mode, iters):
    print("case consumers %s nside=%d lmax=%d iters=%d" % (mode, NSIDE, LMAX, iters), flush=True)
    # the HMC modes write their chain file into the working directory
    with tempfile.TemporaryDirectory() as cwd:
        r = subprocess.run([BIN, mode, str(NSIDE), str(LMAX), str(iters)], stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=240, cwd=cwd)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    return json.loads(line)


@pytest.fixture(scope="module")
def problem():
    """The inputs ref_consumers builds (same seeds), and the exact minimiser."""
    cl = cm.cl_shape(LMAX)
    b = cm.beam(LMAX, 5.0)
    a = cm.simulate_alm(cl, LMAX, 501)
    mp = cm.Map(NSIDE, LMAX)
    mp.set_beam(b)
    d = mp.forward(a) + cm.simulate_white_noise(NSIDE, 10.0, 502)
    mp.close()
    w = cm.weights_from_mask(cm.mask_latitude(NSIDE, math.radians(20.0)), 10.0)
    n = (LMAX + 1) ** 2
    g0 = oracle.neg2lnL(NSIDE, LMAX, np.zeros(n), b, cl, d, w)[1]
    H = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        H[:, i] = oracle.neg2lnL(NSIDE, LMAX, e, b, cl, d, w)[1] - g0
    # monopole/dipole have no prior and may be poorly constrained by the cut sky: H is still SPD
    xstar = np.linalg.solve(H, -g0)
    fstar = oracle.neg2lnL(NSIDE, LMAX, xstar, b, cl, d, w, grad=False)
    return dict(cl=cl, b=b, d=d, w=w, xstar=xstar, fstar=fstar, n=n)


@needs_bin
def test_reference_lbfgs_on_map_likelihood(problem):
    out = run("lbfgs", 500)
    x = np.array(out["x"])
    assert np.linalg.norm(x - problem["xstar"]) <= 1e-6 * np.linalg.norm(problem["xstar"])
    assert abs(out["f"] - problem["fstar"]) <= 1e-10 * abs(problem["fstar"])


@needs_bin
def test_reference_lbfgs_general_on_device_vectors(problem):
    out = run("lbfgs_device", 500)
    x = np.array(out["x"])
    assert np.linalg.norm(x - problem["xstar"]) <= 1e-6 * np.linalg.norm(problem["xstar"])
    assert abs(out["f"] - problem["fstar"]) <= 1e-10 * abs(problem["fstar"])


@needs_bin
def test_reference_hmc_uses_one_evaluation_per_gradient():
    iters = 20
    out = run("hmc", iters)
    assert out["chain_lines"] == iters + 1
    # HMC::calculateDerivs asks for n components per gradient; the cache makes each
    # gradient one device evaluation: at most 1 + iters * (nMax + 1) evaluations.
    assert out["evals"] <= 1 + iters * (10 + 1)


@needs_bin
def test_reference_hmc_general_on_device_vectors(problem):
    out = run("hmc_device", 40)
    assert out["samples"] == 41
    assert out["evals"] <= 1 + 40 * (10 + 2)
    assert np.linalg.norm(out["mean"]) > 0


@needs_bin
def test_reference_conjugate_gradient_on_gibbs_system(problem):
    out = run("cg", 0)
    mask = cm.mask_latitude(NSIDE, math.radians(20.0))
    w0 = cm.simulate_white_noise(NSIDE, 1.0, 601)
    w1 = cm.simulate_white_noise(NSIDE, 1.0, 602)
    b = oracle.gibbs_rhs(NSIDE, LMAX, problem["d"], w0, w1, problem["cl"], problem["b"], mask, 10.0)
    assert np.linalg.norm(np.array(out["b"]) - b) <= 1e-10 * np.linalg.norm(b)
    A = lambda v: oracle.gibbs_apply(NSIDE, LMAX, v, problem["cl"], problem["b"], mask, 10.0)
    M = lambda v: oracle.gibbs_precondition(LMAX, v, problem["cl"], problem["b"], mask, 10.0)
    prec = out["precision"]
    assert abs(prec - 1e-10 * np.linalg.norm(b)) <= 1e-6 * prec
    xo, ito = oracle.cg_solve(A, M, b, precision=prec)
    for key in ("x_cg", "x_dev"):
        x = np.array(out[key])
        assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
        assert np.linalg.norm(A(x) - b) <= 10 * prec
    assert abs(out["iter_cg"] - ito) <= 2 and abs(out["iter_dev"] - ito) <= 2
