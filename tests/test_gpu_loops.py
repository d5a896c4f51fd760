"""GPU: the device-resident L-BFGS and HMC loops (cmbl_lbfgs, cmbl_hmc).

- against the reference: oracle/_ref/ref_consumers (compiled from the
  reference's include/lbfgs_general.hpp, include/line_search.hpp and
  include/hmc_general.hpp) runs LBFGS_General and HMCGeneral on
  cmbl::DeviceLargeVector, then the library's loops from the same start on
  the same map; the per-iteration values, the chains and the final points
  must agree to 1e-10 (the loops are built to match bit for bit, which is
  reported);
- HIP-graph replay against eager launches of the same kernels (bit for bit);
- the L-BFGS minimum against the exact minimiser of the quadratic -2lnL.
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
from cases import Case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "ref_consumers")
needs_bin = pytest.mark.skipif(not os.path.exists(BIN), reason="reference consumers not built (needs the reference tree)")


def run_ref(mode, nside, lmax, iters):
    print("case loops %s nside=%d lmax=%d iters=%d" % (mode, nside, lmax, iters), flush=True)
    with tempfile.TemporaryDirectory() as cwd:
        r = subprocess.run([BIN, mode, str(nside), str(lmax), str(iters)], stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=240, cwd=cwd)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@needs_bin
@pytest.mark.parametrize("nside,lmax,iters", [(8, 16, 60), (16, 40, 40)])
def test_lbfgs_matches_reference_lbfgs_general(nside, lmax, iters):
    out = run_ref("lbfgs_compare", nside, lmax, iters)
    fr, fd = np.array(out["trace_f_ref"]), np.array(out["trace_f_dev"])
    gr, gd = np.array(out["trace_gn_ref"]), np.array(out["trace_gn_dev"])
    xr, xd = np.array(out["x_ref"]), np.array(out["x_dev"])
    exact = fr.size == fd.size and np.array_equal(fr, fd) and np.array_equal(gr, gd) and np.array_equal(xr, xd)
    print("lbfgs_compare nside %d: %d iterations, evaluations ref %d dev %d, bit-exact %s, x rel %.3e"
          % (nside, fr.size - 1, out["evals_ref"], out["evals_dev"], exact, rel(xd, xr)), flush=True)
    assert fr.size == fd.size
    assert np.max(np.abs(fd - fr) / np.abs(fr)) <= 1e-10
    assert np.max(np.abs(gd - gr) / np.maximum(np.abs(gr), 1e-300)) <= 1e-10
    assert rel(xd, xr) <= 1e-10
    assert out["f_dev"] == pytest.approx(out["f_ref"], rel=1e-10)


@needs_bin
def test_hmc_matches_reference_hmc_general():
    out = run_ref("hmc_compare", 8, 16, 30)
    lr, ld = np.array(out["like_ref"]), np.array(out["like_dev"])
    exact = np.array_equal(lr, ld) and np.array_equal(out["x_ref"], out["x_dev"])
    print("hmc_compare: %d samples, accepted %d, bit-exact %s" % (lr.size, out["accepted_dev"], exact), flush=True)
    assert lr.size == ld.size == 31
    assert np.max(np.abs(ld - lr) / np.abs(lr)) <= 1e-10
    assert rel(out["x_dev"], out["x_ref"]) <= 1e-10


def _problem(nside, lmax, seed):
    import torch

    c = Case(nside, lmax, 5.0, 10.0, 20.0, seed=seed)
    mp = c.make_map()
    x0 = torch.zeros(c.x.size, dtype=torch.float64, device="cuda:0")
    return c, mp, x0


def test_lbfgs_graph_replay_equals_eager():
    import torch

    c, mp, x0 = _problem(16, 32, 21)
    out = []
    for graphs in (True, False):
        print("case lbfgs graphs=%s" % graphs, flush=True)
        lb = cm.LBFGS(mp, x0.data_ptr(), m=10, graphs=graphs)
        res = torch.empty_like(x0)
        trace = []
        f, it, ev = lb.minimize(res.data_ptr(), epsilon=1e-14, gnorm_tol=1e-9, max_iter=25,
                                callback=lambda i, fv, gn, x, g, z: trace.append((fv, gn)))
        out.append((f, it, ev, trace, res.cpu().numpy()))
        lb.close()
    (f1, it1, ev1, t1, x1), (f2, it2, ev2, t2, x2) = out
    assert (f1, it1, ev1) == (f2, it2, ev2) and t1 == t2 and np.array_equal(x1, x2)
    mp.close()


def test_lbfgs_reaches_exact_minimiser():
    """The quadratic's minimiser from the oracle's gradients (column by column)."""
    import torch

    c, mp, x0 = _problem(8, 16, 22)
    n = c.x.size
    g0 = oracle.neg2lnL(8, 16, np.zeros(n), c.beam, c.cl, c.d, c.w)[1]
    H = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        H[:, i] = oracle.neg2lnL(8, 16, e, c.beam, c.cl, c.d, c.w)[1] - g0
    xstar = np.linalg.solve(H, -g0)
    lb = cm.LBFGS(mp, x0.data_ptr(), m=10)
    res = torch.empty_like(x0)
    f, it, ev = lb.minimize(res.data_ptr(), epsilon=1e-15, gnorm_tol=1e-9, max_iter=500)
    print("lbfgs: %d iterations, %d evaluations" % (it, ev), flush=True)
    assert rel(res.cpu().numpy(), xstar) <= 1e-6
    g = torch.empty_like(x0)
    lb.gradient(g.data_ptr())
    assert float(torch.linalg.norm(g)) <= 1e-6 * np.linalg.norm(g0)
    lb.close()
    mp.close()


def test_hmc_fixed_steps_and_graph_replay():
    """Fixed step size and count; graph replay and eager launches give the same chain."""
    import torch

    c, mp, x0 = _problem(8, 16, 23)
    l = np.floor(np.sqrt(np.arange(c.x.size))).astype(int)
    masses = torch.from_numpy(np.where(l >= 2, 2.0 / np.maximum(c.cl[l], 1e-300), 1.0)).to("cuda:0")
    chains = []
    for graphs in (True, False):
        print("case hmc fixed graphs=%s" % graphs, flush=True)
        h = cm.HMC(mp, x0.data_ptr(), masses.data_ptr(), 0.3, 5, 91, fixed=True, graphs=graphs)
        likes = []
        h.run(12, callback=lambda x, like: likes.append(like))
        st = h.state()
        assert st["trajectories"] == 12 and st["evaluations"] == 1 + 12 * 5
        pos = torch.empty_like(x0)
        h.position(pos.data_ptr())
        chains.append((likes, pos.cpu().numpy()))
        h.close()
    assert chains[0][0] == chains[1][0] and np.array_equal(chains[0][1], chains[1][1])
    assert len(chains[0][0]) == 13 and all(math.isfinite(v) for v in chains[0][0])
    mp.close()


def test_loop_set_up_copies_are_ordered():
    """The set-up copies of the start point (and of the masses) complete before
    the context's first kernels, at a size where an unordered device-to-device
    copy is still in flight: the L-BFGS starting value equals a direct
    evaluation at x0, and each HMC trajectory's initial kinetic term
    p . M^-1 p = sum noise^2 lies within 5 sigma of n (noise ~ N(0, 1))."""
    import torch

    print("case loop set-up ordering nside=256 lmax=512", flush=True)
    c = Case(256, 512, 40.0 / 60.0, 40.0, 20.0, seed=31)
    mp = c.make_map()
    n = c.x.size
    f_direct, _ = mp.value_and_grad(c.x)
    x0 = torch.from_numpy(c.x).to("cuda:0")
    lb = cm.LBFGS(mp, x0.data_ptr(), m=10)
    assert lb.state()["f"] == f_direct
    lb.close()
    l = np.floor(np.sqrt(np.arange(n))).astype(int)
    masses = torch.from_numpy(np.where(l >= 2, 2.0 / np.maximum(c.cl[l], 1e-300), 1.0) * 1e3).to("cuda:0")
    h = cm.HMC(mp, x0.data_ptr(), masses.data_ptr(), 1e-3, 2, 57, fixed=True)
    k0 = []
    h.run(3, callback=lambda x, like: k0.append(h.last_trajectory()[1]))
    print("kinetic terms at trajectory starts: %s (n = %d)" % (k0[1:], n), flush=True)
    for v in k0[1:]:
        assert abs(v - n) < 5 * math.sqrt(2 * n)
    h.close()
    mp.close()
