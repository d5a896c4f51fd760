"""GPU: batched evaluation of uneven maps and the torch autograd wrapper."""
import numpy as np
import pytest

import cosmopp_amd as cm
import oracle
from cases import Case

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def test_batch_uneven_matches_single_maps():
    import torch

    print("case batch uneven", flush=True)
    geoms = [(16, 32), (64, 128), (32, 64), (8, 20), (64, 100), (128, 256)]
    cases = [Case(n, l, 5400.0 / l / 60.0, 10.0 + 5 * i, 10.0 + 3 * i, seed=10 + i) for i, (n, l) in enumerate(geoms)]
    maps = [c.make_map() for c in cases]
    dev = torch.device("cuda:0")
    xs = [torch.from_numpy(c.x).to(dev) for c in cases]
    gs = [torch.empty_like(x) for x in xs]
    f = torch.empty(len(cases), dtype=torch.float64, device=dev)
    b = cm.Batch(maps)
    stream = torch.cuda.current_stream().cuda_stream
    b.value_and_grad_device([x.data_ptr() for x in xs], f.data_ptr(), [g.data_ptr() for g in gs], stream)
    torch.cuda.synchronize()
    fb = f.cpu().numpy()
    for i, (c, mp) in enumerate(zip(cases, maps)):
        fs, gsn = mp.value_and_grad(c.x)
        assert abs(fb[i] - fs) <= 1e-13 * abs(fs)
        assert rel(gs[i].cpu().numpy(), gsn) <= 1e-13
        fo, go = oracle.neg2lnL(c.nside, c.lmax, c.x, c.beam, c.cl, c.d, c.w)
        assert abs(fb[i] - fo) / abs(fo) < 1e-10
        assert rel(gs[i].cpu().numpy(), go) < 1e-10
    b.close()
    for mp in maps:
        mp.close()


def test_batch_several_launch_groups():
    """70 maps: the a_lm preparation, gradient assembly and reduction run in
    launch groups of up to 64 maps; every map equals its own evaluation."""
    import torch

    print("case batch 70 maps", flush=True)
    geoms = [(2, 4), (4, 8), (4, 12), (8, 16), (8, 24)]
    cases = [Case(*geoms[i % 5], 10.0, 10.0 + (i % 7), 15.0 + (i % 5), seed=200 + i) for i in range(70)]
    maps = [c.make_map() for c in cases]
    dev = torch.device("cuda:0")
    xs = [torch.from_numpy(c.x).to(dev) for c in cases]
    gs = [torch.empty_like(x) for x in xs]
    f = torch.empty(len(cases), dtype=torch.float64, device=dev)
    b = cm.Batch(maps)
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(2):
        b.value_and_grad_device([x.data_ptr() for x in xs], f.data_ptr(), [g.data_ptr() for g in gs], stream)
        torch.cuda.synchronize()
        fb = f.cpu().numpy()
        for i, (c, mp) in enumerate(zip(cases, maps)):
            fs, gsn = mp.value_and_grad(c.x)
            assert abs(fb[i] - fs) <= 1e-13 * abs(fs)
            assert rel(gs[i].cpu().numpy(), gsn) <= 1e-13
    b.close()
    for mp in maps:
        mp.close()


def test_batch_matrix_core_adjoint():
    """Eleven maps of one geometry with one ring pair per lane in the adjoint:
    a full group of eight and a partial group of three run the adjoint on the
    matrix cores (nside 128: four 64-pair blocks per workgroup, polar blocks
    starting at different l).  Each map equals its own evaluation, the batch
    with the matrix-core adjoint switched off (CMBL_BATCH_MMA=0) and the
    oracle."""
    import os

    import torch

    print("case batch matrix-core adjoint nside 128 lmax 256", flush=True)
    cases = [Case(128, 256, 20.0 / 60.0, 20.0 + i, 15.0 + i, seed=400 + i) for i in range(11)]
    maps = []
    for c in cases:
        mp = cm.Map(c.nside, c.lmax, tuning=(1, 1, 4096))
        mp.set_beam(c.beam)
        mp.set_cl(c.cl)
        mp.set_data(c.d, c.w)
        maps.append(mp)
    dev = torch.device("cuda:0")
    xs = [torch.from_numpy(c.x).to(dev) for c in cases]
    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for mma in ("1", "0"):
        os.environ["CMBL_BATCH_MMA"] = mma
        try:
            b = cm.Batch(maps)
        finally:
            del os.environ["CMBL_BATCH_MMA"]
        gs = [torch.empty_like(x) for x in xs]
        f = torch.empty(len(cases), dtype=torch.float64, device=dev)
        for _ in range(2):
            b.value_and_grad_device([x.data_ptr() for x in xs], f.data_ptr(), [g.data_ptr() for g in gs], stream)
        torch.cuda.synchronize()
        out[mma] = (f.cpu().numpy(), [g.cpu().numpy() for g in gs])
        b.close()
    for i, (c, mp) in enumerate(zip(cases, maps)):
        fs, gsn = mp.value_and_grad(c.x)
        for mma in ("1", "0"):
            fb, gb = out[mma]
            assert abs(fb[i] - fs) <= 1e-13 * abs(fs)
            assert rel(gb[i], gsn) <= 1e-13
        fo, go = oracle.neg2lnL(c.nside, c.lmax, c.x, c.beam, c.cl, c.d, c.w)
        assert abs(out["1"][0][i] - fo) / abs(fo) < 1e-10
        assert rel(out["1"][1][i], go) < 1e-10
    for mp in maps:
        mp.close()


def test_torch_autograd_pair():
    import torch

    from cosmopp_amd.torch_fn import synthesize

    print("case torch autograd", flush=True)
    nside, lmax = 32, 64
    mp = cm.Map(nside, lmax)
    b = cm.beam(lmax, 3.0)
    mp.set_beam(b)
    rng = np.random.default_rng(4)
    xh = rng.standard_normal((lmax + 1) ** 2)
    vh = rng.standard_normal(12 * nside * nside)
    dev = torch.device("cuda:0")
    x = torch.tensor(xh, device=dev, requires_grad=True)
    v = torch.tensor(vh, device=dev)
    s = synthesize(x, mp)
    assert rel(s.detach().cpu().numpy(), oracle.synthesis(nside, lmax, xh, b)) < 1e-10
    loss = (s * v).sum()
    loss.backward()
    assert rel(x.grad.cpu().numpy(), oracle.adjoint(nside, lmax, vh, b)) < 1e-10
    # likelihood through autograd: -2lnL data term = sum w (d - s)^2
    c = Case(nside, lmax, 3.0, 10.0, 20.0, seed=21)
    mp2 = c.make_map()
    x2 = torch.tensor(c.x, device=dev, requires_grad=True)
    d = torch.tensor(c.d, device=dev)
    w = torch.tensor(c.w, device=dev)
    s2 = synthesize(x2, mp2)
    chi2 = (w * (d - s2) ** 2).sum()
    chi2.backward()
    f, g = mp2.value_and_grad(c.x)
    cl = torch.tensor(c.cl)
    prior = 0.0
    gprior = np.zeros_like(c.x)
    for l in range(2, lmax + 1):
        blk = c.x[l * l:(l + 1) ** 2]
        prior += (blk[0] ** 2 + 2 * np.sum(blk[1:] ** 2)) / c.cl[l]
        gprior[l * l] = 2 * blk[0] / c.cl[l]
        gprior[l * l + 1:(l + 1) ** 2] = 4 * blk[1:] / c.cl[l]
    assert abs(chi2.item() + prior - f) <= 1e-10 * abs(f)
    assert rel(x2.grad.cpu().numpy() + gprior, g) < 1e-10
    mp.close()
    mp2.close()
