"""GPU: one evaluation (cmbl_value_and_grad_device) captured in a HIP graph and
replayed.  The device path launches kernels only (no allocation, no memset,
no host synchronisation), so a caller may capture it; every replay, with new
inputs and eager work in between, must equal the eager evaluation bit for bit.
"""
import numpy as np
import pytest

from cases import Case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nside,lmax", [(16, 40), (64, 128)])
def test_graph_replay_matches_eager(nside, lmax):
    import torch

    print("case graph nside=%d lmax=%d" % (nside, lmax), flush=True)
    c = Case(nside, lmax, 2.0, 10.0, 20.0, seed=nside + lmax)
    mp = c.make_map()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    inputs = [c.x + 0.05 * rng.standard_normal(c.x.size) for _ in range(4)]
    x = torch.from_numpy(inputs[0]).to(dev)
    g = torch.empty_like(x)
    f = torch.empty(1, dtype=torch.float64, device=dev)
    side = torch.cuda.Stream()
    # eager warm-up on the capture stream (one-time set-up calls happen here)
    with torch.cuda.stream(side):
        mp.value_and_grad_device(x.data_ptr(), f.data_ptr(), g.data_ptr(), side.cuda_stream)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        mp.value_and_grad_device(x.data_ptr(), f.data_ptr(), g.data_ptr(), side.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        for xi in inputs:
            x.copy_(torch.from_numpy(xi))
            graph.replay()
            torch.cuda.synchronize()
            fg, gg = f.item(), g.cpu().numpy()
            fe, ge = mp.value_and_grad(xi)  # eager, host buffers, between replays
            assert fg == fe
            assert np.array_equal(gg, ge)
    del graph
    mp.close()


def test_batch_graph_replay_matches_eager():
    """A batch of uneven maps (two of one geometry, so the shared-recurrence
    synthesis runs too, and three of another with one ring pair per lane in
    the adjoint, which run it on the matrix cores) captured in one HIP graph:
    every replay, with new inputs and eager work in between, equals the eager
    batch bit for bit."""
    import torch

    import cosmopp_amd as cm

    print("case batch graph", flush=True)
    cases = [Case(16, 40, 2.0, 10.0, 20.0, seed=31), Case(32, 64, 2.0, 10.0, 20.0, seed=32),
             Case(32, 64, 2.0, 10.0, 20.0, seed=33)]
    maps = [c.make_map() for c in cases]
    for seed in (34, 35, 36):
        c = Case(64, 96, 2.0, 10.0, 20.0, seed=seed)
        mp = cm.Map(c.nside, c.lmax, tuning=(1, 1, 4096))
        mp.set_beam(c.beam)
        mp.set_cl(c.cl)
        mp.set_data(c.d, c.w)
        cases.append(c)
        maps.append(mp)
    batch = cm.Batch(maps)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(6)
    xs = [torch.from_numpy(c.x).to(dev) for c in cases]
    gs = [torch.empty_like(x) for x in xs]
    f = torch.empty(len(maps), dtype=torch.float64, device=dev)
    xp, gp = [x.data_ptr() for x in xs], [g.data_ptr() for g in gs]
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        batch.value_and_grad_device(xp, f.data_ptr(), gp, side.cuda_stream)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        batch.value_and_grad_device(xp, f.data_ptr(), gp, side.cuda_stream)
    torch.cuda.synchronize()
    for rep in range(3):
        inputs = [c.x + 0.05 * rng.standard_normal(c.x.size) for c in cases]
        for x, xi in zip(xs, inputs):
            x.copy_(torch.from_numpy(xi))
        graph.replay()
        torch.cuda.synchronize()
        fg = f.cpu().numpy().copy()
        gg = [g.cpu().numpy().copy() for g in gs]
        # eager work in between: each map on its own, then the batch eagerly
        for mp, xi in zip(maps, inputs):
            mp.value_and_grad(xi)
        batch.value_and_grad_device(xp, f.data_ptr(), gp, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(fg, f.cpu().numpy())
        for a, g in zip(gg, gs):
            assert np.array_equal(a, g.cpu().numpy())
    del graph
    batch.close()
    for mp in maps:
        mp.close()
