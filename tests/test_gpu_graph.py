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
