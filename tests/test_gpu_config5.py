"""GPU parity at the large + batch mix: one nside 2048 / lmax 4096 map
(FWHM 5 arcmin, sigma 15 uK, 20 degree cut; the adjoint recomputes the
Legendre functions per ring block, no stored P_lm tables) and the 16 chain
maps of nside 256 / lmax 512 (FWHM 40 arcmin, sigma 40 uK) evaluated as one
batch, each compared with the oracle.

The data maps are synthesised by the library (the oracle's synthesis at
nside 2048 alone takes ~20 s); the forward map is compared with the oracle
separately, and the oracle evaluates the likelihood on the same data, so no
GPU result is compared with itself.
"""
import numpy as np
import pytest

import cosmopp_amd as cm
import oracle
from cases import CONFIG5, CONFIG5_CHAIN, device_case

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def main_map():
    print("case config5 main map setup", flush=True)
    mp, inp = device_case(CONFIG5, 501)
    yield mp, inp
    mp.close()


def test_config5_main_forward(main_map):
    mp, inp = main_map
    print("case config5 main forward", flush=True)
    s = mp.forward(inp["x"])
    so = oracle.synthesis(CONFIG5["nside"], CONFIG5["lmax"], inp["x"], inp["beam"])
    print("config5 main: forward rel %.3e" % rel(s, so), flush=True)
    assert rel(s, so) < TOL


def test_config5_main_value_and_grad(main_map):
    mp, inp = main_map
    print("case config5 main value_and_grad", flush=True)
    f, g = mp.value_and_grad(inp["x"])
    fo, go = oracle.neg2lnL(CONFIG5["nside"], CONFIG5["lmax"], inp["x"], inp["beam"], inp["cl"], inp["d"], inp["w"])
    print("config5 main: -2lnL rel %.3e, gradient rel %.3e" % (abs(f - fo) / abs(fo), rel(g, go)), flush=True)
    assert abs(f - fo) / abs(fo) < TOL
    assert rel(g, go) < TOL


def test_config5_chain_batch():
    """16 chain states of independent seeds in one batched evaluation, each
    map against the oracle."""
    import torch

    print("case config5 chain batch", flush=True)
    dev = torch.device("cuda:0")
    maps, inps = [], []
    for i in range(16):
        mp, inp = device_case(CONFIG5_CHAIN, 2000 + 10 * i)
        maps.append(mp)
        inps.append(inp)
    xs = [torch.from_numpy(inp["x"]).to(dev) for inp in inps]
    gs = [torch.empty_like(x) for x in xs]
    f = torch.empty(16, dtype=torch.float64, device=dev)
    batch = cm.Batch(maps)
    sp = torch.cuda.current_stream().cuda_stream
    batch.value_and_grad_device([x.data_ptr() for x in xs], f.data_ptr(), [g.data_ptr() for g in gs], sp)
    torch.cuda.synchronize()
    fb = f.cpu().numpy()
    worst_f = worst_g = 0.0
    for i, inp in enumerate(inps):
        fo, go = oracle.neg2lnL(CONFIG5_CHAIN["nside"], CONFIG5_CHAIN["lmax"], inp["x"], inp["beam"], inp["cl"],
                                inp["d"], inp["w"])
        worst_f = max(worst_f, abs(fb[i] - fo) / abs(fo))
        worst_g = max(worst_g, rel(gs[i].cpu().numpy(), go))
    print("config5 chains: worst -2lnL rel %.3e, worst gradient rel %.3e" % (worst_f, worst_g), flush=True)
    assert worst_f < TOL and worst_g < TOL
    batch.close()
    for mp in maps:
        mp.close()
