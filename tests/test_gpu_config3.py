"""GPU parity for the batched uneven workload: 64 independent maps with lmax
drawn from {256, 512, 768, 1024, 1536}, nside = 2^round(log2(lmax/2)),
FWHM 5400/lmax arcmin, sigma ~ U[10, 50] uK, latitude cut ~ U[10, 30] deg,
independent seeds, evaluated in one batched call (one launch per stage over
all maps); every map's -2lnL and gradient against the oracle, one test per
lmax class.  Data maps are synthesised by the library and the oracle
evaluates the likelihood on the same data.
"""
import numpy as np
import pytest

import cosmopp_amd as cm
import oracle
from cases import batch_configs, device_case

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def batch64():
    import torch

    print("case config3 batch setup (64 maps)", flush=True)
    cfgs = batch_configs()
    dev = torch.device("cuda:0")
    maps, inps = [], []
    for c in cfgs:
        mp, inp = device_case(c, c["seed"])
        maps.append(mp)
        inps.append(inp)
    xs = [torch.from_numpy(inp["x"]).to(dev) for inp in inps]
    gs = [torch.empty_like(x) for x in xs]
    f = torch.empty(len(maps), dtype=torch.float64, device=dev)
    batch = cm.Batch(maps)
    sp = torch.cuda.current_stream().cuda_stream
    batch.value_and_grad_device([x.data_ptr() for x in xs], f.data_ptr(), [g.data_ptr() for g in gs], sp)
    torch.cuda.synchronize()
    out = dict(cfgs=cfgs, inps=inps, f=f.cpu().numpy(), g=[g.cpu().numpy() for g in gs])
    batch.close()
    for mp in maps:
        mp.close()
    yield out


@pytest.mark.parametrize("lmax", [256, 512, 768, 1024, 1536])
def test_config3_batch_class(batch64, lmax):
    print("case config3 lmax class %d" % lmax, flush=True)
    idx = [i for i, c in enumerate(batch64["cfgs"]) if c["lmax"] == lmax]
    assert idx, "the 64-map draw has no map of lmax %d" % lmax
    worst_f = worst_g = 0.0
    for i in idx:
        c, inp = batch64["cfgs"][i], batch64["inps"][i]
        fo, go = oracle.neg2lnL(c["nside"], c["lmax"], inp["x"], inp["beam"], inp["cl"], inp["d"], inp["w"])
        worst_f = max(worst_f, abs(batch64["f"][i] - fo) / abs(fo))
        worst_g = max(worst_g, rel(batch64["g"][i], go))
    print("config3 lmax %d (%d maps, nside %d): worst -2lnL rel %.3e, worst gradient rel %.3e"
          % (lmax, len(idx), batch64["cfgs"][idx[0]]["nside"], worst_f, worst_g), flush=True)
    assert worst_f < TOL and worst_g < TOL
