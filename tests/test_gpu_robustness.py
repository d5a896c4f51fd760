"""GPU: error reporting and workspace isolation of the device entry points.

- a non-finite -2lnL from the asynchronous device entry point raises the
  map's device flag: cmbl_map_check_finite reports it, the next device call
  reports it if the check did not, and either clears it;
- the same per map of a batch (cmbl_batch_check_finite names the map);
- a map may not appear twice in one batch;
- dot products queued on two streams at once use separate workspaces;
- in-place BLAS-1 (y = c y, y += c y);
- the likelihood cache is invalidated by set_beam / set_cl / set_data.
"""
import numpy as np
import pytest

import cosmopp_amd as cm
from cases import Case

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def test_device_nonfinite_flag():
    import torch

    print("case device nonfinite flag", flush=True)
    c = Case(8, 16, 2.0, 10.0, 20.0, seed=5)
    mp = c.make_map()
    x = c.x.copy()
    x[7] = np.nan
    dx, gx = _dev(x), _dev(np.zeros_like(x))
    f = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    sp = torch.cuda.current_stream().cuda_stream
    mp.value_and_grad_device(dx.data_ptr(), f.data_ptr(), gx.data_ptr(), sp)
    with pytest.raises(cm.CmblError) as e:
        mp.check_finite(sp)
    assert e.value.code == cm.CMBL_ERR_NONFINITE
    mp.check_finite(sp)  # cleared
    # without a check, the next device call reports it
    mp.value_and_grad_device(dx.data_ptr(), f.data_ptr(), gx.data_ptr(), sp)
    torch.cuda.synchronize()
    good = _dev(c.x)
    with pytest.raises(cm.CmblError) as e:
        mp.value_and_grad_device(good.data_ptr(), f.data_ptr(), gx.data_ptr(), sp)
    assert e.value.code == cm.CMBL_ERR_NONFINITE
    mp.value_and_grad_device(good.data_ptr(), f.data_ptr(), gx.data_ptr(), sp)
    mp.check_finite(sp)
    assert np.isfinite(f.item())
    mp.close()


def test_batch_nonfinite_flag_and_duplicates():
    import torch

    print("case batch nonfinite flag", flush=True)
    cases = [Case(8, 16, 2.0, 10.0, 20.0, seed=10 + i) for i in range(3)]
    maps = [c.make_map() for c in cases]
    with pytest.raises(cm.CmblError) as e:
        cm.Batch([maps[0], maps[1], maps[0]])
    assert e.value.code == cm.CMBL_ERR_ARG
    b = cm.Batch(maps)
    xs = [_dev(c.x) for c in cases]
    xs[1][3] = float("inf")
    gs = [torch.zeros_like(x) for x in xs]
    f = torch.zeros(3, dtype=torch.float64, device="cuda:0")
    sp = torch.cuda.current_stream().cuda_stream
    b.value_and_grad_device([x.data_ptr() for x in xs], f.data_ptr(), [g.data_ptr() for g in gs], sp)
    with pytest.raises(cm.CmblError) as e:
        b.check_finite(sp)
    assert e.value.code == cm.CMBL_ERR_NONFINITE and "map 1" in str(e.value)
    b.check_finite(sp)
    fv = f.cpu().numpy()
    assert np.isfinite(fv[0]) and np.isfinite(fv[2]) and not np.isfinite(fv[1])
    b.close()
    for mp in maps:
        mp.close()


def test_dot_workspaces_per_stream():
    import torch

    print("case dot per stream", flush=True)
    lib = cm.lib()
    n = 1 << 22
    rng = np.random.default_rng(1)
    a, b, c, d = (rng.standard_normal(n) for _ in range(4))
    da, db, dc, dd = (_dev(v) for v in (a, b, c, d))
    out = torch.zeros(2, dtype=torch.float64, device="cuda:0")
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for _ in range(20):
        with torch.cuda.stream(s1):
            assert lib.cmbl_vec_dot_device(da.data_ptr(), db.data_ptr(), n, out.data_ptr(), s1.cuda_stream) == 0
        with torch.cuda.stream(s2):
            assert lib.cmbl_vec_dot_device(dc.data_ptr(), dd.data_ptr(), n, out.data_ptr() + 8,
                                           s2.cuda_stream) == 0
        torch.cuda.synchronize()
        r = out.cpu().numpy()
        assert abs(r[0] - np.dot(a, b)) <= 1e-12 * np.linalg.norm(a) * np.linalg.norm(b)
        assert abs(r[1] - np.dot(c, d)) <= 1e-12 * np.linalg.norm(c) * np.linalg.norm(d)
    assert lib.cmbl_vec_release_workspaces() == 0
    # usable again after the release
    assert lib.cmbl_vec_dot_device(da.data_ptr(), db.data_ptr(), n, out.data_ptr(), s1.cuda_stream) == 0
    torch.cuda.synchronize()
    assert abs(out[0].item() - np.dot(a, b)) <= 1e-12 * np.linalg.norm(a) * np.linalg.norm(b)


def test_vec_in_place():
    import torch

    lib = cm.lib()
    v = np.arange(1.0, 1001.0)
    dv = _dev(v)
    sp = torch.cuda.current_stream().cuda_stream
    assert lib.cmbl_vec_copy(dv.data_ptr(), dv.data_ptr(), -2.0, v.size, sp) == 0
    assert lib.cmbl_vec_axpy(dv.data_ptr(), dv.data_ptr(), 0.5, v.size, sp) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dv.cpu().numpy(), -3.0 * v)


def test_likelihood_cache_follows_inputs():
    print("case likelihood cache generation", flush=True)
    c = Case(8, 16, 2.0, 10.0, 20.0, seed=6)
    mp = c.make_map()
    like = cm.MapLikelihood(mp)
    f0 = like.evaluate(c.x)
    assert like.evaluate(c.x) == f0 and like.evaluations == 1
    mp.set_cl(2.0 * c.cl)
    f1 = like.evaluate(c.x)
    assert like.evaluations == 2 and f1 != f0
    mp.set_data(c.d, 2.0 * c.w)
    assert like.evaluate(c.x) != f1 and like.evaluations == 3
    mp.close()
