"""GPU: cmbl_vec_dot_device_static, the dot product that allocates nothing
(scratch in a static device array), against cmbl_vec_dot_device through the
C ABI: two calls in a row, the second on a smaller input, each equal bit for
bit to the allocating entry point (same kernels and summation order)."""
import numpy as np
import pytest

import cosmopp_amd as cm

pytestmark = pytest.mark.gpu


def test_static_scratch_dot_matches_allocating_dot():
    import torch

    print("case static-scratch dot", flush=True)
    lib = cm.lib()
    sp = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(2024)
    out = torch.zeros(2, dtype=torch.float64, device="cuda:0")
    for n in (4198401, 16641):
        x = torch.from_numpy(rng.standard_normal(n)).to("cuda:0")
        y = torch.from_numpy(rng.standard_normal(n)).to("cuda:0")
        assert lib.cmbl_vec_dot_device_static(x.data_ptr(), y.data_ptr(), n, out.data_ptr(), sp) == 0
        assert lib.cmbl_vec_dot_device(x.data_ptr(), y.data_ptr(), n, out.data_ptr() + 8, sp) == 0
        torch.cuda.synchronize()
        r = out.cpu().numpy()
        ref = float(np.dot(x.cpu().numpy(), y.cpu().numpy()))
        print("n=%d static %.17g allocating %.17g numpy %.17g" % (n, r[0], r[1], ref), flush=True)
        assert r[0] == r[1]
        assert abs(r[0] - ref) <= 1e-12 * float(torch.linalg.norm(x)) * float(torch.linalg.norm(y))
    # invalid arguments are rejected before any launch
    assert lib.cmbl_vec_dot_device_static(None, None, 10, out.data_ptr(), sp) == cm.CMBL_ERR_ARG
    assert lib.cmbl_vec_dot_device_static(out.data_ptr(), out.data_ptr(), -1, out.data_ptr(), sp) == cm.CMBL_ERR_ARG
