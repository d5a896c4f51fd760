"""GPU: spin-2 synthesis cmbl_alm2map_pol (HEALPix alm2map_pol as called at
source/likelihood.cpp:585 and source/c_matrix_generator.cpp:577-578) through
the C ABI, against the oracle (pinned on the CPU in tests/test_oracle_pol.py).

Bar: T, Q, U within 1e-10 norm-wise relative of the oracle; at sizes the
oracle does not reach, the size-independent properties: T equals the scalar
synthesis bit for bit, (Q, U)[B = a] = (-U, Q)[E = a] bit for bit,
linearity, l < 2 of E and B ignored.
"""
import numpy as np
import pytest

import cosmopp_amd as cm
import oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def spectrum_scaled(rng, L, k=3):
    """Random packed a_lm with a falling spectrum (|a_lm| ~ 1/(l+1))."""
    l = oracle.packed_l(L)
    return rng.standard_normal((k, (L + 1) ** 2)) / (l + 1.0)


@pytest.mark.parametrize("nside,L", [(2, 5), (4, 8), (8, 16), (16, 40), (32, 64), (64, 128), (128, 256)])
def test_pol_matches_oracle(nside, L):
    print("case pol nside=%d L=%d" % (nside, L), flush=True)
    rng = np.random.default_rng(7 * nside + L)
    t, e, b = spectrum_scaled(rng, L)
    mp = cm.Map(nside, L)
    gt, gq, gu = mp.alm2map_pol(t, e, b)
    ot, oq, ou = oracle.alm2map_pol(nside, L, t, e, b)
    assert rel(gt, ot) < 1e-10
    assert rel(gq, oq) < 1e-10
    assert rel(gu, ou) < 1e-10
    # a second call on the same workspace with other inputs
    t2, e2, b2 = spectrum_scaled(rng, L)
    gt, gq, gu = mp.alm2map_pol(t2, e2, b2)
    ot, oq, ou = oracle.alm2map_pol(nside, L, t2, e2, b2)
    assert rel(gq, oq) < 1e-10 and rel(gu, ou) < 1e-10 and rel(gt, ot) < 1e-10
    mp.close()


@pytest.mark.parametrize("nside,L", [(2048, 48), (1024, 64)])
def test_pol_split_rings(nside, L):
    """Polar rings whose Bluestein length would exceed 4096 take the split ring
    transforms (at nside 2048 also in a launch of their own)."""
    print("case pol split nside=%d L=%d" % (nside, L), flush=True)
    rng = np.random.default_rng(nside + L)
    t, e, b = spectrum_scaled(rng, L)
    mp = cm.Map(nside, L)
    gt, gq, gu = mp.alm2map_pol(t, e, b)
    ot, oq, ou = oracle.alm2map_pol(nside, L, t, e, b)
    assert rel(gq, oq) < 1e-10 and rel(gu, ou) < 1e-10 and rel(gt, ot) < 1e-10
    mp.close()


@pytest.mark.parametrize("nside,L", [(512, 1024), (1024, 2048)])
def test_pol_properties_full_size(nside, L):
    print("case pol properties nside=%d L=%d" % (nside, L), flush=True)
    rng = np.random.default_rng(nside)
    t, a, a2 = spectrum_scaled(rng, L)
    z = np.zeros_like(a)
    mp = cm.Map(nside, L)
    gt, qe, ue = mp.alm2map_pol(t, a, z)
    assert np.array_equal(gt, mp.forward(t))  # unit beam: the scalar synthesis
    _, qb, ub = mp.alm2map_pol(t, z, a)
    assert np.array_equal(qb + 0.0, -ue + 0.0) and np.array_equal(ub + 0.0, qe + 0.0)
    _, q2, u2 = mp.alm2map_pol(t, a2, a)
    _, q12, u12 = mp.alm2map_pol(t, a + a2, a)
    # rounding only (the 1/sin^2 factors amplify it near the poles): the parity bar
    assert rel(q12 - q2, qe) < 1e-10 and rel(u12 - u2, ue) < 1e-10
    low = np.zeros_like(a)
    low[:4] = rng.standard_normal(4)
    _, q, u = mp.alm2map_pol(t, a + low, z)
    assert np.array_equal(q, qe) and np.array_equal(u, ue)
    mp.close()


def test_pol_device_entry_point():
    import torch

    nside, L = 64, 128
    print("case pol device nside=%d L=%d" % (nside, L), flush=True)
    rng = np.random.default_rng(11)
    t, e, b = spectrum_scaled(rng, L)
    mp = cm.Map(nside, L)
    ht, hq, hu = mp.alm2map_pol(t, e, b)
    dev = torch.device("cuda:0")
    ins = [torch.from_numpy(v).to(dev) for v in (t, e, b)]
    outs = [torch.empty(mp.npix, dtype=torch.float64, device=dev) for _ in range(3)]
    stream = torch.cuda.current_stream().cuda_stream
    mp.alm2map_pol_device(*(v.data_ptr() for v in ins), *(o.data_ptr() for o in outs), stream)
    torch.cuda.synchronize()
    for o, h in zip(outs, (ht, hq, hu)):
        assert np.array_equal(o.cpu().numpy(), h)
    with pytest.raises(cm.CmblError) as err:
        mp.alm2map_pol_device(0, ins[1].data_ptr(), ins[2].data_ptr(), *(o.data_ptr() for o in outs), stream)
    assert err.value.code == cm.CMBL_ERR_ARG
    mp.close()
