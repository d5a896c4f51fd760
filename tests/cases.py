"""Seeded likelihood inputs shared by the GPU tests, bench.py and smoke().

The construction follows the reference's own recipe: a_lm from
Simulate::simulateAlm with C_l = 1e3/(l(l+1)) (C_0 = C_1 = 0), data
d = Y B a_true + simulateWhiteNoise, a latitude cut (source/test_like_high.cpp:
123-130) plus optional point-source holes (Utils::maskRegions with centres drawn
as source/test_like_high.cpp:111-121), w = mask / sigma^2
(source/cmb_gibbs.cpp:383-386), and the evaluation point
x = a_true + 0.1 * simulateAlm(seed_x).  The data map is synthesised by the
oracle so that GPU results are never compared with themselves.
"""
import math

import numpy as np

import cosmopp_amd as cm


class Case:
    def __init__(self, nside, lmax, fwhm_deg, sigma, cut_deg, seed=1, holes=0, hole_deg=0.5, synth=None):
        self.nside, self.lmax = nside, lmax
        self.npix = 12 * nside * nside
        self.cl = cm.cl_shape(lmax)
        self.beam = cm.beam(lmax, fwhm_deg)
        self.a_true = cm.simulate_alm(self.cl, lmax, 1000 + seed)
        if synth is None:
            import oracle

            synth = oracle.synthesis
        self.d = synth(nside, lmax, self.a_true, self.beam) + cm.simulate_white_noise(nside, sigma, 2000 + seed)
        mask = cm.mask_latitude(nside, math.radians(cut_deg))
        if holes:
            th = np.array(_uniform(3000 + seed, math.pi / 50, math.pi - math.pi / 50, holes))
            ph = np.array(_uniform(3001 + seed, 0.0, 2 * math.pi, holes))
            mask = cm.mask_regions(nside, th, ph, np.full(holes, math.radians(hole_deg)), mask)
        self.mask = mask
        self.w = cm.weights_from_mask(mask, sigma)
        self.x = self.a_true + 0.1 * cm.simulate_alm(self.cl, lmax, 4000 + seed)

    def make_map(self):
        mp = cm.Map(self.nside, self.lmax)
        mp.set_beam(self.beam)
        mp.set_cl(self.cl)
        mp.set_data(self.d, self.w)
        return mp


def _uniform(seed, lo, hi, n):
    """UniformRealGenerator(seed, lo, hi) draws (mt19937 + uniform_real_distribution)."""
    # std::uniform_real_distribution<double> on mt19937 consumes two 32-bit draws per value
    # (generate_canonical with 53 bits); numpy's MT19937 has a different seeding, so the
    # draws come from a small local mt19937 restatement.
    gen = _MT19937(seed)
    out = []
    for _ in range(n):
        lo32 = gen.next()
        hi32 = gen.next()
        u = (lo32 + hi32 * 4294967296.0) / 18446744073709551616.0
        if u >= 1.0:
            u = math.nextafter(1.0, 0.0)
        out.append(lo + (hi - lo) * u)
    return out


class _MT19937:
    def __init__(self, seed):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.idx = 624

    def next(self):
        if self.idx >= 624:
            for i in range(624):
                y = (self.mt[i] & 0x80000000) | (self.mt[(i + 1) % 624] & 0x7FFFFFFF)
                v = self.mt[(i + 397) % 624] ^ (y >> 1)
                if y & 1:
                    v ^= 0x9908B0DF
                self.mt[i] = v
            self.idx = 0
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


# The workloads: nside 64 / lmax 128 (CPU-runnable), nside 512 / lmax 1024
# with 200 point-source holes (finite-difference check), nside 1024 /
# lmax 2048 (the headline workload), nside 2048 / lmax 4096 with 16 chain
# maps of nside 256 / lmax 512, and the 64-map uneven batch below.
CONFIG1 = dict(nside=64, lmax=128, fwhm_deg=2.0, sigma=10.0, cut_deg=20.0)
CONFIG2 = dict(nside=512, lmax=1024, fwhm_deg=20.0 / 60.0, sigma=30.0, cut_deg=20.0, holes=200, hole_deg=0.5)
CONFIG4 = dict(nside=1024, lmax=2048, fwhm_deg=10.0 / 60.0, sigma=20.0, cut_deg=20.0)
CONFIG5 = dict(nside=2048, lmax=4096, fwhm_deg=5.0 / 60.0, sigma=15.0, cut_deg=20.0)
# The chain maps are evaluated 16 at a time (cmbl_batch) with one ring pair
# per lane in the adjoint, so that the batch runs it on the matrix cores in
# two groups of eight maps that share the recurrence.
CONFIG5_CHAIN = dict(nside=256, lmax=512, fwhm_deg=40.0 / 60.0, sigma=40.0, cut_deg=20.0, tuning=(2, 1, 4096))


def batch_configs(nmaps=64, seed=3):
    """64 maps of uneven size: lmax ~ U{256, 512, 768, 1024, 1536}, nside =
    2^round(log2(lmax/2)), FWHM = 5400/lmax arcmin, sigma ~ U[10, 50] uK,
    cut ~ U[10, 30] deg, independent seeds.  Built as batch members: the
    tuned table's synthesis K (1 at lmax 256, else 2) and one ring pair per
    lane in the adjoint, so that the batch runs the adjoint of each lmax
    class on the matrix cores in groups of up to eight maps."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nmaps):
        lmax = int(rng.choice([256, 512, 768, 1024, 1536]))
        nside = 1 << int(round(math.log2(lmax / 2)))
        out.append(dict(nside=nside, lmax=lmax, fwhm_deg=5400.0 / lmax / 60.0, sigma=float(rng.uniform(10, 50)),
                        cut_deg=float(rng.uniform(10, 30)), seed=1000 + 10 * i,
                        tuning=(1 if lmax <= 256 else 2, 1, 4096)))
    return out


def device_case(cfg, seed):
    """Inputs of one workload map with the data synthesised by the library
    (for sizes where the oracle's synthesis would dominate a test's time; the
    oracle then evaluates the likelihood on the same data): returns the map
    (beam, C_l and data set) and a dict with cl, beam, d, w, x."""
    nside, lmax = cfg["nside"], cfg["lmax"]
    cl = cm.cl_shape(lmax)
    b = cm.beam(lmax, cfg["fwhm_deg"])
    a_true = cm.simulate_alm(cl, lmax, seed)
    mp = cm.Map(nside, lmax, tuning=cfg.get("tuning"))
    mp.set_beam(b)
    mp.set_cl(cl)
    d = mp.forward(a_true) + cm.simulate_white_noise(nside, cfg["sigma"], seed + 1)
    mask = cm.mask_latitude(nside, math.radians(cfg["cut_deg"]))
    holes = cfg.get("holes", 0)
    if holes:
        th = np.array(_uniform(seed + 3, math.pi / 50, math.pi - math.pi / 50, holes))
        ph = np.array(_uniform(seed + 4, 0.0, 2 * math.pi, holes))
        mask = cm.mask_regions(nside, th, ph, np.full(holes, math.radians(cfg.get("hole_deg", 0.5))), mask)
    w = cm.weights_from_mask(mask, cfg["sigma"])
    mp.set_data(d, w)
    x = a_true + 0.1 * cm.simulate_alm(cl, lmax, seed + 2)
    return mp, dict(cl=cl, beam=b, d=d, w=w, x=x)


# The Gibbs sampler subtests of source/test_cmb_gibbs.cpp:64-99 (0: nside 16,
# Gibbs lmax 30, FWHM 10 deg, sigma 1; 1: nside 32, lmax 50, FWHM 5 deg,
# sigma 2; 4: as 1 with the reference's mask1.fits).  The simulated sky has
# lmax 64 / 128 as there; its spectrum is C_l = 2pi 1000 / (l(l+1)) (the
# reference's CLASS LCDM spectrum is not available), the beam is the Gaussian
# of the FWHM (the reference's also carries HEALPix's pixel window).
GIBBS_SUBTESTS = {
    0: dict(nside=16, lsim=64, lmax=30, fwhm=10.0, sigma=1.0, scale=0.4, mask=False),
    1: dict(nside=32, lsim=128, lmax=50, fwhm=5.0, sigma=2.0, scale=0.4, mask=False),
    4: dict(nside=32, lsim=128, lmax=50, fwhm=5.0, sigma=2.0, scale=5.0, mask=True),
}


def gibbs_case(i, seed=21, ring_mask=None, synth=None):
    """Inputs of Gibbs subtest i: dict(nside, lmax, map, noise, mask, sigma,
    fwhm, beam, cl0, cl_original, seed, scale).  ring_mask supplies mask1 in
    RING order for subtest 4; synth(nside, L, a, beam) makes the sky map
    (default: the oracle's synthesis)."""
    if synth is None:
        import oracle

        synth = oracle.synthesis
    c = GIBBS_SUBTESTS[i]
    nside, lsim, lmax = c["nside"], c["lsim"], c["lmax"]
    cl_true = np.zeros(lsim + 1)
    l = np.arange(2, lsim + 1)
    cl_true[2:] = 2 * np.pi * 1000.0 / (l * (l + 1))
    a = cm.simulate_alm(cl_true, lsim, seed)
    m = synth(nside, lsim, a, cm.beam(lsim, c["fwhm"]))
    noise = cm.simulate_white_noise(nside, c["sigma"], seed + 1)
    mask = np.ones(12 * nside * nside) if not c["mask"] else np.asarray(ring_mask, dtype=np.float64)
    cl0 = np.full(lmax + 1, 1000.0)                  # source/test_cmb_gibbs.cpp:143-146
    ll = np.arange(1, lmax + 1)
    cl0[1:] *= 2 * np.pi / (ll * (ll + 1))
    cl0[0] = cl0[1]
    orig = np.array([(a[k * k] ** 2 + 2 * np.sum(a[k * k + 1:(k + 1) ** 2] ** 2)) / (2 * k + 1)
                     for k in range(lmax + 1)])
    return dict(nside=nside, lmax=lmax, map=m, noise=noise, mask=mask, sigma=c["sigma"], fwhm=c["fwhm"],
                beam=cm.beam(lmax, c["fwhm"]), cl0=cl0, cl_original=orig, seed=seed + 2, scale=c["scale"])


def gibbs_acceptance(case, sigma_l_chain):
    """source/test_cmb_gibbs.cpp:206-229: for l = 2..lmax the chain's C_l
    must match the realisation's within the relative scale (Math::areEqual,
    include/numerics.hpp:17-25: |est - orig| / |orig| < scale).
    The chain's C_l estimate is the mean over samples of sigma_l / (2l+1), the
    mode of the inverse-gamma draw of generateCl, in place of the peak of the
    smoothed posterior.  Returns the list of failing l."""
    s = np.asarray(sigma_l_chain, dtype=np.float64)
    L = case["lmax"]
    bad = []
    for l in range(2, L + 1):
        est = s[:, l].mean() / (2 * l + 1)
        orig = case["cl_original"][l]
        if not abs(est - orig) < case["scale"] * abs(orig):
            bad.append((l, orig, est))
    return bad
