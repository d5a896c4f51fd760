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
CONFIG5_CHAIN = dict(nside=256, lmax=512, fwhm_deg=40.0 / 60.0, sigma=40.0, cut_deg=20.0)


def batch_configs(nmaps=64, seed=3):
    """64 maps of uneven size: lmax ~ U{256, 512, 768, 1024, 1536}, nside =
    2^round(log2(lmax/2)), FWHM = 5400/lmax arcmin, sigma ~ U[10, 50] uK,
    cut ~ U[10, 30] deg, independent seeds."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nmaps):
        lmax = int(rng.choice([256, 512, 768, 1024, 1536]))
        nside = 1 << int(round(math.log2(lmax / 2)))
        out.append(dict(nside=nside, lmax=lmax, fwhm_deg=5400.0 / lmax / 60.0, sigma=float(rng.uniform(10, 50)),
                        cut_deg=float(rng.uniform(10, 30)), seed=1000 + 10 * i))
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
    mp = cm.Map(nside, lmax)
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
