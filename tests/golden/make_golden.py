"""Generate the golden fixtures in tests/golden/ from the reference's own code.

Runs in the build container only: it needs oracle/_ref/libref.so, built by
`make -C oracle ref` from the header-only reference numerics under
/root/reference/include (SphericalHarmonics, GaussianGenerator,
UniformRealGenerator).  Pixel positions come from the oracle's pix2ang_ring
(pinned separately against the ring tables).

Outputs (small .npz files, data only):
  sh_lambda.npz      lambda_lm(theta) = Re Y_lm(theta, 0) for l <= 96 at 6 thetas
  brute_n8_l16.npz   brute-force map sum_lm a_lm Y_lm(theta_p, phi_p), nside 8, lmax 16
  brute_n16_l40.npz  same, nside 16, lmax 40, with a Gaussian beam
  rng.npz            GaussianGenerator / UniformRealGenerator draws for fixed seeds
  w3j.npz            (l1 l2 l3; 0 0 0)^2 for l's <= 24 from Math::Wigner3JZeroM
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def brute(nside, lmax, seed, fwhm):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((lmax + 1) ** 2)
    for l in range(lmax + 1):
        x[l * l:(l + 1) * (l + 1)] *= 1.0 / (1.0 + l)
    b = oracle.beam(lmax, fwhm)
    npix = 12 * nside * nside
    th = np.zeros(npix)
    ph = np.zeros(npix)
    for p in range(npix):
        th[p], ph[p] = oracle.pix2ang_ring(nside, p)
    m = oracle.ref_brute_synthesis(lmax, x, b, th, ph)
    return dict(nside=nside, lmax=lmax, x=x, beam=b, theta=th, phi=ph, map=m)


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref.so missing: run `make -C oracle ref` first")
    thetas = np.array([0.05, 0.7, 1.3, np.pi / 2, 2.2, 3.0])
    L = 96
    tab = np.stack([oracle.ref_lambda_table(L, t) for t in thetas])
    np.savez_compressed(os.path.join(HERE, "sh_lambda.npz"), lmax=L, theta=thetas, lam=tab)

    np.savez_compressed(os.path.join(HERE, "brute_n8_l16.npz"), **brute(8, 16, 1, 0.0))
    np.savez_compressed(os.path.join(HERE, "brute_n16_l40.npz"), **brute(16, 40, 2, 4.0))

    np.savez_compressed(
        os.path.join(HERE, "rng.npz"),
        gauss_seed=np.array([12345]), gauss=oracle.ref_gauss(12345, 0.0, 1.0, 4000),
        gauss_sigma_seed=np.array([777]), gauss_sigma=oracle.ref_gauss(777, 0.0, 20.0, 1000),
        uni_seed=np.array([1000000]), uni=oracle.ref_uniform(1000000, np.pi / 50, np.pi - np.pi / 50, 200),
    )
    np.savez_compressed(os.path.join(HERE, "w3j.npz"), lmax=np.array([24]), w3j2=oracle.ref_w3j2_table(24))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
