/*
 * CPU FP64 restatement of the Cosmo++ map-level CMB likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (cosmopp_amd) never does.
 *
 * What it restates (reference = aslanyan/cosmopp 3.1):
 *   - operator order of CmbGibbsCGTreats::multiplyByMatrix
 *     (source/cmb_gibbs.cpp:359-405): pack -> B_l -> alm2map -> mask/sigma^2
 *     -> map2alm with ring weight Npix/4pi (exact transpose) -> B_l -> unpack;
 *   - packing almToVector / vectorToAlm (source/cmb_gibbs.cpp:320-357);
 *   - the arithmetic of HEALPix alm2map / map2alm (external, not vendored:
 *     HEALPix C++ 3.20 + libsharp, cmake_settings_template.txt:23-24),
 *     restated from the published algorithm: RING geometry of Gorski et al.
 *     2005, normalised associated-Legendre recursion with Condon-Shortley
 *     phase (same normalisation as include/spherical_harmonics.hpp:29-78),
 *     per-ring phase folding and ring DFTs;
 *   - Utils::beamFunction (source/utils.cpp:54-64), Utils::maskRegions
 *     (source/utils.cpp:235-253), latitude cut (source/test_like_high.cpp:123-130).
 *
 * Parity status per function is listed in the header comment of
 * cmbl_oracle.c.
 */
#ifndef CMBL_ORACLE_H
#define CMBL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Ring of a HEALPix RING-scheme map (0-based ring index 0 .. 4*nside-2). */
typedef struct {
    double z;      /* cos(theta) */
    double sint;   /* sin(theta) */
    double phi0;   /* longitude of the first pixel */
    long nph;      /* pixels in the ring */
    long ofs;      /* index of the first pixel */
} orc_ring;

int  orc_nring(int nside);
void orc_ring_info(int nside, int ring, orc_ring* out);
/* Pixel centre, chealpix pix2ang_ring convention (per pixel, independent of
 * orc_ring_info; used to pin the ring tables). */
void orc_pix2ang_ring(int nside, long pix, double* theta, double* phi);

/* lambda_lm(x) = Y_lm(theta, 0) for l = m..L into out[0 .. L-m].
 * Values below 2^-200 (relative to O(1)) are returned as 0. */
void orc_lambda(int L, int m, double x, double sint, double* out);

/* In-place complex DFT, unnormalised.  sign = -1: sum_j x_j e^{-2 pi i jk/n};
 * sign = +1: e^{+2 pi i jk/n}.  Any n >= 1. */
void orc_fft(int n, double* re, double* im, int sign);
/* Naive O(n^2) DFT, same convention; used only to pin orc_fft. */
void orc_dft(int n, const double* re, const double* im, double* ore, double* oim, int sign);

/* s = Y B a : packed real a_lm (n = (L+1)^2, almToVector layout) -> RING map. */
void orc_synthesis(int nside, int L, const double* x_packed, const double* beam, double* map);
/* orc_synthesis restricted to the rings listed (0-based ring indices): the
 * pixels of ring rings[i] follow those of rings[0..i-1] in out. */
void orc_synthesis_rings(int nside, int L, const double* x_packed, const double* beam, int nsel,
                         const int* rings, double* out);
/* HEALPix alm2map_pol: packed T, E, B -> RING maps T, Q, U (no beam). */
void orc_alm2map_pol(int nside, int L, const double* t, const double* e, const double* b, double* mt,
                     double* mq, double* mu);
/* Exact transpose of orc_synthesis: out_packed = J^T map where J = d s / d x.
 * With u_lm = sum_p map_p lambda_lm(theta_p) e^{-i m phi_p}:
 *   out[l^2] = B_l Re u_l0, out[l^2+2m-1] = 2 B_l Re u_lm, out[l^2+2m] = 2 B_l Im u_lm. */
void orc_adjoint(int nside, int L, const double* map, const double* beam, double* out_packed);

/* -2 ln L = sum_p w_p (d_p - s_p)^2 + sum_{l>=2, C_l>0} (a_l0^2 + 2 sum_m |a_lm|^2) / C_l
 * and (if g != 0) its gradient w.r.t. the packed a_lm. */
double orc_neg2lnL(int nside, int L, const double* x_packed, const double* beam,
                   const double* cl, const double* d, const double* w, double* g);

/* Utils::beamFunction(l, fwhm) for l = 0..L (fwhm in degrees). */
void orc_beam(int L, double fwhm_deg, double* beam);
/* mask_p = 0 where |pi/2 - theta_p| < cut, else unchanged (mask must be initialised). */
void orc_mask_latitude(int nside, double cut_rad, double* mask);
/* Utils::maskRegions: mask_p = 0 where dir_p . dir_j > cos(angle_j). */
void orc_mask_regions(int nside, int nreg, const double* theta, const double* phi,
                      const double* angle, double* mask);

/* n draws of Math::GaussianGenerator(seed, 0, 1) (include/random.hpp:32-52). */
void orc_gauss(long seed, long n, double* out);
int orc_num_threads(void);
/* OpenMP threads of the calling thread's later parallel regions (timing). */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
