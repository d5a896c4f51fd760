/*
 * cmbl.h - C ABI of the MI355X-native map-level CMB likelihood
 * (package cosmopp_amd, library cosmopp_amd/libcosmopp_amd.so).
 *
 * The path: -2lnL(a_lm) and its gradient for a HEALPix RING map,
 *   a_lm -> B_l -> spherical-harmonic synthesis -> r = d - s -> w r
 *   -> adjoint analysis (exact transpose) -> gradient in the packed layout,
 * as composed by Cosmo++ around HEALPix alm2map / map2alm in
 * CmbGibbsCGTreats::multiplyByMatrix (source/cmb_gibbs.cpp:359-405).
 *
 * Conventions (all entry points):
 *   - return CMBL_OK (0) or a CMBL_ERR_* code; cmbl_last_error() gives the
 *     message of the last failure on the calling thread.  No C++ exception
 *     crosses this boundary (the reference throws StandardException,
 *     include/exception_handler.hpp:10-30; the C++ wrapper rethrows).
 *   - "packed" a_lm vectors use the almToVector layout
 *     (source/cmb_gibbs.cpp:320-357): for l = 0..lmax: Re a_l0, then for
 *     m = 1..l: Re a_lm, Im a_lm.  Length (lmax+1)^2.
 *   - maps are FP64, RING order, 12*nside^2 pixels.
 *   - buffers are owned by the caller; *_device entry points take device
 *     pointers and an explicit hipStream_t (passed as void*) and are
 *     asynchronous on that stream; the others take host pointers and block.
 *   - a context is not thread-safe; different contexts are independent.
 */
#ifndef CMBL_H
#define CMBL_H

#ifdef __cplusplus
extern "C" {
#endif

#define CMBL_OK 0
#define CMBL_ERR_ARG 1       /* invalid argument (sizes, null pointers, ranges) */
#define CMBL_ERR_HIP 2       /* HIP runtime error or no usable GPU */
#define CMBL_ERR_ALLOC 3     /* device or host allocation failed */
#define CMBL_ERR_STATE 4     /* call out of order (e.g. no data set yet) */
#define CMBL_ERR_NONFINITE 5 /* NaN/Inf in the likelihood value */
#define CMBL_ERR_NOCONV 6    /* iteration limit reached before convergence */

typedef struct cmbl_map cmbl_map;     /* one map: geometry, tables, data, workspace */
typedef struct cmbl_batch cmbl_batch; /* many maps evaluated in one launch per stage */
typedef struct cmbl_gibbs cmbl_gibbs; /* Gibbs constrained-realisation system */

const char* cmbl_last_error(void);
/* Library version string. */
const char* cmbl_version(void);

/* ---------------------------------------------------------------- map context */
/* Replaces the HEALPix map/alm objects and the geometry set-up that
 * CmbGibbsCGTreats holds (source/cmb_gibbs.cpp:303-318): builds the RING
 * geometry, the Legendre recurrence tables and the seeds on the current
 * device.  1 <= nside (power of two), 0 <= lmax <= 4*nside, lmax >= 2. */
int cmbl_map_create(cmbl_map** out, int nside, int lmax);
void cmbl_map_destroy(cmbl_map* map);

/* Launch shapes of a map's kernels, chosen per geometry by the autotuner
 * (no counterpart in the reference, whose transforms are HEALPix's):
 *   k_synth, k_adj  ring pairs per lane of the Legendre synthesis and
 *                   adjoint (1, 2, 4 or 8; a work item is one degree m of
 *                   64 k ring pairs);
 *   split_m         Bluestein ring transforms longer than this (2048, 4096
 *                   or 8192) run as two half-length transforms.
 * cmbl_map_create takes the geometry's entry of the tuning table (the file
 * tuning_<arch>.txt next to the library, or CMBL_TUNING_FILE, read at the
 * first map creation, plus cmbl_tune / cmbl_tuning_set / cmbl_tuning_load
 * results), else a fixed rule.  Results do not depend on the tuning beyond
 * rounding (the parity bar holds for every choice). */
typedef struct cmbl_tuning {
    int k_synth;
    int k_adj;
    int split_m;
} cmbl_tuning;
int cmbl_map_create_tuned(cmbl_map** out, int nside, int lmax, const cmbl_tuning* tuning);
int cmbl_map_get_tuning(const cmbl_map* map, cmbl_tuning* tuning);
/* Autotuner: times reps evaluations (after two untimed ones) of candidate
 * maps of this geometry on the current device -- k_synth = k_adj = 1, 2, 4, 8
 * (each stage timed separately), then, from nside 512, split_m = 2048, 4096,
 * 8192 -- and picks per parameter the first candidate within 2% of the
 * fastest (the fixed rule's choice first, so timing noise does not flip the
 * choice).  The choice enters the tuning table.  times (may be NULL) receives
 * 11 average stage times in ms: synthesis at k = 1, 2, 4, 8; adjoint at
 * k = 1, 2, 4, 8; ring stage at split_m = 2048, 4096, 8192 (-1 if not timed). */
int cmbl_tune(int nside, int lmax, int reps, cmbl_tuning* best, double* times);
int cmbl_tuning_set(int nside, int lmax, const cmbl_tuning* tuning);
int cmbl_tuning_load(const char* path);
int cmbl_tuning_save(const char* path);
int cmbl_map_nside(const cmbl_map* map);
int cmbl_map_lmax(const cmbl_map* map);
long cmbl_map_npix(const cmbl_map* map);
long cmbl_map_nparams(const cmbl_map* map); /* (lmax+1)^2 */

/* Beam B_l (lmax+1 values), the per-l scaling of source/cmb_gibbs.cpp:370-375. */
int cmbl_map_set_beam(cmbl_map* map, const double* beam);
/* Prior spectrum C_l (lmax+1 values); l with C_l <= 0 carry no prior term
 * (the Gibbs sampler sets C_0 = C_1 = 0, source/cmb_gibbs.cpp:180-181). */
int cmbl_map_set_cl(cmbl_map* map, const double* cl);
/* Data map d and per-pixel inverse-noise weight w (= mask/sigma^2, the
 * weighting of source/cmb_gibbs.cpp:381-387), host pointers, npix each. */
int cmbl_map_set_data(cmbl_map* map, const double* d, const double* w);
int cmbl_map_set_data_device(cmbl_map* map, const double* d_d, const double* d_w, void* stream);
/* Count of cmbl_map_set_* calls so far: a cache of (x, -2lnL, gradient) is
 * valid only while it is unchanged (the reference's function objects are
 * rebuilt when their inputs change; here the map is updated in place). */
long cmbl_map_generation(const cmbl_map* map);

/* ---------------------------------------------------------------- likelihood */
/* -2lnL and its gradient at the packed a_lm x.  This is the body of the
 * function-with-gradient pair Math::RealFunctionMultiDim::evaluate /
 * Math::RealFunctionMultiToMulti::evaluate (include/function.hpp:138,168)
 * as adapted by BasicLBFGSFunc::value / derivative (include/lbfgs.hpp:91-93)
 * and Math::LikelihoodWithDerivs::calculate / calculateDeriv
 * (include/likelihood_function.hpp:39,46).  g may be NULL (value only). */
int cmbl_value_and_grad(cmbl_map* map, const double* x, double* f, double* g);
/* Same, device pointers, asynchronous on stream; *d_f is written on device.
 * A non-finite -2lnL raises the map's non-finite flag on the device (the
 * reference's CHECKS_ON check, include/macros.hpp:113-129, cannot run inside
 * an asynchronous call): this entry point returns CMBL_ERR_NONFINITE when an
 * earlier evaluation of the map that has completed raised it, and
 * cmbl_map_check_finite waits for the stream and reports it.  Either clears it. */
int cmbl_value_and_grad_device(cmbl_map* map, const double* d_x, double* d_f, double* d_g,
                               void* stream);
/* Synchronise stream; CMBL_ERR_NONFINITE if an evaluation of the map since
 * the last check produced a non-finite -2lnL (the flag is cleared). */
int cmbl_map_check_finite(cmbl_map* map, void* stream);

/* Forward half of the autograd pair: map = Y B a(x)  (HEALPix alm2map with
 * the beam applied, source/cmb_gibbs.cpp:370-378). */
int cmbl_forward(cmbl_map* map, const double* x, double* out_map);
int cmbl_forward_device(cmbl_map* map, const double* d_x, double* d_map, void* stream);
/* Backward half: g = J^T v with J = d map / d x, i.e. map2alm with ring
 * weight Npix/4pi (exact transpose, source/cmb_gibbs.cpp:390-399) followed
 * by the beam and the packing: g[l^2] = B_l Re u_l0,
 * g[l^2+2m-1] = 2 B_l Re u_lm, g[l^2+2m] = 2 B_l Im u_lm. */
int cmbl_backward(cmbl_map* map, const double* v, double* g);
int cmbl_backward_device(cmbl_map* map, const double* d_v, double* d_g, void* stream);

/* ---------------------------------------------------------------- stage timing */
/* Diagnostics with no counterpart in the reference (it times whole calls
 * with Timer, include/timer.hpp:9-49).  When enabled, every evaluation of the
 * map records HIP events between its stages on the evaluation stream (no
 * host synchronisation); cmbl_map_stage_times synchronises and returns the
 * accumulated milliseconds of each stage since the last reset:
 *   [0] a_lm preparation  [1] Legendre synthesis  [2] ring transforms
 *   [3] adjoint Legendre  [4] gradient assembly   [5] -2lnL reduction */
#define CMBL_NSTAGES 6
int cmbl_map_set_timing(cmbl_map* map, int on); /* on != 0 enables and resets */
int cmbl_map_stage_times(cmbl_map* map, double* ms, long* nevals);

/* ---------------------------------------------------------------- batch */
/* Many independent maps (uneven nside/lmax) evaluated with one launch per
 * stage over a cost-ordered work list.  The maps must stay alive. */
int cmbl_batch_create(cmbl_batch** out, cmbl_map* const* maps, int nmaps);
void cmbl_batch_destroy(cmbl_batch* batch);
/* d_x[i], d_g[i]: device pointers for map i; d_f: device array of nmaps.
 * A map may appear once per batch.  Non-finite values raise each map's flag
 * as in cmbl_value_and_grad_device; a flag raised by an earlier completed
 * evaluation makes this call return CMBL_ERR_NONFINITE. */
int cmbl_batch_value_and_grad_device(cmbl_batch* batch, const double* const* d_x, double* d_f,
                                     double* const* d_g, void* stream);
/* Synchronise stream; CMBL_ERR_NONFINITE if any map's flag is raised
 * (*first_map, may be NULL, receives the first such map, else -1).  Clears them. */
int cmbl_batch_check_finite(cmbl_batch* batch, void* stream, int* first_map);

/* ---------------------------------------------------------------- Gibbs CG */
/* The constrained-realisation system of CMBGibbsSampler::generateSignal,
 * A x = b with A = I + D Y^T N^-1 Y D, D = B_l sqrt(C_l), N^-1 = mask/sigma^2
 * (mask > 0.5), Y^T = map2alm with ring weight Npix/4pi, in the HEALPix
 * complex-a_lm convention and the packed layout: the operator of
 * CmbGibbsCGTreats (source/cmb_gibbs.cpp:303-434).  C_l >= 0, sigma > 0,
 * at least one pixel with mask > 0.5. */
int cmbl_gibbs_create(cmbl_gibbs** out, int nside, int lmax, const double* cl, const double* beam,
                      const double* mask, double pixel_noise);
void cmbl_gibbs_destroy(cmbl_gibbs* gibbs);
/* y = A x: CmbGibbsCGTreats::multiplyByMatrix (source/cmb_gibbs.cpp:359-405). */
int cmbl_gibbs_apply(cmbl_gibbs* gibbs, const double* x, double* y);
int cmbl_gibbs_apply_device(cmbl_gibbs* gibbs, const double* d_x, double* d_y, void* stream);
/* y_lm = x_lm / (1 + C_l B_l^2 / N_l), N_l = sigma^2 4pi / (pixels with mask > 0.5):
 * CmbGibbsCGTreats::preconditioner (source/cmb_gibbs.cpp:407-423). */
int cmbl_gibbs_precondition(cmbl_gibbs* gibbs, const double* x, double* y);
int cmbl_gibbs_precondition_device(cmbl_gibbs* gibbs, const double* d_x, double* d_y, void* stream);
/* Right-hand side of generateSignal (source/cmb_gibbs.cpp:437-487):
 * b = D Y^T (N^-1 d + N^-1/2 omega1) + sqrt(4pi/Npix) Y^T omega0, for the
 * data map d (monopole and dipole already removed, as :449) and the two
 * white-noise maps omega0, omega1 (npix each). */
int cmbl_gibbs_rhs(cmbl_gibbs* gibbs, const double* d, const double* omega0, const double* omega1, double* b);
int cmbl_gibbs_rhs_device(cmbl_gibbs* gibbs, const double* d_d, const double* d_omega0, const double* d_omega1,
                          double* d_b, void* stream);
/* Math::ConjugateGradient::solve (include/conjugate_gradient.hpp:45-102) on
 * the device: x0 = 0, preconditioned CG, stop when |r|^2 <= precision^2.
 * max_iter > 0 bounds the iterations (the reference has no bound):
 * CMBL_ERR_NOCONV if reached, x then holds the last iterate.  *iterations
 * (may be NULL) receives the count.  The device variant blocks on each
 * iteration's scalars, as the reference's host loop does. */
int cmbl_gibbs_solve(cmbl_gibbs* gibbs, const double* b, double* x, double precision, int max_iter,
                     int* iterations);
int cmbl_gibbs_solve_device(cmbl_gibbs* gibbs, const double* d_b, double* d_x, double precision, int max_iter,
                            int* iterations, void* stream);
/* Replace C_l (lmax+1 values, finite, >= 0): D_l = B_l sqrt(C_l) and the
 * preconditioner's divisors follow, as CmbGibbsCGTreats reads cl_ by
 * reference (source/cmb_gibbs.cpp:306, 372, 414) after generateCl changes it. */
int cmbl_gibbs_set_cl(cmbl_gibbs* gibbs, const double* cl);

/* ---------------------------------------------------------------- Gibbs sampler */
/* CMBGibbsSampler (source/cmb_gibbs.cpp:63-549) with its maps, a_lm and
 * solves on the GPU.  map, noise_map and mask are RING-ordered (npix each;
 * cmbl_reorder converts NESTED maps, as the constructor's swap_scheme does);
 * the data are map + noise_map; s starts as map2alm(map, weight 1) / B_l.
 * beam (lmax+1 values, nonzero) may be NULL: then B_l is the Gaussian of
 * fwhm_deg (the reference multiplies it by the HEALPix pixel window read
 * from HEALPix's data files, source/utils.cpp:67-130; pass that product as
 * beam to match).  starting_cl: lmax+1 values >= 0.  seed 0 takes the time
 * (source/cmb_gibbs.cpp:172-173).  One host generator (mt19937 + standard
 * normal, GaussianGenerator of include/random.hpp:32-52) supplies every draw
 * in the reference's order: 4 per pixel (generateW), 2 per pixel
 * (generateSignal), then 2l - 1 per l = 2..lmax (generateCl). */
typedef struct cmbl_gibbs_sampler cmbl_gibbs_sampler;
int cmbl_gibbs_sampler_create(cmbl_gibbs_sampler** out, int nside, int lmax, const double* map,
                              const double* noise_map, const double* mask, double pixel_noise, double fwhm_deg,
                              const double* beam, const double* starting_cl, long seed);
void cmbl_gibbs_sampler_destroy(cmbl_gibbs_sampler* sampler);
/* CMBGibbsSampler::step (source/cmb_gibbs.cpp:533-538): generateW,
 * generateSignal, generateCl.  The signal's CG stops at |r|^2 <= precision^2
 * (the reference passes 1e-20) or after max_iter iterations (the reference
 * has no bound; CMBL_ERR_NOCONV then, and the step still completes);
 * *iterations (may be NULL) receives its count. */
int cmbl_gibbs_sampler_step(cmbl_gibbs_sampler* sampler, double precision, int max_iter, int* iterations);
int cmbl_gibbs_sampler_cl(const cmbl_gibbs_sampler* sampler, double* cl);           /* getCl, lmax+1 */
int cmbl_gibbs_sampler_sigma_l(const cmbl_gibbs_sampler* sampler, double* sigma_l); /* getSigmaL, lmax+1 */
/* The current signal s (packed a_lm, (lmax+1)^2) and the monopole / dipole
 * amplitudes a00, a1-1, a10, a11 of the last generateW. */
int cmbl_gibbs_sampler_signal(const cmbl_gibbs_sampler* sampler, double* s_packed);
int cmbl_gibbs_sampler_dipole(const cmbl_gibbs_sampler* sampler, double* a4);

/* ---------------------------------------------------------------- device-resident loops */
/* The reference's optimiser and sampler loops with every vector on the GPU
 * and each phase's vector work in fused passes, captured in HIP graphs and
 * replayed (pass CMBL_LOOP_EAGER to launch the same kernels without graphs).
 * Their dot products run the order of cmbl_vec_dot and their updates the
 * operations of the cmbl_vec_* entry points, so the iterates equal, bit for
 * bit, those of the reference's LBFGS_General / HMCGeneral run on device
 * vectors built from the cmbl_vec_* entry points (include/cmbl_likelihood.hpp).
 * Each context owns a stream; calls block until their work is complete. */
#define CMBL_LOOP_EAGER 1 /* no HIP graphs */
#define CMBL_HMC_FIXED 2  /* step size tau_max and n_max steps on every trajectory */

typedef struct cmbl_lbfgs cmbl_lbfgs;
/* LBFGS_General(factory, f, starting, m) with the More-Thuente line search
 * (include/lbfgs_general.hpp:144-200: setStarting evaluates -2lnL and its
 * gradient at d_x0).  The map must have its data set. */
int cmbl_lbfgs_create(cmbl_lbfgs** out, cmbl_map* map, const double* d_x0, int m, int flags);
void cmbl_lbfgs_destroy(cmbl_lbfgs* lb);
/* setStarting (include/lbfgs_general.hpp:167-200). */
int cmbl_lbfgs_set_starting(cmbl_lbfgs* lb, const double* d_x0);
/* Callback of LBFGS_General::minimize (iteration, -2lnL, gradient norm, and
 * device pointers to x, g and the last direction z). */
typedef void (*cmbl_lbfgs_callback)(void* user, int iter, double f, double gnorm, const double* d_x, const double* d_g,
                                    const double* d_z);
/* LBFGS_General::minimize(res, epsilon, gNormTol, maxIter, callback)
 * (include/lbfgs_general.hpp:202-405): d_res (may be NULL) receives the
 * minimum point, *f (may be NULL) the value, *iterations and *evaluations the
 * counts of the call.  cb may be NULL. */
int cmbl_lbfgs_minimize(cmbl_lbfgs* lb, double epsilon, double gnorm_tol, int max_iter, double* d_res, double* f,
                        int* iterations, int* evaluations, cmbl_lbfgs_callback cb, void* user);
/* LBFGS_General::getGradient. */
int cmbl_lbfgs_gradient(cmbl_lbfgs* lb, double* d_g);
/* Current value, gradient norm, total iterations and evaluations. */
int cmbl_lbfgs_state(const cmbl_lbfgs* lb, double* f, double* gnorm, int* iter, long* evaluations);

typedef struct cmbl_hmc cmbl_hmc;
/* HMCGeneral(factory, f, starting, masses, tauMax, nMax, seed)
 * (include/hmc_general.hpp:97-127).  seed != 0 (the reference's 0 means
 * time(0)).  Without CMBL_HMC_FIXED each trajectory draws tau = U tau_max and
 * n = ceil(U n_max) as the reference does (:158-164); with it, tau = tau_max
 * and n = n_max and those two draws are not made.  Momenta are the
 * reference's draws (BasicLBFGSFunc::whitenoise with seed + 1, + 2, ...,
 * source/lbfgs.cpp:401-416), made on host threads up to eight trajectories
 * ahead. */
int cmbl_hmc_create(cmbl_hmc** out, cmbl_map* map, const double* d_start, const double* d_masses, double tau_max,
                    int n_max, int seed, int flags);
void cmbl_hmc_destroy(cmbl_hmc* hm);
/* Callback of HMCGeneral::run: device pointer to the current point and its -2lnL. */
typedef void (*cmbl_hmc_callback)(void* user, const double* d_x, double like);
/* HMCGeneral::run(maxIters, callback) (include/hmc_general.hpp:135-231). */
int cmbl_hmc_run(cmbl_hmc* hm, int max_iters, cmbl_hmc_callback cb, void* user);
/* Current -2lnL and totals: trajectories, accepted, likelihood evaluations. */
int cmbl_hmc_state(const cmbl_hmc* hm, double* like, long* trajectories, long* accepted, long* evaluations);
int cmbl_hmc_position(cmbl_hmc* hm, double* d_out);
/* The last trajectory's -2lnL at its end point and its kinetic terms
 * p . M^-1 p at start and end (the Metropolis test compares
 * like_end + kinetic_end with the current -2lnL + kinetic_start). */
int cmbl_hmc_last_trajectory(const cmbl_hmc* hm, double* like_end, double* kinetic_start, double* kinetic_end);

/* The More-Thuente line search of the device-resident L-BFGS as a state
 * machine (moreThuenteSearch, include/line_search.hpp:160-272, with its
 * step rule moreThuenteStep, :12-157): begin() returns the first trial step;
 * feed() takes the function value and slope along the search direction at
 * the last trial and returns info = 0 with the next step in *stp, or the
 * search's result code (1: sufficient decrease and curvature; 2-6: stopped
 * by the interval width, step bounds, evaluation count or rounding) with the
 * final step in *stp.  The reference passes ftol 0.1, gtol 0.5, xtol 1e-15,
 * stpmin 1e-15, stpmax 1e15, maxfev 100 (include/lbfgs_general.hpp:257-262).
 * Host only; no GPU. */
typedef struct cmbl_line_search cmbl_line_search;
int cmbl_line_search_create(cmbl_line_search** out, double ftol, double gtol, double xtol, double stpmin,
                            double stpmax, int maxfev);
void cmbl_line_search_destroy(cmbl_line_search* ls);
int cmbl_line_search_begin(cmbl_line_search* ls, double f0, double dg0, double stp0, double* stp);
int cmbl_line_search_feed(cmbl_line_search* ls, double f, double dg, double* stp, int* info);

/* ---------------------------------------------------------------- inputs (host) */
/* Simulate::simulateAlm (source/simulate.cpp:162-192): mt19937(seed) +
 * std::normal_distribution, l-outer / m-inner draw order; output packed. */
int cmbl_simulate_alm(const double* cl, int lmax, long seed, double* x_packed);
/* Simulate::simulateWhiteNoise (source/simulate.cpp:194-203), RING order. */
int cmbl_simulate_white_noise(int nside, double sigma, long seed, double* out_map);
/* Utils::beamFunction (source/utils.cpp:54-64) for l = 0..lmax, fwhm in degrees. */
int cmbl_beam(int lmax, double fwhm_deg, double* beam);
/* Latitude cut: mask_p = 0 where |pi/2 - theta_p| < cut (source/test_like_high.cpp:123-130). */
int cmbl_mask_latitude(int nside, double cut_rad, double* mask);
/* Utils::maskRegions (source/utils.cpp:235-253): mask_p = 0 where the pixel
 * centre lies within angle_j of (theta_j, phi_j). */
int cmbl_mask_regions(int nside, int nreg, const double* theta, const double* phi,
                      const double* angle, double* mask);
/* w_p = 1/sigma^2 where mask_p > 0.5, else 0 (source/cmb_gibbs.cpp:383-386). */
int cmbl_weights_from_mask(long npix, const double* mask, double sigma, double* w);
/* Pixel centre of a RING-scheme pixel (HEALPix pix2ang_ring). */
int cmbl_pix2ang_ring(int nside, long pix, double* theta, double* phi);

/* BasicLBFGSFunc::whitenoise for one process (source/lbfgs.cpp:401-416):
 * seed' = ceil(1e6 * UniformRealGenerator(seed).generate()), then n draws of
 * amplitude * GaussianGenerator(seed', 0, 1), into a host array. */
int cmbl_whitenoise(int seed, double amplitude, long n, double* out);

/* ---------------------------------------------------------------- quadrature analysis, MASTER */
/* HEALPix map2alm with ring weight 1 (quadrature weight 4pi/Npix per pixel, no
 * iteration), as Master uses it for the mask and the masked map
 * (source/master.cpp:112-117, 143-147): packed complex a_lm (almToVector
 * layout), no beam. */
int cmbl_map2alm_quadrature(cmbl_map* map, const double* m, double* alm);
int cmbl_map2alm_quadrature_device(cmbl_map* map, const double* d_m, double* d_alm, void* stream);

/* ---------------------------------------------------------------- spin-2 synthesis */
/* HEALPix alm2map_pol(almT, almE, almB, mapT, mapQ, mapU) as called at
 * source/likelihood.cpp:585 and source/c_matrix_generator.cpp:577-578,
 * 645-646, 662-663: packed T, E, B a_lm (almToVector layout, no beam) ->
 * RING maps
 *   T = sum a^T Y,  Q = -sum (a^E W + i a^B X),  U = -sum (a^B W - i a^E X)
 * with W = (2Y + -2Y)/2, X = (2Y - -2Y)/2 (spin-weighted harmonics of Goldberg
 * et al. 1967 with the Condon-Shortley phase; a^E = -(a_2 + a_-2)/2,
 * a^B = i (a_2 - a_-2)/2).  l < 2 of E and B do not contribute.  The first call
 * allocates a coefficient table and three parity-sum buffers on the map's
 * device (5 * 16 * nalm + 3 * 32 * 2 nside (lmax+1) bytes). */
int cmbl_alm2map_pol(cmbl_map* map, const double* t, const double* e, const double* b, double* map_t,
                     double* map_q, double* map_u);
int cmbl_alm2map_pol_device(cmbl_map* map, const double* d_t, const double* d_e, const double* d_b,
                            double* d_map_t, double* d_map_q, double* d_map_u, void* stream);
/* ps(alm, l) = (|a_l0|^2 + 2 sum_m |a_lm|^2) / (2l+1) (source/master.cpp:62-73), l = 0..lmax. */
int cmbl_pseudo_cl(const double* alm, int lmax, double* cl);
/* Mode-coupling matrix of Master::calculateCouplingKernel (source/master.cpp:180-249):
 * M[l1][l2] = (2 l2 + 1)/(4pi) sum_{l3 <= lmax} (2 l3 + 1) W_l3 (l1 l2 l3; 0 0 0)^2,
 * row-major (lmax+1)^2, from the mask pseudo spectrum W (lmax+1 values).  The
 * device variant needs a workspace of (lmax+1)^2 doubles. */
int cmbl_coupling_matrix(const double* w, int lmax, double* M);
int cmbl_coupling_matrix_device(const double* d_w, int lmax, double* d_work, double* d_M, void* stream);

/* MASTER estimator (class Master, source/master.cpp:28-155, 353-471): mask and
 * beam fixed at creation (RING order); bins = nedges increasing l edges (bin b
 * covers [bins[b], bins[b+1]), lmax = bins[nedges-1] - 1, pass lmax 0) or
 * nedges = 0 for every l up to lmax.  The K matrix is factorised once. */
typedef struct cmbl_master cmbl_master;
int cmbl_master_create(cmbl_master** out, int nside, int lmax, const double* mask, const double* beam,
                       const int* bins, int nedges);
void cmbl_master_destroy(cmbl_master* master);
int cmbl_master_nbins(const cmbl_master* master);
int cmbl_master_mask_spectrum(const cmbl_master* master, double* w);   /* W_l, lmax+1 values */
int cmbl_master_coupling(const cmbl_master* master, double* M);        /* (lmax+1)^2 */
/* Master::calculate(map): ps[b] per bin (or per l), pseudo (may be NULL) the
 * masked map's pseudo spectrum, lmax+1 values. */
int cmbl_master_calculate(cmbl_master* master, const double* map, double* ps, double* pseudo);

/* ---------------------------------------------------------------- map files, NEST/RING */
/* HEALPix FITS maps as read and written by HEALPix read_Healpix_map_from_fits /
 * write_Healpix_map_to_fits, which the reference uses for its maps, noise
 * maps and masks (source/cmb_gibbs.cpp:56-58, source/utils.cpp:26-52): a
 * BINTABLE extension with PIXTYPE 'HEALPIX', ORDERING, NSIDE and implicit
 * pixel indexing; columns of float64/float32/int types (TSCAL/TZERO applied).
 * Values are returned in the file's ordering (see *nested). */
int cmbl_fits_map_info(const char* path, int* nside, int* nested, long* npix, int* ncolumns);
int cmbl_fits_read_map(const char* path, int column, double* map, long npix); /* column from 1 */
int cmbl_fits_write_map(const char* path, const double* map, int nside, int nested, int overwrite);
/* HEALPix pixel index conversions (nest2ring / ring2nest of Healpix_Base,
 * as Healpix_Map::Import uses them to change a map's ordering). */
int cmbl_nest2ring(int nside, long ipnest, long* ipring);
int cmbl_ring2nest(int nside, long ipring, long* ipnest);
/* Reorder a whole map on the GPU: to_ring != 0 maps NESTED -> RING, else
 * RING -> NESTED.  in and out must not alias. */
int cmbl_reorder(int nside, const double* in, double* out, int to_ring);
int cmbl_reorder_device(int nside, const double* d_in, double* d_out, int to_ring, void* stream);

/* ---------------------------------------------------------------- device memory */
/* Plain device buffers for callers that hold no HIP headers (the C++ wrapper
 * include/cmbl_likelihood.hpp).  Copies block until complete. */
int cmbl_device_alloc(void** ptr, unsigned long bytes);
int cmbl_device_free(void* ptr);
int cmbl_copy_to_device(void* d_dst, const void* src, unsigned long bytes);
int cmbl_copy_to_host(void* dst, const void* d_src, unsigned long bytes);
int cmbl_copy_device(void* d_dst, const void* d_src, unsigned long bytes, void* stream);
int cmbl_stream_synchronize(void* stream);

/* ---------------------------------------------------------------- device vectors */
/* BLAS-1 on device vectors of n doubles, for a device-resident LargeVector
 * (the operations of BasicLargeVector, source/lbfgs.cpp:323-399).  All are
 * asynchronous on stream except the reductions, which block and return the
 * value through a host pointer.  y and x may be the same vector.  The dot
 * products use a workspace per (device, stream), allocated at the first dot
 * on that stream and kept until cmbl_vec_release_workspaces. */
int cmbl_vec_copy(double* d_y, const double* d_x, double c, long n, void* stream);   /* y = c x */
int cmbl_vec_set(double* d_y, double v, long n, void* stream);                       /* y = v */
int cmbl_vec_axpy(double* d_y, const double* d_x, double c, long n, void* stream);   /* y += c x */
int cmbl_vec_mul(double* d_y, const double* d_x, long n, void* stream);              /* y *= x */
int cmbl_vec_div(double* d_y, const double* d_x, long n, void* stream);              /* y /= x */
int cmbl_vec_pow(double* d_y, double p, long n, void* stream);                       /* y = y^p */
int cmbl_vec_dot(const double* d_x, const double* d_y, long n, double* out, void* stream);
/* Device-side dot product: *d_out = x . y, asynchronous. */
int cmbl_vec_dot_device(const double* d_x, const double* d_y, long n, double* d_out, void* stream);
/* Free every dot-product workspace (waits for each device first). */
int cmbl_vec_release_workspaces(void);
/* cmbl_vec_dot_device without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Its scratch (1024 partial sums,
 * 8 KiB) is a static array in device memory, one copy per GPU, whose
 * address is looked up on the first call on each device; the scratch does
 * not grow with n, so any n >= 0 is accepted.  Calls on one GPU share that
 * scratch: calls that overlap (different streams, or several threads) race
 * and must be ordered by the caller.  The result equals cmbl_vec_dot_device's
 * bit for bit (same kernels, same summation order). */
int cmbl_vec_dot_device_static(const double* d_x, const double* d_y, long n, double* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
