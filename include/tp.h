/*
 * tp.h -- C ABI of libtp.so, the MI355X (gfx950) implementation of the
 * ThreePhotons.jl three-photon correlation c3(k1,k2,k3,alpha,beta) and its
 * Bayes log-likelihood energy, as evaluated on every Monte Carlo step of the
 * rotate_all_at_once annealing.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns 0
 * on success or a negative TP_E_* code (the message is in tp_last_error()),
 * and never throws, aborts or exits.  Calls are synchronous: results are in
 * the caller's buffers when the call returns.  Host pointers are borrowed for
 * the duration of the call only.  A tp_ctx is used by one host thread at a
 * time.
 *
 * Layouts follow the reference's Julia (column-major) arrays:
 *   coefficients  complex128 interleaved (re,im), shell k (1-based) starts at
 *                 element (k-1)*ld, entry (l,m) at seanindex(l,m,LMAX)-1
 *                 (spherical_harmonics.jl:163-170); ld >= LMAX^2.
 *   c3raw         R x klength, R = 2N^2, row r = a + (b-1)N, column kt in the
 *                 order for k1=1:K, k2=1:k1, k3=1:k2 (correlations.jl:179-184,
 *                 return_raw=true).
 *   c3ref         N x 2N x Kr x Kr x Kr float64, element (a,b,k3,k2,k1)
 *                 (datagen.jl:432-475 after loadHistograms).
 */
#ifndef TP_H
#define TP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TP_ABI_VERSION 3

#define TP_OK 0
#define TP_E_ARG (-1)          /* invalid argument or shape */
#define TP_E_HIP (-2)          /* HIP runtime error (device missing, launch failure) */
#define TP_E_OOM (-3)          /* device or host allocation failed */
#define TP_E_STATE (-4)        /* object used with a mismatching basis/context */
#define TP_E_UNSUPPORTED (-5)  /* parameter combination not implemented */

typedef struct tp_ctx tp_ctx;
typedef struct tp_basis tp_basis;
typedef struct tp_ref tp_ref;
typedef struct tp_state tp_state;
typedef struct tp_batch tp_batch;
typedef struct tp_mlp tp_mlp;

/* ABI version of the loaded library (== TP_ABI_VERSION). */
int tp_abi_version(void);

/* Thread-local text of the last error. */
const char *tp_last_error(void);

/* Open device `device` (HIP ordinal) and create a context with its own stream.
 * Replaces CUDA_init (cuda.jl:10-21); a failure lets the caller fall back to
 * the reference's CPU BasisType, as rotation_search does
 * (determination.jl:210-216). */
int tp_init(int device, tp_ctx **out);
void tp_finalize(tp_ctx *ctx);

/* Precompute the basis in FP64 on the host and upload it.
 * Replaces calculate_basis + CUDA_store_basis / CUDA_calculate_basis
 * (correlations.jl:87-153, cuda.jl:45-65).  for_intensity=1 selects even l
 * (lrange = 0:2:L).  lambda dq != 0 is the curved Ewald sphere: the
 * Legendre factors are taken at theta_k = acos(k lambda dq / (4 pi))
 * (correlations.jl:94), which needs |K lambda dq| <= 4 pi; such a basis
 * evaluates in FP64 or TP_PREC_FP16X3_T64 (mixed-precision triple products,
 * G and c3 in FP64) through tp_c3, tp_energy, tp_state_*, tp_anneal,
 * tp_energy_batch, tp_energy_grad (FP64) and the batched runs (tp_batch_*,
 * TP_PREC_FP16X3_T64); the other mixed modes and tp_two_photons (use
 * tp_two_photons_ewald) return TP_E_UNSUPPORTED for it. */
int tp_basis_create(tp_ctx *ctx, int L, int LMAX, int N, int K, double lambda, double dq,
                    int for_intensity, tp_basis **out);
/* The same from a basis the caller already holds: wignerlist[basislen] and
 * indices (9 x basislen, column j = [sean(l1,m1), sean(l2,m2), sean(l3,-m3),
 * l1, m1, l2, m2, l3, -m3]) of the reference's BasisType (correlations.jl:35-63,
 * 87-153); the caller's weights are used as given.  pacombos (9 x ncombos) may
 * be NULL; when given its length must be a multiple of klength.  Replaces
 * CUDA_store_basis (cuda.jl:45-55).  Even l (for_intensity) only. */
int tp_basis_from_host(tp_ctx *ctx, const double *wignerlist, const int64_t *indices, int64_t basislen,
                       const int64_t *pacombos, int64_t ncombos, int N, int L, int LMAX, int K, double lambda,
                       double dq, tp_basis **out);
void tp_basis_destroy(tp_basis *b);

/* basislen, number of l-triples, klength = K(K+1)(K+2)/6, R = 2N^2. */
int tp_basis_info(const tp_basis *b, int64_t *basislen, int64_t *n_lt, int64_t *klength, int64_t *R);

/* Arithmetic of the contraction for every later call on this basis.
 *   TP_PREC_FP64   (default) FP64 MFMA throughout; the parity path.
 *   TP_PREC_FP16X3 c3 = F.As with F and As split into FP16 hi/lo pairs after
 *                  power-of-two scaling; hi.hi + (hi.lo + lo.hi) on FP16 MFMA
 *                  with FP32 accumulation.  Triple products, the histogram
 *                  terms and the log-likelihood sum stay FP64.
 *   TP_PREC_BF16   one BF16 product with FP32 accumulation (comparison mode).
 *   TP_PREC_TRIPLE16 compensated FP16 triple products (as in FP16X3) with the
 *                  FP64 contraction.
 *   TP_PREC_FP16X3_T64 the factored mixed form: FP64 triple products (the
 *                  sorted l-triple form of the FP64 path), G = H.As with H
 *                  and As split into FP16 hi/lo pairs after power-of-two
 *                  scaling (hi.hi + (hi.lo + lo.hi), FP32 accumulation),
 *                  then c3 = Bm.G and the energy in FP64 on the separated
 *                  grid: needs the library's own weights, K <= 32 and
 *                  L <= 20 (TP_E_UNSUPPORTED otherwise).
 * FP16X3, BF16 and TRIPLE16 compute the triple products on FP16 MFMA with
 * hi/lo compensation.  A batch (tp_batch_create) keeps the mode its basis
 * had when it was made.  No counterpart in the reference (whose GPU path is FP32
 * end to end, cuda.jl:79-88). */
#define TP_PREC_FP64 0
#define TP_PREC_FP16X3 1
#define TP_PREC_BF16 2
#define TP_PREC_TRIPLE16 3
#define TP_PREC_FP16X3_T64 4
int tp_basis_set_precision(tp_basis *b, int mode);
int tp_basis_get_precision(const tp_basis *b);

/* c3raw for the coefficients of shells 1..K.
 * Replaces FullCorrelation_parallized(intensity, basis, minimal, normalize,
 * return_raw=true) (correlations.jl:179-184, cuda.jl:74-98). */
int tp_c3(tp_ctx *ctx, const tp_basis *b, const double *coeff, int64_t ld, double *c3_out);

/* Compress a reference histogram for the energy: the slices visited by the
 * K3_range = k3lo:k3hi walk of energy() (correlations.jl:211-217) are
 * normalised (q / sum|q|) and kept on the device.  Kr is the shell extent of
 * the c3ref array.  Mirrors the c3ref compression of new_energy
 * (cuda.jl:127-139).  A visited slice with sum|q| = 0 or a non-finite entry
 * has no normalised form (the reference's energy() would produce NaN): the
 * call fails with TP_E_ARG.  Raw counts go through tp_load_histograms
 * (loadHistograms' clamp) first. */
int tp_ref_create(tp_ctx *ctx, const tp_basis *b, const double *c3ref, int Kr, int k3lo, int k3hi,
                  tp_ref **out);
void tp_ref_destroy(tp_ref *r);

/* Bayes energy of host coefficients (shells 1..K used).
 * Replaces energy(intensity, basis, c3ref, K3_range, measure) with
 * measure "Bayes" (correlations.jl:201-226); measure 0 = "Bayes", any other
 * value returns 0.0 as the reference does (correlations.jl:210). */
int tp_energy(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, const double *coeff, int64_t ld,
              int measure, double *E_out);

/* tp_energy for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments, arithmetic and result as
 * tp_energy (bit for bit), but every per-evaluation device buffer comes from
 * one static array of 32 MiB per GPU that the library holds until the process
 * ends: the call makes no hipMalloc, hipFree or memset.  Bases whose scratch
 * exceeds 32 MiB (the full-size L = 18, K = 26, N = 32 basis needs about 20 MiB
 * in FP64 and 27 MiB in FP16X3) get TP_E_UNSUPPORTED.  All calls on one GPU
 * share the array: calls that overlap on one GPU (other threads or
 * contexts) race and must be serialised by the caller. */
int tp_energy_noalloc(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, const double *coeff, int64_t ld,
                      int measure, double *E_out);

/* ---- device-resident structure (the annealing state) -------------------- */

/* Upload nshell shells of coefficients (nshell >= K; the reference rotates all
 * KMAX shells, determination.jl:236).  Only even l <= L are kept on the
 * device; the rest of the caller's vector is untouched by tp_state_read. */
int tp_state_create(tp_ctx *ctx, const tp_basis *b, const double *coeff, int64_t ld, int nshell,
                    tp_state **out);
void tp_state_destroy(tp_state *s);
/* Write the device coefficients back into the caller's (ld-strided) vectors. */
int tp_state_read(tp_ctx *ctx, const tp_state *s, double *coeff, int64_t ld);
/* Energy / c3raw of the device-resident coefficients (no host transfer of
 * coefficients).  These are the per-step calls of rotate_all_at_once. */
int tp_state_energy(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, tp_state *s, double *E_out);
int tp_state_c3(tp_ctx *ctx, const tp_basis *b, tp_state *s, double *c3_out);

/* Plain Monte Carlo simulated annealing, run entirely on the device.
 * Replaces rotate_all_at_once (determination.jl:281-338) with
 * calculate_temperature (:244-261), get_MC_step (:232-241) and evaluate_step
 * (:491-515); nsteps fixed steps replace the step-size stopping rule.
 * Random numbers: Philox4x32-10 keyed by `seed`.
 * Outputs (host, may be NULL): E_trace[nsteps] (E after step i),
 * Enew_trace[nsteps] (proposal energy), acc_trace[nsteps] (1 = accepted),
 * scalars[4] = {E0, T0, E_final, T_final}.  The state holds the final
 * structure. */
typedef struct tp_anneal_params {
    int nsteps;
    int ntemp;                          /* proposals sampled for T0 (reference: 50) */
    double initial_stepsize;            /* stepsize_l = initial_stepsize*(l-1) */
    double stepsizefactor;
    double initial_temperature_factor;
    double temperature_decay;
    uint64_t seed;
} tp_anneal_params;

int tp_anneal(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, tp_state *s, const tp_anneal_params *p,
              double *E_trace, double *Enew_trace, int32_t *acc_trace, double *scalars);

/* The same annealing in resumable pieces, with the reference's stopping rule.
 * tp_anneal_status is rotate_all_at_once's `state` dict for one resolution L
 * (determination.jl:281-338): step = state["i"], E, T, stepsizes[l/2 - 1] =
 * state["stepsizes"][l] (l = 2, 4, ..., L).  The step-i proposal and
 * acceptance draw from Philox4x32-10 keyed by (seed, i), so a run continued
 * from a saved status and structure repeats the uninterrupted run bit for bit. */
#define TP_MAX_LBLOCKS 32
typedef struct tp_anneal_status {
    int64_t step;                          /* steps taken so far */
    double E, T;                           /* current energy and temperature */
    double E0, T0;                         /* at the start of the run */
    double stepsizes[TP_MAX_LBLOCKS];      /* index l/2 - 1 */
    int32_t nl;                            /* L/2 blocks in use */
    int32_t naccepted;                     /* accepted steps so far */
    int32_t stopped;                       /* 1 once stepsizes[nl-1] <= min_stepsize */
    int32_t reserved;
} tp_anneal_status;
/* A new run: initial energy, stepsize_l = initial_stepsize (l - 1) and T0 from
 * p->ntemp proposals (calculate_temperature, determination.jl:244-261, :291-312);
 * step = 0.  p->nsteps is not used. */
int tp_anneal_begin(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, tp_state *s, const tp_anneal_params *p,
                    tp_anneal_status *st);
/* Up to max_steps further steps from st (updated in place), all on the
 * device.  min_stepsize > 0 is the loop condition of rotate_all_at_once,
 * `while stepsizes[L] > minstepsize` (determination.jl:1 uses 1.5 degrees,
 * :322): no step is taken once the step size of the largest l is at or below
 * it, and st->stopped says so.  Traces (may be NULL) get one entry per step
 * taken; *steps_done (may be NULL) = their number. */
int tp_anneal_steps(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, tp_state *s, const tp_anneal_params *p,
                    tp_anneal_status *st, int max_steps, double min_stepsize, double *E_trace, double *Enew_trace,
                    int32_t *acc_trace, int32_t *steps_done);

/* Checkpoint of an annealing run: the counterpart of saveState / the
 * regular_action every 500 steps and the resume of rotation_search
 * (determination.jl:154-229, 264-278; utilities.jl:278-287), in a versioned
 * binary file of this library's own (the reference serialises Julia dicts).
 * It holds the run's parameters, its status, the basis shape it ran on and
 * nshell shells of ld complex coefficients; the file is written to path.tmp
 * and renamed, and carries a checksum that tp_checkpoint_load verifies. */
typedef struct tp_checkpoint_info {
    tp_anneal_params params;
    tp_anneal_status status;
    double min_stepsize;
    double lambda, dq;
    int32_t L, LMAX, N, K, k3lo, k3hi;
    int32_t nshell;                        /* shells stored */
    int32_t ld;                            /* complex entries per shell stored (>= LMAX^2) */
} tp_checkpoint_info;
int tp_checkpoint_save(const char *path, const tp_checkpoint_info *info, const double *coeff);
/* Reads the header into *info; with coeff != NULL also the coefficients, which
 * need room for info->nshell * info->ld complex values (at most cap_doubles
 * doubles).  TP_E_ARG for a missing, truncated, foreign or corrupted file. */
int tp_checkpoint_load(const char *path, tp_checkpoint_info *info, double *coeff, int64_t cap_doubles);

/* loadHistograms (datagen.jl:432-475) for the triplet histogram: an
 * NA x NB x Kf x Kf x Kf c3 histogram (column-major, as read from a
 * histogram file) becomes c3_out[N, 2N, K3, K3, K3]: an old N x N histogram
 * (NA == NB) is doubled along beta for k1 >= k2 >= k3 (:445-458), every entry
 * is clamped to >= 1e-30 (:460-462), the first K3 shells are kept and divided
 * by their total (:465-468).  Host code (no GPU call).  K3 <= Kf. */
int tp_load_histograms(const double *c3_full, int NA, int NB, int Kf, int K3, double *c3_out);

/* ---- batches: many evaluations in one graph launch ------------------------- */

/* Bayes energies of n independent evaluations, each with its own basis (all
 * sharing L, LMAX, N, context and precision; K may differ) and histogram
 * refs[i] (built on bases[i]).  coeff[i] has K_i shells, ld[i] >= LMAX^2.
 * Mixed precision (tp_basis_set_precision FP16X3 / BF16 on every basis): the
 * n evaluations run as one fixed sequence of batched launches, captured once
 * as a HIP graph and replayed while the same (bases, refs) are passed again.
 * FP64: the evaluations run one after another on the device.
 * The counterpart of calling energy() (correlations.jl:201-226) n times. */
int tp_energy_batch(tp_ctx *ctx, int n, const tp_basis *const *bases, const tp_ref *const *refs,
                    const double *const *coeff, const int64_t *ld, double *E_out);

/* The Bayes energy (FP64, as tp_energy with measure 0) and its gradient with
 * respect to the coefficients of the first K shells:
 *   grad[2 (k ld + s)] + i grad[2 (k ld + s) + 1] = dE/dRe c_k[s] + i dE/dIm c_k[s],
 * every entry treated as an independent complex variable (entries outside
 * even l <= L are zero).  grad_out has K x ld complex entries, the layout of
 * coeff.  No counterpart in the reference, which evaluates energy()
 * (correlations.jl:201-226) only; it is that function's chain rule through
 * the contraction and the triple products (:165-193).  Needs N <= 32, L <= 22
 * and K <= 27, and on the curved Ewald sphere L <= 19 (TP_E_UNSUPPORTED
 * otherwise). */
int tp_energy_grad(tp_ctx *ctx, const tp_basis *basis, const tp_ref *ref, const double *coeff, int64_t ld,
                   double *E_out, double *grad_out);

/* Photon-pair and photon-triplet histograms of sparse images (the data the
 * energy is scored against).  tp_hist_create ≙ the zeroed c1/c2/c3 of
 * generateHistogram (datagen.jl:363-372) with dq = qcut/K2; tp_hist_add ≙
 * histogramCorrelationsInPicture_alltoall (datagen.jl:228-282) applied to
 * npictures pictures: picture i is photons[3 offsets[i] .. 3 offsets[i+1])
 * (x, y, z per photon; offsets[0] = 0; at most 512 photons per picture);
 * tp_hist_read returns c1[K2], c2[N,K2,K2] and c3[N,2N,K3,K3,K3] in the
 * reference's (column-major) layouts, weighted as the reference weights them,
 * and optionally the integer counts behind them (n1/n2/n3, same layouts; any
 * output may be NULL).  Counts are exact 64-bit sums, so repeated or split
 * calls give identical histograms. */
typedef struct tp_hist tp_hist;
int tp_hist_create(tp_ctx *ctx, int N, int K2, int K3, double dq, tp_hist **out);
int tp_hist_add(tp_hist *h, const double *photons, const int64_t *offsets, int npictures);
int tp_hist_read(tp_hist *h, double *c1, double *c2, double *c3, uint64_t *n1, uint64_t *n2, uint64_t *n3);
void tp_hist_destroy(tp_hist *h);

/* Two-photon correlation of a structure: twoPhotons (correlations.jl:229-243)
 * on the flat detector (lambda = 0).  coeff: K2 shells of ld complex (interleaved) coefficients in the
 * seanindex layout; the l range and N are the basis's.  c2_out[N,K2,K2] in the
 * reference's (column-major) layout, c2[a,k2,k1] = sum_{l even <= L}
 * Re(c_k1,l . c_k2,l) P_l(cos alpha_a) / (4 pi), alpha_a = (a + 1/2) pi / N;
 * minimal != 0 computes only k2 <= k1 (0 elsewhere); normalize != 0 divides by
 * sum |c2| instead of by 4 pi.  One kernel launch (timed by tp_set_timing). */
int tp_two_photons(tp_ctx *ctx, const tp_basis *basis, const double *coeff, int64_t ld, int K2, int minimal,
                   int normalize, double *c2_out);
/* twoPhotons with the basis's lambda and the volume's shell spacing dq
 * (dq(volume), correlations.jl:231): the series is taken at P_l(cos
 * alpha_star) (datagen.jl:49-53), cos alpha_star = sin th1 sin th2 cos alpha +
 * cos th1 cos th2 with th_k = acos(k dq lambda / (4 pi)); needs
 * |K2 dq lambda| <= 4 pi.  With lambda = 0 it is tp_two_photons. */
int tp_two_photons_ewald(tp_ctx *ctx, const tp_basis *basis, const double *coeff, int64_t ld, int K2, int minimal,
                         int normalize, double dq, double *c2_out);

/* A start structure from a two-photon histogram: retrieveSolution
 * (correlations.jl:262-319), lambda = 0, host code (no GPU call).  c2[N,K2dim,K2dim]
 * in the reference's layout; shells k2lo..k2hi (1-based) are fitted, with
 * L < (k2hi - k2lo + 1)/2 and L <= LMAX - 1.  coeff_out: k2hi shells of ld
 * complex (interleaved) coefficients, zero outside the fitted shells and the
 * even l <= L.  Eigenvector signs, which LAPACK leaves open, are fixed so that
 * each eigenvector's entry of largest magnitude is positive.  flipped (may be
 * NULL) = 1 if the negativity check changed every sign; negativity (may be
 * NULL) = the summed shell negativity before that change. */
int tp_retrieve_solution(const double *c2, int N, int K2dim, int k2lo, int k2hi, int L, int LMAX,
                         double *coeff_out, int64_t ld, int *flipped, double *negativity);
/* retrieveSolution(c2, L, LMAX, k2lo:k2hi, qmax, lambda) with any lambda: the
 * fit's Legendre columns are taken at alpha_star of each shell pair with
 * dq = qmax / k2hi, and the slices are symmetrised only for lambda = 0
 * (correlations.jl:282-289).  With lambda = 0 it is tp_retrieve_solution. */
int tp_retrieve_solution_ewald(const double *c2, int N, int K2dim, int k2lo, int k2hi, int L, int LMAX, double qmax,
                               double lambda, double *coeff_out, int64_t ld, int *flipped, double *negativity);

/* Rotational fit of a structure to a reference: the scoring inside
 * fitStructures_full / fitStructures_random (data_processing.jl:26-79).
 * tp_fit_create takes the reference (k_hi shells of ld complex coefficients),
 * the shell range k_lo..k_hi (1-based) and a basis for L and LMAX (2 <= L <= 22,
 * LMAX <= 32); tp_fit_set_volume sets the structure to rotate (k_hi shells);
 * tp_fit_similarity writes, for each of n Euler triples (theta, phi, gamma),
 * sc_out[i] = similarity(reference, rotateStructure(volume, theta, phi, gamma,
 * k_hi, 2:2:L), k_lo:k_hi) (spherical_harmonics.jl:426-435, 598-612): the mean
 * over the shells of Pearson's correlation of the real surfaces on the
 * (2 LMAX)^2 grid of the inverse transform.  With timing on, tp_last_timings
 * gives the rotation and similarity kernel times of the last call. */
/* Synthetic sparse pictures: pointsPerOrientation (datagen.jl:88-144) for
 * npictures orientations at once.  The intensity volume is given by its
 * coefficients (K shells of ld complex, shell spacing rmax / K) and sampled
 * through its (2 LMAX)^2 surfaces with the reference's trilinear
 * interpolation (getVolumeInterpolated, spherical_harmonics.jl:484-530) and
 * envelope bound 1.3 max(surface).  Picture i uses the random streams of
 * (seed, first_picture + i): a Haar rotation (returned row-major in
 * rotations[9 i], if not NULL), photons_per_image Gaussian incident photons of
 * width envelope_sigma, one acceptance uniform per photon; with particles > 1
 * and noise, the segments of generateHistogram's loop (datagen.jl:399-407):
 * each particle's photons, then its noise photons from the isotropic noise
 * volume of get_noise_volume (datagen.jl:296-304) with the same rotation
 * (pointsPerOrientation's incident_photon_variance for the particles only;
 * beamstop_only replaces all of them by one beamstop shot, datagen.jl:408-411).
 * Accepted photons, in incident order and at their detector positions (x, y, 0), go to
 * photons[(i capacity + j) 3 ..] for j < capacity; counts[i] is their number,
 * which may exceed capacity (the excess is dropped). */
typedef struct tp_picture_params {
    double qcut, envelope_sigma, lambda, beamstop_width;
    int photons_per_image;
    int capacity;
    uint64_t seed;
    int particles;        /* number_particles (>= 1): one rotation and set of photons each */
    int noise_photons;    /* GaussianNoise photons per particle (0: no noise) */
    double noise_sigma;   /* GaussianNoise sigma: noise volume c00(k) = gaussian(k dq; sigma), envelope 1.05 sigma */
    int incident_photon_variance; /* > 0: each particle's incident count is round(Normal(photons_per_image, this)), >= 0 */
    int beamstop_only;    /* nonzero: beamstop_shot (datagen.jl:147-161) instead of the particles, identity rotation */
} tp_picture_params;
int tp_pictures(tp_ctx *ctx, int LMAX, int K, double rmax, const double *coeff, int64_t ld, const tp_picture_params *p,
                int npictures, int first_picture, double *photons, int32_t *counts, double *rotations);
/* generateHistogram's inner loop (datagen.jl:391-415) on the device: the
 * pictures of tp_pictures (capacity <= 512) histogrammed into h as
 * tp_hist_add does, without leaving the device.  A picture that accepted more
 * than capacity photons cannot be histogrammed whole and is skipped (none of
 * its photons, pairs or triplets is counted); overflow (may be NULL) = the
 * number of pictures skipped. */
int tp_hist_add_generated(tp_hist *h, int LMAX, int K, double rmax, const double *coeff, int64_t ld,
                          const tp_picture_params *p, int npictures, int first_picture, int64_t *overflow);

/* Spherical-harmonics transforms on s2kit's (2 LMAX)^2 grid (theta_j = pi (2j+1) /
 * (4 LMAX), phi_i = 2 pi i / (2 LMAX), samples theta-major), with
 * Ylm = sphPlm e^{i m phi}: tp_sh_surface = getSurfaceVolume (backtransform,
 * spherical_harmonics.jl:139-153), K shells of ld complex coefficients ->
 * surf[K][(2 LMAX)^2] complex; tp_sh_coeff = getSphericalHarmonicsVolume
 * (transform, :127-160) with Driscoll-Healy weights, exact for band-limited
 * surfaces.  tp_sph_bessel = forward (forward != 0) / backward
 * (sphBesselTransform, :310-377): every (l, m) radial profile of K shells
 * through the reference's Hankel quadrature.  LMAX <= 48. */
int tp_sh_surface(tp_ctx *ctx, int LMAX, int K, const double *coeff, int64_t ld, double *surf);
int tp_sh_coeff(tp_ctx *ctx, int LMAX, int K, const double *surf, double *coeff, int64_t ld);
int tp_sph_bessel(tp_ctx *ctx, int LMAX, int K, int forward, const double *coeff, int64_t ld, double *out);

/* Phase retrieval: sRAAR(intensity, iterations, beta0, beta_max, tau; cutoff_factor)
 * (phases.jl:31-95) on the device.  intensity: K shells of ld complex
 * coefficients; phases: the initial random phases [K][(2 LMAX)^2] in the
 * surface sample order, or NULL to draw them uniformly in [0, 2 pi) from seed.
 * coeff_out: the density's K shells of ld complex coefficients; diff_out (may
 * be NULL): the per-iteration real-space change sum|density - tmp_u| / sum|density|. */
int tp_sraar(tp_ctx *ctx, int LMAX, int K, const double *intensity, int64_t ld, int iterations, double beta0,
             double beta_max, double tau, double cutoff_factor, const double *phases, uint64_t seed,
             double *coeff_out, double *diff_out);

typedef struct tp_fit tp_fit;
int tp_fit_create(tp_ctx *ctx, const tp_basis *basis, const double *reference, int64_t ld, int k_lo, int k_hi,
                  tp_fit **out);
int tp_fit_set_volume(tp_fit *f, const double *coeff, int64_t ld);
int tp_fit_similarity(tp_fit *f, int n, const double *euler, double *sc_out);
void tp_fit_destroy(tp_fit *f);

#define TP_MOVE_QUAT 0    /* random slots: an SO(3) rotation per even l (uniform axis, Haar angle) */
#define TP_MOVE_PLANE 1   /* random slots: the reference's move, a rotation in the plane of a
                           * Haar-random 2-frame of R^(2l+1) (default) */

/* Concurrent annealing runs ("groups"), each with its own histogram refs[g]
 * (built on any basis sharing b's L, LMAX, N), shell count K_g = the upper end
 * of refs[g]'s K3_range (<= b's K), start coefficients coeff[g] (K_g shells)
 * and seed.  Every sweep proposes nprop (>= 3)
 * rotations per run, each acting on the l-blocks (even l >= 2) of all shells
 * in FP64, which keeps every two-photon correlation sum_m c_k1,lm^* c_k2,lm
 * as the reference's moves do (get_MC_step, determination.jl:232-241):
 * TP_MOVE_PLANE (default) the reference's move, a rotation by
 * b = amax (pi/4)(l-1) z (z normal, clipped to |b| <= pi/4) in the plane of
 * a Haar-random 2-frame of R^(2l+1); TP_MOVE_QUAT an SO(3) rotation per l
 * (uniform axis, Haar-distributed angle <= amax pi/4) through the real
 * Wigner-D recurrence.  All ngroups * nprop
 * evaluations of a sweep, and the per-run decision, are one replay of a
 * captured HIP graph.  Decision per run: the lowest-energy proposal is
 * accepted by the Metropolis rule on its exact energy difference
 * (evaluate_step, determination.jl:491-515); T decays by `decay` per sweep;
 * amax grows by `factor` (capped at 1) on acceptance and shrinks otherwise.
 * b must be in mixed precision.  hist_cap sweeps of per-run energies are kept. */
int tp_batch_create(tp_ctx *ctx, const tp_basis *b, int ngroups, int nprop, const tp_ref *const *refs,
                    const double *const *coeff, const int64_t *ld, const uint64_t *seeds, int hist_cap,
                    tp_batch **out);
void tp_batch_destroy(tp_batch *B);
/* Time the candidate triple-product work orders and contraction kernels on
 * this batch (reps launches each) and keep the fastest.  ms_out[14] (may be
 * NULL): work orders 0 (by cost), 1 (evaluation-major; in TP_PREC_FP16X3_T64
 * the sorted l-triple items of each evaluation on one XCD, one evaluation
 * after another), then the contraction candidates.  Dense form (FP16X3, BF16): tiles 128x64, 64x64,
 * 128x128, 64x128 with parallel row tiles, the same four with row-walking
 * workgroups, and the same four with the row tiles partitioned over the XCDs
 * (1e30 where a shape does not divide by 8); config_out[4] (may be NULL):
 * chosen tile rows, tile columns, work order, schedule (1 parallel, 2
 * row-walking, 3 XCD-partitioned rows).  Factored form (FP16X3_T64):
 * ms_out[2..5] for the G = H.As kernel with all class rows x 64 columns, all
 * rows x 32 columns, 3 row blocks x 64 columns and 4 row blocks x 16
 * columns per workgroup (each with the separated-grid stage that follows;
 * ms_out[6..13] = 1e30); config_out: class rows and columns per workgroup,
 * work order, 4 + candidate. */
int tp_batch_autotune(tp_batch *B, int reps, double *ms_out, int *config_out);
/* Initial energies and temperatures: slot 0 of every run is the identity, T0 =
 * temperature_factor * std(energies of the other slots) (calculate_temperature,
 * determination.jl:244-261). */
int tp_batch_probe(tp_batch *B, double temperature_factor);
/* Move family of the random slots (TP_MOVE_PLANE or TP_MOVE_QUAT). */
int tp_batch_set_family(tp_batch *B, int family);
/* Angle bound of every run's random proposals, as a fraction of pi/4
 * (0 < amax <= 1; 1 at creation); it then adapts by `factor` per sweep. */
int tp_batch_set_amax(tp_batch *B, double amax);
/* Use contraction candidate `cand` (0..11, 0..3 in TP_PREC_FP16X3_T64; the order of tp_batch_autotune's
 * ms_out[2..13]) instead of the tuned one, re-capturing the sweep graph. */
int tp_batch_set_contraction(tp_batch *B, int cand);
/* nsweeps sweeps (graph replays). */
int tp_batch_run(tp_batch *B, int nsweeps, double decay, double factor);
/* Per-run state: energy, temperature, angle bound, accepted sweeps, sweeps done. */
int tp_batch_group(tp_batch *B, int g, double *E, double *T, double *amax, int32_t *nacc, int32_t *nsweep);
/* Current coefficients of run g written into the caller's vectors (even l <= L
 * entries of its K_g shells; the rest untouched). */
int tp_batch_coeff(tp_batch *B, int g, double *coeff, int64_t ld);
/* Energies after each sweep, [min(n, hist_cap)][ngroups]. */
int tp_batch_history(tp_batch *B, double *Ehist, int n);
/* Evaluate given rotations of the current states: quat[ngroups*nprop][L/2][4]
 * (w,x,y,z per even l >= 2; slot e = g*nprop + p), energies to E_out[ngroups*nprop].
 * No decision is taken. */
int tp_batch_eval_quats(tp_batch *B, const double *quat, double *E_out);
/* Evaluate given rotations of the current states (no decision taken), each
 * even l >= 2 block of all K_g shells turned by a real rotation D_l acting on
 * x = Umat(l) c (spherical_harmonics.jl:445-457), as get_MC_step applies its
 * moves (determination.jl:232-241).  Slot e = g*nprop + p:
 *   p < ngeo: D_l = exp(-theta[e][l/2-1] G_l) with G_l run g's real
 *     skew-symmetric generator of block l ((2l+1)^2 doubles, row-major,
 *     Frobenius norm <= 1), concatenated over l = 2, 4, .., L per run (gen
 *     [ngroups][sum_l (2l+1)^2]): a rotation along the one-parameter subgroup
 *     of G_l (e.g. the normalised Riemannian gradient of the energy on
 *     SO(2l+1)); computed on the device to ~1e-16 (scaling and squaring of
 *     the degree-12 Taylor polynomial); theta finite;
 *   p >= ngeo: D_l = I + (cos b - 1)(u u^T + v v^T) + sin b (v u^T - u v^T)
 *     from uvb[e][l/2-1] = (u[2L+1], v[2L+1], b) (u, v orthonormal, first
 *     2l+1 entries used): the reference's plane rotation (utilities.jl:180-190).
 * Energies (the batch's precision) -> E_out[ngroups*nprop].  No counterpart
 * in the reference; L <= 22. */
int tp_batch_eval_rotations(tp_batch *B, int ngeo, const double *gen, const double *theta, const double *uvb,
                            double *E_out);
/* Replace run g's current structure (its K_g shells, even l <= L entries of
 * coeff) and its current energy, e.g. after a decision taken by the caller. */
int tp_batch_set_state(tp_batch *B, int g, const double *coeff, int64_t ld, double E);

/* Rotate the even l >= 2 blocks of nshell shells by quat[L/2][4] (one unit
 * quaternion per l, w,x,y,z) in FP64, through the real Wigner-D recurrence:
 * the rotated coefficients describe f(R^-1 r).  Entries other than even l <= L
 * are copied.  One quaternion for every l rotates the whole intensity, which
 * leaves c3 unchanged (its SO(3) invariance); rotateStructure
 * (spherical_harmonics.jl:598-612) is the reference's whole-structure form. */
int tp_rotate(tp_ctx *ctx, const tp_basis *b, const double *coeff, int64_t ld, int nshell, const double *quat,
              double *out);

/* ---- learned rotation proposer --------------------------------------------- */
/* No counterpart in the reference, whose moves come only from
 * random_rotation_step (utilities.jl:180-190).  An MLP 192 -> 256 -> 256 -> 256
 * (ReLU) -> 64 in BF16 with FP32 accumulation and FP32 master weights / Adam:
 * input = the 190 real-basis coefficients of one shell (even l <= 18, scaled
 * to norm 8), log10 T, 0; output rows 0..35 = one quaternion (w,x,y,z) per
 * even l = 2..18 (rows 36..63 are held at zero).  Parameters flat, as
 * W1 [256][192], b1 [256], W2 [256][256], b2, W3 [256][256], b3, W4 [64][256],
 * b4 [64] (tp_mlp_nparam() floats).  Training minimises
 * (1/B) sum_b sum_{j<36} (y_bj - t_bj)^2 with Adam (0.9, 0.999, 1e-8) on
 * minibatches drawn with replacement (Philox keyed by seed). */
int tp_mlp_nparam(void);
int tp_mlp_create(tp_ctx *ctx, uint64_t seed, tp_mlp **out);
void tp_mlp_destroy(tp_mlp *m);
int tp_mlp_get_params(tp_mlp *m, float *out);
/* Also resets the Adam state. */
int tp_mlp_set_params(tp_mlp *m, const float *in);
/* Y[n][64] for inputs X[n][192]. */
int tp_mlp_forward(tp_mlp *m, const float *X, int n, float *Y);
/* Loss and its gradient (FP32, flat like the parameters) on n rows (n a
 * multiple of 64); nothing is updated. */
int tp_mlp_loss_grad(tp_mlp *m, const float *X, const float *T, int n, double *loss, float *grad);
/* steps Adam steps on host data (X[n][192], T[n][36]); loss_trace[steps] may be NULL. */
int tp_mlp_train(tp_mlp *m, const float *X, const float *T, int n, int steps, int batch, float lr, uint64_t seed,
                 float *loss_trace);
/* The same on a batch's move log (the data never leaves the device). */
int tp_mlp_train_log(tp_mlp *m, tp_batch *B, int steps, int batch, float lr, uint64_t seed, float *loss_trace);

/* Keep a move log of up to `capacity` samples: after every sweep, each
 * proposal that passes the Metropolis test on its own against its run's current energy
 * adds one sample per shell of the run (that shell's features in the state the
 * move starts from, and the proposal's quaternions with w >= 0; a plane move
 * is recorded as (cos |b|/2, sin |b|/2, 0, 0) of its angle b per l). */
int tp_batch_log(tp_batch *B, int capacity);
int tp_batch_log_count(tp_batch *B, int32_t *n);
int tp_batch_log_read(tp_batch *B, int first, int n, float *X, float *T);
/* The last nlearn (< nprop) slots of every run take the network's quaternions
 * (input: shell j*K_g/nlearn of the run's state and its T), each composed with
 * a random rotation of angle <= noise * amax * pi/4; the other slots stay
 * random.  With TP_MOVE_QUAT the quaternion is the SO(3) rotation of block l;
 * with TP_MOVE_PLANE its angle is the angle of a plane rotation of block l in
 * a Haar-random plane (the reference's move with a learned step).  Acceptance
 * is unchanged (exact energies).  nlearn = 0 switches the network off. */
int tp_batch_set_proposer(tp_batch *B, tp_mlp *m, int nlearn, double noise);

/* ---- synthetic input ------------------------------------------------------ */

/* Spherical-harmonic coefficients of the intensity of point atoms with unit
 * form factor, I(q) = |sum_j exp(i q.r_j)|^2, on shells q_k = k*dq,
 * k = 1..KMAX, by Gauss-Legendre (cos theta) x uniform-phi quadrature with
 * ntheta x nphi points; entries with l <= lfill are written, the rest zeroed.
 * Counterpart of getIntensity_cartesian / createSphericalHarmonicsStructure
 * (structure.jl:74-134) for point atoms. */
int tp_intensity_from_atoms(tp_ctx *ctx, const double *atoms_xyz, int natoms, int KMAX, double dq, int LMAX,
                            int lfill, int ntheta, int nphi, double *coeff_out);

/* Electron density of Gaussian atoms, getDensity_cartesian (structure.jl:74-87),
 * sampled as getDensity does (spherical_harmonics.jl:234-244): shells
 * r_k = k rmax / K, k = 1..K, at the (2 LMAX)^2 points of i_to_ang (:203-210).
 * atoms: natoms rows (x, y, z, R, B), R and B the atom type's form factor
 * (the reference's atomfactorlist; atoms of type "1" left out by the caller);
 * rho(p) = sum over atoms with |p - a|^2 < 5 of B / (R sqrt(2 pi))^3
 * exp(-|p - a|^2 / (2 R^2)).  surf: [K][(2 LMAX)^2] complex (imaginary 0). */
int tp_density_surface(tp_ctx *ctx, const double *atoms, int natoms, int LMAX, int K, double rmax, double *surf);
/* absoluteSquare (spherical_harmonics.jl:393-397): per shell, the coefficients
 * of |f|^2 with f the surface of the shell's coefficients.  LMAX <= 48. */
int tp_absolute_square(tp_ctx *ctx, int LMAX, int K, const double *coeff, int64_t ld, double *out);
/* createSphericalHarmonicsStructure (structure.jl:124-134) after loadPDB, on
 * the device: density = getDensity (sampled as tp_density_surface, then
 * transformed), fourier = forward(density) (sphBesselTransform), intensity =
 * absoluteSquare(fourier); each K shells of ld complex coefficients, any of
 * them may be NULL.  LMAX <= 48, K >= 2. */
int tp_structure_from_atoms(tp_ctx *ctx, const double *atoms, int natoms, int LMAX, int K, double rmax,
                            double *density, double *fourier, double *intensity, int64_t ld);

/* ---- instrumentation ------------------------------------------------------ */

/* Per-stage device time (ms) of the last tp_energy / tp_state_energy call,
 * measured with HIP events on the context stream: [0] triple products,
 * [1] G = H.As contraction, [2] c3 = Bm.G contraction with the energy
 * partial sums, [3] energy finish.  Enabled by tp_set_timing(ctx, 1). */
int tp_set_timing(tp_ctx *ctx, int on);
int tp_last_timings(const tp_ctx *ctx, double *ms_out, int n);

#ifdef __cplusplus
}
#endif

#endif /* TP_H */
