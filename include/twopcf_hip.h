/* twopcf_hip.h — C ABI of libtwopcf_hip, the MI355X (gfx950) pair-counting
 * library behind two_pcf's correlator entry points.
 *
 * The reference has no FFI; its seam is the correlator layer
 * (correlators.h:7-31, correlators_invpweight.h:9-35):
 *
 *   void auto_pair_count_{monopole,sigma_pi,s_mu}(const GalaxyBoxes*, JKHists[2D]*);
 *   void cross_pair_count_{monopole,sigma_pi,s_mu}(const GalaxyBoxes*, const GalaxyBoxes*, JKHists[2D]*);
 *   ... and the *_invpweights / *_angular variants.
 *
 * twopcf_count() replaces the body of those functions: it takes the two
 * catalogues (uploaded once with twopcf_catalog_upload, which replaces the
 * GalaxyBoxes construction of boxes.cpp:10-163), the histogram axes
 * (Axis, hist.cpp:15-36) and the number of jackknife regions, and returns the
 * RAW sums: full[b] = sum of pair weights in bin b, and jk[r][b] = sum over
 * pairs with at least one member in region r.  The reference's
 * invert_jk_hists / multiply_hists(2.0) (hist.cpp:464-482) and normalise()
 * run on the host afterwards, unchanged (see two_pcf_amd/csrc/host/).
 *
 * Conventions: every function returns 0 on success and a negative
 * TWOPCF_E_* code on failure; twopcf_last_error() gives the message.  No C++
 * exception crosses this boundary.  All pointers passed in are host memory
 * owned by the caller; device memory is owned by the library.  One context
 * owns one HIP stream; a context is not thread-safe, distinct contexts are
 * independent.
 */
#ifndef TWOPCF_HIP_H
#define TWOPCF_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TWOPCF_ABI_VERSION 5

enum {
	TWOPCF_OK = 0,
	TWOPCF_E_ARG = -1,      /* invalid argument (null pointer, bad size, region out of range) */
	TWOPCF_E_HIP = -2,      /* a HIP runtime call failed (message carries hipGetErrorString) */
	TWOPCF_E_NODEV = -3,    /* no usable GPU */
	TWOPCF_E_LIMIT = -4,    /* a size limit of this build was exceeded */
	TWOPCF_E_AXIS = -5      /* axis is not monotone under the reference's float pipeline */
};

/* Correlator kinds (correlators.h / correlators_invpweight.h). */
enum {
	TWOPCF_MONOPOLE = 0,   /* s                     correlators.cpp:11-125   */
	TWOPCF_SIGMA_PI = 1,   /* sigma x pi            correlators.cpp:156-278  */
	TWOPCF_S_MU = 2,       /* s x mu                correlators.cpp:309-430  */
	TWOPCF_ANGULAR = 3     /* 1 - cos(theta)        correlators_invpweight.cpp:713-983 */
};

/* Counting methods. */
enum {
	/* The drop-in, and the check for the batched method: one launch per
	 * count, the reference's cell size (rmax) and 27-cell neighbourhood, and
	 * per pair the reference's own mixed FP64/FP32 arithmetic
	 * (correlators.cpp:309-343) evaluated exactly. */
	TWOPCF_PLAIN_FP64 = 0,
	/* Batched: the jobs of the call in persistent launches (one per
	 * correlator kind, weighting, line of sight and LDS use) over a
	 * cell-pair work queue; cells subdivided; pair separations in FP32 from
	 * cell-local offsets; pairs within the proven FP32 error band of a range
	 * limit or bin edge are re-evaluated with the exact pipeline, so counts
	 * are identical to TWOPCF_PLAIN_FP64.  A monopole job over the same
	 * catalogues, weights and s minimum as an s-mu job of the call, neither
	 * with regions, is binned in the s-mu job's pass; its outputs are its
	 * own. */
	TWOPCF_FP32_RECHECK = 1
};

/* Line-of-sight definition of pi (sigma-pi and s-mu). */
enum {
	TWOPCF_PI_ANGULAR = 0,   /* (d . (n1 + n2)) / 2, _ANGULAR_PI_DEFN (makefile:4, the default build) correlators.cpp:319-323 */
	TWOPCF_PI_MIDPOINT = 1   /* d . (x1 + x2) / |x1 + x2|, builds without it          correlators.cpp:325-336 */
};

/* Pair-weight rules (correlators.cpp:340, correlators_invpweight.cpp). */
enum {
	TWOPCF_W_PLAIN = 0,     /* w1*w2 (float product)                         correlators.cpp:340 */
	TWOPCF_W_INVP_3D = 1,   /* (iw1*iw2) * float(up(1-cos)*n_runs/popc)      correlators_invpweight.cpp:507-533 */
	TWOPCF_W_ANGUP_3D = 2,  /* float(float(w1*w2) * up(1-cos))               correlators_invpweight.cpp:574 */
	TWOPCF_W_ANG_IW = 3,    /* iw1*iw2                                       correlators_invpweight.cpp:901,968 */
	TWOPCF_W_ANG_INVP = 4,  /* (iw1*iw2) * ((float)n_runs / popc)            correlators_invpweight.cpp:743-764 */
	TWOPCF_W_ANG_W = 5      /* w1*w2 (angular DR with inverse-p weights)     correlators_invpweight.cpp:853 */
};

typedef struct twopcf_ctx twopcf_ctx;
typedef struct twopcf_cat twopcf_cat;

/* One catalogue, structure of arrays, host memory.  Pass only the objects
 * the reference keeps in front of n_gals (d_comov >= 0, boxes.cpp:16-26).
 * Mirrors Galaxy (galaxy.h:16-35). */
typedef struct {
	int64_t n;
	const double *x, *y, *z;        /* Galaxy::x                                  */
	const double *nx, *ny, *nz;     /* Galaxy::x_norm                             */
	const float *w;                 /* Galaxy::weight        (NULL: all 1)        */
	const int32_t *jk;              /* Galaxy::jk_region     (NULL: all 0)        */
	const float *iw;                /* Galaxy::input_weight  (NULL: same as w)    */
	const uint64_t *masks;          /* Galaxy::bitmask, n_words per object, object-major (NULL: none) */
	int32_t n_words;
	const uint64_t *dither;         /* Galaxy::dither_mask (_USE_DITHER_WEIGHTS, galaxy.h:28-30),
	                                   n_words per object, object-major (NULL: none) */
} twopcf_cat_desc;

/* Axis(min, max, n_bins, log_base) — hist.cpp:15-36.  Values are given as
 * the parameter file holds them (double); the library rounds them to float
 * exactly as the Hist/Hist2D constructors do. */
typedef struct {
	double min, max;
	int32_t n_bins;
	double log_base;
} twopcf_axis;

/* One histogram to fill (one call of an auto_/cross_pair_count_* function). */
typedef struct {
	int32_t kind;                   /* TWOPCF_MONOPOLE ... TWOPCF_ANGULAR               */
	const twopcf_cat *a;            /* first catalogue (cat_1)                          */
	const twopcf_cat *b;            /* second catalogue, NULL for an auto count         */
	twopcf_axis ax[2];              /* ax[1] used by sigma-pi (pi) and s-mu (mu)        */
	int32_t n_jk;                   /* JKHists::n_jk_regions (>= 1)                     */
	int32_t weight_mode;            /* TWOPCF_W_*                                       */
	int32_t n_runs;                 /* Params::n_bitwise_runs (inverse-p modes)         */
	const double *up_bins;          /* angular up-weight DD/DDinvp per bin (modes 1,2)  */
	twopcf_axis up_axis;            /* its axis                                         */
	int32_t pi_def;                 /* TWOPCF_PI_* (0: the reference's default build)   */
	int32_t dither;                 /* modes 1 and 4: numerator popc(dither1 & dither2), 0 -> 1,
	                                   instead of n_runs (_USE_DITHER_WEIGHTS,
	                                   correlators_invpweight.cpp:62-70, 749-757)       */
	/* outputs, caller-allocated host memory */
	double *full;                   /* [nb]        raw sum per bin                      */
	double *jk;                     /* [n_jk * nb] raw sum over pairs touching region r (may be NULL) */
	uint64_t *n_pairs;              /* accepted pairs (may be NULL)                     */
	uint64_t *n_rechecked;          /* pairs re-evaluated exactly (may be NULL)         */
} twopcf_job;

typedef struct {
	double kernel_ms;        /* HIP-event time of the counting kernel(s) on the context stream */
	double prep_ms;          /* gridding + planning (host wall clock)                            */
	double total_ms;         /* whole call (host wall clock)                                     */
	int32_t n_launches;
	int32_t n_work_items;
	uint64_t n_pairs;        /* accepted pairs over all jobs   */
	uint64_t n_rechecked;    /* exact re-evaluations           */
	uint64_t n_staged;       /* objects the work items read from the catalogues: each item's
	                            first-side objects and its stencil rows' second-side objects */
} twopcf_timing;

/* Tuning knobs of TWOPCF_FP32_RECHECK.  A zero field takes the default
 * (2, 12, 256): as many of the 12 region slots as fit in LDS.  A NULL tuning takes the choice twopcf_autotune made for the
 * same job set, or the defaults.  No knob changes a result. */
typedef struct {
	int32_t cell_subdiv;     /* cells per rmax along each axis (1..3; larger values mean 3) */
	int32_t max_slots;       /* jackknife region slots per work item (1..16) */
	int32_t i_chunk;         /* first-side objects per work item (64..256, multiple of 64) */
} twopcf_tuning;

int twopcf_abi_version(void);
int twopcf_ctx_create(int device, twopcf_ctx **out);
int twopcf_ctx_destroy(twopcf_ctx *ctx);
const char *twopcf_last_error(const twopcf_ctx *ctx);
int twopcf_device_info(twopcf_ctx *ctx, char *name, int name_len, int *n_cu, int *lds_per_block);

int twopcf_catalog_upload(twopcf_ctx *ctx, const twopcf_cat_desc *desc, twopcf_cat **out);
int twopcf_catalog_free(twopcf_cat *cat);
int64_t twopcf_catalog_size(const twopcf_cat *cat);

/* Fill every job.  method: TWOPCF_PLAIN_FP64 or TWOPCF_FP32_RECHECK.
 * tuning may be NULL.  timing may be NULL. */
int twopcf_count(twopcf_ctx *ctx, const twopcf_job *jobs, int n_jobs, int method,
                 const twopcf_tuning *tuning, twopcf_timing *timing);

/* twopcf_count without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same jobs,
 * methods, tuning and outputs as twopcf_count, but it calls no hipMalloc,
 * hipFree or allocating library routine: every per-call device buffer (job
 * table, work items, stencil rows, region slots, tables, up-weight tables,
 * output pools, counters) is carved from a library-owned device array of
 * TWOPCF_NOALLOC_SCRATCH_MIB MiB, one per GPU, held until the process ends.
 * A job set that needs more returns TWOPCF_E_LIMIT.  The catalogues must
 * already be gridded for the job set's cell geometry: call twopcf_count once
 * with the same jobs and tuning first (otherwise TWOPCF_E_ARG).  No graph is
 * captured or replayed.  The outputs are cleared on the device at every
 * call; calls that overlap on one GPU (from any context or thread) share the
 * array and race. */
#define TWOPCF_NOALLOC_SCRATCH_MIB 64
int twopcf_count_noalloc(twopcf_ctx *ctx, const twopcf_job *jobs, int n_jobs, int method,
                         const twopcf_tuning *tuning, twopcf_timing *timing);

/* Autotuner of the batched method (no counterpart in the reference, whose
 * cell size is fixed at rmax): times trial runs of the job set on a sample of
 * its work queue, choosing cell subdivision, region slots and first-side
 * chunk size, and records the choice in the context for this job set (the
 * same catalogues, axes, regions and weight rule), where twopcf_count uses it
 * when called with tuning == NULL.  Writes the choice to *best and the
 * estimated full-batch kernel time to *best_ms (either may be NULL).  The
 * jobs' output buffers are not written. */
int twopcf_autotune(twopcf_ctx *ctx, const twopcf_job *jobs, int n_jobs,
                    twopcf_tuning *best, double *best_ms);

/* Host-only helpers (no GPU needed). ------------------------------------ */

/* The s^2 thresholds the kernels bin by: for a monopole/s-mu s axis (or the
 * sigma-pi range), t[0..n] such that a pair with float s^2 in [t[k], t[k+1])
 * lands in s-bin k and pairs outside [t[0], t[n]) are rejected — the exact
 * image of correlators.cpp:315-317 + hist.cpp:46-53, 486-490 under glibc
 * log1pf.  kind selects the range rule (correlators.cpp:13-14, 192-195). */
int twopcf_s2_thresholds(int kind, const twopcf_axis *ax1, const twopcf_axis *ax2, float *t);
/* Value-domain thresholds of one axis: bin(v) = k  iff  t[k] <= v < t[k+1]. */
int twopcf_axis_thresholds(const twopcf_axis *ax, float *t);

/* Comoving-distance table + lookup and Galaxy::set, as the reference's
 * readers apply them (cosmology.cpp:11-100, dataset.cpp:285-300,
 * galaxy.cpp:8-27).  eq_to_cart writes positions and unit vectors. */
int twopcf_comoving_distance(double omega_m, double h, double z_min, double z_max,
                             int64_t n, const double *z, double *d);
int twopcf_eq_to_cart(int64_t n, const double *ra, const double *dec, const double *d,
                      double *x, double *y, double *z, double *nx, double *ny, double *nz);

#ifdef __cplusplus
}
#endif
#endif
