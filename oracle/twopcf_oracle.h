/* twopcf_oracle.h — CPU restatement of the two_pcf pair-counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the CPU baseline.  The product (two_pcf_amd / libtwopcf_hip) never links,
 * loads or calls it.
 *
 * Every function restates one piece of the reference and cites it.  How each
 * piece is pinned (golden vectors from the reference binary, brute force,
 * closed forms) is listed at the top of twopcf_oracle.c.
 */
#ifndef TWOPCF_ORACLE_H
#define TWOPCF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Axis: hist.h:5-19.  All constants are float, as in the reference. */
typedef struct {
	float min, minus_min, max, step, inv_step, base, ilb, ils;
	int n_bins;
	int log_binning;
} or_axis;

/* One catalogue in structure-of-arrays form.  Only objects with
 * d_comov >= 0 are passed (the reference moves the others behind n_gals,
 * boxes.cpp:16-26). */
typedef struct {
	long n;
	const double *x, *y, *z;     /* Galaxy::x */
	const double *nx, *ny, *nz;  /* Galaxy::x_norm */
	const float *w;              /* Galaxy::weight */
	const int *jk;               /* Galaxy::jk_region */
	const float *iw;             /* Galaxy::input_weight (inverse-p builds), may be NULL */
	const uint64_t *masks;       /* Galaxy::bitmask, n_words per object, may be NULL */
	int n_words;
	const uint64_t *dither;      /* Galaxy::dither_mask (_USE_DITHER_WEIGHTS), n_words per object, may be NULL */
} or_cat;

/* Pair-weight rules of the inverse-p build (correlators_invpweight.cpp). */
enum {
	OR_W_PLAIN = 0,        /* w1*w2                                   correlators.cpp:340 */
	OR_W_INVP_3D = 1,      /* (iw1*iw2)*float(interp*n_runs/pc)       correlators_invpweight.cpp:507-533 */
	OR_W_ANGUP_3D = 2,     /* float(float(w1*w2)*interp)              correlators_invpweight.cpp:569-574 */
	OR_W_ANG_IW = 3,       /* iw1*iw2 (angular, parent)               correlators_invpweight.cpp:901, 968 */
	OR_W_ANG_INVP = 4,     /* (iw1*iw2)*(float)n_runs/pc              correlators_invpweight.cpp:743-764 */
	OR_W_ANG_W = 5         /* w1*w2 (angular DR with inverse-p)       correlators_invpweight.cpp:853 */
};

/* Line-of-sight definition of pi: the angular one (_ANGULAR_PI_DEFN,
 * makefile:4, the reference's default build) or the pair midpoint. */
enum { OR_PI_ANGULAR = 0, OR_PI_MIDPOINT = 1 };

typedef struct {
	int mode;
	int n_runs;
	const double *up_bins;   /* angular up-weight table (JKHists::full_hist.bins of DD/DDinvp) */
	or_axis up_axis;
	int pi_def;              /* OR_PI_* */
	int dither;              /* _USE_DITHER_WEIGHTS: the numerator of OR_W_INVP_3D / OR_W_ANG_INVP is the
	                            popcount of dither1 & dither2 (0 counts as 1) instead of n_runs,
	                            correlators_invpweight.cpp:62-70, 749-757 */
} or_wopt;

enum { OR_MONO = 0, OR_SIGMA_PI = 1, OR_S_MU = 2, OR_ANGULAR = 3 };

void or_axis_init(or_axis *a, float mn, float mx, int n, float base);
float or_axis_bin_num(const or_axis *a, float v);
double or_axis_value_from_bin(const or_axis *a, double b);
double or_axis_bin_centre(const or_axis *a, int b);
double or_axis_bin_width(const or_axis *a, int b);
double or_interp(const or_axis *a, const double *bins, double value);

/* Bin one pair under the reference's float pipeline.  Returns 1 and the
 * flattened bin if the pair lands in the histogram, else 0. */
int or_pair_bin(int kind, const double *x1, const double *n1, const double *x2, const double *n2,
                float min_r_sq, float max_r_sq, const or_axis *ax1, const or_axis *ax2, int *bin);
/* The same with the line-of-sight definition chosen (OR_PI_*). */
int or_pair_bin_pi(int kind, const double *x1, const double *n1, const double *x2, const double *n2,
                   float min_r_sq, float max_r_sq, const or_axis *ax1, const or_axis *ax2, int pi_def, int *bin);
void or_range_sq(int kind, const or_axis *ax1, const or_axis *ax2, float *min_r_sq, float *max_r_sq);

/* Raw pair sums: full[nb] and sub[n_jk*nb] = sum over pairs touching each
 * region, before invert_jk_hists / multiply_hists.  b == NULL means an auto
 * count.  method 0 = brute force over all pairs, 1 = cell lists. */
int or_count(int kind, const or_cat *a, const or_cat *b, const or_axis *ax1, const or_axis *ax2,
             int n_jk, const or_wopt *wo, int method, int nthreads,
             double *full, double *sub, unsigned long long *n_accepted);

/* Same as or_count with method 1 but restricted to the first-catalogue cells
 * whose index is selected by a seeded Bernoulli(frac) draw (a bounded sample
 * of the cell-pair work); *frac_out returns the fraction of candidate pairs
 * covered. */
int or_count_sample(int kind, const or_cat *a, const or_cat *b, const or_axis *ax1, const or_axis *ax2,
                    int n_jk, int nthreads, double frac, unsigned long long seed,
                    double *full, double *sub, unsigned long long *n_accepted, double *frac_out);

/* invert_jk_hists + multiply_hists(2.0) for auto counts (hist.cpp:464-482). */
void or_finish(int is_auto, long nb, int n_jk, double *full, double *sub);
void or_normalise_auto(long nb, int n_jk, double *full, double *sub, const float *w, const int *jk, long n);
void or_normalise_cross(long nb, int n_jk, double *full, double *sub,
                        const float *w1, const int *jk1, long n1, const float *w2, const int *jk2, long n2);
double or_ls(double dd, double dr, double rr);

/* Cosmology table and lookup (cosmology.cpp:11-100, dataset.cpp:254-300). */
int or_cosmo_len(double z_min, double z_max, double z_save_step);
void or_cosmo_table(double omega_m, double h, double z_min, double z_max, double z_save_step,
                    int z_oversample, double *y);
double or_comov_dist(const double *y, int len, double z_min, double z_save_step, double z);
/* Galaxy::set (galaxy.cpp:8-27) over arrays. */
void or_eq_to_cart(long n, const double *ra, const double *dec, const double *d,
                   double *x, double *y, double *z, double *nx, double *ny, double *nz);
/* Cartesian reader geometry (io.cpp:393-399). */
void or_cart_norm(long n, const double *x, const double *y, const double *z,
                  double *d, double *nx, double *ny, double *nz);

/* Exhaustive checks of threshold tables against the float pipeline
 * (non-negative float ranges [lo, hi]); return the number of mismatches. */
long or_scan_s2_table(int kind, const or_axis *ax1, const or_axis *ax2, const float *t,
                      float lo, float hi, int nthreads);
long or_scan_value_table(const or_axis *a, const float *t, float lo, float hi, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
