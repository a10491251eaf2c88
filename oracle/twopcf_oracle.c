/* twopcf_oracle.c — CPU restatement of the two_pcf pair-counting path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker and the CPU baseline.  The product
 * never links, loads or calls this file.
 *
 * What it restates (reference file:line) and what pins each part:
 *
 *   or_axis_init / or_axis_bin_num / value_from_bin   hist.cpp:15-62
 *       pinned: Axis constants compared bit-for-bit with the reference
 *       binary's (the golden JSON "axis1"/"axis2" blocks); the log index is
 *       glibc log1pf exactly as the reference calls it.
 *   or_pair_bin (monopole, sigma-pi, s-mu)           correlators.cpp:31-41,
 *       156-188, 309-343, hist.cpp:246-260, 484-502
 *       pinned: golden raw histograms written by the reference binary
 *       (exact integer counts and exactly-representable weighted sums) on
 *       seeded catalogues including duplicates, lattice points on bin edges,
 *       radial (mu = 1) pairs and log axes with min > 0; plus brute force.
 *   or_pair_bin_pi, pair-midpoint line of sight       correlators.cpp:172-183,
 *       325-336 (builds without _ANGULAR_PI_DEFN) — pinned: golden vectors
 *       of the reference compiled without that macro (oracle/_ref/
 *       ref_counts_mid), sigma-pi and s-mu.
 *   or_count (cells vs brute force)                   correlators.cpp:11-125,
 *       190-278, 345-430 — pinned: both methods agree bit-for-bit and agree
 *       with the golden vectors; sum over bins == accepted pair count.
 *   or_finish / or_normalise_* / or_ls                hist.cpp:270-367,
 *       464-589, io.cpp:643-651 — pinned: golden "norm" blocks.
 *   cosmology / eq_to_cart / cart_norm                cosmology.cpp:7-100,
 *       dataset.cpp:254-300, galaxy.cpp:8-27, io.cpp:393-399
 *       pinned: golden catalogue dumps (positions and unit vectors as hex).
 *   inverse-p pair weights and angular counts         correlators_invpweight.cpp
 *       :476-577, 713-983; hist.cpp:369-391 — parity unpinned against the
 *       reference binary (its inverse-p build needs libhdf5); pinned instead
 *       by brute-force popcount (Python int.bit_count) and closed forms in
 *       tests/test_oracle_invp.py.
 *   dither-weight numerator (_USE_DITHER_WEIGHTS)     correlators_invpweight.cpp
 *       :62-70, 749-757 — parity unpinned against the reference binary (same
 *       libhdf5 reason); pinned by the same Python brute force, with dither
 *       masks whose pair popcounts include 0 (the 0 -> 1 rule).
 *
 * Compiled with -ffp-contract=off and no fast-math so that every float
 * operation rounds exactly as in the reference's x86-64 g++ build.
 */
#define _GNU_SOURCE
#include "twopcf_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- Axis -- */

/* hist.cpp:15-36.  log(float) in the reference resolves to the float
 * overload (logf); pow(float,int) promotes to double. */
void or_axis_init(or_axis *a, float mn, float mx, int n, float base) {
	a->min = mn;
	a->minus_min = -mn;
	a->max = mx;
	a->n_bins = n;
	a->step = (mx - mn) / (float)n;
	a->inv_step = (float)(1.0 / (double)a->step);
	a->base = base;
	if ((double)base > 1.02) {
		a->ilb = (float)(1.0 / (double)logf(base));
		a->ils = (float)((pow((double)base, (double)n) - 1.0) / (double)(mx - mn));
		a->log_binning = 1;
	} else {
		a->ilb = 0.0f;
		a->ils = 0.0f;
		a->log_binning = 0;
	}
}

/* hist.cpp:46-53 */
float or_axis_bin_num(const or_axis *a, float v) {
	if (a->log_binning) return a->ilb * log1pf(a->ils * (v - a->min));
	return (v + a->minus_min) * a->inv_step;
}

/* hist.cpp:55-62 */
double or_axis_value_from_bin(const or_axis *a, double b) {
	if (a->log_binning) return (exp(b / (double)a->ilb) - 1.0) / (double)a->ils + (double)a->min;
	return (double)a->min + b * (double)a->step;
}

/* hist.cpp:38-44 */
double or_axis_bin_centre(const or_axis *a, int b) {
	return (or_axis_value_from_bin(a, b + 1.0) + or_axis_value_from_bin(a, (double)b)) / 2.0;
}
double or_axis_bin_width(const or_axis *a, int b) {
	return or_axis_value_from_bin(a, b + 1.0) - or_axis_value_from_bin(a, (double)b);
}

/* float -> int as the reference's implicit conversion on x86-64
 * (truncation; NaN and out-of-range give INT_MIN). */
static inline int f2i(float f) {
	return (f > -2147483904.0f && f < 2147483648.0f) ? (int)f : INT_MIN;
}

/* JKHists::interp, hist.cpp:369-391 */
double or_interp(const or_axis *a, const double *bins, double value) {
	if (value < (double)a->min) return bins[0];
	if (value > (double)a->max) return bins[a->n_bins - 1];
	float bn = (float)((double)or_axis_bin_num(a, (float)value) - 0.5);
	if (bn < 0) return bins[0];
	if (bn > a->n_bins - 1) return bins[a->n_bins - 1];
	int k = f2i(bn);
	return bins[k] + (bins[k + 1] - bins[k]) * (value - or_axis_bin_centre(a, k)) /
		(or_axis_bin_centre(a, k + 1) - or_axis_bin_centre(a, k));
}

/* --------------------------------------------------------- pair binning -- */

/* Range limits as the correlators set them: correlators.cpp:13-14 (monopole),
 * 192-195 (sigma-pi), 347-348 (s-mu).  Angular counts have no range test. */
void or_range_sq(int kind, const or_axis *ax1, const or_axis *ax2, float *mn, float *mx) {
	if (kind == OR_SIGMA_PI) {
		*mx = ax1->max * ax1->max + ax2->max * ax2->max;
		*mn = ax1->min * ax1->min;
		if (ax1->min > ax2->min) *mn = ax2->min * ax2->min;
	} else {
		*mx = ax1->max * ax1->max;
		*mn = ax1->min * ax1->min;
	}
}

static inline int accept_idx(const or_axis *a, float f, int *b) {
	int k = f2i(f);
	if (k < a->n_bins && k >= 0) { *b = k; return 1; }
	return 0;
}

/* Angular-definition line-of-sight separation, correlators.cpp:319-323:
 * each term is a double product of the float difference and the double sum
 * of unit vectors; the running value is rounded to float after every step. */
static inline float los_pi(float d0, float d1, float d2, const double *n1, const double *n2) {
	float p = (float)((double)d0 * (n1[0] + n2[0]));
	p = (float)((double)p + (double)d1 * (n1[1] + n2[1]));
	p = (float)((double)p + (double)d2 * (n1[2] + n2[2]));
	p = (float)((double)p * 0.5);
	return p;
}

/* Pair-midpoint line-of-sight separation, correlators.cpp:172-183 / 325-336
 * (builds without _ANGULAR_PI_DEFN): sum_vec = float(x1 + x2), its squared
 * norm summed in float, the inverse norm as 1.0 / sqrtf(...) in double
 * rounded to float, then pi = (d0 sum0) inv + (d1 sum1) inv + (d2 sum2) inv
 * in float, left to right. */
static inline float mid_pi(float d0, float d1, float d2, const double *x1, const double *x2) {
	float sv[3], inv = 0.0f;
	for (int k = 0; k < 3; ++k) {
		sv[k] = (float)(x1[k] + x2[k]);
		inv += sv[k] * sv[k];
	}
	inv = (float)(1.0 / (double)sqrtf(inv));
	float p = d0 * sv[0] * inv;
	p += d1 * sv[1] * inv;
	p += d2 * sv[2] * inv;
	return p;
}

static inline __attribute__((always_inline)) int pair_bin_k(int kind,
		const double *x1, const double *n1, const double *x2, const double *n2,
		float min_r_sq, float max_r_sq, const or_axis *ax1, const or_axis *ax2, int pi_def, int *bin) {
	if (kind == OR_ANGULAR) {
		double v = 1.0;
		v -= n1[0] * n2[0];
		v -= n1[1] * n2[1];
		v -= n1[2] * n2[2];
		return accept_idx(ax1, or_axis_bin_num(ax1, (float)v), bin);
	}
	float d0 = (float)(x1[0] - x2[0]);
	float d1 = (float)(x1[1] - x2[1]);
	float d2 = (float)(x1[2] - x2[2]);
	float s = d0 * d0 + d1 * d1 + d2 * d2;
	if (!(s < max_r_sq && s > min_r_sq)) return 0;
	if (kind == OR_MONO) return accept_idx(ax1, or_axis_bin_num(ax1, sqrtf(s)), bin);
	float p = pi_def == OR_PI_MIDPOINT ? mid_pi(d0, d1, d2, x1, x2) : los_pi(d0, d1, d2, n1, n2);
	int b1, b2;
	if (kind == OR_SIGMA_PI) {
		float sig = sqrtf(s - p * p);
		if (!accept_idx(ax1, or_axis_bin_num(ax1, sig), &b1)) return 0;
		if (!accept_idx(ax2, or_axis_bin_num(ax2, fabsf(p)), &b2)) return 0;
	} else {
		float r = sqrtf(s);
		if (!accept_idx(ax1, or_axis_bin_num(ax1, r), &b1)) return 0;
		if (!accept_idx(ax2, or_axis_bin_num(ax2, fabsf(p) / r), &b2)) return 0;
	}
	*bin = b1 * ax2->n_bins + b2;
	return 1;
}

int or_pair_bin(int kind, const double *x1, const double *n1, const double *x2, const double *n2,
                float min_r_sq, float max_r_sq, const or_axis *ax1, const or_axis *ax2, int *bin) {
	return pair_bin_k(kind, x1, n1, x2, n2, min_r_sq, max_r_sq, ax1, ax2, OR_PI_ANGULAR, bin);
}

int or_pair_bin_pi(int kind, const double *x1, const double *n1, const double *x2, const double *n2,
                   float min_r_sq, float max_r_sq, const or_axis *ax1, const or_axis *ax2, int pi_def, int *bin) {
	return pair_bin_k(kind, x1, n1, x2, n2, min_r_sq, max_r_sq, ax1, ax2, pi_def, bin);
}

/* ------------------------------------------------------------ weights -- */

static inline int popc_and(const uint64_t *m1, const uint64_t *m2, int nw) {
	int c = 0;
	for (int k = 0; k < nw; ++k) c += __builtin_popcountll(m1[k] & m2[k]);
	return c;
}

/* The numerator of the inverse-p rules: n_runs, or with dither weights the
 * float popcount of the AND of the dither masks, 0 counting as 1
 * (correlators_invpweight.cpp:62-68, 749-755). */
static inline float invp_numerator(const or_wopt *wo, const or_cat *a, long i, const or_cat *b, long j) {
	if (!wo->dither) return (float)wo->n_runs;
	float dw = (float)popc_and(a->dither + i * a->n_words, b->dither + j * b->n_words, a->n_words);
	if (dw < 0.5f) dw += 1.0f;
	return dw;
}

/* Pair weight under each rule; see the enum in twopcf_oracle.h. */
static inline float pair_weight(const or_wopt *wo, const or_cat *a, long i, const or_cat *b, long j,
                                const double *n1, const double *n2) {
	if (!wo || wo->mode == OR_W_PLAIN) return a->w[i] * b->w[j];
	double omc;
	float pc;
	switch (wo->mode) {
	case OR_W_INVP_3D:
		omc = 1.0;
		omc -= n1[0] * n2[0];
		omc -= n1[1] * n2[1];
		omc -= n1[2] * n2[2];
		pc = (float)popc_and(a->masks + i * a->n_words, b->masks + j * b->n_words, a->n_words);
		if (pc < 0.5f) pc += 1.0f;
		/* interp(omc) * numerator / pc in double, rounded to float once; the
		 * numerator is an integer (n_runs, or a float holding a popcount) */
		pc = (float)(or_interp(&wo->up_axis, wo->up_bins, omc) * (double)invp_numerator(wo, a, i, b, j) /
		             (double)pc);
		return a->iw[i] * b->iw[j] * pc;
	case OR_W_ANGUP_3D:
		omc = 1.0;
		omc -= n1[0] * n2[0];
		omc -= n1[1] * n2[1];
		omc -= n1[2] * n2[2];
		return (float)((double)(a->w[i] * b->w[j]) * or_interp(&wo->up_axis, wo->up_bins, omc));
	case OR_W_ANG_IW:
		return a->iw[i] * b->iw[j];
	case OR_W_ANG_INVP:
		pc = (float)popc_and(a->masks + i * a->n_words, b->masks + j * b->n_words, a->n_words);
		if (pc < 0.5f) pc += 1.0f;
		pc = invp_numerator(wo, a, i, b, j) / pc;
		return a->iw[i] * b->iw[j] * pc;
	case OR_W_ANG_W:
	default:
		return a->w[i] * b->w[j];
	}
}

/* ------------------------------------------------------------- counting -- */

typedef struct {
	double *full, *sub;
	unsigned long long acc;
} priv_hist;

/* JKHists(2D)::add_pair, hist.cpp:246-260 / 484-502 */
static inline void add_pair(priv_hist *h, long nb, int bin, int jk1, int jk2, float w) {
	h->full[bin] += w;
	h->sub[(long)jk1 * nb + bin] += w;
	if (jk1 != jk2) h->sub[(long)jk2 * nb + bin] += w;
	h->acc++;
}

static inline void load_pos(const or_cat *c, long i, double *x, double *n) {
	x[0] = c->x[i]; x[1] = c->y[i]; x[2] = c->z[i];
	n[0] = c->nx[i]; n[1] = c->ny[i]; n[2] = c->nz[i];
}

static int check_jk(const or_cat *c, int n_jk) {
	for (long i = 0; i < c->n; ++i)
		if (c->jk[i] < 0 || c->jk[i] >= n_jk) return -1;
	return 0;
}

/* Cell grid shared by both catalogues (union bounds as boxes.cpp:29-57, but
 * with true bounds and a side a hair above rmax).  Counting sort. */
typedef struct {
	int dim[3];
	double lo[3], h;
	long ncell;
	long *start_a, *start_b;   /* ncell+1 */
	long *ord_a, *ord_b;       /* cell-sorted object indices */
} or_grid;

static void cell_of(const or_grid *g, double x, double y, double z, int *c) {
	double v[3] = {x, y, z};
	for (int k = 0; k < 3; ++k) {
		long q = (long)((v[k] - g->lo[k]) / g->h);
		if (q < 0) q = 0;
		if (q >= g->dim[k]) q = g->dim[k] - 1;
		c[k] = (int)q;
	}
}

static void sort_cat(const or_grid *g, const or_cat *c, int angular, long **start, long **ord) {
	long *cnt = calloc(g->ncell + 1, sizeof(long));
	long *cid = malloc((c->n > 0 ? c->n : 1) * sizeof(long));
	for (long i = 0; i < c->n; ++i) {
		int q[3];
		if (angular) cell_of(g, c->nx[i], c->ny[i], c->nz[i], q);
		else cell_of(g, c->x[i], c->y[i], c->z[i], q);
		cid[i] = ((long)q[0] * g->dim[1] + q[1]) * g->dim[2] + q[2];
		cnt[cid[i] + 1]++;
	}
	for (long k = 0; k < g->ncell; ++k) cnt[k + 1] += cnt[k];
	long *o = malloc((c->n > 0 ? c->n : 1) * sizeof(long));
	long *fill = malloc(g->ncell * sizeof(long));
	memcpy(fill, cnt, g->ncell * sizeof(long));
	for (long i = 0; i < c->n; ++i) o[fill[cid[i]]++] = i;
	free(fill);
	free(cid);
	*start = cnt;
	*ord = o;
}

static void grid_build(or_grid *g, const or_cat *a, const or_cat *b, double rmax, int angular) {
	double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
	const or_cat *cs[2] = {a, b};
	for (int c = 0; c < 2; ++c) {
		if (!cs[c]) continue;
		for (long i = 0; i < cs[c]->n; ++i) {
			double v[3];
			if (angular) { v[0] = cs[c]->nx[i]; v[1] = cs[c]->ny[i]; v[2] = cs[c]->nz[i]; }
			else { v[0] = cs[c]->x[i]; v[1] = cs[c]->y[i]; v[2] = cs[c]->z[i]; }
			for (int k = 0; k < 3; ++k) {
				if (v[k] < lo[k]) lo[k] = v[k];
				if (v[k] > hi[k]) hi[k] = v[k];
			}
		}
	}
	g->h = rmax * (1.0 + 1e-6) + 1e-12;
	g->ncell = 1;
	for (int k = 0; k < 3; ++k) {
		if (hi[k] < lo[k]) { lo[k] = 0; hi[k] = 0; }
		g->lo[k] = lo[k];
		long d = (long)((hi[k] - lo[k]) / g->h) + 1;
		if (d > 1024) d = 1024;
		g->dim[k] = (int)d;
		g->ncell *= d;
	}
	sort_cat(g, a, angular, &g->start_a, &g->ord_a);
	if (b) sort_cat(g, b, angular, &g->start_b, &g->ord_b);
	else { g->start_b = g->start_a; g->ord_b = g->ord_a; }
}

static void grid_free(or_grid *g) {
	if (g->start_b != g->start_a) { free(g->start_b); free(g->ord_b); }
	free(g->start_a);
	free(g->ord_a);
}

static double pair_rmax(int kind, const or_axis *ax1, const or_axis *ax2) {
	if (kind == OR_SIGMA_PI) return sqrt((double)ax1->max * ax1->max + (double)ax2->max * ax2->max);
	if (kind == OR_ANGULAR) return sqrt(2.0 * (double)ax1->max) + 1e-9; /* chord length */
	return (double)ax1->max;
}

/* Inner double loop over one cell pair (or the whole catalogue in brute
 * force), restating correlators.cpp:360-381 / 406-423. */
static inline __attribute__((always_inline)) void loop_block(int kind, const or_cat *a, const long *ia, long na,
		const or_cat *b, const long *ib, long nb_, int triangle, float mn, float mx,
		const or_axis *ax1, const or_axis *ax2, long nbins, const or_wopt *wo, priv_hist *h) {
	for (long p = 0; p < na; ++p) {
		long i = ia[p];
		double x1[3], n1[3];
		load_pos(a, i, x1, n1);
		for (long q = triangle ? p + 1 : 0; q < nb_; ++q) {
			long j = ib[q];
			double x2[3], n2[3];
			load_pos(b, j, x2, n2);
			int bin;
			if (pair_bin_k(kind, x1, n1, x2, n2, mn, mx, ax1, ax2, wo ? wo->pi_def : OR_PI_ANGULAR, &bin))
				add_pair(h, nbins, bin, a->jk[i], b->jk[j], pair_weight(wo, a, i, b, j, n1, n2));
		}
	}
}

static int count_impl(int kind, const or_cat *a, const or_cat *b, const or_axis *ax1, const or_axis *ax2,
                      int n_jk, const or_wopt *wo, int method, int nthreads, double frac,
                      unsigned long long seed, double *full, double *sub, unsigned long long *n_acc,
                      double *frac_out) {
	if (n_jk < 1 || !a) return -1;
	if (check_jk(a, n_jk) || (b && check_jk(b, n_jk))) return -2;
	if (wo && wo->mode != OR_W_PLAIN) {
		if ((wo->mode == OR_W_INVP_3D || wo->mode == OR_W_ANG_INVP) &&
		    (!a->masks || !(b ? b : a)->masks || a->n_words != (b ? b : a)->n_words)) return -3;
		if (wo->dither && (wo->mode == OR_W_INVP_3D || wo->mode == OR_W_ANG_INVP) &&
		    (!a->dither || !(b ? b : a)->dither)) return -3;
		if ((wo->mode == OR_W_INVP_3D || wo->mode == OR_W_ANG_IW || wo->mode == OR_W_ANG_INVP) &&
		    (!a->iw || !(b ? b : a)->iw)) return -3;
		if ((wo->mode == OR_W_INVP_3D || wo->mode == OR_W_ANGUP_3D) && !wo->up_bins) return -3;
	}
	long nbins = (long)ax1->n_bins * (kind == OR_MONO || kind == OR_ANGULAR ? 1 : ax2->n_bins);
	const or_axis *ax2u = (kind == OR_MONO || kind == OR_ANGULAR) ? ax1 : ax2;
	float mn = 0.0f, mx = 0.0f;
	if (kind != OR_ANGULAR) or_range_sq(kind, ax1, ax2, &mn, &mx);
	int is_auto = (b == NULL);
	const or_cat *bb = is_auto ? a : b;
	memset(full, 0, nbins * sizeof(double));
	memset(sub, 0, (size_t)n_jk * nbins * sizeof(double));
	unsigned long long acc = 0;
	double cand_all = 0.0, cand_sel = 0.0;
#ifdef _OPENMP
	if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
	if (method == 0) {
		long *ia = malloc((a->n > 0 ? a->n : 1) * sizeof(long));
		long *ib = malloc((bb->n > 0 ? bb->n : 1) * sizeof(long));
		for (long i = 0; i < a->n; ++i) ia[i] = i;
		for (long i = 0; i < bb->n; ++i) ib[i] = i;
#pragma omp parallel
		{
			priv_hist h = {calloc(nbins, sizeof(double)), calloc((size_t)n_jk * nbins, sizeof(double)), 0};
#pragma omp for schedule(dynamic, 16)
			for (long p = 0; p < a->n; ++p) {
				if (is_auto)
					loop_block(kind, a, ia + p, 1, bb, ib + p + 1, bb->n - p - 1, 0, mn, mx, ax1, ax2u, nbins, wo, &h);
				else
					loop_block(kind, a, ia + p, 1, bb, ib, bb->n, 0, mn, mx, ax1, ax2u, nbins, wo, &h);
			}
#pragma omp critical
			{
				for (long k = 0; k < nbins; ++k) full[k] += h.full[k];
				for (long k = 0; k < (long)n_jk * nbins; ++k) sub[k] += h.sub[k];
				acc += h.acc;
			}
			free(h.full);
			free(h.sub);
		}
		free(ia);
		free(ib);
		cand_all = cand_sel = 1.0;
	} else {
		or_grid g;
		grid_build(&g, a, b, pair_rmax(kind, ax1, ax2), kind == OR_ANGULAR);
		/* Seeded Bernoulli selection of first-catalogue cells (sample mode). */
		unsigned char *sel = malloc(g.ncell);
		unsigned long long st = seed * 0x9E3779B97F4A7C15ull + 1;
		for (long c = 0; c < g.ncell; ++c) {
			st ^= st << 13; st ^= st >> 7; st ^= st << 17;
			sel[c] = (frac >= 1.0) || ((double)(st >> 11) * (1.0 / 9007199254740992.0) < frac);
		}
#pragma omp parallel reduction(+:cand_all, cand_sel)
		{
			priv_hist h = {calloc(nbins, sizeof(double)), calloc((size_t)n_jk * nbins, sizeof(double)), 0};
#pragma omp for schedule(dynamic)
			for (long c = 0; c < g.ncell; ++c) {
				long na = g.start_a[c + 1] - g.start_a[c];
				if (na == 0) continue;
				int cx = (int)(c / ((long)g.dim[1] * g.dim[2]));
				int cy = (int)((c / g.dim[2]) % g.dim[1]);
				int cz = (int)(c % g.dim[2]);
				for (int dx = -1; dx <= 1; ++dx)
				for (int dy = -1; dy <= 1; ++dy)
				for (int dz = -1; dz <= 1; ++dz) {
					int qx = cx + dx, qy = cy + dy, qz = cz + dz;
					if (qx < 0 || qy < 0 || qz < 0 || qx >= g.dim[0] || qy >= g.dim[1] || qz >= g.dim[2]) continue;
					long d = ((long)qx * g.dim[1] + qy) * g.dim[2] + qz;
					if (is_auto && d < c) continue;
					long nbq = g.start_b[d + 1] - g.start_b[d];
					if (nbq == 0) continue;
					double cand = (is_auto && d == c) ? 0.5 * na * (na - 1) : (double)na * nbq;
					cand_all += cand;
					if (!sel[c]) continue;
					cand_sel += cand;
					loop_block(kind, a, g.ord_a + g.start_a[c], na, bb, g.ord_b + g.start_b[d], nbq,
					           is_auto && d == c, mn, mx, ax1, ax2u, nbins, wo, &h);
				}
			}
#pragma omp critical
			{
				for (long k = 0; k < nbins; ++k) full[k] += h.full[k];
				for (long k = 0; k < (long)n_jk * nbins; ++k) sub[k] += h.sub[k];
				acc += h.acc;
			}
			free(h.full);
			free(h.sub);
		}
		free(sel);
		grid_free(&g);
	}
	if (n_acc) *n_acc = acc;
	if (frac_out) *frac_out = cand_all > 0 ? cand_sel / cand_all : 1.0;
	return 0;
}

int or_count(int kind, const or_cat *a, const or_cat *b, const or_axis *ax1, const or_axis *ax2,
             int n_jk, const or_wopt *wo, int method, int nthreads,
             double *full, double *sub, unsigned long long *n_accepted) {
	return count_impl(kind, a, b, ax1, ax2, n_jk, wo, method, nthreads, 1.0, 0, full, sub, n_accepted, NULL);
}

int or_count_sample(int kind, const or_cat *a, const or_cat *b, const or_axis *ax1, const or_axis *ax2,
                    int n_jk, int nthreads, double frac, unsigned long long seed,
                    double *full, double *sub, unsigned long long *n_accepted, double *frac_out) {
	return count_impl(kind, a, b, ax1, ax2, n_jk, NULL, 1, nthreads, frac, seed, full, sub, n_accepted, frac_out);
}

/* -------------------------------------------------- finish / normalise -- */

/* invert_jk_hists (hist.cpp:464-471) then multiply_hists(2.0) for auto
 * counts (hist.cpp:473-482; correlators.cpp:391-392). */
void or_finish(int is_auto, long nb, int n_jk, double *full, double *sub) {
	for (int r = 0; r < n_jk; ++r)
		for (long k = 0; k < nb; ++k) sub[(long)r * nb + k] = full[k] - sub[(long)r * nb + k];
	if (is_auto) {
		for (long k = 0; k < nb; ++k) full[k] *= 2.0;
		for (long k = 0; k < (long)n_jk * nb; ++k) sub[k] *= 2.0;
	}
}

/* JKHists2D::normalise(cat, n_gals), hist.cpp:512-543 (and the 1-D twin
 * hist.cpp:290-321), including the (1 - 1/n_in_region) factor. */
void or_normalise_auto(long nb, int n_jk, double *full, double *sub, const float *w, const int *jk, long n) {
	double sw = 0.0;
	double *swj = calloc(n_jk, sizeof(double));
	int *nj = calloc(n_jk, sizeof(int));
	for (long i = 0; i < n; ++i) {
		sw += w[i];
		swj[jk[i]] += w[i];
		++nj[jk[i]];
	}
	for (int r = 0; r < n_jk; ++r) swj[r] = sw - swj[r];
	sw = sw * sw * (1.0 - 1.0 / (double)n);
	for (int r = 0; r < n_jk; ++r) swj[r] = swj[r] * swj[r] * (1.0 - 1.0 / nj[r]);
	for (long k = 0; k < nb; ++k) full[k] /= sw;
	for (int r = 0; r < n_jk; ++r)
		for (long k = 0; k < nb; ++k) sub[(long)r * nb + k] /= swj[r];
	free(swj);
	free(nj);
}

/* JKHists2D::normalise(cat_1, n_1, cat_2, n_2), hist.cpp:545-589 */
void or_normalise_cross(long nb, int n_jk, double *full, double *sub,
                        const float *w1, const int *jk1, long n1, const float *w2, const int *jk2, long n2) {
	double s1 = 0.0, s2 = 0.0;
	double *j1 = calloc(n_jk, sizeof(double)), *j2 = calloc(n_jk, sizeof(double));
	for (long i = 0; i < n1; ++i) { s1 += w1[i]; j1[jk1[i]] += w1[i]; }
	for (int r = 0; r < n_jk; ++r) j1[r] = s1 - j1[r];
	for (long i = 0; i < n2; ++i) { s2 += w2[i]; j2[jk2[i]] += w2[i]; }
	for (int r = 0; r < n_jk; ++r) j2[r] = s2 - j2[r];
	s1 *= s2;
	for (int r = 0; r < n_jk; ++r) j1[r] *= j2[r];
	for (long k = 0; k < nb; ++k) full[k] /= s1;
	for (int r = 0; r < n_jk; ++r)
		for (long k = 0; k < nb; ++k) sub[(long)r * nb + k] /= j1[r];
	free(j1);
	free(j2);
}

/* io.cpp:643-651 */
double or_ls(double dd, double dr, double rr) {
	if (rr == 0) return 0;
	return dd / rr - 2.0 * dr / rr + 1.0;
}

/* ----------------------------------------------------------- cosmology -- */

/* cosmology.cpp:7-9 */
static double codist_integrand(double om, double ol, double opz) {
	return 1.0 / (sqrt(om * opz * opz * opz + ol));
}

/* cosmology.cpp:18 (table length) */
int or_cosmo_len(double z_min, double z_max, double step) {
	return (int)(((z_max - z_min) / step) + 1);
}

/* Comoving-distance table, cosmology.cpp:11-65.  The pre-loop passes its
 * arguments as (omega_l, omega_m) in the second term (cosmology.cpp:42),
 * which cancels exactly when z_min == 0. */
void or_cosmo_table(double om, double h, double z_min, double z_max, double step, int overs, double *y) {
	double ol = 1.0 - om;
	double c = 299792.458;
	double hubble = c / (h * 100.0);
	double z = 0.0, cur = 0.0, prev = 0.0;
	double fine = step / (double)overs;
	int nb = or_cosmo_len(z_min, z_max, step);
	while (z <= z_min) {
		prev = cur;
		cur += codist_integrand(om, ol, 1.0 + z) + codist_integrand(ol, om, 1.0 + z + fine);
		z += fine;
	}
	cur -= ((z - z_min) / fine) * (cur - prev);
	z = z_min;
	for (int i = 0; i < nb; ++i) {
		y[i] = 0.5 * fine * hubble * cur;
		for (int j = 0; j < overs; ++j) {
			cur += codist_integrand(om, ol, 1.0 + z) + codist_integrand(om, ol, 1 + z + fine);
			z += fine;
		}
	}
}

/* DatasetFixedXY::lin_interp, dataset.cpp:274-300; outside the table it
 * returns the x end-point, as the reference does. */
double or_comov_dist(const double *y, int len, double z_min, double step, double z) {
	double x0 = z_min, xn = x0 + (len - 1) * step;
	if (z <= x0) return x0;
	if (z >= xn) return xn;
	int k = (int)((z - x0) / step);
	return y[k] + ((z - (x0 + k * step)) / step) * (y[k + 1] - y[k]);
}

/* Galaxy::set, galaxy.cpp:8-27.  The reference's g++ -O3 build fuses the
 * sin and cos of each angle into one glibc sincos call, whose sine differs
 * from sin() in the last bit for some arguments; sincos is called here
 * explicitly so the unit vectors match that build bit for bit. */
void or_eq_to_cart(long n, const double *ra, const double *dec, const double *d,
                   double *x, double *y, double *z, double *nx, double *ny, double *nz) {
	const double d2r = M_PI / 180.0;
	for (long i = 0; i < n; ++i) {
		double sr, cr, sd, cd;
		sincos(d2r * ra[i], &sr, &cr);
		sincos(d2r * dec[i], &sd, &cd);
		nx[i] = cd * cr;
		ny[i] = cd * sr;
		nz[i] = sd;
		if (d[i] >= 0.0) { x[i] = d[i] * nx[i]; y[i] = d[i] * ny[i]; z[i] = d[i] * nz[i]; }
		else { x[i] = 0.0; y[i] = 0.0; z[i] = 0.0; }
	}
}

/* io.cpp:393-399 */
void or_cart_norm(long n, const double *x, const double *y, const double *z,
                  double *d, double *nx, double *ny, double *nz) {
	for (long i = 0; i < n; ++i) {
		d[i] = sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
		nx[i] = x[i] / d[i];
		ny[i] = y[i] / d[i];
		nz[i] = z[i] / d[i];
	}
}

/* ------------------------------------------------ threshold-table check -- */

/* Scan every float s2 in [lo, hi] and compare the reference's accept/bin
 * decision for the s axis (range test, sqrtf, Axis::bin_num, truncation;
 * correlators.cpp:315-317, 338; hist.cpp:46-53, 486-490) with the table
 * rule "bin k iff t[k] <= s2 < t[k+1], rejected outside [t[0], t[n])".
 * Returns the number of disagreeing floats. */
long or_scan_s2_table(int kind, const or_axis *ax1, const or_axis *ax2, const float *t,
                      float lo, float hi, int nthreads) {
	float mn, mx;
	or_range_sq(kind, ax1, ax2, &mn, &mx);
	uint32_t klo, khi;
	memcpy(&klo, &lo, 4);
	memcpy(&khi, &hi, 4);
	const int n = ax1->n_bins;
	long bad = 0;
#ifdef _OPENMP
	if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static, 1 << 16) reduction(+:bad)
	for (long k = (long)klo; k <= (long)khi; ++k) {
		uint32_t b = (uint32_t)k;
		float s2;
		memcpy(&s2, &b, 4);
		int ref_ok = 0, ref_bin = -1, bb;
		if (s2 < mx && s2 > mn && accept_idx(ax1, or_axis_bin_num(ax1, sqrtf(s2)), &bb)) { ref_ok = 1; ref_bin = bb; }
		int tab_ok = (s2 >= t[0] && s2 < t[n]);
		int tab_bin = -1;
		if (tab_ok) {
			int lo_ = 0, hi_ = n - 1;
			while (lo_ < hi_) { int mid = (lo_ + hi_ + 1) / 2; if (t[mid] <= s2) lo_ = mid; else hi_ = mid - 1; }
			tab_bin = lo_;
		}
		if (ref_ok != tab_ok || (ref_ok && ref_bin != tab_bin)) bad++;
	}
	return bad;
}

/* Same for a value-domain table of one axis over all floats in [lo, hi]. */
long or_scan_value_table(const or_axis *a, const float *t, float lo, float hi, int nthreads) {
	uint32_t klo, khi;
	memcpy(&klo, &lo, 4);
	memcpy(&khi, &hi, 4);
	const int n = a->n_bins;
	long bad = 0;
#ifdef _OPENMP
	if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static, 1 << 16) reduction(+:bad)
	for (long k = (long)klo; k <= (long)khi; ++k) {
		uint32_t b = (uint32_t)k;
		float v;
		memcpy(&v, &b, 4);
		int bb, ref_ok = accept_idx(a, or_axis_bin_num(a, v), &bb);
		int tab_ok = (v >= t[0] && v < t[n]);
		int tab_bin = -1;
		if (tab_ok) {
			int lo_ = 0, hi_ = n - 1;
			while (lo_ < hi_) { int mid = (lo_ + hi_ + 1) / 2; if (t[mid] <= v) lo_ = mid; else hi_ = mid - 1; }
			tab_bin = lo_;
		}
		if (ref_ok != tab_ok || (ref_ok && bb != tab_bin)) bad++;
	}
	return bad;
}
