// twopcf_ref_api.cpp — see twopcf_ref_api.hpp.  Compiled with
// -ffp-contract=off so the host-side float/double arithmetic (Axis constants,
// bin centres, normalisation, LS estimator) rounds as in the reference build.
#include "twopcf_ref_api.hpp"

#include "h5lib.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <thread>

// ------------------------------------------------------------------ context --

twopcf_ctx *twopcf_default_ctx() {
	static twopcf_ctx *ctx = nullptr;
	if (!ctx) {
		const char *dev = std::getenv("TWOPCF_DEVICE");
		int rc = twopcf_ctx_create(dev ? std::atoi(dev) : 0, &ctx);
		if (rc) {
			std::fprintf(stderr, "twopcf: no usable GPU (twopcf_ctx_create returned %d); there is no CPU path\n", rc);
			std::exit(1);
		}
	}
	return ctx;
}

int twopcf_default_method() {
	const char *m = std::getenv("TWOPCF_METHOD");
	if (m && std::string(m) == "fp64") return TWOPCF_PLAIN_FP64;
	return TWOPCF_FP32_RECHECK;
}

static void check(int rc, const char *what) {
	if (rc) {
		std::fprintf(stderr, "twopcf: %s failed (%d): %s\n", what, rc, twopcf_last_error(twopcf_default_ctx()));
		std::exit(1);
	}
}

// ------------------------------------------------------------------- galaxy --

// galaxy.cpp:8-27 (the library's eq_to_cart is the bit-exact restatement)
void Galaxy::set(double ra, double dec) {
	double d = d_comov;
	twopcf_eq_to_cart(1, &ra, &dec, &d, &x[0], &x[1], &x[2], &x_norm[0], &x_norm[1], &x_norm[2]);
}

GalaxyCatalogue::GalaxyCatalogue(int n, int n_jk) : n_jk_regions(n_jk), n_gals(n) { gals = new Galaxy[n > 0 ? n : 1]; }
GalaxyCatalogue::~GalaxyCatalogue() { delete[] gals; }

// --------------------------------------------------------------------- hist --

// hist.cpp:15-36
Axis::Axis(float input_min, float input_max, int input_n_bins, float log_base) {
	min = input_min;
	minus_min = -input_min;
	max = input_max;
	n_bins = input_n_bins;
	step = (max - min) / (float)n_bins;
	inv_step = (float)(1.0 / (double)step);
	base = log_base;
	if ((double)base > 1.02) {
		inv_log_of_base = (float)(1.0 / (double)logf(log_base));
		inv_log_step = (float)((std::pow((double)log_base, (double)input_n_bins) - 1.0) / (double)(input_max - input_min));
		log_binning = true;
	} else {
		inv_log_of_base = 0.0f;
		inv_log_step = 0.0f;
		log_binning = false;
	}
}

double Axis::bin_centre(const int b) { return (value_from_bin(b + 1.0) + value_from_bin((double)b)) / 2.0; }
double Axis::bin_width(const int b) { return value_from_bin(b + 1.0) - value_from_bin((double)b); }

float Axis::bin_num(const float value) {
	if (log_binning) return inv_log_of_base * log1pf(inv_log_step * (value - min));
	return (value + minus_min) * inv_step;
}

double Axis::value_from_bin(const double b) {
	if (log_binning) return (std::exp(b / (double)inv_log_of_base) - 1.0) / (double)inv_log_step + (double)min;
	return (double)min + b * (double)step;
}

twopcf_axis Axis::desc() const { return twopcf_axis{(double)min, (double)max, n_bins, (double)base}; }

Hist::Hist(float mn, float mx, int n, float b) : ax(mn, mx, n, b) { bins = (double *)calloc(ax.n_bins, sizeof(double)); }
Hist::Hist(const Hist &o) : ax(o.ax.min, o.ax.max, o.ax.n_bins, o.ax.base) {
	bins = (double *)calloc(ax.n_bins, sizeof(double));
	std::memcpy(bins, o.bins, sizeof(double) * ax.n_bins);
}
Hist::~Hist() { free(bins); }

JKHists::JKHists(int n_jk, const Hist &t) : n_jk_regions(n_jk), full_hist(t) {
	sub_hist = (double *)calloc((size_t)n_jk * full_hist.ax.n_bins, sizeof(double));
}
JKHists::~JKHists() { free(sub_hist); }

// hist.cpp:270-288
void JKHists::invert_jk_hists() {
	const int nb = full_hist.ax.n_bins;
	for (int r = 0; r < n_jk_regions; ++r)
		for (int b = 0; b < nb; ++b) sub_hist[r * nb + b] = full_hist.bins[b] - sub_hist[r * nb + b];
}
void JKHists::multiply_hists(double f) {
	const int nb = full_hist.ax.n_bins;
	for (int b = 0; b < nb; ++b) full_hist.bins[b] *= f;
	for (int k = 0; k < n_jk_regions * nb; ++k) sub_hist[k] *= f;
}
// hist.cpp:224-231
void JKHists::divide_by_hist(const JKHists &o) {
	const int nb = full_hist.ax.n_bins;
	for (int b = 0; b < nb; ++b) full_hist.bins[b] /= o.full_hist.bins[b];
	for (int k = 0; k < n_jk_regions * nb; ++k) sub_hist[k] /= o.sub_hist[k];
}

// Weight sums over the first n objects (hist.cpp:296-303, 518-525).
static void weight_sums(const GalaxyCatalogue *cat, int n, int n_jk, double &sw, std::vector<double> &swj,
                        std::vector<int> &nj) {
	sw = 0.0;
	swj.assign(n_jk, 0.0);
	nj.assign(n_jk, 0);
	for (int i = 0; i < n; ++i) {
		sw += cat->gals[i].weight;
		swj[cat->gals[i].jk_region] += cat->gals[i].weight;
		++nj[cat->gals[i].jk_region];
	}
	for (int r = 0; r < n_jk; ++r) swj[r] = sw - swj[r];
}

// hist.cpp:290-321 / 512-543 (the region factor uses the count inside the
// region, as the reference does)
static void normalise_auto(double *full, double *sub, int nb, int n_jk, const GalaxyCatalogue *cat, int n) {
	double sw;
	std::vector<double> swj;
	std::vector<int> nj;
	weight_sums(cat, n, n_jk, sw, swj, nj);
	sw = sw * sw * (1.0 - 1.0 / (double)n);
	for (int r = 0; r < n_jk; ++r) swj[r] = swj[r] * swj[r] * (1.0 - 1.0 / nj[r]);
	for (int b = 0; b < nb; ++b) full[b] /= sw;
	for (int r = 0; r < n_jk; ++r)
		for (int b = 0; b < nb; ++b) sub[r * nb + b] /= swj[r];
}

// hist.cpp:323-367 / 545-589
static void normalise_cross(double *full, double *sub, int nb, int n_jk, const GalaxyCatalogue *c1, int n1,
                            const GalaxyCatalogue *c2, int n2) {
	double s1, s2;
	std::vector<double> j1, j2;
	std::vector<int> k1, k2;
	weight_sums(c1, n1, n_jk, s1, j1, k1);
	weight_sums(c2, n2, n_jk, s2, j2, k2);
	s1 *= s2;
	for (int r = 0; r < n_jk; ++r) j1[r] *= j2[r];
	for (int b = 0; b < nb; ++b) full[b] /= s1;
	for (int r = 0; r < n_jk; ++r)
		for (int b = 0; b < nb; ++b) sub[r * nb + b] /= j1[r];
}

void JKHists::normalise(const GalaxyCatalogue *cat, int n) {
	normalise_auto(full_hist.bins, sub_hist, full_hist.ax.n_bins, n_jk_regions, cat, n);
}
void JKHists::normalise(const GalaxyCatalogue *c1, int n1, const GalaxyCatalogue *c2, int n2) {
	normalise_cross(full_hist.bins, sub_hist, full_hist.ax.n_bins, n_jk_regions, c1, n1, c2, n2);
}

// hist.cpp:369-391
double JKHists::interp(const double value) {
	Axis &a = full_hist.ax;
	if (value < a.min) return full_hist.bins[0];
	if (value > a.max) return full_hist.bins[a.n_bins - 1];
	float bn = (float)((double)a.bin_num((float)value) - 0.5);
	if (bn < 0) return full_hist.bins[0];
	if (bn > a.n_bins - 1) return full_hist.bins[a.n_bins - 1];
	const int k = (int)bn;
	return full_hist.bins[k] + (full_hist.bins[k + 1] - full_hist.bins[k]) * (value - a.bin_centre(k)) /
	                               (a.bin_centre(k + 1) - a.bin_centre(k));
}

Hist2D::Hist2D(float mn1, float mx1, int n1, float b1, float mn2, float mx2, int n2, float b2)
	: ax1(mn1, mx1, n1, b1), ax2(mn2, mx2, n2, b2) {
	total_bins = ax1.n_bins * ax2.n_bins;
	bins = (double *)calloc(total_bins, sizeof(double));
}
Hist2D::Hist2D(const Hist2D &o)
	: ax1(o.ax1.min, o.ax1.max, o.ax1.n_bins, o.ax1.base), ax2(o.ax2.min, o.ax2.max, o.ax2.n_bins, o.ax2.base) {
	total_bins = ax1.n_bins * ax2.n_bins;
	bins = (double *)calloc(total_bins, sizeof(double));
	std::memcpy(bins, o.bins, sizeof(double) * total_bins);
}
Hist2D::~Hist2D() { free(bins); }

JKHists2D::JKHists2D(int n_jk, const Hist2D &t) : n_jk_regions(n_jk), full_hist(t) {
	sub_hist = (double *)calloc((size_t)n_jk * full_hist.total_bins, sizeof(double));
}
JKHists2D::~JKHists2D() { free(sub_hist); }
void JKHists2D::invert_jk_hists() {
	const int nb = full_hist.total_bins;
	for (int r = 0; r < n_jk_regions; ++r)
		for (int b = 0; b < nb; ++b) sub_hist[r * nb + b] = full_hist.bins[b] - sub_hist[r * nb + b];
}
void JKHists2D::multiply_hists(double f) {
	const int nb = full_hist.total_bins;
	for (int b = 0; b < nb; ++b) full_hist.bins[b] *= f;
	for (int k = 0; k < n_jk_regions * nb; ++k) sub_hist[k] *= f;
}
void JKHists2D::normalise(const GalaxyCatalogue *cat, int n) {
	normalise_auto(full_hist.bins, sub_hist, full_hist.total_bins, n_jk_regions, cat, n);
}
void JKHists2D::normalise(const GalaxyCatalogue *c1, int n1, const GalaxyCatalogue *c2, int n2) {
	normalise_cross(full_hist.bins, sub_hist, full_hist.total_bins, n_jk_regions, c1, n1, c2, n2);
}

// -------------------------------------------------------------------- boxes --

// Objects without a usable distance (d_comov < 0) go behind n_gals
// (boxes.cpp:16-26); the first n_gals are uploaded to the GPU.
static int partition_valid(Galaxy *g, int n) {
	Galaxy *mid = std::stable_partition(g, g + n, [](const Galaxy &a) { return !(a.d_comov < 0.0); });
	return (int)(mid - g);
}

static twopcf_cat *upload(Galaxy *g, int n, const std::vector<uint64_t> *masks, int n_words,
                          const std::vector<uint64_t> *dither = nullptr) {
	std::vector<double> c[6];
	for (auto &v : c) v.resize(n);
	std::vector<float> w(n), iw(n);
	std::vector<int32_t> jk(n);
	for (int i = 0; i < n; ++i) {
		for (int k = 0; k < 3; ++k) {
			c[k][i] = g[i].x[k];
			c[3 + k][i] = g[i].x_norm[k];
		}
		w[i] = g[i].weight;
		iw[i] = g[i].input_weight;
		jk[i] = g[i].jk_region;
	}
	twopcf_cat_desc d{};
	d.n = n;
	d.x = c[0].data();
	d.y = c[1].data();
	d.z = c[2].data();
	d.nx = c[3].data();
	d.ny = c[4].data();
	d.nz = c[5].data();
	d.w = w.data();
	d.jk = jk.data();
	d.iw = iw.data();
	if (masks && n_words > 0) {
		d.masks = masks->data();
		d.n_words = n_words;
		if (dither && !dither->empty()) d.dither = dither->data();
	}
	twopcf_cat *cat = nullptr;
	check(twopcf_catalog_upload(twopcf_default_ctx(), &d, &cat), "twopcf_catalog_upload");
	return cat;
}

GalaxyBoxes::GalaxyBoxes(Galaxy in[], int n, Galaxy *, int, double r) : rmax(r) {
	n_gals = partition_valid(in, n);
	cat = upload(in, n_gals, nullptr, 0);
}

// Move the objects with d_comov < 0 behind the others, keeping the masks
// aligned with their objects; returns the number in front.
static int partition_catalogue(GalaxyCatalogue *in) {
	const int n = in->n_gals, nw = in->n_words;
	std::vector<int> order(n);
	for (int i = 0; i < n; ++i) order[i] = i;
	auto mid = std::stable_partition(order.begin(), order.end(), [&](int i) { return !(in->gals[i].d_comov < 0.0); });
	const int n_valid = (int)(mid - order.begin());
	std::vector<Galaxy> g(n);
	const bool dith = !in->dither.empty();
	std::vector<uint64_t> m((size_t)n * nw), dm(dith ? (size_t)n * nw : 0);
	for (int i = 0; i < n; ++i) {
		g[i] = in->gals[order[i]];
		for (int k = 0; k < nw; ++k) m[(size_t)i * nw + k] = in->masks[(size_t)order[i] * nw + k];
		if (dith)
			for (int k = 0; k < nw; ++k) dm[(size_t)i * nw + k] = in->dither[(size_t)order[i] * nw + k];
	}
	std::copy(g.begin(), g.end(), in->gals);
	in->masks.swap(m);
	in->dither.swap(dm);
	return n_valid;
}

GalaxyBoxes::GalaxyBoxes(GalaxyCatalogue *in, GalaxyCatalogue *, double r) : rmax(r), source(in) {
	n_gals = partition_catalogue(in);
	cat = upload(in->gals, n_gals, &in->masks, in->n_words, &in->dither);
}

// boxes.cpp:215-240: with require_redshift the objects without a distance go
// to the back and are left out; otherwise every object takes part (angular
// counts use only the unit vectors)
GalaxyPixels::GalaxyPixels(Galaxy in[], int n, int, double t, bool require_redshift) : theta_max(t) {
	n_gals = require_redshift ? partition_valid(in, n) : n;
	cat = upload(in, n_gals, nullptr, 0);
}

GalaxyPixels::GalaxyPixels(GalaxyCatalogue *in, int, double t, bool require_redshift) : theta_max(t) {
	n_gals = require_redshift ? partition_catalogue(in) : in->n_gals;
	cat = upload(in->gals, n_gals, &in->masks, in->n_words, &in->dither);
}

GalaxyPixels::~GalaxyPixels() {
	if (cat) twopcf_catalog_free(cat);
}

GalaxyBoxes::~GalaxyBoxes() {
	if (cat) twopcf_catalog_free(cat);
}

// -------------------------------------------------------------- correlators --

namespace {

struct Out {
	double *full;
	double *sub;
	int nb, n_jk;
	bool is_auto;
};

// Run a batch of counts in one call and leave each histogram as the
// reference's pair-count functions do: jackknife sums inverted
// (hist.cpp:464-471) and, for auto counts, multiplied by 2 (hist.cpp:473-482).
void run(std::vector<twopcf_job> &jobs, std::vector<Out> &outs) {
	std::vector<std::vector<double>> raw(jobs.size());
	for (size_t q = 0; q < jobs.size(); ++q) {
		raw[q].assign((size_t)outs[q].n_jk * outs[q].nb, 0.0);
		jobs[q].full = outs[q].full;
		jobs[q].jk = raw[q].data();
	}
	check(twopcf_count(twopcf_default_ctx(), jobs.data(), (int)jobs.size(), twopcf_default_method(), nullptr, nullptr),
	      "twopcf_count");
	for (size_t q = 0; q < jobs.size(); ++q) {
		const Out &o = outs[q];
		std::memcpy(o.sub, raw[q].data(), sizeof(double) * raw[q].size());
		for (int r = 0; r < o.n_jk; ++r)
			for (int b = 0; b < o.nb; ++b) o.sub[r * o.nb + b] = o.full[b] - o.sub[r * o.nb + b];
		if (o.is_auto) {
			for (int b = 0; b < o.nb; ++b) o.full[b] *= 2.0;
			for (int k = 0; k < o.n_jk * o.nb; ++k) o.sub[k] *= 2.0;
		}
	}
}

twopcf_job job1(int kind, const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists *h) {
	twopcf_job j{};
	j.kind = kind;
	j.a = a->cat;
	j.b = b ? b->cat : nullptr;
	j.ax[0] = h->full_hist.ax.desc();
	j.ax[1] = twopcf_axis{0.0, 1.0, 1, 0.0};
	j.n_jk = h->n_jk_regions;
	j.weight_mode = TWOPCF_W_PLAIN;
	return j;
}

twopcf_job job2(int kind, const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists2D *h) {
	twopcf_job j{};
	j.pi_def = twopcf_pi_definition();
	j.kind = kind;
	j.a = a->cat;
	j.b = b ? b->cat : nullptr;
	j.ax[0] = h->full_hist.ax1.desc();
	j.ax[1] = h->full_hist.ax2.desc();
	j.n_jk = h->n_jk_regions;
	j.weight_mode = TWOPCF_W_PLAIN;
	return j;
}

Out out1(JKHists *h, bool a) { return Out{h->full_hist.bins, h->sub_hist, h->full_hist.ax.n_bins, h->n_jk_regions, a}; }
Out out2(JKHists2D *h, bool a) {
	return Out{h->full_hist.bins, h->sub_hist, h->full_hist.total_bins, h->n_jk_regions, a};
}

void count1(int kind, const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists *h) {
	std::vector<twopcf_job> j{job1(kind, a, b, h)};
	std::vector<Out> o{out1(h, b == nullptr)};
	run(j, o);
}
void count2(int kind, const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists2D *h) {
	std::vector<twopcf_job> j{job2(kind, a, b, h)};
	std::vector<Out> o{out2(h, b == nullptr)};
	run(j, o);
}

}  // namespace

void auto_pair_count_monopole(const GalaxyBoxes *c, JKHists *h) { count1(TWOPCF_MONOPOLE, c, nullptr, h); }
void cross_pair_count_monopole(const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists *h) { count1(TWOPCF_MONOPOLE, a, b, h); }
void auto_pair_count_sigma_pi(const GalaxyBoxes *c, JKHists2D *h) { count2(TWOPCF_SIGMA_PI, c, nullptr, h); }
void cross_pair_count_sigma_pi(const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists2D *h) { count2(TWOPCF_SIGMA_PI, a, b, h); }
void auto_pair_count_s_mu(const GalaxyBoxes *c, JKHists2D *h) { count2(TWOPCF_S_MU, c, nullptr, h); }
void cross_pair_count_s_mu(const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists2D *h) { count2(TWOPCF_S_MU, a, b, h); }

// correlators.cpp:127-154 — DD, DR and RR in one batched call
void auto_correlate_monopole(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params p) {
	std::cout << "Calculating autocorrelation monopole\n";
	Hist t(p.monopole_min, p.monopole_max, p.monopole_n_bins, p.monopole_log_base);
	JKHists DD(p.n_jk_regions, t), DR(p.n_jk_regions, t), RR(p.n_jk_regions, t);
	GalaxyBoxes db(data, randoms, p.monopole_max);
	GalaxyBoxes rb(randoms, data, p.monopole_max);
	std::vector<twopcf_job> jobs{job1(TWOPCF_MONOPOLE, &db, nullptr, &DD), job1(TWOPCF_MONOPOLE, &db, &rb, &DR),
	                             job1(TWOPCF_MONOPOLE, &rb, nullptr, &RR)};
	std::vector<Out> outs{out1(&DD, true), out1(&DR, false), out1(&RR, true)};
	run(jobs, outs);
	DD.normalise(data, db.n_gals);
	DR.normalise(data, db.n_gals, randoms, rb.n_gals);
	RR.normalise(randoms, rb.n_gals);
	output_xi_1D(p.monopole_filename, p.monopole_output_type, &DD, &DR, &RR, data, randoms);
}

// correlators.cpp:280-307
void auto_correlate_sigma_pi(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params p) {
	std::cout << "Calculating autocorrelation sigma-pi decomposition\n";
	Hist2D t(p.sigma_min, p.sigma_max, p.sigma_n_bins, p.sigma_log_base, p.pi_min, p.pi_max, p.pi_n_bins,
	         p.pi_log_base);
	JKHists2D DD(p.n_jk_regions, t), DR(p.n_jk_regions, t), RR(p.n_jk_regions, t);
	const double rmax = std::sqrt(p.sigma_max * p.sigma_max + p.pi_max * p.pi_max);
	GalaxyBoxes db(data, randoms, rmax);
	GalaxyBoxes rb(randoms, data, rmax);
	std::vector<twopcf_job> jobs{job2(TWOPCF_SIGMA_PI, &db, nullptr, &DD), job2(TWOPCF_SIGMA_PI, &db, &rb, &DR),
	                             job2(TWOPCF_SIGMA_PI, &rb, nullptr, &RR)};
	std::vector<Out> outs{out2(&DD, true), out2(&DR, false), out2(&RR, true)};
	run(jobs, outs);
	DD.normalise(data, db.n_gals);
	DR.normalise(data, db.n_gals, randoms, rb.n_gals);
	RR.normalise(randoms, rb.n_gals);
	output_xi_2D(p.sigma_pi_filename, p.sigma_pi_output_type, "sigma", "pi", &DD, &DR, &RR, data, randoms);
}

// correlators.cpp:432-458
void auto_correlate_s_mu(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params p) {
	std::cout << "Calculating autocorrelation s-mu decomposition\n";
	Hist2D t(p.s_min, p.s_max, p.s_n_bins, p.s_log_base, 0.0, 1.0, p.mu_n_bins, 0);
	JKHists2D DD(p.n_jk_regions, t), DR(p.n_jk_regions, t), RR(p.n_jk_regions, t);
	GalaxyBoxes db(data, randoms, p.s_max);
	GalaxyBoxes rb(randoms, data, p.s_max);
	std::vector<twopcf_job> jobs{job2(TWOPCF_S_MU, &db, nullptr, &DD), job2(TWOPCF_S_MU, &db, &rb, &DR),
	                             job2(TWOPCF_S_MU, &rb, nullptr, &RR)};
	std::vector<Out> outs{out2(&DD, true), out2(&DR, false), out2(&RR, true)};
	run(jobs, outs);
	DD.normalise(data, db.n_gals);
	DR.normalise(data, db.n_gals, randoms, rb.n_gals);
	RR.normalise(randoms, rb.n_gals);
	output_xi_2D(p.s_mu_filename, p.s_mu_output_type, "s", "mu", &DD, &DR, &RR, data, randoms);
}

// ------------------------------------------------------------- inverse-p --

namespace {

// Angular histograms by file name (in place of the HDF5 files)
struct StoredHist {
	float min, max, base;
	int n_bins, n_jk;
	std::vector<double> full, sub;
};
std::map<std::string, StoredHist> &angular_store() {
	static std::map<std::string, StoredHist> m;
	return m;
}

twopcf_job job_ang(int mode, const GalaxyPixels *a, const GalaxyPixels *b, JKHists *h, int n_runs) {
	twopcf_job j{};
	j.kind = TWOPCF_ANGULAR;
	j.a = a->cat;
	j.b = b ? b->cat : nullptr;
	j.ax[0] = h->full_hist.ax.desc();
	j.ax[1] = twopcf_axis{0.0, 1.0, 1, 0.0};
	j.n_jk = h->n_jk_regions;
	j.weight_mode = mode;
	j.n_runs = n_runs;
	j.dither = mode == TWOPCF_W_ANG_INVP && twopcf_dither_weights_enabled();
	return j;
}

void count_ang(int mode, const GalaxyPixels *a, const GalaxyPixels *b, JKHists *h, int n_runs) {
	std::vector<twopcf_job> j{job_ang(mode, a, b, h, n_runs)};
	std::vector<Out> o{out1(h, b == nullptr)};
	run(j, o);
}

// The angular up-weight of the 3-D counts: the parent-weight angular
// histogram divided by the inverse-p one (correlators_invpweight.cpp:20-24),
// bin by bin; JKHists::interp reads its full histogram.
struct UpWeight {
	JKHists *h = nullptr;
	explicit UpWeight(const std::string &num, const std::string &den) {
		h = read_angular_hist(num);
		JKHists *d = read_angular_hist(den);
		h->divide_by_hist(*d);
		delete d;
	}
	~UpWeight() { delete h; }
	void apply(twopcf_job &j, int mode, int n_runs) const {
		j.weight_mode = mode;
		j.n_runs = n_runs;
		j.dither = mode == TWOPCF_W_INVP_3D && twopcf_dither_weights_enabled();
		j.up_bins = h->full_hist.bins;
		j.up_axis = h->full_hist.ax.desc();
	}
};

}  // namespace

void store_angular_hist(const std::string &fn, const JKHists &h) {
	StoredHist s;
	s.min = h.full_hist.ax.min;
	s.max = h.full_hist.ax.max;
	s.base = h.full_hist.ax.base;
	s.n_bins = h.full_hist.ax.n_bins;
	s.n_jk = h.n_jk_regions;
	s.full.assign(h.full_hist.bins, h.full_hist.bins + s.n_bins);
	s.sub.assign(h.sub_hist, h.sub_hist + (size_t)s.n_jk * s.n_bins);
	angular_store()[fn] = s;
}

JKHists *read_angular_hist(const std::string &fn) {
	auto it = angular_store().find(fn);
	if (it == angular_store().end()) {
		// not computed in this run: the file an earlier run wrote
		if (!std::ifstream(fn).good()) {
			std::fprintf(stderr, "twopcf: angular histogram %s was neither computed in this run nor found as a file "
			                     "(set the matching calculate_angular_* key)\n", fn.c_str());
			std::exit(1);
		}
		JKHists *h = nullptr;
		read_hdf5_hist(fn, &h);
		return h;
	}
	const StoredHist &s = it->second;
	Hist t(s.min, s.max, s.n_bins, s.base);
	JKHists *h = new JKHists(s.n_jk, t);
	std::copy(s.full.begin(), s.full.end(), h->full_hist.bins);
	std::copy(s.sub.begin(), s.sub.end(), h->sub_hist);
	return h;
}

// correlators_invpweight.cpp:870-936, 938-983, 713-821, 823-868
void auto_pair_count_angular(const GalaxyPixels *c, JKHists *h) { count_ang(TWOPCF_W_ANG_IW, c, nullptr, h, 0); }
void cross_pair_count_angular(const GalaxyPixels *a, const GalaxyPixels *b, JKHists *h) {
	count_ang(TWOPCF_W_ANG_IW, a, b, h, 0);
}
void auto_pair_count_angular_invpweights(const GalaxyPixels *c, JKHists *h, Params p) {
	count_ang(TWOPCF_W_ANG_INVP, c, nullptr, h, p.n_bitwise_runs);
}
void cross_pair_count_angular_invpweights(const GalaxyPixels *a, const GalaxyPixels *b, JKHists *h) {
	count_ang(TWOPCF_W_ANG_W, a, b, h, 0);
}

// correlators_invpweight.cpp:985-1033: each histogram is written to its
// HDF5 file (JKHists::output_hdf5) and kept in the process for the 3-D
// counts; without a libhdf5 it is only kept in the process.
void calc_angular_upweight(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params p) {
	auto store_angular_hist = [](const std::string &fn, JKHists &h) {
		::store_angular_hist(fn, h);
		if (h5::available())
			h.output_hdf5(fn);
		else
			std::cout << "No libhdf5: " << fn << " is kept in this process only\n";
	};
	Hist t(0.0, 1.0 - std::cos(p.theta_max), p.theta_n_bins, p.theta_log_base);
	if (p.calculate_angular_dd) {
		std::cout << "Calculating angular DDs\n";
		GalaxyPixels dp(data, p.healpix_order, p.theta_max, false);
		JKHists DD(p.n_jk_regions, t);
		auto_pair_count_angular(&dp, &DD);
		store_angular_hist(p.angular_dd_filename, DD);
	}
	if (p.calculate_angular_dd_invpweights) {
		std::cout << "Calculating angular DDs with inverse p-weights\n";
		GalaxyPixels dp(data, p.healpix_order, p.theta_max, true);
		JKHists DD(p.n_jk_regions, t);
		auto_pair_count_angular_invpweights(&dp, &DD, p);
		store_angular_hist(p.angular_dd_invpweights_filename, DD);
	}
	if (p.calculate_angular_dr) {
		std::cout << "Calculating angular DRs\n";
		GalaxyPixels dp(data, p.healpix_order, p.theta_max, false);
		GalaxyPixels rp(randoms, p.healpix_order, p.theta_max, false);
		JKHists DR(p.n_jk_regions, t);
		cross_pair_count_angular(&dp, &rp, &DR);
		store_angular_hist(p.angular_dr_filename, DR);
	}
	if (p.calculate_angular_dr_invpweights) {
		std::cout << "Calculating angular DRs with inverse weights\n";
		GalaxyPixels dp(data, p.healpix_order, p.theta_max, true);
		GalaxyPixels rp(randoms, p.healpix_order, p.theta_max, true);
		JKHists DR(p.n_jk_regions, t);
		cross_pair_count_angular_invpweights(&dp, &rp, &DR);
		store_angular_hist(p.angular_dr_invpweights_filename, DR);
	}
}

// correlators_invpweight.cpp:14-146, 148-208 (and the sigma-pi, s-mu
// counterparts at 339-443, 579-680)
void auto_pair_count_monopole_invpweights(const GalaxyBoxes *c, JKHists *h, Params p) {
	UpWeight u(p.angular_dd_filename, p.angular_dd_invpweights_filename);
	std::vector<twopcf_job> j{job1(TWOPCF_MONOPOLE, c, nullptr, h)};
	u.apply(j[0], TWOPCF_W_INVP_3D, p.n_bitwise_runs);
	std::vector<Out> o{out1(h, true)};
	run(j, o);
}
void cross_pair_count_monopole_invpweights(const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists *h, Params p) {
	UpWeight u(p.angular_dr_filename, p.angular_dr_invpweights_filename);
	std::vector<twopcf_job> j{job1(TWOPCF_MONOPOLE, a, b, h)};
	u.apply(j[0], TWOPCF_W_ANGUP_3D, 0);
	std::vector<Out> o{out1(h, false)};
	run(j, o);
}
void auto_pair_count_sigma_pi_invpweights(const GalaxyBoxes *c, JKHists2D *h, Params p) {
	UpWeight u(p.angular_dd_filename, p.angular_dd_invpweights_filename);
	std::vector<twopcf_job> j{job2(TWOPCF_SIGMA_PI, c, nullptr, h)};
	u.apply(j[0], TWOPCF_W_INVP_3D, p.n_bitwise_runs);
	std::vector<Out> o{out2(h, true)};
	run(j, o);
}
void cross_pair_count_sigma_pi_invpweights(const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists2D *h, Params p) {
	UpWeight u(p.angular_dr_filename, p.angular_dr_invpweights_filename);
	std::vector<twopcf_job> j{job2(TWOPCF_SIGMA_PI, a, b, h)};
	u.apply(j[0], TWOPCF_W_ANGUP_3D, 0);
	std::vector<Out> o{out2(h, false)};
	run(j, o);
}
void auto_pair_count_s_mu_invpweights(const GalaxyBoxes *c, JKHists2D *h, Params p) {
	UpWeight u(p.angular_dd_filename, p.angular_dd_invpweights_filename);
	std::vector<twopcf_job> j{job2(TWOPCF_S_MU, c, nullptr, h)};
	u.apply(j[0], TWOPCF_W_INVP_3D, p.n_bitwise_runs);
	std::vector<Out> o{out2(h, true)};
	run(j, o);
}
void cross_pair_count_s_mu_invpweights(const GalaxyBoxes *a, const GalaxyBoxes *b, JKHists2D *h, Params p) {
	UpWeight u(p.angular_dr_filename, p.angular_dr_invpweights_filename);
	std::vector<twopcf_job> j{job2(TWOPCF_S_MU, a, b, h)};
	u.apply(j[0], TWOPCF_W_ANGUP_3D, 0);
	std::vector<Out> o{out2(h, false)};
	run(j, o);
}

// correlators_invpweight.cpp:210-237, 445-474, 682-711: DD with inverse-p
// weights, DR with the angular up-weight, RR plain — here in one batched call
void auto_correlate_monopole_invpweights(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params p) {
	std::cout << "Calculating autocorrelation monopole with inverse p weights\n";
	Hist t(p.monopole_min, p.monopole_max, p.monopole_n_bins, p.monopole_log_base);
	JKHists DD(p.n_jk_regions, t), DR(p.n_jk_regions, t), RR(p.n_jk_regions, t);
	GalaxyBoxes db(data, randoms, p.monopole_max);
	GalaxyBoxes rb(randoms, data, p.monopole_max);
	UpWeight udd(p.angular_dd_filename, p.angular_dd_invpweights_filename);
	UpWeight udr(p.angular_dr_filename, p.angular_dr_invpweights_filename);
	std::vector<twopcf_job> jobs{job1(TWOPCF_MONOPOLE, &db, nullptr, &DD), job1(TWOPCF_MONOPOLE, &db, &rb, &DR),
	                             job1(TWOPCF_MONOPOLE, &rb, nullptr, &RR)};
	udd.apply(jobs[0], TWOPCF_W_INVP_3D, p.n_bitwise_runs);
	udr.apply(jobs[1], TWOPCF_W_ANGUP_3D, 0);
	std::vector<Out> outs{out1(&DD, true), out1(&DR, false), out1(&RR, true)};
	run(jobs, outs);
	DD.normalise(data, db.n_gals);
	DR.normalise(data, db.n_gals, randoms, rb.n_gals);
	RR.normalise(randoms, rb.n_gals);
	output_xi_1D(p.monopole_filename, p.monopole_output_type, &DD, &DR, &RR, data, randoms);
}

void auto_correlate_sigma_pi_invpweights(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params p) {
	std::cout << "Calculating autocorrelation sigma-pi decomposition with inverse p weights\n";
	Hist2D t(p.sigma_min, p.sigma_max, p.sigma_n_bins, p.sigma_log_base, p.pi_min, p.pi_max, p.pi_n_bins,
	         p.pi_log_base);
	JKHists2D DD(p.n_jk_regions, t), DR(p.n_jk_regions, t), RR(p.n_jk_regions, t);
	const double rmax = std::sqrt(p.sigma_max * p.sigma_max + p.pi_max * p.pi_max);
	GalaxyBoxes db(data, randoms, rmax);
	GalaxyBoxes rb(randoms, data, rmax);
	UpWeight udd(p.angular_dd_filename, p.angular_dd_invpweights_filename);
	UpWeight udr(p.angular_dr_filename, p.angular_dr_invpweights_filename);
	std::vector<twopcf_job> jobs{job2(TWOPCF_SIGMA_PI, &db, nullptr, &DD), job2(TWOPCF_SIGMA_PI, &db, &rb, &DR),
	                             job2(TWOPCF_SIGMA_PI, &rb, nullptr, &RR)};
	udd.apply(jobs[0], TWOPCF_W_INVP_3D, p.n_bitwise_runs);
	udr.apply(jobs[1], TWOPCF_W_ANGUP_3D, 0);
	std::vector<Out> outs{out2(&DD, true), out2(&DR, false), out2(&RR, true)};
	run(jobs, outs);
	DD.normalise(data, db.n_gals);
	DR.normalise(data, db.n_gals, randoms, rb.n_gals);
	RR.normalise(randoms, rb.n_gals);
	output_xi_2D(p.sigma_pi_filename, p.sigma_pi_output_type, "sigma", "pi", &DD, &DR, &RR, data, randoms);
}

void auto_correlate_s_mu_invpweights(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params p) {
	std::cout << "Calculating autocorrelation s-mu decomposition with inverse p weights\n";
	Hist2D t(p.s_min, p.s_max, p.s_n_bins, p.s_log_base, 0.0, 1.0, p.mu_n_bins, 0);
	JKHists2D DD(p.n_jk_regions, t), DR(p.n_jk_regions, t), RR(p.n_jk_regions, t);
	GalaxyBoxes db(data, randoms, p.s_max);
	GalaxyBoxes rb(randoms, data, p.s_max);
	UpWeight udd(p.angular_dd_filename, p.angular_dd_invpweights_filename);
	UpWeight udr(p.angular_dr_filename, p.angular_dr_invpweights_filename);
	std::vector<twopcf_job> jobs{job2(TWOPCF_S_MU, &db, nullptr, &DD), job2(TWOPCF_S_MU, &db, &rb, &DR),
	                             job2(TWOPCF_S_MU, &rb, nullptr, &RR)};
	udd.apply(jobs[0], TWOPCF_W_INVP_3D, p.n_bitwise_runs);
	udr.apply(jobs[1], TWOPCF_W_ANGUP_3D, 0);
	std::vector<Out> outs{out2(&DD, true), out2(&DR, false), out2(&RR, true)};
	run(jobs, outs);
	DD.normalise(data, db.n_gals);
	DR.normalise(data, db.n_gals, randoms, rb.n_gals);
	RR.normalise(randoms, rb.n_gals);
	output_xi_2D(p.s_mu_filename, p.s_mu_output_type, "s", "mu", &DD, &DR, &RR, data, randoms);
}

static int g_pi_def = TWOPCF_PI_ANGULAR;
void twopcf_set_pi_definition(int pi_def) { g_pi_def = pi_def; }
int twopcf_pi_definition() { return g_pi_def; }

static bool g_inverse_p = false;
void twopcf_enable_inverse_p(bool on) { g_inverse_p = on; }
bool twopcf_inverse_p_enabled() { return g_inverse_p; }

static bool g_dither = false;
void twopcf_enable_dither_weights(bool on) { g_dither = on; }
bool twopcf_dither_weights_enabled() { return g_dither; }

// io.cpp:521-588: the input weight keeps the parent weight; the weight is
// scaled by n_bitwise_runs / total, or with dither weights by
// dither_total / total (io.cpp:549-574), when total != n_bitwise_runs
void apply_bitwise_weights(GalaxyCatalogue *cat, const Params &p) {
	const int nw = cat->n_words;
	const bool dith = !cat->dither.empty();
	for (int i = 0; i < cat->n_gals; ++i) {
		int total = 0, dither_total = 0;
		for (int k = 0; k < nw; ++k) total += __builtin_popcountll(cat->masks[(size_t)i * nw + k]);
		if (dith)
			for (int k = 0; k < nw; ++k) dither_total += __builtin_popcountll(cat->dither[(size_t)i * nw + k]);
		Galaxy &g = cat->gals[i];
		g.input_weight = g.weight;
		if (total == 0) total = 1;
		if (total != p.n_bitwise_runs)
			g.weight *= (dith ? (double)dither_total : (double)p.n_bitwise_runs) / (double)total;
	}
}

// ----------------------------------------------------------------------- io --

static std::string strip_spaces(std::string s) {
	s.erase(std::remove(s.begin(), s.end(), ' '), s.end());
	return s;
}

// io.cpp:16-237: "key = value", '#' comments, spaces removed
Params::Params(std::string fn) : param_filename(fn) {
	std::cout << "Reading parameter file: " << fn << "\n";
	std::ifstream file(fn);
	std::string line, key, value;
	while (std::getline(file, line)) {
		if (line.empty() || line[0] == '#' || line.find('=') == std::string::npos) continue;
		std::stringstream nt(line);
		std::getline(nt, line, '#');
		std::stringstream ss(line);
		std::getline(ss, key, '=');
		std::getline(ss, value);
		key = strip_spaces(key);
		value = strip_spaces(value);
		auto I = [&](int &v) { v = std::stoi(value); };
		auto B = [&](bool &v) { v = std::stoi(value) != 0; };
		auto D = [&](double &v) { v = std::stod(value); };
		if (key == "data_filename") data_filename = value;
		else if (key == "data_file_type") data_file_type = value;
		else if (key == "random_filename") random_filename = value;
		else if (key == "random_file_type") random_file_type = value;
		else if (key == "coord_system") coord_system = value;
		else if (key == "ra_x_dataset_name") ra_x_dataset_name = value;
		else if (key == "dec_y_dataset_name") dec_y_dataset_name = value;
		else if (key == "z_z_dataset_name") z_z_dataset_name = value;
		else if (key == "weight_dataset_name") weight_dataset_name = value;
		else if (key == "jk_dataset_name") jk_dataset_name = value;
		else if (key == "use_weights") B(use_weights);
		else if (key == "n_threads") I(n_threads);
		else if (key == "n_jk_regions") I(n_jk_regions);
		else if (key == "omega_m") D(omega_m);
		else if (key == "h") D(h);
		else if (key == "z_min") D(z_min);
		else if (key == "z_max") D(z_max);
		else if (key == "plot_monopole") B(plot_monopole);
		else if (key == "monopole_filename") monopole_filename = value;
		else if (key == "monopole_output_type") monopole_output_type = value;
		else if (key == "monopole_log_base") D(monopole_log_base);
		else if (key == "monopole_min") D(monopole_min);
		else if (key == "monopole_max") D(monopole_max);
		else if (key == "monopole_n_bins") I(monopole_n_bins);
		else if (key == "plot_sigma_pi") B(plot_sigma_pi);
		else if (key == "sigma_pi_filename") sigma_pi_filename = value;
		else if (key == "sigma_pi_output_type") sigma_pi_output_type = value;
		else if (key == "sigma_log_base") D(sigma_log_base);
		else if (key == "sigma_min") D(sigma_min);
		else if (key == "sigma_max") D(sigma_max);
		else if (key == "sigma_n_bins") I(sigma_n_bins);
		else if (key == "pi_log_base") D(pi_log_base);
		else if (key == "pi_min") D(pi_min);
		else if (key == "pi_max") D(pi_max);
		else if (key == "pi_n_bins") I(pi_n_bins);
		else if (key == "plot_s_mu") B(plot_s_mu);
		else if (key == "s_mu_filename") s_mu_filename = value;
		else if (key == "s_mu_output_type") s_mu_output_type = value;
		else if (key == "s_log_base") D(s_log_base);
		else if (key == "s_min") D(s_min);
		else if (key == "s_max") D(s_max);
		else if (key == "s_n_bins") I(s_n_bins);
		else if (key == "mu_n_bins") I(mu_n_bins);
		else if (key == "calculate_angular_dd") B(calculate_angular_dd);
		else if (key == "angular_dd_filename") angular_dd_filename = value;
		else if (key == "calculate_angular_dd_invpweights") B(calculate_angular_dd_invpweights);
		else if (key == "angular_dd_invpweights_filename") angular_dd_invpweights_filename = value;
		else if (key == "calculate_angular_dr") B(calculate_angular_dr);
		else if (key == "angular_dr_filename") angular_dr_filename = value;
		else if (key == "calculate_angular_dr_invpweights") B(calculate_angular_dr_invpweights);
		else if (key == "angular_dr_invpweights_filename") angular_dr_invpweights_filename = value;
		else if (key == "theta_max") D(theta_max);
		else if (key == "theta_n_bins") I(theta_n_bins);
		else if (key == "theta_log_base") D(theta_log_base);
		else if (key == "healpix_order") I(healpix_order);
		else if (key == "bitwise_weight_dataset_name") bitwise_weight_dataset_name = value;
		else if (key == "dither_weight_dataset_name") dither_weight_dataset_name = value;
		else if (key == "n_bitwise_runs") I(n_bitwise_runs);
		else std::cout << "Key: " << key << " KEY NOT FOUND\n";
	}
	if (n_jk_regions <= 1) n_jk_regions = 1;
}

// One field of an ascii line as the reference's reader takes it
// (io.cpp:287-303, std::getline on the separator): the next non-empty part
// between separators, spaces removed, copied into buf.  Returns false when
// the line has no further part.
static bool next_field(const char *&p, const char *e, char sep, char *buf, size_t cap) {
	while (p < e) {
		const char *q = (const char *)std::memchr(p, sep, (size_t)(e - p));
		if (!q) q = e;
		const char *b = p;
		p = q < e ? q + 1 : e;
		if (q == b) continue;   // empty part
		size_t n = 0;
		for (const char *c = b; c < q; ++c)
			if (*c != ' ' && n + 1 < cap) buf[n++] = *c;
		buf[n] = '\0';
		return true;
	}
	return false;
}

namespace {

// Parse failures of a line, reported by the caller with its line number.
enum { PARSE_OK = 0, PARSE_MISSING = 1, PARSE_BAD = 2 };

struct Columns {
	std::vector<double> c1, c2, c3;
	std::vector<float> w;
	std::vector<int> jk;
	std::vector<uint64_t> masks;
};

// The numeric conversions of std::stod / std::stof / std::stoi / std::stoll /
// std::stoull (strtod, strtof, strtol, strtoll, strtoull on the field).
int parse_line(const char *b, const char *e, const Params &p, int n_words, size_t i, Columns &c) {
	char sep = ' ';
	if (std::memchr(b, '\t', (size_t)(e - b))) sep = '\t';
	if (std::memchr(b, ',', (size_t)(e - b))) sep = ',';
	char buf[128];
	char *end;
	double *xyz[3] = {&c.c1[i], &c.c2[i], &c.c3[i]};
	for (double *v : xyz) {
		if (!next_field(b, e, sep, buf, sizeof buf)) return PARSE_MISSING;
		*v = std::strtod(buf, &end);
		if (end == buf) return PARSE_BAD;
	}
	c.w[i] = 1.0f;
	if (p.use_weights) {
		if (!next_field(b, e, sep, buf, sizeof buf)) return PARSE_MISSING;
		c.w[i] = std::strtof(buf, &end);
		if (end == buf) return PARSE_BAD;
	}
	c.jk[i] = 0;
	if (p.n_jk_regions > 1) {
		if (!next_field(b, e, sep, buf, sizeof buf)) return PARSE_MISSING;
		const long v = std::strtol(buf, &end, 10);
		if (end == buf) return PARSE_BAD;
		c.jk[i] = (int)v;
	}
	for (int k = 0; k < n_words; ++k) {
		if (!next_field(b, e, sep, buf, sizeof buf)) return PARSE_MISSING;
		uint64_t m = buf[0] == '-' ? (uint64_t)std::strtoll(buf, &end, 10) : (uint64_t)std::strtoull(buf, &end, 10);
		if (end == buf) return PARSE_BAD;
		c.masks[i * (size_t)n_words + k] = m;
	}
	return PARSE_OK;
}

}  // namespace

// Run work(t) for t = 0 .. n_threads - 1 on the calling thread and
// n_threads - 1 others.
static const int n_threads = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
template <class F>
static void parallel_chunks(F work) {
	std::vector<std::thread> pool;
	for (int t = 1; t < n_threads; ++t) pool.emplace_back(work, t);
	work(0);
	for (auto &th : pool) th.join();
}

// Mask words per object in the inverse-p build (0 otherwise).
static int mask_words(const Params &p) {
	if (twopcf_inverse_p_enabled() && p.n_bitwise_runs <= 0) {
		std::fprintf(stderr, "twopcf: the inverse-p pipeline needs n_bitwise_runs > 0\n");
		std::exit(1);
	}
	return twopcf_inverse_p_enabled() ? (p.n_bitwise_runs + 63) / 64 : 0;
}

// The part of both readers after the columns are read (io.cpp:323-415,
// 454-588): weights and regions, the bit masks and their IIP weights, then
// positions: comoving distances and Galaxy::set for equatorial catalogues,
// the columns themselves for cartesian ones.  Regions outside
// [0, n_jk_regions) are an error (the reference writes out of bounds); the
// message names the object by its `unit` ("line" or "entry") number.
static GalaxyCatalogue *finish_catalogue(GalaxyCatalogue *cat, Columns &c, int n_words, int n_cols, const Params &p,
                                         const std::string &filename, const char *unit) {
	const int n = cat->n_gals;
	const bool eq = p.coord_system == "equatorial";
	for (int i = 0; i < n; ++i) {
		Galaxy &g = cat->gals[i];
		g.weight = c.w[i];
		g.jk_region = c.jk[i];
		g.input_weight = g.weight;
		if (g.jk_region < 0 || g.jk_region >= p.n_jk_regions) {
			std::fprintf(stderr, "twopcf: %s %s %d: jackknife region %d outside [0, %d)\n", filename.c_str(), unit, i + 1,
			             g.jk_region, p.n_jk_regions);
			std::exit(1);
		}
	}
	cat->n_words = n_words;
	if (n_cols > n_words) {
		cat->masks.resize((size_t)n * n_words);
		cat->dither.resize((size_t)n * n_words);
		for (size_t i = 0; i < (size_t)n; ++i)
			for (int k = 0; k < n_words; ++k) {
				cat->masks[i * n_words + k] = c.masks[i * n_cols + k];
				cat->dither[i * n_words + k] = c.masks[i * n_cols + n_words + k];
			}
	} else {
		cat->masks.swap(c.masks);
	}
	if (n_words > 0) apply_bitwise_weights(cat, p);
	if (eq) {
		// comoving distances and Galaxy::set, in chunks over the threads
		// (both are per-object functions of the library)
		std::vector<double> d(n);
		std::vector<double> pos[6];
		for (auto &v : pos) v.resize(n);
		std::vector<int> rc(n_threads, 0);
		parallel_chunks([&](int t) {
			const size_t b = (size_t)n * t / n_threads, e = (size_t)n * (t + 1) / n_threads;
			const int64_t m = (int64_t)(e - b);
			if (m <= 0) return;
			rc[t] = twopcf_comoving_distance(p.omega_m, p.h, p.z_min, p.z_max, m, c.c3.data() + b, d.data() + b);
			if (!rc[t])
				rc[t] = twopcf_eq_to_cart(m, c.c1.data() + b, c.c2.data() + b, d.data() + b, pos[0].data() + b,
				                          pos[1].data() + b, pos[2].data() + b, pos[3].data() + b, pos[4].data() + b,
				                          pos[5].data() + b);
		});
		for (int r : rc) check(r, "twopcf_comoving_distance / twopcf_eq_to_cart");
		for (int i = 0; i < n; ++i) cat->gals[i].d_comov = d[i];
		for (int i = 0; i < n; ++i) {
			Galaxy &g = cat->gals[i];
			for (int k = 0; k < 3; ++k) {
				g.x[k] = pos[k][i];
				g.x_norm[k] = pos[3 + k][i];
			}
		}
	} else {
		for (int i = 0; i < n; ++i) {
			Galaxy &g = cat->gals[i];
			g.x[0] = c.c1[i];
			g.x[1] = c.c2[i];
			g.x[2] = c.c3[i];
			g.d_comov = std::sqrt(g.x[0] * g.x[0] + g.x[1] * g.x[1] + g.x[2] * g.x[2]);
			for (int k = 0; k < 3; ++k) g.x_norm[k] = g.x[k] / g.d_comov;
		}
	}
	return cat;
}

// io.cpp:240-415 on a whole-file buffer: lines (not empty, not starting with
// '#') are found in one pass, then parsed by up to 16 threads into
// per-object columns; values are the reference's conversions of the same
// fields.
GalaxyCatalogue *read_data_ascii(std::string filename, const Params &p) {
	std::ifstream file(filename, std::ios::binary);
	if (!file.is_open()) {
		std::cout << "File not open\n";
		return new GalaxyCatalogue(0, p.n_jk_regions);
	}
	file.seekg(0, std::ios::end);
	std::string data((size_t)std::max<std::streamoff>(file.tellg(), 0), '\0');
	file.seekg(0, std::ios::beg);
	file.read(&data[0], (std::streamsize)data.size());
	std::vector<std::pair<size_t, size_t>> lines;
	for (size_t pos = 0; pos < data.size();) {
		const char *nl = (const char *)std::memchr(data.data() + pos, '\n', data.size() - pos);
		const size_t end = nl ? (size_t)(nl - data.data()) : data.size();
		if (end > pos && data[pos] != '#') lines.emplace_back(pos, end);
		pos = end + 1;
	}
	const int n = (int)lines.size();
	auto *cat = new GalaxyCatalogue(n, p.n_jk_regions);
	const bool eq = p.coord_system == "equatorial";
	if (!eq && p.coord_system != "cartesian") {
		std::cout << "Coord system not found\n";
		return cat;
	}
	// inverse-p build: bit masks as extra integer columns; with dither
	// weights, n_words more columns follow: the dither mask
	const int n_words = mask_words(p);
	const int n_cols = n_words * (twopcf_dither_weights_enabled() ? 2 : 1);
	Columns c;
	c.c1.resize(n);
	c.c2.resize(n);
	c.c3.resize(n);
	c.w.resize(n);
	c.jk.resize(n);
	c.masks.assign((size_t)n * n_cols, 0);
	std::vector<int> fail_at(n_threads, -1), fail_code(n_threads, PARSE_OK);
	auto work = [&](int t) {
		const size_t b = (size_t)n * t / n_threads, e = (size_t)n * (t + 1) / n_threads;
		for (size_t i = b; i < e; ++i) {
			const int rc = parse_line(data.data() + lines[i].first, data.data() + lines[i].second, p, n_cols, i, c);
			if (rc) {
				fail_at[t] = (int)i;
				fail_code[t] = rc;
				return;
			}
		}
	};
	parallel_chunks(work);
	for (int t = 0; t < n_threads; ++t)
		if (fail_at[t] >= 0) {
			std::fprintf(stderr, "twopcf: %s catalogue line %d: %s\n", filename.c_str(), fail_at[t] + 1,
			             fail_code[t] == PARSE_MISSING ? "too few columns" : "not a number");
			std::exit(1);
		}
	return finish_catalogue(cat, c, n_words, n_cols, p, filename, "line");
}

namespace {

[[noreturn]] void h5_fail(const std::string &file, const std::string &what) {
	std::fprintf(stderr, "twopcf: %s: HDF5 %s failed\n", file.c_str(), what.c_str());
	std::exit(1);
}

// One dataset of the catalogue file read whole into dst as `type`; it must
// hold n values (n < 0: any number, returned).
long long h5_read_column(h5::hid_t f, const std::string &file, const std::string &name, h5::hid_t type, void *dst,
                         long long n) {
	const h5::Lib &H = h5::lib();
	h5::hid_t d = H.Dopen2(f, name.c_str(), h5::P_DEFAULT);
	if (d < 0) h5_fail(file, "open of dataset '" + name + "'");
	h5::hid_t sp = H.Dget_space(d);
	const long long m = sp < 0 ? -1 : (long long)H.Sget_simple_extent_npoints(sp);
	if (sp >= 0) H.Sclose(sp);
	if (m < 0) h5_fail(file, "size of dataset '" + name + "'");
	if (n >= 0 && m != n) {
		std::fprintf(stderr, "twopcf: %s: dataset '%s' holds %lld values, expected %lld\n", file.c_str(), name.c_str(),
		             m, n);
		std::exit(1);
	}
	if (dst && H.Dread(d, type, h5::S_ALL, h5::S_ALL, h5::P_DEFAULT, dst) < 0)
		h5_fail(file, "read of dataset '" + name + "'");
	H.Dclose(d);
	return m;
}

}  // namespace

// io.cpp:417-590: the columns come from the datasets the parameter file
// names (ra_x, dec_y, z_z as double; the weight as double, converted to
// float; the region as int; in the inverse-p build the bit masks from
// <bitwise_weight_dataset_name>0, 1, ... as long, and with dither weights the
// dither masks from <dither_weight_dataset_name>0, 1, ...), one value per
// object; the object count is the ra_x dataset's size.  A file or dataset
// that cannot be read, or a dataset of another size, is an error.
GalaxyCatalogue *read_data_hdf5(std::string filename, const Params &p) {
	const h5::Lib &H = h5::lib();
	h5::hid_t f = H.Fopen(filename.c_str(), h5::ACC_RDONLY, h5::P_DEFAULT);
	if (f < 0) h5_fail(filename, "open");
	std::cout << "File successfully opened\n";
	const long long n64 = h5_read_column(f, filename, p.ra_x_dataset_name, H.NATIVE_DOUBLE, nullptr, -1);
	if (n64 > 0x7fffffffLL) h5_fail(filename, "read (more than 2^31 - 1 objects)");
	const int n = (int)n64;
	std::cout << "N points in file: " << n << "\n";
	auto *cat = new GalaxyCatalogue(n, p.n_jk_regions);
	if (p.coord_system != "equatorial" && p.coord_system != "cartesian") {
		std::cout << "Cannot find coordinate system\n";
		H.Fclose(f);
		return cat;
	}
	const int n_words = mask_words(p);
	const bool dith = twopcf_dither_weights_enabled();
	const int n_cols = n_words * (dith ? 2 : 1);
	Columns c;
	c.c1.resize(n);
	c.c2.resize(n);
	c.c3.resize(n);
	c.w.assign(n, 1.0f);
	c.jk.assign(n, 0);
	c.masks.assign((size_t)n * n_cols, 0);
	h5_read_column(f, filename, p.ra_x_dataset_name, H.NATIVE_DOUBLE, c.c1.data(), n);
	h5_read_column(f, filename, p.dec_y_dataset_name, H.NATIVE_DOUBLE, c.c2.data(), n);
	h5_read_column(f, filename, p.z_z_dataset_name, H.NATIVE_DOUBLE, c.c3.data(), n);
	if (p.use_weights) {
		std::vector<double> w(n);
		h5_read_column(f, filename, p.weight_dataset_name, H.NATIVE_DOUBLE, w.data(), n);
		for (int i = 0; i < n; ++i) c.w[i] = (float)w[i];
	}
	if (p.n_jk_regions > 1) h5_read_column(f, filename, p.jk_dataset_name, H.NATIVE_INT, c.jk.data(), n);
	std::vector<long long> col(n_words > 0 ? n : 0);
	for (int set = 0; set < (dith ? 2 : 1) && n_words > 0; ++set) {
		const std::string &base = set ? p.dither_weight_dataset_name : p.bitwise_weight_dataset_name;
		for (int k = 0; k < n_words; ++k) {
			h5_read_column(f, filename, base + std::to_string(k), H.NATIVE_LONG, col.data(), n);
			for (int i = 0; i < n; ++i) c.masks[(size_t)i * n_cols + set * n_words + k] = (uint64_t)col[i];
		}
	}
	H.Fclose(f);
	return finish_catalogue(cat, c, n_words, n_cols, p, filename, "entry");
}

// io.cpp:601-627
GalaxyCatalogue *read_file(std::string filename, std::string file_type, const Params &p) {
	std::cout << "\nReading file: " << filename << "\n";
	if (file_type == "ascii") {
		std::cout << "File type ASCII\n";
		return read_data_ascii(filename, p);
	}
	if (file_type == "hdf5") {
		std::cout << "File type hdf5\n";
		return read_data_hdf5(filename, p);
	}
	std::cout << "Can't find file type for " << filename << "\n";
	return new GalaxyCatalogue(0, p.n_jk_regions);
}

// io.cpp:630-641
std::string generate_sub_filename(std::string filename, long long sub_number) {
	size_t dot = filename.find_last_of('.');
	if (dot == std::string::npos) return filename + std::to_string(sub_number);
	std::string ext = filename.substr(dot);
	std::string r = filename;
	r.erase(dot);
	r.append(std::to_string(sub_number));
	r.append(ext);
	return r;
}

// io.cpp:643-651
double ls_estimator(double DD, double DR, double RR) {
	if (RR == 0) return 0;
	return DD / RR - 2.0 * DR / RR + 1.0;
}

// io.cpp:724-737 (default stream precision, tab separated)
static void single_1d(std::string fn, Axis *ax, double *DD, double *DR, double *RR) {
	std::ofstream f(fn);
	for (int i = 0; i < ax->n_bins; ++i) {
		f << i << "\t" << ax->bin_centre(i) << "\t" << ax->bin_width(i) << "\t" << ls_estimator(DD[i], DR[i], RR[i])
		  << "\t" << DD[i] << "\t" << DR[i] << "\t" << RR[i] << "\n";
	}
}

// io.cpp:836-859
static void single_2d(std::string fn, Axis *a1, Axis *a2, double *DD, double *DR, double *RR) {
	std::ofstream f(fn);
	for (int i = 0; i < a1->n_bins; ++i)
		for (int j = 0; j < a2->n_bins; ++j) {
			int b = i * a2->n_bins + j;
			f << i << "\t" << j << "\t" << a1->bin_centre(i) << "\t" << a2->bin_centre(j) << "\t" << a1->bin_width(i)
			  << "\t" << a2->bin_width(j) << "\t" << ls_estimator(DD[b], DR[b], RR[b]) << "\t" << DD[b] << "\t" << DR[b]
			  << "\t" << RR[b] << "\n";
		}
}

namespace {

// io.cpp:653-697: scalar attributes.  A string attribute has the type
// H5T_C_S1, one character, so the file holds the first character of the
// name, as the reference's files do.
void h5_attr(h5::hid_t loc, const char *name, h5::hid_t type, const void *v, const std::string &file) {
	const h5::Lib &H = h5::lib();
	h5::hid_t sp = H.Screate(h5::S_SCALAR);
	h5::hid_t a = H.Acreate2(loc, name, type, sp, h5::P_DEFAULT, h5::P_DEFAULT);
	if (a < 0 || H.Awrite(a, type, v) < 0) h5_fail(file, std::string("write of attribute '") + name + "'");
	H.Aclose(a);
	H.Sclose(sp);
}

// io.cpp:699-722: a 1-d dataset of the native double type, or a 2-d one of
// a copy of it set little-endian, written from row-major doubles.
void h5_dataset(h5::hid_t loc, const char *name, const double *v, int rank, const h5::hsize_t *dims,
                const std::string &file) {
	const h5::Lib &H = h5::lib();
	h5::hid_t type = H.NATIVE_DOUBLE;
	if (rank == 2) {
		type = H.Tcopy(H.NATIVE_DOUBLE);
		H.Tset_order(type, h5::T_ORDER_LE);
	}
	h5::hid_t sp = H.Screate_simple(rank, dims, nullptr);
	h5::hid_t d = H.Dcreate2(loc, name, type, sp, h5::P_DEFAULT, h5::P_DEFAULT, h5::P_DEFAULT);
	if (d < 0 || H.Dwrite(d, H.NATIVE_DOUBLE, h5::S_ALL, h5::S_ALL, h5::P_DEFAULT, v) < 0)
		h5_fail(file, std::string("write of dataset '") + name + "'");
	H.Dclose(d);
	H.Sclose(sp);
	if (rank == 2) H.Tclose(type);
}

h5::hid_t h5_create(const std::string &file) {
	h5::hid_t f = h5::lib().Fcreate(file.c_str(), h5::ACC_TRUNC, h5::P_DEFAULT, h5::P_DEFAULT);
	if (f < 0) h5_fail(file, "create");
	return f;
}

h5::hid_t h5_group(h5::hid_t f, const std::string &name, const std::string &file) {
	h5::hid_t g = h5::lib().Gcreate2(f, name.c_str(), h5::P_DEFAULT, h5::P_DEFAULT, h5::P_DEFAULT);
	if (g < 0) h5_fail(file, "create of group '" + name + "'");
	return g;
}

std::string jk_group(int r) { return "jk_reg" + std::to_string((long long)r); }

// The LS estimate of every bin of one histogram triple.
std::vector<double> xi_of(const double *DD, const double *DR, const double *RR, int nb) {
	std::vector<double> xi(nb);
	for (int b = 0; b < nb; ++b) xi[b] = ls_estimator(DD[b], DR[b], RR[b]);
	return xi;
}

// io.cpp:762-830: group full_result (attribute Axis_name "s", datasets
// Bin_centre, Bin_width, attribute n_jk_regions, datasets DD, DR, RR, xi),
// then a group jk_reg<k> (DD, DR, RR, xi) per region.
void xi_1d_hdf5(const std::string &fn, JKHists *DD, JKHists *DR, JKHists *RR) {
	const h5::Lib &H = h5::lib();
	Axis &ax = DD->full_hist.ax;
	const int nb = ax.n_bins;
	const h5::hsize_t dims[1] = {(h5::hsize_t)nb};
	h5::hid_t f = h5_create(fn);
	h5::hid_t g = h5_group(f, "full_result", fn);
	h5_attr(g, "Axis_name", H.C_S1, "s", fn);
	std::vector<double> axis(nb);
	for (int i = 0; i < nb; ++i) axis[i] = ax.bin_centre(i);
	h5_dataset(g, "Bin_centre", axis.data(), 1, dims, fn);
	for (int i = 0; i < nb; ++i) axis[i] = ax.bin_width(i);
	h5_dataset(g, "Bin_width", axis.data(), 1, dims, fn);
	h5_attr(g, "n_jk_regions", H.NATIVE_INT, &DD->n_jk_regions, fn);
	h5_dataset(g, "DD", DD->full_hist.bins, 1, dims, fn);
	h5_dataset(g, "DR", DR->full_hist.bins, 1, dims, fn);
	h5_dataset(g, "RR", RR->full_hist.bins, 1, dims, fn);
	h5_dataset(g, "xi", xi_of(DD->full_hist.bins, DR->full_hist.bins, RR->full_hist.bins, nb).data(), 1, dims, fn);
	H.Gclose(g);
	for (int r = 0; r < DD->n_jk_regions; ++r) {
		const size_t o = (size_t)r * nb;
		g = h5_group(f, jk_group(r), fn);
		h5_dataset(g, "DD", DD->sub_hist + o, 1, dims, fn);
		h5_dataset(g, "DR", DR->sub_hist + o, 1, dims, fn);
		h5_dataset(g, "RR", RR->sub_hist + o, 1, dims, fn);
		h5_dataset(g, "xi", xi_of(DD->sub_hist + o, DR->sub_hist + o, RR->sub_hist + o, nb).data(), 1, dims, fn);
		H.Gclose(g);
	}
	H.Fclose(f);
}

// io.cpp:882-971: group full_result (attributes Axis_1_name, Axis_2_name,
// datasets Axis_1/2_bin_centre and _bin_width, attribute n_jk_regions, 2-d
// datasets DD, DR, RR, xi of n1 x n2), then jk_reg<k> groups (DD, DR, RR, xi).
void xi_2d_hdf5(const std::string &fn, const std::string &name1, const std::string &name2, JKHists2D *DD,
                JKHists2D *DR, JKHists2D *RR) {
	const h5::Lib &H = h5::lib();
	Axis &a1 = DD->full_hist.ax1, &a2 = DD->full_hist.ax2;
	const int nb = DD->full_hist.total_bins;
	const h5::hsize_t dims[2] = {(h5::hsize_t)a1.n_bins, (h5::hsize_t)a2.n_bins};
	h5::hid_t f = h5_create(fn);
	h5::hid_t g = h5_group(f, "full_result", fn);
	h5_attr(g, "Axis_1_name", H.C_S1, name1.c_str(), fn);
	h5_attr(g, "Axis_2_name", H.C_S1, name2.c_str(), fn);
	Axis *axes[2] = {&a1, &a2};
	const char *names[2][2] = {{"Axis_1_bin_centre", "Axis_1_bin_width"}, {"Axis_2_bin_centre", "Axis_2_bin_width"}};
	for (int k = 0; k < 2; ++k) {
		const int n = axes[k]->n_bins;
		const h5::hsize_t d1[1] = {(h5::hsize_t)n};
		std::vector<double> axis(n);
		for (int i = 0; i < n; ++i) axis[i] = axes[k]->bin_centre(i);
		h5_dataset(g, names[k][0], axis.data(), 1, d1, fn);
		for (int i = 0; i < n; ++i) axis[i] = axes[k]->bin_width(i);
		h5_dataset(g, names[k][1], axis.data(), 1, d1, fn);
	}
	h5_attr(g, "n_jk_regions", H.NATIVE_INT, &DD->n_jk_regions, fn);
	h5_dataset(g, "DD", DD->full_hist.bins, 2, dims, fn);
	h5_dataset(g, "DR", DR->full_hist.bins, 2, dims, fn);
	h5_dataset(g, "RR", RR->full_hist.bins, 2, dims, fn);
	h5_dataset(g, "xi", xi_of(DD->full_hist.bins, DR->full_hist.bins, RR->full_hist.bins, nb).data(), 2, dims, fn);
	H.Gclose(g);
	for (int r = 0; r < DD->n_jk_regions; ++r) {
		const size_t o = (size_t)r * nb;
		g = h5_group(f, jk_group(r), fn);
		h5_dataset(g, "DD", DD->sub_hist + o, 2, dims, fn);
		h5_dataset(g, "DR", DR->sub_hist + o, 2, dims, fn);
		h5_dataset(g, "RR", RR->sub_hist + o, 2, dims, fn);
		h5_dataset(g, "xi", xi_of(DD->sub_hist + o, DR->sub_hist + o, RR->sub_hist + o, nb).data(), 2, dims, fn);
		H.Gclose(g);
	}
	H.Fclose(f);
}

}  // namespace

// io.cpp:741-834
void output_xi_1D(std::string fn, std::string type, JKHists *DD, JKHists *DR, JKHists *RR, GalaxyCatalogue *,
                  GalaxyCatalogue *) {
	if (type == "hdf5") {
		xi_1d_hdf5(fn, DD, DR, RR);
		return;
	}
	if (type != "ascii") {
		std::cout << "Output file type not found\n";
		return;
	}
	const int nb = DD->full_hist.ax.n_bins;
	single_1d(fn, &DD->full_hist.ax, DD->full_hist.bins, DR->full_hist.bins, RR->full_hist.bins);
	for (int r = 0; r < DD->n_jk_regions; ++r)
		single_1d(generate_sub_filename(fn, r), &DD->full_hist.ax, &DD->sub_hist[r * nb], &DR->sub_hist[r * nb],
		          &RR->sub_hist[r * nb]);
}

// io.cpp:861-975
void output_xi_2D(std::string fn, std::string type, std::string name1, std::string name2, JKHists2D *DD,
                  JKHists2D *DR, JKHists2D *RR, GalaxyCatalogue *, GalaxyCatalogue *) {
	if (type == "hdf5") {
		xi_2d_hdf5(fn, name1, name2, DD, DR, RR);
		return;
	}
	if (type != "ascii") {
		std::cout << "File type not found\n";
		return;
	}
	const int nb = DD->full_hist.total_bins;
	single_2d(fn, &DD->full_hist.ax1, &DD->full_hist.ax2, DD->full_hist.bins, DR->full_hist.bins, RR->full_hist.bins);
	for (int r = 0; r < DD->n_jk_regions; ++r)
		single_2d(generate_sub_filename(fn, r), &DD->full_hist.ax1, &DD->full_hist.ax2, &DD->sub_hist[r * nb],
		          &DR->sub_hist[r * nb], &RR->sub_hist[r * nb]);
}

// hist.cpp:180-221: min, max and base are written as float attributes
void JKHists::output_hdf5(std::string fn) {
	const h5::Lib &H = h5::lib();
	const int nb = full_hist.ax.n_bins;
	const h5::hsize_t dims[1] = {(h5::hsize_t)nb};
	h5::hid_t f = h5_create(fn);
	h5::hid_t g = h5_group(f, "full_result", fn);
	h5_attr(g, "n_jk_regions", H.NATIVE_INT, &n_jk_regions, fn);
	h5_attr(g, "n_bins", H.NATIVE_INT, &full_hist.ax.n_bins, fn);
	h5_attr(g, "min", H.NATIVE_FLOAT, &full_hist.ax.min, fn);
	h5_attr(g, "max", H.NATIVE_FLOAT, &full_hist.ax.max, fn);
	h5_attr(g, "base", H.NATIVE_FLOAT, &full_hist.ax.base, fn);
	h5_dataset(g, "hist", full_hist.bins, 1, dims, fn);
	std::vector<double> axis(nb);
	for (int i = 0; i < nb; ++i) axis[i] = full_hist.ax.bin_centre(i);
	h5_dataset(g, "Bin_centre", axis.data(), 1, dims, fn);
	for (int i = 0; i < nb; ++i) axis[i] = full_hist.ax.bin_width(i);
	h5_dataset(g, "Bin_width", axis.data(), 1, dims, fn);
	H.Gclose(g);
	for (int r = 0; r < n_jk_regions; ++r) {
		g = h5_group(f, jk_group(r), fn);
		h5_dataset(g, "hist", sub_hist + (size_t)r * nb, 1, dims, fn);
		H.Gclose(g);
	}
	H.Fclose(f);
}

// hist.cpp:115-178: the attributes are read as int / double (the float
// attributes convert exactly) and the Hist is built from them
void read_hdf5_hist(std::string fn, JKHists **hist) {
	std::cout << "\nReading file: " << fn << "\n";
	const h5::Lib &H = h5::lib();
	h5::hid_t f = H.Fopen(fn.c_str(), h5::ACC_RDONLY, h5::P_DEFAULT);
	if (f < 0) h5_fail(fn, "open");
	h5::hid_t g = H.Gopen2(f, "full_result", h5::P_DEFAULT);
	if (g < 0) h5_fail(fn, "open of group 'full_result'");
	auto attr = [&](const char *name, h5::hid_t type, void *v) {
		h5::hid_t a = H.Aopen(g, name, h5::P_DEFAULT);
		if (a < 0 || H.Aread(a, type, v) < 0) h5_fail(fn, std::string("read of attribute '") + name + "'");
		H.Aclose(a);
	};
	int n_jk = 0, n_bins = 0;
	double mn = 0, mx = 0, base = 0;
	attr("n_jk_regions", H.NATIVE_INT, &n_jk);
	attr("n_bins", H.NATIVE_INT, &n_bins);
	attr("min", H.NATIVE_DOUBLE, &mn);
	attr("max", H.NATIVE_DOUBLE, &mx);
	attr("base", H.NATIVE_DOUBLE, &base);
	if (n_bins <= 0 || n_jk < 0) h5_fail(fn, "read (bad n_bins or n_jk_regions)");
	Hist t((float)mn, (float)mx, n_bins, (float)base);
	JKHists *h = new JKHists(n_jk, t);
	h5_read_column(g, fn, "hist", H.NATIVE_DOUBLE, h->full_hist.bins, n_bins);
	H.Gclose(g);
	for (int r = 0; r < n_jk; ++r) {
		g = H.Gopen2(f, jk_group(r).c_str(), h5::P_DEFAULT);
		if (g < 0) h5_fail(fn, "open of group '" + jk_group(r) + "'");
		h5_read_column(g, fn, "hist", H.NATIVE_DOUBLE, h->sub_hist + (size_t)r * n_bins, n_bins);
		H.Gclose(g);
	}
	H.Fclose(f);
	*hist = h;
}
