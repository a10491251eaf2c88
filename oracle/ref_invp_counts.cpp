// ref_invp_counts — a driver of this project's own that links the two_pcf
// reference sources compiled with _USE_INV_WEIGHTS (and, in one variant,
// _USE_DITHER_WEIGHTS) and _HAVE_HDF5 (see oracle/Makefile), and prints what
// the inverse-p pipeline computes, at full precision, as JSON:
//
//   - each catalogue as read by the reference's hdf5 reader (io.cpp:417-590):
//     positions, unit vectors, the IIP-adjusted weight and the input weight;
//   - the four angular histograms calc_angular_upweight writes
//     (correlators_invpweight.cpp:985-1033), read back with read_hdf5_hist
//     (hist.cpp:115-178);
//   - the up-weight tables DD/DDinvp and DR/DRinvp (divide_by_hist,
//     hist.cpp:224-231) and JKHists::interp (hist.cpp:369-391) of both at the
//     probe values of an optional file (one double per line, C99 hex or
//     decimal);
//   - for each correlator the parameter file asks for, the raw histograms of
//     auto_pair_count_*_invpweights (DD), cross_pair_count_*_invpweights (DR)
//     and auto_pair_count_* (RR), as auto_correlate_*_invpweights
//     (correlators_invpweight.cpp:210-233, 447-470, 684-706) leaves them
//     before normalise(), then the normalised ones.
//
// With "--main" instead of an output file it runs main.cpp:49-62's sequence
// and writes the output files the parameter file names.  A probe file of
// "none" evaluates no probes; a fourth argument 0 leaves the catalogues out
// of the JSON (timing runs).
//
// TEST INFRASTRUCTURE ONLY.  Used in this container to write golden fixtures
// (tests/golden/) and, on the GPU host, as the reference CPU baseline of the
// inverse-p workload.  Nothing in the product links or runs it.

#include "io.h"
#include "hist.h"
#include "boxes.h"
#include "galaxy.h"
#include "cosmology.h"
#include "correlators.h"
#include "correlators_invpweight.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef _USE_INV_WEIGHTS
#error "ref_invp_counts needs the _USE_INV_WEIGHTS build"
#endif

static void put_arr(FILE *f, const double *v, long n) {
	fputc('[', f);
	for (long i = 0; i < n; ++i) {
		if (std::isfinite(v[i])) fprintf(f, "%s%.17g", i ? "," : "", v[i]);
		else fprintf(f, "%s\"%s\"", i ? "," : "", std::isnan(v[i]) ? "nan" : (v[i] > 0 ? "inf" : "-inf"));
	}
	fputc(']', f);
}

static void put_jk1(FILE *f, const char *name, const JKHists &h) {
	fprintf(f, "\"%s\": {\"full\": ", name);
	put_arr(f, h.full_hist.bins, h.full_hist.ax.n_bins);
	fprintf(f, ", \"sub\": ");
	put_arr(f, h.sub_hist, (long)h.n_jk_regions * h.full_hist.ax.n_bins);
	fputc('}', f);
}

static void put_jk2(FILE *f, const char *name, const JKHists2D &h) {
	fprintf(f, "\"%s\": {\"full\": ", name);
	put_arr(f, h.full_hist.bins, h.full_hist.total_bins);
	fprintf(f, ", \"sub\": ");
	put_arr(f, h.sub_hist, (long)h.n_jk_regions * h.full_hist.total_bins);
	fputc('}', f);
}

static double now_s() {
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void put_cat(FILE *f, const char *nm, const GalaxyCatalogue *c) {
	fprintf(f, ", \"%s_cat\": {", nm);
	const char *fld[9] = {"x", "y", "z", "nx", "ny", "nz", "d", "w", "iw"};
	for (int k = 0; k < 9; ++k) {
		fprintf(f, "%s\"%s\": [", k ? ", " : "", fld[k]);
		for (int i = 0; i < c->n_gals; ++i) {
			const Galaxy &g = c->gals[i];
			double v = k < 3 ? g.x[k] : k < 6 ? g.x_norm[k - 3] : k == 6 ? g.d_comov
			         : k == 7 ? (double)g.weight : (double)g.input_weight;
			fprintf(f, "%s\"%a\"", i ? "," : "", v);
		}
		fputc(']', f);
	}
	fputc('}', f);
}

// read_hdf5_hist of one angular file; NULL if the file was not asked for
static JKHists *read_angular(bool asked, const std::string &fn) {
	if (!asked) return NULL;
	JKHists *h = NULL;
	read_hdf5_hist(fn, &h);
	return h;
}

int main(int argc, char **argv) {
	if (argc < 3) {
		fprintf(stderr, "usage: ref_invp_counts <param-file> <out.json | --main> [probe-file|none] [dump_catalog=1|0]\n");
		return 1;
	}
	Params params(argv[1]);
#ifdef _OPENMP
	if (params.n_threads <= 0) params.n_threads = omp_get_max_threads();
	omp_set_num_threads(params.n_threads);
#endif
	Cosmology *cosmo = NULL;
	if (params.coord_system == "equatorial")
		cosmo = new Cosmology(params.omega_m, params.h, params.z_min, params.z_max, 0.0001, 100);
	GalaxyCatalogue *data = read_file(params.data_filename, params.data_file_type, cosmo, params);
	GalaxyCatalogue *randoms = read_file(params.random_filename, params.random_file_type, cosmo, params);
	delete cosmo;

	if (std::strcmp(argv[2], "--main") == 0) {
		// main.cpp:49-62 under _USE_INV_WEIGHTS
		calc_angular_upweight(data, randoms, params);
		if (params.plot_monopole) auto_correlate_monopole_invpweights(data, randoms, params);
		if (params.plot_sigma_pi) auto_correlate_sigma_pi_invpweights(data, randoms, params);
		if (params.plot_s_mu) auto_correlate_s_mu_invpweights(data, randoms, params);
		return 0;
	}

	FILE *f = fopen(argv[2], "w");
	fprintf(f, "{\"n_threads\": %d, \"n_data\": %d, \"n_randoms\": %d, \"n_mask_ints\": %d",
		params.n_threads, data->n_gals, randoms->n_gals, (int)_N_MASK_INTS);
	// catalogues as read, before GalaxyPixels / GalaxyBoxes reorder them
	if (argc <= 4 || std::atoi(argv[4]) != 0) {
		put_cat(f, "data", data);
		put_cat(f, "randoms", randoms);
	}

	double t0 = now_s();
	calc_angular_upweight(data, randoms, params);
	fprintf(f, ", \"angular_seconds\": %.6f", now_s() - t0);
	JKHists *add = read_angular(params.calculate_angular_dd, params.angular_dd_filename);
	JKHists *addi = read_angular(params.calculate_angular_dd_invpweights, params.angular_dd_invpweights_filename);
	JKHists *adr = read_angular(params.calculate_angular_dr, params.angular_dr_filename);
	JKHists *adri = read_angular(params.calculate_angular_dr_invpweights, params.angular_dr_invpweights_filename);
	fputs(", \"angular\": {", f);
	const char *anames[4] = {"DD", "DDinvp", "DR", "DRinvp"};
	JKHists *ah[4] = {add, addi, adr, adri};
	bool first = true;
	for (int k = 0; k < 4; ++k) {
		if (!ah[k]) continue;
		if (!first) fputs(", ", f);
		put_jk1(f, anames[k], *ah[k]);
		first = false;
	}
	fputc('}', f);

	// up-weights and their interpolation (as the 3-D correlators form them)
	std::vector<double> probes;
	if (argc > 3 && std::strcmp(argv[3], "none") != 0) {
		FILE *pf = fopen(argv[3], "r");
		char line[128];
		while (pf && fgets(line, sizeof line, pf)) probes.push_back(strtod(line, NULL));
		if (pf) fclose(pf);
	}
	JKHists *ups[2] = {NULL, NULL};
	if (add && addi) { ups[0] = add; add->divide_by_hist(*addi); }
	if (adr && adri) { ups[1] = adr; adr->divide_by_hist(*adri); }
	const char *unames[2] = {"up_dd", "up_dr"};
	for (int k = 0; k < 2; ++k) {
		if (!ups[k]) continue;
		fprintf(f, ", \"%s\": {\"bins\": ", unames[k]);
		put_arr(f, ups[k]->full_hist.bins, ups[k]->full_hist.ax.n_bins);
		std::vector<double> v(probes.size());
		for (size_t i = 0; i < probes.size(); ++i) v[i] = ups[k]->interp(probes[i]);
		fputs(", \"interp\": ", f);
		put_arr(f, v.data(), (long)v.size());
		fputc('}', f);
	}
	fputs(", \"probes\": [", f);
	for (size_t i = 0; i < probes.size(); ++i) fprintf(f, "%s\"%a\"", i ? "," : "", probes[i]);
	fputc(']', f);
	for (int k = 0; k < 4; ++k) delete ah[k];

	if (params.plot_monopole) {
		Hist t(params.monopole_min, params.monopole_max, params.monopole_n_bins, params.monopole_log_base);
		JKHists DD(params.n_jk_regions, t), DR(params.n_jk_regions, t), RR(params.n_jk_regions, t);
		GalaxyBoxes db(data->gals, data->n_gals, randoms->gals, randoms->n_gals, params.monopole_max);
		GalaxyBoxes rb(randoms->gals, randoms->n_gals, data->gals, data->n_gals, params.monopole_max);
		double a = now_s();
		auto_pair_count_monopole_invpweights(&db, &DD, params);
		double b = now_s();
		cross_pair_count_monopole_invpweights(&db, &rb, &DR, params);
		double c = now_s();
		auto_pair_count_monopole(&rb, &RR);
		double d = now_s();
		fprintf(f, ", \"monopole\": {\"n_data_used\": %d, \"n_randoms_used\": %d, \"seconds\": [%.6f, %.6f, %.6f], ",
			db.n_gals, rb.n_gals, b - a, c - b, d - c);
		put_jk1(f, "DD", DD); fputs(", ", f); put_jk1(f, "DR", DR); fputs(", ", f); put_jk1(f, "RR", RR);
		DD.normalise(data, db.n_gals);
		DR.normalise(data, db.n_gals, randoms, rb.n_gals);
		RR.normalise(randoms, rb.n_gals);
		fputs(", \"norm\": {", f);
		put_jk1(f, "DD", DD); fputs(", ", f); put_jk1(f, "DR", DR); fputs(", ", f); put_jk1(f, "RR", RR);
		fputs("}}", f);
	}
	if (params.plot_sigma_pi) {
		Hist2D t(params.sigma_min, params.sigma_max, params.sigma_n_bins, params.sigma_log_base,
			params.pi_min, params.pi_max, params.pi_n_bins, params.pi_log_base);
		JKHists2D DD(params.n_jk_regions, t), DR(params.n_jk_regions, t), RR(params.n_jk_regions, t);
		double rmax = sqrt(params.sigma_max * params.sigma_max + params.pi_max * params.pi_max);
		GalaxyBoxes db(data->gals, data->n_gals, randoms->gals, randoms->n_gals, rmax);
		GalaxyBoxes rb(randoms->gals, randoms->n_gals, data->gals, data->n_gals, rmax);
		double a = now_s();
		auto_pair_count_sigma_pi_invpweights(&db, &DD, params);
		double b = now_s();
		cross_pair_count_sigma_pi_invpweights(&db, &rb, &DR, params);
		double c = now_s();
		auto_pair_count_sigma_pi(&rb, &RR);
		double d = now_s();
		fprintf(f, ", \"sigma_pi\": {\"n_data_used\": %d, \"n_randoms_used\": %d, \"seconds\": [%.6f, %.6f, %.6f], ",
			db.n_gals, rb.n_gals, b - a, c - b, d - c);
		put_jk2(f, "DD", DD); fputs(", ", f); put_jk2(f, "DR", DR); fputs(", ", f); put_jk2(f, "RR", RR);
		DD.normalise(data, db.n_gals);
		DR.normalise(data, db.n_gals, randoms, rb.n_gals);
		RR.normalise(randoms, rb.n_gals);
		fputs(", \"norm\": {", f);
		put_jk2(f, "DD", DD); fputs(", ", f); put_jk2(f, "DR", DR); fputs(", ", f); put_jk2(f, "RR", RR);
		fputs("}}", f);
	}
	if (params.plot_s_mu) {
		Hist2D t(params.s_min, params.s_max, params.s_n_bins, params.s_log_base, 0.0, 1.0, params.mu_n_bins, 0);
		JKHists2D DD(params.n_jk_regions, t), DR(params.n_jk_regions, t), RR(params.n_jk_regions, t);
		GalaxyBoxes db(data->gals, data->n_gals, randoms->gals, randoms->n_gals, params.s_max);
		GalaxyBoxes rb(randoms->gals, randoms->n_gals, data->gals, data->n_gals, params.s_max);
		double a = now_s();
		auto_pair_count_s_mu_invpweights(&db, &DD, params);
		double b = now_s();
		cross_pair_count_s_mu_invpweights(&db, &rb, &DR, params);
		double c = now_s();
		auto_pair_count_s_mu(&rb, &RR);
		double d = now_s();
		fprintf(f, ", \"s_mu\": {\"n_data_used\": %d, \"n_randoms_used\": %d, \"seconds\": [%.6f, %.6f, %.6f], ",
			db.n_gals, rb.n_gals, b - a, c - b, d - c);
		put_jk2(f, "DD", DD); fputs(", ", f); put_jk2(f, "DR", DR); fputs(", ", f); put_jk2(f, "RR", RR);
		DD.normalise(data, db.n_gals);
		DR.normalise(data, db.n_gals, randoms, rb.n_gals);
		RR.normalise(randoms, rb.n_gals);
		fputs(", \"norm\": {", f);
		put_jk2(f, "DD", DD); fputs(", ", f); put_jk2(f, "DR", DR); fputs(", ", f); put_jk2(f, "RR", RR);
		fputs("}}", f);
	}
	fputs("}\n", f);
	fclose(f);
	return 0;
}
