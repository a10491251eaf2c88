// ref_counts — a driver of this project's own that links the two_pcf
// reference sources (compiled where they lie, see oracle/Makefile) and prints
// the raw pair-count histograms at full double precision as JSON.
//
// TEST INFRASTRUCTURE ONLY.  It is used in this container to write golden
// fixtures (tests/golden/) and, as `cpu_baseline.kind = "reference"`, to time
// the reference's OpenMP path on a bounded sample.  Nothing in the product
// links or runs it.
//
// The call sequence mirrors auto_correlate_{monopole,sigma_pi,s_mu}
// (correlators.cpp:127-154, 280-307, 432-458) up to, but not including,
// normalise(): the histograms printed are exactly what
// auto_/cross_pair_count_* leave behind (already inverted and, for auto
// counts, multiplied by 2).  The normalised values are printed too.

#include "io.h"
#include "hist.h"
#include "boxes.h"
#include "galaxy.h"
#include "cosmology.h"
#include "correlators.h"
#include <cstdio>
#include <cstring>
#include <cmath>
#include <string>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

static void put_axis(FILE *f, const char *name, const Axis &a) {
	fprintf(f, "\"%s\": {\"n_bins\": %d, \"log\": %d, \"min\": \"%a\", \"max\": \"%a\", "
		"\"step\": \"%a\", \"inv_step\": \"%a\", \"ilb\": \"%a\", \"ils\": \"%a\"}",
		name, a.n_bins, (int)a.log_binning, (double)a.min, (double)a.max, (double)a.step,
		(double)a.inv_step, (double)a.inv_log_of_base, (double)a.inv_log_step);
}

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

int main(int argc, char **argv) {
	if (argc < 3) {
		fprintf(stderr, "usage: ref_counts <param-file> <out.json> [dump_catalog=0|1] [normalised=0|1]\n");
		return 1;
	}
	int dump_cat = argc > 3 ? atoi(argv[3]) : 0;
	int want_norm = argc > 4 ? atoi(argv[4]) : 1;
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

	FILE *f = fopen(argv[2], "w");
	fprintf(f, "{\"n_threads\": %d, \"n_data\": %d, \"n_randoms\": %d", params.n_threads, data->n_gals, randoms->n_gals);
	if (dump_cat) {
		// Catalogue as read (before any box reordering): positions and unit vectors.
		const GalaxyCatalogue *cats[2] = {data, randoms};
		const char *nm[2] = {"data", "randoms"};
		for (int c = 0; c < 2; ++c) {
			fprintf(f, ", \"%s_cat\": {", nm[c]);
			const char *fld[7] = {"x", "y", "z", "nx", "ny", "nz", "d"};
			for (int k = 0; k < 7; ++k) {
				fprintf(f, "%s\"%s\": [", k ? ", " : "", fld[k]);
				for (int i = 0; i < cats[c]->n_gals; ++i) {
					const Galaxy &g = cats[c]->gals[i];
					double v = k < 3 ? g.x[k] : (k < 6 ? g.x_norm[k - 3] : g.d_comov);
					fprintf(f, "%s\"%a\"", i ? "," : "", v);
				}
				fputc(']', f);
			}
			fputc('}', f);
		}
	}

	if (params.plot_monopole) {
		Hist t(params.monopole_min, params.monopole_max, params.monopole_n_bins, params.monopole_log_base);
		JKHists DD(params.n_jk_regions, t), DR(params.n_jk_regions, t), RR(params.n_jk_regions, t);
		GalaxyBoxes db(data->gals, data->n_gals, randoms->gals, randoms->n_gals, params.monopole_max);
		GalaxyBoxes rb(randoms->gals, randoms->n_gals, data->gals, data->n_gals, params.monopole_max);
		double t0 = now_s();
		auto_pair_count_monopole(&db, &DD);
		double t1 = now_s();
		cross_pair_count_monopole(&db, &rb, &DR);
		double t2 = now_s();
		auto_pair_count_monopole(&rb, &RR);
		double t3 = now_s();
		fprintf(f, ", \"monopole\": {\"n_data_used\": %d, \"n_randoms_used\": %d, \"seconds\": [%.6f, %.6f, %.6f], ",
			db.n_gals, rb.n_gals, t1 - t0, t2 - t1, t3 - t2);
		put_axis(f, "axis", DD.full_hist.ax);
		fputs(", ", f); put_jk1(f, "DD", DD);
		fputs(", ", f); put_jk1(f, "DR", DR);
		fputs(", ", f); put_jk1(f, "RR", RR);
		if (want_norm) {
			DD.normalise(data, db.n_gals);
			DR.normalise(data, db.n_gals, randoms, rb.n_gals);
			RR.normalise(randoms, rb.n_gals);
			fputs(", \"norm\": {", f);
			put_jk1(f, "DD", DD); fputs(", ", f); put_jk1(f, "DR", DR); fputs(", ", f); put_jk1(f, "RR", RR);
			fputc('}', f);
		}
		fputc('}', f);
	}
	if (params.plot_sigma_pi) {
		Hist2D t(params.sigma_min, params.sigma_max, params.sigma_n_bins, params.sigma_log_base,
			params.pi_min, params.pi_max, params.pi_n_bins, params.pi_log_base);
		JKHists2D DD(params.n_jk_regions, t), DR(params.n_jk_regions, t), RR(params.n_jk_regions, t);
		double rmax = sqrt(params.sigma_max * params.sigma_max + params.pi_max * params.pi_max);
		GalaxyBoxes db(data->gals, data->n_gals, randoms->gals, randoms->n_gals, rmax);
		GalaxyBoxes rb(randoms->gals, randoms->n_gals, data->gals, data->n_gals, rmax);
		double t0 = now_s();
		auto_pair_count_sigma_pi(&db, &DD);
		double t1 = now_s();
		cross_pair_count_sigma_pi(&db, &rb, &DR);
		double t2 = now_s();
		auto_pair_count_sigma_pi(&rb, &RR);
		double t3 = now_s();
		fprintf(f, ", \"sigma_pi\": {\"n_data_used\": %d, \"n_randoms_used\": %d, \"seconds\": [%.6f, %.6f, %.6f], ",
			db.n_gals, rb.n_gals, t1 - t0, t2 - t1, t3 - t2);
		put_axis(f, "axis1", DD.full_hist.ax1); fputs(", ", f);
		put_axis(f, "axis2", DD.full_hist.ax2);
		fputs(", ", f); put_jk2(f, "DD", DD);
		fputs(", ", f); put_jk2(f, "DR", DR);
		fputs(", ", f); put_jk2(f, "RR", RR);
		if (want_norm) {
			DD.normalise(data, db.n_gals);
			DR.normalise(data, db.n_gals, randoms, rb.n_gals);
			RR.normalise(randoms, rb.n_gals);
			fputs(", \"norm\": {", f);
			put_jk2(f, "DD", DD); fputs(", ", f); put_jk2(f, "DR", DR); fputs(", ", f); put_jk2(f, "RR", RR);
			fputc('}', f);
		}
		fputc('}', f);
	}
	if (params.plot_s_mu) {
		Hist2D t(params.s_min, params.s_max, params.s_n_bins, params.s_log_base, 0.0, 1.0, params.mu_n_bins, 0);
		JKHists2D DD(params.n_jk_regions, t), DR(params.n_jk_regions, t), RR(params.n_jk_regions, t);
		GalaxyBoxes db(data->gals, data->n_gals, randoms->gals, randoms->n_gals, params.s_max);
		GalaxyBoxes rb(randoms->gals, randoms->n_gals, data->gals, data->n_gals, params.s_max);
		double t0 = now_s();
		auto_pair_count_s_mu(&db, &DD);
		double t1 = now_s();
		cross_pair_count_s_mu(&db, &rb, &DR);
		double t2 = now_s();
		auto_pair_count_s_mu(&rb, &RR);
		double t3 = now_s();
		fprintf(f, ", \"s_mu\": {\"n_data_used\": %d, \"n_randoms_used\": %d, \"seconds\": [%.6f, %.6f, %.6f], ",
			db.n_gals, rb.n_gals, t1 - t0, t2 - t1, t3 - t2);
		put_axis(f, "axis1", DD.full_hist.ax1); fputs(", ", f);
		put_axis(f, "axis2", DD.full_hist.ax2);
		fputs(", ", f); put_jk2(f, "DD", DD);
		fputs(", ", f); put_jk2(f, "DR", DR);
		fputs(", ", f); put_jk2(f, "RR", RR);
		if (want_norm) {
			DD.normalise(data, db.n_gals);
			DR.normalise(data, db.n_gals, randoms, rb.n_gals);
			RR.normalise(randoms, rb.n_gals);
			fputs(", \"norm\": {", f);
			put_jk2(f, "DD", DD); fputs(", ", f); put_jk2(f, "DR", DR); fputs(", ", f); put_jk2(f, "RR", RR);
			fputc('}', f);
		}
		fputc('}', f);
	}
	fputs("}\n", f);
	fclose(f);
	return 0;
}
