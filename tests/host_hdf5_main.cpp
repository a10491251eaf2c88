// Test harness for the host layer's HDF5 writers and histogram reader.
//   host_hdf5 <dir> <n_jk>
// Fills JKHists / JKHists2D with fixed values (some RR bins zero), writes
// <dir>/mono.h5 (output_xi_1D), <dir>/sp.h5 (output_xi_2D, sigma-pi),
// <dir>/smu.h5 (output_xi_2D, s-mu) and <dir>/ang.h5 (JKHists::output_hdf5),
// reads ang.h5 back with read_hdf5_hist, and prints every array written and
// read back as "<name> <hex double> ...", one array per line.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../two_pcf_amd/csrc/host/twopcf_ref_api.hpp"

static void put(const char *name, const double *v, int n) {
	std::printf("%s", name);
	for (int i = 0; i < n; ++i) std::printf(" %a", v[i]);
	std::printf("\n");
}

static double value(int tag, int k) {
	// zero on every seventh bin of RR (tag 2), so xi takes its RR == 0 branch
	if (tag == 2 && k % 7 == 3) return 0.0;
	return (double)((k * 7919 + tag * 104729) % 1009) / 13.0 + 0.5 * tag;
}

template <class H>
static void fill(H &h, int tag, int nb) {
	for (int b = 0; b < nb; ++b) h.full_hist.bins[b] = value(tag, b);
	for (int k = 0; k < h.n_jk_regions * nb; ++k) h.sub_hist[k] = value(tag, nb + k);
}

static void put_axis(const char *tag, Axis &ax) {
	std::string c = std::string(tag) + "_centre", w = std::string(tag) + "_width";
	double *vc = new double[ax.n_bins], *vw = new double[ax.n_bins];
	for (int i = 0; i < ax.n_bins; ++i) {
		vc[i] = ax.bin_centre(i);
		vw[i] = ax.bin_width(i);
	}
	put(c.c_str(), vc, ax.n_bins);
	put(w.c_str(), vw, ax.n_bins);
	delete[] vc;
	delete[] vw;
}

template <class H>
static void put_hists(const char *stem, H *hs[3], int nb) {
	const char *names[3] = {"DD", "DR", "RR"};
	for (int t = 0; t < 3; ++t) {
		std::string s = std::string(stem) + "_" + names[t];
		put((s + "_full").c_str(), hs[t]->full_hist.bins, nb);
		put((s + "_sub").c_str(), hs[t]->sub_hist, hs[t]->n_jk_regions * nb);
	}
}

int main(int argc, char **argv) {
	if (argc != 3) return 2;
	const std::string dir = argv[1];
	const int n_jk = std::atoi(argv[2]);

	Hist h1(0.0f, 60.0f, 12, 1.3f);
	JKHists DD(n_jk, h1), DR(n_jk, h1), RR(n_jk, h1);
	JKHists *m[3] = {&DD, &DR, &RR};
	for (int t = 0; t < 3; ++t) fill(*m[t], t, 12);
	output_xi_1D(dir + "/mono.h5", "hdf5", &DD, &DR, &RR, nullptr, nullptr);
	put_axis("mono", DD.full_hist.ax);
	put_hists("mono", m, 12);

	Hist2D sp(0.0f, 40.0f, 8, 1.0f, 0.0f, 40.0f, 5, 1.0f);
	Hist2D smu(0.0f, 100.0f, 9, 1.3f, 0.0f, 1.0f, 4, 0.0f);
	Hist2D *tpl[2] = {&sp, &smu};
	const char *stems[2] = {"sp", "smu"};
	const char *names[2][2] = {{"sigma", "pi"}, {"s", "mu"}};
	for (int k = 0; k < 2; ++k) {
		JKHists2D A(n_jk, *tpl[k]), B(n_jk, *tpl[k]), Cc(n_jk, *tpl[k]);
		JKHists2D *h[3] = {&A, &B, &Cc};
		const int nb = A.full_hist.total_bins;
		for (int t = 0; t < 3; ++t) fill(*h[t], t, nb);
		output_xi_2D(dir + "/" + stems[k] + ".h5", "hdf5", names[k][0], names[k][1], &A, &B, &Cc, nullptr, nullptr);
		put_axis((std::string(stems[k]) + "1").c_str(), A.full_hist.ax1);
		put_axis((std::string(stems[k]) + "2").c_str(), A.full_hist.ax2);
		put_hists(stems[k], h, nb);
	}

	Hist ha(0.0f, 1.0f - 0.98472f, 40, 1.6f);
	JKHists ang(n_jk, ha);
	fill(ang, 5, 40);
	ang.output_hdf5(dir + "/ang.h5");
	put_axis("ang", ang.full_hist.ax);
	put("ang_full", ang.full_hist.bins, 40);
	put("ang_sub", ang.sub_hist, n_jk * 40);
	const double axc[3] = {ang.full_hist.ax.min, ang.full_hist.ax.max, ang.full_hist.ax.base};
	put("ang_axis", axc, 3);

	JKHists *back = nullptr;
	read_hdf5_hist(dir + "/ang.h5", &back);
	std::printf("back_n %d %d\n", back->n_jk_regions, back->full_hist.ax.n_bins);
	put("back_full", back->full_hist.bins, back->full_hist.ax.n_bins);
	put("back_sub", back->sub_hist, back->n_jk_regions * back->full_hist.ax.n_bins);
	Axis &bx = back->full_hist.ax;
	const double bc[8] = {bx.min, bx.max, bx.base, bx.step, bx.inv_step, bx.inv_log_of_base, bx.inv_log_step,
	                      (double)bx.log_binning};
	put("back_axis", bc, 8);
	Axis &ox = ang.full_hist.ax;
	const double oc[8] = {ox.min, ox.max, ox.base, ox.step, ox.inv_step, ox.inv_log_of_base, ox.inv_log_step,
	                      (double)ox.log_binning};
	put("orig_axis", oc, 8);
	delete back;
	return 0;
}
