// twopcf_ref_api.hpp — the reference's host-side interface (galaxy.h, hist.h,
// boxes.h, io.h, correlators.h) restated on top of libtwopcf_hip.
//
// The names, argument meanings and post-conditions follow the reference so
// that a caller of auto_/cross_pair_count_* or auto_correlate_* can switch
// without changes; the pair loops themselves run on the GPU through the C ABI
// in include/twopcf_hip.h.  Differences from the reference, on purpose:
//   * GalaxyBoxes does not cell-sort the caller's catalogue (the GPU grids its
//     own copy); it only moves objects with d_comov < 0 behind n_gals, as
//     boxes.cpp:16-26 does, which is all normalise() relies on.
//   * Params members are zero-initialised instead of left undefined when a
//     key is missing (io.cpp:16-237 leaves them uninitialised).
//   * Catalogues and outputs are ascii or hdf5, as in a reference build with
//     _HAVE_HDF5; libhdf5 is opened at run time on the first hdf5 file
//     (h5lib.hpp), so ascii runs never need it.  In the inverse-p build the
//     hdf5 reader takes the bit masks from the datasets <name>0, <name>1, ...
//     as the reference does, and the ascii reader takes them as
//     ceil(n_bitwise_runs / 64) extra integer columns.
//   * The inverse-p build (_USE_INV_WEIGHTS) is a run-time switch,
//     twopcf_enable_inverse_p(), set by the twopcf_invp driver, and so are
//     dither weights (_USE_DITHER_WEIGHTS), twopcf_enable_dither_weights(),
//     which the driver sets when TWOPCF_DITHER=1.  The angular histograms
//     calc_angular_upweight produces are written to their HDF5 files as in
//     the reference and also kept in the process under their file names, so
//     the 3-D counts of the same run do not read them back.
//   * GalaxyPixels accepts healpix_order and ignores it: angular candidates
//     come from a grid on the unit vectors, a superset of the pixel discs,
//     so the counts are the same.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/twopcf_hip.h"

// galaxy.h:16-35 (the bitmask lives in GalaxyCatalogue::masks, n_words per object)
class Galaxy {
public:
	double x[3];
	float weight;
	int jk_region;
	double d_comov;
	double x_norm[3];
	float input_weight;
	void set(double ra, double dec);   // galaxy.cpp:8-27
};

class GalaxyCatalogue {
public:
	int n_jk_regions;
	int n_gals;
	Galaxy *gals;
	int n_words = 0;
	std::vector<uint64_t> masks;   // n_gals * n_words
	std::vector<uint64_t> dither;  // Galaxy::dither_mask, n_gals * n_words (empty: none)
	GalaxyCatalogue(int n_gals, int n_jk_regions);
	~GalaxyCatalogue();
	GalaxyCatalogue(const GalaxyCatalogue &) = delete;
	GalaxyCatalogue &operator=(const GalaxyCatalogue &) = delete;
};

// hist.h:5-19
class Axis {
public:
	int n_bins;
	bool log_binning;
	float min, minus_min, max, step, inv_step, base, inv_log_of_base, inv_log_step;
	Axis(float input_min, float input_max, int input_n_bins, float log_base);
	double bin_centre(const int bin_num);
	double bin_width(const int bin_num);
	double value_from_bin(const double bin_num);
	float bin_num(const float value);
	twopcf_axis desc() const;
};

class Hist {
public:
	Axis ax;
	double *bins;
	Hist(float input_min, float input_max, int input_n_bins, float log_base);
	Hist(const Hist &copy_obj);
	~Hist();
};

class JKHists {
public:
	int n_jk_regions;
	Hist full_hist;
	double *sub_hist;
	JKHists(int jk_regions, const Hist &template_hist);
	~JKHists();
	void invert_jk_hists();
	void multiply_hists(double factor);
	void divide_by_hist(const JKHists &input_hist);
	double get(const int bin_num) { return full_hist.bins[bin_num]; }
	double get(const int jk_region, const int bin_num) { return sub_hist[jk_region * full_hist.ax.n_bins + bin_num]; }
	void normalise(const GalaxyCatalogue *cat, int n_gals);
	void normalise(const GalaxyCatalogue *cat_1, int n_gals_1, const GalaxyCatalogue *cat_2, int n_gals_2);
	double interp(const double value);
	// hist.cpp:180-221: groups full_result (attributes n_jk_regions, n_bins,
	// min, max, base; datasets hist, Bin_centre, Bin_width) and jk_reg<k>
	void output_hdf5(std::string filename);
};

// hist.cpp:115-178: a JKHists from a file JKHists::output_hdf5 wrote; *hist
// is a new object (caller deletes).  Exits with a message if the file cannot
// be opened.
void read_hdf5_hist(std::string filename, JKHists **hist);

class Hist2D {
public:
	Axis ax1, ax2;
	double *bins;
	int total_bins;
	Hist2D(float input_min_1, float input_max_1, int input_n_bins_1, float axis_1_log_base, float input_min_2,
	       float input_max_2, int input_n_bins_2, float axis_2_log_base);
	Hist2D(const Hist2D &copy_obj);
	~Hist2D();
};

class JKHists2D {
public:
	int n_jk_regions;
	Hist2D full_hist;
	double *sub_hist;
	JKHists2D(int jk_regions, const Hist2D &template_hist);
	~JKHists2D();
	void invert_jk_hists();
	void multiply_hists(double factor);
	double get(const int b1, const int b2) { return full_hist.bins[b1 * full_hist.ax2.n_bins + b2]; }
	void normalise(const GalaxyCatalogue *cat, int n_gals);
	void normalise(const GalaxyCatalogue *cat_1, int n_gals_1, const GalaxyCatalogue *cat_2, int n_gals_2);
};

// boxes.h:16-36: here the device-resident catalogue of the first n_gals objects
class GalaxyBoxes {
public:
	double rmax;
	int n_gals;
	twopcf_cat *cat = nullptr;
	const GalaxyCatalogue *source = nullptr;
	GalaxyBoxes(Galaxy input_galaxies[], int input_n_gals, Galaxy other_cat[], int other_cat_n_gals,
	            double input_rmax);
	GalaxyBoxes(GalaxyCatalogue *input, GalaxyCatalogue *other, double input_rmax);
	~GalaxyBoxes();
	GalaxyBoxes(const GalaxyBoxes &) = delete;
	GalaxyBoxes &operator=(const GalaxyBoxes &) = delete;
};

// boxes.h:51-67: the device-resident catalogue of an angular count
class GalaxyPixels {
public:
	int n_gals;
	double theta_max;
	twopcf_cat *cat = nullptr;
	GalaxyPixels(Galaxy input_galaxies[], int input_n_gals, int healpix_order, double input_theta_max,
	             bool require_redshift);
	GalaxyPixels(GalaxyCatalogue *input, int healpix_order, double input_theta_max, bool require_redshift);
	~GalaxyPixels();
	GalaxyPixels(const GalaxyPixels &) = delete;
	GalaxyPixels &operator=(const GalaxyPixels &) = delete;
};

// io.h:10-86
class Params {
public:
	std::string param_filename, data_filename, data_file_type, random_filename, random_file_type, coord_system;
	std::string ra_x_dataset_name, dec_y_dataset_name, z_z_dataset_name, weight_dataset_name, jk_dataset_name;
	bool use_weights = false;
	int n_threads = 0;
	int n_jk_regions = 1;
	double omega_m = 0, h = 0, z_min = 0, z_max = 0;
	bool plot_monopole = false;
	std::string monopole_filename, monopole_output_type;
	double monopole_log_base = 0, monopole_min = 0, monopole_max = 0;
	int monopole_n_bins = 0;
	bool plot_sigma_pi = false;
	std::string sigma_pi_filename, sigma_pi_output_type;
	double sigma_log_base = 0, sigma_min = 0, sigma_max = 0;
	int sigma_n_bins = 0;
	double pi_log_base = 0, pi_min = 0, pi_max = 0;
	int pi_n_bins = 0;
	bool plot_s_mu = false;
	std::string s_mu_filename, s_mu_output_type;
	double s_log_base = 0, s_min = 0, s_max = 0;
	int s_n_bins = 0, mu_n_bins = 0;
	bool calculate_angular_dd = false, calculate_angular_dd_invpweights = false;
	bool calculate_angular_dr = false, calculate_angular_dr_invpweights = false;
	std::string angular_dd_filename, angular_dd_invpweights_filename, angular_dr_filename,
	    angular_dr_invpweights_filename;
	double theta_max = 0;
	int theta_n_bins = 0;
	double theta_log_base = 0;
	int healpix_order = 0;
	std::string bitwise_weight_dataset_name, dither_weight_dataset_name;
	int n_bitwise_runs = 0;
	Params(std::string input_filename);
};

GalaxyCatalogue *read_data_ascii(std::string filename, const Params &params);
GalaxyCatalogue *read_data_hdf5(std::string filename, const Params &params);
GalaxyCatalogue *read_file(std::string filename, std::string file_type, const Params &params);
double ls_estimator(double DD, double DR, double RR);
std::string generate_sub_filename(std::string filename, long long sub_number);
void output_xi_1D(std::string filename, std::string file_type, JKHists *DD, JKHists *DR, JKHists *RR,
                  GalaxyCatalogue *data_cat, GalaxyCatalogue *random_cat);
void output_xi_2D(std::string filename, std::string file_type, std::string axis_1_name, std::string axis_2_name,
                  JKHists2D *DD, JKHists2D *DR, JKHists2D *RR, GalaxyCatalogue *data_cat,
                  GalaxyCatalogue *random_cat);

// correlators.h:7-31 — same names and post-conditions: histograms hold the
// inverted jackknife sums, multiplied by 2 for auto counts, not normalised.
void auto_pair_count_monopole(const GalaxyBoxes *cat, JKHists *hist);
void cross_pair_count_monopole(const GalaxyBoxes *cat_1, const GalaxyBoxes *cat_2, JKHists *hist);
void auto_pair_count_sigma_pi(const GalaxyBoxes *cat, JKHists2D *hist);
void cross_pair_count_sigma_pi(const GalaxyBoxes *cat_1, const GalaxyBoxes *cat_2, JKHists2D *hist);
void auto_pair_count_s_mu(const GalaxyBoxes *cat, JKHists2D *hist);
void cross_pair_count_s_mu(const GalaxyBoxes *cat_1, const GalaxyBoxes *cat_2, JKHists2D *hist);
void auto_correlate_monopole(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params params);
void auto_correlate_sigma_pi(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params params);
void auto_correlate_s_mu(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params params);

// correlators_invpweight.h:9-33 and correlators.h:29-31 — the inverse-p
// pipeline, same names and post-conditions.
void auto_pair_count_angular(const GalaxyPixels *cat, JKHists *hist);
void cross_pair_count_angular(const GalaxyPixels *cat_1, const GalaxyPixels *cat_2, JKHists *hist);
void auto_pair_count_angular_invpweights(const GalaxyPixels *cat, JKHists *hist, Params params);
void cross_pair_count_angular_invpweights(const GalaxyPixels *cat_1, const GalaxyPixels *cat_2, JKHists *hist);
void calc_angular_upweight(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params params);
void auto_pair_count_monopole_invpweights(const GalaxyBoxes *cat, JKHists *hist, Params params);
void cross_pair_count_monopole_invpweights(const GalaxyBoxes *cat_1, const GalaxyBoxes *cat_2, JKHists *hist,
                                           Params params);
void auto_pair_count_sigma_pi_invpweights(const GalaxyBoxes *cat, JKHists2D *hist, Params params);
void cross_pair_count_sigma_pi_invpweights(const GalaxyBoxes *cat_1, const GalaxyBoxes *cat_2, JKHists2D *hist,
                                           Params params);
void auto_pair_count_s_mu_invpweights(const GalaxyBoxes *cat, JKHists2D *hist, Params params);
void cross_pair_count_s_mu_invpweights(const GalaxyBoxes *cat_1, const GalaxyBoxes *cat_2, JKHists2D *hist,
                                       Params params);
void auto_correlate_monopole_invpweights(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params params);
void auto_correlate_sigma_pi_invpweights(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params params);
void auto_correlate_s_mu_invpweights(GalaxyCatalogue *data, GalaxyCatalogue *randoms, Params params);

// The line-of-sight definition of sigma-pi and s-mu (TWOPCF_PI_*).  The
// reference fixes it when compiled (_ANGULAR_PI_DEFN, makefile:4, set in its
// default build); here the drivers take it from TWOPCF_PI (angular by
// default, "midpoint" for builds without that macro).
void twopcf_set_pi_definition(int pi_def);
int twopcf_pi_definition();

// The inverse-p build switch: the readers then take bit masks and apply the
// IIP weights (io.cpp:521-588).
void twopcf_enable_inverse_p(bool on);
bool twopcf_inverse_p_enabled();

// io.cpp:521-588: input_weight = weight; weight *= n_runs / popcount(masks)
// when the popcount differs from n_runs (0 counts as 1).
void apply_bitwise_weights(GalaxyCatalogue *cat, const Params &params);
// _USE_DITHER_WEIGHTS (makefile:6) as a run-time switch of the inverse-p
// pipeline: catalogues carry dither masks (ascii: n_words more integer
// columns after the bit masks), the read-time weight uses dither_total /
// total and the inverse-p pair weights the dither popcount as numerator.
void twopcf_enable_dither_weights(bool on);
bool twopcf_dither_weights_enabled();

// The angular histograms of calc_angular_upweight, by file name, within one
// process.  read_angular_hist returns a new object (caller deletes): the one
// stored under the name in this process, else the HDF5 file of that name
// (read_hdf5_hist), as the reference's *_invpweights correlators read it.
void store_angular_hist(const std::string &filename, const JKHists &hist);
JKHists *read_angular_hist(const std::string &filename);

// The context every entry point uses (device TWOPCF_DEVICE, default 0;
// counting method TWOPCF_METHOD = fp32 (default) or fp64).  Aborts with a
// message when no GPU is usable: there is no CPU fallback.
twopcf_ctx *twopcf_default_ctx();
int twopcf_default_method();
