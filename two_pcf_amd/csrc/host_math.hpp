// host_math.hpp — host-side pieces of the counting path that must reproduce
// the reference's float/double arithmetic exactly: Axis constants, the
// threshold tables the GPU bins by, the comoving-distance table and
// Galaxy::set.
#pragma once
#include <cstdint>
#include <vector>

namespace twopcf {

// Axis, hist.h:5-19 / hist.cpp:15-36.  Constants are float.
struct AxisF {
	float min, minus_min, max, step, inv_step, base, ilb, ils;
	int n;
	bool log;
	AxisF() = default;
	AxisF(float mn, float mx, int n_bins, float log_base);
	// Axis::bin_num (hist.cpp:46-53), glibc log1pf for log axes.
	float index(float v) const;
	double value_from_bin(double b) const;   // hist.cpp:55-62
	double centre(int b) const;              // hist.cpp:38-40
	double width(int b) const;               // hist.cpp:42-44
};

enum Kind { MONOPOLE = 0, SIGMA_PI = 1, S_MU = 2, ANGULAR = 3 };

// Range limits of the correlators (correlators.cpp:13-14, 192-195, 347-348).
void range_sq(int kind, const AxisF &a1, const AxisF &a2, float *min_r_sq, float *max_r_sq);

// Value-domain thresholds t[0..n]: the reference bins v into k iff
// t[k] <= v < t[k+1] and drops v outside [t[0], t[n]) (NaN included).
// Returns false if the index is not monotone near a threshold.
bool value_thresholds(const AxisF &a, std::vector<float> &t);

// Segment thresholds of JKHists::interp (hist.cpp:378-389): with
// bn = float(double(index(float v)) - 0.5), t[k] (k < n) is the first float
// with bn >= k and t[n] the first with bn > n - 1.
bool interp_thresholds(const AxisF &a, std::vector<float> &t);

// sqrtf-domain thresholds: X[k] = the smallest float x with sqrtf(x) >= t[k]
// (x below every X[k] includes the negative floats, whose sqrtf is NaN);
// -0 for t[k] <= 0, +inf for t[k] = +inf.  Returns false if sqrtf is not
// monotone near a threshold.
bool sqrt_thresholds(const std::vector<float> &t, std::vector<float> &X);

// s^2-domain thresholds of a monopole / s-mu s axis, folding in the range
// test and sqrtf (correlators.cpp:315-317, 338).  Same convention.
bool s2_thresholds(int kind, const AxisF &a1, const AxisF &a2, std::vector<float> &t);

// Comoving distance: the trapezoid table of cosmology.cpp:11-65 (including
// its z_min pre-loop) and DatasetFixedXY::lin_interp (dataset.cpp:285-300).
struct Cosmology {
	double z_min, step;
	std::vector<double> y;
	Cosmology(double omega_m, double h, double z_min, double z_max, double z_save_step = 1e-4, int z_oversample = 100);
	double comov_dist(double z) const;
};

// Galaxy::set (galaxy.cpp:8-27).
void eq_to_cart(int64_t n, const double *ra, const double *dec, const double *d,
                double *x, double *y, double *z, double *nx, double *ny, double *nz);

}  // namespace twopcf
