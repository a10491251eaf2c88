// host_math.cpp — see host_math.hpp.  Compiled with -ffp-contract=off and
// plain x86-64 code generation so every float operation rounds as in the
// reference's g++ build.
#include "host_math.hpp"

#include <cmath>
#include <cstring>
#include <limits>

namespace twopcf {

AxisF::AxisF(float mn, float mx, int n_bins, float log_base) {
	min = mn;
	minus_min = -mn;
	max = mx;
	n = n_bins;
	step = (mx - mn) / (float)n_bins;
	inv_step = (float)(1.0 / (double)step);
	base = log_base;
	if ((double)log_base > 1.02) {
		// log(float) binds to the float overload; pow(float, int) promotes.
		ilb = (float)(1.0 / (double)logf(log_base));
		ils = (float)((std::pow((double)log_base, (double)n_bins) - 1.0) / (double)(mx - mn));
		log = true;
	} else {
		ilb = 0.0f;
		ils = 0.0f;
		log = false;
	}
}

float AxisF::index(float v) const {
	if (log) return ilb * log1pf(ils * (v - min));
	return (v + minus_min) * inv_step;
}

double AxisF::value_from_bin(double b) const {
	if (log) return (std::exp(b / (double)ilb) - 1.0) / (double)ils + (double)min;
	return (double)min + b * (double)step;
}

double AxisF::centre(int b) const { return (value_from_bin(b + 1.0) + value_from_bin((double)b)) / 2.0; }
double AxisF::width(int b) const { return value_from_bin(b + 1.0) - value_from_bin((double)b); }

void range_sq(int kind, const AxisF &a1, const AxisF &a2, float *mn, float *mx) {
	if (kind == SIGMA_PI) {
		*mx = a1.max * a1.max + a2.max * a2.max;
		*mn = a1.min * a1.min;
		if (a1.min > a2.min) *mn = a2.min * a2.min;
	} else {
		*mx = a1.max * a1.max;
		*mn = a1.min * a1.min;
	}
}

// ---------------------------------------------------------------------------
// Threshold search over float bit patterns.  Keys order all non-NaN floats.

static inline uint32_t fkey(float f) {
	uint32_t b;
	std::memcpy(&b, &f, 4);
	return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
static inline float keyf(uint32_t k) {
	uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
	float f;
	std::memcpy(&f, &b, 4);
	return f;
}

// Smallest key in [lo, hi] where pred holds (pred monotone false -> true);
// hi + 1 if none.
template <class P>
static uint32_t first_true(uint32_t lo, uint32_t hi, P pred) {
	if (!pred(keyf(hi))) return hi + 1;
	while (lo < hi) {
		uint32_t mid = lo + (hi - lo) / 2;
		if (pred(keyf(mid))) hi = mid;
		else lo = mid + 1;
	}
	return lo;
}

// Truncation semantics of the reference (float -> int, NaN dropped):
// trunc(x) >= 0  <=>  x > -1,   trunc(x) >= k (k >= 1)  <=>  x >= k.
static inline bool idx_at_least(float x, int k) { return k == 0 ? (x > -1.0f) : (x >= (float)k); }

// Verify the predicate switches exactly once within `win` keys around t.
template <class P>
static bool check_switch(uint32_t t, uint32_t lo, uint32_t hi, P pred, uint32_t win = 4096) {
	uint32_t a = (t > lo + win) ? t - win : lo;
	uint32_t b = (t + win < hi) ? t + win : hi;
	for (uint32_t k = a; k <= b && k != 0xffffffffu; ++k)
		if (pred(keyf(k)) != (k >= t)) return false;
	return true;
}

bool value_thresholds(const AxisF &a, std::vector<float> &t) {
	const uint32_t lo = fkey(-std::numeric_limits<float>::max());
	const uint32_t hi = fkey(std::numeric_limits<float>::max());
	t.assign(a.n + 1, 0.0f);
	for (int k = 0; k <= a.n; ++k) {
		auto pred = [&](float v) { return idx_at_least(a.index(v), k); };
		uint32_t key = first_true(lo, hi, pred);
		if (key > hi) {
			t[k] = std::numeric_limits<float>::infinity();
			continue;
		}
		if (!check_switch(key, lo, hi, pred)) return false;
		t[k] = keyf(key);
	}
	for (int k = 1; k <= a.n; ++k)
		if (t[k] < t[k - 1]) return false;
	return true;
}

bool interp_thresholds(const AxisF &a, std::vector<float> &t) {
	const uint32_t lo = fkey(-std::numeric_limits<float>::max());
	const uint32_t hi = fkey(std::numeric_limits<float>::max());
	t.assign(a.n + 1, 0.0f);
	for (int k = 0; k <= a.n; ++k) {
		auto pred = [&](float v) {
			const float bn = (float)((double)a.index(v) - 0.5);
			return k < a.n ? (bn >= (float)k) : (bn > (float)(a.n - 1));
		};
		uint32_t key = first_true(lo, hi, pred);
		if (key > hi) {
			t[k] = std::numeric_limits<float>::infinity();
			continue;
		}
		if (!check_switch(key, lo, hi, pred)) return false;
		t[k] = keyf(key);
	}
	for (int k = 1; k <= a.n; ++k)
		if (t[k] < t[k - 1]) return false;
	return true;
}

bool sqrt_thresholds(const std::vector<float> &t, std::vector<float> &X) {
	const uint32_t lo = fkey(0.0f);
	const uint32_t hi = fkey(std::numeric_limits<float>::max());
	X.assign(t.size(), 0.0f);
	for (size_t k = 0; k < t.size(); ++k) {
		const float tk = t[k];
		if (!(tk > 0.0f)) {
			X[k] = -0.0f;   // sqrtf(x) >= tk for every x >= -0, NaN below
			continue;
		}
		auto pred = [&](float x) { return sqrtf(x) >= tk; };
		uint32_t key = first_true(lo, hi, pred);
		if (key > hi) {
			X[k] = std::numeric_limits<float>::infinity();
			continue;
		}
		if (!check_switch(key, lo, hi, pred)) return false;
		X[k] = keyf(key);
	}
	for (size_t k = 1; k < X.size(); ++k)
		if (X[k] < X[k - 1]) return false;
	return true;
}

bool s2_thresholds(int kind, const AxisF &a1, const AxisF &a2, std::vector<float> &t) {
	float mn, mx;
	range_sq(kind, a1, a2, &mn, &mx);
	const uint32_t lo = fkey(0.0f);
	const uint32_t hi = fkey(std::numeric_limits<float>::max());
	const int n = a1.n;
	t.assign(n + 1, 0.0f);
	auto at = [&](int k) {
		return [&, k](float s2) { return idx_at_least(a1.index(sqrtf(s2)), k); };
	};
	// t[0]: first s2 above min_r_sq that also has index > -1.
	{
		auto p0 = at(0);
		auto pred = [&](float s2) { return s2 > mn && p0(s2); };
		uint32_t key = first_true(lo, hi, pred);
		if (key > hi) return false;
		if (!check_switch(key, lo, hi, pred)) return false;
		t[0] = keyf(key);
	}
	for (int k = 1; k <= n; ++k) {
		auto pred = at(k);
		uint32_t key = first_true(lo, hi, pred);
		float v = std::numeric_limits<float>::infinity();
		if (key <= hi) {
			if (!check_switch(key, lo, hi, pred)) return false;
			v = keyf(key);
		}
		t[k] = v;
	}
	// Upper range limit (s2 < max_r_sq) closes the last bin.
	if (mx < t[n]) t[n] = mx;
	for (int k = 1; k < n; ++k) {
		if (t[k] < t[0]) t[k] = t[0];
		if (t[k] > t[n]) t[k] = t[n];
	}
	for (int k = 1; k <= n; ++k)
		if (t[k] < t[k - 1]) return false;
	return true;
}

// ---------------------------------------------------------------------------
// Cosmology, cosmology.cpp:7-100 and dataset.cpp:254-300.

static double codist_integrand(double om, double ol, double opz) {
	return 1.0 / (std::sqrt(om * opz * opz * opz + ol));
}

Cosmology::Cosmology(double om, double h, double zmin, double z_max, double z_save_step, int overs)
	: z_min(zmin), step(z_save_step) {
	double ol = 1.0 - om;
	double c = 299792.458;
	double hubble = c / (h * 100.0);
	double z = 0.0, cur = 0.0, prev = 0.0;
	double fine = z_save_step / (double)overs;
	int nb = (int)(((z_max - zmin) / z_save_step) + 1);
	y.assign(nb, 0.0);
	// Pre-loop up to z_min; its second term passes (omega_l, omega_m)
	// (cosmology.cpp:42), which the correction below cancels when z_min == 0.
	while (z <= zmin) {
		prev = cur;
		cur += codist_integrand(om, ol, 1.0 + z) + codist_integrand(ol, om, 1.0 + z + fine);
		z += fine;
	}
	cur -= ((z - zmin) / fine) * (cur - prev);
	z = zmin;
	for (int i = 0; i < nb; ++i) {
		y[i] = 0.5 * fine * hubble * cur;
		for (int j = 0; j < overs; ++j) {
			cur += codist_integrand(om, ol, 1.0 + z) + codist_integrand(om, ol, 1 + z + fine);
			z += fine;
		}
	}
}

// Outside the table the reference returns the x end-point (dataset.cpp:287-292).
double Cosmology::comov_dist(double zz) const {
	double x0 = z_min, xn = x0 + (double)((int)y.size() - 1) * step;
	if (zz <= x0) return x0;
	if (zz >= xn) return xn;
	int k = (int)((zz - x0) / step);
	return y[k] + ((zz - (x0 + k * step)) / step) * (y[k + 1] - y[k]);
}

// The reference's g++ -O3 build evaluates the sine and cosine of each angle
// with one glibc sincos call (its sine differs from sin() in the last bit for
// some arguments); sincos is called here so positions agree bit for bit.
void eq_to_cart(int64_t n, const double *ra, const double *dec, const double *d,
                double *x, double *y, double *z, double *nx, double *ny, double *nz) {
	const double d2r = M_PI / 180.0;
	for (int64_t i = 0; i < n; ++i) {
		double sr, cr, sd, cd;
		sincos(d2r * ra[i], &sr, &cr);
		sincos(d2r * dec[i], &sd, &cd);
		nx[i] = cd * cr;
		ny[i] = cd * sr;
		nz[i] = sd;
		if (d[i] >= 0.0) {
			x[i] = d[i] * nx[i];
			y[i] = d[i] * ny[i];
			z[i] = d[i] * nz[i];
		} else {
			x[i] = y[i] = z[i] = 0.0;
		}
	}
}

}  // namespace twopcf
