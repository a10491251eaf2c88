// pairs.hpp — device-side pair arithmetic shared by the counting kernels.
//
// exact_*: the reference's per-pair pipeline evaluated with the same IEEE
// operations in the same order (correlators.cpp:31-41, 156-188, 309-343;
// hist.cpp:46-53, 484-502).  The library is compiled with
// -ffp-contract=off, so no multiply-add is fused here; sqrt and division are
// the correctly rounded IEEE operations (__builtin_sqrtf, '/').  The log
// index of log axes is never evaluated on the device: bins come from
// threshold tables computed on the host with glibc log1pf (host_math.cpp).
#pragma once
#include <hip/hip_runtime.h>

namespace twopcf {

enum { K_MONO = 0, K_SIGMA_PI = 1, K_S_MU = 2, K_ANGULAR = 3 };

// Largest k in [0, n-1] with t[k] <= v, given t[0] <= v and t ascending.
__device__ __forceinline__ int thr_bin(const float *t, int n, float v) {
	int k = 0;
	for (int step = 1 << 11; step; step >>= 1)
		if (step < n && k + step < n && t[k + step] <= v) k += step;
	return k;
}

// Linear-axis index as the reference computes it: (v + minus_min) * inv_step
// in float, truncated; accepted iff it truncates into [0, n).
__device__ __forceinline__ bool lin_bin(float v, float minus_min, float inv_step, int n, int &b) {
	float x = (v + minus_min) * inv_step;
	if (!(x > -1.0f && x < (float)n)) return false;   // NaN fails both
	b = (int)x;
	return true;
}

// Value-domain threshold bin: accepted iff t[0] <= v < t[n].
__device__ __forceinline__ bool tab_bin(const float *t, int n, float v, int &b) {
	if (!(v >= t[0] && v < t[n])) return false;
	b = thr_bin(t, n, v);
	return true;
}

struct AxisDev {
	int n;
	int log;            // bins from the value-threshold table when set
	float minus_min, inv_step;
};

// Float difference of two double coordinates (correlators.cpp:311-314).
__device__ __forceinline__ float ref_diff(double a, double b) { return (float)(a - b); }

// s^2 summed in float, left to right (correlators.cpp:315).
__device__ __forceinline__ float ref_s2(float d0, float d1, float d2) {
	float a = d0 * d0;
	float b = d1 * d1;
	float c = d2 * d2;
	return (a + b) + c;
}

// Line-of-sight separation, angular definition (correlators.cpp:319-323).
__device__ __forceinline__ float ref_pi(float d0, float d1, float d2, const double *n1, const double *n2) {
	float p = (float)((double)d0 * (n1[0] + n2[0]));
	p = (float)((double)p + (double)d1 * (n1[1] + n2[1]));
	p = (float)((double)p + (double)d2 * (n1[2] + n2[2]));
	p = (float)((double)p * 0.5);
	return p;
}

// Line-of-sight separation, pair-midpoint definition (builds without
// _ANGULAR_PI_DEFN, correlators.cpp:325-336): sum = float(x1 + x2), its norm
// squared summed in float, 1.0 / sqrtf(.) divided in double and rounded to
// float, then the three float products added left to right.
__device__ __forceinline__ float ref_pi_mid(float d0, float d1, float d2, const double *x1, const double *x2) {
	const float s0 = (float)(x1[0] + x2[0]), s1 = (float)(x1[1] + x2[1]), s2 = (float)(x1[2] + x2[2]);
	float q = 0.0f;
	q += s0 * s0;
	q += s1 * s1;
	q += s2 * s2;
	const float inv = (float)(1.0 / (double)__builtin_sqrtf(q));
	float p = d0 * s0 * inv;
	p += d1 * s1 * inv;
	p += d2 * s2 * inv;
	return p;
}

// 1 - cos(theta) as the angular correlators form it: double, running
// subtraction of the three products (correlators_invpweight.cpp:738-741).
__device__ __forceinline__ double ref_one_minus_cos(const double *n1, const double *n2) {
	double v = 1.0;
	v -= n1[0] * n2[0];
	v -= n1[1] * n2[1];
	v -= n1[2] * n2[2];
	return v;
}

// One pair through the reference pipeline.  t_s2: s^2 thresholds (monopole,
// s-mu); t1/t2: value thresholds of the sigma / pi axes (sigma-pi, log axes);
// a1/a2: linear-axis constants.  Returns true and the flattened bin if the
// pair lands in the histogram.
template <int KIND>
__device__ __forceinline__ bool exact_bin(const double *x1, const double *n1, const double *x2, const double *n2,
                                          float min_r_sq, float max_r_sq, const float *t_s2, const float *t1,
                                          const float *t2, const AxisDev &a1, const AxisDev &a2, int &bin,
                                          int pi_mid = 0) {
	if constexpr (KIND == K_ANGULAR) {
		// value rounded to float for Axis::bin_num (hist.cpp:246-250); no range test
		return tab_bin(t1, a1.n, (float)ref_one_minus_cos(n1, n2), bin);
	}
	float d0 = ref_diff(x1[0], x2[0]);
	float d1 = ref_diff(x1[1], x2[1]);
	float d2 = ref_diff(x1[2], x2[2]);
	float s2 = ref_s2(d0, d1, d2);
	if (KIND == K_MONO || KIND == K_S_MU) {
		if (!(s2 >= t_s2[0] && s2 < t_s2[a1.n])) return false;
		int b1 = thr_bin(t_s2, a1.n, s2);
		if (KIND == K_MONO) {
			bin = b1;
			return true;
		}
		float p = pi_mid ? ref_pi_mid(d0, d1, d2, x1, x2) : ref_pi(d0, d1, d2, n1, n2);
		float r = __builtin_sqrtf(s2);
		float mu = fabsf(p) / r;
		int b2;
		if (!lin_bin(mu, a2.minus_min, a2.inv_step, a2.n, b2)) return false;
		bin = b1 * a2.n + b2;
		return true;
	} else {
		if (!(s2 < max_r_sq && s2 > min_r_sq)) return false;
		float p = pi_mid ? ref_pi_mid(d0, d1, d2, x1, x2) : ref_pi(d0, d1, d2, n1, n2);
		float sig = __builtin_sqrtf(s2 - p * p);
		int b1, b2;
		if (a1.log ? !tab_bin(t1, a1.n, sig, b1) : !lin_bin(sig, a1.minus_min, a1.inv_step, a1.n, b1)) return false;
		float ap = fabsf(p);
		if (a2.log ? !tab_bin(t2, a2.n, ap, b2) : !lin_bin(ap, a2.minus_min, a2.inv_step, a2.n, b2)) return false;
		bin = b1 * a2.n + b2;
		return true;
	}
}

}  // namespace twopcf
