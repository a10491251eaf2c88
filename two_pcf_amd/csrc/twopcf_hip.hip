// twopcf_hip.hip — libtwopcf_hip: gfx950 pair counting behind the two_pcf
// correlator entry points.  C ABI in include/twopcf_hip.h.
//
// Layout on the device (per catalogue and cell geometry):
//   sorted SoA   double x,y,z,nx,ny,nz; float w, iw; int jk   (cell-major,
//                sorted along z within a cell: a counting sort by cell, then
//                by 2^ZB_BITS z buckets of the cell and catalogue index, so
//                that a row of cells along z is sorted along z as a whole up
//                to ties within one bucket)
//   cell_start   int[n_cell + 1]
// Work: a work item is (job, up to 256 first-side objects of one cell, a set
// of neighbour-cell rows, up to S region slots).  A 1024-thread block (16
// waves: 4 groups over the first-side objects x 4 groups over each tile) owns
// one item at a time: each lane holds one first-side object, the second-side
// objects of each row are staged through LDS in tiles of 1024 and every tile
// is shared by all 16 waves.  Per-item histograms live in LDS (u32 counts or
// f64 sums) and are flushed once per item with global atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/twopcf_hip.h"
#include "host_math.hpp"
#include "pairs.hpp"

using namespace twopcf;

namespace {

constexpr int WAVE = 64;
constexpr int IG = 4;                       // wave groups over first-side objects
constexpr int JG = 4;                       // wave groups over each staged tile
constexpr int BLOCK = WAVE * IG * JG;       // 1024 threads, 16 waves
constexpr int ICHUNK = WAVE * IG;           // first-side objects per work item
constexpr int TILE_FAST = 1024;             // staged second-side objects (fp32 path)
constexpr int TILE_EXACT = 512;             // staged second-side objects (fp64 path)
constexpr int RING = 512;                   // per-wave queue of in-range pairs, u16 entries (fp32 path)
// pairs per lane in one binning pass of the fp32 path (three per lane
// measured slower than two on the s-mu jackknife batch).  Sigma-pi bins one:
// its decision (two axes, per-pair margins) holds more registers, and with two
// pairs per lane the weighted kernel spilled (equatorial 155.3 -> 151.4 ms
// with one).
__host__ __device__ constexpr int np_of(int kind, bool) { return kind == K_SIGMA_PI ? 1 : 2; }
constexpr int NP_MAX = 2;
static_assert(RING >= NP_MAX * 64 + 4 * 64, "ring: pending pairs plus one screening step");
static_assert(TILE_FAST <= BLOCK && TILE_FAST <= 1024, "tile: one entry per thread, 10-bit ring index");
static_assert(ICHUNK <= 256, "first-side index in one byte of the recheck queue");
// the ring holds < NP * 64 pending pairs plus one screening step's 256; the
// recheck queue < 64 plus one binning pass's NP * 64
constexpr int RQ = WAVE * (NP_MAX + 1);     // per-wave queue of pairs to re-evaluate exactly (fp32 path)
constexpr int REGION_MAP = 512;             // region -> slot code table (fp32 path, n_jk up to this)
constexpr int TILE_BLOCKS = TILE_FAST / 16; // 16-entry screening blocks per tile (one lane each)
static_assert(TILE_BLOCKS <= 64, "one lane per tile block");
// s_waitcnt immediate of gfx9-family targets: vmcnt 0 (bits 3:0 and 15:14),
// expcnt 7 and lgkmcnt 15 (no wait on those)
constexpr unsigned VMCNT0 = 0x0F70u;
constexpr int ROW_SPANS = 64;              // stencil rows of a work item with their second-side spans (fp32 path)
// tile + (fp32 path) first-side chunk with its two screening thresholds,
// per-wave pair queues, the region map, the z bounds of the tile blocks and
// the item's row spans
__host__ __device__ constexpr int tile_bytes(bool fast) {
	return fast ? TILE_FAST * (2 * 16 + 4) + ICHUNK * (32 + 8) + (BLOCK / WAVE) * (RING * 2 + RQ * 5) + REGION_MAP * 4 +
	                  TILE_BLOCKS * 8 + (ROW_SPANS + 1) * 8
	            : TILE_EXACT * (6 * 8 + 4 + 4 + 4);
}
constexpr int MAX_SLOTS = 16;
// item queues of a persistent launch: one per XCD (the dispatcher places
// workgroup b on XCD b mod 8), each over a contiguous range of the items
constexpr int NQ = 8;
constexpr int MAX_CLASSES = 64;             // launch classes of one call
// histogram layouts of the fp32 kernel's jobs (its RG template parameter)
constexpr int RG_ANY = 0, RG_NONE = 1, RG_JK = 2;
// a second job binned in the pass of another (fp32 path, its FZ template
// parameter): none; a monopole in an s-mu job's pass; a twin, an angular job
// of the same catalogues and axis with another pair-weight rule
constexpr int FZ_NONE = 0, FZ_MONO = 1, FZ_TWIN = 2;
constexpr int MAX_ROWS = 64;                // stencil rows per work item (a u64 row mask)
static_assert(MAX_ROWS <= ROW_SPANS, "one LDS span per stencil row");
constexpr int MAX_SUBDIV = 3;               // cells per rmax; 4 needs more than MAX_ROWS rows
constexpr int MAX_AXIS_BINS = 4096;

// ------------------------------------------------------------ device data --

struct CatDev {
	const double *x, *y, *z, *nx, *ny, *nz;
	const float *w, *iw;
	const int *jk;
	const uint64_t *masks;
	int m_stride;       // mask words per object: the bit mask, then the dither mask if any
	const int *start;   // cell starts [n_cell + 1]
};

struct JobDev {
	CatDev A, B;        // A: first side (lanes), B: second side (staged)
	int dim[3];
	double lo[3];
	double h;
	int kind, is_auto, n1, n2, nb, n_jk, weighted;
	float min_r_sq, max_r_sq;
	AxisDev ax1, ax2;
	int t_s2_off, t1_off, t2_off, tA_off, tB_off;   // offsets into the table pool (floats)
	int tq_off;                                      // (A_k, B_k+1) per bin: bin k sure for A_k <= s^2 < B_k+1
	float s_min, s_ilb_ln2, s_ils, s_inv_step;       // fast s-index estimate
	int s_log;
	float mu_h0, mu_c2;                              // fast mu test: 1/2 - delta = h0 - c2 * rsq(s2)
	int f_shift;                                     // fp32 path: full histogram in 2^f_shift lane copies
	int full_global;                                 // full histogram too large for LDS: added to the pool
	int x1_off;                                      // sigma-pi fp32: sqrtf-domain thresholds of the sigma axis
	float p_min, p_ilb_ln2, p_ils, p_inv_step;       // sigma-pi fp32: pi-axis index estimate
	int p_log;
	float e_c1, e_c2;                                // sigma-pi fp32: E(S) = e_c1 S + e_c2 sqrt(S)
	float pi_c1, pi_c2;                              // sigma-pi fp32: |pi_fp32 - pi_ref| <= pi_c1 s + pi_c2
	int pi_def;                                      // TWOPCF_PI_* (line of sight of sigma-pi and s-mu)
	float mid_c;                                     // pair-midpoint pi: extra |dpi| / s per 1 / |x1 + x2|
	int row_off, n_rows;                             // stencil rows in the row pool
	int wmode, n_runs, n_words;                      // pair-weight rule (TWOPCF_W_*)
	int dither;                                      // inverse-p numerator from the dither masks
	int mstage;                                      // fp32 path: first-side mask rows in LDS, words per row (0: none)
	float scr_eps;                                   // fp32 screening: bound on |s^2_mfma - s^2_fp32|
	int up_n;                                        // angular up-weight table
	double up_min, up_max;
	const double *up_bins, *up_centre;
	const float *up_thr;                             // [up_n + 1] segment thresholds
	unsigned long long *full_u, *jk_u;
	double *full_d, *jk_d;
	unsigned long long *counters;                    // [0] accepted, [1] rechecked
	// fp32 path, s-mu jobs that also bin a monopole job of the same
	// catalogues, weights and s minimum on the same pairs (no regions): its
	// axis, tables, lane copies and outputs; its estimate terms are stored
	// with the unused ones zero (inv_step on a log axis, ils and ilb on a
	// linear one)
	int m_nb, m_f_shift, m_tq_off, m_t_s2_off, m_tA_off, m_tB_off;
	// fp32 path, an angular job with a twin (same catalogues, axis and pair
	// set, no regions; the twin's outputs in the m_* fields): the twin's
	// weight is the product of the other staged weights (w when this job's
	// rule stages iw, and the reverse; staged in the region-code places),
	// or with t_w = 0 the same product of staged weights as this job's
	int t_w;
	float m_s_ilb_ln2, m_s_ils, m_s_inv_step;
	AxisDev m_ax1;
	unsigned long long *m_full_u;
	double *m_full_d;
	unsigned long long *m_counters;
};

struct WorkItem {
	int job;
	int i_begin, i_count, i_cell;
	unsigned long long row_mask;
	int slot_base, n_slots;
	int all_slotted;   // fp32 path: every region the item meets has an LDS slot
	// fp32 path: the region of every first-side object of the item (its cell
	// lies in one region), else -1
	int uf_region;
};

}  // namespace

// Device memory of twopcf_count_noalloc: the per-call buffers (job table,
// work items, stencil rows, region slots, float tables, up-weight tables,
// output pools, counters, queue counters) are carved from this array instead
// of hipMalloc.  Each GPU the library runs on holds it until the process
// ends.  Host code never passes the array's name to a kernel or a copy: it
// uses the address hipGetSymbolAddress gives on each device
// (noalloc_scratch_ptr), resolved once per device in twopcf_ctx_create.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES];   // per device; null until resolved
std::mutex g_scratch_mu;

// Resolve the scratch array's address on the current device (once per
// device).  Returns 0, or a HIP error from hipGetSymbolAddress.
hipError_t resolve_scratch(int device) {
	if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipSuccess;   // twopcf_count_noalloc then refuses
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (g_scratch_ptr[device]) return hipSuccess;
	void *p = nullptr;
	hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	if (e == hipSuccess) g_scratch_ptr[device] = p;
	return e;
}

void *noalloc_scratch_ptr(int device) {
	if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	return g_scratch_ptr[device];
}

// Bump allocation of a plan's device buffers from the scratch array.
struct Arena {
	char *base;
	size_t cap, used;
};

// -------------------------------------------------------------- kernels --

// Gridding: a counting sort by cell, then by z within each cell.  Objects
// of a cell are ordered by ZB_BITS z buckets of the cell's height and, inside
// a bucket, by their index in the catalogue; a row of cells along z is then
// sorted along z up to ties within one bucket (h 2^-ZB_BITS).
constexpr int ZB_BITS = 10;
constexpr int ZB = 1 << ZB_BITS;
constexpr int SORT_BLOCK = 1024;            // threads of the scan and per-cell kernels (one per z bucket)
static_assert(SORT_BLOCK == ZB, "one thread per z bucket");
static_assert(ZB_BITS >= 9, "the counting kernel's z window slack, h 2^-8, covers two z buckets");

// cell and z bucket of every object; count[cell] += 1
__global__ void k_cell_bucket(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, int64_t n, double lo0, double lo1, double lo2, double h,
                              int d0, int d1, int d2, int *__restrict__ cell, unsigned short *__restrict__ zb,
                              unsigned *__restrict__ count) {
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const double fz = (z[i] - lo2) / h;
	int c0 = (int)floor((x[i] - lo0) / h), c1 = (int)floor((y[i] - lo1) / h), c2 = (int)floor(fz);
	c0 = min(max(c0, 0), d0 - 1);
	c1 = min(max(c1, 0), d1 - 1);
	c2 = min(max(c2, 0), d2 - 1);
	const double q = floor((fz - c2) * (double)ZB);
	const int c = (c0 * d1 + c1) * d2 + c2;
	cell[i] = c;
	zb[i] = (unsigned short)min(max(q, 0.0), (double)(ZB - 1));
	atomicAdd(&count[c], 1u);
}

// Exclusive scan of v over a block of SORT_BLOCK threads (16 waves); *total
// gets the block's sum.  wsum: 16 LDS words.
__device__ unsigned block_exclusive_scan(unsigned v, unsigned *wsum, unsigned *total) {
	constexpr int NW = SORT_BLOCK / WAVE;
	const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
	unsigned x = v;
	for (int d = 1; d < WAVE; d <<= 1) {
		const unsigned y = __shfl_up(x, d);
		if (lane >= d) x += y;
	}
	if (lane == WAVE - 1) wsum[w] = x;
	__syncthreads();
	if (w == 0) {
		unsigned t = lane < NW ? wsum[lane] : 0u;
		for (int d = 1; d < NW; d <<= 1) {
			const unsigned y = __shfl_up(t, d);
			if (lane >= d) t += y;
		}
		if (lane < NW) wsum[lane] = t;
	}
	__syncthreads();
	const unsigned before = w ? wsum[w - 1] : 0u;
	*total = wsum[NW - 1];
	__syncthreads();
	return before + x - v;
}

// scan in tiles of SORT_BLOCK: out = exclusive prefix within the tile,
// sums[tile] = the tile's total
__global__ void __launch_bounds__(SORT_BLOCK) k_scan_tiles(const unsigned *__restrict__ in, unsigned *__restrict__ out,
                                                          int64_t n, unsigned *__restrict__ sums) {
	__shared__ unsigned wsum[SORT_BLOCK / WAVE];
	const int64_t i = (int64_t)blockIdx.x * SORT_BLOCK + threadIdx.x;
	unsigned total;
	const unsigned ex = block_exclusive_scan(i < n ? in[i] : 0u, wsum, &total);
	if (i < n) out[i] = ex;
	if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan of the tile sums in place (one block, carried across chunks)
__global__ void __launch_bounds__(SORT_BLOCK) k_scan_sums(unsigned *__restrict__ sums, int n_tiles) {
	__shared__ unsigned wsum[SORT_BLOCK / WAVE];
	unsigned carry = 0;
	for (int b = 0; b < n_tiles; b += SORT_BLOCK) {
		const int k = b + (int)threadIdx.x;
		unsigned total;
		const unsigned ex = block_exclusive_scan(k < n_tiles ? sums[k] : 0u, wsum, &total);
		if (k < n_tiles) sums[k] = carry + ex;
		carry += total;
	}
}

__global__ void __launch_bounds__(SORT_BLOCK) k_scan_add(unsigned *__restrict__ out, int64_t n,
                                                        const unsigned *__restrict__ sums) {
	const int64_t i = (int64_t)blockIdx.x * SORT_BLOCK + threadIdx.x;
	if (i < n) out[i] += sums[blockIdx.x];
}

// objects to their cell's segment (in any order inside it)
__global__ void k_scatter_cells(const int *__restrict__ cell, const unsigned short *__restrict__ zb, int64_t n,
                                const unsigned *__restrict__ start, unsigned *__restrict__ cursor,
                                int *__restrict__ tperm, unsigned short *__restrict__ tzb) {
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int c = cell[i];
	const unsigned pos = start[c] + atomicAdd(&cursor[c], 1u);
	tperm[pos] = (int)i;
	tzb[pos] = zb[i];
}

// Each cell's segment by z bucket (a counting sort in LDS, one thread per
// bucket), then each bucket by catalogue index (an insertion sort by the
// bucket's thread), so the order is deterministic.  Blocks stride over the
// cells; every loop bound is uniform over the block.
__global__ void __launch_bounds__(SORT_BLOCK) k_zsort_cells(const unsigned *__restrict__ start, int n_cell,
                                                           const int *__restrict__ tperm,
                                                           const unsigned short *__restrict__ tzb,
                                                           int *__restrict__ perm) {
	__shared__ unsigned hist[ZB];
	__shared__ unsigned wsum[SORT_BLOCK / WAVE];
	const int tid = threadIdx.x;
	for (int c = blockIdx.x; c < n_cell; c += gridDim.x) {
		const unsigned b = start[c], e = start[c + 1];
		if (e - b <= 1u) {
			if (tid == 0 && e > b) perm[b] = tperm[b];
			continue;
		}
		hist[tid] = 0u;
		__syncthreads();
		for (unsigned k = b + tid; k < e; k += SORT_BLOCK) atomicAdd(&hist[tzb[k]], 1u);
		__syncthreads();
		const unsigned cnt = hist[tid];
		unsigned total;
		const unsigned off = block_exclusive_scan(cnt, wsum, &total);
		hist[tid] = off;
		__syncthreads();
		for (unsigned k = b + tid; k < e; k += SORT_BLOCK) perm[b + atomicAdd(&hist[tzb[k]], 1u)] = tperm[k];
		__syncthreads();   // the block's global writes are visible to the block
		const unsigned lo = b + off, hi = lo + cnt;
		for (unsigned i = lo + 1; i < hi; ++i) {
			const int key = perm[i];
			unsigned j = i;
			while (j > lo && perm[j - 1] > key) {
				perm[j] = perm[j - 1];
				--j;
			}
			perm[j] = key;
		}
		__syncthreads();
	}
}

__global__ void k_gather(const int *__restrict__ perm, int64_t n, const double *x, const double *y, const double *z,
                         const double *nx, const double *ny, const double *nz, const float *w, const float *iw,
                         const int *jk, const uint64_t *masks, int n_words, double *ox, double *oy, double *oz,
                         double *onx, double *ony, double *onz, float *ow, float *oiw, int *ojk, uint64_t *omasks) {
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int s = perm[i];
	ox[i] = x[s];
	oy[i] = y[s];
	oz[i] = z[s];
	onx[i] = nx[s];
	ony[i] = ny[s];
	onz[i] = nz[s];
	ow[i] = w[s];
	oiw[i] = iw[s];
	ojk[i] = jk[s];
	for (int k = 0; k < n_words; ++k) omasks[i * n_words + k] = masks[(int64_t)s * n_words + k];
}

__global__ void k_fill_u32(unsigned *p, int64_t n, unsigned v) {
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v) {
	int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

// LDS layout of the counting kernel (bytes), see lds_layout().
struct Lds {
	int tile, tables, slots, hist, total;
	int masks;   // first-side masks of the popcount rules (fp32 path; 0: not staged)
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Global-memory views of device pointers.  Loads and atomics through them
// compile to global_* instructions, which only the vector-memory counter
// tracks; through generic pointers they are flat_* instructions, which the
// LDS counter tracks too, so that every LDS wait after one also waits for
// the global access to complete.
#define GLOBAL_AS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const T GLOBAL_AS *gld(const T *p) { return (const T GLOBAL_AS *)p; }
__device__ __forceinline__ void gadd(unsigned long long *p, unsigned long long v) {
	__hip_atomic_fetch_add((unsigned long long GLOBAL_AS *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gadd(double *p, double v) {
	__hip_atomic_fetch_add((double GLOBAL_AS *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool WEIGHTED>
struct HistT;
template <>
struct HistT<false> {
	using T = unsigned int;
};
template <>
struct HistT<true> {
	using T = double;
};

// JKHists::interp (hist.cpp:369-391) on the up-weight table.  The segment
// (int)(bin_num(float(v)) - 0.5) comes from host-built float thresholds
// (glibc log1pf is not evaluated on the device); the interpolation is the
// reference's double expression in the same order.
__device__ double interp_up(const JobDev &J, double v) {
	const int n = J.up_n;
	const double GLOBAL_AS *b = gld(J.up_bins), *c = gld(J.up_centre);
	if (v < J.up_min) return b[0];
	if (v > J.up_max) return b[n - 1];
	const float fv = (float)v;
	const float GLOBAL_AS *thr = gld(J.up_thr);
	// the number of thresholds <= fv, by bisection (the thresholds are
	// monotone; a NaN fv counts none, as in a scan)
	int lo = 0, hi = n + 1;   // thr[k] <= fv for k < lo, > fv for k >= hi
	while (lo < hi) {
		const int mid = (lo + hi) >> 1;
		if (thr[mid] <= fv) lo = mid + 1;
		else hi = mid;
	}
	const int seg = lo - 1;
	if (seg < 0) return b[0];
	if (seg >= n - 1) return b[n - 1];
	return b[seg] + (b[seg + 1] - b[seg]) * (v - c[seg]) / (c[seg + 1] - c[seg]);
}

#define LDS_AS __attribute__((address_space(3)))
template <class P>
struct VecOf;
template <>
struct VecOf<const uint64_t GLOBAL_AS *> {
	using type = const u32x4 GLOBAL_AS *;
};
template <>
struct VecOf<const uint64_t LDS_AS *> {
	using type = const u32x4 LDS_AS *;
};

// popcount(m1 & m2) over nw words; m1 is a row in global memory or in LDS.
// With vec (both rows 16-byte aligned): CH words of each side per step in
// 16-byte loads, all issued before any is used.
template <int CH, class P1>
__device__ __forceinline__ int popc_rows(P1 m1, const uint64_t GLOBAL_AS *m2, int nw, bool vec) {
	int c = 0, k = 0;
	if (vec) {
		for (; k + CH <= nw; k += CH) {
			u32x4 a[CH / 2], b[CH / 2];
#pragma unroll
			for (int t = 0; t < CH / 2; ++t) {
				a[t] = ((typename VecOf<P1>::type)(m1 + k))[t];
				b[t] = ((const u32x4 GLOBAL_AS *)(m2 + k))[t];
			}
#pragma unroll
			for (int t = 0; t < CH / 2; ++t)
				c += __popc(a[t].x & b[t].x) + __popc(a[t].y & b[t].y) + __popc(a[t].z & b[t].z) +
				     __popc(a[t].w & b[t].w);
		}
	}
	for (; k < nw; ++k) c += __popcll(m1[k] & m2[k]);
	return c;
}

// float(popcount(m1 & m2)) over n_words words at word offset off of each
// object's masks (the bit mask at 0, the dither mask after it), 0 counting
// as 1 (correlators_invpweight.cpp:55-59, 62-68).  The first side's row is
// read from global memory, or from its LDS copy m1l when use_l (rows of
// J.mstage words).  CH words of each side per step (2 or 4); the batched
// s-mu kernel takes 2 to stay within its registers.
template <int CH = 4>
__device__ __forceinline__ float popc_and(const JobDev &J, int gi, int gj, int off, const uint64_t LDS_AS *m1l = nullptr,
                                          bool use_l = false) {
	const uint64_t GLOBAL_AS *m2 = gld(J.B.masks) + (size_t)gj * J.B.m_stride + off;
	const int c = use_l ? popc_rows<CH>(m1l + off, m2, J.n_words, ((J.mstage | J.B.m_stride | off) & 1) == 0)
	                    : popc_rows<CH>(gld(J.A.masks) + (size_t)gi * J.A.m_stride + off, m2, J.n_words,
	                                    ((J.A.m_stride | J.B.m_stride | off) & 1) == 0);
	float pc = (float)c;
	if (pc < 0.5f) pc += 1.0f;
	return pc;
}

template <int CH = 4>
__device__ __forceinline__ float popc_weight(const JobDev &J, int gi, int gj, const uint64_t LDS_AS *m1l = nullptr,
                                             bool use_l = false) {
	return popc_and<CH>(J, gi, gj, 0, m1l, use_l);
}

// The numerator of the inverse-p rules: n_runs, or under dither weights the
// dither popcount (correlators_invpweight.cpp:62-70, 749-757).  Both are
// integers, exact in float.
template <int CH = 4>
__device__ __forceinline__ float invp_numerator(const JobDev &J, int gi, int gj, const uint64_t LDS_AS *m1l = nullptr,
                                                bool use_l = false) {
	return J.dither ? popc_and<CH>(J, gi, gj, J.n_words, m1l, use_l) : (float)J.n_runs;
}

// Pair weight under the job's rule, in the reference's arithmetic
// (correlators.cpp:340; correlators_invpweight.cpp:507-533, 574, 743-764,
// 853, 901).  Every rule is symmetric in the two objects.
__device__ float pair_weight(const JobDev &J, int gi, int gj, const double *ni, const double *nj) {
	switch (J.wmode) {
	case TWOPCF_W_INVP_3D: {
		const float pc = popc_weight(J, gi, gj);
		double t = interp_up(J, ref_one_minus_cos(ni, nj)) * (double)invp_numerator(J, gi, gj);
		t = t / (double)pc;
		return J.A.iw[gi] * J.B.iw[gj] * (float)t;
	}
	case TWOPCF_W_ANGUP_3D: {
		const float p = J.A.w[gi] * J.B.w[gj];
		return (float)((double)p * interp_up(J, ref_one_minus_cos(ni, nj)));
	}
	case TWOPCF_W_ANG_IW:
		return J.A.iw[gi] * J.B.iw[gj];
	case TWOPCF_W_ANG_INVP: {
		const float pc = popc_weight(J, gi, gj);
		return J.A.iw[gi] * J.B.iw[gj] * (invp_numerator(J, gi, gj) / pc);
	}
	case TWOPCF_W_ANG_W:
	case TWOPCF_W_PLAIN:
	default:
		return J.A.w[gi] * J.B.w[gj];
	}
}

// The rules of pair_weight that need more than the two staged weights, given
// their float product wp (w1 * w2, or iw1 * iw2 for the input-weight rules):
// the same operations in the same order.
template <int CH>
__device__ float pair_weight_staged(const JobDev &J, int gi, int gj, float wp, const uint64_t LDS_AS *m1l,
                                    bool use_l) {
	// 1 - cos of the pair's unit vectors, read only by the rules that need it
	auto one_minus_cos = [&]() {
		const double ni[3] = {gld(J.A.nx)[gi], gld(J.A.ny)[gi], gld(J.A.nz)[gi]};
		const double nj[3] = {gld(J.B.nx)[gj], gld(J.B.ny)[gj], gld(J.B.nz)[gj]};
		return ref_one_minus_cos(ni, nj);
	};
	switch (J.wmode) {
	case TWOPCF_W_INVP_3D: {
		const float pc = popc_weight<CH>(J, gi, gj, m1l, use_l);
		double t = interp_up(J, one_minus_cos()) * (double)invp_numerator<CH>(J, gi, gj, m1l, use_l);
		t = t / (double)pc;
		return wp * (float)t;
	}
	case TWOPCF_W_ANGUP_3D:
		return (float)((double)wp * interp_up(J, one_minus_cos()));
	case TWOPCF_W_ANG_INVP: {
		const float pc = popc_weight<CH>(J, gi, gj, m1l, use_l);
		return wp * (invp_numerator<CH>(J, gi, gj, m1l, use_l) / pc);
	}
	default:
		return wp;
	}
}

template <int KIND, bool WEIGHTED>
__device__ void process_item(const JobDev &J, const WorkItem &W, const int4 *__restrict__ rows,
                             const int *__restrict__ slot_regions, char *smem, const Lds &L) {
	using HT = typename HistT<WEIGHTED>::T;
	constexpr int TILE = TILE_EXACT;
	const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
	const int ig = wave % IG, jg = wave / IG;
	const int nb = J.nb;
	const int n_slots = W.n_slots;
	const bool jk = J.n_jk > 1;

	// LDS carve-up: 6 doubles per tile entry, then w, gj, slot
	double *tx = (double *)(smem + L.tile);
	float *tw = (float *)(tx + 6 * TILE);
	int *tgj = (int *)(tw + TILE);
	int *tslot = tgj + TILE;
	float *tab = (float *)(smem + L.tables);
	int *sreg = (int *)(smem + L.slots);
	// the full histogram (none in LDS when it is kept in the global pool),
	// then the region slots
	const bool full_g = J.full_global != 0;
	HT *F = (HT *)(smem + L.hist);
	HT *H = F + (full_g ? 0 : nb);

	// tables and slot regions
	const int n_tab = L.slots - L.tables;
	(void)n_tab;
	for (int k = tid; k < n_slots; k += BLOCK) sreg[k] = slot_regions[W.slot_base + k];

	// first-side object of this lane
	const int li = ig * WAVE + lane;
	const bool ivalid = li < W.i_count;
	const int gi = W.i_begin + min(li, W.i_count - 1);
	const int d1 = J.dim[1], d2 = J.dim[2];
	const int cx = W.i_cell / (d1 * d2), cy = (W.i_cell / d2) % d1, cz = W.i_cell % d2;
	double xi[3], ni[3];
	xi[0] = J.A.x[gi];
	xi[1] = J.A.y[gi];
	xi[2] = J.A.z[gi];
	ni[0] = J.A.nx[gi];
	ni[1] = J.A.ny[gi];
	ni[2] = J.A.nz[gi];
	const float wi = J.A.w[gi];
	const int jki = J.A.jk[gi];
	__syncthreads();
	// region code: an LDS slot s >= 0, or -(region + 1) for a region without a
	// slot in this item, whose sums go straight to global memory
	int slot_i = 0;
	if (jk) {
		slot_i = -(jki + 1);
		for (int s = 0; s < n_slots; ++s)
			if (sreg[s] == jki) slot_i = s;
	}

	const float *t_s2 = tab + J.t_s2_off;
	const float *t1 = tab + J.t1_off;
	const float *t2 = tab + J.t2_off;

	unsigned int n_acc = 0;

	auto add_region = [&](int code, int bin, float w) {
		if (code >= 0) {
			if (WEIGHTED) atomicAdd((double *)&H[code * nb + bin], (double)w);
			else atomicAdd((unsigned int *)&H[code * nb + bin], 1u);
		} else {
			const size_t o = (size_t)(-code - 1) * nb + bin;
			if (WEIGHTED) gadd(&J.jk_d[o], (double)w);
			else gadd(&J.jk_u[o], 1ull);
		}
	};
	auto add = [&](int bin, float w, int slot_j) {
		if (full_g) {
			if (WEIGHTED) gadd(&J.full_d[bin], (double)w);
			else gadd(&J.full_u[bin], 1ull);
		} else if (WEIGHTED) {
			atomicAdd((double *)&F[bin], (double)w);
		} else {
			atomicAdd((unsigned int *)&F[bin], 1u);
		}
		if (jk) {
			add_region(slot_i, bin, w);
			if (slot_j != slot_i) add_region(slot_j, bin, w);
		}
	};

	for (int r = 0; r < J.n_rows; ++r) {
		if (!((W.row_mask >> r) & 1ull)) continue;
		const int4 row = rows[J.row_off + r];
		const int qx = cx + row.x, qy = cy + row.y;
		if (qx < 0 || qy < 0 || qx >= J.dim[0] || qy >= d1) continue;
		if (J.is_auto && (row.x < 0 || (row.x == 0 && row.y < 0))) continue;
		int z0 = max(cz + row.z, 0), z1 = min(cz + row.w, d2 - 1);
		if (J.is_auto && row.x == 0 && row.y == 0) z0 = max(z0, cz);
		if (z0 > z1) continue;
		const int base = (qx * d1 + qy) * d2;
		const int jb = J.B.start[base + z0], je = J.B.start[base + z1 + 1];
		for (int t0 = jb; t0 < je; t0 += TILE) {
			const int nt = min(TILE, je - t0);
			if (tid < nt) {
				const int gj = t0 + tid;
				const double bx = J.B.x[gj], by = J.B.y[gj], bz = J.B.z[gj];
				const double bnx = J.B.nx[gj], bny = J.B.ny[gj], bnz = J.B.nz[gj];
				const float bw = J.B.w[gj];
				int sl = 0;
				if (jk) {
					const int rj = J.B.jk[gj];
					sl = -(rj + 1);
					for (int s = 0; s < n_slots; ++s)
						if (sreg[s] == rj) sl = s;
				}
				tx[6 * tid + 0] = bx;
				tx[6 * tid + 1] = by;
				tx[6 * tid + 2] = bz;
				tx[6 * tid + 3] = bnx;
				tx[6 * tid + 4] = bny;
				tx[6 * tid + 5] = bnz;
				tw[tid] = bw;
				tgj[tid] = gj;
				tslot[tid] = sl;
			}
			__syncthreads();
			if (ivalid) {
				for (int q = jg; q < nt; q += JG) {
					int bin;
					const int gj = tgj[q];
					if (J.is_auto && gj <= gi) continue;
					const double *e = tx + 6 * q;
					if (!exact_bin<KIND>(xi, ni, e, e + 3, J.min_r_sq, J.max_r_sq, t_s2, t1, t2, J.ax1, J.ax2, bin, J.pi_def))
						continue;
					++n_acc;
					add(bin, J.wmode == TWOPCF_W_PLAIN ? wi * tw[q] : pair_weight(J, gi, gj, ni, e + 3), tslot[q]);
				}
			}
			__syncthreads();
		}
	}

	// flush and clear the LDS histograms
	for (int b = tid; b < (full_g ? 0 : nb); b += BLOCK) {
		HT v = F[b];
		if (v != (HT)0) {
			if (WEIGHTED) gadd(&J.full_d[b], (double)v);
			else gadd(&J.full_u[b], (unsigned long long)v);
			F[b] = (HT)0;
		}
	}
	if (jk) {
		for (int e = tid; e < n_slots * nb; e += BLOCK) {
			HT v = H[e];
			if (v != (HT)0) {
				const int s = e / nb, b = e - s * nb;
				const size_t o = (size_t)sreg[s] * nb + b;
				if (WEIGHTED) gadd(&J.jk_d[o], (double)v);
				else gadd(&J.jk_u[o], (unsigned long long)v);
				H[e] = (HT)0;
			}
		}
	}
	// pair statistics
	for (int off = 32; off; off >>= 1) n_acc += __shfl_xor(n_acc, off);
	if (lane == 0 && n_acc) gadd(&J.counters[0], (unsigned long long)n_acc);
	__syncthreads();
}

// FP32 path with per-wave compaction.  Screening: every wave tests its
// first-side objects (one per lane) against its share of the staged tile with
// FP32 s^2 from cell-local offsets; pairs that may be in range are appended
// to the wave's LDS ring.  Binning: whenever 64 pairs are queued, each lane
// takes one, recomputes the same s^2, bins it and proves the bin with the
// host-computed margins.  Pairs the margins cannot decide go to the wave's
// recheck queue (global indices), which is re-evaluated 64 at a time with
// the exact pipeline, away from the screening loop.
// MID: the job's pi is the pair-midpoint definition (sigma-pi and s-mu),
// chosen per item so that the per-pair code has no branch on it.
// RG: the jobs' histogram layout, fixed per launch so that the per-pair code
// has no branch on it either: RG_NONE no regions, RG_JK regions, both with
// the full histogram in LDS; RG_ANY read from the job (any layout, the full
// histogram possibly in the global pool).
template <int KIND, bool WEIGHTED, bool MID, int FZ, int RG>
__device__ void process_item_fast(const JobDev &J, const WorkItem &W, const int4 *__restrict__ rows,
                                  const int *__restrict__ slot_regions, char *smem, const Lds &L) {
	constexpr bool FUSE = FZ == FZ_MONO, TWIN = FZ == FZ_TWIN;
	static_assert(!FUSE || KIND == K_S_MU, "a monopole is fused into s-mu jobs only");
	static_assert(!TWIN || (KIND == K_ANGULAR && WEIGHTED && RG == RG_NONE), "twins: weighted angular jobs only");
	using HT = typename HistT<WEIGHTED>::T;
	constexpr int TILE = TILE_FAST;
	constexpr int NP = np_of(KIND, WEIGHTED);   // pairs per lane in binning (MID: one)
	const int tid = threadIdx.x, lane = tid & (WAVE - 1);
	// wave index, made visibly uniform so that the loops below run on the
	// scalar unit
	const int wave = __builtin_amdgcn_readfirstlane(tid / WAVE);
	// wave groups: ig_n groups over the first-side objects (one object per
	// lane), jg_n groups over each staged tile; small items use fewer
	// first-side groups so that no wave idles
	const int ig_n = W.i_count > 2 * WAVE ? IG : (W.i_count > WAVE ? 2 : 1);
	const int jg_n = (BLOCK / WAVE) / ig_n;
	// The waves of a block are placed on the CU's four SIMDs in turn (wave w
	// on SIMD w mod 4).  Each group's waves are spread over the four SIMDs
	// (groups take runs of four consecutive waves), so that a group with more
	// pairs than the others loads every SIMD alike instead of one.
	const int ig = (wave >> 2) % ig_n, jg = (wave & 3) | (((wave >> 2) / ig_n) << 2);
	// The item's row blocks of 16 first-side objects are dealt to the groups
	// in turn (row block b to group b mod ig_n, as its (b / ig_n)-th), so that
	// the groups of a partial item differ by one row block at most.  The LDS
	// chunk keeps each group's row blocks together: object i of the item is
	// in slot 64 g + 16 k + (i mod 16), and ring entries name slots.
	const int n_rbk = (W.i_count + 15) >> 4;
	auto slot_of = [&](int i) { const int b = i >> 4; return ((b % ig_n) << 6) + ((b / ig_n) << 4) + (i & 15); };
	auto item_of = [&](int s) { return ((((s >> 4) & 3) * ig_n + (s >> 6)) << 4) + (s & 15); };
	const int nb = J.nb, n1 = J.n1, n2 = J.n2;
	const bool is_auto = J.is_auto != 0;
	const bool jk = RG == RG_ANY ? J.n_jk > 1 : RG == RG_JK;
	const int n_slots = W.n_slots;
	// every region the item meets has an LDS slot: no lane's pair needs the
	// global pool (the region test per pair is skipped)
	const bool all_slotted = W.all_slotted != 0;
	// Every first-side object of the item is in region R = W.uf_region: the
	// full histogram (in LDS) then takes every pair, which is also region R's
	// touching count of this item, and a pair adds to a slot histogram only
	// when its second object lies in another region.  The item's R sums are
	// the full histogram's, added at the end.
	const bool uf = jk && W.uf_region >= 0;
	// weights staged with the objects: the input weights for the rules built
	// on them, else the weights
	const int wmode = J.wmode;
	const bool stage_iw = wmode == TWOPCF_W_ANG_IW || wmode == TWOPCF_W_ANG_INVP || wmode == TWOPCF_W_INVP_3D;
	const bool needs_more = wmode == TWOPCF_W_ANG_INVP || wmode == TWOPCF_W_INVP_3D || wmode == TWOPCF_W_ANGUP_3D;

	float4 *tfa = (float4 *)(smem + L.tile);   // (o0, o1, o2, |o|^2)
	float4 *tfb = tfa + TILE;                  // (n0, n1, n2, w)
	int *tslot = (int *)(tfb + TILE);
	float4 *cia = (float4 *)(tslot + TILE);    // first side: (o0, o1, o2, w)
	float4 *cib = cia + ICHUNK;                // (n0, n1, n2, region code)
	float *cth = (float *)(cib + ICHUNK);      // screening thresholds per first-side object: upper
	float *ctl = cth + ICHUNK;                 //   and lower
	// ring entries: (tile entry << 6) | first-side object within the wave's 64
	unsigned short *ring = (unsigned short *)(ctl + ICHUNK) + wave * RING;
	// recheck queue: second-side object (global index) and first-side object
	int *rq_g = (int *)((unsigned short *)(ctl + ICHUNK) + (BLOCK / WAVE) * RING) + wave * RQ;
	unsigned char *rq_l = (unsigned char *)((int *)((unsigned short *)(ctl + ICHUNK) + (BLOCK / WAVE) * RING) +
	                                        (BLOCK / WAVE) * RQ) + wave * RQ;
	// region -> code (the LDS byte address of slot s's histogram for LDS slot
	// s, else -(region + 1)) for this item
	int *rmap = (int *)((unsigned char *)((int *)((unsigned short *)(ctl + ICHUNK) + (BLOCK / WAVE) * RING) +
	                                      (BLOCK / WAVE) * RQ) + (BLOCK / WAVE) * RQ);
	// z of the first and of the last entry of each 16-entry tile block
	float *tbz = (float *)(rmap + REGION_MAP);
	float *tbe = tbz + TILE_BLOCKS;
	// per stencil row of the item: the second side's objects [x, y) (empty
	// for rows outside the grid or the auto count's excluded half); entry
	// ROW_SPANS: the end of the first-side cell (auto counts)
	int2 *rspan = (int2 *)(tbe + TILE_BLOCKS);
	float *tab = (float *)(smem + L.tables);
	int *sreg = (int *)(smem + L.slots);
	// the full histogram: 2^f_shift copies, bin-major, lane l adding to copy
	// l mod 2^f_shift (small histograms: lanes of one atomic never collide)
	// (none in LDS when the histogram is kept in the global pool)
	const int f_shift = J.f_shift;
	const int f_lane = lane & ((1 << f_shift) - 1);
	const bool full_g = RG == RG_ANY && J.full_global != 0;
	HT *F = (HT *)(smem + L.hist);
	HT *H = F + (full_g ? 0 : (nb << f_shift));
	// region codes of slots are LDS byte addresses: slot s's bin b is at
	// code + (b << hb), one address operation per atomic
	constexpr int hb = WEIGHTED ? 3 : 2;
	const int h_lds = (int)(uintptr_t)(HT LDS_AS *)H;
	const int slot_bytes = nb << hb;

	for (int k = tid; k < n_slots; k += BLOCK) sreg[k] = slot_regions[W.slot_base + k];
	const bool use_map = jk && J.n_jk <= REGION_MAP;
	// the region map: every region without a slot here, then (after a
	// barrier) each slot's region
	if (use_map)
		for (int r = tid; r < J.n_jk; r += BLOCK) rmap[r] = -(r + 1);
	auto code_of = [&](int r) {
		if (use_map) return rmap[r];
		int code = -(r + 1);
		for (int k = 0; k < n_slots; ++k)
			if (sreg[k] == r) code = h_lds + k * slot_bytes;
		return code;
	};
	// below s2_lo: certainly out of range; at or above s2_hi: certainly out
	// (of both ranges, with a fused monopole)
	float s2_lo_v = tab[J.tB_off], s2_hi_v = tab[J.tA_off + n1];
	if constexpr (FUSE) {
		s2_lo_v = fminf(s2_lo_v, tab[J.m_tB_off]);
		s2_hi_v = fmaxf(s2_hi_v, tab[J.m_tA_off + J.m_nb]);
	}
	const float s2_lo = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s2_lo_v)));
	const float s2_hi = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s2_hi_v)));
	const bool check_lo = s2_lo > 0.0f;       // with s_min = 0 every s^2 >= 0 passes
	const float s2_lo_eff = check_lo ? s2_lo : -INFINITY;
	const int d1 = J.dim[1], d2 = J.dim[2];
	const int cx = W.i_cell / (d1 * d2), cy = (W.i_cell / d2) % d1, cz = W.i_cell % d2;
	const double c0 = J.lo[0] + (cx + 0.5) * J.h, c1 = J.lo[1] + (cy + 0.5) * J.h, c2 = J.lo[2] + (cz + 0.5) * J.h;
	// s-mu with the angular pi stages its normals times n2 / 2, so that
	// mu n2 = |d.(n1 + n2)| rsq(s2) takes no scale multiply per pair
	const float nscale = (KIND == K_S_MU && !MID) ? 0.5f * (float)n2 : 1.0f;
	// the item's row spans, one thread per stencil row, read from the cell
	// starts once per item instead of once per row and wave
	if (tid < J.n_rows || tid == ROW_SPANS) {
		int2 sp = make_int2(0, 0);
		if (tid == ROW_SPANS) {
			if (is_auto) sp.x = gld(J.B.start)[W.i_cell + 1];
		} else if ((W.row_mask >> tid) & 1ull) {
			const int4 row = rows[J.row_off + tid];
			const int qx = cx + row.x, qy = cy + row.y;
			if (qx >= 0 && qy >= 0 && qx < J.dim[0] && qy < d1 && !(is_auto && (row.x < 0 || (row.x == 0 && row.y < 0)))) {
				int z0 = max(cz + row.z, 0);
				const int z1 = min(cz + row.w, d2 - 1);
				if (is_auto && row.x == 0 && row.y == 0) z0 = max(z0, cz);
				const int base = (qx * d1 + qy) * d2;
				if (z0 <= z1) sp = make_int2(gld(J.B.start)[base + z0], gld(J.B.start)[base + z1 + 1]);
			}
		}
		rspan[tid] = sp;
	}
	__syncthreads();
	if (use_map) {
		for (int k = tid; k < n_slots; k += BLOCK) rmap[sreg[k]] = h_lds + k * slot_bytes;
		__syncthreads();
	}
	if (tid < n_rbk * 16) {
		const int g = W.i_begin + min(tid, W.i_count - 1);
		const int s = slot_of(tid);
		// every global load first, so that one latency covers them all
		// (angular counts work on the unit vectors: 1 - n1.n2 = |n1 - n2|^2 / 2)
		const double GLOBAL_AS *px = gld(KIND == K_ANGULAR ? J.A.nx : J.A.x);
		const double GLOBAL_AS *py = gld(KIND == K_ANGULAR ? J.A.ny : J.A.y);
		const double GLOBAL_AS *pz = gld(KIND == K_ANGULAR ? J.A.nz : J.A.z);
		const double xg = px[g], yg = py[g], zg = pz[g];
		const double nxg = gld(J.A.nx)[g], nyg = gld(J.A.ny)[g], nzg = gld(J.A.nz)[g];
		const float wg = stage_iw ? gld(J.A.iw)[g] : gld(J.A.w)[g];
		// a twin's other weight takes the place of the region code
		const float wg2 = TWIN && J.t_w ? (stage_iw ? gld(J.A.w)[g] : gld(J.A.iw)[g]) : 0.0f;
		const int rg = jk ? gld(J.A.jk)[g] : 0;
		const int code = jk ? code_of(rg) : TWIN ? __float_as_int(wg2) : 0;
		const float ox = (float)(xg - c0), oy = (float)(yg - c1), oz = (float)(zg - c2);
		cia[s] = make_float4(ox, oy, oz, wg);
		cib[s] = make_float4(nscale * (float)nxg, nscale * (float)nyg, nscale * (float)nzg, __int_as_float(code));
		// s^2 = |a|^2 + (|b|^2 - 2 a.b) may be in range iff the bracket lies
		// within [s2_lo - eps - |a|^2, s2_hi + eps - |a|^2); objects past the
		// end of the item never pass
		const float aa = fmaf(oz, oz, fmaf(oy, oy, ox * ox));
		const bool ok = tid < W.i_count;
		cth[s] = ok ? (s2_hi + J.scr_eps) - aa : -1e30f;
		ctl[s] = ok ? (s2_lo - J.scr_eps) - aa : 1e30f;
	}
	// the popcount rules read the first side's masks from an LDS copy when
	// the launch has room for it: rows of J.mstage words (bit mask, then the
	// dither mask), copied in 16-byte units.  Only the angular and monopole
	// kernels (small histograms) have the room; the s-mu and sigma-pi ones
	// keep their registers for binning.
	constexpr bool can_stage = WEIGHTED && (KIND == K_ANGULAR || KIND == K_MONO);
	const bool masks_in_lds = can_stage && L.masks != 0 && J.mstage != 0;
	const uint64_t LDS_AS *m_lds = (const uint64_t LDS_AS *)(smem + L.masks);
	if (masks_in_lds) {
		const int ms = J.A.m_stride, upo = ms >> 1;
		const u32x4 GLOBAL_AS *src = (const u32x4 GLOBAL_AS *)(gld(J.A.masks) + (size_t)W.i_begin * ms);
		uint64_t LDS_AS *dst = (uint64_t LDS_AS *)(smem + L.masks);
		for (int e = tid; e < W.i_count * upo; e += BLOCK) {
			const int o = e / upo, u = e - o * upo;
			*(u32x4 LDS_AS *)(dst + (size_t)slot_of(o) * J.mstage + 2 * u) = src[e];
		}
	}
	__syncthreads();
	// row blocks of 16 first-side objects this wave screens (uniform)
	const int rb_n = ig < n_rbk ? min(4, (n_rbk - ig + ig_n - 1) / ig_n) : 0;

	const float2 *tq = (const float2 *)(tab + J.tq_off);

	// accepted and rechecked pairs: per lane (the recheck) and per wave
	// (binning, counted from its lane masks on the scalar unit)
	unsigned int n_acc = 0, n_rc = 0, n_acc2 = 0;
	unsigned int n_acc_w = 0, n_acc2_w = 0;

	// an LDS atomic at byte address a
	auto lds_add = [](int a, float w) {
		if (WEIGHTED)
			__hip_atomic_fetch_add((double LDS_AS *)(uintptr_t)(unsigned)a, (double)w, __ATOMIC_RELAXED,
			                       __HIP_MEMORY_SCOPE_WORKGROUP);
		else
			__hip_atomic_fetch_add((unsigned int LDS_AS *)(uintptr_t)(unsigned)a, 1u, __ATOMIC_RELAXED,
			                       __HIP_MEMORY_SCOPE_WORKGROUP);
	};
	// Region codes: the histogram's LDS byte address for a region with an LDS
	// slot, -(region + 1) for a region without a slot in this item (its sums
	// go straight to global memory).
	auto add_region = [&](int code, int bin, float w) {
		if (code >= 0) {
			lds_add(code + (bin << hb), w);
		} else {
			const size_t o = (size_t)(-code - 1) * nb + bin;
			if (WEIGHTED) gadd(&J.jk_d[o], (double)w);
			else gadd(&J.jk_u[o], 1ull);
		}
	};
	// lane's copy of the full histogram; bin b is f_stride << b bytes further.
	// Unweighted jackknife counts keep in it only the pairs whose two
	// regions differ: each such pair touches two regions, so the full count
	// is the sum of the touching counts over all regions minus this cross
	// count, which the host forms exactly in integers.  A pair within one
	// region then takes one LDS atomic instead of two.
	char *const F_lane = (char *)(F + f_lane);
	const int f_bshift = f_shift + (WEIGHTED ? 3 : 2);
	const bool cross_only = !WEIGHTED && jk;
	auto add_full = [&](int bin, float w) {
		if (full_g) {
			if (WEIGHTED) gadd(&J.full_d[bin], (double)w);
			else gadd(&J.full_u[bin], 1ull);
			return;
		}
		HT *const fp = (HT *)(F_lane + (bin << f_bshift));
		if (WEIGHTED) atomicAdd((double *)fp, (double)w);
		else atomicAdd((unsigned int *)fp, 1u);
	};
	auto add = [&](int bin, float w, int si, int sj) {
		if (uf) {
			add_full(bin, w);
			if (sj != si) {
				if (all_slotted || __builtin_amdgcn_ballot_w64(sj < 0) == 0ull) {
					lds_add(sj + (bin << hb), w);
				} else {
					add_region(sj, bin, w);
					// the cross count of a pair whose second region has no
					// slot (slotted ones are summed from the slots at the end)
					if (cross_only && sj < 0) gadd(&J.full_u[bin], 1ull);
				}
			}
			return;
		}
		if (!cross_only) add_full(bin, w);
		if (jk) {
			if (all_slotted || __builtin_amdgcn_ballot_w64(si < 0 || sj < 0) == 0ull) {
				// every active lane's regions have LDS slots
				lds_add(si + (bin << hb), w);
				if (sj != si) {
					lds_add(sj + (bin << hb), w);
					if (cross_only) add_full(bin, w);
				}
			} else {
				add_region(si, bin, w);
				if (sj != si) {
					add_region(sj, bin, w);
					if (cross_only) add_full(bin, w);
				}
			}
		}
	};
	// the fused monopole's (or the twin's) full histogram follows the job's
	// own (no regions)
	char *const F2_lane = (char *)(F + (nb << f_shift) + (lane & ((1 << J.m_f_shift) - 1)));
	const int f2_bshift = J.m_f_shift + (WEIGHTED ? 3 : 2);
	auto add2 = [&](int bin, float w) {
		HT *const fp = (HT *)(F2_lane + (bin << f2_bshift));
		if (WEIGHTED) atomicAdd((double *)fp, (double)w);
		else atomicAdd((unsigned int *)fp, 1u);
	};
	// wp: the float product of the two staged weights (w, or iw for the
	// rules built on the input weights); rules that need more read it here
	auto weight = [&](int l2, int g2, float wp) {
		if (WEIGHTED && needs_more) {
			const int g1 = W.i_begin + item_of(l2);
			wp = pair_weight_staged<KIND == K_S_MU ? 2 : 4>(J, g1, g2, wp, m_lds + (size_t)l2 * J.mstage,
			                                                 masks_in_lds);
		}
		return wp;
	};

	// recheck queue: the exact pipeline on up to RQ pairs, 64 lanes at a time
	unsigned rq_n = 0;
	auto flush_rq = [&]() {
		__builtin_amdgcn_wave_barrier();
		for (unsigned b = 0; b < rq_n; b += WAVE) {
			if (b + lane < rq_n) {
				++n_rc;
				const int2 e = make_int2(rq_l[b + lane], rq_g[b + lane]);
				const int g1 = W.i_begin + item_of(e.x), g2 = e.y;
				const double xi[3] = {gld(J.A.x)[g1], gld(J.A.y)[g1], gld(J.A.z)[g1]};
				const double ni[3] = {gld(J.A.nx)[g1], gld(J.A.ny)[g1], gld(J.A.nz)[g1]};
				const double xj[3] = {gld(J.B.x)[g2], gld(J.B.y)[g2], gld(J.B.z)[g2]};
				const double nj[3] = {gld(J.B.nx)[g2], gld(J.B.ny)[g2], gld(J.B.nz)[g2]};
				const float *t_s2 = tab + J.t_s2_off;
				const float *t1 = tab + J.t1_off;
				const float *t2 = tab + J.t2_off;
				int bin;
				if (exact_bin<KIND>(xi, ni, xj, nj, J.min_r_sq, J.max_r_sq, t_s2, t1, t2, J.ax1, J.ax2, bin, J.pi_def)) {
					++n_acc;
					const int sj = jk ? code_of(gld(J.B.jk)[g2]) : 0;
					const float w = weight(e.x, g2, cia[e.x].w * (stage_iw ? gld(J.B.iw)[g2] : gld(J.B.w)[g2]));
					add(bin, w, __float_as_int(cib[e.x].w), sj);
					if constexpr (TWIN) {
						++n_acc2;
						add2(bin, J.t_w ? cib[e.x].w * (stage_iw ? gld(J.B.w)[g2] : gld(J.B.iw)[g2])
						                : cia[e.x].w * (stage_iw ? gld(J.B.iw)[g2] : gld(J.B.w)[g2]));
					}
				}
				if constexpr (FUSE) {
					const float *m_t = tab + J.m_t_s2_off;
					if (exact_bin<K_MONO>(xi, ni, xj, nj, J.min_r_sq, J.max_r_sq, m_t, m_t, m_t, J.m_ax1, J.m_ax1, bin)) {
						++n_acc2;
						add2(bin, weight(e.x, g2, cia[e.x].w * (stage_iw ? gld(J.B.iw)[g2] : gld(J.B.w)[g2])));
					}
				}
			}
		}
		rq_n = 0;
		// Every load above has landed before the loop goes on (vmcnt 0):
		// the exact pipeline loads some operands on paths that never use
		// them, and without this wait the compiler keeps them pending into
		// the screening and binning loops, where its waits for them would
		// also wait for every global atomic issued since (regions without
		// an LDS slot).
		__builtin_amdgcn_s_waitcnt(VMCNT0);
		__builtin_amdgcn_wave_barrier();
	};

	// closed-form bin estimates without a per-pair select between log and
	// linear axes: k = (int)(inv_step v + ilb ln2 log2(1 + ils v)), with
	// inv_step = 0 on a log axis and ils = ilb = 0 on a linear one (log2(1) = 0)
	const bool s_log = J.s_log != 0, p_log = J.p_log != 0;
	const float s_min = J.s_min;
	const float s_ilb = s_log ? J.s_ilb_ln2 : 0.0f, s_ils = s_log ? J.s_ils : 0.0f, s_inv = s_log ? 0.0f : J.s_inv_step;
	const float p_ilb = p_log ? J.p_ilb_ln2 : 0.0f, p_ils = p_log ? J.p_ils : 0.0f, p_inv = p_log ? 0.0f : J.p_inv_step;
	// clamped to [0, n - 1] before the conversion (one v_med3); a NaN
	// estimate converts to 0, and any index in range is safe because the
	// thresholds of that bin then decide
	auto bin_estimate = [](float v, float inv, float ilb, float ils, float top) {
		return (int)__builtin_amdgcn_fmed3f(fmaf(inv, v, ilb * __builtin_amdgcn_logf(fmaf(ils, v, 1.0f))), 0.0f, top);
	};
	const float n1_top = (float)(n1 - 1), n2_top = (float)(n2 - 1);
	// the fused monopole's estimate (its s minimum is the s-mu one) and its
	// (A_k, B_k+1) pairs
	const float m_ilb = FUSE ? J.m_s_ilb_ln2 : 0.0f, m_ils = FUSE ? J.m_s_ils : 0.0f;
	const float m_inv = FUSE ? J.m_s_inv_step : 0.0f, m_top = FUSE ? (float)(J.m_nb - 1) : 0.0f;
	const float2 *mtq = (const float2 *)(tab + (FUSE ? J.m_tq_off : 0));
	const float mu_h0 = J.mu_h0, mu_c2 = J.mu_c2;
	// mu_h0 held in a vector register: the margin's multiply-add then reads
	// one scalar operand, with no per-pair copy
	float mu_h0_v;
	asm volatile("v_mov_b32 %0, %1" : "=v"(mu_h0_v) : "s"(mu_h0));
	const float *x1t = tab + J.x1_off, *t2t = tab + J.t2_off;
	// sigma-pi: the outer thresholds of both axes (uniform)
	const float sp_x0 = KIND == K_SIGMA_PI ? x1t[0] : 0.0f, sp_xn = KIND == K_SIGMA_PI ? x1t[n1] : 0.0f;
	const float sp_t0 = KIND == K_SIGMA_PI ? t2t[0] : 0.0f, sp_tn = KIND == K_SIGMA_PI ? t2t[n2] : 0.0f;
	const float rA0 = tab[J.tA_off], rBn = tab[J.tB_off + n1];
	// line of sight in FP32: the angular definition from the float unit
	// vectors, or the pair midpoint from 2c + the two cell-local offsets;
	// inv_mid = 1 / |x1 + x2| for the midpoint (it widens the margins), else 0
	const float mid_c = J.mid_c;
	const float c2x = (float)(2.0 * c0), c2y = (float)(2.0 * c1), c2z = (float)(2.0 * c2);
	// (mid: std::integral_constant<bool, MID>)
	auto fp32_pi = [&](auto mid, float dx, float dy, float dz, const float4 &ia, const float4 &ib, const float4 &a,
	                   const float4 &b, float &inv_mid) {
		if constexpr (decltype(mid)::value) {
			const float sx = c2x + (ia.x + a.x), sy = c2y + (ia.y + a.y), sz = c2z + (ia.z + a.z);
			inv_mid = __builtin_amdgcn_rsqf(fmaf(sz, sz, fmaf(sy, sy, sx * sx)));
			return fmaf(dz, sz, fmaf(dy, sy, dx * sx)) * inv_mid;
		}
		inv_mid = 0.0f;
		return 0.5f * fmaf(dz, ib.z + b.z, fmaf(dy, ib.y + b.y, dx * (ib.x + b.x)));
	};
	// The FP32 decision for the pair of ring slot idx: bin, whether it is
	// in range, and whether the margins prove both (sure); the ring holds
	// pairs of the current tile, whose first entry is object t_base of the
	// second side.  The decisions are kept as lane masks (scalar registers):
	// the lanes whose pair is not sure, and those whose pair is in range.
	// Every lane of the wave evaluates a pair (uniform control flow).
	int t_base = 0;
	struct PairEval {
		unsigned long long m_ns, m_in;
		int bin, bin2, l2, q;   // bin2: the fused monopole's bin
		float wp;     // product of the staged weights
		int si, sj;   // region codes of the two objects
	};
	auto eval = [&](auto mid, int l2, int q, const float4 &ia, const float4 &ib, const float4 &a, const float4 &b) {
		PairEval r;
		r.l2 = l2;
		r.q = q;
		r.wp = WEIGHTED ? ia.w * b.w : 1.0f;
		r.si = __float_as_int(ib.w);
		r.sj = WEIGHTED ? tslot[q] : __float_as_int(b.w);
		const float dx = ia.x - a.x, dy = ia.y - a.y, dz = ia.z - a.z;
		const float s2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
		// a ballot of each comparison is that comparison's own lane mask; the
		// masks are combined on the scalar unit
		unsigned long long m_ns, m_in = ~0ull;
		int bin;
		if constexpr (KIND == K_SIGMA_PI) {
			// range test on s^2, sigma in the x = s^2 - pi^2 domain
			// against sqrtf-domain thresholds, |pi| against its value
			// thresholds; margins per pair from the bounds on s^2 and pi.
			// Every threshold is loaded and every comparison made (no
			// branches); the decisions are combined as lane masks.
			float inv_mid;
			const float p = fp32_pi(mid, dx, dy, dz, ia, ib, a, b, inv_mid);
			const float ap = fabsf(p);
			const float s = __builtin_amdgcn_sqrtf(s2);
			float dpi = fmaf(J.pi_c1, s, J.pi_c2);
			if constexpr (decltype(mid)::value) dpi += mid_c * s * inv_mid;
			const float x = fmaf(-p, p, s2);
			const float ex = fmaf(J.e_c1, s2, J.e_c2 * s) + dpi * fmaf(2.0f, ap, dpi) + 3.6e-7f * fmaf(p, p, s2);
			const float sg = __builtin_amdgcn_sqrtf(fmaxf(x, 0.0f)) - s_min;
			const int k1 = bin_estimate(sg, s_inv, s_ilb, s_ils, n1_top);
			const float pv = ap - J.p_min;
			const int k2 = bin_estimate(pv, p_inv, p_ilb, p_ils, n2_top);
			const float x_lo = x1t[k1], x_hi = x1t[k1 + 1], t_lo = t2t[k2], t_hi = t2t[k2 + 1];
			// r_in: min_r_sq < s2_ref < max_r_sq
			const unsigned long long m_rin = __builtin_amdgcn_ballot_w64(s2 >= rA0) & __builtin_amdgcn_ballot_w64(s2 < rBn);
			// certainly outside the histogram's sigma or pi range
			const unsigned long long m_out =
			    __builtin_amdgcn_ballot_w64(x < sp_x0 - ex) | __builtin_amdgcn_ballot_w64(x >= sp_xn + ex) |
			    __builtin_amdgcn_ballot_w64(ap < sp_t0 - dpi) | __builtin_amdgcn_ballot_w64(ap >= sp_tn + dpi);
			// both bins proved
			const unsigned long long m_sure12 =
			    __builtin_amdgcn_ballot_w64(x >= x_lo + ex) & __builtin_amdgcn_ballot_w64(x < x_hi - ex) &
			    __builtin_amdgcn_ballot_w64(ap >= t_lo + dpi) & __builtin_amdgcn_ballot_w64(ap < t_hi - dpi);
			// within the screening's range (not from its error shell)
			const unsigned long long m_shell =
			    __builtin_amdgcn_ballot_w64(s2 < s2_hi) & __builtin_amdgcn_ballot_w64(s2 >= s2_lo_eff);
			// in range: not from the shell and not in range but certainly
			// outside the histogram; sure: out of range, or in range with
			// both bins proved
			m_in = m_shell & ~(m_rin & m_out);
			m_ns = m_shell & ~(m_rin & (m_out | m_sure12));
			bin = __mul24(k1, n2) + k2;
		} else {
			// bin: closed-form estimate, checked against the exact thresholds,
			// proved by the error margins; the axis variable is s, or
			// 1 - cos = chord^2 / 2 for angular counts.  Pairs the screening
			// let through from its error shell (s^2 within eps outside the
			// range) are not sure in any bin: the first bin's lower threshold
			// lies above s2_lo and the last bin's upper one below s2_hi, so
			// the recheck drops them, which costs less than a range test on
			// every pair.  With a fused monopole the screening range is the
			// union of two: pairs out of the s-mu range are dropped here.
			if constexpr (FUSE)
				m_in = __builtin_amdgcn_ballot_w64(s2 < s2_hi) & __builtin_amdgcn_ballot_w64(s2 >= s2_lo_eff);
			// s-mu takes s for its estimate as s2 rsq(s2).  At s2 = 0 that is
			// 0 * inf = NaN: the estimate clamps to a valid bin, whose lower
			// threshold T_k + E > 0 fails, and mu is NaN, so the pair is never
			// sure (it is out of range, or rechecked when s_min = 0)
			const float rs = KIND == K_S_MU ? __builtin_amdgcn_rsqf(s2) : 0.0f;
			const float sv = KIND == K_ANGULAR ? fmaf(0.5f, s2, -s_min)
			               : KIND == K_S_MU    ? fmaf(s2, rs, -s_min)
			                                   : __builtin_amdgcn_sqrtf(s2) - s_min;
			const int k = bin_estimate(sv, s_inv, s_ilb, s_ils, n1_top);
			const float2 q = tq[k];
			// sure: T_k + E <= s2 < T_k+1 - E
			m_ns = __builtin_amdgcn_ballot_w64(!(s2 >= q.x)) | __builtin_amdgcn_ballot_w64(!(s2 < q.y));
			if (KIND == K_S_MU) {
				const float nmu = (float)n2;
				// mu n2 = |pi| rsq(s2) n2: the midpoint pi times n2, or for the
				// angular pi d.(n1 + n2) from normals staged times n2 / 2
				float inv_mid, xm;
				if constexpr (decltype(mid)::value) {
					xm = fabsf(fp32_pi(mid, dx, dy, dz, ia, ib, a, b, inv_mid)) * rs * nmu;
				} else {
					inv_mid = 0.0f;
					xm = fabsf(fmaf(dz, ib.z + b.z, fmaf(dy, ib.y + b.y, dx * (ib.x + b.x)))) * rs;
				}
				// Bin km of mu is proved when xm lies at least the margin
				// delta = c1 + c2 rsq(s2) (+ the midpoint term) inside it:
				// |fract(xm) - 1/2| <= 1/2 - delta.  The host's h0 = 1/2 - c1 -
				// 2^-21 leaves room for this test's own roundings.  km < n2
				// follows: mu <= 1, so xm exceeds n2 by at most its error, c1,
				// and then fract(xm) <= c1 fails the test.  At mu = 0 the lower
				// edge is not a bin edge; the test still asks for the margin
				// there, which only rechecks pairs with mu below about 1e-5.
				float hm = fmaf(-mu_c2, rs, mu_h0_v);
				if constexpr (decltype(mid)::value) hm -= nmu * mid_c * inv_mid;
				const int km = (int)xm;
				const float fr = __builtin_amdgcn_fractf(xm);
				m_ns |= __builtin_amdgcn_ballot_w64(!(fabsf(fr - 0.5f) <= hm));
				bin = __mul24(k, n2) + km;
				if constexpr (FUSE) {
					// the monopole bin from the same s: a pair goes to the
					// histograms only when both bins are proved, else both
					// decisions are rechecked
					const int k2 = bin_estimate(sv, m_inv, m_ilb, m_ils, m_top);
					const float2 qm = mtq[k2];
					m_ns |= __builtin_amdgcn_ballot_w64(!(s2 >= qm.x)) | __builtin_amdgcn_ballot_w64(!(s2 < qm.y));
					r.bin2 = k2;
				}
			} else {
				bin = k;
			}
			// out of the s-mu range: certainly not counted
			if constexpr (FUSE) m_ns &= m_in;
		}
		r.m_ns = m_ns;
		r.m_in = m_in;
		r.bin = bin;
		return r;
	};
	// bin n <= 2 * WAVE queued pairs, two per lane (independent chains for
	// the scheduler to interleave); pairs the margins cannot decide go to the
	// recheck queue
	auto push = [&](unsigned long long mr, const PairEval &p) {
		if (mr) {
			if (__builtin_amdgcn_inverse_ballot_w64(mr)) {
				const unsigned off = __builtin_amdgcn_mbcnt_hi((unsigned)(mr >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mr, rq_n));
				rq_l[off] = (unsigned char)p.l2;
				rq_g[off] = t_base + p.q;
			}
			rq_n += (unsigned)__popcll(mr);
		}
	};
	const std::integral_constant<bool, MID> mid;
	// Binning raises the wave's priority: a wave whose pairs' operands are in
	// flight is issued ahead of the waves that screen, which shortens the
	// dependent LDS chains of binning.
	auto consume = [&](unsigned head, unsigned n) {
		__builtin_amdgcn_s_setprio(2);
		__builtin_amdgcn_wave_barrier();
		if constexpr (MID) {
			// the midpoint's longer per-pair code: one pair per lane at a time
			for (unsigned h = 0; h < n; h += WAVE) {
				const bool v = h + lane < n;
				const unsigned e = ring[(head & (RING - 1)) + h + lane];
				const int l = ig * WAVE + (int)(e & 63u), q = min((int)(e >> 6), TILE - 1);
				const PairEval p = eval(mid, l, q, cia[l], cib[l], tfa[q], tfb[q]);
				const unsigned long long m_v = __builtin_amdgcn_ballot_w64(v);
				const unsigned long long m_acc = m_v & p.m_in & ~p.m_ns, m_rc = m_v & p.m_ns;
				if (__builtin_amdgcn_inverse_ballot_w64(m_acc)) {
					const float w = weight(p.l2, t_base + p.q, p.wp);
					add(p.bin, w, p.si, p.sj);
					if constexpr (FUSE) add2(p.bin2, w);
				}
				n_acc_w += (unsigned)__popcll(m_acc);
				if constexpr (FUSE) n_acc2_w += (unsigned)__popcll(m_acc);
				push(m_rc, p);
				__builtin_amdgcn_wave_barrier();
			}
		} else {
			// NP pairs per lane, all their operands loaded before any is
			// evaluated (independent chains for the scheduler to interleave)
			int l[NP], q[NP];
			float4 ia[NP], ib[NP], a[NP], b[NP];
#pragma unroll
			for (int k = 0; k < NP; ++k) {
				const unsigned e = ring[(head & (RING - 1)) + k * WAVE + lane];
				l[k] = ig * WAVE + (int)(e & 63u);
				q[k] = min((int)(e >> 6), TILE - 1);
			}
#pragma unroll
			for (int k = 0; k < NP; ++k) {
				ia[k] = cia[l[k]];
				a[k] = tfa[q[k]];
			}
#pragma unroll
			for (int k = 0; k < NP; ++k) {
				ib[k] = cib[l[k]];
				b[k] = tfb[q[k]];
			}
			PairEval p[NP];
#pragma unroll
			for (int k = 0; k < NP; ++k) p[k] = eval(mid, l[k], q[k], ia[k], ib[k], a[k], b[k]);
			// the lanes whose pair is accepted and those whose pair goes to the
			// recheck, taken where the decisions are made: the branches and the
			// counts below use the masks (scalar registers)
			unsigned long long m_acc[NP], m_rc[NP];
#pragma unroll
			for (int k = 0; k < NP; ++k) {
				// lanes with a queued pair (n is uniform)
				const unsigned nk = n - min(n, (unsigned)(k * WAVE));
				const unsigned long long m_v = nk >= (unsigned)WAVE ? ~0ull : (1ull << nk) - 1ull;
				m_acc[k] = m_v & p[k].m_in & ~p[k].m_ns;
				m_rc[k] = m_v & p[k].m_ns;
			}
#pragma unroll
			for (int k = 0; k < NP; ++k) {
				if (__builtin_amdgcn_inverse_ballot_w64(m_acc[k])) {
					const float w = weight(p[k].l2, t_base + p[k].q, p[k].wp);
					add(p[k].bin, w, p[k].si, p[k].sj);
					if constexpr (FUSE) add2(p[k].bin2, w);
					// a twin: the same bin, its weight from its staged pair
					if constexpr (TWIN) add2(p[k].bin, J.t_w ? __int_as_float(p[k].si) * __int_as_float(p[k].sj) : p[k].wp);
				}
				n_acc_w += (unsigned)__popcll(m_acc[k]);
				if constexpr (FUSE || TWIN) n_acc2_w += (unsigned)__popcll(m_acc[k]);
			}
#pragma unroll
			for (int k = 0; k < NP; ++k) push(m_rc[k], p[k]);
		}
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_s_setprio(0);
		if (rq_n >= WAVE) flush_rq();
	};

	// append the lanes of mask m with their ring entries.  head and tail
	// start at 0 with every tile, and head moves by NP * 64 except at the
	// tile's end: every binning pass reads a window that starts at a multiple
	// of NP * 64 and does not wrap around the ring.
	static_assert(RING % (NP_MAX * WAVE) == 0, "binning windows do not wrap");
	unsigned head = 0, tail = 0;
	auto append = [&](unsigned long long m, unsigned entry) {
		if (m == 0ull) return;
		if (__builtin_amdgcn_inverse_ballot_w64(m)) {
			const unsigned off = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, tail));
			ring[off & (RING - 1)] = (unsigned short)entry;
		}
		tail += (unsigned)__popcll(m);
	};
	// matrix-core screening: a 16 x 16 block of pairs per MFMA, rows = 16
	// first-side objects (A: -2 o, 1), columns = 16 tile entries (B: o, |o|^2),
	// D = |b|^2 - 2 a.b.  Lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
	// and receives D[4 (l / 16) + r][l % 16] in register r.
	const int lrow = lane & 15, lk = lane >> 4;
	const unsigned lane_code = ((unsigned)lrow << 6) | (unsigned)(4 * lk);
	const int lane_d = lrow - 4 * lk;   // (column - row) within the block, for the order test
	const float *tff = (const float *)tfa;
	// z window slack: float offsets, and sort ties within one z bucket of the
	// gridding (h 2^-ZB_BITS) on both sides of a block's first and last entry
	const float zpad = (float)(J.h * 0x1p-8);

	const int own_cell_end = __builtin_amdgcn_readfirstlane(rspan[ROW_SPANS].x);
	for (int r = 0; r < J.n_rows; ++r) {
		if (!((W.row_mask >> r) & 1ull)) continue;
		const int jb = __builtin_amdgcn_readfirstlane(rspan[r].x);
		const int je = __builtin_amdgcn_readfirstlane(rspan[r].y);
		if (jb >= je) continue;
		const int4 row = rows[J.row_off + r];
		// A pair can pass the screening only if its z separation is at most
		// reach: the first-side cell and this row's cells are at least gx, gy
		// apart across x and y, and the screening passes s^2 < s2_hi + eps.
		const float gx = (float)(max(abs(row.x) - 1, 0) * J.h), gy = (float)(max(abs(row.y) - 1, 0) * J.h);
		const float reach = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(
		    __builtin_amdgcn_sqrtf(fmaxf(s2_hi + J.scr_eps - fmaf(gx, gx, gy * gy), 0.0f)) * (1.0f + 0x1p-10f) + zpad)));
		// only the row through the first-side cell of an auto count holds
		// pairs of the excluded half (cell-major order)
		const int own_end = is_auto && row.x == 0 && row.y == 0 ? own_cell_end : jb;
		for (int t0 = jb; t0 < je; t0 += TILE) {
			const int nt = min(TILE, je - t0);
			t_base = t0;
			if (tid < nt) {
				const int gj = t0 + tid;
				// every global load first, so that one latency covers them all
				const double GLOBAL_AS *px = gld(KIND == K_ANGULAR ? J.B.nx : J.B.x);
				const double GLOBAL_AS *py = gld(KIND == K_ANGULAR ? J.B.ny : J.B.y);
				const double GLOBAL_AS *pz = gld(KIND == K_ANGULAR ? J.B.nz : J.B.z);
				const double xg = px[gj], yg = py[gj], zg = pz[gj];
				const double nxg = gld(J.B.nx)[gj], nyg = gld(J.B.ny)[gj], nzg = gld(J.B.nz)[gj];
				const float wg = WEIGHTED ? (stage_iw ? gld(J.B.iw)[gj] : gld(J.B.w)[gj]) : 0.0f;
				// a twin's other weight takes the place of the region code
				const float wg2 = TWIN && J.t_w ? (stage_iw ? gld(J.B.w)[gj] : gld(J.B.iw)[gj]) : 0.0f;
				const int rj = jk ? gld(J.B.jk)[gj] : 0;
				const int sl = jk ? code_of(rj) : TWIN ? __float_as_int(wg2) : 0;
				const float ox = (float)(xg - c0), oy = (float)(yg - c1), oz = (float)(zg - c2);
				tfa[tid] = make_float4(ox, oy, oz, fmaf(oz, oz, fmaf(oy, oy, ox * ox)));
				// unweighted jobs keep the region code where the weight would be
				tfb[tid] = make_float4(nscale * (float)nxg, nscale * (float)nyg, nscale * (float)nzg,
				                       WEIGHTED ? wg : __int_as_float(sl));
				tslot[tid] = sl;
				// the row is sorted along z: each block's z range is its first
				// and last entry (up to sort ties within the slack)
				if ((tid & 15) == 0) tbz[tid >> 4] = oz;
				if ((tid & 15) == 15 || tid == nt - 1) tbe[tid >> 4] = oz;
			} else if (tid < ((nt + 15) & ~15)) {
				// the last 16-entry block is padded with entries that are
				// never in range (|b|^2 = 1e30)
				tfa[tid] = make_float4(0.0f, 0.0f, 0.0f, 1e30f);
			}
			__syncthreads();
			if (rb_n > 0) {
				const bool order = t0 < own_end;
				const int nblk = (nt + 15) >> 4;
				const unsigned long long valid = nblk >= 64 ? ~0ull : (1ull << nblk) - 1ull;
				// The (row block, tile block) units of this wave group, row
				// block by row block, are dealt to its jg_n waves in turn: unit
				// u goes to wave u mod jg_n.  A row block's window of kb..ke
				// then starts at this wave's next unit, and the waves differ by
				// one unit at most over the whole tile (instead of by one per
				// row block).
				int dealt = 0;
				// z span of tile block `lane` (lanes past the tile's blocks read
				// stale bounds, which the valid mask drops)
				const float tz0 = tbz[lane], tz1 = tbe[lane];
				for (int rb = 0; rb < rb_n; ++rb) {
					const int o = ig * WAVE + rb * 16;   // first object of the row block
					// The first-side objects are sorted along z too: the row
					// block spans cia[o].z .. cia[o + 15].z (entries past the
					// item repeat its last object).  Tile blocks wholly below
					// or above that span widened by reach hold no pair to
					// screen; screening runs from the first block not below to
					// the last block not above (tile block `lane` spans
					// tbz[lane] .. tbe[lane]).
					const float zlo = cia[o].z - reach, zhi = cia[o + 15].z + reach;
					const unsigned long long below = __builtin_amdgcn_ballot_w64(tz1 < zlo) & valid;
					const unsigned long long need = ~__builtin_amdgcn_ballot_w64(tz0 > zhi) & valid;
					if (need == 0ull || below == valid) continue;
					const int kb = __builtin_ctzll(~below), ke = 64 - __builtin_clzll(need);
					if (kb >= ke) continue;
					const int first = (jg - dealt) & (jg_n - 1);   // this wave's first unit (jg_n: 4, 8 or 16)
					dealt += ke - kb;
					// A[row][lk]: -2 o of the row's object for lk < 3, else 1 (one
					// load per lane, no branch on lk)
					const float ca = ((const float *)cia)[4 * (o + lrow) + min(lk, 2)];
					const float av = lk < 3 ? -2.0f * ca : 1.0f;
					const float4 th = *(const float4 *)&cth[o + 4 * lk];
					const float4 tl = *(const float4 *)&ctl[o + 4 * lk];
					// gj > gi  <=>  lane_d > i_begin + o + r - t0 - qb
					const int ord0 = W.i_begin + ((rb * ig_n + ig) << 4) - t0;
					// The next block's operand is loaded while this block's
					// product is formed and its masks appended (a product
					// started a block ahead as well, which needs its result
					// copied between registers every block, measured 1.7 %
					// slower).  The tile's last block is padded with entries
					// never in range; the load past it stays inside the tile
					// arrays (the tile is followed by its normals, TILE
					// entries of 16 B) and is not used.
					auto b_load = [&](int qb) { return tff[4 * (qb + lrow) + lk]; };
					const int step = jg_n * 16;
					auto product = [&](float bv) {
						return __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
					};
					// this wave's blocks of the window: kb + first, then every jg_n-th
					const int q0 = 16 * (kb + first), q1 = 16 * ke;
					float b_next = b_load(q0);
					for (int qb = q0; qb < q1; qb += step) {
						const f32x4 d = product(b_next);
						b_next = b_load(qb + step);
						const unsigned base = ((unsigned)qb << 6) + (unsigned)(rb * 16) + lane_code;
						unsigned long long m[4];
						m[0] = __builtin_amdgcn_ballot_w64(d[0] < th.x);
						m[1] = __builtin_amdgcn_ballot_w64(d[1] < th.y);
						m[2] = __builtin_amdgcn_ballot_w64(d[2] < th.z);
						m[3] = __builtin_amdgcn_ballot_w64(d[3] < th.w);
						if (check_lo) {
							m[0] &= __builtin_amdgcn_ballot_w64(d[0] >= tl.x);
							m[1] &= __builtin_amdgcn_ballot_w64(d[1] >= tl.y);
							m[2] &= __builtin_amdgcn_ballot_w64(d[2] >= tl.z);
							m[3] &= __builtin_amdgcn_ballot_w64(d[3] >= tl.w);
						}
						if (order) {
#pragma unroll
							for (int r = 0; r < 4; ++r) m[r] &= __builtin_amdgcn_ballot_w64(lane_d > ord0 + r - qb);
						}
#pragma unroll
						for (int r = 0; r < 4; ++r) append(m[r], base + (unsigned)r);
						while (tail - head >= NP * WAVE) {
							consume(head, NP * WAVE);
							head += NP * WAVE;
						}
					}
				}
				if (tail != head) consume(head, tail - head);
				head = tail = 0;
			}
			__syncthreads();
		}
	}
	if (rq_n) flush_rq();
	__syncthreads();   // every wave's rechecked pairs are in the LDS histograms

	for (int b = tid; b < (full_g ? 0 : nb); b += BLOCK) {
		HT v = (HT)0;
		// (only the copies that hold something are cleared)
		for (int r = 0; r < (1 << f_shift); ++r) {
			const HT x = F[(b << f_shift) + r];
			if (x != (HT)0) {
				v += x;
				F[(b << f_shift) + r] = (HT)0;
			}
		}
		if (v != (HT)0) {
			if (uf) {
				// every pair of the item: region R's touching count (and,
				// weighted, the full sums)
				const size_t o = (size_t)W.uf_region * nb + b;
				if (WEIGHTED) {
					gadd(&J.jk_d[o], (double)v);
					gadd(&J.full_d[b], (double)v);
				} else {
					gadd(&J.jk_u[o], (unsigned long long)v);
				}
			} else if (WEIGHTED) {
				gadd(&J.full_d[b], (double)v);
			} else {
				gadd(&J.full_u[b], (unsigned long long)v);
			}
		}
	}
	if (uf && cross_only) {
		// the slots hold the item's cross pairs, each once (its second
		// region): their sum per bin is the item's cross count
		for (int b = tid; b < nb; b += BLOCK) {
			unsigned long long x = 0;
			for (int s_ = 0; s_ < n_slots; ++s_) {
				const HT v = H[s_ * nb + b];
				if (v != (HT)0) {
					gadd(&J.jk_u[(size_t)sreg[s_] * nb + b], (unsigned long long)v);
					H[s_ * nb + b] = (HT)0;
					x += (unsigned long long)v;
				}
			}
			if (x) gadd(&J.full_u[b], x);
		}
	} else if (jk) {
		for (int e = tid; e < n_slots * nb; e += BLOCK) {
			HT v = H[e];
			if (v != (HT)0) {
				const int s_ = e / nb, b = e - s_ * nb;
				const size_t o = (size_t)sreg[s_] * nb + b;
				if (WEIGHTED) gadd(&J.jk_d[o], (double)v);
				else gadd(&J.jk_u[o], (unsigned long long)v);
				H[e] = (HT)0;
			}
		}
	}
	if constexpr (FUSE || TWIN) {
		HT *F2 = F + (nb << f_shift);
		const int mc = 1 << J.m_f_shift;
		for (int b = tid; b < J.m_nb; b += BLOCK) {
			HT v = (HT)0;
			for (int r = 0; r < mc; ++r) {
				v += F2[b * mc + r];
				F2[b * mc + r] = (HT)0;
			}
			if (v != (HT)0) {
				if (WEIGHTED) gadd(&J.m_full_d[b], (double)v);
				else gadd(&J.m_full_u[b], (unsigned long long)v);
			}
		}
	}
	for (int off = 32; off; off >>= 1) {
		n_acc += __shfl_xor(n_acc, off);
		n_rc += __shfl_xor(n_rc, off);
		if (FUSE || TWIN) n_acc2 += __shfl_xor(n_acc2, off);
	}
	n_acc += n_acc_w;
	n_acc2 += n_acc2_w;
	if (lane == 0) {
		if (n_acc) gadd(&J.counters[0], (unsigned long long)n_acc);
		if (n_rc) gadd(&J.counters[1], (unsigned long long)n_rc);
		if ((FUSE || TWIN) && n_acc2) gadd(&J.m_counters[0], (unsigned long long)n_acc2);
		if ((FUSE || TWIN) && n_rc) gadd(&J.m_counters[1], (unsigned long long)n_rc);
	}
	__syncthreads();
}

// One launch: item = blockIdx.x (plain) or pulled from the queues (persistent,
// queue != nullptr: NQ counters over ranges of the items).  Every block
// leaves when every range is drained, so the grid always finishes.
template <int KIND, bool FAST, bool WEIGHTED, bool MID = false, int FZ = FZ_NONE, int RG = 0>
__global__ void __launch_bounds__(BLOCK) k_count(const JobDev *__restrict__ jobs, const WorkItem *__restrict__ items,
                                                 int n_items, unsigned int *queue, const int4 *__restrict__ rows,
                                                 const int *__restrict__ slot_regions, const float *__restrict__ tables,
                                                 int n_tables, Lds L) {
	extern __shared__ __align__(16) char smem[];
	__shared__ int s_item;
	using HT = typename HistT<WEIGHTED>::T;
	// tables into LDS once per block, histograms cleared once
	float *tab = (float *)(smem + L.tables);
	for (int k = threadIdx.x; k < n_tables; k += BLOCK) tab[k] = tables[k];
	HT *hist = (HT *)(smem + L.hist);
	const int n_hist = ((L.masks ? L.masks : L.total) - L.hist) / (int)sizeof(HT);
	for (int k = threadIdx.x; k < n_hist; k += BLOCK) hist[k] = (HT)0;
	// Persistent launch: NQ queues over contiguous ranges of the items; a
	// block takes items from its own XCD's range first, so that the items an
	// XCD runs at one time are neighbours sharing their cells in its L2, and
	// then from the other ranges in turn.  Every block leaves when all are
	// drained.
	// Range q holds (n_items - q + nq - 1) / nq items (item_ranges()).
	const int nq = gridDim.x >= NQ && n_items >= NQ ? NQ : 1;
	const int q_base = n_items / nq, q_extra = n_items % nq;
	int tried = 0;   // (thread 0) ranges found drained
	int it = blockIdx.x;
	for (;;) {
		if (queue) {
			if (threadIdx.x == 0) {
				int got = n_items;
				while (tried < nq) {
					const int q = (int)((blockIdx.x + tried) % nq);
					const int lo = q * q_base + min(q, q_extra), hi = lo + q_base + (q < q_extra ? 1 : 0);
					const int k = lo + (int)atomicAdd(queue + q, 1u);
					if (k < hi) {
						got = k;
						break;
					}
					++tried;
				}
				s_item = got;
			}
			__syncthreads();
			it = __builtin_amdgcn_readfirstlane(s_item);
			__syncthreads();
		}
		if (it >= n_items) break;
		const WorkItem W = items[it];
		if constexpr (FAST) {
			// MID: every job of this launch has the pair-midpoint pi
			process_item_fast<KIND, WEIGHTED, MID, FZ, RG>(jobs[W.job], W, rows, slot_regions, smem, L);
		} else {
			process_item<KIND, WEIGHTED>(jobs[W.job], W, rows, slot_regions, smem, L);
		}
		if (!queue) break;
	}
}

// ------------------------------------------------------------ host side --

double now_ms() {
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

using KernelFn = void (*)(const JobDev *, const WorkItem *, int, unsigned *, const int4 *, const int *, const float *,
                          int, Lds);

// the fp32 kernel of one histogram layout; a fused monopole or a twin comes
// only with jobs without regions whose histograms fit in LDS
template <int KIND, bool WEIGHTED, bool MID, int FZ>
KernelFn fast_kernel_of(int rg) {
	if constexpr (FZ != FZ_NONE) return k_count<KIND, true, WEIGHTED, MID, FZ, RG_NONE>;
	if (rg == RG_NONE) return k_count<KIND, true, WEIGHTED, MID, FZ_NONE, RG_NONE>;
	if (rg == RG_JK) return k_count<KIND, true, WEIGHTED, MID, FZ_NONE, RG_JK>;
	return k_count<KIND, true, WEIGHTED, MID, FZ_NONE, RG_ANY>;
}

struct PlanLaunch {
	KernelFn fn;
	int grid;
	Lds L;
	const WorkItem *items;
	int n_items;
	unsigned *queue;
};

struct PlanOut {
	bool weighted;
	size_t full_off, jk_off, full_n, jk_n;
	// the full pool holds the cross-region count (unweighted jackknife jobs
	// of the FP32 kernel): full = sum over regions of touching - cross
	bool cross;
};

// Everything one twopcf_count call needs on the device, kept for replays.
struct Plan {
	std::vector<uint64_t> cat_ids;
	std::vector<void *> owned;
	JobDev *d_jobs = nullptr;
	int4 *d_rows = nullptr;
	int *d_slots = nullptr;
	float *d_tab = nullptr;
	unsigned long long *d_pool_u = nullptr, *d_cnt = nullptr;
	double *d_pool_d = nullptr;
	unsigned int *d_queue = nullptr;
	size_t pool_u = 0, pool_d = 0;
	int n_jobs = 0, n_tables = 0, n_items = 0;
	uint64_t n_staged = 0;   // objects read by the work items: first sides and their stencil rows' second sides
	double prep_ms = 0;
	std::vector<PlanLaunch> launches;
	bool concurrent = false;   // launches forked over the context's streams (batched method)
	std::vector<PlanOut> outs;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	void release() {
		if (exec) (void)hipGraphExecDestroy(exec);
		if (graph) (void)hipGraphDestroy(graph);
		exec = nullptr;
		graph = nullptr;
		for (void *p : owned) (void)hipFree(p);
		owned.clear();
	}
	~Plan() { release(); }
};

}  // namespace

// side streams of a context: with the context stream, as many hardware queues
// as a process gets by default
constexpr int N_SIDE = 3;

struct twopcf_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	// the batched method's launch classes run concurrently: forked from the
	// context stream onto the side streams and joined back
	hipStream_t side[N_SIDE] = {};
	hipEvent_t fork = nullptr, join[N_SIDE] = {};
	std::string err;
	int n_cu = 0;
	size_t lds_max = 65536;
	std::string name;
	std::vector<twopcf_cat *> cats;
	std::map<std::string, std::unique_ptr<Plan>> plans;
	std::map<std::string, twopcf_tuning> tuned;   // autotuned choice per job set
	uint64_t next_cat_id = 1;
	// temporaries of the gridding sort, kept between catalogues and grown
	// as needed (every use is on the context stream, in order)
	void *grid_tmp = nullptr;
	size_t grid_tmp_bytes = 0;
};

struct SortedCat {
	double *x = nullptr, *y = nullptr, *z = nullptr, *nx = nullptr, *ny = nullptr, *nz = nullptr;
	float *w = nullptr, *iw = nullptr;
	int *jk = nullptr;
	uint64_t *masks = nullptr;
	int m_stride = 0;
	int *start = nullptr;
	void *block = nullptr;   // one device allocation holding every array above
	std::vector<int> start_h;
	// regions present in each cell, CSR over cells (n_region > 1 only)
	std::vector<int> reg_off, reg_ids;
	void release() {
		if (block) (void)hipFree(block);
		*this = SortedCat();
	}
};

struct GridKey {
	double lo[3];
	double h;
	int dim[3];
	int n_region;
	int on_norm;   // cells over the unit vectors (angular counts) instead of positions
	bool operator<(const GridKey &o) const {
		return std::tie(lo[0], lo[1], lo[2], h, dim[0], dim[1], dim[2], n_region, on_norm) <
		       std::tie(o.lo[0], o.lo[1], o.lo[2], o.h, o.dim[0], o.dim[1], o.dim[2], o.n_region, o.on_norm);
	}
};

struct twopcf_cat {
	twopcf_ctx *ctx = nullptr;
	uint64_t id = 0;
	int64_t n = 0;
	int n_words = 0;
	int m_stride = 0;                 // mask words per object: bit mask, then the dither mask if any
	double *x = nullptr, *y = nullptr, *z = nullptr, *nx = nullptr, *ny = nullptr, *nz = nullptr;
	float *w = nullptr, *iw = nullptr;
	int *jk = nullptr;
	uint64_t *masks = nullptr;
	double lo[3], hi[3], nlo[3], nhi[3];
	bool unit_w = true, has_iw = false, has_dither = false;
	int jk_min = 0, jk_max = 0;
	std::map<GridKey, SortedCat> grids;
};

namespace {

int fail(twopcf_ctx *ctx, int code, const std::string &msg) {
	if (ctx) ctx->err = msg;
	return code;
}

#define HIPCK(ctx, call)                                                                           \
	do {                                                                                           \
		hipError_t e_ = (call);                                                                    \
		if (e_ != hipSuccess)                                                                      \
			return fail((ctx), TWOPCF_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
	} while (0)

template <class T>
int dalloc(twopcf_ctx *ctx, T **p, size_t n) {
	HIPCK(ctx, hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
	return 0;
}

// Arrays laid out one after another in one device allocation, each at a
// 256-byte boundary: add() them, then place() them at the base address.
struct Carve {
	size_t at = 0;
	std::vector<std::pair<void **, size_t>> parts;
	template <class T>
	void add(T **p, size_t count) {
		parts.emplace_back((void **)p, at);
		at += (std::max<size_t>(count, 1) * sizeof(T) + 255) & ~size_t(255);
	}
	void place(void *base) {
		for (auto &pt : parts) *pt.first = (char *)base + pt.second;
	}
};

// Sort one catalogue into the cell geometry g (cached per geometry).  With
// may_alloc false only a cached grid is taken (twopcf_count_noalloc).
int grid_catalog(twopcf_ctx *ctx, twopcf_cat *c, const GridKey &g, SortedCat **out, bool may_alloc = true) {
	auto it = c->grids.find(g);
	if (it != c->grids.end()) {
		*out = &it->second;
		return 0;
	}
	if (!may_alloc)
		return fail(ctx, TWOPCF_E_ARG,
		            "catalogue not yet gridded for this job set's cell geometry (twopcf_count_noalloc allocates "
		            "nothing): call twopcf_count once with the same jobs and tuning first");
	SortedCat s;
	const int64_t n = c->n;
	const int n_cell = g.dim[0] * g.dim[1] * g.dim[2];
	const int64_t n_scan = (int64_t)n_cell + 1;
	const int n_tiles = (int)((n_scan + SORT_BLOCK - 1) / SORT_BLOCK);
	// Arrays are carved out of one allocation each (the sorted catalogue,
	// and the sort's temporaries, kept in the context): one hipMalloc per
	// catalogue instead of one per array.
	int *cell = nullptr, *tperm = nullptr, *perm = nullptr;
	unsigned short *zb = nullptr, *tzb = nullptr;
	unsigned *count = nullptr, *sums = nullptr;
	Carve tc;
	tc.add(&cell, n);
	tc.add(&zb, n);
	tc.add(&tperm, n);
	tc.add(&tzb, n);
	tc.add(&perm, n);
	tc.add(&count, n_scan);
	tc.add(&sums, n_tiles);
	if (tc.at > ctx->grid_tmp_bytes) {
		if (ctx->grid_tmp) (void)hipFree(ctx->grid_tmp);
		ctx->grid_tmp = nullptr;
		ctx->grid_tmp_bytes = 0;
		HIPCK(ctx, hipMalloc(&ctx->grid_tmp, tc.at));
		ctx->grid_tmp_bytes = tc.at;
	}
	tc.place(ctx->grid_tmp);
	Carve sc;
	sc.add(&s.x, n);
	sc.add(&s.y, n);
	sc.add(&s.z, n);
	sc.add(&s.nx, n);
	sc.add(&s.ny, n);
	sc.add(&s.nz, n);
	sc.add(&s.w, n);
	sc.add(&s.iw, n);
	sc.add(&s.jk, n);
	sc.add(&s.masks, (size_t)n * c->m_stride);
	sc.add(&s.start, (size_t)n_cell + 1);
	HIPCK(ctx, hipMalloc(&s.block, sc.at));
	sc.place(s.block);
	// the block is freed on every error return
	struct Own {
		SortedCat *s;
		bool keep = false;
		~Own() {
			if (!keep) s->release();
		}
	} own{&s};
	int rc = 0;
	s.m_stride = c->m_stride;
	// counting sort by cell (count, scan into the cell starts, scatter), then
	// by z bucket and catalogue index within each cell
	unsigned *start = (unsigned *)s.start;
	const int nblk = (int)((n + 255) / 256);
	auto check = [&]() -> int {
		hipError_t e = hipGetLastError();
		if (e != hipSuccess) return fail(ctx, TWOPCF_E_HIP, std::string("gridding: ") + hipGetErrorString(e));
		return 0;
	};
	k_fill_u32<<<(unsigned)((n_scan + 255) / 256), 256, 0, ctx->stream>>>(count, n_scan, 0u);
	if (n > 0)
		k_cell_bucket<<<nblk, 256, 0, ctx->stream>>>(g.on_norm ? c->nx : c->x, g.on_norm ? c->ny : c->y,
		                                             g.on_norm ? c->nz : c->z, n, g.lo[0], g.lo[1], g.lo[2], g.h,
		                                             g.dim[0], g.dim[1], g.dim[2], cell, zb, count);
	k_scan_tiles<<<n_tiles, SORT_BLOCK, 0, ctx->stream>>>(count, start, n_scan, sums);
	k_scan_sums<<<1, SORT_BLOCK, 0, ctx->stream>>>(sums, n_tiles);
	k_scan_add<<<n_tiles, SORT_BLOCK, 0, ctx->stream>>>(start, n_scan, sums);
	if ((rc = check())) return rc;
	if (n > 0) {
		k_fill_u32<<<(unsigned)((n_scan + 255) / 256), 256, 0, ctx->stream>>>(count, n_scan, 0u);   // cursors
		k_scatter_cells<<<nblk, 256, 0, ctx->stream>>>(cell, zb, n, start, count, tperm, tzb);
		k_zsort_cells<<<std::max(1, std::min(n_cell, ctx->n_cu * 8)), SORT_BLOCK, 0, ctx->stream>>>(start, n_cell, tperm,
		                                                                                           tzb, perm);
		k_gather<<<nblk, 256, 0, ctx->stream>>>(perm, n, c->x, c->y, c->z, c->nx, c->ny, c->nz, c->w, c->iw, c->jk,
		                                        c->masks, c->m_stride, s.x, s.y, s.z, s.nx, s.ny, s.nz, s.w, s.iw,
		                                        s.jk, s.masks);
		if ((rc = check())) return rc;
	}
	s.start_h.resize((size_t)n_cell + 1);
	HIPCK(ctx, hipMemcpyAsync(s.start_h.data(), s.start, ((size_t)n_cell + 1) * sizeof(int), hipMemcpyDeviceToHost,
	                          ctx->stream));
	std::vector<int> jk_h;
	if (g.n_region > 1 && n > 0) {
		jk_h.resize(n);
		HIPCK(ctx, hipMemcpyAsync(jk_h.data(), s.jk, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	}
	HIPCK(ctx, hipStreamSynchronize(ctx->stream));
	if (g.n_region > 1) {
		s.reg_off.assign((size_t)n_cell + 1, 0);
		std::vector<int> seen((size_t)g.n_region, -1);   // last cell that listed the region
		for (int cc = 0; cc < n_cell; ++cc) {
			s.reg_off[cc] = (int)s.reg_ids.size();
			for (int i = s.start_h[cc]; i < s.start_h[cc + 1]; ++i)
				if (seen[jk_h[i]] != cc) {
					seen[jk_h[i]] = cc;
					s.reg_ids.push_back(jk_h[i]);
				}
		}
		s.reg_off[n_cell] = (int)s.reg_ids.size();
	}
	auto res = c->grids.emplace(g, s);
	own.keep = true;
	*out = &res.first->second;
	return 0;
}

// Stencil rows (dx, dy, dz_lo, dz_hi) of neighbour cells that can hold a
// pair within rmax when cells have side rmax/k: cells whose closest points
// are at least rmax apart are left out.
std::vector<int4> stencil_rows(int k) {
	std::vector<int4> rows;
	auto gap = [](int d) { return std::max(0, std::abs(d) - 1); };
	for (int dx = -k; dx <= k; ++dx)
		for (int dy = -k; dy <= k; ++dy) {
			int g = gap(dx) * gap(dx) + gap(dy) * gap(dy);
			if (g >= k * k) continue;
			int m = 0;
			for (int dz = 0; dz <= k; ++dz)
				if (g + gap(dz) * gap(dz) < k * k) m = dz;
			rows.push_back(make_int4(dx, dy, -m, m));
		}
	return rows;
}

struct Prepared {
	JobDev jd;
	SortedCat *A, *B;
	bool swapped;
	int nb;
	size_t full_n, jk_n;
	size_t full_off, jk_off;   // offsets into the device output pools
	bool weighted;
	int f_elems = 0;           // LDS elements of the full histogram (lane copies included)
	int mask_bytes = 0;        // LDS copy of the first side's masks (popcount rules, fp32 path; 0: none)
};

// Error model of the FP32 path.  Both objects are held as float offsets
// from the centre of the first object's cell, each coordinate within R of
// it, so each offset carries a rounding error of at most 2^-24 R.  The FP32
// difference, the multiply-adds of s^2 and the reference's own rounding of
// d = (float)(x1 - x2) and of s^2 = (d0^2 + d1^2) + d2^2 then keep the FP32
// s^2 within E(S) of the reference's s^2 at true squared separation S.
inline double e_s2(double S, double R) { return std::ldexp(1.0, -23) * (11.0 * S + 4.0 * R * std::sqrt(S)); }

float round_up(double v) {
	float f = (float)v;
	if ((double)f < v) f = std::nextafter(f, INFINITY);
	return f;
}
float round_dn(double v) {
	float f = (float)v;
	if ((double)f > v) f = std::nextafter(f, -INFINITY);
	return f;
}

}  // namespace

// ================================================================== ABI ==

namespace {

// Runs one entry point's body: a C++ exception (host memory exhausted, a
// thread that could not start) becomes an error code and message, so that
// none crosses the C boundary.
template <class F>
int guarded(twopcf_ctx *ctx, F &&body) {
	try {
		return body();
	} catch (const std::bad_alloc &) {
		return fail(ctx, TWOPCF_E_LIMIT, "host memory exhausted");
	} catch (const std::exception &e) {
		return fail(ctx, TWOPCF_E_LIMIT, std::string("host error: ") + e.what());
	} catch (...) {
		return fail(ctx, TWOPCF_E_LIMIT, "host error");
	}
}

}  // namespace

extern "C" {

int twopcf_abi_version(void) { return TWOPCF_ABI_VERSION; }

int twopcf_ctx_create(int device, twopcf_ctx **out) {
	return guarded(nullptr, [&]() -> int {
		if (!out) return TWOPCF_E_ARG;
		*out = nullptr;
		int n = 0;
		if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return TWOPCF_E_NODEV;
		if (device < 0 || device >= n) return TWOPCF_E_ARG;
		auto *ctx = new twopcf_ctx();
		ctx->device = device;
		hipDeviceProp_t prop;
		bool ok = hipSetDevice(device) == hipSuccess && hipGetDeviceProperties(&prop, device) == hipSuccess &&
		          hipStreamCreate(&ctx->stream) == hipSuccess && hipEventCreate(&ctx->ev0) == hipSuccess &&
		          hipEventCreate(&ctx->ev1) == hipSuccess &&
		          hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) == hipSuccess;
		for (int k = 0; k < N_SIDE && ok; ++k)
			ok = hipStreamCreate(&ctx->side[k]) == hipSuccess &&
			     hipEventCreateWithFlags(&ctx->join[k], hipEventDisableTiming) == hipSuccess;
		if (!ok) {
			twopcf_ctx_destroy(ctx);
			return TWOPCF_E_HIP;
		}
		ctx->n_cu = prop.multiProcessorCount;
		ctx->lds_max = std::max(prop.sharedMemPerBlock, prop.sharedMemPerBlockOptin);
		ctx->name = prop.gcnArchName;
		// the no-allocation entry point's scratch on this device (device current)
		if (resolve_scratch(device) != hipSuccess) {
			twopcf_ctx_destroy(ctx);
			return TWOPCF_E_HIP;
		}
		*out = ctx;
		return 0;
	});
}

int twopcf_ctx_destroy(twopcf_ctx *ctx) {
	if (!ctx) return TWOPCF_E_ARG;
	(void)hipSetDevice(ctx->device);
	ctx->plans.clear();
	while (!ctx->cats.empty()) twopcf_catalog_free(ctx->cats.back());
	if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
	if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
	if (ctx->fork) (void)hipEventDestroy(ctx->fork);
	for (int k = 0; k < N_SIDE; ++k) {
		if (ctx->join[k]) (void)hipEventDestroy(ctx->join[k]);
		if (ctx->side[k]) (void)hipStreamDestroy(ctx->side[k]);
	}
	if (ctx->grid_tmp) (void)hipFree(ctx->grid_tmp);
	if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
	delete ctx;
	return 0;
}

const char *twopcf_last_error(const twopcf_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int twopcf_device_info(twopcf_ctx *ctx, char *name, int name_len, int *n_cu, int *lds_per_block) {
	if (!ctx) return TWOPCF_E_ARG;
	if (name && name_len > 0) {
		std::strncpy(name, ctx->name.c_str(), name_len - 1);
		name[name_len - 1] = 0;
	}
	if (n_cu) *n_cu = ctx->n_cu;
	if (lds_per_block) *lds_per_block = (int)ctx->lds_max;
	return 0;
}

int twopcf_catalog_upload(twopcf_ctx *ctx, const twopcf_cat_desc *d, twopcf_cat **out) {
	return guarded(ctx, [&]() -> int {
		if (!ctx || !d || !out) return fail(ctx, TWOPCF_E_ARG, "null argument");
		*out = nullptr;
		if (d->n < 0 || d->n > INT32_MAX - 1) return fail(ctx, TWOPCF_E_LIMIT, "catalogue size out of range");
		if (d->n > 0 && (!d->x || !d->y || !d->z || !d->nx || !d->ny || !d->nz))
			return fail(ctx, TWOPCF_E_ARG, "positions and unit vectors are required");
		if (d->n_words < 0 || d->n_words > 64 || (d->n_words > 0 && !d->masks) || (d->dither && d->n_words < 1))
			return fail(ctx, TWOPCF_E_ARG, "bad mask description");
		HIPCK(ctx, hipSetDevice(ctx->device));
		auto *c = new twopcf_cat();
		c->ctx = ctx;
		c->id = ctx->next_cat_id++;
		c->n = d->n;
		c->n_words = d->n_words;
		c->has_dither = d->dither != nullptr;
		c->m_stride = d->n_words * (c->has_dither ? 2 : 1);
		const int64_t n = d->n;
		// masks on the device: per object the bit mask words, then the dither mask words
		std::vector<uint64_t> mk;
		if (c->has_dither) {
			mk.resize((size_t)n * c->m_stride);
			for (int64_t i = 0; i < n; ++i) {
				std::copy(d->masks + i * d->n_words, d->masks + (i + 1) * d->n_words, mk.begin() + i * c->m_stride);
				std::copy(d->dither + i * d->n_words, d->dither + (i + 1) * d->n_words,
				          mk.begin() + i * c->m_stride + d->n_words);
			}
		}
		std::vector<float> w(n, 1.0f), iw;
		std::vector<int> jk(n, 0);
		for (int k = 0; k < 3; ++k) {
			c->lo[k] = c->nlo[k] = INFINITY;
			c->hi[k] = c->nhi[k] = -INFINITY;
		}
		const double *P[3] = {d->x, d->y, d->z}, *N[3] = {d->nx, d->ny, d->nz};
		for (int64_t i = 0; i < n; ++i)
			for (int k = 0; k < 3; ++k) {
				if (!std::isfinite(P[k][i])) {
					delete c;
					return fail(ctx, TWOPCF_E_ARG, "non-finite position");
				}
				c->lo[k] = std::min(c->lo[k], P[k][i]);
				c->hi[k] = std::max(c->hi[k], P[k][i]);
				c->nlo[k] = std::min(c->nlo[k], N[k][i]);
				c->nhi[k] = std::max(c->nhi[k], N[k][i]);
			}
		if (d->w) std::copy(d->w, d->w + n, w.begin());
		for (int64_t i = 0; i < n; ++i) c->unit_w &= (w[i] == 1.0f);
		if (d->jk) std::copy(d->jk, d->jk + n, jk.begin());
		if (n > 0) {
			c->jk_min = *std::min_element(jk.begin(), jk.end());
			c->jk_max = *std::max_element(jk.begin(), jk.end());
		}
		c->has_iw = d->iw != nullptr;
		iw.assign(d->iw ? d->iw : w.data(), (d->iw ? d->iw : w.data()) + n);
		int rc = 0;
		if ((rc = dalloc(ctx, &c->x, n)) || (rc = dalloc(ctx, &c->y, n)) || (rc = dalloc(ctx, &c->z, n)) ||
		    (rc = dalloc(ctx, &c->nx, n)) || (rc = dalloc(ctx, &c->ny, n)) || (rc = dalloc(ctx, &c->nz, n)) ||
		    (rc = dalloc(ctx, &c->w, n)) || (rc = dalloc(ctx, &c->iw, n)) || (rc = dalloc(ctx, &c->jk, n)) ||
		    (rc = dalloc(ctx, &c->masks, (size_t)n * c->m_stride))) {
			ctx->cats.push_back(c);
			twopcf_catalog_free(c);
			return rc;
		}
		ctx->cats.push_back(c);
		auto up = [&](void *dst, const void *src, size_t bytes) -> int {
			if (bytes) HIPCK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
			return 0;
		};
		if ((rc = up(c->x, d->x, n * 8)) || (rc = up(c->y, d->y, n * 8)) || (rc = up(c->z, d->z, n * 8)) ||
		    (rc = up(c->nx, d->nx, n * 8)) || (rc = up(c->ny, d->ny, n * 8)) || (rc = up(c->nz, d->nz, n * 8)) ||
		    (rc = up(c->w, w.data(), n * 4)) || (rc = up(c->iw, iw.data(), n * 4)) || (rc = up(c->jk, jk.data(), n * 4)) ||
		    (rc = up(c->masks, c->has_dither ? (const void *)mk.data() : (const void *)d->masks,
		             (size_t)n * c->m_stride * 8))) {
			twopcf_catalog_free(c);
			return rc;
		}
		HIPCK(ctx, hipStreamSynchronize(ctx->stream));
		*out = c;
		return 0;
	});
}

int twopcf_catalog_free(twopcf_cat *c) {
	if (!c) return TWOPCF_E_ARG;
	twopcf_ctx *ctx = c->ctx;
	if (ctx) {
		for (auto it = ctx->plans.begin(); it != ctx->plans.end();) {
			const auto &ids = it->second->cat_ids;
			if (std::find(ids.begin(), ids.end(), c->id) != ids.end()) it = ctx->plans.erase(it);
			else ++it;
		}
	}
	for (auto &kv : c->grids) kv.second.release();
	for (void *p : {(void *)c->x, (void *)c->y, (void *)c->z, (void *)c->nx, (void *)c->ny, (void *)c->nz,
	                (void *)c->w, (void *)c->iw, (void *)c->jk, (void *)c->masks})
		if (p) (void)hipFree(p);
	if (ctx) ctx->cats.erase(std::remove(ctx->cats.begin(), ctx->cats.end(), c), ctx->cats.end());
	delete c;
	return 0;
}

int64_t twopcf_catalog_size(const twopcf_cat *c) { return c ? c->n : -1; }

}  // extern "C"

namespace {

// Build the device-side plan of one twopcf_count call: grids, tables, work
// items, output pools and the launch list.  Cached per context.
// sample_stride > 1 keeps every sample_stride-th item of the ordered
// queue (autotuning trials; the outputs are then partial).
// arena non-null: every device buffer of the plan comes from it (the
// scratch array of twopcf_count_noalloc) and no catalogue is gridded anew.
int build_plan(twopcf_ctx *ctx, const twopcf_job *jobs, int n_jobs, int method, const twopcf_tuning *tuning,
               Plan &plan, int sample_stride = 1, Arena *arena = nullptr) {
	const bool fast = method == TWOPCF_FP32_RECHECK;
	// device buffer of the plan: from the arena, or hipMalloc owned by the plan
	auto dev_alloc = [&](void **p, size_t bytes) -> int {
		bytes = std::max<size_t>(bytes, 16);
		if (arena) {
			const size_t at = (arena->used + 255) & ~size_t(255);
			if (at > arena->cap || bytes > arena->cap - at)
				return fail(ctx, TWOPCF_E_LIMIT,
				            "job set needs more than the " + std::to_string(arena->cap >> 20) +
				                " MiB scratch of twopcf_count_noalloc");
			*p = arena->base + at;
			arena->used = at + bytes;
			return 0;
		}
		hipError_t e = hipMalloc(p, bytes);
		if (e != hipSuccess) return fail(ctx, TWOPCF_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
		plan.owned.push_back(*p);
		return 0;
	};
	int subdiv = fast ? 2 : 1;
	// region slots: 12 requested on the batched path (as many as fit in LDS)
	int max_slots = fast ? 12 : 8;
	int i_chunk = ICHUNK;
	if (tuning) {
		if (fast && tuning->cell_subdiv > 0) subdiv = std::min(tuning->cell_subdiv, MAX_SUBDIV);
		if (tuning->max_slots > 0) max_slots = std::min(tuning->max_slots, MAX_SLOTS);
		if (fast && tuning->i_chunk > 0) i_chunk = std::max(WAVE, std::min(ICHUNK, tuning->i_chunk / WAVE * WAVE));
	}

	// ---- validate and build per-job host descriptions ----
	struct HostJob {
		const twopcf_job *j;
		twopcf_cat *a, *b;   // after the optional swap (a = lanes)
		bool swapped;
		AxisF ax1, ax2;
		double rmax;
		std::vector<float> t_s2, t1, t2, tA, tB, up_thr;
		std::vector<double> up_bins, up_centre;
		AxisF up_ax;
		float min_r_sq, max_r_sq;
		int nb;
		bool weighted;
		GridKey g;
	};
	std::vector<HostJob> hj(n_jobs);
	// The float thresholds depend only on the kind and the axes, and each
	// search checks a window of floats around every threshold (milliseconds
	// per axis): jobs with the same kind and axes take the first one's.
	std::map<std::string, int> axes_first;
	for (int q = 0; q < n_jobs; ++q) {
		const twopcf_job &J = jobs[q];
		HostJob &H = hj[q];
		H.j = &J;
		if (!J.a || !J.full) return fail(ctx, TWOPCF_E_ARG, "job needs a catalogue and a full[] output");
		if (J.kind < TWOPCF_MONOPOLE || J.kind > TWOPCF_ANGULAR) return fail(ctx, TWOPCF_E_ARG, "bad correlator kind");
		if (J.weight_mode < TWOPCF_W_PLAIN || J.weight_mode > TWOPCF_W_ANG_W)
			return fail(ctx, TWOPCF_E_ARG, "bad weight mode");
		if (J.pi_def != TWOPCF_PI_ANGULAR && J.pi_def != TWOPCF_PI_MIDPOINT)
			return fail(ctx, TWOPCF_E_ARG, "bad line-of-sight definition");
		const bool needs_masks = J.weight_mode == TWOPCF_W_INVP_3D || J.weight_mode == TWOPCF_W_ANG_INVP;
		const bool needs_up = J.weight_mode == TWOPCF_W_INVP_3D || J.weight_mode == TWOPCF_W_ANGUP_3D;
		if (needs_masks) {
			const twopcf_cat *bb = J.b ? J.b : J.a;
			if (J.a->n_words < 1 || J.a->n_words != bb->n_words)
				return fail(ctx, TWOPCF_E_ARG, "inverse-p weights need bit masks of equal width in both catalogues");
			if (J.n_runs <= 0) return fail(ctx, TWOPCF_E_ARG, "n_runs must be positive");
			if (J.dither && (!J.a->has_dither || !bb->has_dither))
				return fail(ctx, TWOPCF_E_ARG, "dither weights need dither masks in both catalogues");
		} else if (J.dither) {
			return fail(ctx, TWOPCF_E_ARG, "dither weights apply to the inverse-p rules only");
		}
		if (needs_up) {
			if (!J.up_bins || J.up_axis.n_bins < 2 || J.up_axis.n_bins > MAX_AXIS_BINS)
				return fail(ctx, TWOPCF_E_ARG, "angular up-weight table missing or too short");
			H.up_ax = AxisF((float)J.up_axis.min, (float)J.up_axis.max, J.up_axis.n_bins, (float)J.up_axis.log_base);
			if (!interp_thresholds(H.up_ax, H.up_thr)) return fail(ctx, TWOPCF_E_AXIS, "up-weight axis not monotone");
			H.up_bins.assign(J.up_bins, J.up_bins + J.up_axis.n_bins);
			for (int k = 0; k < H.up_ax.n; ++k) H.up_centre.push_back(H.up_ax.centre(k));
		}
		if (J.n_jk < 1) return fail(ctx, TWOPCF_E_ARG, "n_jk must be >= 1");
		twopcf_cat *a = const_cast<twopcf_cat *>(J.a), *b = const_cast<twopcf_cat *>(J.b);
		for (twopcf_cat *c : {a, b}) {
			if (!c) continue;
			if (c->ctx != ctx) return fail(ctx, TWOPCF_E_ARG, "catalogue belongs to another context");
			if (J.n_jk > 1 && c->n > 0 && (c->jk_min < 0 || c->jk_max >= J.n_jk))
				return fail(ctx, TWOPCF_E_ARG, "jackknife region out of [0, n_jk)");
		}
		const bool two_axes = J.kind == TWOPCF_SIGMA_PI || J.kind == TWOPCF_S_MU;
		for (int k = 0; k < (two_axes ? 2 : 1); ++k)
			if (J.ax[k].n_bins < 1 || J.ax[k].n_bins > MAX_AXIS_BINS)
				return fail(ctx, TWOPCF_E_LIMIT, "axis bins out of [1, 4096]");
		H.ax1 = AxisF((float)J.ax[0].min, (float)J.ax[0].max, J.ax[0].n_bins, (float)J.ax[0].log_base);
		if (J.kind == TWOPCF_S_MU) H.ax2 = AxisF(0.0f, 1.0f, J.ax[1].n_bins, 0.0f);
		else if (J.kind == TWOPCF_SIGMA_PI)
			H.ax2 = AxisF((float)J.ax[1].min, (float)J.ax[1].max, J.ax[1].n_bins, (float)J.ax[1].log_base);
		else H.ax2 = AxisF(0.0f, 1.0f, 1, 0.0f);
		H.nb = H.ax1.n * (two_axes ? H.ax2.n : 1);
		range_sq(J.kind, H.ax1, H.ax2, &H.min_r_sq, &H.max_r_sq);
		std::string akey((const char *)&J.kind, sizeof J.kind);
		for (const twopcf_axis &x : J.ax) {
			akey.append((const char *)&x.min, sizeof x.min);
			akey.append((const char *)&x.max, sizeof x.max);
			akey.append((const char *)&x.n_bins, sizeof x.n_bins);
			akey.append((const char *)&x.log_base, sizeof x.log_base);
		}
		const auto seen = axes_first.find(akey);
		const HostJob *same = seen != axes_first.end() ? &hj[seen->second] : nullptr;
		if (!same) axes_first.emplace(akey, q);
		if (J.kind == TWOPCF_ANGULAR) {
			// cells over the unit vectors, side = the largest chord whose
			// 1 - cos can still land in the histogram
			if (same) H.t1 = same->t1;
			else if (!value_thresholds(H.ax1, H.t1)) return fail(ctx, TWOPCF_E_AXIS, "axis index not monotone");
			const double vmax = std::max(0.0, (double)H.t1[H.ax1.n]);
			H.rmax = std::sqrt(2.0 * std::min(vmax, 2.0) * (1.0 + 1e-6) + 1e-14);
			H.max_r_sq = (float)(H.rmax * H.rmax);
		} else if (J.kind == TWOPCF_SIGMA_PI) {
			H.rmax = std::sqrt(J.ax[0].max * J.ax[0].max + J.ax[1].max * J.ax[1].max);
			if (same) {
				H.t1 = same->t1;
				H.t2 = same->t2;
			} else if (!value_thresholds(H.ax1, H.t1) || !value_thresholds(H.ax2, H.t2)) {
				return fail(ctx, TWOPCF_E_AXIS, "axis index not monotone");
			}
		} else {
			H.rmax = J.ax[0].max;
			if (same) H.t_s2 = same->t_s2;
			else if (!s2_thresholds(J.kind, H.ax1, H.ax2, H.t_s2)) return fail(ctx, TWOPCF_E_AXIS, "axis index not monotone");
		}
		if (!(H.rmax > 0) || !std::isfinite(H.rmax)) return fail(ctx, TWOPCF_E_ARG, "rmax must be positive");
		// The range test is on float s^2; the grid must cover every float s^2
		// below max_r_sq.
		H.rmax = std::max(H.rmax, std::sqrt((double)H.max_r_sq)) * (1.0 + 1e-6);
		H.weighted = !(a->unit_w && (!b || b->unit_w)) || J.weight_mode != TWOPCF_W_PLAIN;
		// A cross count is symmetric: in the batched method the smaller
		// catalogue takes the first side (the lanes), so that the larger one
		// fills the staged tiles (fewer, fuller tiles per pair than the other
		// way round); the plain method, one block per item, is faster with the
		// larger one first.
		H.swapped = b && (fast ? b->n < a->n : b->n > a->n);
		H.a = H.swapped ? b : a;
		H.b = H.swapped ? a : b;
	}

	// ---- shared cell geometry per rmax over all catalogues of the call ----
	const double t_prep0 = now_ms();
	std::vector<int4> row_pool;
	std::map<double, std::pair<int, int>> rows_of;   // rmax -> (row_off, n_rows)
	{
		auto rows = stencil_rows(subdiv);
		if ((int)rows.size() > MAX_ROWS) return fail(ctx, TWOPCF_E_LIMIT, "stencil too large");
	}
	for (auto &H : hj) {
		double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
		for (auto &G : hj) {
			if (G.rmax != H.rmax) continue;
			if ((G.j->kind == TWOPCF_ANGULAR) != (H.j->kind == TWOPCF_ANGULAR)) continue;
			for (twopcf_cat *c : {G.a, G.b}) {
				if (!c || c->n == 0) continue;
				const bool ang = H.j->kind == TWOPCF_ANGULAR;
				for (int k = 0; k < 3; ++k) {
					lo[k] = std::min(lo[k], ang ? c->nlo[k] : c->lo[k]);
					hi[k] = std::max(hi[k], ang ? c->nhi[k] : c->hi[k]);
				}
			}
		}
		for (int k = 0; k < 3; ++k)
			if (!(lo[k] <= hi[k])) lo[k] = hi[k] = 0.0;
		H.g.h = H.rmax / subdiv;
		long long ncell = 1;
		for (int k = 0; k < 3; ++k) {
			H.g.lo[k] = lo[k];
			long long dk = (long long)((hi[k] - lo[k]) / H.g.h) + 1;
			if (dk > (1 << 20)) return fail(ctx, TWOPCF_E_LIMIT, "grid too fine");
			H.g.dim[k] = (int)dk;
			ncell *= dk;
		}
		if (ncell > (1LL << 28)) return fail(ctx, TWOPCF_E_LIMIT, "too many cells");
		H.g.n_region = H.j->n_jk > 1 ? H.j->n_jk : 1;
		H.g.on_norm = H.j->kind == TWOPCF_ANGULAR ? 1 : 0;
		if (!rows_of.count(H.rmax)) {
			auto rows = stencil_rows(subdiv);
			rows_of[H.rmax] = {(int)row_pool.size(), (int)rows.size()};
			row_pool.insert(row_pool.end(), rows.begin(), rows.end());
		}
	}

	// ---- grids, tables, outputs, work items ----
	std::vector<Prepared> prep(n_jobs);
	std::vector<float> tables;
	std::map<std::string, int> table_at;
	std::vector<int> slot_pool;
	std::vector<WorkItem> items;
	std::vector<double> item_cost;
	uint64_t n_staged = 0;
	size_t pool_u = 0, pool_d = 0;
	int max_tab_n1 = 0;
	// table sizes are known up front: every job pushes at most 3 * (n1 + 1)
	// + (n2 + 1) floats
	size_t tab_floats = 1;
	{
		// per distinct (kind, axes): at most 7 * (n1 + 1) + (n2 + 1) floats
		// plus alignment; jobs on the same axes share their tables
		std::vector<std::string> seen;
		for (auto &H : hj) {
			std::string k((const char *)&H.j->kind, sizeof H.j->kind);
			k.append((const char *)H.j->ax, sizeof H.j->ax);
			if (std::find(seen.begin(), seen.end(), k) != seen.end()) continue;
			seen.push_back(k);
			tab_floats += 9 * (H.ax1.n + 1) + (H.ax2.n + 1) + 32;
		}
	}
	// LDS besides the histograms: the method's tile (with the fp32 path's
	// first-side chunk and pair queues), the tables and the slot list, as
	// the launch layout places them
	const int lds_fixed = (int)(((tile_bytes(fast) + 15) & ~15) + ((tab_floats * 4 + 15) & ~15) +
	                            ((MAX_SLOTS * 4 + 15) & ~15));
	const size_t lds_budget = ctx->lds_max;
	// The popcount rules of weighted angular and monopole jobs (small
	// histograms) read the first side's masks from an LDS copy when it fits
	// beside one region slot: rows of an odd number of 16-byte units, so
	// that lanes reading different objects spread over the banks.  These
	// jobs form their own launch class; they keep at most 16 lane copies of
	// the full histogram and leave room for the largest copy of their kind
	// (stage_bytes).
	auto mask_row = [](int ms) { return ((ms >> 1) & 1) ? ms : ms + 2; };
	auto wants_copy = [&](const HostJob &H) {
		const int wm = H.j->weight_mode, ms = H.a->m_stride;
		return fast && H.weighted && (H.j->kind == TWOPCF_ANGULAR || H.j->kind == TWOPCF_MONOPOLE) &&
		       (wm == TWOPCF_W_ANG_INVP || wm == TWOPCF_W_INVP_3D) && ms > 0 && (ms & 1) == 0;
	};
	int stage_bytes[4] = {0, 0, 0, 0};
	for (int q = 0; q < n_jobs; ++q) {
		const HostJob &H = hj[q];
		if (!wants_copy(H)) continue;
		const int bytes = ICHUNK * mask_row(H.a->m_stride) * 8;
		const int hb = H.nb * 8;
		const int fs = (size_t)H.nb * WAVE * 8 <= 32768 ? 4 : 0;
		if ((long long)lds_fixed + (hb << fs) + hb + bytes + 16 <= (long long)ctx->lds_max)
			stage_bytes[H.j->kind] = std::max(stage_bytes[H.j->kind], bytes);
	}
	// A monopole job over the same catalogues, weights and s minimum as an
	// s-mu job, without regions, is binned in the s-mu job's pass (fp32
	// path): the pairs are screened and compacted once.  fused[s] = m for the
	// s-mu job s, absorbed[m] = true for the monopole job m, which gets no
	// work items of its own.
	std::vector<int> fused(n_jobs, -1);
	std::vector<char> absorbed(n_jobs, 0);
	auto copies_of = [](int nb, bool weighted) { return (size_t)nb * WAVE * (weighted ? 8 : 4) <= 32768 ? 6 : 0; };
	for (int s_ = 0; s_ < n_jobs && fast; ++s_) {
		const HostJob &S = hj[s_];
		if (S.j->kind != TWOPCF_S_MU || S.j->n_jk > 1 || S.j->weight_mode != TWOPCF_W_PLAIN) continue;
		for (int m = 0; m < n_jobs; ++m) {
			const HostJob &M = hj[m];
			if (absorbed[m] || M.j->kind != TWOPCF_MONOPOLE || M.j->n_jk > 1 || M.j->weight_mode != TWOPCF_W_PLAIN)
				continue;
			if (M.a != S.a || M.b != S.b || M.weighted != S.weighted || M.rmax != S.rmax) continue;
			if (M.ax1.min != S.ax1.min) continue;
			const int hs = S.nb * (S.weighted ? 8 : 4), hm = M.nb * (M.weighted ? 8 : 4);
			if ((long long)lds_fixed + ((long long)hs << copies_of(S.nb, S.weighted)) +
			        ((long long)hm << copies_of(M.nb, M.weighted)) + 16 >
			    (long long)ctx->lds_max)
				continue;
			fused[s_] = m;
			absorbed[m] = 1;
			break;
		}
	}
	// Twins: an angular job with parent weights (the products of the weights
	// or of the input weights) and another of the same catalogues, axis and
	// pair set with the other product, neither with regions, are counted in
	// one pass (fp32 path): each pair is screened, compacted and binned once
	// and added to both histograms with its two weights.  twin[p] = t for the
	// job p whose pass bins job t, absorbed[t] = true.
	std::vector<int> twin(n_jobs, -1);
	auto parent_rule = [](int m) { return m == TWOPCF_W_ANG_IW || m == TWOPCF_W_ANG_W; };
	for (int p = 0; p < n_jobs && fast; ++p) {
		const HostJob &P1 = hj[p];
		if (absorbed[p] || fused[p] >= 0 || P1.j->kind != TWOPCF_ANGULAR || P1.j->n_jk > 1 || !P1.weighted ||
		    !parent_rule(P1.j->weight_mode))
			continue;
		for (int t = 0; t < n_jobs; ++t) {
			const HostJob &T = hj[t];
			if (t == p || absorbed[t] || fused[t] >= 0 || twin[t] >= 0 || T.j->kind != TWOPCF_ANGULAR ||
			    T.j->n_jk > 1 || !T.weighted || !parent_rule(T.j->weight_mode))
				continue;
			if (T.a != P1.a || T.b != P1.b || T.rmax != P1.rmax || std::memcmp(T.j->ax, P1.j->ax, sizeof T.j->ax) != 0)
				continue;
			const long long hp = (long long)P1.nb * 8 << copies_of(P1.nb, true);
			const long long ht = (long long)T.nb * 8 << copies_of(T.nb, true);
			if ((long long)lds_fixed + hp + ht + 16 > (long long)ctx->lds_max) continue;
			twin[p] = t;
			absorbed[t] = 1;
			break;
		}
	}
	auto fz_of = [&](int q) { return fused[q] >= 0 ? FZ_MONO : twin[q] >= 0 ? FZ_TWIN : FZ_NONE; };
	// Work items are built after this loop, on host threads: each job leaves
	// a builder of the items of a range of its cells, and the ranges' items,
	// slot lists and counts are appended in job and cell order, so that the
	// item list is the one a single thread would build.
	struct RangeOut {
		std::vector<WorkItem> items;
		std::vector<double> cost;
		std::vector<int> slots;
		uint64_t n_staged = 0;
		bool too_large = false;
	};
	std::vector<std::function<void(int, int, RangeOut &)>> builders;
	std::vector<int> builder_cells;
	for (int q = 0; q < n_jobs; ++q) {
		HostJob &H = hj[q];
		Prepared &P = prep[q];
		SortedCat *A = nullptr, *B = nullptr;
		int rc;
		if ((rc = grid_catalog(ctx, H.a, H.g, &A, !arena))) return rc;
		if (H.b) {
			if ((rc = grid_catalog(ctx, H.b, H.g, &B, !arena))) return rc;
		} else {
			B = A;
		}
		P.A = A;
		P.B = B;
		P.swapped = H.swapped;
		P.nb = H.nb;
		P.weighted = H.weighted;
		JobDev &D = P.jd;
		std::memset(&D, 0, sizeof(D));
		auto fill_cat = [](CatDev &cd, SortedCat *s) {
			cd.x = s->x; cd.y = s->y; cd.z = s->z;
			cd.nx = s->nx; cd.ny = s->ny; cd.nz = s->nz;
			cd.w = s->w; cd.iw = s->iw; cd.jk = s->jk; cd.masks = s->masks; cd.m_stride = s->m_stride;
			cd.start = s->start;
		};
		fill_cat(D.A, A);
		fill_cat(D.B, B);
		for (int k = 0; k < 3; ++k) {
			D.dim[k] = H.g.dim[k];
			D.lo[k] = H.g.lo[k];
		}
		D.h = H.g.h;
		// screening forms s^2 as |a|^2 + (|b|^2 - 2 a.b) with the matrix cores;
		// offsets are within R per axis, so every term is below 6 R^2 and the
		// FP32 roundings of this form and of the difference form stay far
		// below 2^-15 R^2
		D.scr_eps = (float)std::ldexp(std::pow(1.01 * (subdiv + 1) * H.g.h, 2), -15);
		D.kind = H.j->kind;
		D.is_auto = H.b == nullptr;
		D.n1 = H.ax1.n;
		D.n2 = H.ax2.n;
		D.nb = H.nb;
		D.n_jk = H.j->n_jk;
		D.weighted = H.weighted;
		D.min_r_sq = H.min_r_sq;
		D.max_r_sq = H.max_r_sq;
		D.ax1 = AxisDev{H.ax1.n, H.ax1.log ? 1 : 0, H.ax1.minus_min, H.ax1.inv_step};
		D.ax2 = AxisDev{H.ax2.n, H.ax2.log ? 1 : 0, H.ax2.minus_min, H.ax2.inv_step};
		// identical tables (jobs on the same axes) are stored once; every
		// table starts 16-byte aligned
		// bin k is sure for A_k <= s2 < B_k+1 (A = T + E rounded up, B = T - E
		// rounded down): the pairs (A_k, B_k+1), one 8-byte load per pair
		auto sure_pairs = [](const std::vector<float> &A, const std::vector<float> &B) {
			std::vector<float> P(2 * A.size());
			for (size_t k = 0; k < A.size(); ++k) {
				P[2 * k] = A[k];
				P[2 * k + 1] = k + 1 < B.size() ? B[k + 1] : -INFINITY;
			}
			return P;
		};
		auto push = [&](const std::vector<float> &v) {
			std::string k((const char *)v.data(), v.size() * sizeof(float));
			auto f = table_at.find(k);
			if (f != table_at.end()) return f->second;
			while (tables.size() % 4) tables.push_back(0.0f);
			int off = (int)tables.size();
			tables.insert(tables.end(), v.begin(), v.end());
			table_at[k] = off;
			return off;
		};
		if (H.j->kind == TWOPCF_ANGULAR) {
			D.t1_off = push(H.t1);
			if (fast) {
				// chord^2 = 2 (1 - cos): thresholds 2 t (exact), margins from
				// the FP32 chord^2 bound plus the reference's own roundings of
				// 1 - n1.n2 in double (about 1e-15 each for the products and
				// subtractions and for the unit vectors' norm error; 1e-14 is
				// allowed in the chord^2 domain) and to float (2^-24 relative,
				// 2^-21 allowed)
				const double R = 1.01 * (subdiv + 1) * H.g.h;
				const size_t nt = H.t1.size();
				std::vector<float> T2(nt), A_(nt), B_(nt);
				for (size_t k = 0; k < nt; ++k) {
					const double T = 2.0 * (double)H.t1[k];
					T2[k] = (float)T;
					if (!std::isfinite(T)) {
						A_[k] = B_[k] = (float)T;
						continue;
					}
					const double E = 1.01 * e_s2(std::max(T, 0.0) * 1.001, R) + 1e-14 + std::ldexp(std::fabs(T), -21) +
					                 1e-30;
					A_[k] = round_up(T + E);
					B_[k] = round_dn(T - E);
				}
				D.t_s2_off = push(T2);
				D.tA_off = push(A_);
				D.tB_off = push(B_);
				D.tq_off = push(sure_pairs(A_, B_));
				D.s_min = H.ax1.min;
				D.s_log = H.ax1.log ? 1 : 0;
				D.s_ilb_ln2 = (float)((double)H.ax1.ilb * std::log(2.0));
				D.s_ils = H.ax1.ils;
				D.s_inv_step = H.ax1.inv_step;
			}
		} else if (H.j->kind == TWOPCF_SIGMA_PI) {
			D.t1_off = push(H.t1);
			D.t2_off = push(H.t2);
			if (fast) {
				// sigma decided on x = s^2 - pi^2 against the smallest x whose
				// sqrtf reaches each sigma threshold; range limits with the
				// s^2 margins (A/B at [0] and [n1]); per-pair margins from
				// E(S) and |dpi| <= 2^-23 (20 s + 8 R) (generous for the
				// offsets', the normals' and both pipelines' roundings)
				std::vector<float> X1;
				if (!sqrt_thresholds(H.t1, X1)) return fail(ctx, TWOPCF_E_AXIS, "sigma axis not monotone under sqrtf");
				D.x1_off = push(X1);
				const double R = 1.01 * (subdiv + 1) * H.g.h;
				const int n1 = H.ax1.n;
				std::vector<float> A_(n1 + 1, 0.0f), B_(n1 + 1, 0.0f);
				const double mn = H.min_r_sq, mx = H.max_r_sq;
				const double Emn = 1.01 * e_s2(std::max(mn, 0.0) * 1.001, R) + 1e-30;
				const double Emx = 1.01 * e_s2(mx * 1.001, R) + 1e-30;
				B_[0] = round_dn(mn - Emn);
				A_[0] = round_up(mn + Emn);
				A_[n1] = round_up(mx + Emx);
				B_[n1] = round_dn(mx - Emx);
				D.tA_off = push(A_);
				D.tB_off = push(B_);
				D.s_min = H.ax1.min;
				D.s_log = H.ax1.log ? 1 : 0;
				D.s_ilb_ln2 = (float)((double)H.ax1.ilb * std::log(2.0));
				D.s_ils = H.ax1.ils;
				D.s_inv_step = H.ax1.inv_step;
				D.p_min = H.ax2.min;
				D.p_log = H.ax2.log ? 1 : 0;
				D.p_ilb_ln2 = (float)((double)H.ax2.ilb * std::log(2.0));
				D.p_ils = H.ax2.ils;
				D.p_inv_step = H.ax2.inv_step;
				D.e_c1 = (float)(1.02 * std::ldexp(11.0, -23));
				D.e_c2 = (float)(1.02 * std::ldexp(4.0 * R, -23));
				const double mid = H.j->pi_def == TWOPCF_PI_MIDPOINT ? 2.0 : 1.0;
				D.pi_c1 = (float)(1.02 * std::ldexp(20.0 * mid, -23));
				D.pi_c2 = (float)(1.02 * std::ldexp(8.0 * R, -23));
				D.mid_c = H.j->pi_def == TWOPCF_PI_MIDPOINT ? (float)(1.02 * std::ldexp(8.0 * R, -23)) : 0.0f;
			}
		} else {
			D.t_s2_off = push(H.t_s2);
			// fast-path margins around each s^2 threshold
			const double R = 1.01 * (subdiv + 1) * H.g.h;
			std::vector<float> A_(H.t_s2.size()), B_(H.t_s2.size());
			for (size_t k = 0; k < H.t_s2.size(); ++k) {
				double T = H.t_s2[k];
				if (!std::isfinite(T)) {
					A_[k] = B_[k] = (float)T;
					continue;
				}
				double E = 1.01 * e_s2(std::max(T, 0.0) * 1.001, R) + 1e-30;
				A_[k] = round_up(T + E);
				B_[k] = round_dn(T - E);
			}
			D.tA_off = push(A_);
			D.tB_off = push(B_);
			D.tq_off = push(sure_pairs(A_, B_));
			D.s_min = H.ax1.min;
			D.s_log = H.ax1.log ? 1 : 0;
			D.s_ilb_ln2 = (float)((double)H.ax1.ilb * std::log(2.0));
			D.s_ils = H.ax1.ils;
			D.s_inv_step = H.ax1.inv_step;
			const double n2 = H.ax2.n;
			// the pair-midpoint line of sight doubles the relative part and
			// adds 1.02 2^-23 8 R s / |x1 + x2| to the pi bound
			const double mid = H.j->pi_def == TWOPCF_PI_MIDPOINT ? 2.0 : 1.0;
			// relative part: 41 ulp (2^-23) for the pi and s chain, doubled
			// for the midpoint; the angular definition's normals staged times
			// n2 / 2 add two roundings, taken as 2 ulp more
			const double c1 = n2 * std::ldexp(1.0, -23) * (mid > 1.0 ? 82.0 : 43.0);
			D.mu_h0 = (float)(0.5 - c1 - std::ldexp(1.0, -21));
			D.mu_c2 = (float)(n2 * std::ldexp(1.0, -23) * 4.75 * R * 1.01);
			D.mid_c = H.j->pi_def == TWOPCF_PI_MIDPOINT ? (float)(1.02 * std::ldexp(8.0 * R, -23)) : 0.0f;
			max_tab_n1 = std::max(max_tab_n1, H.ax1.n);
		}
		auto rr = rows_of[H.rmax];
		D.row_off = rr.first;
		D.n_rows = rr.second;
		// small histograms of the fp32 path get one copy per lane
		D.f_shift = (fast && (size_t)H.nb * WAVE * (H.weighted ? 8 : 4) <= 32768)
		                ? 6 : 0;
		// first-side mask copy in LDS (see stage_bytes); the jobs that take
		// it form their own launch class
		if (wants_copy(H) && stage_bytes[H.j->kind] > 0) {
			D.f_shift = std::min(D.f_shift, 4);
			P.mask_bytes = stage_bytes[H.j->kind];
			D.mstage = mask_row(H.a->m_stride);
		}
		P.f_elems = H.nb << D.f_shift;
		D.pi_def = H.j->pi_def;
		D.wmode = H.j->weight_mode;
		D.n_runs = H.j->n_runs;
		D.n_words = H.a->n_words;
		D.dither = H.j->dither ? 1 : 0;
		if (!H.up_bins.empty()) {
			double *ub = nullptr, *uc = nullptr;
			float *ut = nullptr;
			const size_t n = H.up_bins.size();
			for (void **p : {(void **)&ub, (void **)&uc, (void **)&ut}) {
				if (int rc = dev_alloc(p, (n + 1) * 8)) {
					plan.release();
					return rc;
				}
			}
			HIPCK(ctx, hipMemcpy(ub, H.up_bins.data(), n * 8, hipMemcpyHostToDevice));
			HIPCK(ctx, hipMemcpy(uc, H.up_centre.data(), n * 8, hipMemcpyHostToDevice));
			HIPCK(ctx, hipMemcpy(ut, H.up_thr.data(), (n + 1) * 4, hipMemcpyHostToDevice));
			D.up_bins = ub;
			D.up_centre = uc;
			D.up_thr = ut;
			D.up_n = (int)n;
			D.up_min = (double)H.up_ax.min;
			D.up_max = (double)H.up_ax.max;
		}
		P.full_n = H.nb;
		P.jk_n = H.j->n_jk > 1 ? (size_t)H.j->n_jk * H.nb : 0;
		if (H.weighted) {
			P.full_off = pool_d;
			P.jk_off = pool_d + P.full_n;
			pool_d += P.full_n + P.jk_n;
		} else {
			P.full_off = pool_u;
			P.jk_off = pool_u + P.full_n;
			pool_u += P.full_n + P.jk_n;
		}

		// work items: chunks of <= 64 first-side objects per cell, rows
		// grouped so that each item sees at most slot_cap regions; the slot
		// cap also keeps (1 + slots) histograms inside the LDS budget
		if (absorbed[q]) continue;   // binned in its s-mu job's pass
		if (fused[q] >= 0) {
			const HostJob &M = hj[fused[q]];
			D.m_nb = M.nb;
			D.m_f_shift = copies_of(M.nb, M.weighted);
			P.f_elems += M.nb << D.m_f_shift;
		}
		if (twin[q] >= 0) {
			const HostJob &T = hj[twin[q]];
			D.m_nb = T.nb;
			D.m_f_shift = copies_of(T.nb, true);
			P.f_elems += T.nb << D.m_f_shift;
			// the twin's product is of the other staged weights unless its
			// rule stages the same ones
			D.t_w = T.j->weight_mode != H.j->weight_mode ? 1 : 0;
		}
		// LDS for the histograms: the full one (with its lane copies) and, with
		// regions, at least one slot histogram.  When they do not fit, the
		// lane copies go first, then the full histogram moves to the global
		// pool (its pairs are added there with atomics); regions without a
		// slot are added to the global pool as well, and with no room for a
		// slot every region is.
		const bool jk = H.j->n_jk > 1;
		const int hist_bytes = H.nb * (H.weighted ? 8 : 4);
		const long long avail = (long long)lds_budget - lds_fixed - (P.mask_bytes ? P.mask_bytes + 16 : 0);
		long long f_bytes = (long long)P.f_elems * (H.weighted ? 8 : 4);
		if (fz_of(q) == FZ_NONE && f_bytes + (jk ? hist_bytes : 0) > avail && D.f_shift > 0) {
			D.f_shift = 0;
			P.f_elems = H.nb;
			f_bytes = hist_bytes;
		}
		// With regions, the full histogram (every pair of an item in a
		// one-region cell, the cross pairs of the others) takes the most
		// atomics: two lane copies when eight region slots (or the tuning's
		// maximum, if smaller) still fit beside them.  Jackknife batch 378.1 ->
		// 373.3 ms with 8 slots instead of 9; four copies (6 slots) 384.2.
		if (fast && jk && fz_of(q) == FZ_NONE && P.mask_bytes == 0 && D.f_shift == 0) {
			const int want = std::min(max_slots, 8);
			if (2LL * hist_bytes + (long long)want * hist_bytes <= avail) {
				D.f_shift = 1;
				P.f_elems = H.nb << 1;
				f_bytes = 2LL * hist_bytes;
			}
		}
		// Without regions the LDS beside the fixed part holds only the full
		// histograms: a large one gets as many lane copies (a power of two, up
		// to one per lane) as fit, so that lanes adding to one bin in the
		// same atomic hit different addresses.
		if (fast && !jk && P.mask_bytes == 0 && f_bytes <= avail) {
			const long long es = H.weighted ? 8 : 4;
			const long long other = f_bytes - ((long long)H.nb << D.f_shift) * es;   // a fused or twin job's
			int fs = D.f_shift;
			while (fs < 6 && (((long long)H.nb << (fs + 1)) * es) + other <= avail) ++fs;
			if (fs > D.f_shift) {
				P.f_elems += (H.nb << fs) - (H.nb << D.f_shift);
				D.f_shift = fs;
				f_bytes = (long long)P.f_elems * es;
			}
		}
		if (f_bytes + (jk ? hist_bytes : 0) > avail) {
			if (fz_of(q) != FZ_NONE) return fail(ctx, TWOPCF_E_LIMIT, "fused histograms do not fit in LDS");
			D.full_global = 1;
			P.f_elems = 0;
			f_bytes = 0;
		}
		if (avail < f_bytes) return fail(ctx, TWOPCF_E_LIMIT, "tables do not fit in LDS");
		const int slot_cap = jk ? (int)std::min<long long>(max_slots, (avail - f_bytes) / std::max(hist_bytes, 1)) : 0;
		// group rows by the regions they bring only when there are slots to fill
		const bool group_regions = jk && slot_cap > 0;
		const int n_cell = H.g.dim[0] * H.g.dim[1] * H.g.dim[2];
		const int d1 = H.g.dim[1], d2 = H.g.dim[2];
		std::vector<int> rws;
		for (int r = 0; r < D.n_rows; ++r) rws.push_back(r);
		builder_cells.push_back(n_cell);
		builders.emplace_back([=, &H, &D, &row_pool](int c_lo, int c_hi, RangeOut &out) {
			// scratch reused across cells and chunks (no allocation per row)
			std::vector<int> regs_i, regs, add;
			struct RowInfo {
				int r, jb, je;
			};
			std::vector<RowInfo> ri;
			for (int c = c_lo; c < c_hi; ++c) {
				const int ib = A->start_h[c], ie = A->start_h[c + 1];
				if (ib == ie) continue;
				const int cx = c / (d1 * d2), cy = (c / d2) % d1, cz = c % d2;
				// per row: j range, cost
				ri.clear();
				for (int r : rws) {
					const int4 row = row_pool[D.row_off + r];
					const int qx = cx + row.x, qy = cy + row.y;
					if (qx < 0 || qy < 0 || qx >= H.g.dim[0] || qy >= d1) continue;
					if (D.is_auto && (row.x < 0 || (row.x == 0 && row.y < 0))) continue;
					int z0 = std::max(cz + row.z, 0), z1 = std::min(cz + row.w, d2 - 1);
					if (D.is_auto && row.x == 0 && row.y == 0) z0 = std::max(z0, cz);
					if (z0 > z1) continue;
					const int base = (qx * d1 + qy) * d2;
					const int jb = B->start_h[base + z0], je = B->start_h[base + z1 + 1];
					if (jb == je) continue;
					ri.push_back({r, jb, je});
				}
				if (ri.empty()) continue;
				// a cell in one region (fp32 path, full histogram in LDS): its
				// items count every pair in the full histogram, which is also
				// that region's touching count, and need no slot for it
				const int uf_r = fast && group_regions && !D.full_global && A->reg_off[c + 1] - A->reg_off[c] == 1
				                     ? A->reg_ids[A->reg_off[c]] : -1;
				for (int i0 = ib; i0 < ie; i0 += i_chunk) {
					const int ni = std::min(i_chunk, ie - i0);
					// group rows by region budget
					regs_i.clear();
					if (group_regions && uf_r < 0) {
						// regions of this chunk: those present in its cell (a
						// superset of the chunk's regions)
						for (int k = A->reg_off[c]; k < A->reg_off[c + 1]; ++k) regs_i.push_back(A->reg_ids[k]);
					}
					size_t g0 = 0;
					while (g0 < ri.size()) {
						regs.assign(regs_i.begin(), regs_i.end());
						unsigned long long mask = 0;
						double cost = 0;
						size_t g1 = g0;
						for (; g1 < ri.size(); ++g1) {
							if (group_regions) {
								add.clear();
								const int4 row = row_pool[D.row_off + ri[g1].r];
								const int qx = cx + row.x, qy = cy + row.y;
								const int base = (qx * d1 + qy) * d2;
								int z0 = std::max(cz + row.z, 0), z1 = std::min(cz + row.w, d2 - 1);
								if (D.is_auto && row.x == 0 && row.y == 0) z0 = std::max(z0, cz);
								for (int z = z0; z <= z1; ++z)
									for (int k = B->reg_off[base + z]; k < B->reg_off[base + z + 1]; ++k)
										if (B->reg_ids[k] != uf_r && std::find(regs.begin(), regs.end(), B->reg_ids[k]) == regs.end() &&
										    std::find(add.begin(), add.end(), B->reg_ids[k]) == add.end())
											add.push_back(B->reg_ids[k]);
								if ((int)(regs.size() + add.size()) > slot_cap && g1 > g0) break;
								regs.insert(regs.end(), add.begin(), add.end());
							}
							mask |= 1ull << ri[g1].r;
							cost += (double)ni * (ri[g1].je - ri[g1].jb);
							out.n_staged += (uint64_t)(ri[g1].je - ri[g1].jb);
						}
						if (cost * 1.0 >= 4.0e9) {
							out.too_large = true;
							return;
						}
						WorkItem W;
						W.job = q;
						W.i_begin = i0;
						W.i_count = ni;
						W.i_cell = c;
						W.row_mask = mask;
						W.slot_base = (int)out.slots.size();   // within this range's list
						// regs holds every region of the chunk's cell and of the
						// rows' cells (but the cell's own with uf_r): when they
						// all fit, every pair's two regions have slots
						W.all_slotted = group_regions && (int)regs.size() <= slot_cap ? 1 : 0;
						W.uf_region = uf_r;
						if ((int)regs.size() > slot_cap) regs.resize(slot_cap);   // the rest: global atomics
						W.n_slots = group_regions ? (int)regs.size() : 0;
						if (group_regions) out.slots.insert(out.slots.end(), regs.begin(), regs.end());
						out.items.push_back(W);
						out.cost.push_back(cost);
						out.n_staged += (uint64_t)ni;
						g0 = g1;
					}
				}
			}
		});
	}
	{
		// each job's cells in up to 16 ranges, the ranges pulled by the threads
		struct Task {
			int b, lo, hi;
		};
		std::vector<Task> tasks;
		for (size_t b = 0; b < builders.size(); ++b) {
			const int nc = builder_cells[b];
			const int nr = std::max(1, std::min(16, nc / 256 + 1));
			for (int t = 0; t < nr; ++t)
				tasks.push_back({(int)b, (int)((long long)nc * t / nr), (int)((long long)nc * (t + 1) / nr)});
		}
		std::vector<RangeOut> outs(tasks.size());
		std::atomic<size_t> next{0};
		std::atomic<bool> failed{false};
		auto work = [&]() {
			try {
				for (size_t k; (k = next.fetch_add(1)) < tasks.size();) builders[tasks[k].b](tasks[k].lo, tasks[k].hi, outs[k]);
			} catch (...) {
				failed = true;   // host memory: the call fails once every thread has joined
				next = tasks.size();
			}
		};
		const int n_thr = std::max(1, std::min<int>({16, (int)std::thread::hardware_concurrency(), (int)tasks.size()}));
		std::vector<std::thread> thr;
		for (int t = 1; t < n_thr; ++t) {
			try {
				thr.emplace_back(work);
			} catch (const std::system_error &) {
				break;   // fewer threads pull the same tasks
			}
		}
		work();
		for (auto &th : thr) th.join();
		if (failed) throw std::bad_alloc();
		for (RangeOut &o : outs) {
			if (o.too_large) return fail(ctx, TWOPCF_E_LIMIT, "work item too large for 32-bit counters");
			const int sb = (int)slot_pool.size();
			for (WorkItem &W : o.items) W.slot_base += sb;
			items.insert(items.end(), o.items.begin(), o.items.end());
			item_cost.insert(item_cost.end(), o.cost.begin(), o.cost.end());
			slot_pool.insert(slot_pool.end(), o.slots.begin(), o.slots.end());
			n_staged += o.n_staged;
		}
	}
	if (slot_pool.empty()) slot_pool.push_back(0);
	if (tables.empty()) tables.push_back(0.0f);
	if (row_pool.empty()) row_pool.push_back(make_int4(0, 0, 0, 0));

	// ---- device buffers for this call ----
	JobDev *d_jobs = nullptr;
	WorkItem *d_items = nullptr;
	int4 *d_rows = nullptr;
	int *d_slots = nullptr;
	float *d_tab = nullptr;
	unsigned long long *d_pool_u = nullptr, *d_cnt = nullptr;
	double *d_pool_d = nullptr;
	unsigned int *d_queue = nullptr;
	auto release = [&]() { plan.release(); };
	auto alloc = [&](void **p, size_t bytes) -> int { return dev_alloc(p, bytes); };
	int rc = 0;
	if ((rc = alloc((void **)&d_jobs, sizeof(JobDev) * std::max(n_jobs, 1))) ||
	    (rc = alloc((void **)&d_items, sizeof(WorkItem) * std::max<size_t>(items.size(), 1))) ||
	    (rc = alloc((void **)&d_rows, sizeof(int4) * row_pool.size())) ||
	    (rc = alloc((void **)&d_slots, sizeof(int) * slot_pool.size())) ||
	    (rc = alloc((void **)&d_tab, sizeof(float) * tables.size())) ||
	    (rc = alloc((void **)&d_pool_u, sizeof(unsigned long long) * std::max<size_t>(pool_u, 1))) ||
	    (rc = alloc((void **)&d_pool_d, sizeof(double) * std::max<size_t>(pool_d, 1))) ||
	    (rc = alloc((void **)&d_cnt, sizeof(unsigned long long) * 2 * std::max(n_jobs, 1))) ||
	    (rc = alloc((void **)&d_queue, sizeof(unsigned int) * NQ * MAX_CLASSES))) {
		release();
		return rc;
	}
	for (int q = 0; q < n_jobs; ++q) {
		Prepared &P = prep[q];
		if (P.weighted) {
			P.jd.full_d = d_pool_d + P.full_off;
			P.jd.jk_d = d_pool_d + P.jk_off;
		} else {
			P.jd.full_u = d_pool_u + P.full_off;
			P.jd.jk_u = d_pool_u + P.jk_off;
		}
		P.jd.counters = d_cnt + 2 * q;
	}
	for (int q = 0; q < n_jobs; ++q) {
		if (twin[q] < 0) continue;
		JobDev &D = prep[q].jd;
		const JobDev &T = prep[twin[q]].jd;
		D.m_full_u = T.full_u;
		D.m_full_d = T.full_d;
		D.m_counters = T.counters;
	}
	for (int q = 0; q < n_jobs; ++q) {
		if (fused[q] < 0) continue;
		JobDev &D = prep[q].jd;
		const JobDev &M = prep[fused[q]].jd;
		D.m_tq_off = M.tq_off;
		D.m_t_s2_off = M.t_s2_off;
		D.m_tA_off = M.tA_off;
		D.m_tB_off = M.tB_off;
		// the estimate's unused terms zero: inv_step on a log axis, ils and
		// ilb on a linear one
		D.m_s_ilb_ln2 = M.s_log ? M.s_ilb_ln2 : 0.0f;
		D.m_s_ils = M.s_log ? M.s_ils : 0.0f;
		D.m_s_inv_step = M.s_log ? 0.0f : M.s_inv_step;
		D.m_ax1 = M.ax1;
		D.m_full_u = M.full_u;
		D.m_full_d = M.full_d;
		D.m_counters = M.counters;
	}
	std::vector<JobDev> jd_h(n_jobs);
	for (int q = 0; q < n_jobs; ++q) jd_h[q] = prep[q].jd;
	// Queue order of the batched method.  The items of the 3-D kinds stay in
	// build order (job by job, cells in order): items that run close in time
	// then share their neighbour cells in the caches, which halves the HBM
	// traffic of the jackknife batch (1.11 -> 0.48 B fetched per pair) and
	// shortens it 1.5 %.  Angular items are ordered by cost, largest first,
	// which measured faster for them.
	if (fast) {
		std::vector<size_t> ord, rest;
		ord.reserve(items.size());
		for (size_t k = 0; k < items.size(); ++k)
			(hj[items[k].job].j->kind == TWOPCF_ANGULAR ? ord : rest).push_back(k);
		std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return item_cost[x] > item_cost[y]; });
		ord.insert(ord.end(), rest.begin(), rest.end());
		std::vector<WorkItem> s(items.size());
		for (size_t k = 0; k < ord.size(); ++k) s[k] = items[ord[k]];
		if (sample_stride > 1) {
			std::vector<WorkItem> t;
			for (size_t k = 0; k < s.size(); k += (size_t)sample_stride) t.push_back(s[k]);
			s.swap(t);
		}
		items.swap(s);
	}
	auto h2d = [&](void *dst, const void *src, size_t bytes) -> int {
		if (!bytes) return 0;
		HIPCK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
		return 0;
	};
	if ((rc = h2d(d_jobs, jd_h.data(), sizeof(JobDev) * n_jobs)) ||
	    (rc = fast ? 0 : h2d(d_items, items.data(), sizeof(WorkItem) * items.size())) ||   // batched: by class below
	    (rc = h2d(d_rows, row_pool.data(), sizeof(int4) * row_pool.size())) ||
	    (rc = h2d(d_slots, slot_pool.data(), sizeof(int) * slot_pool.size())) ||
	    (rc = h2d(d_tab, tables.data(), sizeof(float) * tables.size()))) {
		release();
		return rc;
	}
	const double t_prep1 = now_ms();
	plan.prep_ms = t_prep1 - t_prep0;
	plan.d_jobs = d_jobs;
	plan.d_rows = d_rows;
	plan.d_slots = d_slots;
	plan.d_tab = d_tab;
	plan.d_pool_u = d_pool_u;
	plan.d_pool_d = d_pool_d;
	plan.d_cnt = d_cnt;
	plan.d_queue = d_queue;
	plan.pool_u = pool_u;
	plan.pool_d = pool_d;
	plan.n_jobs = n_jobs;
	plan.n_tables = (int)tables.size();
	plan.n_items = (int)items.size();
	plan.n_staged = n_staged;
	for (int q = 0; q < n_jobs; ++q)
		plan.outs.push_back(PlanOut{prep[q].weighted, prep[q].full_off, prep[q].jk_off, prep[q].full_n, prep[q].jk_n,
		                            fast && !prep[q].weighted && prep[q].jk_n > 0});

	// ---- launches ----
	// LDS: tile + tables + slot regions + histograms
	int max_nb = 1, max_tab = (int)tables.size();
	for (auto &H : hj) max_nb = std::max(max_nb, H.nb);
	(void)max_tab_n1;
	// h_elems: histogram elements of the launch's largest job: its full
	// histogram in LDS (with lane copies; none when kept in the global pool)
	// and its region slots
	auto layout = [&](bool fast_kernel, bool weighted, int h_elems, int mask_bytes) {
		Lds L;
		L.tile = 0;
		L.tables = (tile_bytes(fast_kernel) + 15) & ~15;
		L.slots = L.tables + (((int)tables.size() * 4 + 15) & ~15);
		L.hist = L.slots + ((MAX_SLOTS * 4 + 15) & ~15);
		L.total = L.hist + h_elems * (weighted ? 8 : 4);
		L.masks = 0;
		if (mask_bytes > 0) {
			L.masks = (L.total + 15) & ~15;
			L.total = L.masks + mask_bytes;
		}
		return L;
	};
	auto launch = [&](int kind, bool weighted, bool mid, int fz, int rg, const WorkItem *it, int n_it, int h_elems,
	                  unsigned *queue, int grid, int mask_bytes) -> int {
		const bool fast_kernel = fast;
		Lds L = layout(fast_kernel, weighted, h_elems, mask_bytes);
		if ((size_t)L.total > ctx->lds_max)
			return fail(ctx, TWOPCF_E_LIMIT, "histogram does not fit in LDS (" + std::to_string(L.total) + " B)");
		KernelFn fn = nullptr;
		if (fz != FZ_NONE && rg != RG_NONE)
			return fail(ctx, TWOPCF_E_LIMIT, "fused monopole or twin with regions or a global histogram");
#define PICK(K, F, W) fn = k_count<K, F, W>
#define PICK_FAST(K, W, M, FZ) fn = fast_kernel_of<K, W, M, FZ>(rg)
#define PICK_PI(K, W, FZ) if (mid) PICK_FAST(K, W, true, FZ); else PICK_FAST(K, W, false, FZ)
		if (fast) {
			if (kind == K_MONO) { if (weighted) PICK_FAST(K_MONO, true, false, FZ_NONE); else PICK_FAST(K_MONO, false, false, FZ_NONE); }
			else if (kind == K_S_MU && fz == FZ_MONO) {
				// s-mu jobs with a fused monopole
				if (weighted) { PICK_PI(K_S_MU, true, FZ_MONO); } else { PICK_PI(K_S_MU, false, FZ_MONO); }
			}
			else if (kind == K_S_MU) { if (weighted) { PICK_PI(K_S_MU, true, FZ_NONE); } else { PICK_PI(K_S_MU, false, FZ_NONE); } }
			else if (kind == K_SIGMA_PI) { if (weighted) { PICK_PI(K_SIGMA_PI, true, FZ_NONE); } else { PICK_PI(K_SIGMA_PI, false, FZ_NONE); } }
			else if (fz == FZ_TWIN) {
				// weighted angular jobs with a twin
				if (!weighted) return fail(ctx, TWOPCF_E_ARG, "twin of an unweighted job");
				PICK_FAST(K_ANGULAR, true, false, FZ_TWIN);
			}
			else { if (weighted) PICK_FAST(K_ANGULAR, true, false, FZ_NONE); else PICK_FAST(K_ANGULAR, false, false, FZ_NONE); }
		} else {
			if (kind == K_MONO) { if (weighted) PICK(K_MONO, false, true); else PICK(K_MONO, false, false); }
			else if (kind == K_S_MU) { if (weighted) PICK(K_S_MU, false, true); else PICK(K_S_MU, false, false); }
			else if (kind == K_SIGMA_PI) { if (weighted) PICK(K_SIGMA_PI, false, true); else PICK(K_SIGMA_PI, false, false); }
			else { if (weighted) PICK(K_ANGULAR, false, true); else PICK(K_ANGULAR, false, false); }
		}
#undef PICK
#undef PICK_FAST
#undef PICK_PI
		HIPCK(ctx, hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, L.total));
		if (n_it <= 0) return 0;
		plan.launches.push_back(PlanLaunch{fn, grid, L, it, n_it, queue});
		return 0;
	};
	if (!fast) {
		// plain: one launch per job, one work item per block, in cell order
		size_t k0 = 0;
		for (int q = 0; q < n_jobs; ++q) {
			size_t k1 = k0;
			while (k1 < items.size() && items[k1].job == q) ++k1;
			int slots = 0;
			for (size_t k = k0; k < k1; ++k) slots = std::max(slots, items[k].n_slots);
			if ((rc = launch(hj[q].j->kind, prep[q].weighted, false, FZ_NONE, RG_ANY, d_items + k0, (int)(k1 - k0),
			                 prep[q].f_elems + prep[q].nb * slots, nullptr, (int)(k1 - k0), 0))) {
				release();
				return rc;
			}
			k0 = k1;
		}
	} else {
		// batched: one persistent launch per (kind, weighted, pi definition,
		// fused monopole, LDS mask copy, histogram layout) class over the
		// ordered queue of all jobs in that class; the classes run
		// concurrently unless TWOPCF_SERIAL_CLASSES is set
		plan.concurrent = getenv("TWOPCF_SERIAL_CLASSES") == nullptr;
		auto mid_of = [&](int q) {
			const int k = hj[q].j->kind;
			return (k == TWOPCF_SIGMA_PI || k == TWOPCF_S_MU) && hj[q].j->pi_def == TWOPCF_PI_MIDPOINT ? 1 : 0;
		};
		// histogram layout: regions or none, full histogram in LDS; jobs with
		// the full histogram in the global pool take the generic kernel
		auto rg_of = [&](int q) {
			if (prep[q].jd.full_global) return RG_ANY;
			return hj[q].j->n_jk > 1 ? RG_JK : RG_NONE;
		};
		// launch class of each job with items, numbered in key order
		auto key_of = [&](int q) {
			return std::make_tuple(hj[q].j->kind, prep[q].weighted ? 1 : 0,
			                       mid_of(q) + 2 * fz_of(q) + (prep[q].jd.mstage > 0 ? 8 : 0) + 16 * rg_of(q));
		};
		std::vector<char> has_items(n_jobs, 0);
		for (const WorkItem &W : items) has_items[W.job] = 1;
		std::map<std::tuple<int, int, int>, int> cls;
		for (int q = 0; q < n_jobs; ++q)
			if (has_items[q]) cls.emplace(key_of(q), 0);
		if (cls.size() > MAX_CLASSES) {
			release();
			return fail(ctx, TWOPCF_E_LIMIT, "more than 64 launch classes in one call");
		}
		int n_cls = 0;
		for (auto &kv : cls) kv.second = n_cls++;
		std::vector<int> job_cls(n_jobs, -1);
		for (int q = 0; q < n_jobs; ++q)
			if (has_items[q]) job_cls[q] = cls[key_of(q)];
		// each class's items in queue order
		std::vector<std::vector<uint32_t>> members(n_cls);
		for (size_t k = 0; k < items.size(); ++k) members[job_cls[items[k].job]].push_back((uint32_t)k);
		size_t off = 0;
		std::vector<WorkItem> flat;
		flat.reserve(items.size());
		std::vector<std::tuple<int, int, size_t, size_t, int>> spans;
		for (auto &kv : cls) {
			const std::vector<uint32_t> &mem = members[kv.second];
			const size_t n_it = mem.size();
			spans.emplace_back(std::get<0>(kv.first), std::get<1>(kv.first), off, n_it, std::get<2>(kv.first));
			if (std::get<0>(kv.first) == K_ANGULAR && n_it >= (size_t)NQ) {
				// angular items are ordered by cost: dealt to the launch's NQ
				// queue ranges in turn, so that each range runs its items
				// largest first and the ranges carry about the same cost
				const size_t base = n_it / NQ, extra = n_it % NQ;
				for (size_t q = 0; q < (size_t)NQ; ++q)
					for (size_t j = 0; j < base + (q < extra ? 1 : 0); ++j) flat.push_back(items[mem[j * NQ + q]]);
			} else {
				for (uint32_t k : mem) flat.push_back(items[k]);
			}
			off += n_it;
		}
		if ((rc = h2d(d_items, flat.data(), sizeof(WorkItem) * flat.size()))) {
			release();
			return rc;
		}
		int qi = 0;
		for (auto &sp : spans) {
			const int kind = std::get<0>(sp);
			const bool weighted = std::get<1>(sp) != 0;
			const int mid = std::get<4>(sp) & 1;
			const int fz = (std::get<4>(sp) >> 1) & 3;   // s-mu jobs with a fused monopole, angular jobs with a twin
			const bool st = (std::get<4>(sp) & 8) != 0;   // jobs with the LDS mask copy
			const int rg = std::get<4>(sp) >> 4;            // histogram layout
			// per job: its largest slot count over the launch's items
			std::vector<int> slots_of(n_jobs, 0);
			for (size_t k = std::get<2>(sp); k < std::get<2>(sp) + std::get<3>(sp); ++k)
				slots_of[flat[k].job] = std::max(slots_of[flat[k].job], flat[k].n_slots);
			int h_elems = 1, mask_bytes = 0;
			for (int q = 0; q < n_jobs; ++q)
				if (!absorbed[q] && hj[q].j->kind == kind && prep[q].weighted == weighted && mid_of(q) == mid &&
				    fz_of(q) == fz && (prep[q].jd.mstage > 0) == st && rg_of(q) == rg) {
					h_elems = std::max(h_elems, prep[q].f_elems + hj[q].nb * slots_of[q]);
					mask_bytes = std::max(mask_bytes, prep[q].mask_bytes);
				}
			// jobs of other rules in the same launch may leave no room for the
			// mask copy: then every job of it reads its masks from global memory
			if ((size_t)layout(true, weighted, h_elems, mask_bytes).total > ctx->lds_max) mask_bytes = 0;
			Lds L = layout(true, weighted, h_elems, mask_bytes);
			const int per_cu = std::max(1, std::min(8, (int)(ctx->lds_max / std::max(L.total, 1))));
			const int grid = std::min<long long>((long long)ctx->n_cu * per_cu, (long long)std::get<3>(sp));
			if ((rc = launch(kind, weighted, mid != 0, fz, rg, d_items + std::get<2>(sp), (int)std::get<3>(sp), h_elems,
			                 d_queue + NQ * qi, std::max(grid, 1), mask_bytes))) {
				release();
				return rc;
			}
			++qi;
		}
	}
	return 0;
}

// Every input that shapes a plan except the tuning: method, catalogues,
// axes, regions, weight rule and up-weight table of every job.
std::string job_set_key(const twopcf_job *jobs, int n_jobs, int method, std::vector<uint64_t> *ids) {
	std::string key;
	auto put = [&](const void *p, size_t n) { key.append((const char *)p, n); };
	put(&method, sizeof method);
	for (int q = 0; q < n_jobs; ++q) {
		const twopcf_job &J = jobs[q];
		const uint64_t ia = J.a->id, ib = J.b ? J.b->id : 0;
		if (ids) {
			ids->push_back(ia);
			if (J.b) ids->push_back(ib);
		}
		put(&J.kind, sizeof J.kind);
		put(&ia, 8);
		put(&ib, 8);
		put(J.ax, sizeof J.ax);
		put(&J.n_jk, sizeof J.n_jk);
		put(&J.weight_mode, sizeof J.weight_mode);
		put(&J.n_runs, sizeof J.n_runs);
		put(&J.pi_def, sizeof J.pi_def);
		put(&J.dither, sizeof J.dither);
		put(&J.up_axis, sizeof J.up_axis);
		if (J.up_bins && J.up_axis.n_bins > 0) put(J.up_bins, sizeof(double) * J.up_axis.n_bins);
	}
	return key;
}

// Enqueue one execution of a plan on the context stream: clear the outputs
// with fill kernels (a graph may hold them; no memset nodes) and launch.
// The batched method's launch classes write disjoint outputs and pull from
// their own queues, so they run concurrently: launch k goes to the context
// stream (k = 0) or side stream (k - 1) mod N_SIDE, each side stream waits
// for the fills, and the context stream waits for every side stream; the
// tail of one persistent launch then overlaps the next class.  Captured into
// a graph, the fork and join become its dependencies.
int enqueue_plan(twopcf_ctx *ctx, Plan &plan) {
	auto fill = [&](unsigned long long *p, size_t n) {
		if (n) k_fill_u64<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(p, (int64_t)n, 0ull);
	};
	fill(plan.d_pool_u, plan.pool_u);
	fill((unsigned long long *)plan.d_pool_d, plan.pool_d);   // +0.0 has all bits clear
	fill(plan.d_cnt, 2 * (size_t)plan.n_jobs);
	fill((unsigned long long *)plan.d_queue, NQ * MAX_CLASSES / 2);
	HIPCK(ctx, hipGetLastError());
	const int n = (int)plan.launches.size();
	const int n_side = plan.concurrent ? std::min(N_SIDE, n - 1) : 0;
	if (n_side > 0) HIPCK(ctx, hipEventRecord(ctx->fork, ctx->stream));
	for (int k = 0; k < n_side; ++k) HIPCK(ctx, hipStreamWaitEvent(ctx->side[k], ctx->fork, 0));
	for (int k = 0; k < n; ++k) {
		const PlanLaunch &l = plan.launches[k];
		hipStream_t s = (n_side > 0 && k > 0) ? ctx->side[(k - 1) % n_side] : ctx->stream;
		hipLaunchKernelGGL(l.fn, dim3(l.grid), dim3(BLOCK), l.L.total, s, plan.d_jobs, l.items, l.n_items, l.queue,
		                   plan.d_rows, plan.d_slots, plan.d_tab, plan.n_tables, l.L);
		HIPCK(ctx, hipGetLastError());
	}
	for (int k = 0; k < n_side; ++k) {
		HIPCK(ctx, hipEventRecord(ctx->join[k], ctx->side[k]));
		HIPCK(ctx, hipStreamWaitEvent(ctx->stream, ctx->join[k], 0));
	}
	return 0;
}

// Copy a plan's pools and counters back into the jobs' host outputs (the
// cross-region full histograms formed exactly in u64) and fill *timing.
int read_results(twopcf_ctx *ctx, const Plan &plan_, const twopcf_job *jobs, int n_jobs, bool built, double t_start,
                 twopcf_timing *timing) {
	const Plan *plan = &plan_;
	const size_t pool_u = plan->pool_u, pool_d = plan->pool_d;
	std::vector<unsigned long long> hu(pool_u), hcnt(2 * (size_t)n_jobs);
	std::vector<double> hd(pool_d);
	if (pool_u) HIPCK(ctx, hipMemcpyAsync(hu.data(), plan->d_pool_u, pool_u * 8, hipMemcpyDeviceToHost, ctx->stream));
	if (pool_d) HIPCK(ctx, hipMemcpyAsync(hd.data(), plan->d_pool_d, pool_d * 8, hipMemcpyDeviceToHost, ctx->stream));
	if (n_jobs) HIPCK(ctx, hipMemcpyAsync(hcnt.data(), plan->d_cnt, hcnt.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
	HIPCK(ctx, hipStreamSynchronize(ctx->stream));
	float kms = 0.0f;
	HIPCK(ctx, hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
	uint64_t tot_acc = 0, tot_rc = 0;
	for (int q = 0; q < n_jobs; ++q) {
		const twopcf_job &J = jobs[q];
		const PlanOut &P = plan->outs[q];
		if (P.cross) {
			// exact in u64: sum_r touching[r] = same + 2 cross
			for (size_t b = 0; b < P.full_n; ++b) {
				unsigned long long t = 0;
				for (size_t e = b; e < P.jk_n; e += P.full_n) t += hu[P.jk_off + e];
				J.full[b] = (double)(t - hu[P.full_off + b]);
			}
		} else {
			for (size_t b = 0; b < P.full_n; ++b)
				J.full[b] = P.weighted ? hd[P.full_off + b] : (double)hu[P.full_off + b];
		}
		if (J.jk) {
			if (J.n_jk > 1) {
				for (size_t e = 0; e < P.jk_n; ++e) J.jk[e] = P.weighted ? hd[P.jk_off + e] : (double)hu[P.jk_off + e];
			} else {
				for (size_t b = 0; b < P.full_n; ++b) J.jk[b] = J.full[b];   // every pair touches region 0
			}
		}
		if (J.n_pairs) *J.n_pairs = hcnt[2 * q];
		if (J.n_rechecked) *J.n_rechecked = hcnt[2 * q + 1];
		tot_acc += hcnt[2 * q];
		tot_rc += hcnt[2 * q + 1];
	}
	if (timing) {
		timing->kernel_ms = kms;
		timing->prep_ms = built ? plan->prep_ms : 0.0;
		timing->total_ms = now_ms() - t_start;
		timing->n_launches = (int)plan->launches.size();
		timing->n_work_items = plan->n_items;
		timing->n_pairs = tot_acc;
		timing->n_rechecked = tot_rc;
		timing->n_staged = plan->n_staged;
	}
	return 0;
}

}  // namespace

extern "C" {

int twopcf_count(twopcf_ctx *ctx, const twopcf_job *jobs, int n_jobs, int method, const twopcf_tuning *tuning,
                 twopcf_timing *timing) {
	return guarded(ctx, [&]() -> int {
		const double t_start = now_ms();
		if (!ctx || (!jobs && n_jobs > 0) || n_jobs < 0) return fail(ctx, TWOPCF_E_ARG, "null argument");
		if (method != TWOPCF_PLAIN_FP64 && method != TWOPCF_FP32_RECHECK) return fail(ctx, TWOPCF_E_ARG, "bad method");
		for (int q = 0; q < n_jobs; ++q)
			if (!jobs[q].a || !jobs[q].full) return fail(ctx, TWOPCF_E_ARG, "job needs a catalogue and a full[] output");
		HIPCK(ctx, hipSetDevice(ctx->device));
		// no tuning given: the autotuned choice for this job set, if any
		std::vector<uint64_t> ids;
		const std::string jkey = job_set_key(jobs, n_jobs, method, &ids);
		twopcf_tuning tu{};
		if (tuning) {
			tu = *tuning;
		} else if (method == TWOPCF_FP32_RECHECK) {
			auto t = ctx->tuned.find(jkey);
			if (t != ctx->tuned.end()) tu = t->second;
		}
		tuning = &tu;
		// plan cache key: the job set and the tuning
		std::string key = jkey;
		key.append((const char *)&tu, sizeof tu);
		Plan *plan = nullptr;
		bool built = false;
		auto it = ctx->plans.find(key);
		if (it != ctx->plans.end()) {
			plan = it->second.get();
		} else {
			auto np = std::make_unique<Plan>();
			int rc = build_plan(ctx, jobs, n_jobs, method, tuning, *np);
			if (rc) return rc;
			np->cat_ids = ids;
			// capture the launch sequence once; replays run the whole batch as one graph
			if (!getenv("TWOPCF_NO_GRAPH")) {
				hipGraph_t g = nullptr;
				HIPCK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
				rc = enqueue_plan(ctx, *np);
				hipError_t e = hipStreamEndCapture(ctx->stream, &g);
				if (rc) return rc;
				if (e != hipSuccess) return fail(ctx, TWOPCF_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
				np->graph = g;
				HIPCK(ctx, hipGraphInstantiate(&np->exec, g, nullptr, nullptr, 0));
			}
			plan = np.get();
			ctx->plans[key] = std::move(np);
			built = true;
		}
		HIPCK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
		if (plan->exec) {
			HIPCK(ctx, hipGraphLaunch(plan->exec, ctx->stream));
		} else {
			int rc = enqueue_plan(ctx, *plan);
			if (rc) return rc;
		}
		HIPCK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
		return read_results(ctx, *plan, jobs, n_jobs, built, t_start, timing);
	});
}

// twopcf_count without device allocation: the plan is built for this call
// with its device buffers in the scratch array (through the address
// resolved at context creation, never the array's name), launched without a
// graph, and its outputs read back before returning.  The catalogues must
// already be gridded for the job set's geometry (cached by an earlier
// twopcf_count of the same jobs and tuning).  The output pools and counters
// are cleared by fill kernels on every call; calls that overlap on one GPU
// share the array and race.
int twopcf_count_noalloc(twopcf_ctx *ctx, const twopcf_job *jobs, int n_jobs, int method,
                         const twopcf_tuning *tuning, twopcf_timing *timing) {
	return guarded(ctx, [&]() -> int {
		const double t_start = now_ms();
		if (!ctx || (!jobs && n_jobs > 0) || n_jobs < 0) return fail(ctx, TWOPCF_E_ARG, "null argument");
		if (method != TWOPCF_PLAIN_FP64 && method != TWOPCF_FP32_RECHECK) return fail(ctx, TWOPCF_E_ARG, "bad method");
		for (int q = 0; q < n_jobs; ++q)
			if (!jobs[q].a || !jobs[q].full) return fail(ctx, TWOPCF_E_ARG, "job needs a catalogue and a full[] output");
		HIPCK(ctx, hipSetDevice(ctx->device));
		void *scr = noalloc_scratch_ptr(ctx->device);
		if (!scr) return fail(ctx, TWOPCF_E_ARG, "no scratch array on this device");
		twopcf_tuning tu{};
		if (tuning) {
			tu = *tuning;
		} else if (method == TWOPCF_FP32_RECHECK) {
			auto t = ctx->tuned.find(job_set_key(jobs, n_jobs, method, nullptr));
			if (t != ctx->tuned.end()) tu = t->second;
		}
		Arena arena{(char *)scr, NOALLOC_SCRATCH_BYTES, 0};
		Plan plan;   // owns no device memory and no graph
		int rc = build_plan(ctx, jobs, n_jobs, method, &tu, plan, 1, &arena);
		if (rc) return rc;
		HIPCK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
		if ((rc = enqueue_plan(ctx, plan))) return rc;
		HIPCK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
		return read_results(ctx, plan, jobs, n_jobs, true, t_start, timing);
	});
}

// Autotuning of the batched method: trials on every 8th item of the
// queue, timed with events on the context stream, over cell
// subdivision, then region slots (jobs with regions), then first-side chunk
// size, one knob at a time from the defaults.  The choice is kept in the
// context for this job set and used by twopcf_count when it gets no tuning.
int twopcf_autotune(twopcf_ctx *ctx, const twopcf_job *jobs, int n_jobs, twopcf_tuning *best, double *best_ms) {
	return guarded(ctx, [&]() -> int {
		if (!ctx || !jobs || n_jobs <= 0) return fail(ctx, TWOPCF_E_ARG, "null argument");
		for (int q = 0; q < n_jobs; ++q)
			if (!jobs[q].a || !jobs[q].full) return fail(ctx, TWOPCF_E_ARG, "job needs a catalogue and a full[] output");
		HIPCK(ctx, hipSetDevice(ctx->device));
		const int method = TWOPCF_FP32_RECHECK;
		const int stride = 8;
		bool any_jk = false;
		for (int q = 0; q < n_jobs; ++q) any_jk = any_jk || jobs[q].n_jk > 1;
		auto trial = [&](const twopcf_tuning &t, double *ms) -> int {
			Plan p;
			int rc = build_plan(ctx, jobs, n_jobs, method, &t, p, stride);
			if (rc) return rc;
			double m = INFINITY;
			for (int rep = 0; rep < 2; ++rep) {
				HIPCK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
				if ((rc = enqueue_plan(ctx, p))) return rc;
				HIPCK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
				HIPCK(ctx, hipStreamSynchronize(ctx->stream));
				float k = 0.0f;
				HIPCK(ctx, hipEventElapsedTime(&k, ctx->ev0, ctx->ev1));
				m = std::min(m, (double)k);
			}
			*ms = m;
			return 0;
		};
		twopcf_tuning cur{2, 12, ICHUNK};
		double cur_ms = INFINITY;
		int rc = trial(cur, &cur_ms);
		if (rc) return rc;
		auto try_knob = [&](int32_t twopcf_tuning::*f, std::initializer_list<int> vals) -> int {
			const twopcf_tuning base = cur;
			for (int v : vals) {
				if (v == base.*f) continue;
				twopcf_tuning t = base;
				t.*f = v;
				double ms;
				int r = trial(t, &ms);
				if (r == TWOPCF_E_LIMIT) continue;   // not feasible for these jobs (stencil, LDS)
				if (r) return r;
				if (ms < 0.99 * cur_ms) {   // a change must gain more than 1 % (timing noise)
					cur = t;
					cur_ms = ms;
				}
			}
			return 0;
		};
		ctx->err.clear();
		if ((rc = try_knob(&twopcf_tuning::cell_subdiv, {1, 3}))) return rc;
		if (any_jk && (rc = try_knob(&twopcf_tuning::max_slots, {4, 8, 16}))) return rc;
		if ((rc = try_knob(&twopcf_tuning::i_chunk, {128}))) return rc;
		ctx->tuned[job_set_key(jobs, n_jobs, method, nullptr)] = cur;
		if (best) *best = cur;
		if (best_ms) *best_ms = cur_ms * stride;
		return 0;
	});
}

// ------------------------------------------------------ host-only helpers --

int twopcf_s2_thresholds(int kind, const twopcf_axis *ax1, const twopcf_axis *ax2, float *t) {
	return guarded(nullptr, [&]() -> int {
		if (!ax1 || !t) return TWOPCF_E_ARG;
		AxisF a1((float)ax1->min, (float)ax1->max, ax1->n_bins, (float)ax1->log_base);
		AxisF a2 = ax2 ? AxisF((float)ax2->min, (float)ax2->max, ax2->n_bins, (float)ax2->log_base) : a1;
		if (kind == TWOPCF_SIGMA_PI) {
			float mn, mx;
			range_sq(kind, a1, a2, &mn, &mx);
			t[0] = std::nextafter(mn, INFINITY);
			t[1] = mx;
			return 0;
		}
		if (kind != TWOPCF_MONOPOLE && kind != TWOPCF_S_MU) return TWOPCF_E_ARG;
		std::vector<float> v;
		if (!s2_thresholds(kind, a1, a2, v)) return TWOPCF_E_AXIS;
		std::copy(v.begin(), v.end(), t);
		return 0;
	});
}

int twopcf_axis_thresholds(const twopcf_axis *ax, float *t) {
	return guarded(nullptr, [&]() -> int {
		if (!ax || !t) return TWOPCF_E_ARG;
		AxisF a((float)ax->min, (float)ax->max, ax->n_bins, (float)ax->log_base);
		std::vector<float> v;
		if (!value_thresholds(a, v)) return TWOPCF_E_AXIS;
		std::copy(v.begin(), v.end(), t);
		return 0;
	});
}

int twopcf_comoving_distance(double omega_m, double h, double z_min, double z_max, int64_t n, const double *z,
                             double *d) {
	return guarded(nullptr, [&]() -> int {
		if ((n > 0 && (!z || !d)) || n < 0) return TWOPCF_E_ARG;
		Cosmology c(omega_m, h, z_min, z_max);
		for (int64_t i = 0; i < n; ++i) d[i] = c.comov_dist(z[i]);
		return 0;
	});
}

int twopcf_eq_to_cart(int64_t n, const double *ra, const double *dec, const double *d, double *x, double *y,
                      double *z, double *nx, double *ny, double *nz) {
	if (n < 0 || (n > 0 && (!ra || !dec || !d || !x || !y || !z || !nx || !ny || !nz))) return TWOPCF_E_ARG;
	eq_to_cart(n, ra, dec, d, x, y, z, nx, ny, nz);
	return 0;
}

}  // extern "C"
