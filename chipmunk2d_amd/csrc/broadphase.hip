// broadphase.hip -- candidate pairs for the narrowphase.
//
// Replaces the reference's cpBBTree reindex query (src/cpBBTree.c:638-654)
// with a linear BVH rebuilt every step:
//   Morton keys of the shapes' box centres, prefixed by the world index, so
//   every world is one subtree of a single hierarchy;
//   an LSD radix sort whose block-local ranking uses LDS counters and
//   wavefront ballots (stable, 8-bit digits, digits that do not vary skipped);
//   the Karras (2012) radix-tree construction, one internal node per lane;
//   a bottom-up refit of node boxes;
//   one traversal lane per dynamic shape against the dynamic tree (pairs
//   reported once, by the lower leaf) and the static tree.
// Node boxes are unions (min/max) of the shapes' tight boxes, so culling
// with the reference's inclusive cpBBIntersects test loses no pair; every
// reported pair satisfies the exact QueryReject predicate
// (src/cpSpaceStep.c:204-232).
#include "cpgpu_internal.h"

namespace CPGPU_NS {

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 4;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int MORTON_BITS = 16;   // per axis: 65536 cells across a world's extent
constexpr int WORLD_SHIFT = 2 * MORTON_BITS;
constexpr int TRAVERSAL_STACK = 48;

// order-preserving integer image of a box coordinate (as a double in both builds)
__device__ __forceinline__ unsigned long long ord_of(double x)
{
	unsigned long long u = (unsigned long long)__double_as_longlong(x);
	return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double double_of(unsigned long long u)
{
	u = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
	return __longlong_as_double((long long)u);
}

__global__ void k_zero_count(int32_t *count) { *count = 0; }

__global__ void k_bounds_reset(unsigned long long *wb, int n_worlds)
{
	const int w = blockIdx.x * blockDim.x + threadIdx.x;
	if (w >= n_worlds) return;
	wb[4 * w + 0] = ~0ull;
	wb[4 * w + 1] = ~0ull;
	wb[4 * w + 2] = 0ull;
	wb[4 * w + 3] = 0ull;
}

__device__ __forceinline__ V2 box_centre(real4 b) { return v2(0.5 * (b.x + b.z), 0.5 * (b.y + b.w)); }

// Per-world bounds of the box centres (order-preserving integer images of the
// doubles, so atomic min / max give the exact extremes).  Waves and then
// workgroups whose lanes share one world combine before the atomics.
__global__ void __launch_bounds__(256) k_bounds(const DevWorld W, const int32_t *shapes, int n, unsigned long long *wb)
{
	__shared__ int ww[4];
	__shared__ unsigned long long wv[4][4];
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const bool valid = i < n;
	int w = -1;
	unsigned long long x0 = ~0ull, y0 = ~0ull, x1 = 0ull, y1 = 0ull;
	if (valid) {
		const int s = shapes[i];
		const V2 c = box_centre(W.s_bb[s]);
		w = W.s_world[s];
		// a centre that is not finite (a body whose state overflowed) bounds nothing
		if (__builtin_isfinite(c.x) && __builtin_isfinite(c.y)) {
			x0 = x1 = ord_of(c.x);
			y0 = y1 = ord_of(c.y);
		}
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int w_first = __shfl(w, 0);
	const bool uniform = __ballot(valid && w != w_first) == 0ull;
	if (uniform) {
		for (int off = 32; off > 0; off >>= 1) {
			x0 = min(x0, (unsigned long long)__shfl_xor(x0, off));
			y0 = min(y0, (unsigned long long)__shfl_xor(y0, off));
			x1 = max(x1, (unsigned long long)__shfl_xor(x1, off));
			y1 = max(y1, (unsigned long long)__shfl_xor(y1, off));
		}
	}
	if (lane == 0) {
		ww[wave] = uniform ? w_first : -2;
		wv[wave][0] = x0; wv[wave][1] = y0; wv[wave][2] = x1; wv[wave][3] = y1;
	}
	__syncthreads();
	const bool block_uniform = ww[0] >= 0 && ww[1] == ww[0] && ww[2] == ww[0] && ww[3] == ww[0];
	if (block_uniform) {
		if (threadIdx.x == 0) {
			for (int q = 1; q < 4; q++) {
				x0 = min(x0, wv[q][0]); y0 = min(y0, wv[q][1]);
				x1 = max(x1, wv[q][2]); y1 = max(y1, wv[q][3]);
			}
			atomicMin(&wb[4 * w_first + 0], x0);
			atomicMin(&wb[4 * w_first + 1], y0);
			atomicMax(&wb[4 * w_first + 2], x1);
			atomicMax(&wb[4 * w_first + 3], y1);
		}
	} else if (uniform) {
		if (lane == 0 && w_first >= 0) {
			atomicMin(&wb[4 * w_first + 0], x0);
			atomicMin(&wb[4 * w_first + 1], y0);
			atomicMax(&wb[4 * w_first + 2], x1);
			atomicMax(&wb[4 * w_first + 3], y1);
		}
	} else if (valid) {
		atomicMin(&wb[4 * w + 0], x0);
		atomicMin(&wb[4 * w + 1], y0);
		atomicMax(&wb[4 * w + 2], x1);
		atomicMax(&wb[4 * w + 3], y1);
	}
}

__device__ __forceinline__ uint64_t spread22(uint32_t x)
{
	uint64_t v = x & 0x3fffffull;
	v = (v | (v << 16)) & 0x0000ffff0000ffffull;
	v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
	v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
	v = (v | (v << 2)) & 0x3333333333333333ull;
	v = (v | (v << 1)) & 0x5555555555555555ull;
	return v;
}

__device__ __forceinline__ uint32_t quantise(double x, double lo, double hi)
{
	const double scale = (hi > lo) ? (double)((1u << MORTON_BITS) - 1) / (hi - lo) : 0.0;
	double q = (x - lo) * scale;
	if (!(q > 0.0)) q = 0.0;
	if (q > (double)((1u << MORTON_BITS) - 1)) q = (double)((1u << MORTON_BITS) - 1);
	return (uint32_t)q;
}

constexpr int KEYS_BLOCKS = 1024;

// Morton keys prefixed by the world index (grid-stride), and the OR / AND of
// all keys, which tell the radix sort which digits vary: one atomic pair per
// workgroup.
__global__ void __launch_bounds__(256) k_keys(const DevWorld W, const int32_t *shapes, int n,
                                             const unsigned long long *wb, uint64_t *keys, int32_t *vals,
                                             unsigned long long *orand)
{
	__shared__ unsigned long long red[2][4];
	unsigned long long o = 0ull, a = ~0ull;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const int s = shapes[i];
		const int w = W.s_world[s];
		const V2 c = box_centre(W.s_bb[s]);
		const uint32_t qx = quantise(c.x, double_of(wb[4 * w + 0]), double_of(wb[4 * w + 2]));
		const uint32_t qy = quantise(c.y, double_of(wb[4 * w + 1]), double_of(wb[4 * w + 3]));
		const uint64_t key = ((uint64_t)w << WORLD_SHIFT) | spread22(qx) | (spread22(qy) << 1);
		keys[i] = key;
		vals[i] = s;
		o |= key;
		a &= key;
	}
	for (int off = 32; off > 0; off >>= 1) {
		o |= __shfl_xor(o, off);
		a &= __shfl_xor(a, off);
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	if (lane == 0) { red[0][wave] = o; red[1][wave] = a; }
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int q = 1; q < 4; q++) { o |= red[0][q]; a &= red[1][q]; }
		atomicOr(&orand[0], o);
		atomicAnd(&orand[1], a);
	}
}

// ---- radix sort ----
// LSD passes over the bytes the key width can hold; the element count may
// come from the device (n_dev), the launch is sized for the capacity.  A
// pass whose digit is the same for every key (the keys' OR and AND, reduced
// on the device while the keys were made, agree on it) is the identity
// permutation of a stable sort: it becomes a copy, so the ping-pong of the
// buffers stays fixed and no host round trip is needed.

__device__ __forceinline__ bool rs_digit_constant(const unsigned long long *orand, int shift)
{
	return (((orand[0] ^ orand[1]) >> shift) & 255ull) == 0ull;
}

__device__ __forceinline__ int rs_count(const int32_t *n_dev, int n) { return n_dev ? *n_dev : n; }

__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint64_t *keys, const int32_t *n_dev, int n_cap,
                                                      int shift, uint32_t *hist, int nblocks,
                                                      const unsigned long long *orand)
{
	if (rs_digit_constant(orand, shift)) return;
	const int n = rs_count(n_dev, n_cap);
	__shared__ uint32_t h[256];
	const int tid = threadIdx.x;
	h[tid] = 0;
	__syncthreads();
	for (int r = 0; r < RS_ITEMS; r++) {
		const int i = blockIdx.x * RS_TILE + r * RS_THREADS + tid;
		if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1u);
	}
	__syncthreads();
	hist[tid * nblocks + blockIdx.x] = h[tid];
}

// exclusive scan of one value per lane over a 256-lane workgroup; *total gets the sum
__device__ __forceinline__ uint32_t block_exclusive_scan256(uint32_t v, uint32_t *lds4, uint32_t *total)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	uint32_t x = v;
	for (int off = 1; off < 64; off <<= 1) {
		const uint32_t t = __shfl_up(x, off);
		if (lane >= off) x += t;
	}
	if (lane == 63) lds4[wave] = x;
	__syncthreads();
	uint32_t pre = 0, tot = 0;
	for (int q = 0; q < 4; q++) {
		if (q < wave) pre += lds4[q];
		tot += lds4[q];
	}
	__syncthreads();
	*total = tot;
	return pre + x - v;
}

// exclusive scan of each digit's row of block counts (one workgroup per digit);
// the row total goes to bin_total[digit]
__global__ void __launch_bounds__(256) k_rs_scan_rows(uint32_t *hist, int nblocks, uint32_t *bin_total, int shift,
                                                     const unsigned long long *orand)
{
	if (rs_digit_constant(orand, shift)) return;
	__shared__ uint32_t lds4[4];
	uint32_t *row = hist + (size_t)blockIdx.x * nblocks;
	uint32_t run = 0;
	for (int base = 0; base < nblocks; base += 256) {
		const int i = base + threadIdx.x;
		const uint32_t v = i < nblocks ? row[i] : 0u;
		uint32_t tot;
		const uint32_t ex = block_exclusive_scan256(v, lds4, &tot);
		if (i < nblocks) row[i] = run + ex;
		run += tot;
	}
	if (threadIdx.x == 0) bin_total[blockIdx.x] = run;
}

// exclusive scan of the 256 digit totals, in place
__global__ void __launch_bounds__(256) k_rs_scan_bins(uint32_t *bin_total, int shift, const unsigned long long *orand)
{
	if (rs_digit_constant(orand, shift)) return;
	__shared__ uint32_t lds4[4];
	uint32_t tot;
	const uint32_t ex = block_exclusive_scan256(bin_total[threadIdx.x], lds4, &tot);
	bin_total[threadIdx.x] = ex;
}

__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint64_t *kin, const int32_t *vin, uint64_t *kout,
                                                         int32_t *vout, const int32_t *n_dev, int n_cap, int shift,
                                                         const uint32_t *offsets, int nblocks,
                                                         const uint32_t *bin_base, const unsigned long long *orand)
{
	const int n = rs_count(n_dev, n_cap);
	if (rs_digit_constant(orand, shift)) {
		for (int r = 0; r < RS_ITEMS; r++) {
			const int i = blockIdx.x * RS_TILE + r * RS_THREADS + threadIdx.x;
			if (i < n) { kout[i] = kin[i]; vout[i] = vin[i]; }
		}
		return;
	}
	__shared__ uint32_t base[256];
	__shared__ uint32_t run[256];
	__shared__ uint32_t wcnt[RS_THREADS / 64][256];
	const int tid = threadIdx.x;
	const int wave = tid >> 6, lane = tid & 63;
	const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	base[tid] = offsets[tid * nblocks + blockIdx.x] + bin_base[tid];
	run[tid] = 0;
	for (int r = 0; r < RS_ITEMS; r++) {
		const int i = blockIdx.x * RS_TILE + r * RS_THREADS + tid;
		const bool valid = i < n;
		uint64_t k = 0;
		int32_t v = 0;
		uint32_t digit = 0;
		if (valid) { k = kin[i]; v = vin[i]; digit = (uint32_t)((k >> shift) & 255); }
		for (int q = 0; q < RS_THREADS / 64; q++) wcnt[q][tid] = 0;
		__syncthreads();
		// lanes of this wave holding the same digit
		uint64_t peers = __ballot(valid);
		for (int bit = 0; bit < 8; bit++) {
			const bool b = (digit >> bit) & 1;
			const uint64_t m = __ballot(b);
			peers &= b ? m : ~m;
		}
		const uint32_t rank = (uint32_t)__popcll(peers & lt);
		if (valid && rank == 0) wcnt[wave][digit] = (uint32_t)__popcll(peers);
		__syncthreads();
		if (valid) {
			uint32_t pre = run[digit];
			for (int q = 0; q < wave; q++) pre += wcnt[q][digit];
			const uint32_t pos = base[digit] + pre + rank;
			kout[pos] = k;
			vout[pos] = v;
		}
		__syncthreads();
		uint32_t add = 0;
		for (int q = 0; q < RS_THREADS / 64; q++) add += wcnt[q][tid];
		run[tid] += add;
		__syncthreads();
	}
}

// ---- Karras hierarchy ----

__device__ __forceinline__ int kdelta(const uint64_t *k, int n, int i, int j)
{
	if (j < 0 || j >= n) return -1;
	const uint64_t a = k[i], b = k[j];
	if (a == b) return 64 + __clz((unsigned int)(i ^ j));
	return __clzll((long long)(a ^ b));
}

__global__ void k_world_ranges(const uint64_t *keys, int n, int32_t *first, int32_t *last, int32_t *root)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int w = (int)(keys[i] >> WORLD_SHIFT);
	const bool f = (i == 0) || (int)(keys[i - 1] >> WORLD_SHIFT) != w;
	const bool l = (i == n - 1) || (int)(keys[i + 1] >> WORLD_SHIFT) != w;
	if (f) first[w] = i;
	if (l) last[w] = i;
	if (f && l) root[w] = (n - 1) + i;
}

// a build's resets in one launch: no world has a root yet, the sort's key OR
// and AND start at 0 and all ones, the whole tree's root has no parent
__global__ void k_bvh_reset(int32_t *world_root, int n_worlds, unsigned long long *orand, int32_t *parent)
{
	for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < n_worlds; w += gridDim.x * blockDim.x) world_root[w] = -1;
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		orand[0] = 0ull;
		orand[1] = ~0ull;
		parent[0] = -1;
	}
}

__global__ void k_karras(const uint64_t *keys, int n, int32_t *child, int32_t *parent, int32_t *range,
                         const int32_t *first, const int32_t *last, int32_t *root, PairNode *node)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n - 1) return;
	const int d = kdelta(keys, n, i, i + 1) > kdelta(keys, n, i, i - 1) ? 1 : -1;
	const int dmin = kdelta(keys, n, i, i - d);
	int lmax = 2;
	while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
	int l = 0;
	for (int t = lmax >> 1; t >= 1; t >>= 1)
		if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
	const int j = i + l * d;
	const int dnode = kdelta(keys, n, i, j);
	int s = 0, t = l;
	do {
		t = (t + 1) >> 1;
		if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
	} while (t > 1);
	const int gamma = i + s * d + min(d, 0);
	const int lo = min(i, j), hi = max(i, j);
	const int left = (lo == gamma) ? (n - 1) + gamma : gamma;
	const int right = (hi == gamma + 1) ? (n - 1) + gamma + 1 : gamma + 1;
	child[2 * i] = left;
	child[2 * i + 1] = right;
	parent[left] = i;
	parent[right] = i;
	range[2 * i] = lo;
	range[2 * i + 1] = hi;
	// the traversal's record: the left child ends at leaf gamma, the right one at hi (the refit adds the boxes)
	node[i].child[0] = left;
	node[i].child[1] = right;
	node[i].hi[0] = gamma;
	node[i].hi[1] = hi;
	const int w = (int)(keys[lo] >> WORLD_SHIFT);
	if (first[w] == lo && last[w] == hi) root[w] = i;
}

// min and max that drop a NaN operand: a leaf box with a NaN side (a body whose
// state went NaN; such a box overlaps nothing) leaves the boxes of the nodes
// above it as the union of the others, so no other pair is pruned away
__device__ __forceinline__ real min_nn(real a, real b) { return (a < b || b != b) ? a : b; }
__device__ __forceinline__ real max_nn(real a, real b) { return (a > b || b != b) ? a : b; }

__device__ __forceinline__ real4 box_union(real4 a, real4 b)
{
	return make_real4(min_nn(a.x, b.x), min_nn(a.y, b.y), max_nn(a.z, b.z), max_nn(a.w, b.w));
}

// node x's box, also into its parent's traversal record
__device__ __forceinline__ void put_box(real4 *box, PairNode *node, const int32_t *child, const int32_t *parent, int x,
                                        real4 b)
{
	box[x] = b;
	const int p = parent[x];
	if (p >= 0) node[p].box[child[2 * p] == x ? 0 : 1] = b;
}

__device__ __forceinline__ real4 shfl_up4(real4 v, int off)
{
	return make_real4(__shfl_up(v.x, off), __shfl_up(v.y, off), __shfl_up(v.z, off), __shfl_up(v.w, off));
}
__device__ __forceinline__ real4 shfl_down4(real4 v, int off)
{
	return make_real4(__shfl_down(v.x, off), __shfl_down(v.y, off), __shfl_down(v.z, off), __shfl_down(v.w, off));
}

// Box refit in two fence-free passes over chunks of REFIT_CHUNK sorted leaves.
// Pass 1 (one workgroup per chunk): leaf boxes into LDS; the nodes whose leaf
// range lies inside the chunk are refitted bottom-up in LDS (the second lane
// to reach a node combines its children, workgroup-scope ordering); the
// chunk's inclusive prefix and suffix unions of leaf boxes and its total box
// are written out.
constexpr int REFIT_CHUNK = 256;

__global__ void __launch_bounds__(REFIT_CHUNK) k_refit_local(const DevWorld W, int n, const int32_t *leaf_shape,
                                                           const int32_t *child, const int32_t *parent,
                                                           const int32_t *range, real4 *box, PairNode *node,
                                                           real4 *pre, real4 *suf, real4 *cbox)
{
	__shared__ real4 lb[REFIT_CHUNK];
	__shared__ real4 ib[REFIT_CHUNK];
	__shared__ uint32_t vis[REFIT_CHUNK];
	__shared__ real4 wtot[REFIT_CHUNK / 64];
	const int t = threadIdx.x, lo = blockIdx.x * REFIT_CHUNK, i = lo + t;
	const int hi = min(lo + REFIT_CHUNK - 1, n - 1);
	const bool valid = i < n;
	const real inf = __builtin_inf();
	const real4 b = valid ? W.s_bb[leaf_shape[i]] : make_real4(inf, inf, -inf, -inf);
	lb[t] = b;
	vis[t] = 0;
	if (valid) {
		if (n > 1) put_box(box, node, child, parent, (n - 1) + i, b);
		else box[(n - 1) + i] = b;
	}
	const int lane = t & 63, wave = t >> 6;
	real4 pf = b, sf = b;
	for (int off = 1; off < 64; off <<= 1) {
		const real4 u = shfl_up4(pf, off);
		const real4 d = shfl_down4(sf, off);
		if (lane >= off) pf = box_union(pf, u);
		if (lane + off < 64) sf = box_union(sf, d);
	}
	if (lane == 63) wtot[wave] = pf;
	__syncthreads();
	for (int q = 0; q < REFIT_CHUNK / 64; q++) {
		if (q < wave) pf = box_union(pf, wtot[q]);
		if (q > wave) sf = box_union(sf, wtot[q]);
	}
	if (valid) { pre[i] = pf; suf[i] = sf; }
	if (t == REFIT_CHUNK - 1) cbox[blockIdx.x] = pf;
	if (!valid || n < 2) return;
	// a node's children lie inside its leaf range, and Karras node k is an end of its own range
	int p = parent[(n - 1) + i];
	while (p >= 0) {
		if (range[2 * p] < lo || range[2 * p + 1] > hi) break;
		if (__hip_atomic_fetch_add(&vis[p - lo], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u) break;
		const int c0 = child[2 * p], c1 = child[2 * p + 1];
		const real4 b0 = c0 >= n - 1 ? lb[c0 - (n - 1) - lo] : ib[c0 - lo];
		const real4 b1 = c1 >= n - 1 ? lb[c1 - (n - 1) - lo] : ib[c1 - lo];
		const real4 u = box_union(b0, b1);
		ib[p - lo] = u;
		put_box(box, node, child, parent, p, u);
		p = parent[p];
	}
}

// Pass 2: a node spanning chunks gets the union of the suffix of its first
// chunk, the whole chunks between, and the prefix of its last chunk.  Nodes
// above the world roots are never read and are skipped.
__global__ void k_refit_span(const uint64_t *keys, int n, const int32_t *range, const int32_t *child,
                             const int32_t *parent, const real4 *pre, const real4 *suf, const real4 *cbox, real4 *box,
                             PairNode *node)
{
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= n - 1) return;
	const int lo = range[2 * p], hi = range[2 * p + 1];
	const int cl = lo / REFIT_CHUNK, ch = hi / REFIT_CHUNK;
	if (cl == ch) return;
	if ((keys[lo] >> WORLD_SHIFT) != (keys[hi] >> WORLD_SHIFT)) return;
	real4 u = box_union(suf[lo], pre[hi]);
	// the whole chunks between, eight independent loads at a time (the top nodes span hundreds)
	int c = cl + 1;
	for (; c + 8 <= ch; c += 8) {
		real4 b[8];
#pragma unroll
		for (int k = 0; k < 8; k++) b[k] = cbox[c + k];
#pragma unroll
		for (int k = 0; k < 4; k++) b[k] = box_union(b[k], b[k + 4]);
		u = box_union(u, box_union(box_union(b[0], b[1]), box_union(b[2], b[3])));
	}
	for (; c < ch; c++) u = box_union(u, cbox[c]);
	put_box(box, node, child, parent, p, u);
}

// ---- traversal ----

__device__ __forceinline__ bool box_overlap(real4 a, real4 b)
{
	return a.x <= b.z && b.x <= a.z && a.y <= b.w && b.y <= a.w;
}

// QueryReject without the box test (cpSpaceStep.c:219-232), plus the
// both-infinite-mass and sensor rejections of cpSpaceCollideShapes (:268-281):
// with the default handler neither kind of pair is ever solved.
__device__ __forceinline__ bool pair_ok(const DevWorld &W, int s, int t)
{
	const int ba = W.s_body[s], bb = W.s_body[t];
	if (ba == bb) return false;
	const uint64_t ga = W.s_group[s], gb = W.s_group[t];
	if ((ga != 0 && ga == gb) || (W.s_cats[s] & W.s_mask[t]) == 0 || (W.s_cats[t] & W.s_mask[s]) == 0) return false;
	for (int k = W.nc_start[ba]; k < W.nc_start[ba + 1]; k++)
		if (W.nc_list[k] == bb) return false;
	// a pair never solved (cpSpaceStep.c:268-272) matters only to handlers that see it
	if ((W.s_sensor && (W.s_sensor[s] | W.s_sensor[t])) ||
	    (W.b_m[ba] == __builtin_inf() && W.b_m[bb] == __builtin_inf()))
		return W.n_handlers > 0 &&
		       (pair_events(W, s, t) & (CPGPU_EVENT_BEGIN | CPGPU_EVENT_PRE_SOLVE | CPGPU_EVENT_SEPARATE)) != 0;
	return true;
}

// Candidate pairs are staged per wave in LDS and written out PAIR_STAGE at a
// time with one atomic on the global count.  Pairs are found inside divergent
// traversal loops, a few lanes at a time: an atomic per find puts millions
// of same-address atomics per step on one counter.
constexpr int PAIR_STAGE = 128;

struct PairStage {
	int2 *buf;              // this wave's PAIR_STAGE slots
	volatile int32_t *n;    // pairs staged (read by every lane, written by one)
	int2 *pairs;
	int cap;
	int32_t *count;
};

// the active lanes copy the staged pairs out; one atomic reserves their range
__device__ __forceinline__ void stage_flush(const PairStage &st, uint64_t active, int staged)
{
	const int lane = __lane_id();
	const int leader = __ffsll((long long)active) - 1;
	const int n_act = __popcll(active);
	const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	const int rank = __popcll(active & lt);
	int base = 0;
	if (lane == leader) base = atomicAdd(st.count, staged);
	base = __shfl(base, leader);
	for (int i = rank; i < staged; i += n_act)
		if (base + i < st.cap) st.pairs[base + i] = st.buf[i];
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

__device__ __forceinline__ void emit_pair(const DevWorld &W, int s, int t, const PairStage &st)
{
	// orientation of cpCollide (cpCollision.c:707-710); same types: lower index first
	const int ts = W.s_type[s], tt = W.s_type[t];
	int a = s, b = t;
	if (ts > tt || (ts == tt && s > t)) { a = t; b = s; }
	const uint64_t active = __ballot(1);
	const int lane = __lane_id();
	const int leader = __ffsll((long long)active) - 1;
	const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	const int rank = __popcll(active & lt);
	const int n_act = __popcll(active);
	int staged = *st.n;
	if (staged + n_act > PAIR_STAGE) {
		stage_flush(st, active, staged);
		staged = 0;
	}
	st.buf[staged + rank] = make_int2(a, b);
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
	if (lane == leader) *st.n = staged + n_act;
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}


// one dynamic shape's queries: pairs with later leaves of its world's dynamic
// tree and with its world's static tree; a popped node's children come in one
// record (PairNode)
__device__ __forceinline__ void traverse_pairs(const DevWorld &W, int q, int nd, const uint64_t *dkeys,
                                               const int32_t *dshape, const PairNode *dnode, const int32_t *drange,
                                               const real4 *dbox, const int32_t *droot, int ns, const int32_t *sshape,
                                               const PairNode *snode, const real4 *sbox, const int32_t *sroot,
                                               int32_t *stk, const PairStage &st)
{
	const int s = dshape[q];
	const real4 bb = W.s_bb[s];
	const int w = (int)(dkeys[q] >> WORLD_SHIFT);
	int sp = 0;
	// dynamic tree: each pair is reported by its lower leaf, so only leaves after q
	{
		const int root = droot[w];
		if (root >= nd - 1) {
			// a one-leaf world: nothing after q
		} else if (drange[2 * root + 1] > q && box_overlap(bb, dbox[root])) {
			stk[256 * sp++] = root;
		}
	}
	while (sp > 0) {
		const PairNode nd_rec = dnode[stk[256 * --sp]];
		for (int side = 0; side < 2; side++) {
			const int c = nd_rec.child[side];
			if (nd_rec.hi[side] <= q) continue;   // a leaf at or before q, or a subtree ending there
			if (!box_overlap(bb, nd_rec.box[side])) continue;
			if (c >= nd - 1) {
				const int t = dshape[c - (nd - 1)];
				if (pair_ok(W, s, t)) emit_pair(W, s, t, st);
			} else {
				if (sp >= TRAVERSAL_STACK) { raise_err(W.err, CPGPU_ERR_TRAVERSAL_STACK); continue; }
				stk[256 * sp++] = c;
			}
		}
	}
	// static tree
	if (ns > 0 && sroot[w] >= 0) {
		const int root = sroot[w];
		if (root >= ns - 1) {
			const int t = sshape[root - (ns - 1)];
			if (box_overlap(bb, sbox[root]) && pair_ok(W, s, t)) emit_pair(W, s, t, st);
		} else if (box_overlap(bb, sbox[root])) {
			stk[256 * sp++] = root;
		}
		while (sp > 0) {
			const PairNode nd_rec = snode[stk[256 * --sp]];
			for (int side = 0; side < 2; side++) {
				const int c = nd_rec.child[side];
				if (!box_overlap(bb, nd_rec.box[side])) continue;
				if (c >= ns - 1) {
					const int t = sshape[c - (ns - 1)];
					if (pair_ok(W, s, t)) emit_pair(W, s, t, st);
				} else {
					if (sp >= TRAVERSAL_STACK) { raise_err(W.err, CPGPU_ERR_TRAVERSAL_STACK); continue; }
					stk[256 * sp++] = c;
				}
			}
		}
	}
}

__global__ void __launch_bounds__(256) k_pairs(const DevWorld W, int nd, const uint64_t *dkeys,
                                               const int32_t *dshape, const PairNode *dnode,
                                               const int32_t *drange, const real4 *dbox, const int32_t *droot,
                                               int ns, const int32_t *sshape, const PairNode *snode,
                                               const real4 *sbox, const int32_t *sroot, int2 *pairs, int cap,
                                               int32_t *count)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	// per-lane stack in LDS, laid out [depth][lane] so lanes at one depth hit distinct banks
	__shared__ int32_t stk[TRAVERSAL_STACK][256];
	__shared__ int2 stage_buf[256 / 64][PAIR_STAGE];
	__shared__ int32_t stage_n[256 / 64];
	const int lane = threadIdx.x;
	const int wave = threadIdx.x >> 6;
	if (__lane_id() == 0) stage_n[wave] = 0;
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
	const PairStage st{stage_buf[wave], &stage_n[wave], pairs, cap, count};
	if (q < nd) traverse_pairs(W, q, nd, dkeys, dshape, dnode, drange, dbox, droot, ns, sshape, snode, sbox, sroot,
	                           &stk[0][lane], st);
	// every lane of the wave is here again: write out what is left
	const int staged = *st.n;
	if (staged > 0) stage_flush(st, __ballot(1), staged);
}

inline int blocks_for(int n, int t) { return (n + t - 1) / t; }

} // namespace

// ---------------------------------------------------------------------------

hipError_t radix_sort_reserve(RadixSortBuffers &b, int n)
{
	if (n <= b.cap) return hipSuccess;
	radix_sort_free(b);
	int cap = 1024;
	while (cap < n) cap *= 2;
	const int nblocks = blocks_for(cap, RS_TILE);
	// on a failed allocation every buffer is released and cap stays 0, so the
	// next reserve tries again and no kernel is given a stale pointer
	hipError_t e = hipMalloc(&b.keys_alt, sizeof(uint64_t) * cap);
	if (e == hipSuccess) e = hipMalloc(&b.vals_alt, sizeof(int32_t) * cap);
	if (e == hipSuccess) e = hipMalloc(&b.block_hist, sizeof(uint32_t) * 256 * nblocks);
	if (e == hipSuccess) e = hipMalloc(&b.orand, sizeof(uint64_t) * 2);
	if (e == hipSuccess) e = hipMalloc(&b.bin_total, sizeof(uint32_t) * 256);
	if (e != hipSuccess) {
		radix_sort_free(b);
		return e;
	}
	b.cap = cap;
	return hipSuccess;
}

void radix_sort_free(RadixSortBuffers &b)
{
	if (b.keys_alt) (void)hipFree(b.keys_alt);
	if (b.vals_alt) (void)hipFree(b.vals_alt);
	if (b.block_hist) (void)hipFree(b.block_hist);
	if (b.orand) (void)hipFree(b.orand);
	if (b.bin_total) (void)hipFree(b.bin_total);
	b = RadixSortBuffers();
}

hipError_t radix_sort_pairs(uint64_t *keys, int32_t *vals, int n_cap, const int32_t *n_dev, int key_bits,
                            RadixSortBuffers &b, hipStream_t stream, KernelTimer *timer)
{
	if (n_cap <= 1) return hipSuccess;
	const int nblocks = blocks_for(n_cap, RS_TILE);
	const unsigned long long *orand = (const unsigned long long *)b.orand;
	uint64_t *kin = keys, *kout = b.keys_alt;
	int32_t *vin = vals, *vout = b.vals_alt;
	for (int shift = 0; shift < key_bits; shift += 8) {
		TIMED(timer, "radix_sort", stream, {
			k_rs_hist<<<nblocks, RS_THREADS, 0, stream>>>(kin, n_dev, n_cap, shift, b.block_hist, nblocks, orand);
			k_rs_scan_rows<<<256, 256, 0, stream>>>(b.block_hist, nblocks, b.bin_total, shift, orand);
			k_rs_scan_bins<<<1, 256, 0, stream>>>(b.bin_total, shift, orand);
			k_rs_scatter<<<nblocks, RS_THREADS, 0, stream>>>(kin, vin, kout, vout, n_dev, n_cap, shift, b.block_hist,
			                                                 nblocks, b.bin_total, orand);
		});
		HIP_TRY(hipGetLastError());
		uint64_t *tk = kin; kin = kout; kout = tk;
		int32_t *tv = vin; vin = vout; vout = tv;
	}
	if (kin != keys) {
		HIP_TRY(hipMemcpyAsync(keys, kin, sizeof(uint64_t) * n_cap, hipMemcpyDeviceToDevice, stream));
		HIP_TRY(hipMemcpyAsync(vals, vin, sizeof(int32_t) * n_cap, hipMemcpyDeviceToDevice, stream));
	}
	return hipSuccess;
}

hipError_t lbvh_reserve(Lbvh &t, int n, int n_worlds)
{
	if (n <= t.cap && t.world_root) return hipSuccess;
	lbvh_free(t);
	int cap = 1024;
	while (cap < n) cap *= 2;
	hipError_t e = hipMalloc(&t.keys, sizeof(uint64_t) * cap);
	if (e == hipSuccess) e = hipMalloc(&t.leaf_shape, sizeof(int32_t) * cap);
	if (e == hipSuccess) e = hipMalloc(&t.child, sizeof(int32_t) * 2 * cap);
	if (e == hipSuccess) e = hipMalloc(&t.parent, sizeof(int32_t) * 2 * cap);
	if (e == hipSuccess) e = hipMalloc(&t.range, sizeof(int32_t) * 2 * cap);
	if (e == hipSuccess) e = hipMalloc(&t.box, sizeof(real4) * 2 * cap);
	if (e == hipSuccess) e = hipMalloc(&t.node, sizeof(PairNode) * cap);
	if (e == hipSuccess) e = hipMalloc(&t.pre, sizeof(real4) * cap);
	if (e == hipSuccess) e = hipMalloc(&t.suf, sizeof(real4) * cap);
	if (e == hipSuccess) e = hipMalloc(&t.cbox, sizeof(real4) * (cap / REFIT_CHUNK + 1));
	if (e == hipSuccess) e = hipMalloc(&t.world_root, sizeof(int32_t) * n_worlds);
	if (e == hipSuccess) e = hipMalloc(&t.world_first, sizeof(int32_t) * n_worlds);
	if (e == hipSuccess) e = hipMalloc(&t.world_last, sizeof(int32_t) * n_worlds);
	if (e != hipSuccess) {
		lbvh_free(t);   // cap stays 0 and world_root null: the next reserve tries again
		return e;
	}
	t.cap = cap;
	return hipSuccess;
}

void lbvh_free(Lbvh &t)
{
	void *ptrs[] = {t.keys, t.leaf_shape, t.child, t.parent, t.range, t.box, t.node, t.pre, t.suf, t.cbox, t.world_root,
	                t.world_first, t.world_last};
	for (void *p : ptrs) if (p) (void)hipFree(p);
	t = Lbvh();
}

hipError_t lbvh_build(Lbvh &t, const DevWorld &W, const int32_t *shapes, int n, int n_worlds,
                      RadixSortBuffers &sortbuf, real4 *world_bounds, hipStream_t stream, KernelTimer *timer)
{
	t.n = n;
	HIP_TRY(radix_sort_reserve(sortbuf, std::max(n, 1)));   // before the reset: it writes the sort's key OR / AND
	k_bvh_reset<<<std::min(blocks_for(n_worlds, 256), 64), 256, 0, stream>>>(
	    t.world_root, n_worlds, (unsigned long long *)sortbuf.orand, t.parent);
	HIP_TRY(hipGetLastError());
	if (n == 0) return hipSuccess;
	unsigned long long *wb = (unsigned long long *)world_bounds;
	TIMED(timer, "bvh_keys", stream, {
		k_bounds_reset<<<blocks_for(n_worlds, 256), 256, 0, stream>>>(wb, n_worlds);
		k_bounds<<<blocks_for(n, 256), 256, 0, stream>>>(W, shapes, n, wb);
		k_keys<<<min(blocks_for(n, 256), KEYS_BLOCKS), 256, 0, stream>>>(W, shapes, n, wb, t.keys, t.leaf_shape,
		                                                (unsigned long long *)sortbuf.orand);
	});
	HIP_TRY(hipGetLastError());
	// world index above the 2 * MORTON_BITS Morton bits
	int world_bits = 0;
	while ((1 << world_bits) < n_worlds) world_bits++;
	HIP_TRY(radix_sort_pairs(t.keys, t.leaf_shape, n, nullptr, WORLD_SHIFT + world_bits, sortbuf, stream, timer));
	TIMED(timer, "bvh_build", stream, {
		k_world_ranges<<<blocks_for(n, 256), 256, 0, stream>>>(t.keys, n, t.world_first, t.world_last,
		                                                        t.world_root);
		if (n > 1)
			k_karras<<<blocks_for(n - 1, 256), 256, 0, stream>>>(t.keys, n, t.child, t.parent, t.range,
			                                                      t.world_first, t.world_last, t.world_root, t.node);
		k_refit_local<<<blocks_for(n, REFIT_CHUNK), REFIT_CHUNK, 0, stream>>>(W, n, t.leaf_shape, t.child, t.parent,
		                                                                    t.range, t.box, t.node, t.pre, t.suf,
		                                                                    t.cbox);
		if (n > REFIT_CHUNK)
			k_refit_span<<<blocks_for(n - 1, 256), 256, 0, stream>>>(t.keys, n, t.range, t.child, t.parent, t.pre,
			                                                         t.suf, t.cbox, t.box, t.node);
	});
	return hipGetLastError();
}

hipError_t broadphase_pairs(const Lbvh &dyn, const Lbvh &stat, const DevWorld &W, int2 *pairs, int cap,
                            int32_t *count, hipStream_t stream, KernelTimer *timer)
{
	k_zero_count<<<1, 1, 0, stream>>>(count);   // a kernel, not a memset: the step may be captured in a graph
	if (dyn.n == 0) return hipGetLastError();
	TIMED(timer, "bvh_pairs", stream, {
		k_pairs<<<blocks_for(dyn.n, 256), 256, 0, stream>>>(W, dyn.n, dyn.keys, dyn.leaf_shape, dyn.node, dyn.range,
		                                                    dyn.box, dyn.world_root, stat.n, stat.leaf_shape,
		                                                    stat.node, stat.box, stat.world_root, pairs, cap, count);
	});
	return hipGetLastError();
}

} // namespace CPGPU_NS
