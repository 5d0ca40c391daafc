// cpgpu_internal.h -- host-side launchers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "cpgpu_device.h"

namespace CPGPU_NS {

// ---- per-kernel device timing with HIP events on the launch stream ----
struct KernelTimer {
	bool enabled = false;
	bool solve_only = false;   // record only the solve kernels (names starting "solver" or "reverse")
	struct Rec { const char *name; hipEvent_t a, b; };
	std::vector<Rec> pending;
	std::vector<hipEvent_t> pool;
	std::vector<int> open;     // per open region: its record in pending, or -1 when not recorded
	std::map<std::string, std::pair<double, long>> acc;   // name -> (ms, launches)
	hipEvent_t take() {
		if (pool.empty()) { hipEvent_t e; (void)hipEventCreate(&e); return e; }
		hipEvent_t e = pool.back(); pool.pop_back(); return e;
	}
	void begin(const char *name, hipStream_t s) {
		if (!enabled) return;
		if (solve_only && strncmp(name, "solver", 6) != 0 && strncmp(name, "reverse", 7) != 0) {
			open.push_back(-1);
			return;
		}
		Rec r; r.name = name; r.a = take(); r.b = take();
		(void)hipEventRecord(r.a, s);
		open.push_back((int)pending.size());
		pending.push_back(r);
	}
	void end(hipStream_t s) {
		if (!enabled || open.empty()) return;
		const int k = open.back();
		open.pop_back();
		if (k >= 0) (void)hipEventRecord(pending[k].b, s);
	}
	// after a stream sync: fold finished records into the totals; records of
	// another stream that has not reached them yet stay pending
	void collect() {
		std::vector<Rec> keep;
		for (auto &r : pending) {
			if (hipEventQuery(r.b) == hipErrorNotReady) {
				keep.push_back(r);
				continue;
			}
			float ms = 0.0f;
			if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
				auto &x = acc[r.name];
				x.first += ms; x.second += 1;
			}
			pool.push_back(r.a); pool.push_back(r.b);
		}
		(void)hipGetLastError();   // a not-ready query is no error
		pending.swap(keep);
	}
	~KernelTimer() {
		for (auto &r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
		for (auto e : pool) (void)hipEventDestroy(e);
	}
};

#define TIMED(timer, name, stream, ...) do { \
	if (timer) (timer)->begin(name, stream); \
	__VA_ARGS__; \
	if (timer) (timer)->end(stream); \
} while (0)

// ---- radix sort (broadphase.hip) ----
struct RadixSortBuffers {
	uint64_t *keys_alt = nullptr;
	int32_t *vals_alt = nullptr;
	uint32_t *block_hist = nullptr;   // [256][n_blocks]
	uint64_t *orand = nullptr;        // [2]: OR and AND of all keys
	uint32_t *bin_total = nullptr;    // [256] per-digit totals, then their exclusive scan
	int cap = 0;
};
hipError_t radix_sort_reserve(RadixSortBuffers &b, int n);
void radix_sort_free(RadixSortBuffers &b);
// Sorts (keys, vals) ascending by key, stable, over the low key_bits bits:
// *n_dev elements (n_dev == nullptr: n_cap), at most n_cap.  b.orand holds
// the keys' OR and AND (reduced on the device when they were made); digits
// they agree on are copied, not sorted.  No host round trip; the result ends
// in keys/vals.
hipError_t radix_sort_pairs(uint64_t *keys, int32_t *vals, int n_cap, const int32_t *n_dev, int key_bits,
                            RadixSortBuffers &b, hipStream_t stream, KernelTimer *timer);

// ---- LBVH (broadphase.hip) ----
// an internal node as the pair traversal reads it: both children's boxes,
// indices and last leaves (a leaf child's own leaf index), in one record
struct PairNode {
	real4 box[2];
	int32_t child[2];
	int32_t hi[2];
};
struct Lbvh {
	int n = 0;                      // leaves
	int cap = 0;
	uint64_t *keys = nullptr;       // sorted keys
	int32_t *leaf_shape = nullptr;  // sorted leaf -> shape index
	int32_t *child = nullptr;       // [n-1][2] children (>= n-1: leaf n-1+i)
	int32_t *parent = nullptr;      // [2n-1]
	int32_t *range = nullptr;       // [n-1][2] leaf range of internal nodes
	real4 *box = nullptr;         // [2n-1]
	PairNode *node = nullptr;     // [n-1]
	real4 *pre = nullptr;         // [n] refit: union of the chunk's leaf boxes up to each leaf
	real4 *suf = nullptr;         // [n] ... and from each leaf to the chunk's end
	real4 *cbox = nullptr;        // [n / chunk + 1] union of each chunk's leaf boxes
	int32_t *world_root = nullptr;  // [n_worlds] root node per world (-1: none)
	int32_t *world_first = nullptr; // [n_worlds]
	int32_t *world_last = nullptr;  // [n_worlds]
};
hipError_t lbvh_reserve(Lbvh &t, int n, int n_worlds);
void lbvh_free(Lbvh &t);
// Build over the shapes in `shapes` (n entries): keys from bb centres
// quantised in per-world bounds, sorted, Karras hierarchy, boxes refitted.
hipError_t lbvh_build(Lbvh &t, const DevWorld &W, const int32_t *shapes, int n, int n_worlds,
                      RadixSortBuffers &sortbuf, real4 *world_bounds, hipStream_t stream, KernelTimer *timer);

// Candidate pairs of the exact QueryReject predicate: dynamic x dynamic
// (each pair once) and dynamic x static.  Pairs are written oriented for
// cpCollide (lower shape type first, then lower index) into pairs[cap][2];
// the total count (possibly > cap) goes to *count.
hipError_t broadphase_pairs(const Lbvh &dyn, const Lbvh &stat, const DevWorld &W, int2 *pairs, int cap,
                            int32_t *count, hipStream_t stream, KernelTimer *timer);

// ---- spatial queries (query.hip) ----
// the two trees of the last step as the queries walk them
struct QueryTrees {
	int nd, ns;                          // leaves of the dynamic and the static tree
	const int32_t *droot, *sroot;        // per space: root node (-1: none)
	const int32_t *dshape, *sshape;      // leaf -> shape
	const PairNode *dnode, *snode;
	const real4 *dbox, *sbox;
};
// a batch of queries: space index (NULL: 0), points or segment starts a [n][2]
// (boxes [n][4] for box queries), segment ends b [n][2], the per-query scalar
// (max distance or segment radius), filters (NULL: all)
struct QueryIn {
	const int32_t *space;
	const double *a, *b, *scalar;
	const cpgpuShapeFilter *filter;
	const int32_t *shape;          // shape queries: the query shapes (shapes of the world)
};
// nearest point (segment == false) or first along a segment: shape_out [n], info_out [n][5]
hipError_t query_nearest_or_first(const DevWorld &W, const QueryTrees &T, int n, const QueryIn &in, bool segment,
                                  int32_t *shape_out, double *info_out, hipStream_t stream);
// every hit: kind 0 point, 1 segment, 2 box; (query, shape) pairs and their five numbers appended
// up to cap, the total in *count
// kind 3: cpSpaceShapeQuery of in.shape[q]: (query, shape) pairs with counts[k] and sets[k][12]
hipError_t query_all(const DevWorld &W, const QueryTrees &T, int n, const QueryIn &in, int kind, int2 *hits,
                     double *info, int64_t cap, unsigned long long *count, hipStream_t stream,
                     int32_t *n_pts = nullptr, double *sets = nullptr);

#define HIP_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return e_; } while (0)

} // namespace CPGPU_NS
