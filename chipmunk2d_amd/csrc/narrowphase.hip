// narrowphase.hip -- collision and arbiter persistence, one candidate pair per lane.
//
// The reference keeps two caches between steps: the BBTree pairs, which hold
// the GJK warm-start id of every collided pair until one of the two leaves is
// rebuilt (src/cpBBTree.c:69-78,206-223,498-517), and the cachedArbiters
// hash set (src/cpSpaceStep.c:248-251, src/cpHashSet.c), which holds the
// arbiters with their contacts' accumulated impulses for
// collisionPersistence steps (cpSpaceArbiterSetFilter, cpSpaceStep.c:292-325).
// Here both live in one table of pair records keyed by the shape pair:
//   {key, GJK id, step the id was written, arbiter reference}.
// Each step the records of the previous step are hashed (open addressing,
// linear probing), every candidate pair looks its record up, collides,
// and appends a new record; arbiters touched this step go to the solver's
// array, untouched ones within persistence to a cached array.
#include "cpgpu_internal.h"
#include "narrowphase.h"
#include "cpgpu_step.h"

#include <type_traits>

namespace CPGPU_NS {

namespace {

constexpr uint64_t EMPTY_KEY = ~0ull;

__global__ void k_table_clear(uint64_t *tkeys, int32_t *tvals, int tcap)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= tcap) return;
	tkeys[i] = EMPTY_KEY;
	tvals[i] = -1;
}

__global__ void k_table_insert(const PairRec *recs, const int32_t *n_recs_ptr, uint64_t *tkeys, int32_t *tvals,
                               int tmask, uint8_t *seen, int32_t *err)
{
	const int n = *n_recs_ptr;
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
		seen[r] = 0;
		const uint64_t key = recs[r].key;
		uint64_t h = mix64(key) & (uint64_t)tmask;
		bool done = false;
		for (int probe = 0; probe <= tmask; probe++) {
			const unsigned long long prev = atomicCAS((unsigned long long *)&tkeys[h], (unsigned long long)EMPTY_KEY,
			                                          (unsigned long long)key);
			if (prev == EMPTY_KEY || prev == key) { tvals[h] = r; done = true; break; }
			h = (h + 1) & (uint64_t)tmask;
		}
		if (!done) raise_err(err, CPGPU_ERR_TABLE_FULL);
	}
}

__device__ __forceinline__ int table_find(const uint64_t *tkeys, const int32_t *tvals, int tmask, uint64_t key)
{
	uint64_t h = mix64(key) & (uint64_t)tmask;
	for (int probe = 0; probe <= tmask; probe++) {
		const uint64_t k = tkeys[h];
		if (k == key) return tvals[h];
		if (k == EMPTY_KEY) return -1;
		h = (h + 1) & (uint64_t)tmask;
	}
	return -1;
}

// wave-aggregated append: one atomic per wave for the lanes that call it
__device__ __forceinline__ int append_index(int32_t *counter)
{
	const uint64_t active = __ballot(1);
	const int lane = __lane_id();
	const int leader = __ffsll((long long)active) - 1;
	const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	int base = 0;
	if (lane == leader) base = atomicAdd(counter, __popcll(active));
	base = __shfl(base, leader);
	return base + __popcll(active & lt);
}

__device__ __forceinline__ const Arbiter *arb_ref(const StepBuffers &B, int32_t ref)
{
	return (ref & CPGPU_CACHED_REF) ? &B.prev_cached[ref & ~CPGPU_CACHED_REF] : &B.prev_act[ref];
}

// Keep an untouched arbiter for collisionPersistence steps as CACHED
// (cpSpaceArbiterSetFilter, cpSpaceStep.c:292-325).  Returns the new
// reference or -1 when evicted.
__device__ __forceinline__ int32_t carry_arbiter(const StepBuffers &B, const DevWorld &W, int32_t ref)
{
	const Arbiter *old = arb_ref(B, ref);
	const WorldStep ws = W.ws[old->world];
	const uint32_t ticks = ws.stamp - old->stamp;
	// touched the step before, not this one: separate (cpSpaceStep.c:306-310)
	if (ticks == 1u && (old->events & CPGPU_EVENT_SEPARATE) && B.seps) {
		const int k = atomicAdd(B.n_seps, 1);
		if (k < B.cap_seps) {
			B.seps[3 * k] = old->world; B.seps[3 * k + 1] = old->sa; B.seps[3 * k + 2] = old->sb;
		}
	}
	if (ticks >= ws.persistence) return -1;
	const int m = append_index(B.n_cached);
	if (m >= B.cap_cached) { raise_err(W.err, CPGPU_ERR_ARBITERS_OVERFLOW); return -1; }
	// copied in 16- (or 8-) byte pieces straight from the old arbiter (a local copy of the
	// whole record would sit in scratch memory), then marked cached
	using Piece = typename std::conditional<sizeof(Arbiter) % 16 == 0, uint4, uint2>::type;
	static_assert(sizeof(Arbiter) % sizeof(Piece) == 0, "arbiters copy in whole pieces");
	const Piece *src = reinterpret_cast<const Piece *>(old);
	Piece *dst = reinterpret_cast<Piece *>(&B.cached[m]);
#pragma unroll
	for (int i = 0; i < (int)(sizeof(Arbiter) / sizeof(Piece)); i++) dst[i] = src[i];
	B.cached[m].state = CPGPU_ARB_CACHED;
	return m | CPGPU_CACHED_REF;
}

// a sensor's pair or one of two infinite masses: never solved (cpSpaceStep.c:268-272)
__device__ __forceinline__ bool never_solved(const DevWorld &W, int s, int t)
{
	return (W.s_sensor && (W.s_sensor[s] | W.s_sensor[t])) ||
	       (W.b_m[W.s_body[s]] == __builtin_inf() && W.b_m[W.s_body[t]] == __builtin_inf());
}

__device__ __forceinline__ void write_record(const StepBuffers &B, const DevWorld &W, uint64_t key, uint32_t id,
                                             uint32_t id_stamp, int32_t arb)
{
	const int k = append_index(B.n_recs);
	if (k >= B.cap_recs) { raise_err(W.err, CPGPU_ERR_RECORDS_OVERFLOW); return; }
	PairRec r;
	r.key = key; r.id = id; r.id_stamp = id_stamp; r.arb = arb; r._pad = 0;
	B.recs[k] = r;
}

// n_pairs bounds the pairs from above; the count itself is read from n_dev
// (at most n_pairs of them were stored)
__global__ void __launch_bounds__(256) k_narrowphase(const DevWorld W, const StepBuffers B, const int2 *pairs,
                                                     int n_pairs, const int32_t *n_dev)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= min(n_pairs, *n_dev)) return;
	const int2 pr = pairs[i];
	const int ia = pr.x, ib = pr.y;
	const uint64_t key = pair_key(ia, ib);
	const int world = W.s_world[ia];
	const WorldStep ws = W.ws[world];

	// the record of the previous step: warm-start id and arbiter
	uint32_t id = 0;
	int32_t old = -1;
	const int r = table_find(B.tkeys, B.tvals, B.tmask, key);
	if (r >= 0) {
		const PairRec rec = B.prev_recs[r];
		B.seen[r] = 1;
		if (rec.id_stamp >= W.s_rebuild[ia] && rec.id_stamp >= W.s_rebuild[ib]) id = rec.id;
		old = rec.arb;
	}

	const ShapeView sa = load_shape(W, ia);
	const ShapeView sb = load_shape(W, ib);
	CollInfo info;
	collide(sa, sb, id, info, W.err);

	int32_t arb_out = -1;
	if (info.count > 0) {
		// cpArbiterUpdate (cpArbiter.c:356-414) on a new or persisting arbiter
		const int m = append_index(B.n_act);
		if (m >= B.cap_act) {
			raise_err(W.err, CPGPU_ERR_ARBITERS_OVERFLOW);
		} else {
			const Arbiter *po = old >= 0 ? arb_ref(B, old) : nullptr;
			Arbiter a;
			a.key = key;
			a.sa = ia; a.sb = ib;
			a.ba = W.s_body[ia]; a.bb = W.s_body[ib];
			a.world = world;
			a.stamp = ws.stamp;
			a.id_in = id;
			a.events = W.n_handlers > 0 ? (int32_t)pair_events(W, ia, ib) : 0;
			// state: an arbiter solved last step was reset to NORMAL at the start of
			// this step (cpSpaceStep.c:351-360), an ignored one stays ignored while its
			// shapes touch; a cached one becomes FIRST_COLLISION
			if (po == nullptr) a.state = CPGPU_ARB_FIRST_COLLISION;
			else if (po->stamp == ws.stamp - 1u) a.state = po->state == CPGPU_ARB_IGNORE ? CPGPU_ARB_IGNORE : CPGPU_ARB_NORMAL;
			else a.state = CPGPU_ARB_FIRST_COLLISION;
			const real2 pA = W.b_p[a.ba], pB = W.b_p[a.bb];
			const int old_count = po ? po->count : 0;
			for (int c = 0; c < 2; c++) {
				Contact con;
				if (c < info.count) {
					con.r1x = info.p1[c].x - pA.x; con.r1y = info.p1[c].y - pA.y;
					con.r2x = info.p2[c].x - pB.x; con.r2y = info.p2[c].y - pB.y;
					con.hash = info.hash[c];
					con.jnAcc = 0.0; con.jtAcc = 0.0;
					con.src = -1;
					for (int j = 0; j < old_count; j++) {
						if (con.hash == po->c[j].hash) {
							con.jnAcc = po->c[j].jnAcc; con.jtAcc = po->c[j].jtAcc;
							con.src = po->c[j].tape;
						}
					}
				} else {
					con.r1x = con.r1y = con.r2x = con.r2y = 0.0;
					con.hash = 0; con.jnAcc = 0.0; con.jtAcc = 0.0;
					con.src = -1;
				}
				con.tape = -1;
				con.nMass = con.tMass = con.bounce = con.jBias = con.bias = 0.0;
				a.c[c] = con;
			}
			a.count = info.count;
			a.nx = info.n.x; a.ny = info.n.y;
			a.e = W.s_e[ia] * W.s_e[ib];
			a.u = W.s_u[ia] * W.s_u[ib];
			const real2 svA = W.s_sv[ia], svB = W.s_sv[ib];
			const V2 surface_vr = v2(svB.x - svA.x, svB.y - svA.y);
			const V2 svr = vsub(surface_vr, vmult(info.n, vdot(surface_vr, info.n)));
			a.svx = svr.x; a.svy = svr.y;
			// begin (on a first collision) and preSolve are the host's (cpSpaceStep.c:255-266);
			// an ignored arbiter, or one never solved that no callback sees, keeps no contacts
			// (:276-282)
			const bool due = a.state != CPGPU_ARB_IGNORE &&
			                 ((a.events & CPGPU_EVENT_PRE_SOLVE) ||
			                  (a.state == CPGPU_ARB_FIRST_COLLISION && (a.events & CPGPU_EVENT_BEGIN)));
			if (a.state == CPGPU_ARB_IGNORE || (!due && never_solved(W, ia, ib))) {
				a.count = 0;
				if (a.state != CPGPU_ARB_IGNORE) a.state = CPGPU_ARB_NORMAL;
			}
			B.act[m] = a;
			arb_out = m;
			if (due && B.cb) {
				const int k = append_index(B.n_cb);
				if (k < B.cap_act) B.cb[k] = m;
				else raise_err(W.err, CPGPU_ERR_ARBITERS_OVERFLOW);
			}
		}
	} else if (old >= 0) {
		arb_out = carry_arbiter(B, W, old);
	}
	write_record(B, W, key, info.id, ws.stamp, arb_out);

	// cumulative arbiter / contact counts (count is 0, 1 or 2), one atomic per wave
	const unsigned long long na = (unsigned long long)__popcll(__ballot(info.count > 0));
	const unsigned long long nc = na + (unsigned long long)__popcll(__ballot(info.count > 1));
	if (__lane_id() == __ffsll((long long)__ballot(1)) - 1 && na > 0) {
		atomicAdd(&W.totals[0], na);
		atomicAdd(&W.totals[1], nc);
	}
}

// Records of the previous step whose pair was not a candidate this step.
__global__ void k_carry(const DevWorld W, const StepBuffers B)
{
	const int n = *B.n_prev_recs;
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
		if (B.seen[r]) continue;
		const PairRec rec = B.prev_recs[r];
		const int32_t ia = (int32_t)(rec.key >> 32), ib = (int32_t)(rec.key & 0xffffffffu);
		const bool id_ok = rec.id_stamp >= W.s_rebuild[ia] && rec.id_stamp >= W.s_rebuild[ib];
		int32_t arb = -1;
		if (rec.arb >= 0) arb = carry_arbiter(B, W, rec.arb);
		if (arb < 0 && !id_ok) continue;
		write_record(B, W, rec.key, rec.id, rec.id_stamp, arb);
	}
}

// ---- collision handler callbacks (cpSpaceStep.c:253-284) ----

// the due arbiters' default decisions: solved, values unchanged
__global__ void k_callbacks_init(const StepBuffers B, cpgpuCallbackDecision *dec)
{
	const int n = min(*B.n_cb, B.cap_act);
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
		const Arbiter &a = B.act[B.cb[k]];
		cpgpuCallbackDecision d;
		d.slot = k; d.solve = 1; d.ignore = 0; d._pad = 0;
		d.e = (double)a.e; d.u = (double)a.u;
		d.surface_vr[0] = (double)a.svx; d.surface_vr[1] = (double)a.svy;
		dec[k] = d;
	}
}

// the arbiters as the host's callbacks see them
__global__ void k_callbacks_pack(const StepBuffers B, cpgpuCallbackArbiter *out)
{
	const int n = min(*B.n_cb, B.cap_act);
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
		const Arbiter &a = B.act[B.cb[k]];
		cpgpuCallbackArbiter v;
		v.slot = k; v.space = a.world; v.shape_a = a.sa; v.shape_b = a.sb;
		v.state = a.state; v.events = a.events; v.count = a.count; v._pad = 0;
		v.normal[0] = (double)a.nx; v.normal[1] = (double)a.ny;
		v.e = (double)a.e; v.u = (double)a.u;
		v.surface_vr[0] = (double)a.svx; v.surface_vr[1] = (double)a.svy;
		for (int c = 0; c < 2; c++) {
			const Contact &q = a.c[c];
			const bool live = c < a.count;
			v.contact[c][0] = live ? (double)q.r1x : 0.0; v.contact[c][1] = live ? (double)q.r1y : 0.0;
			v.contact[c][2] = live ? (double)q.r2x : 0.0; v.contact[c][3] = live ? (double)q.r2y : 0.0;
			v.contact[c][4] = live ? (double)q.jnAcc : 0.0; v.contact[c][5] = live ? (double)q.jtAcc : 0.0;
			v.hash[c] = live ? q.hash : 0;
		}
		out[k] = v;
	}
}

// what begin and preSolve decided: an ignored arbiter stays ignored; one not solved keeps no
// contacts and, unless ignored, counts as normal (:268-282)
__global__ void k_callbacks_apply(const DevWorld W, const StepBuffers B, const cpgpuCallbackDecision *dec)
{
	const int n = min(*B.n_cb, B.cap_act);
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
		Arbiter &a = B.act[B.cb[k]];
		const cpgpuCallbackDecision d = dec[k];
		if (d.ignore) a.state = CPGPU_ARB_IGNORE;
		if (d.solve && !d.ignore && !never_solved(W, a.sa, a.sb)) {
			a.e = (real)d.e; a.u = (real)d.u;
			a.svx = (real)d.surface_vr[0]; a.svy = (real)d.surface_vr[1];
		} else {
			a.count = 0;
			if (a.state != CPGPU_ARB_IGNORE) a.state = CPGPU_ARB_NORMAL;
		}
	}
}

__global__ void k_fill_i32(int32_t *p, int n, int32_t v)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

// ---- grouping the candidate pairs by shape types ----
// One lane per pair runs the whole cpCollide dispatch; a wave pays for every
// type path any of its lanes takes.  Pairs grouped by (type_a, type_b) before
// the narrowphase fill waves with one path each.  The order of the pairs does
// not change any result: every arbiter is found by its shape pair's key.

__device__ __forceinline__ int pair_class(const DevWorld &W, int2 pr)
{
	return W.s_type[pr.x] * 3 + W.s_type[pr.y];
}

__global__ void __launch_bounds__(256) k_pair_class_count(const DevWorld W, const int2 *pairs, int n,
                                                          const int32_t *n_dev, int32_t *counts)
{
	n = min(n, *n_dev);
	__shared__ int32_t h[PAIR_CLASSES];
	if (threadIdx.x < PAIR_CLASSES) h[threadIdx.x] = 0;
	__syncthreads();
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
		atomicAdd(&h[pair_class(W, pairs[i])], 1);
	__syncthreads();
	if (threadIdx.x < PAIR_CLASSES && h[threadIdx.x]) atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

// counts -> class starts (cursors), the costliest classes (polygon pairs,
// highest class numbers) first, so their long-running waves start first
__global__ void k_pair_class_scan(int32_t *counts)
{
	if (threadIdx.x != 0) return;
	int run = 0;
	for (int c = PAIR_CLASSES - 1; c >= 0; c--) {
		const int k = counts[c];
		counts[PAIR_CLASSES + c] = run;
		run += k;
	}
}

// each workgroup counts its pairs per class in LDS, reserves one range per
// class, then writes its pairs into them
__global__ void __launch_bounds__(256) k_pair_class_scatter(const DevWorld W, const int2 *pairs, int n,
                                                            const int32_t *n_dev, int32_t *cursors, int2 *out)
{
	__shared__ int32_t h[PAIR_CLASSES], base[PAIR_CLASSES];
	n = min(n, *n_dev);
	const int per_block = (n + gridDim.x - 1) / gridDim.x;
	const int lo = blockIdx.x * per_block, hi = min(n, lo + per_block);
	if (threadIdx.x < PAIR_CLASSES) h[threadIdx.x] = 0;
	__syncthreads();
	for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) atomicAdd(&h[pair_class(W, pairs[i])], 1);
	__syncthreads();
	if (threadIdx.x < PAIR_CLASSES) {
		base[threadIdx.x] = h[threadIdx.x] ? atomicAdd(&cursors[threadIdx.x], h[threadIdx.x]) : 0;
		h[threadIdx.x] = 0;
	}
	__syncthreads();
	for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
		const int2 pr = pairs[i];
		const int c = pair_class(W, pr);
		out[base[c] + atomicAdd(&h[c], 1)] = pr;
	}
}

// cpShapesCollide for a list of shape pairs: per pair the count and the
// contact point set (shapes_collide_set)
__global__ void __launch_bounds__(256) k_shapes_collide(const DevWorld W, const int2 *pairs, int n, int32_t *count,
                                                        double *sets)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int2 pr = pairs[i];
	count[i] = shapes_collide_set(W, pr.x, pr.y, sets + 12 * (size_t)i, W.err);
}

inline int blocks_for(int n, int t) { return (n + t - 1) / t; }

} // namespace

hipError_t narrowphase_table(const DevWorld &W, const StepBuffers &B, int n_prev_recs_cap, hipStream_t stream,
                             KernelTimer *timer)
{
	TIMED(timer, "pair_table", stream, {
		k_table_clear<<<blocks_for(B.tmask + 1, 256), 256, 0, stream>>>(B.tkeys, B.tvals, B.tmask + 1);
		if (n_prev_recs_cap > 0)
			k_table_insert<<<std::min(blocks_for(n_prev_recs_cap, 256), 2048), 256, 0, stream>>>(
			    B.prev_recs, B.n_prev_recs, B.tkeys, B.tvals, B.tmask, B.seen, W.err);
	});
	return hipGetLastError();
}

hipError_t narrowphase_step(const DevWorld &W, const StepBuffers &B, const int2 *pairs, const int32_t *n_pairs_dev,
                            int n_pairs_host, int n_prev_recs_host, hipStream_t stream, KernelTimer *timer)
{
	// the step's three counters are consecutive: records, solved arbiters, cached arbiters
	// (cleared by a kernel: the step may be captured in a graph); then the callbacks' two
	k_fill_i32<<<1, 32, 0, stream>>>(B.n_recs, 3, 0);
	if (B.n_cb) k_fill_i32<<<1, 32, 0, stream>>>(B.n_cb, 2, 0);
	if (n_pairs_host > 0) {
		const int g = std::min(blocks_for(n_pairs_host, 256), 1024);
		TIMED(timer, "pair_classes", stream, {
			k_fill_i32<<<1, 2 * PAIR_CLASSES, 0, stream>>>(B.type_counts, 2 * PAIR_CLASSES, 0);
			k_pair_class_count<<<g, 256, 0, stream>>>(W, pairs, n_pairs_host, n_pairs_dev, B.type_counts);
			k_pair_class_scan<<<1, 64, 0, stream>>>(B.type_counts);
			k_pair_class_scatter<<<g, 256, 0, stream>>>(W, pairs, n_pairs_host, n_pairs_dev,
			                                            B.type_counts + PAIR_CLASSES, B.pairs_by_type);
		});
		TIMED(timer, "narrowphase", stream,
		      k_narrowphase<<<blocks_for(n_pairs_host, 256), 256, 0, stream>>>(W, B, B.pairs_by_type, n_pairs_host,
		                                                                      n_pairs_dev));
		HIP_TRY(hipGetLastError());
	}
	if (n_prev_recs_host > 0) {
		TIMED(timer, "arbiter_carry", stream,
		      k_carry<<<blocks_for(n_prev_recs_host, 256), 256, 0, stream>>>(W, B));
		HIP_TRY(hipGetLastError());
	}
	return hipSuccess;
}

hipError_t callbacks_init(const StepBuffers &B, int cap, cpgpuCallbackDecision *dec, hipStream_t stream)
{
	if (cap <= 0) return hipSuccess;
	k_callbacks_init<<<std::min(blocks_for(cap, 256), 1024), 256, 0, stream>>>(B, dec);
	return hipGetLastError();
}

hipError_t callbacks_pack(const StepBuffers &B, int cap, cpgpuCallbackArbiter *out, hipStream_t stream)
{
	if (cap <= 0) return hipSuccess;
	k_callbacks_pack<<<std::min(blocks_for(cap, 256), 1024), 256, 0, stream>>>(B, out);
	return hipGetLastError();
}

hipError_t callbacks_apply(const DevWorld &W, const StepBuffers &B, int cap, const cpgpuCallbackDecision *dec,
                           hipStream_t stream)
{
	if (cap <= 0) return hipSuccess;
	k_callbacks_apply<<<std::min(blocks_for(cap, 256), 1024), 256, 0, stream>>>(W, B, dec);
	return hipGetLastError();
}

hipError_t shapes_collide(const DevWorld &W, const int2 *pairs, int n, int32_t *count, double *sets,
                          hipStream_t stream)
{
	if (n <= 0) return hipSuccess;
	k_shapes_collide<<<blocks_for(n, 256), 256, 0, stream>>>(W, pairs, n, count, sets);
	return hipGetLastError();
}

} // namespace CPGPU_NS
