// world.hip -- the C ABI of include/cpgpu.h: device state, the step driver,
// the integration and shape-caching kernels, and readback.
//
// Step order follows cpSpaceStep (src/cpSpaceStep.c:335-445):
//   integrate positions (cpBodyUpdatePosition, cpBody.c:511-522)
//   cache shape geometry (cpShapeCacheBB, cpShape.c:210-220) and update the
//     dynamic leaves' fattened boxes (cpBBTree.c:498-517)
//   broadphase (broadphase.hip), narrowphase + arbiter persistence
//     (narrowphase.hip)
//   solver: colouring, pre-steps, velocity integration, cached impulses,
//     iterations (solver.hip)
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "cpgpu_errors.h"
#include "cpgpu_step.h"

#include "cpgpu_adjoint.h"

namespace CPGPU_NS {

namespace {

int fail(int code, const std::string &msg)
{
	cpgpu_last_error = msg;
	return code;
}

#define HIP_OR_FAIL(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
	return fail(CPGPU_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); } while (0)

inline int blocks_for(long n, int t) { return (int)((n + t - 1) / t); }

// the device's pair_key on the host: unordered shape pair, lower index high
inline uint64_t pair_key_host(int32_t a, int32_t b)
{
	const uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
	return ((uint64_t)lo << 32) | hi;
}

template <typename T>
hipError_t dalloc(T *&p, size_t n)
{
	p = nullptr;
	return hipMalloc((void **)&p, sizeof(T) * (n ? n : 1));
}

template <typename T>
hipError_t upload(T *&p, const T *src, size_t n)
{
	hipError_t e = dalloc(p, n);
	if (e != hipSuccess) return e;
	if (n && src) return hipMemcpy(p, src, sizeof(T) * n, hipMemcpyHostToDevice);
	if (n) return hipMemset(p, 0, sizeof(T) * n);
	return hipSuccess;
}

// ---- kernels: integration and shape caching ----

// cpBodyUpdatePosition + SetTransform (cpBody.c:511-522, 347-357)
__global__ void k_update_positions(const DevWorld W, const int32_t *bodies, int n)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n) return;
	const int b = bodies[q];
	const real dt = W.ws[W.b_world[b]].dt;
	BodyVel bv = W.b_vel[b];
	const real2 p0 = W.b_p[b];
	const V2 p = vadd(v2(p0.x, p0.y), vmult(vadd(v2(bv.vx, bv.vy), v2(bv.vbx, bv.vby)), dt));
	const real a = W.b_a[b] + (bv.w + bv.wb) * dt;
	double sn_d, cs_d;
	cp_sincos(a, sn_d, cs_d);
	const real sn = (real)sn_d, cs = (real)cs_d;
	const real2 c = W.b_cog[b];
	XF T;
	T.a = cs; T.b = sn; T.c = -sn; T.d = cs;
	T.tx = p.x - (c.x * cs - c.y * sn);
	T.ty = p.y - (c.x * sn + c.y * cs);
	W.b_p[b] = make_real2(p.x, p.y);
	W.b_a[b] = a;
	W.b_T[b] = T;
	bv.vbx = real(0.0); bv.vby = real(0.0); bv.wb = real(0.0);
	W.b_vel[b] = bv;
	// the sanity check after the update (cpBody.c:521)
	const real2 f = W.b_f[b];
	if (!body_sane(W.b_m[b], bv.m_inv, bv.i_inv, p.x, p.y, bv.vx, bv.vy, f.x, f.y, a, bv.w, W.b_t[b]))
		raise_body_fault(W.err, CPGPU_FAULT_BODY_STATE, b);
}

__global__ void k_set_joint_bias(Joint *joints, const real *bc, int n)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c < n) joints[c].bias_coef = bc[c];
}

// cacheData of the three shape classes (cpShape.c:291-296, 378-405; cpPolyShape.c:39-64)
__device__ void cache_shape(const DevWorld &W, int s)
{
	const XF T = W.b_T[W.s_body[s]];
	const int type = W.s_type[s];
	const real r = W.s_r[s];
	const real *g = W.s_geom + 10 * (size_t)s;
	real *wg = W.s_wgeom + 6 * (size_t)s;
	real4 bb;
	if (type == CPGPU_SHAPE_CIRCLE) {
		const V2 c = xpoint(T, v2(g[0], g[1]));
		wg[0] = c.x; wg[1] = c.y;
		bb = make_real4(c.x - r, c.y - r, c.x + r, c.y + r);
	} else if (type == CPGPU_SHAPE_SEGMENT) {
		const V2 ta = xpoint(T, v2(g[0], g[1]));
		const V2 tb = xpoint(T, v2(g[2], g[3]));
		const V2 tn = xvect(T, v2(g[4], g[5]));
		wg[0] = ta.x; wg[1] = ta.y; wg[2] = tb.x; wg[3] = tb.y; wg[4] = tn.x; wg[5] = tn.y;
		real l, rr, bt, tp;
		if (ta.x < tb.x) { l = ta.x; rr = tb.x; } else { l = tb.x; rr = ta.x; }
		if (ta.y < tb.y) { bt = ta.y; tp = tb.y; } else { bt = tb.y; tp = ta.y; }
		bb = make_real4(l - r, bt - r, rr + r, tp + r);
	} else {
		const int st = W.s_pstart[s], cnt = W.s_pcount[s];
		real l = real(__builtin_inf()), rr = real(-__builtin_inf()), bt = real(__builtin_inf()), tp = real(-__builtin_inf());
		for (int i = 0; i < cnt; i++) {
			const real2 lv = W.pv_local[st + i], ln = W.pn_local[st + i];
			const V2 v = xpoint(T, v2(lv.x, lv.y));
			const V2 nn = xvect(T, v2(ln.x, ln.y));
			W.pv_world[st + i] = make_real2(v.x, v.y);
			W.pn_world[st + i] = make_real2(nn.x, nn.y);
			l = fminc(l, v.x); rr = fmaxc(rr, v.x); bt = fminc(bt, v.y); tp = fmaxc(tp, v.y);
		}
		bb = make_real4(l - r, bt - r, rr + r, tp + r);
	}
	W.s_bb[s] = bb;
}

// GetBB (cpBBTree.c:82-98) with the body velocity (cpSpace.c:106)
__device__ real4 fat_box(const DevWorld &W, int s, real4 bb)
{
	const BodyVel bv = W.b_vel[W.s_body[s]];
	const V2 v = vmult(v2(bv.vx, bv.vy), real(0.1));
	const real x = (bb.z - bb.x) * real(0.1);
	const real y = (bb.w - bb.y) * real(0.1);
	return make_real4(bb.x + fminc(-x, v.x), bb.y + fminc(-y, v.y), bb.z + fmaxc(x, v.x), bb.w + fmaxc(y, v.y));
}

__global__ void k_cache_shapes(const DevWorld W, const int32_t *shapes, int n, int leaf_update, int init_fat)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n) return;
	const int s = shapes[q];
	cache_shape(W, s);
	if (init_fat) {
		W.s_fat[s] = fat_box(W, s, W.s_bb[s]);
	} else if (leaf_update) {
		// LeafUpdate (cpBBTree.c:498-517): rebuild the leaf box when the shape escaped it
		const real4 f = W.s_fat[s], b = W.s_bb[s];
		const bool contains = f.x <= b.x && f.z >= b.z && f.y <= b.y && f.w >= b.w;
		if (!contains) {
			W.s_fat[s] = fat_box(W, s, b);
			W.s_rebuild[s] = W.ws[W.s_world[s]].stamp;
		}
	}
}

// ---- host-side deterministic colouring of the joints (same rule as the device arbiters) ----

uint64_t mix64_host(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

int colour_host(int n_bodies, const std::vector<int> &ba, const std::vector<int> &bb, const std::vector<uint64_t> &prio,
                std::vector<int> &colour)
{
	const size_t n = ba.size();
	colour.assign(n, -1);
	if (n == 0) return 0;
	std::vector<uint64_t> claim(n_bodies, ~0ull);
	std::vector<int> taken(n_bodies, -1);
	size_t left = n;
	int c = 0;
	while (left > 0) {
		for (int r = 0; r < COLOUR_SUBROUNDS; r++) {
			for (size_t k = 0; k < n; k++) {
				if (ba[k] >= 0) claim[ba[k]] = ~0ull;
				if (bb[k] >= 0) claim[bb[k]] = ~0ull;
			}
			auto eligible = [&](size_t k) {
				return colour[k] < 0 && !(ba[k] >= 0 && taken[ba[k]] == c) && !(bb[k] >= 0 && taken[bb[k]] == c);
			};
			for (size_t k = 0; k < n; k++) {
				if (!eligible(k)) continue;
				if (ba[k] >= 0) claim[ba[k]] = std::min(claim[ba[k]], prio[k]);
				if (bb[k] >= 0) claim[bb[k]] = std::min(claim[bb[k]], prio[k]);
			}
			std::vector<size_t> won;
			for (size_t k = 0; k < n; k++) {
				if (!eligible(k)) continue;
				if ((ba[k] < 0 || claim[ba[k]] == prio[k]) && (bb[k] < 0 || claim[bb[k]] == prio[k])) won.push_back(k);
			}
			for (size_t k : won) {
				colour[k] = c;
				if (ba[k] >= 0) taken[ba[k]] = c;
				if (bb[k] >= 0) taken[bb[k]] = c;
				left--;
			}
		}
		c++;
	}
	return c;
}

} // namespace

// ---------------------------------------------------------------------------

struct cpgpuWorld {
	int32_t precision = sizeof(real) == 4 ? CPGPU_FP32 : CPGPU_FP64;   // first: read by the C ABI dispatch
	int device = 0;
	hipStream_t stream = nullptr;
	int nw = 0, nb = 0, ns = 0, nc = 0, npv = 0;
	std::vector<cpgpuSpaceParams> sp;
	std::vector<int32_t> body_type;
	DevWorld W{};
	// owned device arrays (freed at destroy)
	std::vector<void *> owned;
	// index lists
	int32_t *d_dyn_bodies = nullptr;
	int n_dyn_bodies = 0;
	int32_t *d_dyn_shapes = nullptr;
	int n_dyn_shapes = 0;
	int32_t *d_static_shapes = nullptr;
	int n_static_shapes = 0;
	int32_t *d_all_shapes = nullptr;
	WorldStep *d_ws = nullptr;
	WorldStep *h_ws = nullptr;          // pinned
	real *d_joint_bias_coef = nullptr;
	double last_joint_dt = -1.0;
	std::vector<double> joint_error_bias;
	bool other_joints = false;   // a joint the adjoint has no reverse for (pin, groove, rotary ones)
	// broadphase
	Lbvh dyn_tree, static_tree;
	RadixSortBuffers sortbuf;
	real4 *d_world_bounds = nullptr;
	int2 *d_pairs = nullptr;
	int pair_cap = 0;
	int2 *d_pairs_by_type = nullptr;    // the pairs grouped by shape types (StepBuffers::pairs_by_type)
	int pairs_by_type_cap = 0;
	int32_t *d_type_counts = nullptr;
	int32_t *d_counters = nullptr;      // [0] n_pairs
	// persistence (double-buffered)
	PairRec *recs[2] = {nullptr, nullptr};
	int rec_cap[2] = {0, 0};
	Arbiter *act[2] = {nullptr, nullptr};
	int act_cap[2] = {0, 0};
	Arbiter *cached[2] = {nullptr, nullptr};
	int cached_cap[2] = {0, 0};
	int32_t *d_step_counts = nullptr;   // [2][3]: n_recs, n_act, n_cached per slot
	int cur = 0;
	// collision handler callbacks, answered between cpgpuWorldStepBegin and cpgpuWorldStepEnd
	int n_handlers = 0;
	bool in_step = false;                   // a step begun and not ended
	bool zero_step = false;                 // ... with dt 0: nothing to do
	int32_t *d_cb = nullptr;                // [cb_cap] solved-list slots of the due arbiters
	cpgpuCallbackDecision *d_dec = nullptr; // [cb_cap]
	cpgpuCallbackArbiter *d_cbview = nullptr;
	int cb_cap = 0, dec_cap = 0, cbview_cap = 0;
	int32_t *d_seps = nullptr;              // [seps_words / 3][3]
	int seps_words = 0;
	int32_t *d_hcounts = nullptr;           // [0] due arbiters, [1] separations
	int n_cb = 0;
	int step_pairs = 0, step_prev_recs = 0; // bounds of the begun step, for its second half
	uint64_t *tkeys = nullptr;
	int32_t *tvals = nullptr;
	int tcap = 0;
	uint8_t *seen = nullptr;
	int seen_cap = 0;
	// solver
	int32_t *colour = nullptr, *order = nullptr;
	int2 *cbody = nullptr;              // colouring scratch (SolveBuffers::cbody / cprio / ccol)
	unsigned long long *cprio = nullptr;
	int32_t *ccol = nullptr;
	int colour_cap = 0;
	unsigned long long *claim = nullptr;
	int32_t *taken = nullptr;
	int32_t *d_jorder = nullptr, *d_jcol_off = nullptr;
	int32_t *d_jw_order = nullptr, *d_jw_col = nullptr, *d_jw_cbase = nullptr, *d_jw_ncol = nullptr;
	int2 *d_jw_body = nullptr;
	// small / large worlds
	int small_world_max_bodies = 4096;
	uint8_t *d_world_small = nullptr;
	int n_small = 0, n_large_worlds = 0;
	std::vector<int32_t> h_world_nb;      // bodies per world (static included)
	std::vector<uint8_t> h_small;
	// small worlds in size buckets (see WorldBucket), each on its own stream
	int n_buckets = 0;
	WorldBucket buckets[CPGPU_MAX_BUCKETS] = {};
	int32_t *bucket_lists[CPGPU_MAX_BUCKETS] = {};
	hipStream_t aux[CPGPU_MAX_BUCKETS] = {};
	hipEvent_t fork_join[CPGPU_MAX_BUCKETS + 3] = {};   // the solver's fork and joins; the strips' fork and join
	int32_t *d_large_bodies = nullptr;
	int n_large_bodies = 0;
	int max_iterations_large = 0;
	int32_t *arb_by_world = nullptr;
	uint64_t *world_keys = nullptr;
	int32_t *warb_off = nullptr, *large_first = nullptr;
	int n_large_joints = 0;
	// rollout tape for the adjoint (cpgpuWorldTapeBegin)
	struct Tape {
		bool on = false;
		int every = 10;
		int steps = 0;
		Arbiter *arb = nullptr;         // every recorded arbiter: those alive at the start, then per step
		long cap = 0, n = 0;
		std::vector<long> off;          // per step: first recorded arbiter of the step
		std::vector<int> count;         // per step: solved arbiters
		std::vector<std::vector<WorldStep>> ws;
		std::vector<BodyState *> ckpt;  // bodies at steps 0, every, 2 every, ...
		std::vector<Joint *> jckpt;     // joints at the same steps
	} tape;
	// a finished tape's buffers, kept for the next tape: a loop that tapes every
	// rollout allocates only on its first (the sizes depend only on the world)
	std::vector<BodyState *> spare_ckpt;
	std::vector<Joint *> spare_jckpt;
	Arbiter *spare_arb = nullptr;
	long spare_arb_cap = 0;
	std::vector<double> h_joint_grads;  // [nc * 8] parameter adjoints of the last backward pass
	int32_t *d_replay_n = nullptr;
	// solve units of the grid set
	int grid_flow = 1;
	int32_t *udone = nullptr;
	int32_t *d_large_index = nullptr, *phist = nullptr, *bpart = nullptr, *bslot = nullptr, *pcount = nullptr, *sbody = nullptr;
	int32_t *pcol = nullptr;
	unsigned long long *pbound = nullptr;
	int2 *pinfo = nullptr;
	int n_jcolours = 0;
	std::vector<int32_t> h_jcolour;
	GridBarrier *bar = nullptr;
	int solver_grid = 0;
	int grid_parts = GRID_PARTS_MIN;   // strips of every large world's grid set
	int max_iterations = 0;
	// host mirror of counters after the last sync
	int32_t *h_counts = nullptr;        // pinned: [0..5] step counts, [6] n_pairs, [7] err, [8] first failing body
	int32_t faults_reported = 0;        // the CPGPU_FAULT_* a synchronising call has already returned
	int64_t steps = 0;
	int64_t last_pairs = 0;
	int64_t total_pairs = 0;
	bool pending = false;
	KernelTimer timer;
	// steps captured in HIP graphs (cpgpuWorldSetGraph): one executable graph per arbiter slot
	bool graph_on = false;
	bool graph_dirty = true;
	double graph_dt = 0.0;
	hipGraphExec_t gexec[2] = {nullptr, nullptr};
	hipEvent_t graph_done = nullptr;        // after the last captured step
	double *d_curr_dt = nullptr;            // per world: the previous step's dt
	unsigned long long *d_graph_pairs = nullptr;   // candidate pairs summed over captured steps
	long graph_max_pairs = 0;
	int graph_rec_cap = 0, graph_cached_cap = 0;
	int64_t graph_steps = 0;
	bool dyn_tree_built = false;            // the dynamic tree over the shapes' cached boxes
};

namespace {

template <typename T>
hipError_t own_upload(cpgpuWorld *w, T *&p, const T *src, size_t n)
{
	hipError_t e = upload(p, src, n);
	if (e == hipSuccess) w->owned.push_back((void *)p);
	return e;
}

// Upload n_scalars doubles of the description as the build's real type (D is
// real or real2; both builds read the same double description).
template <typename D>
hipError_t own_upload_real(cpgpuWorld *w, D *&p, const double *src, size_t n_scalars)
{
	std::vector<real> h(n_scalars);
	for (size_t i = 0; i < n_scalars; i++) h[i] = (real)src[i];
	return own_upload(w, p, reinterpret_cast<const D *>(h.data()), n_scalars * sizeof(real) / sizeof(D));
}

// Download n device reals as doubles.
static hipError_t download_real(double *dst, const real *src, size_t n)
{
	std::vector<real> h(n);
	hipError_t e = hipMemcpy(h.data(), src, sizeof(real) * n, hipMemcpyDeviceToHost);
	for (size_t i = 0; e == hipSuccess && i < n; i++) dst[i] = (double)h[i];
	return e;
}

static hipError_t upload_real_into(real *dst, const double *src, size_t n)
{
	std::vector<real> h(n);
	for (size_t i = 0; i < n; i++) h[i] = (real)src[i];
	return hipMemcpy(dst, h.data(), sizeof(real) * n, hipMemcpyHostToDevice);
}

template <typename T>
hipError_t own_alloc(cpgpuWorld *w, T *&p, size_t n)
{
	hipError_t e = dalloc(p, n);
	if (e == hipSuccess) w->owned.push_back((void *)p);
	return e;
}

int pow2_at_least(long n)
{
	long c = 64;
	while (c < n) c *= 2;
	return (int)c;
}

template <typename T>
hipError_t regrow(T *&p, int &cap, long need)
{
	if (need <= cap && p) return hipSuccess;
	if (p) (void)hipFree(p);
	long c = std::max<long>(need + need / 2, 1024);
	cap = (int)c;
	return dalloc(p, (size_t)c);
}

int validate(const cpgpuWorldDesc *d)
{
	if (!d) return fail(CPGPU_ERR_INVALID, "null description");
	if (d->n_worlds < 1) return fail(CPGPU_ERR_INVALID, "n_worlds must be >= 1");
	if (!d->space || !d->world_body_start || !d->world_shape_start)
		return fail(CPGPU_ERR_INVALID, "missing space parameters or world ranges");
	if (d->world_body_start[0] != 0 || d->world_body_start[d->n_worlds] != d->n_bodies)
		return fail(CPGPU_ERR_INVALID, "world_body_start must run from 0 to n_bodies");
	if (d->world_shape_start[0] != 0 || d->world_shape_start[d->n_worlds] != d->n_shapes)
		return fail(CPGPU_ERR_INVALID, "world_shape_start must run from 0 to n_shapes");
	if (d->n_constraints > 0 && (!d->world_constraint_start || d->world_constraint_start[0] != 0 ||
	                             d->world_constraint_start[d->n_worlds] != d->n_constraints))
		return fail(CPGPU_ERR_INVALID, "world_constraint_start must run from 0 to n_constraints");
	for (int w = 0; w < d->n_worlds; w++) {
		if (d->world_body_start[w + 1] < d->world_body_start[w] || d->world_shape_start[w + 1] < d->world_shape_start[w])
			return fail(CPGPU_ERR_INVALID, "world ranges must be non-decreasing");
		if (d->space[w].iterations < 0) return fail(CPGPU_ERR_INVALID, "iterations must be >= 0");
	}
	if (d->n_worlds >= (1 << 20)) return fail(CPGPU_ERR_INVALID, "at most 2^20 - 1 worlds");
	if (d->n_handlers < 0 || (d->n_handlers > 0 && !d->handlers))
		return fail(CPGPU_ERR_INVALID, "bad collision handler list");
	if (d->n_handlers > 256) return fail(CPGPU_ERR_INVALID, "at most 256 collision handlers");
	for (int k = 0; k < d->n_handlers; k++) {
		const cpgpuHandlerDesc &h = d->handlers[k];
		if (h.space < -1 || h.space >= d->n_worlds)
			return fail(CPGPU_ERR_INVALID, "collision handler " + std::to_string(k) + " names space " +
			                                   std::to_string(h.space));
		if (h.type_a == CPGPU_WILDCARD && h.type_b != CPGPU_WILDCARD)
			return fail(CPGPU_ERR_INVALID, "a wildcard handler gives its type first");
	}
	if (d->n_handlers > 0 && d->shape_ctype)
		for (int s = 0; s < d->n_shapes; s++)
			if (d->shape_ctype[s] == CPGPU_WILDCARD)
				return fail(CPGPU_ERR_INVALID, "a shape's collision type is the wildcard");
	std::vector<int> body_world(d->n_bodies);
	for (int w = 0; w < d->n_worlds; w++)
		for (int b = d->world_body_start[w]; b < d->world_body_start[w + 1]; b++) body_world[b] = w;
	for (int b = 0; b < d->n_bodies; b++) {
		const int t = d->body_type[b];
		if (t < 0 || t > 2) return fail(CPGPU_ERR_INVALID, "body type out of range");
	}
	// shapes and worlds walked together: world w owns [world_shape_start[w], world_shape_start[w + 1])
	int sw = 0;
	for (int s = 0; s < d->n_shapes; s++) {
		while (s >= d->world_shape_start[sw + 1]) sw++;
		const int t = d->shape_type[s];
		if (t < 0 || t > 2) return fail(CPGPU_ERR_INVALID, "shape type out of range");
		const int b = d->shape_body[s];
		if (b < 0 || b >= d->n_bodies) return fail(CPGPU_ERR_INVALID, "shape body index out of range");
		if (body_world[b] != sw) return fail(CPGPU_ERR_INVALID, "shape and its body belong to different worlds");
		if (t == CPGPU_SHAPE_POLY) {
			const int c = d->shape_poly_count[s], st = d->shape_poly_start[s];
			if (c < 1 || c > CPGPU_MAX_POLY_VERTS)
				return fail(CPGPU_ERR_INVALID, "polygon vertex count must be in [1, CPGPU_MAX_POLY_VERTS]");
			if (st < 0 || st + c > d->n_poly_verts) return fail(CPGPU_ERR_INVALID, "polygon vertex range out of range");
		}
	}
	for (int c = 0; c < d->n_constraints; c++) {
		const int t = d->cons_type[c];
		if (t < 0 || t >= CPGPU_JOINT_TYPES) return fail(CPGPU_ERR_INVALID, "constraint type out of range");
		if (d->cons_a[c] < 0 || d->cons_a[c] >= d->n_bodies || d->cons_b[c] < 0 || d->cons_b[c] >= d->n_bodies)
			return fail(CPGPU_ERR_INVALID, "constraint body index out of range");
		if (body_world[d->cons_a[c]] != body_world[d->cons_b[c]])
			return fail(CPGPU_ERR_INVALID, "constraint joins bodies of different worlds");
	}
	return CPGPU_OK;
}

// The error word's conditions other than the faults: the call that waited fails.
int check_device_error(cpgpuWorld *w)
{
	const int32_t e = w->h_counts[7] & ~CPGPU_FAULTS;
	if (!e) return CPGPU_OK;
	std::string m = "device error during step:";
	if (e & CPGPU_ERR_PAIRS_OVERFLOW) m += " pair buffer overflow;";
	if (e & CPGPU_ERR_ARBITERS_OVERFLOW) m += " arbiter buffer or colour overflow;";
	if (e & CPGPU_ERR_RECORDS_OVERFLOW) m += " pair-record overflow;";
	if (e & CPGPU_ERR_TRAVERSAL_STACK) m += " BVH traversal stack overflow;";
	if (e & CPGPU_ERR_BARRIER_TIMEOUT) m += " solver grid barrier timeout;";
	if (e & CPGPU_ERR_TABLE_FULL) m += " pair table full;";
	if (e & CPGPU_ERR_EPA_HULL) m += " EPA hull overflow;";
	if (e & CPGPU_ERR_ADJOINT) m += " adjoint: a contact's collision re-run differs from the rollout's;";
	return fail(CPGPU_ERR_DEVICE, m);
}

// The reference's debug-build checks (CPGPU_FAULT_*) raised since the last
// report, as of the last wait: returned once, by a step when it has enqueued
// all its work or by cpgpuWorldSync, never by a reader.
int report_faults(cpgpuWorld *w)
{
	const int32_t e = w->h_counts[7] & CPGPU_FAULTS & ~w->faults_reported;
	if (!e) return CPGPU_OK;
	w->faults_reported |= e;
	std::string m = "the step failed the reference's checks:";
	if (e & CPGPU_FAULT_UNSOLVABLE) m += " Unsolvable collision or constraint (an effective mass of zero);";
	if (e & CPGPU_FAULT_BODY_MASS) m += " Body's mass and moment must be positive to simulate;";
	if (e & CPGPU_FAULT_BODY_STATE) m += " Body's state is invalid (NaN or infinite, or a negative mass or moment);";
	if (e & (CPGPU_FAULT_BODY_MASS | CPGPU_FAULT_BODY_STATE))
		m += " first body: " + std::to_string(0x7fffffff - w->h_counts[8]) + ";";
	return fail(CPGPU_ERR_FAULT, m);
}

// queue a copy of the step counters and the error word into pinned memory
void queue_counters(cpgpuWorld *w)
{
	(void)hipMemcpyAsync(w->h_counts, w->d_step_counts, 6 * sizeof(int32_t), hipMemcpyDeviceToHost, w->stream);
	(void)hipMemcpyAsync(w->h_counts + 6, w->d_counters, sizeof(int32_t), hipMemcpyDeviceToHost, w->stream);
	(void)hipMemcpyAsync(w->h_counts + 7, w->W.err, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, w->stream);
}

int sync_world(cpgpuWorld *w)
{
	queue_counters(w);
	HIP_OR_FAIL(hipStreamSynchronize(w->stream));
	w->timer.collect();
	w->pending = false;
	if (w->graph_on && w->graph_steps > 0) w->last_pairs = w->h_counts[6];   // a captured step's candidate pairs
	return check_device_error(w);
}

// sync_world for the readers behind the C ABI
int wait_world(cpgpuWorld *w)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	return sync_world(w);
}

} // namespace

// ---- the C ABI of this build (dispatch.hip forwards each call) ----

// Distribute the small worlds over the size buckets (first bucket whose
// max_bodies holds the world), largest first within a bucket, and size each
// bucket's LDS for its largest world (worlds beyond the device limit keep
// their bodies in global memory).
static hipError_t build_buckets(cpgpuWorld *w)
{
	const int max_lds = solver_max_lds_bytes();
	std::vector<std::vector<int32_t>> lists(w->n_buckets);
	for (int k = 0; k < w->nw; k++) {
		if (!w->h_small[k]) continue;
		int q = 0;
		while (q < w->n_buckets - 1 && w->h_world_nb[k] > w->buckets[q].max_bodies) q++;
		lists[q].push_back(k);
	}
	for (int q = 0; q < w->n_buckets; q++) {
		auto &L = lists[q];
		std::stable_sort(L.begin(), L.end(), [&](int a, int b) { return w->h_world_nb[a] > w->h_world_nb[b]; });
		if (w->bucket_lists[q]) { (void)hipFree(w->bucket_lists[q]); w->bucket_lists[q] = nullptr; }
		WorldBucket &B = w->buckets[q];
		B.n = (int)L.size();
		B.lds_cap = 0;
		B.worlds = nullptr;
		if (L.empty()) continue;
		HIP_TRY(hipMalloc(&w->bucket_lists[q], sizeof(int32_t) * L.size()));
		HIP_TRY(hipMemcpy(w->bucket_lists[q], L.data(), sizeof(int32_t) * L.size(), hipMemcpyHostToDevice));
		B.worlds = w->bucket_lists[q];
		const int largest = w->h_world_nb[L[0]];
		B.lds_cap = std::min(largest, max_lds / (int)sizeof(BodyVel));
	}
	return hipSuccess;
}

int cpgpuWorldCreate(const cpgpuWorldDesc *d, int32_t precision, cpgpuWorld **out)
{
	if (!out) return fail(CPGPU_ERR_INVALID, "null output pointer");
	*out = nullptr;
	int rc = validate(d);
	if (rc != CPGPU_OK) return rc;
	if (precision != (sizeof(real) == 4 ? CPGPU_FP32 : CPGPU_FP64))
		return fail(CPGPU_ERR_INVALID, "precision must be CPGPU_FP64 or CPGPU_FP32");
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CPGPU_ERR_HIP, "no HIP device available");

	cpgpuWorld *w = new cpgpuWorld();
	w->precision = precision;
	HIP_OR_FAIL(hipGetDevice(&w->device));
	HIP_OR_FAIL(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
	for (int q = 0; q < CPGPU_MAX_BUCKETS; q++) HIP_OR_FAIL(hipStreamCreateWithFlags(&w->aux[q], hipStreamNonBlocking));
	for (int q = 0; q < CPGPU_MAX_BUCKETS + 3; q++)
		HIP_OR_FAIL(hipEventCreateWithFlags(&w->fork_join[q], hipEventDisableTiming));
	w->nw = d->n_worlds; w->nb = d->n_bodies; w->ns = d->n_shapes; w->nc = d->n_constraints;
	w->npv = d->n_poly_verts;
	w->sp.assign(d->space, d->space + d->n_worlds);
	w->body_type.assign(d->body_type, d->body_type + d->n_bodies);
	DevWorld &W = w->W;
	W.n_worlds = w->nw; W.n_bodies = w->nb; W.n_shapes = w->ns; W.n_joints = w->nc;
	const int nb = w->nb, ns = w->ns, nc = w->nc;

	// bodies
	std::vector<int32_t> body_world(nb);
	for (int k = 0; k < w->nw; k++)
		for (int b = d->world_body_start[k]; b < d->world_body_start[k + 1]; b++) body_world[b] = k;
	std::vector<BodyVel> vel(nb);
	std::vector<XF> T(nb);
	for (int b = 0; b < nb; b++) {
		vel[b] = {(real)d->body_v[2 * b], (real)d->body_v[2 * b + 1], (real)d->body_w[b], (real)d->body_v_bias[2 * b],
		          (real)d->body_v_bias[2 * b + 1], (real)d->body_w_bias[b], (real)d->body_m_inv[b],
		          (real)d->body_i_inv[b]};
		const double *t = d->body_transform + 6 * b;
		T[b] = {(real)t[0], (real)t[1], (real)t[2], (real)t[3], (real)t[4], (real)t[5]};
	}
	int32_t *p_btype, *p_bworld;
	real *p_bm;
	HIP_OR_FAIL(own_upload(w, p_btype, d->body_type, nb)); W.b_type = p_btype;
	HIP_OR_FAIL(own_upload(w, p_bworld, body_world.data(), nb)); W.b_world = p_bworld;
	HIP_OR_FAIL(own_upload_real(w, p_bm, d->body_m, nb)); W.b_m = p_bm;
	HIP_OR_FAIL(own_upload_real(w, W.b_p, d->body_p, 2 * (size_t)nb));
	HIP_OR_FAIL(own_upload_real(w, W.b_a, d->body_a, nb));
	HIP_OR_FAIL(own_upload_real(w, W.b_f, d->body_f, 2 * (size_t)nb));
	HIP_OR_FAIL(own_upload_real(w, W.b_t, d->body_t, nb));
	HIP_OR_FAIL(own_upload_real(w, W.b_cog, d->body_cog, 2 * (size_t)nb));
	HIP_OR_FAIL(own_upload(w, W.b_T, T.data(), nb));
	HIP_OR_FAIL(own_upload(w, W.b_vel, vel.data(), nb));

	// shapes
	std::vector<int32_t> shape_world(ns);
	std::vector<uint8_t> shape_static(ns);
	std::vector<int32_t> dyn_shapes, static_shapes, all_shapes(ns);
	for (int s = 0; s < ns; s++) {
		shape_world[s] = body_world[d->shape_body[s]];
		shape_static[s] = d->body_type[d->shape_body[s]] == CPGPU_BODY_STATIC;
		(shape_static[s] ? static_shapes : dyn_shapes).push_back(s);
		all_shapes[s] = s;
	}
	{
		int32_t *pt, *pb, *pw, *pps, *ppc;
		uint64_t *ph, *pg;
		real *pe, *pu, *pr, *pgeom;
		real2 *psv;
		uint32_t *pcat, *pmask;
		uint8_t *pst;
		HIP_OR_FAIL(own_upload(w, pt, d->shape_type, ns)); W.s_type = pt;
		HIP_OR_FAIL(own_upload(w, pb, d->shape_body, ns)); W.s_body = pb;
		HIP_OR_FAIL(own_upload(w, pw, shape_world.data(), ns)); W.s_world = pw;
		HIP_OR_FAIL(own_upload(w, ph, d->shape_hashid, ns)); W.s_hashid = ph;
		HIP_OR_FAIL(own_upload_real(w, pe, d->shape_e, ns)); W.s_e = pe;
		HIP_OR_FAIL(own_upload_real(w, pu, d->shape_u, ns)); W.s_u = pu;
		HIP_OR_FAIL(own_upload_real(w, psv, d->shape_surface_v, 2 * (size_t)ns)); W.s_sv = psv;
		HIP_OR_FAIL(own_upload(w, pg, d->shape_group, ns)); W.s_group = pg;
		HIP_OR_FAIL(own_upload(w, pcat, d->shape_categories, ns)); W.s_cats = pcat;
		HIP_OR_FAIL(own_upload(w, pmask, d->shape_mask, ns)); W.s_mask = pmask;
		W.s_sensor = nullptr;
		if (d->shape_sensor) {
			bool any = false;
			for (int s = 0; s < ns && !any; s++) any = d->shape_sensor[s] != 0;
			if (any) {
				uint8_t *psens;
				HIP_OR_FAIL(own_upload(w, psens, d->shape_sensor, ns));
				W.s_sensor = psens;
			}
		}
		w->n_handlers = (d->n_handlers > 0 && d->handlers) ? d->n_handlers : 0;
		W.n_handlers = w->n_handlers;
		W.handlers = nullptr;
		W.s_ctype = nullptr;
		if (w->n_handlers > 0) {
			cpgpuHandlerDesc *phd;
			HIP_OR_FAIL(own_upload(w, phd, d->handlers, (size_t)w->n_handlers));
			W.handlers = phd;
			std::vector<uint32_t> ct(ns, 0u);
			if (d->shape_ctype) ct.assign(d->shape_ctype, d->shape_ctype + ns);
			uint32_t *pct;
			HIP_OR_FAIL(own_upload(w, pct, ct.data(), (size_t)ns));
			W.s_ctype = pct;
			HIP_OR_FAIL(own_upload(w, w->d_hcounts, (const int32_t *)nullptr, 2));
		}
		HIP_OR_FAIL(own_upload_real(w, pr, d->shape_radius, ns)); W.s_r = pr;
		HIP_OR_FAIL(own_upload_real(w, pgeom, d->shape_geom, 10 * (size_t)ns)); W.s_geom = pgeom;
		std::vector<int32_t> pstart(ns, 0), pcount(ns, 0);
		for (int s = 0; s < ns; s++)
			if (d->shape_type[s] == CPGPU_SHAPE_POLY) { pstart[s] = d->shape_poly_start[s]; pcount[s] = d->shape_poly_count[s]; }
		HIP_OR_FAIL(own_upload(w, pps, pstart.data(), ns)); W.s_pstart = pps;
		HIP_OR_FAIL(own_upload(w, ppc, pcount.data(), ns)); W.s_pcount = ppc;
		HIP_OR_FAIL(own_upload(w, pst, shape_static.data(), ns)); W.s_static = pst;
		real2 *pvl, *pnl;
		HIP_OR_FAIL(own_upload_real(w, pvl, d->poly_verts, 2 * (size_t)w->npv)); W.pv_local = pvl;
		HIP_OR_FAIL(own_upload_real(w, pnl, d->poly_normals, 2 * (size_t)w->npv)); W.pn_local = pnl;
		HIP_OR_FAIL(own_alloc(w, W.pv_world, w->npv));
		HIP_OR_FAIL(own_alloc(w, W.pn_world, w->npv));
		HIP_OR_FAIL(own_alloc(w, W.s_wgeom, 6 * (size_t)ns));
		HIP_OR_FAIL(own_alloc(w, W.s_bb, ns));
		HIP_OR_FAIL(own_alloc(w, W.s_fat, ns));
		HIP_OR_FAIL(own_upload(w, W.s_rebuild, (const uint32_t *)nullptr, ns));
	}
	w->n_dyn_shapes = (int)dyn_shapes.size();
	w->n_static_shapes = (int)static_shapes.size();
	HIP_OR_FAIL(own_upload(w, w->d_dyn_shapes, dyn_shapes.data(), dyn_shapes.size()));
	HIP_OR_FAIL(own_upload(w, w->d_static_shapes, static_shapes.data(), static_shapes.size()));
	HIP_OR_FAIL(own_upload(w, w->d_all_shapes, all_shapes.data(), all_shapes.size()));
	std::vector<int32_t> dyn_bodies;
	for (int b = 0; b < nb; b++) if (d->body_type[b] != CPGPU_BODY_STATIC) dyn_bodies.push_back(b);
	w->n_dyn_bodies = (int)dyn_bodies.size();
	HIP_OR_FAIL(own_upload(w, w->d_dyn_bodies, dyn_bodies.data(), dyn_bodies.size()));

	// no-collide partners (QueryRejectConstraint, cpSpaceStep.c:204-216)
	{
		std::vector<std::vector<int32_t>> part(nb);
		for (int c = 0; c < nc; c++) {
			if (d->cons_collide_bodies[c]) continue;
			part[d->cons_a[c]].push_back(d->cons_b[c]);
			part[d->cons_b[c]].push_back(d->cons_a[c]);
		}
		std::vector<int32_t> start(nb + 1, 0), list;
		for (int b = 0; b < nb; b++) {
			start[b] = (int32_t)list.size();
			list.insert(list.end(), part[b].begin(), part[b].end());
		}
		start[nb] = (int32_t)list.size();
		int32_t *pst, *pl;
		HIP_OR_FAIL(own_upload(w, pst, start.data(), nb + 1)); W.nc_start = pst;
		HIP_OR_FAIL(own_upload(w, pl, list.data(), list.size())); W.nc_list = pl;
		int32_t *pw0, *pb0;
		HIP_OR_FAIL(own_upload(w, pw0, d->world_shape_start, w->nw + 1)); W.w_shape0 = pw0;
		HIP_OR_FAIL(own_upload(w, pb0, d->world_body_start, w->nw + 1)); W.w_body0 = pb0;
	}

	// joints, coloured once
	{
		std::vector<Joint> J(nc);
		std::vector<int> ba(nc), bbv(nc);
		std::vector<uint64_t> prio(nc);
		w->joint_error_bias.resize(nc);
		for (int c = 0; c < nc; c++) {
			Joint &j = J[c];
			memset(&j, 0, sizeof(j));
			j.type = d->cons_type[c]; j.a = d->cons_a[c]; j.b = d->cons_b[c]; j.world = body_world[j.a];
			j.maxForce = (real)d->cons_max_force[c]; j.errorBias = (real)d->cons_error_bias[c];
			j.maxBias = (real)d->cons_max_bias[c];
			for (int k = 0; k < 8; k++) j.P[k] = (real)d->cons_params[8 * c + k];
			const double j0 = d->cons_jacc ? d->cons_jacc[2 * c] : 0.0, j1 = d->cons_jacc ? d->cons_jacc[2 * c + 1] : 0.0;
			j.jAx = (real)j0; j.jAy = (real)j1; j.jAcc = (real)j0; j.jnAcc = (real)j0;
			j.angle = (real)j1;
			if (j.type > CPGPU_JOINT_GEAR) w->other_joints = true;
			w->joint_error_bias[c] = d->cons_error_bias[c];
			ba[c] = d->body_type[j.a] == CPGPU_BODY_DYNAMIC ? j.a : -1;
			bbv[c] = d->body_type[j.b] == CPGPU_BODY_DYNAMIC ? j.b : -1;
			prio[c] = mix64_host((uint64_t)(uint32_t)(c - d->world_constraint_start[j.world]) | 0x8000000000000000ull);
		}
		// worlds are coloured independently (disjoint bodies), so one pass colours them all
		const int max_col = colour_host(nb, ba, bbv, prio, w->h_jcolour);
		(void)max_col;

		// small worlds (one workgroup each) and large worlds (the grid solve)
		std::vector<uint8_t> small(w->nw);
		std::vector<int32_t> large_bodies;
		w->h_world_nb.assign(w->nw, 0);
		w->n_small = 0;
		for (int k = 0; k < w->nw; k++) {
			int ndyn = 0;
			for (int b = d->world_body_start[k]; b < d->world_body_start[k + 1]; b++)
				ndyn += d->body_type[b] != CPGPU_BODY_STATIC;
			small[k] = ndyn <= w->small_world_max_bodies;
			w->n_small += small[k];
			w->h_world_nb[k] = d->world_body_start[k + 1] - d->world_body_start[k];
		}
		w->h_small = small;
		for (int b = 0; b < nb; b++)
			if (d->body_type[b] != CPGPU_BODY_STATIC && !small[body_world[b]]) large_bodies.push_back(b);
		w->n_large_worlds = w->nw - w->n_small;
		w->n_large_bodies = (int)large_bodies.size();
		// one strip count for every large world, from the largest grid set (its dynamic bodies)
		long max_grid = 0;
		for (int k = 0; k < w->nw; k++) {
			if (small[k]) continue;
			long n = 0;
			for (int b = d->world_body_start[k]; b < d->world_body_start[k + 1]; b++) n += d->body_type[b] == CPGPU_BODY_DYNAMIC;
			max_grid = std::max(max_grid, n);
		}
		w->grid_parts = grid_parts_for(max_grid);
		w->max_iterations_large = 0;
		for (int k = 0; k < w->nw; k++)
			if (!small[k]) w->max_iterations_large = std::max(w->max_iterations_large, (int)w->sp[k].iterations);
		HIP_OR_FAIL(own_upload(w, w->d_world_small, small.data(), small.size()));
		const int default_buckets[4][2] = {{64, 64}, {256, 128}, {1024, 256}, {1 << 30, 512}};
		w->n_buckets = 4;
		for (int q = 0; q < 4; q++) {
			w->buckets[q].max_bodies = default_buckets[q][0];
			w->buckets[q].threads = default_buckets[q][1];
		}
		HIP_OR_FAIL(build_buckets(w));
		HIP_OR_FAIL(own_upload(w, w->d_large_bodies, large_bodies.data(), large_bodies.size()));
		std::vector<int32_t> large_index(w->nw, -1);
		for (int k = 0, l = 0; k < w->nw; k++)
			if (!small[k]) large_index[k] = l++;
		HIP_OR_FAIL(own_upload(w, w->d_large_index, large_index.data(), large_index.size()));
		const int n_large = std::max(w->n_large_worlds, 1);
		HIP_OR_FAIL(own_alloc(w, w->pbound, 2 * (size_t)n_large));
		HIP_OR_FAIL(own_alloc(w, w->phist, (size_t)PART_BINS * n_large));
		const int n_units = grid_units(n_large, w->grid_parts);
		HIP_OR_FAIL(own_alloc(w, w->pcol, (size_t)n_units * (CPGPU_MAX_COLOURS + 1)));
		HIP_OR_FAIL(own_alloc(w, w->pinfo, (size_t)n_units));
		HIP_OR_FAIL(own_alloc(w, w->udone, (size_t)n_units));
		HIP_OR_FAIL(own_alloc(w, w->pcount, (size_t)w->grid_parts * n_large));
		HIP_OR_FAIL(own_alloc(w, w->sbody, (size_t)w->grid_parts * n_large * GRID_RES_BODIES));

		// large worlds' joints: by colour
		int ncol_large = 0;
		for (int c = 0; c < nc; c++)
			if (!small[J[c].world]) ncol_large = std::max(ncol_large, w->h_jcolour[c] + 1);
		std::vector<int32_t> jorder, joff(ncol_large + 1, 0);
		for (int c = 0; c < nc; c++)
			if (!small[J[c].world]) joff[w->h_jcolour[c] + 1]++;
		for (int q = 0; q < ncol_large; q++) joff[q + 1] += joff[q];
		jorder.assign(joff[ncol_large], 0);
		{
			std::vector<int32_t> cursor(joff.begin(), joff.end() - 1);
			for (int c = 0; c < nc; c++)
				if (!small[J[c].world]) jorder[cursor[w->h_jcolour[c]]++] = c;
		}
		w->n_jcolours = ncol_large;
		w->n_large_joints = (int)jorder.size();
		// every world's joints: by (world, colour), with per-world colour offsets
		std::vector<int32_t> jw_order(nc), jw_col, jw_cbase(w->nw), jw_ncol(w->nw);
		{
			int pos = 0;
			for (int k = 0; k < w->nw; k++) {
				const int c0 = d->world_constraint_start ? d->world_constraint_start[k] : 0;
				const int c1 = d->world_constraint_start ? d->world_constraint_start[k + 1] : 0;
				int ncol = 0;
				for (int c = c0; c < c1; c++) ncol = std::max(ncol, w->h_jcolour[c] + 1);
				jw_ncol[k] = ncol;
				jw_cbase[k] = (int32_t)jw_col.size();
				std::vector<int32_t> cnt(ncol + 1, 0);
				for (int c = c0; c < c1; c++) cnt[w->h_jcolour[c] + 1]++;
				for (int q = 0; q < ncol; q++) cnt[q + 1] += cnt[q];
				for (int q = 0; q <= ncol; q++) jw_col.push_back(pos + cnt[q]);
				std::vector<int32_t> cursor(cnt.begin(), cnt.end() - 1);
				for (int c = c0; c < c1; c++) jw_order[pos + cursor[w->h_jcolour[c]]++] = c;
				pos += c1 - c0;
			}
		}
		HIP_OR_FAIL(own_upload(w, W.joints, J.data(), nc));
		HIP_OR_FAIL(own_upload(w, w->d_jorder, jorder.data(), jorder.size()));
		HIP_OR_FAIL(own_upload(w, w->d_jcol_off, joff.data(), joff.size()));
		HIP_OR_FAIL(own_upload(w, w->d_jw_order, jw_order.data(), jw_order.size()));
		// beside each, its bodies (~b for a body that is not dynamic): body types are fixed
		std::vector<int2> jw_body(nc);
		for (int p = 0; p < nc; p++) {
			const Joint &j = J[jw_order[p]];
			const bool da = d->body_type[j.a] == CPGPU_BODY_DYNAMIC, db = d->body_type[j.b] == CPGPU_BODY_DYNAMIC;
			jw_body[p] = make_int2(da ? j.a : ~j.a, db ? j.b : ~j.b);
		}
		HIP_OR_FAIL(own_upload(w, w->d_jw_body, jw_body.data(), jw_body.size()));
		HIP_OR_FAIL(own_upload(w, w->d_jw_col, jw_col.data(), jw_col.size()));
		HIP_OR_FAIL(own_upload(w, w->d_jw_cbase, jw_cbase.data(), jw_cbase.size()));
		HIP_OR_FAIL(own_upload(w, w->d_jw_ncol, jw_ncol.data(), jw_ncol.size()));
	}

	// per-world step parameters, counters, error word
	HIP_OR_FAIL(own_alloc(w, w->d_ws, w->nw));
	W.ws = w->d_ws;
	HIP_OR_FAIL(hipHostMalloc((void **)&w->h_ws, sizeof(WorldStep) * w->nw, hipHostMallocDefault));
	HIP_OR_FAIL(hipHostMalloc((void **)&w->h_counts, sizeof(int32_t) * 16, hipHostMallocDefault));
	memset(w->h_counts, 0, sizeof(int32_t) * 16);
	HIP_OR_FAIL(own_upload(w, W.err, (const int32_t *)nullptr, 2));
	HIP_OR_FAIL(own_upload(w, W.totals, (const unsigned long long *)nullptr, 2 + 2 * CPGPU_SOLVE_UNITS));
	HIP_OR_FAIL(own_upload(w, w->d_counters, (const int32_t *)nullptr, 4));
	HIP_OR_FAIL(own_upload(w, w->d_step_counts, (const int32_t *)nullptr, 6));
	// four ordered 64-bit keys per world (k_bounds_reset / k_bounds), whatever the size of
	// real: one real4 in FP64, two in FP32
	HIP_OR_FAIL(own_alloc(w, w->d_world_bounds, (size_t)w->nw * (4 * sizeof(unsigned long long) / sizeof(real4))));
	HIP_OR_FAIL(own_alloc(w, w->d_joint_bias_coef, nc));
	HIP_OR_FAIL(own_upload(w, w->bar, (const GridBarrier *)nullptr, 1));
	HIP_OR_FAIL(own_upload(w, w->warb_off, (const int32_t *)nullptr, w->nw + 2));
	HIP_OR_FAIL(own_upload(w, w->large_first, (const int32_t *)nullptr, 3));
	HIP_OR_FAIL(own_alloc(w, w->bpart, nb));
	HIP_OR_FAIL(own_alloc(w, w->bslot, nb));
	HIP_OR_FAIL(own_alloc(w, w->claim, CLAIM_SETS * 2 * (size_t)nb));
	HIP_OR_FAIL(own_alloc(w, w->taken, CLAIM_SETS * (size_t)nb));
	w->max_iterations = 0;
	for (auto &s : w->sp) w->max_iterations = std::max(w->max_iterations, (int)s.iterations);

	// initial cache of every shape with its body's transform (cpSpaceAddShape, cpSpace.c:433)
	// and the dynamic leaves' first fattened boxes (cpBBTreeInsert, cpBBTree.c:605-616)
	if (ns > 0)
		k_cache_shapes<<<blocks_for(ns, 256), 256, 0, w->stream>>>(W, w->d_all_shapes, ns, 0, 0);
	if (w->n_dyn_shapes > 0)
		k_cache_shapes<<<blocks_for(w->n_dyn_shapes, 256), 256, 0, w->stream>>>(W, w->d_dyn_shapes, w->n_dyn_shapes, 0, 1);
	HIP_OR_FAIL(hipGetLastError());
	// the static tree, built once
	if (w->n_static_shapes > 0) {
		HIP_OR_FAIL(lbvh_reserve(w->static_tree, w->n_static_shapes, w->nw));
		HIP_OR_FAIL(lbvh_build(w->static_tree, W, w->d_static_shapes, w->n_static_shapes, w->nw, w->sortbuf,
		                       w->d_world_bounds, w->stream, nullptr));
	}
	HIP_OR_FAIL(lbvh_reserve(w->dyn_tree, std::max(w->n_dyn_shapes, 1), w->nw));
	HIP_OR_FAIL(hipStreamSynchronize(w->stream));

	// initial capacities
	const long base = std::max<long>(4L * w->n_dyn_shapes, 4096);
	HIP_OR_FAIL(regrow(w->d_pairs, w->pair_cap, base));
	HIP_OR_FAIL(own_alloc(w, w->d_type_counts, 2 * PAIR_CLASSES));
	for (int k = 0; k < 2; k++) {
		HIP_OR_FAIL(regrow(w->recs[k], w->rec_cap[k], 2 * base));
		HIP_OR_FAIL(regrow(w->act[k], w->act_cap[k], base));
		HIP_OR_FAIL(regrow(w->cached[k], w->cached_cap[k], base));
	}
	// the grid solve: one workgroup per column (a strip and the seam to its right) of one large
	// world, as many as are co-resident; few large workgroups keep its barriers cheap
	w->solver_grid = std::min(solver_max_grid(), w->grid_parts);
	if (w->solver_grid < 1) w->solver_grid = 1;
	*out = w;
	return CPGPU_OK;
}

static void tape_clear(cpgpuWorld *w, bool release);

void cpgpuWorldDestroy(cpgpuWorld *w)
{
	if (!w) return;
	(void)hipStreamSynchronize(w->stream);
	for (void *p : w->owned) (void)hipFree(p);
	lbvh_free(w->dyn_tree);
	lbvh_free(w->static_tree);
	radix_sort_free(w->sortbuf);
	void *extra[] = {w->d_pairs, w->d_pairs_by_type, w->recs[0], w->recs[1], w->act[0], w->act[1], w->cached[0], w->cached[1],
	                 w->tkeys, w->tvals, w->seen, w->colour, w->order, w->arb_by_world, w->world_keys,
	                 w->cbody, w->cprio, w->ccol, w->d_cb, w->d_dec, w->d_cbview, w->d_seps};
	for (void *p : extra) if (p) (void)hipFree(p);
	for (int slot = 0; slot < 2; slot++)
		if (w->gexec[slot]) (void)hipGraphExecDestroy(w->gexec[slot]);
	if (w->graph_done) (void)hipEventDestroy(w->graph_done);
	if (w->d_curr_dt) (void)hipFree(w->d_curr_dt);
	if (w->d_graph_pairs) (void)hipFree(w->d_graph_pairs);
	if (w->h_ws) (void)hipHostFree(w->h_ws);
	if (w->h_counts) (void)hipHostFree(w->h_counts);
	tape_clear(w, true);
	if (w->d_replay_n) (void)hipFree(w->d_replay_n);
	for (int q = 0; q < CPGPU_MAX_BUCKETS; q++)
		if (w->bucket_lists[q]) (void)hipFree(w->bucket_lists[q]);
	for (int q = 0; q < CPGPU_MAX_BUCKETS; q++)
		if (w->aux[q]) (void)hipStreamDestroy(w->aux[q]);
	for (int q = 0; q < CPGPU_MAX_BUCKETS + 3; q++)
		if (w->fork_join[q]) (void)hipEventDestroy(w->fork_join[q]);
	(void)hipStreamDestroy(w->stream);
	delete w;
}

// The solver's buffers over an arbiter array (the step's, or a replay's).
static SolveBuffers solve_buffers(cpgpuWorld *w, Arbiter *act, const int32_t *n_act, int32_t *colour, int32_t *order,
                                  int32_t *warb_off)
{
	SolveBuffers S;
	S.act = act; S.n_act = n_act;
	S.colour = colour; S.order = order;
	S.cbody = w->cbody; S.cprio = w->cprio; S.ccol = w->ccol;
	// with collision handlers the arbiters a step does not solve are sorted out of every set
	const bool grouped = w->nw > 1 || w->n_large_worlds > 0 || w->n_handlers > 0;
	S.arb_by_world = grouped ? w->arb_by_world : nullptr;
	S.keys = w->world_keys;
	S.n_large_joints = w->n_large_joints;
	S.warb_off = warb_off; S.large_first = w->large_first;
	S.world_small = w->d_world_small;
	S.claim = w->claim; S.taken = w->taken;
	S.jorder = w->d_jorder; S.jcol_off = w->d_jcol_off; S.n_jcolours = w->n_jcolours;
	S.jw_order = w->d_jw_order; S.jw_body = w->d_jw_body; S.jw_col = w->d_jw_col; S.jw_cbase = w->d_jw_cbase; S.jw_ncol = w->d_jw_ncol;
	S.bar = w->bar;
	S.max_iterations = w->max_iterations_large;
	S.n_bodies = w->nb;
	S.large_bodies = w->d_large_bodies; S.n_large_bodies = w->n_large_bodies;
	S.large_index = w->d_large_index; S.n_large_worlds = w->n_large_worlds; S.parts = w->grid_parts;
	S.pbound = w->pbound; S.phist = w->phist; S.bpart = w->bpart; S.pcol = w->pcol; S.pinfo = w->pinfo;
	S.udone = w->udone; S.grid_flow = w->grid_flow;
	S.n_worlds = w->nw;
	S.bslot = w->bslot; S.pcount = w->pcount; S.sbody = w->sbody;
	S.rec = nullptr; S.rec_stride = 0; S.jrec = nullptr;
	return S;
}

static hipError_t tape_reserve(cpgpuWorld *w, long n)
{
	auto &T = w->tape;
	if (!T.arb && w->spare_arb) {   // the last tape's records buffer (holds nothing yet)
		T.arb = w->spare_arb;
		T.cap = w->spare_arb_cap;
		w->spare_arb = nullptr;
		w->spare_arb_cap = 0;
	}
	if (n <= T.cap) return hipSuccess;
	long cap = std::max<long>(1024, T.cap);
	while (cap < n) cap *= 2;
	Arbiter *p = nullptr;
	HIP_TRY(hipMalloc(&p, sizeof(Arbiter) * (size_t)cap));
	if (T.n > 0) HIP_TRY(hipMemcpyAsync(p, T.arb, sizeof(Arbiter) * (size_t)T.n, hipMemcpyDeviceToDevice, w->stream));
	HIP_TRY(hipStreamSynchronize(w->stream));
	if (T.arb) (void)hipFree(T.arb);
	T.arb = p;
	T.cap = cap;
	return hipSuccess;
}

// end the tape, its buffers kept for the next one (release: freed)
static void tape_clear(cpgpuWorld *w, bool release = false)
{
	auto &T = w->tape;
	if (T.arb && (w->spare_arb || release)) (void)hipFree(T.arb);
	else if (T.arb) { w->spare_arb = T.arb; w->spare_arb_cap = T.cap; }
	w->spare_ckpt.insert(w->spare_ckpt.end(), T.ckpt.begin(), T.ckpt.end());
	w->spare_jckpt.insert(w->spare_jckpt.end(), T.jckpt.begin(), T.jckpt.end());
	T = cpgpuWorld::Tape();
	if (release) {
		if (w->spare_arb) (void)hipFree(w->spare_arb);
		w->spare_arb = nullptr;
		w->spare_arb_cap = 0;
		for (BodyState *c : w->spare_ckpt) (void)hipFree(c);
		for (Joint *c : w->spare_jckpt) (void)hipFree(c);
		w->spare_ckpt.clear();
		w->spare_jckpt.clear();
	}
}

// a checkpoint buffer of n elements, a kept one if there is one
template <class E> static hipError_t take_spare(std::vector<E *> &spare, E *&p, size_t n)
{
	if (!spare.empty()) {
		p = spare.back();
		spare.pop_back();
		return hipSuccess;
	}
	return hipMalloc(&p, sizeof(E) * n);
}

// bias_coef (chipmunk_private.h:264-268) per joint for step dt, libm pow on the
// host (once per distinct errorBias), then scattered into the joints on the device
static hipError_t set_joint_bias(cpgpuWorld *w, double dt)
{
	std::vector<real> bc(w->nc);
	double last_eb = 0.0;
	real last_bc = 0.0;
	bool have = false;
	for (int c = 0; c < w->nc; c++) {
		const double eb = w->joint_error_bias[c];
		if (!have || eb != last_eb) { last_eb = eb; last_bc = (real)(1.0 - pow(eb, dt)); have = true; }
		bc[c] = last_bc;
	}
	HIP_TRY(hipMemcpyAsync(w->d_joint_bias_coef, bc.data(), sizeof(real) * w->nc, hipMemcpyHostToDevice, w->stream));
	k_set_joint_bias<<<blocks_for(w->nc, 256), 256, 0, w->stream>>>(w->W.joints, w->d_joint_bias_coef, w->nc);
	HIP_TRY(hipGetLastError());
	// the host vector must outlive the copy
	HIP_TRY(hipStreamSynchronize(w->stream));
	w->last_joint_dt = dt;
	return hipSuccess;
}

// The per-world scalars of this step (cpSpaceStep.c:341-344, 383-384, 399, 407),
// pow on the host.  advance: this is the step itself (stamp++, dt becomes the
// space's curr_dt); otherwise the scalars a captured step starts from (the
// stamp and dt of the last step, advanced on the device by k_advance_worlds).
static void fill_world_steps(cpgpuWorld *w, double dt, bool advance)
{
	for (int k = 0; k < w->nw; k++) {
		cpgpuSpaceParams &s = w->sp[k];
		const double prev_dt = s.curr_dt;
		if (advance) {
			s.stamp++;
			s.curr_dt = dt;
		}
		WorldStep &ws = w->h_ws[k];
		ws.dt = advance ? dt : prev_dt;
		ws.dt_coef = (prev_dt == 0.0 ? 0.0 : dt / prev_dt);
		ws.damping = pow(s.damping, dt);
		ws.bias_coef = 1.0 - pow(s.collision_bias, dt);
		ws.slop = s.collision_slop;
		ws.gx = s.gravity_x; ws.gy = s.gravity_y;
		ws.iterations = s.iterations;
		ws.stamp = s.stamp;
		ws.persistence = s.collision_persistence;
	}
}

// Integration, shape caching and the broadphase of a step (cpSpaceStep.c:364-372):
// launch sizes depend only on the world, not on this step's counts.
static hipError_t enqueue_front(cpgpuWorld *w, KernelTimer *timer)
{
	DevWorld &W = w->W;
	// integrate positions, cache the dynamic shapes, update leaf boxes
	if (w->n_dyn_bodies > 0)
		TIMED(timer, "integrate_positions", w->stream,
		      k_update_positions<<<blocks_for(w->n_dyn_bodies, 256), 256, 0, w->stream>>>(W, w->d_dyn_bodies,
		                                                                                 w->n_dyn_bodies));
	if (w->n_dyn_shapes > 0)
		TIMED(timer, "cache_shapes", w->stream,
		      k_cache_shapes<<<blocks_for(w->n_dyn_shapes, 256), 256, 0, w->stream>>>(W, w->d_dyn_shapes,
		                                                                             w->n_dyn_shapes, 1, 0));
	HIP_TRY(hipGetLastError());

	// the large worlds' strips from this step's positions, on an auxiliary stream beside the
	// broadphase and the narrowphase
	const int cur = w->cur;
	if (w->n_large_worlds > 0 && w->n_large_bodies > 0) {
		hipEvent_t strips_fork = w->fork_join[CPGPU_MAX_BUCKETS + 1], strips_join = w->fork_join[CPGPU_MAX_BUCKETS + 2];
		hipStream_t aux = w->aux[CPGPU_MAX_BUCKETS - 1];
		HIP_TRY(hipEventRecord(strips_fork, w->stream));
		HIP_TRY(hipStreamWaitEvent(aux, strips_fork, 0));
		const SolveBuffers S0 = solve_buffers(w, w->act[cur], nullptr, w->colour, w->order, w->warb_off);
		HIP_TRY(solver_strips(W, S0, aux, timer));
		HIP_TRY(hipEventRecord(strips_join, aux));
	}

	// broadphase
	HIP_TRY(lbvh_build(w->dyn_tree, W, w->d_dyn_shapes, w->n_dyn_shapes, w->nw, w->sortbuf, w->d_world_bounds,
	                   w->stream, timer));
	return broadphase_pairs(w->dyn_tree, w->static_tree, W, w->d_pairs, w->pair_cap, w->d_counters, w->stream, timer);
}

// The buffers of the step in slot w->cur.
static StepBuffers step_buffers(cpgpuWorld *w)
{
	const int prev = w->cur ^ 1, cur = w->cur;
	StepBuffers B;
	B.tkeys = w->tkeys; B.tvals = w->tvals; B.tmask = w->tcap - 1;
	B.prev_recs = w->recs[prev]; B.n_prev_recs = w->d_step_counts + 3 * prev + 0;
	B.seen = w->seen;
	B.prev_act = w->act[prev]; B.prev_cached = w->cached[prev];
	B.recs = w->recs[cur]; B.n_recs = w->d_step_counts + 3 * cur + 0; B.cap_recs = w->rec_cap[cur];   // n_recs, n_act and
	// n_cached are consecutive (narrowphase_step clears the three at once)
	B.act = w->act[cur]; B.n_act = w->d_step_counts + 3 * cur + 1; B.cap_act = w->act_cap[cur];
	B.cached = w->cached[cur]; B.n_cached = w->d_step_counts + 3 * cur + 2; B.cap_cached = w->cached_cap[cur];
	B.pairs_by_type = w->d_pairs_by_type;
	B.type_counts = w->d_type_counts;
	B.cb = nullptr; B.n_cb = nullptr; B.seps = nullptr; B.n_seps = nullptr; B.cap_seps = 0;
	if (w->n_handlers > 0) {
		B.cb = w->d_cb; B.n_cb = w->d_hcounts;
		B.seps = w->d_seps; B.n_seps = w->d_hcounts + 1; B.cap_seps = w->seps_words / 3;
	}
	return B;
}

// The narrowphase and arbiter persistence of a step (cpSpaceStep.c:372-380):
// up to where the collision handlers' begin and preSolve are due.  n_pairs and
// n_prev_recs bound this step's candidate pairs and last step's records from
// above: the kernels read the counts.
static hipError_t enqueue_narrow(cpgpuWorld *w, int n_pairs, int n_prev_recs, KernelTimer *timer)
{
	DevWorld &W = w->W;
	const StepBuffers B = step_buffers(w);
	HIP_TRY(narrowphase_table(W, B, n_prev_recs, w->stream, timer));
	HIP_TRY(narrowphase_step(W, B, w->d_pairs, w->d_counters, n_pairs, n_prev_recs, w->stream, timer));
	if (w->n_handlers > 0) HIP_TRY(callbacks_init(B, w->act_cap[w->cur], w->d_dec, w->stream));
	return hipSuccess;
}

// The solve of a step (cpSpaceStep.c:382-440), after the handlers' decisions.
static hipError_t enqueue_solve(cpgpuWorld *w, int n_pairs, KernelTimer *timer)
{
	DevWorld &W = w->W;
	const int cur = w->cur;
	const StepBuffers B = step_buffers(w);
	if (w->n_handlers > 0) HIP_TRY(callbacks_apply(W, B, w->act_cap[cur], w->d_dec, w->stream));
	const bool grouped = w->nw > 1 || w->n_large_worlds > 0 || w->n_handlers > 0;
	SolveBuffers S = solve_buffers(w, w->act[cur], B.n_act, w->colour, w->order, w->warb_off);
	if (w->n_large_worlds > 0 && w->n_large_bodies > 0)
		HIP_TRY(hipStreamWaitEvent(w->stream, w->fork_join[CPGPU_MAX_BUCKETS + 2], 0));
	if (grouped) {
		// the solved arbiters by world and grid unit; launches sized for the capacity (at most one solved
		// arbiter per candidate pair), the kernels read the count
		HIP_TRY(solver_group(W, S, std::max(n_pairs, 1), w->nw, w->n_large_worlds > 0, w->sortbuf, w->stream, timer));
	}
	return solver_step(W, S, w->buckets, w->n_buckets, w->n_large_worlds > 0 ? w->solver_grid : 0, w->stream, w->aux,
	                   w->fork_join, timer);
}

// The narrowphase, arbiter persistence and the solve of a step (cpSpaceStep.c:372-440).
static hipError_t enqueue_back(cpgpuWorld *w, int n_pairs, int n_prev_recs, KernelTimer *timer)
{
	HIP_TRY(enqueue_narrow(w, n_pairs, n_prev_recs, timer));
	return enqueue_solve(w, n_pairs, timer);
}

// Buffers of a step's outputs in arbiter slot `slot`: records, solved and
// cached arbiters (with the colouring's buffers and the pairs grouped by type
// beside the solved ones), and the table over last step's prev_recs records.
static hipError_t reserve_step(cpgpuWorld *w, int slot, long recs, long act, long cached, long prev_recs)
{
	HIP_TRY(regrow(w->recs[slot], w->rec_cap[slot], recs));
	HIP_TRY(regrow(w->act[slot], w->act_cap[slot], act));
	HIP_TRY(regrow(w->cached[slot], w->cached_cap[slot], cached));
	const int tcap = pow2_at_least(2L * prev_recs + 64);
	if (tcap > w->tcap) {
		if (w->tkeys) (void)hipFree(w->tkeys);
		if (w->tvals) (void)hipFree(w->tvals);
		w->tkeys = nullptr; w->tvals = nullptr; w->tcap = 0;
		HIP_TRY(dalloc(w->tkeys, tcap));
		HIP_TRY(dalloc(w->tvals, tcap));
		w->tcap = tcap;
	}
	HIP_TRY(regrow(w->seen, w->seen_cap, prev_recs + 1));
	if (w->n_handlers > 0) {
		HIP_TRY(regrow(w->d_cb, w->cb_cap, w->act_cap[slot]));
		HIP_TRY(regrow(w->d_dec, w->dec_cap, w->act_cap[slot]));
		// only the arbiters touched last step can separate: at most its solved list
		HIP_TRY(regrow(w->d_seps, w->seps_words, 3L * (w->act_cap[slot ^ 1] + 1)));
	}
	if (w->act_cap[slot] > w->colour_cap) {
		void *old[] = {w->colour, w->order, w->arb_by_world, w->world_keys, w->cbody, w->cprio, w->ccol};
		for (void *p : old) if (p) (void)hipFree(p);
		w->colour_cap = w->act_cap[slot];
		HIP_TRY(dalloc(w->colour, w->colour_cap));
		HIP_TRY(dalloc(w->order, w->colour_cap));
		HIP_TRY(dalloc(w->arb_by_world, w->colour_cap));
		HIP_TRY(dalloc(w->world_keys, w->colour_cap));
		HIP_TRY(dalloc(w->cbody, w->colour_cap));
		HIP_TRY(dalloc(w->cprio, w->colour_cap));
		HIP_TRY(dalloc(w->ccol, w->colour_cap));
	}
	return regrow(w->d_pairs_by_type, w->pairs_by_type_cap, act);
}

static int graph_step(cpgpuWorld *w, double dt);

// The first half of a step (cpSpaceStep.c:341-380): up to where the collision
// handlers' begin and preSolve are due.
static int step_begin(cpgpuWorld *w, double dt)
{
	DevWorld &W = w->W;
	KernelTimer *timer = w->timer.enabled ? &w->timer : nullptr;

	fill_world_steps(w, dt, true);
	HIP_OR_FAIL(hipMemcpyAsync(w->d_ws, w->h_ws, sizeof(WorldStep) * w->nw, hipMemcpyHostToDevice, w->stream));
	if (w->tape.on) {
		if (w->tape.steps > 0 && w->tape.steps % w->tape.every == 0) {
			BodyState *ck = nullptr;
			HIP_OR_FAIL(take_spare(w->spare_ckpt, ck, (size_t)w->nb));
			w->tape.ckpt.push_back(ck);
			HIP_OR_FAIL(adjoint_state_save(W, ck, w->stream));
		}
		w->tape.ws.emplace_back(w->h_ws, w->h_ws + w->nw);
	}
	if (w->nc > 0 && dt != w->last_joint_dt) HIP_OR_FAIL(set_joint_bias(w, dt));
	if (w->tape.on && w->nc > 0 && w->tape.steps % w->tape.every == 0) {
		// the joints at the checkpoint, with this step's bias coefficients
		Joint *jk = nullptr;
		HIP_OR_FAIL(take_spare(w->spare_jckpt, jk, (size_t)w->nc));
		w->tape.jckpt.push_back(jk);
		HIP_OR_FAIL(hipMemcpyAsync(jk, W.joints, sizeof(Joint) * (size_t)w->nc, hipMemcpyDeviceToDevice, w->stream));
	}

	HIP_OR_FAIL(enqueue_front(w, timer));
	w->dyn_tree_built = true;
	// the pair count, read back to size this step's buffers
	int rc = sync_world(w);
	if (rc != CPGPU_OK) return rc;
	int n_pairs = w->h_counts[6];
	if (n_pairs > w->pair_cap) {
		HIP_OR_FAIL(regrow(w->d_pairs, w->pair_cap, n_pairs));
		HIP_OR_FAIL(broadphase_pairs(w->dyn_tree, w->static_tree, W, w->d_pairs, w->pair_cap, w->d_counters,
		                             w->stream, timer));
		rc = sync_world(w);
		if (rc != CPGPU_OK) return rc;
		n_pairs = w->h_counts[6];
	}
	w->last_pairs = n_pairs;
	w->total_pairs += n_pairs;

	// capacities of this step's outputs from last step's counts
	const int prev = w->cur ^ 1, cur = w->cur;
	const int n_prev_recs = w->steps > 0 ? w->h_counts[3 * prev + 0] : 0;
	const int n_prev_act = w->steps > 0 ? w->h_counts[3 * prev + 1] : 0;
	const int n_prev_cached = w->steps > 0 ? w->h_counts[3 * prev + 2] : 0;
	HIP_OR_FAIL(reserve_step(w, cur, (long)n_pairs + n_prev_recs + 1, (long)n_pairs + 1, (long)n_prev_act + n_prev_cached + 1,
	                         n_prev_recs));
	if (w->steps == 0) HIP_OR_FAIL(hipMemsetAsync(w->d_step_counts + 3 * prev, 0, 3 * sizeof(int32_t), w->stream));
	HIP_OR_FAIL(enqueue_narrow(w, n_pairs, n_prev_recs, timer));
	w->step_pairs = n_pairs;
	w->step_prev_recs = n_prev_recs;
	return CPGPU_OK;
}

// The second half: the solve, with the decisions of the handlers, and the tape.
static int step_end(cpgpuWorld *w)
{
	KernelTimer *timer = w->timer.enabled ? &w->timer : nullptr;
	const int cur = w->cur;
	HIP_OR_FAIL(enqueue_solve(w, w->step_pairs, timer));

	if (w->tape.on) {
		// record the solved arbiters (their contacts learn where they were recorded)
		const int32_t *n_act_dev = w->d_step_counts + 3 * cur + 1;
		HIP_OR_FAIL(hipMemcpyAsync(w->h_counts + 10, n_act_dev, sizeof(int32_t), hipMemcpyDeviceToHost, w->stream));
		HIP_OR_FAIL(hipStreamSynchronize(w->stream));
		const int n_act = w->h_counts[10];
		HIP_OR_FAIL(tape_reserve(w, w->tape.n + n_act));
		HIP_OR_FAIL(adjoint_tape_record(w->act[cur], nullptr, n_act, w->tape.arb, (int)w->tape.n, w->stream));
		w->tape.off.push_back(w->tape.n);
		w->tape.count.push_back(n_act);
		w->tape.n += n_act;
		w->tape.steps++;
	}

	w->cur ^= 1;
	w->steps++;
	w->pending = true;
	return CPGPU_OK;
}

int cpgpuWorldStep(cpgpuWorld *w, double dt)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (w->in_step) return fail(CPGPU_ERR_INVALID, "a step begun with cpgpuWorldStepBegin is not ended");
	if (dt == 0.0) return CPGPU_OK;   // cpSpaceStep.c:339
	if (w->graph_on) return graph_step(w, dt);
	int rc = step_begin(w, dt);
	if (rc == CPGPU_OK) rc = step_end(w);
	if (rc != CPGPU_OK) return rc;
	return report_faults(w);   // faults seen by the step's wait, now that it is enqueued in full
}

// ---- collision handler callbacks: a step in two halves around the host ----

int cpgpuWorldStepBegin(cpgpuWorld *w, double dt, int64_t *n_callbacks)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (n_callbacks) *n_callbacks = 0;
	if (w->in_step) return fail(CPGPU_ERR_INVALID, "a step begun with cpgpuWorldStepBegin is not ended");
	if (w->graph_on) return fail(CPGPU_ERR_INVALID, "a step with callbacks is stepped without graphs");
	w->n_cb = 0;
	if (dt == 0.0) {   // cpSpaceStep.c:339: nothing happens
		w->in_step = true;
		w->zero_step = true;
		return CPGPU_OK;
	}
	int rc = step_begin(w, dt);
	if (rc != CPGPU_OK) return rc;
	w->in_step = true;
	w->zero_step = false;
	if (w->n_handlers > 0) {
		HIP_OR_FAIL(hipMemcpyAsync(w->h_counts + 11, w->d_hcounts, sizeof(int32_t), hipMemcpyDeviceToHost, w->stream));
		rc = sync_world(w);
		if (rc != CPGPU_OK) return rc;
		w->n_cb = std::min(w->h_counts[11], w->act_cap[w->cur]);
	}
	if (n_callbacks) *n_callbacks = w->n_cb;
	return CPGPU_OK;
}

int cpgpuWorldReadCallbacks(cpgpuWorld *w, int64_t cap, int64_t *n_out, cpgpuCallbackArbiter *out)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (!w->in_step) return fail(CPGPU_ERR_INVALID, "no step begun with cpgpuWorldStepBegin");
	if (cap < 0) return fail(CPGPU_ERR_INVALID, "negative capacity");
	if (cap > 0 && !out) return fail(CPGPU_ERR_INVALID, "null output array");
	if (n_out) *n_out = w->n_cb;
	const int n = (int)std::min<int64_t>(cap, w->n_cb);
	if (n <= 0 || w->zero_step) return CPGPU_OK;
	HIP_OR_FAIL(regrow(w->d_cbview, w->cbview_cap, w->n_cb));
	HIP_OR_FAIL(callbacks_pack(step_buffers(w), w->n_cb, w->d_cbview, w->stream));
	HIP_OR_FAIL(hipStreamSynchronize(w->stream));
	HIP_OR_FAIL(hipMemcpy(out, w->d_cbview, sizeof(cpgpuCallbackArbiter) * (size_t)n, hipMemcpyDeviceToHost));
	return CPGPU_OK;
}

int cpgpuWorldWriteDecisions(cpgpuWorld *w, int64_t n, const cpgpuCallbackDecision *decisions)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (!w->in_step) return fail(CPGPU_ERR_INVALID, "no step begun with cpgpuWorldStepBegin");
	if (n < 0 || (n > 0 && !decisions)) return fail(CPGPU_ERR_INVALID, "bad decision list");
	for (int64_t k = 0; k < n; k++)
		if (decisions[k].slot < 0 || decisions[k].slot >= w->n_cb)
			return fail(CPGPU_ERR_INVALID, "decision " + std::to_string(k) + " names slot " +
			                                   std::to_string(decisions[k].slot) + " of " + std::to_string(w->n_cb));
	if (n == 0) return CPGPU_OK;
	std::vector<cpgpuCallbackDecision> h((size_t)w->n_cb);
	HIP_OR_FAIL(hipStreamSynchronize(w->stream));
	HIP_OR_FAIL(hipMemcpy(h.data(), w->d_dec, sizeof(cpgpuCallbackDecision) * h.size(), hipMemcpyDeviceToHost));
	for (int64_t k = 0; k < n; k++) {
		cpgpuCallbackDecision d = decisions[k];
		d.solve = d.solve ? 1 : 0;
		d.ignore = d.ignore ? 1 : 0;
		h[(size_t)d.slot] = d;
	}
	HIP_OR_FAIL(hipMemcpy(w->d_dec, h.data(), sizeof(cpgpuCallbackDecision) * h.size(), hipMemcpyHostToDevice));
	return CPGPU_OK;
}

int cpgpuWorldStepEnd(cpgpuWorld *w)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (!w->in_step) return fail(CPGPU_ERR_INVALID, "no step begun with cpgpuWorldStepBegin");
	w->in_step = false;
	if (w->zero_step) return CPGPU_OK;
	const int rc = step_end(w);
	if (rc != CPGPU_OK) return rc;
	return report_faults(w);
}

int cpgpuWorldReadSeparations(cpgpuWorld *w, int64_t cap, int64_t *n_out, int32_t *pairs)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	int rc = sync_world(w);
	if (rc != CPGPU_OK) return rc;
	if (cap < 0) return fail(CPGPU_ERR_INVALID, "negative capacity");
	if (cap > 0 && !pairs) return fail(CPGPU_ERR_INVALID, "null output array");
	if (w->n_handlers == 0 || w->steps == 0) {
		if (n_out) *n_out = 0;
		return CPGPU_OK;
	}
	int32_t n = 0;
	HIP_OR_FAIL(hipMemcpy(&n, w->d_hcounts + 1, sizeof(int32_t), hipMemcpyDeviceToHost));
	if (n > w->seps_words / 3) return fail(CPGPU_ERR_ARBITERS_OVERFLOW, "more separations than their buffer holds");
	if (n_out) *n_out = n;
	const int64_t m = std::min<int64_t>(cap, n);
	if (m > 0) HIP_OR_FAIL(hipMemcpy(pairs, w->d_seps, sizeof(int32_t) * 3 * (size_t)m, hipMemcpyDeviceToHost));
	return CPGPU_OK;
}

// ---- steps captured in HIP graphs (cpgpuWorldSetGraph) ----

namespace {

// The per-step scalars of cpSpaceStep (cpSpaceStep.c:341-344) advanced on the
// device for a captured step: stamp + 1, dt_coef = dt / the previous dt (0
// after a first step), and dt, whose pow scalars the host wrote once.
__global__ void k_advance_worlds(WorldStep *ws, double *curr_dt, int n, double dt)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const double prev = curr_dt[k];
	ws[k].dt = (real)dt;
	ws[k].dt_coef = (real)(prev == 0.0 ? 0.0 : dt / prev);
	ws[k].stamp += 1u;
	curr_dt[k] = dt;
}

// A captured step sizes its launches for its buffers' capacities: more
// candidate pairs than that raise the error word; the pair totals of captured
// steps add up on the device.
__global__ void k_graph_pairs(const int32_t *n_pairs, int cap, int32_t *err, unsigned long long *total)
{
	const int n = *n_pairs;
	if (n > cap) raise_err(err, CPGPU_ERR_PAIRS_OVERFLOW);
	*total += (unsigned long long)n;
}

} // namespace

// Capacities of a captured step: room for twice the most pairs, records and
// cached arbiters seen so far (the counts of the last finished step), and for
// two pairs per dynamic shape at least.
static hipError_t graph_reserve(cpgpuWorld *w)
{
	const int prev = w->cur ^ 1;
	const long last_recs = w->steps > 0 ? w->h_counts[3 * prev] : 0;
	const long last_cached = w->steps > 0 ? w->h_counts[3 * prev + 2] : 0;
	const long seen_pairs = std::max<long>(w->last_pairs, w->graph_max_pairs);
	const long pairs = std::max<long>({2 * seen_pairs + 1024, (long)w->pair_cap, 2L * w->n_dyn_shapes});
	const long recs = std::max<long>({2 * pairs, 2 * last_recs + 1024, (long)w->graph_rec_cap});
	const long cached = std::max<long>({pairs, 2 * last_cached + 1024, (long)w->graph_cached_cap});
	HIP_TRY(regrow(w->d_pairs, w->pair_cap, pairs));
	for (int slot = 0; slot < 2; slot++) HIP_TRY(reserve_step(w, slot, recs, (long)w->pair_cap + 1, cached, recs));
	w->graph_rec_cap = std::min(w->rec_cap[0], w->rec_cap[1]);
	w->graph_cached_cap = std::min(w->cached_cap[0], w->cached_cap[1]);
	return radix_sort_reserve(w->sortbuf, std::max(w->pair_cap, std::max(w->n_dyn_shapes, 1)));
}

static int graph_capture(cpgpuWorld *w, double dt)
{
	int rc = sync_world(w);
	if (rc != CPGPU_OK) return rc;
	for (int slot = 0; slot < 2; slot++)
		if (w->gexec[slot]) { (void)hipGraphExecDestroy(w->gexec[slot]); w->gexec[slot] = nullptr; }
	HIP_OR_FAIL(graph_reserve(w));
	if (w->nc > 0 && dt != w->last_joint_dt) HIP_OR_FAIL(set_joint_bias(w, dt));
	const int prev = w->cur ^ 1;
	if (w->steps == 0) HIP_OR_FAIL(hipMemsetAsync(w->d_step_counts + 3 * prev, 0, 3 * sizeof(int32_t), w->stream));
	// the scalars the first captured step advances from
	fill_world_steps(w, dt, false);
	std::vector<double> curr(w->nw);
	for (int k = 0; k < w->nw; k++) curr[k] = w->sp[k].curr_dt;
	HIP_OR_FAIL(hipMemcpyAsync(w->d_ws, w->h_ws, sizeof(WorldStep) * w->nw, hipMemcpyHostToDevice, w->stream));
	HIP_OR_FAIL(hipMemcpyAsync(w->d_curr_dt, curr.data(), sizeof(double) * w->nw, hipMemcpyHostToDevice, w->stream));
	HIP_OR_FAIL(hipStreamSynchronize(w->stream));
	// one graph per arbiter slot: step k reads slot k % 2's predecessor and writes its own
	const int cur0 = w->cur;
	for (int q = 0; q < 2; q++) {
		w->cur = cur0 ^ q;
		hipGraph_t g = nullptr;
		HIP_OR_FAIL(hipStreamBeginCapture(w->stream, hipStreamCaptureModeRelaxed));
		hipError_t e = hipSuccess;
		k_advance_worlds<<<blocks_for(w->nw, 256), 256, 0, w->stream>>>(w->d_ws, w->d_curr_dt, w->nw, dt);
		e = hipGetLastError();
		if (e == hipSuccess) e = enqueue_front(w, nullptr);
		if (e == hipSuccess) {
			k_graph_pairs<<<1, 1, 0, w->stream>>>(w->d_counters, w->pair_cap, w->W.err, w->d_graph_pairs);
			e = hipGetLastError();
		}
		if (e == hipSuccess) e = enqueue_back(w, w->pair_cap, w->rec_cap[w->cur ^ 1], nullptr);
		if (e == hipSuccess) queue_counters(w);
		hipError_t e2 = hipStreamEndCapture(w->stream, &g);
		if (e == hipSuccess) e = e2;
		if (e == hipSuccess) e = hipGraphInstantiate(&w->gexec[w->cur], g, nullptr, nullptr, 0);
		if (g) (void)hipGraphDestroy(g);
		if (e != hipSuccess) {
			w->cur = cur0;
			(void)hipGetLastError();
			return fail(CPGPU_ERR_HIP, std::string("capturing the step in a graph: ") + hipGetErrorString(e));
		}
	}
	w->cur = cur0;
	w->graph_dt = dt;
	w->graph_dirty = false;
	return CPGPU_OK;
}

static int graph_step(cpgpuWorld *w, double dt)
{
	// the counts of the last finished captured step, copied by its last nodes: capacities with
	// less than half to spare are grown, and the step captured again
	if (w->graph_done && hipEventQuery(w->graph_done) == hipSuccess) {
		const int last = w->cur ^ 1;
		const long pairs = w->h_counts[6], recs = w->h_counts[3 * last], cached = w->h_counts[3 * last + 2];
		w->graph_max_pairs = std::max<long>(w->graph_max_pairs, pairs);
		if (2 * pairs > w->pair_cap || 2 * recs > w->graph_rec_cap || 2 * cached > w->graph_cached_cap)
			w->graph_dirty = true;
	}
	(void)hipGetLastError();   // a not-ready query is no error
	if (w->graph_dirty || dt != w->graph_dt || !w->gexec[w->cur]) {
		const int rc = graph_capture(w, dt);
		if (rc != CPGPU_OK) return rc;
	}
	for (int k = 0; k < w->nw; k++) {
		w->sp[k].stamp++;
		w->sp[k].curr_dt = dt;
	}
	HIP_OR_FAIL(hipGraphLaunch(w->gexec[w->cur], w->stream));
	w->dyn_tree_built = true;
	HIP_OR_FAIL(hipEventRecord(w->graph_done, w->stream));
	w->cur ^= 1;
	w->steps++;
	w->graph_steps++;
	w->pending = true;
	return report_faults(w);   // faults seen by a capture's wait
}

int cpgpuWorldSetGraph(cpgpuWorld *w, int32_t enable)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (enable) {
		if (w->n_handlers > 0) return fail(CPGPU_ERR_INVALID, "spaces with collision handlers are stepped without graphs");
		if (w->in_step) return fail(CPGPU_ERR_INVALID, "a step begun with cpgpuWorldStepBegin is not ended");
		if (w->tape.on) return fail(CPGPU_ERR_INVALID, "a taped rollout is stepped without graphs");
		if (w->timer.enabled) return fail(CPGPU_ERR_INVALID, "timed steps (cpgpuWorldSetTiming) are stepped without graphs");
		if (w->n_large_worlds > 0)
			return fail(CPGPU_ERR_INVALID, "spaces of more than 4096 dynamic bodies (the grid solve) are stepped "
			                               "without graphs");
	}
	int rc = sync_world(w);
	if (rc != CPGPU_OK) return rc;
	if (enable && !w->graph_done) {
		HIP_OR_FAIL(hipEventCreateWithFlags(&w->graph_done, hipEventDisableTiming));
		HIP_OR_FAIL(dalloc(w->d_curr_dt, w->nw));
		HIP_OR_FAIL(upload(w->d_graph_pairs, (const unsigned long long *)nullptr, 1));
	}
	w->graph_on = enable != 0;
	w->graph_dirty = true;
	return CPGPU_OK;
}

int cpgpuWorldStepN(cpgpuWorld *w, double dt, int32_t n_steps)
{
	for (int k = 0; k < n_steps; k++) {
		const int rc = cpgpuWorldStep(w, dt);
		if (rc != CPGPU_OK) return rc;
	}
	return CPGPU_OK;
}

int cpgpuWorldSync(cpgpuWorld *w)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	const int rc = sync_world(w);
	return rc != CPGPU_OK ? rc : report_faults(w);
}

int cpgpuWorldGetFaults(cpgpuWorld *w, int32_t *faults, int32_t *first_body)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (faults) *faults = w->h_counts[7] & CPGPU_FAULTS;
	if (first_body) *first_body = w->h_counts[8] > 0 ? 0x7fffffff - w->h_counts[8] : -1;
	return CPGPU_OK;
}

int cpgpuWorldReadBodies(cpgpuWorld *w, double *p, double *v, double *a, double *wv, double *v_bias, double *w_bias,
                         double *transform)
{
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	const int nb = w->nb;
	if (p) HIP_OR_FAIL(download_real(p, reinterpret_cast<const real *>(w->W.b_p), 2 * (size_t)nb));
	if (a) HIP_OR_FAIL(download_real(a, w->W.b_a, nb));
	if (transform) HIP_OR_FAIL(download_real(transform, reinterpret_cast<const real *>(w->W.b_T), 6 * (size_t)nb));
	if (v || wv || v_bias || w_bias) {
		std::vector<BodyVel> vel(nb);
		HIP_OR_FAIL(hipMemcpy(vel.data(), w->W.b_vel, sizeof(BodyVel) * nb, hipMemcpyDeviceToHost));
		for (int b = 0; b < nb; b++) {
			if (v) { v[2 * b] = vel[b].vx; v[2 * b + 1] = vel[b].vy; }
			if (wv) wv[b] = vel[b].w;
			if (v_bias) { v_bias[2 * b] = vel[b].vbx; v_bias[2 * b + 1] = vel[b].vby; }
			if (w_bias) w_bias[b] = vel[b].wb;
		}
	}
	return CPGPU_OK;
}

int cpgpuWorldWriteBodies(cpgpuWorld *w, const double *v, const double *wv, const double *f, const double *t)
{
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	const int nb = w->nb;
	if (v || wv) {
		std::vector<BodyVel> vel(nb);
		HIP_OR_FAIL(hipMemcpy(vel.data(), w->W.b_vel, sizeof(BodyVel) * nb, hipMemcpyDeviceToHost));
		for (int b = 0; b < nb; b++) {
			if (v) { vel[b].vx = (real)v[2 * b]; vel[b].vy = (real)v[2 * b + 1]; }
			if (wv) vel[b].w = (real)wv[b];
		}
		HIP_OR_FAIL(hipMemcpy(w->W.b_vel, vel.data(), sizeof(BodyVel) * nb, hipMemcpyHostToDevice));
	}
	if (f) HIP_OR_FAIL(upload_real_into(reinterpret_cast<real *>(w->W.b_f), f, 2 * (size_t)nb));
	if (t) HIP_OR_FAIL(upload_real_into(w->W.b_t, t, nb));
	return CPGPU_OK;
}

// Rewrite the listed entries of a device array of `width` scalars per body.
static hipError_t edit_rows(real *dev, int width, int nb, int32_t n, const int32_t *index, const double *src)
{
	std::vector<real> h((size_t)width * nb);
	HIP_TRY(hipMemcpy(h.data(), dev, sizeof(real) * h.size(), hipMemcpyDeviceToHost));
	for (int k = 0; k < n; k++)
		for (int c = 0; c < width; c++) h[(size_t)width * index[k] + c] = (real)src[(size_t)width * k + c];
	return hipMemcpy(dev, h.data(), sizeof(real) * h.size(), hipMemcpyHostToDevice);
}

int cpgpuWorldWriteBodiesAt(cpgpuWorld *w, int32_t n, const int32_t *index, const double *p, const double *v,
                            const double *f, const double *a, const double *wv, const double *t,
                            const double *transform)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (n < 0) return fail(CPGPU_ERR_INVALID, "negative body count");
	if (n == 0) return CPGPU_OK;
	if (!index) return fail(CPGPU_ERR_INVALID, "null body index list");
	const bool pose = p || a || transform;
	for (int k = 0; k < n; k++) {
		if (index[k] < 0 || index[k] >= w->nb)
			return fail(CPGPU_ERR_INVALID, "body index " + std::to_string(index[k]) + " out of range");
		// static shapes are cached once, when the world is made (cpSpaceReindexStatic's role)
		if (pose && w->body_type[index[k]] == CPGPU_BODY_STATIC)
			return fail(CPGPU_ERR_INVALID, "a static body's pose changes only with a new world (body " +
			                                   std::to_string(index[k]) + ")");
	}
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	const int nb = w->nb;
	if (p) HIP_OR_FAIL(edit_rows(reinterpret_cast<real *>(w->W.b_p), 2, nb, n, index, p));
	if (a) HIP_OR_FAIL(edit_rows(w->W.b_a, 1, nb, n, index, a));
	if (f) HIP_OR_FAIL(edit_rows(reinterpret_cast<real *>(w->W.b_f), 2, nb, n, index, f));
	if (t) HIP_OR_FAIL(edit_rows(w->W.b_t, 1, nb, n, index, t));
	if (transform) HIP_OR_FAIL(edit_rows(reinterpret_cast<real *>(w->W.b_T), 6, nb, n, index, transform));
	if (v || wv) {
		std::vector<BodyVel> vel(nb);
		HIP_OR_FAIL(hipMemcpy(vel.data(), w->W.b_vel, sizeof(BodyVel) * nb, hipMemcpyDeviceToHost));
		for (int k = 0; k < n; k++) {
			BodyVel &b = vel[index[k]];
			if (v) { b.vx = (real)v[2 * k]; b.vy = (real)v[2 * k + 1]; }
			if (wv) b.w = (real)wv[k];
		}
		HIP_OR_FAIL(hipMemcpy(w->W.b_vel, vel.data(), sizeof(BodyVel) * nb, hipMemcpyHostToDevice));
	}
	return CPGPU_OK;
}

int cpgpuWorldSetSpaceParams(cpgpuWorld *w, int32_t space, const cpgpuSpaceParams *params)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (!params) return fail(CPGPU_ERR_INVALID, "null space parameters");
	if (space < 0 || space >= w->nw) return fail(CPGPU_ERR_INVALID, "space index " + std::to_string(space) + " out of range");
	if (params->iterations < 0) return fail(CPGPU_ERR_INVALID, "iterations must be >= 0");
	if (w->tape.on) return fail(CPGPU_ERR_INVALID, "space parameters are fixed during a taped rollout");
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	w->sp[space] = *params;
	w->graph_dirty = true;
	w->max_iterations = 0;
	w->max_iterations_large = 0;
	for (int k = 0; k < w->nw; k++) {
		w->max_iterations = std::max(w->max_iterations, (int)w->sp[k].iterations);
		if (!w->h_small[k]) w->max_iterations_large = std::max(w->max_iterations_large, (int)w->sp[k].iterations);
	}
	return CPGPU_OK;
}

int cpgpuWorldCreateFrom(const cpgpuWorldDesc *d, int32_t precision, cpgpuWorld *old, const int32_t *shape_from,
                         cpgpuWorld **out)
{
	if (!out) return fail(CPGPU_ERR_INVALID, "null output pointer");
	*out = nullptr;
	if (!old) return fail(CPGPU_ERR_INVALID, "null previous world");
	if (old->precision != precision) return fail(CPGPU_ERR_INVALID, "the new world must keep the old world's precision");
	if (old->tape.on) return fail(CPGPU_ERR_INVALID, "a taped world cannot be packed again (cpgpuWorldTapeEnd first)");
	int rc = validate(d);
	if (rc != CPGPU_OK) return rc;
	if (d->n_shapes > 0 && !shape_from) return fail(CPGPU_ERR_INVALID, "null shape map");
	// new -> old and old -> new shape indices
	std::vector<int32_t> inv(old->ns, -1);
	for (int s = 0; s < d->n_shapes; s++) {
		const int o = shape_from[s];
		if (o < -1 || o >= old->ns) return fail(CPGPU_ERR_INVALID, "shape map entry " + std::to_string(s) + " out of range");
		if (o < 0) continue;
		if (inv[o] >= 0) return fail(CPGPU_ERR_INVALID, "shape map names old shape " + std::to_string(o) + " twice");
		inv[o] = s;
	}
	rc = wait_world(old);
	if (rc != CPGPU_OK) return rc;
	cpgpuWorld *w = nullptr;
	rc = cpgpuWorldCreate(d, precision, &w);
	if (rc != CPGPU_OK) return rc;
	auto abandon = [&](int code) { cpgpuWorldDestroy(w); return code; };
#define MIGRATE_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
	return abandon(fail(CPGPU_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_))); } while (0)
	const int ns = w->ns, nso = old->ns;
	std::vector<int32_t> shape_body_old(nso), shape_world(ns);
	std::vector<uint8_t> static_old(nso), static_new(ns);
	if (nso) MIGRATE_TRY(hipMemcpy(shape_body_old.data(), old->W.s_body, sizeof(int32_t) * nso, hipMemcpyDeviceToHost));
	if (nso) MIGRATE_TRY(hipMemcpy(static_old.data(), old->W.s_static, nso, hipMemcpyDeviceToHost));
	if (ns) MIGRATE_TRY(hipMemcpy(shape_world.data(), w->W.s_world, sizeof(int32_t) * ns, hipMemcpyDeviceToHost));
	if (ns) MIGRATE_TRY(hipMemcpy(static_new.data(), w->W.s_static, ns, hipMemcpyDeviceToHost));
	// a shape that stays in the dynamic tree keeps its fattened leaf box and the step of its
	// last rebuild, so warm-start ids stay valid exactly as long as without the new pack
	if (ns > 0 && nso > 0) {
		std::vector<real4> fat_old(nso), fat(ns);
		std::vector<uint32_t> reb_old(nso), reb(ns);
		MIGRATE_TRY(hipMemcpy(fat_old.data(), old->W.s_fat, sizeof(real4) * nso, hipMemcpyDeviceToHost));
		MIGRATE_TRY(hipMemcpy(reb_old.data(), old->W.s_rebuild, sizeof(uint32_t) * nso, hipMemcpyDeviceToHost));
		MIGRATE_TRY(hipMemcpy(fat.data(), w->W.s_fat, sizeof(real4) * ns, hipMemcpyDeviceToHost));
		MIGRATE_TRY(hipMemcpy(reb.data(), w->W.s_rebuild, sizeof(uint32_t) * ns, hipMemcpyDeviceToHost));
		for (int s = 0; s < ns; s++) {
			const int o = shape_from[s];
			if (o < 0 || static_new[s] || static_old[o]) continue;
			fat[s] = fat_old[o];
			reb[s] = reb_old[o];
		}
		MIGRATE_TRY(hipMemcpy(w->W.s_fat, fat.data(), sizeof(real4) * ns, hipMemcpyHostToDevice));
		MIGRATE_TRY(hipMemcpy(w->W.s_rebuild, reb.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice));
	}
	// the persistent pair records and arbiters of the old world's last step, for every pair
	// whose two shapes are both packed again, under their new indices
	if (old->steps > 0) {
		const int L = old->cur ^ 1;
		const int n_recs = old->h_counts[3 * L], n_act = old->h_counts[3 * L + 1], n_cached = old->h_counts[3 * L + 2];
		std::vector<PairRec> recs(n_recs);
		std::vector<Arbiter> act(n_act), cached(n_cached);
		if (n_recs) MIGRATE_TRY(hipMemcpy(recs.data(), old->recs[L], sizeof(PairRec) * n_recs, hipMemcpyDeviceToHost));
		if (n_act) MIGRATE_TRY(hipMemcpy(act.data(), old->act[L], sizeof(Arbiter) * n_act, hipMemcpyDeviceToHost));
		if (n_cached)
			MIGRATE_TRY(hipMemcpy(cached.data(), old->cached[L], sizeof(Arbiter) * n_cached, hipMemcpyDeviceToHost));
		auto remap = [&](std::vector<Arbiter> &v, std::vector<int32_t> &where) {
			where.assign(v.size(), -1);
			size_t keep = 0;
			for (size_t m = 0; m < v.size(); m++) {
				Arbiter a = v[m];
				const int sa = inv[a.sa], sb = inv[a.sb];
				if (sa < 0 || sb < 0 || shape_world[sa] != shape_world[sb]) continue;
				a.sa = sa; a.sb = sb;
				a.key = pair_key_host(sa, sb);
				a.ba = d->shape_body[sa]; a.bb = d->shape_body[sb];
				a.world = shape_world[sa];
				for (int c = 0; c < 2; c++) { a.c[c].tape = -1; a.c[c].src = -1; }
				where[m] = (int32_t)keep;
				v[keep++] = a;
			}
			v.resize(keep);
		};
		std::vector<int32_t> act_at, cached_at;
		remap(act, act_at);
		remap(cached, cached_at);
		size_t keep = 0;
		for (int r = 0; r < n_recs; r++) {
			PairRec rec = recs[r];
			const int sa = inv[(int32_t)(rec.key >> 32)], sb = inv[(int32_t)(rec.key & 0xffffffffu)];
			if (sa < 0 || sb < 0 || shape_world[sa] != shape_world[sb]) continue;
			rec.key = pair_key_host(sa, sb);
			if (rec.arb >= 0 && (rec.arb & CPGPU_CACHED_REF)) {
				const int32_t m = cached_at[rec.arb & ~CPGPU_CACHED_REF];
				rec.arb = m >= 0 ? (m | CPGPU_CACHED_REF) : -1;
			} else if (rec.arb >= 0) {
				rec.arb = act_at[rec.arb];
			}
			recs[keep++] = rec;
		}
		recs.resize(keep);
		const int P = w->cur ^ 1;
		MIGRATE_TRY(regrow(w->recs[P], w->rec_cap[P], (long)recs.size() + 1));
		MIGRATE_TRY(regrow(w->act[P], w->act_cap[P], (long)act.size() + 1));
		MIGRATE_TRY(regrow(w->cached[P], w->cached_cap[P], (long)cached.size() + 1));
		if (!recs.empty())
			MIGRATE_TRY(hipMemcpy(w->recs[P], recs.data(), sizeof(PairRec) * recs.size(), hipMemcpyHostToDevice));
		if (!act.empty())
			MIGRATE_TRY(hipMemcpy(w->act[P], act.data(), sizeof(Arbiter) * act.size(), hipMemcpyHostToDevice));
		if (!cached.empty())
			MIGRATE_TRY(hipMemcpy(w->cached[P], cached.data(), sizeof(Arbiter) * cached.size(), hipMemcpyHostToDevice));
		const int32_t counts[3] = {(int32_t)recs.size(), (int32_t)act.size(), (int32_t)cached.size()};
		MIGRATE_TRY(hipMemcpy(w->d_step_counts + 3 * P, counts, sizeof(counts), hipMemcpyHostToDevice));
		for (int k = 0; k < 3; k++) w->h_counts[3 * P + k] = counts[k];
		w->steps = old->steps;
		w->total_pairs = old->total_pairs;
		MIGRATE_TRY(hipMemcpy(w->W.totals, old->W.totals, sizeof(unsigned long long) * (2 + 2 * CPGPU_SOLVE_UNITS),
		                      hipMemcpyDeviceToDevice));
	}
	// the old world's tiling and grid schedule
	if (old->n_buckets == w->n_buckets) {
		for (int q = 0; q < w->n_buckets; q++) {
			w->buckets[q].max_bodies = old->buckets[q].max_bodies;
			w->buckets[q].threads = old->buckets[q].threads;
		}
		MIGRATE_TRY(build_buckets(w));
	}
	w->grid_flow = old->grid_flow;
#undef MIGRATE_TRY
	*out = w;
	return CPGPU_OK;
}

// ---- snapshots of the state a step carries forward (cpgpuWorldSaveState / LoadState) ----

namespace {

struct SnapshotHeader {
	char magic[8];               // "cpgpuSS1"
	int32_t version;             // 1
	int32_t precision;           // CPGPU_FP64 / CPGPU_FP32
	int32_t n_worlds, n_bodies, n_shapes, n_constraints;
	int32_t n_recs, n_act, n_cached;   // the last step's pair records, solved and cached arbiters
	int32_t real_bytes;
	int32_t sizes[4];            // sizeof PairRec, Arbiter, Joint, BodyVel of the build that wrote it
	int32_t pad;
	int64_t steps, total_pairs, last_pairs;
};

constexpr char SNAP_MAGIC[8] = {'c', 'p', 'g', 'p', 'u', 'S', 'S', '1'};
constexpr int SNAP_TOTALS = 2 + 2 * CPGPU_SOLVE_UNITS;

// the sections after the header, in order: {device pointer, bytes}
struct SnapSection {
	void *dev;
	size_t bytes;
};

std::vector<SnapSection> snap_sections(cpgpuWorld *w, int n_recs, int n_act, int n_cached)
{
	const int P = w->cur ^ 1;
	DevWorld &W = w->W;
	const size_t nb = (size_t)w->nb, ns = (size_t)w->ns;
	return {
	    {W.b_p, sizeof(real2) * nb},         {W.b_a, sizeof(real) * nb},       {W.b_f, sizeof(real2) * nb},
	    {W.b_t, sizeof(real) * nb},          {W.b_T, sizeof(XF) * nb},         {W.b_vel, sizeof(BodyVel) * nb},
	    {W.s_fat, sizeof(real4) * ns},       {W.s_rebuild, sizeof(uint32_t) * ns},
	    {W.joints, sizeof(Joint) * (size_t)w->nc},
	    {w->recs[P], sizeof(PairRec) * (size_t)n_recs}, {w->act[P], sizeof(Arbiter) * (size_t)n_act},
	    {w->cached[P], sizeof(Arbiter) * (size_t)n_cached}, {W.totals, sizeof(unsigned long long) * SNAP_TOTALS},
	};
}

} // namespace

int cpgpuWorldSaveState(cpgpuWorld *w, void *buf, int64_t cap, int64_t *size_out)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (cap < 0) return fail(CPGPU_ERR_INVALID, "negative capacity");
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	const int L = w->cur ^ 1;
	SnapshotHeader h;
	memset(&h, 0, sizeof(h));
	memcpy(h.magic, SNAP_MAGIC, 8);
	h.version = 1;
	h.precision = w->precision;
	h.n_worlds = w->nw; h.n_bodies = w->nb; h.n_shapes = w->ns; h.n_constraints = w->nc;
	h.n_recs = w->steps > 0 ? w->h_counts[3 * L] : 0;
	h.n_act = w->steps > 0 ? w->h_counts[3 * L + 1] : 0;
	h.n_cached = w->steps > 0 ? w->h_counts[3 * L + 2] : 0;
	h.real_bytes = (int32_t)sizeof(real);
	h.sizes[0] = sizeof(PairRec); h.sizes[1] = sizeof(Arbiter); h.sizes[2] = sizeof(Joint); h.sizes[3] = sizeof(BodyVel);
	h.steps = w->steps; h.total_pairs = w->total_pairs; h.last_pairs = w->last_pairs;
	const auto sec = snap_sections(w, h.n_recs, h.n_act, h.n_cached);
	size_t size = sizeof(h) + sizeof(cpgpuSpaceParams) * (size_t)w->nw;
	for (const auto &q : sec) size += q.bytes;
	if (size_out) *size_out = (int64_t)size;
	if (!buf) return CPGPU_OK;
	if ((size_t)cap < size)
		return fail(CPGPU_ERR_CAPACITY, "snapshot needs " + std::to_string(size) + " bytes, " + std::to_string(cap) +
		                                    " given");
	char *o = static_cast<char *>(buf);
	memcpy(o, &h, sizeof(h)); o += sizeof(h);
	memcpy(o, w->sp.data(), sizeof(cpgpuSpaceParams) * (size_t)w->nw); o += sizeof(cpgpuSpaceParams) * (size_t)w->nw;
	for (const auto &q : sec) {
		if (q.bytes) HIP_OR_FAIL(hipMemcpy(o, q.dev, q.bytes, hipMemcpyDeviceToHost));
		o += q.bytes;
	}
	return CPGPU_OK;
}

int cpgpuWorldLoadState(cpgpuWorld *w, const void *buf, int64_t size)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (!buf || size < (int64_t)sizeof(SnapshotHeader)) return fail(CPGPU_ERR_INVALID, "snapshot too short");
	if (w->tape.on) return fail(CPGPU_ERR_INVALID, "a taped world cannot load a snapshot (cpgpuWorldTapeEnd first)");
	SnapshotHeader h;
	memcpy(&h, buf, sizeof(h));
	if (memcmp(h.magic, SNAP_MAGIC, 8) != 0 || h.version != 1) return fail(CPGPU_ERR_INVALID, "not a cpgpu snapshot");
	if (h.precision != w->precision || h.real_bytes != (int32_t)sizeof(real))
		return fail(CPGPU_ERR_INVALID, "the snapshot was saved from a world of the other precision");
	if (h.sizes[0] != (int32_t)sizeof(PairRec) || h.sizes[1] != (int32_t)sizeof(Arbiter) ||
	    h.sizes[2] != (int32_t)sizeof(Joint) || h.sizes[3] != (int32_t)sizeof(BodyVel))
		return fail(CPGPU_ERR_INVALID, "the snapshot was saved by another build of the library");
	if (h.n_worlds != w->nw || h.n_bodies != w->nb || h.n_shapes != w->ns || h.n_constraints != w->nc)
		return fail(CPGPU_ERR_INVALID, "the snapshot is of a world of another description (counts differ)");
	if (h.n_recs < 0 || h.n_act < 0 || h.n_cached < 0 || h.steps < 0)
		return fail(CPGPU_ERR_INVALID, "negative counts in the snapshot");
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	const int P = w->cur ^ 1;
	HIP_OR_FAIL(regrow(w->recs[P], w->rec_cap[P], (long)h.n_recs + 1));
	HIP_OR_FAIL(regrow(w->act[P], w->act_cap[P], (long)h.n_act + 1));
	HIP_OR_FAIL(regrow(w->cached[P], w->cached_cap[P], (long)h.n_cached + 1));
	const auto sec = snap_sections(w, h.n_recs, h.n_act, h.n_cached);
	size_t need = sizeof(h) + sizeof(cpgpuSpaceParams) * (size_t)w->nw;
	for (const auto &q : sec) need += q.bytes;
	if ((size_t)size != need)
		return fail(CPGPU_ERR_INVALID, "snapshot of " + std::to_string(size) + " bytes, its header describes " +
		                                   std::to_string(need));
	const char *in = static_cast<const char *>(buf) + sizeof(h);
	std::vector<cpgpuSpaceParams> sp(w->nw);
	memcpy(sp.data(), in, sizeof(cpgpuSpaceParams) * (size_t)w->nw);
	in += sizeof(cpgpuSpaceParams) * (size_t)w->nw;
	for (const auto &q : sec) {
		if (q.bytes) HIP_OR_FAIL(hipMemcpy(q.dev, in, q.bytes, hipMemcpyHostToDevice));
		in += q.bytes;
	}
	const int32_t counts[3] = {h.n_recs, h.n_act, h.n_cached};
	HIP_OR_FAIL(hipMemcpy(w->d_step_counts + 3 * P, counts, sizeof(counts), hipMemcpyHostToDevice));
	for (int k = 0; k < 3; k++) w->h_counts[3 * P + k] = counts[k];
	w->sp = sp;
	w->max_iterations = 0;
	w->max_iterations_large = 0;
	for (int k = 0; k < w->nw; k++) {
		w->max_iterations = std::max(w->max_iterations, (int)w->sp[k].iterations);
		if (!w->h_small[k]) w->max_iterations_large = std::max(w->max_iterations_large, (int)w->sp[k].iterations);
	}
	w->steps = h.steps;
	w->total_pairs = h.total_pairs;
	w->last_pairs = h.last_pairs;
	w->last_joint_dt = -1.0;   // the joints' bias coefficients are recomputed with the next step's dt
	w->graph_dirty = true;
	w->dyn_tree_built = false;
	// the dynamic shapes' cached geometry from the loaded transforms (their leaf boxes are loaded)
	if (w->n_dyn_shapes > 0) {
		k_cache_shapes<<<blocks_for(w->n_dyn_shapes, 256), 256, 0, w->stream>>>(w->W, w->d_dyn_shapes, w->n_dyn_shapes,
		                                                                        0, 0);
		HIP_OR_FAIL(hipGetLastError());
		HIP_OR_FAIL(hipStreamSynchronize(w->stream));
	}
	return CPGPU_OK;
}

int cpgpuWorldGetStats(cpgpuWorld *w, cpgpuStats *out)
{
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	const int last = w->cur ^ 1;
	memset(out, 0, sizeof(*out));
	out->candidate_pairs = w->last_pairs;
	if (w->steps > 0) {
		out->arbiters = w->h_counts[3 * last + 1];
		out->cached_arbiters = w->h_counts[3 * last + 1] + w->h_counts[3 * last + 2];
		// contacts and colours of the solved arbiters (colours count per space; the largest is reported)
		std::vector<Arbiter> arbs(out->arbiters);
		std::vector<int32_t> col(out->arbiters);
		if (out->arbiters) {
			HIP_OR_FAIL(hipMemcpy(arbs.data(), w->act[last], sizeof(Arbiter) * out->arbiters, hipMemcpyDeviceToHost));
			HIP_OR_FAIL(hipMemcpy(col.data(), w->colour, sizeof(int32_t) * out->arbiters, hipMemcpyDeviceToHost));
		}
		// with collision handlers, arbiters touching but not solved have count 0 and no colour
		if (w->n_handlers > 0) {
			out->arbiters = 0;
			for (auto &a : arbs) out->arbiters += a.count > 0;
		}
		for (auto &a : arbs) out->contacts += a.count;
		for (size_t k = 0; k < col.size(); k++)
			if (arbs[k].count > 0) out->arbiter_colours = std::max(out->arbiter_colours, col[k] + 1);
	}
	out->constraint_colours = w->n_jcolours;
	out->steps = w->steps;
	unsigned long long totals[2];
	HIP_OR_FAIL(hipMemcpy(totals, w->W.totals, sizeof(totals), hipMemcpyDeviceToHost));
	out->total_arbiters = (int64_t)totals[0];
	out->total_contacts = (int64_t)totals[1];
	out->total_pairs = w->total_pairs;
	if (w->d_graph_pairs) {
		unsigned long long g = 0;
		HIP_OR_FAIL(hipMemcpy(&g, w->d_graph_pairs, sizeof(g), hipMemcpyDeviceToHost));
		out->total_pairs += (int64_t)g;
	}
	return CPGPU_OK;
}

int cpgpuWorldReadArbiters(cpgpuWorld *w, int64_t cap, int64_t *n_out, int32_t *shapes, int32_t *colour,
                           int32_t *state, int32_t *count, double *normal, double *contact, uint64_t *hash)
{
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	const int last = w->cur ^ 1;
	const int n = w->steps > 0 ? w->h_counts[3 * last + 1] : 0;
	// with collision handlers, arbiters touching but not solved (count 0) are not among the step's
	if (n_out && w->n_handlers == 0) *n_out = n;
	if (n == 0 || (cap <= 0 && w->n_handlers == 0)) {
		if (n_out && n == 0) *n_out = 0;
		return CPGPU_OK;
	}
	std::vector<Arbiter> arbs(n);
	std::vector<int32_t> col(n);
	HIP_OR_FAIL(hipMemcpy(arbs.data(), w->act[last], sizeof(Arbiter) * n, hipMemcpyDeviceToHost));
	HIP_OR_FAIL(hipMemcpy(col.data(), w->colour, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
	std::vector<int> idx;
	idx.reserve(n);
	for (int k = 0; k < n; k++)
		if (arbs[k].count > 0) idx.push_back(k);
	if (n_out) *n_out = (int64_t)idx.size();
	std::sort(idx.begin(), idx.end(), [&](int x, int y) {
		if (col[x] != col[y]) return col[x] < col[y];
		return arbs[x].key < arbs[y].key;
	});
	for (int q = 0; q < (int)idx.size() && q < cap; q++) {
		const Arbiter &a = arbs[idx[q]];
		if (shapes) { shapes[2 * q] = a.sa; shapes[2 * q + 1] = a.sb; }
		if (colour) colour[q] = col[idx[q]];
		if (state) state[q] = a.state;
		if (count) count[q] = a.count;
		if (normal) { normal[2 * q] = a.nx; normal[2 * q + 1] = a.ny; }
		for (int i = 0; i < 2; i++) {
			const Contact &c = a.c[i];
			const bool valid = i < a.count;
			if (contact) {
				double *o = contact + 24 * q + 12 * i;
				const double vals[12] = {c.r1x, c.r1y, c.r2x, c.r2y, c.nMass, c.tMass, c.bounce,
				                         c.jnAcc, c.jtAcc, c.jBias, c.bias, 0.0};
				for (int k = 0; k < 12; k++) o[k] = valid ? vals[k] : 0.0;
			}
			if (hash) hash[2 * q + i] = valid ? c.hash : 0;
		}
	}
	return CPGPU_OK;
}

int cpgpuWorldReadPairs(cpgpuWorld *w, int64_t cap, int64_t *n_out, int32_t *pairs)
{
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	const int64_t n = w->steps > 0 ? w->last_pairs : 0;
	if (n_out) *n_out = n;
	if (n == 0 || cap <= 0 || !pairs) return CPGPU_OK;
	std::vector<int2> h((size_t)n);
	HIP_OR_FAIL(hipMemcpy(h.data(), w->d_pairs, sizeof(int2) * (size_t)n, hipMemcpyDeviceToHost));
	// unordered pairs as (lower index, higher index), sorted: the device's order depends on the
	// tree traversal's timing, the set does not
	for (auto &p : h)
		if (p.x > p.y) std::swap(p.x, p.y);
	std::sort(h.begin(), h.end(), [](const int2 &x, const int2 &y) { return x.x != y.x ? x.x < y.x : x.y < y.y; });
	for (int64_t q = 0; q < n && q < cap; q++) {
		pairs[2 * q] = h[q].x;
		pairs[2 * q + 1] = h[q].y;
	}
	return CPGPU_OK;
}

// Device bytes that cpgpuWorldShapesCollideIn stages n pairs in: the pairs,
// the counts and the contact point sets, each 256-byte aligned.
size_t cpgpuShapesCollideScratchBytes(int64_t n)
{
	auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
	return up(sizeof(int2) * (size_t)n) + up(sizeof(int32_t) * (size_t)n) + up(sizeof(double) * 12 * (size_t)n);
}

// cpShapesCollide for n pairs through the caller's device memory `scratch`
// (at least cpgpuShapesCollideScratchBytes(n) bytes on the world's device):
// the pairs are staged there, the kernel writes the results there, and they
// are copied back before the call returns.  Allocates nothing.
int cpgpuWorldShapesCollideIn(cpgpuWorld *w, int64_t n, const int32_t *pairs, int32_t *count, double *sets,
                              void *scratch, size_t scratch_bytes)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (n < 0) return fail(CPGPU_ERR_INVALID, "negative pair count");
	if (n == 0) return CPGPU_OK;
	if (!pairs || !count || !sets) return fail(CPGPU_ERR_INVALID, "null pair, count or contact set array");
	if (n > (int64_t)INT32_MAX / 12) return fail(CPGPU_ERR_INVALID, "too many pairs in one call");
	if (!scratch || scratch_bytes < cpgpuShapesCollideScratchBytes(n))
		return fail(CPGPU_ERR_CAPACITY, "staging memory too small for " + std::to_string(n) + " pairs");
	for (int64_t k = 0; k < 2 * n; k++)
		if (pairs[k] < 0 || pairs[k] >= w->ns)
			return fail(CPGPU_ERR_INVALID, "shape index " + std::to_string(pairs[k]) + " out of range (pair " +
			                                   std::to_string(k / 2) + ")");
	int cur = -1;
	HIP_OR_FAIL(hipGetDevice(&cur));
	if (cur != w->device) return fail(CPGPU_ERR_INVALID, "call on the device the world was created on");
	auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
	char *base = static_cast<char *>(scratch);
	int2 *d_pairs = reinterpret_cast<int2 *>(base);
	int32_t *d_count = reinterpret_cast<int32_t *>(base + up(sizeof(int2) * (size_t)n));
	double *d_sets = reinterpret_cast<double *>(base + up(sizeof(int2) * (size_t)n) + up(sizeof(int32_t) * (size_t)n));
	// queued behind the world's steps on its stream: the shapes' cached geometry is the last step's
	HIP_OR_FAIL(hipMemcpyAsync(d_pairs, pairs, sizeof(int2) * (size_t)n, hipMemcpyHostToDevice, w->stream));
	HIP_OR_FAIL(shapes_collide(w->W, d_pairs, (int)n, d_count, d_sets, w->stream));
	HIP_OR_FAIL(hipMemcpyAsync(count, d_count, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, w->stream));
	HIP_OR_FAIL(hipMemcpyAsync(sets, d_sets, sizeof(double) * 12 * (size_t)n, hipMemcpyDeviceToHost, w->stream));
	return sync_world(w);
}

int cpgpuWorldShapesCollide(cpgpuWorld *w, int64_t n, const int32_t *pairs, int32_t *count, double *sets)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (n <= 0) return n < 0 ? fail(CPGPU_ERR_INVALID, "negative pair count") : CPGPU_OK;
	// this form stages the call in device memory of its own
	const size_t bytes = cpgpuShapesCollideScratchBytes(n);
	void *dev = nullptr;
	HIP_OR_FAIL(hipMalloc(&dev, bytes));
	const int rc = cpgpuWorldShapesCollideIn(w, n, pairs, count, sets, dev, bytes);
	(void)hipFree(dev);
	return rc;
}

// ---- spatial queries (cpSpaceQuery.c) on the last step's trees ----

static int query_trees(cpgpuWorld *w, QueryTrees &T)
{
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	if (!w->dyn_tree_built && w->n_dyn_shapes > 0) {
		// before the first step: the tree over the boxes cached at creation
		HIP_OR_FAIL(lbvh_build(w->dyn_tree, w->W, w->d_dyn_shapes, w->n_dyn_shapes, w->nw, w->sortbuf,
		                       w->d_world_bounds, w->stream, nullptr));
		HIP_OR_FAIL(hipStreamSynchronize(w->stream));
		w->dyn_tree_built = true;
	}
	T.nd = w->n_dyn_shapes;
	T.ns = w->n_static_shapes;
	T.droot = w->dyn_tree.world_root; T.sroot = w->static_tree.world_root;
	T.dshape = w->dyn_tree.leaf_shape; T.sshape = w->static_tree.leaf_shape;
	T.dnode = w->dyn_tree.node; T.snode = w->static_tree.node;
	T.dbox = w->dyn_tree.box; T.sbox = w->static_tree.box;
	return CPGPU_OK;
}

// Stages a batch of queries in one device allocation: inputs (space, a [n][2
// or 4], b [n][2], scalar [n], filter [n]) then outputs (out_bytes).
struct QueryStage {
	char *base = nullptr;
	QueryIn in{};
	char *out = nullptr;
	~QueryStage() { if (base) (void)hipFree(base); }
};

static int query_stage(cpgpuWorld *w, int32_t n, const int32_t *space, const double *a, int a_width, const double *b,
                       const double *scalar, const cpgpuShapeFilter *filter, size_t out_bytes, QueryStage &S)
{
	if (n < 0) return fail(CPGPU_ERR_INVALID, "negative query count");
	if (n > 0 && (!a || (a_width == 2 && !scalar)))
		return fail(CPGPU_ERR_INVALID, "null query points, segments, boxes or distances");
	for (int32_t q = 0; space && q < n; q++)
		if (space[q] < 0 || space[q] >= w->nw)
			return fail(CPGPU_ERR_INVALID, "query " + std::to_string(q) + " names space " + std::to_string(space[q]) +
			                                   " of " + std::to_string(w->nw));
	auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
	const size_t nn = (size_t)std::max(n, 1);
	const size_t sz_space = up(sizeof(int32_t) * nn), sz_a = up(sizeof(double) * a_width * nn);
	const size_t sz_b = up(sizeof(double) * 2 * nn), sz_s = up(sizeof(double) * nn);
	const size_t sz_f = up(sizeof(cpgpuShapeFilter) * nn);
	HIP_OR_FAIL(hipMalloc(&S.base, sz_space + sz_a + sz_b + sz_s + sz_f + up(out_bytes)));
	char *p = S.base;
	int32_t *d_space = space ? reinterpret_cast<int32_t *>(p) : nullptr; p += sz_space;
	double *d_a = reinterpret_cast<double *>(p); p += sz_a;
	double *d_b = b ? reinterpret_cast<double *>(p) : nullptr; p += sz_b;
	double *d_s = scalar ? reinterpret_cast<double *>(p) : nullptr; p += sz_s;
	cpgpuShapeFilter *d_f = filter ? reinterpret_cast<cpgpuShapeFilter *>(p) : nullptr; p += sz_f;
	S.out = p;
	if (n > 0) {
		if (space) HIP_OR_FAIL(hipMemcpy(d_space, space, sizeof(int32_t) * n, hipMemcpyHostToDevice));
		HIP_OR_FAIL(hipMemcpy(d_a, a, sizeof(double) * a_width * n, hipMemcpyHostToDevice));
		if (b) HIP_OR_FAIL(hipMemcpy(d_b, b, sizeof(double) * 2 * n, hipMemcpyHostToDevice));
		if (scalar) HIP_OR_FAIL(hipMemcpy(d_s, scalar, sizeof(double) * n, hipMemcpyHostToDevice));
		if (filter) HIP_OR_FAIL(hipMemcpy(d_f, filter, sizeof(cpgpuShapeFilter) * n, hipMemcpyHostToDevice));
	}
	S.in.space = d_space; S.in.a = d_a; S.in.b = d_b; S.in.scalar = d_s; S.in.filter = d_f;
	return CPGPU_OK;
}

static int query_one_each(cpgpuWorld *w, int32_t n, const int32_t *space, const double *a, const double *b,
                          const double *scalar, const cpgpuShapeFilter *filter, bool segment, int32_t *shape_out,
                          double *info_out)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (n > 0 && (!shape_out || !info_out)) return fail(CPGPU_ERR_INVALID, "null query outputs");
	if (segment && n > 0 && !b) return fail(CPGPU_ERR_INVALID, "null segment ends");
	QueryTrees T;
	int rc = query_trees(w, T);
	if (rc != CPGPU_OK) return rc;
	QueryStage S;
	rc = query_stage(w, n, space, a, 2, b, scalar, filter, (sizeof(int32_t) + 5 * sizeof(double)) * (size_t)std::max(n, 1) + 256, S);
	if (rc != CPGPU_OK || n == 0) return rc;
	int32_t *d_shape = reinterpret_cast<int32_t *>(S.out);
	double *d_info = reinterpret_cast<double *>(S.out + ((sizeof(int32_t) * (size_t)n + 255) & ~(size_t)255));
	HIP_OR_FAIL(query_nearest_or_first(w->W, T, n, S.in, segment, d_shape, d_info, w->stream));
	rc = sync_world(w);
	if (rc != CPGPU_OK) return rc;
	HIP_OR_FAIL(hipMemcpy(shape_out, d_shape, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
	HIP_OR_FAIL(hipMemcpy(info_out, d_info, sizeof(double) * 5 * n, hipMemcpyDeviceToHost));
	return CPGPU_OK;
}

static int query_every(cpgpuWorld *w, int32_t n, const int32_t *space, const double *a, int a_width, const double *b,
                       const double *scalar, const cpgpuShapeFilter *filter, int kind, int64_t cap, int64_t *n_out,
                       int32_t *hits, double *info)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (cap < 0) return fail(CPGPU_ERR_INVALID, "negative hit capacity");
	if (cap > 0 && !hits) return fail(CPGPU_ERR_INVALID, "null hit array");
	if (kind == 1 && n > 0 && !b) return fail(CPGPU_ERR_INVALID, "null segment ends");
	QueryTrees T;
	int rc = query_trees(w, T);
	if (rc != CPGPU_OK) return rc;
	// the device keeps room for at least 4096 hits, so a first call with cap 0 learns the total
	const int64_t dcap = std::max<int64_t>(cap, 4096);
	const size_t hit_bytes = (sizeof(int2) * (size_t)dcap + 255) & ~(size_t)255;
	const size_t info_bytes = (sizeof(double) * 5 * (size_t)dcap + 255) & ~(size_t)255;
	QueryStage S;
	rc = query_stage(w, n, space, a, a_width, b, scalar, filter, hit_bytes + info_bytes + 256, S);
	if (rc != CPGPU_OK) return rc;
	if (n == 0) {
		if (n_out) *n_out = 0;
		return CPGPU_OK;
	}
	int2 *d_hits = reinterpret_cast<int2 *>(S.out);
	double *d_info = reinterpret_cast<double *>(S.out + hit_bytes);
	unsigned long long *d_count = reinterpret_cast<unsigned long long *>(S.out + hit_bytes + info_bytes);
	HIP_OR_FAIL(hipMemsetAsync(d_count, 0, sizeof(unsigned long long), w->stream));
	HIP_OR_FAIL(query_all(w->W, T, n, S.in, kind, d_hits, kind == 2 ? nullptr : d_info, dcap, d_count, w->stream));
	rc = sync_world(w);
	if (rc != CPGPU_OK) return rc;
	unsigned long long total = 0;
	HIP_OR_FAIL(hipMemcpy(&total, d_count, sizeof(total), hipMemcpyDeviceToHost));
	if (n_out) *n_out = (int64_t)total;
	const int64_t got = std::min<int64_t>((int64_t)total, dcap);
	if (cap == 0 || got == 0) return CPGPU_OK;
	std::vector<int2> h(got);
	std::vector<double> hi(kind == 2 ? 0 : 5 * got);
	HIP_OR_FAIL(hipMemcpy(h.data(), d_hits, sizeof(int2) * got, hipMemcpyDeviceToHost));
	if (kind != 2) HIP_OR_FAIL(hipMemcpy(hi.data(), d_info, sizeof(double) * 5 * got, hipMemcpyDeviceToHost));
	// by query, then shape: the device appended them in no fixed order
	std::vector<int64_t> order(got);
	for (int64_t k = 0; k < got; k++) order[k] = k;
	std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
		return h[x].x != h[y].x ? h[x].x < h[y].x : h[x].y < h[y].y;
	});
	for (int64_t k = 0; k < got && k < cap; k++) {
		hits[2 * k] = h[order[k]].x;
		hits[2 * k + 1] = h[order[k]].y;
		if (info && kind != 2)
			for (int i = 0; i < 5; i++) info[5 * k + i] = hi[5 * order[k] + i];
	}
	return CPGPU_OK;
}

int cpgpuWorldQueryPointNearest(cpgpuWorld *w, int32_t n, const int32_t *space, const double *points,
                                const double *max_distance, const cpgpuShapeFilter *filter, int32_t *shape_out,
                                double *info_out)
{
	return query_one_each(w, n, space, points, nullptr, max_distance, filter, false, shape_out, info_out);
}

int cpgpuWorldQuerySegmentFirst(cpgpuWorld *w, int32_t n, const int32_t *space, const double *starts,
                                const double *ends, const double *radius, const cpgpuShapeFilter *filter,
                                int32_t *shape_out, double *info_out)
{
	return query_one_each(w, n, space, starts, ends, radius, filter, true, shape_out, info_out);
}

int cpgpuWorldQueryPointAll(cpgpuWorld *w, int32_t n, const int32_t *space, const double *points,
                            const double *max_distance, const cpgpuShapeFilter *filter, int64_t cap, int64_t *n_out,
                            int32_t *hits, double *info)
{
	return query_every(w, n, space, points, 2, nullptr, max_distance, filter, 0, cap, n_out, hits, info);
}

int cpgpuWorldQuerySegmentAll(cpgpuWorld *w, int32_t n, const int32_t *space, const double *starts,
                              const double *ends, const double *radius, const cpgpuShapeFilter *filter, int64_t cap,
                              int64_t *n_out, int32_t *hits, double *info)
{
	return query_every(w, n, space, starts, 2, ends, radius, filter, 1, cap, n_out, hits, info);
}

int cpgpuWorldQueryBBAll(cpgpuWorld *w, int32_t n, const int32_t *space, const double *boxes,
                         const cpgpuShapeFilter *filter, int64_t cap, int64_t *n_out, int32_t *hits)
{
	return query_every(w, n, space, boxes, 4, nullptr, nullptr, filter, 2, cap, n_out, hits, nullptr);
}

int cpgpuWorldQueryShapes(cpgpuWorld *w, int32_t n, const int32_t *shapes, int64_t cap, int64_t *n_out, int32_t *hits,
                          int32_t *counts, double *sets)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (n < 0) return fail(CPGPU_ERR_INVALID, "negative query count");
	if (cap < 0) return fail(CPGPU_ERR_INVALID, "negative hit capacity");
	if (n > 0 && !shapes) return fail(CPGPU_ERR_INVALID, "null query shapes");
	if (cap > 0 && (!hits || !counts || !sets)) return fail(CPGPU_ERR_INVALID, "null hit, count or contact set array");
	for (int32_t q = 0; q < n; q++)
		if (shapes[q] < 0 || shapes[q] >= w->ns)
			return fail(CPGPU_ERR_INVALID, "query " + std::to_string(q) + " names shape " + std::to_string(shapes[q]) +
			                                   " of " + std::to_string(w->ns));
	QueryTrees T;
	int rc = query_trees(w, T);
	if (rc != CPGPU_OK) return rc;
	if (n == 0) {
		if (n_out) *n_out = 0;
		return CPGPU_OK;
	}
	const int64_t dcap = std::max<int64_t>(cap, 4096);
	auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
	const size_t b_q = up(sizeof(int32_t) * (size_t)n), b_hit = up(sizeof(int2) * (size_t)dcap);
	const size_t b_cnt = up(sizeof(int32_t) * (size_t)dcap), b_set = up(sizeof(double) * 12 * (size_t)dcap);
	char *base = nullptr;
	HIP_OR_FAIL(hipMalloc(&base, b_q + b_hit + b_cnt + b_set + 256));
	struct Free { char *p; ~Free() { (void)hipFree(p); } } release{base};
	int32_t *d_q = reinterpret_cast<int32_t *>(base);
	int2 *d_hits = reinterpret_cast<int2 *>(base + b_q);
	int32_t *d_cnt = reinterpret_cast<int32_t *>(base + b_q + b_hit);
	double *d_set = reinterpret_cast<double *>(base + b_q + b_hit + b_cnt);
	unsigned long long *d_count = reinterpret_cast<unsigned long long *>(base + b_q + b_hit + b_cnt + b_set);
	HIP_OR_FAIL(hipMemcpy(d_q, shapes, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
	HIP_OR_FAIL(hipMemsetAsync(d_count, 0, sizeof(unsigned long long), w->stream));
	QueryIn in{};
	in.shape = d_q;
	HIP_OR_FAIL(query_all(w->W, T, n, in, 3, d_hits, nullptr, dcap, d_count, w->stream, d_cnt, d_set));
	rc = sync_world(w);
	if (rc != CPGPU_OK) return rc;
	unsigned long long total = 0;
	HIP_OR_FAIL(hipMemcpy(&total, d_count, sizeof(total), hipMemcpyDeviceToHost));
	if (n_out) *n_out = (int64_t)total;
	const int64_t got = std::min<int64_t>((int64_t)total, dcap);
	if (cap == 0 || got == 0) return CPGPU_OK;
	std::vector<int2> h(got);
	std::vector<int32_t> hc(got);
	std::vector<double> hs(12 * got);
	HIP_OR_FAIL(hipMemcpy(h.data(), d_hits, sizeof(int2) * got, hipMemcpyDeviceToHost));
	HIP_OR_FAIL(hipMemcpy(hc.data(), d_cnt, sizeof(int32_t) * got, hipMemcpyDeviceToHost));
	HIP_OR_FAIL(hipMemcpy(hs.data(), d_set, sizeof(double) * 12 * got, hipMemcpyDeviceToHost));
	// by query, then shape: the device appended them in no fixed order
	std::vector<int64_t> order(got);
	for (int64_t k = 0; k < got; k++) order[k] = k;
	std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
		return h[x].x != h[y].x ? h[x].x < h[y].x : h[x].y < h[y].y;
	});
	for (int64_t k = 0; k < got && k < cap; k++) {
		const int64_t o = order[k];
		hits[2 * k] = h[o].x;
		hits[2 * k + 1] = h[o].y;
		counts[k] = hc[o];
		for (int i = 0; i < 12; i++) sets[12 * k + i] = hs[12 * o + i];
	}
	return CPGPU_OK;
}

int cpgpuWorldReadConstraints(cpgpuWorld *w, double *jacc, int32_t *colour)
{
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	std::vector<Joint> J(w->nc);
	if (w->nc) HIP_OR_FAIL(hipMemcpy(J.data(), w->W.joints, sizeof(Joint) * w->nc, hipMemcpyDeviceToHost));
	for (int c = 0; c < w->nc; c++) {
		const Joint &j = J[c];
		if (jacc) {
			if (j.type == CPGPU_JOINT_PIVOT || j.type == CPGPU_JOINT_GROOVE) { jacc[2 * c] = j.jAx; jacc[2 * c + 1] = j.jAy; }
			else if (j.type == CPGPU_JOINT_SLIDE || j.type == CPGPU_JOINT_PIN) { jacc[2 * c] = j.jnAcc; jacc[2 * c + 1] = 0.0; }
			else if (j.type == CPGPU_JOINT_RATCHET) { jacc[2 * c] = j.jAcc; jacc[2 * c + 1] = j.angle; }
			else { jacc[2 * c] = j.jAcc; jacc[2 * c + 1] = 0.0; }
		}
		if (colour) colour[c] = w->h_jcolour[c];
	}
	return CPGPU_OK;
}

int cpgpuWorldSetTiming(cpgpuWorld *w, int32_t enable)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	if (enable < CPGPU_TIMING_OFF || enable > CPGPU_TIMING_SOLVE)
		return fail(CPGPU_ERR_INVALID, "unknown timing mode " + std::to_string(enable));
	if (enable != CPGPU_TIMING_OFF && w->graph_on)
		return fail(CPGPU_ERR_INVALID, "steps captured in graphs are timed as a whole (cpgpuWorldSetGraph 0 to time "
		                               "each launch)");
	w->timer.enabled = enable != CPGPU_TIMING_OFF;
	w->timer.solve_only = enable == CPGPU_TIMING_SOLVE;
	w->timer.acc.clear();
	return CPGPU_OK;
}

int cpgpuWorldSetGridSchedule(cpgpuWorld *w, int32_t schedule)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (schedule != CPGPU_GRID_FLOW && schedule != CPGPU_GRID_PHASES)
		return fail(CPGPU_ERR_INVALID, "unknown grid schedule " + std::to_string(schedule));
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	w->grid_flow = schedule == CPGPU_GRID_FLOW;
	w->graph_dirty = true;
	return CPGPU_OK;
}

int cpgpuWorldGetBuckets(cpgpuWorld *w, int32_t *max_bodies, int32_t *threads, int32_t *n_worlds, int32_t *lds_bodies)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	for (int q = 0; q < CPGPU_MAX_BUCKETS; q++) {
		const bool on = q < w->n_buckets;
		if (max_bodies) max_bodies[q] = on ? w->buckets[q].max_bodies : 0;
		if (threads) threads[q] = on ? w->buckets[q].threads : 0;
		if (n_worlds) n_worlds[q] = on ? w->buckets[q].n : 0;
		if (lds_bodies) lds_bodies[q] = on ? w->buckets[q].lds_cap : 0;
	}
	return CPGPU_OK;
}

static bool valid_threads(int t) { return t == 64 || t == 128 || t == 256 || t == 512 || t == 1024; }

int cpgpuWorldSetBuckets(cpgpuWorld *w, int32_t n, const int32_t *max_bodies, const int32_t *threads)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (n < 1 || n > CPGPU_MAX_BUCKETS) return fail(CPGPU_ERR_INVALID, "bucket count out of range");
	for (int q = 0; q < n; q++) {
		if (!valid_threads(threads[q])) return fail(CPGPU_ERR_INVALID, "bucket threads must be 64..1024, a power of 2");
		if (q > 0 && max_bodies[q] <= max_bodies[q - 1]) return fail(CPGPU_ERR_INVALID, "bucket limits must rise");
	}
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	w->n_buckets = n;
	for (int q = 0; q < n; q++) {
		w->buckets[q].max_bodies = max_bodies[q];
		w->buckets[q].threads = threads[q];
	}
	HIP_OR_FAIL(build_buckets(w));
	w->graph_dirty = true;
	return CPGPU_OK;
}

int cpgpuWorldAutotune(cpgpuWorld *w, double dt, int32_t trial_steps, int32_t *threads_out)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (trial_steps < 1) return fail(CPGPU_ERR_INVALID, "trial_steps must be positive");
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	static const int candidates[] = {64, 128, 256, 512, 1024};
	static const char *names[CPGPU_MAX_BUCKETS] = {"solver_worlds_b0", "solver_worlds_b1", "solver_worlds_b2",
	                                               "solver_worlds_b3"};
	const bool was_timing = w->timer.enabled;
	const bool was_graph = w->graph_on;
	w->graph_on = false;   // trials time every bucket's launches: eager steps
	w->graph_dirty = true;
	auto saved = w->timer.acc;
	double best[CPGPU_MAX_BUCKETS];
	int best_t[CPGPU_MAX_BUCKETS];
	for (int q = 0; q < CPGPU_MAX_BUCKETS; q++) { best[q] = 1e300; best_t[q] = q < w->n_buckets ? w->buckets[q].threads : 0; }
	w->timer.enabled = true;
	for (int t : candidates) {
		// every bucket tries the same workgroup size at once (the buckets run concurrently)
		for (int q = 0; q < w->n_buckets; q++) w->buckets[q].threads = t;
		w->timer.acc.clear();
		for (int k = 0; k < trial_steps; k++) {
			rc = cpgpuWorldStep(w, dt);
			if (rc != CPGPU_OK) { w->timer.enabled = was_timing; w->graph_on = was_graph; return rc; }
		}
		rc = wait_world(w);
		if (rc != CPGPU_OK) { w->timer.enabled = was_timing; w->graph_on = was_graph; return rc; }
		for (int q = 0; q < w->n_buckets; q++) {
			auto it = w->timer.acc.find(names[q]);
			if (it == w->timer.acc.end() || it->second.second == 0) continue;
			const double per = it->second.first / it->second.second;
			if (per < best[q]) { best[q] = per; best_t[q] = t; }
		}
	}
	for (int q = 0; q < w->n_buckets; q++) w->buckets[q].threads = best_t[q];
	if (threads_out)
		for (int q = 0; q < CPGPU_MAX_BUCKETS; q++) threads_out[q] = q < w->n_buckets ? best_t[q] : 0;
	w->timer.acc = saved;
	w->timer.enabled = was_timing;
	w->graph_on = was_graph;
	return CPGPU_OK;
}

int cpgpuWorldTapeBegin(cpgpuWorld *w, int32_t checkpoint_every)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (checkpoint_every < 1) return fail(CPGPU_ERR_INVALID, "checkpoint_every must be positive");
	if (sizeof(real) != 8) return fail(CPGPU_ERR_INVALID, "the adjoint needs a CPGPU_FP64 world");
	if (w->graph_on) return fail(CPGPU_ERR_INVALID, "a taped rollout is stepped without graphs (cpgpuWorldSetGraph 0)");
	if (w->other_joints)
		return fail(CPGPU_ERR_INVALID, "the adjoint covers pivot, damped-spring, slide and gear joints only");
	if (w->n_handlers > 0)
		return fail(CPGPU_ERR_INVALID, "the adjoint covers spaces without collision handlers only");
	if (w->n_large_worlds > 0)
		return fail(CPGPU_ERR_INVALID, "the adjoint supports spaces solved one workgroup each (at most " +
		                                   std::to_string(w->small_world_max_bodies) + " dynamic bodies)");
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	tape_clear(w);
	auto &T = w->tape;
	T.every = checkpoint_every;
	// the arbiters alive now (last step's solved and cached ones) are recorded
	// first, so that warm starts from them find a recorded source
	if (w->steps > 0) {
		const int prev = w->cur ^ 1;
		const int n_act = w->h_counts[3 * prev + 1], n_cached = w->h_counts[3 * prev + 2];
		HIP_OR_FAIL(tape_reserve(w, (long)n_act + n_cached));
		HIP_OR_FAIL(adjoint_tape_record(w->act[prev], nullptr, n_act, T.arb, 0, w->stream));
		HIP_OR_FAIL(adjoint_tape_record(w->cached[prev], nullptr, n_cached, T.arb, n_act, w->stream));
		T.n = (long)n_act + n_cached;
	}
	BodyState *ck = nullptr;
	HIP_OR_FAIL(take_spare(w->spare_ckpt, ck, (size_t)w->nb));
	T.ckpt.push_back(ck);
	HIP_OR_FAIL(adjoint_state_save(w->W, ck, w->stream));
	if (!w->d_replay_n) HIP_OR_FAIL(hipMalloc(&w->d_replay_n, sizeof(int32_t)));
	T.on = true;
	return wait_world(w);
}

int cpgpuWorldTapeEnd(cpgpuWorld *w)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	tape_clear(w);
	return CPGPU_OK;
}

// Replay step t of the tape from the bodies now in W: position update, the
// recorded arbiters with their warm starts, the solve with pre-state records.
static int replay_step(cpgpuWorld *w, int t, Arbiter *act, int32_t *colour, int32_t *order, int32_t *warb_off,
                       ArbRec *rec, JointRec *jrec, KernelTimer *timer)
{
	auto &T = w->tape;
	DevWorld &W = w->W;
	HIP_OR_FAIL(hipMemcpyAsync(w->d_ws, T.ws[t].data(), sizeof(WorldStep) * w->nw, hipMemcpyHostToDevice,
	                           w->stream));
	if (w->n_dyn_bodies > 0)
		k_update_positions<<<blocks_for(w->n_dyn_bodies, 256), 256, 0, w->stream>>>(W, w->d_dyn_bodies,
		                                                                           w->n_dyn_bodies);
	const int n = T.count[t];
	HIP_OR_FAIL(adjoint_replay_prepare(T.arb, (int)T.off[t], n, act, w->stream));
	HIP_OR_FAIL(hipMemcpyAsync(w->d_replay_n, &T.count[t], sizeof(int32_t), hipMemcpyHostToDevice, w->stream));
	SolveBuffers S = solve_buffers(w, act, w->d_replay_n, colour, order, warb_off);
	S.rec = rec;
	S.rec_stride = n;
	S.jrec = jrec;
	if (w->nw > 1) HIP_OR_FAIL(solver_group(W, S, n, w->nw, false, w->sortbuf, w->stream, nullptr));
	TIMED(timer, "replay_solve", w->stream,
	      HIP_OR_FAIL(solver_step(W, S, w->buckets, w->n_buckets, 0, w->stream, w->aux, w->fork_join, nullptr)));
	HIP_OR_FAIL(hipStreamSynchronize(w->stream));
	return CPGPU_OK;
}

int cpgpuWorldBackward(cpgpuWorld *w, const double *g_p, const double *g_a, const double *g_v, const double *g_w,
                       double *d_p0, double *d_a0, double *d_v0, double *d_w0, double *d_m_inv, double *d_i_inv,
                       double *d_u)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	auto &T = w->tape;
	if (!T.on) return fail(CPGPU_ERR_INVALID, "no tape: call cpgpuWorldTapeBegin before the rollout");
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	DevWorld &W = w->W;
	KernelTimer *timer = w->timer.enabled ? &w->timer : nullptr;
	const int nb = w->nb, ns = w->ns;
	int max_n = 1, max_it = 1;
	for (int c : T.count) max_n = std::max(max_n, c);
	for (auto &s : w->sp) max_it = std::max(max_it, (int)s.iterations);
	const int K = T.every;
	// device buffers of the pass
	BodyAdj *gb = nullptr;
	double *gminv = nullptr, *giinv = nullptr, *gu = nullptr;
	double2 *gJ = nullptr;
	BodyState *live = nullptr, *start = nullptr;
	Arbiter *act = nullptr;
	int32_t *colour = nullptr, *order = nullptr, *warb = nullptr;
	ArbRec *rec = nullptr;
	void *garb = nullptr;
	const int nc = w->nc;
	JointRec *jrec = nullptr;
	Joint *jsnap = nullptr, *live_joints = nullptr;
	void *gjoint = nullptr;
	double2 *gJacc = nullptr;
	double *gP = nullptr;
	const double live_joint_dt = w->last_joint_dt;
	unsigned long long *totals_saved = nullptr;
	std::vector<void *> bufs;
	auto grab = [&](void **p, size_t bytes) -> hipError_t {
		hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
		if (e == hipSuccess) bufs.push_back(*p);
		return e;
	};
	auto release = [&]() { for (void *p : bufs) (void)hipFree(p); bufs.clear(); };
#define ADJ_TRY(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { release(); \
	return fail(CPGPU_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); } } while (0)
	ADJ_TRY(grab((void **)&gb, sizeof(BodyAdj) * nb));
	ADJ_TRY(grab((void **)&gminv, sizeof(double) * nb));
	ADJ_TRY(grab((void **)&giinv, sizeof(double) * nb));
	ADJ_TRY(grab((void **)&gu, sizeof(double) * ns));
	ADJ_TRY(grab((void **)&gJ, sizeof(double2) * 2 * (size_t)std::max<long>(T.n, 1)));
	ADJ_TRY(grab((void **)&live, sizeof(BodyState) * nb));
	ADJ_TRY(grab((void **)&start, sizeof(BodyState) * nb * (size_t)K));
	ADJ_TRY(grab((void **)&act, sizeof(Arbiter) * (size_t)max_n * K));
	ADJ_TRY(grab((void **)&colour, sizeof(int32_t) * (size_t)max_n * K));
	ADJ_TRY(grab((void **)&order, sizeof(int32_t) * (size_t)max_n * K));
	ADJ_TRY(grab((void **)&warb, sizeof(int32_t) * (size_t)(w->nw + 2) * K));
	ADJ_TRY(grab((void **)&rec, sizeof(ArbRec) * (size_t)max_n * max_it * K));
	ADJ_TRY(grab(&garb, adjoint_arb_scratch_bytes() * (size_t)max_n));
	ADJ_TRY(grab((void **)&jrec, sizeof(JointRec) * (size_t)nc * (max_it + 1) * K));
	ADJ_TRY(grab((void **)&jsnap, sizeof(Joint) * (size_t)nc * K));
	ADJ_TRY(grab((void **)&live_joints, sizeof(Joint) * (size_t)nc));
	ADJ_TRY(grab(&gjoint, adjoint_joint_scratch_bytes() * (size_t)nc));
	ADJ_TRY(grab((void **)&gJacc, sizeof(double2) * (size_t)std::max(nc, 1)));
	ADJ_TRY(grab((void **)&gP, sizeof(double) * 8 * (size_t)std::max(nc, 1)));
	ADJ_TRY(grab((void **)&totals_saved, sizeof(unsigned long long) * (2 + 2 * CPGPU_SOLVE_UNITS)));
	// replay scratch sized for the largest step
	if (max_n > w->colour_cap) {
		void *old[] = {w->colour, w->order, w->arb_by_world, w->world_keys, w->cbody, w->cprio, w->ccol};
		for (void *p : old) if (p) (void)hipFree(p);
		w->colour_cap = max_n;
		ADJ_TRY(dalloc(w->colour, w->colour_cap));
		ADJ_TRY(dalloc(w->order, w->colour_cap));
		ADJ_TRY(dalloc(w->arb_by_world, w->colour_cap));
		ADJ_TRY(dalloc(w->world_keys, w->colour_cap));
		ADJ_TRY(dalloc(w->cbody, w->colour_cap));
		ADJ_TRY(dalloc(w->cprio, w->colour_cap));
		ADJ_TRY(dalloc(w->ccol, w->colour_cap));
	}
	// seeds: the adjoint of the state after the last step
	{
		std::vector<BodyAdj> h(nb);
		for (int b = 0; b < nb; b++) {
			BodyAdj &x = h[b];
			x.p[0] = g_p ? g_p[2 * b] : 0.0; x.p[1] = g_p ? g_p[2 * b + 1] : 0.0;
			x.a = g_a ? g_a[b] : 0.0;
			x.v[0] = g_v ? g_v[2 * b] : 0.0; x.v[1] = g_v ? g_v[2 * b + 1] : 0.0;
			x.w = g_w ? g_w[b] : 0.0;
			x.vb[0] = x.vb[1] = x.wb = 0.0;
		}
		ADJ_TRY(hipMemcpy(gb, h.data(), sizeof(BodyAdj) * nb, hipMemcpyHostToDevice));
	}
	ADJ_TRY(hipMemset(gminv, 0, sizeof(double) * nb));
	ADJ_TRY(hipMemset(giinv, 0, sizeof(double) * nb));
	ADJ_TRY(hipMemset(gu, 0, sizeof(double) * ns));
	ADJ_TRY(hipMemset(gJ, 0, sizeof(double2) * 2 * (size_t)std::max<long>(T.n, 1)));
	ADJ_TRY(hipMemset(gJacc, 0, sizeof(double2) * (size_t)std::max(nc, 1)));
	ADJ_TRY(hipMemset(gP, 0, sizeof(double) * 8 * (size_t)std::max(nc, 1)));
	if (nc > 0)
		ADJ_TRY(hipMemcpyAsync(live_joints, W.joints, sizeof(Joint) * nc, hipMemcpyDeviceToDevice, w->stream));
	ADJ_TRY(adjoint_state_save(W, live, w->stream));
	ADJ_TRY(hipMemcpyAsync(totals_saved, W.totals, sizeof(unsigned long long) * (2 + 2 * CPGPU_SOLVE_UNITS),
	                       hipMemcpyDeviceToDevice, w->stream));
	const int n_seg = (T.steps + K - 1) / K;
	int result = CPGPU_OK;
	for (int seg = n_seg - 1; seg >= 0 && result == CPGPU_OK; seg--) {
		const int t0 = seg * K, t1 = std::min(t0 + K, T.steps);
		ADJ_TRY(adjoint_state_load(W, T.ckpt[seg], w->stream));
		double replay_dt = T.ws[t0][0].dt;   // the joint checkpoint holds this step's bias coefficients
		if (nc > 0)
			ADJ_TRY(hipMemcpyAsync(W.joints, T.jckpt[seg], sizeof(Joint) * nc, hipMemcpyDeviceToDevice, w->stream));
		for (int t = t0; t < t1 && result == CPGPU_OK; t++) {
			const int q = t - t0;
			if (nc > 0 && T.ws[t][0].dt != replay_dt) {
				replay_dt = T.ws[t][0].dt;
				ADJ_TRY(set_joint_bias(w, replay_dt));
			}
			ADJ_TRY(adjoint_state_save(W, start + (size_t)nb * q, w->stream));
			result = replay_step(w, t, act + (size_t)max_n * q, colour + (size_t)max_n * q, order + (size_t)max_n * q,
			                     warb + (size_t)(w->nw + 2) * q, rec + (size_t)max_n * max_it * q,
			                     nc > 0 ? jrec + (size_t)nc * (max_it + 1) * q : nullptr, timer);
			if (nc > 0)
				ADJ_TRY(hipMemcpyAsync(jsnap + (size_t)nc * q, W.joints, sizeof(Joint) * nc, hipMemcpyDeviceToDevice,
				                       w->stream));
		}
		for (int t = t1 - 1; t >= t0 && result == CPGPU_OK; t--) {
			const int q = t - t0;
			// stream-ordered: the previous step's reverse may still be reading the parameters
			ADJ_TRY(hipMemcpyAsync(w->d_ws, T.ws[t].data(), sizeof(WorldStep) * w->nw, hipMemcpyHostToDevice,
			                       w->stream));
			AdjStep A;
			A.n = T.count[t];
			A.act = act + (size_t)max_n * q;
			A.order = order + (size_t)max_n * q;
			A.colour = colour + (size_t)max_n * q;
			A.warb_off = w->nw > 1 ? warb + (size_t)(w->nw + 2) * q : nullptr;
			A.rec = rec + (size_t)max_n * max_it * q;
			A.start = start + (size_t)nb * q;
			A.tape = T.arb;
			A.garb = garb;
			A.gb = gb; A.gminv = gminv; A.giinv = giinv; A.gu = gu; A.gJ = gJ;
			A.nc = nc;
			A.joints = jsnap + (size_t)nc * q;
			A.jrec = jrec + (size_t)nc * (max_it + 1) * q;
			A.jw_order = w->d_jw_order; A.jw_col = w->d_jw_col; A.jw_cbase = w->d_jw_cbase; A.jw_ncol = w->d_jw_ncol;
			A.gjoint = gjoint; A.gJacc = gJacc; A.gP = gP;
			TIMED(timer, "reverse_solve", w->stream, ADJ_TRY(adjoint_reverse_solve(W, A, w->nw, w->stream)));
			TIMED(timer, "reverse_geometry", w->stream, ADJ_TRY(adjoint_reverse_geometry(W, A, w->nw, w->stream)));
		}
		if (result == CPGPU_OK) result = sync_world(w);
	}
	// results, then the world as it was
	if (result == CPGPU_OK) {
		std::vector<BodyAdj> h(nb);
		ADJ_TRY(hipMemcpy(h.data(), gb, sizeof(BodyAdj) * nb, hipMemcpyDeviceToHost));
		for (int b = 0; b < nb; b++) {
			if (d_p0) { d_p0[2 * b] = h[b].p[0]; d_p0[2 * b + 1] = h[b].p[1]; }
			if (d_a0) d_a0[b] = h[b].a;
			if (d_v0) { d_v0[2 * b] = h[b].v[0]; d_v0[2 * b + 1] = h[b].v[1]; }
			if (d_w0) d_w0[b] = h[b].w;
		}
		if (d_m_inv) ADJ_TRY(hipMemcpy(d_m_inv, gminv, sizeof(double) * nb, hipMemcpyDeviceToHost));
		if (d_i_inv) ADJ_TRY(hipMemcpy(d_i_inv, giinv, sizeof(double) * nb, hipMemcpyDeviceToHost));
		if (d_u) ADJ_TRY(hipMemcpy(d_u, gu, sizeof(double) * ns, hipMemcpyDeviceToHost));
		w->h_joint_grads.assign(8 * (size_t)nc, 0.0);
		if (nc > 0)
			ADJ_TRY(hipMemcpy(w->h_joint_grads.data(), gP, sizeof(double) * 8 * nc, hipMemcpyDeviceToHost));
	}
	if (nc > 0) {
		ADJ_TRY(hipMemcpyAsync(W.joints, live_joints, sizeof(Joint) * nc, hipMemcpyDeviceToDevice, w->stream));
		w->last_joint_dt = live_joint_dt;
	}
	ADJ_TRY(adjoint_state_load(W, live, w->stream));
	ADJ_TRY(hipMemcpyAsync(W.totals, totals_saved, sizeof(unsigned long long) * (2 + 2 * CPGPU_SOLVE_UNITS),
	                       hipMemcpyDeviceToDevice, w->stream));
	ADJ_TRY(hipMemcpyAsync(w->d_ws, w->h_ws, sizeof(WorldStep) * w->nw, hipMemcpyHostToDevice, w->stream));
	ADJ_TRY(hipStreamSynchronize(w->stream));
	release();
#undef ADJ_TRY
	return result;
}

int cpgpuWorldGetSolveCounts(cpgpuWorld *w, int64_t *contacts, int64_t *bodies)
{
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	unsigned long long t[2 + 2 * CPGPU_SOLVE_UNITS];
	HIP_OR_FAIL(hipMemcpy(t, w->W.totals, sizeof(t), hipMemcpyDeviceToHost));
	for (int u = 0; u < CPGPU_SOLVE_UNITS; u++) {
		if (contacts) contacts[u] = (int64_t)t[2 + 2 * u];
		if (bodies) bodies[u] = (int64_t)t[3 + 2 * u];
	}
	return CPGPU_OK;
}

int cpgpuWorldGetJointGrads(cpgpuWorld *w, double *d_params)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (w->h_joint_grads.size() != 8 * (size_t)w->nc)
		return fail(CPGPU_ERR_INVALID, "no joint gradients: call cpgpuWorldBackward first");
	if (d_params) std::copy(w->h_joint_grads.begin(), w->h_joint_grads.end(), d_params);
	return CPGPU_OK;
}

int cpgpuWorldGetTiming(cpgpuWorld *w, int32_t cap, int32_t *n_out, double *ms, int64_t *launches, char *names,
                        int32_t names_cap)
{
	int rc = wait_world(w);
	if (rc != CPGPU_OK) return rc;
	std::string all;
	int k = 0;
	for (auto &kv : w->timer.acc) {
		if (k < cap) {
			if (ms) ms[k] = kv.second.first;
			if (launches) launches[k] = kv.second.second;
		}
		all += kv.first;
		all += "\n";
		k++;
	}
	if (n_out) *n_out = k;
	if (names && names_cap > 0) {
		const size_t m = std::min<size_t>(all.size(), (size_t)names_cap - 1);
		memcpy(names, all.data(), m);
		names[m] = 0;
	}
	return CPGPU_OK;
}

} // namespace CPGPU_NS
