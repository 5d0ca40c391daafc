// query.hip -- spatial queries on the step's LBVH (src/cpSpaceQuery.c), one
// query per lane.
//
// The reference asks its spatial indexes for candidates and checks each
// exactly with the shape's own query (cpShapePointQuery, cpShapeSegmentQuery,
// src/cpShape.c:223-257, the class functions of cpShape.c and
// cpPolyShape.c).  Here a lane walks the dynamic and the static tree of its
// query's space, built by the last step over the shapes' tight boxes, and
// runs the same exact checks, restated operation for operation so points,
// distances, gradients, normals and alphas are the reference's bits.  The
// trees only prune: every shape whose exact check could pass is visited, so
// the results are those of checking every shape.  Of equal nearest
// distances (first alphas) the lowest shape index is kept; the reference
// keeps whichever its tree met first.
#include "cpgpu_internal.h"
#include "narrowphase.h"

namespace CPGPU_NS {

namespace {

constexpr int QUERY_STACK = 48;
constexpr int QUERY_THREADS = 256;

struct PQ { V2 point; real d; V2 grad; };
struct SQ { int hit; V2 point, normal; real alpha; };

#define Q_MAGIC_EPSILON 1e-5   // chipmunk_private.h:32

// cpClosetPointOnSegment (include/chipmunk/chipmunk.h:182-188)
CPD V2 closest_on_segment(V2 p, V2 a, V2 b)
{
	const V2 delta = vsub(a, b);
	const real t = fclamp01(vdot(delta, vsub(p, b)) / vlengthsq(delta));
	return vadd(b, vmult(delta, t));
}

// cpShapePointQuery (cpShape.c:223-234): circle cpShape.c:299-312, segment
// cpShape.c:408-424, polygon cpPolyShape.c:67-103
CPD PQ point_query(const DevWorld &W, const ShapeView &s, int idx, V2 p)
{
	PQ q;
	if (s.type == CPGPU_SHAPE_CIRCLE) {
		const V2 delta = vsub(p, s.tc);
		const real d = vlength(delta), r = s.r;
		const real r_over_d = d > real(0.0) ? r / d : r;
		q.point = vadd(s.tc, vmult(delta, r_over_d));
		q.d = d - r;
		q.grad = d > real(Q_MAGIC_EPSILON) ? vmult(delta, real(1.0) / d) : v2(0.0, 1.0);
	} else if (s.type == CPGPU_SHAPE_SEGMENT) {
		const V2 closest = closest_on_segment(p, s.ta, s.tb);
		const V2 delta = vsub(p, closest);
		const real d = vlength(delta), r = s.r;
		const V2 g = vmult(delta, real(1.0) / d);
		q.point = d != real(0.0) ? vadd(closest, vmult(g, r)) : closest;
		q.d = d - r;
		// the segment's body-local normal, as the reference reads seg->n
		const real *lg = W.s_geom + 10 * (size_t)idx;
		q.grad = d > real(Q_MAGIC_EPSILON) ? g : v2(lg[4], lg[5]);
	} else {
		const int count = s.count;
		V2 v0 = s.wv[count - 1];
		real minDist = real(INFINITY);
		V2 closestPoint = v2(0.0, 0.0), closestNormal = v2(0.0, 0.0);
		bool outside = false;
		for (int i = 0; i < count; i++) {
			const V2 v1 = s.wv[i];
			outside = outside || (vdot(s.wn[i], vsub(p, v1)) > real(0.0));
			const V2 closest = closest_on_segment(p, v0, v1);
			const real dist = vlength(vsub(p, closest));
			if (dist < minDist) { minDist = dist; closestPoint = closest; closestNormal = s.wn[i]; }
			v0 = v1;
		}
		const real dist = outside ? minDist : -minDist;
		const V2 g = vmult(vsub(p, closestPoint), real(1.0) / dist);
		q.point = vadd(closestPoint, vmult(g, s.r));
		q.d = dist - s.r;
		q.grad = minDist > real(Q_MAGIC_EPSILON) ? g : closestNormal;
	}
	return q;
}

// CircleSegmentQuery (chipmunk_private.h:121-142)
CPD void circle_segment_query(V2 center, real r1, V2 a, V2 b, real r2, SQ &info)
{
	const V2 da = vsub(a, center), db = vsub(b, center);
	const real rsum = r1 + r2;
	const real qa = vdot(da, da) - real(2.0) * vdot(da, db) + vdot(db, db);
	const real qb = vdot(da, db) - vdot(da, da);
	const real det = qb * qb - qa * (vdot(da, da) - rsum * rsum);
	if (det >= real(0.0)) {
		const real t = (-qb - csqrt(det)) / qa;
		if (real(0.0) <= t && t <= real(1.0)) {
			const V2 n = vnormalize(vlerp(da, db, t));
			info.hit = 1;
			info.point = vsub(vlerp(a, b, t), vmult(n, r2));
			info.normal = n;
			info.alpha = t;
		}
	}
}

CPD SQ sq_blank(V2 b)
{
	SQ q;
	q.hit = 0; q.point = b; q.normal = v2(0.0, 0.0); q.alpha = real(1.0);
	return q;
}

// cpShapeSegmentQuery (cpShape.c:238-257): circle cpShape.c:315-318, segment
// cpShape.c:426-466, polygon cpPolyShape.c:106-145
CPD SQ segment_query(const DevWorld &W, const ShapeView &s, int idx, V2 a, V2 b, real radius)
{
	SQ info = sq_blank(b);
	const PQ nearest = point_query(W, s, idx, a);
	if (nearest.d <= radius) {
		info.hit = 1;
		info.alpha = real(0.0);
		info.normal = vnormalize(vsub(a, nearest.point));
		return info;
	}
	if (s.type == CPGPU_SHAPE_CIRCLE) {
		circle_segment_query(s.tc, s.r, a, b, radius, info);
	} else if (s.type == CPGPU_SHAPE_SEGMENT) {
		const V2 n = s.tn;
		const real d = vdot(vsub(s.ta, a), n);
		const real r = s.r + radius;
		const V2 flipped_n = d > real(0.0) ? vneg(n) : n;
		const V2 seg_offset = vsub(vmult(flipped_n, r), a);
		const V2 seg_a = vadd(s.ta, seg_offset), seg_b = vadd(s.tb, seg_offset);
		const V2 delta = vsub(b, a);
		if (vcross(delta, seg_a) * vcross(delta, seg_b) <= real(0.0)) {
			const real d_offset = d + (d > real(0.0) ? -r : r);
			const real ad = -d_offset;
			const real bd = vdot(delta, n) - d_offset;
			if (ad * bd < real(0.0)) {
				const real t = ad / (ad - bd);
				info.hit = 1;
				info.point = vsub(vlerp(a, b, t), vmult(flipped_n, radius));
				info.normal = flipped_n;
				info.alpha = t;
			}
		} else if (r != real(0.0)) {
			SQ info1 = sq_blank(b), info2 = sq_blank(b);
			circle_segment_query(s.ta, s.r, a, b, radius, info1);
			circle_segment_query(s.tb, s.r, a, b, radius, info2);
			info = info1.alpha < info2.alpha ? info1 : info2;
		}
	} else {
		const int count = s.count;
		const real r = s.r, rsum = r + radius;
		for (int i = 0; i < count; i++) {
			const V2 n = s.wn[i];
			const real an = vdot(a, n);
			const real d = an - vdot(s.wv[i], n) - rsum;
			if (d < real(0.0)) continue;
			const real bn = vdot(b, n);
			const real t = d / fmaxc(an - bn, CP_REAL_MIN);
			if (t < real(0.0) || real(1.0) < t) continue;
			const V2 point = vlerp(a, b, t);
			const real dt = vcross(n, point);
			const real dtMin = vcross(n, s.wv[(i - 1 + count) % count]);
			const real dtMax = vcross(n, s.wv[i]);
			if (dtMin <= dt && dt <= dtMax) {
				info.hit = 1;
				info.point = vsub(vlerp(a, b, t), vmult(n, radius));
				info.normal = n;
				info.alpha = t;
			}
		}
		if (rsum > real(0.0)) {
			for (int i = 0; i < count; i++) {
				SQ circle_info = sq_blank(b);
				circle_segment_query(s.wv[i], r, a, b, radius, circle_info);
				if (circle_info.alpha < info.alpha) info = circle_info;
			}
		}
	}
	return info;
}

// cpShapeFilterReject (chipmunk_private.h:144-155) of a shape and a query's filter
CPD bool filter_reject(const DevWorld &W, int s, const cpgpuShapeFilter &f)
{
	const uint64_t g = W.s_group[s];
	return (g != 0 && g == f.group) || (W.s_cats[s] & f.mask) == 0 || (f.categories & W.s_mask[s]) == 0;
}

CPD bool box_meets(real4 a, real4 b)   // cpBBIntersects (cpBB.h:58-61)
{
	return a.x <= b.z && b.x <= a.z && a.y <= b.w && b.y <= a.w;
}

// Every leaf of space w's two trees whose box meets qbox: visit(shape).  The
// stack lives in LDS as [depth][lane].
template <class Visit>
CPD void walk_trees(const QueryTrees &T, int w, real4 qbox, int32_t *stk, int32_t *err, Visit visit)
{
	for (int tree = 0; tree < 2; tree++) {
		const int n = tree ? T.ns : T.nd;
		if (n == 0) continue;
		const int root = (tree ? T.sroot : T.droot)[w];
		if (root < 0) continue;
		const PairNode *node = tree ? T.snode : T.dnode;
		const real4 *box = tree ? T.sbox : T.dbox;
		const int32_t *leaf = tree ? T.sshape : T.dshape;
		if (root >= n - 1) {
			if (box_meets(qbox, box[root])) visit(leaf[root - (n - 1)]);
			continue;
		}
		if (!box_meets(qbox, box[root])) continue;
		int sp = 0;
		stk[QUERY_THREADS * sp++] = root;
		while (sp > 0) {
			const PairNode rec = node[stk[QUERY_THREADS * --sp]];
			for (int side = 0; side < 2; side++) {
				if (!box_meets(qbox, rec.box[side])) continue;
				const int c = rec.child[side];
				if (c >= n - 1) {
					visit(leaf[c - (n - 1)]);
				} else if (sp >= QUERY_STACK) {
					raise_err(err, CPGPU_ERR_TRAVERSAL_STACK);
				} else {
					stk[QUERY_THREADS * sp++] = c;
				}
			}
		}
	}
}

CPD cpgpuShapeFilter query_filter(const cpgpuShapeFilter *filters, int q)
{
	if (filters) return filters[q];
	cpgpuShapeFilter all;
	all.group = 0; all.categories = 0xffffffffu; all.mask = 0xffffffffu;
	return all;
}

// a conservative box around a segment of the given radius
CPD real4 segment_box(V2 a, V2 b, real radius)
{
	// shapes within the radius lie in the segment's box grown by it; a little more keeps
	// the box's rounding from dropping one that grazes it
	const real m = radius + real(1e-6) * (real(1.0) + fmaxc(fmaxc(fabs(a.x), fabs(a.y)), fmaxc(fabs(b.x), fabs(b.y))));
	real4 r;
	r.x = fminc(a.x, b.x) - m; r.y = fminc(a.y, b.y) - m;
	r.z = fmaxc(a.x, b.x) + m; r.w = fmaxc(a.y, b.y) + m;
	return r;
}

__global__ void __launch_bounds__(QUERY_THREADS) k_point_nearest(const DevWorld W, const QueryTrees T, int n,
                                                                  const QueryIn in, int32_t *shape_out,
                                                                  double *info_out)
{
	__shared__ int32_t stk[QUERY_STACK][QUERY_THREADS];
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n) return;
	const int w = in.space ? in.space[q] : 0;
	const V2 p = v2((real)in.a[2 * q], (real)in.a[2 * q + 1]);
	const real maxd = (real)in.scalar[q];
	const cpgpuShapeFilter f = query_filter(in.filter, q);
	const real ext = fmaxc(maxd, real(0.0));   // cpBBNewForCircle(point, cpfmax(maxDistance, 0))
	real4 qbox;
	qbox.x = p.x - ext; qbox.y = p.y - ext; qbox.z = p.x + ext; qbox.w = p.y + ext;
	int best = -1;
	PQ out;
	out.point = v2(0.0, 0.0); out.d = maxd; out.grad = v2(0.0, 0.0);
	walk_trees(T, w, qbox, &stk[0][threadIdx.x], W.err, [&](int s) {
		if (filter_reject(W, s, f) || (W.s_sensor && W.s_sensor[s])) return;   // cpSpaceQuery.c:64
		const PQ r = point_query(W, load_shape(W, s), s, p);
		if (r.d < out.d || (best >= 0 && r.d == out.d && s < best)) { out = r; best = s; }
	});
	shape_out[q] = best;
	double *o = info_out + 5 * (size_t)q;
	o[0] = (double)out.point.x; o[1] = (double)out.point.y; o[2] = (double)out.d;
	o[3] = (double)out.grad.x; o[4] = (double)out.grad.y;
}

__global__ void __launch_bounds__(QUERY_THREADS) k_segment_first(const DevWorld W, const QueryTrees T, int n,
                                                                  const QueryIn in, int32_t *shape_out,
                                                                  double *info_out)
{
	__shared__ int32_t stk[QUERY_STACK][QUERY_THREADS];
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n) return;
	const int w = in.space ? in.space[q] : 0;
	const V2 a = v2((real)in.a[2 * q], (real)in.a[2 * q + 1]);
	const V2 b = v2((real)in.b[2 * q], (real)in.b[2 * q + 1]);
	const real radius = (real)in.scalar[q];
	const cpgpuShapeFilter f = query_filter(in.filter, q);
	int best = -1;
	SQ out = sq_blank(b);
	walk_trees(T, w, segment_box(a, b, radius), &stk[0][threadIdx.x], W.err, [&](int s) {
		if (filter_reject(W, s, f) || (W.s_sensor && W.s_sensor[s])) return;   // cpSpaceQuery.c:145
		const SQ r = segment_query(W, load_shape(W, s), s, a, b, radius);
		if (r.hit && (r.alpha < out.alpha || (best >= 0 && r.alpha == out.alpha && s < best))) { out = r; best = s; }
	});
	shape_out[q] = best;
	double *o = info_out + 5 * (size_t)q;
	o[0] = (double)out.point.x; o[1] = (double)out.point.y; o[2] = (double)out.normal.x;
	o[3] = (double)out.normal.y; o[4] = (double)out.alpha;
}

// every hit of every query, appended (query, shape) with its five numbers, or
// (shape queries) its contact count and contact point set
struct HitsOut {
	int2 *hit;
	double *info;
	int64_t cap;
	unsigned long long *count;
	int32_t *n_pts;
	double *set;
};

CPD void append_hit(const HitsOut &H, int q, int s, const double *v)
{
	const unsigned long long k = atomicAdd(H.count, 1ull);
	if ((int64_t)k >= H.cap) return;
	H.hit[k] = make_int2(q, s);
	if (H.info && v)
		for (int i = 0; i < 5; i++) H.info[5 * k + i] = v[i];
}

CPD void append_contact_hit(const HitsOut &H, int q, int s, int n, const double *set)
{
	const unsigned long long k = atomicAdd(H.count, 1ull);
	if ((int64_t)k >= H.cap) return;
	H.hit[k] = make_int2(q, s);
	H.n_pts[k] = n;
	for (int i = 0; i < 12; i++) H.set[12 * k + i] = set[i];
}

// kind 0: cpSpacePointQuery (cpSpaceQuery.c:33-57); 1: cpSpaceSegmentQuery
// (cpSpaceQuery.c:107-135); 2: cpSpaceBBQuery (cpSpaceQuery.c:186-206);
// 3: cpSpaceShapeQuery (cpSpaceQuery.c:231-246) of a shape of the world, put
// at its body's present transform as cpShapeUpdate does there (the other
// shapes stay at their cached pose): every other shape of its space that
// passes the two shapes' filters (cpShapeFilterReject(a->filter, b->filter))
// and touches it, with the contact point set of cpShapesCollide(query shape,
// shape)
__global__ void __launch_bounds__(QUERY_THREADS) k_query_all(const DevWorld W, const QueryTrees T, int n,
                                                              const QueryIn in, int kind, HitsOut H)
{
	__shared__ int32_t stk[QUERY_STACK][QUERY_THREADS];
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n) return;
	int32_t *stack = &stk[0][threadIdx.x];
	if (kind == 3) {
		const int a = in.shape[q];
		const uint64_t ga = W.s_group[a];
		const uint32_t ca = W.s_cats[a], ma = W.s_mask[a];
		PosedPoly P;
		const ShapeView va = pose_shape(W, a, P);
		const real4 qbox = make_real4(va.bb[0], va.bb[1], va.bb[2], va.bb[3]);
		walk_trees(T, W.s_world[a], qbox, stack, W.err, [&](int s) {
			if (s == a) return;
			const uint64_t gb = W.s_group[s];
			if ((ga != 0 && ga == gb) || (ca & W.s_mask[s]) == 0 || (W.s_cats[s] & ma) == 0) return;
			double set[12];
			const int m = shapes_collide_views(va, load_shape(W, s), set, W.err);
			if (m > 0) append_contact_hit(H, q, s, m, set);
		});
		return;
	}
	const int w = in.space ? in.space[q] : 0;
	const cpgpuShapeFilter f = query_filter(in.filter, q);
	if (kind == 0) {
		const V2 p = v2((real)in.a[2 * q], (real)in.a[2 * q + 1]);
		const real maxd = (real)in.scalar[q];
		const real ext = fmaxc(maxd, real(0.0));
		real4 qbox;
		qbox.x = p.x - ext; qbox.y = p.y - ext; qbox.z = p.x + ext; qbox.w = p.y + ext;
		walk_trees(T, w, qbox, stack, W.err, [&](int s) {
			if (filter_reject(W, s, f)) return;
			const PQ r = point_query(W, load_shape(W, s), s, p);
			if (!(r.d < maxd)) return;
			const double v[5] = {(double)r.point.x, (double)r.point.y, (double)r.d, (double)r.grad.x,
			                     (double)r.grad.y};
			append_hit(H, q, s, v);
		});
	} else if (kind == 1) {
		const V2 a = v2((real)in.a[2 * q], (real)in.a[2 * q + 1]);
		const V2 b = v2((real)in.b[2 * q], (real)in.b[2 * q + 1]);
		const real radius = (real)in.scalar[q];
		walk_trees(T, w, segment_box(a, b, radius), stack, W.err, [&](int s) {
			if (filter_reject(W, s, f)) return;
			const SQ r = segment_query(W, load_shape(W, s), s, a, b, radius);
			if (!r.hit) return;
			const double v[5] = {(double)r.point.x, (double)r.point.y, (double)r.normal.x, (double)r.normal.y,
			                     (double)r.alpha};
			append_hit(H, q, s, v);
		});
	} else {
		real4 qbox;
		qbox.x = (real)in.a[4 * q]; qbox.y = (real)in.a[4 * q + 1];
		qbox.z = (real)in.a[4 * q + 2]; qbox.w = (real)in.a[4 * q + 3];
		walk_trees(T, w, qbox, stack, W.err, [&](int s) {
			if (filter_reject(W, s, f)) return;
			if (!box_meets(qbox, W.s_bb[s])) return;   // the shape's own box (cpSpaceQuery.c:191)
			append_hit(H, q, s, nullptr);
		});
	}
}

inline int blocks_for(int n, int t) { return (n + t - 1) / t; }

} // namespace

hipError_t query_nearest_or_first(const DevWorld &W, const QueryTrees &T, int n, const QueryIn &in, bool segment,
                                  int32_t *shape_out, double *info_out, hipStream_t stream)
{
	if (n <= 0) return hipSuccess;
	if (segment)
		k_segment_first<<<blocks_for(n, QUERY_THREADS), QUERY_THREADS, 0, stream>>>(W, T, n, in, shape_out, info_out);
	else
		k_point_nearest<<<blocks_for(n, QUERY_THREADS), QUERY_THREADS, 0, stream>>>(W, T, n, in, shape_out, info_out);
	return hipGetLastError();
}

hipError_t query_all(const DevWorld &W, const QueryTrees &T, int n, const QueryIn &in, int kind, int2 *hits,
                     double *info, int64_t cap, unsigned long long *count, hipStream_t stream, int32_t *n_pts,
                     double *sets)
{
	if (n <= 0) return hipSuccess;
	HitsOut H{hits, info, cap, count, n_pts, sets};
	k_query_all<<<blocks_for(n, QUERY_THREADS), QUERY_THREADS, 0, stream>>>(W, T, n, in, kind, H);
	return hipGetLastError();
}

} // namespace CPGPU_NS
