// narrowphase.h -- device narrowphase, one shape pair per lane.
//
// Restates cpCollide and its pair routines (src/cpCollision.c:44-726,
// src/cpRobust.c) over the SoA shape arrays.  The steps of GJK, EPA and
// contact clipping follow the reference exactly, because contact points and
// the integer feature hashes must match it bit for bit; the device form
// differs in representation only: recursion becomes loops, and the EPA hull
// keeps 16-bit Minkowski point ids (vertex index pairs) instead of points,
// re-deriving a point from its id (the id determines the support vertices,
// so the recomputed point is bit-identical).
//
// Every routine is generic over the scalar T: the build's real for the step, and the
// forward-mode Dual of cpgpu_dual.h for the adjoint, which re-runs a pair's
// collision with the step's GJK id to differentiate its contacts with
// respect to the two bodies' positions and angles (same branches, so the
// contact features are the step's).
#pragma once

#include "cpgpu_device.h"

namespace CPGPU_NS {

#define NP_MAX_GJK_ITERATIONS 30
#define NP_MAX_EPA_ITERATIONS 30
#define NP_EPA_HULL_MAX 36

template <class T>
struct ShapeViewT {
	int type, count;
	real r;
	uint64_t hashid;
	V2T<T> tc;                 // circle centre (world)
	V2T<T> ta, tb, tn;         // segment (world)
	V2 a_tan, b_tan;           // segment neighbour tangents (body-local)
	V2T<T> rot;                // body rotation (transform.a, transform.b)
	const V2T<T> *wv, *wn;     // polygon vertices and edge normals (world)
	real bb[4];
};
using ShapeView = ShapeViewT<real>;

CPD ShapeView load_shape(const DevWorld &W, int s)
{
	ShapeView v;
	v.type = W.s_type[s];
	v.r = W.s_r[s];
	v.hashid = W.s_hashid[s];
	const XF T = W.b_T[W.s_body[s]];
	v.rot = v2(T.a, T.b);
	const real *g = W.s_wgeom + 6 * (size_t)s;
	const real4 bb = W.s_bb[s];
	v.bb[0] = bb.x; v.bb[1] = bb.y; v.bb[2] = bb.z; v.bb[3] = bb.w;
	v.count = 0;
	v.wv = nullptr; v.wn = nullptr;
	if (v.type == CPGPU_SHAPE_CIRCLE) {
		v.tc = v2(g[0], g[1]);
	} else if (v.type == CPGPU_SHAPE_SEGMENT) {
		v.ta = v2(g[0], g[1]); v.tb = v2(g[2], g[3]); v.tn = v2(g[4], g[5]);
		const real *lg = W.s_geom + 10 * (size_t)s;
		v.a_tan = v2(lg[6], lg[7]); v.b_tan = v2(lg[8], lg[9]);
	} else {
		const int st = W.s_pstart[s];
		v.count = W.s_pcount[s];
		// real2 and V2 share their layout {x, y}
		v.wv = reinterpret_cast<const V2 *>(W.pv_world + st);
		v.wn = reinterpret_cast<const V2 *>(W.pn_world + st);
	}
	return v;
}

template <class T> struct SupportPtT { V2T<T> p; uint32_t index; };
template <class T> struct MinkPtT { V2T<T> a, b, ab; uint32_t id; };
template <class T> struct ClosestPtsT { V2T<T> a, b, n; T d; uint32_t id; };
template <class T> struct EdgePtT { V2T<T> p; uint64_t hash; };
template <class T> struct EdgeT { EdgePtT<T> a, b; real r; V2T<T> n; };
template <class T> struct CollInfoT {
	int count;
	uint32_t id;
	V2T<T> n;
	V2T<T> p1[2], p2[2];
	uint64_t hash[2];
};
using CollInfo = CollInfoT<real>;

// PolySupportPointIndex (cpCollision.c:62-77)
template <class T> CPD int poly_support_index(const ShapeViewT<T> &s, V2T<T> n)
{
	T max = T(-__builtin_inf());
	int index = 0;
	for (int i = 0; i < s.count; i++) {
		const T d = vdot(s.wv[i], n);
		if (d > max) { max = d; index = i; }
	}
	return index;
}

// Circle/Segment/PolySupportPoint (cpCollision.c:95-114)
template <class T> CPD SupportPtT<T> shape_support(const ShapeViewT<T> &s, V2T<T> n)
{
	SupportPtT<T> r;
	if (s.type == CPGPU_SHAPE_CIRCLE) { r.p = s.tc; r.index = 0; }
	else if (s.type == CPGPU_SHAPE_SEGMENT) {
		if (vdot(s.ta, n) > vdot(s.tb, n)) { r.p = s.ta; r.index = 0; }
		else { r.p = s.tb; r.index = 1; }
	} else {
		const int i = poly_support_index(s, n);
		r.p = s.wv[i]; r.index = (uint32_t)i;
	}
	return r;
}

// ShapePoint (cpCollision.c:395-413)
template <class T> CPD SupportPtT<T> shape_point(const ShapeViewT<T> &s, int i)
{
	SupportPtT<T> r;
	if (s.type == CPGPU_SHAPE_CIRCLE) { r.p = s.tc; r.index = 0; }
	else if (s.type == CPGPU_SHAPE_SEGMENT) { r.p = (i == 0 ? s.ta : s.tb); r.index = (uint32_t)i; }
	else {
		const int index = (i < s.count ? i : 0);
		r.p = s.wv[index]; r.index = (uint32_t)index;
	}
	return r;
}

// MinkowskiPointNew (cpCollision.c:127-132)
template <class T> CPD MinkPtT<T> mink_new(SupportPtT<T> a, SupportPtT<T> b)
{
	MinkPtT<T> m;
	m.a = a.p; m.b = b.p; m.ab = vsub(b.p, a.p);
	m.id = ((a.index & 0xFF) << 8) | (b.index & 0xFF);
	return m;
}

template <class T> CPD MinkPtT<T> mink_from_id(const ShapeViewT<T> &s1, const ShapeViewT<T> &s2, uint32_t id)
{
	return mink_new(shape_point(s1, (int)((id >> 8) & 0xFF)), shape_point(s2, (int)(id & 0xFF)));
}

// Support (cpCollision.c:140-145)
template <class T> CPD MinkPtT<T> support(const ShapeViewT<T> &s1, const ShapeViewT<T> &s2, V2T<T> n)
{
	return mink_new(shape_support(s1, vneg(n)), shape_support(s2, n));
}

// cpCheckPointGreater / cpCheckAxis (cpRobust.c:4-13)
template <class T> CPD bool check_point_greater(V2T<T> a, V2T<T> b, V2T<T> c)
{
	return (b.y - a.y) * (a.x + b.x - T(2.0) * c.x) > (b.x - a.x) * (a.y + b.y - T(2.0) * c.y);
}
template <class T> CPD bool check_axis(V2T<T> v0, V2T<T> v1, V2T<T> p, V2T<T> n)
{
	return vdot(p, n) <= fmaxc(vdot(v0, n), vdot(v1, n));
}

// ClosestT / LerpT (cpCollision.c:193-207)
template <class T> CPD T closest_t(V2T<T> a, V2T<T> b)
{
	const V2T<T> delta = vsub(b, a);
	return -fclampc(vdot(delta, vadd(a, b)) / (vlengthsq(delta) + CP_REAL_MIN), -1.0, 1.0);
}
template <class T> CPD V2T<T> lerp_t(V2T<T> a, V2T<T> b, T t)
{
	const T ht = T(0.5) * t;
	return vadd(vmult(a, T(0.5) - ht), vmult(b, T(0.5) + ht));
}

// ClosestPointsNew (cpCollision.c:223-258)
template <class T> CPD ClosestPtsT<T> closest_points_new(const MinkPtT<T> &v0, const MinkPtT<T> &v1)
{
	const T t = closest_t(v0.ab, v1.ab);
	const V2T<T> p = lerp_t(v0.ab, v1.ab, t);
	ClosestPtsT<T> r;
	r.a = lerp_t(v0.a, v1.a, t);
	r.b = lerp_t(v0.b, v1.b, t);
	r.id = ((v0.id & 0xFFFF) << 16) | (v1.id & 0xFFFF);
	const V2T<T> delta = vsub(v1.ab, v0.ab);
	const V2T<T> n = vnormalize(vrperp(delta));
	const T d = vdot(n, p);
	if (d <= T(0.0) || (T(-1.0) < t && t < T(1.0))) { r.n = n; r.d = d; }
	else {
		const T d2 = vlength(p);
		r.n = vmult(p, T(1.0) / (d2 + CP_REAL_MIN));
		r.d = d2;
	}
	return r;
}

// ClosestDist (cpCollision.c:262-266)
template <class T> CPD T closest_dist(V2T<T> v0, V2T<T> v1) { return vlengthsq(lerp_t(v0, v1, closest_t(v0, v1))); }

// EPA / EPARecurse (cpCollision.c:270-343) as a loop over id-only hulls.
template <class T>
CPD ClosestPtsT<T> epa(const ShapeViewT<T> &s1, const ShapeViewT<T> &s2, const MinkPtT<T> &v0, const MinkPtT<T> &v1,
                       const MinkPtT<T> &v2, int32_t *err)
{
	uint32_t hullA[NP_EPA_HULL_MAX], hullB[NP_EPA_HULL_MAX];
	uint32_t *hull = hullA, *hull2 = hullB;
	int count = 3;
	hull[0] = v0.id; hull[1] = v1.id; hull[2] = v2.id;
	for (int iteration = 1;; iteration++) {
		int mini = 0;
		T minDist = T(__builtin_inf());
		MinkPtT<T> hi = mink_from_id(s1, s2, hull[count - 1]);
		for (int j = 0; j < count; j++) {
			const MinkPtT<T> hj = mink_from_id(s1, s2, hull[j]);
			const T d = closest_dist(hi.ab, hj.ab);
			if (d < minDist) { minDist = d; mini = (j == 0 ? count - 1 : j - 1); }
			hi = hj;
		}
		const MinkPtT<T> e0 = mink_from_id(s1, s2, hull[mini]);
		const MinkPtT<T> e1 = mink_from_id(s1, s2, hull[(mini + 1) % count]);
		const MinkPtT<T> p = support(s1, s2, vperp(vsub(e1.ab, e0.ab)));
		const bool duplicate = (p.id == e0.id || p.id == e1.id);
		if (!duplicate && check_point_greater(e0.ab, e1.ab, p.ab) && iteration < NP_MAX_EPA_ITERATIONS) {
			if (count + 1 > NP_EPA_HULL_MAX) { raise_err(err, CPGPU_ERR_EPA_HULL); return closest_points_new(e0, e1); }
			int count2 = 1;
			hull2[0] = p.id;
			for (int i = 0; i < count; i++) {
				const int index = (mini + 1 + i) % count;
				const V2T<T> h0 = mink_from_id(s1, s2, hull2[count2 - 1]).ab;
				const V2T<T> h1 = mink_from_id(s1, s2, hull[index]).ab;
				const V2T<T> h2 = (i + 1 < count ? mink_from_id(s1, s2, hull[(index + 1) % count]) : p).ab;
				if (check_point_greater(h0, h2, h1)) hull2[count2++] = hull[index];
			}
			uint32_t *tmp = hull; hull = hull2; hull2 = tmp;
			count = count2;
		} else {
			return closest_points_new(e0, e1);
		}
	}
}

// GJKRecurse (cpCollision.c:348-392) as a loop
template <class T>
CPD ClosestPtsT<T> gjk_recurse(const ShapeViewT<T> &s1, const ShapeViewT<T> &s2, MinkPtT<T> v0, MinkPtT<T> v1,
                               int32_t *err)
{
	const V2T<T> origin = v2t<T>(T(0.0), T(0.0));
	int iteration = 1;
	for (;;) {
		if (iteration > NP_MAX_GJK_ITERATIONS) return closest_points_new(v0, v1);
		if (check_point_greater(v1.ab, v0.ab, origin)) {
			const MinkPtT<T> tmp = v0; v0 = v1; v1 = tmp;
			continue;
		}
		const T t = closest_t(v0.ab, v1.ab);
		const V2T<T> n = (T(-1.0) < t && t < T(1.0)) ? vperp(vsub(v1.ab, v0.ab)) : vneg(lerp_t(v0.ab, v1.ab, t));
		const MinkPtT<T> p = support(s1, s2, n);
		if (check_point_greater(p.ab, v0.ab, origin) && check_point_greater(v1.ab, p.ab, origin)) {
			return epa(s1, s2, v0, p, v1, err);
		}
		if (check_axis(v0.ab, v1.ab, p.ab, n)) return closest_points_new(v0, v1);
		if (closest_dist(v0.ab, p.ab) < closest_dist(p.ab, v1.ab)) v1 = p;
		else v0 = p;
		iteration++;
	}
}

CPD V2 bb_center(const real *bb) { return vlerp(v2(bb[0], bb[1]), v2(bb[2], bb[3]), real(0.5)); }

// GJK (cpCollision.c:416-472)
template <class T>
CPD ClosestPtsT<T> gjk(const ShapeViewT<T> &s1, const ShapeViewT<T> &s2, uint32_t &id, int32_t *err)
{
	MinkPtT<T> v0, v1;
	if (id) {
		v0 = mink_new(shape_point(s1, (int)((id >> 24) & 0xFF)), shape_point(s2, (int)((id >> 16) & 0xFF)));
		v1 = mink_new(shape_point(s1, (int)((id >> 8) & 0xFF)), shape_point(s2, (int)(id & 0xFF)));
	} else {
		const V2T<T> axis = vcast<T>(vperp(vsub(bb_center(s1.bb), bb_center(s2.bb))));
		v0 = support(s1, s2, axis);
		v1 = support(s1, s2, vneg(axis));
	}
	const ClosestPtsT<T> points = gjk_recurse(s1, s2, v0, v1, err);
	id = points.id;
	return points;
}

// cpCollisionInfoPushContact (cpCollision.c:44-55)
template <class T> CPD void push_contact(CollInfoT<T> &info, V2T<T> p1, V2T<T> p2, uint64_t hash)
{
	info.p1[info.count] = p1;
	info.p2[info.count] = p2;
	info.hash[info.count] = hash;
	info.count++;
}

// ContactPoints (cpCollision.c:477-518)
template <class T>
CPD void contact_points(const EdgeT<T> &e1, const EdgeT<T> &e2, const ClosestPtsT<T> &points, CollInfoT<T> &info)
{
	const real mindist = e1.r + e2.r;
	if (points.d <= mindist) {
		const V2T<T> n = info.n = points.n;
		const T d_e1_a = vcross(e1.a.p, n);
		const T d_e1_b = vcross(e1.b.p, n);
		const T d_e2_a = vcross(e2.a.p, n);
		const T d_e2_b = vcross(e2.b.p, n);
		const T e1_denom = T(1.0) / (d_e1_b - d_e1_a + CP_REAL_MIN);
		const T e2_denom = T(1.0) / (d_e2_b - d_e2_a + CP_REAL_MIN);
		{
			const V2T<T> p1 = vadd(vmult(n, e1.r), vlerp(e1.a.p, e1.b.p, fclamp01((d_e2_b - d_e1_a) * e1_denom)));
			const V2T<T> p2 = vadd(vmult(n, -e2.r), vlerp(e2.a.p, e2.b.p, fclamp01((d_e1_a - d_e2_a) * e2_denom)));
			const T dist = vdot(vsub(p2, p1), n);
			if (dist <= T(0.0)) push_contact(info, p1, p2, hash_pair(e1.a.hash, e2.b.hash));
		}
		{
			const V2T<T> p1 = vadd(vmult(n, e1.r), vlerp(e1.a.p, e1.b.p, fclamp01((d_e2_a - d_e1_a) * e1_denom)));
			const V2T<T> p2 = vadd(vmult(n, -e2.r), vlerp(e2.a.p, e2.b.p, fclamp01((d_e1_b - d_e2_a) * e2_denom)));
			const T dist = vdot(vsub(p2, p1), n);
			if (dist <= T(0.0)) push_contact(info, p1, p2, hash_pair(e1.b.hash, e2.a.hash));
		}
	}
}

// SupportEdgeForPoly (cpCollision.c:160-178)
template <class T> CPD EdgeT<T> support_edge_poly(const ShapeViewT<T> &s, V2T<T> n)
{
	const int count = s.count;
	const int i1 = poly_support_index(s, n);
	const int i0 = (i1 - 1 + count) % count;
	const int i2 = (i1 + 1) % count;
	const uint64_t h = s.hashid;
	EdgeT<T> e;
	e.r = s.r;
	if (vdot(n, s.wn[i1]) > vdot(n, s.wn[i2])) {
		e.a.p = s.wv[i0]; e.a.hash = hash_pair(h, (uint64_t)i0);
		e.b.p = s.wv[i1]; e.b.hash = hash_pair(h, (uint64_t)i1);
		e.n = s.wn[i1];
	} else {
		e.a.p = s.wv[i1]; e.a.hash = hash_pair(h, (uint64_t)i1);
		e.b.p = s.wv[i2]; e.b.hash = hash_pair(h, (uint64_t)i2);
		e.n = s.wn[i2];
	}
	return e;
}

// SupportEdgeForSegment (cpCollision.c:180-191)
template <class T> CPD EdgeT<T> support_edge_segment(const ShapeViewT<T> &s, V2T<T> n)
{
	const uint64_t h = s.hashid;
	EdgeT<T> e;
	e.r = s.r;
	if (vdot(s.tn, n) > T(0.0)) {
		e.a.p = s.ta; e.a.hash = hash_pair(h, 0);
		e.b.p = s.tb; e.b.hash = hash_pair(h, 1);
		e.n = s.tn;
	} else {
		e.a.p = s.tb; e.a.hash = hash_pair(h, 1);
		e.b.p = s.ta; e.b.hash = hash_pair(h, 0);
		e.n = vneg(s.tn);
	}
	return e;
}

// CircleToCircle (cpCollision.c:525-537)
template <class T> CPD void circle_circle(const ShapeViewT<T> &c1, const ShapeViewT<T> &c2, CollInfoT<T> &info)
{
	const real mindist = c1.r + c2.r;
	const V2T<T> delta = vsub(c2.tc, c1.tc);
	const T distsq = vlengthsq(delta);
	if (distsq < mindist * mindist) {
		const T dist = csqrt(distsq);
		const V2T<T> n = info.n = (dist != T(0.0) ? vmult(delta, T(1.0) / dist) : v2t<T>(T(1.0), T(0.0)));
		push_contact(info, vadd(c1.tc, vmult(n, c1.r)), vadd(c2.tc, vmult(n, -c2.r)), 0);
	}
}

// CircleToSegment (cpCollision.c:539-569)
template <class T> CPD void circle_segment(const ShapeViewT<T> &circle, const ShapeViewT<T> &seg, CollInfoT<T> &info)
{
	const V2T<T> seg_a = seg.ta, seg_b = seg.tb, center = circle.tc;
	const V2T<T> seg_delta = vsub(seg_b, seg_a);
	const T closest_tt = fclamp01(vdot(seg_delta, vsub(center, seg_a)) / vlengthsq(seg_delta));
	const V2T<T> closest = vadd(seg_a, vmult(seg_delta, closest_tt));
	const real mindist = circle.r + seg.r;
	const V2T<T> delta = vsub(closest, center);
	const T distsq = vlengthsq(delta);
	if (distsq < mindist * mindist) {
		const T dist = csqrt(distsq);
		const V2T<T> n = info.n = (dist != T(0.0) ? vmult(delta, T(1.0) / dist) : seg.tn);
		const V2T<T> rot = seg.rot;
		if ((closest_tt != T(0.0) || vdot(n, vrotate(vcast<T>(seg.a_tan), rot)) >= T(0.0)) &&
		    (closest_tt != T(1.0) || vdot(n, vrotate(vcast<T>(seg.b_tan), rot)) >= T(0.0))) {
			push_contact(info, vadd(center, vmult(n, circle.r)), vadd(closest, vmult(n, -seg.r)), 0);
		}
	}
}

// SegmentToSegment (cpCollision.c:571-604)
template <class T>
CPD void segment_segment(const ShapeViewT<T> &s1, const ShapeViewT<T> &s2, CollInfoT<T> &info, int32_t *err)
{
	const ClosestPtsT<T> points = gjk(s1, s2, info.id, err);
	const V2T<T> n = points.n;
	const V2T<T> rot1 = s1.rot, rot2 = s2.rot;
	if (points.d <= (s1.r + s2.r) &&
	    (!veql(points.a, s1.ta) || vdot(n, vrotate(vcast<T>(s1.a_tan), rot1)) <= T(0.0)) &&
	    (!veql(points.a, s1.tb) || vdot(n, vrotate(vcast<T>(s1.b_tan), rot1)) <= T(0.0)) &&
	    (!veql(points.b, s2.ta) || vdot(n, vrotate(vcast<T>(s2.a_tan), rot2)) >= T(0.0)) &&
	    (!veql(points.b, s2.tb) || vdot(n, vrotate(vcast<T>(s2.b_tan), rot2)) >= T(0.0))) {
		contact_points(support_edge_segment(s1, n), support_edge_segment(s2, vneg(n)), points, info);
	}
}

// PolyToPoly (cpCollision.c:606-627)
template <class T>
CPD void poly_poly(const ShapeViewT<T> &p1, const ShapeViewT<T> &p2, CollInfoT<T> &info, int32_t *err)
{
	const ClosestPtsT<T> points = gjk(p1, p2, info.id, err);
	if (points.d - p1.r - p2.r <= T(0.0)) {
		contact_points(support_edge_poly(p1, points.n), support_edge_poly(p2, vneg(points.n)), points, info);
	}
}

// SegmentToPoly (cpCollision.c:629-659)
template <class T>
CPD void segment_poly(const ShapeViewT<T> &seg, const ShapeViewT<T> &poly, CollInfoT<T> &info, int32_t *err)
{
	const ClosestPtsT<T> points = gjk(seg, poly, info.id, err);
	const V2T<T> n = points.n;
	const V2T<T> rot = seg.rot;
	if (points.d - seg.r - poly.r <= T(0.0) &&
	    (!veql(points.a, seg.ta) || vdot(n, vrotate(vcast<T>(seg.a_tan), rot)) <= T(0.0)) &&
	    (!veql(points.a, seg.tb) || vdot(n, vrotate(vcast<T>(seg.b_tan), rot)) <= T(0.0))) {
		contact_points(support_edge_segment(seg, n), support_edge_poly(poly, vneg(n)), points, info);
	}
}

// CircleToPoly (cpCollision.c:661-679)
template <class T>
CPD void circle_poly(const ShapeViewT<T> &circle, const ShapeViewT<T> &poly, CollInfoT<T> &info, int32_t *err)
{
	const ClosestPtsT<T> points = gjk(circle, poly, info.id, err);
	if (points.d <= circle.r + poly.r) {
		const V2T<T> n = info.n = points.n;
		push_contact(info, vadd(points.a, vmult(n, circle.r)), vadd(points.b, vmult(n, -poly.r)), 0);
	}
}

// cpCollide dispatch (cpCollision.c:688-726); a.type <= b.type
template <class T>
CPD void collide(const ShapeViewT<T> &a, const ShapeViewT<T> &b, uint32_t id, CollInfoT<T> &info, int32_t *err)
{
	info.count = 0;
	info.id = id;
	info.n = v2t<T>(T(0.0), T(0.0));
	switch (a.type + b.type * 3) {
	case 0: circle_circle(a, b, info); break;
	case 3: circle_segment(a, b, info); break;
	case 4: segment_segment(a, b, info, err); break;
	case 6: circle_poly(a, b, info, err); break;
	case 7: segment_poly(a, b, info, err); break;
	case 8: poly_poly(a, b, info, err); break;
	default: break;
	}
}

// cpShapeUpdate(shape, body->transform) (cpShape.c:217-220) into a view of
// its own, leaving the world's cached copy alone: the shape at its body's
// present transform, with the cacheData arithmetic of cache_shape (cpShape.c:
// 291-296, 378-405; cpPolyShape.c:39-64); a polygon's world vertices and
// normals go to P
struct PosedPoly {
	V2 v[CPGPU_MAX_POLY_VERTS], n[CPGPU_MAX_POLY_VERTS];
};

CPD ShapeView pose_shape(const DevWorld &W, int s, PosedPoly &P)
{
	ShapeView v = load_shape(W, s);
	const XF T = W.b_T[W.s_body[s]];
	const real r = v.r;
	const real *g = W.s_geom + 10 * (size_t)s;
	if (v.type == CPGPU_SHAPE_CIRCLE) {
		v.tc = xpoint(T, v2(g[0], g[1]));
		v.bb[0] = v.tc.x - r; v.bb[1] = v.tc.y - r; v.bb[2] = v.tc.x + r; v.bb[3] = v.tc.y + r;
	} else if (v.type == CPGPU_SHAPE_SEGMENT) {
		v.ta = xpoint(T, v2(g[0], g[1]));
		v.tb = xpoint(T, v2(g[2], g[3]));
		v.tn = xvect(T, v2(g[4], g[5]));
		real l, rr, bt, tp;
		if (v.ta.x < v.tb.x) { l = v.ta.x; rr = v.tb.x; } else { l = v.tb.x; rr = v.ta.x; }
		if (v.ta.y < v.tb.y) { bt = v.ta.y; tp = v.tb.y; } else { bt = v.tb.y; tp = v.ta.y; }
		v.bb[0] = l - r; v.bb[1] = bt - r; v.bb[2] = rr + r; v.bb[3] = tp + r;
	} else {
		const int st = W.s_pstart[s];
		real l = real(__builtin_inf()), rr = real(-__builtin_inf()), bt = real(__builtin_inf()), tp = real(-__builtin_inf());
		for (int i = 0; i < v.count; i++) {
			const real2 lv = W.pv_local[st + i], ln = W.pn_local[st + i];
			P.v[i] = xpoint(T, v2(lv.x, lv.y));
			P.n[i] = xvect(T, v2(ln.x, ln.y));
			l = fminc(l, P.v[i].x); rr = fmaxc(rr, P.v[i].x); bt = fminc(bt, P.v[i].y); tp = fmaxc(tp, P.v[i].y);
		}
		v.wv = P.v;
		v.wn = P.n;
		v.bb[0] = l - r; v.bb[1] = bt - r; v.bb[2] = rr + r; v.bb[3] = tp + r;
	}
	return v;
}

// cpShapesCollide (cpShape.c:260-283) of shapes a and b, as its contact point
// set: cpCollide with id 0 on the pair sorted by type (cpCollision.c:707-710);
// when that swapped the shapes the normal is negated and each contact's
// points are given back in the caller's order; distance = (r2 - r1) . normal
// with the unswapped points.  Writes o[12] = {normal.x, normal.y, then per
// point pointA.x, pointA.y, pointB.x, pointB.y, distance} (zeros past the
// count) and returns the count.
CPD int shapes_collide_views(const ShapeView &a, const ShapeView &b, double *o, int32_t *err)
{
	const bool swapped = a.type > b.type;
	CollInfo info;
	if (swapped) collide(b, a, 0u, info, err);
	else collide(a, b, 0u, info, err);
	const V2 nrm = swapped ? vneg(info.n) : info.n;
	o[0] = (double)nrm.x;
	o[1] = (double)nrm.y;
	for (int c = 0; c < 2; c++) {
		double *q = o + 2 + 5 * c;
		if (c < info.count) {
			const V2 p1 = info.p1[c], p2 = info.p2[c];
			const V2 pa = swapped ? p2 : p1, pb = swapped ? p1 : p2;
			q[0] = (double)pa.x; q[1] = (double)pa.y;
			q[2] = (double)pb.x; q[3] = (double)pb.y;
			q[4] = (double)vdot(vsub(p2, p1), nrm);
		} else {
			q[0] = q[1] = q[2] = q[3] = q[4] = 0.0;
		}
	}
	return info.count;
}

// shapes_collide_views of shapes x and y on their cached geometry
CPD int shapes_collide_set(const DevWorld &W, int x, int y, double *o, int32_t *err)
{
	return shapes_collide_views(load_shape(W, x), load_shape(W, y), o, err);
}

} // namespace CPGPU_NS
