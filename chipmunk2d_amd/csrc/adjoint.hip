// adjoint.hip -- reverse-mode derivative of a taped rollout of cpSpaceStep.
//
// The reference has no derivative; this is the library's own extension.
// A taped rollout records, for every step, the solved arbiters after the
// solve (their contacts, with the recorded contact each warm start came
// from) and, every few steps, the bodies' state (checkpoints).  The backward
// pass walks the checkpoint segments from the last: it replays a segment's
// steps from its checkpoint with the step's own kernels (the replayed values
// are bit-identical to the rollout's), recording each arbiter's pre-state in
// every sweep, then reverses the segment's steps one by one with the kernels
// below, one workgroup per space:
//   impulse sweeps in reverse (cpArbiterApplyImpulse, cpArbiter.c:459-498),
//   cached impulses (cpArbiterApplyCachedImpulse, :441-455),
//   velocity integration (cpBodyUpdateVelocity, cpBody.c:493-509),
//   arbiter pre-step (cpArbiterPreStep, cpArbiter.c:416-439),
//   contact geometry: the pair's collision re-run on dual numbers with the
//     rollout's GJK id (same branches, so the contact features stay those of
//     the rollout) gives the derivatives of normal and contact points with
//     respect to both bodies' positions and angles,
//   position integration (cpBodyUpdatePosition, cpBody.c:511-522).
// Warm starts carry adjoints from a contact back to the recorded contact its
// accumulated impulses came from.  Derivatives at the clamps of the
// projected Gauss-Seidel follow the branch the rollout took.
#include "cpgpu_adjoint.h"
#include "cpgpu_dual.h"
#include "narrowphase.h"

namespace CPGPU_NS {

namespace {

constexpr int ADJ_THREADS = 256;
constexpr int DUAL_MAX_VERTS = 8;
constexpr int ADJ_GU_LDS = 4096;   // shapes of a space whose friction adjoints are summed in LDS

__device__ __forceinline__ bool dyn(const DevWorld &W, int b) { return W.b_type[b] == CPGPU_BODY_DYNAMIC; }

// ---- tape records ----

__global__ void k_tape_record(Arbiter *arbs, const int32_t *n_ptr, int n_host, Arbiter *tape, int off)
{
	const int n = n_ptr ? *n_ptr : n_host;
	for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
		Arbiter a = arbs[m];
		for (int c = 0; c < 2; c++) a.c[c].tape = 2 * (off + m) + c;
		arbs[m].c[0].tape = a.c[0].tape;
		arbs[m].c[1].tape = a.c[1].tape;
		tape[off + m] = a;
	}
}

// the rollout's arbiters of one step with their accumulated impulses reset to
// the warm start they started from
__global__ void k_replay_prepare(const Arbiter *tape, int off, int n, Arbiter *act)
{
	const int m = blockIdx.x * blockDim.x + threadIdx.x;
	if (m >= n) return;
	Arbiter a = tape[off + m];
	for (int c = 0; c < 2; c++) {
		const int src = a.c[c].src;
		if (c < a.count && src >= 0) {
			a.c[c].jnAcc = tape[src >> 1].c[src & 1].jnAcc;
			a.c[c].jtAcc = tape[src >> 1].c[src & 1].jtAcc;
		} else {
			a.c[c].jnAcc = 0.0;
			a.c[c].jtAcc = 0.0;
		}
	}
	act[m] = a;
}

__global__ void k_state_save(const DevWorld W, BodyState *out)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= W.n_bodies) return;
	BodyState s;
	s.p = W.b_p[b]; s.a = W.b_a[b]; s.vel = W.b_vel[b]; s.f = W.b_f[b]; s.t = W.b_t[b]; s.T = W.b_T[b];
	out[b] = s;
}

__global__ void k_state_load(const DevWorld W, const BodyState *in)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= W.n_bodies) return;
	const BodyState s = in[b];
	W.b_p[b] = s.p; W.b_a[b] = s.a; W.b_vel[b] = s.vel; W.b_f[b] = s.f; W.b_t[b] = s.t; W.b_T[b] = s.T;
}

// ---- the reverse of one contact's impulse (cpArbiter.c:459-498) ----

struct BodyV {
	V2 v, vb;
	double w, wb;
};

struct ContactAdj {
	V2 r1, r2;
	double nMass, tMass, bias, bounce;
};

// per arbiter: adjoints of its parameters, accumulated over the sweeps, and
// of its contacts' accumulated impulses
struct ArbAdj {
	ContactAdj c[2];
	V2 n, svr;
	double u;
	double jn[2], jt[2], jb[2];
	int gu_s[2];          // the two shapes (space-local) and their friction adjoint terms
	double gu_v[2];
};

__device__ __forceinline__ BodyV bodyv_from(const real *r)
{
	BodyV b;
	b.v = v2(r[0], r[1]); b.w = r[2]; b.vb = v2(r[3], r[4]); b.wb = r[5];
	return b;
}

// The forward of contact i of arb on (a, b) with accumulators (jn, jt, jb):
// the same operations as arb_apply_impulse.
__device__ void contact_forward(const Arbiter &arb, int i, double mA, double iA, double mB, double iB, BodyV &a,
                                BodyV &b, double &jnAcc, double &jtAcc, double &jBias)
{
	const Contact &con = arb.c[i];
	const V2 n = v2(arb.nx, arb.ny);
	const V2 r1 = v2(con.r1x, con.r1y), r2 = v2(con.r2x, con.r2y);
	const V2 vb1 = vadd(a.vb, vmult(vperp(r1), a.wb));
	const V2 vb2 = vadd(b.vb, vmult(vperp(r2), b.wb));
	const V2 vr = vadd(vsub(vadd(b.v, vmult(vperp(r2), b.w)), vadd(a.v, vmult(vperp(r1), a.w))), v2(arb.svx, arb.svy));
	const double vbn = vdot(vsub(vb2, vb1), n);
	const double vrn = vdot(vr, n);
	const double vrt = vdot(vr, vperp(n));
	const double jbn = (con.bias - vbn) * con.nMass;
	const double jbnOld = jBias;
	jBias = fmaxc(jbnOld + jbn, 0.0);
	const double jn = -(con.bounce + vrn) * con.nMass;
	const double jnOld = jnAcc;
	jnAcc = fmaxc(jnOld + jn, 0.0);
	const double jtMax = arb.u * jnAcc;
	const double jt = -vrt * con.tMass;
	const double jtOld = jtAcc;
	jtAcc = fclampc(jtOld + jt, -jtMax, jtMax);
	const V2 jb = vmult(n, jBias - jbnOld);
	a.vb = vadd(a.vb, vmult(vneg(jb), mA)); a.wb += iA * vcross(r1, vneg(jb));
	b.vb = vadd(b.vb, vmult(jb, mB)); b.wb += iB * vcross(r2, jb);
	const V2 j = vrotate(n, v2(jnAcc - jnOld, jtAcc - jtOld));
	a.v = vadd(a.v, vmult(vneg(j), mA)); a.w += iA * vcross(r1, vneg(j));
	b.v = vadd(b.v, vmult(j, mB)); b.w += iB * vcross(r2, j);
}

// Reverse of contact_forward.  (a, b, jnOld, jtOld, jbOld): the pre-state;
// ga, gb: adjoints of the bodies after the contact on entry, before it on
// exit; G.jn/jt/jb[i]: adjoints of the contact's accumulators after it on
// entry, before it on exit.  Parameter adjoints add into G and G.c[i], the
// bodies' inverse mass and moment adjoints into gm (mA, iA, mB, iB).
__device__ void contact_reverse(const Arbiter &arb, int i, double mA, double iA, double mB, double iB,
                                const BodyV &a, const BodyV &b, double jnOld, double jtOld, double jbOld, BodyV &ga,
                                BodyV &gb, ArbAdj &G, double gm[4])
{
	const Contact &con = arb.c[i];
	ContactAdj &gc = G.c[i];
	const V2 n = v2(arb.nx, arb.ny);
	const V2 r1 = v2(con.r1x, con.r1y), r2 = v2(con.r2x, con.r2y);
	// forward intermediates
	const V2 vb1 = vadd(a.vb, vmult(vperp(r1), a.wb));
	const V2 vb2 = vadd(b.vb, vmult(vperp(r2), b.wb));
	const V2 vr = vadd(vsub(vadd(b.v, vmult(vperp(r2), b.w)), vadd(a.v, vmult(vperp(r1), a.w))), v2(arb.svx, arb.svy));
	const double vbn = vdot(vsub(vb2, vb1), n);
	const double vrn = vdot(vr, n);
	const double vrt = vdot(vr, vperp(n));
	const double jbn = (con.bias - vbn) * con.nMass;
	const double z = jbOld + jbn;
	const double jbNew = fmaxc(z, 0.0);
	const double jn = -(con.bounce + vrn) * con.nMass;
	const double y = jnOld + jn;
	const double jnNew = fmaxc(y, 0.0);
	const double M = arb.u * jnNew;
	const double jt = -vrt * con.tMass;
	const double x = jtOld + jt;
	const double t1 = (x > -M) ? x : -M;
	const double jtNew = (t1 < M) ? t1 : M;
	const double db = jbNew - jbOld;
	const V2 jbv = vmult(n, db);
	const V2 d = v2(jnNew - jnOld, jtNew - jtOld);
	const V2 J = vrotate(n, d);
	V2 gn = v2(0.0, 0.0);
	// apply_impulses
	const V2 gJ = vadd(vsub(vmult(gb.v, mB), vmult(ga.v, mA)),
	                   vsub(vmult(vperp(r2), iB * gb.w), vmult(vperp(r1), iA * ga.w)));
	gc.r1 = vsub(gc.r1, vmult(vrperp(J), iA * ga.w));
	gc.r2 = vadd(gc.r2, vmult(vrperp(J), iB * gb.w));
	gm[0] -= vdot(ga.v, J); gm[1] -= ga.w * vcross(r1, J);
	gm[2] += vdot(gb.v, J); gm[3] += gb.w * vcross(r2, J);
	const V2 gd = v2(gJ.x * n.x + gJ.y * n.y, -gJ.x * n.y + gJ.y * n.x);
	gn = vadd(gn, v2(gJ.x * d.x + gJ.y * d.y, -gJ.x * d.y + gJ.y * d.x));
	// bias impulses
	const V2 gjbv = vadd(vsub(vmult(gb.vb, mB), vmult(ga.vb, mA)),
	                     vsub(vmult(vperp(r2), iB * gb.wb), vmult(vperp(r1), iA * ga.wb)));
	gc.r1 = vsub(gc.r1, vmult(vrperp(jbv), iA * ga.wb));
	gc.r2 = vadd(gc.r2, vmult(vrperp(jbv), iB * gb.wb));
	gm[0] -= vdot(ga.vb, jbv); gm[1] -= ga.wb * vcross(r1, jbv);
	gm[2] += vdot(gb.vb, jbv); gm[3] += gb.wb * vcross(r2, jbv);
	const double gdb = vdot(gjbv, n);
	gn = vadd(gn, vmult(gjbv, db));
	// adjoints of the new accumulators: later uses plus this contact's own
	const double g_jtNew = G.jt[i] + gd.y;
	double gx = 0.0, gM = 0.0;
	if (t1 < M) { if (x > -M) gx = g_jtNew; else gM -= g_jtNew; }
	else gM += g_jtNew;
	const double gjt = gx;
	const double gvrt = -gjt * con.tMass;
	gc.tMass += -gjt * vrt;
	G.u += gM * jnNew;
	const double g_jnNew = G.jn[i] + gd.x + gM * arb.u;
	const double gy = (y > 0.0) ? g_jnNew : 0.0;
	const double gjn = gy;
	gc.bounce += -gjn * con.nMass;
	const double gvrn = -gjn * con.nMass;
	gc.nMass += -gjn * (con.bounce + vrn);
	const double g_jbNew = G.jb[i] + gdb;
	const double gz = (z > 0.0) ? g_jbNew : 0.0;
	const double gjbn = gz;
	gc.bias += gjbn * con.nMass;
	const double gvbn = -gjbn * con.nMass;
	gc.nMass += gjbn * (con.bias - vbn);
	// accumulators before the contact
	G.jn[i] = gy - gd.x;
	G.jt[i] = gx - gd.y;
	G.jb[i] = gz - gdb;
	// relative velocities
	const V2 gvr = vadd(vmult(n, gvrn), vmult(vperp(n), gvrt));
	gn = vadd(gn, vadd(vmult(vr, gvrn), vmult(vrperp(vr), gvrt)));
	const V2 gvb2 = vmult(n, gvbn), gvb1 = vmult(n, -gvbn);
	gn = vadd(gn, vmult(vsub(vb2, vb1), gvbn));
	gb.v = vadd(gb.v, gvr); gb.w += vdot(gvr, vperp(r2)); gc.r2 = vadd(gc.r2, vmult(vrperp(gvr), b.w));
	ga.v = vsub(ga.v, gvr); ga.w -= vdot(gvr, vperp(r1)); gc.r1 = vsub(gc.r1, vmult(vrperp(gvr), a.w));
	G.svr = vadd(G.svr, gvr);
	ga.vb = vadd(ga.vb, gvb1); ga.wb += vdot(gvb1, vperp(r1)); gc.r1 = vadd(gc.r1, vmult(vrperp(gvb1), a.wb));
	gb.vb = vadd(gb.vb, gvb2); gb.wb += vdot(gvb2, vperp(r2)); gc.r2 = vadd(gc.r2, vmult(vrperp(gvb2), b.wb));
	G.n = vadd(G.n, gn);
}

// ---- contact geometry on dual numbers ----

struct DualShape {
	ShapeViewT<Dual> v;
	V2T<Dual> wv[DUAL_MAX_VERTS], wn[DUAL_MAX_VERTS];
};

// body position and angle after the step's position update (cpBody.c:511-522, the
// operations of k_update_positions), from the step's start state
__device__ __forceinline__ void updated_pose(const BodyState &s, double dt, V2 &p, double &a)
{
	p = vadd(v2(s.p.x, s.p.y), vmult(vadd(v2(s.vel.vx, s.vel.vy), v2(s.vel.vbx, s.vel.vby)), dt));
	a = s.a + (s.vel.w + s.vel.wb) * dt;
}

// the body transform on duals; seed k0 >= 0 makes (p.x, p.y, a) the inputs k0..k0+2
__device__ XFT<Dual> dual_transform(const DevWorld &W, const BodyState *start, int b, double dt, int k0)
{
	XFT<Dual> T;
	if (W.b_type[b] == CPGPU_BODY_STATIC) {
		const XF t = W.b_T[b];
		T.a = Dual(t.a); T.b = Dual(t.b); T.c = Dual(t.c); T.d = Dual(t.d); T.tx = Dual(t.tx); T.ty = Dual(t.ty);
		return T;
	}
	V2 p;
	double a;
	updated_pose(start[b], dt, p, a);
	double sn, cs;
	cp_sincos(a, sn, cs);
	Dual px(p.x), py(p.y), dsn(sn), dcs(cs);
	if (k0 >= 0) {
		px.d[k0] = 1.0;
		py.d[k0 + 1] = 1.0;
		dsn.d[k0 + 2] = cs;
		dcs.d[k0 + 2] = -sn;
	}
	const real2 c = W.b_cog[b];
	T.a = dcs; T.b = dsn; T.c = -dsn; T.d = dcs;
	T.tx = px - (c.x * dcs - c.y * dsn);
	T.ty = py - (c.x * dsn + c.y * dcs);
	return T;
}

// cacheData on duals (cpShape.c:291-296, 378-405; cpPolyShape.c:39-64), with the
// box of k_cache_shapes from the value parts (it seeds GJK when the id is 0)
__device__ bool dual_shape(const DevWorld &W, int s, const XFT<Dual> &T, DualShape &out)
{
	ShapeViewT<Dual> &v = out.v;
	v.type = W.s_type[s];
	v.r = W.s_r[s];
	v.hashid = W.s_hashid[s];
	v.rot = v2t<Dual>(T.a, T.b);
	v.count = 0;
	v.wv = out.wv; v.wn = out.wn;
	const double r = v.r;
	const real *g = W.s_geom + 10 * (size_t)s;
	if (v.type == CPGPU_SHAPE_CIRCLE) {
		v.tc = xpoint(T, v2(g[0], g[1]));
		const double cx = v.tc.x.v, cy = v.tc.y.v;
		v.bb[0] = cx - r; v.bb[1] = cy - r; v.bb[2] = cx + r; v.bb[3] = cy + r;
	} else if (v.type == CPGPU_SHAPE_SEGMENT) {
		v.ta = xpoint(T, v2(g[0], g[1]));
		v.tb = xpoint(T, v2(g[2], g[3]));
		v.tn = xvect(T, v2(g[4], g[5]));
		v.a_tan = v2(g[6], g[7]); v.b_tan = v2(g[8], g[9]);
		const double ax = v.ta.x.v, ay = v.ta.y.v, bx = v.tb.x.v, by = v.tb.y.v;
		double l, rr, bt, tp;
		if (ax < bx) { l = ax; rr = bx; } else { l = bx; rr = ax; }
		if (ay < by) { bt = ay; tp = by; } else { bt = by; tp = ay; }
		v.bb[0] = l - r; v.bb[1] = bt - r; v.bb[2] = rr + r; v.bb[3] = tp + r;
	} else {
		const int st = W.s_pstart[s], cnt = W.s_pcount[s];
		if (cnt > DUAL_MAX_VERTS) return false;
		v.count = cnt;
		double l = __builtin_inf(), rr = -__builtin_inf(), bt = __builtin_inf(), tp = -__builtin_inf();
		for (int i = 0; i < cnt; i++) {
			const real2 lv = W.pv_local[st + i], ln = W.pn_local[st + i];
			out.wv[i] = xpoint(T, v2(lv.x, lv.y));
			out.wn[i] = xvect(T, v2(ln.x, ln.y));
			const double x = out.wv[i].x.v, y = out.wv[i].y.v;
			l = fminc(l, x); rr = fmaxc(rr, x); bt = fminc(bt, y); tp = fmaxc(tp, y);
		}
		v.bb[0] = l - r; v.bb[1] = bt - r; v.bb[2] = rr + r; v.bb[3] = tp + r;
	}
	return true;
}

// ---- the reverse of the joints (cpPivotJoint.c:24-70, cpDampedSpring.c:24-77,
// cpSlideJoint.c:24-89, cpGearJoint.c:24-69) ----

// per joint: adjoints of its pre-step results, accumulated over the sweeps,
// and of its accumulators (pivot jAcc; spring x: target_vrn; slide, gear x)
struct JointAdj {
	V2 r1, r2, n, bias;        // bias: pivot (bx, by); slide and gear in x
	double k[4];               // pivot k11 k12 k21 k22
	double nMass, v_coef, iSum;
	V2 acc;
};

struct JV {
	V2 v;
	double w;
};

__device__ __forceinline__ JV jv_from(const real *r)
{
	JV x;
	x.v = v2(r[0], r[1]); x.w = r[2];
	return x;
}

__device__ __forceinline__ V2 jv_relvel(const JV &a, const JV &b, V2 r1, V2 r2)
{
	return vsub(vadd(b.v, vmult(vperp(r2), b.w)), vadd(a.v, vmult(vperp(r1), a.w)));
}

// Adjoint of apply_impulses(a, b, r1, r2, J) (chipmunk_private.h:187-199):
// returns dL/dJ; adds to the anchors' and inverse masses' adjoints.  The
// velocity adjoints pass through unchanged (the impulse is added).
__device__ __forceinline__ V2 impulse_adj(const JV &ga, const JV &gb, double mA, double iA, double mB, double iB,
                                          V2 r1, V2 r2, V2 J, V2 &g_r1, V2 &g_r2, double gm[4])
{
	gm[0] -= vdot(ga.v, J); gm[1] -= ga.w * vcross(r1, J);
	gm[2] += vdot(gb.v, J); gm[3] += gb.w * vcross(r2, J);
	g_r1 = vsub(g_r1, vmult(vrperp(J), iA * ga.w));
	g_r2 = vadd(g_r2, vmult(vrperp(J), iB * gb.w));
	return vadd(vsub(vmult(gb.v, mB), vmult(ga.v, mA)), vsub(vmult(vperp(r2), iB * gb.w), vmult(vperp(r1), iA * ga.w)));
}

// adjoint of vr = (vB + perp(r2) wB) - (vA + perp(r1) wA) for dL/dvr = gvr
__device__ __forceinline__ void relvel_adj(const JV &a, const JV &b, V2 r1, V2 r2, V2 gvr, JV &ga, JV &gb, V2 &g_r1,
                                           V2 &g_r2)
{
	gb.v = vadd(gb.v, gvr); gb.w += vdot(gvr, vperp(r2)); g_r2 = vadd(g_r2, vmult(vrperp(gvr), b.w));
	ga.v = vsub(ga.v, gvr); ga.w -= vdot(gvr, vperp(r1)); g_r1 = vsub(g_r1, vmult(vrperp(gvr), a.w));
}

// y = vclamp(x, L) (cpVect.h): the identity, or x L / |x| when |x| > L
__device__ __forceinline__ V2 vclamp_adj(V2 x, double L, V2 gy)
{
	if (!(vdot(x, x) > L * L)) return gy;
	const double len = csqrt(vdot(x, x));
	const V2 xh = vmult(x, 1.0 / len);
	return vmult(vsub(gy, vmult(xh, vdot(xh, gy))), L / len);
}

// d fclampc(f, lo, hi) / df along the branch taken
__device__ __forceinline__ double fclamp_d(double f, double lo, double hi)
{
	const double m = (f > lo) ? f : lo;
	return (f > lo && m < hi) ? 1.0 : 0.0;
}

__device__ __forceinline__ double k_scalar_body_d(double m_inv, double i_inv, V2 r, V2 n)
{
	const double rcn = vcross(r, n);
	return m_inv + i_inv * rcn * rcn;
}

__device__ __forceinline__ JV gb_load(const BodyAdj &g)
{
	JV x;
	x.v = v2(g.v[0], g.v[1]); x.w = g.w;
	return x;
}

__device__ __forceinline__ void gb_store(BodyAdj &g, const JV &x)
{
	g.v[0] = x.v.x; g.v[1] = x.v.y; g.w = x.w;
}

// Reverse of joint_apply_impulse from its pre-state R.  ga, gb: adjoints of
// the bodies after the application on entry, before it on exit; G.acc: of
// the accumulators after it on entry, before it on exit.
__device__ void joint_impulse_reverse(const Joint &j, double dt, const JointRec &R, double mA, double iA, double mB,
                                      double iB, JV &ga, JV &gb, JointAdj &G, double gm[4])
{
	const JV a = jv_from(R.a), b = jv_from(R.b);
	const V2 r1 = v2(j.r1x, j.r1y), r2 = v2(j.r2x, j.r2y);
	if (j.type == CPGPU_JOINT_PIVOT) {
		const V2 jOld = v2(R.acc[0], R.acc[1]);
		const V2 d = vsub(v2(j.bx, j.by), jv_relvel(a, b, r1, r2));
		const V2 jj = v2(d.x * j.k11 + d.y * j.k12, d.x * j.k21 + d.y * j.k22);
		const V2 x = vadd(jOld, jj);
		const double L = j.maxForce * dt;
		const V2 D = vsub(vclamp(x, L), jOld);
		const V2 gD = impulse_adj(ga, gb, mA, iA, mB, iB, r1, r2, D, G.r1, G.r2, gm);
		const V2 gx = vclamp_adj(x, L, vadd(G.acc, gD));
		const V2 gd = v2(gx.x * j.k11 + gx.y * j.k21, gx.x * j.k12 + gx.y * j.k22);
		G.k[0] += gx.x * d.x; G.k[1] += gx.x * d.y; G.k[2] += gx.y * d.x; G.k[3] += gx.y * d.y;
		G.bias = vadd(G.bias, gd);
		relvel_adj(a, b, r1, r2, vneg(gd), ga, gb, G.r1, G.r2);
		G.acc = vsub(gx, gD);
	} else if (j.type == CPGPU_JOINT_DAMPED_SPRING) {
		const V2 n = v2(j.nx, j.ny);
		const double tOld = R.acc[0];
		const V2 vr = jv_relvel(a, b, r1, r2);
		const double vrn = vdot(vr, n);
		const double v_damp = (tOld - vrn) * j.v_coef;
		const double j_damp = v_damp * j.nMass;
		const V2 gJ = impulse_adj(ga, gb, mA, iA, mB, iB, r1, r2, vmult(n, j_damp), G.r1, G.r2, gm);
		const double g_jd = vdot(gJ, n);
		G.n = vadd(G.n, vmult(gJ, j_damp));
		// target_vrn' = vrn + v_damp, j_damp = v_damp nMass, v_damp = (target_vrn - vrn) v_coef
		const double g_vd = g_jd * j.nMass + G.acc.x;
		G.nMass += g_jd * v_damp;
		const double g_vrn = G.acc.x - g_vd * j.v_coef;
		G.v_coef += g_vd * (tOld - vrn);
		G.n = vadd(G.n, vmult(vr, g_vrn));
		relvel_adj(a, b, r1, r2, vmult(n, g_vrn), ga, gb, G.r1, G.r2);
		G.acc = v2(g_vd * j.v_coef, 0.0);
	} else if (j.type == CPGPU_JOINT_SLIDE) {
		const V2 n = v2(j.nx, j.ny);
		if (n.x == 0.0 && n.y == 0.0) return;
		const double jnOld = R.acc[0];
		const V2 vr = jv_relvel(a, b, r1, r2);
		const double vrn = vdot(vr, n);
		const double jn = (j.bias - vrn) * j.nMass;
		const double x = jnOld + jn;
		const double lo = -j.maxForce * dt;
		const double D = fclampc(x, lo, 0.0) - jnOld;
		const V2 gJ = impulse_adj(ga, gb, mA, iA, mB, iB, r1, r2, vmult(n, D), G.r1, G.r2, gm);
		const double gD = vdot(gJ, n);
		G.n = vadd(G.n, vmult(gJ, D));
		const double gx = (G.acc.x + gD) * fclamp_d(x, lo, 0.0);
		G.bias.x += gx * j.nMass;
		G.nMass += gx * (j.bias - vrn);
		const double g_vrn = -gx * j.nMass;
		G.n = vadd(G.n, vmult(vr, g_vrn));
		relvel_adj(a, b, r1, r2, vmult(n, g_vrn), ga, gb, G.r1, G.r2);
		G.acc.x = gx - gD;
	} else {
		const double ratio = j.P[1], ratio_inv = j.P[2];
		const double jOld = R.acc[0];
		const double wr = b.w * ratio - a.w;
		const double jMax = j.maxForce * dt;
		const double x = jOld + (j.bias - wr) * j.iSum;
		const double D = fclampc(x, -jMax, jMax) - jOld;
		// a.w -= D iA ratio_inv, b.w += D iB
		const double gD = -ga.w * iA * ratio_inv + gb.w * iB;
		gm[1] -= ga.w * D * ratio_inv; gm[3] += gb.w * D;
		const double gx = (G.acc.x + gD) * fclamp_d(x, -jMax, jMax);
		G.bias.x += gx * j.iSum;
		G.iSum += gx * (j.bias - wr);
		const double g_wr = -gx * j.iSum;
		gb.w += g_wr * ratio;
		ga.w -= g_wr;
		G.acc.x = gx - gD;
	}
}

// Reverse of joint_apply_cached with the warm-start accumulators acc.
__device__ void joint_cached_reverse(const Joint &j, double dt_coef, const real *acc, double mA, double iA, double mB,
                                     double iB, const JV &ga, const JV &gb, JointAdj &G, double gm[4])
{
	const V2 r1 = v2(j.r1x, j.r1y), r2 = v2(j.r2x, j.r2y);
	if (j.type == CPGPU_JOINT_PIVOT) {
		const V2 gJ = impulse_adj(ga, gb, mA, iA, mB, iB, r1, r2, vmult(v2(acc[0], acc[1]), dt_coef), G.r1, G.r2, gm);
		G.acc = vadd(G.acc, vmult(gJ, dt_coef));
	} else if (j.type == CPGPU_JOINT_SLIDE) {
		const V2 n = v2(j.nx, j.ny);
		const double s = acc[0] * dt_coef;
		const V2 gJ = impulse_adj(ga, gb, mA, iA, mB, iB, r1, r2, vmult(n, s), G.r1, G.r2, gm);
		G.acc.x += vdot(gJ, n) * dt_coef;
		G.n = vadd(G.n, vmult(gJ, s));
	} else if (j.type == CPGPU_JOINT_GEAR) {
		const double ratio_inv = j.P[2];
		const double jj = acc[0] * dt_coef;
		const double g_jj = -ga.w * iA * ratio_inv + gb.w * iB;
		gm[1] -= ga.w * jj * ratio_inv; gm[3] += gb.w * jj;
		G.acc.x += g_jj * dt_coef;
	}
}

// rotation (cos, sin) of body b after the step's position update
__device__ __forceinline__ void updated_rot(const DevWorld &W, const BodyState &s, int b, double dt, double &cs,
                                            double &sn)
{
	if (W.b_type[b] == CPGPU_BODY_STATIC) {
		const XF t = W.b_T[b];
		cs = t.a; sn = t.b;
		return;
	}
	V2 p;
	double a;
	updated_pose(s, dt, p, a);
	cp_sincos(a, sn, cs);
}

// Reverse of joint_prestep (the spring's impulse included): the adjoints of
// its results in G go to the bodies' updated positions and angles, their
// inverse masses, the parameters gP[8] and, through the spring impulse, read
// the velocity adjoints after it.
__device__ void joint_prestep_reverse(const DevWorld &W, const AdjStep &A, const WorldStep &ws, const Joint &j,
                                      JointAdj &G, double gP[8], V2 &gpA, double &gaA, V2 &gpB, double &gaB,
                                      double gm[4])
{
	const BodyState &sa = A.start[j.a], &sb = A.start[j.b];
	const double mA = sa.vel.m_inv, iA = sa.vel.i_inv, mB = sb.vel.m_inv, iB = sb.vel.i_inv;
	const double dt = ws.dt, bc = j.bias_coef;
	V2 pA, pB;
	double angA = 0.0, angB = 0.0;
	if (W.b_type[j.a] == CPGPU_BODY_STATIC) pA = v2(sa.p.x, sa.p.y); else updated_pose(sa, dt, pA, angA);
	if (W.b_type[j.b] == CPGPU_BODY_STATIC) pB = v2(sb.p.x, sb.p.y); else updated_pose(sb, dt, pB, angB);
	const V2 r1 = v2(j.r1x, j.r1y), r2 = v2(j.r2x, j.r2y);
	const V2 delta = vsub(vadd(pB, r2), vadd(pA, r1));
	// k_scalar(a, b, r1, r2, n) (chipmunk_private.h:212-226) with adjoint g_k
	auto k_scalar_adj = [&](V2 n, double g_k) {
		const double rcn1 = vcross(r1, n), rcn2 = vcross(r2, n);
		gm[0] += g_k; gm[1] += g_k * rcn1 * rcn1;
		gm[2] += g_k; gm[3] += g_k * rcn2 * rcn2;
		G.r1 = vadd(G.r1, vmult(vrperp(n), g_k * 2.0 * iA * rcn1));
		G.r2 = vadd(G.r2, vmult(vrperp(n), g_k * 2.0 * iB * rcn2));
		G.n = vadd(G.n, vadd(vmult(vperp(r1), g_k * 2.0 * iA * rcn1), vmult(vperp(r2), g_k * 2.0 * iB * rcn2)));
	};
	V2 gdelta = v2(0.0, 0.0);
	if (j.type == CPGPU_JOINT_PIVOT) {
		// bias = vclamp(delta (-bias_coef / dt), maxBias)
		gdelta = vmult(vclamp_adj(vmult(delta, -bc / dt), j.maxBias, G.bias), -bc / dt);
		// (k11 k12; k21 k22) = M^-1 (k_tensor, chipmunk_private.h:228-262): dM = -K^T dK K^T
		const double K[4] = {j.k11, j.k12, j.k21, j.k22};
		double P[4];
		for (int r = 0; r < 2; r++)
			for (int q = 0; q < 2; q++) P[2 * r + q] = K[r] * G.k[q] + K[2 + r] * G.k[2 + q];
		double gM[4];
		for (int r = 0; r < 2; r++)
			for (int q = 0; q < 2; q++) gM[2 * r + q] = -(P[2 * r] * K[2 * q] + P[2 * r + 1] * K[2 * q + 1]);
		const double g12 = gM[1] + gM[2];
		gm[0] += gM[0] + gM[3]; gm[2] += gM[0] + gM[3];
		gm[1] += gM[0] * r1.y * r1.y - g12 * r1.x * r1.y + gM[3] * r1.x * r1.x;
		gm[3] += gM[0] * r2.y * r2.y - g12 * r2.x * r2.y + gM[3] * r2.x * r2.x;
		G.r1 = vadd(G.r1, v2(iA * (2.0 * gM[3] * r1.x - g12 * r1.y), iA * (2.0 * gM[0] * r1.y - g12 * r1.x)));
		G.r2 = vadd(G.r2, v2(iB * (2.0 * gM[3] * r2.x - g12 * r2.y), iB * (2.0 * gM[0] * r2.y - g12 * r2.x)));
	} else if (j.type == CPGPU_JOINT_DAMPED_SPRING) {
		const double rest = j.P[4], stiffness = j.P[5], damping = j.P[6];
		const V2 n = v2(j.nx, j.ny);
		const double dist = vlength(delta);
		const double k = k_scalar_body_d(mA, iA, r1, n) + k_scalar_body_d(mB, iB, r2, n);
		// the spring impulse n (rest - dist) stiffness dt, applied to the velocities
		const double js = (rest - dist) * stiffness * dt;
		const JV ga = gb_load(A.gb[j.a]), gb = gb_load(A.gb[j.b]);
		const V2 gJ = impulse_adj(ga, gb, mA, iA, mB, iB, r1, r2, vmult(n, js), G.r1, G.r2, gm);
		const double g_f = vdot(gJ, n) * dt;
		G.n = vadd(G.n, vmult(gJ, js));
		gP[4] += g_f * stiffness;
		gP[5] += g_f * (rest - dist);
		double g_dist = -g_f * stiffness;
		// v_coef = 1 - exp(-damping dt k), nMass = 1 / k
		const double e = cp_exp(-damping * dt * k);
		const double g_arg = -G.v_coef * e;
		gP[6] += g_arg * (-dt * k);
		const double g_k = g_arg * (-damping * dt) - G.nMass / (k * k);
		k_scalar_adj(n, g_k);
		// n = delta / dist
		if (dist != 0.0) gdelta = vadd(vmult(vsub(G.n, vmult(n, vdot(n, G.n))), 1.0 / dist), vmult(n, g_dist));
	} else if (j.type == CPGPU_JOINT_SLIDE) {
		const V2 n = v2(j.nx, j.ny);
		if (!(n.x == 0.0 && n.y == 0.0)) {
			const double mn = j.P[4], mx = j.P[5];
			const double dist = vlength(delta);
			const double s = dist > mx ? 1.0 : -1.0;
			const double pdist = s > 0.0 ? dist - mx : mn - dist;
			// bias = clamp(-bias_coef pdist / dt, +-maxBias)
			const double xb = -bc * pdist / dt;
			const double g_pdist = G.bias.x * fclamp_d(xb, -j.maxBias, j.maxBias) * (-bc / dt);
			if (s > 0.0) gP[5] -= g_pdist; else gP[4] += g_pdist;
			const double g_dist = s * g_pdist;
			const double k = k_scalar_body_d(mA, iA, r1, n) + k_scalar_body_d(mB, iB, r2, n);
			k_scalar_adj(n, -G.nMass / (k * k));
			// n = s delta / |delta|
			const V2 nh = vmult(n, s);
			gdelta = vadd(vmult(vsub(G.n, vmult(nh, vdot(nh, G.n))), s / dist), vmult(nh, g_dist));
		}
	} else {
		// iSum = 1 / (iA ratio_inv + ratio iB); bias = clamp(-bias_coef (aB ratio - aA - phase) / dt, +-maxBias)
		const double ratio = j.P[1], ratio_inv = j.P[2];
		const double g_den = -G.iSum * j.iSum * j.iSum;
		gm[1] += g_den * ratio_inv; gm[3] += g_den * ratio;
		const double aA = W.b_type[j.a] == CPGPU_BODY_STATIC ? (double)sa.a : angA;
		const double aB = W.b_type[j.b] == CPGPU_BODY_STATIC ? (double)sb.a : angB;
		const double xb = -bc * (aB * ratio - aA - j.P[0]) / dt;
		const double g_x = G.bias.x * fclamp_d(xb, -j.maxBias, j.maxBias);
		gaB += g_x * (-bc / dt) * ratio;
		gaA += g_x * (bc / dt);
		gP[0] += g_x * (bc / dt);
		return;
	}
	// delta = (pB + r2) - (pA + r1)
	gpB = vadd(gpB, gdelta); gpA = vsub(gpA, gdelta);
	G.r2 = vadd(G.r2, gdelta); G.r1 = vsub(G.r1, gdelta);
	// r = R(angle) (anchor - cog): dr/dangle = perp(r), dr/danchor = R
	gaA += vdot(G.r1, vperp(r1));
	gaB += vdot(G.r2, vperp(r2));
	double cs, sn;
	updated_rot(W, sa, j.a, dt, cs, sn);
	gP[0] += cs * G.r1.x + sn * G.r1.y; gP[1] += -sn * G.r1.x + cs * G.r1.y;
	updated_rot(W, sb, j.b, dt, cs, sn);
	gP[2] += cs * G.r2.x + sn * G.r2.y; gP[3] += -sn * G.r2.x + cs * G.r2.y;
}

// ---- adjoint state ----

__device__ __forceinline__ void add2(double *p, V2 v) { p[0] += v.x; p[1] += v.y; }

// colour boundaries of space w's colour order, positions [lo, hi); returns
// the colour count (col_start[0..n] in LDS)
__device__ int reverse_colours(const AdjStep &A, int lo, int hi, int32_t *col_start, int32_t *n_col)
{
	const int tid = threadIdx.x, stride = blockDim.x;
	if (tid == 0) *n_col = hi > lo ? A.colour[A.order[hi - 1]] + 1 : 0;
	for (int pos = lo + tid; pos < hi; pos += stride) {
		const int c = A.colour[A.order[pos]];
		if (pos == lo || A.colour[A.order[pos - 1]] != c) col_start[c] = pos;
	}
	__syncthreads();
	if (tid == 0) col_start[*n_col] = hi;
	__syncthreads();
	return *n_col;
}

// The reverse of one step for one small space (one workgroup), in two
// kernels: the impulses (sweeps, cached impulses, velocity integration) here,
// with few registers; the pre-steps and the contact geometry on dual numbers
// (k_reverse_geometry) after it.  The per-arbiter and per-joint adjoints pass
// between them in garb / gjoint.
__global__ void __launch_bounds__(ADJ_THREADS) k_reverse_solve(const DevWorld W, const AdjStep A)
{
	__shared__ int32_t col_start[CPGPU_MAX_COLOURS + 1];
	__shared__ int32_t n_col;
	const int w = blockIdx.x;
	const int tid = threadIdx.x, stride = blockDim.x;
	ArbAdj *garb = reinterpret_cast<ArbAdj *>(A.garb);
	const int lo = A.warb_off ? A.warb_off[w] : 0;
	const int hi = A.warb_off ? A.warb_off[w + 1] : A.n;
	const WorldStep ws = W.ws[w];
	const int b0 = W.w_body0[w], b1 = W.w_body0[w + 1];
	reverse_colours(A, lo, hi, col_start, &n_col);
	// per-arbiter adjoints: the final accumulators' adjoints come from later warm starts
	for (int pos = lo + tid; pos < hi; pos += stride) {
		const Arbiter &arb = A.act[A.order[pos]];
		ArbAdj G;
		for (int i = 0; i < 2; i++) {
			G.c[i].r1 = v2(0.0, 0.0); G.c[i].r2 = v2(0.0, 0.0);
			G.c[i].nMass = G.c[i].tMass = G.c[i].bias = G.c[i].bounce = 0.0;
			const bool on = i < arb.count && arb.c[i].tape >= 0;
			G.jn[i] = on ? A.gJ[arb.c[i].tape].x : 0.0;
			G.jt[i] = on ? A.gJ[arb.c[i].tape].y : 0.0;
			G.jb[i] = 0.0;
		}
		G.n = v2(0.0, 0.0); G.svr = v2(0.0, 0.0); G.u = 0.0;
		garb[pos] = G;
	}
	// the space's joints by colour (positions in jw_order); their accumulators' adjoints come from the
	// next step's warm start (a spring's accumulator does not carry over)
	const int32_t *jcol = A.nc > 0 ? A.jw_col + A.jw_cbase[w] : nullptr;
	const int njc = A.nc > 0 ? A.jw_ncol[w] : 0;
	JointAdj *gjt = reinterpret_cast<JointAdj *>(A.gjoint);
	if (njc > 0)
		for (int pos = jcol[0] + tid; pos < jcol[njc]; pos += stride) {
			const int c = A.jw_order[pos];
			JointAdj G;
			G.r1 = G.r2 = G.n = G.bias = v2(0.0, 0.0);
			G.k[0] = G.k[1] = G.k[2] = G.k[3] = 0.0;
			G.nMass = G.v_coef = G.iSum = 0.0;
			G.acc = A.joints[c].type == CPGPU_JOINT_DAMPED_SPRING ? v2(0.0, 0.0) : v2(A.gJacc[c].x, A.gJacc[c].y);
			gjt[pos] = G;
		}
	__syncthreads();
	const int ncol = n_col;
	// one joint application reversed (sweep slot: 0 cached, it + 1 sweep it)
	auto joint_step_reverse = [&](int pos, int slot) {
		const Joint &j = A.joints[A.jw_order[pos]];
		const double mA = W.b_vel[j.a].m_inv, iA = W.b_vel[j.a].i_inv;
		const double mB = W.b_vel[j.b].m_inv, iB = W.b_vel[j.b].i_inv;
		BodyAdj &GA = A.gb[j.a];
		BodyAdj &GB = A.gb[j.b];
		JV ga = gb_load(GA), gb = gb_load(GB);
		JointAdj G = gjt[pos];
		double gm[4] = {0.0, 0.0, 0.0, 0.0};
		const JointRec &R = A.jrec[(size_t)slot * A.nc + pos];
		if (slot == 0) joint_cached_reverse(j, ws.dt_coef, R.acc, mA, iA, mB, iB, ga, gb, G, gm);
		else joint_impulse_reverse(j, ws.dt, R, mA, iA, mB, iB, ga, gb, G, gm);
		gjt[pos] = G;
		if (dyn(W, j.a)) { gb_store(GA, ga); A.gminv[j.a] += gm[0]; A.giinv[j.a] += gm[1]; }
		if (dyn(W, j.b)) { gb_store(GB, gb); A.gminv[j.b] += gm[2]; A.giinv[j.b] += gm[3]; }
	};

	// ---- impulse sweeps, last first (in a sweep the joints follow the arbiters) ----
	for (int it = ws.iterations - 1; it >= 0; it--) {
		for (int jc = njc - 1; jc >= 0; jc--) {
			for (int pos = jcol[jc] + tid; pos < jcol[jc + 1]; pos += stride) joint_step_reverse(pos, it + 1);
			__syncthreads();
		}
		for (int c = ncol - 1; c >= 0; c--) {
			for (int pos = col_start[c] + tid; pos < col_start[c + 1]; pos += stride) {
				const Arbiter &arb = A.act[A.order[pos]];
				const ArbRec &rec = A.rec[(size_t)it * A.n + pos];
				const int ba = arb.ba, bb = arb.bb;
				const double mA = W.b_vel[ba].m_inv, iA = W.b_vel[ba].i_inv;
				const double mB = W.b_vel[bb].m_inv, iB = W.b_vel[bb].i_inv;
				BodyAdj &GA = A.gb[ba];
				BodyAdj &GB = A.gb[bb];
				BodyV ga, gb;
				ga.v = v2(GA.v[0], GA.v[1]); ga.w = GA.w; ga.vb = v2(GA.vb[0], GA.vb[1]); ga.wb = GA.wb;
				gb.v = v2(GB.v[0], GB.v[1]); gb.w = GB.w; gb.vb = v2(GB.vb[0], GB.vb[1]); gb.wb = GB.wb;
				ArbAdj G = garb[pos];
				double gm[4] = {0.0, 0.0, 0.0, 0.0};
				// pre-states: before contact 0 (recorded) and before contact 1 (recomputed)
				const BodyV a0 = bodyv_from(rec.a), b0v = bodyv_from(rec.b);
				if (arb.count > 1) {
					BodyV a1 = a0, b1v = b0v;
					double jn = rec.acc[0][0], jt = rec.acc[0][1], jb = rec.acc[0][2];
					contact_forward(arb, 0, mA, iA, mB, iB, a1, b1v, jn, jt, jb);
					contact_reverse(arb, 1, mA, iA, mB, iB, a1, b1v, rec.acc[1][0], rec.acc[1][1], rec.acc[1][2], ga,
					                gb, G, gm);
				}
				if (arb.count > 0)
					contact_reverse(arb, 0, mA, iA, mB, iB, a0, b0v, rec.acc[0][0], rec.acc[0][1], rec.acc[0][2], ga,
					                gb, G, gm);
				garb[pos] = G;
				if (dyn(W, ba)) {
					GA.v[0] = ga.v.x; GA.v[1] = ga.v.y; GA.w = ga.w; GA.vb[0] = ga.vb.x; GA.vb[1] = ga.vb.y; GA.wb = ga.wb;
					A.gminv[ba] += gm[0]; A.giinv[ba] += gm[1];
				}
				if (dyn(W, bb)) {
					GB.v[0] = gb.v.x; GB.v[1] = gb.v.y; GB.w = gb.w; GB.vb[0] = gb.vb.x; GB.vb[1] = gb.vb.y; GB.wb = gb.wb;
					A.gminv[bb] += gm[2]; A.giinv[bb] += gm[3];
				}
			}
			__syncthreads();
		}
	}

	// ---- cached impulses: the joints' (after the arbiters'), then the arbiters' (cpArbiterApplyCachedImpulse,
	// skipped on a first collision) ----
	for (int jc = njc - 1; jc >= 0; jc--) {
		for (int pos = jcol[jc] + tid; pos < jcol[jc + 1]; pos += stride) joint_step_reverse(pos, 0);
		__syncthreads();
	}
	// the joints' warm starts: to the accumulators at the end of the step before (a spring's pre-step
	// overwrites its accumulator, an inactive slide's clears it)
	if (njc > 0)
		for (int pos = jcol[0] + tid; pos < jcol[njc]; pos += stride) {
			const int c = A.jw_order[pos];
			const Joint &j = A.joints[c];
			const bool carried = j.type == CPGPU_JOINT_PIVOT || j.type == CPGPU_JOINT_GEAR ||
			                     (j.type == CPGPU_JOINT_SLIDE && !(j.nx == 0.0 && j.ny == 0.0));
			const V2 g = carried ? gjt[pos].acc : v2(0.0, 0.0);
			A.gJacc[c] = make_double2(g.x, g.y);
		}
	for (int c = ncol - 1; c >= 0; c--) {
		for (int pos = col_start[c] + tid; pos < col_start[c + 1]; pos += stride) {
			const Arbiter &arb = A.act[A.order[pos]];
			ArbAdj G = garb[pos];
			if (arb.state != CPGPU_ARB_FIRST_COLLISION) {
				const int ba = arb.ba, bb = arb.bb;
				const double mA = W.b_vel[ba].m_inv, iA = W.b_vel[ba].i_inv;
				const double mB = W.b_vel[bb].m_inv, iB = W.b_vel[bb].i_inv;
				BodyAdj &GA = A.gb[ba];
				BodyAdj &GB = A.gb[bb];
				const V2 gva = v2(GA.v[0], GA.v[1]), gvb = v2(GB.v[0], GB.v[1]);
				const double gwa = GA.w, gwb = GB.w;
				const V2 n = v2(arb.nx, arb.ny);
				double gm[4] = {0.0, 0.0, 0.0, 0.0};
				for (int i = 0; i < arb.count; i++) {
					const Contact &con = arb.c[i];
					const V2 r1 = v2(con.r1x, con.r1y), r2 = v2(con.r2x, con.r2y);
					const int src = con.src;
					const double jn0 = src >= 0 ? A.tape[src >> 1].c[src & 1].jnAcc : 0.0;
					const double jt0 = src >= 0 ? A.tape[src >> 1].c[src & 1].jtAcc : 0.0;
					const V2 q = v2(jn0, jt0);
					const V2 J = vmult(vrotate(n, q), ws.dt_coef);
					const V2 gJ = vadd(vsub(vmult(gvb, mB), vmult(gva, mA)),
					                   vsub(vmult(vperp(r2), iB * gwb), vmult(vperp(r1), iA * gwa)));
					G.c[i].r1 = vsub(G.c[i].r1, vmult(vrperp(J), iA * gwa));
					G.c[i].r2 = vadd(G.c[i].r2, vmult(vrperp(J), iB * gwb));
					gm[0] -= vdot(gva, J); gm[1] -= gwa * vcross(r1, J);
					gm[2] += vdot(gvb, J); gm[3] += gwb * vcross(r2, J);
					const V2 gR = vmult(gJ, ws.dt_coef);
					G.jn[i] += gR.x * n.x + gR.y * n.y;
					G.jt[i] += -gR.x * n.y + gR.y * n.x;
					G.n = vadd(G.n, v2(gR.x * q.x + gR.y * q.y, -gR.x * q.y + gR.y * q.x));
				}
				if (dyn(W, ba)) { A.gminv[ba] += gm[0]; A.giinv[ba] += gm[1]; }
				if (dyn(W, bb)) { A.gminv[bb] += gm[2]; A.giinv[bb] += gm[3]; }
			}
			garb[pos] = G;
		}
		__syncthreads();
	}
	// The warm starts send the adjoints of the initial accumulators to the
	// contacts they came from.  An arbiter's sources are contacts of its own
	// shape pair's earlier arbiter, so each arbiter's lane adds them alone.
	// The friction adjoint goes to both shapes (u = uA uB); a shape has many
	// arbiters, so those sums run in a fixed order on one lane (gradients are
	// reproducible bit for bit), into LDS when the space's shapes fit.
	__shared__ double lgu[ADJ_GU_LDS];
	const int s0 = W.w_shape0[w], ns_w = W.w_shape0[w + 1] - s0;
	const bool gu_lds = ns_w <= ADJ_GU_LDS;
	for (int pos = lo + tid; pos < hi; pos += stride) {
		const Arbiter &arb = A.act[A.order[pos]];
		ArbAdj &G = garb[pos];
		for (int i = 0; i < arb.count; i++) {
			const int src = arb.c[i].src;
			if (src >= 0) {
				A.gJ[src].x += G.jn[i];
				A.gJ[src].y += G.jt[i];
			}
		}
		G.gu_s[0] = arb.sa - s0; G.gu_s[1] = arb.sb - s0;
		G.gu_v[0] = G.u * W.s_u[arb.sb];
		G.gu_v[1] = G.u * W.s_u[arb.sa];
	}
	if (gu_lds)
		for (int q = tid; q < ns_w; q += stride) lgu[q] = 0.0;
	__syncthreads();
	if (tid == 0) {
		double *gu = gu_lds ? lgu : A.gu + s0;
		int pos = lo;
		// four arbiters' terms are loaded ahead of their sums
		for (; pos + 4 <= hi; pos += 4) {
			int sh[8];
			double v[8];
			for (int q = 0; q < 4; q++) {
				const ArbAdj &G = garb[pos + q];
				sh[2 * q] = G.gu_s[0]; sh[2 * q + 1] = G.gu_s[1];
				v[2 * q] = G.gu_v[0]; v[2 * q + 1] = G.gu_v[1];
			}
			for (int q = 0; q < 8; q++) gu[sh[q]] += v[q];
		}
		for (; pos < hi; pos++) {
			const ArbAdj &G = garb[pos];
			gu[G.gu_s[0]] += G.gu_v[0];
			gu[G.gu_s[1]] += G.gu_v[1];
		}
	}
	__syncthreads();
	if (gu_lds)
		for (int q = tid; q < ns_w; q += stride) A.gu[s0 + q] += lgu[q];
	__syncthreads();

	// ---- velocity integration (cpBodyUpdateVelocity): v = v damping + (g + f m_inv) dt ----
	for (int b = b0 + tid; b < b1; b += stride) {
		if (!dyn(W, b)) continue;
		BodyAdj &GB = A.gb[b];
		const BodyState &s = A.start[b];
		A.gminv[b] += (GB.v[0] * s.f.x + GB.v[1] * s.f.y) * ws.dt;
		A.giinv[b] += GB.w * s.t * ws.dt;
		GB.v[0] *= ws.damping; GB.v[1] *= ws.damping;
		GB.w *= ws.damping;
	}
	__syncthreads();

}

__global__ void __launch_bounds__(ADJ_THREADS) k_reverse_geometry(const DevWorld W, const AdjStep A)
{
	__shared__ int32_t col_start[CPGPU_MAX_COLOURS + 1];
	__shared__ int32_t n_col;
	const int w = blockIdx.x;
	const int tid = threadIdx.x, stride = blockDim.x;
	ArbAdj *garb = reinterpret_cast<ArbAdj *>(A.garb);
	const int lo = A.warb_off ? A.warb_off[w] : 0;
	const int hi = A.warb_off ? A.warb_off[w + 1] : A.n;
	const WorldStep ws = W.ws[w];
	const int b0 = W.w_body0[w], b1 = W.w_body0[w + 1];
	const int ncol = reverse_colours(A, lo, hi, col_start, &n_col);
	const int32_t *jcol = A.nc > 0 ? A.jw_col + A.jw_cbase[w] : nullptr;
	const int njc = A.nc > 0 ? A.jw_ncol[w] : 0;
	JointAdj *gjt = reinterpret_cast<JointAdj *>(A.gjoint);

	// ---- joint pre-steps (after the arbiters' pre-steps, before the velocity integration) ----
	for (int jc = njc - 1; jc >= 0; jc--) {
		for (int pos = jcol[jc] + tid; pos < jcol[jc + 1]; pos += stride) {
			const int c = A.jw_order[pos];
			const Joint &j = A.joints[c];
			JointAdj G = gjt[pos];
			double gP[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
			double gm[4] = {0.0, 0.0, 0.0, 0.0};
			V2 gpA = v2(0.0, 0.0), gpB = v2(0.0, 0.0);
			double gaA = 0.0, gaB = 0.0;
			joint_prestep_reverse(W, A, ws, j, G, gP, gpA, gaA, gpB, gaB, gm);
			for (int k = 0; k < 8; k++) A.gP[8 * (size_t)c + k] += gP[k];
			if (dyn(W, j.a)) {
				BodyAdj &GA = A.gb[j.a];
				GA.p[0] += gpA.x; GA.p[1] += gpA.y; GA.a += gaA;
				A.gminv[j.a] += gm[0]; A.giinv[j.a] += gm[1];
			}
			if (dyn(W, j.b)) {
				BodyAdj &GB = A.gb[j.b];
				GB.p[0] += gpB.x; GB.p[1] += gpB.y; GB.a += gaB;
				A.gminv[j.b] += gm[2]; A.giinv[j.b] += gm[3];
			}
		}
		__syncthreads();
	}

	// ---- pre-step and contact geometry, colour by colour (bodies of one colour are distinct) ----
	for (int c = ncol - 1; c >= 0; c--) {
		for (int pos = col_start[c] + tid; pos < col_start[c + 1]; pos += stride) {
			const Arbiter &arb = A.act[A.order[pos]];
			ArbAdj G = garb[pos];
			const int ba = arb.ba, bb = arb.bb;
			const BodyState &sa = A.start[ba], &sb = A.start[bb];
			const double iA = sa.vel.i_inv, iB = sb.vel.i_inv;
			const V2 n = v2(arb.nx, arb.ny);
			V2 pA, pB;
			double angA, angB;
			if (W.b_type[ba] == CPGPU_BODY_STATIC) pA = v2(sa.p.x, sa.p.y); else updated_pose(sa, ws.dt, pA, angA);
			if (W.b_type[bb] == CPGPU_BODY_STATIC) pB = v2(sb.p.x, sb.p.y); else updated_pose(sb, ws.dt, pB, angB);
			const V2 body_delta = vsub(pB, pA);
			// velocities the pre-step saw (before the velocity integration)
			const V2 vA = v2(sa.vel.vx, sa.vel.vy), vB = v2(sb.vel.vx, sb.vel.vy);
			const double wA = sa.vel.w, wB = sb.vel.w;
			double gm[4] = {0.0, 0.0, 0.0, 0.0};
			V2 gpA = v2(0.0, 0.0), gpB = v2(0.0, 0.0);
			V2 gvA = v2(0.0, 0.0), gvB = v2(0.0, 0.0);
			double gwA = 0.0, gwB = 0.0;
			for (int i = 0; i < arb.count; i++) {
				const Contact &con = arb.c[i];
				ContactAdj &gc = G.c[i];
				const V2 r1 = v2(con.r1x, con.r1y), r2 = v2(con.r2x, con.r2y);
				// nMass = 1 / k(n), tMass = 1 / k(perp n)  (k_scalar, chipmunk_private.h:212-226)
				const double kn_bar = -gc.nMass * con.nMass * con.nMass;
				const double rcn1 = vcross(r1, n), rcn2 = vcross(r2, n);
				gm[0] += kn_bar; gm[1] += kn_bar * rcn1 * rcn1;
				gm[2] += kn_bar; gm[3] += kn_bar * rcn2 * rcn2;
				gc.r1 = vadd(gc.r1, vmult(vrperp(n), kn_bar * 2.0 * iA * rcn1));
				gc.r2 = vadd(gc.r2, vmult(vrperp(n), kn_bar * 2.0 * iB * rcn2));
				G.n = vadd(G.n, vadd(vmult(vperp(r1), kn_bar * 2.0 * iA * rcn1), vmult(vperp(r2), kn_bar * 2.0 * iB * rcn2)));
				const V2 t = vperp(n);
				const double kt_bar = -gc.tMass * con.tMass * con.tMass;
				const double rct1 = vcross(r1, t), rct2 = vcross(r2, t);
				gm[0] += kt_bar; gm[1] += kt_bar * rct1 * rct1;
				gm[2] += kt_bar; gm[3] += kt_bar * rct2 * rct2;
				gc.r1 = vadd(gc.r1, vmult(vrperp(t), kt_bar * 2.0 * iA * rct1));
				gc.r2 = vadd(gc.r2, vmult(vrperp(t), kt_bar * 2.0 * iB * rct2));
				const V2 gt = vadd(vmult(vperp(r1), kt_bar * 2.0 * iA * rct1), vmult(vperp(r2), kt_bar * 2.0 * iB * rct2));
				G.n = vadd(G.n, v2(gt.y, -gt.x));
				// bias = -bias_coef min(0, dist + slop) / dt, dist = (r2 - r1 + pB - pA) . n
				const V2 dvec = vadd(vsub(r2, r1), body_delta);
				const double dist = vdot(dvec, n);
				const double xb = dist + ws.slop;
				if (!(0.0 < xb)) {
					const double gdist = gc.bias * (-ws.bias_coef / ws.dt);
					gc.r2 = vadd(gc.r2, vmult(n, gdist));
					gc.r1 = vsub(gc.r1, vmult(n, gdist));
					gpB = vadd(gpB, vmult(n, gdist));
					gpA = vsub(gpA, vmult(n, gdist));
					G.n = vadd(G.n, vmult(dvec, gdist));
				}
				// bounce = (relative velocity . n) e
				const double gvrn = gc.bounce * arb.e;
				if (gvrn != 0.0) {
					const V2 rv = vsub(vadd(vB, vmult(vperp(r2), wB)), vadd(vA, vmult(vperp(r1), wA)));
					gvB = vadd(gvB, vmult(n, gvrn)); gwB += gvrn * vdot(vperp(r2), n);
					gc.r2 = vadd(gc.r2, vmult(vrperp(n), wB * gvrn));
					gvA = vsub(gvA, vmult(n, gvrn)); gwA -= gvrn * vdot(vperp(r1), n);
					gc.r1 = vsub(gc.r1, vmult(vrperp(n), wA * gvrn));
					G.n = vadd(G.n, vmult(rv, gvrn));
				}
			}
			// surface velocity: svr = s - n (s . n) with s the shapes' surface velocity difference
			{
				const real2 svA = W.s_sv[arb.sa], svB = W.s_sv[arb.sb];
				const V2 sv = v2(svB.x - svA.x, svB.y - svA.y);
				G.n = vsub(G.n, vadd(vmult(G.svr, vdot(sv, n)), vmult(sv, vdot(n, G.svr))));
			}
			// contact geometry: r1 = p1 - pA, r2 = p2 - pB with (n, p1, p2) from the collision on duals
			{
				DualShape dsa, dsb;
				const XFT<Dual> TA = dual_transform(W, A.start, ba, ws.dt, dyn(W, ba) ? 0 : -1);
				const XFT<Dual> TB = dual_transform(W, A.start, bb, ws.dt, dyn(W, bb) ? 3 : -1);
				const bool ok = dual_shape(W, arb.sa, TA, dsa) && dual_shape(W, arb.sb, TB, dsb);
				CollInfoT<Dual> info;
				if (ok) collide(dsa.v, dsb.v, arb.id_in, info, W.err);
				bool same = ok && info.count == arb.count && info.n.x.v == arb.nx && info.n.y.v == arb.ny;
				for (int i = 0; same && i < arb.count; i++) {
					same = info.p1[i].x.v - pA.x == arb.c[i].r1x && info.p1[i].y.v - pA.y == arb.c[i].r1y &&
					       info.p2[i].x.v - pB.x == arb.c[i].r2x && info.p2[i].y.v - pB.y == arb.c[i].r2y;
				}
				if (!same) {
					raise_err(W.err, CPGPU_ERR_ADJOINT);
				} else {
					double gin[DUAL_N];
					for (int k = 0; k < DUAL_N; k++) gin[k] = G.n.x * info.n.x.d[k] + G.n.y * info.n.y.d[k];
					for (int i = 0; i < arb.count; i++) {
						const ContactAdj &gc = G.c[i];
						for (int k = 0; k < DUAL_N; k++)
							gin[k] += gc.r1.x * info.p1[i].x.d[k] + gc.r1.y * info.p1[i].y.d[k] +
							          gc.r2.x * info.p2[i].x.d[k] + gc.r2.y * info.p2[i].y.d[k];
						gpA = vsub(gpA, gc.r1);
						gpB = vsub(gpB, gc.r2);
					}
					if (dyn(W, ba)) {
						BodyAdj &GA = A.gb[ba];
						GA.p[0] += gpA.x + gin[0]; GA.p[1] += gpA.y + gin[1]; GA.a += gin[2];
						GA.v[0] += gvA.x; GA.v[1] += gvA.y; GA.w += gwA;
						A.gminv[ba] += gm[0]; A.giinv[ba] += gm[1];
					}
					if (dyn(W, bb)) {
						BodyAdj &GB = A.gb[bb];
						GB.p[0] += gpB.x + gin[3]; GB.p[1] += gpB.y + gin[4]; GB.a += gin[5];
						GB.v[0] += gvB.x; GB.v[1] += gvB.y; GB.w += gwB;
						A.gminv[bb] += gm[2]; A.giinv[bb] += gm[3];
					}
				}
			}
		}
		__syncthreads();
	}

	// ---- position integration (cpBodyUpdatePosition): p' = p + (v + v_bias) dt, a' = a + (w + w_bias) dt ----
	// The bias velocities the solve started from were zero (reset by the update), so
	// the adjoint of the step's incoming bias velocities is the update's alone.
	for (int b = b0 + tid; b < b1; b += stride) {
		if (!dyn(W, b)) continue;
		BodyAdj &GB = A.gb[b];
		GB.v[0] += GB.p[0] * ws.dt; GB.v[1] += GB.p[1] * ws.dt;
		GB.vb[0] = GB.p[0] * ws.dt; GB.vb[1] = GB.p[1] * ws.dt;
		GB.w += GB.a * ws.dt;
		GB.wb = GB.a * ws.dt;
	}
}

inline int blocks_for(int n, int t) { return (n + t - 1) / t; }

} // namespace

hipError_t adjoint_tape_record(Arbiter *arbs, const int32_t *n_dev, int n_host, Arbiter *tape, int off,
                               hipStream_t stream)
{
	if (n_host <= 0) return hipSuccess;
	k_tape_record<<<std::min(blocks_for(n_host, 256), 1024), 256, 0, stream>>>(arbs, n_dev, n_host, tape, off);
	return hipGetLastError();
}

hipError_t adjoint_replay_prepare(const Arbiter *tape, int off, int n, Arbiter *act, hipStream_t stream)
{
	if (n <= 0) return hipSuccess;
	k_replay_prepare<<<blocks_for(n, 256), 256, 0, stream>>>(tape, off, n, act);
	return hipGetLastError();
}

hipError_t adjoint_state_save(const DevWorld &W, BodyState *out, hipStream_t stream)
{
	if (W.n_bodies <= 0) return hipSuccess;
	k_state_save<<<blocks_for(W.n_bodies, 256), 256, 0, stream>>>(W, out);
	return hipGetLastError();
}

hipError_t adjoint_state_load(const DevWorld &W, const BodyState *in, hipStream_t stream)
{
	if (W.n_bodies <= 0) return hipSuccess;
	k_state_load<<<blocks_for(W.n_bodies, 256), 256, 0, stream>>>(W, in);
	return hipGetLastError();
}

size_t adjoint_arb_scratch_bytes() { return sizeof(ArbAdj); }
size_t adjoint_joint_scratch_bytes() { return sizeof(JointAdj); }

hipError_t adjoint_reverse_solve(const DevWorld &W, const AdjStep &A, int n_worlds, hipStream_t stream)
{
	k_reverse_solve<<<n_worlds, ADJ_THREADS, 0, stream>>>(W, A);
	return hipGetLastError();
}

hipError_t adjoint_reverse_geometry(const DevWorld &W, const AdjStep &A, int n_worlds, hipStream_t stream)
{
	k_reverse_geometry<<<n_worlds, ADJ_THREADS, 0, stream>>>(W, A);
	return hipGetLastError();
}

} // namespace CPGPU_NS
