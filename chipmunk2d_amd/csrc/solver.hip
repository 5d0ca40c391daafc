// solver.hip -- the velocity solver of cpSpaceStep: persistent kernels.
//
// Reference: the second half of cpSpaceStep (src/cpSpaceStep.c:380-427):
// arbiter pre-step, constraint pre-step, velocity integration, cached
// impulses, then `iterations` sweeps of cpArbiterApplyImpulse and the joints'
// applyImpulse.  The reference sweeps its arbiter list sequentially
// (Gauss-Seidel).  Here the arbiters of one solve unit (a small world, or a
// strip, seam or the wide set of a large world's grid set) are coloured so
// that no two arbiters of a colour share a dynamic body; a sweep
// runs colour after colour with a barrier between colours, so the result is
// exactly the sequential sweep in colour order (arbiters of one colour touch
// disjoint bodies, so their order inside the colour does not matter).
// Units that run concurrently share no dynamic body.  Static and kinematic
// bodies have zero inverse mass, are never written, and do not constrain
// the colouring.
//
// The colouring is deterministic (independent of thread timing): colour c is
// built in COLOUR_SUBROUNDS rounds in which every uncoloured arbiter whose
// dynamic bodies are not yet taken by colour c claims them with an atomic min
// of its priority (a bijective hash of its shape-pair key); arbiters holding
// all their claims take colour c.
#include "cpgpu_step.h"

namespace CPGPU_NS {

namespace {

constexpr int GRID_THREADS = 1024;  // grid solve: one workgroup per column of strips and seams

// ---- grid barrier (all blocks co-resident: cooperative launch) ----
// The arrival count starts at 0 with each launch (k_bar_reset) and only
// grows: barrier e of the launch completes when it reaches e * gridDim.x, so
// the last arrival has nothing to reset or publish.  One lane per workgroup:
// a release fence (writes back this XCD's L2) before arriving, a relaxed
// spin on the count, one acquire fence (invalidates L2) when it is reached.
// Acquire loads inside the spin would invalidate L2 on every poll while
// other workgroups still work.
__device__ void grid_sync(GridBarrier *bar, int32_t *err, unsigned int &epoch)
{
	__syncthreads();
	if (threadIdx.x == 0) {
		const unsigned int target = (++epoch) * gridDim.x;
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
		const unsigned int arrived = __hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (arrived + 1 != target) {
			long spins = 0;
			while (__hip_atomic_load(&bar->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
				__builtin_amdgcn_s_sleep(1);
				if (++spins > (1L << 25)) {
					// a barrier that never completes ends the kernel instead of hanging it
					atomicOr(err, CPGPU_ERR_BARRIER_TIMEOUT);
					break;
				}
				if ((spins & 1023) == 0 &&
				    (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & CPGPU_ERR_BARRIER_TIMEOUT))
					break;
			}
		}
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
	}
	__syncthreads();
}

__global__ void k_bar_reset(GridBarrier *bar)
{
	bar->count = 0;
	bar->gen = 0;
}

__device__ __forceinline__ bool is_dynamic(const DevWorld &W, int b) { return W.b_type[b] == CPGPU_BODY_DYNAMIC; }

// ---- body helpers (chipmunk_private.h:172-226) ----
struct BV {
	V2 v, vb;
	real w, wb, m_inv, i_inv;
};

__device__ __forceinline__ BV load_bv(const DevWorld &W, int b)
{
	const BodyVel q = W.b_vel[b];
	BV r;
	r.v = v2(q.vx, q.vy); r.w = q.w; r.vb = v2(q.vbx, q.vby); r.wb = q.wb; r.m_inv = q.m_inv; r.i_inv = q.i_inv;
	return r;
}

__device__ __forceinline__ void store_bv(const DevWorld &W, int b, const BV &r)
{
	BodyVel q;
	q.vx = r.v.x; q.vy = r.v.y; q.w = r.w; q.vbx = r.vb.x; q.vby = r.vb.y; q.wb = r.wb; q.m_inv = r.m_inv;
	q.i_inv = r.i_inv;
	W.b_vel[b] = q;
}

// Where the impulse loops read and write body velocities: global memory, or
// a copy of one world's contiguous body range staged in LDS.
struct GlobalBodies {
	const DevWorld *W;
	__device__ __forceinline__ BV load(int b) const { return load_bv(*W, b); }
	__device__ __forceinline__ void store(int b, const BV &r) const { store_bv(*W, b, r); }
};
struct LdsBodies {
	BodyVel *v;
	int base;
	__device__ __forceinline__ BV load(int b) const
	{
		const BodyVel &q = v[b - base];
		BV r;
		r.v = v2(q.vx, q.vy); r.w = q.w; r.vb = v2(q.vbx, q.vby); r.wb = q.wb; r.m_inv = q.m_inv; r.i_inv = q.i_inv;
		return r;
	}
	__device__ __forceinline__ void store(int b, const BV &r) const
	{
		BodyVel &q = v[b - base];
		q.vx = r.v.x; q.vy = r.v.y; q.w = r.w; q.vbx = r.vb.x; q.vby = r.vb.y; q.wb = r.wb;
	}
};

__device__ __forceinline__ V2 rel_vel(const BV &a, const BV &b, V2 r1, V2 r2)
{
	const V2 v1 = vadd(a.v, vmult(vperp(r1), a.w));
	const V2 v2s = vadd(b.v, vmult(vperp(r2), b.w));
	return vsub(v2s, v1);
}

__device__ __forceinline__ void apply_impulse(BV &body, V2 j, V2 r)
{
	body.v = vadd(body.v, vmult(j, body.m_inv));
	body.w += body.i_inv * vcross(r, j);
}

__device__ __forceinline__ void apply_impulses(BV &a, BV &b, V2 r1, V2 r2, V2 j)
{
	apply_impulse(a, vneg(j), r1);
	apply_impulse(b, j, r2);
}

__device__ __forceinline__ void apply_bias_impulse(BV &body, V2 j, V2 r)
{
	body.vb = vadd(body.vb, vmult(j, body.m_inv));
	body.wb += body.i_inv * vcross(r, j);
}

__device__ __forceinline__ real k_scalar_body(const BV &b, V2 r, V2 n)
{
	const real rcn = vcross(r, n);
	return b.m_inv + b.i_inv * rcn * rcn;
}

__device__ __forceinline__ real k_scalar(const BV &a, const BV &b, V2 r1, V2 r2, V2 n)
{
	return k_scalar_body(a, r1, n) + k_scalar_body(b, r2, n);
}

// k_scalar's "Unsolvable collision or constraint." check (chipmunk_private.h:223)
__device__ __forceinline__ real solvable(const DevWorld &W, real k)
{
	if (k == real(0.0)) raise_err(W.err, CPGPU_FAULT_UNSOLVABLE);
	return k;
}

// ---- arbiters: cpArbiterPreStep / ApplyCachedImpulse / ApplyImpulse (cpArbiter.c:416-498) ----

__device__ void arb_prestep(const DevWorld &W, Arbiter &arb)
{
	const WorldStep ws = W.ws[arb.world];
	const BV a = load_bv(W, arb.ba), b = load_bv(W, arb.bb);
	const real2 pa = W.b_p[arb.ba], pb = W.b_p[arb.bb];
	const V2 n = v2(arb.nx, arb.ny);
	const V2 body_delta = v2(pb.x - pa.x, pb.y - pa.y);
	for (int i = 0; i < arb.count; i++) {
		Contact &con = arb.c[i];
		const V2 r1 = v2(con.r1x, con.r1y), r2 = v2(con.r2x, con.r2y);
		con.nMass = 1.0 / solvable(W, k_scalar(a, b, r1, r2, n));
		con.tMass = 1.0 / solvable(W, k_scalar(a, b, r1, r2, vperp(n)));
		const real dist = vdot(vadd(vsub(r2, r1), body_delta), n);
		con.bias = -ws.bias_coef * fminc(real(0.0), dist + ws.slop) / ws.dt;
		con.jBias = 0.0;
		con.bounce = vdot(rel_vel(a, b, r1, r2), n) * arb.e;
	}
}

// The impulse applications take the arbiter's body indices and whether they
// are dynamic as arguments, so a caller that has them at hand (the grid
// solve keeps them in LDS) loads the bodies beside the arbiter instead of
// after it.
// ARB: Arbiter, or the grid solve's ArbSolve (the same fields the sweeps use, kept in LDS)
template <class BA, class ARB>
__device__ void arb_apply_cached(const BA &bodies, const ARB &arb, int ia, int ib, bool da, bool db, real dt_coef)
{
	// every load is issued before the state and count are known
	BV a = bodies.load(ia), b = bodies.load(ib);
	const auto c0 = arb.c[0], c1 = arb.c[1];
	const int count = arb.count;
	if (arb.state == CPGPU_ARB_FIRST_COLLISION) return;
	const V2 n = v2(arb.nx, arb.ny);
	if (count > 0)
		apply_impulses(a, b, v2(c0.r1x, c0.r1y), v2(c0.r2x, c0.r2y), vmult(vrotate(n, v2(c0.jnAcc, c0.jtAcc)), dt_coef));
	if (count > 1)
		apply_impulses(a, b, v2(c1.r1x, c1.r1y), v2(c1.r2x, c1.r2y), vmult(vrotate(n, v2(c1.jnAcc, c1.jtAcc)), dt_coef));
	if (da) bodies.store(ia, a);
	if (db) bodies.store(ib, b);
}

// one contact's impulses (cpArbiterApplyImpulse's loop body): reads the
// contact from `in`, writes its accumulated impulses to `out`
template <class CIN, class COUT>
__device__ __forceinline__ void contact_impulse(BV &a, BV &b, V2 n, V2 surface_vr, real friction, const CIN &in,
                                                COUT &out)
{
	const real nMass = in.nMass;
	const V2 r1 = v2(in.r1x, in.r1y), r2 = v2(in.r2x, in.r2y);
	const V2 vb1 = vadd(a.vb, vmult(vperp(r1), a.wb));
	const V2 vb2 = vadd(b.vb, vmult(vperp(r2), b.wb));
	const V2 vr = vadd(rel_vel(a, b, r1, r2), surface_vr);
	const real vbn = vdot(vsub(vb2, vb1), n);
	const real vrn = vdot(vr, n);
	const real vrt = vdot(vr, vperp(n));
	const real jbn = (in.bias - vbn) * nMass;
	const real jbnOld = in.jBias;
	const real jBias = fmaxc(jbnOld + jbn, real(0.0));
	const real jn = -(in.bounce + vrn) * nMass;
	const real jnOld = in.jnAcc;
	const real jnAcc = fmaxc(jnOld + jn, real(0.0));
	const real jtMax = friction * jnAcc;
	const real jt = -vrt * in.tMass;
	const real jtOld = in.jtAcc;
	const real jtAcc = fclampc(jtOld + jt, -jtMax, jtMax);
	out.jBias = jBias;
	out.jnAcc = jnAcc;
	out.jtAcc = jtAcc;
	const V2 jb = vmult(n, jBias - jbnOld);
	apply_bias_impulse(a, vneg(jb), r1);
	apply_bias_impulse(b, jb, r2);
	apply_impulses(a, b, r1, r2, vrotate(n, v2(jnAcc - jnOld, jtAcc - jtOld)));
}

// Both contact slots are read before the count is known, so that every load
// of the arbiter is in flight at once (a slot past the count is read and
// left unused).
template <class BA, class ARB>
__device__ void arb_apply_impulse(const DevWorld &W, const BA &bodies, ARB &arb, int ia, int ib, bool da, bool db)
{
	BV a = bodies.load(ia), b = bodies.load(ib);
	const auto c0 = arb.c[0], c1 = arb.c[1];
	const int count = arb.count;
	const V2 n = v2(arb.nx, arb.ny);
	const V2 surface_vr = v2(arb.svx, arb.svy);
	const real friction = arb.u;
	if (count > 0) contact_impulse(a, b, n, surface_vr, friction, c0, arb.c[0]);
	if (count > 1) contact_impulse(a, b, n, surface_vr, friction, c1, arb.c[1]);
	if (da) bodies.store(ia, a);
	if (db) bodies.store(ib, b);
}
template <class BA>
__device__ __forceinline__ void arb_apply_impulse(const DevWorld &W, const BA &bodies, Arbiter &arb)
{
	arb_apply_impulse(W, bodies, arb, arb.ba, arb.bb, is_dynamic(W, arb.ba), is_dynamic(W, arb.bb));
}

// ---- joints (cpPivotJoint.c:24-70, cpDampedSpring.c:24-77, cpSlideJoint.c:24-89, cpGearJoint.c:24-69,
// cpPinJoint.c:24-71, cpGrooveJoint.c:24-93, cpRotaryLimitJoint.c:24-83, cpRatchetJoint.c:24-89,
// cpSimpleMotor.c:24-67, cpDampedRotarySpring.c:24-72) ----

// k_tensor (chipmunk_private.h:228-262): the inverse 2x2 effective mass of a
// point joint, with its "Unsolvable constraint." check (:255)
__device__ __forceinline__ void k_tensor(const DevWorld &W, const BV &a, const BV &b, V2 r1, V2 r2, Joint &j)
{
	const real m_sum = a.m_inv + b.m_inv;
	real k11 = m_sum, k12 = 0.0, k21 = 0.0, k22 = m_sum;
	const real a_i_inv = a.i_inv;
	const real r1xsq = r1.x * r1.x * a_i_inv;
	const real r1ysq = r1.y * r1.y * a_i_inv;
	const real r1nxy = -r1.x * r1.y * a_i_inv;
	k11 += r1ysq; k12 += r1nxy; k21 += r1nxy; k22 += r1xsq;
	const real b_i_inv = b.i_inv;
	const real r2xsq = r2.x * r2.x * b_i_inv;
	const real r2ysq = r2.y * r2.y * b_i_inv;
	const real r2nxy = -r2.x * r2.y * b_i_inv;
	k11 += r2ysq; k12 += r2nxy; k21 += r2nxy; k22 += r2xsq;
	const real det = k11 * k22 - k12 * k21;
	if (det == real(0.0)) raise_err(W.err, CPGPU_FAULT_UNSOLVABLE);
	const real det_inv = 1.0 / det;
	j.k11 = k22 * det_inv; j.k12 = -k12 * det_inv;
	j.k21 = -k21 * det_inv; j.k22 = k11 * det_inv;
}

// an angular impulse jj between the bodies (rotary limit, ratchet, simple motor)
__device__ __forceinline__ void apply_angular(BV &a, BV &b, real jj)
{
	a.w -= jj * a.i_inv;
	b.w += jj * b.i_inv;
}

__device__ void joint_prestep(const DevWorld &W, Joint &j)
{
	const real dt = W.ws[j.world].dt;
	BV a = load_bv(W, j.a), b = load_bv(W, j.b);
	const XF Ta = W.b_T[j.a], Tb = W.b_T[j.b];
	const real2 cogA = W.b_cog[j.a], cogB = W.b_cog[j.b];
	const real2 pA = W.b_p[j.a], pB = W.b_p[j.b];
	const V2 anchorA = v2(j.P[0], j.P[1]), anchorB = v2(j.P[2], j.P[3]);
	const real bc = j.bias_coef;
	if (j.type == CPGPU_JOINT_PIVOT) {
		const V2 r1 = xvect(Ta, vsub(anchorA, v2(cogA.x, cogA.y)));
		const V2 r2 = xvect(Tb, vsub(anchorB, v2(cogB.x, cogB.y)));
		j.r1x = r1.x; j.r1y = r1.y; j.r2x = r2.x; j.r2y = r2.y;
		k_tensor(W, a, b, r1, r2, j);
		const V2 delta = vsub(vadd(v2(pB.x, pB.y), r2), vadd(v2(pA.x, pA.y), r1));
		const V2 bias = vclamp(vmult(delta, -bc / dt), j.maxBias);
		j.bx = bias.x; j.by = bias.y;
	} else if (j.type == CPGPU_JOINT_DAMPED_SPRING) {
		const real restLength = j.P[4], stiffness = j.P[5], damping = j.P[6];
		const V2 r1 = xvect(Ta, vsub(anchorA, v2(cogA.x, cogA.y)));
		const V2 r2 = xvect(Tb, vsub(anchorB, v2(cogB.x, cogB.y)));
		j.r1x = r1.x; j.r1y = r1.y; j.r2x = r2.x; j.r2y = r2.y;
		const V2 delta = vsub(vadd(v2(pB.x, pB.y), r2), vadd(v2(pA.x, pA.y), r1));
		const real dist = vlength(delta);
		const V2 n = vmult(delta, 1.0 / (dist != 0.0 ? dist : __builtin_inf()));
		j.nx = n.x; j.ny = n.y;
		const real k = solvable(W, k_scalar(a, b, r1, r2, n));
		j.nMass = 1.0 / k;
		j.target_vrn = 0.0;
		j.v_coef = 1.0 - cp_exp(-damping * dt * k);
		const real f_spring = (restLength - dist) * stiffness;
		const real j_spring = j.jAcc = f_spring * dt;
		apply_impulses(a, b, r1, r2, vmult(n, j_spring));
		if (is_dynamic(W, j.a)) store_bv(W, j.a, a);
		if (is_dynamic(W, j.b)) store_bv(W, j.b, b);
	} else if (j.type == CPGPU_JOINT_SLIDE) {
		const real mn = j.P[4], mx = j.P[5];
		const V2 r1 = xvect(Ta, vsub(anchorA, v2(cogA.x, cogA.y)));
		const V2 r2 = xvect(Tb, vsub(anchorB, v2(cogB.x, cogB.y)));
		j.r1x = r1.x; j.r1y = r1.y; j.r2x = r2.x; j.r2y = r2.y;
		const V2 delta = vsub(vadd(v2(pB.x, pB.y), r2), vadd(v2(pA.x, pA.y), r1));
		const real dist = vlength(delta);
		real pdist = 0.0;
		V2 n;
		if (dist > mx) { pdist = dist - mx; n = vnormalize(delta); }
		else if (dist < mn) { pdist = mn - dist; n = vneg(vnormalize(delta)); }
		else { n = v2(0.0, 0.0); j.jnAcc = 0.0; }
		j.nx = n.x; j.ny = n.y;
		j.nMass = 1.0 / solvable(W, k_scalar(a, b, r1, r2, n));
		const real maxBias = j.maxBias;
		j.bias = fclampc(-bc * pdist / dt, -maxBias, maxBias);
	} else if (j.type == CPGPU_JOINT_GEAR) {
		const real phase = j.P[0], ratio = j.P[1], ratio_inv = j.P[2];
		j.iSum = 1.0 / (a.i_inv * ratio_inv + ratio * b.i_inv);
		const real maxBias = j.maxBias;
		j.bias = fclampc(-bc * (W.b_a[j.b] * ratio - W.b_a[j.a] - phase) / dt, -maxBias, maxBias);
	} else if (j.type == CPGPU_JOINT_PIN) {
		const V2 r1 = xvect(Ta, vsub(anchorA, v2(cogA.x, cogA.y)));
		const V2 r2 = xvect(Tb, vsub(anchorB, v2(cogB.x, cogB.y)));
		j.r1x = r1.x; j.r1y = r1.y; j.r2x = r2.x; j.r2y = r2.y;
		const V2 delta = vsub(vadd(v2(pB.x, pB.y), r2), vadd(v2(pA.x, pA.y), r1));
		const real dist = vlength(delta);
		const V2 n = vmult(delta, 1.0 / (dist != 0.0 ? dist : __builtin_inf()));
		j.nx = n.x; j.ny = n.y;
		j.nMass = 1.0 / solvable(W, k_scalar(a, b, r1, r2, n));
		const real maxBias = j.maxBias;
		j.bias = fclampc(-bc * (dist - j.P[4]) / dt, -maxBias, maxBias);
	} else if (j.type == CPGPU_JOINT_GROOVE) {
		// P: the groove's ends on body A, the anchor on body B, the groove's normal
		const V2 ta = xpoint(Ta, v2(j.P[0], j.P[1])), tb = xpoint(Ta, v2(j.P[2], j.P[3]));
		const V2 n = xvect(Ta, v2(j.P[6], j.P[7]));
		const real d = vdot(ta, n);
		j.nx = n.x; j.ny = n.y;
		const V2 r2 = xvect(Tb, vsub(v2(j.P[4], j.P[5]), v2(cogB.x, cogB.y)));
		const V2 p_a = v2(pA.x, pA.y);
		const real td = vcross(vadd(v2(pB.x, pB.y), r2), n);
		V2 r1;
		if (td <= vcross(ta, n)) { j.clamp = 1.0; r1 = vsub(ta, p_a); }
		else if (td >= vcross(tb, n)) { j.clamp = -1.0; r1 = vsub(tb, p_a); }
		else { j.clamp = 0.0; r1 = vsub(vadd(vmult(vperp(n), -td), vmult(n, d)), p_a); }
		j.r1x = r1.x; j.r1y = r1.y; j.r2x = r2.x; j.r2y = r2.y;
		k_tensor(W, a, b, r1, r2, j);
		const V2 delta = vsub(vadd(v2(pB.x, pB.y), r2), vadd(p_a, r1));
		const V2 bias = vclamp(vmult(delta, -bc / dt), j.maxBias);
		j.bx = bias.x; j.by = bias.y;
	} else if (j.type == CPGPU_JOINT_ROTARY_LIMIT || j.type == CPGPU_JOINT_RATCHET) {
		const real delta = W.b_a[j.b] - W.b_a[j.a];
		real pdist = 0.0;
		if (j.type == CPGPU_JOINT_ROTARY_LIMIT) {
			if (delta > j.P[1]) pdist = j.P[1] - delta;
			else if (delta < j.P[0]) pdist = j.P[0] - delta;
		} else {
			const real phase = j.P[0], ratchet = j.P[1];
			const real diff = j.angle - delta;
			if (diff * ratchet > 0.0) pdist = diff;
			else j.angle = floor((delta - phase) / ratchet) * ratchet + phase;
		}
		j.iSum = 1.0 / (a.i_inv + b.i_inv);
		const real maxBias = j.maxBias;
		j.bias = fclampc(-bc * pdist / dt, -maxBias, maxBias);
		if (!j.bias) j.jAcc = 0.0;
	} else if (j.type == CPGPU_JOINT_SIMPLE_MOTOR) {
		j.iSum = 1.0 / (a.i_inv + b.i_inv);
	} else {
		// damped rotary spring, the default torque (relativeAngle - restAngle) stiffness
		const real restAngle = j.P[0], stiffness = j.P[1], damping = j.P[2];
		const real moment = a.i_inv + b.i_inv;
		j.iSum = 1.0 / moment;
		j.v_coef = 1.0 - cp_exp(-damping * dt * moment);
		j.target_vrn = 0.0;
		const real j_spring = ((W.b_a[j.a] - W.b_a[j.b]) - restAngle) * stiffness * dt;
		j.jAcc = j_spring;
		a.w -= j_spring * a.i_inv;
		b.w += j_spring * b.i_inv;
		if (is_dynamic(W, j.a)) store_bv(W, j.a, a);
		if (is_dynamic(W, j.b)) store_bv(W, j.b, b);
	}
}

// dt_coef: the joint's space's cached-impulse scale; ia, ib: its bodies, da, db: whether
// they are dynamic (given by a caller that has them at hand, so the bodies load beside the
// joint instead of after it)
template <class BA>
__device__ void joint_apply_cached(const BA &bodies, const Joint &j, real dt_coef, int ia, int ib, bool da, bool db)
{
	BV a = bodies.load(ia), b = bodies.load(ib);
	if (j.type == CPGPU_JOINT_DAMPED_SPRING || j.type == CPGPU_JOINT_DAMPED_ROTARY_SPRING) return;
	const V2 r1 = v2(j.r1x, j.r1y), r2 = v2(j.r2x, j.r2y);
	if (j.type == CPGPU_JOINT_PIVOT || j.type == CPGPU_JOINT_GROOVE) {
		apply_impulses(a, b, r1, r2, vmult(v2(j.jAx, j.jAy), dt_coef));
	} else if (j.type == CPGPU_JOINT_SLIDE || j.type == CPGPU_JOINT_PIN) {
		apply_impulses(a, b, r1, r2, vmult(v2(j.nx, j.ny), j.jnAcc * dt_coef));
	} else if (j.type == CPGPU_JOINT_GEAR) {
		const real ratio_inv = j.P[2];
		const real jj = j.jAcc * dt_coef;
		a.w -= jj * a.i_inv * ratio_inv;
		b.w += jj * b.i_inv;
	} else {
		apply_angular(a, b, j.jAcc * dt_coef);
	}
	if (da) bodies.store(ia, a);
	if (db) bodies.store(ib, b);
}
template <class BA>
__device__ __forceinline__ void joint_apply_cached(const DevWorld &W, const BA &bodies, const Joint &j, real dt_coef)
{
	joint_apply_cached(bodies, j, dt_coef, j.a, j.b, is_dynamic(W, j.a), is_dynamic(W, j.b));
}

// dt: the joint's space's time step; ia, ib, da, db as for joint_apply_cached
template <class BA>
__device__ void joint_apply_impulse(const BA &bodies, Joint &j, real dt, int ia, int ib, bool da, bool db)
{
	BV a = bodies.load(ia), b = bodies.load(ib);
	const V2 r1 = v2(j.r1x, j.r1y), r2 = v2(j.r2x, j.r2y);
	if (j.type == CPGPU_JOINT_PIVOT) {
		const V2 vr = rel_vel(a, b, r1, r2);
		const V2 d = vsub(v2(j.bx, j.by), vr);
		V2 jj = v2(d.x * j.k11 + d.y * j.k12, d.x * j.k21 + d.y * j.k22);
		const V2 jOld = v2(j.jAx, j.jAy);
		const V2 jAcc = vclamp(vadd(jOld, jj), j.maxForce * dt);
		j.jAx = jAcc.x; j.jAy = jAcc.y;
		jj = vsub(jAcc, jOld);
		apply_impulses(a, b, r1, r2, jj);
	} else if (j.type == CPGPU_JOINT_DAMPED_SPRING) {
		const V2 n = v2(j.nx, j.ny);
		const real vrn = vdot(rel_vel(a, b, r1, r2), n);
		const real v_damp = (j.target_vrn - vrn) * j.v_coef;
		j.target_vrn = vrn + v_damp;
		const real j_damp = v_damp * j.nMass;
		j.jAcc += j_damp;
		apply_impulses(a, b, r1, r2, vmult(n, j_damp));
	} else if (j.type == CPGPU_JOINT_SLIDE) {
		const V2 n = v2(j.nx, j.ny);
		if (veql(n, v2(0.0, 0.0))) return;
		const V2 vr = rel_vel(a, b, r1, r2);
		const real vrn = vdot(vr, n);
		real jn = (j.bias - vrn) * j.nMass;
		const real jnOld = j.jnAcc;
		j.jnAcc = fclampc(jnOld + jn, -j.maxForce * dt, 0.0);
		jn = j.jnAcc - jnOld;
		apply_impulses(a, b, r1, r2, vmult(n, jn));
	} else if (j.type == CPGPU_JOINT_GEAR) {
		const real ratio = j.P[1], ratio_inv = j.P[2];
		const real wr = b.w * ratio - a.w;
		const real jMax = j.maxForce * dt;
		real jj = (j.bias - wr) * j.iSum;
		const real jOld = j.jAcc;
		j.jAcc = fclampc(jOld + jj, -jMax, jMax);
		jj = j.jAcc - jOld;
		a.w -= jj * a.i_inv * ratio_inv;
		b.w += jj * b.i_inv;
	} else if (j.type == CPGPU_JOINT_PIN) {
		const V2 n = v2(j.nx, j.ny);
		const real vrn = vdot(rel_vel(a, b, r1, r2), n);
		const real jnMax = j.maxForce * dt;
		real jn = (j.bias - vrn) * j.nMass;
		const real jnOld = j.jnAcc;
		j.jnAcc = fclampc(jnOld + jn, -jnMax, jnMax);
		jn = j.jnAcc - jnOld;
		apply_impulses(a, b, r1, r2, vmult(n, jn));
	} else if (j.type == CPGPU_JOINT_GROOVE) {
		const V2 vr = rel_vel(a, b, r1, r2);
		const V2 d = vsub(v2(j.bx, j.by), vr);
		const V2 jj = v2(d.x * j.k11 + d.y * j.k12, d.x * j.k21 + d.y * j.k22);
		const V2 jOld = v2(j.jAx, j.jAy);
		// grooveConstrain (cpGrooveJoint.c:74-79)
		const V2 jc = vadd(jOld, jj), n = v2(j.nx, j.ny);
		const V2 jClamp = (j.clamp * vcross(jc, n) > 0.0) ? jc : vmult(n, vdot(jc, n) / vdot(n, n));
		const V2 jAcc = vclamp(jClamp, j.maxForce * dt);
		j.jAx = jAcc.x; j.jAy = jAcc.y;
		apply_impulses(a, b, r1, r2, vsub(jAcc, jOld));
	} else if (j.type == CPGPU_JOINT_ROTARY_LIMIT || j.type == CPGPU_JOINT_RATCHET) {
		if (!j.bias) return;
		const real wr = b.w - a.w;
		const real jMax = j.maxForce * dt;
		const real jj = -(j.bias + wr) * j.iSum;
		const real jOld = j.jAcc;
		if (j.type == CPGPU_JOINT_RATCHET) {
			const real ratchet = j.P[1];
			j.jAcc = fclampc((jOld + jj) * ratchet, 0.0, jMax * fabs(ratchet)) / ratchet;
		} else if (j.bias < 0.0) {
			j.jAcc = fclampc(jOld + jj, 0.0, jMax);
		} else {
			j.jAcc = fclampc(jOld + jj, -jMax, 0.0);
		}
		apply_angular(a, b, j.jAcc - jOld);
	} else if (j.type == CPGPU_JOINT_SIMPLE_MOTOR) {
		const real wr = b.w - a.w + j.P[0];
		const real jMax = j.maxForce * dt;
		const real jj = -wr * j.iSum;
		const real jOld = j.jAcc;
		j.jAcc = fclampc(jOld + jj, -jMax, jMax);
		apply_angular(a, b, j.jAcc - jOld);
	} else {
		const real wrn = a.w - b.w;
		const real w_damp = (j.target_vrn - wrn) * j.v_coef;
		j.target_vrn = wrn + w_damp;
		const real j_damp = w_damp * j.iSum;
		j.jAcc += j_damp;
		a.w += j_damp * a.i_inv;
		b.w -= j_damp * b.i_inv;
	}
	if (da) bodies.store(ia, a);
	if (db) bodies.store(ib, b);
}
template <class BA>
__device__ __forceinline__ void joint_apply_impulse(const DevWorld &W, const BA &bodies, Joint &j, real dt)
{
	joint_apply_impulse(bodies, j, dt, j.a, j.b, is_dynamic(W, j.a), is_dynamic(W, j.b));
}

// ---- the solve routine, shared by the workgroup-per-world and grid kernels ----

__device__ __forceinline__ int dyn_or_neg(const DevWorld &W, int b) { return is_dynamic(W, b) ? b : -1; }

// cpBodyUpdateVelocity (cpBody.c:493-509) with the space's damping and
// gravity (cpSpaceStep.c:399-404); kinematic and static bodies keep their velocity
__device__ __forceinline__ void integrate_velocity(const DevWorld &W, int b)
{
	if (W.b_type[b] != CPGPU_BODY_DYNAMIC) return;
	const WorldStep ws = W.ws[W.b_world[b]];
	BodyVel bv = W.b_vel[b];
	const real2 f = W.b_f[b];
	const real m = W.b_m[b];
	// "Body's mass and moment must be positive to simulate." (cpBody.c:498); the
	// moment is positive when its inverse lies in [0, inf)
	if (!(m > real(0.0) && bv.i_inv >= real(0.0) && bv.i_inv < __builtin_inf()))
		raise_body_fault(W.err, CPGPU_FAULT_BODY_MASS, b);
	const V2 v = vadd(vmult(v2(bv.vx, bv.vy), ws.damping),
	                  vmult(vadd(v2(ws.gx, ws.gy), vmult(v2(f.x, f.y), bv.m_inv)), ws.dt));
	bv.vx = v.x; bv.vy = v.y;
	bv.w = bv.w * ws.damping + W.b_t[b] * bv.i_inv * ws.dt;
	W.b_vel[b] = bv;
	W.b_f[b] = make_real2(0.0, 0.0);
	W.b_t[b] = 0.0;
	// the sanity check after the update (cpBody.c:508): position and angle are
	// as the position update checked them this step, force and torque now zero
	if (!body_sane(m, bv.m_inv, bv.i_inv, real(0.0), real(0.0), bv.vx, bv.vy, real(0.0), real(0.0), real(0.0), bv.w,
	               real(0.0)))
		raise_body_fault(W.err, CPGPU_FAULT_BODY_STATE, b);
}

// colouring priority: a bijective hash of the shape pair's indices within its
// own space, so a space is coloured (and stepped) identically alone or in a batch
__device__ __forceinline__ unsigned long long arbiter_priority(const DevWorld &W, const Arbiter &arb)
{
	const int s0 = W.w_shape0[arb.world];
	return mix64(pair_key(arb.sa - s0, arb.sb - s0));
}

// Synchronisation of the threads taking part: one workgroup, or the whole
// co-resident grid.
// sweep_tid() numbers the threads for the joint sweeps.  On the grid it
// interleaves the workgroups (thread t of workgroup g is g + t * gridDim.x),
// so the joints of a colour, however few, are spread over every workgroup
// instead of filling the first ones: a sweep is bound by each CU's scattered
// joint and body loads.
struct BlockSync {
	int tid, stride;
	__device__ void sync() { __syncthreads(); }
	__device__ int sweep_tid() const { return tid; }
};
struct GridSync {
	int tid, stride;
	GridBarrier *bar;
	int32_t *err;
	unsigned int epoch;   // barriers passed (used by lane 0)
	__device__ void sync() { grid_sync(bar, err, epoch); }
	__device__ int sweep_tid() const { return (int)(blockIdx.x + threadIdx.x * gridDim.x); }
};

// The elements one solve covers.  Arbiters are list[0..n) (indices into
// S.act; list == nullptr means 0..n-1).  col_off / col_cursor / left live in
// LDS.
struct SolveSet {
	const int32_t *list;
	int n;
	int32_t *order;            // output: the n arbiters in colour order
	int32_t *col_off;          // [CPGPU_MAX_COLOURS + 1]
	int32_t *col_cursor;       // [CPGPU_MAX_COLOURS]
	int32_t *left;
	int body0, nb;             // bodies to integrate: [body0, body0 + nb)
	const int32_t *jorder;     // joints in colour order
	const int32_t *jcol_off;   // [njcol + 1]
	int njcol;
	int iterations;            // sweeps of the set's space
	real dt_coef;              // its cached-impulse scale (dt / previous dt, 0 on the space's first step)
	real dt;                   // its time step
	BodyVel *lds_bodies;       // LDS room for lds_cap bodies, or nullptr
	int lds_cap;
	int unit;                  // solve unit whose contacts / bodies are counted (W.totals[2 + 2 unit ...])
	int order_base;            // position of order[0] in S.order (records are indexed by it)
	int cset;                  // claim arrays used by the colouring (units coloured concurrently
	                           // with shared bodies use different ones)
	// colouring with the claim arrays in LDS (lclaim != nullptr): the set's dynamic bodies
	// are numbered 0 .. by slot[body], or by body - slot_base when slot is nullptr
	unsigned long long *lclaim;   // [2][lcap]
	int32_t *ltaken;              // [lcap]
	int lcap;
	const int32_t *slot;
	int slot_base;
};

// a wave-reduced 64-bit add to a counter
__device__ __forceinline__ void wave_add(unsigned long long *counter, unsigned long long v)
{
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
	if (__lane_id() == 0 && v) atomicAdd(counter, v);
}

// One arbiter's part of a colouring round (see the file comment): its
// dynamic bodies bd (-1 for none), priority and colour so far.
// Claim phase: an uncoloured arbiter none of whose bodies colour c has taken
// claims them with an atomic min of its priority.
__device__ __forceinline__ void colour_claim(const int32_t *taken, int2 bd, unsigned long long prio, int col, int c,
                                             unsigned long long *cl)
{
	if (col >= 0) return;
	if ((bd.x >= 0 && taken[bd.x] == c) || (bd.y >= 0 && taken[bd.y] == c)) return;
	if (bd.x >= 0) atomicMin(&cl[bd.x], prio);
	if (bd.y >= 0) atomicMin(&cl[bd.y], prio);
}
// Resolve phase: an arbiter holding all its claims takes colour c and its
// bodies; the others reset their bodies' claims for the next round.
// Returns 1 when the arbiter took colour c.
__device__ __forceinline__ int colour_resolve(int32_t *taken, int2 bd, unsigned long long prio, int &col, int c,
                                              unsigned long long *cl, unsigned long long *other)
{
	if (col >= 0) return 0;
	const bool win = (bd.x < 0 || __hip_atomic_load(&cl[bd.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == prio) &&
	                 (bd.y < 0 || __hip_atomic_load(&cl[bd.y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == prio);
	if (win) {
		col = c;
		if (bd.x >= 0) taken[bd.x] = c;
		if (bd.y >= 0) taken[bd.y] = c;
		return 1;
	}
	if (bd.x >= 0) other[bd.x] = ~0ull;
	if (bd.y >= 0) other[bd.y] = ~0ull;
	return 0;
}

// Run f(k) for every position k of colours [0, n_colours) in colour order,
// synchronising between colours.
template <class Sync, class F>
__device__ void for_colours(Sync &sy, const int32_t *col_off, int n_colours, F f)
{
	for (int q = 0; q < n_colours; q++) {
		for (int k = col_off[q] + sy.sweep_tid(); k < col_off[q + 1]; k += sy.stride) f(k);
		sy.sync();
	}
}

template <class BA> __device__ __forceinline__ void record_arbiter(ArbRec *r, const BA &bodies, const Arbiter &arb)
{
	const BV a = bodies.load(arb.ba), b = bodies.load(arb.bb);
	r->a[0] = a.v.x; r->a[1] = a.v.y; r->a[2] = a.w; r->a[3] = a.vb.x; r->a[4] = a.vb.y; r->a[5] = a.wb;
	r->b[0] = b.v.x; r->b[1] = b.v.y; r->b[2] = b.w; r->b[3] = b.vb.x; r->b[4] = b.vb.y; r->b[5] = b.wb;
	for (int i = 0; i < 2; i++) {
		r->acc[i][0] = arb.c[i].jnAcc;
		r->acc[i][1] = arb.c[i].jtAcc;
		r->acc[i][2] = arb.c[i].jBias;
	}
}

template <class BA> __device__ __forceinline__ void record_joint(JointRec *r, const BA &bodies, const Joint &j)
{
	const BV a = bodies.load(j.a), b = bodies.load(j.b);
	r->a[0] = a.v.x; r->a[1] = a.v.y; r->a[2] = a.w;
	r->b[0] = b.v.x; r->b[1] = b.v.y; r->b[2] = b.w;
	if (j.type == CPGPU_JOINT_PIVOT) { r->acc[0] = j.jAx; r->acc[1] = j.jAy; }
	else if (j.type == CPGPU_JOINT_DAMPED_SPRING) { r->acc[0] = j.target_vrn; r->acc[1] = j.jAcc; }
	else if (j.type == CPGPU_JOINT_SLIDE) { r->acc[0] = j.jnAcc; r->acc[1] = 0.0; }
	else { r->acc[0] = j.jAcc; r->acc[1] = 0.0; }
}

// cached impulses (cpSpaceStep.c:407-415), then the impulse iterations
// (cpSpaceStep.c:418-427): arbiters colour by colour, then joints colour by colour
template <class Sync, class BA>
__device__ void solve_impulses(const DevWorld &W, const SolveBuffers &S, const SolveSet &T, Sync &sy, int n_colours,
                               const BA &bodies)
{
	const int tid = sy.sweep_tid(), stride = sy.stride;
	// the sorted arbiters' bodies, written by the colour sort
	const int2 *obody = S.cbody + T.order_base;
	auto body_of = [](int e) { return e >= 0 ? e : ~e; };
	for_colours(sy, T.col_off, n_colours, [&](int k) {
		const int2 e = obody[k];
		arb_apply_cached(bodies, S.act[T.order[k]], body_of(e.x), body_of(e.y), e.x >= 0, e.y >= 0, T.dt_coef);
	});
	for (int jc = 0; jc < T.njcol; jc++) {
		for (int k = T.jcol_off[jc] + tid; k < T.jcol_off[jc + 1]; k += stride) {
			const int2 e = S.jw_body[k];
			if (S.jrec) record_joint(S.jrec + k, bodies, W.joints[T.jorder[k]]);
			joint_apply_cached(bodies, W.joints[T.jorder[k]], T.dt_coef, body_of(e.x), body_of(e.y), e.x >= 0, e.y >= 0);
		}
		sy.sync();
	}
	for (int it = 0; it < T.iterations; it++) {
		// every arbiter of the set is of its one space, whose iterations bound the loop
		for_colours(sy, T.col_off, n_colours, [&](int k) {
			const int2 e = obody[k];
			Arbiter &arb = S.act[T.order[k]];
			if (S.rec) record_arbiter(S.rec + (size_t)it * S.rec_stride + T.order_base + k, bodies, arb);
			arb_apply_impulse(W, bodies, arb, body_of(e.x), body_of(e.y), e.x >= 0, e.y >= 0);
		});
		for (int jc = 0; jc < T.njcol; jc++) {
			for (int k = T.jcol_off[jc] + tid; k < T.jcol_off[jc + 1]; k += stride) {
				const int2 e = S.jw_body[k];
				Joint &j = W.joints[T.jorder[k]];
				if (S.jrec) record_joint(S.jrec + (size_t)(it + 1) * W.n_joints + k, bodies, j);
				joint_apply_impulse(bodies, j, T.dt, body_of(e.x), body_of(e.y), e.x >= 0, e.y >= 0);
			}
			sy.sync();
		}
	}
}

// Colour the set's arbiters and sort them into colour order (T.order,
// T.col_off) on one workgroup; returns the colour count.  Thread t handles
// positions t, t + blockDim.x, ...; the first COLOUR_REGS of them keep their
// bodies, priority and colour in registers through the rounds, the rest in
// the position-indexed arrays S.cbody / S.cprio / S.ccol.
constexpr int COLOUR_REGS = 4;
__device__ __forceinline__ int colour_set(const DevWorld &W, const SolveBuffers &S, const SolveSet &T)
{
	const int tid = threadIdx.x, stride = blockDim.x, n = T.n;
	unsigned long long *claim0, *claim1;
	int32_t *taken;
	if (T.lclaim) {
		claim0 = T.lclaim;
		claim1 = T.lclaim + T.lcap;
		taken = T.ltaken;
	} else {
		claim0 = S.claim + (size_t)T.cset * 2 * S.n_bodies;
		claim1 = claim0 + S.n_bodies;
		taken = S.taken + (size_t)T.cset * S.n_bodies;
	}
	auto arb_index = [&](int k) { return T.list ? T.list[k] : k; };
	int2 *cbody = S.cbody + T.order_base;
	unsigned long long *cprio = S.cprio + T.order_base;
	int32_t *ccol = S.ccol + T.order_base;
	int2 rb[COLOUR_REGS];
	unsigned long long rp[COLOUR_REGS];
	int rc[COLOUR_REGS];
	const int k_mem = tid + COLOUR_REGS * stride;   // first position kept in memory

	// ---- colouring ----
	auto init = [&](int k, int2 &bd, unsigned long long &prio) {
		const Arbiter &arb = S.act[arb_index(k)];
		bd = make_int2(dyn_or_neg(W, arb.ba), dyn_or_neg(W, arb.bb));
		if (T.lclaim) {
			auto slot_of = [&](int b) { return T.slot ? T.slot[b] : b - T.slot_base; };
			bd = make_int2(bd.x >= 0 ? slot_of(bd.x) : -1, bd.y >= 0 ? slot_of(bd.y) : -1);
		}
		prio = arbiter_priority(W, arb);
		if (bd.x >= 0) { claim0[bd.x] = ~0ull; claim1[bd.x] = ~0ull; taken[bd.x] = -1; }
		if (bd.y >= 0) { claim0[bd.y] = ~0ull; claim1[bd.y] = ~0ull; taken[bd.y] = -1; }
	};
#pragma unroll
	for (int r = 0; r < COLOUR_REGS; r++) {
		rc[r] = -1;
		rb[r] = make_int2(-1, -1);
		rp[r] = 0;
		if (tid + r * stride < n) init(tid + r * stride, rb[r], rp[r]);
	}
	for (int k = k_mem; k < n; k += stride) {
		int2 bd;
		unsigned long long prio;
		init(k, bd, prio);
		cbody[k] = bd;
		cprio[k] = prio;
		ccol[k] = -1;
	}
	for (int k = tid; k <= CPGPU_MAX_COLOURS; k += stride) T.col_off[k] = 0;
	if (tid == 0) *T.left = n;
	__syncthreads();

	int c = 0, sub = 0;
	while (true) {
		const int left = __hip_atomic_load(T.left, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		if (left == 0) break;
		if (c >= CPGPU_MAX_COLOURS) { if (tid == 0) raise_err(W.err, CPGPU_ERR_ARBITERS_OVERFLOW); break; }
		unsigned long long *cl = (sub & 1) ? claim1 : claim0;
		unsigned long long *other = (sub & 1) ? claim0 : claim1;
#pragma unroll
		for (int r = 0; r < COLOUR_REGS; r++)
			if (tid + r * stride < n) colour_claim(taken, rb[r], rp[r], rc[r], c, cl);
		for (int k = k_mem; k < n; k += stride) colour_claim(taken, cbody[k], cprio[k], ccol[k], c, cl);
		__syncthreads();
		int won = 0;
#pragma unroll
		for (int r = 0; r < COLOUR_REGS; r++)
			if (tid + r * stride < n) won += colour_resolve(taken, rb[r], rp[r], rc[r], c, cl, other);
		for (int k = k_mem; k < n; k += stride) {
			int col = ccol[k];
			if (colour_resolve(taken, cbody[k], cprio[k], col, c, cl, other)) {
				ccol[k] = col;
				won++;
			}
		}
		if (won) atomicSub(T.left, won);
		__syncthreads();
		if (++sub % COLOUR_SUBROUNDS == 0) c++;
	}
	const int n_colours = (sub % COLOUR_SUBROUNDS == 0) ? c : c + 1;

	// ---- colour sort ----
	// f(position, colour) for this thread's positions
	auto each = [&](auto f) {
#pragma unroll
		for (int r = 0; r < COLOUR_REGS; r++)
			if (tid + r * stride < n) f(tid + r * stride, rc[r]);
		for (int k = k_mem; k < n; k += stride) f(k, ccol[k]);
	};
	unsigned long long n_contacts = 0;
	each([&](int k, int col) {
		const int a = arb_index(k);
		S.colour[a] = col;
		if (col >= 0) atomicAdd(&T.col_off[col], 1);
		n_contacts += S.act[a].count;
	});
	wave_add(&W.totals[2 + 2 * T.unit], n_contacts);
	__syncthreads();
	if (tid == 0) {
		int run = 0;
		for (int q = 0; q < n_colours; q++) {
			const int cnt = T.col_off[q];
			T.col_off[q] = run;
			T.col_cursor[q] = run;
			run += cnt;
		}
		T.col_off[n_colours] = run;
	}
	__syncthreads();
	// beside each sorted arbiter, its bodies (~b for one that is not dynamic) in the
	// colouring's body array, which the colouring no longer reads
	each([&](int k, int col) {
		if (col >= 0) {
			const int a = arb_index(k);
			const Arbiter &arb = S.act[a];
			const int pos = atomicAdd(&T.col_cursor[col], 1);
			T.order[pos] = a;
			cbody[pos] = make_int2(is_dynamic(W, arb.ba) ? arb.ba : ~arb.ba, is_dynamic(W, arb.bb) ? arb.bb : ~arb.bb);
		}
	});
	return n_colours;
}

template <class Sync>
__device__ void solve_set(const DevWorld &W, const SolveBuffers &S, const SolveSet &Tin, Sync &sy)
{
	const int tid = sy.tid, stride = sy.stride, n = Tin.n;
	auto arb_index = [&](int k) { return Tin.list ? Tin.list[k] : k; };
	SolveSet T = Tin;
	if (T.lds_bodies && T.nb <= T.lds_cap) {
		// the world's bodies are staged in LDS only after the colouring: until then the claim
		// arrays live there (20 bytes a body, no more than a staged body's)
		T.lclaim = reinterpret_cast<unsigned long long *>(T.lds_bodies);
		T.lcap = T.nb;
		T.ltaken = reinterpret_cast<int32_t *>(T.lclaim + 2 * T.nb);
		T.slot = nullptr;
		T.slot_base = T.body0;
	}
	static_assert(sizeof(BodyVel) >= 2 * sizeof(unsigned long long) + sizeof(int32_t), "claims fit a body's room");
	const int n_colours = colour_set(W, S, T);
	// ---- pre-steps (cpSpaceStep.c:383-396) ----
	for (int k = tid; k < n; k += stride) arb_prestep(W, S.act[arb_index(k)]);
	sy.sync();
	for (int jc = 0; jc < T.njcol; jc++) {
		for (int k = T.jcol_off[jc] + tid; k < T.jcol_off[jc + 1]; k += stride) joint_prestep(W, W.joints[T.jorder[k]]);
		sy.sync();
	}
	// ---- velocity integration (cpSpaceStep.c:399-404, cpBody.c:493-509) ----
	unsigned long long n_int = 0;
	for (int q = tid; q < T.nb; q += stride) {
		integrate_velocity(W, T.body0 + q);
		n_int++;
	}
	wave_add(&W.totals[3 + 2 * T.unit], n_int);
	sy.sync();
	// ---- impulses: on a copy of the bodies in LDS when the set is one world's range that fits ----
	if (T.lds_bodies && T.nb <= T.lds_cap) {
		for (int q = tid; q < T.nb; q += stride) T.lds_bodies[q] = W.b_vel[T.body0 + q];
		sy.sync();
		solve_impulses(W, S, T, sy, n_colours, LdsBodies{T.lds_bodies, T.body0});
		for (int q = tid; q < T.nb; q += stride)
			if (is_dynamic(W, T.body0 + q)) W.b_vel[T.body0 + q] = T.lds_bodies[q];
		sy.sync();
	} else {
		solve_impulses(W, S, T, sy, n_colours, GlobalBodies{&W});
	}
}

__device__ __forceinline__ int lower_bound_u64(const uint64_t *keys, int lo, int hi, uint64_t key)
{
	while (lo < hi) {
		const int mid = (lo + hi) >> 1;
		if (keys[mid] < key) lo = mid + 1; else hi = mid;
	}
	return lo;
}

// One workgroup per small world: every synchronisation is a workgroup barrier.
template <int THREADS>
// At least 4 waves a SIMD for the 1024-thread size (which needs them); the smaller sizes,
// whose occupancy their spaces' LDS bodies bound, may take the registers to run unspilled.
__global__ void __launch_bounds__(THREADS, THREADS >= 1024 ? 4 : 2)
	k_solve_worlds(const DevWorld W, const SolveBuffers S, const int32_t *worlds, int lds_cap, int unit)
{
	extern __shared__ BodyVel lds_bodies[];
	__shared__ int32_t col_off[CPGPU_MAX_COLOURS + 1];
	__shared__ int32_t col_cursor[CPGPU_MAX_COLOURS];
	__shared__ int32_t left;
	SolveSet T = {};
	T.col_off = col_off;
	T.col_cursor = col_cursor;
	T.left = &left;
	T.cset = 0;
	T.lds_bodies = lds_cap > 0 ? lds_bodies : nullptr;
	T.lds_cap = lds_cap;
	T.unit = unit;
	BlockSync sy{(int)threadIdx.x, (int)blockDim.x};
	{
		const int w = worlds[blockIdx.x];
		if (S.arb_by_world) {
			T.list = S.arb_by_world + S.warb_off[w];
			T.n = S.warb_off[w + 1] - S.warb_off[w];
			T.order_base = S.warb_off[w];
		} else {   // a single space: every solved arbiter is its own
			T.list = nullptr;
			T.n = *S.n_act;
			T.order_base = 0;
		}
		T.order = S.order + T.order_base;
		T.body0 = W.w_body0[w];
		T.nb = W.w_body0[w + 1] - W.w_body0[w];
		T.jorder = S.jw_order;
		T.jcol_off = S.jw_col + S.jw_cbase[w];
		T.njcol = S.jw_ncol[w];
		T.iterations = W.ws[w].iterations;
		T.dt_coef = W.ws[w].dt_coef;
		T.dt = W.ws[w].dt;
		solve_set(W, S, T, sy);
	}
}

// Sort key of grid-set unit u (the strips, even seams and odd seams of every
// large world, then the wide set), the unit of an arbiter whose dynamic
// bodies lie in strips pa and pb (-1: none) of the l-th large world, and the
// claim arrays of a unit's colouring.
__device__ __forceinline__ uint64_t excluded_key(const SolveBuffers &S)
{
	return (uint64_t)S.n_worlds + (uint64_t)grid_units(S.n_large_worlds, S.parts);
}
__device__ __forceinline__ uint64_t unit_key(const SolveBuffers &S, int u)
{
	return (uint64_t)S.n_worlds + (uint64_t)u;
}
__device__ __forceinline__ int unit_of(int l, int pa, int pb, int n_large, int parts)
{
	if (pa < 0 || pb < 0 || pa == pb) return l * parts + max(pa, pb);
	if (abs(pa - pb) != 1) return grid_units(n_large, parts) - 1;
	const int q = min(pa, pb), e0 = n_large * parts;
	return q % 2 == 0 ? e0 + l * even_seams(parts) + q / 2
	                  : e0 + n_large * even_seams(parts) + l * odd_seams(parts) + q / 2;
}
__device__ __forceinline__ int unit_claims(int u, int n_large, int parts)
{
	if (u < n_large * parts) return 0;
	if (u < n_large * (parts + even_seams(parts))) return 1;
	if (u < n_large * (2 * parts - 1)) return 2;
	return 3;
}

// Completion counters of the grid-set units (passes finished): thread 0
// waits until a unit has finished `need` passes, giving up (and flagging a
// barrier time-out) after about a second, or when another wait gave up.
__device__ void unit_wait(const int32_t *done, int need, int32_t *err)
{
	long spins = 0;
	while (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
		__builtin_amdgcn_s_sleep(1);
		if (++spins > (1L << 25)) {
			atomicOr(err, CPGPU_ERR_BARRIER_TIMEOUT);
			break;
		}
		if ((spins & 1023) == 0 &&
		    (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & CPGPU_ERR_BARRIER_TIMEOUT))
			break;
	}
}

constexpr int GRID_LDS_COLS = 2;   // columns per workgroup whose units' colour offsets stay in LDS
constexpr int GRID_LDS_POS = 1024; // arbiters per workgroup whose index and bodies stay in LDS (colours not kept)
constexpr int GRID_RES_COLOURS = 64;   // colours of the first column's units that may be kept in LDS
constexpr int GRID_LDS_BODIES = 1536;   // strips of at most this many bodies are coloured with claims in LDS

// The part of an arbiter the sweeps read and write (the names of Arbiter's
// fields, so the same impulse code runs on it), kept in LDS by the grid solve
// for the colours it keeps of a workgroup's first column.  ia / ib: the LDS
// slot of a body of the workgroup's strip kept in LDS, else ~(its global
// index); flags: bit 0 / 1 when body a / b is dynamic.
struct ContactSolve {
	real r1x, r1y, r2x, r2y;
	real nMass, tMass, bounce, jnAcc, jtAcc, jBias, bias;
};
struct ArbSolve {
	ContactSolve c[2];
	real nx, ny, u, svx, svy;
	int32_t count, state;
	int32_t ia, ib;
	int32_t flags;
};

// bodies by ArbSolve's encoding: an LDS slot, or ~(global index)
struct MixedBodies {
	BodyVel *lds;
	const DevWorld *W;
	__device__ __forceinline__ BV load(int e) const
	{
		if (e < 0) return load_bv(*W, ~e);
		const BodyVel &q = lds[e];
		BV r;
		r.v = v2(q.vx, q.vy); r.w = q.w; r.vb = v2(q.vbx, q.vby); r.wb = q.wb; r.m_inv = q.m_inv; r.i_inv = q.i_inv;
		return r;
	}
	__device__ __forceinline__ void store(int e, const BV &r) const
	{
		if (e < 0) { store_bv(*W, ~e, r); return; }
		BodyVel &q = lds[e];
		q.vx = r.v.x; q.vy = r.v.y; q.w = r.w; q.vbx = r.vb.x; q.vby = r.vb.y; q.wb = r.wb;
	}
};

// Dynamic LDS of k_solve_grid: first the colouring's claim arrays; after the
// colouring the same bytes hold the first column's strip bodies (BodyVel and
// global index by slot, GRID_RES_BODIES of them) and then ArbSolve records.
__host__ __device__ constexpr size_t grid_claim_bytes()
{
	return 2 * GRID_LDS_BODIES * sizeof(unsigned long long) + GRID_LDS_BODIES * sizeof(int32_t);
}
__host__ __device__ constexpr size_t grid_body_bytes()
{
	return ((GRID_RES_BODIES * (sizeof(BodyVel) + sizeof(int32_t)) + 15) / 16) * 16;
}

// The grid set of all large worlds on the co-resident grid.  Column c =
// l * parts + q holds strip q of the l-th large world and seam q to its
// right; workgroup g colours and sweeps columns g, g + gridDim.x, ... (the
// last workgroup also the wide set).  A sweep runs the strips, the even
// seams, the odd seams, the wide set, then the joints, as in cpSpaceStep.
// Without joints and wide set, units wait only for the units they share
// bodies with (a strip for its two seams' previous sweep, an even seam for
// its two strips, an odd seam for the even seams beside it) through
// completion counters; otherwise grid barriers separate the phases.
// The first column of each workgroup keeps what fits of its units in LDS
// through the sweeps: a strip whose arbiters all fit with its bodies keeps
// both (the bodies loaded before each of its sweeps and written back after
// it, for the seams); otherwise the smallest colours of the seam and the strip
// keep their arbiters.  Kept arbiters' accumulated impulses are written back
// after the last sweep.  Every value is read and written in the same order as
// from global memory.
__global__ void __launch_bounds__(GRID_THREADS) k_solve_grid(const DevWorld W, const SolveBuffers S, int lds_bytes)
{
	extern __shared__ __align__(16) unsigned char dyn[];
	constexpr int SCRATCH = 2 * GRID_LDS_COLS;   // LDS slot of units kept in global memory
	__shared__ int32_t lcol[SCRATCH + 1][CPGPU_MAX_COLOURS + 1];
	__shared__ int2 linfo[SCRATCH];
	__shared__ int32_t lcursor[CPGPU_MAX_COLOURS];
	__shared__ int32_t lleft;
	// the units' arbiters in colour order: index, and body indices (~b for a body that is not dynamic)
	__shared__ int32_t lpos[GRID_LDS_POS];
	__shared__ int2 lpb[GRID_LDS_POS];
	__shared__ int32_t lbase[SCRATCH];   // a unit's first entry in lpos / lpb, -1 when not kept there
	__shared__ int32_t liter[SCRATCH];   // iterations of the unit's space (a strip or seam lies in one space)
	__shared__ real ldt[SCRATCH];        // its cached-impulse scale
	// first column's strip (0) and seam (1): each colour's first ArbSolve record, -1 for a colour
	// whose arbiters stay in global memory
	__shared__ int32_t lres[2][GRID_RES_COLOURS];
	__shared__ int32_t lnb;              // first column's strip bodies kept in LDS (0: none)
	unsigned long long *lclaim = reinterpret_cast<unsigned long long *>(dyn);
	int32_t *ltaken = reinterpret_cast<int32_t *>(dyn + 2 * GRID_LDS_BODIES * sizeof(unsigned long long));
	BodyVel *lbody = reinterpret_cast<BodyVel *>(dyn);
	int32_t *lbidx = reinterpret_cast<int32_t *>(dyn + GRID_RES_BODIES * sizeof(BodyVel));
	int lds_used = 0;
	const int n_act = S.large_first[2];   // the solved arbiters' end (unsolved ones sort after)
	const int grid_first = S.large_first[0], wide_first = S.large_first[1];
	const int n_large = S.n_large_worlds;
	const int parts = S.parts;
	const int n_units = grid_units(n_large, parts), n_cols = n_large * parts;
	const int e0 = n_cols, o0 = e0 + n_large * even_seams(parts);
	const int stat = CPGPU_MAX_BUCKETS;
	const int g = blockIdx.x, G = gridDim.x;
	GridSync sy{(int)(blockIdx.x * blockDim.x + threadIdx.x), (int)(gridDim.x * blockDim.x), S.bar, W.err, 0u};
	const bool claims_fit = lds_bytes >= (int)grid_claim_bytes();

	// the seam unit of column c, -1 for a world's last strip
	auto seam_of = [&](int c) {
		const int l = c / parts, q = c % parts;
		if (q == parts - 1) return -1;
		return (q % 2 == 0 ? e0 + l * even_seams(parts) : o0 + l * odd_seams(parts)) + q / 2;
	};
	auto slot = [&](int i, int seam) { return i < GRID_LDS_COLS ? 2 * i + seam : SCRATCH; };

	// ---- every unit coloured by its workgroup ----
	auto colour_unit = [&](int u, int s) {
		const uint64_t key = unit_key(S, u);
		const int start = lower_bound_u64(S.keys, grid_first, n_act, key);
		const int end = u == n_units - 1 ? n_act : lower_bound_u64(S.keys, start, n_act, key + 1);
		int32_t *col = lcol[s];
		int nc = 0;
		if (end > start) {
			SolveSet P = {};
			P.list = S.arb_by_world + start;
			P.n = end - start;
			P.order = S.order + start;
			P.order_base = start;
			P.col_off = col;
			P.col_cursor = lcursor;
			P.left = &lleft;
			P.unit = stat;
			P.cset = unit_claims(u, n_large, parts);
			if (u < n_cols && S.pcount[u] <= GRID_LDS_BODIES && claims_fit) {
				P.lclaim = lclaim;
				P.ltaken = ltaken;
				P.lcap = GRID_LDS_BODIES;
				P.slot = S.bslot;
			}
			nc = colour_set(W, S, P);
			__syncthreads();
		}
		if (s == SCRATCH) {
			int32_t *pc = S.pcol + (size_t)u * (CPGPU_MAX_COLOURS + 1);
			for (int q = threadIdx.x; q <= nc; q += blockDim.x) pc[q] = col[q];
			if (threadIdx.x == 0) S.pinfo[u] = make_int2(start, nc);
		} else {
			const int n = end - start;
			const bool keep = lds_used + n <= GRID_LDS_POS;
			if (threadIdx.x == 0) {
				linfo[s] = make_int2(start, nc);
				lbase[s] = keep ? lds_used : -1;
				const int wd = n > 0 ? S.act[S.arb_by_world[start]].world : 0;
				liter[s] = W.ws[wd].iterations;
				ldt[s] = W.ws[wd].dt_coef;
			}
			if (keep) {
				for (int k = threadIdx.x; k < n; k += blockDim.x) {
					const int a = S.order[start + k];
					const Arbiter &arb = S.act[a];
					lpos[lds_used + k] = a;
					lpb[lds_used + k] = make_int2(is_dynamic(W, arb.ba) ? arb.ba : ~arb.ba,
					                              is_dynamic(W, arb.bb) ? arb.bb : ~arb.bb);
				}
				lds_used += n;
			}
		}
		if (threadIdx.x == 0) S.udone[u] = 0;
		__syncthreads();
	};
	for (int c = g, i = 0; c < n_cols; c += G, i++) {
		colour_unit(c, slot(i, 0));
		if (seam_of(c) >= 0) colour_unit(seam_of(c), slot(i, 1));
	}
	if (g == G - 1) colour_unit(n_units - 1, SCRATCH);

	// ---- the first column's units into LDS (the claim arrays are done with) ----
	// A strip whose bodies and arbiters all fit keeps both (its bodies are its own: no other
	// unit of its sweep touches them).  Otherwise colours are kept from the last, the smallest,
	// back: the seam's, then the strip's, while records fit; the bodies stay in global memory.
	const int c0 = g < n_cols ? g : -1;
	if (threadIdx.x == 0) {
		lnb = 0;
		for (int s = 0; s < 2; s++)
			for (int q = 0; q < GRID_RES_COLOURS; q++) lres[s][q] = -1;
		if (c0 >= 0) {
			const int nc0 = linfo[0].y, nc1 = seam_of(c0) >= 0 ? linfo[1].y : 0;
			const int n0 = nc0 > 0 ? lcol[0][nc0] : 0;
			const int nb = S.pcount[c0];
			const int rec = (int)sizeof(ArbSolve);
			const int cap_b = lds_bytes >= (int)grid_body_bytes() ? (lds_bytes - (int)grid_body_bytes()) / rec : 0;
			int used = 0, budget = lds_bytes / rec;
			if (S.sbody && lds_bytes >= (int)grid_body_bytes() && nb > 0 && nb <= GRID_RES_BODIES && n0 > 0 &&
			    nc0 <= GRID_RES_COLOURS && n0 <= cap_b) {
				lnb = nb;
				budget = cap_b;
				for (int q = 0; q < nc0; q++) lres[0][q] = lcol[0][q];
				used = n0;
			}
			for (int s = 1; s >= 0; s--) {
				if (s == 0 && lnb > 0) continue;
				for (int q = min(s ? nc1 : nc0, GRID_RES_COLOURS) - 1; q >= 0; q--) {
					const int sz = lcol[s][q + 1] - lcol[s][q];
					if (used + sz > budget) break;
					lres[s][q] = used;
					used += sz;
				}
			}
		}
	}
	__syncthreads();
	const int nbr = lnb;
	ArbSolve *larb = reinterpret_cast<ArbSolve *>(dyn + (nbr > 0 ? grid_body_bytes() : 0));
	for (int q = threadIdx.x; q < nbr; q += blockDim.x) lbidx[q] = S.sbody[(size_t)c0 * GRID_RES_BODIES + q];
	// f(record, arbiter index) for every kept arbiter
	auto each_kept = [&](auto f) {
		for (int s = 0; s < 2; s++)
			for (int q = 0; q < GRID_RES_COLOURS; q++) {
				if (lres[s][q] < 0) continue;
				const int k0 = lcol[s][q], k1 = lcol[s][q + 1], first = lres[s][q], start = linfo[s].x;
				for (int k = k0 + (int)threadIdx.x; k < k1; k += (int)blockDim.x)
					f(larb[first + k - k0], S.order[start + k], s);
			}
	};
	each_kept([&](ArbSolve &r, int a, int s) {
		const Arbiter &arb = S.act[a];
		for (int q = 0; q < 2; q++) {
			const Contact &c = arb.c[q];
			r.c[q].r1x = c.r1x; r.c[q].r1y = c.r1y; r.c[q].r2x = c.r2x; r.c[q].r2y = c.r2y;
			r.c[q].nMass = c.nMass; r.c[q].tMass = c.tMass; r.c[q].bounce = c.bounce;
			r.c[q].jnAcc = c.jnAcc; r.c[q].jtAcc = c.jtAcc; r.c[q].jBias = c.jBias; r.c[q].bias = c.bias;
		}
		r.nx = arb.nx; r.ny = arb.ny; r.u = arb.u; r.svx = arb.svx; r.svy = arb.svy;
		r.count = arb.count; r.state = arb.state;
		const bool da = is_dynamic(W, arb.ba), db = is_dynamic(W, arb.bb);
		// a strip arbiter's dynamic bodies are the strip's: in LDS when the strip's bodies are
		r.ia = (s == 0 && nbr > 0 && da) ? S.bslot[arb.ba] : ~arb.ba;
		r.ib = (s == 0 && nbr > 0 && db) ? S.bslot[arb.bb] : ~arb.bb;
		r.flags = (da ? 1 : 0) | (db ? 2 : 0);
	});
	sy.sync();

	// ---- joint pre-steps (cpSpaceStep.c:389-396) and, with joints, the velocity
	// integration of the grid set (cpSpaceStep.c:399-404; without joints it ran before) ----
	for (int jc = 0; jc < S.n_jcolours; jc++) {
		for (int k = S.jcol_off[jc] + sy.tid; k < S.jcol_off[jc + 1]; k += sy.stride)
			joint_prestep(W, W.joints[S.jorder[k]]);
		sy.sync();
	}
	if (S.n_large_joints > 0) {
		unsigned long long n_int = 0;
		for (int q = sy.tid; q < S.n_large_bodies; q += sy.stride) {
			integrate_velocity(W, S.large_bodies[q]);
			n_int++;
		}
		wave_add(&W.totals[3 + 2 * stat], n_int);
		sy.sync();
	}

	// ---- sweeps ----
	const GlobalBodies bodies{&W};
	const MixedBodies mixed{lbody, &W};
	// the first column's strip bodies into LDS before its sweep, and back after it
	auto strip_bodies_in = [&]() {
		for (int q = threadIdx.x; q < nbr; q += blockDim.x) lbody[q] = W.b_vel[lbidx[q]];
		__syncthreads();
	};
	auto strip_bodies_out = [&]() {
		__syncthreads();
		for (int q = threadIdx.x; q < nbr; q += blockDim.x) W.b_vel[lbidx[q]] = lbody[q];
	};
	// f(arbiter, its bodies as kept in lpb, the unit's LDS slot: SCRATCH for a unit kept in
	// global memory, whose arbiters' spaces are then read per arbiter) for the unit's
	// arbiters, colour by colour; fl(record, slot) for the arbiters of a colour kept in LDS
	auto sweep_unit = [&](int u, int s, auto f, auto fl) {
		int2 info;
		const int32_t *col;
		int base = -1;
		if (s < SCRATCH) {
			info = linfo[s];
			col = lcol[s];
			base = lbase[s];
		} else {
			info = S.pinfo[u];
			col = S.pcol + (size_t)u * (CPGPU_MAX_COLOURS + 1);
		}
		for (int q = 0; q < info.y; q++) {
			const int c1 = col[q + 1];
			const bool kept = s < 2 && q < GRID_RES_COLOURS && lres[s][q] >= 0;
			const int rofs = kept ? lres[s][q] - col[q] : 0;
			for (int k = col[q] + (int)threadIdx.x; k < c1; k += (int)blockDim.x) {
				if (kept) {
					fl(larb[rofs + k], s);
				} else if (base >= 0) {
					f(lpos[base + k], lpb[base + k], s);
				} else {
					const int a = S.order[info.x + k];
					const Arbiter &arb = S.act[a];
					f(a, make_int2(is_dynamic(W, arb.ba) ? arb.ba : ~arb.ba, is_dynamic(W, arb.bb) ? arb.bb : ~arb.bb),
					  s);
				}
			}
			__syncthreads();
		}
	};
	// a strip's sweep: its bodies through LDS when it is the first column's and they are kept there
	auto sweep_strip = [&](int c, int i, auto f, auto fl) {
		const bool in_lds = i == 0 && nbr > 0;
		if (in_lds) strip_bodies_in();
		sweep_unit(c, slot(i, 0), f, fl);
		if (in_lds) strip_bodies_out();
	};
	// units a and b (-1: none) have finished `need` passes
	auto wait_units = [&](int a, int b, int need) {
		if (threadIdx.x == 0) {
			if (a >= 0) unit_wait(&S.udone[a], need, W.err);
			if (b >= 0) unit_wait(&S.udone[b], need, W.err);
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
		}
		__syncthreads();
	};
	auto finish_unit = [&](int u, int passes) {
		__syncthreads();
		if (threadIdx.x == 0) {
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
			__hip_atomic_store(&S.udone[u], passes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
	};
	const bool wide = wide_first < n_act;
	const bool flow = S.grid_flow && !wide && S.n_jcolours == 0;
	// sweep p (0: cached impulses) of the arbiters with f / fl, then of the joints with fj
	auto sweep = [&](int p, auto f, auto fl, auto fj) {
		if (flow) {
			for (int c = g, i = 0; c < n_cols; c += G, i++) {
				if (p > 0) wait_units(c % parts > 0 ? seam_of(c - 1) : -1, seam_of(c), p);
				sweep_strip(c, i, f, fl);
				finish_unit(c, p + 1);
			}
			for (int parity = 0; parity < 2; parity++) {
				for (int c = g, i = 0; c < n_cols; c += G, i++) {
					const int su = seam_of(c);
					if (su < 0 || c % parts % 2 != parity) continue;
					// an even seam waits for its two strips; an odd seam for the even seams beside
					// it, which waited for its strips -- or, beside a space's last strip (an odd
					// strip count), for that strip itself
					if (parity == 0) wait_units(c, c + 1, p + 1);
					else wait_units(seam_of(c - 1), seam_of(c + 1) >= 0 ? seam_of(c + 1) : c + 1, p + 1);
					sweep_unit(su, slot(i, 1), f, fl);
					finish_unit(su, p + 1);
				}
			}
		} else {
			for (int c = g, i = 0; c < n_cols; c += G, i++) sweep_strip(c, i, f, fl);
			sy.sync();
			for (int parity = 0; parity < 2; parity++) {
				for (int c = g, i = 0; c < n_cols; c += G, i++)
					if (seam_of(c) >= 0 && c % parts % 2 == parity) sweep_unit(seam_of(c), slot(i, 1), f, fl);
				sy.sync();
			}
			if (wide) {
				if (g == G - 1) sweep_unit(n_units - 1, SCRATCH, f, fl);
				sy.sync();
			}
		}
		for (int jc = 0; jc < S.n_jcolours; jc++) {
			for (int k = S.jcol_off[jc] + sy.sweep_tid(); k < S.jcol_off[jc + 1]; k += sy.stride) fj(W.joints[S.jorder[k]]);
			sy.sync();
		}
	};
	// cached impulses (cpSpaceStep.c:407-415)
	sweep(0, [&](int a, int2 e, int s) {
		const Arbiter &arb = S.act[a];
		arb_apply_cached(bodies, arb, e.x >= 0 ? e.x : ~e.x, e.y >= 0 ? e.y : ~e.y, e.x >= 0, e.y >= 0,
		                 s < SCRATCH ? ldt[s] : W.ws[arb.world].dt_coef);
	}, [&](const ArbSolve &r, int s) {
		arb_apply_cached(mixed, r, r.ia, r.ib, r.flags & 1, r.flags & 2, ldt[s]);
	}, [&](Joint &j) { joint_apply_cached(W, bodies, j, W.ws[j.world].dt_coef); });
	// the impulse iterations (cpSpaceStep.c:418-427)
	for (int it = 0; it < S.max_iterations; it++) {
		sweep(1 + it, [&](int a, int2 e, int s) {
			Arbiter &arb = S.act[a];
			if (it < (s < SCRATCH ? liter[s] : W.ws[arb.world].iterations))
				arb_apply_impulse(W, bodies, arb, e.x >= 0 ? e.x : ~e.x, e.y >= 0 ? e.y : ~e.y, e.x >= 0, e.y >= 0);
		}, [&](ArbSolve &r, int s) {
			if (it < liter[s]) arb_apply_impulse(W, mixed, r, r.ia, r.ib, r.flags & 1, r.flags & 2);
		}, [&](Joint &j) {
			if (it < W.ws[j.world].iterations) joint_apply_impulse(W, bodies, j, W.ws[j.world].dt);
		});
	}

	// ---- the kept arbiters' accumulated impulses back to the arbiters ----
	__syncthreads();
	each_kept([&](const ArbSolve &r, int a, int) {
		Arbiter &arb = S.act[a];
		for (int q = 0; q < 2; q++) {
			arb.c[q].jnAcc = r.c[q].jnAcc; arb.c[q].jtAcc = r.c[q].jtAcc; arb.c[q].jBias = r.c[q].jBias;
		}
	});
}

// ---- strips of the grid set ----

// a double's bits as an unsigned key with the same order (x + 0.0: -0.0 becomes +0.0)
__device__ __forceinline__ unsigned long long x_key(double x)
{
	const unsigned long long u = (unsigned long long)__double_as_longlong(x + 0.0);
	return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double key_x(unsigned long long k)
{
	return __longlong_as_double((long long)((k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k));
}

// histogram bin of x in [lo, hi]
__device__ __forceinline__ int part_bin(double x, double lo, double hi)
{
	if (!(hi > lo)) return 0;
	const double t = (x - lo) * ((double)PART_BINS / (hi - lo));
	if (!(t >= 0.0)) return 0;
	return t >= (double)PART_BINS ? PART_BINS - 1 : (int)t;
}

// the grid set: the dynamic bodies of the large worlds
__device__ __forceinline__ bool in_grid_set(const DevWorld &W, const SolveBuffers &S, int b)
{
	return is_dynamic(W, b);
}

// ---- grouping the solved arbiters: strips and keys ----

constexpr int GROUP_BLOCKS = 1024;

// clear the strips' bounds, histograms and body counts
__global__ void k_part_init(const SolveBuffers S)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	const int n_large = S.n_large_worlds;
	for (int i = q; i < n_large * PART_BINS; i += gridDim.x * blockDim.x) {
		S.phist[i] = 0;
		if (i < n_large * S.parts) S.pcount[i] = 0;
		if (i < n_large) { S.pbound[2 * i] = ~0ull; S.pbound[2 * i + 1] = 0ull; }
	}
}

// The grid set's least and greatest x per large world: waves reduce their
// lanes, the workgroup its waves in LDS (its bodies lie in at most two large
// worlds, as in k_part_assign), then one atomic pair per world.
constexpr int PART_BOUNDS_THREADS = 1024;
__global__ void __launch_bounds__(PART_BOUNDS_THREADS) k_part_bounds(const DevWorld W, const SolveBuffers S)
{
	__shared__ unsigned long long lmin[2], lmax[2];
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	const int l0 = S.large_index[W.b_world[S.large_bodies[blockIdx.x * blockDim.x]]];
	if (threadIdx.x < 2) { lmin[threadIdx.x] = ~0ull; lmax[threadIdx.x] = 0ull; }
	__syncthreads();
	int rel = -1;
	unsigned long long k = 0;
	if (q < S.n_large_bodies) {
		const int b = S.large_bodies[q];
		if (in_grid_set(W, S, b)) {
			rel = S.large_index[W.b_world[b]] - l0;
			k = x_key((double)W.b_p[b].x);
		}
	}
	for (int r = 0; r < 2; r++) {
		unsigned long long mn = rel == r ? k : ~0ull, mx = rel == r ? k : 0ull;
		for (int off = 32; off > 0; off >>= 1) {
			const unsigned long long a = __shfl_xor(mn, off), b = __shfl_xor(mx, off);
			mn = a < mn ? a : mn;
			mx = b > mx ? b : mx;
		}
		if (__lane_id() == 0 && mn <= mx) {
			atomicMin(&lmin[r], mn);
			atomicMax(&lmax[r], mx);
		}
	}
	__syncthreads();
	if (threadIdx.x < 2 && lmin[threadIdx.x] <= lmax[threadIdx.x]) {
		atomicMin(&S.pbound[2 * (l0 + threadIdx.x)], lmin[threadIdx.x]);
		atomicMax(&S.pbound[2 * (l0 + threadIdx.x) + 1], lmax[threadIdx.x]);
	}
}

// histogram of the grid set's x per large world
__global__ void k_part_hist(const DevWorld W, const SolveBuffers S)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= S.n_large_bodies) return;
	const int b = S.large_bodies[q];
	if (!in_grid_set(W, S, b)) return;
	const int l = S.large_index[W.b_world[b]];
	const int bin = part_bin((double)W.b_p[b].x, key_x(S.pbound[2 * l]), key_x(S.pbound[2 * l + 1]));
	atomicAdd(&S.phist[l * PART_BINS + bin], 1);
}

// one workgroup (PART_BINS threads) per large world: each bin's strip is
// `parts` times the fraction of the set's bodies in the bins before it
__global__ void __launch_bounds__(PART_BINS) k_part_scan(const SolveBuffers S)
{
	__shared__ int32_t sc[PART_BINS];
	int32_t *h = S.phist + blockIdx.x * PART_BINS;
	const int i = threadIdx.x;
	const int own = h[i];
	sc[i] = own;
	__syncthreads();
	for (int off = 1; off < PART_BINS; off <<= 1) {
		const int add = i >= off ? sc[i - off] : 0;
		__syncthreads();
		sc[i] += add;
		__syncthreads();
	}
	const long long total = sc[PART_BINS - 1], before = sc[i] - own;
	h[i] = total > 0 ? (int)(before * S.parts / total) : 0;
}

// each grid-set body's strip, and its slot among the strip's bodies: the
// workgroup counts its bodies per strip in LDS and reserves each strip's
// slots with one atomic.  The large bodies are in body order and a large
// world has more than a workgroup's bodies, so a workgroup's bodies lie in
// at most two large worlds.
constexpr int PART_ASSIGN_THREADS = 1024;
__global__ void __launch_bounds__(PART_ASSIGN_THREADS) k_part_assign(const DevWorld W, const SolveBuffers S)
{
	__shared__ int32_t cnt[2][GRID_PARTS_MAX], base[2][GRID_PARTS_MAX];
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	const int parts = S.parts;
	for (int i = threadIdx.x; i < 2 * parts; i += blockDim.x) cnt[i / parts][i % parts] = 0;
	const int l0 = S.large_index[W.b_world[S.large_bodies[blockIdx.x * blockDim.x]]];
	__syncthreads();
	int b = -1, rel = -1, strip = 0, local = 0;
	if (q < S.n_large_bodies) {
		b = S.large_bodies[q];
		if (in_grid_set(W, S, b)) {
			const int l = S.large_index[W.b_world[b]];
			const int bin = part_bin((double)W.b_p[b].x, key_x(S.pbound[2 * l]), key_x(S.pbound[2 * l + 1]));
			strip = S.phist[l * PART_BINS + bin];
			S.bpart[b] = strip;
			rel = l - l0;
			local = atomicAdd(&cnt[rel][strip], 1);
		}
	}
	__syncthreads();
	for (int i = threadIdx.x; i < 2 * parts; i += blockDim.x) {
		const int r = i / parts, st = i % parts;
		if (cnt[r][st] > 0) base[r][st] = atomicAdd(&S.pcount[(l0 + r) * parts + st], cnt[r][st]);
	}
	__syncthreads();
	if (rel >= 0) {
		const int sl = base[rel][strip] + local;
		S.bslot[b] = sl;
		if (sl < GRID_RES_BODIES) S.sbody[((size_t)(l0 + rel) * parts + strip) * GRID_RES_BODIES + sl] = b;
	}
}

// the key OR and AND of the world sort start at 0 and all ones
__global__ void k_orand_reset(unsigned long long *orand)
{
	orand[0] = 0ull;
	orand[1] = ~0ull;
}

// Sort keys of the solved arbiters; the pre-steps of the large worlds'
// arbiters (they only read velocities, before the joints' pre-steps and the
// velocity integration, as in cpSpaceStep); OR / AND of the keys for the
// radix sort.
__global__ void __launch_bounds__(256) k_arb_keys(const DevWorld W, const SolveBuffers S, unsigned long long *orand)
{
	__shared__ unsigned long long red[2][4];
	const int n = *S.n_act;
	unsigned long long o = 0ull, a = ~0ull;
	for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
		Arbiter &arb = S.act[k];
		const int w = arb.world;
		unsigned long long key;
		if (arb.count == 0) {
			// touching but not solved (a collision handler's or a sensor's): after every set
			key = excluded_key(S);
		} else if (S.world_small[w]) {
			key = (unsigned long long)w;
		} else {
			// the strip of its dynamic bodies, the seam of two neighbouring strips, or the wide set
			const int pa = is_dynamic(W, arb.ba) ? S.bpart[arb.ba] : -1;
			const int pb = is_dynamic(W, arb.bb) ? S.bpart[arb.bb] : -1;
			key = unit_key(S, unit_of(S.large_index[w], pa, pb, S.n_large_worlds, S.parts));
			arb_prestep(W, arb);
		}
		S.keys[k] = key;
		S.arb_by_world[k] = k;
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

// velocity integration of the large worlds' bodies when they have no joints
// (after the pre-steps above, before any impulse; with joints the grid solve
// integrates them after the joints' pre-steps)
__global__ void k_large_integrate(const DevWorld W, const SolveBuffers S)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= S.n_large_bodies) return;
	integrate_velocity(W, S.large_bodies[q]);
}

// warb_off[w] = first sorted position of small world w (n_worlds + 1 entries);
// large_first = first position of the grid set, of the wide set, and the end of
// the solved arbiters (the unsolved ones sort after them)
__global__ void k_world_offsets(const SolveBuffers S, int n_worlds)
{
	const int n = *S.n_act;
	const int w = blockIdx.x * blockDim.x + threadIdx.x;
	if (w <= n_worlds) S.warb_off[w] = lower_bound_u64(S.keys, 0, n, (uint64_t)w);
	if (w == n_worlds + 1) {
		S.large_first[0] = lower_bound_u64(S.keys, 0, n, unit_key(S, 0));
		S.large_first[1] = lower_bound_u64(S.keys, 0, n, unit_key(S, grid_units(S.n_large_worlds, S.parts) - 1));
		S.large_first[2] = lower_bound_u64(S.keys, 0, n, excluded_key(S));
	}
}

} // namespace

// dynamic LDS of k_solve_grid: what a workgroup may hold beside the kernel's static LDS
// (asked once: the library runs on one kind of device)
static int grid_lds_bytes()
{
	static int cached = -1;
	if (cached >= 0) return cached;
	int dev = 0, lds = 0;
	(void)hipGetDevice(&dev);
	(void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
	hipFuncAttributes attr;
	if (hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(k_solve_grid)) != hipSuccess) return cached = 0;
	const int dyn = lds - (int)attr.sharedSizeBytes;
	return cached = dyn > 0 ? dyn & ~15 : 0;
}

int solver_max_grid()
{
	int dev = 0, per_cu = 0;
	hipDeviceProp_t prop;
	(void)hipGetDevice(&dev);
	(void)hipGetDeviceProperties(&prop, dev);
	(void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_solve_grid),
	                                                   GRID_THREADS, (size_t)grid_lds_bytes());
	if (per_cu < 1) per_cu = 1;
	return per_cu * prop.multiProcessorCount;
}

hipError_t solver_strips(const DevWorld &W, const SolveBuffers &S, hipStream_t stream, KernelTimer *timer)
{
	const int nb = S.n_large_bodies;
	if (S.n_large_worlds == 0 || nb == 0) return hipSuccess;
	TIMED(timer, "strips", stream, {
		k_part_init<<<std::max(1, std::min((S.n_large_worlds * PART_BINS + 255) / 256, GROUP_BLOCKS)), 256, 0, stream>>>(
		    S);
		k_part_bounds<<<(nb + PART_BOUNDS_THREADS - 1) / PART_BOUNDS_THREADS, PART_BOUNDS_THREADS, 0, stream>>>(W, S);
		k_part_hist<<<(nb + 255) / 256, 256, 0, stream>>>(W, S);
		k_part_scan<<<S.n_large_worlds, PART_BINS, 0, stream>>>(S);
		k_part_assign<<<(nb + PART_ASSIGN_THREADS - 1) / PART_ASSIGN_THREADS, PART_ASSIGN_THREADS, 0, stream>>>(W, S);
	});
	return hipGetLastError();
}

hipError_t solver_group(const DevWorld &W, const SolveBuffers &S, int n_act, int n_worlds, bool large,
                        RadixSortBuffers &sortbuf, hipStream_t stream, KernelTimer *timer)
{
	// n_act: an upper bound of *S.n_act (the kernels read the count)
	const int nb = S.n_large_bodies;
	if (n_act > 0) {
		HIP_TRY(radix_sort_reserve(sortbuf, n_act));
		k_orand_reset<<<1, 1, 0, stream>>>((unsigned long long *)sortbuf.orand);
		TIMED(timer, "world_sort", stream,
		      k_arb_keys<<<std::min((n_act + 255) / 256, GROUP_BLOCKS), 256, 0, stream>>>(
		          W, S, (unsigned long long *)sortbuf.orand));
		HIP_TRY(hipGetLastError());
		// keys up to n_worlds + grid_units (the unsolved arbiters' key)
		const uint64_t key_end = (uint64_t)n_worlds + (uint64_t)grid_units(S.n_large_worlds, S.parts);
		int key_bits = 1;
		while ((1ull << key_bits) <= key_end) key_bits++;
		HIP_TRY(radix_sort_pairs(S.keys, S.arb_by_world, n_act, S.n_act, key_bits, sortbuf, stream, timer));
	}
	if (large && nb > 0 && S.n_large_joints == 0) {
		TIMED(timer, "large_integrate", stream,
		      k_large_integrate<<<(nb + 255) / 256, 256, 0, stream>>>(W, S));
	}
	k_world_offsets<<<(n_worlds + 2 + 255) / 256, 256, 0, stream>>>(S, n_worlds);
	return hipGetLastError();
}

template <int THREADS>
static void launch_worlds(const DevWorld &W, const SolveBuffers &S, const int32_t *worlds, int blocks, int lds_cap,
                          int unit, hipStream_t st)
{
	k_solve_worlds<THREADS><<<blocks, THREADS, (size_t)lds_cap * sizeof(BodyVel), st>>>(W, S, worlds, lds_cap, unit);
}

static const char *bucket_names[CPGPU_MAX_BUCKETS] = {"solver_worlds_b0", "solver_worlds_b1", "solver_worlds_b2",
                                                      "solver_worlds_b3"};

int solver_max_lds_bytes()
{
	int dev = 0, lds = 0;
	(void)hipGetDevice(&dev);
	(void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
	// the static LDS of k_solve_worlds: colour offsets and cursors
	return lds - (int)(sizeof(int32_t) * (2 * CPGPU_MAX_COLOURS + 4));
}

hipError_t solver_step(const DevWorld &W, const SolveBuffers &S, const WorldBucket *buckets, int n_buckets,
                       int grid, hipStream_t stream, const hipStream_t *aux,
                       const hipEvent_t *events, KernelTimer *timer)
{
	// fork: the buckets' streams start after everything queued so far; the
	// buckets of the largest spaces (the longest workgroups) are queued first
	HIP_TRY(hipEventRecord(events[0], stream));
	for (int q = n_buckets - 1; q >= 0; q--) {
		const WorldBucket &B = buckets[q];
		if (B.n == 0) continue;
		HIP_TRY(hipStreamWaitEvent(aux[q], events[0], 0));
		TIMED(timer, bucket_names[q], aux[q], {
			switch (B.threads) {
			case 64: launch_worlds<64>(W, S, B.worlds, B.n, B.lds_cap, q, aux[q]); break;
			case 128: launch_worlds<128>(W, S, B.worlds, B.n, B.lds_cap, q, aux[q]); break;
			case 256: launch_worlds<256>(W, S, B.worlds, B.n, B.lds_cap, q, aux[q]); break;
			case 512: launch_worlds<512>(W, S, B.worlds, B.n, B.lds_cap, q, aux[q]); break;
			default: launch_worlds<1024>(W, S, B.worlds, B.n, B.lds_cap, q, aux[q]); break;
			}
		});
		HIP_TRY(hipGetLastError());
	}
	if (grid > 0) {
		int lds_bytes = grid_lds_bytes();
		void *args[] = {(void *)&W, (void *)&S, (void *)&lds_bytes};
		hipError_t e = hipSuccess;
		for (int attempt = 0; attempt < 2; attempt++) {
			k_bar_reset<<<1, 1, 0, stream>>>(S.bar);
			TIMED(timer, "solver", stream,
			      e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_solve_grid), dim3(grid),
			                                     dim3(GRID_THREADS), args, (size_t)lds_bytes, stream));
			if (e == hipSuccess || lds_bytes == 0) break;
			// a device that refuses the dynamic LDS: the same solve, claims and arbiters in global memory
			(void)hipGetLastError();
			lds_bytes = 0;
		}
		HIP_TRY(e);
	}
	// join
	for (int q = 0; q < n_buckets; q++) {
		if (buckets[q].n == 0) continue;
		HIP_TRY(hipEventRecord(events[1 + q], aux[q]));
		HIP_TRY(hipStreamWaitEvent(stream, events[1 + q], 0));
	}
	return hipSuccess;
}

} // namespace CPGPU_NS
