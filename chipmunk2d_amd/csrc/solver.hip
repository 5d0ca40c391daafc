// solver.hip -- the velocity solver of cpSpaceStep as one persistent kernel.
//
// Reference: the second half of cpSpaceStep (src/cpSpaceStep.c:380-427):
// arbiter pre-step, constraint pre-step, velocity integration, cached
// impulses, then `iterations` sweeps of cpArbiterApplyImpulse and the joints'
// applyImpulse.  The reference sweeps its arbiter list sequentially
// (Gauss-Seidel).  Here the arbiters of one step are coloured so that no two
// arbiters of a colour share a dynamic body; a sweep runs colour after colour
// with a grid barrier between colours, so the result is exactly the
// sequential sweep in colour order (arbiters of one colour touch disjoint
// bodies, so their order inside the colour does not matter).  Static and
// kinematic bodies have zero inverse mass, are never written, and do not
// constrain the colouring.
//
// The colouring is deterministic (independent of thread timing): colour c is
// built in COLOUR_SUBROUNDS rounds in which every uncoloured arbiter whose
// dynamic bodies are not yet taken by colour c claims them with an atomic min
// of its priority (a bijective hash of its shape-pair key); arbiters holding
// all their claims take colour c.
#include "cpgpu_step.h"

namespace {

constexpr int SOLVER_THREADS = 256;

// ---- grid barrier (all blocks co-resident: cooperative launch) ----
__device__ void grid_sync(GridBarrier *bar, int32_t *err)
{
	__syncthreads();
	if (threadIdx.x == 0) {
		const unsigned int nblocks = gridDim.x;
		const unsigned int g = __hip_atomic_load(&bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
		__threadfence();
		const unsigned int arrived = __hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		if (arrived == nblocks - 1) {
			__hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			__hip_atomic_fetch_add(&bar->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
		} else {
			long spins = 0;
			while (__hip_atomic_load(&bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
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
		__threadfence();
	}
	__syncthreads();
}

__device__ __forceinline__ bool is_dynamic(const DevWorld &W, int b) { return W.b_type[b] == CPGPU_BODY_DYNAMIC; }

// ---- body helpers (chipmunk_private.h:172-226) ----
struct BV {
	V2 v, vb;
	double w, wb, m_inv, i_inv;
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

__device__ __forceinline__ double k_scalar_body(const BV &b, V2 r, V2 n)
{
	const double rcn = vcross(r, n);
	return b.m_inv + b.i_inv * rcn * rcn;
}

__device__ __forceinline__ double k_scalar(const BV &a, const BV &b, V2 r1, V2 r2, V2 n)
{
	return k_scalar_body(a, r1, n) + k_scalar_body(b, r2, n);
}

// ---- arbiters: cpArbiterPreStep / ApplyCachedImpulse / ApplyImpulse (cpArbiter.c:416-498) ----

__device__ void arb_prestep(const DevWorld &W, Arbiter &arb)
{
	const WorldStep ws = W.ws[arb.world];
	const BV a = load_bv(W, arb.ba), b = load_bv(W, arb.bb);
	const double2 pa = W.b_p[arb.ba], pb = W.b_p[arb.bb];
	const V2 n = v2(arb.nx, arb.ny);
	const V2 body_delta = v2(pb.x - pa.x, pb.y - pa.y);
	for (int i = 0; i < arb.count; i++) {
		Contact &con = arb.c[i];
		const V2 r1 = v2(con.r1x, con.r1y), r2 = v2(con.r2x, con.r2y);
		con.nMass = 1.0 / k_scalar(a, b, r1, r2, n);
		con.tMass = 1.0 / k_scalar(a, b, r1, r2, vperp(n));
		const double dist = vdot(vadd(vsub(r2, r1), body_delta), n);
		con.bias = -ws.bias_coef * fminc(0.0, dist + ws.slop) / ws.dt;
		con.jBias = 0.0;
		con.bounce = vdot(rel_vel(a, b, r1, r2), n) * arb.e;
	}
}

__device__ void arb_apply_cached(const DevWorld &W, const Arbiter &arb)
{
	if (arb.state == CPGPU_ARB_FIRST_COLLISION) return;
	const double dt_coef = W.ws[arb.world].dt_coef;
	BV a = load_bv(W, arb.ba), b = load_bv(W, arb.bb);
	const V2 n = v2(arb.nx, arb.ny);
	for (int i = 0; i < arb.count; i++) {
		const Contact &con = arb.c[i];
		const V2 j = vrotate(n, v2(con.jnAcc, con.jtAcc));
		apply_impulses(a, b, v2(con.r1x, con.r1y), v2(con.r2x, con.r2y), vmult(j, dt_coef));
	}
	if (is_dynamic(W, arb.ba)) store_bv(W, arb.ba, a);
	if (is_dynamic(W, arb.bb)) store_bv(W, arb.bb, b);
}

__device__ void arb_apply_impulse(const DevWorld &W, Arbiter &arb)
{
	BV a = load_bv(W, arb.ba), b = load_bv(W, arb.bb);
	const V2 n = v2(arb.nx, arb.ny);
	const V2 surface_vr = v2(arb.svx, arb.svy);
	const double friction = arb.u;
	for (int i = 0; i < arb.count; i++) {
		Contact &con = arb.c[i];
		const double nMass = con.nMass;
		const V2 r1 = v2(con.r1x, con.r1y), r2 = v2(con.r2x, con.r2y);
		const V2 vb1 = vadd(a.vb, vmult(vperp(r1), a.wb));
		const V2 vb2 = vadd(b.vb, vmult(vperp(r2), b.wb));
		const V2 vr = vadd(rel_vel(a, b, r1, r2), surface_vr);
		const double vbn = vdot(vsub(vb2, vb1), n);
		const double vrn = vdot(vr, n);
		const double vrt = vdot(vr, vperp(n));
		const double jbn = (con.bias - vbn) * nMass;
		const double jbnOld = con.jBias;
		con.jBias = fmaxc(jbnOld + jbn, 0.0);
		const double jn = -(con.bounce + vrn) * nMass;
		const double jnOld = con.jnAcc;
		con.jnAcc = fmaxc(jnOld + jn, 0.0);
		const double jtMax = friction * con.jnAcc;
		const double jt = -vrt * con.tMass;
		const double jtOld = con.jtAcc;
		con.jtAcc = fclampc(jtOld + jt, -jtMax, jtMax);
		const V2 jb = vmult(n, con.jBias - jbnOld);
		apply_bias_impulse(a, vneg(jb), r1);
		apply_bias_impulse(b, jb, r2);
		apply_impulses(a, b, r1, r2, vrotate(n, v2(con.jnAcc - jnOld, con.jtAcc - jtOld)));
	}
	if (is_dynamic(W, arb.ba)) store_bv(W, arb.ba, a);
	if (is_dynamic(W, arb.bb)) store_bv(W, arb.bb, b);
}

// ---- joints (cpPivotJoint.c:24-70, cpDampedSpring.c:24-77, cpSlideJoint.c:24-89, cpGearJoint.c:24-69) ----

__device__ void joint_prestep(const DevWorld &W, Joint &j)
{
	const double dt = W.ws[j.world].dt;
	BV a = load_bv(W, j.a), b = load_bv(W, j.b);
	const XF Ta = W.b_T[j.a], Tb = W.b_T[j.b];
	const double2 cogA = W.b_cog[j.a], cogB = W.b_cog[j.b];
	const double2 pA = W.b_p[j.a], pB = W.b_p[j.b];
	const V2 anchorA = v2(j.P[0], j.P[1]), anchorB = v2(j.P[2], j.P[3]);
	const double bc = j.bias_coef;
	if (j.type == CPGPU_JOINT_PIVOT) {
		const V2 r1 = xvect(Ta, vsub(anchorA, v2(cogA.x, cogA.y)));
		const V2 r2 = xvect(Tb, vsub(anchorB, v2(cogB.x, cogB.y)));
		j.r1x = r1.x; j.r1y = r1.y; j.r2x = r2.x; j.r2y = r2.y;
		const double m_sum = a.m_inv + b.m_inv;
		double k11 = m_sum, k12 = 0.0, k21 = 0.0, k22 = m_sum;
		const double a_i_inv = a.i_inv;
		const double r1xsq = r1.x * r1.x * a_i_inv;
		const double r1ysq = r1.y * r1.y * a_i_inv;
		const double r1nxy = -r1.x * r1.y * a_i_inv;
		k11 += r1ysq; k12 += r1nxy; k21 += r1nxy; k22 += r1xsq;
		const double b_i_inv = b.i_inv;
		const double r2xsq = r2.x * r2.x * b_i_inv;
		const double r2ysq = r2.y * r2.y * b_i_inv;
		const double r2nxy = -r2.x * r2.y * b_i_inv;
		k11 += r2ysq; k12 += r2nxy; k21 += r2nxy; k22 += r2xsq;
		const double det = k11 * k22 - k12 * k21;
		const double det_inv = 1.0 / det;
		j.k11 = k22 * det_inv; j.k12 = -k12 * det_inv;
		j.k21 = -k21 * det_inv; j.k22 = k11 * det_inv;
		const V2 delta = vsub(vadd(v2(pB.x, pB.y), r2), vadd(v2(pA.x, pA.y), r1));
		const V2 bias = vclamp(vmult(delta, -bc / dt), j.maxBias);
		j.bx = bias.x; j.by = bias.y;
	} else if (j.type == CPGPU_JOINT_DAMPED_SPRING) {
		const double restLength = j.P[4], stiffness = j.P[5], damping = j.P[6];
		const V2 r1 = xvect(Ta, vsub(anchorA, v2(cogA.x, cogA.y)));
		const V2 r2 = xvect(Tb, vsub(anchorB, v2(cogB.x, cogB.y)));
		j.r1x = r1.x; j.r1y = r1.y; j.r2x = r2.x; j.r2y = r2.y;
		const V2 delta = vsub(vadd(v2(pB.x, pB.y), r2), vadd(v2(pA.x, pA.y), r1));
		const double dist = vlength(delta);
		const V2 n = vmult(delta, 1.0 / (dist != 0.0 ? dist : __builtin_inf()));
		j.nx = n.x; j.ny = n.y;
		const double k = k_scalar(a, b, r1, r2, n);
		j.nMass = 1.0 / k;
		j.target_vrn = 0.0;
		j.v_coef = 1.0 - cp_exp(-damping * dt * k);
		const double f_spring = (restLength - dist) * stiffness;
		const double j_spring = j.jAcc = f_spring * dt;
		apply_impulses(a, b, r1, r2, vmult(n, j_spring));
		if (is_dynamic(W, j.a)) store_bv(W, j.a, a);
		if (is_dynamic(W, j.b)) store_bv(W, j.b, b);
	} else if (j.type == CPGPU_JOINT_SLIDE) {
		const double mn = j.P[4], mx = j.P[5];
		const V2 r1 = xvect(Ta, vsub(anchorA, v2(cogA.x, cogA.y)));
		const V2 r2 = xvect(Tb, vsub(anchorB, v2(cogB.x, cogB.y)));
		j.r1x = r1.x; j.r1y = r1.y; j.r2x = r2.x; j.r2y = r2.y;
		const V2 delta = vsub(vadd(v2(pB.x, pB.y), r2), vadd(v2(pA.x, pA.y), r1));
		const double dist = vlength(delta);
		double pdist = 0.0;
		V2 n;
		if (dist > mx) { pdist = dist - mx; n = vnormalize(delta); }
		else if (dist < mn) { pdist = mn - dist; n = vneg(vnormalize(delta)); }
		else { n = v2(0.0, 0.0); j.jnAcc = 0.0; }
		j.nx = n.x; j.ny = n.y;
		j.nMass = 1.0 / k_scalar(a, b, r1, r2, n);
		const double maxBias = j.maxBias;
		j.bias = fclampc(-bc * pdist / dt, -maxBias, maxBias);
	} else {
		const double phase = j.P[0], ratio = j.P[1], ratio_inv = j.P[2];
		j.iSum = 1.0 / (a.i_inv * ratio_inv + ratio * b.i_inv);
		const double maxBias = j.maxBias;
		j.bias = fclampc(-bc * (W.b_a[j.b] * ratio - W.b_a[j.a] - phase) / dt, -maxBias, maxBias);
	}
}

__device__ void joint_apply_cached(const DevWorld &W, const Joint &j)
{
	if (j.type == CPGPU_JOINT_DAMPED_SPRING) return;
	const double dt_coef = W.ws[j.world].dt_coef;
	BV a = load_bv(W, j.a), b = load_bv(W, j.b);
	const V2 r1 = v2(j.r1x, j.r1y), r2 = v2(j.r2x, j.r2y);
	if (j.type == CPGPU_JOINT_PIVOT) {
		apply_impulses(a, b, r1, r2, vmult(v2(j.jAx, j.jAy), dt_coef));
	} else if (j.type == CPGPU_JOINT_SLIDE) {
		apply_impulses(a, b, r1, r2, vmult(v2(j.nx, j.ny), j.jnAcc * dt_coef));
	} else {
		const double ratio_inv = j.P[2];
		const double jj = j.jAcc * dt_coef;
		a.w -= jj * a.i_inv * ratio_inv;
		b.w += jj * b.i_inv;
	}
	if (is_dynamic(W, j.a)) store_bv(W, j.a, a);
	if (is_dynamic(W, j.b)) store_bv(W, j.b, b);
}

__device__ void joint_apply_impulse(const DevWorld &W, Joint &j)
{
	const double dt = W.ws[j.world].dt;
	BV a = load_bv(W, j.a), b = load_bv(W, j.b);
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
		const double vrn = vdot(rel_vel(a, b, r1, r2), n);
		const double v_damp = (j.target_vrn - vrn) * j.v_coef;
		j.target_vrn = vrn + v_damp;
		const double j_damp = v_damp * j.nMass;
		j.jAcc += j_damp;
		apply_impulses(a, b, r1, r2, vmult(n, j_damp));
	} else if (j.type == CPGPU_JOINT_SLIDE) {
		const V2 n = v2(j.nx, j.ny);
		if (veql(n, v2(0.0, 0.0))) return;
		const V2 vr = rel_vel(a, b, r1, r2);
		const double vrn = vdot(vr, n);
		double jn = (j.bias - vrn) * j.nMass;
		const double jnOld = j.jnAcc;
		j.jnAcc = fclampc(jnOld + jn, -j.maxForce * dt, 0.0);
		jn = j.jnAcc - jnOld;
		apply_impulses(a, b, r1, r2, vmult(n, jn));
	} else {
		const double ratio = j.P[1], ratio_inv = j.P[2];
		const double wr = b.w * ratio - a.w;
		const double jMax = j.maxForce * dt;
		double jj = (j.bias - wr) * j.iSum;
		const double jOld = j.jAcc;
		j.jAcc = fclampc(jOld + jj, -jMax, jMax);
		jj = j.jAcc - jOld;
		a.w -= jj * a.i_inv * ratio_inv;
		b.w += jj * b.i_inv;
	}
	if (is_dynamic(W, j.a)) store_bv(W, j.a, a);
	if (is_dynamic(W, j.b)) store_bv(W, j.b, b);
}

// ---- the persistent kernel ----

__device__ __forceinline__ int dyn_or_neg(const DevWorld &W, int b) { return is_dynamic(W, b) ? b : -1; }

__global__ void __launch_bounds__(SOLVER_THREADS) k_solve(const DevWorld W, const SolveBuffers S)
{
	const int tid = blockIdx.x * blockDim.x + threadIdx.x;
	const int stride = gridDim.x * blockDim.x;
	const int n = *S.n_act;
	unsigned long long *claim0 = S.claim, *claim1 = S.claim + S.n_bodies;

	// ---- colouring ----
	for (int k = tid; k < n; k += stride) {
		S.colour[k] = -1;
		const int ba = dyn_or_neg(W, S.act[k].ba), bb = dyn_or_neg(W, S.act[k].bb);
		if (ba >= 0) { claim0[ba] = ~0ull; claim1[ba] = ~0ull; S.taken[ba] = -1; }
		if (bb >= 0) { claim0[bb] = ~0ull; claim1[bb] = ~0ull; S.taken[bb] = -1; }
	}
	for (int k = tid; k <= CPGPU_MAX_COLOURS; k += stride) S.col_off[k] = 0;
	if (tid == 0) *S.left = n;
	grid_sync(S.bar, W.err);

	int c = 0, sub = 0;
	while (true) {
		if (__hip_atomic_load(S.left, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
		if (c >= CPGPU_MAX_COLOURS) { if (tid == 0) raise_err(W.err, CPGPU_ERR_ARBITERS_OVERFLOW); break; }
		unsigned long long *cl = (sub & 1) ? claim1 : claim0;
		unsigned long long *other = (sub & 1) ? claim0 : claim1;
		for (int k = tid; k < n; k += stride) {
			if (S.colour[k] >= 0) continue;
			const Arbiter &arb = S.act[k];
			const int ba = dyn_or_neg(W, arb.ba), bb = dyn_or_neg(W, arb.bb);
			if ((ba >= 0 && S.taken[ba] == c) || (bb >= 0 && S.taken[bb] == c)) continue;
			const unsigned long long prio = mix64(arb.key);
			if (ba >= 0) atomicMin(&cl[ba], prio);
			if (bb >= 0) atomicMin(&cl[bb], prio);
		}
		grid_sync(S.bar, W.err);
		int won = 0;
		for (int k = tid; k < n; k += stride) {
			if (S.colour[k] >= 0) continue;
			const Arbiter &arb = S.act[k];
			const int ba = dyn_or_neg(W, arb.ba), bb = dyn_or_neg(W, arb.bb);
			const unsigned long long prio = mix64(arb.key);
			const bool win = (ba < 0 || __hip_atomic_load(&cl[ba], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == prio) &&
			                 (bb < 0 || __hip_atomic_load(&cl[bb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == prio);
			if (win) {
				S.colour[k] = c;
				if (ba >= 0) S.taken[ba] = c;
				if (bb >= 0) S.taken[bb] = c;
				won++;
			} else {
				if (ba >= 0) other[ba] = ~0ull;
				if (bb >= 0) other[bb] = ~0ull;
			}
		}
		if (won) atomicSub(S.left, won);
		grid_sync(S.bar, W.err);
		if (++sub % COLOUR_SUBROUNDS == 0) c++;
	}
	const int n_colours = (sub % COLOUR_SUBROUNDS == 0) ? c : c + 1;

	// ---- colour sort ----
	for (int k = tid; k < n; k += stride)
		if (S.colour[k] >= 0) atomicAdd(&S.col_off[S.colour[k]], 1);
	grid_sync(S.bar, W.err);
	if (tid == 0) {
		int run = 0;
		for (int q = 0; q < n_colours; q++) {
			const int cnt = S.col_off[q];
			S.col_off[q] = run;
			S.col_cursor[q] = run;
			run += cnt;
		}
		S.col_off[n_colours] = run;
		*S.n_colours = n_colours;
	}
	grid_sync(S.bar, W.err);
	for (int k = tid; k < n; k += stride) {
		if (S.colour[k] < 0) continue;
		const int pos = atomicAdd(&S.col_cursor[S.colour[k]], 1);
		S.order[pos] = k;
	}
	// ---- pre-steps (cpSpaceStep.c:383-396) ----
	for (int k = tid; k < n; k += stride) arb_prestep(W, S.act[k]);
	grid_sync(S.bar, W.err);
	for (int jc = 0; jc < S.n_jcolours; jc++) {
		for (int k = S.jcol_off[jc] + tid; k < S.jcol_off[jc + 1]; k += stride) joint_prestep(W, W.joints[S.jorder[k]]);
		grid_sync(S.bar, W.err);
	}
	// ---- velocity integration (cpSpaceStep.c:399-404, cpBody.c:493-509) ----
	for (int q = tid; q < S.n_dyn_bodies; q += stride) {
		const int b = S.dyn_bodies[q];
		if (W.b_type[b] != CPGPU_BODY_DYNAMIC) continue;   // kinematic bodies keep their velocity
		const WorldStep ws = W.ws[W.b_world[b]];
		BodyVel bv = W.b_vel[b];
		const double2 f = W.b_f[b];
		const V2 v = vadd(vmult(v2(bv.vx, bv.vy), ws.damping),
		                  vmult(vadd(v2(ws.gx, ws.gy), vmult(v2(f.x, f.y), bv.m_inv)), ws.dt));
		bv.vx = v.x; bv.vy = v.y;
		bv.w = bv.w * ws.damping + W.b_t[b] * bv.i_inv * ws.dt;
		W.b_vel[b] = bv;
		W.b_f[b] = make_double2(0.0, 0.0);
		W.b_t[b] = 0.0;
	}
	grid_sync(S.bar, W.err);
	// ---- cached impulses (cpSpaceStep.c:407-415) ----
	for (int q = 0; q < n_colours; q++) {
		for (int k = S.col_off[q] + tid; k < S.col_off[q + 1]; k += stride) arb_apply_cached(W, S.act[S.order[k]]);
		grid_sync(S.bar, W.err);
	}
	for (int jc = 0; jc < S.n_jcolours; jc++) {
		for (int k = S.jcol_off[jc] + tid; k < S.jcol_off[jc + 1]; k += stride)
			joint_apply_cached(W, W.joints[S.jorder[k]]);
		grid_sync(S.bar, W.err);
	}
	// ---- the impulse iterations (cpSpaceStep.c:418-427) ----
	for (int it = 0; it < S.max_iterations; it++) {
		for (int q = 0; q < n_colours; q++) {
			for (int k = S.col_off[q] + tid; k < S.col_off[q + 1]; k += stride) {
				Arbiter &arb = S.act[S.order[k]];
				if (it < W.ws[arb.world].iterations) arb_apply_impulse(W, arb);
			}
			grid_sync(S.bar, W.err);
		}
		for (int jc = 0; jc < S.n_jcolours; jc++) {
			for (int k = S.jcol_off[jc] + tid; k < S.jcol_off[jc + 1]; k += stride) {
				Joint &j = W.joints[S.jorder[k]];
				if (it < W.ws[j.world].iterations) joint_apply_impulse(W, j);
			}
			grid_sync(S.bar, W.err);
		}
	}
}

} // namespace

int solver_max_grid()
{
	int dev = 0, per_cu = 0;
	hipDeviceProp_t prop;
	(void)hipGetDevice(&dev);
	(void)hipGetDeviceProperties(&prop, dev);
	(void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_solve),
	                                                   SOLVER_THREADS, 0);
	if (per_cu < 1) per_cu = 1;
	return per_cu * prop.multiProcessorCount;
}

hipError_t solver_step(const DevWorld &W, const SolveBuffers &S, int grid, hipStream_t stream, KernelTimer *timer)
{
	void *args[] = {(void *)&W, (void *)&S};
	hipError_t e = hipSuccess;
	TIMED(timer, "solver", stream,
	      e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_solve), dim3(grid), dim3(SOLVER_THREADS),
	                                     args, 0, stream));
	return e;
}
