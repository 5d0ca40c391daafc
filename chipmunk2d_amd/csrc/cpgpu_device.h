// cpgpu_device.h -- device-side math and state layout of the gfx950 step library.
//
// All kernels are compiled with -ffp-contract=off: every + - * / sqrt below is
// one IEEE operation of the build's real type.  In the FP64 build (real =
// double) that is the sequence the reference performs, so results are
// bit-identical to the reference's strict build; the FP32 build (real =
// float) runs the same sequence in single precision.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cpgpu.h"

// The library is compiled twice: the FP64 step (namespace cpgpu64, real =
// double) and the FP32 step (namespace cpgpu32, real = float); dispatch.hip
// routes each C ABI call to the world's own build.
#ifndef CPGPU_NS
#define CPGPU_NS cpgpu64
#endif
#ifndef CPGPU_REAL
#define CPGPU_REAL double
#endif

namespace CPGPU_NS {

// The arithmetic type of the device state and the step.
typedef CPGPU_REAL real;
typedef HIP_vector_type<real, 2> real2;
typedef HIP_vector_type<real, 4> real4;
__host__ __device__ inline real2 make_real2(real x, real y)
{
	real2 r;
	r.x = x; r.y = y;
	return r;
}
__host__ __device__ inline real4 make_real4(real x, real y, real z, real w)
{
	real4 r;
	r.x = x; r.y = y; r.z = z; r.w = w;
	return r;
}
// smallest positive normal value (CPFLOAT_MIN, chipmunk_types.h)
#define CP_REAL_MIN (sizeof(real) == 4 ? (real)1.17549435e-38f : (real)2.2250738585072014e-308)

#define CPD __device__ __forceinline__

// ---------------------------------------------------------------------------
// vectors (include/chipmunk/cpVect.h, chipmunk_types.h:96-147)

// Generic over the scalar: double for the step, a forward-mode dual number
// (cpgpu_dual.h) where the adjoint needs contact-geometry derivatives.  The
// value part of every operation is the same IEEE double operation.
template <class T> struct V2T { T x, y; };
using V2 = V2T<real>;

CPD double csqrt(double x) { return __builtin_sqrt(x); }
CPD float csqrt(float x) { return __builtin_sqrtf(x); }

CPD V2 v2(real x, real y) { V2 r; r.x = x; r.y = y; return r; }
template <class T> CPD V2T<T> v2t(T x, T y) { V2T<T> r; r.x = x; r.y = y; return r; }
template <class T, class S> CPD V2T<T> vcast(V2T<S> a) { return v2t<T>(T(a.x), T(a.y)); }
template <class T> CPD V2T<T> vadd(V2T<T> a, V2T<T> b) { return v2t<T>(a.x + b.x, a.y + b.y); }
template <class T> CPD V2T<T> vsub(V2T<T> a, V2T<T> b) { return v2t<T>(a.x - b.x, a.y - b.y); }
template <class T> CPD V2T<T> vneg(V2T<T> a) { return v2t<T>(-a.x, -a.y); }
template <class T, class S> CPD V2T<T> vmult(V2T<T> a, S s) { return v2t<T>(a.x * s, a.y * s); }
template <class T> CPD T vdot(V2T<T> a, V2T<T> b) { return a.x * b.x + a.y * b.y; }
template <class T> CPD T vcross(V2T<T> a, V2T<T> b) { return a.x * b.y - a.y * b.x; }
template <class T> CPD V2T<T> vperp(V2T<T> a) { return v2t<T>(-a.y, a.x); }
template <class T> CPD V2T<T> vrperp(V2T<T> a) { return v2t<T>(a.y, -a.x); }
template <class T> CPD V2T<T> vrotate(V2T<T> a, V2T<T> b)
{
	return v2t<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class T> CPD T vlengthsq(V2T<T> a) { return vdot(a, a); }
template <class T> CPD T vlength(V2T<T> a) { return csqrt(vdot(a, a)); }
template <class T> CPD V2T<T> vnormalize(V2T<T> a) { return vmult(a, T(1.0) / (vlength(a) + CP_REAL_MIN)); }
template <class T, class S> CPD V2T<T> vlerp(V2T<T> a, V2T<T> b, S t)
{
	return vadd(vmult(a, S(1.0) - t), vmult(b, t));
}
template <class T> CPD bool veql(V2T<T> a, V2T<T> b) { return a.x == b.x && a.y == b.y; }
template <class T, class S> CPD V2T<T> vclamp(V2T<T> v, S len)
{
	return (vdot(v, v) > len * len) ? vmult(vnormalize(v), len) : v;
}
template <class T> CPD T fmaxc(T a, T b) { return (a > b) ? a : b; }
template <class T> CPD T fminc(T a, T b) { return (a < b) ? a : b; }
template <class T, class L, class H> CPD T fclampc(T f, L lo, H hi) { return fminc(fmaxc(f, T(lo)), T(hi)); }
template <class T> CPD T fclamp01(T f) { return fmaxc(T(0.0), fminc(f, T(1.0))); }

// cpTransform {a, b, c, d, tx, ty} (cpTransform.h)
template <class T> struct XFT { T a, b, c, d, tx, ty; };
using XF = XFT<real>;
template <class T, class S> CPD V2T<T> xpoint(const XFT<T> &t, V2T<S> p)
{
	return v2t<T>(t.a * p.x + t.c * p.y + t.tx, t.b * p.x + t.d * p.y + t.ty);
}
template <class T, class S> CPD V2T<T> xvect(const XFT<T> &t, V2T<S> v)
{
	return v2t<T>(t.a * v.x + t.c * v.y, t.b * v.x + t.d * v.y);
}

// CP_HASH_PAIR (chipmunk_private.h:28-29), 64-bit cpHashValue
#define CP_HASH_COEF 3344921057ull
CPD uint64_t hash_pair(uint64_t a, uint64_t b) { return a * CP_HASH_COEF ^ b * CP_HASH_COEF; }

// splitmix64 finalizer: a bijection, used for hashing keys and colouring priorities
CPD uint64_t mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

CPD uint64_t pair_key(int32_t a, int32_t b)
{
	uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
	return ((uint64_t)lo << 32) | hi;
}

// ---------------------------------------------------------------------------
// deterministic sin/cos and exp (bit-identical on the host oracle): Cody-Waite
// reduction in three FMA steps, Taylor series with 1/n! coefficients.

CPD void cp_sincos(double x, double &s_out, double &c_out)
{
	if (!(__builtin_fabs(x) < 1e15)) { s_out = x - x; c_out = x - x; return; }
	const double k = __builtin_rint(x * 0x1.45f306dc9c883p-1);
	double r = __builtin_fma(-k, 0x1.921fb54442d18p+0, x);
	r = __builtin_fma(-k, 0x1.1a62633145c07p-54, r);
	r = __builtin_fma(-k, -0x1.f1976b7ed8fbcp-110, r);
	const double z = r * r;
	double ps = -0x1.2f49b46814157p-57;
	ps = ps * z + 0x1.952c77030ad4ap-49;
	ps = ps * z - 0x1.ae7f3e733b81fp-41;
	ps = ps * z + 0x1.6124613a86d09p-33;
	ps = ps * z - 0x1.ae64567f544e4p-26;
	ps = ps * z + 0x1.71de3a556c734p-19;
	ps = ps * z - 0x1.a01a01a01a01ap-13;
	ps = ps * z + 0x1.1111111111111p-7;
	ps = ps * z - 0x1.5555555555555p-3;
	const double sn = r + r * (z * ps);
	double qc = -0x1.6827863b97d97p-53;
	qc = qc * z + 0x1.ae7f3e733b81fp-45;
	qc = qc * z - 0x1.93974a8c07c9dp-37;
	qc = qc * z + 0x1.1eed8eff8d898p-29;
	qc = qc * z - 0x1.27e4fb7789f5cp-22;
	qc = qc * z + 0x1.a01a01a01a01ap-16;
	qc = qc * z - 0x1.6c16c16c16c17p-10;
	qc = qc * z + 0x1.5555555555555p-5;
	const double hz = 0.5 * z;
	const double wc = 1.0 - hz;
	const double cs = wc + (((1.0 - wc) - hz) + (z * z) * qc);
	const long q = (long)k & 3;
	if (q == 0) { s_out = sn; c_out = cs; }
	else if (q == 1) { s_out = cs; c_out = -sn; }
	else if (q == 2) { s_out = -sn; c_out = -cs; }
	else { s_out = -cs; c_out = sn; }
}

CPD double cp_exp(double x)
{
	if (x != x) return x;
	if (x > 709.782712893384) return __builtin_inf();
	if (x < -745.2) return 0.0;
	const double k = __builtin_rint(x * 0x1.71547652b82fep+0);
	double r = __builtin_fma(-k, 0x1.62e42fefa39efp-1, x);
	r = __builtin_fma(-k, 0x1.abc9e3b39803fp-56, r);
	double p = 0x1.6124613a86d09p-33;
	p = p * r + 0x1.1eed8eff8d898p-29;
	p = p * r + 0x1.ae64567f544e4p-26;
	p = p * r + 0x1.27e4fb7789f5cp-22;
	p = p * r + 0x1.71de3a556c734p-19;
	p = p * r + 0x1.a01a01a01a01ap-16;
	p = p * r + 0x1.a01a01a01a01ap-13;
	p = p * r + 0x1.6c16c16c16c17p-10;
	p = p * r + 0x1.1111111111111p-7;
	p = p * r + 0x1.5555555555555p-5;
	p = p * r + 0x1.5555555555555p-3;
	p = p * r + 0.5;
	const double e = 1.0 + (r + (r * r) * p);
	return __builtin_ldexp(e, (int)k);
}

// ---------------------------------------------------------------------------
// device state

// Per-world step parameters, refreshed by the host every step.
struct WorldStep {
	real dt, dt_coef, damping, bias_coef, slop;
	real gx, gy;
	int32_t iterations;
	uint32_t stamp;
	uint32_t persistence;
	int32_t _pad;
};

// Velocity record of a body: the solver's working set (64 bytes).
struct BodyVel {
	real vx, vy, w, vbx, vby, wb, m_inv, i_inv;
};

// A contact of an arbiter (struct cpContact, chipmunk_structs.h:101-111).
struct Contact {
	real r1x, r1y, r2x, r2y;
	real nMass, tMass, bounce, jnAcc, jtAcc, jBias, bias;
	uint64_t hash;
	// rollout tape (adjoint): where this contact's accumulated impulses were
	// recorded after its last solve, and the recorded contact its warm start
	// came from (-1: none)
	int32_t tape, src;
};

// An arbiter (struct cpArbiter, chipmunk_structs.h:124-145), device form.
struct Arbiter {
	uint64_t key;
	int32_t sa, sb, ba, bb;
	int32_t count, state;
	uint32_t stamp;
	int32_t world;
	uint32_t id_in;            // GJK warm-start id the collision started from
	int32_t events;            // the pair's collision handler events (CPGPU_EVENT_*)
	real nx, ny, e, u, svx, svy;
	Contact c[2];
};

// Persistent pair record: the warm-start GJK id of a collided pair and the
// arbiter it owns (index into the arbiter array, -1 for none).
struct PairRec {
	uint64_t key;
	uint32_t id, id_stamp;
	int32_t arb;
	int32_t _pad;
};

// Joint (struct cpConstraint + joint subclasses), device form.
struct Joint {
	int32_t type, a, b, world;
	real maxForce, errorBias, maxBias;
	real P[8];
	real r1x, r1y, r2x, r2y, nx, ny;
	real k11, k12, k21, k22;
	real bx, by, bias, nMass, jAx, jAy, jAcc, jnAcc, target_vrn, v_coef, iSum;
	real bias_coef;
	real clamp;   // groove: the groove end the anchor rests on (+1, -1), or 0 between them
	real angle;   // ratchet: the angle of the last notch passed, carried between steps
};

struct DevWorld {
	int32_t n_worlds, n_bodies, n_shapes, n_joints;
	// bodies
	const int32_t *b_type;
	const int32_t *b_world;
	const real *b_m;             // mass (INFINITY marks static/kinematic)
	real2 *b_p;                  // position
	real *b_a;                   // angle
	real2 *b_f;
	real *b_t;
	real2 *b_cog;
	XF *b_T;                     // transform
	BodyVel *b_vel;
	// shapes
	const int32_t *s_type, *s_body, *s_world;
	const uint64_t *s_hashid;
	const real *s_e, *s_u;
	const real2 *s_sv;
	const uint64_t *s_group;
	const uint32_t *s_cats, *s_mask;
	const uint8_t *s_sensor;     // 1 for a sensor shape; nullptr when the world has none
	const uint32_t *s_ctype;     // collision type; nullptr when the description has no handlers
	const cpgpuHandlerDesc *handlers;
	int n_handlers;
	const real *s_r;
	const real *s_geom;          // [n][10] body-local geometry
	const int32_t *s_pstart, *s_pcount;
	const real2 *pv_local, *pn_local;
	real2 *pv_world, *pn_world;
	real *s_wgeom;               // [n][6]: circle tc; segment ta, tb, tn
	real4 *s_bb;                 // l, b, r, t
	real4 *s_fat;                // dynamic-tree leaf boxes
	uint32_t *s_rebuild;         // step of the last leaf rebuild
	const uint8_t *s_static;     // shape belongs to a static body
	// no-collide constraint partners (CSR over bodies)
	const int32_t *nc_start, *nc_list;
	// first shape / body index of every world ([n_worlds + 1])
	const int32_t *w_shape0;
	const int32_t *w_body0;
	// per-world step parameters
	const WorldStep *ws;
	// joints
	Joint *joints;
	// device error word
	int32_t *err;
	// cumulative counters: [0] arbiters solved, [1] contacts solved
	unsigned long long *totals;
};

// The events of the collision handlers that match shapes s and t: the pair's own
// handler and the wildcard handlers of either type (cpSpaceLookupHandler,
// cpSpace.c:382-399; cpArbiterUpdate, cpArbiter.c:402-409; the default
// handlers call the wildcard ones, cpSpace.c:63-103).
__device__ __forceinline__ uint32_t pair_events(const DevWorld &W, int s, int t)
{
	uint32_t ev = 0;
	const uint32_t ta = W.s_ctype[s], tb = W.s_ctype[t];
	const int space = W.s_world[s];
	for (int k = 0; k < W.n_handlers; k++) {
		const cpgpuHandlerDesc h = W.handlers[k];
		if (h.space >= 0 && h.space != space) continue;
		bool match;
		if (h.type_b == CPGPU_WILDCARD) match = h.type_a == CPGPU_WILDCARD || h.type_a == ta || h.type_a == tb;
		else match = (h.type_a == ta && h.type_b == tb) || (h.type_a == tb && h.type_b == ta);
		if (match) ev |= h.events;
	}
	return ev;
}

#define CPGPU_ERR_PAIRS_OVERFLOW 1
#define CPGPU_ERR_ARBITERS_OVERFLOW 2
#define CPGPU_ERR_RECORDS_OVERFLOW 4
#define CPGPU_ERR_TRAVERSAL_STACK 8
#define CPGPU_ERR_BARRIER_TIMEOUT 16
#define CPGPU_ERR_TABLE_FULL 32
#define CPGPU_ERR_EPA_HULL 64
#define CPGPU_ERR_ADJOINT 256

CPD void raise_err(int32_t *err, int32_t code) { atomicOr(err, code); }

// The reference's debug-build step checks (CPGPU_FAULT_*) share the word.
#define CPGPU_FAULTS (CPGPU_FAULT_UNSOLVABLE | CPGPU_FAULT_BODY_STATE | CPGPU_FAULT_BODY_MASS)

// a body failed one of them: err[1] keeps the lowest such index b as INT32_MAX - b (0: none)
CPD void raise_body_fault(int32_t *err, int32_t code, int b)
{
	atomicOr(err, code);
	atomicMax(err + 1, 0x7fffffff - b);
}

// cpBodySanityCheck (cpBody.c:110-125) over what the device keeps of a body:
// its moment only as the inverse, which carries the same NaN and sign tests
CPD bool body_sane(real m, real m_inv, real i_inv, real px, real py, real vx, real vy, real fx, real fy, real a,
                   real w, real t)
{
	return m == m && m_inv == m_inv && i_inv == i_inv && m >= real(0.0) && i_inv >= real(0.0) &&
	       __builtin_isfinite(px) && __builtin_isfinite(py) && __builtin_isfinite(vx) && __builtin_isfinite(vy) &&
	       __builtin_isfinite(fx) && __builtin_isfinite(fy) && __builtin_isfinite(a) && __builtin_isfinite(w) &&
	       __builtin_isfinite(t);
}

} // namespace CPGPU_NS
