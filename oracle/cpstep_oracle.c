/*
 * cpstep_oracle.c -- CPU restatement of Chipmunk2D 7.0.3's cpSpaceStep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP step library; see the
 * header).  Compiled with gcc -O2 -ffp-contract=off -fno-fast-math so every
 * + - * / sqrt is a single IEEE num operation, the same as the reference's
 * strict build and the device library.
 *
 * Every routine cites the reference file:line it restates.  How each part is
 * pinned (tests/test_oracle_pins.py, against the reference compiled from
 * /root/reference into oracle/_ref by oracle/Makefile):
 *   integration, shape caching   bit-exact vs cpBodyUpdatePosition/Velocity and
 *                                cpShapeUpdate of the compiled reference, plus the
 *                                closed form of BodyTest.m:147-184
 *   candidate pairs              equal as a set to the reference's own
 *                                cpSpatialIndexReindexQuery + QueryReject, and to
 *                                O(n^2) brute force on small worlds
 *   narrowphase                  bit-exact vs the reference's cpCollide on the same
 *                                shapes, orientation and GJK id (all 6 pair types)
 *   one full step                bit-exact vs cpSpaceStep when replayed in the
 *                                reference's arbiter order with libm trig
 *   joints                       bit-exact vs one cpSpaceStep with joints replayed
 *                                in the reference's constraint order
 *   cp_sincos / cp_exp           within 1 ulp of glibc over dense sweeps
 * The deterministic graph colouring (the device solve order) has no reference
 * counterpart; it is pinned by its defining properties (a proper colouring,
 * every arbiter coloured, identical to an independent Python restatement).
 * So has the split of a large space into solve units (strips, seams, the
 * wide set), pinned against an independent numpy restatement, each unit's
 * colouring against the Python claim rounds.
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <stdio.h>

#include "cpstep_oracle.h"

/* The scalar of the step's state and arithmetic: double here.  dual_step.cpp
 * compiles this file again as C++ with a dual number (value and tangent) in
 * its place, for the forward-mode derivative of the same step; NUM_VAL(x) is
 * the value part, for the decisions and outputs that take a double. */
#ifndef ORACLE_NUM
typedef double num;
#define NUM_VAL(x) (x)
#else
typedef ORACLE_NUM num;
#endif

/* ------------------------------------------------------------------ */
/* scalar and vector helpers: include/chipmunk/chipmunk_types.h:96-147,
 * cpVect.h, cpBB.h, cpTransform.h                                     */

typedef struct { num x, y; } V;

static inline V v2(num x, num y) { V r = {x, y}; return r; }
static inline V vadd(V a, V b) { return v2(a.x + b.x, a.y + b.y); }
static inline V vsub(V a, V b) { return v2(a.x - b.x, a.y - b.y); }
static inline V vneg(V a) { return v2(-a.x, -a.y); }
static inline V vmult(V a, num s) { return v2(a.x * s, a.y * s); }
static inline num vdot(V a, V b) { return a.x * b.x + a.y * b.y; }
static inline num vcross(V a, V b) { return a.x * b.y - a.y * b.x; }
static inline V vperp(V a) { return v2(-a.y, a.x); }
static inline V vrperp(V a) { return v2(a.y, -a.x); }
static inline V vrotate(V a, V b) { return v2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
static inline num vlengthsq(V a) { return vdot(a, a); }
static inline num vlength(V a) { return sqrt(vdot(a, a)); }
static inline V vnormalize(V a) { return vmult(a, 1.0 / (vlength(a) + DBL_MIN)); }
static inline V vlerp(V a, V b, num t) { return vadd(vmult(a, 1.0 - t), vmult(b, t)); }
static inline int veql(V a, V b) { return a.x == b.x && a.y == b.y; }
static inline V vclamp(V v, num len) { return (vdot(v, v) > len * len) ? vmult(vnormalize(v), len) : v; }
static inline num fmaxc(num a, num b) { return (a > b) ? a : b; }
static inline num fminc(num a, num b) { return (a < b) ? a : b; }
static inline num fclampc(num f, num lo, num hi) { return fminc(fmaxc(f, lo), hi); }
static inline num fclamp01(num f) { return fmaxc(0.0, fminc(f, 1.0)); }

typedef struct { num a, b, c, d, tx, ty; } XF;
static inline V xpoint(XF t, V p) { return v2(t.a * p.x + t.c * p.y + t.tx, t.b * p.x + t.d * p.y + t.ty); }
static inline V xvect(XF t, V v) { return v2(t.a * v.x + t.c * v.y, t.b * v.x + t.d * v.y); }

/* CP_HASH_PAIR, include/chipmunk/chipmunk_private.h:28-29 (cpHashValue is 64-bit uintptr_t). */
#define HASH_COEF 3344921057ull
static inline uint64_t hash_pair(uint64_t a, uint64_t b) { return a * HASH_COEF ^ b * HASH_COEF; }

/* ------------------------------------------------------------------ */
/* Deterministic transcendentals shared (as an algorithm) with the device
 * library.  sin/cos: Cody-Waite reduction by pi/2 in three FMA steps, then
 * the Taylor series of sin and cos on [-pi/4, pi/4] (coefficients 1/n!
 * rounded to num), with the 1 - z/2 rounding error of cos carried
 * separately.  exp: reduction by ln 2 and the Taylor series of exp on
 * [-ln2/2, ln2/2], scaled by ldexp.  Pure IEEE operations: identical bits
 * on every IEEE-754 num machine. */

#define PIO2_1 0x1.921fb54442d18p+0
#define PIO2_2 0x1.1a62633145c07p-54
#define PIO2_3 -0x1.f1976b7ed8fbcp-110
#define TWO_OVER_PI 0x1.45f306dc9c883p-1

void cp_sincos_oracle(double x, double *s_out, double *c_out)
{
	if (!(fabs(x) < 1e15)) { *s_out = x - x; *c_out = x - x; return; }
	double k = rint(x * TWO_OVER_PI);
	double r = fma(-k, PIO2_1, x);
	r = fma(-k, PIO2_2, r);
	r = fma(-k, PIO2_3, r);
	double z = r * r;
	/* sin r = r + r z (-1/3! + z (1/5! - ...)), through r^19 */
	double ps = -0x1.2f49b46814157p-57;          /* -1/19! */
	ps = ps * z + 0x1.952c77030ad4ap-49;         /*  1/17! */
	ps = ps * z - 0x1.ae7f3e733b81fp-41;         /* -1/15! */
	ps = ps * z + 0x1.6124613a86d09p-33;         /*  1/13! */
	ps = ps * z - 0x1.ae64567f544e4p-26;         /* -1/11! */
	ps = ps * z + 0x1.71de3a556c734p-19;         /*  1/9!  */
	ps = ps * z - 0x1.a01a01a01a01ap-13;         /* -1/7!  */
	ps = ps * z + 0x1.1111111111111p-7;          /*  1/5!  */
	ps = ps * z - 0x1.5555555555555p-3;          /* -1/3!  */
	double sn = r + r * (z * ps);
	/* cos r = 1 - z/2 + z^2 (1/4! - z/6! + ...), through r^18 */
	double qc = -0x1.6827863b97d97p-53;          /* -1/18! */
	qc = qc * z + 0x1.ae7f3e733b81fp-45;         /*  1/16! */
	qc = qc * z - 0x1.93974a8c07c9dp-37;         /* -1/14! */
	qc = qc * z + 0x1.1eed8eff8d898p-29;         /*  1/12! */
	qc = qc * z - 0x1.27e4fb7789f5cp-22;         /* -1/10! */
	qc = qc * z + 0x1.a01a01a01a01ap-16;         /*  1/8!  */
	qc = qc * z - 0x1.6c16c16c16c17p-10;         /* -1/6!  */
	qc = qc * z + 0x1.5555555555555p-5;          /*  1/4!  */
	double hz = 0.5 * z;
	double wc = 1.0 - hz;
	double cs = wc + (((1.0 - wc) - hz) + (z * z) * qc);
	long q = (long)k & 3;
	if (q == 0) { *s_out = sn; *c_out = cs; }
	else if (q == 1) { *s_out = cs; *c_out = -sn; }
	else if (q == 2) { *s_out = -sn; *c_out = -cs; }
	else { *s_out = -cs; *c_out = sn; }
}

#define LN2_HI 0x1.62e42fefa39efp-1
#define LN2_LO 0x1.abc9e3b39803fp-56
#define LOG2E 0x1.71547652b82fep+0

double cp_exp_oracle(double x)
{
	if (x != x) return x;
	if (x > 709.782712893384) return INFINITY;
	if (x < -745.2) return 0.0;
	double k = rint(x * LOG2E);
	double r = fma(-k, LN2_HI, x);
	r = fma(-k, LN2_LO, r);
	double p = 0x1.6124613a86d09p-33;            /* 1/13! */
	p = p * r + 0x1.1eed8eff8d898p-29;           /* 1/12! */
	p = p * r + 0x1.ae64567f544e4p-26;           /* 1/11! */
	p = p * r + 0x1.27e4fb7789f5cp-22;           /* 1/10! */
	p = p * r + 0x1.71de3a556c734p-19;           /* 1/9!  */
	p = p * r + 0x1.a01a01a01a01ap-16;           /* 1/8!  */
	p = p * r + 0x1.a01a01a01a01ap-13;           /* 1/7!  */
	p = p * r + 0x1.6c16c16c16c17p-10;           /* 1/6!  */
	p = p * r + 0x1.1111111111111p-7;            /* 1/5!  */
	p = p * r + 0x1.5555555555555p-5;            /* 1/4!  */
	p = p * r + 0x1.5555555555555p-3;            /* 1/3!  */
	p = p * r + 0.5;
	double e = 1.0 + (r + (r * r) * p);
	return ldexp(e, (int)k);
}

/* ------------------------------------------------------------------ */
/* state                                                               */

typedef struct {
	int type, world;
	num m, m_inv, i, i_inv;
	V cog, p, v, f;
	num a, w, t;
	XF T;
	V v_bias;
	num w_bias;
} OBody;

typedef struct {
	int type, body, world;
	uint64_t hashid;
	num e, u;
	V sv;
	uint64_t group;
	uint32_t cats, mask;
	int sensor;                              /* cpShape.sensor */
	uint32_t ctype;                          /* cpShape.type (collision type) */
	num r;
	V c, tc;                                 /* circle */
	V a, b, n, ta, tb, tn, a_tan, b_tan;     /* segment */
	int count;                               /* poly */
	const V *lv, *ln;
	V *wv, *wn;
	num bb[4];                            /* l, b, r, t */
	num fat[4];                           /* velocity-fattened leaf box of the dynamic BBTree */
	uint32_t rebuild_stamp;                  /* step of the last leaf rebuild */
} OShape;

typedef struct {
	V r1, r2;
	num nMass, tMass, bounce, jnAcc, jtAcc, jBias, bias;
	uint64_t hash;
} OContact;

typedef struct {
	int32_t sa, sb, ba, bb;
	uint64_t key, prio;
	num e, u;
	V svr, n;
	int count;
	OContact con[2];
	uint32_t id, stamp;
	int state, colour;
	int unit;   /* solve unit of a large space (grid_units) */
	uint32_t events;   /* the pair's collision handler events (CPGPU_EVENT_*) */
} OArb;

typedef struct {
	int type, a, b;
	num maxForce, errorBias, maxBias;
	int collide;
	num P[8];
	/* solver state (chipmunk_structs.h:270-391) */
	V r1, r2, n, bias2, jAcc2;
	num k11, k12, k21, k22;
	num nMass, bias, jAcc, jnAcc, target_vrn, v_coef, iSum;
	num clamp;   /* groove: which end of the groove the anchor rests on (+1, -1) or 0 */
	num angle;   /* ratchet: the angle of the last notch passed (carried between steps) */
	int colour;
	uint64_t prio;
} OJoint;

typedef struct {
	int b0, b1, s0, s1, c0, c1;
	cpgpuSpaceParams sp;
	OArb *arbs;
	int64_t narb, arbcap;
	int64_t *tab;
	int64_t tabmask;
	int64_t *active;       /* indices into arbs, in solve order */
	int64_t nactive;
	int32_t *jorder;       /* constraint solve order */
	int ncol_arb, ncol_cons;
	uint64_t *nocollide;   /* sorted body-pair keys of collideBodies == false constraints */
	int64_t n_nocollide;
	int64_t npairs, ncontacts;
	/* GJK warm-start ids of collided pairs (the Pair.id cache of cpBBTree.c:69-78) */
	uint64_t *pid_key;
	uint32_t *pid_id, *pid_stamp;
	int64_t pid_n, pid_mask;
} OWorldPart;

struct oworld {
	int nw;
	OWorldPart *W;
	int nb, ns, nc;
	OBody *B;
	OShape *S;
	OJoint *J;
	V *polymem;
	int libm_trig;
	int64_t next_n;
	int32_t *next_pairs;
	int64_t ncoll;
	int32_t *coll;      /* replayed collisions {a, b, id}, sorted by pair key */
	uint64_t *coll_key;
	int joint_list_order;
	int64_t steps;
	int grid_parts;     /* strips of every large space's grid set (grid_parts_of) */
	/* collision handlers: the description's table, the host's answers */
	cpgpuHandlerDesc *handlers;
	int n_handlers;
	ow_handler_fn hfn;
	void *hctx;
	int32_t *seps;      /* the last step's separations {space, a, b} */
	int64_t nseps, sepcap;
	/* the conditions a debug build of the reference asserts during a step
	 * (CPGPU_FAULT_*), kept from step to step, and the lowest body that failed */
	int32_t faults, fault_body;
};

static int grid_parts_of(const oworld *w);

/* splitmix64 finalizer: a bijection on 64-bit integers; the colouring priority. */
static inline uint64_t mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

static inline uint64_t pair_key(int32_t a, int32_t b)
{
	uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
	return ((uint64_t)lo << 32) | hi;
}

/* ------------------------------------------------------------------ */
/* construction                                                        */

static XF xf_from(const double *t) { XF x = {t[0], t[1], t[2], t[3], t[4], t[5]}; return x; }

static int cmp_u64(const void *a, const void *b)
{
	uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
	return x < y ? -1 : (x > y ? 1 : 0);
}

static void cache_shape(oworld *w, OShape *s);
static void fat_bb(oworld *w, OShape *s);

oworld *ow_new(const cpgpuWorldDesc *d)
{
	oworld *w = (oworld *)calloc(1, sizeof(oworld));
	w->fault_body = -1;
	w->nw = d->n_worlds;
	w->nb = d->n_bodies;
	w->ns = d->n_shapes;
	w->nc = d->n_constraints;
	w->B = (OBody *)calloc((size_t)(w->nb ? w->nb : 1), sizeof(OBody));
	w->S = (OShape *)calloc((size_t)(w->ns ? w->ns : 1), sizeof(OShape));
	w->J = (OJoint *)calloc((size_t)(w->nc ? w->nc : 1), sizeof(OJoint));
	w->W = (OWorldPart *)calloc((size_t)w->nw, sizeof(OWorldPart));
	w->polymem = (V *)calloc((size_t)(4 * (d->n_poly_verts + 1)), sizeof(V));

	for (int wi = 0; wi < w->nw; wi++) {
		OWorldPart *P = &w->W[wi];
		P->b0 = d->world_body_start[wi]; P->b1 = d->world_body_start[wi + 1];
		P->s0 = d->world_shape_start[wi]; P->s1 = d->world_shape_start[wi + 1];
		P->c0 = d->world_constraint_start ? d->world_constraint_start[wi] : 0;
		P->c1 = d->world_constraint_start ? d->world_constraint_start[wi + 1] : 0;
		P->sp = d->space[wi];
		for (int b = P->b0; b < P->b1; b++) w->B[b].world = wi;
	}
	for (int i = 0; i < w->nb; i++) {
		OBody *b = &w->B[i];
		b->type = d->body_type[i];
		b->m = d->body_m[i]; b->m_inv = d->body_m_inv[i];
		b->i = d->body_i[i]; b->i_inv = d->body_i_inv[i];
		b->cog = v2(d->body_cog[2 * i], d->body_cog[2 * i + 1]);
		b->p = v2(d->body_p[2 * i], d->body_p[2 * i + 1]);
		b->v = v2(d->body_v[2 * i], d->body_v[2 * i + 1]);
		b->f = v2(d->body_f[2 * i], d->body_f[2 * i + 1]);
		b->a = d->body_a[i]; b->w = d->body_w[i]; b->t = d->body_t[i];
		b->v_bias = v2(d->body_v_bias[2 * i], d->body_v_bias[2 * i + 1]);
		b->w_bias = d->body_w_bias[i];
		b->T = xf_from(d->body_transform + 6 * i);
	}
	V *lvbase = w->polymem, *lnbase = w->polymem + d->n_poly_verts;
	V *wvbase = w->polymem + 2 * d->n_poly_verts, *wnbase = w->polymem + 3 * d->n_poly_verts;
	for (int k = 0; k < d->n_poly_verts; k++) {
		lvbase[k] = v2(d->poly_verts[2 * k], d->poly_verts[2 * k + 1]);
		lnbase[k] = v2(d->poly_normals[2 * k], d->poly_normals[2 * k + 1]);
	}
	for (int i = 0; i < w->ns; i++) {
		OShape *s = &w->S[i];
		const double *g = d->shape_geom + 10 * i;
		s->type = d->shape_type[i];
		s->body = d->shape_body[i];
		s->world = w->B[s->body].world;
		s->hashid = d->shape_hashid[i];
		s->e = d->shape_e[i]; s->u = d->shape_u[i];
		s->sv = v2(d->shape_surface_v[2 * i], d->shape_surface_v[2 * i + 1]);
		s->group = d->shape_group[i]; s->cats = d->shape_categories[i]; s->mask = d->shape_mask[i];
		s->sensor = d->shape_sensor ? d->shape_sensor[i] != 0 : 0;
		s->ctype = d->shape_ctype ? d->shape_ctype[i] : 0u;
		s->r = d->shape_radius[i];
		if (s->type == CPGPU_SHAPE_CIRCLE) {
			s->c = v2(g[0], g[1]);
		} else if (s->type == CPGPU_SHAPE_SEGMENT) {
			s->a = v2(g[0], g[1]); s->b = v2(g[2], g[3]); s->n = v2(g[4], g[5]);
			s->a_tan = v2(g[6], g[7]); s->b_tan = v2(g[8], g[9]);
		} else {
			int st = d->shape_poly_start[i];
			s->count = d->shape_poly_count[i];
			s->lv = lvbase + st; s->ln = lnbase + st;
			s->wv = wvbase + st; s->wn = wnbase + st;
		}
		/* cpSpaceAddShape caches every shape once with its body's transform (cpSpace.c:433)
		 * and inserts it into the dynamic index with a fattened box (cpBBTree.c:605-616). */
		cache_shape(w, s);
		if (w->B[s->body].type != CPGPU_BODY_STATIC) fat_bb(w, s);
	}
	for (int i = 0; i < w->nc; i++) {
		OJoint *j = &w->J[i];
		j->type = d->cons_type[i];
		j->a = d->cons_a[i]; j->b = d->cons_b[i];
		j->maxForce = d->cons_max_force[i]; j->errorBias = d->cons_error_bias[i]; j->maxBias = d->cons_max_bias[i];
		j->collide = d->cons_collide_bodies[i];
		for (int q = 0; q < 8; q++) j->P[q] = d->cons_params[8 * i + q];
		num j0 = d->cons_jacc ? d->cons_jacc[2 * i] : 0.0, j1 = d->cons_jacc ? d->cons_jacc[2 * i + 1] : 0.0;
		j->jAcc2 = v2(j0, j1);
		j->jAcc = j0; j->jnAcc = j0;
		j->angle = j1;
		/* colouring priority from the constraint's index within its own space */
		j->prio = mix64((uint64_t)(uint32_t)(i - w->W[w->B[j->a].world].c0) | 0x8000000000000000ull);
	}
	for (int wi = 0; wi < w->nw; wi++) {
		OWorldPart *P = &w->W[wi];
		int64_t n = 0;
		for (int c = P->c0; c < P->c1; c++) if (!w->J[c].collide) n++;
		P->nocollide = (uint64_t *)malloc((size_t)(n ? n : 1) * sizeof(uint64_t));
		n = 0;
		for (int c = P->c0; c < P->c1; c++) if (!w->J[c].collide) P->nocollide[n++] = pair_key(w->J[c].a, w->J[c].b);
		qsort(P->nocollide, (size_t)n, sizeof(uint64_t), cmp_u64);
		P->n_nocollide = n;
		P->arbcap = 64;
		P->arbs = (OArb *)calloc((size_t)P->arbcap, sizeof(OArb));
		P->active = (int64_t *)calloc((size_t)P->arbcap, sizeof(int64_t));
		P->tabmask = 127;
		P->tab = (int64_t *)malloc(128 * sizeof(int64_t));
		for (int k = 0; k < 128; k++) P->tab[k] = -1;
		P->jorder = (int32_t *)calloc((size_t)(P->c1 - P->c0 + 1), sizeof(int32_t));
		for (int c = P->c0; c < P->c1; c++) P->jorder[c - P->c0] = c;
	}
	w->grid_parts = grid_parts_of(w);
	if (d->n_handlers > 0 && d->handlers) {
		w->n_handlers = d->n_handlers;
		w->handlers = (cpgpuHandlerDesc *)malloc((size_t)d->n_handlers * sizeof(cpgpuHandlerDesc));
		memcpy(w->handlers, d->handlers, (size_t)d->n_handlers * sizeof(cpgpuHandlerDesc));
	}
	return w;
}

void ow_set_handler_fn(oworld *w, ow_handler_fn fn, void *ctx)
{
	w->hfn = fn;
	w->hctx = ctx;
}

/* The events of the handlers that match a pair of shapes: the pair's own
 * handler and the wildcard handlers of either type (cpSpaceLookupHandler,
 * cpSpace.c:382-399; cpArbiterUpdate, cpArbiter.c:402-409; the default
 * handlers call the wildcard ones, cpSpace.c:63-103). */
static uint32_t pair_events(const oworld *w, const OShape *a, const OShape *b)
{
	uint32_t ev = 0;
	const uint32_t ta = a->ctype, tb = b->ctype;
	for (int k = 0; k < w->n_handlers; k++) {
		const cpgpuHandlerDesc *h = &w->handlers[k];
		if (h->space >= 0 && h->space != a->world) continue;
		int match;
		if (h->type_b == CPGPU_WILDCARD) match = h->type_a == CPGPU_WILDCARD || h->type_a == ta || h->type_a == tb;
		else match = (h->type_a == ta && h->type_b == tb) || (h->type_a == tb && h->type_b == ta);
		if (match) ev |= h->events;
	}
	return ev;
}

static void sep_push(oworld *w, int32_t space, int32_t a, int32_t b)
{
	if (w->nseps == w->sepcap) {
		w->sepcap = w->sepcap ? 2 * w->sepcap : 64;
		w->seps = (int32_t *)realloc(w->seps, (size_t)w->sepcap * 3 * sizeof(int32_t));
	}
	int32_t *e = w->seps + 3 * w->nseps++;
	e[0] = space; e[1] = a; e[2] = b;
}

int64_t ow_read_separations(oworld *w, int64_t cap, int32_t *pairs)
{
	for (int64_t k = 0; k < w->nseps && k < cap; k++)
		for (int q = 0; q < 3; q++) pairs[3 * k + q] = w->seps[3 * k + q];
	return w->nseps;
}

void ow_free(oworld *w)
{
	if (!w) return;
	for (int wi = 0; wi < w->nw; wi++) {
		free(w->W[wi].arbs); free(w->W[wi].active); free(w->W[wi].tab);
		free(w->W[wi].nocollide); free(w->W[wi].jorder);
		free(w->W[wi].pid_key); free(w->W[wi].pid_id); free(w->W[wi].pid_stamp);
	}
	free(w->W); free(w->B); free(w->S); free(w->J); free(w->polymem); free(w->next_pairs);
	free(w->coll); free(w->coll_key);
	free(w->handlers); free(w->seps);
	free(w);
}

void ow_set_libm_trig(oworld *w, int enable) { w->libm_trig = enable; }
void ow_set_joint_list_order(oworld *w, int enable) { w->joint_list_order = enable; }

static int64_t *g_sort_coll_idx;
static const uint64_t *g_coll_keys;
static int cmp_coll(const void *x, const void *y)
{
	uint64_t a = g_coll_keys[*(const int64_t *)x], b = g_coll_keys[*(const int64_t *)y];
	return a < b ? -1 : (a > b ? 1 : 0);
}

void ow_set_next_collisions(oworld *w, int64_t n, const int32_t *coll)
{
	free(w->coll); free(w->coll_key);
	uint64_t *keys = (uint64_t *)malloc((size_t)(n + 1) * sizeof(uint64_t));
	int64_t *idx = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
	for (int64_t k = 0; k < n; k++) { keys[k] = pair_key(coll[3 * k], coll[3 * k + 1]); idx[k] = k; }
	g_coll_keys = keys;
	g_sort_coll_idx = idx;
	qsort(idx, (size_t)n, sizeof(int64_t), cmp_coll);
	w->coll = (int32_t *)malloc((size_t)(3 * n + 3) * sizeof(int32_t));
	w->coll_key = (uint64_t *)malloc((size_t)(n + 1) * sizeof(uint64_t));
	for (int64_t k = 0; k < n; k++) {
		memcpy(w->coll + 3 * k, coll + 3 * idx[k], 3 * sizeof(int32_t));
		w->coll_key[k] = keys[idx[k]];
	}
	w->ncoll = n;
	free(keys); free(idx);
}

static const int32_t *coll_find(oworld *w, uint64_t key)
{
	int64_t lo = 0, hi = w->ncoll;
	while (lo < hi) {
		int64_t mid = (lo + hi) / 2;
		if (w->coll_key[mid] < key) lo = mid + 1; else hi = mid;
	}
	return (lo < w->ncoll && w->coll_key[lo] == key) ? w->coll + 3 * lo : NULL;
}

void ow_set_next_order(oworld *w, int64_t n, const int32_t *pairs)
{
	free(w->next_pairs);
	w->next_pairs = (int32_t *)malloc((size_t)(2 * n + 2) * sizeof(int32_t));
	memcpy(w->next_pairs, pairs, (size_t)(2 * n) * sizeof(int32_t));
	w->next_n = n;
}

/* ------------------------------------------------------------------ */
/* integration: cpBodyUpdatePosition / SetTransform (cpBody.c:511-522, 347-357)
 * and cpBodyUpdateVelocity (cpBody.c:493-509)                          */

/* cpBodySanityCheck (cpBody.c:110-125): what a debug build of the reference
 * asserts of a body after each of its integrations (cpBody.c:508,521) */
static int fin(num x) { return isfinite(NUM_VAL(x)); }

static int body_sane(const OBody *b)
{
	return NUM_VAL(b->m) == NUM_VAL(b->m) && NUM_VAL(b->m_inv) == NUM_VAL(b->m_inv) &&
	       NUM_VAL(b->i) == NUM_VAL(b->i) && NUM_VAL(b->i_inv) == NUM_VAL(b->i_inv) &&
	       NUM_VAL(b->m) >= 0.0 && NUM_VAL(b->i) >= 0.0 &&
	       fin(b->p.x) && fin(b->p.y) && fin(b->v.x) && fin(b->v.y) && fin(b->f.x) && fin(b->f.y) &&
	       fin(b->a) && fin(b->w) && fin(b->t);
}

static void body_fault(oworld *w, int32_t code, int body)
{
	w->faults |= code;
	if (w->fault_body < 0 || body < w->fault_body) w->fault_body = body;
}

static void update_position(oworld *w, OBody *b, num dt)
{
	b->p = vadd(b->p, vmult(vadd(b->v, b->v_bias), dt));
	b->a = b->a + (b->w + b->w_bias) * dt;
	num sn, cs;
	if (w->libm_trig) { cs = cos(b->a); sn = sin(b->a); }
	else cp_sincos_oracle(b->a, &sn, &cs);
	V c = b->cog;
	/* cpTransformNewTranspose(rot.x, -rot.y, tx, rot.y, rot.x, ty) */
	b->T.a = cs; b->T.b = sn; b->T.c = -sn; b->T.d = cs;
	b->T.tx = b->p.x - (c.x * cs - c.y * sn);
	b->T.ty = b->p.y - (c.x * sn + c.y * cs);
	b->v_bias = v2(0.0, 0.0);
	b->w_bias = 0.0;
	if (!body_sane(b)) body_fault(w, CPGPU_FAULT_BODY_STATE, (int)(b - w->B));
}

static void update_velocity(oworld *w, OBody *b, V g, num damping, num dt)
{
	if (b->type == CPGPU_BODY_KINEMATIC) return;
	/* :498 "Body's mass and moment must be positive to simulate." */
	if (!(NUM_VAL(b->m) > 0.0 && NUM_VAL(b->i) > 0.0)) body_fault(w, CPGPU_FAULT_BODY_MASS, (int)(b - w->B));
	b->v = vadd(vmult(b->v, damping), vmult(vadd(g, vmult(b->f, b->m_inv)), dt));
	b->w = b->w * damping + b->t * b->i_inv * dt;
	b->f = v2(0.0, 0.0);
	b->t = 0.0;
	if (!body_sane(b)) body_fault(w, CPGPU_FAULT_BODY_STATE, (int)(b - w->B));
}

/* ------------------------------------------------------------------ */
/* shape caching: cpCircleShapeCacheData (cpShape.c:291-296),
 * cpSegmentShapeCacheData (cpShape.c:378-405), cpPolyShapeCacheData
 * (cpPolyShape.c:39-64)                                                 */

static void cache_shape(oworld *w, OShape *s)
{
	XF T = w->B[s->body].T;
	if (s->type == CPGPU_SHAPE_CIRCLE) {
		V c = s->tc = xpoint(T, s->c);
		s->bb[0] = c.x - s->r; s->bb[1] = c.y - s->r; s->bb[2] = c.x + s->r; s->bb[3] = c.y + s->r;
	} else if (s->type == CPGPU_SHAPE_SEGMENT) {
		s->ta = xpoint(T, s->a); s->tb = xpoint(T, s->b); s->tn = xvect(T, s->n);
		num l, r, bt, tp;
		if (s->ta.x < s->tb.x) { l = s->ta.x; r = s->tb.x; } else { l = s->tb.x; r = s->ta.x; }
		if (s->ta.y < s->tb.y) { bt = s->ta.y; tp = s->tb.y; } else { bt = s->tb.y; tp = s->ta.y; }
		num rad = s->r;
		s->bb[0] = l - rad; s->bb[1] = bt - rad; s->bb[2] = r + rad; s->bb[3] = tp + rad;
	} else {
		num l = INFINITY, r = -INFINITY, bt = INFINITY, tp = -INFINITY;
		for (int i = 0; i < s->count; i++) {
			V v = xpoint(T, s->lv[i]);
			V n = xvect(T, s->ln[i]);
			s->wv[i] = v; s->wn[i] = n;
			l = fminc(l, v.x); r = fmaxc(r, v.x); bt = fminc(bt, v.y); tp = fmaxc(tp, v.y);
		}
		num rad = s->r;
		s->bb[0] = l - rad; s->bb[1] = bt - rad; s->bb[2] = r + rad; s->bb[3] = tp + rad;
	}
}

/* GetBB (cpBBTree.c:82-98) with ShapeVelocityFunc (cpSpace.c:106): the leaf
 * box of a dynamic shape, and LeafUpdate's containment test (cpBBTree.c:
 * 498-517).  A rebuilt leaf clears its cached pairs, so the warm-start GJK
 * id of every pair of that shape restarts at 0 (cpBBTree.c:211,434). */
static void fat_bb(oworld *w, OShape *s)
{
	const num *bb = s->bb;
	V v = vmult(w->B[s->body].v, 0.1);
	num x = (bb[2] - bb[0]) * 0.1;
	num y = (bb[3] - bb[1]) * 0.1;
	s->fat[0] = bb[0] + fminc(-x, v.x);
	s->fat[1] = bb[1] + fminc(-y, v.y);
	s->fat[2] = bb[2] + fmaxc(x, v.x);
	s->fat[3] = bb[3] + fmaxc(y, v.y);
}

static void leaf_update(oworld *w, OShape *s)
{
	const num *f = s->fat, *b = s->bb;
	int contains = f[0] <= b[0] && f[2] >= b[2] && f[1] <= b[1] && f[3] >= b[3];
	if (!contains) {
		fat_bb(w, s);
		s->rebuild_stamp = w->W[s->world].sp.stamp;
	}
}

void ow_cache_all_shapes(oworld *w) { for (int i = 0; i < w->ns; i++) cache_shape(w, &w->S[i]); }

void ow_read_shape_bbs(oworld *w, double *bb)
{
	for (int i = 0; i < w->ns; i++)
		for (int q = 0; q < 4; q++) bb[4 * i + q] = NUM_VAL(w->S[i].bb[q]);
}

void ow_set_transforms(oworld *w, const double *t)
{
	for (int i = 0; i < w->nb; i++) w->B[i].T = xf_from(t + 6 * i);
}

/* ------------------------------------------------------------------ */
/* narrowphase: src/cpCollision.c                                      */

typedef struct { V p; uint32_t index; } SupportPt;                  /* :81-91 */
typedef struct { V a, b, ab; uint32_t id; } MinkPt;                  /* :117-132 */
typedef struct { const OShape *s1, *s2; } Ctx;                       /* :134-137 */
typedef struct { V p; uint64_t hash; } EdgePt;                       /* :147-151 */
typedef struct { EdgePt a, b; num r; V n; } Edge;                 /* :154-158 */
typedef struct { V a, b, n; num d; uint32_t id; } ClosestPts;     /* :211-221 */
typedef struct { int count; uint32_t id; V n; V p1[2], p2[2]; uint64_t hash[2]; } CollInfo;

static inline int poly_support_index(const OShape *s, V n)            /* :62-77 */
{
	num max = -INFINITY;
	int index = 0;
	for (int i = 0; i < s->count; i++) {
		num d = vdot(s->wv[i], n);
		if (d > max) { max = d; index = i; }
	}
	return index;
}

static inline SupportPt shape_support(const OShape *s, V n)          /* :95-114 */
{
	SupportPt r;
	if (s->type == CPGPU_SHAPE_CIRCLE) { r.p = s->tc; r.index = 0; }
	else if (s->type == CPGPU_SHAPE_SEGMENT) {
		if (vdot(s->ta, n) > vdot(s->tb, n)) { r.p = s->ta; r.index = 0; }
		else { r.p = s->tb; r.index = 1; }
	} else {
		int i = poly_support_index(s, n);
		r.p = s->wv[i]; r.index = (uint32_t)i;
	}
	return r;
}

static inline MinkPt mink_new(SupportPt a, SupportPt b)               /* :127-132 */
{
	MinkPt m;
	m.a = a.p; m.b = b.p; m.ab = vsub(b.p, a.p);
	m.id = ((a.index & 0xFF) << 8) | (b.index & 0xFF);
	return m;
}

static inline MinkPt support(const Ctx *ctx, V n)                     /* :140-145 */
{
	SupportPt a = shape_support(ctx->s1, vneg(n));
	SupportPt b = shape_support(ctx->s2, n);
	return mink_new(a, b);
}

static inline int check_point_greater(V a, V b, V c)                 /* cpRobust.c:4-8 */
{
	return (b.y - a.y) * (a.x + b.x - 2 * c.x) > (b.x - a.x) * (a.y + b.y - 2 * c.y);
}

static inline int check_axis(V v0, V v1, V p, V n)                    /* cpRobust.c:10-13 */
{
	return vdot(p, n) <= fmaxc(vdot(v0, n), vdot(v1, n));
}

static inline num closest_t(V a, V b)                              /* :193-199 */
{
	V delta = vsub(b, a);
	return -fclampc(vdot(delta, vadd(a, b)) / (vlengthsq(delta) + DBL_MIN), -1.0, 1.0);
}

static inline V lerp_t(V a, V b, num t)                            /* :201-207 */
{
	num ht = 0.5 * t;
	return vadd(vmult(a, 0.5 - ht), vmult(b, 0.5 + ht));
}

static ClosestPts closest_points_new(MinkPt v0, MinkPt v1)            /* :223-258 */
{
	num t = closest_t(v0.ab, v1.ab);
	V p = lerp_t(v0.ab, v1.ab, t);
	V pa = lerp_t(v0.a, v1.a, t);
	V pb = lerp_t(v0.b, v1.b, t);
	uint32_t id = ((v0.id & 0xFFFF) << 16) | (v1.id & 0xFFFF);
	V delta = vsub(v1.ab, v0.ab);
	V n = vnormalize(vrperp(delta));
	num d = vdot(n, p);
	ClosestPts r;
	r.a = pa; r.b = pb; r.id = id;
	if (d <= 0.0 || (-1.0 < t && t < 1.0)) { r.n = n; r.d = d; }
	else {
		num d2 = vlength(p);
		r.n = vmult(p, 1.0 / (d2 + DBL_MIN));
		r.d = d2;
	}
	return r;
}

static inline num closest_dist(V v0, V v1)                          /* :262-266 */
{
	return vlengthsq(lerp_t(v0, v1, closest_t(v0, v1)));
}

#define MAX_GJK_ITERATIONS 30
#define MAX_EPA_ITERATIONS 30
#define EPA_HULL_MAX 40

static ClosestPts epa(const Ctx *ctx, MinkPt v0, MinkPt v1, MinkPt v2)  /* :270-343 (loop form) */
{
	MinkPt hullA[EPA_HULL_MAX], hullB[EPA_HULL_MAX];
	MinkPt *hull = hullA, *hull2 = hullB;
	int count = 3;
	hull[0] = v0; hull[1] = v1; hull[2] = v2;
	for (int iteration = 1;; iteration++) {
		int mini = 0;
		num minDist = INFINITY;
		for (int j = 0, i = count - 1; j < count; i = j, j++) {
			num d = closest_dist(hull[i].ab, hull[j].ab);
			if (d < minDist) { minDist = d; mini = i; }
		}
		MinkPt e0 = hull[mini];
		MinkPt e1 = hull[(mini + 1) % count];
		MinkPt p = support(ctx, vperp(vsub(e1.ab, e0.ab)));
		int duplicate = (p.id == e0.id || p.id == e1.id);
		if (!duplicate && check_point_greater(e0.ab, e1.ab, p.ab) && iteration < MAX_EPA_ITERATIONS) {
			int count2 = 1;
			hull2[0] = p;
			for (int i = 0; i < count; i++) {
				int index = (mini + 1 + i) % count;
				V h0 = hull2[count2 - 1].ab;
				V h1 = hull[index].ab;
				V h2 = (i + 1 < count ? hull[(index + 1) % count] : p).ab;
				if (check_point_greater(h0, h2, h1)) hull2[count2++] = hull[index];
			}
			MinkPt *tmp = hull; hull = hull2; hull2 = tmp;
			count = count2;
		} else {
			return closest_points_new(e0, e1);
		}
	}
}

static ClosestPts gjk_recurse(const Ctx *ctx, MinkPt v0, MinkPt v1)    /* :348-392 (loop form) */
{
	int iteration = 1;
	for (;;) {
		if (iteration > MAX_GJK_ITERATIONS) return closest_points_new(v0, v1);
		if (check_point_greater(v1.ab, v0.ab, v2(0.0, 0.0))) {
			MinkPt tmp = v0; v0 = v1; v1 = tmp;
			continue;
		}
		num t = closest_t(v0.ab, v1.ab);
		V n = (-1.0 < t && t < 1.0) ? vperp(vsub(v1.ab, v0.ab)) : vneg(lerp_t(v0.ab, v1.ab, t));
		MinkPt p = support(ctx, n);
		if (check_point_greater(p.ab, v0.ab, v2(0.0, 0.0)) && check_point_greater(v1.ab, p.ab, v2(0.0, 0.0))) {
			return epa(ctx, v0, p, v1);
		}
		if (check_axis(v0.ab, v1.ab, p.ab, n)) return closest_points_new(v0, v1);
		if (closest_dist(v0.ab, p.ab) < closest_dist(p.ab, v1.ab)) v1 = p;
		else v0 = p;
		iteration++;
	}
}

static SupportPt shape_point(const OShape *s, int i)                   /* :395-413 */
{
	SupportPt r;
	if (s->type == CPGPU_SHAPE_CIRCLE) { r.p = s->tc; r.index = 0; }
	else if (s->type == CPGPU_SHAPE_SEGMENT) { r.p = (i == 0 ? s->ta : s->tb); r.index = (uint32_t)i; }
	else {
		int index = (i < s->count ? i : 0);
		r.p = s->wv[index]; r.index = (uint32_t)index;
	}
	return r;
}

static inline V bb_center(const num *bb) { return vlerp(v2(bb[0], bb[1]), v2(bb[2], bb[3]), 0.5); }

static ClosestPts gjk(const Ctx *ctx, uint32_t *id)                    /* :416-472 */
{
	MinkPt v0, v1;
	if (*id) {
		v0 = mink_new(shape_point(ctx->s1, (*id >> 24) & 0xFF), shape_point(ctx->s2, (*id >> 16) & 0xFF));
		v1 = mink_new(shape_point(ctx->s1, (*id >> 8) & 0xFF), shape_point(ctx->s2, (*id) & 0xFF));
	} else {
		V axis = vperp(vsub(bb_center(ctx->s1->bb), bb_center(ctx->s2->bb)));
		v0 = support(ctx, axis);
		v1 = support(ctx, vneg(axis));
	}
	ClosestPts points = gjk_recurse(ctx, v0, v1);
	*id = points.id;
	return points;
}

static inline void push_contact(CollInfo *info, V p1, V p2, uint64_t hash)  /* :44-55 */
{
	info->p1[info->count] = p1;
	info->p2[info->count] = p2;
	info->hash[info->count] = hash;
	info->count++;
}

static void contact_points(Edge e1, Edge e2, ClosestPts points, CollInfo *info)  /* :477-518 */
{
	num mindist = e1.r + e2.r;
	if (points.d <= mindist) {
		V n = info->n = points.n;
		num d_e1_a = vcross(e1.a.p, n);
		num d_e1_b = vcross(e1.b.p, n);
		num d_e2_a = vcross(e2.a.p, n);
		num d_e2_b = vcross(e2.b.p, n);
		num e1_denom = 1.0 / (d_e1_b - d_e1_a + DBL_MIN);
		num e2_denom = 1.0 / (d_e2_b - d_e2_a + DBL_MIN);
		{
			V p1 = vadd(vmult(n, e1.r), vlerp(e1.a.p, e1.b.p, fclamp01((d_e2_b - d_e1_a) * e1_denom)));
			V p2 = vadd(vmult(n, -e2.r), vlerp(e2.a.p, e2.b.p, fclamp01((d_e1_a - d_e2_a) * e2_denom)));
			num dist = vdot(vsub(p2, p1), n);
			if (dist <= 0.0) push_contact(info, p1, p2, hash_pair(e1.a.hash, e2.b.hash));
		}
		{
			V p1 = vadd(vmult(n, e1.r), vlerp(e1.a.p, e1.b.p, fclamp01((d_e2_a - d_e1_a) * e1_denom)));
			V p2 = vadd(vmult(n, -e2.r), vlerp(e2.a.p, e2.b.p, fclamp01((d_e1_b - d_e2_a) * e2_denom)));
			num dist = vdot(vsub(p2, p1), n);
			if (dist <= 0.0) push_contact(info, p1, p2, hash_pair(e1.b.hash, e2.a.hash));
		}
	}
}

static Edge support_edge_poly(const OShape *s, V n)                    /* :160-178 */
{
	int count = s->count;
	int i1 = poly_support_index(s, n);
	int i0 = (i1 - 1 + count) % count;
	int i2 = (i1 + 1) % count;
	uint64_t h = s->hashid;
	Edge e;
	e.r = s->r;
	if (vdot(n, s->wn[i1]) > vdot(n, s->wn[i2])) {
		e.a.p = s->wv[i0]; e.a.hash = hash_pair(h, (uint64_t)i0);
		e.b.p = s->wv[i1]; e.b.hash = hash_pair(h, (uint64_t)i1);
		e.n = s->wn[i1];
	} else {
		e.a.p = s->wv[i1]; e.a.hash = hash_pair(h, (uint64_t)i1);
		e.b.p = s->wv[i2]; e.b.hash = hash_pair(h, (uint64_t)i2);
		e.n = s->wn[i2];
	}
	return e;
}

static Edge support_edge_segment(const OShape *s, V n)                 /* :180-191 */
{
	uint64_t h = s->hashid;
	Edge e;
	e.r = s->r;
	if (vdot(s->tn, n) > 0.0) {
		e.a.p = s->ta; e.a.hash = hash_pair(h, 0);
		e.b.p = s->tb; e.b.hash = hash_pair(h, 1);
		e.n = s->tn;
	} else {
		e.a.p = s->tb; e.a.hash = hash_pair(h, 1);
		e.b.p = s->ta; e.b.hash = hash_pair(h, 0);
		e.n = vneg(s->tn);
	}
	return e;
}

static inline V body_rot(oworld *w, const OShape *s) { XF T = w->B[s->body].T; return v2(T.a, T.b); }

static void circle_circle(const OShape *c1, const OShape *c2, CollInfo *info)  /* :525-537 */
{
	num mindist = c1->r + c2->r;
	V delta = vsub(c2->tc, c1->tc);
	num distsq = vlengthsq(delta);
	if (distsq < mindist * mindist) {
		num dist = sqrt(distsq);
		V n = info->n = (dist != 0.0 ? vmult(delta, 1.0 / dist) : v2(1.0, 0.0));
		push_contact(info, vadd(c1->tc, vmult(n, c1->r)), vadd(c2->tc, vmult(n, -c2->r)), 0);
	}
}

static void circle_segment(oworld *w, const OShape *circle, const OShape *seg, CollInfo *info)  /* :539-569 */
{
	V seg_a = seg->ta, seg_b = seg->tb, center = circle->tc;
	V seg_delta = vsub(seg_b, seg_a);
	num closest_tt = fclamp01(vdot(seg_delta, vsub(center, seg_a)) / vlengthsq(seg_delta));
	V closest = vadd(seg_a, vmult(seg_delta, closest_tt));
	num mindist = circle->r + seg->r;
	V delta = vsub(closest, center);
	num distsq = vlengthsq(delta);
	if (distsq < mindist * mindist) {
		num dist = sqrt(distsq);
		V n = info->n = (dist != 0.0 ? vmult(delta, 1.0 / dist) : seg->tn);
		V rot = body_rot(w, seg);
		if ((closest_tt != 0.0 || vdot(n, vrotate(seg->a_tan, rot)) >= 0.0) &&
		    (closest_tt != 1.0 || vdot(n, vrotate(seg->b_tan, rot)) >= 0.0)) {
			push_contact(info, vadd(center, vmult(n, circle->r)), vadd(closest, vmult(n, -seg->r)), 0);
		}
	}
}

static void segment_segment(oworld *w, const OShape *s1, const OShape *s2, CollInfo *info)  /* :571-604 */
{
	Ctx ctx = {s1, s2};
	ClosestPts points = gjk(&ctx, &info->id);
	V n = points.n;
	V rot1 = body_rot(w, s1), rot2 = body_rot(w, s2);
	if (points.d <= (s1->r + s2->r) &&
	    (!veql(points.a, s1->ta) || vdot(n, vrotate(s1->a_tan, rot1)) <= 0.0) &&
	    (!veql(points.a, s1->tb) || vdot(n, vrotate(s1->b_tan, rot1)) <= 0.0) &&
	    (!veql(points.b, s2->ta) || vdot(n, vrotate(s2->a_tan, rot2)) >= 0.0) &&
	    (!veql(points.b, s2->tb) || vdot(n, vrotate(s2->b_tan, rot2)) >= 0.0)) {
		contact_points(support_edge_segment(s1, n), support_edge_segment(s2, vneg(n)), points, info);
	}
}

static void poly_poly(const OShape *p1, const OShape *p2, CollInfo *info)  /* :606-627 */
{
	Ctx ctx = {p1, p2};
	ClosestPts points = gjk(&ctx, &info->id);
	if (points.d - p1->r - p2->r <= 0.0) {
		contact_points(support_edge_poly(p1, points.n), support_edge_poly(p2, vneg(points.n)), points, info);
	}
}

static void segment_poly(oworld *w, const OShape *seg, const OShape *poly, CollInfo *info)  /* :629-659 */
{
	Ctx ctx = {seg, poly};
	ClosestPts points = gjk(&ctx, &info->id);
	V n = points.n;
	V rot = body_rot(w, seg);
	if (points.d - seg->r - poly->r <= 0.0 &&
	    (!veql(points.a, seg->ta) || vdot(n, vrotate(seg->a_tan, rot)) <= 0.0) &&
	    (!veql(points.a, seg->tb) || vdot(n, vrotate(seg->b_tan, rot)) <= 0.0)) {
		contact_points(support_edge_segment(seg, n), support_edge_poly(poly, vneg(n)), points, info);
	}
}

static void circle_poly(const OShape *circle, const OShape *poly, CollInfo *info)  /* :661-679 */
{
	Ctx ctx = {circle, poly};
	ClosestPts points = gjk(&ctx, &info->id);
	if (points.d <= circle->r + poly->r) {
		V n = info->n = points.n;
		push_contact(info, vadd(points.a, vmult(n, circle->r)), vadd(points.b, vmult(n, -poly->r)), 0);
	}
}

/* cpCollide (cpCollision.c:701-726): a and b already type-ordered by the caller. */
static void collide(oworld *w, int32_t ia, int32_t ib, uint32_t id, CollInfo *info)
{
	const OShape *a = &w->S[ia], *b = &w->S[ib];
	info->count = 0;
	info->id = id;
	info->n = v2(0.0, 0.0);
	int code = a->type + b->type * 3;
	switch (code) {
	case 0: circle_circle(a, b, info); break;
	case 3: circle_segment(w, a, b, info); break;
	case 4: segment_segment(w, a, b, info); break;
	case 6: circle_poly(a, b, info); break;
	case 7: segment_poly(w, a, b, info); break;
	case 8: poly_poly(a, b, info); break;
	default: fprintf(stderr, "oracle: shape types not sorted\n"); abort();
	}
}

int ow_collide(oworld *w, int32_t a, int32_t b, uint32_t id, int32_t *ab_out, uint32_t *id_out,
               double *n_out, double *points_out, uint64_t *hash_out)
{
	if (w->S[a].type > w->S[b].type) { int32_t t = a; a = b; b = t; }
	CollInfo info;
	collide(w, a, b, id, &info);
	ab_out[0] = a; ab_out[1] = b;
	*id_out = info.id;
	n_out[0] = NUM_VAL(info.n.x); n_out[1] = NUM_VAL(info.n.y);
	for (int i = 0; i < info.count; i++) {
		points_out[4 * i + 0] = NUM_VAL(info.p1[i].x); points_out[4 * i + 1] = NUM_VAL(info.p1[i].y);
		points_out[4 * i + 2] = NUM_VAL(info.p2[i].x); points_out[4 * i + 3] = NUM_VAL(info.p2[i].y);
		hash_out[i] = info.hash[i];
	}
	return info.count;
}

/* cpShapesCollide (cpShape.c:260-283): cpCollide with id 0 on the pair
 * sorted by type (cpCollision.c:707-710); when that swapped the shapes the
 * normal is negated and each point pair is given back in the caller's order;
 * distance = (r2 - r1) . normal with the unswapped points. */
int ow_shapes_collide(oworld *w, int32_t a, int32_t b, double *n_out, double *points_out)
{
	int32_t x = a, y = b;
	if (w->S[a].type > w->S[b].type) { x = b; y = a; }
	const int swapped = x != a;
	CollInfo info;
	collide(w, x, y, 0, &info);
	const V n = swapped ? vneg(info.n) : info.n;
	n_out[0] = NUM_VAL(n.x); n_out[1] = NUM_VAL(n.y);
	for (int i = 0; i < info.count; i++) {
		const V p1 = info.p1[i], p2 = info.p2[i];
		const V pa = swapped ? p2 : p1, pb = swapped ? p1 : p2;
		points_out[5 * i + 0] = NUM_VAL(pa.x); points_out[5 * i + 1] = NUM_VAL(pa.y);
		points_out[5 * i + 2] = NUM_VAL(pb.x); points_out[5 * i + 3] = NUM_VAL(pb.y);
		points_out[5 * i + 4] = NUM_VAL(vdot(vsub(p2, p1), n));
	}
	return info.count;
}

/* ------------------------------------------------------------------ */
/* broadphase semantics: the exact QueryReject predicate
 * (cpSpaceStep.c:204-232, cpShapeFilterReject chipmunk_private.h:144-155)
 * over dynamic x dynamic and dynamic x static shapes (cpBBTree.c:638-654,
 * cpSpatialIndex.c:61-68), by sort-and-sweep on the left edge.          */

static inline int bb_intersects(const num *a, const num *b)   /* cpBB.h:58-61 */
{
	return a[0] <= b[2] && b[0] <= a[2] && a[1] <= b[3] && b[1] <= a[3];
}

static inline int filter_reject(const OShape *a, const OShape *b)
{
	return (a->group != 0 && a->group == b->group) || (a->cats & b->mask) == 0 || (b->cats & a->mask) == 0;
}

static int nocollide_has(const OWorldPart *P, int ba, int bb)
{
	uint64_t k = pair_key(ba, bb);
	int64_t lo = 0, hi = P->n_nocollide;
	while (lo < hi) {
		int64_t mid = (lo + hi) / 2;
		if (P->nocollide[mid] < k) lo = mid + 1; else hi = mid;
	}
	return lo < P->n_nocollide && P->nocollide[lo] == k;
}

/* Pair admitted by the device library: QueryReject passes, not both bodies
 * of infinite mass and neither shape a sensor (cpSpaceStep.c:268-281; with
 * the default handler such pairs never reach the solver, so they are dropped
 * before the narrowphase). */
static int pair_ok(oworld *w, const OWorldPart *P, int i, int j)
{
	const OShape *a = &w->S[i], *b = &w->S[j];
	const OBody *ba = &w->B[a->body], *bb = &w->B[b->body];
	if (ba->type == CPGPU_BODY_STATIC && bb->type == CPGPU_BODY_STATIC) return 0;
	if (!bb_intersects(a->bb, b->bb)) return 0;
	if (a->body == b->body) return 0;
	if (filter_reject(a, b)) return 0;
	if (nocollide_has(P, a->body, b->body)) return 0;
	/* a pair that is never solved (cpSpaceStep.c:268-272) matters only to handlers
	 * that see it: begin, preSolve or separate */
	if ((ba->m == INFINITY && bb->m == INFINITY) || a->sensor || b->sensor)
		return w->n_handlers > 0 &&
		       (pair_events(w, a, b) & (CPGPU_EVENT_BEGIN | CPGPU_EVENT_PRE_SOLVE | CPGPU_EVENT_SEPARATE)) != 0;
	return 1;
}

/* the arbiter as a handler callback sees it */
static void cb_view(const oworld *w, int world, const OArb *arb, cpgpuCallbackArbiter *v)
{
	memset(v, 0, sizeof(*v));
	v->slot = -1;
	v->space = world;
	v->shape_a = arb->sa; v->shape_b = arb->sb;
	v->state = arb->state;
	v->events = (int32_t)arb->events;
	v->count = arb->count;
	v->normal[0] = NUM_VAL(arb->n.x); v->normal[1] = NUM_VAL(arb->n.y);
	v->e = NUM_VAL(arb->e); v->u = NUM_VAL(arb->u);
	v->surface_vr[0] = NUM_VAL(arb->svr.x); v->surface_vr[1] = NUM_VAL(arb->svr.y);
	for (int c = 0; c < arb->count; c++) {
		const OContact *con = &arb->con[c];
		double *o = v->contact[c];
		o[0] = NUM_VAL(con->r1.x); o[1] = NUM_VAL(con->r1.y); o[2] = NUM_VAL(con->r2.x); o[3] = NUM_VAL(con->r2.y);
		o[4] = NUM_VAL(con->jnAcc); o[5] = NUM_VAL(con->jtAcc);
		v->hash[c] = con->hash;
	}
	(void)w;
}

static const num *g_sort_bbs;
static int cmp_by_left(const void *x, const void *y)
{
	int i = *(const int *)x, j = *(const int *)y;
	num a = g_sort_bbs[4 * i], b = g_sort_bbs[4 * j];
	if (a < b) return -1;
	if (a > b) return 1;
	return i < j ? -1 : (i > j ? 1 : 0);
}

typedef struct { int32_t a, b; } IPair;
static int cmp_pair(const void *x, const void *y)
{
	const IPair *p = (const IPair *)x, *q = (const IPair *)y;
	if (p->a != q->a) return p->a < q->a ? -1 : 1;
	if (p->b != q->b) return p->b < q->b ? -1 : 1;
	return 0;
}

/* Candidate pairs of one world, (a < b), sorted. */
static IPair *world_pairs(oworld *w, OWorldPart *P, int64_t *n_out)
{
	int n = P->s1 - P->s0;
	num *bbs = (num *)malloc((size_t)(4 * (n ? n : 1)) * sizeof(num));
	int *idx = (int *)malloc((size_t)(n ? n : 1) * sizeof(int));
	int m = 0;
	for (int k = 0; k < n; k++) {
		const num *bb = w->S[P->s0 + k].bb;
		memcpy(bbs + 4 * k, bb, 4 * sizeof(num));
		/* a box with a NaN side overlaps nothing (cpBBIntersects compares
		 * every side), so it stays out of the sweep, whose order it would break */
		if (!isnan(NUM_VAL(bb[0])) && !isnan(NUM_VAL(bb[1])) && !isnan(NUM_VAL(bb[2])) && !isnan(NUM_VAL(bb[3])))
			idx[m++] = k;
	}
	n = m;
	g_sort_bbs = bbs;
	qsort(idx, (size_t)n, sizeof(int), cmp_by_left);
	int64_t cap = 1024, np = 0;
	IPair *pairs = (IPair *)malloc((size_t)cap * sizeof(IPair));
	for (int x = 0; x < n; x++) {
		int i = idx[x];
		num right = bbs[4 * i + 2];
		for (int y = x + 1; y < n; y++) {
			int j = idx[y];
			if (!(bbs[4 * j] <= right)) break;
			int gi = P->s0 + i, gj = P->s0 + j;
			if (!pair_ok(w, P, gi, gj)) continue;
			if (np == cap) { cap *= 2; pairs = (IPair *)realloc(pairs, (size_t)cap * sizeof(IPair)); }
			pairs[np].a = gi < gj ? gi : gj;
			pairs[np].b = gi < gj ? gj : gi;
			np++;
		}
	}
	qsort(pairs, (size_t)np, sizeof(IPair), cmp_pair);
	free(bbs); free(idx);
	*n_out = np;
	return pairs;
}

int64_t ow_candidate_pairs(oworld *w, int64_t cap, int32_t *out)
{
	int64_t total = 0;
	for (int wi = 0; wi < w->nw; wi++) {
		int64_t np;
		IPair *p = world_pairs(w, &w->W[wi], &np);
		for (int64_t k = 0; k < np; k++, total++) {
			if (total < cap) { out[2 * total] = p[k].a; out[2 * total + 1] = p[k].b; }
		}
		free(p);
	}
	return total;
}

/* ------------------------------------------------------------------ */
/* spatial queries: cpSpaceQuery.c over the shapes as the last step cached
 * them, with the shape queries of cpShape.c and cpPolyShape.c.  The
 * reference finds candidates through its spatial indexes and checks each
 * exactly; here every shape of the space is checked, which gives the same
 * results (the exact checks decide) -- except that a segment query with a
 * radius also finds shapes within the radius whose boxes the zero-width
 * segment misses, which the reference's BBTree prunes (cpBBTree.c
 * SubtreeSegmentQuery tests node boxes with cpBBSegmentQuery).  Of equal
 * nearest distances or first alphas the lowest shape index is kept. */

#define MAGIC_EPSILON 1e-5   /* chipmunk_private.h:32 */

typedef struct { int hit; V point; num d; V grad; } PQInfo;            /* cpPointQueryInfo */
typedef struct { int hit; V point, normal; num alpha; } SQInfo;        /* cpSegmentQueryInfo */

static inline V closest_on_segment(V p, V a, V b)                      /* chipmunk.h:182-188 */
{
	V delta = vsub(a, b);
	num t = fclamp01(vdot(delta, vsub(p, b)) / vlengthsq(delta));
	return vadd(b, vmult(delta, t));
}

/* cpShapePointQuery (cpShape.c:223-234) through the class functions:
 * circle cpShape.c:299-312, segment cpShape.c:408-424, poly cpPolyShape.c:67-103 */
static PQInfo point_query(const OShape *s, V p)
{
	PQInfo q;
	q.hit = 1;
	if (s->type == CPGPU_SHAPE_CIRCLE) {
		V delta = vsub(p, s->tc);
		num d = vlength(delta), r = s->r;
		num r_over_d = d > 0.0 ? r / d : r;
		q.point = vadd(s->tc, vmult(delta, r_over_d));
		q.d = d - r;
		q.grad = d > MAGIC_EPSILON ? vmult(delta, 1.0 / d) : v2(0.0, 1.0);
	} else if (s->type == CPGPU_SHAPE_SEGMENT) {
		V closest = closest_on_segment(p, s->ta, s->tb);
		V delta = vsub(p, closest);
		num d = vlength(delta), r = s->r;
		V g = vmult(delta, 1.0 / d);
		q.point = d != 0.0 ? vadd(closest, vmult(g, r)) : closest;
		q.d = d - r;
		q.grad = d > MAGIC_EPSILON ? g : s->n;   /* the segment's body-local normal, as the reference reads it */
	} else {
		int count = s->count;
		V v0 = s->wv[count - 1];
		num minDist = INFINITY;
		V closestPoint = v2(0.0, 0.0), closestNormal = v2(0.0, 0.0);
		int outside = 0;
		for (int i = 0; i < count; i++) {
			V v1 = s->wv[i];
			outside = outside || (vdot(s->wn[i], vsub(p, v1)) > 0.0);
			V closest = closest_on_segment(p, v0, v1);
			num dist = vlength(vsub(p, closest));
			if (dist < minDist) { minDist = dist; closestPoint = closest; closestNormal = s->wn[i]; }
			v0 = v1;
		}
		num dist = outside ? minDist : -minDist;
		V g = vmult(vsub(p, closestPoint), 1.0 / dist);
		q.point = vadd(closestPoint, vmult(g, s->r));
		q.d = dist - s->r;
		q.grad = minDist > MAGIC_EPSILON ? g : closestNormal;
	}
	return q;
}

/* CircleSegmentQuery (chipmunk_private.h:121-142) */
static void circle_segment_query(V center, num r1, V a, V b, num r2, SQInfo *info)
{
	V da = vsub(a, center), db = vsub(b, center);
	num rsum = r1 + r2;
	num qa = vdot(da, da) - 2.0 * vdot(da, db) + vdot(db, db);
	num qb = vdot(da, db) - vdot(da, da);
	num det = qb * qb - qa * (vdot(da, da) - rsum * rsum);
	if (det >= 0.0) {
		num t = (-qb - sqrt(det)) / qa;
		if (0.0 <= t && t <= 1.0) {
			V n = vnormalize(vlerp(da, db, t));
			info->hit = 1;
			info->point = vsub(vlerp(a, b, t), vmult(n, r2));
			info->normal = n;
			info->alpha = t;
		}
	}
}

static inline SQInfo sq_blank(V b) { SQInfo q; q.hit = 0; q.point = b; q.normal = v2(0.0, 0.0); q.alpha = 1.0; return q; }

/* cpShapeSegmentQuery (cpShape.c:238-257) through the class functions:
 * circle cpShape.c:315-318, segment cpShape.c:426-466, poly cpPolyShape.c:106-145 */
static SQInfo segment_query(const OShape *s, V a, V b, num radius)
{
	SQInfo info = sq_blank(b);
	PQInfo nearest = point_query(s, a);
	if (nearest.d <= radius) {
		info.hit = 1;
		info.alpha = 0.0;
		info.normal = vnormalize(vsub(a, nearest.point));
		return info;
	}
	if (s->type == CPGPU_SHAPE_CIRCLE) {
		circle_segment_query(s->tc, s->r, a, b, radius, &info);
	} else if (s->type == CPGPU_SHAPE_SEGMENT) {
		V n = s->tn;
		num d = vdot(vsub(s->ta, a), n);
		num r = s->r + radius;
		V flipped_n = d > 0.0 ? vneg(n) : n;
		V seg_offset = vsub(vmult(flipped_n, r), a);
		V seg_a = vadd(s->ta, seg_offset), seg_b = vadd(s->tb, seg_offset);
		V delta = vsub(b, a);
		if (vcross(delta, seg_a) * vcross(delta, seg_b) <= 0.0) {
			num d_offset = d + (d > 0.0 ? -r : r);
			num ad = -d_offset;
			num bd = vdot(delta, n) - d_offset;
			if (ad * bd < 0.0) {
				num t = ad / (ad - bd);
				info.hit = 1;
				info.point = vsub(vlerp(a, b, t), vmult(flipped_n, radius));
				info.normal = flipped_n;
				info.alpha = t;
			}
		} else if (r != 0.0) {
			SQInfo info1 = sq_blank(b), info2 = sq_blank(b);
			circle_segment_query(s->ta, s->r, a, b, radius, &info1);
			circle_segment_query(s->tb, s->r, a, b, radius, &info2);
			info = info1.alpha < info2.alpha ? info1 : info2;
		}
	} else {
		int count = s->count;
		num r = s->r, rsum = r + radius;
		for (int i = 0; i < count; i++) {
			V n = s->wn[i];
			num an = vdot(a, n);
			num d = an - vdot(s->wv[i], n) - rsum;
			if (d < 0.0) continue;
			num bn = vdot(b, n);
			num t = d / fmaxc(an - bn, DBL_MIN);
			if (t < 0.0 || 1.0 < t) continue;
			V point = vlerp(a, b, t);
			num dt = vcross(n, point);
			num dtMin = vcross(n, s->wv[(i - 1 + count) % count]);
			num dtMax = vcross(n, s->wv[i]);
			if (dtMin <= dt && dt <= dtMax) {
				info.hit = 1;
				info.point = vsub(vlerp(a, b, t), vmult(n, radius));
				info.normal = n;
				info.alpha = t;
			}
		}
		if (rsum > 0.0) {
			for (int i = 0; i < count; i++) {
				SQInfo circle_info = sq_blank(b);
				circle_segment_query(s->wv[i], r, a, b, radius, &circle_info);
				if (circle_info.alpha < info.alpha) info = circle_info;
			}
		}
	}
	return info;
}

static inline int query_filter_reject(const OShape *s, uint64_t group, uint32_t cats, uint32_t mask)
{                                                                      /* chipmunk_private.h:144-155 */
	return (s->group != 0 && s->group == group) || (s->cats & mask) == 0 || (cats & s->mask) == 0;
}

static const OShape *world_shapes(oworld *w, int world, int *s0, int *s1)
{
	*s0 = w->W[world].s0;
	*s1 = w->W[world].s1;
	return w->S;
}

/* cpSpacePointQueryNearest (cpSpaceQuery.c:72-98): the shape (-1: none) and
 * {point.x, point.y, distance, gradient.x, gradient.y} */
int32_t ow_point_query_nearest(oworld *w, int32_t world, const double *p, double max_distance, uint64_t group,
                               uint32_t cats, uint32_t mask, double *info)
{
	int s0, s1;
	const OShape *S = world_shapes(w, world, &s0, &s1);
	V pt = v2(p[0], p[1]);
	int32_t best = -1;
	PQInfo out;
	out.hit = 0; out.point = v2(0.0, 0.0); out.d = max_distance; out.grad = v2(0.0, 0.0);
	for (int i = s0; i < s1; i++) {
		if (query_filter_reject(&S[i], group, cats, mask) || S[i].sensor) continue;   /* cpSpaceQuery.c:64 */
		PQInfo q = point_query(&S[i], pt);
		if (q.d < out.d) { out = q; best = i; }
	}
	info[0] = NUM_VAL(out.point.x); info[1] = NUM_VAL(out.point.y); info[2] = NUM_VAL(out.d);
	info[3] = NUM_VAL(out.grad.x); info[4] = NUM_VAL(out.grad.y);
	return best;
}

/* cpSpaceSegmentQueryFirst (cpSpaceQuery.c:154-175): the shape (-1: none)
 * and {point.x, point.y, normal.x, normal.y, alpha} */
int32_t ow_segment_query_first(oworld *w, int32_t world, const double *a, const double *b, double radius,
                               uint64_t group, uint32_t cats, uint32_t mask, double *info)
{
	int s0, s1;
	const OShape *S = world_shapes(w, world, &s0, &s1);
	V va = v2(a[0], a[1]), vb = v2(b[0], b[1]);
	int32_t best = -1;
	SQInfo out = sq_blank(vb);
	for (int i = s0; i < s1; i++) {
		if (query_filter_reject(&S[i], group, cats, mask) || S[i].sensor) continue;   /* cpSpaceQuery.c:145 */
		SQInfo q = segment_query(&S[i], va, vb, radius);
		if (q.hit && q.alpha < out.alpha) { out = q; best = i; }
	}
	info[0] = NUM_VAL(out.point.x); info[1] = NUM_VAL(out.point.y); info[2] = NUM_VAL(out.normal.x);
	info[3] = NUM_VAL(out.normal.y); info[4] = NUM_VAL(out.alpha);
	return best;
}

/* cpSpacePointQuery (cpSpaceQuery.c:33-57): every shape closer than
 * max_distance, in index order, with {point, distance, gradient}; returns
 * the count (up to cap written) */
int64_t ow_point_query(oworld *w, int32_t world, const double *p, double max_distance, uint64_t group, uint32_t cats,
                       uint32_t mask, int64_t cap, int32_t *shapes, double *info)
{
	int s0, s1;
	const OShape *S = world_shapes(w, world, &s0, &s1);
	V pt = v2(p[0], p[1]);
	int64_t n = 0;
	for (int i = s0; i < s1; i++) {
		if (query_filter_reject(&S[i], group, cats, mask)) continue;
		PQInfo q = point_query(&S[i], pt);
		if (!(q.d < max_distance)) continue;
		if (n < cap) {
			shapes[n] = i;
			double *o = info + 5 * n;
			o[0] = NUM_VAL(q.point.x); o[1] = NUM_VAL(q.point.y); o[2] = NUM_VAL(q.d);
			o[3] = NUM_VAL(q.grad.x); o[4] = NUM_VAL(q.grad.y);
		}
		n++;
	}
	return n;
}

/* cpSpaceSegmentQuery (cpSpaceQuery.c:107-135): every shape the segment hits */
int64_t ow_segment_query(oworld *w, int32_t world, const double *a, const double *b, double radius, uint64_t group,
                         uint32_t cats, uint32_t mask, int64_t cap, int32_t *shapes, double *info)
{
	int s0, s1;
	const OShape *S = world_shapes(w, world, &s0, &s1);
	V va = v2(a[0], a[1]), vb = v2(b[0], b[1]);
	int64_t n = 0;
	for (int i = s0; i < s1; i++) {
		if (query_filter_reject(&S[i], group, cats, mask)) continue;
		SQInfo q = segment_query(&S[i], va, vb, radius);
		if (!q.hit) continue;
		if (n < cap) {
			shapes[n] = i;
			double *o = info + 5 * n;
			o[0] = NUM_VAL(q.point.x); o[1] = NUM_VAL(q.point.y); o[2] = NUM_VAL(q.normal.x);
			o[3] = NUM_VAL(q.normal.y); o[4] = NUM_VAL(q.alpha);
		}
		n++;
	}
	return n;
}

/* cpSpaceBBQuery (cpSpaceQuery.c:186-206): every shape whose box meets bb {l, b, r, t} */
int64_t ow_bb_query(oworld *w, int32_t world, const double *bb, uint64_t group, uint32_t cats, uint32_t mask,
                    int64_t cap, int32_t *shapes)
{
	int s0, s1;
	const OShape *S = world_shapes(w, world, &s0, &s1);
	int64_t n = 0;
	for (int i = s0; i < s1; i++) {
		if (query_filter_reject(&S[i], group, cats, mask)) continue;
		if (!(bb[0] <= S[i].bb[2] && S[i].bb[0] <= bb[2] && bb[1] <= S[i].bb[3] && S[i].bb[1] <= bb[3])) continue;
		if (n < cap) shapes[n] = i;
		n++;
	}
	return n;
}

/* cpSpaceShapeQuery (cpSpaceQuery.c:231-246) of shape s, every shape
 * checked: s is first put at its body's present transform (cpShapeUpdate,
 * cpShape.c:217-220; the other shapes keep their cached pose), then the shapes
 * of its space other than s whose box meets s's new box, that pass
 * cpShapeFilterReject(s->filter, shape->filter) (chipmunk_private.h:144-155)
 * and that touch it, in shape order, with cpShapesCollide(s, shape)'s count
 * and contact point set {normal, then per point pointA, pointB, distance}.
 * The reference's dynamic tree prunes by a shape's fattened leaf box instead
 * of its own; the two differ only where cpCollide reports a contact between
 * shapes whose boxes are apart, which degenerate input does (collinear
 * segments: a zero normal).  Unlike the reference, s's cached pose is given
 * back afterwards. */
int64_t ow_shape_query(oworld *w, int32_t s, int64_t cap, int32_t *shapes, int32_t *counts, double *sets)
{
	OShape *a = &w->S[s];
	const OShape saved = *a;
	V wv[CPGPU_MAX_POLY_VERTS], wn[CPGPU_MAX_POLY_VERTS];
	const int poly = a->type == CPGPU_SHAPE_POLY;
	if (poly)
		for (int i = 0; i < a->count; i++) { wv[i] = a->wv[i]; wn[i] = a->wn[i]; }
	cache_shape(w, a);
	int s0, s1;
	world_shapes(w, w->B[a->body].world, &s0, &s1);
	int64_t n = 0;
	for (int i = s0; i < s1; i++) {
		const OShape *b = &w->S[i];
		if (i == s || !bb_intersects(a->bb, b->bb)) continue;
		if ((a->group != 0 && a->group == b->group) || (a->cats & b->mask) == 0 || (b->cats & a->mask) == 0) continue;
		double set[12] = {0};
		const int m = ow_shapes_collide(w, s, i, set, set + 2);
		if (m == 0) continue;
		if (n < cap) {
			shapes[n] = i;
			counts[n] = m;
			for (int k = 0; k < 12; k++) sets[12 * n + k] = set[k];
		}
		n++;
	}
	if (poly)
		for (int i = 0; i < a->count; i++) { a->wv[i] = wv[i]; a->wn[i] = wn[i]; }
	*a = saved;
	return n;
}

/* ------------------------------------------------------------------ */
/* arbiter persistence: the cachedArbiters hash set of cpSpace
 * (cpSpaceStep.c:248-251, cpHashSet.c:152-176) keyed by the unordered
 * shape pair, eviction by cpSpaceArbiterSetFilter (cpSpaceStep.c:292-325). */

static void tab_rebuild(OWorldPart *P)
{
	int64_t need = 2 * P->narb + 16, cap = 128;
	while (cap < need) cap *= 2;
	if (cap - 1 != P->tabmask) { free(P->tab); P->tab = (int64_t *)malloc((size_t)cap * sizeof(int64_t)); P->tabmask = cap - 1; }
	for (int64_t k = 0; k < cap; k++) P->tab[k] = -1;
	for (int64_t i = 0; i < P->narb; i++) {
		uint64_t h = mix64(P->arbs[i].key) & (uint64_t)P->tabmask;
		while (P->tab[h] >= 0) h = (h + 1) & (uint64_t)P->tabmask;
		P->tab[h] = i;
	}
}

static int64_t tab_find(OWorldPart *P, uint64_t key)
{
	uint64_t h = mix64(key) & (uint64_t)P->tabmask;
	while (P->tab[h] >= 0) {
		if (P->arbs[P->tab[h]].key == key) return P->tab[h];
		h = (h + 1) & (uint64_t)P->tabmask;
	}
	return -1;
}

static int64_t arb_append(OWorldPart *P)
{
	if (P->narb == P->arbcap) {
		P->arbcap *= 2;
		P->arbs = (OArb *)realloc(P->arbs, (size_t)P->arbcap * sizeof(OArb));
		P->active = (int64_t *)realloc(P->active, (size_t)P->arbcap * sizeof(int64_t));
	}
	memset(&P->arbs[P->narb], 0, sizeof(OArb));
	return P->narb++;
}

/* Warm-start ids.  The reference caches the id returned by cpCollide in the
 * BBTree Pair of the two leaves; a leaf rebuilt by LeafUpdate drops its
 * pairs, so the pair restarts from id 0 (cpBBTree.c:206-223,434,504-511).
 * Here an id is kept per collided pair with the step it was written, and is
 * valid while neither shape's leaf has been rebuilt since. */
static void pid_rebuild(oworld *w, OWorldPart *P, int keep_all)
{
	int64_t n = 0;
	for (int64_t k = 0; k <= P->pid_mask && P->pid_key; k++) if (P->pid_key[k] != UINT64_MAX) n++;
	int64_t cap = 256;
	while (cap < 2 * n + 64) cap *= 2;
	uint64_t *ok = P->pid_key; uint32_t *oi = P->pid_id, *os = P->pid_stamp;
	int64_t omask = P->pid_mask;
	P->pid_key = (uint64_t *)malloc((size_t)cap * sizeof(uint64_t));
	P->pid_id = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
	P->pid_stamp = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
	for (int64_t k = 0; k < cap; k++) P->pid_key[k] = UINT64_MAX;
	P->pid_mask = cap - 1;
	P->pid_n = 0;
	for (int64_t k = 0; ok && k <= omask; k++) {
		if (ok[k] == UINT64_MAX) continue;
		int32_t a = (int32_t)(ok[k] >> 32), b = (int32_t)(ok[k] & 0xffffffffu);
		if (!keep_all && (os[k] < w->S[a].rebuild_stamp || os[k] < w->S[b].rebuild_stamp)) continue;
		uint64_t h = mix64(ok[k]) & (uint64_t)P->pid_mask;
		while (P->pid_key[h] != UINT64_MAX) h = (h + 1) & (uint64_t)P->pid_mask;
		P->pid_key[h] = ok[k]; P->pid_id[h] = oi[k]; P->pid_stamp[h] = os[k];
		P->pid_n++;
	}
	free(ok); free(oi); free(os);
}

static uint32_t pid_get(oworld *w, OWorldPart *P, uint64_t key, int32_t a, int32_t b)
{
	if (!P->pid_key) return 0;
	uint64_t h = mix64(key) & (uint64_t)P->pid_mask;
	while (P->pid_key[h] != UINT64_MAX) {
		if (P->pid_key[h] == key) {
			uint32_t st = P->pid_stamp[h];
			if (st < w->S[a].rebuild_stamp || st < w->S[b].rebuild_stamp) return 0;
			return P->pid_id[h];
		}
		h = (h + 1) & (uint64_t)P->pid_mask;
	}
	return 0;
}

static void pid_put(OWorldPart *P, uint64_t key, uint32_t id)
{
	uint64_t h = mix64(key) & (uint64_t)P->pid_mask;
	while (P->pid_key[h] != UINT64_MAX && P->pid_key[h] != key) h = (h + 1) & (uint64_t)P->pid_mask;
	if (P->pid_key[h] == UINT64_MAX) P->pid_n++;
	P->pid_key[h] = key; P->pid_id[h] = id; P->pid_stamp[h] = P->sp.stamp;
}

/* ------------------------------------------------------------------ */
/* deterministic graph colouring (the device solve order).
 * Conflict nodes are dynamic bodies only (static and kinematic bodies have
 * zero inverse mass and are never written).  Colour c is built in
 * COLOUR_SUBROUNDS claim rounds: every uncoloured element none of whose
 * dynamic bodies is taken by colour c claims those bodies with an atomic
 * min of its priority; elements holding all their claims take colour c and
 * mark their bodies taken.  The result is independent of processing order. */

#define COLOUR_SUBROUNDS 3

static int colour_elements(oworld *w, int64_t n, const int32_t *ba, const int32_t *bb, const uint64_t *prio,
                           int32_t *colour)
{
	if (n == 0) return 0;
	uint64_t *claim = (uint64_t *)malloc((size_t)w->nb * sizeof(uint64_t));
	int32_t *taken = (int32_t *)malloc((size_t)w->nb * sizeof(int32_t));
	for (int i = 0; i < w->nb; i++) taken[i] = -1;
	for (int64_t k = 0; k < n; k++) colour[k] = -1;
	int64_t left = n;
	int c = 0;
	while (left > 0) {
		for (int r = 0; r < COLOUR_SUBROUNDS; r++) {
			for (int64_t k = 0; k < n; k++) {
				if (ba[k] >= 0) claim[ba[k]] = UINT64_MAX;
				if (bb[k] >= 0) claim[bb[k]] = UINT64_MAX;
			}
			for (int64_t k = 0; k < n; k++) {
				if (colour[k] >= 0) continue;
				if ((ba[k] >= 0 && taken[ba[k]] == c) || (bb[k] >= 0 && taken[bb[k]] == c)) continue;
				if (ba[k] >= 0 && prio[k] < claim[ba[k]]) claim[ba[k]] = prio[k];
				if (bb[k] >= 0 && prio[k] < claim[bb[k]]) claim[bb[k]] = prio[k];
			}
			for (int64_t k = 0; k < n; k++) {
				if (colour[k] >= 0) continue;
				if ((ba[k] >= 0 && taken[ba[k]] == c) || (bb[k] >= 0 && taken[bb[k]] == c)) continue;
				if ((ba[k] < 0 || claim[ba[k]] == prio[k]) && (bb[k] < 0 || claim[bb[k]] == prio[k])) {
					colour[k] = c;
					left--;
				}
			}
			for (int64_t k = 0; k < n; k++) {
				if (colour[k] != c) continue;
				if (ba[k] >= 0) taken[ba[k]] = c;
				if (bb[k] >= 0) taken[bb[k]] = c;
			}
		}
		c++;
	}
	free(claim); free(taken);
	return c;
}

static OArb *g_sort_arbs;
static int cmp_solve(const void *x, const void *y)
{
	const OArb *a = &g_sort_arbs[*(const int64_t *)x], *b = &g_sort_arbs[*(const int64_t *)y];
	if (a->unit != b->unit) return a->unit < b->unit ? -1 : 1;
	if (a->colour != b->colour) return a->colour < b->colour ? -1 : 1;
	return a->key < b->key ? -1 : (a->key > b->key ? 1 : 0);
}
static OJoint *g_sort_joints;
static int cmp_jsolve(const void *x, const void *y)
{
	int32_t i = *(const int32_t *)x, j = *(const int32_t *)y;
	const OJoint *a = &g_sort_joints[i], *b = &g_sort_joints[j];
	if (a->colour != b->colour) return a->colour < b->colour ? -1 : 1;
	return i < j ? -1 : (i > j ? 1 : 0);
}

/* ------------------------------------------------------------------ */
/* the device's solve units of a large space (no reference counterpart).
 * A space with more than SMALL_WORLD_MAX_BODIES non-static bodies is solved
 * as one grid set: its dynamic bodies are cut into P strips along x
 * holding equal shares of them: a PART_BINS-bin histogram of x between the
 * bodies' least and greatest x, each bin going to strip
 * floor(P * (bodies in earlier bins) / (bodies)).  P is one count for every
 * large space of the description: about GRID_STRIP_BODIES dynamic bodies a
 * strip in the largest, between GRID_PARTS_MIN and GRID_PARTS_MAX.  An arbiter
 * belongs to the strip of its dynamic bodies, to seam q when they lie in
 * strips q and q + 1, or to the wide set when their strips are further apart.
 * Each unit is coloured on its own and the space is solved unit by unit: the
 * strips (1 + q), the even seams (1 + P + q / 2), the odd seams
 * (1 + P + P / 2 + q / 2), the wide set (1 + P + P / 2 + (P - 1) / 2).  A small
 * space is one unit, 0. */

#define SMALL_WORLD_MAX_BODIES 4096
#define GRID_PARTS_MIN 64
#define GRID_PARTS_MAX 256
#define GRID_STRIP_BODIES 384
#define PART_BINS 1024
#define UNIT_EVEN(P) (1 + (P))
#define UNIT_ODD(P) (UNIT_EVEN(P) + (P) / 2)
#define UNIT_WIDE(P) (UNIT_ODD(P) + ((P) - 1) / 2)

/* strips of every large space of the description, from its largest grid set */
static int grid_parts_of(const oworld *w)
{
	long max_grid = 0;
	for (int wi = 0; wi < w->nw; wi++) {
		const OWorldPart *P = &w->W[wi];
		long n_ns = 0, n_dyn = 0;
		for (int b = P->b0; b < P->b1; b++) {
			n_ns += w->B[b].type != CPGPU_BODY_STATIC;
			n_dyn += w->B[b].type == CPGPU_BODY_DYNAMIC;
		}
		if (n_ns > SMALL_WORLD_MAX_BODIES && n_dyn > max_grid) max_grid = n_dyn;
	}
	long p = (max_grid + GRID_STRIP_BODIES - 1) / GRID_STRIP_BODIES;
	return (int)(p < GRID_PARTS_MIN ? GRID_PARTS_MIN : (p > GRID_PARTS_MAX ? GRID_PARTS_MAX : p));
}

/* a num's bits as an unsigned key of the same order (x + 0.0 turns -0.0 into +0.0) */
static uint64_t x_key(double x)
{
	double y = x + 0.0;
	uint64_t u;
	memcpy(&u, &y, sizeof u);
	return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
static double key_x(uint64_t k)
{
	uint64_t u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
	double x;
	memcpy(&x, &u, sizeof x);
	return x;
}
static int part_bin(double x, double lo, double hi)
{
	if (!(hi > lo)) return 0;
	double t = (x - lo) * ((double)PART_BINS / (hi - lo));
	if (!(t >= 0.0)) return 0;
	return t >= (double)PART_BINS ? PART_BINS - 1 : (int)t;
}

/* unit of every active arbiter (OArb.unit); returns 0 for a small space (all units 0) */
static int grid_units(oworld *w, OWorldPart *P)
{
	int n_ns = 0;
	for (int b = P->b0; b < P->b1; b++) n_ns += w->B[b].type != CPGPU_BODY_STATIC;
	for (int64_t k = 0; k < P->nactive; k++) P->arbs[P->active[k]].unit = 0;
	if (n_ns <= SMALL_WORLD_MAX_BODIES) return 0;
	const int nbw = P->b1 - P->b0;
	int32_t *part = (int32_t *)malloc((size_t)nbw * sizeof(int32_t));
	int32_t hist[PART_BINS];
#define DYN(b) (w->B[b].type == CPGPU_BODY_DYNAMIC)
	/* the strips of the dynamic bodies */
	uint64_t lo_k = UINT64_MAX, hi_k = 0;
	int n_grid = 0;
	for (int i = 0; i < nbw; i++) {
		part[i] = -1;
		if (!DYN(P->b0 + i)) continue;
		part[i] = 0;
		uint64_t key = x_key(NUM_VAL(w->B[P->b0 + i].p.x));
		if (key < lo_k) lo_k = key;
		if (key > hi_k) hi_k = key;
		n_grid++;
	}
	double lo = key_x(lo_k), hi = key_x(hi_k);
	memset(hist, 0, sizeof hist);
	for (int i = 0; i < nbw; i++)
		if (part[i] >= 0) hist[part_bin(NUM_VAL(w->B[P->b0 + i].p.x), lo, hi)]++;
	for (int q = 0, before = 0; q < PART_BINS; q++) {
		int cnt = hist[q];
		hist[q] = n_grid > 0 ? (int)((int64_t)before * w->grid_parts / n_grid) : 0;
		before += cnt;
	}
	for (int i = 0; i < nbw; i++)
		if (part[i] >= 0) part[i] = hist[part_bin(NUM_VAL(w->B[P->b0 + i].p.x), lo, hi)];
	for (int64_t k = 0; k < P->nactive; k++) {
		OArb *arb = &P->arbs[P->active[k]];
		int pa = DYN(arb->ba) ? part[arb->ba - P->b0] : -1, pb = DYN(arb->bb) ? part[arb->bb - P->b0] : -1;
		if (pa < 0 || pb < 0 || pa == pb) arb->unit = 1 + (pa > pb ? pa : pb);
		else if (pa - pb == 1 || pb - pa == 1) {
			int q = pa < pb ? pa : pb;
			arb->unit = (q % 2 == 0 ? UNIT_EVEN(w->grid_parts) : UNIT_ODD(w->grid_parts)) + q / 2;
		} else arb->unit = UNIT_WIDE(w->grid_parts);
	}
#undef DYN
	free(part);
	return 1;
}

/* ------------------------------------------------------------------ */
/* contact solver: cpArbiterPreStep / ApplyCachedImpulse / ApplyImpulse
 * (cpArbiter.c:416-498) with the helpers of chipmunk_private.h:172-226   */

static inline V rel_vel(const OBody *a, const OBody *b, V r1, V r2)   /* :172-178 */
{
	V v1 = vadd(a->v, vmult(vperp(r1), a->w));
	V v2s = vadd(b->v, vmult(vperp(r2), b->w));
	return vsub(v2s, v1);
}

static inline void apply_impulse(OBody *body, V j, V r)                /* :186-190 */
{
	body->v = vadd(body->v, vmult(j, body->m_inv));
	body->w += body->i_inv * vcross(r, j);
}

static inline void apply_impulses(OBody *a, OBody *b, V r1, V r2, V j)  /* :192-197 */
{
	apply_impulse(a, vneg(j), r1);
	apply_impulse(b, j, r2);
}

static inline void apply_bias_impulse(OBody *body, V j, V r)           /* :199-203 */
{
	body->v_bias = vadd(body->v_bias, vmult(j, body->m_inv));
	body->w_bias += body->i_inv * vcross(r, j);
}

static inline num k_scalar_body(const OBody *b, V r, V n)           /* :212-217 */
{
	num rcn = vcross(r, n);
	return b->m_inv + b->i_inv * rcn * rcn;
}

/* :219-226, with its "Unsolvable collision or constraint." check (:223) */
static inline num k_scalar(oworld *w, const OBody *a, const OBody *b, V r1, V r2, V n)
{
	num value = k_scalar_body(a, r1, n) + k_scalar_body(b, r2, n);
	if (NUM_VAL(value) == 0.0) w->faults |= CPGPU_FAULT_UNSOLVABLE;
	return value;
}

static void arb_prestep(oworld *w, OArb *arb, num dt, num slop, num bias)
{
	const OBody *a = &w->B[arb->ba], *b = &w->B[arb->bb];
	V n = arb->n;
	V body_delta = vsub(b->p, a->p);
	for (int i = 0; i < arb->count; i++) {
		OContact *con = &arb->con[i];
		con->nMass = 1.0 / k_scalar(w, a, b, con->r1, con->r2, n);
		con->tMass = 1.0 / k_scalar(w, a, b, con->r1, con->r2, vperp(n));
		num dist = vdot(vadd(vsub(con->r2, con->r1), body_delta), n);
		con->bias = -bias * fminc(0.0, dist + slop) / dt;
		con->jBias = 0.0;
		con->bounce = vdot(rel_vel(a, b, con->r1, con->r2), n) * arb->e;
	}
}

static void arb_apply_cached(oworld *w, OArb *arb, num dt_coef)
{
	if (arb->state == CPGPU_ARB_FIRST_COLLISION) return;
	OBody *a = &w->B[arb->ba], *b = &w->B[arb->bb];
	V n = arb->n;
	for (int i = 0; i < arb->count; i++) {
		OContact *con = &arb->con[i];
		V j = vrotate(n, v2(con->jnAcc, con->jtAcc));
		apply_impulses(a, b, con->r1, con->r2, vmult(j, dt_coef));
	}
}

static void arb_apply_impulse(oworld *w, OArb *arb)
{
	OBody *a = &w->B[arb->ba], *b = &w->B[arb->bb];
	V n = arb->n;
	V surface_vr = arb->svr;
	num friction = arb->u;
	for (int i = 0; i < arb->count; i++) {
		OContact *con = &arb->con[i];
		num nMass = con->nMass;
		V r1 = con->r1, r2 = con->r2;
		V vb1 = vadd(a->v_bias, vmult(vperp(r1), a->w_bias));
		V vb2 = vadd(b->v_bias, vmult(vperp(r2), b->w_bias));
		V vr = vadd(rel_vel(a, b, r1, r2), surface_vr);
		num vbn = vdot(vsub(vb2, vb1), n);
		num vrn = vdot(vr, n);
		num vrt = vdot(vr, vperp(n));
		num jbn = (con->bias - vbn) * nMass;
		num jbnOld = con->jBias;
		con->jBias = fmaxc(jbnOld + jbn, 0.0);
		num jn = -(con->bounce + vrn) * nMass;
		num jnOld = con->jnAcc;
		con->jnAcc = fmaxc(jnOld + jn, 0.0);
		num jtMax = friction * con->jnAcc;
		num jt = -vrt * con->tMass;
		num jtOld = con->jtAcc;
		con->jtAcc = fclampc(jtOld + jt, -jtMax, jtMax);
		V jb = vmult(n, con->jBias - jbnOld);
		apply_bias_impulse(a, vneg(jb), r1);
		apply_bias_impulse(b, jb, r2);
		apply_impulses(a, b, r1, r2, vrotate(n, v2(con->jnAcc - jnOld, con->jtAcc - jtOld)));
	}
}

/* ------------------------------------------------------------------ */
/* joints: src/cpPivotJoint.c:24-70, cpDampedSpring.c:24-77,
 * cpSlideJoint.c:24-89, cpGearJoint.c:24-69, cpPinJoint.c:24-71,
 * cpGrooveJoint.c:24-93, cpRotaryLimitJoint.c:24-83, cpRatchetJoint.c:24-89,
 * cpSimpleMotor.c:24-67, cpDampedRotarySpring.c:24-72; bias_coef
 * chipmunk_private.h:264-268 (libm pow, evaluated on the host);
 * k_tensor chipmunk_private.h:228-262                                   */

static num bias_coef_host(num errorBias, num dt) { return 1.0 - pow(NUM_VAL(errorBias), NUM_VAL(dt)); }

/* k_tensor (chipmunk_private.h:228-262): the inverse of the 2x2 effective
 * mass, with its "Unsolvable constraint." check (:255) */
static void k_tensor(oworld *w, const OBody *a, const OBody *b, V r1, V r2, OJoint *j)
{
	num m_sum = a->m_inv + b->m_inv;
	num k11 = m_sum, k12 = 0.0, k21 = 0.0, k22 = m_sum;
	num a_i_inv = a->i_inv;
	num r1xsq = r1.x * r1.x * a_i_inv;
	num r1ysq = r1.y * r1.y * a_i_inv;
	num r1nxy = -r1.x * r1.y * a_i_inv;
	k11 += r1ysq; k12 += r1nxy; k21 += r1nxy; k22 += r1xsq;
	num b_i_inv = b->i_inv;
	num r2xsq = r2.x * r2.x * b_i_inv;
	num r2ysq = r2.y * r2.y * b_i_inv;
	num r2nxy = -r2.x * r2.y * b_i_inv;
	k11 += r2ysq; k12 += r2nxy; k21 += r2nxy; k22 += r2xsq;
	num det = k11 * k22 - k12 * k21;
	if (NUM_VAL(det) == 0.0) w->faults |= CPGPU_FAULT_UNSOLVABLE;
	num det_inv = 1.0 / det;
	j->k11 = k22 * det_inv; j->k12 = -k12 * det_inv;
	j->k21 = -k21 * det_inv; j->k22 = k11 * det_inv;
}

/* the angular impulse j between a and b (rotary limit, ratchet, simple motor:
 * a->w -= j a->i_inv, b->w += j b->i_inv) */
static inline void apply_angular(OBody *a, OBody *b, num jj)
{
	a->w -= jj * a->i_inv;
	b->w += jj * b->i_inv;
}

static void joint_prestep(oworld *w, OJoint *j, num dt)
{
	OBody *a = &w->B[j->a], *b = &w->B[j->b];
	V anchorA = v2(j->P[0], j->P[1]), anchorB = v2(j->P[2], j->P[3]);
	if (j->type == CPGPU_JOINT_PIVOT) {
		j->r1 = xvect(a->T, vsub(anchorA, a->cog));
		j->r2 = xvect(b->T, vsub(anchorB, b->cog));
		k_tensor(w, a, b, j->r1, j->r2, j);
		V delta = vsub(vadd(b->p, j->r2), vadd(a->p, j->r1));
		j->bias2 = vclamp(vmult(delta, -bias_coef_host(j->errorBias, dt) / dt), j->maxBias);
	} else if (j->type == CPGPU_JOINT_DAMPED_SPRING) {
		num restLength = j->P[4], stiffness = j->P[5], damping = j->P[6];
		j->r1 = xvect(a->T, vsub(anchorA, a->cog));
		j->r2 = xvect(b->T, vsub(anchorB, b->cog));
		V delta = vsub(vadd(b->p, j->r2), vadd(a->p, j->r1));
		num dist = vlength(delta);
		j->n = vmult(delta, 1.0 / (dist != 0.0 ? dist : INFINITY));
		num k = k_scalar(w, a, b, j->r1, j->r2, j->n);
		j->nMass = 1.0 / k;
		j->target_vrn = 0.0;
		j->v_coef = 1.0 - (w->libm_trig ? exp(-damping * dt * k) : cp_exp_oracle(-damping * dt * k));
		num f_spring = (restLength - dist) * stiffness;
		num j_spring = j->jAcc = f_spring * dt;
		apply_impulses(a, b, j->r1, j->r2, vmult(j->n, j_spring));
	} else if (j->type == CPGPU_JOINT_SLIDE) {
		num mn = j->P[4], mx = j->P[5];
		j->r1 = xvect(a->T, vsub(anchorA, a->cog));
		j->r2 = xvect(b->T, vsub(anchorB, b->cog));
		V delta = vsub(vadd(b->p, j->r2), vadd(a->p, j->r1));
		num dist = vlength(delta);
		num pdist = 0.0;
		if (dist > mx) { pdist = dist - mx; j->n = vnormalize(delta); }
		else if (dist < mn) { pdist = mn - dist; j->n = vneg(vnormalize(delta)); }
		else { j->n = v2(0.0, 0.0); j->jnAcc = 0.0; }
		j->nMass = 1.0 / k_scalar(w, a, b, j->r1, j->r2, j->n);
		num maxBias = j->maxBias;
		j->bias = fclampc(-bias_coef_host(j->errorBias, dt) * pdist / dt, -maxBias, maxBias);
	} else if (j->type == CPGPU_JOINT_GEAR) {
		num phase = j->P[0], ratio = j->P[1], ratio_inv = j->P[2];
		j->iSum = 1.0 / (a->i_inv * ratio_inv + ratio * b->i_inv);
		num maxBias = j->maxBias;
		j->bias = fclampc(-bias_coef_host(j->errorBias, dt) * (b->a * ratio - a->a - phase) / dt, -maxBias, maxBias);
	} else if (j->type == CPGPU_JOINT_PIN) {
		num dist_rest = j->P[4];
		j->r1 = xvect(a->T, vsub(anchorA, a->cog));
		j->r2 = xvect(b->T, vsub(anchorB, b->cog));
		V delta = vsub(vadd(b->p, j->r2), vadd(a->p, j->r1));
		num dist = vlength(delta);
		j->n = vmult(delta, 1.0 / (dist != 0.0 ? dist : INFINITY));
		j->nMass = 1.0 / k_scalar(w, a, b, j->r1, j->r2, j->n);
		num maxBias = j->maxBias;
		j->bias = fclampc(-bias_coef_host(j->errorBias, dt) * (dist - dist_rest) / dt, -maxBias, maxBias);
	} else if (j->type == CPGPU_JOINT_GROOVE) {
		/* P: groove ends {a.x, a.y, b.x, b.y} and anchorB on body B, groove normal n */
		V grv_a = v2(j->P[0], j->P[1]), grv_b = v2(j->P[2], j->P[3]);
		V anchor = v2(j->P[4], j->P[5]), grv_n = v2(j->P[6], j->P[7]);
		V ta = xpoint(a->T, grv_a), tb = xpoint(a->T, grv_b);
		V n = xvect(a->T, grv_n);
		num d = vdot(ta, n);
		j->n = n;
		j->r2 = xvect(b->T, vsub(anchor, b->cog));
		num td = vcross(vadd(b->p, j->r2), n);
		if (td <= vcross(ta, n)) {
			j->clamp = 1.0;
			j->r1 = vsub(ta, a->p);
		} else if (td >= vcross(tb, n)) {
			j->clamp = -1.0;
			j->r1 = vsub(tb, a->p);
		} else {
			j->clamp = 0.0;
			j->r1 = vsub(vadd(vmult(vperp(n), -td), vmult(n, d)), a->p);
		}
		k_tensor(w, a, b, j->r1, j->r2, j);
		V delta = vsub(vadd(b->p, j->r2), vadd(a->p, j->r1));
		j->bias2 = vclamp(vmult(delta, -bias_coef_host(j->errorBias, dt) / dt), j->maxBias);
	} else if (j->type == CPGPU_JOINT_ROTARY_LIMIT) {
		num mn = j->P[0], mx = j->P[1];
		num dist = b->a - a->a;
		num pdist = 0.0;
		if (dist > mx) pdist = mx - dist;
		else if (dist < mn) pdist = mn - dist;
		j->iSum = 1.0 / (a->i_inv + b->i_inv);
		num maxBias = j->maxBias;
		j->bias = fclampc(-bias_coef_host(j->errorBias, dt) * pdist / dt, -maxBias, maxBias);
		if (!j->bias) j->jAcc = 0.0;
	} else if (j->type == CPGPU_JOINT_RATCHET) {
		num phase = j->P[0], ratchet = j->P[1];
		num delta = b->a - a->a;
		num diff = j->angle - delta;
		num pdist = 0.0;
		if (diff * ratchet > 0.0) pdist = diff;
		else j->angle = floor((delta - phase) / ratchet) * ratchet + phase;
		j->iSum = 1.0 / (a->i_inv + b->i_inv);
		num maxBias = j->maxBias;
		j->bias = fclampc(-bias_coef_host(j->errorBias, dt) * pdist / dt, -maxBias, maxBias);
		if (!j->bias) j->jAcc = 0.0;
	} else if (j->type == CPGPU_JOINT_SIMPLE_MOTOR) {
		j->iSum = 1.0 / (a->i_inv + b->i_inv);
	} else {
		/* damped rotary spring with the default torque (relativeAngle - restAngle) stiffness */
		num restAngle = j->P[0], stiffness = j->P[1], damping = j->P[2];
		num moment = a->i_inv + b->i_inv;
		j->iSum = 1.0 / moment;
		j->v_coef = 1.0 - (w->libm_trig ? exp(-damping * dt * moment) : cp_exp_oracle(-damping * dt * moment));
		j->target_vrn = 0.0;
		num j_spring = ((a->a - b->a) - restAngle) * stiffness * dt;
		j->jAcc = j_spring;
		a->w -= j_spring * a->i_inv;
		b->w += j_spring * b->i_inv;
	}
}

static void joint_apply_cached(oworld *w, OJoint *j, num dt_coef)
{
	OBody *a = &w->B[j->a], *b = &w->B[j->b];
	if (j->type == CPGPU_JOINT_PIVOT) {
		apply_impulses(a, b, j->r1, j->r2, vmult(j->jAcc2, dt_coef));
	} else if (j->type == CPGPU_JOINT_SLIDE) {
		V jj = vmult(j->n, j->jnAcc * dt_coef);
		apply_impulses(a, b, j->r1, j->r2, jj);
	} else if (j->type == CPGPU_JOINT_GEAR) {
		num ratio_inv = j->P[2];
		num jj = j->jAcc * dt_coef;
		a->w -= jj * a->i_inv * ratio_inv;
		b->w += jj * b->i_inv;
	} else if (j->type == CPGPU_JOINT_PIN) {
		apply_impulses(a, b, j->r1, j->r2, vmult(j->n, j->jnAcc * dt_coef));
	} else if (j->type == CPGPU_JOINT_GROOVE) {
		apply_impulses(a, b, j->r1, j->r2, vmult(j->jAcc2, dt_coef));
	} else if (j->type != CPGPU_JOINT_DAMPED_SPRING && j->type != CPGPU_JOINT_DAMPED_ROTARY_SPRING) {
		apply_angular(a, b, j->jAcc * dt_coef);
	}
}

static void joint_apply_impulse(oworld *w, OJoint *j, num dt)
{
	OBody *a = &w->B[j->a], *b = &w->B[j->b];
	if (j->type == CPGPU_JOINT_PIVOT) {
		V r1 = j->r1, r2 = j->r2;
		V vr = rel_vel(a, b, r1, r2);
		V d = vsub(j->bias2, vr);
		V jj = v2(d.x * j->k11 + d.y * j->k12, d.x * j->k21 + d.y * j->k22);
		V jOld = j->jAcc2;
		j->jAcc2 = vclamp(vadd(j->jAcc2, jj), j->maxForce * dt);
		jj = vsub(j->jAcc2, jOld);
		apply_impulses(a, b, j->r1, j->r2, jj);
	} else if (j->type == CPGPU_JOINT_DAMPED_SPRING) {
		V n = j->n;
		num vrn = vdot(rel_vel(a, b, j->r1, j->r2), n);
		num v_damp = (j->target_vrn - vrn) * j->v_coef;
		j->target_vrn = vrn + v_damp;
		num j_damp = v_damp * j->nMass;
		j->jAcc += j_damp;
		apply_impulses(a, b, j->r1, j->r2, vmult(j->n, j_damp));
	} else if (j->type == CPGPU_JOINT_SLIDE) {
		if (veql(j->n, v2(0.0, 0.0))) return;
		V n = j->n;
		V vr = rel_vel(a, b, j->r1, j->r2);
		num vrn = vdot(vr, n);
		num jn = (j->bias - vrn) * j->nMass;
		num jnOld = j->jnAcc;
		j->jnAcc = fclampc(jnOld + jn, -j->maxForce * dt, 0.0);
		jn = j->jnAcc - jnOld;
		apply_impulses(a, b, j->r1, j->r2, vmult(n, jn));
	} else if (j->type == CPGPU_JOINT_GEAR) {
		num ratio = j->P[1], ratio_inv = j->P[2];
		num wr = b->w * ratio - a->w;
		num jMax = j->maxForce * dt;
		num jj = (j->bias - wr) * j->iSum;
		num jOld = j->jAcc;
		j->jAcc = fclampc(jOld + jj, -jMax, jMax);
		jj = j->jAcc - jOld;
		a->w -= jj * a->i_inv * ratio_inv;
		b->w += jj * b->i_inv;
	} else if (j->type == CPGPU_JOINT_PIN) {
		V n = j->n;
		num vrn = vdot(rel_vel(a, b, j->r1, j->r2), n);
		num jnMax = j->maxForce * dt;
		num jn = (j->bias - vrn) * j->nMass;
		num jnOld = j->jnAcc;
		j->jnAcc = fclampc(jnOld + jn, -jnMax, jnMax);
		jn = j->jnAcc - jnOld;
		apply_impulses(a, b, j->r1, j->r2, vmult(n, jn));
	} else if (j->type == CPGPU_JOINT_GROOVE) {
		V vr = rel_vel(a, b, j->r1, j->r2);
		V d = vsub(j->bias2, vr);
		V jj = v2(j->k11 * d.x + j->k12 * d.y, j->k21 * d.x + j->k22 * d.y);
		V jOld = j->jAcc2;
		/* grooveConstrain (cpGrooveJoint.c:74-79) */
		V jc = vadd(jOld, jj);
		V n = j->n;
		V jClamp = (j->clamp * vcross(jc, n) > 0.0) ? jc : vmult(n, vdot(jc, n) / vdot(n, n));
		j->jAcc2 = vclamp(jClamp, j->maxForce * dt);
		jj = vsub(j->jAcc2, jOld);
		apply_impulses(a, b, j->r1, j->r2, jj);
	} else if (j->type == CPGPU_JOINT_ROTARY_LIMIT || j->type == CPGPU_JOINT_RATCHET) {
		if (!j->bias) return;
		num wr = b->w - a->w;
		num jMax = j->maxForce * dt;
		num jj = -(j->bias + wr) * j->iSum;
		num jOld = j->jAcc;
		if (j->type == CPGPU_JOINT_RATCHET) {
			num ratchet = j->P[1];
			j->jAcc = fclampc((jOld + jj) * ratchet, 0.0, jMax * fabs(ratchet)) / ratchet;
		} else if (j->bias < 0.0) {
			j->jAcc = fclampc(jOld + jj, 0.0, jMax);
		} else {
			j->jAcc = fclampc(jOld + jj, -jMax, 0.0);
		}
		apply_angular(a, b, j->jAcc - jOld);
	} else if (j->type == CPGPU_JOINT_SIMPLE_MOTOR) {
		num wr = b->w - a->w + j->P[0];
		num jMax = j->maxForce * dt;
		num jj = -wr * j->iSum;
		num jOld = j->jAcc;
		j->jAcc = fclampc(jOld + jj, -jMax, jMax);
		apply_angular(a, b, j->jAcc - jOld);
	} else {
		num wrn = a->w - b->w;
		num w_damp = (j->target_vrn - wrn) * j->v_coef;
		j->target_vrn = wrn + w_damp;
		num j_damp = w_damp * j->iSum;
		j->jAcc += j_damp;
		a->w += j_damp * a->i_inv;
		b->w -= j_damp * b->i_inv;
	}
}

/* ------------------------------------------------------------------ */
/* the step: cpSpaceStep (cpSpaceStep.c:335-445)                       */

static void step_world(oworld *w, OWorldPart *P, num dt, const int32_t *order, int64_t order_n)
{
	cpgpuSpaceParams *sp = &P->sp;
	sp->stamp++;
	num prev_dt = sp->curr_dt;
	sp->curr_dt = NUM_VAL(dt);

	/* :351-360 reset the arbiters solved last step */
	for (int64_t k = 0; k < P->nactive; k++) P->arbs[P->active[k]].state = CPGPU_ARB_NORMAL;
	P->nactive = 0;

	/* :364-367 integrate positions of dynamic and kinematic bodies */
	for (int b = P->b0; b < P->b1; b++) if (w->B[b].type != CPGPU_BODY_STATIC) update_position(w, &w->B[b], dt);
	/* :371 cache the dynamic-index shapes */
	for (int s = P->s0; s < P->s1; s++) if (w->B[w->S[s].body].type != CPGPU_BODY_STATIC) cache_shape(w, &w->S[s]);
	/* :372 cpBBTreeReindexQuery first updates every leaf (cpBBTree.c:644) */
	for (int s = P->s0; s < P->s1; s++) if (w->B[w->S[s].body].type != CPGPU_BODY_STATIC) leaf_update(w, &w->S[s]);

	/* :372 broadphase + cpSpaceCollideShapes (:235-289) */
	int64_t np;
	IPair *pairs = world_pairs(w, P, &np);
	P->npairs = np;
	P->ncontacts = 0;
	pid_rebuild(w, P, 0);
	while (2 * (P->pid_n + np) + 64 > P->pid_mask + 1) {
		int64_t need = 2 * (P->pid_n + np) + 64, cap = P->pid_mask + 1;
		while (cap < need) cap *= 2;
		/* grow: reinsert into a larger table */
		uint64_t *ok = P->pid_key; uint32_t *oi = P->pid_id, *os = P->pid_stamp;
		int64_t omask = P->pid_mask;
		P->pid_key = (uint64_t *)malloc((size_t)cap * sizeof(uint64_t));
		P->pid_id = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
		P->pid_stamp = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
		for (int64_t k = 0; k < cap; k++) P->pid_key[k] = UINT64_MAX;
		P->pid_mask = cap - 1;
		P->pid_n = 0;
		for (int64_t k = 0; k <= omask; k++) if (ok[k] != UINT64_MAX) {
			uint64_t h = mix64(ok[k]) & (uint64_t)P->pid_mask;
			while (P->pid_key[h] != UINT64_MAX) h = (h + 1) & (uint64_t)P->pid_mask;
			P->pid_key[h] = ok[k]; P->pid_id[h] = oi[k]; P->pid_stamp[h] = os[k];
			P->pid_n++;
		}
		free(ok); free(oi); free(os);
	}
	for (int64_t k = 0; k < np; k++) {
		int32_t ia = pairs[k].a, ib = pairs[k].b;
		/* narrowphase orientation: cpCollide sorts by type (cpCollision.c:707-710);
		 * same-type pairs take the lower shape index first */
		if (w->S[ia].type > w->S[ib].type) { int32_t t = ia; ia = ib; ib = t; }
		if (order) {
			for (int64_t q = 0; q < order_n; q++) {
				if (order[2 * q] == ib && order[2 * q + 1] == ia && w->S[ia].type == w->S[ib].type) {
					int32_t t = ia; ia = ib; ib = t;
					break;
				}
			}
		}
		uint64_t key = pair_key(ia, ib);
		int64_t ai = tab_find(P, key);
		uint32_t id = pid_get(w, P, key, ia, ib);
		if (w->ncoll > 0) {
			const int32_t *c = coll_find(w, key);
			if (c) { ia = c[0]; ib = c[1]; id = (uint32_t)c[2]; }
		}
		CollInfo info;
		collide(w, ia, ib, id, &info);
		pid_put(P, key, info.id);   /* cpSpaceCollideShapes returns the new id to the tree (:245) */
		if (info.count == 0) continue;
		if (ai < 0) {
			/* cpArbiterInit (cpArbiter.c:316-350) */
			ai = arb_append(P);
			OArb *na = &P->arbs[ai];
			na->key = key;
			/* colouring priority from the shape indices within the space, so a
			 * space colours the same alone or inside a batch */
			na->prio = mix64(pair_key(ia - P->s0, ib - P->s0));
			na->count = 0;
			na->state = CPGPU_ARB_FIRST_COLLISION;
			na->stamp = 0;
		}
		OArb *arb = &P->arbs[ai];
		/* cpArbiterUpdate (cpArbiter.c:356-414) */
		const OShape *sa = &w->S[ia], *sb = &w->S[ib];
		arb->sa = ia; arb->sb = ib; arb->ba = sa->body; arb->bb = sb->body;
		OContact nc[2];
		for (int i = 0; i < info.count; i++) {
			OContact *con = &nc[i];
			memset(con, 0, sizeof(*con));
			con->r1 = vsub(info.p1[i], w->B[sa->body].p);
			con->r2 = vsub(info.p2[i], w->B[sb->body].p);
			con->hash = info.hash[i];
			con->jnAcc = con->jtAcc = 0.0;
			for (int j = 0; j < arb->count; j++) {
				OContact *old = &arb->con[j];
				if (con->hash == old->hash) { con->jnAcc = old->jnAcc; con->jtAcc = old->jtAcc; }
			}
		}
		arb->con[0] = nc[0];
		if (info.count > 1) arb->con[1] = nc[1];
		arb->count = info.count;
		arb->n = info.n;
		arb->e = sa->e * sb->e;
		arb->u = sa->u * sb->u;
		V surface_vr = vsub(sb->sv, sa->sv);
		arb->svr = vsub(surface_vr, vmult(info.n, vdot(surface_vr, info.n)));
		if (arb->state == CPGPU_ARB_CACHED) arb->state = CPGPU_ARB_FIRST_COLLISION;
		/* :255-284 begin (on a first collision; false ignores the pair until it
		 * separates), preSolve, then admission: not ignored, accepted, no sensor,
		 * not two infinite masses */
		arb->events = w->n_handlers > 0 ? pair_events(w, sa, sb) : 0u;
		int solve = 1;
		const int due_begin = arb->state == CPGPU_ARB_FIRST_COLLISION && (arb->events & CPGPU_EVENT_BEGIN);
		const int due_pre = arb->state != CPGPU_ARB_IGNORE && (arb->events & CPGPU_EVENT_PRE_SOLVE);
		if (arb->state == CPGPU_ARB_IGNORE) {
			solve = 0;
		} else if (w->hfn && (due_begin || due_pre)) {
			cpgpuCallbackArbiter v;
			cpgpuCallbackDecision dec = {-1, 1, 0, 0, NUM_VAL(arb->e), NUM_VAL(arb->u),
			                             {NUM_VAL(arb->svr.x), NUM_VAL(arb->svr.y)}};
			if (due_begin) {
				cb_view(w, (int)(P - w->W), arb, &v);
				w->hfn(w->hctx, CPGPU_EVENT_BEGIN, &v, &dec);
				if (!dec.solve || dec.ignore) arb->state = CPGPU_ARB_IGNORE;
			}
			if (arb->state != CPGPU_ARB_IGNORE && (arb->events & CPGPU_EVENT_PRE_SOLVE)) {
				arb->e = dec.e; arb->u = dec.u; arb->svr.x = dec.surface_vr[0]; arb->svr.y = dec.surface_vr[1];
				cb_view(w, (int)(P - w->W), arb, &v);
				dec.solve = 1;
				w->hfn(w->hctx, CPGPU_EVENT_PRE_SOLVE, &v, &dec);
				if (dec.ignore) arb->state = CPGPU_ARB_IGNORE;
			}
			arb->e = dec.e; arb->u = dec.u; arb->svr.x = dec.surface_vr[0]; arb->svr.y = dec.surface_vr[1];
			if (!dec.solve || arb->state == CPGPU_ARB_IGNORE) solve = 0;
		}
		if (sa->sensor || sb->sensor || (w->B[sa->body].m == INFINITY && w->B[sb->body].m == INFINITY)) solve = 0;
		if (solve) {
			P->active[P->nactive++] = ai;
			P->ncontacts += info.count;
		} else {
			/* :276-282 the contacts are dropped; a pair not ignored counts as normal */
			arb->count = 0;
			if (arb->state != CPGPU_ARB_IGNORE) arb->state = CPGPU_ARB_NORMAL;
		}
		/* :287 stamp */
		arb->stamp = sp->stamp;
		/* keep the lookup index current for arbiters created this step */
		if (P->narb * 2 + 16 > P->tabmask + 1) tab_rebuild(P);
		else if (tab_find(P, key) < 0) {
			uint64_t h = mix64(key) & (uint64_t)P->tabmask;
			while (P->tab[h] >= 0) h = (h + 1) & (uint64_t)P->tabmask;
			P->tab[h] = ai;
		}
	}
	free(pairs);

	/* :380 cpHashSetFilter(cachedArbiters, cpSpaceArbiterSetFilter) */
	{
		int64_t *remap = (int64_t *)malloc((size_t)(P->narb ? P->narb : 1) * sizeof(int64_t));
		int64_t keep = 0;
		for (int64_t i = 0; i < P->narb; i++) {
			OArb *arb = &P->arbs[i];
			uint32_t ticks = sp->stamp - arb->stamp;
			const OBody *a = &w->B[arb->ba], *b = &w->B[arb->bb];
			int evict = 0;
			if (a->type == CPGPU_BODY_STATIC && b->type == CPGPU_BODY_STATIC) {
				evict = 0;
			} else {
				if (ticks >= 1 && arb->state != CPGPU_ARB_CACHED) {
					/* :306-310 separate */
					arb->state = CPGPU_ARB_CACHED;
					if (arb->events & CPGPU_EVENT_SEPARATE) {
						sep_push(w, (int32_t)(P - w->W), arb->sa, arb->sb);
						if (w->hfn) {
							cpgpuCallbackArbiter v;
							cpgpuCallbackDecision dec;
							memset(&dec, 0, sizeof(dec));
							cb_view(w, (int)(P - w->W), arb, &v);
							w->hfn(w->hctx, CPGPU_EVENT_SEPARATE, &v, &dec);
						}
					}
				}
				if (ticks >= sp->collision_persistence) evict = 1;
			}
			if (evict) { remap[i] = -1; continue; }
			remap[i] = keep;
			if (keep != i) P->arbs[keep] = *arb;
			keep++;
		}
		P->narb = keep;
		for (int64_t k = 0; k < P->nactive; k++) P->active[k] = remap[P->active[k]];
		free(remap);
		tab_rebuild(P);
	}

	/* solve order: deterministic colours (or an explicit replay order) */
	if (order) {
		int64_t n = 0;
		for (int64_t q = 0; q < order_n; q++) {
			uint64_t key = pair_key(order[2 * q], order[2 * q + 1]);
			int64_t ai = tab_find(P, key);
			if (ai >= 0 && P->arbs[ai].stamp == sp->stamp) { P->arbs[ai].colour = (int)q; P->active[n++] = ai; }
		}
		P->nactive = n;
		P->ncol_arb = (int)n;
	} else {
		int64_t n = P->nactive;
		int32_t *ba = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
		int32_t *bb = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
		uint64_t *pr = (uint64_t *)malloc((size_t)(n + 1) * sizeof(uint64_t));
		int32_t *col = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
		for (int64_t k = 0; k < n; k++) {
			OArb *arb = &P->arbs[P->active[k]];
			ba[k] = w->B[arb->ba].type == CPGPU_BODY_DYNAMIC ? arb->ba : -1;
			bb[k] = w->B[arb->bb].type == CPGPU_BODY_DYNAMIC ? arb->bb : -1;
			pr[k] = arb->prio;
		}
		if (!grid_units(w, P)) {
			P->ncol_arb = colour_elements(w, n, ba, bb, pr, col);
			for (int64_t k = 0; k < n; k++) P->arbs[P->active[k]].colour = col[k];
		} else {
			/* each unit coloured on its own */
			int64_t *idx = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
			int32_t *uba = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
			int32_t *ubb = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
			uint64_t *upr = (uint64_t *)malloc((size_t)(n + 1) * sizeof(uint64_t));
			P->ncol_arb = 0;
			/* the active arbiters grouped by unit, in active order within a unit */
			const int n_units = UNIT_WIDE(w->grid_parts) + 1;
			int64_t *ustart = (int64_t *)calloc((size_t)n_units + 1, sizeof(int64_t));
			int64_t *byunit = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
			for (int64_t k = 0; k < n; k++) ustart[P->arbs[P->active[k]].unit + 1]++;
			for (int u = 0; u < n_units; u++) ustart[u + 1] += ustart[u];
			for (int64_t k = 0; k < n; k++) {
				const int u = P->arbs[P->active[k]].unit;
				byunit[ustart[u]++] = k;
			}
			for (int u = n_units; u > 0; u--) ustart[u] = ustart[u - 1];
			ustart[0] = 0;
			for (int u = 0; u < n_units; u++) {
				int64_t m = 0;
				for (int64_t j = ustart[u]; j < ustart[u + 1]; j++) {
					const int64_t k = byunit[j];
					idx[m] = k; uba[m] = ba[k]; ubb[m] = bb[k]; upr[m] = pr[k]; m++;
				}
				int nc = colour_elements(w, m, uba, ubb, upr, col);
				if (nc > P->ncol_arb) P->ncol_arb = nc;
				for (int64_t q = 0; q < m; q++) P->arbs[P->active[idx[q]]].colour = col[q];
			}
			free(ustart); free(byunit);
			free(idx); free(uba); free(ubb); free(upr);
		}
		g_sort_arbs = P->arbs;
		qsort(P->active, (size_t)n, sizeof(int64_t), cmp_solve);
		free(ba); free(bb); free(pr); free(col);
	}
	{
		int64_t n = P->c1 - P->c0;
		int32_t *ba = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
		int32_t *bb = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
		uint64_t *pr = (uint64_t *)malloc((size_t)(n + 1) * sizeof(uint64_t));
		int32_t *col = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
		for (int64_t k = 0; k < n; k++) {
			OJoint *j = &w->J[P->c0 + k];
			ba[k] = w->B[j->a].type == CPGPU_BODY_DYNAMIC ? j->a : -1;
			bb[k] = w->B[j->b].type == CPGPU_BODY_DYNAMIC ? j->b : -1;
			pr[k] = j->prio;
		}
		if (w->joint_list_order) {
			/* replay the reference's constraint list order */
			for (int64_t k = 0; k < n; k++) { w->J[P->c0 + k].colour = (int)k; P->jorder[k] = P->c0 + (int32_t)k; }
			P->ncol_cons = (int)n;
		} else {
			P->ncol_cons = colour_elements(w, n, ba, bb, pr, col);
			for (int64_t k = 0; k < n; k++) { w->J[P->c0 + k].colour = col[k]; P->jorder[k] = P->c0 + (int32_t)k; }
			g_sort_joints = w->J;
			qsort(P->jorder, (size_t)n, sizeof(int32_t), cmp_jsolve);
		}
		free(ba); free(bb); free(pr); free(col);
	}

	/* :383-387 prestep the arbiters */
	num slop = sp->collision_slop;
	num biasCoef = 1.0 - pow(sp->collision_bias, NUM_VAL(dt));
	for (int64_t k = 0; k < P->nactive; k++) arb_prestep(w, &P->arbs[P->active[k]], dt, slop, biasCoef);
	/* :389-396 prestep the constraints (in solve order) */
	int64_t nj = P->c1 - P->c0;
	for (int64_t k = 0; k < nj; k++) joint_prestep(w, &w->J[P->jorder[k]], dt);
	/* :399-404 integrate velocities */
	num damping = pow(sp->damping, NUM_VAL(dt));
	V g = v2(sp->gravity_x, sp->gravity_y);
	for (int b = P->b0; b < P->b1; b++) if (w->B[b].type != CPGPU_BODY_STATIC) update_velocity(w, &w->B[b], g, damping, dt);
	/* :407-415 cached impulses */
	num dt_coef = (prev_dt == 0.0 ? 0.0 : dt / prev_dt);
	for (int64_t k = 0; k < P->nactive; k++) arb_apply_cached(w, &P->arbs[P->active[k]], dt_coef);
	for (int64_t k = 0; k < nj; k++) joint_apply_cached(w, &w->J[P->jorder[k]], dt_coef);
	/* :418-427 the impulse solver */
	for (int it = 0; it < sp->iterations; it++) {
		for (int64_t k = 0; k < P->nactive; k++) arb_apply_impulse(w, &P->arbs[P->active[k]]);
		for (int64_t k = 0; k < nj; k++) joint_apply_impulse(w, &w->J[P->jorder[k]], dt);
	}

	/* :437-442 postSolve of the solved arbiters */
	if (w->hfn)
		for (int64_t k = 0; k < P->nactive; k++) {
			const OArb *arb = &P->arbs[P->active[k]];
			if (!(arb->events & CPGPU_EVENT_POST_SOLVE)) continue;
			cpgpuCallbackArbiter v;
			cpgpuCallbackDecision dec;
			memset(&dec, 0, sizeof(dec));
			cb_view(w, (int)(P - w->W), arb, &v);
			w->hfn(w->hctx, CPGPU_EVENT_POST_SOLVE, &v, &dec);
		}
}

int32_t ow_faults(const oworld *w, int32_t *first_body)
{
	if (first_body) *first_body = w->fault_body;
	return w->faults;
}

void ow_step(oworld *w, double dt)
{
	if (dt == 0.0) return;   /* cpSpaceStep.c:339 */
	w->nseps = 0;
	for (int wi = 0; wi < w->nw; wi++) {
		step_world(w, &w->W[wi], dt, w->next_n > 0 ? w->next_pairs : NULL, w->next_n > 0 ? w->next_n : 0);
	}
	w->next_n = 0;
	w->ncoll = 0;
	w->steps++;
}

/* ------------------------------------------------------------------ */
/* readback                                                            */

int ow_read_bodies(oworld *w, double *p, double *v, double *a, double *wv,
                   double *v_bias, double *w_bias, double *transform)
{
	for (int i = 0; i < w->nb; i++) {
		OBody *b = &w->B[i];
		if (p) { p[2 * i] = NUM_VAL(b->p.x); p[2 * i + 1] = NUM_VAL(b->p.y); }
		if (v) { v[2 * i] = NUM_VAL(b->v.x); v[2 * i + 1] = NUM_VAL(b->v.y); }
		if (a) a[i] = NUM_VAL(b->a);
		if (wv) wv[i] = NUM_VAL(b->w);
		if (v_bias) { v_bias[2 * i] = NUM_VAL(b->v_bias.x); v_bias[2 * i + 1] = NUM_VAL(b->v_bias.y); }
		if (w_bias) w_bias[i] = NUM_VAL(b->w_bias);
		if (transform) {
			double *t = transform + 6 * i;
			t[0] = NUM_VAL(b->T.a); t[1] = NUM_VAL(b->T.b); t[2] = NUM_VAL(b->T.c);
			t[3] = NUM_VAL(b->T.d); t[4] = NUM_VAL(b->T.tx); t[5] = NUM_VAL(b->T.ty);
		}
	}
	return 0;
}

int ow_write_bodies(oworld *w, const double *v, const double *wv, const double *f, const double *t)
{
	for (int i = 0; i < w->nb; i++) {
		OBody *b = &w->B[i];
		if (v) b->v = v2(v[2 * i], v[2 * i + 1]);
		if (wv) b->w = wv[i];
		if (f) b->f = v2(f[2 * i], f[2 * i + 1]);
		if (t) b->t = t[i];
	}
	return 0;
}

/* The setters between steps (cpBodySetPosition / SetVelocity / SetForce /
 * SetAngle / SetAngularVelocity / SetTorque, cpBody.c:375-470, which also set
 * the transform): the listed bodies' fields from n-entry arrays. */
int ow_write_bodies_at(oworld *w, int32_t n, const int32_t *index, const double *p, const double *v,
                       const double *f, const double *a, const double *wv, const double *t, const double *transform)
{
	for (int k = 0; k < n; k++)
		if (index[k] < 0 || index[k] >= w->nb) return -1;
	for (int k = 0; k < n; k++) {
		OBody *b = &w->B[index[k]];
		if (p) b->p = v2(p[2 * k], p[2 * k + 1]);
		if (v) b->v = v2(v[2 * k], v[2 * k + 1]);
		if (f) b->f = v2(f[2 * k], f[2 * k + 1]);
		if (a) b->a = a[k];
		if (wv) b->w = wv[k];
		if (t) b->t = t[k];
		if (transform) b->T = xf_from(transform + 6 * k);
	}
	return 0;
}

/* The cpSpace parameter setters (cpSpace.c:240-300). */
int ow_set_space_params(oworld *w, int32_t space, const cpgpuSpaceParams *sp)
{
	if (space < 0 || space >= w->nw) return -1;
	w->W[space].sp = *sp;
	return 0;
}

/* A new oracle world from the spaces packed again after adds and removes
 * (cpSpaceAdd* / cpSpaceRemove*, cpSpace.c:417-571), keeping the old world's
 * collision state under the new shape indices: every cached arbiter and GJK
 * warm-start id whose two shapes are packed again, and each dynamic shape's
 * fattened leaf box and rebuild step.  Arbiters of removed shapes are
 * dropped, as cpSpaceRemoveShape does through cpSpaceFilterArbiters
 * (cpSpace.c:513-541).  shape_from[s] is the old index of shape s, or -1. */
oworld *ow_new_from(const cpgpuWorldDesc *d, oworld *old, const int32_t *shape_from)
{
	oworld *w = ow_new(d);
	int32_t *inv = (int32_t *)malloc((size_t)(old->ns ? old->ns : 1) * sizeof(int32_t));
	for (int o = 0; o < old->ns; o++) inv[o] = -1;
	for (int s = 0; s < w->ns; s++) if (shape_from[s] >= 0) inv[shape_from[s]] = s;
	for (int s = 0; s < w->ns; s++) {
		const int o = shape_from[s];
		if (o < 0 || w->B[w->S[s].body].type == CPGPU_BODY_STATIC || old->B[old->S[o].body].type == CPGPU_BODY_STATIC)
			continue;
		memcpy(w->S[s].fat, old->S[o].fat, sizeof(w->S[s].fat));
		w->S[s].rebuild_stamp = old->S[o].rebuild_stamp;
	}
	/* the new space of an old pair, or -1 when a shape is gone */
#define NEW_PAIR_WORLD(a, b) ((inv[(a)] < 0 || inv[(b)] < 0 || w->S[inv[(a)]].world != w->S[inv[(b)]].world) \
                              ? -1 : w->S[inv[(a)]].world)
	int64_t *pid_count = (int64_t *)calloc((size_t)w->nw, sizeof(int64_t));
	for (int wi = 0; wi < old->nw; wi++) {
		OWorldPart *Po = &old->W[wi];
		int64_t *at = (int64_t *)malloc((size_t)(Po->narb ? Po->narb : 1) * sizeof(int64_t));
		for (int64_t i = 0; i < Po->narb; i++) {
			const OArb *oa = &Po->arbs[i];
			const int nw = NEW_PAIR_WORLD(oa->sa, oa->sb);
			at[i] = -1;
			if (nw < 0) continue;
			OWorldPart *Pn = &w->W[nw];
			const int64_t j = arb_append(Pn);
			OArb *na = &Pn->arbs[j];
			*na = *oa;
			na->sa = inv[oa->sa]; na->sb = inv[oa->sb];
			na->ba = w->S[na->sa].body; na->bb = w->S[na->sb].body;
			na->key = pair_key(na->sa, na->sb);
			na->prio = mix64(pair_key(na->sa - Pn->s0, na->sb - Pn->s0));
			at[i] = j;
		}
		for (int64_t k = 0; k < Po->nactive; k++) {
			const int64_t i = Po->active[k];
			if (at[i] < 0) continue;
			OWorldPart *Pn = &w->W[w->S[inv[Po->arbs[i].sa]].world];
			Pn->active[Pn->nactive++] = at[i];
		}
		free(at);
		for (int64_t k = 0; Po->pid_key && k <= Po->pid_mask; k++) {
			if (Po->pid_key[k] == UINT64_MAX) continue;
			const int nw = NEW_PAIR_WORLD((int32_t)(Po->pid_key[k] >> 32), (int32_t)(Po->pid_key[k] & 0xffffffffu));
			if (nw >= 0) pid_count[nw]++;
		}
	}
	for (int wi = 0; wi < w->nw; wi++) {
		OWorldPart *Pn = &w->W[wi];
		int64_t cap = 256;
		while (cap < 2 * pid_count[wi] + 64) cap *= 2;
		Pn->pid_key = (uint64_t *)malloc((size_t)cap * sizeof(uint64_t));
		Pn->pid_id = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
		Pn->pid_stamp = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
		for (int64_t k = 0; k < cap; k++) Pn->pid_key[k] = UINT64_MAX;
		Pn->pid_mask = cap - 1;
		Pn->pid_n = 0;
	}
	for (int wi = 0; wi < old->nw; wi++) {
		OWorldPart *Po = &old->W[wi];
		for (int64_t k = 0; Po->pid_key && k <= Po->pid_mask; k++) {
			if (Po->pid_key[k] == UINT64_MAX) continue;
			const int32_t a = (int32_t)(Po->pid_key[k] >> 32), b = (int32_t)(Po->pid_key[k] & 0xffffffffu);
			const int nw = NEW_PAIR_WORLD(a, b);
			if (nw < 0) continue;
			OWorldPart *Pn = &w->W[nw];
			const uint64_t key = pair_key(inv[a], inv[b]);
			uint64_t h = mix64(key) & (uint64_t)Pn->pid_mask;
			while (Pn->pid_key[h] != UINT64_MAX) h = (h + 1) & (uint64_t)Pn->pid_mask;
			Pn->pid_key[h] = key; Pn->pid_id[h] = Po->pid_id[k]; Pn->pid_stamp[h] = Po->pid_stamp[k];
			Pn->pid_n++;
		}
	}
#undef NEW_PAIR_WORLD
	for (int wi = 0; wi < w->nw; wi++) tab_rebuild(&w->W[wi]);
	free(pid_count);
	free(inv);
	w->steps = old->steps;
	return w;
}

int64_t ow_read_arbiters(oworld *w, int64_t cap, int32_t *shapes, int32_t *colour, int32_t *state,
                         int32_t *count, double *normal, double *contact, uint64_t *hash)
{
	int64_t n = 0;
	for (int wi = 0; wi < w->nw; wi++) {
		OWorldPart *P = &w->W[wi];
		for (int64_t k = 0; k < P->nactive; k++, n++) {
			if (n >= cap) continue;
			OArb *arb = &P->arbs[P->active[k]];
			if (shapes) { shapes[2 * n] = arb->sa; shapes[2 * n + 1] = arb->sb; }
			if (colour) colour[n] = arb->colour;
			if (state) state[n] = arb->state;
			if (count) count[n] = arb->count;
			if (normal) { normal[2 * n] = NUM_VAL(arb->n.x); normal[2 * n + 1] = NUM_VAL(arb->n.y); }
			for (int i = 0; i < 2; i++) {
				OContact *c = &arb->con[i];
				int valid = i < arb->count;
				if (contact) {
					double *o = contact + 24 * n + 12 * i;
					num vals[12] = {c->r1.x, c->r1.y, c->r2.x, c->r2.y, c->nMass, c->tMass, c->bounce,
					                   c->jnAcc, c->jtAcc, c->jBias, c->bias, 0.0};
					for (int q = 0; q < 12; q++) o[q] = valid ? NUM_VAL(vals[q]) : 0.0;
				}
				if (hash) hash[2 * n + i] = valid ? c->hash : 0;
			}
		}
	}
	return n;
}

/* solve unit of every arbiter of the last step, in ow_read_arbiters' order */
int64_t ow_read_units(oworld *w, int64_t cap, int32_t *unit)
{
	int64_t n = 0;
	for (int wi = 0; wi < w->nw; wi++) {
		OWorldPart *P = &w->W[wi];
		for (int64_t k = 0; k < P->nactive; k++, n++)
			if (n < cap) unit[n] = P->arbs[P->active[k]].unit;
	}
	return n;
}

int ow_read_constraints(oworld *w, double *jacc, int32_t *colour)
{
	for (int i = 0; i < w->nc; i++) {
		OJoint *j = &w->J[i];
		if (jacc) {
			if (j->type == CPGPU_JOINT_PIVOT) { jacc[2 * i] = NUM_VAL(j->jAcc2.x); jacc[2 * i + 1] = NUM_VAL(j->jAcc2.y); }
			else if (j->type == CPGPU_JOINT_GROOVE) { jacc[2 * i] = NUM_VAL(j->jAcc2.x); jacc[2 * i + 1] = NUM_VAL(j->jAcc2.y); }
			else if (j->type == CPGPU_JOINT_SLIDE || j->type == CPGPU_JOINT_PIN) { jacc[2 * i] = NUM_VAL(j->jnAcc); jacc[2 * i + 1] = 0.0; }
			else if (j->type == CPGPU_JOINT_RATCHET) { jacc[2 * i] = NUM_VAL(j->jAcc); jacc[2 * i + 1] = NUM_VAL(j->angle); }
			else { jacc[2 * i] = NUM_VAL(j->jAcc); jacc[2 * i + 1] = 0.0; }
		}
		if (colour) colour[i] = j->colour;
	}
	return 0;
}

void ow_stats(oworld *w, cpgpuStats *out)
{
	memset(out, 0, sizeof(*out));
	for (int wi = 0; wi < w->nw; wi++) {
		OWorldPart *P = &w->W[wi];
		out->candidate_pairs += P->npairs;
		out->arbiters += P->nactive;
		out->contacts += P->ncontacts;
		out->cached_arbiters += P->narb;
		if (P->ncol_arb > out->arbiter_colours) out->arbiter_colours = P->ncol_arb;
		if (P->ncol_cons > out->constraint_colours) out->constraint_colours = P->ncol_cons;
	}
	out->steps = w->steps;
}
