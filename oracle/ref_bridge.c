/*
 * ref_bridge.c -- test-only bridge into the compiled reference (Chipmunk2D 7.0.3).
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
 * the unmodified reference sources under /root/reference/src into
 * oracle/_ref/libcpref.so.  The tests use it to pin the oracle
 * (cpstep_oracle.c) against the real cpSpaceStep, cpCollide, cpBBTree and
 * body integrators, and bench.py times the reference's cpSpaceStep /
 * cpHastySpaceStep through it as the CPU baseline.  It builds a real cpSpace
 * from the same packed description (include/cpgpu.h) through the public API,
 * then writes the few private fields (chipmunk_structs.h) that the public
 * setters would recompute through libm (body transform, inverse masses).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "chipmunk/chipmunk_private.h"
#include "chipmunk/cpHastySpace.h"
#include "../include/cpgpu.h"

typedef struct rbWorld {
	cpSpace *space;
	int hasty;
	int owns_space;
	int nb, ns, nc;
	int b0, s0, c0;
	cpBody **bodies;
	cpShape **shapes;
	cpConstraint **cons;
	int *ctype;
	int wi;
	/* collision handler callbacks answered by the caller (the oracle's ow_handler_fn form) */
	void (*hfn)(void *ctx, int32_t event, const cpgpuCallbackArbiter *arb, cpgpuCallbackDecision *dec);
	void *hctx;
} rbWorld;

static cpTransform xf(const double *t)
{
	cpTransform x = {t[0], t[1], t[2], t[3], t[4], t[5]};
	return x;
}

/* Build world `wi` of the description as a cpSpace (hasty != 0: cpHastySpace
 * with that many threads). */
static rbWorld *fill_space(rbWorld *w, const cpgpuWorldDesc *d, int wi);

rbWorld *rb_new(const cpgpuWorldDesc *d, int wi, int hasty)
{
	rbWorld *w = (rbWorld *)calloc(1, sizeof(rbWorld));
	w->hasty = hasty;
	w->owns_space = 1;
	if (hasty) {
		w->space = cpHastySpaceNew();
		cpHastySpaceSetThreads(w->space, (unsigned long)hasty);
	} else {
		w->space = cpSpaceNew();
	}
	return fill_space(w, d, wi);
}

/* Build world `wi` of the description, through the public API, into a space
 * the caller made (for instance a cpGPUSpace) and keeps. */
rbWorld *rb_new_in(cpSpace *space, const cpgpuWorldDesc *d, int wi)
{
	rbWorld *w = (rbWorld *)calloc(1, sizeof(rbWorld));
	w->space = space;
	return fill_space(w, d, wi);
}

/* ---- collision handlers: each callback hands the arbiter to w->hfn ---- */

static void rb_view(rbWorld *w, cpArbiter *arb, cpgpuCallbackArbiter *v)
{
	memset(v, 0, sizeof(*v));
	v->slot = -1;
	v->space = w->wi;
	v->shape_a = (int32_t)(intptr_t)arb->a->userData;
	v->shape_b = (int32_t)(intptr_t)arb->b->userData;
	v->state = arb->state;
	v->count = arb->count;
	v->normal[0] = arb->n.x; v->normal[1] = arb->n.y;
	v->e = arb->e; v->u = arb->u;
	v->surface_vr[0] = arb->surface_vr.x; v->surface_vr[1] = arb->surface_vr.y;
	for (int c = 0; c < arb->count; c++) {
		const struct cpContact *con = &arb->contacts[c];
		double *o = v->contact[c];
		o[0] = con->r1.x; o[1] = con->r1.y; o[2] = con->r2.x; o[3] = con->r2.y;
		o[4] = con->jnAcc; o[5] = con->jtAcc;
		v->hash[c] = (uint64_t)con->hash;
	}
}

static cpBool rb_decide(cpArbiter *arb, int32_t event, rbWorld *w)
{
	if (!w->hfn) return cpTrue;
	cpgpuCallbackArbiter v;
	rb_view(w, arb, &v);
	cpgpuCallbackDecision dec = {-1, 1, 0, 0, arb->e, arb->u, {arb->surface_vr.x, arb->surface_vr.y}};
	w->hfn(w->hctx, event, &v, &dec);
	cpArbiterSetRestitution(arb, dec.e);
	cpArbiterSetFriction(arb, dec.u);
	cpArbiterSetSurfaceVelocity(arb, cpv(dec.surface_vr[0], dec.surface_vr[1]));
	if (dec.ignore) cpArbiterIgnore(arb);
	return dec.solve ? cpTrue : cpFalse;
}

static cpBool rb_begin(cpArbiter *arb, cpSpace *space, cpDataPointer data)
{
	(void)space;
	return rb_decide(arb, CPGPU_EVENT_BEGIN, (rbWorld *)data);
}
static cpBool rb_pre_solve(cpArbiter *arb, cpSpace *space, cpDataPointer data)
{
	(void)space;
	return rb_decide(arb, CPGPU_EVENT_PRE_SOLVE, (rbWorld *)data);
}
static void rb_notify(cpArbiter *arb, int32_t event, rbWorld *w)
{
	if (!w->hfn) return;
	cpgpuCallbackArbiter v;
	cpgpuCallbackDecision dec;
	memset(&dec, 0, sizeof(dec));
	rb_view(w, arb, &v);
	w->hfn(w->hctx, event, &v, &dec);
}
static void rb_post_solve(cpArbiter *arb, cpSpace *space, cpDataPointer data)
{
	(void)space;
	rb_notify(arb, CPGPU_EVENT_POST_SOLVE, (rbWorld *)data);
}
static void rb_separate(cpArbiter *arb, cpSpace *space, cpDataPointer data)
{
	(void)space;
	rb_notify(arb, CPGPU_EVENT_SEPARATE, (rbWorld *)data);
}

/* the description's handlers of space wi, through cpSpaceAddCollisionHandler,
 * cpSpaceAddWildcardHandler and cpSpaceAddDefaultCollisionHandler */
static void rb_add_handlers(rbWorld *w, const cpgpuWorldDesc *d, int wi)
{
	for (int k = 0; k < d->n_handlers; k++) {
		const cpgpuHandlerDesc *h = &d->handlers[k];
		if (h->space >= 0 && h->space != wi) continue;
		cpCollisionHandler *ch;
		if (h->type_b == CPGPU_WILDCARD && h->type_a == CPGPU_WILDCARD) ch = cpSpaceAddDefaultCollisionHandler(w->space);
		else if (h->type_b == CPGPU_WILDCARD) ch = cpSpaceAddWildcardHandler(w->space, (cpCollisionType)h->type_a);
		else ch = cpSpaceAddCollisionHandler(w->space, (cpCollisionType)h->type_a, (cpCollisionType)h->type_b);
		ch->userData = w;
		if (h->events & CPGPU_EVENT_BEGIN) ch->beginFunc = rb_begin;
		if (h->events & CPGPU_EVENT_PRE_SOLVE) ch->preSolveFunc = rb_pre_solve;
		if (h->events & CPGPU_EVENT_POST_SOLVE) ch->postSolveFunc = rb_post_solve;
		if (h->events & CPGPU_EVENT_SEPARATE) ch->separateFunc = rb_separate;
	}
}

void rb_set_handler_fn(rbWorld *w, void (*fn)(void *, int32_t, const cpgpuCallbackArbiter *, cpgpuCallbackDecision *),
                       void *ctx)
{
	w->hfn = fn;
	w->hctx = ctx;
}

static rbWorld *fill_space(rbWorld *w, const cpgpuWorldDesc *d, int wi)
{
	const cpgpuSpaceParams *sp = &d->space[wi];
	w->wi = wi;
	cpSpace *space = w->space;
	cpSpaceSetIterations(space, sp->iterations);
	cpSpaceSetGravity(space, cpv(sp->gravity_x, sp->gravity_y));
	cpSpaceSetDamping(space, sp->damping);
	cpSpaceSetCollisionSlop(space, sp->collision_slop);
	cpSpaceSetCollisionBias(space, sp->collision_bias);
	cpSpaceSetCollisionPersistence(space, sp->collision_persistence);
	space->curr_dt = sp->curr_dt;
	space->stamp = sp->stamp;

	w->b0 = d->world_body_start[wi];
	w->nb = d->world_body_start[wi + 1] - w->b0;
	w->s0 = d->world_shape_start[wi];
	w->ns = d->world_shape_start[wi + 1] - w->s0;
	w->c0 = d->world_constraint_start ? d->world_constraint_start[wi] : 0;
	w->nc = d->world_constraint_start ? d->world_constraint_start[wi + 1] - w->c0 : 0;
	w->bodies = (cpBody **)calloc((size_t)w->nb + 1, sizeof(cpBody *));
	w->shapes = (cpShape **)calloc((size_t)w->ns + 1, sizeof(cpShape *));
	w->cons = (cpConstraint **)calloc((size_t)w->nc + 1, sizeof(cpConstraint *));
	w->ctype = (int *)calloc((size_t)w->nc + 1, sizeof(int));

	for (int k = 0; k < w->nb; k++) {
		int i = w->b0 + k;
		cpBody *b;
		if (d->body_type[i] == CPGPU_BODY_STATIC) b = cpBodyNewStatic();
		else if (d->body_type[i] == CPGPU_BODY_KINEMATIC) b = cpBodyNewKinematic();
		else b = cpBodyNew(d->body_m[i], d->body_i[i]);
		b->m = d->body_m[i]; b->m_inv = d->body_m_inv[i];
		b->i = d->body_i[i]; b->i_inv = d->body_i_inv[i];
		b->cog = cpv(d->body_cog[2 * i], d->body_cog[2 * i + 1]);
		b->p = cpv(d->body_p[2 * i], d->body_p[2 * i + 1]);
		b->v = cpv(d->body_v[2 * i], d->body_v[2 * i + 1]);
		b->f = cpv(d->body_f[2 * i], d->body_f[2 * i + 1]);
		b->a = d->body_a[i]; b->w = d->body_w[i]; b->t = d->body_t[i];
		b->v_bias = cpv(d->body_v_bias[2 * i], d->body_v_bias[2 * i + 1]);
		b->w_bias = d->body_w_bias[i];
		b->transform = xf(d->body_transform + 6 * i);
		b->userData = (cpDataPointer)(intptr_t)i;
		cpSpaceAddBody(space, b);
		w->bodies[k] = b;
	}
	for (int k = 0; k < w->ns; k++) {
		int i = w->s0 + k;
		cpBody *body = w->bodies[d->shape_body[i] - w->b0];
		const double *g = d->shape_geom + 10 * i;
		double r = d->shape_radius[i];
		cpShape *s;
		if (d->shape_type[i] == CPGPU_SHAPE_CIRCLE) {
			s = cpCircleShapeNew(body, r, cpv(g[0], g[1]));
		} else if (d->shape_type[i] == CPGPU_SHAPE_SEGMENT) {
			s = cpSegmentShapeNew(body, cpv(g[0], g[1]), cpv(g[2], g[3]), r);
			cpSegmentShape *seg = (cpSegmentShape *)s;
			seg->n = cpv(g[4], g[5]);
			seg->a_tangent = cpv(g[6], g[7]);
			seg->b_tangent = cpv(g[8], g[9]);
		} else {
			int st = d->shape_poly_start[i], cnt = d->shape_poly_count[i];
			cpVect *verts = (cpVect *)malloc((size_t)cnt * sizeof(cpVect));
			for (int q = 0; q < cnt; q++) verts[q] = cpv(d->poly_verts[2 * (st + q)], d->poly_verts[2 * (st + q) + 1]);
			s = cpPolyShapeNewRaw(body, cnt, verts, r);
			cpPolyShape *poly = (cpPolyShape *)s;
			for (int q = 0; q < cnt; q++) {
				poly->planes[q + cnt].n = cpv(d->poly_normals[2 * (st + q)], d->poly_normals[2 * (st + q) + 1]);
			}
			free(verts);
		}
		cpShapeSetElasticity(s, d->shape_e[i]);
		cpShapeSetFriction(s, d->shape_u[i]);
		cpShapeSetSurfaceVelocity(s, cpv(d->shape_surface_v[2 * i], d->shape_surface_v[2 * i + 1]));
		cpShapeSetFilter(s, cpShapeFilterNew((cpGroup)d->shape_group[i], d->shape_categories[i], d->shape_mask[i]));
		if (d->shape_sensor && d->shape_sensor[i]) cpShapeSetSensor(s, cpTrue);
		if (d->shape_ctype) cpShapeSetCollisionType(s, (cpCollisionType)d->shape_ctype[i]);
		s->userData = (cpDataPointer)(intptr_t)i;
		cpSpaceAddShape(space, s);
		if (s->hashid != d->shape_hashid[i]) {
			fprintf(stderr, "ref_bridge: shape %d hashid %lu != description %lu\n", i,
			        (unsigned long)s->hashid, (unsigned long)d->shape_hashid[i]);
			abort();
		}
		w->shapes[k] = s;
	}
	/* cpSpaceAddShape caches shapes with the body transform; the body-type
	 * mass accumulation is bypassed (shapes carry zero mass), so restore the
	 * description's mass fields in case a public call touched them. */
	for (int k = 0; k < w->nb; k++) {
		int i = w->b0 + k;
		cpBody *b = w->bodies[k];
		b->m = d->body_m[i]; b->m_inv = d->body_m_inv[i];
		b->i = d->body_i[i]; b->i_inv = d->body_i_inv[i];
		b->cog = cpv(d->body_cog[2 * i], d->body_cog[2 * i + 1]);
		b->p = cpv(d->body_p[2 * i], d->body_p[2 * i + 1]);
		b->transform = xf(d->body_transform + 6 * i);
	}
	for (int k = 0; k < w->nc; k++) {
		int i = w->c0 + k;
		cpBody *a = w->bodies[d->cons_a[i] - w->b0], *b = w->bodies[d->cons_b[i] - w->b0];
		const double *P = d->cons_params + 8 * i;
		cpConstraint *c;
		int t = d->cons_type[i];
		if (t == CPGPU_JOINT_PIVOT) c = cpPivotJointNew2(a, b, cpv(P[0], P[1]), cpv(P[2], P[3]));
		else if (t == CPGPU_JOINT_DAMPED_SPRING) c = cpDampedSpringNew(a, b, cpv(P[0], P[1]), cpv(P[2], P[3]), P[4], P[5], P[6]);
		else if (t == CPGPU_JOINT_SLIDE) c = cpSlideJointNew(a, b, cpv(P[0], P[1]), cpv(P[2], P[3]), P[4], P[5]);
		else if (t == CPGPU_JOINT_PIN) {
			c = cpPinJointNew(a, b, cpv(P[0], P[1]), cpv(P[2], P[3]));
			cpPinJointSetDist(c, P[4]);
		}
		/* the groove's normal is the reference's own (cpGrooveJointInit), not the description's */
		else if (t == CPGPU_JOINT_GROOVE) c = cpGrooveJointNew(a, b, cpv(P[0], P[1]), cpv(P[2], P[3]), cpv(P[4], P[5]));
		else if (t == CPGPU_JOINT_ROTARY_LIMIT) c = cpRotaryLimitJointNew(a, b, P[0], P[1]);
		else if (t == CPGPU_JOINT_RATCHET) c = cpRatchetJointNew(a, b, P[0], P[1]);
		else if (t == CPGPU_JOINT_SIMPLE_MOTOR) c = cpSimpleMotorNew(a, b, P[0]);
		else if (t == CPGPU_JOINT_DAMPED_ROTARY_SPRING) c = cpDampedRotarySpringNew(a, b, P[0], P[1], P[2]);
		else {
			c = cpGearJointNew(a, b, P[0], P[1]);
			((cpGearJoint *)c)->ratio_inv = P[2];
		}
		cpConstraintSetMaxForce(c, d->cons_max_force[i]);
		cpConstraintSetErrorBias(c, d->cons_error_bias[i]);
		cpConstraintSetMaxBias(c, d->cons_max_bias[i]);
		cpConstraintSetCollideBodies(c, (cpBool)d->cons_collide_bodies[i]);
		{
			/* the carried state; no description state is zero (a ratchet's angle too) */
			double j0 = d->cons_jacc ? d->cons_jacc[2 * i] : 0.0, j1 = d->cons_jacc ? d->cons_jacc[2 * i + 1] : 0.0;
			if (t == CPGPU_JOINT_PIVOT) ((cpPivotJoint *)c)->jAcc = cpv(j0, j1);
			else if (t == CPGPU_JOINT_DAMPED_SPRING) ((cpDampedSpring *)c)->jAcc = j0;
			else if (t == CPGPU_JOINT_SLIDE) ((cpSlideJoint *)c)->jnAcc = j0;
			else if (t == CPGPU_JOINT_PIN) ((cpPinJoint *)c)->jnAcc = j0;
			else if (t == CPGPU_JOINT_GROOVE) ((cpGrooveJoint *)c)->jAcc = cpv(j0, j1);
			else if (t == CPGPU_JOINT_ROTARY_LIMIT) ((cpRotaryLimitJoint *)c)->jAcc = j0;
			else if (t == CPGPU_JOINT_RATCHET) { ((cpRatchetJoint *)c)->jAcc = j0; cpRatchetJointSetAngle(c, j1); }
			else if (t == CPGPU_JOINT_SIMPLE_MOTOR) ((cpSimpleMotor *)c)->jAcc = j0;
			else if (t == CPGPU_JOINT_DAMPED_ROTARY_SPRING) ((cpDampedRotarySpring *)c)->jAcc = j0;
			else ((cpGearJoint *)c)->jAcc = j0;
		}
		c->userData = (cpDataPointer)(intptr_t)i;
		cpSpaceAddConstraint(space, c);
		w->cons[k] = c;
		w->ctype[k] = t;
	}
	if (d->n_handlers > 0 && d->handlers) rb_add_handlers(w, d, wi);
	return w;
}

void rb_free(rbWorld *w)
{
	if (!w) return;
	for (int k = 0; k < w->nc; k++) { cpSpaceRemoveConstraint(w->space, w->cons[k]); cpConstraintFree(w->cons[k]); }
	for (int k = 0; k < w->ns; k++) {
		if (!w->shapes[k]) continue;
		cpSpaceRemoveShape(w->space, w->shapes[k]);
		cpShapeFree(w->shapes[k]);
	}
	for (int k = 0; k < w->nb; k++) {
		if (!w->bodies[k]) continue;
		cpSpaceRemoveBody(w->space, w->bodies[k]);
		cpBodyFree(w->bodies[k]);
	}
	if (w->owns_space) {
		if (w->hasty) cpHastySpaceFree(w->space); else cpSpaceFree(w->space);
	}
	free(w->bodies); free(w->shapes); free(w->cons); free(w->ctype);
	free(w);
}

void rb_step(rbWorld *w, double dt, int n)
{
	for (int k = 0; k < n; k++) {
		if (w->hasty) cpHastySpaceStep(w->space, dt);
		else cpSpaceStep(w->space, dt);
	}
}

/* Body state of this world, arrays indexed by local body index. */
void rb_read_bodies(rbWorld *w, double *p, double *v, double *a, double *wv, double *v_bias, double *w_bias,
                    double *transform)
{
	for (int k = 0; k < w->nb; k++) {
		cpBody *b = w->bodies[k];
		if (!b) continue;   /* removed (rb_remove_body) */
		if (p) { p[2 * k] = b->p.x; p[2 * k + 1] = b->p.y; }
		if (v) { v[2 * k] = b->v.x; v[2 * k + 1] = b->v.y; }
		if (a) a[k] = b->a;
		if (wv) wv[k] = b->w;
		if (v_bias) { v_bias[2 * k] = b->v_bias.x; v_bias[2 * k + 1] = b->v_bias.y; }
		if (w_bias) w_bias[k] = b->w_bias;
		if (transform) {
			cpTransform t = b->transform;
			double *o = transform + 6 * k;
			o[0] = t.a; o[1] = t.b; o[2] = t.c; o[3] = t.d; o[4] = t.tx; o[5] = t.ty;
		}
	}
}

/* cpSpace.arbiters after the last step, in solve order; same layout as
 * cpgpuWorldReadArbiters (shape indices are global description indices). */
int64_t rb_read_arbiters(rbWorld *w, int64_t cap, int32_t *shapes, int32_t *state, int32_t *count,
                         double *normal, double *contact, uint64_t *hash)
{
	cpArray *arr = w->space->arbiters;
	int64_t n = arr->num;
	for (int64_t k = 0; k < n && k < cap; k++) {
		cpArbiter *arb = (cpArbiter *)arr->arr[k];
		if (shapes) {
			shapes[2 * k] = (int32_t)(intptr_t)arb->a->userData;
			shapes[2 * k + 1] = (int32_t)(intptr_t)arb->b->userData;
		}
		if (state) state[k] = (int32_t)arb->state;
		if (count) count[k] = arb->count;
		if (normal) { normal[2 * k] = arb->n.x; normal[2 * k + 1] = arb->n.y; }
		for (int i = 0; i < 2; i++) {
			int valid = i < arb->count;
			struct cpContact *c = valid ? &arb->contacts[i] : NULL;
			if (contact) {
				double *o = contact + 24 * k + 12 * i;
				if (valid) {
					double vals[12] = {c->r1.x, c->r1.y, c->r2.x, c->r2.y, c->nMass, c->tMass, c->bounce,
					                   c->jnAcc, c->jtAcc, c->jBias, c->bias, 0.0};
					memcpy(o, vals, sizeof(vals));
				} else memset(o, 0, 12 * sizeof(double));
			}
			if (hash) hash[2 * k + i] = valid ? (uint64_t)c->hash : 0;
		}
	}
	return n;
}

int64_t rb_num_cached_arbiters(rbWorld *w) { return cpHashSetCount(w->space->cachedArbiters); }

/* Constraint accumulated impulses, [nc][2], description order. */
void rb_read_constraints(rbWorld *w, double *jacc)
{
	for (int k = 0; k < w->nc; k++) {
		cpConstraint *c = w->cons[k];
		double j0 = 0, j1 = 0;
		int t = w->ctype[k];
		if (t == CPGPU_JOINT_PIVOT) { j0 = ((cpPivotJoint *)c)->jAcc.x; j1 = ((cpPivotJoint *)c)->jAcc.y; }
		else if (t == CPGPU_JOINT_DAMPED_SPRING) j0 = ((cpDampedSpring *)c)->jAcc;
		else if (t == CPGPU_JOINT_SLIDE) j0 = ((cpSlideJoint *)c)->jnAcc;
		else if (t == CPGPU_JOINT_PIN) j0 = ((cpPinJoint *)c)->jnAcc;
		else if (t == CPGPU_JOINT_GROOVE) { j0 = ((cpGrooveJoint *)c)->jAcc.x; j1 = ((cpGrooveJoint *)c)->jAcc.y; }
		else if (t == CPGPU_JOINT_ROTARY_LIMIT) j0 = ((cpRotaryLimitJoint *)c)->jAcc;
		else if (t == CPGPU_JOINT_RATCHET) { j0 = ((cpRatchetJoint *)c)->jAcc; j1 = cpRatchetJointGetAngle(c); }
		else if (t == CPGPU_JOINT_SIMPLE_MOTOR) j0 = ((cpSimpleMotor *)c)->jAcc;
		else if (t == CPGPU_JOINT_DAMPED_ROTARY_SPRING) j0 = ((cpDampedRotarySpring *)c)->jAcc;
		else j0 = ((cpGearJoint *)c)->jAcc;
		jacc[2 * k] = j0; jacc[2 * k + 1] = j1;
	}
}

/* Set body transforms (local indices) and re-cache every shape with them
 * through the public cpShapeUpdate (cpShape.c:216-220). */
void rb_set_transforms(rbWorld *w, const double *transform)
{
	for (int k = 0; k < w->nb; k++) w->bodies[k]->transform = xf(transform + 6 * k);
	for (int k = 0; k < w->ns; k++) cpShapeUpdate(w->shapes[k], w->shapes[k]->body->transform);
}

void rb_read_shape_bbs(rbWorld *w, double *bb)
{
	for (int k = 0; k < w->ns; k++) {
		cpBB b = w->shapes[k]->bb;
		bb[4 * k] = b.l; bb[4 * k + 1] = b.b; bb[4 * k + 2] = b.r; bb[4 * k + 3] = b.t;
	}
}

/* cpCollide on two shapes (local indices) as they are cached. */
int rb_collide(rbWorld *w, int32_t a, int32_t b, uint32_t id, int32_t *ab_out, uint32_t *id_out, double *n_out,
               double *points_out, uint64_t *hash_out)
{
	struct cpContact contacts[CP_MAX_CONTACTS_PER_ARBITER];
	struct cpCollisionInfo info = cpCollide(w->shapes[a], w->shapes[b], id, contacts);
	ab_out[0] = (int32_t)(intptr_t)info.a->userData;
	ab_out[1] = (int32_t)(intptr_t)info.b->userData;
	*id_out = info.id;
	n_out[0] = info.n.x; n_out[1] = info.n.y;
	for (int i = 0; i < info.count; i++) {
		points_out[4 * i + 0] = contacts[i].r1.x; points_out[4 * i + 1] = contacts[i].r1.y;
		points_out[4 * i + 2] = contacts[i].r2.x; points_out[4 * i + 3] = contacts[i].r2.y;
		hash_out[i] = (uint64_t)contacts[i].hash;
	}
	return info.count;
}

/* The reference's own cpShapesCollide (cpShape.c:260-283) on world-local
 * shapes a and b: writes the contact point set's normal and, per point,
 * {pointA.x, pointA.y, pointB.x, pointB.y, distance}; returns the count. */
int rb_shapes_collide(rbWorld *w, int32_t a, int32_t b, double *n_out, double *points_out)
{
	cpContactPointSet set = cpShapesCollide(w->shapes[a], w->shapes[b]);
	n_out[0] = set.normal.x; n_out[1] = set.normal.y;
	for (int i = 0; i < set.count; i++) {
		points_out[5 * i + 0] = set.points[i].pointA.x; points_out[5 * i + 1] = set.points[i].pointA.y;
		points_out[5 * i + 2] = set.points[i].pointB.x; points_out[5 * i + 3] = set.points[i].pointB.y;
		points_out[5 * i + 4] = set.points[i].distance;
	}
	return set.count;
}

/* The reference broadphase: cpSpatialIndexReindexQuery over the dynamic
 * BBTree (cpBBTree.c:638-654, which also collides against the static tree)
 * filtered by the same QueryReject predicate cpSpaceCollideShapes applies
 * (cpSpaceStep.c:204-232).  Shapes must have been cached first.  Writes
 * (min, max) global index pairs; returns the count. */
typedef struct { rbWorld *w; int64_t cap, n; int32_t *out; } PairCtx;

static cpBool reject_constraint(cpBody *a, cpBody *b)
{
	CP_BODY_FOREACH_CONSTRAINT(a, constraint) {
		if (!constraint->collideBodies &&
		    ((constraint->a == a && constraint->b == b) || (constraint->a == b && constraint->b == a))) return cpTrue;
	}
	return cpFalse;
}

static cpCollisionID pair_recorder(cpShape *a, cpShape *b, cpCollisionID id, void *data)
{
	PairCtx *ctx = (PairCtx *)data;
	if (!cpBBIntersects(a->bb, b->bb) || a->body == b->body || cpShapeFilterReject(a->filter, b->filter) ||
	    reject_constraint(a->body, b->body)) return id;
	if (a->body->m == INFINITY && b->body->m == INFINITY) return id;
	int32_t ia = (int32_t)(intptr_t)a->userData, ib = (int32_t)(intptr_t)b->userData;
	if (ctx->n < ctx->cap) {
		ctx->out[2 * ctx->n] = ia < ib ? ia : ib;
		ctx->out[2 * ctx->n + 1] = ia < ib ? ib : ia;
	}
	ctx->n++;
	return id;
}

int64_t rb_candidate_pairs(rbWorld *w, int64_t cap, int32_t *out)
{
	PairCtx ctx = {w, cap, 0, out};
	cpSpatialIndexReindexQuery(w->space->dynamicShapes, (cpSpatialIndexQueryFunc)pair_recorder, &ctx);
	return ctx.n;
}

/* The reference integrators on one body (local index) with libm trig. */
void rb_body_update_position(rbWorld *w, int k, double dt) { cpBodyUpdatePosition(w->bodies[k], dt); }
void rb_body_update_velocity(rbWorld *w, int k, double gx, double gy, double damping, double dt)
{
	cpBodyUpdateVelocity(w->bodies[k], cpv(gx, gy), damping, dt);
}

int64_t rb_space_stamp(rbWorld *w) { return (int64_t)w->space->stamp; }

/* ---- collision log: every cpCollide made by cpSpaceCollideShapes
 * (cpSpaceStep.c:242), linked through -Wl,--wrap=cpCollide.  Records the
 * oriented shapes (after cpCollide's type sort), the warm-start id the
 * BBTree pair supplied, and the id returned. */
struct cpCollisionInfo __real_cpCollide(const cpShape *a, const cpShape *b, cpCollisionID id, struct cpContact *contacts);

static int g_log_on;
static int64_t g_log_n, g_log_cap;
static int32_t *g_log;

struct cpCollisionInfo __wrap_cpCollide(const cpShape *a, const cpShape *b, cpCollisionID id, struct cpContact *contacts)
{
	struct cpCollisionInfo info = __real_cpCollide(a, b, id, contacts);
	if (g_log_on) {
		if (g_log_n == g_log_cap) {
			g_log_cap = g_log_cap ? 2 * g_log_cap : 4096;
			g_log = (int32_t *)realloc(g_log, (size_t)g_log_cap * 4 * sizeof(int32_t));
		}
		int32_t *e = g_log + 4 * g_log_n++;
		e[0] = (int32_t)(intptr_t)info.a->userData;
		e[1] = (int32_t)(intptr_t)info.b->userData;
		e[2] = (int32_t)id;
		e[3] = (int32_t)info.id;
	}
	return info;
}

void rb_log_collisions(int on) { g_log_on = on; g_log_n = 0; }

/* Copies up to cap entries of {a, b, id_in, id_out}; returns the count. */
int64_t rb_read_collision_log(int64_t cap, int32_t *out)
{
	int64_t n = g_log_n < cap ? g_log_n : cap;
	if (out && n) memcpy(out, g_log, (size_t)n * 4 * sizeof(int32_t));
	return g_log_n;
}

/* cpSpaceRemoveShape / cpSpaceRemoveBody (cpSpace.c:523-555) of one object of
 * the description, which is then freed and forgotten. */
void rb_remove_shape(rbWorld *w, int k)
{
	cpSpaceRemoveShape(w->space, w->shapes[k]);
	cpShapeFree(w->shapes[k]);
	w->shapes[k] = NULL;
}

void rb_remove_body(rbWorld *w, int k)
{
	cpSpaceRemoveBody(w->space, w->bodies[k]);
	cpBodyFree(w->bodies[k]);
	w->bodies[k] = NULL;
}

/* ---- objects and raw state for the cpGPUSpace tests ---- */

cpSpace *rb_space(rbWorld *w) { return w->space; }
cpBody *rb_body(rbWorld *w, int k) { return w->bodies[k]; }
cpShape *rb_shape(rbWorld *w, int k) { return w->shapes[k]; }
cpConstraint *rb_constraint(rbWorld *w, int k) { return w->cons[k]; }

/* The body's private state fields: p, v, f (2 each), a, w, t, v_bias (2),
 * w_bias, transform (6): 17 doubles. */
void rb_body_raw(const cpBody *b, double *out)
{
	out[0] = b->p.x; out[1] = b->p.y; out[2] = b->v.x; out[3] = b->v.y; out[4] = b->f.x; out[5] = b->f.y;
	out[6] = b->a; out[7] = b->w; out[8] = b->t; out[9] = b->v_bias.x; out[10] = b->v_bias.y; out[11] = b->w_bias;
	out[12] = b->transform.a; out[13] = b->transform.b; out[14] = b->transform.c;
	out[15] = b->transform.d; out[16] = b->transform.tx; out[17] = b->transform.ty;
}

/* A constraint's accumulated impulse as the description stores it. */
void rb_constraint_jacc(const cpConstraint *c, double *out)
{
	out[0] = out[1] = 0.0;
	if (cpConstraintIsPivotJoint(c)) { out[0] = ((cpPivotJoint *)c)->jAcc.x; out[1] = ((cpPivotJoint *)c)->jAcc.y; }
	else if (cpConstraintIsDampedSpring(c)) out[0] = ((cpDampedSpring *)c)->jAcc;
	else if (cpConstraintIsSlideJoint(c)) out[0] = ((cpSlideJoint *)c)->jnAcc;
	else if (cpConstraintIsGearJoint(c)) out[0] = ((cpGearJoint *)c)->jAcc;
	else if (cpConstraintIsPinJoint(c)) out[0] = ((cpPinJoint *)c)->jnAcc;
	else if (cpConstraintIsGrooveJoint(c)) { out[0] = ((cpGrooveJoint *)c)->jAcc.x; out[1] = ((cpGrooveJoint *)c)->jAcc.y; }
	else if (cpConstraintIsRotaryLimitJoint(c)) out[0] = ((cpRotaryLimitJoint *)c)->jAcc;
	else if (cpConstraintIsRatchetJoint(c)) { out[0] = ((cpRatchetJoint *)c)->jAcc; out[1] = cpRatchetJointGetAngle(c); }
	else if (cpConstraintIsSimpleMotor(c)) out[0] = ((cpSimpleMotor *)c)->jAcc;
	else if (cpConstraintIsDampedRotarySpring(c)) out[0] = ((cpDampedRotarySpring *)c)->jAcc;
}

typedef struct ArbList {
	int n, cap;
	const cpShape **a, **b;
	int *count, *swapped;
	double *jn, *jt, *impulse;
	double *points;   /* per arbiter: normal, then per point pointA, pointB, distance */
} ArbList;

static void arb_visit(cpBody *body, cpArbiter *arb, void *data)
{
	ArbList *L = (ArbList *)data;
	const int k = L->n++;
	if (k >= L->cap) return;
	CP_ARBITER_GET_SHAPES(arb, sa, sb);
	L->a[k] = sa; L->b[k] = sb;
	L->count[k] = cpArbiterGetCount(arb);
	L->swapped[k] = arb->swapped;
	L->jn[k] = L->jt[k] = 0.0;
	for (int i = 0; i < arb->count; i++) { L->jn[k] += arb->contacts[i].jnAcc; L->jt[k] += arb->contacts[i].jtAcc; }
	cpVect j = cpArbiterTotalImpulse(arb);
	L->impulse[2 * k] = j.x; L->impulse[2 * k + 1] = j.y;
	cpContactPointSet set = cpArbiterGetContactPointSet(arb);
	double *q = L->points + 12 * k;
	memset(q, 0, 12 * sizeof(double));
	q[0] = set.normal.x; q[1] = set.normal.y;
	for (int i = 0; i < set.count; i++) {
		q[2 + 5 * i] = set.points[i].pointA.x; q[3 + 5 * i] = set.points[i].pointA.y;
		q[4 + 5 * i] = set.points[i].pointB.x; q[5 + 5 * i] = set.points[i].pointB.y;
		q[6 + 5 * i] = set.points[i].distance;
	}
}

/* cpBodyEachArbiter (cpBody.c) over one body: each arbiter's shapes in the
 * body's orientation (CP_ARBITER_GET_SHAPES), count, summed jnAcc / jtAcc,
 * cpArbiterTotalImpulse and cpArbiterGetContactPointSet.  Returns the count. */
int rb_body_arbiters(cpBody *body, int cap, const cpShape **a, const cpShape **b, int *count, int *swapped, double *jn,
                     double *jt, double *impulse, double *points)
{
	ArbList L = {0, cap, a, b, count, swapped, jn, jt, impulse, points};
	cpBodyEachArbiter(body, arb_visit, &L);
	return L.n;
}

/* ---- spatial queries through the reference's public API (cpSpaceQuery.c) ---- */

static cpShapeFilter query_filter(uint64_t group, uint32_t cats, uint32_t mask)
{
	return cpShapeFilterNew((cpGroup)(uintptr_t)group, cats, mask);
}

/* cpSpacePointQueryNearest: description index of the shape (-1: none), info
 * {point.x, point.y, distance, gradient.x, gradient.y} */
int32_t rb_point_query_nearest(rbWorld *w, const double *p, double max_distance, uint64_t group, uint32_t cats,
                               uint32_t mask, double *info)
{
	cpPointQueryInfo q;
	cpShape *s = cpSpacePointQueryNearest(w->space, cpv(p[0], p[1]), max_distance, query_filter(group, cats, mask), &q);
	info[0] = q.point.x; info[1] = q.point.y; info[2] = q.distance; info[3] = q.gradient.x; info[4] = q.gradient.y;
	return s ? (int32_t)(intptr_t)s->userData : -1;
}

/* cpSpaceSegmentQueryFirst: shape index (-1: none), info {point, normal, alpha} */
int32_t rb_segment_query_first(rbWorld *w, const double *a, const double *b, double radius, uint64_t group,
                               uint32_t cats, uint32_t mask, double *info)
{
	cpSegmentQueryInfo q;
	cpShape *s = cpSpaceSegmentQueryFirst(w->space, cpv(a[0], a[1]), cpv(b[0], b[1]), radius,
	                                      query_filter(group, cats, mask), &q);
	info[0] = q.point.x; info[1] = q.point.y; info[2] = q.normal.x; info[3] = q.normal.y; info[4] = q.alpha;
	return s ? (int32_t)(intptr_t)s->userData : -1;
}

typedef struct QueryHits {
	int64_t n, cap;
	int32_t *shapes;
	double *info;
} QueryHits;

static void point_hit(cpShape *shape, cpVect point, cpFloat distance, cpVect gradient, void *data)
{
	QueryHits *h = (QueryHits *)data;
	if (h->n < h->cap) {
		h->shapes[h->n] = (int32_t)(intptr_t)shape->userData;
		double *o = h->info + 5 * h->n;
		o[0] = point.x; o[1] = point.y; o[2] = distance; o[3] = gradient.x; o[4] = gradient.y;
	}
	h->n++;
}

static void segment_hit(cpShape *shape, cpVect point, cpVect normal, cpFloat alpha, void *data)
{
	QueryHits *h = (QueryHits *)data;
	if (h->n < h->cap) {
		h->shapes[h->n] = (int32_t)(intptr_t)shape->userData;
		double *o = h->info + 5 * h->n;
		o[0] = point.x; o[1] = point.y; o[2] = normal.x; o[3] = normal.y; o[4] = alpha;
	}
	h->n++;
}

static void bb_hit(cpShape *shape, void *data)
{
	QueryHits *h = (QueryHits *)data;
	if (h->n < h->cap) h->shapes[h->n] = (int32_t)(intptr_t)shape->userData;
	h->n++;
}

/* cpSpacePointQuery / cpSpaceSegmentQuery / cpSpaceBBQuery: every shape the
 * callback saw, in the order it saw them; returns the count */
int64_t rb_point_query(rbWorld *w, const double *p, double max_distance, uint64_t group, uint32_t cats, uint32_t mask,
                       int64_t cap, int32_t *shapes, double *info)
{
	QueryHits h = {0, cap, shapes, info};
	cpSpacePointQuery(w->space, cpv(p[0], p[1]), max_distance, query_filter(group, cats, mask), point_hit, &h);
	return h.n;
}

int64_t rb_segment_query(rbWorld *w, const double *a, const double *b, double radius, uint64_t group, uint32_t cats,
                         uint32_t mask, int64_t cap, int32_t *shapes, double *info)
{
	QueryHits h = {0, cap, shapes, info};
	cpSpaceSegmentQuery(w->space, cpv(a[0], a[1]), cpv(b[0], b[1]), radius, query_filter(group, cats, mask),
	                    segment_hit, &h);
	return h.n;
}

typedef struct ShapeHits {
	int64_t n, cap;
	int32_t *shapes, *counts;
	double *sets;
} ShapeHits;

static void shape_hit(cpShape *shape, cpContactPointSet *set, void *data)
{
	ShapeHits *h = (ShapeHits *)data;
	if (h->n < h->cap) {
		h->shapes[h->n] = (int32_t)(intptr_t)shape->userData;
		h->counts[h->n] = set->count;
		double *o = h->sets + 12 * h->n;
		memset(o, 0, 12 * sizeof(double));
		o[0] = set->normal.x; o[1] = set->normal.y;
		for (int i = 0; i < set->count; i++) {
			o[2 + 5 * i + 0] = set->points[i].pointA.x; o[2 + 5 * i + 1] = set->points[i].pointA.y;
			o[2 + 5 * i + 2] = set->points[i].pointB.x; o[2 + 5 * i + 3] = set->points[i].pointB.y;
			o[2 + 5 * i + 4] = set->points[i].distance;
		}
	}
	h->n++;
}

/* cpSpaceShapeQuery with the world's shape k (local index): every shape the
 * callback saw, in the order it saw them; *any: its return value */
int64_t rb_shape_query(rbWorld *w, int32_t k, int64_t cap, int32_t *shapes, int32_t *counts, double *sets,
                       int32_t *any)
{
	ShapeHits h = {0, cap, shapes, counts, sets};
	const cpBool r = cpSpaceShapeQuery(w->space, w->shapes[k], shape_hit, &h);
	if (any) *any = r ? 1 : 0;
	return h.n;
}

int64_t rb_bb_query(rbWorld *w, const double *bb, uint64_t group, uint32_t cats, uint32_t mask, int64_t cap,
                    int32_t *shapes)
{
	QueryHits h = {0, cap, shapes, NULL};
	cpSpaceBBQuery(w->space, cpBBNew(bb[0], bb[1], bb[2], bb[3]), query_filter(group, cats, mask), bb_hit, &h);
	return h.n;
}
