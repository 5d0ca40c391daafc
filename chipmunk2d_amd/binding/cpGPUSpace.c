/*
 * cpGPUSpace.c -- cpGPUSpace and cpGPUBatch (include/cpGPUSpace.h): Chipmunk
 * spaces stepped by the gfx950 step library (include/cpgpu.h).
 *
 * The reference-side binding a Chipmunk maintainer adds beside cpHastySpace.c.
 * It reads the private structs (chipmunk_structs.h) the way cpHastySpace
 * does, packs them into the library's structure-of-arrays description, and
 * writes device state back into them on cpGPUSpaceSyncHost.  It is compiled
 * against Chipmunk's headers and linked with Chipmunk and libcpgpu.so.
 *
 * A pack collects, per space: the bodies of space->dynamicBodies, then of
 * space->staticBodies (static and kinematic, cpSpace.c:405-415), then the
 * space's built-in static body if a shape or constraint uses it; the shapes of
 * both spatial indexes in the order they were added (their hashid,
 * cpSpace.c:432); the constraints in space->constraints order.
 *
 * Before every step the spaces are collected again and compared with the
 * pack:
 *   - the same objects with the same constant fields (masses, geometry,
 *     materials, filters, joint parameters): the bodies whose state the caller
 *     changed since the device state was last written back (a field that no
 *     longer equals the value the binding last gave it) are written to the
 *     device, and changed space parameters are set;
 *   - anything else (objects added or removed, a constant changed, a static
 *     body moved): the spaces are packed again with the device's body state
 *     and joint impulses for every field the caller did not change, and the
 *     new device world takes over the old one's cached arbiters and warm-start
 *     ids (cpgpuWorldCreateFrom).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "chipmunk/chipmunk_private.h"
#include "cpGPUSpace.h"

/* ---- errors (cpAssertHard's role, include/chipmunk/chipmunk.h:48-58) ---- */

static int abort_on_error = 1;
static char last_error[512];

void cpGPUSetAbortOnError(int enable) { abort_on_error = enable; }
const char *cpGPUGetError(void) { return last_error; }
void cpGPUClearError(void) { last_error[0] = 0; }

static int report(const char *what, const char *detail)
{
	snprintf(last_error, sizeof(last_error), "%s%s%s", what, detail ? ": " : "", detail ? detail : "");
	if (abort_on_error) {
		fprintf(stderr, "Aborting due to Chipmunk error: %s\n", last_error);
		abort();
	}
	return -1;
}

static int lib_error(const char *call, int rc)
{
	char buf[400];
	snprintf(buf, sizeof(buf), "%s returned %d: %s", call, rc, cpgpuGetLastError());
	return report("cpGPUSpace", buf);
}

/* ---- growable arrays ---- */

#define GROW(ptr, cap, need) do { \
	if ((need) > (cap)) { \
		size_t c_ = (cap) ? (size_t)(cap) : 16; \
		while (c_ < (size_t)(need)) c_ *= 2; \
		(ptr) = realloc((ptr), c_ * sizeof(*(ptr))); \
		(cap) = (int)c_; \
	} \
} while (0)

/* ---- pointer -> index map (open addressing) ---- */

typedef struct PtrMap {
	const void **key;
	int *val;
	int mask;
} PtrMap;

static uint64_t ptr_hash(const void *p)
{
	uint64_t z = (uint64_t)(uintptr_t)p;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}

static void map_reset(PtrMap *m, int n)
{
	int cap = 64;
	while (cap < 2 * n + 16) cap *= 2;
	if (cap - 1 != m->mask) {
		free(m->key);
		free(m->val);
		m->key = (const void **)malloc((size_t)cap * sizeof(void *));
		m->val = (int *)malloc((size_t)cap * sizeof(int));
		m->mask = cap - 1;
	}
	memset(m->key, 0, (size_t)cap * sizeof(void *));
}

static void map_put(PtrMap *m, const void *p, int v)
{
	uint64_t h = ptr_hash(p) & (uint64_t)m->mask;
	while (m->key[h] && m->key[h] != p) h = (h + 1) & (uint64_t)m->mask;
	m->key[h] = p;
	m->val[h] = v;
}

static int map_get(const PtrMap *m, const void *p)
{
	if (!m->key || !p) return -1;
	uint64_t h = ptr_hash(p) & (uint64_t)m->mask;
	while (m->key[h]) {
		if (m->key[h] == p) return m->val[h];
		h = (h + 1) & (uint64_t)m->mask;
	}
	return -1;
}

static void map_free(PtrMap *m)
{
	free(m->key);
	free(m->val);
	memset(m, 0, sizeof(*m));
}

/* ---- the packed description and the objects behind it ---- */

enum { BODY_P, BODY_V, BODY_F, BODY_A, BODY_W, BODY_T, BODY_VB, BODY_WB, BODY_XF, BODY_FIELDS };
static const int body_width[BODY_FIELDS] = {2, 2, 2, 1, 1, 1, 2, 1, 6};

typedef struct Desc {
	int nw, nb, ns, nc, npv;
	int cap_b, cap_s, cap_c, cap_pv, cap_w;
	cpBody **bodies;
	cpShape **shapes;
	cpConstraint **cons;
	cpgpuSpaceParams *sp;
	int32_t *wbs, *wss, *wcs;
	int32_t *body_type;
	double *body_m, *body_m_inv, *body_i, *body_i_inv, *body_cog;
	double *state[BODY_FIELDS];          /* body state, body_width[f] scalars a body */
	int32_t *shape_type, *shape_body, *shape_pstart, *shape_pcount;
	uint64_t *shape_hashid, *shape_group;
	uint32_t *shape_cats, *shape_mask;
	uint8_t *shape_sensor;
	uint32_t *shape_ctype;               /* collision type, as an index into types */
	double *shape_e, *shape_u, *shape_sv, *shape_r, *shape_geom;
	double *pverts, *pnorms;
	int32_t *cons_type, *cons_a, *cons_b, *cons_collide;
	double *cons_max_force, *cons_error_bias, *cons_max_bias, *cons_params, *cons_jacc;
	/* collision handlers: the ones that are not Chipmunk's do-nothing defaults, with
	 * their types as indices into types */
	cpgpuHandlerDesc *handlers;
	int nh, cap_h;
	cpCollisionType *types;
	int ntypes, cap_t;
	cpgpuWorldDesc d;
} Desc;

static void desc_reserve(Desc *D, int nw, int nb, int ns, int nc, int npv)
{
	if (nw + 1 > D->cap_w) {
		int c = D->cap_w;
		GROW(D->sp, c, nw + 1);
		c = D->cap_w; GROW(D->wbs, c, nw + 1);
		c = D->cap_w; GROW(D->wss, c, nw + 1);
		c = D->cap_w; GROW(D->wcs, c, nw + 1);
		D->cap_w = c;
	}
	if (nb > D->cap_b) {
		int c = D->cap_b;
		GROW(D->bodies, c, nb);
		D->body_type = realloc(D->body_type, (size_t)c * sizeof(int32_t));
		D->body_m = realloc(D->body_m, (size_t)c * sizeof(double));
		D->body_m_inv = realloc(D->body_m_inv, (size_t)c * sizeof(double));
		D->body_i = realloc(D->body_i, (size_t)c * sizeof(double));
		D->body_i_inv = realloc(D->body_i_inv, (size_t)c * sizeof(double));
		D->body_cog = realloc(D->body_cog, (size_t)c * 2 * sizeof(double));
		for (int f = 0; f < BODY_FIELDS; f++)
			D->state[f] = realloc(D->state[f], (size_t)c * body_width[f] * sizeof(double));
		D->cap_b = c;
	}
	if (ns > D->cap_s) {
		int c = D->cap_s;
		GROW(D->shapes, c, ns);
		D->shape_type = realloc(D->shape_type, (size_t)c * sizeof(int32_t));
		D->shape_body = realloc(D->shape_body, (size_t)c * sizeof(int32_t));
		D->shape_pstart = realloc(D->shape_pstart, (size_t)c * sizeof(int32_t));
		D->shape_pcount = realloc(D->shape_pcount, (size_t)c * sizeof(int32_t));
		D->shape_hashid = realloc(D->shape_hashid, (size_t)c * sizeof(uint64_t));
		D->shape_group = realloc(D->shape_group, (size_t)c * sizeof(uint64_t));
		D->shape_cats = realloc(D->shape_cats, (size_t)c * sizeof(uint32_t));
		D->shape_mask = realloc(D->shape_mask, (size_t)c * sizeof(uint32_t));
		D->shape_sensor = realloc(D->shape_sensor, (size_t)c * sizeof(uint8_t));
		D->shape_ctype = realloc(D->shape_ctype, (size_t)c * sizeof(uint32_t));
		D->shape_e = realloc(D->shape_e, (size_t)c * sizeof(double));
		D->shape_u = realloc(D->shape_u, (size_t)c * sizeof(double));
		D->shape_sv = realloc(D->shape_sv, (size_t)c * 2 * sizeof(double));
		D->shape_r = realloc(D->shape_r, (size_t)c * sizeof(double));
		D->shape_geom = realloc(D->shape_geom, (size_t)c * 10 * sizeof(double));
		D->cap_s = c;
	}
	if (nc > D->cap_c) {
		int c = D->cap_c;
		GROW(D->cons, c, nc);
		D->cons_type = realloc(D->cons_type, (size_t)c * sizeof(int32_t));
		D->cons_a = realloc(D->cons_a, (size_t)c * sizeof(int32_t));
		D->cons_b = realloc(D->cons_b, (size_t)c * sizeof(int32_t));
		D->cons_collide = realloc(D->cons_collide, (size_t)c * sizeof(int32_t));
		D->cons_max_force = realloc(D->cons_max_force, (size_t)c * sizeof(double));
		D->cons_error_bias = realloc(D->cons_error_bias, (size_t)c * sizeof(double));
		D->cons_max_bias = realloc(D->cons_max_bias, (size_t)c * sizeof(double));
		D->cons_params = realloc(D->cons_params, (size_t)c * 8 * sizeof(double));
		D->cons_jacc = realloc(D->cons_jacc, (size_t)c * 2 * sizeof(double));
		D->cap_c = c;
	}
	if (npv > D->cap_pv) {
		int c = D->cap_pv;
		GROW(D->pverts, c, npv);
		D->pnorms = realloc(D->pnorms, (size_t)c * 2 * sizeof(double));
		D->pverts = realloc(D->pverts, (size_t)c * 2 * sizeof(double));
		D->cap_pv = c;
	}
}

static void desc_free(Desc *D)
{
	void *p[] = {D->bodies, D->shapes, D->cons, D->sp, D->wbs, D->wss, D->wcs, D->body_type, D->body_m, D->body_m_inv,
	             D->body_i, D->body_i_inv, D->body_cog, D->shape_type, D->shape_body, D->shape_pstart, D->shape_pcount,
	             D->shape_hashid, D->shape_group, D->shape_cats, D->shape_mask, D->shape_sensor, D->shape_e, D->shape_u,
	             D->shape_sv,
	             D->shape_r, D->shape_geom, D->pverts, D->pnorms, D->cons_type, D->cons_a, D->cons_b, D->cons_collide,
	             D->cons_max_force, D->cons_error_bias, D->cons_max_bias, D->cons_params, D->cons_jacc,
	             D->shape_ctype, D->handlers, D->types};
	for (size_t k = 0; k < sizeof(p) / sizeof(p[0]); k++) free(p[k]);
	for (int f = 0; f < BODY_FIELDS; f++) free(D->state[f]);
	memset(D, 0, sizeof(*D));
}

static void desc_link(Desc *D)
{
	cpgpuWorldDesc *d = &D->d;
	memset(d, 0, sizeof(*d));
	d->n_worlds = D->nw;
	d->space = D->sp;
	d->world_body_start = D->wbs;
	d->world_shape_start = D->wss;
	d->world_constraint_start = D->wcs;
	d->n_bodies = D->nb;
	d->body_type = D->body_type;
	d->body_m = D->body_m; d->body_m_inv = D->body_m_inv;
	d->body_i = D->body_i; d->body_i_inv = D->body_i_inv;
	d->body_cog = D->body_cog;
	d->body_p = D->state[BODY_P]; d->body_v = D->state[BODY_V]; d->body_f = D->state[BODY_F];
	d->body_a = D->state[BODY_A]; d->body_w = D->state[BODY_W]; d->body_t = D->state[BODY_T];
	d->body_v_bias = D->state[BODY_VB]; d->body_w_bias = D->state[BODY_WB];
	d->body_transform = D->state[BODY_XF];
	d->n_shapes = D->ns;
	d->shape_type = D->shape_type; d->shape_body = D->shape_body; d->shape_hashid = D->shape_hashid;
	d->shape_e = D->shape_e; d->shape_u = D->shape_u; d->shape_surface_v = D->shape_sv;
	d->shape_group = D->shape_group; d->shape_categories = D->shape_cats; d->shape_mask = D->shape_mask;
	d->shape_sensor = D->shape_sensor;
	d->shape_ctype = D->nh > 0 ? D->shape_ctype : NULL;
	d->n_handlers = D->nh;
	d->handlers = D->nh > 0 ? D->handlers : NULL;
	d->shape_radius = D->shape_r; d->shape_geom = D->shape_geom;
	d->shape_poly_start = D->shape_pstart; d->shape_poly_count = D->shape_pcount;
	d->n_poly_verts = D->npv;
	d->poly_verts = D->pverts; d->poly_normals = D->pnorms;
	d->n_constraints = D->nc;
	d->cons_type = D->cons_type; d->cons_a = D->cons_a; d->cons_b = D->cons_b;
	d->cons_max_force = D->cons_max_force; d->cons_error_bias = D->cons_error_bias; d->cons_max_bias = D->cons_max_bias;
	d->cons_collide_bodies = D->cons_collide;
	d->cons_params = D->cons_params;
	d->cons_jacc = D->cons_jacc;
}

/* ---- reading one space ---- */

/* the damped springs' default force and torque, taken from freshly made objects
 * (their functions are private to cpDampedSpring.c and cpDampedRotarySpring.c) */
/* the defaults that only call the wildcard handlers: DefaultBegin ... DefaultSeparate (cpSpace.c:66-90) */
static cpCollisionBeginFunc wild_begin;
static cpCollisionPreSolveFunc wild_pre_solve;
static cpCollisionPostSolveFunc wild_post_solve;
static cpCollisionSeparateFunc wild_separate;
static cpDampedSpringForceFunc default_spring_force;
static cpDampedRotarySpringTorqueFunc default_spring_torque;
static int defaults_known;

static void learn_defaults(void)
{
	if (defaults_known) return;
	cpSpace *s = cpSpaceNew();
	const cpCollisionHandler *h = cpSpaceAddCollisionHandler(s, 1, 2);
	wild_begin = h->beginFunc;
	wild_pre_solve = h->preSolveFunc;
	wild_post_solve = h->postSolveFunc;
	wild_separate = h->separateFunc;
	cpSpaceFree(s);
	cpConstraint *c = cpDampedSpringNew(NULL, NULL, cpvzero, cpvzero, 0.0, 0.0, 0.0);
	default_spring_force = ((cpDampedSpring *)c)->springForceFunc;
	cpConstraintFree(c);
	c = cpDampedRotarySpringNew(NULL, NULL, 0.0, 0.0, 0.0);
	default_spring_torque = ((cpDampedRotarySpring *)c)->springTorqueFunc;
	cpConstraintFree(c);
	defaults_known = 1;
}


typedef struct Collect {
	cpShape **list;
	int n, cap;
} Collect;

static void collect_shape(void *obj, void *data)
{
	Collect *c = (Collect *)data;
	GROW(c->list, c->cap, c->n + 1);
	c->list[c->n++] = (cpShape *)obj;
}

static int cmp_hashid(const void *x, const void *y)
{
	const cpShape *a = *(cpShape *const *)x, *b = *(cpShape *const *)y;
	return a->hashid < b->hashid ? -1 : a->hashid > b->hashid;
}

typedef struct Maps {
	PtrMap body, shape, cons;
} Maps;

typedef struct Pack {
	int n_spaces;
	cpSpace **spaces;
	int precision;
	cpgpuWorld *world;
	int packs;
	Desc cur;                  /* the description the device world was made from */
	Desc fresh;                /* the spaces as collected before a step */
	Maps cur_at, fresh_at;     /* their objects -> index */
	Collect scratch;
	int32_t *shape_from;       /* last pack: old index of each shape, -1 new */
	int cap_from;
	/* the value of every body field and joint impulse as the binding last made the
	 * host struct and the device agree on it */
	Desc shadow;
	/* host arbiters of the last SyncHost */
	cpArbiter *arbs;
	struct cpContact *contacts;
	int narb, cap_arb;
	/* readback buffers */
	double *rb[BODY_FIELDS];
	int cap_rb;
	double *rb_jacc;
	int cap_rb_jacc;
	int32_t *idx;
	int cap_idx;
	double *wbuf[BODY_FIELDS];
	int cap_wbuf;
} Pack;

static int body_index(Desc *D, PtrMap *m, cpSpace *space, cpBody *body)
{
	int k = map_get(m, body);
	if (k >= 0) return k;
	if (body != space->staticBody) return -1;
	/* the space's built-in static body, packed when something uses it */
	desc_reserve(D, D->nw, D->nb + 1, D->ns, D->nc, D->npv);
	k = D->nb++;
	D->bodies[k] = body;
	map_put(m, body, k);
	return k;
}

static void read_body(Desc *D, int k, cpBody *b)
{
	D->body_type[k] = (int32_t)cpBodyGetType(b);
	D->body_m[k] = b->m; D->body_m_inv[k] = b->m_inv;
	D->body_i[k] = b->i; D->body_i_inv[k] = b->i_inv;
	D->body_cog[2 * k] = b->cog.x; D->body_cog[2 * k + 1] = b->cog.y;
	double *s;
	s = D->state[BODY_P] + 2 * k; s[0] = b->p.x; s[1] = b->p.y;
	s = D->state[BODY_V] + 2 * k; s[0] = b->v.x; s[1] = b->v.y;
	s = D->state[BODY_F] + 2 * k; s[0] = b->f.x; s[1] = b->f.y;
	D->state[BODY_A][k] = b->a;
	D->state[BODY_W][k] = b->w;
	D->state[BODY_T][k] = b->t;
	s = D->state[BODY_VB] + 2 * k; s[0] = b->v_bias.x; s[1] = b->v_bias.y;
	D->state[BODY_WB][k] = b->w_bias;
	s = D->state[BODY_XF] + 6 * k;
	s[0] = b->transform.a; s[1] = b->transform.b; s[2] = b->transform.c;
	s[3] = b->transform.d; s[4] = b->transform.tx; s[5] = b->transform.ty;
}

static void set_body_field(cpBody *b, int f, const double *v);

static void write_body(const Desc *D, int k, cpBody *b)
{
	for (int f = 0; f < BODY_FIELDS; f++) set_body_field(b, f, D->state[f] + (size_t)body_width[f] * k);
}

/* the index of a collision type in D->types (added when new) */
static uint32_t type_index(Desc *D, cpCollisionType t)
{
	if (t == CP_WILDCARD_COLLISION_TYPE) return CPGPU_WILDCARD;
	for (int i = 0; i < D->ntypes; i++)
		if (D->types[i] == t) return (uint32_t)i;
	GROW(D->types, D->cap_t, D->ntypes + 1);
	D->types[D->ntypes] = t;
	return (uint32_t)D->ntypes++;
}

/* The callbacks of a handler that are the caller's: neither Chipmunk's do-nothing
 * ones (cpSpace.c:100-106) nor the defaults that only call the wildcard handlers
 * (cpSpace.c:66-90), whose own entries bring their events. */
static uint32_t handler_events(const cpCollisionHandler *h)
{
	const cpCollisionHandler *none = &cpCollisionHandlerDoNothing;
	uint32_t ev = 0;
	if (h->beginFunc != none->beginFunc && h->beginFunc != wild_begin) ev |= CPGPU_EVENT_BEGIN;
	if (h->preSolveFunc != none->preSolveFunc && h->preSolveFunc != wild_pre_solve) ev |= CPGPU_EVENT_PRE_SOLVE;
	if (h->postSolveFunc != none->postSolveFunc && h->postSolveFunc != wild_post_solve) ev |= CPGPU_EVENT_POST_SOLVE;
	if (h->separateFunc != none->separateFunc && h->separateFunc != wild_separate) ev |= CPGPU_EVENT_SEPARATE;
	return ev;
}

static void add_handler(Desc *D, int space, uint32_t ta, uint32_t tb, uint32_t events)
{
	if (!events) return;
	GROW(D->handlers, D->cap_h, D->nh + 1);
	cpgpuHandlerDesc *h = &D->handlers[D->nh++];
	h->space = space; h->type_a = ta; h->type_b = tb; h->events = events;
}

struct HandlerCtx {
	Desc *D;
	int space;
};

/* cpSpaceAddCollisionHandler and cpSpaceAddWildcardHandler entries of cpSpace.collisionHandlers */
static void collect_handler(void *elt, void *data)
{
	const struct HandlerCtx *c = (const struct HandlerCtx *)data;
	const cpCollisionHandler *h = (const cpCollisionHandler *)elt;
	cpCollisionType a = h->typeA, b = h->typeB;
	if (a == CP_WILDCARD_COLLISION_TYPE && b != CP_WILDCARD_COLLISION_TYPE) { a = h->typeB; b = h->typeA; }
	add_handler(c->D, c->space, type_index(c->D, a), type_index(c->D, b), handler_events(h));
}

static int read_shape(Desc *D, int k, cpShape *sh, int body)
{
	memset(D->shape_geom + 10 * k, 0, 10 * sizeof(double));
	D->shape_type[k] = (int32_t)sh->klass->type;
	D->shape_body[k] = body;
	D->shape_hashid[k] = (uint64_t)sh->hashid;
	D->shape_e[k] = sh->e;
	D->shape_u[k] = sh->u;
	D->shape_sv[2 * k] = sh->surfaceV.x; D->shape_sv[2 * k + 1] = sh->surfaceV.y;
	D->shape_group[k] = (uint64_t)sh->filter.group;
	D->shape_cats[k] = (uint32_t)sh->filter.categories;
	D->shape_mask[k] = (uint32_t)sh->filter.mask;
	D->shape_sensor[k] = sh->sensor ? 1 : 0;
	D->shape_ctype[k] = type_index(D, sh->type);
	D->shape_pstart[k] = 0;
	D->shape_pcount[k] = 0;
	double *g = D->shape_geom + 10 * k;
	switch (sh->klass->type) {
	case CP_CIRCLE_SHAPE: {
		const cpCircleShape *c = (const cpCircleShape *)sh;
		D->shape_r[k] = c->r;
		g[0] = c->c.x; g[1] = c->c.y;
		break;
	}
	case CP_SEGMENT_SHAPE: {
		const cpSegmentShape *s = (const cpSegmentShape *)sh;
		D->shape_r[k] = s->r;
		g[0] = s->a.x; g[1] = s->a.y; g[2] = s->b.x; g[3] = s->b.y; g[4] = s->n.x; g[5] = s->n.y;
		g[6] = s->a_tangent.x; g[7] = s->a_tangent.y; g[8] = s->b_tangent.x; g[9] = s->b_tangent.y;
		break;
	}
	case CP_POLY_SHAPE: {
		const cpPolyShape *p = (const cpPolyShape *)sh;
		if (p->count > CPGPU_MAX_POLY_VERTS) return report("cpGPUSpace: polygon with too many vertices", NULL);
		D->shape_r[k] = p->r;
		desc_reserve(D, D->nw, D->nb, D->ns, D->nc, D->npv + p->count);
		D->shape_pstart[k] = D->npv;
		D->shape_pcount[k] = p->count;
		/* the untransformed planes follow the transformed ones (cpPolyShape.c:147-165) */
		for (int i = 0; i < p->count; i++) {
			const struct cpSplittingPlane *pl = &p->planes[p->count + i];
			D->pverts[2 * (D->npv + i)] = pl->v0.x; D->pverts[2 * (D->npv + i) + 1] = pl->v0.y;
			D->pnorms[2 * (D->npv + i)] = pl->n.x; D->pnorms[2 * (D->npv + i) + 1] = pl->n.y;
		}
		D->npv += p->count;
		break;
	}
	default:
		return report("cpGPUSpace: unknown shape type", NULL);
	}
	return 0;
}

static int read_constraint(Desc *D, int k, cpConstraint *c, int a, int b)
{
	if (c->preSolve || c->postSolve)
		return report("cpGPUSpace: constraint pre/post-solve callbacks are not supported on the device", NULL);
	D->cons_a[k] = a;
	D->cons_b[k] = b;
	D->cons_max_force[k] = c->maxForce;
	D->cons_error_bias[k] = c->errorBias;
	D->cons_max_bias[k] = c->maxBias;
	D->cons_collide[k] = c->collideBodies ? 1 : 0;
	double *P = D->cons_params + 8 * k, *J = D->cons_jacc + 2 * k;
	memset(P, 0, 8 * sizeof(double));
	J[0] = J[1] = 0.0;
	if (cpConstraintIsPivotJoint(c)) {
		const cpPivotJoint *j = (const cpPivotJoint *)c;
		D->cons_type[k] = CPGPU_JOINT_PIVOT;
		P[0] = j->anchorA.x; P[1] = j->anchorA.y; P[2] = j->anchorB.x; P[3] = j->anchorB.y;
		J[0] = j->jAcc.x; J[1] = j->jAcc.y;
	} else if (cpConstraintIsDampedSpring(c)) {
		const cpDampedSpring *j = (const cpDampedSpring *)c;
		if (j->springForceFunc != default_spring_force)
			return report("cpGPUSpace: a custom damped-spring force is not supported on the device", NULL);
		D->cons_type[k] = CPGPU_JOINT_DAMPED_SPRING;
		P[0] = j->anchorA.x; P[1] = j->anchorA.y; P[2] = j->anchorB.x; P[3] = j->anchorB.y;
		P[4] = j->restLength; P[5] = j->stiffness; P[6] = j->damping;
		J[0] = j->jAcc;
	} else if (cpConstraintIsSlideJoint(c)) {
		const cpSlideJoint *j = (const cpSlideJoint *)c;
		D->cons_type[k] = CPGPU_JOINT_SLIDE;
		P[0] = j->anchorA.x; P[1] = j->anchorA.y; P[2] = j->anchorB.x; P[3] = j->anchorB.y;
		P[4] = j->min; P[5] = j->max;
		J[0] = j->jnAcc;
	} else if (cpConstraintIsGearJoint(c)) {
		const cpGearJoint *j = (const cpGearJoint *)c;
		D->cons_type[k] = CPGPU_JOINT_GEAR;
		P[0] = j->phase; P[1] = j->ratio; P[2] = j->ratio_inv;
		J[0] = j->jAcc;
	} else if (cpConstraintIsPinJoint(c)) {
		const cpPinJoint *j = (const cpPinJoint *)c;
		D->cons_type[k] = CPGPU_JOINT_PIN;
		P[0] = j->anchorA.x; P[1] = j->anchorA.y; P[2] = j->anchorB.x; P[3] = j->anchorB.y;
		P[4] = j->dist;
		J[0] = j->jnAcc;
	} else if (cpConstraintIsGrooveJoint(c)) {
		const cpGrooveJoint *j = (const cpGrooveJoint *)c;
		D->cons_type[k] = CPGPU_JOINT_GROOVE;
		P[0] = j->grv_a.x; P[1] = j->grv_a.y; P[2] = j->grv_b.x; P[3] = j->grv_b.y;
		P[4] = j->anchorB.x; P[5] = j->anchorB.y; P[6] = j->grv_n.x; P[7] = j->grv_n.y;
		J[0] = j->jAcc.x; J[1] = j->jAcc.y;
	} else if (cpConstraintIsRotaryLimitJoint(c)) {
		const cpRotaryLimitJoint *j = (const cpRotaryLimitJoint *)c;
		D->cons_type[k] = CPGPU_JOINT_ROTARY_LIMIT;
		P[0] = j->min; P[1] = j->max;
		J[0] = j->jAcc;
	} else if (cpConstraintIsRatchetJoint(c)) {
		const cpRatchetJoint *j = (const cpRatchetJoint *)c;
		D->cons_type[k] = CPGPU_JOINT_RATCHET;
		P[0] = j->phase; P[1] = j->ratchet;
		J[0] = j->jAcc; J[1] = j->angle;
	} else if (cpConstraintIsSimpleMotor(c)) {
		const cpSimpleMotor *j = (const cpSimpleMotor *)c;
		D->cons_type[k] = CPGPU_JOINT_SIMPLE_MOTOR;
		P[0] = j->rate;
		J[0] = j->jAcc;
	} else if (cpConstraintIsDampedRotarySpring(c)) {
		const cpDampedRotarySpring *j = (const cpDampedRotarySpring *)c;
		if (j->springTorqueFunc != default_spring_torque)
			return report("cpGPUSpace: a custom damped-rotary-spring torque is not supported on the device", NULL);
		D->cons_type[k] = CPGPU_JOINT_DAMPED_ROTARY_SPRING;
		P[0] = j->restAngle; P[1] = j->stiffness; P[2] = j->damping;
		J[0] = j->jAcc;
	} else {
		return report("cpGPUSpace: a constraint class the device does not run", NULL);
	}
	return 0;
}

/* the carried state J of a constraint of packed type `type` into its struct */
static void write_jacc(int type, const double *J, cpConstraint *c)
{
	switch (type) {
	case CPGPU_JOINT_PIVOT: ((cpPivotJoint *)c)->jAcc = cpv(J[0], J[1]); break;
	case CPGPU_JOINT_DAMPED_SPRING: ((cpDampedSpring *)c)->jAcc = J[0]; break;
	case CPGPU_JOINT_SLIDE: ((cpSlideJoint *)c)->jnAcc = J[0]; break;
	case CPGPU_JOINT_GEAR: ((cpGearJoint *)c)->jAcc = J[0]; break;
	case CPGPU_JOINT_PIN: ((cpPinJoint *)c)->jnAcc = J[0]; break;
	case CPGPU_JOINT_GROOVE: ((cpGrooveJoint *)c)->jAcc = cpv(J[0], J[1]); break;
	case CPGPU_JOINT_ROTARY_LIMIT: ((cpRotaryLimitJoint *)c)->jAcc = J[0]; break;
	case CPGPU_JOINT_RATCHET: ((cpRatchetJoint *)c)->jAcc = J[0]; ((cpRatchetJoint *)c)->angle = J[1]; break;
	case CPGPU_JOINT_SIMPLE_MOTOR: ((cpSimpleMotor *)c)->jAcc = J[0]; break;
	case CPGPU_JOINT_DAMPED_ROTARY_SPRING: ((cpDampedRotarySpring *)c)->jAcc = J[0]; break;
	}
}

/* Collect every space into D (objects, constants and the host structs' state). */
static int collect(Pack *pk, Desc *D, Maps *M)
{
	PtrMap *bm = &M->body, *sm = &M->shape, *cm = &M->cons;
	learn_defaults();
	D->nw = pk->n_spaces;
	D->nh = 0;
	D->ntypes = 0;
	D->nb = D->ns = D->nc = D->npv = 0;
	desc_reserve(D, D->nw, 16, 16, 16, 16);
	int nb = 0, ns = 0, nc = 0;
	for (int w = 0; w < pk->n_spaces; w++) {
		cpSpace *sp = pk->spaces[w];
		nb += sp->dynamicBodies->num + sp->staticBodies->num + 1;
		ns += cpSpatialIndexCount(sp->dynamicShapes) + cpSpatialIndexCount(sp->staticShapes);
		nc += sp->constraints->num;
	}
	map_reset(bm, nb);
	map_reset(sm, ns);
	map_reset(cm, nc);
	for (int w = 0; w < pk->n_spaces; w++) {
		cpSpace *sp = pk->spaces[w];
		if (sp->sleepTimeThreshold != INFINITY) return report("cpGPUSpace: sleeping is not supported on the device", NULL);
		/* the collision handlers (cpSpace.c:382-413): pair and wildcard ones, then the default */
		struct HandlerCtx hc = {D, w};
		cpHashSetEach(sp->collisionHandlers, collect_handler, &hc);
		add_handler(D, w, CPGPU_WILDCARD, CPGPU_WILDCARD, handler_events(&sp->defaultHandler));
		if (sp->locked) return report("cpGPUSpace: the space is locked", NULL);
		cpgpuSpaceParams *P = &D->sp[w];
		memset(P, 0, sizeof(*P));
		P->iterations = sp->iterations;
		P->collision_persistence = (uint32_t)sp->collisionPersistence;
		P->gravity_x = sp->gravity.x; P->gravity_y = sp->gravity.y;
		P->damping = sp->damping;
		P->collision_slop = sp->collisionSlop;
		P->collision_bias = sp->collisionBias;
		P->curr_dt = sp->curr_dt;
		P->stamp = (uint32_t)sp->stamp;
		/* bodies: dynamic (and kinematic) array, static array, the built-in static body on demand */
		D->wbs[w] = D->nb;
		cpArray *arrays[2] = {sp->dynamicBodies, sp->staticBodies};
		for (int q = 0; q < 2; q++) {
			desc_reserve(D, D->nw, D->nb + arrays[q]->num, D->ns, D->nc, D->npv);
			for (int i = 0; i < arrays[q]->num; i++) {
				cpBody *b = (cpBody *)arrays[q]->arr[i];
				if (b->velocity_func != cpBodyUpdateVelocity || b->position_func != cpBodyUpdatePosition)
					return report("cpGPUSpace: custom body integrators are not supported on the device", NULL);
				D->bodies[D->nb] = b;
				map_put(bm, b, D->nb);
				D->nb++;
			}
		}
		/* shapes in the order they were added */
		pk->scratch.n = 0;
		cpSpatialIndexEach(sp->dynamicShapes, collect_shape, &pk->scratch);
		cpSpatialIndexEach(sp->staticShapes, collect_shape, &pk->scratch);
		qsort(pk->scratch.list, (size_t)pk->scratch.n, sizeof(cpShape *), cmp_hashid);
		D->wss[w] = D->ns;
		for (int i = 0; i < pk->scratch.n; i++) {
			cpShape *sh = pk->scratch.list[i];
			const int b = body_index(D, bm, sp, sh->body);
			if (b < 0) return report("cpGPUSpace: a shape's body is not in its space", NULL);
			desc_reserve(D, D->nw, D->nb, D->ns + 1, D->nc, D->npv);
			D->shapes[D->ns] = sh;
			map_put(sm, sh, D->ns);
			if (read_shape(D, D->ns, sh, b)) return -1;
			D->ns++;
		}
		D->wcs[w] = D->nc;
		desc_reserve(D, D->nw, D->nb, D->ns, D->nc + sp->constraints->num, D->npv);
		for (int i = 0; i < sp->constraints->num; i++) {
			cpConstraint *c = (cpConstraint *)sp->constraints->arr[i];
			const int a = body_index(D, bm, sp, c->a), b = body_index(D, bm, sp, c->b);
			if (a < 0 || b < 0) return report("cpGPUSpace: a constraint's body is not in its space", NULL);
			desc_reserve(D, D->nw, D->nb, D->ns, D->nc + 1, D->npv);
			D->cons[D->nc] = c;
			map_put(cm, c, D->nc);
			if (read_constraint(D, D->nc, c, a, b)) return -1;
			D->nc++;
		}
		/* the built-in static body may have been appended after this space's shapes were read:
		 * bodies are read last */
		for (int k = D->wbs[w]; k < D->nb; k++) read_body(D, k, D->bodies[k]);
	}
	D->wbs[D->nw] = D->nb;
	D->wss[D->nw] = D->ns;
	D->wcs[D->nw] = D->nc;
	desc_link(D);
	return 0;
}

static int same_i32(const int32_t *a, const int32_t *b, int n) { return n == 0 || !memcmp(a, b, (size_t)n * sizeof(*a)); }
static int same_u64(const uint64_t *a, const uint64_t *b, int n) { return n == 0 || !memcmp(a, b, (size_t)n * sizeof(*a)); }
static int same_u32(const uint32_t *a, const uint32_t *b, int n) { return n == 0 || !memcmp(a, b, (size_t)n * sizeof(*a)); }
static int same_f64(const double *a, const double *b, int n)
{
	for (int i = 0; i < n; i++)
		if (!(a[i] == b[i]) && !(isnan(a[i]) && isnan(b[i]))) return 0;
	return 1;
}
static int same_ptr(void *const *a, void *const *b, int n) { return n == 0 || !memcmp(a, b, (size_t)n * sizeof(*a)); }

/* Same objects with the same constants (everything but body state, joint
 * impulses and the space parameters)? */
static int same_structure(const Desc *A, const Desc *B)
{
	if (A->nw != B->nw || A->nb != B->nb || A->ns != B->ns || A->nc != B->nc || A->npv != B->npv) return 0;
	return same_i32(A->wbs, B->wbs, A->nw + 1) && same_i32(A->wss, B->wss, A->nw + 1) &&
	       same_i32(A->wcs, B->wcs, A->nw + 1) && same_ptr((void *const *)A->bodies, (void *const *)B->bodies, A->nb) &&
	       same_ptr((void *const *)A->shapes, (void *const *)B->shapes, A->ns) &&
	       same_ptr((void *const *)A->cons, (void *const *)B->cons, A->nc) && same_i32(A->body_type, B->body_type, A->nb) &&
	       same_f64(A->body_m, B->body_m, A->nb) && same_f64(A->body_m_inv, B->body_m_inv, A->nb) &&
	       same_f64(A->body_i, B->body_i, A->nb) && same_f64(A->body_i_inv, B->body_i_inv, A->nb) &&
	       same_f64(A->body_cog, B->body_cog, 2 * A->nb) && same_i32(A->shape_type, B->shape_type, A->ns) &&
	       same_i32(A->shape_body, B->shape_body, A->ns) && same_u64(A->shape_hashid, B->shape_hashid, A->ns) &&
	       same_f64(A->shape_e, B->shape_e, A->ns) && same_f64(A->shape_u, B->shape_u, A->ns) &&
	       same_f64(A->shape_sv, B->shape_sv, 2 * A->ns) && same_u64(A->shape_group, B->shape_group, A->ns) &&
	       same_u32(A->shape_cats, B->shape_cats, A->ns) && same_u32(A->shape_mask, B->shape_mask, A->ns) &&
	       (A->ns == 0 || !memcmp(A->shape_sensor, B->shape_sensor, (size_t)A->ns)) &&
	       same_u32(A->shape_ctype, B->shape_ctype, A->ns) && A->nh == B->nh &&
	       (A->nh == 0 || !memcmp(A->handlers, B->handlers, (size_t)A->nh * sizeof(cpgpuHandlerDesc))) &&
	       same_f64(A->shape_r, B->shape_r, A->ns) && same_f64(A->shape_geom, B->shape_geom, 10 * A->ns) &&
	       same_i32(A->shape_pstart, B->shape_pstart, A->ns) && same_i32(A->shape_pcount, B->shape_pcount, A->ns) &&
	       same_f64(A->pverts, B->pverts, 2 * A->npv) && same_f64(A->pnorms, B->pnorms, 2 * A->npv) &&
	       same_i32(A->cons_type, B->cons_type, A->nc) && same_i32(A->cons_a, B->cons_a, A->nc) &&
	       same_i32(A->cons_b, B->cons_b, A->nc) && same_i32(A->cons_collide, B->cons_collide, A->nc) &&
	       same_f64(A->cons_max_force, B->cons_max_force, A->nc) &&
	       same_f64(A->cons_error_bias, B->cons_error_bias, A->nc) &&
	       same_f64(A->cons_max_bias, B->cons_max_bias, A->nc) && same_f64(A->cons_params, B->cons_params, 8 * A->nc);
}

static int same_params(const cpgpuSpaceParams *a, const cpgpuSpaceParams *b)
{
	return a->iterations == b->iterations && a->collision_persistence == b->collision_persistence &&
	       a->gravity_x == b->gravity_x && a->gravity_y == b->gravity_y && a->damping == b->damping &&
	       a->collision_slop == b->collision_slop && a->collision_bias == b->collision_bias;
}

/* body k's field f differs in the two descriptions */
static int field_changed(const Desc *A, int ka, const Desc *B, int kb, int f)
{
	const int n = body_width[f];
	return !same_f64(A->state[f] + (size_t)n * ka, B->state[f] + (size_t)n * kb, n);
}

/* copy the shadow's layout (every body and constraint of `from`) */
static void shadow_take(Pack *pk, const Desc *from)
{
	Desc *S = &pk->shadow;
	desc_reserve(S, from->nw, from->nb, 1, from->nc, 1);
	S->nw = from->nw; S->nb = from->nb; S->nc = from->nc;
	for (int f = 0; f < BODY_FIELDS; f++)
		memcpy(S->state[f], from->state[f], (size_t)from->nb * body_width[f] * sizeof(double));
	memcpy(S->cons_jacc, from->cons_jacc, (size_t)from->nc * 2 * sizeof(double));
	memcpy(S->sp, from->sp, (size_t)from->nw * sizeof(cpgpuSpaceParams));
}

static void rb_reserve(Pack *pk, int nb, int nc)
{
	if (nb > pk->cap_rb) {
		int c = pk->cap_rb;
		GROW(pk->idx, c, nb);
		for (int f = 0; f < BODY_FIELDS; f++) {
			pk->rb[f] = realloc(pk->rb[f], (size_t)c * body_width[f] * sizeof(double));
			pk->wbuf[f] = realloc(pk->wbuf[f], (size_t)c * body_width[f] * sizeof(double));
		}
		pk->cap_rb = c;
	}
	if (nc > pk->cap_rb_jacc) {
		int c = pk->cap_rb_jacc;
		GROW(pk->rb_jacc, c, nc);
		pk->rb_jacc = realloc(pk->rb_jacc, (size_t)c * 2 * sizeof(double));
		pk->cap_rb_jacc = c;
	}
}

/* the device's body state and joint impulses into pk->rb / rb_jacc */
static int read_device(Pack *pk)
{
	const Desc *C = &pk->cur;
	rb_reserve(pk, C->nb, C->nc);
	int rc = cpgpuWorldReadBodies(pk->world, pk->rb[BODY_P], pk->rb[BODY_V], pk->rb[BODY_A], pk->rb[BODY_W],
	                              pk->rb[BODY_VB], pk->rb[BODY_WB], pk->rb[BODY_XF]);
	if (rc) return lib_error("cpgpuWorldReadBodies", rc);
	/* forces and torques are zero on the device after a step (cpBodyUpdateVelocity,
	 * cpBody.c:504-506) except on kinematic bodies, which keep theirs; the binding
	 * keeps the host's, which equal the device's unless the caller changed them */
	memcpy(pk->rb[BODY_F], pk->shadow.state[BODY_F], (size_t)C->nb * 2 * sizeof(double));
	memcpy(pk->rb[BODY_T], pk->shadow.state[BODY_T], (size_t)C->nb * sizeof(double));
	if (C->nc) {
		int32_t *colour = (int32_t *)malloc((size_t)C->nc * sizeof(int32_t));
		rc = cpgpuWorldReadConstraints(pk->world, pk->rb_jacc, colour);
		free(colour);
		if (rc) return lib_error("cpgpuWorldReadConstraints", rc);
	}
	return 0;
}

static void swap_desc(Desc *a, Desc *b)
{
	Desc t = *a;
	*a = *b;
	*b = t;
	desc_link(a);
	desc_link(b);
}

/* A new device world from the spaces as collected in pk->fresh. */
static int repack(Pack *pk)
{
	Desc *F = &pk->fresh, *C = &pk->cur, *S = &pk->shadow;
	if (pk->world) {
		if (read_device(pk)) return -1;
		/* each field of a body packed before: the caller's value where the host struct
		 * no longer holds the value last given to it, else the device's */
		for (int k = 0; k < F->nb; k++) {
			const int o = map_get(&pk->cur_at.body, F->bodies[k]);
			if (o < 0) continue;
			for (int f = 0; f < BODY_FIELDS; f++) {
				const int n = body_width[f];
				if (!field_changed(F, k, S, o, f)) memcpy(F->state[f] + n * k, pk->rb[f] + n * o, n * sizeof(double));
			}
		}
		for (int k = 0; k < F->nc; k++) {
			const int o = map_get(&pk->cur_at.cons, F->cons[k]);
			if (o < 0 || F->cons_type[k] != C->cons_type[o]) continue;
			if (same_f64(F->cons_jacc + 2 * k, S->cons_jacc + 2 * o, 2))
				memcpy(F->cons_jacc + 2 * k, pk->rb_jacc + 2 * o, 2 * sizeof(double));
		}
	}
	GROW(pk->shape_from, pk->cap_from, F->ns + 1);
	for (int k = 0; k < F->ns; k++) pk->shape_from[k] = pk->world ? map_get(&pk->cur_at.shape, F->shapes[k]) : -1;
	cpgpuWorld *nw = NULL;
	int rc = pk->world ? cpgpuWorldCreateFrom(&F->d, pk->precision, pk->world, pk->shape_from, &nw)
	                   : cpgpuWorldCreate(&F->d, pk->precision, &nw);
	if (rc) return lib_error(pk->world ? "cpgpuWorldCreateFrom" : "cpgpuWorldCreate", rc);
	if (pk->world) cpgpuWorldDestroy(pk->world);
	pk->world = nw;
	pk->packs++;
	/* the host structs take the state the device now holds */
	for (int k = 0; k < F->nb; k++) write_body(F, k, F->bodies[k]);
	for (int k = 0; k < F->nc; k++) write_jacc(F->cons_type[k], F->cons_jacc + 2 * k, F->cons[k]);
	swap_desc(C, F);
	Maps t = pk->cur_at;
	pk->cur_at = pk->fresh_at;
	pk->fresh_at = t;
	shadow_take(pk, C);
	return 0;
}

/* ---- before a step: what the caller changed ---- */

static void body_field(const cpBody *b, int f, double *out)
{
	switch (f) {
	case BODY_P: out[0] = b->p.x; out[1] = b->p.y; break;
	case BODY_V: out[0] = b->v.x; out[1] = b->v.y; break;
	case BODY_F: out[0] = b->f.x; out[1] = b->f.y; break;
	case BODY_A: out[0] = b->a; break;
	case BODY_W: out[0] = b->w; break;
	case BODY_T: out[0] = b->t; break;
	case BODY_VB: out[0] = b->v_bias.x; out[1] = b->v_bias.y; break;
	case BODY_WB: out[0] = b->w_bias; break;
	case BODY_XF:
		out[0] = b->transform.a; out[1] = b->transform.b; out[2] = b->transform.c;
		out[3] = b->transform.d; out[4] = b->transform.tx; out[5] = b->transform.ty;
		break;
	}
}

static void set_body_field(cpBody *b, int f, const double *v)
{
	switch (f) {
	case BODY_P: b->p = cpv(v[0], v[1]); break;
	case BODY_V: b->v = cpv(v[0], v[1]); break;
	case BODY_F: b->f = cpv(v[0], v[1]); break;
	case BODY_A: b->a = v[0]; break;
	case BODY_W: b->w = v[0]; break;
	case BODY_T: b->t = v[0]; break;
	case BODY_VB: b->v_bias = cpv(v[0], v[1]); break;
	case BODY_WB: b->w_bias = v[0]; break;
	case BODY_XF:
		b->transform.a = v[0]; b->transform.b = v[1]; b->transform.c = v[2];
		b->transform.d = v[3]; b->transform.tx = v[4]; b->transform.ty = v[5];
		break;
	}
}

/* Same objects and constants as the device world: does any change still need a
 * new pack (a static body moved, bias velocities or joint impulses set)? */
static int needs_repack(Pack *pk)
{
	const Desc *F = &pk->fresh, *S = &pk->shadow;
	for (int k = 0; k < F->nb; k++) {
		if (field_changed(F, k, S, k, BODY_VB) || field_changed(F, k, S, k, BODY_WB)) return 1;
		if (F->body_type[k] == CP_BODY_TYPE_STATIC &&
		    (field_changed(F, k, S, k, BODY_P) || field_changed(F, k, S, k, BODY_A) || field_changed(F, k, S, k, BODY_XF)))
			return 1;
	}
	return !same_f64(F->cons_jacc, S->cons_jacc, 2 * F->nc);
}

/* The fields the caller changed since the binding last wrote them, one
 * cpgpuWorldWriteBodiesAt per field; the space parameters. */
static int upload_writes(Pack *pk)
{
	Desc *F = &pk->fresh, *S = &pk->shadow;
	static const int writable[] = {BODY_P, BODY_V, BODY_F, BODY_A, BODY_W, BODY_T, BODY_XF};
	rb_reserve(pk, F->nb, F->nc);
	for (size_t q = 0; q < sizeof(writable) / sizeof(writable[0]); q++) {
		const int f = writable[q], n = body_width[f];
		int m = 0;
		for (int k = 0; k < F->nb; k++) {
			if (!field_changed(F, k, S, k, f)) continue;
			pk->idx[m] = k;
			memcpy(pk->wbuf[f] + (size_t)n * m, F->state[f] + (size_t)n * k, n * sizeof(double));
			memcpy(S->state[f] + (size_t)n * k, F->state[f] + (size_t)n * k, n * sizeof(double));
			m++;
		}
		if (!m) continue;
		const double *in[BODY_FIELDS] = {0};
		in[f] = pk->wbuf[f];
		const int rc = cpgpuWorldWriteBodiesAt(pk->world, m, pk->idx, in[BODY_P], in[BODY_V], in[BODY_F], in[BODY_A],
		                                       in[BODY_W], in[BODY_T], in[BODY_XF]);
		if (rc) return lib_error("cpgpuWorldWriteBodiesAt", rc);
	}
	for (int w = 0; w < F->nw; w++) {
		if (same_params(&F->sp[w], &S->sp[w])) continue;
		const int rc = cpgpuWorldSetSpaceParams(pk->world, w, &F->sp[w]);
		if (rc) return lib_error("cpgpuWorldSetSpaceParams", rc);
		S->sp[w] = F->sp[w];
	}
	return 0;
}

/* Bring the device world up to date with the spaces. */
static int prepare(Pack *pk)
{
	if (collect(pk, &pk->fresh, &pk->fresh_at)) return -1;
	if (!pk->world || !same_structure(&pk->fresh, &pk->cur) || needs_repack(pk)) return repack(pk);
	return upload_writes(pk);
}

/* What a step changes on the host without device data: the spaces' stamp and
 * dt (cpSpaceStep.c:341-344) and the dynamic bodies' forces and torques
 * (cpBodyUpdateVelocity, cpBody.c:504-506). */
static void host_after_steps(Pack *pk, double dt, int steps)
{
	for (int w = 0; w < pk->n_spaces; w++) {
		pk->spaces[w]->stamp += (cpTimestamp)steps;
		pk->spaces[w]->curr_dt = dt;
		pk->shadow.sp[w].stamp = (uint32_t)pk->spaces[w]->stamp;
		pk->shadow.sp[w].curr_dt = dt;
	}
	const Desc *C = &pk->cur;
	for (int k = 0; k < C->nb; k++) {
		if (C->body_type[k] != CP_BODY_TYPE_DYNAMIC) continue;
		cpBody *b = C->bodies[k];
		b->f = cpvzero;
		b->t = 0.0;
		pk->shadow.state[BODY_F][2 * k] = pk->shadow.state[BODY_F][2 * k + 1] = 0.0;
		pk->shadow.state[BODY_T][k] = 0.0;
	}
}

/* ---- collision handler callbacks around a device step ---- */

/* cpSpaceLookupHandler (cpArbiter.c:348-354) */
static cpCollisionHandler *lookup_handler(cpSpace *space, cpCollisionType a, cpCollisionType b, cpCollisionHandler *dflt)
{
	cpCollisionType types[] = {a, b};
	cpCollisionHandler *h = (cpCollisionHandler *)cpHashSetFind(space->collisionHandlers, CP_HASH_PAIR(a, b), types);
	return h ? h : dflt;
}

/* An arbiter's handlers as cpArbiterUpdate chooses them (cpArbiter.c:399-409). */
static void set_handlers(cpArbiter *arb)
{
	const cpShape *a = arb->a, *b = arb->b;
	cpSpace *space = a->space;
	cpCollisionHandler *dflt = &space->defaultHandler;
	cpCollisionHandler *h = arb->handler = lookup_handler(space, a->type, b->type, dflt);
	const cpBool swapped = arb->swapped = (a->type != h->typeA && h->typeA != CP_WILDCARD_COLLISION_TYPE);
	arb->handlerA = arb->handlerB = &cpCollisionHandlerDoNothing;
	if (h != dflt || space->usesWildcards) {
		arb->handlerA = lookup_handler(space, swapped ? b->type : a->type, CP_WILDCARD_COLLISION_TYPE,
		                               &cpCollisionHandlerDoNothing);
		arb->handlerB = lookup_handler(space, swapped ? a->type : b->type, CP_WILDCARD_COLLISION_TYPE,
		                               &cpCollisionHandlerDoNothing);
	}
}

/* A cpArbiter for a callback on shapes a and b. */
static void callback_arbiter(cpArbiter *arb, struct cpContact *con, cpShape *a, cpShape *b, int count)
{
	memset(arb, 0, sizeof(*arb));
	memset(con, 0, 2 * sizeof(*con));
	arb->a = a; arb->b = b;
	arb->body_a = a->body; arb->body_b = b->body;
	arb->contacts = con;
	arb->count = count;
	arb->stamp = a->space->stamp;
	set_handlers(arb);
}

/* begin and preSolve of the step's due arbiters, as cpSpaceCollideShapes calls
 * them (cpSpaceStep.c:255-266); their answers go back to the device */
static int run_callbacks(Pack *pk, int64_t n)
{
	const Desc *C = &pk->cur;
	cpgpuCallbackArbiter *due = (cpgpuCallbackArbiter *)malloc((size_t)n * sizeof(*due));
	cpgpuCallbackDecision *dec = (cpgpuCallbackDecision *)malloc((size_t)n * sizeof(*dec));
	int64_t got = 0;
	int rc = cpgpuWorldReadCallbacks(pk->world, n, &got, due);
	if (rc == 0) {
		for (int64_t k = 0; k < n; k++) {
			const cpgpuCallbackArbiter *v = &due[k];
			cpShape *a = C->shapes[v->shape_a], *b = C->shapes[v->shape_b];
			cpSpace *space = a->space;
			cpArbiter arb;
			struct cpContact con[2];
			callback_arbiter(&arb, con, a, b, v->count);
			arb.n = cpv(v->normal[0], v->normal[1]);
			arb.e = v->e; arb.u = v->u;
			arb.surface_vr = cpv(v->surface_vr[0], v->surface_vr[1]);
			arb.state = (enum cpArbiterState)v->state;
			for (int c = 0; c < v->count && c < 2; c++) {
				con[c].r1 = cpv(v->contact[c][0], v->contact[c][1]);
				con[c].r2 = cpv(v->contact[c][2], v->contact[c][3]);
				con[c].jnAcc = v->contact[c][4];
				con[c].jtAcc = v->contact[c][5];
				con[c].hash = (cpHashValue)v->hash[c];
			}
			cpCollisionHandler *h = arb.handler;
			if (arb.state == CP_ARBITER_STATE_FIRST_COLLISION && !h->beginFunc(&arb, space, h->userData))
				cpArbiterIgnore(&arb);
			cpBool solve = cpFalse;
			if (arb.state != CP_ARBITER_STATE_IGNORE)
				solve = h->preSolveFunc(&arb, space, h->userData) && arb.state != CP_ARBITER_STATE_IGNORE;
			cpgpuCallbackDecision *d = &dec[k];
			d->slot = v->slot;
			d->solve = solve ? 1 : 0;
			d->ignore = arb.state == CP_ARBITER_STATE_IGNORE;
			d->_pad = 0;
			d->e = arb.e; d->u = arb.u;
			d->surface_vr[0] = arb.surface_vr.x; d->surface_vr[1] = arb.surface_vr.y;
		}
		rc = cpgpuWorldWriteDecisions(pk->world, n, dec);
		if (rc) rc = lib_error("cpgpuWorldWriteDecisions", rc);
	} else {
		rc = lib_error("cpgpuWorldReadCallbacks", rc);
	}
	free(due);
	free(dec);
	return rc ? -1 : 0;
}

/* separate for the pairs that separated (cpSpaceStep.c:306-310), then postSolve
 * for the solved arbiters (:437-442) */
static int after_callbacks(Pack *pk)
{
	const Desc *C = &pk->cur;
	uint32_t events = 0;
	for (int k = 0; k < C->nh; k++) events |= C->handlers[k].events;
	int64_t n = 0;
	int rc;
	if (events & CPGPU_EVENT_SEPARATE) {
		rc = cpgpuWorldReadSeparations(pk->world, 0, &n, NULL);
		if (rc) return lib_error("cpgpuWorldReadSeparations", rc);
		if (n > 0) {
			int32_t *pairs = (int32_t *)malloc((size_t)n * 3 * sizeof(int32_t));
			rc = cpgpuWorldReadSeparations(pk->world, n, &n, pairs);
			for (int64_t k = 0; rc == 0 && k < n; k++) {
				cpShape *a = C->shapes[pairs[3 * k + 1]], *b = C->shapes[pairs[3 * k + 2]];
				cpArbiter arb;
				struct cpContact con[2];
				callback_arbiter(&arb, con, a, b, 0);
				arb.state = CP_ARBITER_STATE_CACHED;
				arb.handler->separateFunc(&arb, a->space, arb.handler->userData);
			}
			free(pairs);
			if (rc) return lib_error("cpgpuWorldReadSeparations", rc);
		}
	}
	if (!(events & CPGPU_EVENT_POST_SOLVE)) return 0;
	rc = cpgpuWorldReadArbiters(pk->world, 0, &n, NULL, NULL, NULL, NULL, NULL, NULL, NULL);
	if (rc) return lib_error("cpgpuWorldReadArbiters", rc);
	if (n == 0) return 0;
	int32_t *shapes = (int32_t *)malloc((size_t)n * 2 * sizeof(int32_t));
	int32_t *state = (int32_t *)malloc((size_t)n * sizeof(int32_t));
	int32_t *count = (int32_t *)malloc((size_t)n * sizeof(int32_t));
	double *normal = (double *)malloc((size_t)n * 2 * sizeof(double));
	double *contact = (double *)malloc((size_t)n * 24 * sizeof(double));
	uint64_t *hash = (uint64_t *)malloc((size_t)n * 2 * sizeof(uint64_t));
	rc = cpgpuWorldReadArbiters(pk->world, n, &n, shapes, NULL, state, count, normal, contact, hash);
	for (int64_t i = 0; rc == 0 && i < n; i++) {
		cpShape *a = C->shapes[shapes[2 * i]], *b = C->shapes[shapes[2 * i + 1]];
		cpArbiter arb;
		struct cpContact con[2];
		callback_arbiter(&arb, con, a, b, count[i]);
		arb.n = cpv(normal[2 * i], normal[2 * i + 1]);
		arb.e = a->e * b->e;
		arb.u = a->u * b->u;
		arb.state = (enum cpArbiterState)state[i];
		for (int c = 0; c < count[i] && c < 2; c++) {
			const double *q = contact + 24 * i + 12 * c;
			con[c].r1 = cpv(q[0], q[1]); con[c].r2 = cpv(q[2], q[3]);
			con[c].nMass = q[4]; con[c].tMass = q[5]; con[c].bounce = q[6];
			con[c].jnAcc = q[7]; con[c].jtAcc = q[8]; con[c].jBias = q[9]; con[c].bias = q[10];
			con[c].hash = (cpHashValue)hash[2 * i + c];
		}
		arb.handler->postSolveFunc(&arb, a->space, arb.handler->userData);
	}
	free(shapes); free(state); free(count); free(normal); free(contact); free(hash);
	if (rc) return lib_error("cpgpuWorldReadArbiters", rc);
	return 0;
}

static int pack_sync_state(Pack *pk);

static int pack_step(Pack *pk, double dt)
{
	if (dt == 0.0) return 0;   /* cpSpaceStep.c:339 */
	if (prepare(pk)) return -1;
	if (pk->cur.nh == 0) {
		const int rc = cpgpuWorldStep(pk->world, dt);
		if (rc && rc != CPGPU_ERR_FAULT) return lib_error("cpgpuWorldStep", rc);
		host_after_steps(pk, dt, 1);
		/* a fault: the step ran in full and failed a check of the reference's
		 * debug build, reported (by default: aborting) as that build does */
		return rc ? lib_error("cpgpuWorldStep", rc) : 0;
	}
	/* with collision handlers: the step in two halves around the callbacks, which see
	 * the bodies at this step's positions and the spaces' new stamps, with the spaces
	 * locked as cpSpaceStep locks them; unlocking runs the post-step callbacks the
	 * handlers queued (cpSpaceStep.c:368-445, cpSpaceUnlock :69-104) */
	int64_t n = 0;
	int rc = cpgpuWorldStepBegin(pk->world, dt, &n);
	if (rc) return lib_error("cpgpuWorldStepBegin", rc);
	host_after_steps(pk, dt, 1);
	int failed = n > 0 && pack_sync_state(pk);
	for (int w = 0; w < pk->n_spaces; w++) cpSpaceLock(pk->spaces[w]);
	if (!failed && n > 0) failed = run_callbacks(pk, n);
	rc = cpgpuWorldStepEnd(pk->world);
	if (!failed && rc && rc != CPGPU_ERR_FAULT) failed = lib_error("cpgpuWorldStepEnd", rc);
	if (!failed) failed = after_callbacks(pk);
	for (int w = 0; w < pk->n_spaces; w++) cpSpaceUnlock(pk->spaces[w], failed ? cpFalse : cpTrue);
	if (!failed && rc == CPGPU_ERR_FAULT) failed = lib_error("cpgpuWorldStepEnd", rc);   /* as above */
	return failed ? -1 : 0;
}

/* ---- after a step: device state back into the structs ---- */

/* cpBodyPushArbiter (cpSpaceComponent.c:169-181) */
static void push_arbiter(cpBody *body, cpArbiter *arb)
{
	cpArbiter *next = body->arbiterList;
	cpArbiterThreadForBody(arb, body)->next = next;
	if (next) cpArbiterThreadForBody(next, body)->prev = arb;
	body->arbiterList = arb;
}

/* ---- the binding's arbiters in the spaces' arbiter containers ----
 * A step of the reference leaves its arbiters in space->cachedArbiters and
 * space->arbiters (cpSpaceStep.c:250-260, 420-436) as well as threaded onto the
 * bodies.  cpSpaceRemoveShape and cpSpaceRemoveBody unthread an arbiter only
 * through space->cachedArbiters (cpSpaceFilterArbiters, cpSpace.c:483-520), so
 * the arbiters the binding threads are registered there too: otherwise a body
 * removed and freed would stay on its partners' arbiter lists. */

static int is_ours(const Pack *pk, const void *p)
{
	const cpArbiter *a = (const cpArbiter *)p;
	return pk->arbs && a >= pk->arbs && a < pk->arbs + pk->cap_arb;
}

static void drop_ours(cpArray *arr, const Pack *pk)
{
	int j = 0;
	for (int i = 0; i < arr->num; i++)
		if (!is_ours(pk, arr->arr[i])) arr->arr[j++] = arr->arr[i];
	arr->num = j;
}

struct ForgetCtx {
	const Pack *pk;
	cpSpace *space;
};

/* every arbiter leaves the set: the binding's own are dropped; one a step of the
 * reference left behind (a space stepped by cpSpaceStep before it was batched) is
 * retired as cachedArbitersFilter retires it (cpSpace.c:503-505) */
static cpBool forget_arbiter(void *elt, void *data)
{
	const struct ForgetCtx *c = (const struct ForgetCtx *)data;
	cpArbiter *arb = (cpArbiter *)elt;
	if (!is_ours(c->pk, arb)) {
		cpArbiterUnthread(arb);
		cpArrayDeleteObj(c->space->arbiters, arb);
		cpArrayPush(c->space->pooledArbiters, arb);
	}
	return cpFalse;
}

/* The spaces' arbiter containers and the live bodies' arbiter lists emptied of
 * every arbiter (the binding's are about to be replaced or freed). */
static void forget_arbiters(Pack *pk)
{
	for (int w = 0; w < pk->n_spaces; w++) {
		cpSpace *sp = pk->spaces[w];
		struct ForgetCtx c = {pk, sp};
		cpHashSetFilter(sp->cachedArbiters, forget_arbiter, &c);
		drop_ours(sp->arbiters, pk);
		drop_ours(sp->pooledArbiters, pk);
		cpArray *arrays[2] = {sp->dynamicBodies, sp->staticBodies};
		for (int q = 0; q < 2; q++)
			for (int i = 0; i < arrays[q]->num; i++) ((cpBody *)arrays[q]->arr[i])->arbiterList = NULL;
		sp->staticBody->arbiterList = NULL;
	}
	pk->narb = 0;
}

static void *same_arbiter(const void *shapes, void *arb)
{
	(void)shapes;
	return arb;
}

/* The last step's arbiters as cpArbiters threaded onto their bodies, in the
 * device's solve order, as cpSpaceProcessComponents threads space->arbiters
 * (cpSpaceComponent.c:253-267), and entered into their space's cachedArbiters
 * and arbiters. */
static int sync_arbiters(Pack *pk)
{
	const Desc *C = &pk->cur;
	const Maps *live = &pk->fresh_at;
	forget_arbiters(pk);
	int64_t n = 0;
	int rc = cpgpuWorldReadArbiters(pk->world, 0, &n, NULL, NULL, NULL, NULL, NULL, NULL, NULL);
	if (rc) return lib_error("cpgpuWorldReadArbiters", rc);
	if (n == 0) return 0;
	int32_t *shapes = (int32_t *)malloc((size_t)n * 2 * sizeof(int32_t));
	int32_t *state = (int32_t *)malloc((size_t)n * sizeof(int32_t));
	int32_t *count = (int32_t *)malloc((size_t)n * sizeof(int32_t));
	double *normal = (double *)malloc((size_t)n * 2 * sizeof(double));
	double *contact = (double *)malloc((size_t)n * 24 * sizeof(double));
	uint64_t *hash = (uint64_t *)malloc((size_t)n * 2 * sizeof(uint64_t));
	int64_t got = 0;
	rc = cpgpuWorldReadArbiters(pk->world, n, &got, shapes, NULL, state, count, normal, contact, hash);
	if (rc == 0) {
		if (n > pk->cap_arb) {
			int c = pk->cap_arb;
			GROW(pk->arbs, c, (int)n);
			pk->contacts = realloc(pk->contacts, (size_t)c * 2 * sizeof(struct cpContact));
			pk->cap_arb = c;
		}
		for (int64_t i = 0; i < n; i++) {
			cpShape *a = C->shapes[shapes[2 * i]], *b = C->shapes[shapes[2 * i + 1]];
			if (map_get(&live->shape, a) < 0 || map_get(&live->shape, b) < 0) continue;
			cpArbiter *arb = &pk->arbs[pk->narb];
			struct cpContact *con = &pk->contacts[2 * pk->narb];
			pk->narb++;
			memset(arb, 0, sizeof(*arb));
			arb->a = a; arb->b = b;
			arb->body_a = a->body; arb->body_b = b->body;
			arb->count = count[i];
			arb->contacts = con;
			arb->n = cpv(normal[2 * i], normal[2 * i + 1]);
			/* cpArbiterUpdate (cpArbiter.c:392-398) */
			arb->e = a->e * b->e;
			arb->u = a->u * b->u;
			const cpVect surface_vr = cpvsub(b->surfaceV, a->surfaceV);
			arb->surface_vr = cpvsub(surface_vr, cpvmult(arb->n, cpvdot(surface_vr, arb->n)));
			set_handlers(arb);
			arb->stamp = a->space->stamp;
			arb->state = (enum cpArbiterState)state[i];
			for (int c = 0; c < 2; c++) {
				const double *q = contact + 24 * i + 12 * c;
				memset(&con[c], 0, sizeof(con[c]));
				if (c >= count[i]) continue;
				con[c].r1 = cpv(q[0], q[1]);
				con[c].r2 = cpv(q[2], q[3]);
				con[c].nMass = q[4]; con[c].tMass = q[5]; con[c].bounce = q[6];
				con[c].jnAcc = q[7]; con[c].jtAcc = q[8]; con[c].jBias = q[9]; con[c].bias = q[10];
				con[c].hash = (cpHashValue)hash[2 * i + c];
			}
			push_arbiter(arb->body_a, arb);
			push_arbiter(arb->body_b, arb);
			const cpShape *pair[2] = {a, b};
			cpHashSetInsert(a->space->cachedArbiters, CP_HASH_PAIR((cpHashValue)a, (cpHashValue)b), pair, same_arbiter,
			                arb);
			cpArrayPush(a->space->arbiters, arb);
		}
	}
	free(shapes); free(state); free(count); free(normal); free(contact); free(hash);
	if (rc) return lib_error("cpgpuWorldReadArbiters", rc);
	return 0;
}

/* The device's body state and joint impulses into the structs (the fields the
 * caller has not changed), and the shapes' cached data. */
static int pack_sync_state(Pack *pk)
{
	if (!pk->world) return 0;
	/* the objects now in the spaces: only those are written (a removed body may be freed) */
	if (collect(pk, &pk->fresh, &pk->fresh_at)) return -1;
	if (read_device(pk)) return -1;
	const Desc *C = &pk->cur;
	Desc *S = &pk->shadow;
	static const int synced[] = {BODY_P, BODY_V, BODY_A, BODY_W, BODY_VB, BODY_WB, BODY_XF};
	for (int k = 0; k < C->nb; k++) {
		cpBody *b = C->bodies[k];
		if (map_get(&pk->fresh_at.body, b) < 0) continue;
		double host[6];
		/* the device's value for every field the caller has not changed since the binding
		 * last wrote it; a changed field is kept and goes to the device with the next step */
		for (size_t q = 0; q < sizeof(synced) / sizeof(synced[0]); q++) {
			const int f = synced[q], n = body_width[f];
			body_field(b, f, host);
			if (!same_f64(host, S->state[f] + (size_t)n * k, n)) continue;
			memcpy(S->state[f] + (size_t)n * k, pk->rb[f] + (size_t)n * k, n * sizeof(double));
			set_body_field(b, f, pk->rb[f] + (size_t)n * k);
		}
	}
	for (int k = 0; k < C->nc; k++) {
		cpConstraint *c = C->cons[k];
		if (map_get(&pk->fresh_at.cons, c) < 0) continue;
		const int fk = map_get(&pk->fresh_at.cons, c);
		if (!same_f64(pk->fresh.cons_jacc + 2 * fk, S->cons_jacc + 2 * k, 2)) continue;
		memcpy(S->cons_jacc + 2 * k, pk->rb_jacc + 2 * k, 2 * sizeof(double));
		/* the type is the current pack's (the shadow keeps only state) */
		write_jacc(C->cons_type[k], S->cons_jacc + 2 * k, c);
	}
	/* cached shape data and boxes from the new transforms (cpShapeUpdate, cpShape.c:216-220) */
	for (int k = 0; k < C->ns; k++) {
		cpShape *sh = C->shapes[k];
		if (map_get(&pk->fresh_at.shape, sh) < 0) continue;
		if (cpBodyGetType(sh->body) == CP_BODY_TYPE_STATIC) continue;
		cpShapeUpdate(sh, sh->body->transform);
	}
	return 0;
}

static int pack_sync(Pack *pk)
{
	if (pack_sync_state(pk)) return -1;
	return pk->world ? sync_arbiters(pk) : 0;
}

static void pack_release(Pack *pk)
{
	if (pk->world) {
		(void)cpgpuWorldSync(pk->world);
		/* the host arbiters out of the spaces and off the bodies still in them */
		forget_arbiters(pk);
		cpgpuWorldDestroy(pk->world);
		pk->world = NULL;
	}
	desc_free(&pk->cur);
	desc_free(&pk->fresh);
	desc_free(&pk->shadow);
	Maps *m[2] = {&pk->cur_at, &pk->fresh_at};
	for (int q = 0; q < 2; q++) {
		map_free(&m[q]->body);
		map_free(&m[q]->shape);
		map_free(&m[q]->cons);
	}
	free(pk->scratch.list);
	free(pk->shape_from);
	free(pk->arbs);
	free(pk->contacts);
	for (int f = 0; f < BODY_FIELDS; f++) { free(pk->rb[f]); free(pk->wbuf[f]); }
	free(pk->rb_jacc);
	free(pk->idx);
	free(pk->spaces);
	memset(pk, 0, sizeof(*pk));
}

static int valid_precision(int precision)
{
	if (precision == CPGPU_FP64 || precision == CPGPU_FP32) return 1;
	report("cpGPUSpace: precision must be CPGPU_FP64 or CPGPU_FP32", NULL);
	return 0;
}

/* ---- cpGPUSpace: a cpSpace as its first member, like cpHastySpace (cpHastySpace.c:397-416) ---- */

typedef struct cpGPUSpace {
	cpSpace space;
	Pack pack;
} cpGPUSpace;

cpSpace *cpGPUSpaceNewWithPrecision(int precision)
{
	if (!valid_precision(precision)) return NULL;
	cpGPUSpace *g = (cpGPUSpace *)cpcalloc(1, sizeof(cpGPUSpace));
	cpSpaceInit(&g->space);
	g->pack.n_spaces = 1;
	g->pack.spaces = (cpSpace **)malloc(sizeof(cpSpace *));
	g->pack.spaces[0] = &g->space;
	g->pack.precision = precision;
	return &g->space;
}

cpSpace *cpGPUSpaceNew(void) { return cpGPUSpaceNewWithPrecision(CPGPU_FP64); }

void cpGPUSpaceFree(cpSpace *space)
{
	if (!space) return;
	pack_release(&((cpGPUSpace *)space)->pack);
	cpSpaceFree(space);
}

void cpGPUSpaceStep(cpSpace *space, cpFloat dt) { (void)pack_step(&((cpGPUSpace *)space)->pack, dt); }
void cpGPUSpaceSyncHost(cpSpace *space) { (void)pack_sync(&((cpGPUSpace *)space)->pack); }
cpgpuWorld *cpGPUSpaceGetWorld(cpSpace *space) { return ((cpGPUSpace *)space)->pack.world; }
const cpgpuWorldDesc *cpGPUSpaceGetDesc(cpSpace *space)
{
	const Pack *pk = &((cpGPUSpace *)space)->pack;
	return pk->world ? &pk->cur.d : NULL;
}
int cpGPUSpaceBodyIndex(cpSpace *space, const cpBody *body) { return map_get(&((cpGPUSpace *)space)->pack.cur_at.body, body); }
int cpGPUSpaceShapeIndex(cpSpace *space, const cpShape *shape) { return map_get(&((cpGPUSpace *)space)->pack.cur_at.shape, shape); }
int cpGPUSpaceConstraintIndex(cpSpace *space, const cpConstraint *constraint)
{
	return map_get(&((cpGPUSpace *)space)->pack.cur_at.cons, constraint);
}
int cpGPUSpaceGetPackCount(cpSpace *space) { return ((cpGPUSpace *)space)->pack.packs; }

const cpgpuWorldDesc *cpGPUSpaceDescribe(cpSpace *space)
{
	Pack *pk = &((cpGPUSpace *)space)->pack;
	return collect(pk, &pk->fresh, &pk->fresh_at) ? NULL : &pk->fresh.d;
}
const int32_t *cpGPUSpaceGetShapeMap(cpSpace *space) { return ((cpGPUSpace *)space)->pack.shape_from; }

/* ---- cpGPUBatch ---- */

struct cpGPUBatch {
	Pack pack;
};

cpGPUBatch *cpGPUBatchNewWithPrecision(cpSpace **spaces, int n, int precision)
{
	if (!valid_precision(precision)) return NULL;
	if (!spaces || n < 1) {
		report("cpGPUBatch: at least one space", NULL);
		return NULL;
	}
	cpGPUBatch *b = (cpGPUBatch *)calloc(1, sizeof(cpGPUBatch));
	b->pack.n_spaces = n;
	b->pack.spaces = (cpSpace **)malloc((size_t)n * sizeof(cpSpace *));
	memcpy(b->pack.spaces, spaces, (size_t)n * sizeof(cpSpace *));
	b->pack.precision = precision;
	return b;
}

cpGPUBatch *cpGPUBatchNew(cpSpace **spaces, int n) { return cpGPUBatchNewWithPrecision(spaces, n, CPGPU_FP64); }

void cpGPUBatchFree(cpGPUBatch *batch)
{
	if (!batch) return;
	pack_release(&batch->pack);
	free(batch);
}

void cpGPUBatchStep(cpGPUBatch *batch, cpFloat dt) { (void)pack_step(&batch->pack, dt); }
void cpGPUBatchSyncHost(cpGPUBatch *batch) { (void)pack_sync(&batch->pack); }

void cpGPUBatchAutotune(cpGPUBatch *batch, cpFloat dt, int trial_steps)
{
	Pack *pk = &batch->pack;
	if (dt == 0.0 || trial_steps < 1) return;
	if (prepare(pk)) return;
	int32_t threads[CPGPU_MAX_BUCKETS];
	const int rc = cpgpuWorldAutotune(pk->world, dt, trial_steps, threads);
	if (rc) {
		(void)lib_error("cpgpuWorldAutotune", rc);
		return;
	}
	host_after_steps(pk, dt, 5 * trial_steps);
}

cpgpuWorld *cpGPUBatchGetWorld(cpGPUBatch *batch) { return batch->pack.world; }
const cpgpuWorldDesc *cpGPUBatchGetDesc(cpGPUBatch *batch) { return batch->pack.world ? &batch->pack.cur.d : NULL; }
const int32_t *cpGPUBatchGetShapeMap(cpGPUBatch *batch) { return batch->pack.shape_from; }
int cpGPUBatchGetPackCount(cpGPUBatch *batch) { return batch->pack.packs; }

const cpgpuWorldDesc *cpGPUBatchDescribe(cpGPUBatch *batch)
{
	Pack *pk = &batch->pack;
	return collect(pk, &pk->fresh, &pk->fresh_at) ? NULL : &pk->fresh.d;
}
