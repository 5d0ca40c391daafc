/*
 * cpgpu.h -- C ABI of the MI355X (gfx950) step library for Chipmunk2D spaces.
 *
 * The library replaces the body of Chipmunk2D 7.0.3's cpSpaceStep
 * (src/cpSpaceStep.c:335-445) with hand-written HIP kernels.  It does not
 * know Chipmunk's structs: a caller packs its space into the plain
 * structure-of-arrays description below (doubles, integers, plain pointers
 * and sizes) and gets body state back the same way.  The packing code a
 * Chipmunk maintainer adds on the reference side (cpGPUSpaceNew / Step /
 * SyncHost over cpSpace, modelled on cpHastySpace) is described in
 * INTEGRATION.md; the Python host mirror in chipmunk2d_amd/ packs its own
 * Space objects into the same description.
 *
 * Several independent spaces ("worlds") can be packed into one description;
 * they are stepped together, each with its own parameters, and never interact.
 *
 * Errors: every entry point returns CPGPU_OK (0) or a negative code and
 * records a message retrievable with cpgpuGetLastError().  The reference
 * aborts through cpAssertHard (include/chipmunk/chipmunk.h:48-58); this
 * library returns the same condition text as an error instead, so the caller
 * (for instance the cpGPUSpaceStep binding) decides whether to abort.
 */
#ifndef CPGPU_H
#define CPGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPGPU_OK 0
#define CPGPU_ERR_INVALID -1      /* bad argument or unsupported feature in the description */
#define CPGPU_ERR_HIP -2          /* a HIP runtime call failed */
#define CPGPU_ERR_CAPACITY -3     /* a fixed device capacity was exceeded */
#define CPGPU_ERR_DEVICE -4       /* the device-side error word was raised during a step */
#define CPGPU_ERR_FAULT -5        /* the step ran in full and failed one of the checks a debug build of the
                                     reference makes (CPGPU_FAULT_*, cpgpuWorldGetFaults) */

/* Body types, same numbering as cpBodyType (include/chipmunk/cpBody.h:30-45). */
#define CPGPU_BODY_DYNAMIC 0
#define CPGPU_BODY_KINEMATIC 1
#define CPGPU_BODY_STATIC 2

/* Shape types, same numbering as cpShapeType (include/chipmunk/chipmunk_structs.h:166-171). */
#define CPGPU_SHAPE_CIRCLE 0
#define CPGPU_SHAPE_SEGMENT 1
#define CPGPU_SHAPE_POLY 2

/* Constraint types (the joint classes of src/cpPivotJoint.c, cpDampedSpring.c,
 * cpSlideJoint.c, cpGearJoint.c, cpPinJoint.c, cpGrooveJoint.c,
 * cpRotaryLimitJoint.c, cpRatchetJoint.c, cpSimpleMotor.c,
 * cpDampedRotarySpring.c). */
#define CPGPU_JOINT_PIVOT 0
#define CPGPU_JOINT_DAMPED_SPRING 1
#define CPGPU_JOINT_SLIDE 2
#define CPGPU_JOINT_GEAR 3
#define CPGPU_JOINT_PIN 4
#define CPGPU_JOINT_GROOVE 5
#define CPGPU_JOINT_ROTARY_LIMIT 6
#define CPGPU_JOINT_RATCHET 7
#define CPGPU_JOINT_SIMPLE_MOTOR 8
#define CPGPU_JOINT_DAMPED_ROTARY_SPRING 9
#define CPGPU_JOINT_TYPES 10

/* Arbiter states, same numbering as enum cpArbiterState (chipmunk_structs.h:83-95). */
#define CPGPU_ARB_FIRST_COLLISION 0
#define CPGPU_ARB_NORMAL 1
#define CPGPU_ARB_IGNORE 2
#define CPGPU_ARB_CACHED 3

/* Arithmetic of the device state. */
#define CPGPU_FP64 0
#define CPGPU_FP32 1

/* Per-space parameters: the cpSpace fields read by cpSpaceStep
 * (chipmunk_structs.h:396-442; defaults set in src/cpSpace.c:131-156). */
typedef struct cpgpuSpaceParams {
	int32_t iterations;            /* cpSpace.iterations (default 10) */
	uint32_t collision_persistence;/* cpSpace.collisionPersistence (default 3) */
	double gravity_x, gravity_y;   /* cpSpace.gravity */
	double damping;                /* cpSpace.damping (default 1) */
	double collision_slop;         /* cpSpace.collisionSlop (default 0.1) */
	double collision_bias;         /* cpSpace.collisionBias (default pow(0.9, 60)) */
	double curr_dt;                /* cpSpace.curr_dt: dt of the previous step, 0 for a fresh space */
	uint32_t stamp;                /* cpSpace.stamp */
	uint32_t _pad;
} cpgpuSpaceParams;

/* Packed description of one or more spaces.  Arrays marked [n][2] hold x,y
 * pairs; [n][6] transforms are cpTransform {a, b, c, d, tx, ty}
 * (include/chipmunk/chipmunk_types.h:257-259).  Bodies, shapes and
 * constraints of world w occupy the index ranges
 * [world_*_start[w], world_*_start[w+1]). */
typedef struct cpgpuWorldDesc {
	int32_t n_worlds;
	const cpgpuSpaceParams *space;       /* [n_worlds] */
	const int32_t *world_body_start;     /* [n_worlds+1] */
	const int32_t *world_shape_start;    /* [n_worlds+1] */
	const int32_t *world_constraint_start; /* [n_worlds+1]; may be NULL when n_constraints == 0 */

	/* Bodies (struct cpBody, chipmunk_structs.h:35-81). */
	int32_t n_bodies;
	const int32_t *body_type;            /* CPGPU_BODY_* */
	const double *body_m, *body_m_inv;   /* mass and inverse (INFINITY / 0 for static and kinematic) */
	const double *body_i, *body_i_inv;   /* moment and inverse */
	const double *body_cog;              /* [n][2] centre of gravity, body-local */
	const double *body_p;                /* [n][2] position of the centre of gravity */
	const double *body_v;                /* [n][2] */
	const double *body_f;                /* [n][2] force accumulated for the next step */
	const double *body_a, *body_w, *body_t; /* angle, angular velocity, torque */
	const double *body_v_bias;           /* [n][2] pseudo-velocity */
	const double *body_w_bias;
	const double *body_transform;        /* [n][6] */

	/* Shapes (struct cpShape and subclasses, chipmunk_structs.h:177-236). */
	int32_t n_shapes;
	const int32_t *shape_type;           /* CPGPU_SHAPE_* */
	const int32_t *shape_body;           /* global body index */
	const uint64_t *shape_hashid;        /* cpShape.hashid (assigned by cpSpaceAddShape, src/cpSpace.c:432) */
	const double *shape_e, *shape_u;     /* elasticity, friction */
	const double *shape_surface_v;       /* [n][2] */
	const uint64_t *shape_group;         /* cpShapeFilter.group */
	const uint32_t *shape_categories;    /* cpShapeFilter.categories */
	const uint32_t *shape_mask;          /* cpShapeFilter.mask */
	const double *shape_radius;          /* circle r, segment r, poly bevel r */
	/* [n][10] body-local geometry: circle {c.x, c.y}; segment {a.x, a.y, b.x, b.y,
	 * n.x, n.y, a_tangent.x, a_tangent.y, b_tangent.x, b_tangent.y}; unused for polys. */
	const double *shape_geom;
	/* Polygons: vertices and edge normals in body-local space, in the order
	 * cpPolyShape stores them (planes[i+count] of src/cpPolyShape.c:147-165). */
	const int32_t *shape_poly_start;     /* [n] first vertex of the shape in poly_verts */
	const int32_t *shape_poly_count;     /* [n] vertex count (0 for non-polys), at most CPGPU_MAX_POLY_VERTS */
	int32_t n_poly_verts;
	const double *poly_verts;            /* [n_poly_verts][2] */
	const double *poly_normals;          /* [n_poly_verts][2] */

	/* Constraints (struct cpConstraint and joints, chipmunk_structs.h:250-391). */
	int32_t n_constraints;
	const int32_t *cons_type;            /* CPGPU_JOINT_* */
	const int32_t *cons_a, *cons_b;      /* global body indices */
	const double *cons_max_force, *cons_error_bias, *cons_max_bias;
	const int32_t *cons_collide_bodies;  /* 0 or 1 */
	/* [n][8] joint parameters:
	 *   pivot   {anchorA.x, anchorA.y, anchorB.x, anchorB.y}
	 *   spring  {anchorA.x, anchorA.y, anchorB.x, anchorB.y, restLength, stiffness, damping}
	 *   slide   {anchorA.x, anchorA.y, anchorB.x, anchorB.y, min, max}
	 *   gear    {phase, ratio, ratio_inv}
	 *   pin     {anchorA.x, anchorA.y, anchorB.x, anchorB.y, dist}
	 *   groove  {grv_a.x, grv_a.y, grv_b.x, grv_b.y, anchorB.x, anchorB.y, grv_n.x, grv_n.y}
	 *           (grv_n = cpvperp(cpvnormalize(grv_b - grv_a)), as cpGrooveJointInit keeps it)
	 *   rotary limit {min, max}
	 *   ratchet {phase, ratchet}
	 *   simple motor {rate}
	 *   damped rotary spring {restAngle, stiffness, damping} (the default spring torque) */
	const double *cons_params;
	/* [n][2] state carried between steps: the accumulated impulse (pivot and
	 * groove jAcc.x/.y, slide and pin jnAcc, every other joint's jAcc) and, for
	 * a ratchet, its angle in [1] (cpRatchetJointInit starts it at
	 * b->a - a->a); NULL means zero. */
	const double *cons_jacc;

	/* [n_shapes] 1 for a sensor shape (cpShape.sensor, src/cpShape.c:120-130), else 0;
	 * NULL: no sensors.  With the default collision handler a sensor's contacts are
	 * never solved (cpSpaceStep.c:268-270), so its pairs are dropped with the
	 * infinite-mass pairs before the narrowphase (and are not candidate pairs);
	 * the nearest-point and first-segment queries skip sensors (cpSpaceQuery.c:64,145). */
	const uint8_t *shape_sensor;

	/* Collision handlers (cpSpaceAddCollisionHandler, cpSpaceAddWildcardHandler,
	 * cpSpaceAddDefaultCollisionHandler; src/cpSpace.c:382-413).  shape_ctype
	 * [n_shapes]: each shape's collision type (cpShape.type, as 32 bits; NULL:
	 * all 0).  handlers [n_handlers]: which callbacks exist for which types; a
	 * pair of shapes takes the events of every entry that matches it (see
	 * cpgpuHandlerDesc).  A pair with events goes through the narrowphase even
	 * when it is never solved (a sensor, two infinite masses), as in
	 * cpSpaceCollideShapes; its begin and preSolve are answered by the host
	 * between cpgpuWorldStepBegin and cpgpuWorldStepEnd, its postSolve and
	 * separate after the step. */
	const uint32_t *shape_ctype;
	int32_t n_handlers;
	const struct cpgpuHandlerDesc *handlers;
} cpgpuWorldDesc;

/* collision handler events */
#define CPGPU_EVENT_BEGIN 1
#define CPGPU_EVENT_PRE_SOLVE 2
#define CPGPU_EVENT_POST_SOLVE 4
#define CPGPU_EVENT_SEPARATE 8
/* the wildcard collision type (CP_WILDCARD_COLLISION_TYPE) */
#define CPGPU_WILDCARD 0xFFFFFFFFu

/* One handler: the space it belongs to (-1: every space of the description)
 * and its types.  {a, b} matches the pairs of types a and b in either order;
 * {a, CPGPU_WILDCARD} every pair with a shape of type a (a wildcard handler);
 * {CPGPU_WILDCARD, CPGPU_WILDCARD} every pair (a default handler).  events:
 * the CPGPU_EVENT_* callbacks it has that are not the do-nothing defaults. */
typedef struct cpgpuHandlerDesc {
	int32_t space;
	uint32_t type_a, type_b;
	uint32_t events;
} cpgpuHandlerDesc;

#define CPGPU_MAX_POLY_VERTS 64

typedef struct cpgpuWorld cpgpuWorld;

/* Step statistics of the most recent step, summed over worlds. */
typedef struct cpgpuStats {
	int64_t candidate_pairs;   /* pairs that passed the exact QueryReject predicate */
	int64_t arbiters;          /* arbiters solved this step (cpSpace.arbiters) */
	int64_t contacts;          /* contacts solved this step */
	int64_t cached_arbiters;   /* arbiters held in the persistence table */
	int32_t arbiter_colours;   /* colours of the arbiter graph */
	int32_t constraint_colours;/* colours of the constraint graph */
	int64_t steps;             /* steps taken since creation */
	int64_t total_arbiters;    /* arbiters solved, summed over all steps since creation */
	int64_t total_contacts;    /* contacts solved, summed over all steps since creation */
	int64_t total_pairs;       /* candidate pairs, summed over all steps since creation */
} cpgpuStats;

/* ---- lifecycle (replaces cpHastySpaceNew / cpHastySpaceFree, src/cpHastySpace.c:560-592) ---- */

/* Upload a packed description to the current HIP device.  precision is
 * CPGPU_FP64 (bit-for-bit the reference's double arithmetic) or CPGPU_FP32. */
int cpgpuWorldCreate(const cpgpuWorldDesc *desc, int32_t precision, cpgpuWorld **out);
void cpgpuWorldDestroy(cpgpuWorld *world);
/* A new world from a description of the same spaces packed again after
 * bodies, shapes or constraints were added or removed (cpSpaceAdd* /
 * cpSpaceRemove*, src/cpSpace.c:417-571), which keeps the old world's
 * persistent collision state: shape_from[s] is the old world's index of the
 * description's shape s, or -1 for a new shape.  Every arbiter and GJK
 * warm-start id of the old world's last step whose two shapes are both in
 * the new description moves over (cached impulses, states, stamps), and every
 * dynamic shape keeps its fattened tree box and the step of its last rebuild,
 * so the next step warm-starts exactly as cpSpaceStep would after the same
 * adds and removes; arbiters of removed shapes are dropped, as
 * cpSpaceRemoveShape does through cpSpaceFilterArbiters (src/cpSpace.c:513-541).  Body state, joint
 * accumulators and space parameters come from the description.  The old
 * world is left as it was (the caller destroys it); same precision; not
 * during a taped rollout. */
int cpgpuWorldCreateFrom(const cpgpuWorldDesc *desc, int32_t precision, cpgpuWorld *old, const int32_t *shape_from,
                         cpgpuWorld **out);

/* ---- the step (replaces cpSpaceStep, src/cpSpaceStep.c:335-445, and
 * cpHastySpaceStep, src/cpHastySpace.c:594-700) ---- */

/* Advance every world by dt.  Asynchronous: returns once the work is queued. */
int cpgpuWorldStep(cpgpuWorld *world, double dt);
/* Advance every world by dt, n_steps times. */
int cpgpuWorldStepN(cpgpuWorld *world, double dt, int32_t n_steps);

/* ---- collision handler callbacks: a step in two halves around the host ----
 *
 * cpgpuWorldStepBegin runs a step up to the point where cpSpaceCollideShapes
 * (cpSpaceStep.c:253-284) would call begin and preSolve: positions, shapes,
 * broadphase, narrowphase and cpArbiterUpdate.  *n_callbacks gets the number
 * of touching arbiters whose handler has a begin (due on a first collision)
 * or a preSolve; cpgpuWorldReadCallbacks copies them out, the host runs the
 * callbacks, cpgpuWorldWriteDecisions hands back what they decided (an
 * arbiter left out keeps the defaults: solved, values unchanged), and
 * cpgpuWorldStepEnd finishes the step (cpSpaceStep.c:377-445).  An arbiter
 * that is not solved -- rejected by preSolve, ignored, a sensor's, or between
 * two infinite masses -- keeps no contacts (its count is 0, as
 * cpSpaceCollideShapes leaves it), is not among the step's arbiters, and an
 * ignored one stays ignored, with no begin and no preSolve, until its shapes
 * separate.  cpgpuWorldStep is the two halves with the default decisions.
 * After the step, postSolve is due for the solved arbiters of handler pairs
 * with that event (cpgpuWorldReadArbiters gives their impulses), separate for
 * the pairs cpgpuWorldReadSeparations lists. */
typedef struct cpgpuCallbackArbiter {
	int32_t slot;             /* the arbiter's handle for its decision */
	int32_t space;
	int32_t shape_a, shape_b; /* in the arbiter's order (cpCollide's a and b) */
	int32_t state;            /* CPGPU_ARB_FIRST_COLLISION (begin is due) or CPGPU_ARB_NORMAL */
	int32_t events;           /* the pair's CPGPU_EVENT_* */
	int32_t count;
	int32_t _pad;
	double normal[2];
	double e, u, surface_vr[2];
	/* per contact: r1.x, r1.y, r2.x, r2.y (offsets from the bodies' positions),
	 * the warm-started jnAcc, jtAcc */
	double contact[2][6];
	uint64_t hash[2];
} cpgpuCallbackArbiter;

typedef struct cpgpuCallbackDecision {
	int32_t slot;
	int32_t solve;   /* 0: not this step (begin or preSolve returned false) */
	int32_t ignore;  /* 1: cpArbiterIgnore was called (not solved until the shapes separate) */
	int32_t _pad;
	double e, u, surface_vr[2];   /* the arbiter's values for the solve (cpArbiterSet*) */
} cpgpuCallbackDecision;

int cpgpuWorldStepBegin(cpgpuWorld *world, double dt, int64_t *n_callbacks);
int cpgpuWorldReadCallbacks(cpgpuWorld *world, int64_t cap, int64_t *n_out, cpgpuCallbackArbiter *out);
int cpgpuWorldWriteDecisions(cpgpuWorld *world, int64_t n, const cpgpuCallbackDecision *decisions);
int cpgpuWorldStepEnd(cpgpuWorld *world);
/* The pairs whose arbiter separated in the last step (cpSpaceArbiterSetFilter,
 * cpSpaceStep.c:306-310: touched the step before, not in this one) and whose
 * handler has a separate event: pairs [cap][3] {space, shape_a, shape_b}. */
int cpgpuWorldReadSeparations(cpgpuWorld *world, int64_t cap, int64_t *n_out, int32_t *pairs);
/* Wait for queued steps and check the device error word. */
int cpgpuWorldSync(cpgpuWorld *world);

/* The conditions a debug build of the reference asserts during cpSpaceStep
 * (a release build, -DNDEBUG, checks none of them and steps on with
 * infinities and NaNs).  The device checks them at the same points and steps
 * on as the release build does.  A fault is returned once, as CPGPU_ERR_FAULT
 * with the reference's message, by the first step or cpgpuWorldSync after the
 * device raised it, once that call has done all its work: a fault of a step's
 * position update by the step itself, one of its solve (effective masses,
 * velocity update) by the next step or the cpgpuWorldSync before it.  Readers
 * never return it.  The bits stay set for the world's lifetime. */
#define CPGPU_FAULT_UNSOLVABLE 128  /* an effective mass of zero: "Unsolvable collision or constraint."
                                       (k_scalar, include/chipmunk/chipmunk_private.h:223) or "Unsolvable
                                       constraint." (k_tensor, :255) */
#define CPGPU_FAULT_BODY_STATE 512  /* cpBodySanityCheck after a position or velocity update (src/cpBody.c:110-125,
                                       508, 521): a NaN or negative mass or moment, a NaN or infinite
                                       position, velocity, force, angle, angular velocity or torque */
#define CPGPU_FAULT_BODY_MASS 1024  /* "Body's mass and moment must be positive to simulate." (src/cpBody.c:498) */
/* The CPGPU_FAULT_* raised so far (as of the last call that waited for the
 * device) and the lowest body index that failed a body check (-1: none).
 * The device keeps a body's moment only as its inverse and checks it through
 * that. */
int cpgpuWorldGetFaults(cpgpuWorld *world, int32_t *faults, int32_t *first_body);

/* ---- state readback (the role of the cpBody getters after cpSpaceStep) ---- */

/* Copy body state to host arrays sized like the description; any pointer may be NULL. */
int cpgpuWorldReadBodies(cpgpuWorld *world, double *p, double *v, double *a, double *w,
                         double *v_bias, double *w_bias, double *transform);
/* Overwrite body velocities / forces from host arrays (the cpBodySetVelocity /
 * cpBodySetForce setters); any pointer may be NULL. */
int cpgpuWorldWriteBodies(cpgpuWorld *world, const double *v, const double *w,
                          const double *f, const double *t);
/* Overwrite the state of n bodies listed by global index, from host arrays of
 * n entries each: positions p [n][2], velocities v [n][2], forces f [n][2],
 * angles a, angular velocities w, torques t [n], transforms [n][6] (the
 * cpBodySetPosition / SetVelocity / SetForce / SetAngle / SetAngularVelocity
 * / SetTorque setters, src/cpBody.c:375-470, which also set the transform).
 * Any pointer may be NULL.  A static body's pose is not written here: static
 * shapes are cached once per world (cpSpaceReindexStatic's role), so a moved
 * static body needs cpgpuWorldCreateFrom. */
int cpgpuWorldWriteBodiesAt(cpgpuWorld *world, int32_t n, const int32_t *index, const double *p, const double *v,
                            const double *f, const double *a, const double *w, const double *t,
                            const double *transform);
/* Replace one space's parameters (the cpSpaceSetGravity / SetIterations /
 * SetDamping / SetCollisionSlop / SetCollisionBias / SetCollisionPersistence
 * setters, src/cpSpace.c:240-300; stamp and curr_dt as the space holds them). */
int cpgpuWorldSetSpaceParams(cpgpuWorld *world, int32_t space, const cpgpuSpaceParams *params);
int cpgpuWorldGetStats(cpgpuWorld *world, cpgpuStats *out);

/* ---- snapshots (checkpoint and resume; no counterpart in the reference,
 * where a cpSpace is saved by its owner's own serialisation) ----
 * Everything a step carries into the next, copied to a host buffer: the
 * spaces' parameters with their stamps and last dt, the bodies' state
 * (positions, angles, forces, torques, transforms, velocities and bias
 * velocities), the dynamic tree's leaf boxes and rebuild stamps, the joints
 * (parameters and accumulated impulses), the last step's pair records with
 * their GJK warm-start ids and its solved and cached arbiters (contacts,
 * impulses, stamps), and the running totals.  Loaded into a world created
 * from the same description (same precision and build of this library), the
 * next steps are bit for bit those of the world it was saved from.
 * Layout: a header (magic "cpgpuSS1", version 1, precision, the counts and
 * the struct sizes of the build), then cpgpuSpaceParams[n_worlds], then the
 * device arrays in order -- positions, angles, forces, torques, transforms,
 * velocity records, leaf boxes, rebuild stamps, joints, pair records, solved
 * arbiters, cached arbiters, totals -- in the build's own layouts.
 * cpgpuWorldSaveState: with buf NULL only *size_out is set (the bytes
 * needed); otherwise cap must hold them (else CPGPU_ERR_CAPACITY). */
int cpgpuWorldSaveState(cpgpuWorld *world, void *buf, int64_t cap, int64_t *size_out);
int cpgpuWorldLoadState(cpgpuWorld *world, const void *buf, int64_t size);

/* Copy the arbiters solved in the last step (cpSpace.arbiters, in solve
 * order: by colour, then by key).  Arrays have room for `cap` arbiters;
 * returns the number of arbiters through *n_out.  Any pointer may be NULL.
 *   shapes  [cap][2] global shape indices (a, b) in narrowphase orientation
 *   colour  [cap]    solve colour
 *   state   [cap]    CPGPU_ARB_*
 *   count   [cap]    contact count (1 or 2)
 *   normal  [cap][2]
 *   contact [cap][2][12] per contact {r1.x, r1.y, r2.x, r2.y, nMass, tMass,
 *                    bounce, jnAcc, jtAcc, jBias, bias, 0}
 *   hash    [cap][2] contact feature hashes */
int cpgpuWorldReadArbiters(cpgpuWorld *world, int64_t cap, int64_t *n_out,
                           int32_t *shapes, int32_t *colour, int32_t *state, int32_t *count,
                           double *normal, double *contact, uint64_t *hash);
/* Copy the candidate pairs of the last step: the pairs that passed the exact
 * QueryReject predicate (src/cpSpaceStep.c:204-232) and were handed to the
 * narrowphase (cpSpaceCollideShapes, cpSpaceStep.c:235), as the reference's
 * cpSpatialIndexReindexQuery (include/chipmunk/cpSpatialIndex.h:222-225)
 * followed by QueryReject would report them.  Up to cap pairs go to
 * pairs[cap][2] as global shape indices (lower, higher), sorted; the total
 * goes to *n_out. */
int cpgpuWorldReadPairs(cpgpuWorld *world, int64_t cap, int64_t *n_out, int32_t *pairs);
/* cpShapesCollide (include/chipmunk/cpShape.h:95, src/cpShape.c:260-283) for
 * n pairs of the world's shapes, on their geometry as cached by the last step
 * (by creation before the first): pairs[n][2] global shape indices (a, b);
 * count[n] gets each contact point set's count (0..2) and sets[n][12]
 * {normal.x, normal.y, then for each of the two points pointA.x, pointA.y,
 * pointB.x, pointB.y, distance} (zeros past the count), in the caller's (a, b)
 * order as cpShapesCollide returns them.  Blocks until the results are on the
 * host.  Stages the call in device memory it allocates and frees. */
int cpgpuWorldShapesCollide(cpgpuWorld *world, int64_t n, const int32_t *pairs, int32_t *count, double *sets);
/* ---- spatial queries (replace cpSpacePointQueryNearest, cpSpaceSegmentQueryFirst,
 * cpSpacePointQuery, cpSpaceSegmentQuery and cpSpaceBBQuery, src/cpSpaceQuery.c:23-206)
 * on the shapes as the last step cached them (as cached at creation before
 * the first step), n queries at once, one device lane each.  space[n] names
 * each query's space (NULL: space 0); filter[n] is its cpShapeFilter (NULL:
 * CP_SHAPE_FILTER_ALL).  Every shape whose exact check passes is found, so
 * results are those of checking every shape; of equal nearest distances or
 * first alphas the lowest shape index is kept, and a segment with a radius
 * also finds shapes within the radius that the reference's tree prunes (it
 * tests node boxes against the zero-width segment).  Each call blocks until
 * its results are on the host and stages its arrays in device memory it
 * allocates and frees. */
typedef struct cpgpuShapeFilter {
	uint64_t group;        /* cpShapeFilter.group */
	uint32_t categories;   /* cpShapeFilter.categories */
	uint32_t mask;         /* cpShapeFilter.mask */
} cpgpuShapeFilter;
/* cpSpacePointQueryNearest: points [n][2], max_distance [n]; shape_out [n]
 * (global shape index, -1 for none), info_out [n][5] {point.x, point.y,
 * distance, gradient.x, gradient.y} (cpPointQueryInfo; with no shape
 * {0, 0, max_distance, 0, 0}). */
int cpgpuWorldQueryPointNearest(cpgpuWorld *world, int32_t n, const int32_t *space, const double *points,
                                const double *max_distance, const cpgpuShapeFilter *filter, int32_t *shape_out,
                                double *info_out);
/* cpSpaceSegmentQueryFirst: starts, ends [n][2], radius [n]; shape_out [n],
 * info_out [n][5] {point.x, point.y, normal.x, normal.y, alpha}
 * (cpSegmentQueryInfo; with no shape {end, 0, 0, 1}). */
int cpgpuWorldQuerySegmentFirst(cpgpuWorld *world, int32_t n, const int32_t *space, const double *starts,
                                const double *ends, const double *radius, const cpgpuShapeFilter *filter,
                                int32_t *shape_out, double *info_out);
/* cpSpacePointQuery / cpSpaceSegmentQuery / cpSpaceBBQuery: every hit of
 * every query, hits [cap][2] {query, shape} sorted by query then shape, with
 * info [cap][5] as above (point queries: {point, distance, gradient};
 * segment queries: {point, normal, alpha}; box queries: none, info may be
 * NULL).  The total goes to *n_out; when it exceeds cap, call again with
 * room for it.  Boxes are {l, b, r, t} [n][4]. */
int cpgpuWorldQueryPointAll(cpgpuWorld *world, int32_t n, const int32_t *space, const double *points,
                            const double *max_distance, const cpgpuShapeFilter *filter, int64_t cap, int64_t *n_out,
                            int32_t *hits, double *info);
int cpgpuWorldQuerySegmentAll(cpgpuWorld *world, int32_t n, const int32_t *space, const double *starts,
                              const double *ends, const double *radius, const cpgpuShapeFilter *filter, int64_t cap,
                              int64_t *n_out, int32_t *hits, double *info);
int cpgpuWorldQueryBBAll(cpgpuWorld *world, int32_t n, const int32_t *space, const double *boxes,
                         const cpgpuShapeFilter *filter, int64_t cap, int64_t *n_out, int32_t *hits);
/* cpSpaceShapeQuery (cpSpaceQuery.c:231-246) with shapes of the world as the
 * query shapes.  As there, the query shape is first put at its body's
 * present transform (cpShapeUpdate), the other shapes keep the pose the last
 * step cached; unlike there, the world's cached copy of the query shape is
 * left as it was.  For each shapes[q]: every other shape of its space whose
 * box meets the query shape's, that passes the two shapes' filters and that
 * touches it, with the cpShapesCollide(query shape, shape) contact point set
 * -- hits [cap][2] {query, shape} sorted by query then shape, counts [cap],
 * sets [cap][12] {normal.x, normal.y, then per point pointA.x, pointA.y,
 * pointB.x, pointB.y, distance}; the total goes to *n_out.  (The reference's
 * dynamic tree prunes by fattened leaf boxes and may also report degenerate
 * contacts -- zero normal -- with shapes whose own boxes miss the query
 * shape's.)  The reference's return value is true when the query shape is no
 * sensor and the last touching shape it visits is none: without sensors, when
 * there is a hit.  (A shape not in the space is queried by adding it as a
 * sensor: it then touches nothing in the step.) */
int cpgpuWorldQueryShapes(cpgpuWorld *world, int32_t n, const int32_t *shapes, int64_t cap, int64_t *n_out,
                          int32_t *hits, int32_t *counts, double *sets);
/* Copy constraint accumulated impulses ([n][2], see cons_jacc) and colours. */
int cpgpuWorldReadConstraints(cpgpuWorld *world, double *jacc, int32_t *colour);

/* ---- profiling ---- */

/* Record HIP events around kernel launches: CPGPU_TIMING_OFF (the default),
 * CPGPU_TIMING_ALL (every launch; the events cost a few microseconds each),
 * CPGPU_TIMING_SOLVE (only the velocity solver's and the adjoint's launches:
 * cheap enough to time them inside a measured run). */
#define CPGPU_TIMING_OFF 0
#define CPGPU_TIMING_ALL 1
#define CPGPU_TIMING_SOLVE 2
int cpgpuWorldSetTiming(cpgpuWorld *world, int32_t enable);
/* Per-kernel accumulated device time in ms and launch counts since timing was
 * enabled.  names is filled with a '\n'-separated list of kernel names.
 * Returns the number of kernels through *n_out. */
int cpgpuWorldGetTiming(cpgpuWorld *world, int32_t cap, int32_t *n_out, double *ms,
                        int64_t *launches, char *names, int32_t names_cap);

/* ---- batch tiling (no counterpart in the reference: a batch of spaces is
 * this library's own form).  Spaces of at most 4096 dynamic bodies are
 * solved one workgroup each, in up to CPGPU_MAX_BUCKETS size buckets
 * (bucket q takes the spaces with more bodies than bucket q-1's limit and at
 * most max_bodies[q], static bodies included; the last takes the rest).
 * Each bucket has its own workgroup size (64, 128, 256, 512 or 1024
 * threads) and stages its spaces' bodies in LDS when they fit; the buckets
 * run concurrently.  Defaults: limits 64, 256, 1024, unbounded with 64,
 * 128, 256, 512 threads. */
#define CPGPU_MAX_BUCKETS 4
int cpgpuWorldGetBuckets(cpgpuWorld *world, int32_t *max_bodies, int32_t *threads, int32_t *n_worlds,
                         int32_t *lds_bodies);
int cpgpuWorldSetBuckets(cpgpuWorld *world, int32_t n, const int32_t *max_bodies, const int32_t *threads);
/* Autotuner: steps the world 5 x trial_steps times (the state advances),
 * timing every bucket's solve at each workgroup size, and keeps the fastest
 * size per bucket (reported in threads_out[CPGPU_MAX_BUCKETS]). */
int cpgpuWorldAutotune(cpgpuWorld *world, double dt, int32_t trial_steps, int32_t *threads_out);
/* Schedule of the grid set of spaces with more than 4096 dynamic bodies
 * (their dynamic bodies, cut into x-strips, with the seams between them).  CPGPU_GRID_FLOW (the default): without
 * joints or arbiters spanning non-neighbouring strips, every strip and seam
 * waits only for the units it shares bodies with; CPGPU_GRID_PHASES: grid
 * barriers between strips, even seams and odd seams.  Results are the same
 * bit for bit; the schedule only changes the time. */
#define CPGPU_GRID_FLOW 0
#define CPGPU_GRID_PHASES 1
int cpgpuWorldSetGridSchedule(cpgpuWorld *world, int32_t schedule);

/* Steps captured in HIP graphs (the north star's one launch sequence per step
 * for a batch; no counterpart in the reference).  With graphs on, each
 * cpgpuWorldStep launches one executable graph that holds the whole step --
 * the per-space scalars (stamp, dt_coef) advanced on the device, integration,
 * broadphase, narrowphase, persistence and the bucketed solve -- with no host
 * round trip: launches are sized for buffer capacities of twice the largest
 * counts seen, the kernels read the counts on the device.  The graphs are
 * captured on the first step and again when dt, the tiling, a space's
 * parameters or (seen after a step has finished) a count above half its
 * capacity call for it.  A step that still outgrew a capacity raises the
 * device error word (reported by the next synchronising call).  Results are
 * bit for bit those of steps without graphs.  Not with a taped rollout, with
 * per-launch timing, or for spaces of more than 4096 dynamic bodies (the grid
 * solve). */
int cpgpuWorldSetGraph(cpgpuWorld *world, int32_t enable);

/* ---- differentiation (no counterpart in the reference: cpFloat is a plain
 * double there).  cpgpuWorldTapeBegin starts recording the rollout: every
 * later cpgpuWorldStep records its solved arbiters, and the bodies' state is
 * checkpointed every checkpoint_every steps.  cpgpuWorldBackward then takes
 * the adjoint (dL/d state) of the state after the last recorded step --
 * positions g_p [n][2], angles g_a, velocities g_v [n][2], angular velocities
 * g_w (NULL: zero) -- and returns dL/d of the state at the tape's start
 * (d_p0, d_a0, d_v0, d_w0) and of the parameters over the rollout: each
 * body's 1/m and 1/I (d_m_inv, d_i_inv) and each shape's friction
 * coefficient (d_u [n_shapes]).  Any output may be NULL.  The backward pass
 * replays each checkpoint segment with the step's own kernels and runs the
 * hand-written reverse kernels; the world is left as it was.  The contact
 * set and the clamp branches of the rollout are held fixed (derivatives
 * where they change are one-sided).  Joints (pivot, damped spring, slide,
 * gear) are differentiated too; their accumulated impulses carry adjoints
 * between steps like the contacts' warm starts; a world with any other
 * joint type is refused.  Spaces must be solved one
 * workgroup each; polygons of at most 8 vertices; CPGPU_FP64 worlds; no
 * velocity or force writes during a taped rollout. */
int cpgpuWorldTapeBegin(cpgpuWorld *world, int32_t checkpoint_every);
int cpgpuWorldTapeEnd(cpgpuWorld *world);
int cpgpuWorldBackward(cpgpuWorld *world, const double *g_p, const double *g_a, const double *g_v,
                       const double *g_w, double *d_p0, double *d_a0, double *d_v0, double *d_w0,
                       double *d_m_inv, double *d_i_inv, double *d_u);
/* After cpgpuWorldBackward: dL/d of each constraint's parameters over the
 * rollout, [n_constraints][8] in the layout of cpgpuWorldDesc::cons_params
 * (spring: anchors, restLength, stiffness, damping; slide: anchors, min, max;
 * pivot: anchors; gear: phase -- ratio entries are 0).  No counterpart in the
 * reference. */
int cpgpuWorldGetJointGrads(cpgpuWorld *world, double *d_params);

/* Contacts solved and bodies velocity-integrated since creation by each
 * solve unit, in the order of the timing names solver_worlds_b0 ..
 * solver_worlds_b3 (size buckets of small worlds), solver (the grid set of
 * the large spaces): CPGPU_SOLVE_UNITS entries each (instrumentation for
 * per-kernel byte counts; no counterpart in the reference). */
#define CPGPU_SOLVE_UNITS 5
int cpgpuWorldGetSolveCounts(cpgpuWorld *world, int64_t *contacts, int64_t *bodies);

const char *cpgpuGetLastError(void);
const char *cpgpuVersion(void);

#ifdef __cplusplus
}
#endif

#endif /* CPGPU_H */
