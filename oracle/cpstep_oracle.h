/*
 * cpstep_oracle.h -- CPU restatement of Chipmunk2D 7.0.3's cpSpaceStep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP step library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in chipmunk2d_amd/ links, imports or calls it.
 *
 * It steps the same packed description (include/cpgpu.h) the device library
 * takes, in plain C double arithmetic compiled with -ffp-contract=off, and
 * reproduces the device library's deterministic solve order (graph colours),
 * so device results can be compared bit for bit.  See cpstep_oracle.c for
 * the reference file:line each function follows and how it is pinned.
 */
#ifndef CPSTEP_ORACLE_H
#define CPSTEP_ORACLE_H

#include <stdint.h>
#include "../include/cpgpu.h"

typedef struct oworld oworld;

/* Collision handler callbacks (cpCollisionHandler's begin, preSolve,
 * postSolve, separate): event is one CPGPU_EVENT_*.  For begin and preSolve
 * the function reads the arbiter and writes its decision -- dec arrives with
 * solve = 1, ignore = 0 and the arbiter's e, u and surface velocity; dec->solve
 * is the callback's return value.  For postSolve the arbiter carries the
 * solved impulses; for separate only the shapes and the space matter and dec
 * is unused. */
typedef void (*ow_handler_fn)(void *ctx, int32_t event, const cpgpuCallbackArbiter *arb, cpgpuCallbackDecision *dec);

oworld *ow_new(const cpgpuWorldDesc *desc);
void ow_free(oworld *w);

/* One cpSpaceStep over every world. */
void ow_step(oworld *w, double dt);
/* The function that answers the description's collision handlers (NULL: the
 * default answers -- every arbiter solved, values unchanged). */
void ow_set_handler_fn(oworld *w, ow_handler_fn fn, void *ctx);
/* The last step's separations of pairs with a separate event: {space, a, b} [cap][3]. */
int64_t ow_read_separations(oworld *w, int64_t cap, int32_t *pairs);
/* The CPGPU_FAULT_* raised since the world was made (the checks a debug build
 * of the reference makes during a step) and the lowest body that failed a body
 * check (-1: none). */
int32_t ow_faults(const oworld *w, int32_t *first_body);

/* 1: body rotations use libm cos/sin (cpBody.c:350 via cpVect.h:104-107) and
 * the damped spring libm exp (cpDampedSpring.c:47) exactly as the reference
 * does.  0 (default): the deterministic cp_sincos / cp_exp that the device
 * library also uses. */
void ow_set_libm_trig(oworld *w, int enable);

/* 1: solve constraints in description order instead of colour order
 * (the reference's cpSpace.constraints order). */
void ow_set_joint_list_order(oworld *w, int enable);

/* Replace the colour order for the next step only: the arbiters are solved in
 * exactly this order and same-type pairs collide in this orientation.  pairs
 * holds n (shape a, shape b) global indices.  Used to replay the reference's
 * own order. */
void ow_set_next_order(oworld *w, int64_t n, const int32_t *pairs);

/* For the next step only: collide these pairs in the given orientation with
 * the given warm-start id ({a, b, id} triples), as the reference did. */
void ow_set_next_collisions(oworld *w, int64_t n, const int32_t *coll);

int ow_read_bodies(oworld *w, double *p, double *v, double *a, double *wv,
                   double *v_bias, double *w_bias, double *transform);
int ow_write_bodies(oworld *w, const double *v, const double *wv, const double *f, const double *t);

/* Same as cpgpuWorldWriteBodiesAt, cpgpuWorldSetSpaceParams and
 * cpgpuWorldCreateFrom (the old world is left as it was). */
int ow_write_bodies_at(oworld *w, int32_t n, const int32_t *index, const double *p, const double *v,
                       const double *f, const double *a, const double *wv, const double *t, const double *transform);
int ow_set_space_params(oworld *w, int32_t space, const cpgpuSpaceParams *sp);
oworld *ow_new_from(const cpgpuWorldDesc *desc, oworld *old, const int32_t *shape_from);

/* Same layout as cpgpuWorldReadArbiters. */
int64_t ow_read_arbiters(oworld *w, int64_t cap, int32_t *shapes, int32_t *colour, int32_t *state,
                         int32_t *count, double *normal, double *contact, uint64_t *hash);
/* solve unit of each arbiter of the last step (same order): 0 in a small
 * space; in a large space 1 + q for strip q (q < 64), 65 + q / 2 for even
 * seam q, 97 + q / 2 for odd seam q, 128 for the wide set */
int64_t ow_read_units(oworld *w, int64_t cap, int32_t *unit);
/* cpSpaceShapeQuery of shape s (a shape of the world) at its cached pose:
 * the touching shapes in shape order with counts [cap] and contact point
 * sets [cap][12]; returns their number. */
int64_t ow_shape_query(oworld *w, int32_t s, int64_t cap, int32_t *shapes, int32_t *counts, double *sets);
int ow_read_constraints(oworld *w, double *jacc, int32_t *colour);
void ow_stats(oworld *w, cpgpuStats *out);

/* ---- pieces exposed for the pinning tests ---- */

/* Recompute the world-space geometry of every shape from the current body
 * transforms (cpShapeUpdate for all shapes, cpShape.c:216-220). */
void ow_cache_all_shapes(oworld *w);
/* Shape bounding boxes [n_shapes][4] (l, b, r, t) as last cached. */
void ow_read_shape_bbs(oworld *w, double *bb);
/* Set body transforms directly ([n][6]), as cpBody.transform. */
void ow_set_transforms(oworld *w, const double *transform);

/* cpCollide(a, b, id) (cpCollision.c:701-726) on cached shapes.  Writes the
 * oriented shapes, the returned id, the normal, count and per contact
 * {r1.x, r1.y, r2.x, r2.y} (absolute points) and hashes.  Returns count. */
int ow_collide(oworld *w, int32_t a, int32_t b, uint32_t id, int32_t *ab_out, uint32_t *id_out,
               double *n_out, double *points_out, uint64_t *hash_out);

/* cpShapesCollide(a, b) (cpShape.c:260-283) on cached shapes: writes the
 * contact point set's normal and per point {pointA.x, pointA.y, pointB.x,
 * pointB.y, distance}; returns the count. */
int ow_shapes_collide(oworld *w, int32_t a, int32_t b, double *n_out, double *points_out);

/* Candidate pairs of the exact QueryReject predicate (cpSpaceStep.c:204-232)
 * for the current cached geometry: writes up to cap (a, b) pairs with a < b,
 * sorted; returns the total count. */
int64_t ow_candidate_pairs(oworld *w, int64_t cap, int32_t *pairs);

/* Spatial queries (cpSpaceQuery.c) on the shapes as the last step cached
 * them, in space `world`, filter {group, categories, mask}: the nearest
 * shape to a point (cpSpacePointQueryNearest; info {point, distance,
 * gradient}), the first shape along a segment (cpSpaceSegmentQueryFirst;
 * info {point, normal, alpha}), and all shapes near a point, along a segment
 * or meeting a box (cpSpacePointQuery, cpSpaceSegmentQuery, cpSpaceBBQuery),
 * in shape order with up to cap results written.  Every shape is checked. */
int32_t ow_point_query_nearest(oworld *w, int32_t world, const double *p, double max_distance, uint64_t group,
                               uint32_t cats, uint32_t mask, double *info);
int32_t ow_segment_query_first(oworld *w, int32_t world, const double *a, const double *b, double radius,
                               uint64_t group, uint32_t cats, uint32_t mask, double *info);
int64_t ow_point_query(oworld *w, int32_t world, const double *p, double max_distance, uint64_t group, uint32_t cats,
                       uint32_t mask, int64_t cap, int32_t *shapes, double *info);
int64_t ow_segment_query(oworld *w, int32_t world, const double *a, const double *b, double radius, uint64_t group,
                         uint32_t cats, uint32_t mask, int64_t cap, int32_t *shapes, double *info);
int64_t ow_bb_query(oworld *w, int32_t world, const double *bb, uint64_t group, uint32_t cats, uint32_t mask,
                    int64_t cap, int32_t *shapes);

/* The deterministic sin/cos shared with the device library. */
void cp_sincos_oracle(double x, double *s, double *c);
double cp_exp_oracle(double x);

#endif
