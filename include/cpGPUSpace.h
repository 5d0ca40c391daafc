/*
 * cpGPUSpace.h -- the drop-in for Chipmunk2D 7.0.3's cpSpaceStep on an
 * MI355X (gfx950): a cpSpace stepped by the HIP step library of cpgpu.h.
 *
 * This is the reference-side binding: it is compiled against Chipmunk's own
 * headers (it reads the private structs of include/chipmunk/chipmunk_structs.h,
 * as cpHastySpace does) and linked with Chipmunk and libcpgpu.so.  It follows
 * the plugin pattern of cpHastySpace (include/chipmunk/cpHastySpace.h:11-27,
 * src/cpHastySpace.c:397-416,560-592): a cpGPUSpace embeds a cpSpace as its
 * first member, is created with cpGPUSpaceNew, freed with cpGPUSpaceFree and
 * must be stepped with cpGPUSpaceStep.  Everything else -- adding and
 * removing bodies, shapes and constraints, the setters and getters -- is the
 * ordinary cpSpace / cpBody API.
 *
 * Device state is authoritative after a step.  The step is asynchronous:
 * body getters (cpBodyGetPosition, ...), cpBodyEachArbiter and the
 * constraints' impulses read what the last cpGPUSpaceSyncHost wrote back, so
 * call it before reading or relatively changing body state.  Between steps a
 * caller may set velocities, forces and torques (uploaded when they differ
 * from what the device holds), set positions and angles, change the space's
 * parameters, and add or remove bodies, shapes and constraints: such changes
 * are packed again before the next step, and the device keeps its cached
 * arbiters (warm starts and GJK ids) for every shape pair that survives.
 *
 * Supported on the device: dynamic, kinematic and static bodies with the
 * default integrators; circles, segments and convex polygons, sensors among
 * them, with the default collision handler; every joint class of the
 * reference (springs with their default force and torque) without pre/post-
 * solve callbacks; sleeping off.  Anything else is reported as an error at
 * pack time (custom handlers, custom spring functions, constraint callbacks,
 * sleeping), as cpAssertHard would (include/chipmunk/chipmunk.h:48-58).
 *
 * Errors: by default a failed call prints the message and aborts, like
 * cpAssertHard.  cpGPUSetAbortOnError(0) makes calls return instead, with the
 * message in cpGPUGetError().
 */
#ifndef CPGPUSPACE_H
#define CPGPUSPACE_H

#include <chipmunk/chipmunk.h>
#include "cpgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- one space (cpHastySpaceNew / Free / Step, src/cpHastySpace.c:560-700) ---- */

/* A new empty space whose steps run on the current HIP device. precision is
 * CPGPU_FP64 (the default of cpGPUSpaceNew: bit for bit the reference's double
 * arithmetic, solved in the device's colour order) or CPGPU_FP32. */
CP_EXPORT cpSpace *cpGPUSpaceNew(void);
CP_EXPORT cpSpace *cpGPUSpaceNewWithPrecision(int precision);
CP_EXPORT void cpGPUSpaceFree(cpSpace *space);
/* cpSpaceStep (src/cpSpaceStep.c:335-445) on the device; returns once the step is queued. */
CP_EXPORT void cpGPUSpaceStep(cpSpace *space, cpFloat dt);
/* Wait for the queued steps and write the device state back into the space's
 * structs: every body's position, velocity, angle, angular velocity, bias
 * velocities and transform; every shape's cached geometry and box
 * (cpShapeUpdate); every constraint's accumulated impulse; and the arbiters
 * of the last step, threaded onto their bodies' arbiter lists as
 * cpSpaceProcessComponents does (src/cpSpaceComponent.c:253-267), so that
 * cpBodyEachArbiter and the cpArbiter getters work.  They are also entered in
 * the space's cachedArbiters and arbiters, as a step of the reference leaves
 * them, so that cpSpaceRemoveShape / cpSpaceRemoveBody unthread them from the
 * remaining bodies (cpSpaceFilterArbiters, src/cpSpace.c:483-520). */
CP_EXPORT void cpGPUSpaceSyncHost(cpSpace *space);
/* The device world behind the space (NULL before the first step), for the
 * cpgpu.h calls that have no cpSpace counterpart (timing, readbacks). */
CP_EXPORT cpgpuWorld *cpGPUSpaceGetWorld(cpSpace *space);
/* The packed description of the last pack, and the index each body, shape
 * and constraint has in it (-1: not packed). */
CP_EXPORT const cpgpuWorldDesc *cpGPUSpaceGetDesc(cpSpace *space);
CP_EXPORT int cpGPUSpaceBodyIndex(cpSpace *space, const cpBody *body);
CP_EXPORT int cpGPUSpaceShapeIndex(cpSpace *space, const cpShape *shape);
CP_EXPORT int cpGPUSpaceConstraintIndex(cpSpace *space, const cpConstraint *constraint);
/* Packs since creation (the first step's included), and the last pack's
 * shape map: the previous pack's index of each shape, -1 for a new shape
 * (the shape_from argument of cpgpuWorldCreateFrom). */
CP_EXPORT int cpGPUSpaceGetPackCount(cpSpace *space);
CP_EXPORT const int32_t *cpGPUSpaceGetShapeMap(cpSpace *space);
/* The description the space packs into now, collected from its structs
 * without touching the device (NULL, with the error set, when the space uses
 * a feature the device does not run).  Valid until the next call on the
 * space.  Objects keep the order a pack gives them. */
CP_EXPORT const cpgpuWorldDesc *cpGPUSpaceDescribe(cpSpace *space);

/* ---- a batch of independent spaces stepped together ---- */

typedef struct cpGPUBatch cpGPUBatch;
/* The n spaces (plain cpSpaces, created with cpSpaceNew) are stepped as one
 * device world; each keeps its own parameters and never interacts with the
 * others.  The batch does not own the spaces: free the batch before them. */
CP_EXPORT cpGPUBatch *cpGPUBatchNew(cpSpace **spaces, int n);
CP_EXPORT cpGPUBatch *cpGPUBatchNewWithPrecision(cpSpace **spaces, int n, int precision);
CP_EXPORT void cpGPUBatchFree(cpGPUBatch *batch);
/* One cpSpaceStep of every space of the batch. */
CP_EXPORT void cpGPUBatchStep(cpGPUBatch *batch, cpFloat dt);
/* cpGPUSpaceSyncHost for every space of the batch. */
CP_EXPORT void cpGPUBatchSyncHost(cpGPUBatch *batch);
/* The size-bucket autotuner (cpgpuWorldAutotune): steps the batch
 * 5 x trial_steps times and keeps the fastest workgroup size per bucket. */
CP_EXPORT void cpGPUBatchAutotune(cpGPUBatch *batch, cpFloat dt, int trial_steps);
CP_EXPORT cpgpuWorld *cpGPUBatchGetWorld(cpGPUBatch *batch);
CP_EXPORT const cpgpuWorldDesc *cpGPUBatchGetDesc(cpGPUBatch *batch);
CP_EXPORT int cpGPUBatchGetPackCount(cpGPUBatch *batch);
CP_EXPORT const int32_t *cpGPUBatchGetShapeMap(cpGPUBatch *batch);
CP_EXPORT const cpgpuWorldDesc *cpGPUBatchDescribe(cpGPUBatch *batch);

/* ---- errors ---- */

CP_EXPORT void cpGPUSetAbortOnError(int enable);
CP_EXPORT const char *cpGPUGetError(void);
CP_EXPORT void cpGPUClearError(void);

#ifdef __cplusplus
}
#endif

#endif /* CPGPUSPACE_H */
