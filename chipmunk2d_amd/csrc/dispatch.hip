// dispatch.hip -- the C ABI of include/cpgpu.h over the library's two builds.
//
// Every device source is compiled twice: namespace cpgpu64 (real = double,
// the bit-exact FP64 step) and namespace cpgpu32 (real = float, the FP32
// mode).  A world remembers its precision in its first field; each entry
// point forwards to the build the world was created with.
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "../../include/cpgpu.h"
#include "cpgpu_errors.h"

thread_local std::string cpgpu_last_error;

#define CPGPU_DECLARE_API \
	int cpgpuWorldCreate(const cpgpuWorldDesc *desc, int32_t precision, cpgpuWorld **out); \
	void cpgpuWorldDestroy(cpgpuWorld *world); \
	int cpgpuWorldStep(cpgpuWorld *world, double dt); \
	int cpgpuWorldStepN(cpgpuWorld *world, double dt, int32_t n_steps); \
	int cpgpuWorldStepBegin(cpgpuWorld *world, double dt, int64_t *n_callbacks); \
	int cpgpuWorldReadCallbacks(cpgpuWorld *world, int64_t cap, int64_t *n_out, cpgpuCallbackArbiter *out); \
	int cpgpuWorldWriteDecisions(cpgpuWorld *world, int64_t n, const cpgpuCallbackDecision *decisions); \
	int cpgpuWorldStepEnd(cpgpuWorld *world); \
	int cpgpuWorldReadSeparations(cpgpuWorld *world, int64_t cap, int64_t *n_out, int32_t *pairs); \
	int cpgpuWorldSync(cpgpuWorld *world); \
	int cpgpuWorldGetFaults(cpgpuWorld *world, int32_t *faults, int32_t *first_body); \
	int cpgpuWorldReadBodies(cpgpuWorld *world, double *p, double *v, double *a, double *w, double *v_bias, double *w_bias, double *transform); \
	int cpgpuWorldWriteBodies(cpgpuWorld *world, const double *v, const double *w, const double *f, const double *t); \
	int cpgpuWorldWriteBodiesAt(cpgpuWorld *world, int32_t n, const int32_t *index, const double *p, const double *v, const double *f, const double *a, const double *w, const double *t, const double *transform); \
	int cpgpuWorldSetSpaceParams(cpgpuWorld *world, int32_t space, const cpgpuSpaceParams *params); \
	int cpgpuWorldCreateFrom(const cpgpuWorldDesc *desc, int32_t precision, cpgpuWorld *old, const int32_t *shape_from, cpgpuWorld **out); \
	int cpgpuWorldGetStats(cpgpuWorld *world, cpgpuStats *out); \
	int cpgpuWorldReadArbiters(cpgpuWorld *world, int64_t cap, int64_t *n_out, int32_t *shapes, int32_t *colour, int32_t *state, int32_t *count, double *normal, double *contact, uint64_t *hash); \
	int cpgpuWorldReadConstraints(cpgpuWorld *world, double *jacc, int32_t *colour); \
	int cpgpuWorldReadPairs(cpgpuWorld *world, int64_t cap, int64_t *n_out, int32_t *pairs); \
	int cpgpuWorldShapesCollide(cpgpuWorld *world, int64_t n, const int32_t *pairs, int32_t *count, double *sets); \
	int cpgpuWorldShapesCollideIn(cpgpuWorld *world, int64_t n, const int32_t *pairs, int32_t *count, double *sets, void *scratch, size_t scratch_bytes); \
	size_t cpgpuShapesCollideScratchBytes(int64_t n); \
	int cpgpuWorldSetTiming(cpgpuWorld *world, int32_t enable); \
	int cpgpuWorldGetTiming(cpgpuWorld *world, int32_t cap, int32_t *n_out, double *ms, int64_t *launches, char *names, int32_t names_cap); \
	int cpgpuWorldGetBuckets(cpgpuWorld *world, int32_t *max_bodies, int32_t *threads, int32_t *n_worlds, int32_t *lds_bodies); \
	int cpgpuWorldSetBuckets(cpgpuWorld *world, int32_t n, const int32_t *max_bodies, const int32_t *threads); \
	int cpgpuWorldAutotune(cpgpuWorld *world, double dt, int32_t trial_steps, int32_t *threads_out); \
	int cpgpuWorldSetGridSchedule(cpgpuWorld *world, int32_t schedule); \
	int cpgpuWorldSetGraph(cpgpuWorld *world, int32_t enable); \
	int cpgpuWorldSaveState(cpgpuWorld *world, void *buf, int64_t cap, int64_t *size_out); \
	int cpgpuWorldLoadState(cpgpuWorld *world, const void *buf, int64_t size); \
	int cpgpuWorldQueryPointNearest(cpgpuWorld *world, int32_t n, const int32_t *space, const double *points, const double *max_distance, const cpgpuShapeFilter *filter, int32_t *shape_out, double *info_out); \
	int cpgpuWorldQuerySegmentFirst(cpgpuWorld *world, int32_t n, const int32_t *space, const double *starts, const double *ends, const double *radius, const cpgpuShapeFilter *filter, int32_t *shape_out, double *info_out); \
	int cpgpuWorldQueryPointAll(cpgpuWorld *world, int32_t n, const int32_t *space, const double *points, const double *max_distance, const cpgpuShapeFilter *filter, int64_t cap, int64_t *n_out, int32_t *hits, double *info); \
	int cpgpuWorldQuerySegmentAll(cpgpuWorld *world, int32_t n, const int32_t *space, const double *starts, const double *ends, const double *radius, const cpgpuShapeFilter *filter, int64_t cap, int64_t *n_out, int32_t *hits, double *info); \
	int cpgpuWorldQueryBBAll(cpgpuWorld *world, int32_t n, const int32_t *space, const double *boxes, const cpgpuShapeFilter *filter, int64_t cap, int64_t *n_out, int32_t *hits); \
	int cpgpuWorldQueryShapes(cpgpuWorld *world, int32_t n, const int32_t *shapes, int64_t cap, int64_t *n_out, int32_t *hits, int32_t *counts, double *sets); \
	int cpgpuWorldTapeBegin(cpgpuWorld *world, int32_t checkpoint_every); \
	int cpgpuWorldTapeEnd(cpgpuWorld *world); \
	int cpgpuWorldBackward(cpgpuWorld *world, const double *g_p, const double *g_a, const double *g_v, const double *g_w, double *d_p0, double *d_a0, double *d_v0, double *d_w0, double *d_m_inv, double *d_i_inv, double *d_u); \
	int cpgpuWorldGetJointGrads(cpgpuWorld *world, double *d_params); \
	int cpgpuWorldGetSolveCounts(cpgpuWorld *world, int64_t *contacts, int64_t *bodies);

namespace cpgpu64 {
struct cpgpuWorld;
CPGPU_DECLARE_API
} // namespace cpgpu64
namespace cpgpu32 {
struct cpgpuWorld;
CPGPU_DECLARE_API
} // namespace cpgpu32

namespace {
// both builds' world structs start with their precision (CPGPU_FP64 / CPGPU_FP32)
inline bool is_fp32(const cpgpuWorld *w) { return *reinterpret_cast<const int32_t *>(w) == CPGPU_FP32; }
inline int null_world()
{
	cpgpu_last_error = "null world";
	return CPGPU_ERR_INVALID;
}

} // namespace

extern "C" {

int cpgpuWorldCreate(const cpgpuWorldDesc *desc, int32_t precision, cpgpuWorld **out)
{
	if (precision == CPGPU_FP32)
		return cpgpu32::cpgpuWorldCreate(desc, precision, reinterpret_cast<cpgpu32::cpgpuWorld **>(out));
	return cpgpu64::cpgpuWorldCreate(desc, precision, reinterpret_cast<cpgpu64::cpgpuWorld **>(out));
}

int cpgpuWorldCreateFrom(const cpgpuWorldDesc *desc, int32_t precision, cpgpuWorld *old, const int32_t *shape_from,
                         cpgpuWorld **out)
{
	if (out) *out = nullptr;
	if (!old) return null_world();
	if (is_fp32(old))
		return cpgpu32::cpgpuWorldCreateFrom(desc, precision, reinterpret_cast<cpgpu32::cpgpuWorld *>(old), shape_from,
		                                     reinterpret_cast<cpgpu32::cpgpuWorld **>(out));
	return cpgpu64::cpgpuWorldCreateFrom(desc, precision, reinterpret_cast<cpgpu64::cpgpuWorld *>(old), shape_from,
	                                     reinterpret_cast<cpgpu64::cpgpuWorld **>(out));
}

void cpgpuWorldDestroy(cpgpuWorld *world)
{
	if (!world) return;
	if (is_fp32(world)) cpgpu32::cpgpuWorldDestroy(reinterpret_cast<cpgpu32::cpgpuWorld *>(world));
	else cpgpu64::cpgpuWorldDestroy(reinterpret_cast<cpgpu64::cpgpuWorld *>(world));
}

int cpgpuWorldStep(cpgpuWorld *world, double dt)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldStep(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), dt);
	return cpgpu64::cpgpuWorldStep(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), dt);
}

int cpgpuWorldStepN(cpgpuWorld *world, double dt, int32_t n_steps)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldStepN(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), dt, n_steps);
	return cpgpu64::cpgpuWorldStepN(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), dt, n_steps);
}

int cpgpuWorldSync(cpgpuWorld *world)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldSync(reinterpret_cast<cpgpu32::cpgpuWorld *>(world));
	return cpgpu64::cpgpuWorldSync(reinterpret_cast<cpgpu64::cpgpuWorld *>(world));
}

int cpgpuWorldGetFaults(cpgpuWorld *world, int32_t *faults, int32_t *first_body)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldGetFaults(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), faults, first_body);
	return cpgpu64::cpgpuWorldGetFaults(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), faults, first_body);
}

int cpgpuWorldReadBodies(cpgpuWorld *world, double *p, double *v, double *a, double *w, double *v_bias, double *w_bias, double *transform)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldReadBodies(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), p, v, a, w, v_bias, w_bias, transform);
	return cpgpu64::cpgpuWorldReadBodies(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), p, v, a, w, v_bias, w_bias, transform);
}

int cpgpuWorldWriteBodies(cpgpuWorld *world, const double *v, const double *w, const double *f, const double *t)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldWriteBodies(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), v, w, f, t);
	return cpgpu64::cpgpuWorldWriteBodies(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), v, w, f, t);
}

int cpgpuWorldWriteBodiesAt(cpgpuWorld *world, int32_t n, const int32_t *index, const double *p, const double *v,
                            const double *f, const double *a, const double *w, const double *t, const double *transform)
{
	if (!world) return null_world();
	if (is_fp32(world))
		return cpgpu32::cpgpuWorldWriteBodiesAt(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), n, index, p, v, f, a, w, t,
		                                        transform);
	return cpgpu64::cpgpuWorldWriteBodiesAt(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), n, index, p, v, f, a, w, t,
	                                        transform);
}

int cpgpuWorldSetSpaceParams(cpgpuWorld *world, int32_t space, const cpgpuSpaceParams *params)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldSetSpaceParams(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), space, params);
	return cpgpu64::cpgpuWorldSetSpaceParams(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), space, params);
}

int cpgpuWorldGetStats(cpgpuWorld *world, cpgpuStats *out)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldGetStats(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), out);
	return cpgpu64::cpgpuWorldGetStats(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), out);
}

int cpgpuWorldReadArbiters(cpgpuWorld *world, int64_t cap, int64_t *n_out, int32_t *shapes, int32_t *colour, int32_t *state, int32_t *count, double *normal, double *contact, uint64_t *hash)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldReadArbiters(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), cap, n_out, shapes, colour, state, count, normal, contact, hash);
	return cpgpu64::cpgpuWorldReadArbiters(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), cap, n_out, shapes, colour, state, count, normal, contact, hash);
}

int cpgpuWorldReadPairs(cpgpuWorld *world, int64_t cap, int64_t *n_out, int32_t *pairs)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldReadPairs(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), cap, n_out, pairs);
	return cpgpu64::cpgpuWorldReadPairs(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), cap, n_out, pairs);
}

int cpgpuWorldShapesCollide(cpgpuWorld *world, int64_t n, const int32_t *pairs, int32_t *count, double *sets)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldShapesCollide(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), n, pairs, count, sets);
	return cpgpu64::cpgpuWorldShapesCollide(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), n, pairs, count, sets);
}

int cpgpuWorldReadConstraints(cpgpuWorld *world, double *jacc, int32_t *colour)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldReadConstraints(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), jacc, colour);
	return cpgpu64::cpgpuWorldReadConstraints(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), jacc, colour);
}

int cpgpuWorldSetTiming(cpgpuWorld *world, int32_t enable)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldSetTiming(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), enable);
	return cpgpu64::cpgpuWorldSetTiming(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), enable);
}

int cpgpuWorldGetTiming(cpgpuWorld *world, int32_t cap, int32_t *n_out, double *ms, int64_t *launches, char *names, int32_t names_cap)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldGetTiming(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), cap, n_out, ms, launches, names, names_cap);
	return cpgpu64::cpgpuWorldGetTiming(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), cap, n_out, ms, launches, names, names_cap);
}

int cpgpuWorldGetBuckets(cpgpuWorld *world, int32_t *max_bodies, int32_t *threads, int32_t *n_worlds, int32_t *lds_bodies)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldGetBuckets(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), max_bodies, threads, n_worlds, lds_bodies);
	return cpgpu64::cpgpuWorldGetBuckets(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), max_bodies, threads, n_worlds, lds_bodies);
}

int cpgpuWorldSetGridSchedule(cpgpuWorld *world, int32_t schedule)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldSetGridSchedule(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), schedule);
	return cpgpu64::cpgpuWorldSetGridSchedule(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), schedule);
}

int cpgpuWorldSetBuckets(cpgpuWorld *world, int32_t n, const int32_t *max_bodies, const int32_t *threads)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldSetBuckets(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), n, max_bodies, threads);
	return cpgpu64::cpgpuWorldSetBuckets(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), n, max_bodies, threads);
}

int cpgpuWorldAutotune(cpgpuWorld *world, double dt, int32_t trial_steps, int32_t *threads_out)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldAutotune(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), dt, trial_steps, threads_out);
	return cpgpu64::cpgpuWorldAutotune(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), dt, trial_steps, threads_out);
}

int cpgpuWorldSetGraph(cpgpuWorld *world, int32_t enable)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldSetGraph(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), enable);
	return cpgpu64::cpgpuWorldSetGraph(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), enable);
}

int cpgpuWorldTapeBegin(cpgpuWorld *world, int32_t checkpoint_every)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldTapeBegin(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), checkpoint_every);
	return cpgpu64::cpgpuWorldTapeBegin(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), checkpoint_every);
}

int cpgpuWorldTapeEnd(cpgpuWorld *world)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldTapeEnd(reinterpret_cast<cpgpu32::cpgpuWorld *>(world));
	return cpgpu64::cpgpuWorldTapeEnd(reinterpret_cast<cpgpu64::cpgpuWorld *>(world));
}

int cpgpuWorldBackward(cpgpuWorld *world, const double *g_p, const double *g_a, const double *g_v, const double *g_w, double *d_p0, double *d_a0, double *d_v0, double *d_w0, double *d_m_inv, double *d_i_inv, double *d_u)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldBackward(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), g_p, g_a, g_v, g_w, d_p0, d_a0, d_v0, d_w0, d_m_inv, d_i_inv, d_u);
	return cpgpu64::cpgpuWorldBackward(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), g_p, g_a, g_v, g_w, d_p0, d_a0, d_v0, d_w0, d_m_inv, d_i_inv, d_u);
}

int cpgpuWorldGetJointGrads(cpgpuWorld *world, double *d_params)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldGetJointGrads(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), d_params);
	return cpgpu64::cpgpuWorldGetJointGrads(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), d_params);
}

int cpgpuWorldGetSolveCounts(cpgpuWorld *world, int64_t *contacts, int64_t *bodies)
{
	if (!world) return null_world();
	if (is_fp32(world)) return cpgpu32::cpgpuWorldGetSolveCounts(reinterpret_cast<cpgpu32::cpgpuWorld *>(world), contacts, bodies);
	return cpgpu64::cpgpuWorldGetSolveCounts(reinterpret_cast<cpgpu64::cpgpuWorld *>(world), contacts, bodies);
}

#define CPGPU_FORWARD(call64, call32) \
	if (!world) return null_world(); \
	if (is_fp32(world)) return call32; \
	return call64

#define W64 reinterpret_cast<cpgpu64::cpgpuWorld *>(world)
#define W32 reinterpret_cast<cpgpu32::cpgpuWorld *>(world)

int cpgpuWorldStepBegin(cpgpuWorld *world, double dt, int64_t *n_callbacks)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldStepBegin(W64, dt, n_callbacks), cpgpu32::cpgpuWorldStepBegin(W32, dt, n_callbacks));
}

int cpgpuWorldReadCallbacks(cpgpuWorld *world, int64_t cap, int64_t *n_out, cpgpuCallbackArbiter *out)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldReadCallbacks(W64, cap, n_out, out),
	              cpgpu32::cpgpuWorldReadCallbacks(W32, cap, n_out, out));
}

int cpgpuWorldWriteDecisions(cpgpuWorld *world, int64_t n, const cpgpuCallbackDecision *decisions)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldWriteDecisions(W64, n, decisions),
	              cpgpu32::cpgpuWorldWriteDecisions(W32, n, decisions));
}

int cpgpuWorldStepEnd(cpgpuWorld *world)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldStepEnd(W64), cpgpu32::cpgpuWorldStepEnd(W32));
}

int cpgpuWorldReadSeparations(cpgpuWorld *world, int64_t cap, int64_t *n_out, int32_t *pairs)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldReadSeparations(W64, cap, n_out, pairs),
	              cpgpu32::cpgpuWorldReadSeparations(W32, cap, n_out, pairs));
}

int cpgpuWorldQueryShapes(cpgpuWorld *world, int32_t n, const int32_t *shapes, int64_t cap, int64_t *n_out,
                          int32_t *hits, int32_t *counts, double *sets)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldQueryShapes(W64, n, shapes, cap, n_out, hits, counts, sets),
	              cpgpu32::cpgpuWorldQueryShapes(W32, n, shapes, cap, n_out, hits, counts, sets));
}

int cpgpuWorldSaveState(cpgpuWorld *world, void *buf, int64_t cap, int64_t *size_out)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldSaveState(W64, buf, cap, size_out),
	              cpgpu32::cpgpuWorldSaveState(W32, buf, cap, size_out));
}

int cpgpuWorldLoadState(cpgpuWorld *world, const void *buf, int64_t size)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldLoadState(W64, buf, size), cpgpu32::cpgpuWorldLoadState(W32, buf, size));
}

int cpgpuWorldQueryPointNearest(cpgpuWorld *world, int32_t n, const int32_t *space, const double *points,
                                const double *max_distance, const cpgpuShapeFilter *filter, int32_t *shape_out,
                                double *info_out)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldQueryPointNearest(W64, n, space, points, max_distance, filter, shape_out, info_out),
	              cpgpu32::cpgpuWorldQueryPointNearest(W32, n, space, points, max_distance, filter, shape_out, info_out));
}

int cpgpuWorldQuerySegmentFirst(cpgpuWorld *world, int32_t n, const int32_t *space, const double *starts,
                                const double *ends, const double *radius, const cpgpuShapeFilter *filter,
                                int32_t *shape_out, double *info_out)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldQuerySegmentFirst(W64, n, space, starts, ends, radius, filter, shape_out, info_out),
	              cpgpu32::cpgpuWorldQuerySegmentFirst(W32, n, space, starts, ends, radius, filter, shape_out, info_out));
}

int cpgpuWorldQueryPointAll(cpgpuWorld *world, int32_t n, const int32_t *space, const double *points,
                            const double *max_distance, const cpgpuShapeFilter *filter, int64_t cap, int64_t *n_out,
                            int32_t *hits, double *info)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldQueryPointAll(W64, n, space, points, max_distance, filter, cap, n_out, hits, info),
	              cpgpu32::cpgpuWorldQueryPointAll(W32, n, space, points, max_distance, filter, cap, n_out, hits, info));
}

int cpgpuWorldQuerySegmentAll(cpgpuWorld *world, int32_t n, const int32_t *space, const double *starts,
                              const double *ends, const double *radius, const cpgpuShapeFilter *filter, int64_t cap,
                              int64_t *n_out, int32_t *hits, double *info)
{
	CPGPU_FORWARD(
	    cpgpu64::cpgpuWorldQuerySegmentAll(W64, n, space, starts, ends, radius, filter, cap, n_out, hits, info),
	    cpgpu32::cpgpuWorldQuerySegmentAll(W32, n, space, starts, ends, radius, filter, cap, n_out, hits, info));
}

int cpgpuWorldQueryBBAll(cpgpuWorld *world, int32_t n, const int32_t *space, const double *boxes,
                         const cpgpuShapeFilter *filter, int64_t cap, int64_t *n_out, int32_t *hits)
{
	CPGPU_FORWARD(cpgpu64::cpgpuWorldQueryBBAll(W64, n, space, boxes, filter, cap, n_out, hits),
	              cpgpu32::cpgpuWorldQueryBBAll(W32, n, space, boxes, filter, cap, n_out, hits));
}

#undef W64
#undef W32
#undef CPGPU_FORWARD

const char *cpgpuGetLastError(void) { return cpgpu_last_error.c_str(); }
const char *cpgpuVersion(void) { return "cpgpu 0.2 (gfx950; Chipmunk2D 7.0.3 cpSpaceStep semantics; FP64 and FP32 builds)"; }

} // extern "C"
