/* cpgpu_on_oracle.c -- TEST INFRASTRUCTURE, never shipped or measured.
 *
 * The entry points of include/cpgpu.h that the cpGPUSpace binding
 * (chipmunk2d_amd/binding/cpGPUSpace.c) calls, answered by the oracle's CPU
 * step (cpstep_oracle.c) instead of the device.  Linked into
 * oracle/_ref/libcpgpuspace_cpu.so with the binding itself, so that the
 * binding's host logic -- packing Chipmunk's structs, uploading setters,
 * packing again after adds and removes, threading arbiters onto bodies --
 * runs on the CPU, under the compiler's address checks, without a GPU.
 *
 * Only what the binding needs: no timing, graphs, tape, queries or handler
 * callbacks.  Each call
 * validates its arguments as the device library does where the binding could
 * get them wrong (indices, counts), and records a message for
 * cpgpuGetLastError.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cpstep_oracle.h"

struct cpgpuWorld {
	oworld *o;
	int precision;
	int32_t nb, ns, nc;
	int32_t faults_reported;   /* the CPGPU_FAULT_* a step has already returned */
};

static char last_error[256];

static int fail(int code, const char *msg)
{
	snprintf(last_error, sizeof(last_error), "%s", msg);
	return code;
}

const char *cpgpuGetLastError(void) { return last_error; }

/* the checks of the device library's validate() that guard indexing */
static int check_desc(const cpgpuWorldDesc *d)
{
	if (!d || d->n_worlds < 1 || !d->space || !d->world_body_start || !d->world_shape_start)
		return fail(CPGPU_ERR_INVALID, "incomplete description");
	if (d->world_body_start[0] != 0 || d->world_body_start[d->n_worlds] != d->n_bodies)
		return fail(CPGPU_ERR_INVALID, "world_body_start must run from 0 to n_bodies");
	if (d->world_shape_start[0] != 0 || d->world_shape_start[d->n_worlds] != d->n_shapes)
		return fail(CPGPU_ERR_INVALID, "world_shape_start must run from 0 to n_shapes");
	for (int s = 0; s < d->n_shapes; s++) {
		if (d->shape_body[s] < 0 || d->shape_body[s] >= d->n_bodies)
			return fail(CPGPU_ERR_INVALID, "shape body index out of range");
		if (d->shape_type[s] == CPGPU_SHAPE_POLY) {
			const int c = d->shape_poly_count[s], st = d->shape_poly_start[s];
			if (c < 1 || c > CPGPU_MAX_POLY_VERTS || st < 0 || st + c > d->n_poly_verts)
				return fail(CPGPU_ERR_INVALID, "polygon vertex range out of range");
		}
	}
	for (int c = 0; c < d->n_constraints; c++) {
		if (d->cons_type[c] < 0 || d->cons_type[c] >= CPGPU_JOINT_TYPES)
			return fail(CPGPU_ERR_INVALID, "constraint type out of range");
		if (d->cons_a[c] < 0 || d->cons_a[c] >= d->n_bodies || d->cons_b[c] < 0 || d->cons_b[c] >= d->n_bodies)
			return fail(CPGPU_ERR_INVALID, "constraint body index out of range");
	}
	return CPGPU_OK;
}

static int wrap(oworld *o, const cpgpuWorldDesc *d, int32_t precision, cpgpuWorld **out)
{
	if (!o) return fail(CPGPU_ERR_INVALID, "the oracle refused the description");
	cpgpuWorld *w = (cpgpuWorld *)calloc(1, sizeof(cpgpuWorld));
	w->o = o;
	w->precision = precision;
	w->nb = d->n_bodies;
	w->ns = d->n_shapes;
	w->nc = d->n_constraints;
	*out = w;
	return CPGPU_OK;
}

int cpgpuWorldCreate(const cpgpuWorldDesc *d, int32_t precision, cpgpuWorld **out)
{
	if (!out) return fail(CPGPU_ERR_INVALID, "null output pointer");
	*out = NULL;
	const int rc = check_desc(d);
	if (rc != CPGPU_OK) return rc;
	return wrap(ow_new(d), d, precision, out);
}

int cpgpuWorldCreateFrom(const cpgpuWorldDesc *d, int32_t precision, cpgpuWorld *old, const int32_t *shape_from,
                         cpgpuWorld **out)
{
	if (!out) return fail(CPGPU_ERR_INVALID, "null output pointer");
	*out = NULL;
	if (!old) return fail(CPGPU_ERR_INVALID, "null previous world");
	const int rc = check_desc(d);
	if (rc != CPGPU_OK) return rc;
	if (d->n_shapes > 0 && !shape_from) return fail(CPGPU_ERR_INVALID, "null shape map");
	char *used = (char *)calloc((size_t)old->ns + 1, 1);
	for (int s = 0; s < d->n_shapes; s++) {
		const int o = shape_from[s];
		if (o < -1 || o >= old->ns || (o >= 0 && used[o])) {
			free(used);
			return fail(CPGPU_ERR_INVALID, "shape map entry out of range or repeated");
		}
		if (o >= 0) used[o] = 1;
	}
	free(used);
	return wrap(ow_new_from(d, old->o, shape_from), d, precision, out);
}

void cpgpuWorldDestroy(cpgpuWorld *w)
{
	if (!w) return;
	ow_free(w->o);
	free(w);
}

int cpgpuWorldStep(cpgpuWorld *w, double dt)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	ow_step(w->o, dt);
	/* the reference's debug-build checks, returned once as the device library does */
	int32_t first;
	const int32_t f = ow_faults(w->o, &first) & ~w->faults_reported;
	if (!f) return CPGPU_OK;
	w->faults_reported |= f;
	char msg[200];
	snprintf(msg, sizeof(msg), "the step failed the reference's checks:%s%s%s first body: %d;",
	         f & CPGPU_FAULT_UNSOLVABLE ? " Unsolvable collision or constraint (an effective mass of zero);" : "",
	         f & CPGPU_FAULT_BODY_MASS ? " Body's mass and moment must be positive to simulate;" : "",
	         f & CPGPU_FAULT_BODY_STATE ? " Body's state is invalid;" : "", first);
	return fail(CPGPU_ERR_FAULT, msg);
}

int cpgpuWorldSync(cpgpuWorld *w) { return w ? CPGPU_OK : fail(CPGPU_ERR_INVALID, "null world"); }

int cpgpuWorldGetFaults(cpgpuWorld *w, int32_t *faults, int32_t *first_body)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	const int32_t f = ow_faults(w->o, first_body);
	if (faults) *faults = f;
	return CPGPU_OK;
}

int cpgpuWorldReadBodies(cpgpuWorld *w, double *p, double *v, double *a, double *wv, double *v_bias, double *w_bias,
                         double *transform)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	return ow_read_bodies(w->o, p, v, a, wv, v_bias, w_bias, transform);
}

int cpgpuWorldWriteBodiesAt(cpgpuWorld *w, int32_t n, const int32_t *index, const double *p, const double *v,
                            const double *f, const double *a, const double *wv, const double *t,
                            const double *transform)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (n < 0 || (n > 0 && !index)) return fail(CPGPU_ERR_INVALID, "bad body index list");
	for (int k = 0; k < n; k++)
		if (index[k] < 0 || index[k] >= w->nb) return fail(CPGPU_ERR_INVALID, "body index out of range");
	return ow_write_bodies_at(w->o, n, index, p, v, f, a, wv, t, transform) ? fail(CPGPU_ERR_INVALID, "oracle refused")
	                                                                       : CPGPU_OK;
}

int cpgpuWorldSetSpaceParams(cpgpuWorld *w, int32_t space, const cpgpuSpaceParams *params)
{
	if (!w || !params) return fail(CPGPU_ERR_INVALID, "null world or parameters");
	return ow_set_space_params(w->o, space, params) ? fail(CPGPU_ERR_INVALID, "space index out of range") : CPGPU_OK;
}

int cpgpuWorldReadArbiters(cpgpuWorld *w, int64_t cap, int64_t *n_out, int32_t *shapes, int32_t *colour,
                           int32_t *state, int32_t *count, double *normal, double *contact, uint64_t *hash)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (cap < 0) return fail(CPGPU_ERR_INVALID, "negative capacity");
	const int64_t n = ow_read_arbiters(w->o, cap, shapes, colour, state, count, normal, contact, hash);
	if (n_out) *n_out = n;
	return CPGPU_OK;
}

int cpgpuWorldReadConstraints(cpgpuWorld *w, double *jacc, int32_t *colour)
{
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	return ow_read_constraints(w->o, jacc, colour);
}

int cpgpuWorldAutotune(cpgpuWorld *w, double dt, int32_t trial_steps, int32_t *threads_out)
{
	(void)dt;
	(void)trial_steps;
	if (!w) return fail(CPGPU_ERR_INVALID, "null world");
	if (threads_out)
		for (int q = 0; q < CPGPU_MAX_BUCKETS; q++) threads_out[q] = 0;
	return CPGPU_OK;
}

/* Collision handler callbacks need the step in two halves, which the oracle's
 * step does not stop between: the binding's handler flows run over the device
 * library (and its CPU emulation) only. */
int cpgpuWorldStepBegin(cpgpuWorld *w, double dt, int64_t *n_callbacks)
{
	(void)w; (void)dt;
	if (n_callbacks) *n_callbacks = 0;
	return fail(CPGPU_ERR_INVALID, "collision handler callbacks need the device library");
}
int cpgpuWorldReadCallbacks(cpgpuWorld *w, int64_t cap, int64_t *n_out, cpgpuCallbackArbiter *out)
{
	(void)w; (void)cap; (void)out;
	if (n_out) *n_out = 0;
	return fail(CPGPU_ERR_INVALID, "collision handler callbacks need the device library");
}
int cpgpuWorldWriteDecisions(cpgpuWorld *w, int64_t n, const cpgpuCallbackDecision *d)
{
	(void)w; (void)n; (void)d;
	return fail(CPGPU_ERR_INVALID, "collision handler callbacks need the device library");
}
int cpgpuWorldStepEnd(cpgpuWorld *w)
{
	(void)w;
	return fail(CPGPU_ERR_INVALID, "collision handler callbacks need the device library");
}
int cpgpuWorldReadSeparations(cpgpuWorld *w, int64_t cap, int64_t *n_out, int32_t *pairs)
{
	(void)w; (void)cap; (void)pairs;
	if (n_out) *n_out = 0;
	return fail(CPGPU_ERR_INVALID, "collision handler callbacks need the device library");
}
