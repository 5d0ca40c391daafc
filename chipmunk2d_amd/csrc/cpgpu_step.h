// cpgpu_step.h -- per-step buffers shared by the narrowphase and solver launchers.
#pragma once

#include "cpgpu_internal.h"

// Buffers of the persistence table and arbiter arrays for one step.  "prev"
// arrays are last step's output; the others are written this step.
struct StepBuffers {
	uint64_t *tkeys;           // open-addressing table over prev_recs
	int32_t *tvals;
	int tmask;
	const PairRec *prev_recs;
	const int32_t *n_prev_recs;
	uint8_t *seen;             // prev record looked up this step
	const Arbiter *prev_act;
	const Arbiter *prev_cached;
	PairRec *recs;
	int32_t *n_recs;
	int cap_recs;
	Arbiter *act;              // arbiters touched this step: the solver's list
	int32_t *n_act;
	int cap_act;
	Arbiter *cached;           // untouched arbiters within persistence
	int32_t *n_cached;
	int cap_cached;
};

hipError_t narrowphase_step(const DevWorld &W, const StepBuffers &B, const int2 *pairs, const int32_t *n_pairs_dev,
                            int n_pairs_host, int n_prev_recs_host, hipStream_t stream, KernelTimer *timer);

#define CPGPU_MAX_COLOURS 512
#define COLOUR_SUBROUNDS 3

struct GridBarrier {
	unsigned int count;
	unsigned int gen;
};

struct SolveBuffers {
	Arbiter *act;
	const int32_t *n_act;
	int32_t *colour;           // [cap] colour of each active arbiter
	int32_t *order;            // [cap] active arbiters sorted by colour
	int32_t *col_off;          // [CPGPU_MAX_COLOURS + 1] counts, then offsets
	int32_t *col_cursor;       // [CPGPU_MAX_COLOURS]
	int32_t *n_colours;        // [1]
	int32_t *left;             // [1] uncoloured arbiters
	unsigned long long *claim; // [2][n_bodies]
	int32_t *taken;            // [n_bodies]
	// joints, coloured once at creation
	const int32_t *jorder;     // joints in colour order
	const int32_t *jcol_off;   // [n_jcolours + 1]
	int n_jcolours;
	GridBarrier *bar;
	int max_iterations;
	int n_bodies;
	const int32_t *dyn_bodies; // non-static bodies
	int n_dyn_bodies;
};

// Launch the persistent solver: colouring, colour sort, pre-steps,
// velocity integration, cached impulses and the impulse iterations.
hipError_t solver_step(const DevWorld &W, const SolveBuffers &S, int grid, hipStream_t stream, KernelTimer *timer);
// Largest co-resident grid of the solver kernel on the current device.
int solver_max_grid();
