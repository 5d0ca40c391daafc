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

// Worlds are solved in one of two ways, chosen per world by size:
//   small worlds: one workgroup each (k_solve_worlds), every synchronisation
//     a workgroup barrier;
//   large worlds: together on a co-resident grid (k_solve_grid) with grid
//     barriers between colours.
struct SolveBuffers {
	Arbiter *act;
	const int32_t *n_act;
	int32_t *colour;           // [cap] colour of each active arbiter
	int32_t *order;            // [cap] active arbiters in colour order (per small world: its own segment)
	int32_t *arb_by_world;     // [cap] active arbiters sorted by (large, world); nullptr for a single space
	int32_t *warb_off;         // [n_worlds + 1] first sorted position of each small world
	int32_t *large_first;      // [1] first sorted position of a large world's arbiter
	const uint8_t *world_small;
	const int32_t *small_worlds;   // the small worlds (one workgroup each)
	int32_t *col_off;          // [CPGPU_MAX_COLOURS + 1] grid solve: counts, then offsets
	int32_t *col_cursor;       // [CPGPU_MAX_COLOURS]
	int32_t *n_colours;        // [1] grid solve colour count
	int32_t *left;             // [1] uncoloured arbiters (grid solve)
	unsigned long long *claim; // [2][n_bodies]
	int32_t *taken;            // [n_bodies]
	// joints, coloured once at creation: large worlds together ...
	const int32_t *jorder;     // joints of large worlds in colour order
	const int32_t *jcol_off;   // [n_jcolours + 1]
	int n_jcolours;
	// ... and every small world on its own
	const int32_t *jw_order;   // joints sorted by (world, colour)
	const int32_t *jw_col;     // per world: n colours + 1 offsets into jw_order
	const int32_t *jw_cbase;   // [n_worlds] first entry of the world's offsets in jw_col
	const int32_t *jw_ncol;    // [n_worlds]
	GridBarrier *bar;
	// grid solve: colours (and the last uncoloured arbiters) with few arbiters
	// are handled by workgroup 0 alone instead of paying a grid barrier each
	int32_t *tail_list;        // [colour_tail_max]
	int32_t *tail_n;           // [1]
	int32_t *tail_state;       // [2] colour and sub-round after the tail colouring
	int colour_tail_max;
	int solve_tail_max;
	int max_iterations;        // over the large worlds
	int n_bodies;
	const int32_t *large_bodies;   // non-static bodies of large worlds
	int n_large_bodies;
};

// Sort the solved arbiters by (large, world) and find the small worlds'
// ranges (skipped for a single space).
hipError_t solver_sort_by_world(const SolveBuffers &S, int n_act, int n_worlds, RadixSortBuffers &sortbuf,
                                uint64_t *keys, hipStream_t stream, KernelTimer *timer);
// Launch the solvers: colouring, colour sort, pre-steps, velocity
// integration, cached impulses and the impulse iterations.  n_small small
// worlds get a workgroup each; grid > 0 launches the grid solve of the large
// worlds with that many workgroups.
hipError_t solver_step(const DevWorld &W, const SolveBuffers &S, int n_small, int grid, hipStream_t stream,
                       KernelTimer *timer);
// Largest co-resident grid of the grid solve on the current device.
int solver_max_grid();
