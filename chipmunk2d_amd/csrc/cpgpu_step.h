// cpgpu_step.h -- per-step buffers shared by the narrowphase and solver launchers.
#pragma once

#include "cpgpu_internal.h"

namespace CPGPU_NS {

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
	// candidate pairs grouped by their shapes' type pair (the narrowphase's input order)
	int2 *pairs_by_type;       // [pair capacity]
	int32_t *type_counts;      // [2 * PAIR_CLASSES]: counts, then cursors
	// collision handler callbacks: the solved-list slots of arbiters whose begin or preSolve
	// is due (nullptr in a world without handlers), and the pairs that separated
	int32_t *cb;               // [cap_act]
	int32_t *n_cb;
	int32_t *seps;             // [cap_seps][3] {space, shape a, shape b}
	int32_t *n_seps;
	int cap_seps;
};

#define PAIR_CLASSES 9         // type_a * 3 + type_b (type_a <= type_b)

// A pair record's arbiter reference: an index into the previous step's solved
// arbiters, or, with this bit set, into its cached ones.
constexpr int32_t CPGPU_CACHED_REF = 0x40000000;

// The previous step's pair records in an open-addressing table by pair key
// (B.tkeys / tvals / tmask, B.prev_recs, B.n_prev_recs, B.seen); they depend
// only on the previous step, so this can run beside the broadphase.
// n_prev_recs_cap bounds the records from above (the kernels read the count).
hipError_t narrowphase_table(const DevWorld &W, const StepBuffers &B, int n_prev_recs_cap, hipStream_t stream,
                             KernelTimer *timer);
// The narrowphase over the candidate pairs (after narrowphase_table):
// n_pairs_host bounds the pairs from above (the launch size); the kernels
// read the count from n_pairs_dev.
hipError_t narrowphase_step(const DevWorld &W, const StepBuffers &B, const int2 *pairs, const int32_t *n_pairs_dev,
                            int n_pairs_host, int n_prev_recs_host, hipStream_t stream, KernelTimer *timer);

// cpShapesCollide (cpShape.c:260-283) for n shape pairs (device arrays) on the
// shapes' cached geometry: count[n] and sets[n][12] {normal, then per point
// pointA, pointB, distance}, written as doubles for either build.
// collision handler callbacks over the step's due arbiters (B.cb, B.n_cb; launches sized
// for cap, the kernels read the count): default decisions, the host's view, the decisions applied
hipError_t callbacks_init(const StepBuffers &B, int cap, cpgpuCallbackDecision *dec, hipStream_t stream);
hipError_t callbacks_pack(const StepBuffers &B, int cap, cpgpuCallbackArbiter *out, hipStream_t stream);
hipError_t callbacks_apply(const DevWorld &W, const StepBuffers &B, int cap, const cpgpuCallbackDecision *dec,
                           hipStream_t stream);
hipError_t shapes_collide(const DevWorld &W, const int2 *pairs, int n, int32_t *count, double *sets,
                          hipStream_t stream);

#define CPGPU_MAX_COLOURS 512
#define COLOUR_SUBROUNDS 3

struct GridBarrier {
	unsigned int count;
	unsigned int gen;
};

// Worlds are solved in one of two ways, chosen per world by size:
//   small worlds: one workgroup each (k_solve_worlds), every synchronisation
//     a workgroup barrier;
//   large worlds: every arbiter and joint in the grid set, solved on a
//     co-resident grid (k_solve_grid).
// The grid set of each large world (its dynamic bodies) is cut into `parts` strips along x
// (grid_parts_for: about GRID_STRIP_BODIES bodies a strip, between GRID_PARTS_MIN and
// GRID_PARTS_MAX; one count for every large world of the description, from the largest)
// with equal numbers of its dynamic bodies (quantiles of a PART_BINS-bin
// histogram of their x between the set's least and greatest x).  An arbiter
// whose dynamic bodies lie in one strip belongs to that strip; one joining
// strips q and q + 1 to seam q; one joining strips further apart to the wide
// set.  Strips share no dynamic body, and neither do even seams (0, 2, ...)
// nor odd ones, so the solve units of the grid set are swept in four
// phases: all strips, all even seams, all odd seams, the wide set.  Each
// unit is coloured on its own and swept by one workgroup with workgroup
// barriers; grid barriers separate the phases.  The solve order is
// therefore: the strips, the even seams, the odd seams (each in its colour
// order), then the wide set in its colour order.
// Sorted arbiter keys, dense so the radix sort has few digits: small world
// w -> w; unit u of the grid set (grid_units' order) -> n_worlds + u.
#define GRID_PARTS_MIN 64
#define GRID_PARTS_MAX 256
#define GRID_STRIP_BODIES 384
// strips of at most this many bodies keep them in LDS during their sweeps
#define GRID_RES_BODIES 512
#define PART_BINS 1024

// strips of a description whose largest grid set has n dynamic bodies
__host__ __device__ inline int grid_parts_for(long n)
{
	const long p = (n + GRID_STRIP_BODIES - 1) / GRID_STRIP_BODIES;
	return (int)(p < GRID_PARTS_MIN ? GRID_PARTS_MIN : (p > GRID_PARTS_MAX ? GRID_PARTS_MAX : p));
}
// seams between `parts` strips, the even ones (0, 2, ...) and the odd ones
__host__ __device__ inline int even_seams(int parts) { return parts / 2; }
__host__ __device__ inline int odd_seams(int parts) { return (parts - 1) / 2; }
// Claim arrays for concurrent colourings: strips and small worlds, even
// seams, odd seams, the wide set.
#define CLAIM_SETS 4

// Pre-state of one arbiter's impulse application in one sweep (rollout
// tape): both bodies' velocities and bias velocities, and each contact's
// accumulated impulses before the application.
struct ArbRec {
	real a[6], b[6];         // v.x, v.y, w, v_bias.x, v_bias.y, w_bias
	real acc[2][3];          // jnAcc, jtAcc, jBias
};

// Pre-state of one joint's impulse application (rollout tape): both bodies'
// velocities and the joint's accumulators before it -- pivot (jAx, jAy),
// damped spring (target_vrn, jAcc), slide (jnAcc, 0), gear (jAcc, 0).
struct JointRec {
	real a[3], b[3];         // v.x, v.y, w
	real acc[2];
};

struct SolveBuffers {
	Arbiter *act;
	const int32_t *n_act;
	int32_t *colour;           // [cap] colour of each active arbiter
	// colouring scratch, indexed by position in arb_by_world (solve-set position + order_base):
	int2 *cbody;               // [cap] the arbiter's dynamic bodies, -1 for a static or kinematic one
	unsigned long long *cprio; // [cap] its colouring priority
	int32_t *ccol;             // [cap] its colour while colouring
	int32_t *order;            // [cap] active arbiters in colour order (per workgroup solve: its own segment)
	int32_t *arb_by_world;     // [cap] active arbiters sorted by key; nullptr for a single small space
	uint64_t *keys;            // [cap] the sorted keys
	int32_t *warb_off;         // [n_worlds + 1] first sorted position of each small world (the last: the grid set)
	int32_t *large_first;      // [3] first sorted position of the grid set, of the wide set; end of the solved
	int n_worlds;
	int n_large_joints;        // joints of large worlds (jorder[0..n))
	const uint8_t *world_small;
	unsigned long long *claim; // [CLAIM_SETS][2][n_bodies]
	int32_t *taken;            // [CLAIM_SETS][n_bodies]
	// joints, coloured once at creation: large worlds together ...
	const int32_t *jorder;     // joints of large worlds in colour order
	const int32_t *jcol_off;   // [n_jcolours + 1]
	int n_jcolours;
	// ... and every small world on its own
	const int32_t *jw_order;   // joints sorted by (world, colour)
	const int2 *jw_body;       // their bodies in that order (~b for a body that is not dynamic)
	const int32_t *jw_col;     // per world: n colours + 1 offsets into jw_order
	const int32_t *jw_cbase;   // [n_worlds] first entry of the world's offsets in jw_col
	const int32_t *jw_ncol;    // [n_worlds]
	GridBarrier *bar;
	int max_iterations;        // over the large worlds
	int n_bodies;
	const int32_t *large_bodies;   // non-static bodies of large worlds
	int n_large_bodies;
	// solve units of the grid set
	const int32_t *large_index;    // [n_worlds] ordinal among the large worlds, -1 for a small world
	int n_large_worlds;
	int parts;                     // strips of each large world's grid set (grid_parts_for)
	unsigned long long *pbound;    // [2 per large world] least and greatest x of the grid set (ordered keys)
	int32_t *phist;                // [PART_BINS per large world] bin counts, then each bin's strip
	int32_t *bpart;                // [n_bodies] strip of each grid-set body
	int32_t *bslot;                // [n_bodies] its slot among its strip's bodies
	int32_t *pcount;               // [parts per large world] bodies of each strip
	int32_t *sbody;                // [parts per large world][GRID_RES_BODIES] each strip's bodies by slot
	int32_t *pcol;                 // [grid_units][CPGPU_MAX_COLOURS + 1] colour offsets of each unit
	int2 *pinfo;                   // [grid_units] first sorted position and colour count of each unit
	int32_t *udone;                // [grid_units] sweeps finished by each unit (dependency-driven schedule)
	int grid_flow;                 // allow the dependency-driven schedule (else grid barriers between phases)
	// recording (adjoint replay): rec[it * rec_stride + position in order], or nullptr
	ArbRec *rec;
	int rec_stride;
	// joints: jrec[slot * n_joints + position in jw_order], slot 0 before the cached impulses,
	// slot it + 1 before sweep it; or nullptr
	JointRec *jrec;
};

// The strips of the large worlds' grid sets (each dynamic body's strip and
// slot).  They depend only on the step's body positions, so they can run
// beside the broadphase and the narrowphase.
hipError_t solver_strips(const DevWorld &W, const SolveBuffers &S, hipStream_t stream, KernelTimer *timer);
// Group the solved arbiters (after solver_strips): the large worlds' arbiter
// pre-steps (and, without joints, their bodies' velocity integration), then
// the arbiters sorted by key with the small worlds' and the grid units'
// ranges.  Skipped for a single small space.  n_act bounds the solved
// arbiters from above (launch sizes); the kernels read the count *S.n_act.
hipError_t solver_group(const DevWorld &W, const SolveBuffers &S, int n_act, int n_worlds, bool large,
                        RadixSortBuffers &sortbuf, hipStream_t stream, KernelTimer *timer);
// solve units of the grid set: the strips, even seams and odd seams of every
// large world, then the wide set
__host__ __device__ inline int grid_units(int n_large_worlds, int parts)
{
	return n_large_worlds * (2 * parts - 1) + 1;
}

// Small worlds are launched in size buckets, each with its own workgroup
// size and room for the largest of its worlds' bodies in LDS (0: the bodies
// stay in global memory).  Worlds are listed largest first.
static_assert(CPGPU_MAX_BUCKETS == 4, "bucket count of the C ABI");
// solve units counted in DevWorld::totals: the buckets, the grid
static_assert(CPGPU_SOLVE_UNITS == CPGPU_MAX_BUCKETS + 1, "solve units: buckets, grid");
struct WorldBucket {
	int max_bodies;            // worlds with at most this many bodies (static included)
	int threads;               // 64, 128, 256, 512 or 1024
	int lds_cap;
	const int32_t *worlds;     // device list
	int n;
};
// Launch the solvers: colouring, colour sort, pre-steps, velocity
// integration, cached impulses and the impulse iterations.  Each bucket of
// small worlds runs on its own auxiliary stream (all independent worlds);
// grid > 0 launches the grid solve with that many workgroups on the main
// stream; the main stream waits for all of them.  events: CPGPU_MAX_BUCKETS + 2.
hipError_t solver_step(const DevWorld &W, const SolveBuffers &S, const WorldBucket *buckets, int n_buckets,
                       int grid, hipStream_t stream, const hipStream_t *aux,
                       const hipEvent_t *events, KernelTimer *timer);
// Largest dynamic LDS a small-world workgroup may take next to its static LDS.
int solver_max_lds_bytes();
// Largest co-resident grid of the grid solve on the current device.
int solver_max_grid();

} // namespace CPGPU_NS
