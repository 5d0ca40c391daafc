// cpgpu_adjoint.h -- rollout tape and reverse-step launchers (adjoint.hip).
#pragma once

#include "cpgpu_step.h"

namespace CPGPU_NS {

// One body's state at the start of a step (checkpoints, segment replay).
struct BodyState {
	real2 p;
	real a;
	BodyVel vel;
	real2 f;
	real t;
	XF T;
};

// Adjoint of one body's step state.
struct BodyAdj {
	double p[2], a, v[2], w, vb[2], wb;
};

// Everything the reverse of one replayed step reads and writes.
struct AdjStep {
	int n;                      // solved arbiters of the step
	const Arbiter *act;         // after the replayed solve
	const int32_t *order;       // colour order (per space: its own segment)
	const int32_t *colour;      // colour of each arbiter
	const int32_t *warb_off;    // first position of each space, nullptr for a single space
	const ArbRec *rec;          // [iteration][position] pre-states
	const BodyState *start;     // bodies at the start of the step
	const Arbiter *tape;        // every recorded arbiter of the rollout
	void *garb;                 // [n] per-arbiter adjoint scratch (adjoint_arb_scratch_bytes each)
	BodyAdj *gb;                // [n_bodies] adjoint state, updated in place
	double *gminv, *giinv;      // [n_bodies] adjoints of 1/m and 1/I
	double *gu;                 // [n_shapes] adjoint of the friction coefficient
	double2 *gJ;                // [recorded contacts] adjoints of the accumulated impulses (jn, jt)
	// joints (nc 0: none)
	int nc;
	const Joint *joints;        // the joints after the replayed step (pre-step results, parameters)
	const JointRec *jrec;       // [slot][position in jw_order] pre-states, slot 0 before the cached impulses
	const int32_t *jw_order, *jw_col, *jw_cbase, *jw_ncol;   // per space: joints by colour
	void *gjoint;               // [nc] per-joint adjoint scratch (adjoint_joint_scratch_bytes each)
	double2 *gJacc;             // [nc] adjoint of each joint's accumulators at the end of the step on entry,
	                            //      at the end of the step before on exit
	double *gP;                 // [nc * 8] adjoints of the joints' parameters (cpgpuWorldDesc::cons_params)
};

hipError_t adjoint_tape_record(Arbiter *arbs, const int32_t *n_dev, int n_host, Arbiter *tape, int off,
                               hipStream_t stream);
hipError_t adjoint_replay_prepare(const Arbiter *tape, int off, int n, Arbiter *act, hipStream_t stream);
hipError_t adjoint_state_save(const DevWorld &W, BodyState *out, hipStream_t stream);
hipError_t adjoint_state_load(const DevWorld &W, const BodyState *in, hipStream_t stream);
// the reverse of one replayed step: the impulses, then the pre-steps and contact geometry
hipError_t adjoint_reverse_solve(const DevWorld &W, const AdjStep &A, int n_worlds, hipStream_t stream);
hipError_t adjoint_reverse_geometry(const DevWorld &W, const AdjStep &A, int n_worlds, hipStream_t stream);
// bytes of the per-arbiter / per-joint adjoint scratch (AdjStep::garb, gjoint) per element
size_t adjoint_arb_scratch_bytes();
size_t adjoint_joint_scratch_bytes();

} // namespace CPGPU_NS
