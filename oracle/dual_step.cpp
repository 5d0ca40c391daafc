/*
 * dual_step.cpp -- the oracle's step (cpstep_oracle.c) on dual numbers: the
 * forward-mode derivative of a rollout along one direction of its inputs.
 *
 * TEST INFRASTRUCTURE ONLY.  The second check of the library's adjoint
 * (cpgpuWorldBackward): the same restatement of cpSpaceStep, compiled with a
 * Dual (dual.h) in place of double.  Its value parts equal the double
 * oracle's bit for bit (tests check it); its tangents give the directional
 * derivative J e of the final state with respect to a seeded direction e of
 * the initial velocities, masses and friction coefficients, to compare with
 * the adjoint's e . grad.  No finite differences: the two agree to rounding.
 *
 * The transcendentals differentiate analytically around the value the double
 * oracle computes: d sin = cos dx, d cos = -sin dx, d exp = exp dx.
 */
#include "dual.h"

#define ORACLE_NUM Dual
#define NUM_VAL(x) ((double)(x))

#include <math.h>

/* the deterministic sin/cos and exp of the double oracle, differentiated */
void cp_sincos_oracle(double x, double *s_out, double *c_out);
double cp_exp_oracle(double x);
static inline void cp_sincos_oracle(Dual x, Dual *s_out, Dual *c_out)
{
	double s, c;
	cp_sincos_oracle(x.v, &s, &c);
	*s_out = Dual(s, c * x.d);
	*c_out = Dual(c, -s * x.d);
}
static inline Dual cp_exp_oracle(Dual x)
{
	const double e = cp_exp_oracle(x.v);
	return Dual(e, e * x.d);
}

#include "cpstep_oracle.c"

/* ---- the C interface of the dual build ---- */

extern "C" {

oworld *dual_new(const cpgpuWorldDesc *desc) { return ow_new(desc); }
void dual_free(oworld *w) { ow_free(w); }
void dual_step(oworld *w, double dt) { ow_step(w, dt); }
void dual_set_libm_trig(oworld *w, int enable) { ow_set_libm_trig(w, enable); }

/* Seed the tangent direction of the inputs (call before the first step):
 * initial velocities dv [n_bodies][2] and angular velocities dw [n_bodies],
 * masses dm [n_bodies] (a dynamic body's moment scales with its mass, 1/m
 * and 1/i follow), friction coefficients du [n_shapes].  NULL: zero. */
void dual_seed(oworld *w, const double *dv, const double *dw, const double *dm, const double *du)
{
	for (int i = 0; i < w->nb; i++) {
		OBody *b = &w->B[i];
		if (dv) { b->v.x.d = dv[2 * i]; b->v.y.d = dv[2 * i + 1]; }
		if (dw) b->w.d = dw[i];
		if (dm && dm[i] != 0.0 && b->type == CPGPU_BODY_DYNAMIC) {
			/* m -> m + t dm with i = k m: d(1/m) = -dm/m^2, d(1/i) = -k dm / i^2 */
			const double m = b->m.v, k = b->i.v / b->m.v;
			b->m.d = dm[i];
			b->m_inv.d = -dm[i] / (m * m);
			b->i.d = k * dm[i];
			b->i_inv.d = -k * dm[i] / (b->i.v * b->i.v);
		}
	}
	for (int s = 0; du && s < w->ns; s++) w->S[s].u.d = du[s];
}

/* Value and tangent of the body state: p, v [n][2], a, w [n]. */
void dual_read_bodies(oworld *w, double *p, double *dp, double *v, double *dv, double *a, double *da, double *wv,
                      double *dwv)
{
	for (int i = 0; i < w->nb; i++) {
		const OBody *b = &w->B[i];
		if (p) { p[2 * i] = b->p.x.v; p[2 * i + 1] = b->p.y.v; }
		if (dp) { dp[2 * i] = b->p.x.d; dp[2 * i + 1] = b->p.y.d; }
		if (v) { v[2 * i] = b->v.x.v; v[2 * i + 1] = b->v.y.v; }
		if (dv) { dv[2 * i] = b->v.x.d; dv[2 * i + 1] = b->v.y.d; }
		if (a) a[i] = b->a.v;
		if (da) da[i] = b->a.d;
		if (wv) wv[i] = b->w.v;
		if (dwv) dwv[i] = b->w.d;
	}
}

}  /* extern "C" */
