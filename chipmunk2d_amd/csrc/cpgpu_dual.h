// cpgpu_dual.h -- forward-mode dual numbers for contact-geometry Jacobians.
//
// A Dual carries a value and its partial derivatives with respect to
// DUAL_N inputs (the adjoint seeds the positions and angles of the two
// bodies of a shape pair: pA.x, pA.y, aA, pB.x, pB.y, aB).  The value part
// of every operation is exactly the IEEE double operation the step performs,
// so a narrowphase run on Duals takes the same branches as the step's own
// run and its value parts are bit-identical; the derivative parts are the
// chain rule through that fixed path (the contact features frozen).
#pragma once

#include "cpgpu_device.h"

namespace CPGPU_NS {

#define DUAL_N 6

struct Dual {
	double v;
	double d[DUAL_N];
	CPD Dual() {}
	CPD Dual(double x) : v(x)
	{
		for (int k = 0; k < DUAL_N; k++) d[k] = 0.0;
	}
};

CPD Dual dual_seed(double x, int k)
{
	Dual r(x);
	r.d[k] = 1.0;
	return r;
}

CPD Dual operator-(const Dual &a)
{
	Dual r;
	r.v = -a.v;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = -a.d[k];
	return r;
}
CPD Dual operator+(const Dual &a, const Dual &b)
{
	Dual r;
	r.v = a.v + b.v;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = a.d[k] + b.d[k];
	return r;
}
CPD Dual operator-(const Dual &a, const Dual &b)
{
	Dual r;
	r.v = a.v - b.v;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = a.d[k] - b.d[k];
	return r;
}
CPD Dual operator*(const Dual &a, const Dual &b)
{
	Dual r;
	r.v = a.v * b.v;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
	return r;
}
CPD Dual operator/(const Dual &a, const Dual &b)
{
	Dual r;
	r.v = a.v / b.v;
	const double inv = 1.0 / b.v;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
	return r;
}
CPD Dual operator+(const Dual &a, double b) { Dual r = a; r.v = a.v + b; return r; }
CPD Dual operator+(double a, const Dual &b) { Dual r = b; r.v = a + b.v; return r; }
CPD Dual operator-(const Dual &a, double b) { Dual r = a; r.v = a.v - b; return r; }
CPD Dual operator-(double a, const Dual &b)
{
	Dual r;
	r.v = a - b.v;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = -b.d[k];
	return r;
}
CPD Dual operator*(const Dual &a, double b)
{
	Dual r;
	r.v = a.v * b;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = a.d[k] * b;
	return r;
}
CPD Dual operator*(double a, const Dual &b)
{
	Dual r;
	r.v = a * b.v;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = a * b.d[k];
	return r;
}
CPD Dual operator/(const Dual &a, double b)
{
	Dual r;
	r.v = a.v / b;
	const double inv = 1.0 / b;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = a.d[k] * inv;
	return r;
}
CPD Dual operator/(double a, const Dual &b)
{
	Dual r;
	r.v = a / b.v;
	const double inv = 1.0 / b.v;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = -r.v * b.d[k] * inv;
	return r;
}

CPD bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
CPD bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
CPD bool operator<=(const Dual &a, const Dual &b) { return a.v <= b.v; }
CPD bool operator>=(const Dual &a, const Dual &b) { return a.v >= b.v; }
CPD bool operator==(const Dual &a, const Dual &b) { return a.v == b.v; }
CPD bool operator!=(const Dual &a, const Dual &b) { return a.v != b.v; }
CPD bool operator<(const Dual &a, double b) { return a.v < b; }
CPD bool operator>(const Dual &a, double b) { return a.v > b; }
CPD bool operator<=(const Dual &a, double b) { return a.v <= b; }
CPD bool operator>=(const Dual &a, double b) { return a.v >= b; }
CPD bool operator==(const Dual &a, double b) { return a.v == b; }
CPD bool operator!=(const Dual &a, double b) { return a.v != b; }
CPD bool operator<(double a, const Dual &b) { return a < b.v; }
CPD bool operator>(double a, const Dual &b) { return a > b.v; }
CPD bool operator<=(double a, const Dual &b) { return a <= b.v; }
CPD bool operator>=(double a, const Dual &b) { return a >= b.v; }

CPD Dual csqrt(const Dual &a)
{
	Dual r;
	r.v = __builtin_sqrt(a.v);
	const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
	for (int k = 0; k < DUAL_N; k++) r.d[k] = a.d[k] * h;
	return r;
}

CPD double dval(double x) { return x; }
CPD double dval(const Dual &x) { return x.v; }

} // namespace CPGPU_NS
