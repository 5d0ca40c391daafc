/*
 * dual.h -- first-order dual numbers for the forward-mode derivative of the
 * oracle's step (TEST INFRASTRUCTURE ONLY: see dual_step.cpp).
 *
 * A Dual holds a value and a tangent.  Every operation computes the value
 * part with exactly the IEEE double operation the plain oracle performs
 * (one + - * / sqrt or fma each), so a step on Duals reproduces the double
 * step bit for bit in its value parts; the tangent parts carry the
 * derivative along one direction, by the rules of differentiation
 * (chain rule; the derivative of a branch taken is the branch's).
 * Comparisons and conversions look at the value part only.
 */
#ifndef ORACLE_DUAL_H
#define ORACLE_DUAL_H

#include <math.h>
#include <stdint.h>

struct Dual {
	double v, d;
	Dual() = default;
	constexpr Dual(double x) : v(x), d(0.0) {}
	constexpr Dual(double x, double dx) : v(x), d(dx) {}
	explicit operator double() const { return v; }
	explicit operator float() const { return (float)v; }
	explicit operator int() const { return (int)v; }
	explicit operator long() const { return (long)v; }
	explicit operator unsigned() const { return (unsigned)v; }
	explicit operator unsigned long() const { return (unsigned long)v; }
	Dual &operator+=(Dual b) { v = v + b.v; d = d + b.d; return *this; }
	Dual &operator-=(Dual b) { v = v - b.v; d = d - b.d; return *this; }
	Dual &operator*=(Dual b) { d = d * b.v + v * b.d; v = v * b.v; return *this; }
	Dual &operator/=(Dual b) { const double q = v / b.v; d = (d - q * b.d) / b.v; v = q; return *this; }
};

inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { const double q = a.v / b.v; return Dual(q, (a.d - q * b.d) / b.v); }
inline bool operator==(Dual a, Dual b) { return a.v == b.v; }
inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline bool operator!(Dual a) { return !a.v; }

inline Dual sqrt(Dual a)
{
	const double s = ::sqrt(a.v);
	return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
inline Dual fabs(Dual a) { return a.v < 0.0 ? -a : (a.v > 0.0 ? a : Dual(::fabs(a.v), a.d)); }
inline Dual fma(Dual a, Dual b, Dual c) { return Dual(::fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d); }
inline Dual cos(Dual a) { return Dual(::cos(a.v), -::sin(a.v) * a.d); }
inline Dual sin(Dual a) { return Dual(::sin(a.v), ::cos(a.v) * a.d); }
inline Dual exp(Dual a)
{
	const double e = ::exp(a.v);
	return Dual(e, e * a.d);
}
/* piecewise constant: zero derivative between the integers (a ratchet's notch) */
inline Dual floor(Dual a) { return Dual(::floor(a.v), 0.0); }

#endif
