// Dual numbers for forward-mode derivatives in device code: a value and its partial derivatives with respect to N
// chosen inputs.  A function written once for a template scalar T runs on double (the value) or on Dual<N> (the
// value, computed by the same floating-point operations in the same order, and its N derivatives).
#pragma once
#include "g3d_common.hpp"

namespace g3d {

template <int N>
struct Dual {
    double v, d[N];
};
template <int N>
G3D_HD Dual<N> dconst(double v) {
    Dual<N> r;
    r.v = v;
    for (int k = 0; k < N; ++k) r.d[k] = 0;
    return r;
}
template <int N>
G3D_HD Dual<N> dvar(double v, int k) {
    Dual<N> r = dconst<N>(v);
    r.d[k] = 1;
    return r;
}
// r.v = value; r.d = a.d * da + b.d * db
template <int N>
G3D_HD Dual<N> dlin(double value, const Dual<N>& a, double da, const Dual<N>& b, double db) {
    Dual<N> r;
    r.v = value;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * da + b.d[k] * db;
    return r;
}
// the chain rule through f with f(a.v) = fv and f'(a.v) = fd
template <int N>
G3D_HD Dual<N> dchain(const Dual<N>& a, double fv, double fd) {
    Dual<N> r;
    r.v = fv;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * fd;
    return r;
}
template <int N>
G3D_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { return dlin(a.v + b.v, a, 1.0, b, 1.0); }
template <int N>
G3D_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { return dlin(a.v - b.v, a, 1.0, b, -1.0); }
template <int N>
G3D_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { return dlin(a.v * b.v, a, b.v, b, a.v); }
template <int N>
G3D_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    const double q = a.v / b.v;
    return dlin(q, a, 1 / b.v, b, -q / b.v);
}
template <int N> G3D_HD Dual<N> operator+(const Dual<N>& a, double b) { return dchain(a, a.v + b, 1.0); }
template <int N> G3D_HD Dual<N> operator+(double a, const Dual<N>& b) { return dchain(b, a + b.v, 1.0); }
template <int N> G3D_HD Dual<N> operator-(const Dual<N>& a, double b) { return dchain(a, a.v - b, 1.0); }
template <int N> G3D_HD Dual<N> operator-(double a, const Dual<N>& b) { return dchain(b, a - b.v, -1.0); }
template <int N> G3D_HD Dual<N> operator-(const Dual<N>& a) { return dchain(a, -a.v, -1.0); }
template <int N> G3D_HD Dual<N> operator*(const Dual<N>& a, double b) { return dchain(a, a.v * b, b); }
template <int N> G3D_HD Dual<N> operator*(double a, const Dual<N>& b) { return dchain(b, a * b.v, a); }
template <int N> G3D_HD Dual<N> operator/(const Dual<N>& a, double b) { return dchain(a, a.v / b, 1 / b); }
template <int N>
G3D_HD Dual<N> operator/(double a, const Dual<N>& b) {
    const double q = a / b.v;
    return dchain(b, q, -q / b.v);
}
template <int N> G3D_HD Dual<N> tsqrt(const Dual<N>& a) { const double s = sqrt(a.v); return dchain(a, s, 0.5 / s); }
template <int N> G3D_HD Dual<N> texp(const Dual<N>& a) { const double e = exp(a.v); return dchain(a, e, e); }
template <int N> G3D_HD Dual<N> tcos(const Dual<N>& a) { return dchain(a, cos(a.v), -sin(a.v)); }
template <int N> G3D_HD Dual<N> tsin(const Dual<N>& a) { return dchain(a, sin(a.v), cos(a.v)); }
G3D_HD double tsqrt(double a) { return sqrt(a); }
G3D_HD double texp(double a) { return exp(a); }
G3D_HD double tcos(double a) { return cos(a); }
G3D_HD double tsin(double a) { return sin(a); }
G3D_HD double val(double a) { return a; }
template <int N> G3D_HD double val(const Dual<N>& a) { return a.v; }
// a constant of the scalar type T
template <typename T> struct Cst { static G3D_HD T of(double v) { return v; } };
template <int N> struct Cst<Dual<N>> { static G3D_HD Dual<N> of(double v) { return dconst<N>(v); } };
template <typename T> G3D_HD T cst(double v) { return Cst<T>::of(v); }

}  // namespace g3d
