// Forward-mode dual numbers: a value and N directional derivatives, for host and device.
//
// The generated-model surface (Model<T>, Numerics<T>) is written against a scalar template
// parameter and uses +, -, *, /, sqrt, fabs, pow, max and min, so instantiating it on Dual<N>
// yields the exact derivatives of the same arithmetic (branches of max / min / the
// comparisons follow the values, i.e. they are frozen at the point of evaluation).  Used for
// per-face and per-cell Jacobian-transpose products in the discrete adjoint and for the
// tangent-linear reference run of the oracle.
#pragma once
#include <cmath>

#ifndef ZDM_HD
#if defined(__HIP__) || defined(__HIPCC__)
#define ZDM_HD __host__ __device__
#else
#define ZDM_HD
#endif
#endif

template <int N>
struct Dual {
    double v;
    double d[N];
    constexpr ZDM_HD Dual() : v(0), d{} {}
    constexpr ZDM_HD Dual(double x) : v(x), d{} {}
    ZDM_HD static Dual seed(double x, int k) {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
    ZDM_HD explicit operator double() const { return v; }
};

template <int N>
ZDM_HD inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
ZDM_HD inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
ZDM_HD inline Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
ZDM_HD inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
ZDM_HD inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    const double q = a.v / b.v;
    Dual<N> r(q);
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
    return r;
}
template <int N>
ZDM_HD inline Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { return a = a + b; }
template <int N>
ZDM_HD inline Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { return a = a - b; }
template <int N>
ZDM_HD inline Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) { return a = a * b; }
template <int N>
ZDM_HD inline Dual<N>& operator/=(Dual<N>& a, const Dual<N>& b) { return a = a / b; }
template <int N>
ZDM_HD inline Dual<N>& operator+=(Dual<N>& a, double b) { a.v += b; return a; }
template <int N>
ZDM_HD inline Dual<N>& operator-=(Dual<N>& a, double b) { a.v -= b; return a; }
template <int N>
ZDM_HD inline Dual<N>& operator*=(Dual<N>& a, double b) { return a = a * Dual<N>(b); }
template <int N>
ZDM_HD inline Dual<N>& operator/=(Dual<N>& a, double b) { return a = a / Dual<N>(b); }

#define ZDM_DUAL_MIXED(op)                                                                        \
    template <int N>                                                                              \
    ZDM_HD inline Dual<N> operator op(const Dual<N>& a, double b) { return a op Dual<N>(b); }    \
    template <int N>                                                                              \
    ZDM_HD inline Dual<N> operator op(double a, const Dual<N>& b) { return Dual<N>(a) op b; }
ZDM_DUAL_MIXED(+)
ZDM_DUAL_MIXED(-)
ZDM_DUAL_MIXED(*)
ZDM_DUAL_MIXED(/)
#undef ZDM_DUAL_MIXED

#define ZDM_DUAL_CMP(op)                                                                          \
    template <int N>                                                                              \
    ZDM_HD inline bool operator op(const Dual<N>& a, const Dual<N>& b) { return a.v op b.v; }    \
    template <int N>                                                                              \
    ZDM_HD inline bool operator op(const Dual<N>& a, double b) { return a.v op b; }              \
    template <int N>                                                                              \
    ZDM_HD inline bool operator op(double a, const Dual<N>& b) { return a op b.v; }
ZDM_DUAL_CMP(<)
ZDM_DUAL_CMP(>)
ZDM_DUAL_CMP(<=)
ZDM_DUAL_CMP(>=)
ZDM_DUAL_CMP(==)
ZDM_DUAL_CMP(!=)
#undef ZDM_DUAL_CMP

namespace std {
template <int N>
ZDM_HD inline Dual<N> sqrt(const Dual<N>& a) {
    const double s = std::sqrt(a.v);
    Dual<N> r(s);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / (2.0 * s);
    return r;
}
template <int N>
ZDM_HD inline Dual<N> fabs(const Dual<N>& a) { return a.v < 0 ? -a : a; }
template <int N>
ZDM_HD inline bool isfinite(const Dual<N>& a) { return std::isfinite(a.v); }
// x^y for a constant exponent: d(x^y) = y x^(y-1) dx (0 for y = 0)
template <int N>
ZDM_HD inline Dual<N> pow(const Dual<N>& a, double y) {
    Dual<N> r(std::pow(a.v, y));
    const double dy = y == 0.0 ? 0.0 : y * std::pow(a.v, y - 1.0);
    for (int k = 0; k < N; ++k) r.d[k] = dy * a.d[k];
    return r;
}
template <int N>
ZDM_HD inline Dual<N> pow(const Dual<N>& a, int n) { return pow(a, (double)n); }
// a^y for a dual exponent: d(a^y) = a^y ln(a) dy (+ y a^(y-1) da)
template <int N>
ZDM_HD inline Dual<N> pow(double a, const Dual<N>& y) {
    const double p = std::pow(a, y.v);
    Dual<N> r(p);
    const double l = p * std::log(a);
    for (int k = 0; k < N; ++k) r.d[k] = l * y.d[k];
    return r;
}
template <int N>
ZDM_HD inline Dual<N> pow(const Dual<N>& a, const Dual<N>& y) {
    const double p = std::pow(a.v, y.v);
    Dual<N> r(p);
    const double da = y.v == 0.0 ? 0.0 : y.v * std::pow(a.v, y.v - 1.0);
    bool dep = false;
    for (int k = 0; k < N; ++k) dep |= y.d[k] != 0.0;
    const double dl = dep ? p * std::log(a.v) : 0.0;
    for (int k = 0; k < N; ++k) r.d[k] = da * a.d[k] + dl * y.d[k];
    return r;
}
}  // namespace std

// value of a scalar or of a dual number
ZDM_HD inline double zdm_value(double x) { return x; }
template <int N>
ZDM_HD inline double zdm_value(const Dual<N>& x) { return x.v; }
