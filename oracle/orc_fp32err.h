// CPU oracle — test infrastructure only.
//
// Running error analysis of an FP32 evaluation (Higham, "Accuracy and
// Stability of Numerical Algorithms", 2nd ed., section 3.3).  Err carries
// the exact value of an expression (in double) and a bound on the absolute
// difference between it and the same expression evaluated in FP32, operation
// by operation in the order written.  Instantiating the oracle's templated
// BSDF code with T = Err therefore yields, per query, a derived bound on how
// far an FP32 evaluation of the reference's formula can be from the exact
// result:
//
//   * inputs that are FP32 values are exact (e = 0); a decimal constant carries
//     its FP32 rounding |fl32(c) - c|;
//   * +, -, *, / propagate the operands' bounds to first and second order and
//     add the result's rounding, u |r| + eta, u = 2^-24, eta = 2^-149 (the
//     FP32 denormal spacing, for results that underflow);
//   * sqrt, exp, log, pow, sin, cos, asin, atan2, sinh propagate the input
//     bound through the function's range over [x - e, x + e] and add
//     TRANS_ULPS ulps of the result (the FP32 device library's documented
//     accuracy is 1-4 ulps for these);
//   * fabs, max, min are 1-Lipschitz and add nothing;
//   * a comparison whose operands' error intervals overlap could take the
//     other branch in FP32: it sets the thread's `ambiguous` flag, and the
//     caller bounds such queries by other means;
//   * a divisor (or a log argument) whose interval reaches zero gives an
//     unbounded result (e = inf): no finite first-order bound exists there.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>
#include <type_traits>

namespace orc {

struct Err {
    double v = 0.0, e = 0.0;
    static constexpr double U = 5.9604644775390625e-08;   // 2^-24
    static constexpr double ETA = 1.401298464324817e-45;  // 2^-149
    static constexpr double TRANS_ULPS = 4.0;

    constexpr Err() = default;
    constexpr Err(double x, double err) : v(x), e(err) {}
    // A decimal constant: the FP32 program holds fl32(x).
    constexpr Err(double x) : v(x), e(ad(double(float(x)) - x)) {}
    constexpr Err(float x) : v(x), e(0.0) {}
    template <typename I, typename std::enable_if<std::is_integral<I>::value, int>::type = 0>
    constexpr Err(I x) : v(double(x)), e(0.0) {}

    static constexpr double ad(double x) { return x < 0 ? -x : x; }
    static double rnd(double r) { return U * std::fabs(r) + ETA; }
    static double trans(double r) { return TRANS_ULPS * U * std::fabs(r) + ETA; }
    static bool &ambiguous() {
        static thread_local bool flag = false;
        return flag;
    }
};

inline Err operator+(Err a, Err b) {
    const double r = a.v + b.v;
    return {r, a.e + b.e + Err::U * (std::fabs(r) + a.e + b.e) + Err::ETA};
}
inline Err operator-(Err a, Err b) {
    const double r = a.v - b.v;
    return {r, a.e + b.e + Err::U * (std::fabs(r) + a.e + b.e) + Err::ETA};
}
inline Err operator-(Err a) { return {-a.v, a.e}; }
// x * e where an exact zero times an unbounded error is zero.
inline double err_mul(double x, double e) { return x == 0.0 ? 0.0 : x * e; }
inline Err operator*(Err a, Err b) {
    const double r = a.v * b.v;
    const double p = err_mul(std::fabs(a.v), b.e) + err_mul(std::fabs(b.v), a.e) + err_mul(a.e, b.e);
    return {r, p + Err::U * (std::fabs(r) + p) + Err::ETA};
}
inline Err operator/(Err a, Err b) {
    const double r = a.v / b.v;
    const double d = std::fabs(b.v) - b.e;
    if (!(d > 0.0)) return {r, std::numeric_limits<double>::infinity()};
    const double p = (a.e + err_mul(std::fabs(r), b.e)) / d;
    return {r, p + Err::U * (std::fabs(r) + p) + Err::ETA};
}
inline Err &operator+=(Err &a, Err b) { return a = a + b; }
inline Err &operator-=(Err &a, Err b) { return a = a - b; }
inline Err &operator*=(Err &a, Err b) { return a = a * b; }
inline Err &operator/=(Err &a, Err b) { return a = a / b; }

inline void err_note(Err a, Err b) {
    const double s = a.e + b.e;
    if (s > 0.0 && std::fabs(a.v - b.v) <= s) Err::ambiguous() = true;
}
inline bool operator<(Err a, Err b) { err_note(a, b); return a.v < b.v; }
inline bool operator<=(Err a, Err b) { err_note(a, b); return a.v <= b.v; }
inline bool operator>(Err a, Err b) { err_note(a, b); return a.v > b.v; }
inline bool operator>=(Err a, Err b) { err_note(a, b); return a.v >= b.v; }
inline bool operator==(Err a, Err b) { err_note(a, b); return a.v == b.v; }
inline bool operator!=(Err a, Err b) { err_note(a, b); return a.v != b.v; }

// The range of a function monotone on [x - e, x + e] (or evaluated at its ends).
template <class F> inline double err_range(F f, double x, double e, double r) {
    const double lo = f(x - e), hi = f(x + e);
    double d = std::max(std::fabs(lo - r), std::fabs(hi - r));
    return std::isfinite(d) ? d : std::numeric_limits<double>::infinity();
}

} // namespace orc

// Overloads for the std:: calls in the templated oracle code (test build only).
namespace std {
inline orc::Err fabs(orc::Err a) { return {std::fabs(a.v), a.e}; }
inline orc::Err max(orc::Err a, orc::Err b) { return {a.v >= b.v ? a.v : b.v, std::max(a.e, b.e)}; }
inline orc::Err min(orc::Err a, orc::Err b) { return {a.v <= b.v ? a.v : b.v, std::max(a.e, b.e)}; }
inline orc::Err sqrt(orc::Err a) {
    const double r = std::sqrt(a.v);
    double p;
    if (a.v - a.e > 0.0) p = orc::err_range([](double x) { return std::sqrt(x); }, a.v, a.e, r);
    else p = std::sqrt(std::max(0.0, a.v + a.e)) + r;  // the interval reaches zero
    return {r, p + orc::Err::trans(r)};
}
inline orc::Err exp(orc::Err a) {
    const double r = std::exp(a.v);
    const double p = std::exp(a.v + a.e) - r;
    return {r, (std::isfinite(p) ? p : std::numeric_limits<double>::infinity()) + orc::Err::trans(r + p)};
}
inline orc::Err log(orc::Err a) {
    const double r = std::log(a.v);
    if (!(a.v - a.e > 0.0)) return {r, std::numeric_limits<double>::infinity()};
    const double p = r - std::log(a.v - a.e);
    return {r, p + orc::Err::trans(r)};
}
inline orc::Err sinh(orc::Err a) {
    const double r = std::sinh(a.v);
    return {r, std::cosh(std::fabs(a.v) + a.e) * a.e + orc::Err::trans(r)};
}
inline orc::Err sin(orc::Err a) {
    const double r = std::sin(a.v);
    return {r, std::min(a.e, 2.0) + orc::Err::trans(r)};
}
inline orc::Err cos(orc::Err a) {
    const double r = std::cos(a.v);
    return {r, std::min(a.e, 2.0) + orc::Err::trans(r)};
}
inline orc::Err asin(orc::Err a) {
    const double r = std::asin(a.v);
    auto f = [](double x) { return std::asin(std::max(-1.0, std::min(1.0, x))); };
    return {r, orc::err_range(f, a.v, a.e, r) + orc::Err::trans(r)};
}
inline orc::Err atan2(orc::Err y, orc::Err x) {
    const double r = std::atan2(y.v, x.v);
    const double rad = std::sqrt(x.v * x.v + y.v * y.v) - (x.e + y.e);
    if (!(rad > 0.0)) return {r, M_PI};
    const double p = (std::fabs(x.v) * y.e + std::fabs(y.v) * x.e) / (rad * rad);
    return {r, std::min(p, M_PI) + orc::Err::trans(r)};
}
inline orc::Err pow(orc::Err a, orc::Err b) {
    const double r = std::pow(a.v, b.v);
    double p = 0.0;
    if (a.e > 0.0) p += orc::err_range([&](double x) { return std::pow(x, b.v); }, a.v, a.e, r);
    if (b.e > 0.0) {
        if (!(a.v > 0.0)) return {r, std::numeric_limits<double>::infinity()};
        p += std::fabs(r * std::log(a.v)) * b.e * 1.01;
    }
    return {r, p + orc::Err::trans(r)};
}
} // namespace std
