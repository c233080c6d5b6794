// CPU oracle — test infrastructure only.
//
// Small vector library for the CPU restatement of the reference renderer.
// Every arithmetic expression is written in one fixed evaluation order and the
// oracle is compiled with -ffp-contract=off, so that integer/float results that
// the tests require to be bit-exact (ray/triangle hits) are reproducible.
//
// Follows: include/vector.h (Nori TVector), include/frame.h, shaders/GeometryUtils.hlsli.
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>
// Err (FP32 running error analysis) and its std:: overloads must precede the
// templates below, whose qualified std:: calls bind at their definition.
#include "orc_fp32err.h"

namespace orc {

template <typename T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    explicit V3(T a) : x(a), y(a), z(a) {}
    template <typename U> explicit V3(const V3<U> &o) : x((T)o.x), y((T)o.y), z((T)o.z) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename T> inline V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> inline V3<T> &operator+=(V3<T> &a, V3<T> b) { a = a + b; return a; }
template <typename T> inline V3<T> &operator*=(V3<T> &a, V3<T> b) { a = a * b; return a; }
template <typename T> inline V3<T> &operator*=(V3<T> &a, T s) { a = a * s; return a; }
template <typename T> inline V3<T> &operator/=(V3<T> &a, T s) { a = a / s; return a; }

// (ax*bx + ay*by) + az*bz, left to right.
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// HLSL normalize(v) = v * rsqrt(dot(v,v)); the oracle uses v / sqrt(dot(v,v)).
template <typename T> inline V3<T> normalize(V3<T> a) { T l = std::sqrt(dot(a, a)); return a / l; }
template <typename T> inline T maxc(V3<T> a) { return std::max(a.x, std::max(a.y, a.z)); }
template <typename T> inline V3<T> lerp3(V3<T> a, V3<T> b, T t) { return a + (b - a) * t; }
template <typename T> inline T lerp1(T a, T b, T t) { return a + (b - a) * t; }
template <typename T> inline T saturate(T v) { return std::min(std::max(v, T(0)), T(1)); }
template <typename T> inline V3<T> reflect(V3<T> i, V3<T> n) { return i - T(2) * dot(n, i) * n; }
// HLSL refract(i, n, eta): zero vector on total internal reflection.
template <typename T> inline V3<T> refract(V3<T> i, V3<T> n, T eta) {
    T c = dot(n, i);
    T k = T(1) - eta * eta * (T(1) - c * c);
    if (k < T(0)) return V3<T>(T(0));
    return eta * i - (eta * c + std::sqrt(k)) * n;
}

typedef V3<float> V3f;
typedef V3<double> V3d;

template <typename T> struct Consts {
    static constexpr T PI = T(3.14159265358979323846);
    static constexpr T INV_PI = T(0.31830988618379067154);
};

// shaders/GeometryUtils.hlsli:225-228
template <typename T> inline T luminance(V3<T> c) { return T(0.212671) * c.x + T(0.715160) * c.y + T(0.072169) * c.z; }

// shaders/GeometryUtils.hlsli:160-165 — the GPU shading frame (differs from Nori's coordinateSystem).
template <typename T> inline void build_onb(V3<T> n, V3<T> &t, V3<T> &b) {
    V3<T> up = std::fabs(n.y) < T(0.999) ? V3<T>(0, 1, 0) : V3<T>(1, 0, 0);
    t = normalize(cross(up, n));
    b = cross(n, t);
}
template <typename T> inline V3<T> to_local(V3<T> v, V3<T> t, V3<T> b, V3<T> n) { return {dot(v, t), dot(v, b), dot(v, n)}; }
template <typename T> inline V3<T> to_world(V3<T> v, V3<T> t, V3<T> b, V3<T> n) { return t * v.x + b * v.y + n * v.z; }

// shaders/GeometryUtils.hlsli:178-183
template <typename T> inline V3<T> offset_ray_origin(V3<T> p, V3<T> ng, V3<T> n) {
    V3<T> q = p + n * T(0.002);
    q += ng * std::max(T(0), -dot(q - p, ng));
    return q;
}

// shaders/GeometryUtils.hlsli:18-24
template <typename T> inline T balance_heuristic(T a, T b) {
    T s = a + b;
    if (s < T(1e-20)) return T(0);
    return a / s;
}

// shaders/GeometryUtils.hlsli:189-194 (z = sqrt(1-u.x); differs from Nori's warp.cpp:81-86)
template <typename T> inline V3<T> cosine_sample_hemisphere(T ux, T uy) {
    T r = std::sqrt(ux);
    T phi = T(2) * Consts<T>::PI * uy;
    return {r * std::cos(phi), r * std::sin(phi), std::sqrt(std::max(T(0), T(1) - ux))};
}

// shaders/GeometryUtils.hlsli:200-219 (same formula as src/cpu/common.cpp:259-288)
template <typename T> inline T fresnel_dielectric(T cosi, T etaI, T etaT) {
    if (etaI == etaT) return T(0);
    if (cosi < T(0)) { std::swap(etaI, etaT); cosi = -cosi; }
    T eta = etaI / etaT;
    T s2 = eta * eta * (T(1) - cosi * cosi);
    if (s2 > T(1)) return T(1);
    T cost = std::sqrt(T(1) - s2);
    T rs = (etaI * cosi - etaT * cost) / (etaI * cosi + etaT * cost);
    T rp = (etaT * cosi - etaI * cost) / (etaT * cosi + etaI * cost);
    return T(0.5) * (rs * rs + rp * rp);
}

} // namespace orc
