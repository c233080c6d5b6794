// Nori's plugin-surface vocabulary (include/common.h, vector.h, color.h,
// transform.h, ray.h, bbox.h, frame.h of the reference) as small value types:
// the build has no Eigen, and the host side only needs 3-vectors, 4x4
// transforms and 3xN vertex matrices.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#define NORI_NAMESPACE_BEGIN namespace nori {
#define NORI_NAMESPACE_END }

NORI_NAMESPACE_BEGIN

using std::cerr;
using std::cout;
using std::endl;

constexpr float Epsilon = 1e-4f;
constexpr float M_PI_F = 3.14159265358979323846f;
constexpr float INV_PI_F = 0.31830988618379067154f;

// ------------------------------------------------------------- formatting
namespace detail {
inline void fmt_arg(std::ostringstream &os, const std::string &v) { os << v; }
inline void fmt_arg(std::ostringstream &os, const char *v) { os << v; }
template <typename T> void fmt_arg(std::ostringstream &os, const T &v) { os << v; }
inline void format_into(std::ostringstream &os, const char *f) { os << f; }
// printf-style: each %-directive (%s, %i, %d, %f, %u, ...) takes the next argument
// in its natural stream form; "%%" is a literal percent sign.
template <typename T, typename... Rest>
void format_into(std::ostringstream &os, const char *f, const T &v, const Rest &...rest) {
    for (; *f; ++f) {
        if (*f == '%' && f[1] == '%') {
            os << '%';
            ++f;
        } else if (*f == '%') {
            ++f;
            while (*f && std::string("0123456789.-+ #lhz").find(*f) != std::string::npos) ++f;
            fmt_arg(os, v);
            if (*f) ++f;
            format_into(os, f, rest...);
            return;
        } else {
            os << *f;
        }
    }
}
} // namespace detail

template <typename... Args> std::string format(const char *f, const Args &...args) {
    std::ostringstream os;
    detail::format_into(os, f, args...);
    return os.str();
}

/// Exception of the plugin surface (include/common.h:135-140): printf-style message.
class NoriException : public std::runtime_error {
public:
    template <typename... Args>
    NoriException(const char *fmt, const Args &...args) : std::runtime_error(format(fmt, args...)) {}
};

// ------------------------------------------------------------- vectors
struct Vector3f {
    float v[3] = {0.0f, 0.0f, 0.0f};
    Vector3f() = default;
    explicit Vector3f(float s) : v{s, s, s} {}
    Vector3f(float x, float y, float z) : v{x, y, z} {}
    float &x() { return v[0]; }
    float &y() { return v[1]; }
    float &z() { return v[2]; }
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
    float &operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    Vector3f operator+(const Vector3f &o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
    Vector3f operator-(const Vector3f &o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
    Vector3f operator-() const { return {-v[0], -v[1], -v[2]}; }
    Vector3f operator*(float s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
    Vector3f operator/(float s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
    Vector3f &operator+=(const Vector3f &o) { v[0] += o.v[0]; v[1] += o.v[1]; v[2] += o.v[2]; return *this; }
    Vector3f &operator-=(const Vector3f &o) { v[0] -= o.v[0]; v[1] -= o.v[1]; v[2] -= o.v[2]; return *this; }
    Vector3f &operator*=(float s) { v[0] *= s; v[1] *= s; v[2] *= s; return *this; }
    Vector3f &operator/=(float s) { v[0] /= s; v[1] /= s; v[2] /= s; return *this; }
    float dot(const Vector3f &o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
    Vector3f cross(const Vector3f &o) const {
        return {v[1] * o.v[2] - v[2] * o.v[1], v[2] * o.v[0] - v[0] * o.v[2], v[0] * o.v[1] - v[1] * o.v[0]};
    }
    float squaredNorm() const { return dot(*this); }
    float norm() const { return std::sqrt(squaredNorm()); }
    Vector3f normalized() const { return *this / norm(); }
    Vector3f cwiseMin(const Vector3f &o) const {
        return {std::min(v[0], o.v[0]), std::min(v[1], o.v[1]), std::min(v[2], o.v[2])};
    }
    Vector3f cwiseMax(const Vector3f &o) const {
        return {std::max(v[0], o.v[0]), std::max(v[1], o.v[1]), std::max(v[2], o.v[2])};
    }
    std::string toString() const { return format("[%f, %f, %f]", v[0], v[1], v[2]); }
};
inline Vector3f operator*(float s, const Vector3f &a) { return a * s; }
using Point3f = Vector3f;
using Normal3f = Vector3f;

struct Vector2f {
    float v[2] = {0.0f, 0.0f};
    Vector2f() = default;
    explicit Vector2f(float s) : v{s, s} {}
    Vector2f(float x, float y) : v{x, y} {}
    float &x() { return v[0]; }
    float &y() { return v[1]; }
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float &operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    Vector2f operator+(const Vector2f &o) const { return {v[0] + o.v[0], v[1] + o.v[1]}; }
    Vector2f operator-(const Vector2f &o) const { return {v[0] - o.v[0], v[1] - o.v[1]}; }
    Vector2f operator*(float s) const { return {v[0] * s, v[1] * s}; }
    std::string toString() const { return format("[%f, %f]", v[0], v[1]); }
};
using Point2f = Vector2f;

struct Vector2i {
    int v[2] = {0, 0};
    Vector2i() = default;
    Vector2i(int x, int y) : v{x, y} {}
    int &x() { return v[0]; }
    int &y() { return v[1]; }
    int x() const { return v[0]; }
    int y() const { return v[1]; }
};

struct Color3f {
    float c[3] = {0.0f, 0.0f, 0.0f};
    Color3f() = default;
    explicit Color3f(float s) : c{s, s, s} {}
    Color3f(float r, float g, float b) : c{r, g, b} {}
    explicit Color3f(const Vector3f &v) : c{v[0], v[1], v[2]} {}
    float &r() { return c[0]; }
    float &g() { return c[1]; }
    float &b() { return c[2]; }
    float r() const { return c[0]; }
    float g() const { return c[1]; }
    float b() const { return c[2]; }
    float &operator[](int i) { return c[i]; }
    float operator[](int i) const { return c[i]; }
    Color3f operator+(const Color3f &o) const { return {c[0] + o.c[0], c[1] + o.c[1], c[2] + o.c[2]}; }
    Color3f operator*(const Color3f &o) const { return {c[0] * o.c[0], c[1] * o.c[1], c[2] * o.c[2]}; }
    Color3f operator*(float s) const { return {c[0] * s, c[1] * s, c[2] * s}; }
    Color3f operator/(float s) const { return {c[0] / s, c[1] / s, c[2] / s}; }
    Color3f &operator+=(const Color3f &o) { c[0] += o.c[0]; c[1] += o.c[1]; c[2] += o.c[2]; return *this; }
    float maxCoeff() const { return std::max(c[0], std::max(c[1], c[2])); }
    bool isZero() const { return c[0] == 0.0f && c[1] == 0.0f && c[2] == 0.0f; }
    bool isValid() const {
        for (float x : c)
            if (!(x >= 0.0f) || !std::isfinite(x)) return false;
        return true;
    }
    /// Rec. 709 luminance (shaders/GeometryUtils.hlsli:225-228)
    float getLuminance() const { return 0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2]; }
    std::string toString() const { return format("[%f, %f, %f]", c[0], c[1], c[2]); }
};
inline Color3f operator*(float s, const Color3f &a) { return a * s; }

// ------------------------------------------------------------- transforms
/// Homogeneous 4x4 transform with its inverse (include/transform.h).
class Transform {
public:
    Transform() { setIdentity(); }
    explicit Transform(const float m[16]) { set(m); }
    void setIdentity() {
        for (int i = 0; i < 16; ++i) m_m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
        for (int i = 0; i < 16; ++i) m_inv[i] = m_m[i];
    }
    void set(const float m[16]);
    const float *matrix() const { return m_m; } ///< row-major
    const float *inverseMatrix() const { return m_inv; }
    float operator()(int r, int c) const { return m_m[4 * r + c]; }
    Transform inverse() const;
    Transform operator*(const Transform &t) const;
    Point3f transformPoint(const Point3f &p) const;
    Vector3f transformVector(const Vector3f &v) const;
    Normal3f transformNormal(const Normal3f &n) const; ///< by the inverse transpose, not normalised
    std::string toString() const;

private:
    float m_m[16], m_inv[16];
};

// ------------------------------------------------------------- rays, boxes, frames
/// Ray with its valid segment [mint, maxt] (include/ray.h).
struct Ray3f {
    Point3f o;
    Vector3f d, dRcp;
    float mint = Epsilon, maxt = INFINITY;
    Ray3f() = default;
    Ray3f(const Point3f &o, const Vector3f &d) : o(o), d(d) { update(); }
    Ray3f(const Point3f &o, const Vector3f &d, float mint, float maxt) : o(o), d(d), mint(mint), maxt(maxt) {
        update();
    }
    void update() { dRcp = Vector3f(1.0f / d[0], 1.0f / d[1], 1.0f / d[2]); }
    Point3f operator()(float t) const { return o + d * t; }
};

struct BoundingBox3f {
    Point3f min{INFINITY, INFINITY, INFINITY}, max{-INFINITY, -INFINITY, -INFINITY};
    BoundingBox3f() = default;
    explicit BoundingBox3f(const Point3f &p) : min(p), max(p) {}
    BoundingBox3f(const Point3f &a, const Point3f &b) : min(a), max(b) {}
    bool isValid() const { return max[0] >= min[0] && max[1] >= min[1] && max[2] >= min[2]; }
    void expandBy(const Point3f &p) { min = min.cwiseMin(p); max = max.cwiseMax(p); }
    void expandBy(const BoundingBox3f &b) { min = min.cwiseMin(b.min); max = max.cwiseMax(b.max); }
    Point3f getCenter() const { return (min + max) * 0.5f; }
    std::string toString() const { return format("BoundingBox3f[min=%s, max=%s]", min.toString(), max.toString()); }
};

/// Orthonormal frame; the shading frame of the GPU path uses BuildONB
/// (shaders/GeometryUtils.hlsli:160-165).
struct Frame {
    Vector3f s, t;
    Normal3f n;
    Frame() = default;
    Frame(const Vector3f &s, const Vector3f &t, const Normal3f &n) : s(s), t(t), n(n) {}
    explicit Frame(const Normal3f &nn);
    Vector3f toLocal(const Vector3f &v) const { return {v.dot(s), v.dot(t), v.dot(n)}; }
    Vector3f toWorld(const Vector3f &v) const { return s * v[0] + t * v[1] + n * v[2]; }
    std::string toString() const;
};

// ------------------------------------------------------------- dynamic matrices
/// Column-major rows x cols matrix (the 3xN / 2xN vertex arrays of a mesh).
template <typename T> class DynMatrix {
public:
    void resize(size_t rows, size_t cols) {
        m_rows = rows;
        m_cols = cols;
        m_data.assign(rows * cols, T(0));
    }
    size_t rows() const { return m_rows; }
    size_t cols() const { return m_cols; }
    size_t size() const { return m_data.size(); }
    T &operator()(size_t r, size_t c) { return m_data[c * m_rows + r]; }
    const T &operator()(size_t r, size_t c) const { return m_data[c * m_rows + r]; }
    T *data() { return m_data.data(); }
    const T *data() const { return m_data.data(); }
    void setZero() { std::fill(m_data.begin(), m_data.end(), T(0)); }
    Vector3f col3(size_t c) const { return {(float)(*this)(0, c), (float)(*this)(1, c), (float)(*this)(2, c)}; }
    void setCol3(size_t c, const Vector3f &v) {
        (*this)(0, c) = (T)v[0];
        (*this)(1, c) = (T)v[1];
        (*this)(2, c) = (T)v[2];
    }

private:
    size_t m_rows = 0, m_cols = 0;
    std::vector<T> m_data;
};
using MatrixXf = DynMatrix<float>;
using MatrixXu = DynMatrix<uint32_t>;

// ------------------------------------------------------------- string helpers
std::string indent(const std::string &string, int amount = 2);
std::string toLower(const std::string &value);
bool endsWith(const std::string &value, const std::string &ending);
bool toBool(const std::string &str);
int toInt(const std::string &str);
unsigned int toUInt(const std::string &str);
float toFloat(const std::string &str);
Vector3f toVector3f(const std::string &str);
std::vector<std::string> tokenize(const std::string &s, const std::string &delim = ", ", bool includeEmpty = false);
std::string memString(size_t size);
inline float degToRad(float v) { return v * (M_PI_F / 180.0f); }
inline float radToDeg(float v) { return v * (180.0f / M_PI_F); }

/// Directory of the scene file being loaded: relative file names of plugins
/// (OBJ meshes, envmaps, volumes) are resolved against it (the reference's
/// getFileResolver()).
class FileResolver {
public:
    void prepend(const std::string &dir) { m_paths.insert(m_paths.begin(), dir); }
    void clear() { m_paths.clear(); }
    std::string resolve(const std::string &name) const;

private:
    std::vector<std::string> m_paths;
};
FileResolver *getFileResolver();

NORI_NAMESPACE_END
