// Helpers of the plugin surface: 4x4 transforms, frames, string parsing
// (semantics of src/cpu/common.cpp:27-116 of the reference).
#include "common.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <iomanip>

NORI_NAMESPACE_BEGIN

namespace {
// Inverse of a 4x4 row-major matrix by Gauss-Jordan elimination with partial
// pivoting (in double precision).
bool invert4(const float *m, float *out) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c) a[r][c] = c < 4 ? (double)m[4 * r + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (std::fabs(a[p][c]) < 1e-30) return false;
        if (p != c)
            for (int k = 0; k < 8; ++k) std::swap(a[p][k], a[c][k]);
        double inv = 1.0 / a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            double f = a[r][c];
            if (f == 0.0) continue;
            for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[4 * r + c] = (float)a[r][c + 4];
    return true;
}
} // namespace

void Transform::set(const float m[16]) {
    for (int i = 0; i < 16; ++i) m_m[i] = m[i];
    if (!invert4(m_m, m_inv))
        for (int i = 0; i < 16; ++i) m_inv[i] = NAN;
}

Transform Transform::inverse() const {
    Transform t;
    for (int i = 0; i < 16; ++i) {
        t.m_m[i] = m_inv[i];
        t.m_inv[i] = m_m[i];
    }
    return t;
}

Transform Transform::operator*(const Transform &o) const {
    float r[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += m_m[4 * i + k] * o.m_m[4 * k + j];
            r[4 * i + j] = s;
        }
    return Transform(r);
}

Point3f Transform::transformPoint(const Point3f &p) const {
    const float *m = m_m;
    float x = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    float y = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    float z = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
    float w = m[12] * p[0] + m[13] * p[1] + m[14] * p[2] + m[15];
    return Point3f(x, y, z) / w;
}

Vector3f Transform::transformVector(const Vector3f &v) const {
    const float *m = m_m;
    return {m[0] * v[0] + m[1] * v[1] + m[2] * v[2], m[4] * v[0] + m[5] * v[1] + m[6] * v[2],
            m[8] * v[0] + m[9] * v[1] + m[10] * v[2]};
}

Normal3f Transform::transformNormal(const Normal3f &n) const {
    const float *m = m_inv; // inverse transpose
    return {m[0] * n[0] + m[4] * n[1] + m[8] * n[2], m[1] * n[0] + m[5] * n[1] + m[9] * n[2],
            m[2] * n[0] + m[6] * n[1] + m[10] * n[2]};
}

std::string Transform::toString() const {
    std::ostringstream os;
    os << "[";
    for (int r = 0; r < 4; ++r) {
        os << (r ? "\n " : "");
        for (int c = 0; c < 4; ++c) os << (c ? ", " : "") << m_m[4 * r + c];
    }
    os << "]";
    return os.str();
}

Frame::Frame(const Normal3f &nn) : n(nn) {
    // BuildONB (shaders/GeometryUtils.hlsli:160-165): s = normalize(up x n), t = n x s
    Vector3f up = std::fabs(n[1]) < 0.999f ? Vector3f(0, 1, 0) : Vector3f(1, 0, 0);
    s = up.cross(n).normalized();
    t = n.cross(s);
}

std::string Frame::toString() const {
    return format("Frame[\n  s = %s,\n  t = %s,\n  n = %s\n]", s.toString(), t.toString(), n.toString());
}

std::string indent(const std::string &string, int amount) {
    std::istringstream iss(string);
    std::ostringstream oss;
    std::string spacer(amount, ' ');
    bool first = true;
    for (std::string line; std::getline(iss, line);) {
        if (!first) oss << spacer;
        oss << line;
        if (!iss.eof()) oss << '\n';
        first = false;
    }
    return oss.str();
}

std::string toLower(const std::string &value) {
    std::string r(value);
    for (auto &ch : r) ch = (char)std::tolower((unsigned char)ch);
    return r;
}

bool endsWith(const std::string &value, const std::string &ending) {
    return ending.size() <= value.size() && std::equal(ending.rbegin(), ending.rend(), value.rbegin());
}

bool toBool(const std::string &str) {
    std::string v = toLower(str);
    if (v == "false") return false;
    if (v == "true") return true;
    throw NoriException("Could not parse boolean value \"%s\"", str);
}

int toInt(const std::string &str) {
    char *end = nullptr;
    long r = std::strtol(str.c_str(), &end, 10);
    if (str.empty() || *end != '\0') throw NoriException("Could not parse integer value \"%s\"", str);
    return (int)r;
}

unsigned int toUInt(const std::string &str) {
    char *end = nullptr;
    unsigned long r = std::strtoul(str.c_str(), &end, 10);
    if (str.empty() || *end != '\0') throw NoriException("Could not parse integer value \"%s\"", str);
    return (unsigned int)r;
}

float toFloat(const std::string &str) {
    char *end = nullptr;
    float r = std::strtof(str.c_str(), &end);
    if (str.empty() || *end != '\0') throw NoriException("Could not parse floating point value \"%s\"", str);
    return r;
}

Vector3f toVector3f(const std::string &str) {
    std::vector<std::string> t = tokenize(str);
    if (t.size() != 3) throw NoriException("Expected 3 values");
    return {toFloat(t[0]), toFloat(t[1]), toFloat(t[2])};
}

std::vector<std::string> tokenize(const std::string &s, const std::string &delim, bool includeEmpty) {
    std::vector<std::string> out;
    size_t start = 0;
    while (true) {
        size_t pos = s.find_first_of(delim, start);
        std::string tok = s.substr(start, pos == std::string::npos ? std::string::npos : pos - start);
        if (!tok.empty() || includeEmpty) out.push_back(tok);
        if (pos == std::string::npos) break;
        start = pos + 1;
    }
    return out;
}

std::string memString(size_t size) {
    double v = (double)size;
    const char *suffix[] = {"B", "KiB", "MiB", "GiB", "TiB", "PiB"};
    int k = 0;
    while (k < 5 && v > 1024.0) {
        v /= 1024.0;
        ++k;
    }
    std::ostringstream os;
    os << std::fixed << std::setprecision(k == 0 ? 0 : 1) << v << " " << suffix[k];
    return os.str();
}

std::string FileResolver::resolve(const std::string &name) const {
    if (!name.empty() && name[0] == '/') return name;
    for (const auto &dir : m_paths) {
        std::string cand = dir.empty() ? name : dir + "/" + name;
        std::ifstream f(cand);
        if (f.good()) return cand;
    }
    return name;
}

FileResolver *getFileResolver() {
    static FileResolver resolver;
    return &resolver;
}

NORI_NAMESPACE_END
