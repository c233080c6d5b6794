// CPU oracle — test infrastructure only.  See orc_guide.h.
#include "orc_guide.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace orc {

static const double PI_D = 3.14159265358979323846;

float bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x; // inf / nan unchanged
    uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7fffu + lsb) & 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

void guide_features(const float p[3], const float n[3], const float wi[3], float rough, float metal, float *f) {
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 8; ++k) {
            float arg = (float)(1 << k) * 3.14159265358979323846f * p[a];
            f[a * 16 + 2 * k] = std::sin(arg);
            f[a * 16 + 2 * k + 1] = std::cos(arg);
        }
    for (int k = 0; k < 3; ++k) {
        f[48 + k] = n[k];
        f[51 + k] = wi[k];
    }
    f[54] = rough;
    f[55] = metal;
    for (int k = 56; k < 64; ++k) f[k] = 0.0f;
}

void guide_forward(const float *params, const float *feat, int bf16, double *out, double *acts) {
    double h[64], z[64];
    for (int i = 0; i < 64; ++i) h[i] = bf16 ? bf16_round(feat[i]) : feat[i];
    for (int l = 0; l < G_LAYERS; ++l) {
        const float *W = params + l * (64 * 64 + 64), *b = W + 64 * 64;
        for (int o = 0; o < 64; ++o) {
            double s = 0.0;
            for (int i = 0; i < 64; ++i) s += (bf16 ? (double)bf16_round(W[o * 64 + i]) : (double)W[o * 64 + i]) * h[i];
            z[o] = s + (double)b[o];
        }
        if (l == G_LAYERS - 1) {
            for (int o = 0; o < 64; ++o) out[o] = z[o];
        } else {
            for (int o = 0; o < 64; ++o) {
                double r = z[o] > 0.0 ? z[o] : 0.0;
                h[o] = bf16 ? (double)bf16_round((float)r) : r;
                if (acts) acts[l * 64 + o] = z[o];
            }
        }
    }
}

namespace {
struct Decoded {
    double mu[8][3], len[8], kappa[8], raw_k[8], w[8], s, c;
};
Decoded decode(const double *o) {
    Decoded d;
    for (int j = 0; j < 8; ++j) {
        double x = o[3 * j], y = o[3 * j + 1], z = o[3 * j + 2];
        double l = std::sqrt(x * x + y * y + z * z);
        l = std::max(l, 1e-12);
        d.len[j] = l;
        d.mu[j][0] = x / l; d.mu[j][1] = y / l; d.mu[j][2] = z / l;
        d.raw_k[j] = o[24 + j];
        d.kappa[j] = std::exp(std::min(std::max(o[24 + j], -4.0), 9.0));
    }
    double mx = o[32];
    for (int j = 1; j < 8; ++j) mx = std::max(mx, o[32 + j]);
    double sum = 0.0;
    for (int j = 0; j < 8; ++j) { d.w[j] = std::exp(o[32 + j] - mx); sum += d.w[j]; }
    for (int j = 0; j < 8; ++j) d.w[j] /= sum;
    d.s = 1.0 / (1.0 + std::exp(-o[40]));
    d.c = 0.1 + 0.5 * d.s;
    return d;
}
inline double vmf(double k, double m) { return k / (2.0 * PI_D * (1.0 - std::exp(-2.0 * k))) * std::exp(k * (m - 1.0)); }
} // namespace

double guide_pdf(const double *out, const double w[3], double b, double *q_out, double *c_out) {
    Decoded d = decode(out);
    double q = 0.0;
    for (int j = 0; j < 8; ++j) {
        double m = d.mu[j][0] * w[0] + d.mu[j][1] * w[1] + d.mu[j][2] * w[2];
        q += d.w[j] * vmf(d.kappa[j], m);
    }
    if (q_out) *q_out = q;
    if (c_out) *c_out = d.c;
    return d.c * q + (1.0 - d.c) * b;
}

double guide_loss_dout(const double *out, const double w[3], double b, double weight, double *g) {
    std::fill(g, g + 64, 0.0);
    Decoded d = decode(out);
    double v[8], m[8], q = 0.0;
    for (int j = 0; j < 8; ++j) {
        m[j] = d.mu[j][0] * w[0] + d.mu[j][1] * w[1] + d.mu[j][2] * w[2];
        v[j] = vmf(d.kappa[j], m[j]);
        q += d.w[j] * v[j];
    }
    double P = d.c * q + (1.0 - d.c) * b;
    if (!(P > 0.0) || weight == 0.0) return 0.0;
    double dP = -weight / P;
    g[40] = dP * (q - b) * 0.5 * d.s * (1.0 - d.s);
    double dq = dP * d.c;
    for (int j = 0; j < 8; ++j) {
        g[32 + j] = dq * d.w[j] * (v[j] - q);
        double k = d.kappa[j], e = std::exp(-2.0 * k);
        double dlogv_dk = 1.0 / k - 2.0 * e / (1.0 - e) + (m[j] - 1.0);
        double dv = dq * d.w[j];
        if (d.raw_k[j] > -4.0 && d.raw_k[j] < 9.0) g[24 + j] = dv * v[j] * dlogv_dk * k;
        double dm = dv * v[j] * k;
        for (int a = 0; a < 3; ++a) g[3 * j + a] = dm * (w[a] - m[j] * d.mu[j][a]) / d.len[j];
    }
    return -weight * std::log(P);
}

double guide_loss_grad(const float *params, const float *feat, const float *w3, const float *bp, const float *wt, int n,
                       double *grad) {
    std::fill(grad, grad + G_PARAMS, 0.0);
    double total = 0.0;
    std::vector<double> hs((G_LAYERS) * 64), zs((G_LAYERS) * 64);
    for (int s = 0; s < n; ++s) {
        const float *f = feat + 64 * (size_t)s;
        // forward keeping activations (inputs of each layer) and pre-activations
        double h[64], z[64];
        for (int i = 0; i < 64; ++i) { h[i] = f[i]; hs[i] = h[i]; }
        for (int l = 0; l < G_LAYERS; ++l) {
            const float *W = params + l * (64 * 64 + 64), *b = W + 64 * 64;
            for (int o = 0; o < 64; ++o) {
                double acc = 0.0;
                for (int i = 0; i < 64; ++i) acc += (double)W[o * 64 + i] * hs[l * 64 + i];
                z[o] = acc + (double)b[o];
                zs[l * 64 + o] = z[o];
            }
            if (l + 1 < G_LAYERS)
                for (int o = 0; o < 64; ++o) hs[(l + 1) * 64 + o] = z[o] > 0.0 ? z[o] : 0.0;
        }
        double wd[3] = {w3[3 * s], w3[3 * s + 1], w3[3 * s + 2]};
        double g[64];
        total += guide_loss_dout(z, wd, bp[s], wt[s], g);
        // backward
        for (int l = G_LAYERS - 1; l >= 0; --l) {
            double *gW = grad + l * (64 * 64 + 64), *gb = gW + 64 * 64;
            const float *W = params + l * (64 * 64 + 64);
            double dh[64];
            for (int i = 0; i < 64; ++i) dh[i] = 0.0;
            for (int o = 0; o < 64; ++o) {
                if (g[o] == 0.0) continue;
                gb[o] += g[o];
                for (int i = 0; i < 64; ++i) {
                    gW[o * 64 + i] += g[o] * hs[l * 64 + i];
                    dh[i] += (double)W[o * 64 + i] * g[o];
                }
            }
            if (l > 0)
                for (int i = 0; i < 64; ++i) g[i] = zs[(l - 1) * 64 + i] > 0.0 ? dh[i] : 0.0;
        }
    }
    return total;
}

void guide_adam(float *p, float *m, float *v, const double *g, int step, float lr, float b1, float b2, float eps) {
    double c1 = 1.0 - std::pow((double)b1, step), c2 = 1.0 - std::pow((double)b2, step);
    for (int i = 0; i < G_PARAMS; ++i) {
        m[i] = (float)(b1 * (double)m[i] + (1.0 - b1) * g[i]);
        v[i] = (float)(b2 * (double)v[i] + (1.0 - b2) * g[i] * g[i]);
        double mh = m[i] / c1, vh = v[i] / c2;
        p[i] = (float)((double)p[i] - lr * mh / (std::sqrt(vh) + eps));
    }
}

void vmf_sample(const double mu[3], double k, double u1, double u2, double w[3]) {
    double cw = 1.0 + std::log(u1 + (1.0 - u1) * std::exp(-2.0 * k)) / k;
    double sw = std::sqrt(std::max(0.0, 1.0 - cw * cw));
    double phi = 2.0 * PI_D * u2;
    // frame around mu (Duff et al. 2017 branchless ONB)
    double sgn = mu[2] >= 0.0 ? 1.0 : -1.0;
    double a = -1.0 / (sgn + mu[2]), bb = mu[0] * mu[1] * a;
    double t[3] = {1.0 + sgn * mu[0] * mu[0] * a, sgn * bb, -sgn * mu[0]};
    double s[3] = {bb, sgn + mu[1] * mu[1] * a, -mu[1]};
    for (int i = 0; i < 3; ++i) w[i] = cw * mu[i] + sw * (std::cos(phi) * t[i] + std::sin(phi) * s[i]);
}

} // namespace orc

extern "C" {
using namespace orc;
int orc_guide_nparams() { return G_PARAMS; }
void orc_guide_features(int n, const float *pos01, const float *nrm, const float *wi, const float *rough,
                        const float *metal, float *feat) {
    for (int i = 0; i < n; ++i)
        guide_features(pos01 + 3 * i, nrm + 3 * i, wi + 3 * i, rough[i], metal[i], feat + 64 * (size_t)i);
}
void orc_guide_forward(const float *params, int n, const float *feat, int bf16, double *out) {
    for (int i = 0; i < n; ++i) guide_forward(params, feat + 64 * (size_t)i, bf16, out + 64 * (size_t)i, nullptr);
}
void orc_guide_pdf(int n, const double *out, const float *w3, const float *b, double *P, double *q, double *c) {
    for (int i = 0; i < n; ++i) {
        double w[3] = {w3[3 * i], w3[3 * i + 1], w3[3 * i + 2]};
        P[i] = guide_pdf(out + 64 * (size_t)i, w, b[i], q + i, c + i);
    }
}
void orc_guide_loss_dout(int n, const double *out, const float *w3, const float *b, const float *wt, double *g,
                         double *loss) {
    for (int i = 0; i < n; ++i) {
        double w[3] = {w3[3 * i], w3[3 * i + 1], w3[3 * i + 2]};
        loss[i] = guide_loss_dout(out + 64 * (size_t)i, w, b[i], wt[i], g + 64 * (size_t)i);
    }
}
double orc_guide_loss_grad(const float *params, int n, const float *feat, const float *w3, const float *b,
                           const float *wt, double *grad) {
    return guide_loss_grad(params, feat, w3, b, wt, n, grad);
}
void orc_guide_adam(float *p, float *m, float *v, const double *g, int step, float lr, float b1, float b2, float eps) {
    guide_adam(p, m, v, g, step, lr, b1, b2, eps);
}
void orc_vmf_sample(int n, const double *mu, const double *k, const double *u1, const double *u2, double *w) {
    for (int i = 0; i < n; ++i) vmf_sample(mu + 3 * i, k[i], u1[i], u2[i], w + 3 * i);
}
float orc_bf16_round(float x) { return bf16_round(x); }
}
