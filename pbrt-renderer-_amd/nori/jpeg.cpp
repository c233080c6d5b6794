// Baseline JPEG (ITU-T T.81 sequential DCT, Huffman coding, 8-bit samples) to
// RGBA8, for the texture properties of the XML scene format.  The reference
// decodes these files with stb_image (src/gpu/DXRApp_Textures.cpp:37), a
// dependency it does not vendor; this decoder follows the Independent JPEG
// Group's published decoder instead, step for step where the output depends
// on it, so that it matches libjpeg(-turbo) bit for bit:
//   * the "islow" integer inverse DCT (Loeffler-Ligtenberg-Moschytz, 13-bit
//     constants, 2 extra bits between the passes), range-limited to 0..255;
//   * "fancy" chroma upsampling: triangle filters for 2:1 horizontal, 2:1
//     vertical and 2:1 both ways, with alternating rounding bias, the first
//     and last sample rows of a component repeated at the image edges;
//   * YCbCr to RGB (JFIF / T.871) in 16-bit fixed point.
// One or three components (JFIF YCbCr, or RGB when an Adobe marker says so),
// sampling factors 1 or 2, restart intervals.  Progressive, arithmetic-coded,
// 12-bit, lossless, CMYK and multi-scan sequential streams are refused.
#include "image.h"

#include <algorithm>
#include <cstring>

namespace nori {

namespace {

constexpr uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
    // canonical codes, assigned in order of length (T.81 Annex C)
    uint8_t vals[256];
    int maxcode[17], valptr[17], mincode[17];
    bool present = false;
    void build(const uint8_t *counts, const uint8_t *symbols, int nsym) {
        std::memcpy(vals, symbols, (size_t)nsym);
        int code = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            valptr[l] = k;
            mincode[l] = code;
            code += counts[l - 1];
            k += counts[l - 1];
            maxcode[l] = counts[l - 1] ? code - 1 : -1;
            code <<= 1;
        }
        present = true;
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0, td = 0, ta = 0, pred = 0;
    int bw = 0, bh = 0;         // blocks per row / column, whole MCUs
    int dw = 0, dh = 0;         // samples actually in the image at this sampling
    std::vector<uint8_t> plane; // bw*8 x bh*8 samples
    int at(int x, int y) const { return plane[(size_t)y * (size_t)bw * 8 + (size_t)x]; }
};

// islow inverse DCT: fixed-point constants FIX(c) = round(c * 2^13)
constexpr int CONST_BITS = 13, PASS1_BITS = 2;
constexpr int64_t fix13(double c) { return (int64_t)(c * (1 << CONST_BITS) + 0.5); }
constexpr int64_t F0_298631336 = fix13(0.298631336), F0_390180644 = fix13(0.390180644),
                  F0_541196100 = fix13(0.541196100), F0_765366865 = fix13(0.765366865),
                  F0_899976223 = fix13(0.899976223), F1_175875602 = fix13(1.175875602),
                  F1_501321110 = fix13(1.501321110), F1_847759065 = fix13(1.847759065),
                  F1_961570560 = fix13(1.961570560), F2_053119869 = fix13(2.053119869),
                  F2_562915447 = fix13(2.562915447), F3_072711026 = fix13(3.072711026);

inline int64_t descale(int64_t x, int n) { return (x + ((int64_t)1 << (n - 1))) >> n; }

// One 1-D pass over 8 values (before descaling).
inline void idct_1d(const int64_t *in, int64_t out[8]) {
    // even part
    int64_t z2 = in[2], z3 = in[6];
    int64_t z1 = (z2 + z3) * F0_541196100;
    int64_t tmp2 = z1 - z3 * F1_847759065, tmp3 = z1 + z2 * F0_765366865;
    z2 = in[0];
    z3 = in[4];
    int64_t tmp0 = (z2 + z3) * (1 << CONST_BITS), tmp1 = (z2 - z3) * (1 << CONST_BITS);
    const int64_t tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
    // odd part
    tmp0 = in[7];
    tmp1 = in[5];
    tmp2 = in[3];
    tmp3 = in[1];
    z1 = tmp0 + tmp3;
    z2 = tmp1 + tmp2;
    z3 = tmp0 + tmp2;
    int64_t z4 = tmp1 + tmp3;
    const int64_t z5 = (z3 + z4) * F1_175875602;
    tmp0 *= F0_298631336;
    tmp1 *= F2_053119869;
    tmp2 *= F3_072711026;
    tmp3 *= F1_501321110;
    z1 *= -F0_899976223;
    z2 *= -F2_562915447;
    z3 = z3 * -F1_961570560 + z5;
    z4 = z4 * -F0_390180644 + z5;
    tmp0 += z1 + z3;
    tmp1 += z2 + z4;
    tmp2 += z2 + z3;
    tmp3 += z1 + z4;
    out[0] = tmp10 + tmp3;
    out[7] = tmp10 - tmp3;
    out[1] = tmp11 + tmp2;
    out[6] = tmp11 - tmp2;
    out[2] = tmp12 + tmp1;
    out[5] = tmp12 - tmp1;
    out[3] = tmp13 + tmp0;
    out[4] = tmp13 - tmp0;
}

// Dequantised coefficients in natural order -> 8x8 samples (row stride `stride`).
void idct_islow(const int64_t coef[64], uint8_t *out, size_t stride) {
    int64_t ws[64], col[8], res[8];
    for (int x = 0; x < 8; ++x) { // columns
        for (int k = 0; k < 8; ++k) col[k] = coef[k * 8 + x];
        idct_1d(col, res);
        for (int k = 0; k < 8; ++k) ws[k * 8 + x] = descale(res[k], CONST_BITS - PASS1_BITS);
    }
    for (int y = 0; y < 8; ++y) { // rows
        idct_1d(&ws[y * 8], res);
        for (int k = 0; k < 8; ++k) {
            const int64_t v = descale(res[k], CONST_BITS + PASS1_BITS + 3) + 128;
            out[(size_t)y * stride + k] = (uint8_t)std::min<int64_t>(255, std::max<int64_t>(0, v));
        }
    }
}

class Decoder {
public:
    explicit Decoder(const std::vector<uint8_t> &f) : f_(f) {}

    bool run(uint32_t &w, uint32_t &h, std::vector<uint8_t> &rgba, std::string &err) {
        if (f_.size() < 4 || f_[0] != 0xFF || f_[1] != 0xD8) return fail(err, "not a JPEG file");
        p_ = 2;
        bool frame = false;
        for (;;) {
            // a marker: 0xFF fill bytes, then its code
            while (p_ < f_.size() && f_[p_] != 0xFF) ++p_;
            while (p_ < f_.size() && f_[p_] == 0xFF) ++p_;
            if (p_ >= f_.size()) break;
            const uint8_t m = f_[p_++];
            if (m == 0xD9) break; // EOI
            if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
            if (p_ + 2 > f_.size()) return fail(err, "truncated JPEG segment");
            const size_t len = (size_t)f_[p_] << 8 | f_[p_ + 1];
            if (len < 2 || p_ + len > f_.size()) return fail(err, "truncated JPEG segment");
            const uint8_t *d = &f_[p_ + 2];
            const size_t n = len - 2;
            p_ += len;
            if (m == 0xC0 || m == 0xC1) {
                if (frame) return fail(err, "JPEG with two frames");
                if (!sof(d, n, err)) return false;
                frame = true;
            } else if (m >= 0xC2 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) {
                return fail(err, "only baseline sequential JPEG is decoded");
            } else if (m == 0xC4) {
                if (!dht(d, n, err)) return false;
            } else if (m == 0xDB) {
                if (!dqt(d, n, err)) return false;
            } else if (m == 0xDD) {
                if (n < 2) return fail(err, "bad JPEG restart interval");
                restart_ = (int)((size_t)d[0] << 8 | d[1]);
            } else if (m == 0xE0 && n >= 5 && !std::memcmp(d, "JFIF", 5)) {
                jfif_ = true;
            } else if (m == 0xEE && n >= 12 && !std::memcmp(d, "Adobe", 5)) {
                adobe_ = true;
                adobe_transform_ = d[11];
            } else if (m == 0xDA) {
                if (!frame) return fail(err, "JPEG scan before frame header");
                if (scanned_) return fail(err, "only single-scan sequential JPEG is decoded");
                if (!scan(d, n, err)) return false;
            }
        }
        if (!scanned_) return fail(err, "JPEG without image data");
        w = width_;
        h = height_;
        to_rgba(rgba);
        return true;
    }

private:
    const std::vector<uint8_t> &f_;
    size_t p_ = 0;
    uint16_t q_[4][64] = {};
    bool q_present_[4] = {};
    Huffman dc_[4], ac_[4];
    std::vector<Component> comps_;
    uint32_t width_ = 0, height_ = 0;
    int hmax_ = 1, vmax_ = 1, restart_ = 0, adobe_transform_ = 1;
    bool scanned_ = false, jfif_ = false, adobe_ = false;
    uint32_t bits_ = 0;
    int nbits_ = 0;

    static bool fail(std::string &err, const char *m) {
        err = m;
        return false;
    }

    bool sof(const uint8_t *d, size_t n, std::string &err) {
        if (n < 6 || d[0] != 8) return fail(err, "only 8-bit JPEG is decoded");
        height_ = (uint32_t)d[1] << 8 | d[2];
        width_ = (uint32_t)d[3] << 8 | d[4];
        const int nc = d[5];
        if (width_ == 0 || height_ == 0 || width_ > (1u << 15) || height_ > (1u << 15) || (nc != 1 && nc != 3) ||
            n < 6 + 3 * (size_t)nc)
            return fail(err, "unsupported JPEG frame");
        comps_.resize((size_t)nc);
        for (int k = 0; k < nc; ++k) {
            Component &c = comps_[(size_t)k];
            c.id = d[6 + 3 * k];
            c.h = d[7 + 3 * k] >> 4;
            c.v = d[7 + 3 * k] & 15;
            c.tq = d[8 + 3 * k];
            if (c.h < 1 || c.h > 2 || c.v < 1 || c.v > 2 || c.tq > 3)
                return fail(err, "unsupported JPEG sampling factors");
            hmax_ = std::max(hmax_, c.h);
            vmax_ = std::max(vmax_, c.v);
        }
        if (nc == 1) comps_[0].h = comps_[0].v = hmax_ = vmax_ = 1; // one component: one block per MCU
        const uint32_t mcux = (width_ + 8 * hmax_ - 1) / (8 * hmax_), mcuy = (height_ + 8 * vmax_ - 1) / (8 * vmax_);
        for (auto &c : comps_) {
            c.bw = (int)mcux * c.h;
            c.bh = (int)mcuy * c.v;
            c.dw = (int)((width_ * (uint32_t)c.h + hmax_ - 1) / hmax_);
            c.dh = (int)((height_ * (uint32_t)c.v + vmax_ - 1) / vmax_);
            c.plane.assign((size_t)c.bw * c.bh * 64, 0);
        }
        return true;
    }

    bool dht(const uint8_t *d, size_t n, std::string &err) {
        size_t p = 0;
        while (p + 17 <= n) {
            const int tc = d[p] >> 4, th = d[p] & 15;
            if (tc > 1 || th > 3) return fail(err, "bad JPEG Huffman table");
            int nsym = 0;
            for (int k = 0; k < 16; ++k) nsym += d[p + 1 + k];
            if (nsym > 256 || p + 17 + (size_t)nsym > n) return fail(err, "bad JPEG Huffman table");
            (tc ? ac_ : dc_)[th].build(d + p + 1, d + p + 17, nsym);
            p += 17 + (size_t)nsym;
        }
        return true;
    }

    bool dqt(const uint8_t *d, size_t n, std::string &err) {
        size_t p = 0;
        while (p < n) {
            const int pq = d[p] >> 4, tq = d[p] & 15;
            if (tq > 3 || pq > 1 || p + 1 + (pq ? 128 : 64) > n) return fail(err, "bad JPEG quantisation table");
            for (int k = 0; k < 64; ++k)
                q_[tq][kZigzag[k]] = pq ? (uint16_t)(d[p + 1 + 2 * k] << 8 | d[p + 2 + 2 * k]) : d[p + 1 + k];
            q_present_[tq] = true;
            p += 1 + (pq ? 128 : 64);
        }
        return true;
    }

    // entropy-coded bits, 0xFF00 unstuffed; at a marker the stream yields zeros
    int bit() {
        if (nbits_ == 0) {
            uint32_t b = 0;
            if (p_ < f_.size() && f_[p_] != 0xFF) {
                b = f_[p_++];
            } else if (p_ + 1 < f_.size() && f_[p_ + 1] == 0x00) {
                b = 0xFF;
                p_ += 2;
            }
            bits_ = b;
            nbits_ = 8;
        }
        --nbits_;
        return (int)((bits_ >> nbits_) & 1u);
    }
    int receive(int s) {
        int v = 0;
        for (int k = 0; k < s; ++k) v = (v << 1) | bit();
        return v;
    }
    static int extend(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }
    int decode(const Huffman &t) {
        int code = 0;
        for (int l = 1; l <= 16; ++l) {
            code = (code << 1) | bit();
            if (t.maxcode[l] >= 0 && code <= t.maxcode[l]) return t.vals[t.valptr[l] + code - t.mincode[l]];
        }
        return -1;
    }

    bool block(Component &c, int bx, int by, std::string &err) {
        // dequantised in 64 bits: 16-bit quantisation tables times 16-bit
        // coefficients need 32 bits and the IDCT adds more
        int64_t coef[64] = {0};
        const int t = decode(dc_[c.td]);
        if (t < 0 || t > 11) return fail(err, "corrupt JPEG data");
        c.pred += extend(receive(t), t);
        // 8-bit data has |DC| <= 1024; a predictor drifting far past that is
        // corrupt (and would overflow given enough blocks)
        if (c.pred > 32767 || c.pred < -32768) return fail(err, "corrupt JPEG data (DC value out of range)");
        coef[0] = (int64_t)c.pred * (int64_t)q_[c.tq][0];
        for (int k = 1; k < 64;) {
            const int rs = decode(ac_[c.ta]);
            if (rs < 0) return fail(err, "corrupt JPEG data");
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break; // end of block
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) return fail(err, "corrupt JPEG data");
            coef[kZigzag[k]] = (int64_t)extend(receive(s), s) * (int64_t)q_[c.tq][kZigzag[k]];
            ++k;
        }
        idct_islow(coef, &c.plane[((size_t)by * 8 * c.bw + (size_t)bx) * 8], (size_t)c.bw * 8);
        return true;
    }

    bool scan(const uint8_t *d, size_t n, std::string &err) {
        const int ns = d[0];
        if (ns < 1 || n < 1 + 2 * (size_t)ns + 3) return fail(err, "bad JPEG scan header");
        if (ns != (int)comps_.size()) return fail(err, "only single-scan sequential JPEG is decoded");
        std::vector<Component *> sc;
        for (int k = 0; k < ns; ++k) {
            Component *c = nullptr;
            for (auto &q : comps_)
                if (q.id == d[1 + 2 * k]) c = &q;
            if (!c) return fail(err, "JPEG scan refers to an unknown component");
            c->td = d[2 + 2 * k] >> 4;
            c->ta = d[2 + 2 * k] & 15;
            if (c->td > 3 || c->ta > 3 || !dc_[c->td].present || !ac_[c->ta].present || !q_present_[c->tq])
                return fail(err, "JPEG scan refers to a missing table");
            c->pred = 0;
            sc.push_back(c);
        }
        nbits_ = 0;
        const int mcux = comps_[0].bw / comps_[0].h, mcuy = comps_[0].bh / comps_[0].v;
        int todo = restart_;
        for (int my = 0; my < mcuy; ++my)
            for (int mx = 0; mx < mcux; ++mx) {
                if (restart_ && todo == 0) {
                    // byte-align, step over the RSTn marker, reset the DC predictors
                    nbits_ = 0;
                    while (p_ + 1 < f_.size() && !(f_[p_] == 0xFF && f_[p_ + 1] >= 0xD0 && f_[p_ + 1] <= 0xD7)) ++p_;
                    p_ = std::min(p_ + 2, f_.size());
                    for (auto *c : sc) c->pred = 0;
                    todo = restart_;
                }
                for (auto *c : sc)
                    for (int v = 0; v < c->v; ++v)
                        for (int h = 0; h < c->h; ++h)
                            if (!block(*c, mx * c->h + h, my * c->v + v, err)) return false;
                if (restart_) --todo;
            }
        scanned_ = true;
        return true;
    }

    // A component brought to the full sampling grid by libjpeg's "fancy"
    // upsampling; `ow` = its row length (the downsampled width times 1 or 2).
    std::vector<uint8_t> upsample(const Component &c, int &ow) const {
        const int rh = hmax_ / c.h, rv = vmax_ / c.v;
        ow = c.dw * rh;
        std::vector<uint8_t> out((size_t)ow * (size_t)(c.dh * rv));
        for (int y = 0; y < c.dh; ++y)
            for (int vv = 0; vv < rv; ++vv) {
                uint8_t *o = &out[(size_t)(y * rv + vv) * (size_t)ow];
                // 2:1 vertical: 3/4 of this row plus 1/4 of the row above (upper output
                // row) or below (lower); the image's first and last rows stand in for
                // the rows beyond them
                const int yo = std::min(std::max(vv == 0 ? y - 1 : y + 1, 0), c.dh - 1);
                auto colsum = [&](int x) { return rv == 1 ? c.at(x, y) : 3 * c.at(x, y) + c.at(x, yo); };
                if (rh == 2 && c.dw <= 2) {
                    // components at most two samples wide are replicated, not filtered
                    for (int x = 0; x < ow; ++x) o[x] = (uint8_t)c.at(x / 2, y);
                    continue;
                }
                if (rh == 1) {
                    const int bias = vv == 0 ? 1 : 2;
                    for (int x = 0; x < c.dw; ++x) o[x] = (uint8_t)(rv == 1 ? c.at(x, y) : (colsum(x) + bias) >> 2);
                    continue;
                }
                // 2:1 horizontal: 3/4 of this column plus 1/4 of the left (even output)
                // or right (odd output) neighbour; rounding alternates between the
                // two outputs
                const int lo = rv == 1 ? 1 : 8, hi = rv == 1 ? 2 : 7, sh = rv == 1 ? 2 : 4;
                int cur = colsum(0), next = colsum(1);
                size_t k = 0;
                o[k++] = (uint8_t)(rv == 1 ? cur : (cur * 4 + 8) >> 4);
                o[k++] = (uint8_t)((cur * 3 + next + hi) >> sh);
                int last = cur;
                cur = next;
                for (int x = 1; x < c.dw - 1; ++x) {
                    next = colsum(x + 1);
                    o[k++] = (uint8_t)((cur * 3 + last + lo) >> sh);
                    o[k++] = (uint8_t)((cur * 3 + next + hi) >> sh);
                    last = cur;
                    cur = next;
                }
                o[k++] = (uint8_t)((cur * 3 + last + lo) >> sh);
                o[k++] = (uint8_t)(rv == 1 ? cur : (cur * 4 + 7) >> 4);
            }
        return out;
    }

    void to_rgba(std::vector<uint8_t> &rgba) const {
        rgba.assign((size_t)width_ * height_ * 4, 255);
        std::vector<std::vector<uint8_t>> planes;
        std::vector<size_t> widths;
        for (const auto &c : comps_) {
            int ow = 0;
            planes.push_back(upsample(c, ow));
            widths.push_back((size_t)ow);
        }
        if (comps_.size() == 1) {
            for (uint32_t y = 0; y < height_; ++y)
                for (uint32_t x = 0; x < width_; ++x) {
                    uint8_t *o = &rgba[4 * ((size_t)y * width_ + x)];
                    o[0] = o[1] = o[2] = planes[0][y * widths[0] + x];
                }
            return;
        }
        // Adobe transform 0, or components named R, G, B without JFIF: already RGB
        const bool rgb = (adobe_ && adobe_transform_ == 0) ||
                         (!jfif_ && !adobe_ && comps_[0].id == 'R' && comps_[1].id == 'G' && comps_[2].id == 'B');
        // YCbCr -> RGB in 16-bit fixed point, FIX(c) = round(c * 2^16)
        constexpr int SB = 16;
        constexpr int64_t HALF = (int64_t)1 << (SB - 1);
        auto fix16 = [](double c) { return (int64_t)(c * (1 << SB) + 0.5); };
        int64_t cr_r[256], cb_b[256], cr_g[256], cb_g[256];
        for (int i = 0; i < 256; ++i) {
            const int64_t x = i - 128;
            cr_r[i] = (fix16(1.40200) * x + HALF) >> SB;
            cb_b[i] = (fix16(1.77200) * x + HALF) >> SB;
            cr_g[i] = -fix16(0.71414) * x;
            cb_g[i] = -fix16(0.34414) * x + HALF;
        }
        auto clamp = [](int64_t v) { return (uint8_t)std::min<int64_t>(255, std::max<int64_t>(0, v)); };
        for (uint32_t y = 0; y < height_; ++y)
            for (uint32_t x = 0; x < width_; ++x) {
                uint8_t *o = &rgba[4 * ((size_t)y * width_ + x)];
                const int Y = planes[0][y * widths[0] + x], cb = planes[1][y * widths[1] + x],
                          cr = planes[2][y * widths[2] + x];
                if (rgb) {
                    o[0] = (uint8_t)Y;
                    o[1] = (uint8_t)cb;
                    o[2] = (uint8_t)cr;
                    continue;
                }
                o[0] = clamp(Y + cr_r[cr]);
                o[1] = clamp(Y + ((cb_g[cb] + cr_g[cr]) >> SB));
                o[2] = clamp(Y + cb_b[cb]);
            }
    }
};

} // namespace

bool decodeJpegRGBA8(const std::vector<uint8_t> &file, uint32_t &width, uint32_t &height, std::vector<uint8_t> &rgba,
                     std::string &err) {
    Decoder d(file);
    return d.run(width, height, rgba, err);
}

} // namespace nori
