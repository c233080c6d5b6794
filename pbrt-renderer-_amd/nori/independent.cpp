// "independent" sampler (src/cpu/independent.cpp:15-67 semantics): pcg32
// uniforms, sampleCount samples per pixel; prepare() seeds from the block
// offset.  The GPU path draws its own per-pixel pcg32 streams
// (seed(render_seed ^ pixel, PCG32_DEFAULT_STREAM), sample s at s << 20).
#include "sampler.h"

#include <cstring>

NORI_NAMESPACE_BEGIN

namespace {
constexpr uint64_t PCG_MULT = 0x5851f42d4c957f2dULL;
constexpr uint64_t PCG_DEFAULT_STATE = 0x853c49e6748fea9bULL;
constexpr uint64_t PCG_DEFAULT_STREAM = 0xda3e39cb94b95bdbULL;

struct Pcg32 {
    uint64_t state = PCG_DEFAULT_STATE, inc = PCG_DEFAULT_STREAM;
    void seed(uint64_t initstate, uint64_t initseq = 1) {
        state = 0;
        inc = (initseq << 1u) | 1u;
        nextUInt();
        state += initstate;
        nextUInt();
    }
    uint32_t nextUInt() {
        uint64_t old = state;
        state = old * PCG_MULT + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
    }
    float nextFloat() {
        uint32_t u = (nextUInt() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &u, 4);
        return f - 1.0f;
    }
};
} // namespace

class Independent : public Sampler {
public:
    explicit Independent(const PropertyList &p) { m_sampleCount = (size_t)p.getInteger("sampleCount", 1); }
    std::unique_ptr<Sampler> clone() const override {
        auto s = std::make_unique<Independent>(PropertyList());
        s->m_sampleCount = m_sampleCount;
        s->m_random = m_random;
        return s;
    }
    void prepare(const ImageBlock &block) override {
        m_random.seed((uint64_t)block.getOffset().x(), (uint64_t)block.getOffset().y());
    }
    void generate() override {}
    void advance() override {}
    float next1D() override { return m_random.nextFloat(); }
    Point2f next2D() override {
        float a = m_random.nextFloat();
        float b = m_random.nextFloat();
        return Point2f(a, b);
    }
    /// 64 bits from the stream (seeds of per-ray GPU streams).
    uint64_t nextSeed() { return ((uint64_t)m_random.nextUInt() << 32) | m_random.nextUInt(); }
    std::string toString() const override { return format("Independent[sampleCount=%i]", m_sampleCount); }

private:
    Pcg32 m_random;
};

NORI_REGISTER_CLASS(Independent, "independent");

NORI_NAMESPACE_END
