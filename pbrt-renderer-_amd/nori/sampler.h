// Sampler plugin interface (include/sampler.h:20-88 of the reference) and the
// image block it is prepared for (include/block.h).
#pragma once
#include "object.h"

NORI_NAMESPACE_BEGIN

/// A rectangular block of the image (offset and size in pixels).
class ImageBlock {
public:
    ImageBlock(const Vector2i &offset, const Vector2i &size) : m_offset(offset), m_size(size) {}
    const Vector2i &getOffset() const { return m_offset; }
    const Vector2i &getSize() const { return m_size; }

private:
    Vector2i m_offset, m_size;
};

class Sampler : public NoriObject {
public:
    virtual ~Sampler() {}
    virtual std::unique_ptr<Sampler> clone() const = 0;
    virtual void prepare(const ImageBlock &block) = 0;
    virtual void generate() = 0;
    virtual void advance() = 0;
    virtual float next1D() = 0;
    virtual Point2f next2D() = 0;
    virtual size_t getSampleCount() const { return m_sampleCount; }
    EClassType getClassType() const override { return ESampler; }

protected:
    size_t m_sampleCount = 1;
};

NORI_NAMESPACE_END
