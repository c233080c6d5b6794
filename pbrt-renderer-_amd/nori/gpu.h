// Glue between the plugin surface and the C ABI of libnori_gpu.so.
#pragma once
#include "../../include/nori_gpu.h"
#include "bsdf.h"

NORI_NAMESPACE_BEGIN

/// The process-wide GPU context the plugins share (device 0, created on first
/// use).  Throws NoriException when no GPU is present: there is no CPU path.
nori_ctx *gpuContext();

/// Throws NoriException(what: last error) unless rc == NORI_OK.
void gpuCheck(int rc, const char *what, nori_ctx *ctx = nullptr);

/// The 144-byte material record of the GPU path for a BSDF's parameters.
nori_material toGPUMaterial(const BSDFGPUData &d);

/// One BSDF query on the GPU through nori_gpu_bsdf_batch (op 0 eval, 1 pdf,
/// 2 sample with u[4]); returns rgb and writes pdf and, for op 2, wo.
Color3f gpuBsdfQuery(int op, const BSDFGPUData &d, const Vector3f &wi, const Vector3f &wo, const float *u4, float h,
                     float &pdf, Vector3f *wo_out);

/// Four uniforms from a 2D sample: the two given plus two derived from them by
/// a PCG hash (the per-query plugin API carries 2 dimensions; Disney needs 3
/// and hair 4).
void expandSample(const Point2f &s, float u[4]);

NORI_NAMESPACE_END
