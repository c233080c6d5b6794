/* libnori_host.so: the Nori plugin surface (NoriObject / PropertyList /
 * NORI_REGISTER_CLASS, the XML scene format, OBJ meshes, BSDF / emitter /
 * camera / sampler / medium / integrator plugins) in C++ on top of
 * libnori_gpu.so, with a C entry point set for loaders that bind C.
 *
 * It replaces the reference's scene loading and rendering entry:
 *   loadFromXML (src/cpu/parser.cpp:17-337), Scene::activate
 *   (src/cpu/scene.cpp), and the render of the loaded scene that the
 *   reference's GPU application performs (src/gpu/DXRApp*.cpp).
 *
 * Errors: functions return NORI_OK (0) or a negative code and copy a message
 * (NoriException text, with the file and line/column for parse errors) into
 * err[0..errlen).  A scene handle is not thread-safe. */
#ifndef NORI_HOST_H
#define NORI_HOST_H

#include "nori_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nori_host_scene nori_host_scene;

typedef struct nori_xml_info {
    uint32_t width, height;   /* camera output size                               */
    uint32_t spp;             /* sampler sampleCount                              */
    uint32_t meshes, triangles, vertices, emitters;
    uint32_t has_medium;      /* a <medium> was given                             */
    uint32_t envmap_loaded;   /* 1: the envmap file was read; 0: flat 0.5 grey    */
    char integrator[64];      /* toString() of the integrator (or "")             */
} nori_xml_info;

/* Parse and activate a scene file (relative paths resolve against its directory). */
int nori_xml_load(const char *filename, nori_host_scene **out, char *err, int errlen);
/* The same for a document in memory; dir: where relative paths resolve. */
int nori_xml_load_string(const char *text, const char *dir, nori_host_scene **out, char *err, int errlen);
void nori_xml_free(nori_host_scene *s);
int nori_xml_info_get(const nori_host_scene *s, nori_xml_info *info);
/* The flat scene as the GPU path receives it (pointers valid until nori_xml_free)
 * and the GPU camera of the scene's camera. */
int nori_xml_desc(const nori_host_scene *s, nori_scene_desc *desc, nori_camera *camera);
/* Integrator::preprocess + the batch render of the camera view at the sampler's
 * sample count: rgb receives width * height * 3 floats (row 0 at the top). */
int nori_xml_render(nori_host_scene *s, float *rgb, char *err, int errlen);
/* Integrator::Li through the plugin: n rays (o, mint, d, maxt per ray), one
 * Independent-sampler draw sequence seeded with `seed`; rgb: n * 3 floats. */
int nori_xml_li(nori_host_scene *s, const nori_ray *rays, int n, uint64_t seed, float *rgb, char *err, int errlen);
/* Write an RGB float image as .pfm or .exr (uncompressed). */
int nori_image_write(const char *filename, const float *rgb, uint32_t width, uint32_t height, char *err, int errlen);
/* Read an RGB float image (.hdr RGBE or .pfm); *rgb is malloc'ed, release with nori_image_free. */
int nori_image_read(const char *filename, float **rgb, uint32_t *width, uint32_t *height, char *err, int errlen);
void nori_image_free(float *rgb);
/* Linear-HDR comparison of two renders of the same size (the metrics of the
 * reference's validation_tests/compare_exr.py:74-118): per-channel means of
 * both images, their ratio (ours / ref; +inf where ref's mean is 1e-10 or
 * less), mean and max |difference|; linear RMSE over all channels; relMSE =
 * mean((ours - ref)^2 / (ref^2 + 1e-2)); PSNR after the Reinhard display curve
 * (x 2^exposure_ev, max(x, 0), x / (1 + x), ^(1/2.2)), +inf when that MSE is
 * 1e-10 or less.  Images: width * height * 3 floats. */
typedef struct nori_image_metrics {
    double mean_ours[3], mean_ref[3], ratio[3], mean_abs_diff[3], max_abs_diff[3];
    double rmse, rel_mse, psnr_tonemapped;
} nori_image_metrics;
int nori_image_compare(const float *ours, const float *ref, uint32_t width, uint32_t height, float exposure_ev,
                       nori_image_metrics *out);
/* White-furnace statistics of validation_tests/disney/analyze_furnace.py:
 * the pixels whose centre ray (vertical fov_deg, camera at (0, 0, cam_z)
 * looking down -z) meets the sphere of `radius` at the origin, their
 * per-channel mean and mean Rec.709 luminance; *count = their number. */
int nori_furnace_stats(const float *rgb, uint32_t width, uint32_t height, float fov_deg, float cam_z, float radius,
                       double mean[3], double *lum_mean, uint32_t *count);
/* Comma-separated names of the registered plugin classes. */
int nori_plugin_names(char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif
