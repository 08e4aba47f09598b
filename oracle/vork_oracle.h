/*
 * vork_oracle.h - CPU oracle of the path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * estimator (voyager-so/vork, crates/vork-render/shaders/path.slang and its
 * imports) used as the checker for the HIP library.  Only tests/, the smoke
 * entry and the CPU-baseline leg of bench.py may load it; the product never
 * links, calls or falls back to it.
 *
 * Every function cites the reference file:line it follows.  Data formats are
 * the public ones of include/vork_hip.h.
 */
#ifndef VORK_ORACLE_H
#define VORK_ORACLE_H

#include <stdint.h>
#include "../include/vork_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ sampler */
uint32_t vo_pixel_hash(uint32_t seed);
void vo_pixel_hash_many(const uint32_t* in, uint32_t n, uint32_t* out);
uint32_t vo_screen_ray_hash(uint32_t px, uint32_t py, uint32_t index, uint32_t stream);
void vo_sobol(uint32_t index, uint32_t out[4]);
void vo_sobol_directions(uint32_t out[3 * 32]);
void vo_sample_group_bits(uint32_t px, uint32_t py, uint32_t index, uint32_t count, uint32_t stream,
                          uint32_t group, uint32_t out[4]);
void vo_sample_group(uint32_t px, uint32_t py, uint32_t index, uint32_t count, uint32_t stream,
                     uint32_t group, float out[4]);
/* Bits of `n` consecutive sample indices [0, n) (or `indices` when given). */
void vo_sample_group_bits_batch(uint32_t px, uint32_t py, uint32_t count, uint32_t stream, uint32_t group,
                                const uint32_t* indices, uint32_t n, uint32_t* out);
void vo_sample_group_bits_many(const uint32_t* px, const uint32_t* py, const uint32_t* index, uint32_t count,
                               uint32_t stream, const uint32_t* group, uint32_t n, uint32_t* out);
float vo_path_random(uint32_t* state);

/* ------------------------------------------------------- film filter */
#define VO_FILTER_INTERVALS 1024
void vo_filter_table(float table[VO_FILTER_INTERVALS + 1]);
double vo_filter_cdf(double x);
float vo_filter_offset(const float* table, float u);

/* ------------------------------------------------------ intersection */
/* Watertight crossing + Moller-Trumbore t; returns 1 on a hit with
 * tmin < t < tmax, writing t and the clamped barycentrics (u, v). */
int vo_intersect(const float a[3], const float b[3], const float c[3], const float o[3], const float d[3],
                 float tmin, float tmax, int cull_back, float* t, float* u, float* v);
int vo_volume_crosses(const float a[3], const float b[3], const float c[3], const float o[3], const float d[3]);

/* ------------------------------------------------------------- scene */
typedef struct vo_scene vo_scene;

typedef struct vo_mesh_desc {
    const float* positions;   /* xyz, tightly packed */
    const float* normals;     /* xyz or NULL */
    const uint32_t* indices;
    const uint32_t* triangle_materials; /* or NULL */
    uint32_t vertex_count;
    uint32_t triangle_count;
} vo_mesh_desc;

vo_scene* vo_scene_create(const vo_mesh_desc* meshes, uint32_t mesh_count, const vh_instance* instances,
                          uint32_t instance_count, const vh_scene_desc* desc);
void vo_scene_release(vo_scene* scene);

/* World-to-object rows of an object-to-world transform, in double precision
 * rounded to float; returns 0 for a singular or non-finite transform. */
int vo_invert_transform(const float m[12], float inv[12]);

void vo_trace_closest(const vo_scene* scene, const vh_ray* rays, uint32_t count, uint32_t mask, vh_hit* hits,
                      int brute_force);
void vo_trace_any(const vo_scene* scene, const vh_ray* rays, uint32_t count, uint32_t mask, uint8_t* occluded,
                  int brute_force);

/* ----------------------------------------------------------- shading */
/* GGX energy tables (environment.slang:14-120 bake, frame.slang:41-66 lookups). */
void vo_brdf_moments(float nv, float roughness, float out[4]);
float vo_dielectric_albedo(float nv, float roughness, float eta);
float vo_dielectric_directional_albedo(float no, float roughness, float eta, uint32_t samples);
void vo_bake_moments_texel(uint32_t column, uint32_t row, float out[4]);

/* BSDF of one closure at a surface with shading normal `n` (also the
 * smooth receiver) and face normal `face`, for view direction `view`.
 * sample: out = [dir xyz, weight rgb, pdf, event]. */
void vo_bsdf_sample(const vh_material* m, const float n[3], const float face[3], const float view[3],
                    const float random[3], float out[8]);
void vo_bsdf_eval(const vh_material* m, const float n[3], const float face[3], const float view[3],
                  const float dir[3], uint32_t mask, float out[4]);
float vo_modulated_ior(float eta, float weight);
float vo_fresnel(float c, float eta);

/* Light selection (light_tree.slang:40-84) over the scene's lights. */
uint32_t vo_light_pick(const vo_scene* scene, const float point[3], float ux, float uy, float* probability);
float vo_light_probability(const vo_scene* scene, uint32_t light, const float point[3]);
/* Rectangle emitter sample / evaluation (emitter.slang:88-150):
 * out = [dir xyz, distance, radiance rgb, pdf]. */
void vo_rect_sample(const vh_rect_light* l, const float world[3], const float u[2], float out[8]);
void vo_rect_eval(const vh_rect_light* l, const float world[3], const float dir[3], float out[8]);

/* -------------------------------------------------------- integrator */
/* Radiance of one exposure sample of one pixel (path_pixel + integrate). */
void vo_sample_radiance(const vo_scene* scene, const vh_render_params* params, uint32_t px, uint32_t py,
                        uint32_t index, float out[4]);
/* Adds samples [sample_begin, sample_end) of every pixel to the films:
 * `film32` (RGBA, float, summed in sample order as the GPU film is) and
 * `film64` (RGBA, double); either may be NULL.  threads <= 0: all cores. */
void vo_render(const vo_scene* scene, const vh_render_params* params, float* film32, double* film64, int threads);
void vo_render_guides(const vo_scene* scene, const vh_render_params* params, float* film32, double* film64,
                      float* guides32, int threads);
/* Camera ray of one sample: out = [origin xyz, tmax, dir xyz, 0]. */
void vo_camera_ray(const vo_scene* scene, const vh_render_params* params, uint32_t px, uint32_t py,
                   uint32_t index, float out[8]);
void vo_camera_rays(const vo_scene* scene, const vh_render_params* params, const uint32_t* pixels,
                    const uint32_t* indices, uint32_t n, vh_ray* out);
/* Film sums over [sample_begin, sample_end) for a list of row-major pixel
 * indices (4 floats each), and single sample values per (pixel, index). */
void vo_render_pixels(const vo_scene* scene, const vh_render_params* params, const uint32_t* pixels, uint32_t n,
                      float* out, int threads);
void vo_sample_radiance_many(const vo_scene* scene, const vh_render_params* params, const uint32_t* pixels,
                             const uint32_t* indices, uint32_t n, float* out, int threads);
/* Path statistics of the last vo_render call: [paths, extension rays, shadow rays, dropped]. */
void vo_render_stats(const vo_scene* scene, uint64_t out[4]);

#ifdef __cplusplus
}
#endif

#endif
