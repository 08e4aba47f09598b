/*
 * vork_hip.h - C ABI of the MI355X path-tracing hot path.
 *
 * The reference renders with hardware ray queries against a driver-built
 * BLAS/TLAS and a Slang megakernel.  This library replaces exactly that
 * part: the acceleration-structure build, closest-hit / any-hit traversal and
 * the path integrator with its film accumulation, as hand-written HIP kernels
 * for gfx950.  Everything above it (scene.json, profiles, CLI, session) stays
 * with the caller.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Unless a call takes VH_DEVICE_POINTERS,
 *    array arguments are host memory and are copied in (or out) before the
 *    call returns.  With VH_DEVICE_POINTERS they are device addresses on the
 *    library's device and stay owned by the caller.
 *  - Every call returns a status: VH_OK or one of the reference's failure
 *    kinds (crates/vork-render/src/failure.rs; exit codes used by
 *    crates/vork-cli/src/session.rs:101-104): 2 usage, 3 scene, 4 device,
 *    5 io.  vh_last_error() gives the message of the calling thread's last
 *    failure.
 *  - Objects are library-owned and freed by their vh_*_release call.  A BLAS,
 *    TLAS or scene is immutable after it is built and may be read by several
 *    threads at once.  Calls that take a stream are ordered on that stream
 *    (NULL = the device's own stream).
 *  - vh_render and vh_trace_* return once their work is done.  Their
 *    scratch (wavefront queues and path state, traversal overflow stacks)
 *    belongs to the stream: the device creates it on a stream's first call,
 *    grows it to the largest call and frees it with the device.  Calls on
 *    one stream wait for each other; calls on different streams run
 *    concurrently from several threads and their kernels overlap on the
 *    device.  A scene is not written by rendering it: renders of one scene
 *    on different streams run at once (an adaptive exposure in progress
 *    keeps its moments per scene and stream), and vh_render_get_stats reads
 *    the stats of the last render of the scene that finished.
 *    vh_scene_set_camera must not overlap a render of its scene.
 */
#ifndef VORK_HIP_H
#define VORK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VH_ABI_VERSION 2u

/* Status codes (crates/vork-render/src/failure.rs, session.rs:101-104). */
enum {
    VH_OK = 0,
    VH_ERR_USAGE = 2,
    VH_ERR_SCENE = 3,
    VH_ERR_DEVICE = 4,
    VH_ERR_IO = 5
};

/* Flags for calls that take array arguments. */
#define VH_DEVICE_POINTERS 1u

/* Ray-event masks (crates/vork-slang/src/registry.rs:186-190). */
#define VH_RAY_CAMERA 1u
#define VH_RAY_SHADOW 2u
#define VH_RAY_DIFFUSE 4u
#define VH_RAY_SPECULAR 8u
#define VH_RAY_TRANSMISSION 16u

/* Instance flags.  Materials default to double_sided=true
 * (crates/vork-scene/src/v2/material.rs:287-289,500), so culling is opt-in,
 * as disable_triangle_culling is in crates/vork-render/src/gpu_tlas.rs:179-183. */
#define VH_INSTANCE_CULL_BACK_FACES 1u

typedef struct vh_device vh_device;
typedef struct vh_blas vh_blas;
typedef struct vh_tlas vh_tlas;
typedef struct vh_scene vh_scene;

/* ---------------------------------------------------------------- device */

/* Opens HIP device `ordinal` with its own stream.  Every BLAS, TLAS and
 * scene built on the device keeps it open: after vh_device_release the device
 * is closed when the last of them is released, in whatever order they are.  Replaces the wgpu device
 * the reference opens in Renderer::new (crates/vork-render/src/lib.rs:726). */
int vh_device_create(int ordinal, vh_device** out);
void vh_device_release(vh_device* device);
/* Blocks until all work the library queued on the device has finished. */
int vh_device_synchronize(vh_device* device);
/* Copies the calling thread's last error message (NUL-terminated, truncated to
 * `length`).  Returns the status of that failure, VH_OK if there was none. */
int vh_last_error(char* message, size_t length);
uint32_t vh_abi_version(void);

/* ---------------------------------------------------------- acceleration */

/* One triangle mesh in object space.  POSITION is read at `position_stride`
 * bytes (the reference uses the 80-byte Vertex with POSITION at offset 0,
 * crates/vork-render/src/lib.rs:189-199); indices are u32 triples
 * (crates/vork-render/src/rt.rs:644-662).  `normals` (optional) are the
 * shading normals at `normal_stride`; `triangle_materials` (optional) give one
 * material slot per triangle, added to the instance's material base. */
typedef struct vh_geometry {
    const float* positions;
    uint32_t position_stride;
    uint32_t vertex_count;
    const uint32_t* indices;
    uint32_t triangle_count;
    uint32_t normal_stride;
    const float* normals;
    const uint32_t* triangle_materials;
} vh_geometry;

typedef struct vh_blas_info {
    uint32_t triangle_count;
    uint32_t node_count;      /* 8-wide nodes */
    uint32_t binary_node_count;
    uint32_t max_depth;       /* wide levels */
    double sah_cost;          /* binary-tree SAH cost: sum over nodes and leaves of area / root area */
    double build_ms;          /* device time of the build kernels */
    uint64_t device_bytes;    /* nodes + triangles + shading attributes */
} vh_blas_info;

/* Builds one BLAS on the device: Morton order, binary tree by parallel
 * locally-ordered clustering, collapse to 8-wide nodes.  Replaces rt::Rays::new
 * (crates/vork-render/src/rt.rs:1097-1289) and the driver build it requests
 * with BUILD_FLAGS=PREFER_FAST_TRACE (rt.rs:9-10).  `flags` takes
 * VH_DEVICE_POINTERS. */
int vh_blas_build(vh_device* device, const vh_geometry* geometry, uint32_t flags, void* stream, vh_blas** out);
int vh_blas_get_info(const vh_blas* blas, vh_blas_info* info);
void vh_blas_release(vh_blas* blas);

/* One TLAS instance: object-to-world 3x4 rows (row-major [R|t]), the BLAS
 * slot, the 24-bit custom index the reference stores as the object row
 * (gpu_tlas.rs:186-190, rt.rs:1116), the material base, an 8-bit visibility
 * mask and VH_INSTANCE_* flags.  A non-finite or singular transform gets mask
 * 0, as the placement kernel in gpu_tlas.rs:70-118 writes. */
typedef struct vh_instance {
    float transform[12];
    uint32_t blas;
    uint32_t custom_index;
    uint32_t material;
    uint8_t mask;
    uint8_t flags;
    uint16_t reserved;
} vh_instance;

#define VH_TLAS_REBUILD 0u
#define VH_TLAS_REFIT 1u

/* Builds the instance hierarchy over `blases`; VH_TLAS_REFIT keeps the
 * handle and instance order and rebuilds the hierarchy from the moved
 * instances' boxes.  Replaces FrameRays::new / rt::encode / rebuild_tlas
 * (rt.rs:1339-1385,1387-1463,1465) and the descriptor placement kernel
 * (gpu_tlas.rs:70-118).  On VH_TLAS_REFIT `*inout` must be a TLAS built over
 * the same BLAS list and instance count.  `flags` takes VH_DEVICE_POINTERS
 * for `instances`. */
int vh_tlas_build(vh_device* device, vh_blas* const* blases, uint32_t blas_count,
                  const vh_instance* instances, uint32_t instance_count,
                  uint32_t mode, uint32_t flags, void* stream, vh_tlas** inout);
void vh_tlas_release(vh_tlas* tlas);

/* ------------------------------------------------------------------ rays */

/* Ray and hit records of the trace calls.  Hits satisfy tmin < t < tmax.
 * Equal-t candidates resolve to the lowest (instance, primitive). */
typedef struct vh_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} vh_ray;

typedef struct vh_hit {
    float t;              /* < 0: no hit */
    uint32_t instance;    /* TLAS instance index */
    uint32_t primitive;   /* triangle index in its BLAS */
    uint32_t custom_index;
    float u, v;           /* barycentrics of corners 1 and 2 */
} vh_hit;

/* Closest hit per ray; replaces the RayQuery in trace
 * (crates/vork-render/shaders/path.slang:140-200).  Instances whose mask
 * shares no bit with `mask` are skipped.  `flags` takes VH_DEVICE_POINTERS
 * (then `rays` must be 16-byte aligned). */
int vh_trace_closest(const vh_tlas* tlas, const vh_ray* rays, uint32_t count, uint32_t mask,
                     uint32_t flags, void* stream, vh_hit* hits);
/* Any hit per ray (occluded[i] = 1 or 0); replaces the first-hit query of
 * straight_shadow_transmittance (path.slang:630-706). */
int vh_trace_any(const vh_tlas* tlas, const vh_ray* rays, uint32_t count, uint32_t mask,
                 uint32_t flags, void* stream, uint8_t* occluded);

/* vh_trace_closest for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * no counterpart in the reference.  `rays` (16-byte aligned) and `hits`
 * (4-byte aligned) are device memory on the TLAS's device; at most 2^31 rays
 * per call, else VH_ERR_USAGE.  Gives the hits vh_trace_closest gives, with
 * the same kernel, and allocates nothing: its counters and traversal
 * overflow stacks live in a static device array of the library (64 MB per
 * GPU, looked up at vh_device_create, held until the process exits; at most
 * 32768 threads of it).  Every call on a GPU uses that GPU's one array, so
 * the library serializes these calls per GPU and each returns once its work
 * is done; they cannot overlap one another. */
int vh_trace_closest_static(const vh_tlas* tlas, const vh_ray* rays, uint32_t count, uint32_t mask, void* stream,
                            vh_hit* hits);

/* Per-ray traversal work of the closest-hit kernel, four uint32 per ray:
 * [nodes visited, triangles tested, deepest traversal-stack entry, 0].  Used
 * to state algorithmic bytes and to show how deep traversals go (entries
 * beyond the 12 a lane keeps in LDS spill to global memory). */
int vh_trace_closest_counts(const vh_tlas* tlas, const vh_ray* rays, uint32_t count, uint32_t mask,
                            uint32_t flags, void* stream, uint32_t* counts);

/* --------------------------------------------------------------- sampler */

/* The sampler draws the path kernels make, for checking them bit for bit
 * against a CPU restatement, as the reference checks its compiled shader in
 * shader_matches_cpu_reference (crates/vork-render/src/path_sampler.rs:262-318).
 * `queries` holds four uint32 per query: px, py, sample index, group.  For a
 * group, `bits` receives the group's four scrambled 32-bit Sobol values of
 * path_sampler(pixel, index, samples, seed) (path_sampler.slang:54-93; a
 * kernel's float is ((bits >> 9) + 0.5) / 2^23, path_sampler.slang:70).  For
 * group VH_SAMPLER_PATH_RANDOM it receives the white-noise stream instead:
 * screen_ray_hash(px, py, index, seed) (vork_sampling.slang:8-15) and the
 * states after 1, 2 and 3 path_random draws (path_sampler.slang:72).  With
 * VH_DEVICE_POINTERS both arrays are 16-byte aligned device memory. */
#define VH_SAMPLER_PATH_RANDOM 0xffffffffu
int vh_sampler_bits(vh_device* device, const uint32_t* queries, uint32_t count, uint32_t samples, uint32_t seed,
                    uint32_t flags, void* stream, uint32_t* bits);

/* ----------------------------------------------------------------- scene */

/* OpenPBR closure subset the path kernel shades: base colour/metalness,
 * specular weight/roughness/IOR (libraries/std/openpbr_core.slang,
 * crates/vork-render/shaders/bsdf.slang). */
typedef struct vh_material {
    float base_color[3];
    float base_metalness;
    float specular_weight;
    float specular_roughness;
    float specular_ior;
    float reserved;
} vh_material;

/* Rectangle area light (kind 3 of emitter.slang:75-170): centre, full edge
 * vectors u and v, emitting side `normal` (perpendicular to u and v, else
 * vh_scene_create fails with VH_ERR_SCENE), and `radiance` = colour x energy,
 * packed as crates/vork-render/src/lights.rs:55-140 packs it. */
typedef struct vh_rect_light {
    float position[3];
    float reserved0;
    float u[3];
    float reserved1;
    float v[3];
    float reserved2;
    float normal[3];
    float reserved3;
    float radiance[3];
    float reserved4;
} vh_rect_light;

/* Pinhole camera (libraries/std/camera.slang:25-40 calibration): eye, unit
 * forward/right/up, tangents of the half field of view, clip range. */
typedef struct vh_camera {
    float eye[3];
    float tan_half_x;
    float forward[3];
    float tan_half_y;
    float right[3];
    float clip_start;
    float up[3];
    float clip_end;
} vh_camera;

#define VH_SKY_CONSTANT 0u
#define VH_SKY_PROCEDURAL 1u

/* Environment seen by escaping rays.  CONSTANT: `color` in every direction.
 * PROCEDURAL: horizon `color`, zenith `zenith`, ground `ground`, plus a soft
 * sun lobe of `sun_radiance` around `sun_direction` with exponent
 * `sun_exponent`. */
typedef struct vh_sky {
    uint32_t kind;
    float color[3];
    float zenith[3];
    float sun_exponent;
    float ground[3];
    float reserved;
    float sun_direction[3];
    float reserved2;
    float sun_radiance[3];
    float reserved3;
} vh_sky;

typedef struct vh_scene_desc {
    const vh_material* materials;
    uint32_t material_count;
    const vh_rect_light* lights;
    uint32_t light_count;
    vh_camera camera;
    vh_sky sky;
} vh_scene_desc;

/* Binds a TLAS (which must outlive the scene) to its materials, lights, camera
 * and sky: the resident part of InstalledScene (crates/vork-render/src/
 * install.rs:401) that the path pass reads. */
int vh_scene_create(vh_device* device, const vh_tlas* tlas, const vh_scene_desc* desc, vh_scene** out);
/* Replaces the scene's camera for the renders that start after it returns
 * (a camera that moves over one installed scene);
 * it must not overlap a render of the scene. */
int vh_scene_set_camera(vh_scene* scene, const vh_camera* camera);
void vh_scene_release(vh_scene* scene);

/* Transport policy of crates/vork-render/src/passes/path.rs:571-589 with the
 * constants of crates/vork-render/src/constants.rs:256-267. */
typedef struct vh_render_params {
    uint32_t width, height;
    uint32_t samples;        /* exposure sample count (sampler block size) */
    uint32_t sample_begin;   /* first exposure sample rendered by this call */
    uint32_t sample_end;     /* one past the last */
    uint32_t seed;
    uint32_t max_bounces;    /* bounces.x */
    uint32_t diffuse_bounces;
    uint32_t glossy_bounces;
    uint32_t caustics;       /* limits.w bits: 1 glossy, 2 transmission after diffuse */
    float clamp_direct;      /* 0 = off */
    float clamp_indirect;
    float light_threshold;
    uint32_t batch_samples;  /* exposure samples per wavefront batch, 0 = auto: up to 2^27 paths (about
                                262 B each in the stream's scratch), halved while device memory is short */
    uint32_t flags;          /* VH_RENDER_* */
    uint32_t adaptive_min_samples;  /* sampling.w: 0 = max(16, sqrt(samples)) */
    float adaptive_threshold;       /* adaptive.x: 0 = 0.01 */
    uint32_t reserved[2];
} vh_render_params;

/* vh_render_params.flags: record HIP events around every stage launch (their
 * totals land in vh_render_stats), and count the nodes and triangles the
 * closest-hit stage visits (slower; for stating algorithmic bytes).
 *
 * VH_RENDER_ADAPTIVE is the adaptive path_pixel entry (path.slang:1792-1838,
 * 1964-1966; on by default, constants.rs:264-266): per-pixel running mean and
 * variance (48 B per pixel, kept in the scene between calls); from sample
 * index `minimum` a pixel is settled when every RGBA channel's standard error
 * is at most threshold * max(|mean|, 0.01), and it stops for the rest of the
 * exposure once it and its whole 3x3 neighbourhood were settled on the
 * previous sample.  A stopped pixel adds its mean for each remaining sample,
 * so the film still holds the sum of `samples` values.  A call with
 * sample_begin > 0 must continue the exposure the previous adaptive call on
 * this scene rendered (same size, seed and sample count), else VH_ERR_USAGE. */
#define VH_RENDER_PROFILE 1u
#define VH_RENDER_COUNT_VISITS 2u
#define VH_RENDER_ADAPTIVE 4u
/* Shading order (the film is the same either way): SORTED sorts each bounce's
 * hits into queues by closure class (diffuse, dielectric GGX, metal GGX) so a
 * wave shades one closure; ARRIVAL shades the queue as traced.  Neither: the
 * library's tuned default. */
#define VH_RENDER_SHADE_SORTED 8u
#define VH_RENDER_SHADE_ARRIVAL 16u

/* Fills `params` with the reference's production defaults for a
 * width x height x samples render (adaptive sampling on). */
void vh_render_params_default(vh_render_params* params, uint32_t width, uint32_t height, uint32_t samples);

/* Renders exposure samples [sample_begin, sample_end) of every pixel and ADDS
 * their sum to `film` (RGBA32F, width*height*4 floats, row-major from the top
 * row): path_pixel + integrate (path.slang:1313-1969) and the resolve
 * accumulation (resolve.slang:228-238).  `film` is a device pointer when
 * `flags` has VH_DEVICE_POINTERS, else host memory that is read, added to and
 * written back. */
int vh_render(vh_scene* scene, const vh_render_params* params, uint32_t flags, void* stream, float* film);

/* vh_render that also ADDS the denoiser guides of every sample's camera hit
 * to `guides` (width*height*8 floats, row-major like the film): per pixel
 * albedo rgb + coverage, then shading normal xyz + depth t/(t+1), 1 on a
 * miss.  These are path_guides (path.slang:1553-1575,1956-1961) with the
 * depth of field_depth (path.slang:1864,1883) in the normal's fourth
 * channel; a dropped sample adds none.  Replaces the guide writes the
 * denoise pass reads (crates/vork-render/src/denoise.rs). */
int vh_render_guides(vh_scene* scene, const vh_render_params* params, uint32_t flags, void* stream, float* film,
                     float* guides);

/* One job of vh_render_shared: a scene, the sample range to render (plain
 * sampling: VH_RENDER_ADAPTIVE is refused) and its film, plus optional guides,
 * as device memory that the job's samples are ADDED to, as vh_render_guides
 * adds them. */
typedef struct vh_render_job {
    vh_scene* scene;
    vh_render_params params;
    float* film;
    float* guides;
} vh_render_job;

/* Renders up to 16 jobs (scenes on one device, each scene in at most one job)
 * with their paths in one wavefront: every batch gives each job with samples
 * left a share of the path slots, the jobs' camera rays fill one extension
 * queue, their hits one shading queue and their shadow rays one shadow queue,
 * and every path traverses and shades with its own job's scene.  Each film
 * equals what vh_render_guides gives for the job alone.  This is the shared
 * wavefront queue of the serving loop; the reference serves one render at a
 * time (crates/vork-cli/src/session.rs:1-7).  A scene may appear in several
 * jobs.  Each job's scene stats get its paths (the last job's, for a scene
 * in several); rays and device time are those of the whole call. */
int vh_render_shared(const vh_render_job* jobs, uint32_t count, void* stream);

/* ---------------------------------------------------------------- denoiser */

/* The learned denoiser that replaces the OIDN pass of
 * crates/vork-render/src/denoise.rs (its guides are vh_render_guides'):
 * a 5-level U-Net of 3x3 convolutions (32, 64, 64, 96, 128 channels),
 * bf16 weights and activations on MFMA with fp32 accumulation.  `weights`
 * is the version-1 blob vork_amd/denoise.py exports ("VHDN", version,
 * layer count, then per layer cin, cout, bf16 weights [tap][cout][cin] and
 * fp32 biases); it is copied to the device. */
typedef struct vh_denoiser vh_denoiser;
int vh_denoiser_create(vh_device* device, const void* weights, size_t bytes, vh_denoiser** out);

/* Denoises one image: `film` (width*height*4 floats) and `guides`
 * (width*height*8) are the sums of `samples` samples from vh_render_guides;
 * `out` (width*height*4) receives the denoised mean radiance with alpha 1.
 * Pointers are device pointers with VH_DEVICE_POINTERS.  Calls on one
 * denoiser wait for each other (it owns its activation scratch). */
int vh_denoise(vh_denoiser* denoiser, uint32_t width, uint32_t height, float samples, const float* film,
               const float* guides, uint32_t flags, void* stream, float* out);
void vh_denoiser_release(vh_denoiser* denoiser);

/* Training of the denoiser with hand-written kernels (the reference trains
 * none: it ships OIDN's weights; config 4 trains this U-Net on pairs the path
 * tracer renders).  A trainer holds fp32 master weights, their Adam moments
 * and the bf16 copies the MFMA kernels read.  `weights` is a version-1 blob to
 * start from, or NULL for a seeded default initialisation (uniform in
 * +-1/sqrt(fan in)). */
typedef struct vh_denoiser_trainer vh_denoiser_trainer;
int vh_trainer_create(vh_device* device, const void* weights, size_t bytes, uint32_t seed,
                      vh_denoiser_trainer** out);
/* One step over `images` images of width x height (device pointers): film
 * and guide sums of `samples` samples as vh_render_guides writes them, and
 * target mean radiance (3 floats per pixel).  Runs the forward pass keeping
 * every activation, the loss (relative MSE of the radiance plus L1 of its
 * log1p, averaged over the batch, returned in *loss), the backward pass
 * (weight and bias gradients summed over the images) and, unless `flags` has
 * VH_TRAIN_NO_UPDATE, an Adam step with `learning_rate`.  Every launch runs
 * over all images of the batch (at most 2^25 pixels in all). */
#define VH_TRAIN_NO_UPDATE 1u
int vh_trainer_step(vh_denoiser_trainer* trainer, uint32_t images, uint32_t width, uint32_t height, float samples,
                    const float* films, const float* guides, const float* targets, float learning_rate,
                    uint32_t flags, void* stream, float* loss);
/* The last step's fp32 gradients in blob order (per layer [tap][cout][cin]
 * weights; biases); the counts must be the network's. */
int vh_trainer_gradients(vh_denoiser_trainer* trainer, float* weights, size_t weight_count, float* biases,
                         size_t bias_count);
/* Writes the current weights as a version-1 blob (*bytes = its size; with
 * `blob` NULL only the size). */
int vh_trainer_export(vh_denoiser_trainer* trainer, void* blob, size_t capacity, size_t* bytes);
void vh_trainer_release(vh_denoiser_trainer* trainer);

/* Kernel statistics of the last vh_render on this scene. */
typedef struct vh_render_stats {
    uint64_t paths;             /* traced paths: stopped adaptive pixels trace none */
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t dropped_samples;   /* non-finite samples zeroed */
    double device_ms;
    /* with VH_RENDER_PROFILE: summed device time of each stage and its launch count */
    double extend_ms, shade_ms, shadow_ms, other_ms;
    uint64_t extend_launches, shade_launches, shadow_launches;
    /* with VH_RENDER_COUNT_VISITS: wide nodes and triangles the closest-hit stage visited */
    uint64_t extend_node_visits, extend_triangle_tests;
    uint64_t shadow_node_visits, shadow_triangle_tests;
    /* with VH_RENDER_ADAPTIVE: samples of stopped pixels (their mean added, no path) */
    uint64_t stopped_samples;
    /* the camera bounce's part of the extend stage (the first launch of each batch): its device time and
     * launches with VH_RENDER_PROFILE, its rays, nodes and triangles with VH_RENDER_COUNT_VISITS */
    double extend_camera_ms;
    uint64_t extend_camera_launches, extend_camera_rays;
    uint64_t extend_camera_node_visits, extend_camera_triangle_tests;
} vh_render_stats;
int vh_render_get_stats(const vh_scene* scene, vh_render_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* VORK_HIP_H */
