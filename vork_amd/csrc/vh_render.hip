// Wavefront path integrator: path_pixel + integrate of
// crates/vork-render/shaders/path.slang:1313-1969 split into stages that each
// run over a queue of live paths.
//
//   camera  -> one path per (pixel, exposure sample) of the batch
//   extend  -> closest hit of every live path's ray
//   shade   -> sky on a miss, else the closure, a BSDF sample, both direct
//              estimators of next_event (their visibility deferred as shadow
//              rays) and the continuation with caps and roulette
//   shadow  -> any-hit of the pending shadow rays; adds the clamped sum
//   film    -> per pixel, the batch's samples summed in sample order into the
//              RGBA32F film (one thread per pixel, no atomics)
//
// Every path owns its state slot, so results do not depend on queue order.
//
// Adaptive sampling (path.slang:1792-1838,1964-1966) keeps a moment per
// pixel.  No pixel can stop before sample index minimum + 1 (it needs its own
// test and its neighbourhood's of the previous sample, both from `minimum`
// on), so samples up to `minimum` are batched like plain ones and the film
// pass records their tests; each later sample index is a batch of its own,
// preceded by the decide pass, and its camera pass emits paths only for the
// pixels that still run.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <tuple>
#include <vector>

#include "vh_internal.h"
#include "vh_shade.h"
#include "vh_traverse.h"

namespace vh {
namespace {

constexpr int kBlock = VH_BLOCK;

// flags word of a path
constexpr uint32_t kEventShift = 0, kDiffuseShift = 8, kGlossyShift = 12, kBounceShift = 16;
constexpr uint32_t kNonFinite = 1u << 30;

struct Paths {
    float4* thr;      // throughput
    float4* rad;      // radiance
    uint4* info;      // pixel, sample index, rng, flags
    float4* hit;      // t, u, v, leaf-order triangle bits
    uint32_t* hit_inst;
    float4* pend;     // per pending next event q: throughput x estimator 0 (rgb), clamp limit
    float4* pend1;    //   throughput x estimator 1 (rgb), path index bits
    uint8_t* vis;     // per shadow ray id 2q + k: 1 = occluded
    float4* guide_a;  // denoiser guides of the camera hit: albedo, 1 (coverage)
    float4* guide_n;  //   shading normal, depth t / (t + 1)
};

// Ray queues hold 32-byte records, float4 pairs: [2i] origin and tmax,
// [2i+1] direction and an id word.  Extension rays carry path | code << 30,
// the code naming the event of the previous vertex (camera, diffuse,
// specular, transmission), which is the ray's visibility mask; shadow rays
// carry 2 * event slot + estimator.  Paths of one batch are limited to 2^30.
constexpr uint32_t kPathBits = 30, kPathMask = (1u << kPathBits) - 1u;
// path state per path: 8 float4 (throughput, radiance, info, hit, two pending
// contributions, two guides), the hit instance and two visibility bytes
constexpr size_t kPathStateBytes = sizeof(float4) * 8 + sizeof(uint32_t) + 2;
__device__ __forceinline__ uint32_t event_code(uint32_t event) {
    return event == VH_RAY_DIFFUSE ? 1u : event == VH_RAY_SPECULAR ? 2u : event == VH_RAY_TRANSMISSION ? 3u : 0u;
}
__device__ __forceinline__ uint32_t code_mask(uint32_t code) {
    return code == 0u ? VH_RAY_CAMERA : code == 1u ? VH_RAY_DIFFUSE : code == 2u ? VH_RAY_SPECULAR : VH_RAY_TRANSMISSION;
}
__device__ __forceinline__ void put_ray(float4* q, uint32_t i, f3 o, float tmax, f3 d, uint32_t id) {
    q[2 * i] = make_float4(o.x, o.y, o.z, tmax);
    q[2 * i + 1] = make_float4(d.x, d.y, d.z, u2f(id));
}

struct Consts {
    uint32_t width, height, samples, seed;
    uint32_t max_bounces, diffuse_bounces, glossy_bounces, caustics;
    float clamp_direct, clamp_indirect, light_threshold;
    uint32_t light_count, light_nodes, material_count;
    uint32_t batch_first, batch_samples, pixels;
    uint32_t group;             // samples of one pixel in consecutive path slots (see path_slot)
    uint32_t minimum, decide;   // adaptive: first tested sample index; batch follows a decide pass
    float threshold;
    uint32_t guides;            // write the denoiser guides
    vh_camera camera;
    vh_sky sky;
};

// Pixel slot t -> pixel: column-major inside bands of 8 rows, so that
// consecutive slots are neighbouring pixels (64 of them an 8x8 tile).  A
// bijection on [0, width*height) for any image size (the last band may be
// shorter).
__device__ __forceinline__ uint32_t slot_to_pixel(uint32_t t, uint32_t w, uint32_t h) {
    const uint32_t band = t / (8u * w), r = t % (8u * w);
    const uint32_t rows = min(8u, h - band * 8u);
    const uint32_t x = r / rows, y = band * 8u + r % rows;
    return y * w + x;
}

// Samples of one pixel that share consecutive path slots: the largest power of
// two dividing the batch's sample count, at most VH_WAVE_SAMPLES.  A wave of
// camera rays then holds the jittered samples of a few neighbouring pixels,
// which walk nearly the same BVH nodes, so its node fetches touch few cache
// lines and its lanes stay in step.
#ifndef VH_WAVE_SAMPLES
#define VH_WAVE_SAMPLES 64
#endif
inline uint32_t sample_group(uint32_t batch_samples) {
    const uint32_t g = batch_samples & (0u - batch_samples);
    return g < VH_WAVE_SAMPLES ? g : VH_WAVE_SAMPLES;
}
// Path slot of sample s (of the batch) of pixel slot t: groups of `group`
// consecutive slots hold `group` samples of one pixel slot, and a whole
// group of samples covers every pixel slot before the next one starts.
__device__ __forceinline__ uint32_t path_slot(uint32_t s, uint32_t t, uint32_t group, uint32_t pixels) {
    return (s / group) * group * pixels + t * group + s % group;
}
__device__ __forceinline__ void slot_sample(uint32_t p, uint32_t group, uint32_t pixels, uint32_t& s, uint32_t& t) {
    const uint32_t span = group * pixels, r = p % span;
    s = (p / span) * group + r % group;
    t = r / group;
}

// Per-pixel adaptive moment (PathMoment, path.slang:57): running mean and
// variance sum of the RGBA sample values, the settled test of the last two
// sample parities, and the stopped flag.
struct Moment {
    float4 mean, variance;
    uint32_t settled[2], stopped, pad;
};
static_assert(sizeof(Moment) == 48, "moment layout");

// path.slang:1812-1816: every channel's standard error within
// threshold * max(|mean|, 0.01).  Only called with index >= 1.
__device__ __forceinline__ bool moment_settled(const Moment& m, uint32_t index, float threshold) {
    const float a = (float)max(index - 1u, 1u), b = (float)index;
    const float v[4] = {m.variance.x, m.variance.y, m.variance.z, m.variance.w};
    const float mu[4] = {m.mean.x, m.mean.y, m.mean.z, m.mean.w};
    bool ok = true;
    for (int k = 0; k < 4; k++) ok = ok && sqrtf(fmaxf(v[k] / a / b, 0.f)) <= threshold * fmaxf(fabsf(mu[k]), .01f);
    return ok;
}

// path.slang:1964: Welford update with sample `index`, each operation rounded
// as written.
__device__ __forceinline__ void moment_add(Moment& m, float4 v, uint32_t index) {
#pragma clang fp contract(off)
    const float n = (float)(index + 1u);
    float4 d = make_float4(v.x - m.mean.x, v.y - m.mean.y, v.z - m.mean.z, v.w - m.mean.w);
    m.mean = make_float4(m.mean.x + d.x / n, m.mean.y + d.y / n, m.mean.z + d.z / n, m.mean.w + d.w / n);
    m.variance = make_float4(m.variance.x + d.x * (v.x - m.mean.x), m.variance.y + d.y * (v.y - m.mean.y),
                             m.variance.z + d.z * (v.z - m.mean.z), m.variance.w + d.w * (v.w - m.mean.w));
}

struct Scene {
    TlasView tv;
    const vh_material* materials;
    const vh_rect_light* lights;
    const LightNode* tree;
    const float* filter;
    Luts luts;
};

// counters: [0] live count in, [2] estimator-b shadow rays queued,
// [3] dropped non-finite contributions, [4..5] extension rays (u64),
// [6..7] shadow rays (u64), [8] errors, [9] non-finite samples zeroed,
// [10..11] samples of stopped adaptive pixels (u64), [16..19] extension node /
// triangle visits of the later bounces, [20..23] shadow node / triangle
// visits (u64), [24] extension ray claims, [25] shadow ray claims, [28] live
// count out and [29] pending next events, one u64 so that one atomic reserves
// both, [32..35] the camera bounce's node / triangle visits and [36..37] its
// rays (u64)
constexpr int kExtendClaims = 24, kShadowClaims = 25, kLiveOut = 28, kPending = 29, kShadowB = 2;
constexpr int kCameraVisits = 32, kCameraRays = 36, kCounterWords = 64;
// [40..42] hits per closure class in the sorted shading queues
constexpr int kClassCount = 40;

// Material-sorted shading: a classify pass adds the sky for every miss and
// appends each hit's queue entry to the queue of its closure class (diffuse
// only, dielectric GGX, metal GGX), so each wave of the shading kernel runs
// one closure's code.  Otherwise the shading kernel takes the queue in
// arrival order and walks past misses.  This is the default when a render's
// flags choose neither VH_RENDER_SHADE_SORTED nor VH_RENDER_SHADE_ARRIVAL.
#ifndef VH_SHADE_SORT
#define VH_SHADE_SORT 0
#endif
constexpr uint32_t kClasses = 3;

// Paths a wavefront batch holds at most when the caller leaves the batch size
// to the library (the queue length of the wavefront).
#ifndef VH_BATCH_PATHS_LOG2
#define VH_BATCH_PATHS_LOG2 27
#endif
// Workgroups per CU of the persistent and grid-stride launches: the grid
// follows the batch, about VH_GRID_PATHS_PER_THREAD paths per thread, between
// VH_GRID_MIN_BLOCKS_PER_CU and VH_GRID_BLOCKS_PER_CU workgroups per CU.  A
// whole 1920x1080 x 64-spp exposure (2^27 paths) takes the most (later
// workgroups start fresh while the first ones drain their slowest rays); a
// 4-spp exposure is faster with a quarter of them, whose threads each claim
// more rays instead of waiting to be scheduled.
#ifndef VH_GRID_BLOCKS_PER_CU
#define VH_GRID_BLOCKS_PER_CU 8
#endif
#ifndef VH_GRID_MIN_BLOCKS_PER_CU
#define VH_GRID_MIN_BLOCKS_PER_CU 4
#endif
#ifndef VH_GRID_PATHS_PER_THREAD
#define VH_GRID_PATHS_PER_THREAD 16
#endif
inline uint32_t wavefront_grid(size_t paths, uint32_t cu_count) {
    const size_t per_cu = paths / ((size_t)cu_count * kBlock * VH_GRID_PATHS_PER_THREAD);
    const uint32_t b = (uint32_t)std::min<size_t>(std::max<size_t>(per_cu, VH_GRID_MIN_BLOCKS_PER_CU),
                                                  VH_GRID_BLOCKS_PER_CU);
    return std::min<uint32_t>((uint32_t)((paths + kBlock - 1) / kBlock), cu_count * b);
}

// One job of a shared-queue batch (vh_render_shared): a scene's own
// constants, tables and film, its range of the batch's path slots and of the
// film pass's tiles.
struct JobArgs {
    Consts c;
    Scene S;
    float4* film;
    float4* guides;
    uint32_t path_base, paths;
    uint32_t tile_base, tiles;
};
// Jobs of one shared batch at most: a shadow ray carries its job in the top
// 4 bits of its id (2 x event slot + estimator < 2^28 with at most 2^27
// paths).
constexpr uint32_t kMaxJobs = 16, kJobShift = 28, kShadowIdMask = (1u << kJobShift) - 1u;

// Arguments of one batch's kernels, in device memory.  Every kernel of a batch
// takes a pointer to this block (and its bounce), so the launch sequence of a
// batch does not depend on the scene, the buffers or the batch: it is recorded
// once per stream into a graph and replayed for every batch, the block
// rewritten before each replay.
struct LaunchArgs {
    Paths P;
    Consts c;
    Scene S;
    float4* rays[2];   // extension queues: bounce b traces rays[b & 1], its shading appends to rays[(b + 1) & 1]
    float4* shadow_a;  // shadow rays of estimator a (slot q = pending event q)
    float4* shadow_b;  // shadow rays of estimator b
    uint32_t* counters;
    uint32_t* overflow;
    float4* film;
    float4* guides;    // null: no guides
    Moment* moments;   // null unless adaptive
    uint32_t* sorted;  // sorted shading: kClasses queues of `cap` extension-queue entries; null: arrival order
    uint32_t cap;      // paths the wavefront buffers hold
    uint32_t n;        // paths of the batch
    const JobArgs* jobs;  // shared-queue batches: the jobs in path order (P, c, S, film unused)
    uint32_t job_count;
};
#define VH_ARGS const LaunchArgs* __restrict__

// The job whose path slots hold path p (jobs are in path order).
__device__ __forceinline__ uint32_t job_of_path(const LaunchArgs* A, uint32_t p) {
    uint32_t j = 0;
    while (j + 1 < A->job_count && p >= A->jobs[j + 1].path_base) j++;
    return j;
}

__device__ __forceinline__ void reset_claims(uint32_t* counters) {
    counters[kExtendClaims] = 0;
    counters[kShadowClaims] = 0;
}
// Opens an adaptive batch: its camera pass appends to an empty queue.
__global__ void k_open_batch(VH_ARGS A) {
    uint32_t* counters = A->counters;
    if (threadIdx.x < kClasses) counters[kClassCount + threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        counters[0] = 0;
        counters[kLiveOut] = 0;
        counters[kShadowB] = 0;
        counters[kPending] = 0;
        reset_claims(counters);
    }
}

// Adaptive stop decision of sample `index` > minimum (path.slang:1804-1836):
// publish this sample's test in its parity slot, then stop if it and the
// whole clamped 3x3 neighbourhood's test of the previous sample passed.  The
// neighbours' slot is one no thread writes in this pass.
__device__ __forceinline__ void decide_pixel(Moment* M, const Consts& c, uint32_t index, uint32_t pixel) {
    Moment& m = M[pixel];
    if (m.stopped) return;
    const uint32_t slot = index & 1u;
    const bool settled = moment_settled(m, index, c.threshold);
    m.settled[slot] = settled ? 1u : 0u;
    bool done = settled;
    const int x = (int)(pixel % c.width), y = (int)(pixel / c.width);
    const int w = (int)c.width, h = (int)c.height;
    for (int dy = -1; dy <= 1 && done; dy++) {
        for (int dx = -1; dx <= 1 && done; dx++) {
            const int nx = min(max(x + dx, 0), w - 1), ny = min(max(y + dy, 0), h - 1);
            done = M[ny * w + nx].settled[slot ^ 1u] != 0u;
        }
    }
    if (done) m.stopped = 1u;
}
__global__ void k_decide(VH_ARGS A) {
    const Consts& c = A->c;
    Moment* M = A->moments;
    const uint32_t index = c.batch_first;
    for (uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x; pixel < c.pixels; pixel += gridDim.x * blockDim.x)
        decide_pixel(M, c, index, pixel);
}


// Camera rays of the batch's paths.  Plain batches write ray p for path p;
// adaptive ones append the rays of the pixels that still run.
template <bool kAdaptive>
__device__ __forceinline__ void camera_path(const Paths& P, const Consts& c, const Scene& S, float4* rays,
                                            uint32_t* counters, const Moment* M, uint32_t p, uint32_t local) {
    uint32_t s_local, t;
    slot_sample(local, c.group, c.pixels, s_local, t);
    const uint32_t pixel = slot_to_pixel(t, c.width, c.height);
    if (kAdaptive && c.decide && M[pixel].stopped) return;
    uint32_t px = pixel % c.width, py = pixel / c.width;
    uint32_t index = c.batch_first + s_local;
    Sampler smp = make_sampler(px, py, index, c.samples, c.seed);
    float cs[4];
    group_values(smp, 0u, cs);
    // path.slang:1727-1775 with camera_ray_line (vork_prelude.slang:95-110): pinhole
    const vh_camera& cam = c.camera;
    float ndcx = ((float)px + .5f + filter_offset(S.filter, cs[0])) / (float)c.width * 2.f + -1.f;
    float ndcy = ((float)py + .5f + filter_offset(S.filter, cs[1])) / (float)c.height * -2.f + 1.f;
    f3 fwd = load3(cam.forward), right = load3(cam.right), up = load3(cam.up), eye = load3(cam.eye);
    f3 d = normalize(fwd + right * (ndcx * cam.tan_half_x) + up * (ndcy * cam.tan_half_y));
    float fd = dot(d, fwd);
    float near_distance = cam.clip_start / fd, far_distance = cam.clip_end / fd;
    f3 o = eye + d * near_distance;
    float tmax = fmaxf(0.f, far_distance - length(o - eye));
    put_ray(rays, kAdaptive ? atomicAdd(&counters[0], 1u) : p, o, tmax, d, p | (event_code(VH_RAY_CAMERA) << kPathBits));
    P.thr[p] = make_float4(1.f, 1.f, 1.f, 0.f);
    P.rad[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.info[p] = make_uint4(pixel, index, screen_ray_hash(px, py, index, c.seed), VH_RAY_CAMERA << kEventShift);
    if (c.guides) {  // a camera ray that misses keeps zero guides and depth 1
        P.guide_a[p] = make_float4(0.f, 0.f, 0.f, 0.f);
        P.guide_n[p] = make_float4(0.f, 0.f, 0.f, 1.f);
    }
}
template <bool kAdaptive>
__global__ void k_camera(VH_ARGS A) {
    const uint32_t n = A->n;
    uint32_t* counters = A->counters;
    if (!kAdaptive && blockIdx.x == 0 && threadIdx.x < kClasses) counters[kClassCount + threadIdx.x] = 0;
    if (!kAdaptive && blockIdx.x == 0 && threadIdx.x == 0) {
        counters[0] = n;
        counters[kLiveOut] = 0;
        counters[kShadowB] = 0;
        counters[kPending] = 0;
        reset_claims(counters);
    }
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        camera_path<kAdaptive>(A->P, A->c, A->S, A->rays[0], counters, A->moments, p, p);
}
// Camera rays of a shared-queue batch: path p belongs to the job whose slot
// range holds it, and is that job's path p - path_base.
__global__ void k_camera_shared(VH_ARGS A) {
    const uint32_t n = A->n;
    uint32_t* counters = A->counters;
    if (blockIdx.x == 0 && threadIdx.x < kClasses) counters[kClassCount + threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[0] = n;
        counters[kLiveOut] = 0;
        counters[kShadowB] = 0;
        counters[kPending] = 0;
        reset_claims(counters);
    }
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const JobArgs& J = A->jobs[job_of_path(A, p)];
        camera_path<false>(A->P, J.c, J.S, A->rays[0], counters, nullptr, p, p - J.path_base);
    }
}

// Sums per-thread visit counts of a block into two u64 counters.
template <bool kCount>
__device__ __forceinline__ void flush_counts(uint64_t nodes, uint64_t tris, unsigned long long* out) {
    if (!kCount) return;
    __shared__ unsigned long long sn, st_;
    if (threadIdx.x == 0) { sn = 0; st_ = 0; }
    __syncthreads();
    atomicAdd(&sn, (unsigned long long)nodes);
    atomicAdd(&st_, (unsigned long long)tris);
    __syncthreads();
    if (threadIdx.x == 0) { atomicAdd(&out[0], sn); atomicAdd(&out[1], st_); }
}

// Extension rays of the live paths; hits into the path state.
template <bool kCount>
struct ExtendSource {
    Paths P;
    const float4* rays;
    uint32_t n;
    uint64_t* nodes;
    uint64_t* tris;
    const LaunchArgs* A;
    __device__ TlasView view(uint32_t p) const { return A->jobs[job_of_path(A, p)].S.tv; }
    __device__ uint32_t count() const { return n; }
    __device__ void fetch(uint32_t i, RayData& r) const {
        r.a = rays[2 * i];
        r.b = rays[2 * i + 1];
    }
    __device__ void unpack(uint32_t, const RayData& r, f3& o, f3& d, float& t0, float& t1, uint32_t& m,
                           uint32_t& key) const {
        o = mk(r.a.x, r.a.y, r.a.z);
        d = mk(r.b.x, r.b.y, r.b.z);
        t0 = 0.f;
        t1 = r.a.w;
        const uint32_t id = f2u(r.b.w);
        m = code_mask(id >> kPathBits);
        key = id & kPathMask;
    }
    __device__ void store(uint32_t p, const TraceHit& h, const TraceCounts& c) const {
        P.hit[p] = make_float4(h.found ? h.t : -1.f, h.u, h.v, u2f(h.item));
        P.hit_inst[p] = h.inst;
        if (kCount) { *nodes += c.nodes; *tris += c.tris; }
    }
};

template <bool kCount, bool kShared = false>
VH_TRACE_KERNEL void k_extend(VH_ARGS A, uint32_t bounce) {
    __shared__ uint64_t lds_stack[kLdsStack * kBlock];
    uint32_t* counters = A->counters;
    const uint32_t n = counters[0];
    if (n == 0) return;  // the paths ended at an earlier bounce
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st = make_stack(lds_stack, reinterpret_cast<uint64_t*>(A->overflow));
    uint64_t nodes = 0, tris = 0;
    const TlasView tv = A->S.tv;
    ExtendSource<kCount> src{A->P, A->rays[bounce & 1u], n, &nodes, &tris, A};
    if (!trace_persistent<false, kCount, ExtendSource<kCount>, kShared>(tv, src, &counters[kExtendClaims], st))
        atomicOr(&counters[8], kStackOverflow);
    if (gid == 0) {
        atomicAdd((unsigned long long*)&counters[4], (unsigned long long)n);
        if (kCount && bounce == 0) atomicAdd((unsigned long long*)&counters[kCameraRays], (unsigned long long)n);
    }
    flush_counts<kCount>(nodes, tris, (unsigned long long*)&counters[bounce == 0 ? kCameraVisits : 16]);
}

// Camera rays whose waves hold the jittered samples of one pixel: a wave
// traces its 64 queue entries as one packet (trace_packet), with the hits
// the per-lane traversal gives.
#ifndef VH_PACKET_CAMERA
#define VH_PACKET_CAMERA 1
#endif
// Packets when a pixel's samples fill at least this many consecutive slots
// (a wave then holds the samples of at most 64 / VH_PACKET_MIN_GROUP nearby
// pixels).
#ifndef VH_PACKET_MIN_GROUP
#define VH_PACKET_MIN_GROUP 16
#endif
// Waves per SIMD the packet kernel is compiled for: its node and triangle
// data and its walk state sit in SGPRs, so it fits 72 VGPRs at 7 waves (8
// spill a few); the tuned table sets 8.
#ifndef VH_PACKET_WAVES_PER_EU
#define VH_PACKET_WAVES_PER_EU 6
#endif
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(VH_PACKET_WAVES_PER_EU))) void k_extend_packet(
    VH_ARGS A) {
    __shared__ uint64_t stacks[kBlock / 64][kPacketStack];
    const uint32_t wave = threadIdx.x >> 6, lane = __lane_id();
    const uint32_t waves = gridDim.x * (kBlock / 64);
    uint32_t* counters = A->counters;
    const Paths P = A->P;
    const Scene S = A->S;
    const float4* rays = A->rays[0];
    const uint32_t n = counters[0];
    bool ok = true;
    for (uint32_t chunk = blockIdx.x * (kBlock / 64) + wave; (size_t)chunk * 64 < n; chunk += waves) {
        const uint32_t i = chunk * 64 + lane;
        PacketRay r;
        r.live = i < n;
        const float4 a = r.live ? rays[2 * (size_t)i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 b = r.live ? rays[2 * (size_t)i + 1] : make_float4(0.f, 0.f, 1.f, 0.f);
        const uint32_t id = f2u(b.w);
        r.wo = r.o = mk(a.x, a.y, a.z);
        r.wd = r.ld = mk(b.x, b.y, b.z);
        r.inv = box_inv(r.ld);
        r.winv = r.inv;
        r.oct = octant(r.inv);
        r.lsh = make_shear(r.ld);
        r.tmin = 0.f;
        r.tmax = a.w;
        r.mask = code_mask(id >> kPathBits);
        r.hit.found = false; r.hit.t = a.w; r.hit.u = r.hit.v = 0.f; r.hit.inst = r.hit.item = 0xffffffffu;
        if (!trace_packet(S.tv, r, stacks[wave])) ok = false;
        if (r.live) {
            const uint32_t p = id & kPathMask;
            P.hit[p] = make_float4(r.hit.found ? r.hit.t : -1.f, r.hit.u, r.hit.v, u2f(r.hit.item));
            P.hit_inst[p] = r.hit.inst;
        }
    }
    if (!ok && lane == 0) atomicOr(&counters[8], kStackOverflow);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long*)&counters[4], (unsigned long long)n);
}

// material.slang:69-79 with the instance's local bounds
__device__ __forceinline__ f3 surface_position_error(const InstanceRecord& r, const BlasRecord& b, f3 world) {
    const float* model = r.o2w;
    const float* inv = r.w2o;
    f3 magnitude = mk(fmaxf(fabsf(b.lo[0]), fabsf(b.hi[0])), fmaxf(fabsf(b.lo[1]), fabsf(b.hi[1])),
                      fmaxf(fabsf(b.lo[2]), fabsf(b.hi[2])));
    f3 span = mk(fabsf(b.hi[0] - b.lo[0]), fabsf(b.hi[1] - b.lo[1]), fabsf(b.hi[2] - b.lo[2]));
    float extent = fmaxf(span.x, fmaxf(span.y, span.z));
    f3 aw = vabs(world);
    f3 li = mk(fabsf(inv[0]) * aw.x + fabsf(inv[1]) * aw.y + fabsf(inv[2]) * aw.z + fabsf(inv[3]),
               fabsf(inv[4]) * aw.x + fabsf(inv[5]) * aw.y + fabsf(inv[6]) * aw.z + fabsf(inv[7]),
               fabsf(inv[8]) * aw.x + fabsf(inv[9]) * aw.y + fabsf(inv[10]) * aw.z + fabsf(inv[11]));
    f3 le = (magnitude + mk(extent, extent, extent) + li) * kGamma7;
    f3 a = mk(fabsf(model[0]) * le.x + fabsf(model[1]) * le.y + fabsf(model[2]) * le.z,
              fabsf(model[4]) * le.x + fabsf(model[5]) * le.y + fabsf(model[6]) * le.z,
              fabsf(model[8]) * le.x + fabsf(model[9]) * le.y + fabsf(model[10]) * le.z);
    f3 bb = mk(fabsf(model[0]) * magnitude.x + fabsf(model[1]) * magnitude.y + fabsf(model[2]) * magnitude.z + fabsf(model[3]),
               fabsf(model[4]) * magnitude.x + fabsf(model[5]) * magnitude.y + fabsf(model[6]) * magnitude.z + fabsf(model[7]),
               fabsf(model[8]) * magnitude.x + fabsf(model[9]) * magnitude.y + fabsf(model[10]) * magnitude.z + fabsf(model[11]));
    return a + bb * kGamma7;
}

struct Surface {
    f3 world, face, smooth;
    uint32_t material;
};
// hit_geometry.slang:120-135 and the double-sided facing of the hit.  The
// vertex positions, normals and material come from the hit triangle's
// leaf-order record (`item`): one 128-byte line, not the index triple.
__device__ __forceinline__ Surface reconstruct(const Scene& S, uint32_t material_count, const InstanceRecord& r,
                                               const BlasRecord& b, uint32_t item, float u, float v, f3 direction) {
    Surface sf;
    const float4* rec = reinterpret_cast<const float4*>(b.surf + kSurfFloats * (size_t)item);
    const float4 q0 = rec[0], q1 = rec[1], q2 = rec[2], q3 = rec[3], q4 = rec[4];
    f3 p0 = mk(q0.x, q0.y, q0.z), p1 = mk(q0.w, q1.x, q1.y), p2 = mk(q1.z, q1.w, q2.x);
    float b1 = u, b2 = v, b0 = 1.f - b1 - b2;
    f3 p = p0 * b0 + p1 * b1 + p2 * b2;
    sf.world = xform_point(r.o2w, p);
    f3 face_o = cross(p1 - p0, p2 - p0);
    const bool vn = b.surf_normals != 0u;
    f3 n0 = vn ? mk(q2.y, q2.z, q2.w) : face_o;
    f3 n1 = vn ? mk(q3.x, q3.y, q3.z) : face_o;
    f3 n2 = vn ? mk(q3.w, q4.x, q4.y) : face_o;
    face_o = finite_unit_normal(face_o, n0);
    // normal matrix rows = columns of world-to-object
    const float* w = r.w2o;
    f3 face = finite_unit_normal(mk(w[0] * face_o.x + w[4] * face_o.y + w[8] * face_o.z,
                                    w[1] * face_o.x + w[5] * face_o.y + w[9] * face_o.z,
                                    w[2] * face_o.x + w[6] * face_o.y + w[10] * face_o.z),
                                 face_o);
    f3 ns = n0 * b0 + n1 * b1 + n2 * b2;
    ns = mk(w[0] * ns.x + w[4] * ns.y + w[8] * ns.z, w[1] * ns.x + w[5] * ns.y + w[9] * ns.z,
            w[2] * ns.x + w[6] * ns.y + w[10] * ns.z);
    ns = finite_unit_normal(ns, face);
    if (dot(ns, face) < 0.f) ns = -ns;
    if (dot(face, direction) > 0.f) {
        face = -face;
        ns = -ns;
    }
    sf.face = face;
    sf.smooth = ns;
    uint32_t local = f2u(q4.z);
    uint32_t m = r.material + local;
    sf.material = m < material_count ? m : material_count - 1;
    return sf;
}
// path.slang:772-781 (no terminator offset, no hair)
__device__ __forceinline__ f3 path_origin(const Surface& sf, f3 error, f3 direction) {
    bool reflection = dot(sf.smooth, direction) > 0.f;
    return surface_ray_origin(sf.world, error, reflection ? sf.smooth : sf.face, direction);
}

// Closure class of a material for the sorted shading queues: metal GGX
// (the F82 lobe alone), dielectric GGX over a diffuse base, diffuse alone.
__device__ __forceinline__ uint32_t closure_class(const vh_material& m) {
    if (m.base_metalness >= 1.f) return 2u;
    return m.specular_weight > 0.f ? 1u : 0u;
}

// Misses add the sky; hits go to the queue of their closure class, one
// atomic per class and wave (ballot compaction).
__global__ __launch_bounds__(kBlock) void k_classify(VH_ARGS A, uint32_t bounce_parity) {
    uint32_t* counters = A->counters;
    const Paths& P = A->P;
    const Consts& c = A->c;
    const Scene& S = A->S;
    const float4* rays = A->rays[bounce_parity & 1u];
    const uint32_t n = counters[0];
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    uint32_t dropped = 0;
    // whole waves walk the queue together so that the ballots see every lane
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < n; base += stride) {
        const uint32_t i = base + __lane_id();
        uint32_t cls = kClasses;  // none: a miss or past the end
        if (i < n) {
            const float4 rd4 = rays[2 * i + 1];
            const uint32_t p = f2u(rd4.w) & kPathMask;
            const float4 h = P.hit[p];
            if (h.x >= 0.f) {
                const InstanceRecord& ir = S.tv.instances[P.hit_inst[p]];
                const BlasRecord& br = S.tv.blases[ir.blas];
                const uint32_t local = f2u(br.surf[kSurfFloats * (size_t)f2u(h.w) + 18]);
                const uint32_t m = ir.material + local;
                cls = closure_class(S.materials[m < c.material_count ? m : c.material_count - 1]);
            } else {
                const uint4 info = P.info[p];
                const float4 t4 = P.thr[p], r4 = P.rad[p];
                const float limit = ((info.w >> kBounceShift) & 0xffu) == 0 ? c.clamp_direct : c.clamp_indirect;
                const f3 sky = clamp_sample(mk(t4.x, t4.y, t4.z) * sky_radiance(c.sky, mk(rd4.x, rd4.y, rd4.z)),
                                            limit, dropped);
                P.rad[p] = make_float4(r4.x + sky.x, r4.y + sky.y, r4.z + sky.z, r4.w);
            }
        }
        for (uint32_t k = 0; k < kClasses; k++) {
            const uint64_t m = __ballot(cls == k);
            if (!m) continue;
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t slot = 0;
            if ((int)__lane_id() == leader) slot = atomicAdd(&counters[kClassCount + k], (uint32_t)__popcll(m));
            slot = __shfl(slot, leader);
            if (cls == k) A->sorted[(size_t)k * A->cap + slot + (uint32_t)__popcll(m & lt)] = i;
        }
    }
    if (dropped) atomicAdd(&counters[3], dropped);
}

// Waves per SIMD k_shade is compiled for (bounds its VGPRs to 512 / waves).
// Its time goes to waiting on the chain of scene loads, so it runs at the
// most waves that keep its spills small (4: 20 B per lane; 5 spills 164 B).
#ifndef VH_SHADE_WAVES_PER_EU
#define VH_SHADE_WAVES_PER_EU 4
#endif
// With kShared the queue holds the paths of several jobs: each path shades
// with its job's constants and tables, and its shadow rays carry the job.
template <bool kShared = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(VH_SHADE_WAVES_PER_EU))) void k_shade(
    VH_ARGS A, uint32_t bounce_parity) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, threads = gridDim.x * blockDim.x;
    uint32_t* counters = A->counters;
    const Paths& P = A->P;
    const Consts& c_all = A->c;
    const Scene& S_all = A->S;
    const float4* rays = A->rays[bounce_parity & 1u];
    float4* next = A->rays[(bounce_parity + 1u) & 1u];
    float4* shadow_rays = A->shadow_a;
    float4* shadow_b = A->shadow_b;
    // with sorted queues the entries are the hits of the classes, in class
    // order; else every extension ray
    const uint32_t* sorted = A->sorted;
    const uint32_t c0 = sorted ? counters[kClassCount] : 0u;
    const uint32_t c1 = sorted ? c0 + counters[kClassCount + 1] : 0u;
    const uint32_t n = sorted ? c1 + counters[kClassCount + 2] : counters[0];
    uint32_t dropped = 0;
    uint32_t i = gid;
    // the queue entry of the lane's next index, loaded one entry ahead (its
    // load is in flight while the current entry is shaded)
    auto entry = [&](uint32_t k) {
        uint32_t e = k;
        if (sorted) {
            const uint32_t cls = k < c0 ? 0u : k < c1 ? 1u : 2u;
            e = sorted[(size_t)cls * A->cap + (k - (cls == 0 ? 0u : cls == 1 ? c0 : c1))];
        }
        return rays[2 * e + 1];
    };
    float4 ahead = i < n ? entry(i) : make_float4(0.f, 0.f, 0.f, 0.f);
    while (true) {
        // A miss only adds the sky.  Each lane walks its entries (i, i + threads,
        // ...) past misses until it holds a hit, so the wave's shading steps run
        // on hits: on the later bounces most extension rays miss, and a step
        // costs the same however few of its lanes hold a hit.
        float4 rd4, t4, r4, h;
        uint4 info;
        uint32_t p = 0;
        bool held = false;
        while (i < n) {
            rd4 = ahead;
            i += threads;
            if (i < n) ahead = entry(i);
            p = f2u(rd4.w) & kPathMask;
            info = P.info[p];
            t4 = P.thr[p];
            r4 = P.rad[p];
            h = P.hit[p];
            if (h.x >= 0.f) {
                held = true;
                break;
            }
            const Consts& cm = kShared ? A->jobs[job_of_path(A, p)].c : c_all;
            const float sky_limit = ((info.w >> kBounceShift) & 0xffu) == 0 ? cm.clamp_direct : cm.clamp_indirect;
            const f3 sky = clamp_sample(mk(t4.x, t4.y, t4.z) * sky_radiance(cm.sky, mk(rd4.x, rd4.y, rd4.z)),
                                        sky_limit, dropped);
            P.rad[p] = make_float4(r4.x + sky.x, r4.y + sky.y, r4.z + sky.z, r4.w);
        }
        if (!held) break;
        const uint32_t job = kShared ? job_of_path(A, p) : 0u;
        const Consts& c = kShared ? A->jobs[job].c : c_all;
        const Scene& S = kShared ? A->jobs[job].S : S_all;
        const uint32_t job_bits = kShared ? job << kJobShift : 0u;
        uint32_t flags = info.w;
        const uint32_t bounce = (flags >> kBounceShift) & 0xffu;
        uint32_t diffuse = (flags >> kDiffuseShift) & 0xfu, glossy = (flags >> kGlossyShift) & 0xfu;
        f3 throughput = mk(t4.x, t4.y, t4.z), direction = mk(rd4.x, rd4.y, rd4.z);
        const float limit = bounce == 0 ? c.clamp_direct : c.clamp_indirect;
        const uint32_t inst = P.hit_inst[p], item = f2u(h.w);
        const InstanceRecord& ir = S.tv.instances[inst];
        const BlasRecord& br = S.tv.blases[ir.blas];
        Surface sf = reconstruct(S, c.material_count, ir, br, item, h.y, h.z, direction);
        const vh_material& mat = S.materials[sf.material];
        Closure cl;
        cl.base_color = load3(mat.base_color);
        cl.metal = mat.base_metalness;
        cl.spec_weight = mat.specular_weight;
        cl.roughness = mat.specular_roughness;
        cl.ior = mat.specular_ior;
        cl.n = sf.smooth;
        const f3 error = surface_position_error(ir, br, sf.world);
        const uint32_t px = info.x % c.width, py = info.x / c.width;
        Sampler smp = make_sampler(px, py, info.y, c.samples, c.seed);
        const uint32_t vertex = bounce;
        float scatter[4];
        group_values(smp, 1u + vertex * 5u + 0u, scatter);
        const f3 view = -direction;
        const f3 interface_face = dot(sf.smooth, view) > 0.f ? sf.smooth : sf.face;
        Prep pr = transport_prepare(S.luts, cl, view, sf.smooth, interface_face);
        if (c.guides && bounce == 0) {
            const f3 a = guide_albedo(S.luts, cl, view);
            P.guide_a[p] = make_float4(a.x, a.y, a.z, 1.f);
            P.guide_n[p] = make_float4(sf.smooth.x, sf.smooth.y, sf.smooth.z, div_a(h.x, h.x + 1.f));
        }
        BsdfSample bs = transport_sample(cl, view, scatter[0], scatter[1], scatter[2], pr);
        if (!finite3(bs.direction) || !finite3(bs.weight) || !isfinite(bs.pdf)) {
            P.info[p].w = flags | kNonFinite;
            continue;
        }
        // ---- next_event (path.slang:851-952), visibility deferred
        uint32_t rng = info.z;
        f3 a = mk(0.f, 0.f, 0.f), b = mk(0.f, 0.f, 0.f);
        f3 s0o = a, s0d = a, s1o = a, s1d = a;
        float s0t = 0.f, s1t = 0.f;
        if (c.light_count) {
            float u[4];
            group_values(smp, 1u + vertex * 5u + 1u, u);
            float probability;
            uint32_t li = light_pick(S.tree, c.light_nodes, sf.world, u[2], u[3], &probability);
            if (probability > 0.f) {
                const vh_rect_light l = S.lights[li];
                f3 ldir;
                EmitEval src = rect_sample(l, sf.world, u[0], u[1], &ldir);
                f3 power = src.pdf > 0.f ? div_a(src.radiance, src.pdf) : mk(0.f, 0.f, 0.f);
                f3 response = analytic_mis_response(cl, view, ldir, sf.face, src.pdf, pr);
                a = div_a(response * power, probability);
                float threshold = c.light_threshold;
                float importance = maxc(a);
                if (importance < threshold && threshold > 0.f) {
                    float keep = div_a(importance, threshold);
                    if (path_random(rng) >= keep) a = mk(0.f, 0.f, 0.f); else a = div_a(a, keep);
                }
                if (nonzero(a)) {
                    f3 origin = path_origin(sf, error, ldir);
                    f3 cd;
                    float cdist;
                    surface_ray_connection(origin, sf.world, ldir, src.distance, &cd, &cdist);
                    s0o = origin;
                    s0d = cd;
                    s0t = cdist;
                }
                bool continuation = bs.event != 0 && event_hemisphere(bs.event, bs.direction, sf.face, sf.smooth);
                if (continuation && dot(bs.direction, bs.direction) > 0.f) {
                    EmitEval e = rect_eval(l, sf.world, bs.direction);
                    if (e.pdf > 0.f) {
                        float weight = bs.pdf > 0.f ? mis_power(bs.pdf, e.pdf) : 1.f;
                        b = (bs.weight * e.radiance) * div_a(weight, probability);
                        float imp2 = maxc(b);
                        if (imp2 < threshold && threshold > 0.f) {
                            float keep = div_a(imp2, threshold);
                            if (path_random(rng) >= keep) b = mk(0.f, 0.f, 0.f); else b = div_a(b, keep);
                        }
                        if (nonzero(b)) {
                            f3 origin = path_origin(sf, error, bs.direction);
                            f3 cd;
                            float cdist;
                            surface_ray_connection(origin, sf.world, bs.direction, e.distance, &cd, &cdist);
                            s1o = origin;
                            s1d = cd;
                            s1t = cdist;
                        }
                    }
                }
            }
        }
        const bool has_a = nonzero(a), has_b = nonzero(b), pending = has_a || has_b;
        const f3 ta = throughput * a, tb = throughput * b;
        // ---- continuation (path.slang:1680-1705)
        bool alive = true;
        if (bounce == c.max_bounces) alive = false;
        uint32_t event = bs.event;
        if (alive) {
            if (event == VH_RAY_DIFFUSE) {
                if (diffuse++ >= c.diffuse_bounces) alive = false;
            } else if (event == VH_RAY_SPECULAR) {
                if (glossy++ >= c.glossy_bounces || (diffuse > 0 && (c.caustics & 1u) == 0)) alive = false;
            } else {
                alive = false;
            }
        }
        if (alive && !event_hemisphere(event, bs.direction, sf.face, sf.smooth)) alive = false;
        if (alive) {
            throughput = throughput * bs.weight;
            if (!finite3(throughput)) {
                flags |= kNonFinite;
                alive = false;
            } else if (maxc(throughput) <= 0.f) {
                alive = false;
            }
        }
        if (alive && bounce >= kRouletteStart) {
            float survive = clampf(maxc(throughput), kRouletteMin, kRouletteMax);
            float d[4];
            group_values(smp, 1u + vertex * 5u + 4u, d);
            if (d[2] > survive) alive = false;
            else throughput = div_a(throughput, survive);
        }
        flags = (flags & kNonFinite) | (min(diffuse, 15u) << kDiffuseShift) | (min(glossy, 15u) << kGlossyShift) |
                ((bounce + 1) << kBounceShift) | (event << kEventShift);
        P.info[p] = make_uint4(info.x, info.y, rng, flags);
        // One atomic per wave reserves the continuing paths' extension-ray
        // slots (low word) and the pending next events' slots (high word): each
        // waits a device-wide round trip, so the wave does one, not three.
        const uint64_t am = __ballot(alive), pm = __ballot(pending);
        const uint64_t lt = (1ull << __lane_id()) - 1ull;
        uint32_t next_slot = 0, q = 0;
        if (am | pm) {
            const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
            unsigned long long base = 0;
            if ((int)__lane_id() == leader)
                base = atomicAdd(reinterpret_cast<unsigned long long*>(&counters[kLiveOut]),
                                 (unsigned long long)__popcll(am) | ((unsigned long long)__popcll(pm) << 32));
            base = __shfl(base, leader);
            next_slot = (uint32_t)base + (uint32_t)__popcll(am & lt);
            q = (uint32_t)(base >> 32) + (uint32_t)__popcll(pm & lt);
        }
        if (pending) {
            // event q: estimator a is shadow ray q (an empty one, tmax -1, when
            // only b is pending), estimator b goes to the second queue
            P.pend[q] = make_float4(ta.x, ta.y, ta.z, limit);
            P.pend1[q] = make_float4(tb.x, tb.y, tb.z, u2f(p));
            if (has_a) put_ray(shadow_rays, q, s0o, s0t, s0d, (2u * q) | job_bits);
            else put_ray(shadow_rays, q, s1o, -1.f, s1d, (2u * q) | job_bits);
        }
        const uint64_t bm = __ballot(has_b);
        if (bm) {
            const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
            uint32_t base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(&counters[kShadowB], (uint32_t)__popcll(bm));
            base = __shfl(base, leader);
            if (has_b) put_ray(shadow_b, base + (uint32_t)__popcll(bm & lt), s1o, s1t, s1d, (2u * q + 1u) | job_bits);
        }
        if (alive) {
            f3 origin = path_origin(sf, error, bs.direction);
            P.thr[p] = make_float4(throughput.x, throughput.y, throughput.z, 0.f);
            put_ray(next, next_slot, origin, FLT_MAX, bs.direction, p | (event_code(event) << kPathBits));
        }
    }
    if (dropped) atomicAdd(&counters[3], dropped);
}

// Shadow rays of the pending next events: ray id 2q+k is estimator k of
// event q.  The first `na` come from the estimator-a queue (slot q = event
// q), the rest from the estimator-b queue.
template <bool kCount, bool kShared>
struct ShadowSource {
    Paths P;
    const float4* rays;
    const float4* rays_b;
    uint32_t na, n;
    uint64_t* nodes;
    uint64_t* tris;
    const LaunchArgs* A;
    __device__ TlasView view(uint32_t id) const { return A->jobs[id >> kJobShift].S.tv; }
    __device__ uint32_t count() const { return n; }
    __device__ void fetch(uint32_t i, RayData& r) const {
        const float4* q = i < na ? rays + 2 * (size_t)i : rays_b + 2 * (size_t)(i - na);
        r.a = q[0];
        r.b = q[1];
    }
    __device__ void unpack(uint32_t, const RayData& r, f3& o, f3& d, float& t0, float& t1, uint32_t& m,
                           uint32_t& key) const {
        o = mk(r.a.x, r.a.y, r.a.z);
        d = mk(r.b.x, r.b.y, r.b.z);
        t0 = 0.f;
        t1 = r.a.w;
        m = VH_RAY_SHADOW;
        key = f2u(r.b.w);
    }
    __device__ void store(uint32_t r, const TraceHit& h, const TraceCounts& c) const {
        P.vis[kShared ? r & kShadowIdMask : r] = h.found ? 1 : 0;
        if (kCount) { *nodes += c.nodes; *tris += c.tris; }
    }
};

template <bool kCount, bool kShared = false>
VH_ANY_KERNEL void k_shadow(VH_ARGS A) {
    __shared__ uint64_t lds_stack[kLdsStack * kBlock];
    uint32_t* counters = A->counters;
    const uint32_t na = counters[kPending], n = na + counters[kShadowB];
    if (n == 0) return;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    Stack st = make_stack(lds_stack, reinterpret_cast<uint64_t*>(A->overflow));
    uint64_t nodes = 0, tris = 0;
    const TlasView tv = A->S.tv;
    ShadowSource<kCount, kShared> src{A->P, A->shadow_a, A->shadow_b, na, n, &nodes, &tris, A};
    if (!trace_persistent<true, kCount, ShadowSource<kCount, kShared>, kShared>(tv, src, &counters[kShadowClaims], st))
        atomicOr(&counters[8], kStackOverflow);
    if (gid == 0) atomicAdd((unsigned long long*)&counters[6], (unsigned long long)n);
    flush_counts<kCount>(nodes, tris, (unsigned long long*)&counters[20]);
}

// Adds each pending next event: the clamped sum of its unoccluded estimators
// (sample_clamp.slang:7-16 over path.slang:925-950's value).
__global__ void k_resolve(VH_ARGS A) {
    uint32_t* counters = A->counters;
    const Paths& P = A->P;
    const uint32_t n = counters[kPending];
    uint32_t dropped = 0;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const float4 pa = P.pend[q], pb = P.pend1[q];
        const uint32_t p = f2u(pb.w);
        f3 va = mk(pa.x, pa.y, pa.z), vb = mk(pb.x, pb.y, pb.z);
        if (nonzero(va) && P.vis[2 * q]) va = mk(0.f, 0.f, 0.f);
        if (nonzero(vb) && P.vis[2 * q + 1]) vb = mk(0.f, 0.f, 0.f);
        const f3 add = clamp_sample(va + vb, pa.w, dropped);
        const float4 r = P.rad[p];
        P.rad[p] = make_float4(r.x + add.x, r.y + add.y, r.z + add.z, r.w);
    }
    if (dropped) atomicAdd(&counters[3], dropped);
}

// One thread per pixel: adds the batch's samples, in sample order, to the
// film (resolve.slang:228-238).  A sample is (radiance, 1), or zero when it
// was non-finite (path.slang:1948-1952).  When a pixel's samples sit in
// groups of 8 or more consecutive path slots, the block stages 8 samples of
// its block of pixels at a time in LDS with loads that read whole cache lines (a
// thread per pixel would read one 16-byte value per line per sample); the
// per-pixel sums keep sample order.  Adaptive batches also update the pixel's moment, add a stopped
// pixel's mean in place of a sample, and record the settled test of sample
// indices up to `minimum` (later ones are the decide pass's).  With `guides`
// it also sums the camera hits' guides (path.slang:1956-1961): albedo and
// coverage at [2 pixel], normal and depth at [2 pixel + 1]; a dropped sample
// adds none, and a stopped pixel's sums grow by (i + 1) / i so they stay
// sums over every sample (path.slang:1806-1809).
constexpr uint32_t kFilmStage = 8, kFilmPad = kFilmStage + 1;
template <bool kAdaptive>
__device__ __forceinline__ void film_tile(const Paths& P, const Consts& c, float4* film, float4* guides, Moment* M,
                                          uint32_t* counters, uint32_t tile, float4* s_rad, uint32_t* s_flags) {
    const uint32_t slot = tile * kBlock + threadIdx.x;
    const bool valid = slot < c.pixels;
    const bool staged = c.group >= kFilmStage;  // the same for the whole launch
    const uint32_t pixel = valid ? slot_to_pixel(slot, c.width, c.height) : 0u;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 g0 = acc, g1 = acc;
    Moment m;
    if (valid) {
        acc = film[pixel];
        if (guides) {
            g0 = guides[2 * pixel];
            g1 = guides[2 * pixel + 1];
        }
        if (kAdaptive) m = M[pixel];
    }
    uint32_t bad_count = 0, stopped_count = 0;
    for (uint32_t s = 0; s < c.batch_samples; s++) {
        if (staged && s % kFilmStage == 0) {
            __syncthreads();
            const uint32_t t0 = tile * kBlock;
            for (uint32_t j = 0; j < kFilmStage; j++) {
                const uint32_t e = threadIdx.x + kBlock * j, ts = t0 + e / kFilmStage, ss = s + e % kFilmStage;
                if (ts < c.pixels && ss < c.batch_samples) {
                    const uint32_t q = path_slot(ss, ts, c.group, c.pixels);
                    s_rad[(e / kFilmStage) * kFilmPad + e % kFilmStage] = P.rad[q];
                    s_flags[e] = P.info[q].w;
                }
            }
            __syncthreads();
        }
        if (!valid) continue;
        const uint32_t index = c.batch_first + s;
        if (kAdaptive) {
            if (index == 0) {
                m.mean = make_float4(0.f, 0.f, 0.f, 0.f);
                m.variance = m.mean;
                m.stopped = 0u;
            }
            if (c.decide && m.stopped) {
                acc.x += m.mean.x;
                acc.y += m.mean.y;
                acc.z += m.mean.z;
                acc.w += m.mean.w;
                if (guides) {
                    const float grow = (float)(index + 1u) / (float)index;
                    g0 = make_float4(g0.x * grow, g0.y * grow, g0.z * grow, g0.w * grow);
                    g1 = make_float4(g1.x * grow, g1.y * grow, g1.z * grow, g1.w * grow);
                }
                stopped_count++;
                continue;
            }
            if (index <= c.minimum)
                m.settled[index & 1u] = index >= c.minimum && moment_settled(m, index, c.threshold) ? 1u : 0u;
        }
        const uint32_t p = path_slot(s, slot, c.group, c.pixels);
        const float4 r = staged ? s_rad[threadIdx.x * kFilmPad + s % kFilmStage] : P.rad[p];
        const uint32_t flags = staged ? s_flags[threadIdx.x * kFilmStage + s % kFilmStage] : P.info[p].w;
        const bool bad = (flags & kNonFinite) || !isfinite(r.x) || !isfinite(r.y) || !isfinite(r.z);
        const float4 v = bad ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(r.x, r.y, r.z, 1.f);
        bad_count += bad ? 1u : 0u;
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
        if (guides && !bad) {
            const float4 a = P.guide_a[p], n = P.guide_n[p];
            g0 = make_float4(g0.x + a.x, g0.y + a.y, g0.z + a.z, g0.w + a.w);
            g1 = make_float4(g1.x + n.x, g1.y + n.y, g1.z + n.z, g1.w + n.w);
        }
        if (kAdaptive) moment_add(m, v, index);
    }
    if (!valid) return;
    film[pixel] = acc;
    if (guides) {
        guides[2 * pixel] = g0;
        guides[2 * pixel + 1] = g1;
    }
    if (kAdaptive) M[pixel] = m;
    if (bad_count) atomicAdd(&counters[9], bad_count);
    if (stopped_count) atomicAdd((unsigned long long*)&counters[10], (unsigned long long)stopped_count);
}
// Tiles of kBlock pixel slots, strided over the grid (the barriers of a tile
// are uniform: every thread of a block walks the same tiles).
template <bool kAdaptive>
__global__ __launch_bounds__(kBlock) void k_film(VH_ARGS A) {
    __shared__ float4 s_rad[kBlock * kFilmPad];
    __shared__ uint32_t s_flags[kBlock * kFilmStage];
    for (uint32_t tile = blockIdx.x; tile * kBlock < A->c.pixels; tile += gridDim.x)
        film_tile<kAdaptive>(A->P, A->c, A->film, A->guides, A->moments, A->counters, tile, s_rad, s_flags);
}
// The film pass of a shared-queue batch: tiles numbered across the jobs, each
// job's tiles over its own pixels and path slots.
__global__ __launch_bounds__(kBlock) void k_film_shared(VH_ARGS A) {
    __shared__ float4 s_rad[kBlock * kFilmPad];
    __shared__ uint32_t s_flags[kBlock * kFilmStage];
    const JobArgs& last = A->jobs[A->job_count - 1];
    for (uint32_t tile = blockIdx.x; tile < last.tile_base + last.tiles; tile += gridDim.x) {
        uint32_t j = 0;
        while (j + 1 < A->job_count && tile >= A->jobs[j + 1].tile_base) j++;
        const JobArgs& J = A->jobs[j];
        Paths P = A->P;  // the job's slots start at its path base
        P.thr += J.path_base; P.rad += J.path_base; P.info += J.path_base; P.hit += J.path_base;
        P.guide_a += J.path_base; P.guide_n += J.path_base; P.hit_inst += J.path_base;
        film_tile<false>(P, J.c, J.film, J.guides, nullptr, A->counters, tile - J.tile_base, s_rad, s_flags);
    }
}

__global__ void k_swap_counts(VH_ARGS A) {
    uint32_t* counters = A->counters;
    if (threadIdx.x < kClasses) counters[kClassCount + threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        counters[0] = counters[kLiveOut];
        counters[kLiveOut] = 0;
        counters[kShadowB] = 0;
        counters[kPending] = 0;
        reset_claims(counters);
    }
}

// ------------------------------------------------------------------ LUTs
// crates/vork-render/shaders/environment.slang:19-27,78-114
__device__ float brdf_cosine(uint32_t column) {
    float t = (float)column / (float)(kBrdfSize - 1);
    float t2 = t * t;
    return t2 * t2;
}
__device__ float brdf_alpha(uint32_t row) {
    if (row == 0) return kGgxAlphaMin;
    if (row == kBrdfSize - 1) return 1.f;
    return exp2f(log2f(kGgxAlphaMin) * (1.f - (float)row / (float)(kBrdfSize - 1)));
}
__global__ void k_bake_moments(float4* out) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kBrdfSize * kBrdfSize) return;
    uint32_t column = k % kBrdfSize, row = k / kBrdfSize;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    float nv = brdf_cosine(column), alpha = brdf_alpha(row);
    for (uint32_t i = 0; i < kEnvironmentSamples; i++) {
        float u0 = ((float)i + .5f) / (float)kEnvironmentSamples;
        float u1 = (float)__builtin_bitreverse32(i) * kRadicalInverseScale;
        float vh;
        float g = ggx_visible_reflection(nv, alpha, u0, u1, &vh);
        float f = powf(1.f - vh, 5.f);
        s0 += (1.f - f) * g;
        s1 += f * g;
        s2 += vh * powf(1.f - vh, 6.f) * g;
    }
    float average = 0.f;
    if (column == 0) {
        for (uint32_t j = 0; j < kBrdfSize; j++) {
            float nvj = brdf_cosine(j);
            float interval = .5f * (brdf_cosine(min(j + 1, kBrdfSize - 1)) - brdf_cosine(j > 0 ? j - 1 : 0));
            for (uint32_t i = 0; i < kEnvironmentSamples; i++) {
                float u0 = ((float)i + .5f) / (float)kEnvironmentSamples;
                float u1 = (float)__builtin_bitreverse32(i) * kRadicalInverseScale;
                float vh;
                average += interval * 2.f * nvj * ggx_visible_reflection(nvj, alpha, u0, u1, &vh);
            }
        }
        average /= (float)kEnvironmentSamples;
    }
    out[k] = make_float4(s0 / (float)kEnvironmentSamples, s1 / (float)kEnvironmentSamples,
                         s2 / (float)kEnvironmentSamples, average);
}
// openpbr_core.slang:154-166 over the (nv, roughness, eta) grid of environment.slang:40-50
__global__ void k_bake_dielectric(float* out) {
    const uint32_t n = kBrdfAlbedoSize;
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (kBrdfIorSteps + 1) * n * n) return;
    uint32_t slice = k / (n * n), cell = k % (n * n);
    uint32_t cx = cell % n, cy = cell / n;
    float q = (float)slice / (float)kBrdfIorSteps;
    float eta = slice == kBrdfIorSteps ? 0.f : q / fmaxf(1.f - q, 1e-20f);
    float no = fmaxf((float)cx / (float)(n - 1), kGgxAlphaMin);
    float r = (float)cy / (float)(n - 1);
    float result = 0.f;
    if (eta == 1.f) {
        result = 0.f;
    } else if (r == 0.f) {
        result = eta <= 0.f ? 1.f : fresnel(no, eta);
    } else {
        float a = specular_alpha(r);
        for (uint32_t i = 0; i < kDielectricAlbedoSamples; i++) {
            float u0 = ((float)i + .5f) / (float)kDielectricAlbedoSamples;
            float u1 = (float)(__builtin_bitreverse32(i) >> 8) * kUintScale;
            float vh;
            float w = ggx_visible_reflection(no, a, u0, u1, &vh);
            result += (eta <= 0.f ? 1.f : fresnel(vh, eta)) * w;
        }
        result /= (float)kDielectricAlbedoSamples;
    }
    out[k] = result;
}

inline uint32_t blocks(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

// Writes one batch's launch block (a kernel, so the write is ordered on the
// stream with the batch before and after it, whatever the host does next).
__global__ void k_store_args(LaunchArgs* dst, LaunchArgs value) {
    constexpr uint32_t words = sizeof(LaunchArgs) / 4;
    static_assert(sizeof(LaunchArgs) % 4 == 0, "launch block in words");
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&value);
    uint32_t* out = reinterpret_cast<uint32_t*>(dst);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) out[i] = src[i];
}

// Sampler draws for a list of (px, py, index, group) queries, through the
// same group_bits / path_random the path kernels call.
__global__ void k_sampler_bits(const uint4* queries, uint32_t n, uint32_t samples, uint32_t seed, uint4* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 q = queries[i];
    uint32_t b[4];
    if (q.w == VH_SAMPLER_PATH_RANDOM) {
        uint32_t state = screen_ray_hash(q.x, q.y, q.z, seed);
        b[0] = state;
        for (int k = 1; k < 4; k++) {
            (void)path_random(state);
            b[k] = state;
        }
    } else {
        group_bits(make_sampler(q.x, q.y, q.z, samples, seed), q.w, b);
    }
    out[i] = make_uint4(b[0], b[1], b[2], b[3]);
}

}  // namespace

int launch_sampler_bits(hipStream_t s, const uint32_t* queries, uint32_t n, uint32_t samples, uint32_t seed,
                        uint32_t* out) {
    hipLaunchKernelGGL(k_sampler_bits, dim3(blocks(n, kBlock)), dim3(kBlock), 0, s,
                       reinterpret_cast<const uint4*>(queries), n, samples, seed, reinterpret_cast<uint4*>(out));
    VH_HIP(hipGetLastError());
    return VH_OK;
}

int bake_luts(vh_device* dev) {
    hipStream_t s = dev->stream;
    VH_HIP(dev->lut_moments.allocate(sizeof(float4) * kBrdfSize * kBrdfSize));
    VH_HIP(dev->lut_dielectric.allocate(sizeof(float) * (kBrdfIorSteps + 1) * kBrdfAlbedoSize * kBrdfAlbedoSize));
    hipLaunchKernelGGL(k_bake_moments, dim3(blocks(kBrdfSize * kBrdfSize, 64)), dim3(64), 0, s,
                       dev->lut_moments.as<float4>());
    hipLaunchKernelGGL(k_bake_dielectric, dim3(blocks((kBrdfIorSteps + 1) * kBrdfAlbedoSize * kBrdfAlbedoSize, 256)),
                       dim3(256), 0, s, dev->lut_dielectric.as<float>());
    VH_HIP(hipGetLastError());
    VH_HIP(hipStreamSynchronize(s));
    return VH_OK;
}

int ensure_overflow(DeviceBuffer& buf, uint32_t threads);

namespace {
// The launch sequence of one batch, fixed by these values (its kernels read
// everything else from the launch block).
struct BatchShape {
    uint32_t bounces = 0, grid = 0;
    uint32_t paths = 0, pixels = 0;  // one thread per path (camera) and per pixel (decide, film)
    uint32_t mode = 0;  // 0 plain, 1 adaptive up to `minimum`, 2 adaptive after a decide pass
    bool count = false, packet = false, sort = false;
    StreamState::GraphKey key() const {
        return {bounces, grid, paths, pixels, mode, count ? 1u : 0u, packet ? 1u : 0u, sort ? 1u : 0u};
    }
};
// Graphs a stream keeps (the oldest goes first): renders of a few image sizes
// and batch layouts in turn each find theirs.
constexpr size_t kGraphsPerStream = 8;

// A kernel's arguments converted to its parameter types, and the array of
// their addresses that hipLaunchKernel and kernel graph nodes take.
template <class... P>
struct KernelArgs {
    std::tuple<P...> values;
    void* ptrs[sizeof...(P) > 0 ? sizeof...(P) : 1];
    template <class... A>
    explicit KernelArgs(A... a) : values(static_cast<P>(a)...) {
        std::apply([&](auto&... v) {
            int i = 0;
            ((ptrs[i++] = static_cast<void*>(&v)), ...);
        }, values);
    }
    KernelArgs(const KernelArgs&) = delete;
};

// Launches a kernel directly on a stream.
struct StreamLauncher {
    hipStream_t s;
    template <class... P, class... A>
    void operator()(void (*k)(P...), dim3 grid, dim3 block, A... a) {
        KernelArgs<P...> args(a...);
        (void)hipLaunchKernel(reinterpret_cast<const void*>(k), grid, block, args.ptrs, 0, s);
    }
};

// Appends a kernel node to a graph after the previous one: the batch's
// graph is built node by node rather than captured from a stream, so other
// threads' work on other streams (the serving loop's slots, torch's own
// launches) cannot invalidate it while it is being recorded.
struct GraphLauncher {
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
    hipError_t error = hipSuccess;
    template <class... P, class... A>
    void operator()(void (*k)(P...), dim3 grid, dim3 block, A... a) {
        if (error != hipSuccess) return;
        KernelArgs<P...> args(a...);  // copied into the node by hipGraphAddKernelNode
        hipKernelNodeParams np{};
        np.func = reinterpret_cast<void*>(k);
        np.gridDim = grid;
        np.blockDim = block;
        np.sharedMemBytes = 0;
        np.kernelParams = args.ptrs;
        np.extra = nullptr;
        hipGraphNode_t node = nullptr;
        error = hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &np);
        if (error == hipSuccess) last = node;
    }
};

// Issues one batch's kernels through `launch` (a stream or a graph).
// `mark(stage, begin)` is called around the extend (0; 3 for the camera
// bounce), shade (1) and shadow (2) launches.
template <class Launch, class Mark>
int launch_batch(const BatchShape& b, const LaunchArgs* A, Launch& launch, Mark&& mark) {
    const dim3 grid(b.grid), block(kBlock), one(1), wave(64);
    const dim3 per_path(blocks(b.paths, kBlock)), per_pixel(blocks(b.pixels, kBlock));
    if (b.mode == 2) launch(k_decide, per_pixel, block, A);
    if (b.mode != 0) {
        launch(k_open_batch, one, wave, A);
        launch(k_camera<true>, per_path, block, A);
    } else {
        launch(k_camera<false>, per_path, block, A);
    }
    int rc;
    for (uint32_t bounce = 0; bounce < b.bounces; bounce++) {
        const int stage = bounce == 0 ? 3 : 0;
        if ((rc = mark(stage, true))) return rc;
        if (b.count) launch(k_extend<true>, grid, block, A, bounce);
        else if (b.packet && bounce == 0) launch(k_extend_packet, grid, block, A);
        else launch(k_extend<false>, grid, block, A, bounce);
        if ((rc = mark(stage, false))) return rc;
        if ((rc = mark(1, true))) return rc;
        if (b.sort) launch(k_classify, grid, block, A, bounce);
        launch(k_shade<false>, grid, block, A, bounce);
        if ((rc = mark(1, false))) return rc;
        if ((rc = mark(2, true))) return rc;
        if (b.count) launch(k_shadow<true>, grid, block, A);
        else launch(k_shadow<false>, grid, block, A);
        if ((rc = mark(2, false))) return rc;
        launch(k_resolve, grid, block, A);
        launch(k_swap_counts, one, wave, A);
    }
    if (b.mode != 0) launch(k_film<true>, per_pixel, block, A);
    else launch(k_film<false>, per_pixel, block, A);
    return VH_OK;
}

// Replays the stream's graph of this batch shape on `s`, building it first
// if the stream has none yet.
int launch_batch_graph(StreamState* st, const BatchShape& b, const LaunchArgs* A, hipStream_t s) {
    const StreamState::GraphKey key = b.key();
    hipGraphExec_t exec = nullptr;
    for (const auto& g : st->graphs)
        if (g.first == key) exec = static_cast<hipGraphExec_t>(g.second);
    if (!exec) {
        if (st->graphs.size() >= kGraphsPerStream) {
            (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(st->graphs.front().second));
            st->graphs.erase(st->graphs.begin());
        }
        GraphLauncher g;
        VH_HIP(hipGraphCreate(&g.graph, 0));
        int rc = launch_batch(b, A, g, [](int, bool) { return VH_OK; });
        hipError_t e = g.error;
        if (!rc && e == hipSuccess) e = hipGraphInstantiate(&exec, g.graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g.graph);
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail(e, "vh_render: batch graph");
        st->graphs.emplace_back(key, exec);
    }
    VH_HIP(hipGraphLaunch(exec, s));
    return VH_OK;
}
}  // namespace

namespace {
// Grows the stream's wavefront buffers to `want` paths (path state, two
// extension queues, the estimator-a and -b shadow queues); on failure the
// stream holds none.
hipError_t reserve_wavefront(StreamState* st, size_t want) {
    if (st->capacity >= want) return hipSuccess;
    st->capacity = 0;
    hipError_t e = st->paths.allocate(want * kPathStateBytes);
    if (e == hipSuccess) e = st->rays_a.allocate(want * 2 * sizeof(float4));
    if (e == hipSuccess) e = st->rays_b.allocate(want * 2 * sizeof(float4));
    if (e == hipSuccess) e = st->shadow_rays.allocate(2 * want * 2 * sizeof(float4));  // estimator a, then b
    if (e == hipSuccess) {
        st->capacity = want;
        return hipSuccess;
    }
    st->paths.release();
    st->rays_a.release();
    st->rays_b.release();
    st->shadow_rays.release();
    return e;
}
// The path-state arrays of `cap` paths in the stream's buffer.
Paths path_layout(StreamState* st, size_t cap) {
    Paths P;
    char* base = st->paths.as<char>();
    float4** f4[] = {&P.thr, &P.rad, (float4**)&P.info, &P.hit, &P.pend, &P.pend1, &P.guide_a, &P.guide_n};
    size_t off = 0;
    for (auto* q : f4) {
        *q = reinterpret_cast<float4*>(base + off);
        off += cap * sizeof(float4);
    }
    P.hit_inst = reinterpret_cast<uint32_t*>(base + off);
    off += cap * sizeof(uint32_t);
    P.vis = reinterpret_cast<uint8_t*>(base + off);
    return P;
}
Scene scene_tables(const vh_scene* sc) {
    const vh_device* dev = sc->device;
    Scene S;
    S.tv = sc->tlas->view();
    S.materials = sc->materials.as<vh_material>();
    S.lights = sc->lights.as<vh_rect_light>();
    S.tree = sc->light_tree.as<LightNode>();
    S.filter = sc->filter.as<float>();
    S.luts.moments = dev->lut_moments.as<float4>();
    S.luts.dielectric = dev->lut_dielectric.as<float>();
    return S;
}
// A render's constants (the batch fields and the adaptive minimum are set per batch).
Consts scene_consts(const vh_scene* sc, const vh_render_params* rp, bool guides) {
    Consts c;
    std::memset(&c, 0, sizeof(c));
    c.width = rp->width; c.height = rp->height; c.samples = rp->samples; c.seed = rp->seed;
    c.max_bounces = std::min(rp->max_bounces, 254u);
    c.diffuse_bounces = std::min(rp->diffuse_bounces, 14u);
    c.glossy_bounces = std::min(rp->glossy_bounces, 14u);
    c.caustics = rp->caustics;
    c.clamp_direct = rp->clamp_direct; c.clamp_indirect = rp->clamp_indirect; c.light_threshold = rp->light_threshold;
    c.light_count = sc->light_count; c.light_nodes = sc->light_nodes; c.material_count = sc->material_count;
    c.pixels = rp->width * rp->height;
    c.threshold = rp->adaptive_threshold > 0.f ? rp->adaptive_threshold : 0.01f;
    c.guides = guides;
    c.camera = sc->camera;
    c.sky = sc->sky;
    return c;
}
}  // namespace

int render_scene(vh_scene* sc, const vh_render_params* rp, float* film, float* guides, hipStream_t s) {
    vh_device* dev = sc->device;
    const uint32_t pixels = rp->width * rp->height;
    const uint32_t total = rp->sample_end - rp->sample_begin;
    vh_render_stats stats{};
    if (total == 0) {
        std::lock_guard<std::mutex> hs(sc->state_lock);
        sc->stats = stats;
        return VH_OK;
    }
    uint32_t batch = rp->batch_samples;
    if (batch == 0) {
        // Up to 2^VH_BATCH_PATHS_LOG2 paths in flight (262 B each; 2^27 is
        // about 35 GB, a whole 1920x1080 x 64-spp exposure), whole exposure
        // samples per batch.  Every wavefront launch pays a tail while its
        // slowest rays finish, so fewer, longer launches are faster.
        batch = std::max<uint32_t>(1u, (1u << VH_BATCH_PATHS_LOG2) / std::max<uint32_t>(pixels, 1u));
    }
    batch = std::min(batch, total);
    if (pixels > kPathMask) return fail(VH_ERR_USAGE, "vh_render: more than 2^30 - 1 pixels");
    batch = std::min<uint32_t>(batch, kPathMask / pixels);
    StreamState* st = stream_state(dev, s);
    std::lock_guard<std::mutex> hold(st->lock);
    // grow the stream's wavefront state; halve the batch while HBM is short
    while (st->capacity < (size_t)pixels * batch) {
        const hipError_t e = reserve_wavefront(st, (size_t)pixels * batch);
        if (e == hipSuccess) break;
        if (e != hipErrorOutOfMemory || batch == 1) return hip_fail(e, "vh_render: wavefront state");
        batch = (batch + 1) / 2;
    }
    const size_t cap = (size_t)pixels * batch;
    VH_HIP(st->counters.reserve(64 * sizeof(uint32_t)));
    const bool adaptive = (rp->flags & VH_RENDER_ADAPTIVE) != 0;
    // path.slang:1801-1802: minimum max(16, sqrt(samples)) and threshold 0.01
    // unless set; a minimum at or beyond the exposure means no pixel stops.
    const uint32_t minimum = std::min<uint32_t>(
        rp->adaptive_min_samples ? rp->adaptive_min_samples
                                 : std::max<uint32_t>(16u, (uint32_t)std::sqrt((float)rp->samples)),
        rp->samples);
    Moment* moments = nullptr;
    vh_scene::Exposure* ex = nullptr;
    if (adaptive) {
        {
            std::lock_guard<std::mutex> hs(sc->state_lock);
            std::unique_ptr<vh_scene::Exposure>& e = sc->exposures[s];
            if (!e) e.reset(new vh_scene::Exposure());
            ex = e.get();  // used only under the stream's lock, which this call holds
        }
        if (rp->sample_begin == 0) {
            VH_HIP(ex->moments.reserve((size_t)pixels * sizeof(Moment)));
            ex->width = rp->width;
            ex->height = rp->height;
            ex->seed = rp->seed;
            ex->samples = rp->samples;
        } else if (ex->next != rp->sample_begin || ex->width != rp->width || ex->height != rp->height ||
                   ex->seed != rp->seed || ex->samples != rp->samples) {
            return fail(VH_ERR_USAGE, "vh_render: an adaptive sample range must continue the exposure rendered "
                                      "before it on this scene and stream");
        }
        ex->next = UINT32_MAX;  // set again once this range is rendered
        moments = ex->moments.as<Moment>();
    }
    const Paths P = path_layout(st, cap);
    const Scene S = scene_tables(sc);
    Consts c = scene_consts(sc, rp, guides != nullptr);
    c.minimum = minimum;
    const uint32_t grid = wavefront_grid(cap, (uint32_t)dev->cu_count);
    int rc = ensure_overflow(st->overflow, grid * kBlock);
    if (rc) return rc;
    VH_HIP(st->launch_args.reserve(sizeof(LaunchArgs)));
    uint32_t* counters = st->counters.as<uint32_t>();
    VH_HIP(hipMemsetAsync(counters, 0, kCounterWords * sizeof(uint32_t), s));
    const bool profile = (rp->flags & VH_RENDER_PROFILE) != 0;
    const bool count = (rp->flags & VH_RENDER_COUNT_VISITS) != 0;
    LaunchArgs args;
    std::memset(&args, 0, sizeof(args));
    args.P = P;
    args.S = S;
    args.rays[0] = st->rays_a.as<float4>();
    args.rays[1] = st->rays_b.as<float4>();
    args.shadow_a = st->shadow_rays.as<float4>();
    args.shadow_b = args.shadow_a + 2 * cap;
    args.counters = counters;
    args.overflow = st->overflow.as<uint32_t>();
    args.film = reinterpret_cast<float4*>(film);
    args.guides = reinterpret_cast<float4*>(guides);
    args.moments = moments;
    args.cap = (uint32_t)cap;
    const bool sort = (rp->flags & VH_RENDER_SHADE_SORTED) ? true
                      : (rp->flags & VH_RENDER_SHADE_ARRIVAL) ? false : VH_SHADE_SORT != 0;
    if (sort) {
        VH_HIP(st->sorted.reserve(kClasses * cap * sizeof(uint32_t)));
        args.sorted = st->sorted.as<uint32_t>();
    }
    const LaunchArgs* dargs = st->launch_args.as<LaunchArgs>();
    // stage events: (stage, start, stop) triples recorded around each launch
    // when profiling (stage 3: the camera bounce's extend launch, also
    // counted in stage 0)
    struct Mark { int stage; hipEvent_t a, b; };
    std::vector<Mark> marks;
    hipEvent_t e0, e1;
    VH_HIP(hipEventCreate(&e0));
    VH_HIP(hipEventCreate(&e1));
    VH_HIP(hipEventRecord(e0, s));
    for (uint32_t first = rp->sample_begin; first < rp->sample_end;) {
        uint32_t nb = std::min(batch, rp->sample_end - first);
        c.decide = 0;
        if (adaptive) {
            if (first <= minimum) nb = (uint32_t)std::min<uint64_t>(nb, (uint64_t)minimum + 1u - first);
            else { nb = 1; c.decide = 1; }
        }
        c.batch_first = first;
        c.batch_samples = nb;
        c.group = sample_group(nb);
        args.c = c;
        args.n = pixels * nb;
        hipLaunchKernelGGL(k_store_args, dim3(1), dim3(kBlock), 0, s, const_cast<LaunchArgs*>(dargs), args);
        BatchShape shape;
        shape.bounces = c.max_bounces + 1;
        shape.grid = grid;
        shape.paths = args.n;
        shape.pixels = pixels;
        shape.mode = !adaptive ? 0u : c.decide ? 2u : 1u;
        shape.count = count;
        shape.packet = VH_PACKET_CAMERA && !count && c.group >= VH_PACKET_MIN_GROUP;
        shape.sort = sort;
        if (profile) {
            // launched one by one, with events around the stage launches
            StreamLauncher direct{s};
            rc = launch_batch(shape, dargs, direct, [&](int stage, bool begin) -> int {
                if (begin) {
                    marks.push_back(Mark{stage, nullptr, nullptr});
                    VH_HIP(hipEventCreate(&marks.back().a));
                    VH_HIP(hipEventCreate(&marks.back().b));
                    VH_HIP(hipEventRecord(marks.back().a, s));
                } else {
                    VH_HIP(hipEventRecord(marks.back().b, s));
                }
                return VH_OK;
            });
        } else {
            rc = launch_batch_graph(st, shape, dargs, s);
        }
        if (rc) return rc;
        VH_HIP(hipGetLastError());
        first += nb;
    }
    VH_HIP(hipEventRecord(e1, s));
    uint32_t h[kCounterWords];
    VH_HIP(hipMemcpyAsync(h, counters, sizeof(h), hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    double stage_ms[4] = {0, 0, 0, 0};
    uint64_t stage_n[4] = {0, 0, 0, 0};
    for (const Mark& m : marks) {
        float t = 0;
        (void)hipEventElapsedTime(&t, m.a, m.b);
        stage_ms[m.stage] += t;
        stage_n[m.stage]++;
        if (m.stage == 3) {  // the camera bounce is part of the extend stage
            stage_ms[0] += t;
            stage_n[0]++;
        }
        (void)hipEventDestroy(m.a);
        (void)hipEventDestroy(m.b);
    }
    uint64_t ext, shd, stopped, v[4], cv[2], crays;
    std::memcpy(&stopped, &h[10], 8);
    std::memcpy(&ext, &h[4], 8);
    std::memcpy(&shd, &h[6], 8);
    std::memcpy(v, &h[16], 32);
    std::memcpy(cv, &h[kCameraVisits], 16);
    std::memcpy(&crays, &h[kCameraRays], 8);
    v[0] += cv[0];
    v[1] += cv[1];
    stats.extend_camera_ms = stage_ms[3];
    stats.extend_camera_launches = stage_n[3];
    stats.extend_camera_rays = crays;
    stats.extend_camera_node_visits = cv[0];
    stats.extend_camera_triangle_tests = cv[1];
    stats.extend_ms = stage_ms[0];
    stats.shade_ms = stage_ms[1];
    stats.shadow_ms = stage_ms[2];
    stats.other_ms = profile ? ms - stage_ms[0] - stage_ms[1] - stage_ms[2] : 0.0;
    stats.extend_launches = stage_n[0];
    stats.shade_launches = stage_n[1];
    stats.shadow_launches = stage_n[2];
    stats.extend_node_visits = v[0];
    stats.extend_triangle_tests = v[1];
    stats.shadow_node_visits = v[2];
    stats.shadow_triangle_tests = v[3];
    stats.paths = (uint64_t)pixels * total - stopped;
    stats.stopped_samples = stopped;
    stats.extension_rays = ext;
    stats.shadow_rays = shd;
    stats.dropped_samples = h[9];
    stats.device_ms = ms;
    if (h[8] & kStackOverflow) return fail(VH_ERR_DEVICE, "traversal stack overflow during render");
    {
        std::lock_guard<std::mutex> hs(sc->state_lock);
        sc->stats = stats;
    }
    if (adaptive) ex->next = rp->sample_end;
    return VH_OK;
}


// Several jobs in one wavefront (vh_render_shared): every batch gives each job
// with samples left a share of the path slots; the jobs' camera rays fill one
// extension queue, their hits one shading queue and their shadow rays one
// shadow queue, each path finding its job by its slot.  A path's arithmetic
// does not depend on the batch it is in, and each job's film sums its samples
// in sample order, so each film equals the job rendered alone.
int render_shared(const vh_render_job* jobs, uint32_t count, hipStream_t s) {
    vh_device* dev = jobs[0].scene->device;
    const size_t cap = std::min<size_t>((size_t)1 << VH_BATCH_PATHS_LOG2, (size_t)1 << 27);
    for (uint32_t j = 0; j < count; j++) {
        const vh_render_params& rp = jobs[j].params;
        if ((size_t)rp.width * rp.height > cap)
            return fail(VH_ERR_USAGE, "vh_render_shared: job " + std::to_string(j) + " has more pixels than a batch");
    }
    std::vector<vh_render_stats> stats(count);
    StreamState* st = stream_state(dev, s);
    std::lock_guard<std::mutex> hold(st->lock);
    VH_HIP(reserve_wavefront(st, cap));
    VH_HIP(st->counters.reserve(kCounterWords * sizeof(uint32_t)));
    VH_HIP(st->launch_args.reserve(sizeof(LaunchArgs)));
    VH_HIP(st->jobs.reserve(kMaxJobs * sizeof(JobArgs)));
    const uint32_t grid = wavefront_grid(cap, (uint32_t)dev->cu_count);
    int rc = ensure_overflow(st->overflow, grid * kBlock);
    if (rc) return rc;
    uint32_t* counters = st->counters.as<uint32_t>();
    VH_HIP(hipMemsetAsync(counters, 0, kCounterWords * sizeof(uint32_t), s));
    LaunchArgs args;
    std::memset(&args, 0, sizeof(args));
    args.P = path_layout(st, cap);
    args.rays[0] = st->rays_a.as<float4>();
    args.rays[1] = st->rays_b.as<float4>();
    args.shadow_a = st->shadow_rays.as<float4>();
    args.shadow_b = args.shadow_a + 2 * cap;
    args.counters = counters;
    args.overflow = st->overflow.as<uint32_t>();
    args.cap = (uint32_t)cap;
    args.jobs = st->jobs.as<JobArgs>();
    // the first job's tables stand in for the kernel-wide ones the shared
    // kernels do not read
    args.S = scene_tables(jobs[0].scene);
    args.c = scene_consts(jobs[0].scene, &jobs[0].params, false);
    LaunchArgs* dargs = st->launch_args.as<LaunchArgs>();
    std::vector<uint32_t> next(count);
    for (uint32_t j = 0; j < count; j++) next[j] = jobs[j].params.sample_begin;
    std::vector<std::vector<JobArgs>> tables;  // kept until the copies are done
    hipEvent_t e0, e1;
    VH_HIP(hipEventCreate(&e0));
    VH_HIP(hipEventCreate(&e1));
    VH_HIP(hipEventRecord(e0, s));
    while (true) {
        std::vector<uint32_t> live;
        for (uint32_t j = 0; j < count; j++)
            if (next[j] < jobs[j].params.sample_end) live.push_back(j);
        if (live.empty()) break;
        // an equal share of the slots per job, at least one sample each while
        // they fit; jobs that do not fit wait for the next batch
        const size_t share = cap / live.size();
        std::vector<JobArgs> table;
        size_t used = 0;
        uint32_t tiles = 0, bounces = 0;
        for (uint32_t j : live) {
            const vh_render_params& rp = jobs[j].params;
            const size_t pixels = (size_t)rp.width * rp.height;
            uint32_t nb = (uint32_t)std::min<size_t>(rp.sample_end - next[j], std::max<size_t>(1, share / pixels));
            while (nb > 0 && used + pixels * nb > cap) nb--;
            if (nb == 0 || table.size() == kMaxJobs) continue;
            JobArgs J;
            std::memset(&J, 0, sizeof(J));
            J.c = scene_consts(jobs[j].scene, &rp, jobs[j].guides != nullptr);
            J.c.minimum = rp.samples;
            J.c.batch_first = next[j];
            J.c.batch_samples = nb;
            J.c.group = sample_group(nb);
            J.S = scene_tables(jobs[j].scene);
            J.film = reinterpret_cast<float4*>(jobs[j].film);
            J.guides = reinterpret_cast<float4*>(jobs[j].guides);
            J.path_base = (uint32_t)used;
            J.paths = (uint32_t)(pixels * nb);
            J.tile_base = tiles;
            J.tiles = blocks(pixels, kBlock);
            table.push_back(J);
            used += pixels * nb;
            tiles += J.tiles;
            bounces = std::max(bounces, J.c.max_bounces + 1);
            stats[j].paths += pixels * nb;
            next[j] += nb;
        }
        tables.push_back(std::move(table));
        const std::vector<JobArgs>& t = tables.back();
        VH_HIP(hipMemcpyAsync(st->jobs.ptr, t.data(), t.size() * sizeof(JobArgs), hipMemcpyHostToDevice, s));
        args.n = (uint32_t)used;
        args.job_count = (uint32_t)t.size();
        hipLaunchKernelGGL(k_store_args, dim3(1), dim3(kBlock), 0, s, dargs, args);
        const LaunchArgs* A = dargs;
        const dim3 g(grid), block(kBlock), one(1), wave(64);
        hipLaunchKernelGGL(k_camera_shared, dim3(blocks(used, kBlock)), block, 0, s, A);
        for (uint32_t bounce = 0; bounce < bounces; bounce++) {
            hipLaunchKernelGGL((k_extend<false, true>), g, block, 0, s, A, bounce);
            hipLaunchKernelGGL(k_shade<true>, g, block, 0, s, A, bounce);
            hipLaunchKernelGGL((k_shadow<false, true>), g, block, 0, s, A);
            hipLaunchKernelGGL(k_resolve, g, block, 0, s, A);
            hipLaunchKernelGGL(k_swap_counts, one, wave, 0, s, A);
        }
        hipLaunchKernelGGL(k_film_shared, dim3(std::min<uint32_t>(tiles, grid)), block, 0, s, A);
        VH_HIP(hipGetLastError());
    }
    VH_HIP(hipEventRecord(e1, s));
    uint32_t h[kCounterWords];
    VH_HIP(hipMemcpyAsync(h, counters, sizeof(h), hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    uint64_t ext, shd;
    std::memcpy(&ext, &h[4], 8);
    std::memcpy(&shd, &h[6], 8);
    for (uint32_t j = 0; j < count; j++) {
        // rays, drops and time belong to the shared batches as a whole
        stats[j].extension_rays = ext;
        stats[j].shadow_rays = shd;
        stats[j].dropped_samples = h[9];
        stats[j].device_ms = ms;
        std::lock_guard<std::mutex> hs(jobs[j].scene->state_lock);
        jobs[j].scene->stats = stats[j];
    }
    if (h[8] & kStackOverflow) return fail(VH_ERR_DEVICE, "traversal stack overflow during render");
    return VH_OK;
}

}  // namespace vh
