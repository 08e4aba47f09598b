// Library-internal objects behind the opaque handles of include/vork_hip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vork_hip.h"
#include "vh_bvh.h"

namespace vh {

// Status + thread-local message; every C entry point funnels failures here.
int fail(int code, const std::string& message);
int hip_fail(hipError_t e, const char* what);
#define VH_HIP(call)                                   \
    do {                                               \
        hipError_t _e = (call);                        \
        if (_e != hipSuccess) return hip_fail(_e, #call); \
    } while (0)

// A device allocation that frees itself.
// Device blocks of released BLAS, TLAS and scene objects, kept per device for
// the next object's buffers of about the same size (a scene rebuilt every
// frame otherwise pays hipMalloc and hipFree of its whole storage each time).
// A block enters the cache only once no work reads it: those objects are
// released after their calls, which all return synchronised.  pool_alloc
// returns a cached block of n to n + n/4 bytes (the size in *bytes) or a new
// one; pool_free keeps up to 16 GiB per device and frees the rest; pool_trim
// frees a device's cached blocks.
hipError_t pool_alloc(void** p, size_t* bytes, int* device);
void pool_free(void* p, size_t bytes, int device);
void pool_trim(int device);

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool pooled = false;  // allocated from and released to the block cache
    int device = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(bool pooled_) : pooled(pooled_) {}
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    hipError_t allocate(size_t n) {
        release();
        if (n == 0) n = 16;
        if (pooled) {
            size_t b = n;
            const hipError_t e = pool_alloc(&ptr, &b, &device);
            if (e == hipSuccess) bytes = b; else ptr = nullptr;
            return e;
        }
        hipError_t e = hipMalloc(&ptr, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    // Grows to at least n bytes; contents are not preserved.
    hipError_t reserve(size_t n) { return n <= bytes ? hipSuccess : allocate(n); }
    void release() {
        if (ptr) {
            if (pooled) pool_free(ptr, bytes, device);
            else (void)hipFree(ptr);
        }
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Scratch of the BVH builder: every temporary of a BLAS or TLAS build, grown
// to the largest build on the stream and reused (allocating and freeing a few
// hundred MB per build cost more than the build's kernels).
struct BuildScratch {
    DeviceBuffer src, src_nrm, lo, hi, bounds, keys, keys_sorted, ids, ids_sorted, left, right, size, nb, ca, cb, nn,
        flags, offsets, sort_tmp, qa, qb, counters, sah, tile_sums, nodes, order, small, dp_cost, dp_dec;
};

// Wavefront and traversal scratch of one stream.  Calls on one stream are
// ordered on the device, so they share it; `lock` is held for the whole
// (synchronous) call, so calls on one stream from several threads wait for
// each other while calls on different streams overlap.
struct StreamState {
    std::mutex lock;
    DeviceBuffer paths, rays_a, rays_b, shadow_rays, counters, overflow, scratch_u32, launch_args, sorted, jobs;
    size_t capacity = 0;  // paths the wavefront buffers hold
    BuildScratch build;
    // instantiated launch sequences of render batches by shape (hipGraphExec_t),
    // built node by node and replayed on the calling stream
    using GraphKey = std::array<uint32_t, 8>;
    std::vector<std::pair<GraphKey, void*>> graphs;
    StreamState() = default;
    StreamState(const StreamState&) = delete;
    StreamState& operator=(const StreamState&) = delete;
    ~StreamState() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(g.second));
    }
};

}  // namespace vh

struct vh_device {
    // one reference for the caller's handle and one per BLAS, TLAS and scene
    // built on the device: it is destroyed when the last is released
    std::atomic<int> refs{1};
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    // GGX energy tables (moments 128x128 float4, dielectric 65x64x64 float)
    vh::DeviceBuffer lut_moments, lut_dielectric;
    // per-stream scratch, created on a stream's first render or trace call
    // and kept (grown to the largest call) until the device is released
    std::mutex streams_lock;
    std::map<hipStream_t, std::unique_ptr<vh::StreamState>> streams;
    // vh_trace_closest_static: this GPU's address of the library's static
    // scratch array (looked up once, at device creation) and a pinned host
    // word its error flags are read back into
    void* static_scratch = nullptr;
    uint32_t* static_host = nullptr;
};

namespace vh {
void device_retain(vh_device* d);
void device_drop(vh_device* d);

// A counted reference to the device an object was built on.  Declared first
// in each object, so it is dropped after the object's device memory is freed.
class DeviceRef {
  public:
    DeviceRef() = default;
    DeviceRef(const DeviceRef&) = delete;
    DeviceRef& operator=(const DeviceRef&) = delete;
    ~DeviceRef() { reset(nullptr); }
    DeviceRef& operator=(vh_device* d) { reset(d); return *this; }
    operator vh_device*() const { return dev_; }
    vh_device* operator->() const { return dev_; }
  private:
    void reset(vh_device* d) {
        if (d) device_retain(d);
        if (dev_) device_drop(dev_);
        dev_ = d;
    }
    vh_device* dev_ = nullptr;
};
}  // namespace vh

struct vh_blas {
    vh::DeviceRef device;  // first: dropped after the buffers below are freed
    vh::DeviceBuffer nodes{true}, tris{true}, tri_ids{true}, positions{true}, normals{true}, indices{true},
        materials{true}, surf{true};
    vh::BlasRecord record{};
    vh_blas_info info{};
};

struct vh_tlas {
    vh::DeviceRef device;  // first: dropped after the buffers below are freed
    vh::DeviceBuffer nodes{true}, inst_ids{true}, instances{true}, blas_records{true};
    std::vector<vh_blas*> blases;
    uint32_t instance_count = 0;
    uint32_t node_count = 0;
    vh::TlasView view() const {
        vh::TlasView v;
        v.nodes = nodes.as<vh::WideNode>();
        v.inst_ids = inst_ids.as<uint32_t>();
        v.instances = instances.as<vh::InstanceRecord>();
        v.blases = blas_records.as<vh::BlasRecord>();
        v.node_count = node_count;
        v.instance_count = instance_count;
        return v;
    }
};

// Light-tree node (crates/vork-render/shaders/light_tree.slang:3-8).
struct vh_light_node {
    float lower_power[4];
    float upper_infinite[4];
    uint32_t links[4];
    uint32_t source[4];
};

struct vh_scene {
    vh::DeviceRef device;  // first: dropped after the buffers below are freed
    const vh_tlas* tlas = nullptr;
    vh::DeviceBuffer materials{true}, lights{true}, light_tree{true}, filter{true};
    uint32_t material_count = 0, light_count = 0, light_nodes = 0;
    vh_camera camera{};
    vh_sky sky{};
    // The scene's tables are not written after vh_scene_create (but for
    // vh_scene_set_camera), so renders of the scene on different streams run
    // at once.  An adaptive exposure in progress keeps its per-pixel moments
    // (48 B each) and its size, seed, sample count and next sample index per
    // stream: calls on one stream are ordered, so the stream's own exposure
    // needs no lock; `state_lock` guards the map and the stats.
    struct Exposure {
        vh::DeviceBuffer moments;
        uint32_t width = 0, height = 0, seed = 0, samples = 0, next = 0;
    };
    std::mutex state_lock;
    std::map<hipStream_t, std::unique_ptr<Exposure>> exposures;
    vh_render_stats stats{};  // of the last render of the scene that finished
};

namespace vh {

// ---------------------------------------------------------------- build
// Generic BVH over n boxes (Morton order + clustering), collapsed to 8-wide nodes.  `lo`/`hi` are device
// arrays of 3 floats per primitive.  Writes the nodes and the leaf-order
// permutation `order` (leaf item -> primitive) and returns the node count.
struct WideBuild {
    const WideNode* nodes = nullptr;  // in the scratch: valid until the next build on the stream
    const uint32_t* order = nullptr;  // leaf item -> primitive, likewise
    uint32_t node_count = 0;
    uint32_t binary_nodes = 0;
    uint32_t max_depth = 0;
    double sah_cost = 0;
};
int build_wide_bvh(vh_device* dev, hipStream_t stream, BuildScratch& scratch, const float* d_lo, const float* d_hi,
                   uint32_t n, uint32_t leaf_max, WideBuild* out);

// Kernels of the BLAS build that depend on the mesh layout.
int launch_triangle_bounds(hipStream_t s, BuildScratch& scratch, const float* positions, uint32_t stride_floats, const uint32_t* indices,
                           uint32_t n, float* lo, float* hi, float* compact_positions, uint32_t vertex_count);
// Leaf-order copies of the triangles: vertex positions, vertex normals and
// materials (when the mesh has them) and the original primitive index.
int launch_gather_triangles(hipStream_t s, const float* positions, const float* normals, const uint32_t* materials,
                            const uint32_t* indices, const uint32_t* order, uint32_t n, float* tris,
                            float* surf, uint32_t* tri_ids);
int launch_instance_bounds(hipStream_t s, const InstanceRecord* inst, const BlasRecord* blases, uint32_t n, float* lo,
                           float* hi);
int launch_compact3(hipStream_t s, const float* src, uint32_t stride_floats, uint32_t count, float* dst);
int launch_bounds_reduce(hipStream_t s, BuildScratch& scratch, const float* lo, const float* hi, uint32_t n,
                         float out[6]);
int launch_instance_records(hipStream_t s, const vh_instance* inst, uint32_t n, uint32_t blas_count,
                            InstanceRecord* out);

// ---------------------------------------------------------------- trace
// The device address of the static scratch of the allocation-free closest-hit
// entry point on the current device (hipGetSymbolAddress).
int static_scratch_address(void** out);
// Closest hit of `count` device-resident rays into device-resident hits with
// the static scratch only; `dev->static_scratch` must be set.
int launch_trace_static(vh_device* dev, hipStream_t s, const TlasView& tv, const vh_ray* rays, uint32_t count,
                        uint32_t mask, vh_hit* hits);
// Largest ray count one vh_trace_closest_static call takes.
constexpr uint32_t kStaticTraceMaxRays = 1u << 31;
int launch_trace(vh_device* dev, hipStream_t s, const TlasView& tv, const void* rays, uint32_t count, uint32_t mask,
                 void* out, int mode);  // mode 0 closest -> vh_hit, 1 any -> uint8, 2 counts -> 4 uint32 per ray

// -------------------------------------------------------------- sampler
// queries and out: 4 uint32 per query, device memory, 16-byte aligned
int launch_sampler_bits(hipStream_t s, const uint32_t* queries, uint32_t n, uint32_t samples, uint32_t seed,
                        uint32_t* out);

// ------------------------------------------------------------------ LUT
int bake_luts(vh_device* dev);

// --------------------------------------------------------------- render
int render_scene(vh_scene* scene, const vh_render_params* p, float* film, float* guides, hipStream_t s);
// Jobs checked by the caller: on one device, plain (not adaptive), device films.
int render_shared(const vh_render_job* jobs, uint32_t count, hipStream_t s);
// The scratch of stream s on this device, created on first use.
StreamState* stream_state(vh_device* dev, hipStream_t s);

}  // namespace vh
