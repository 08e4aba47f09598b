// Host side of the C ABI in include/vork_hip.h: object lifetimes, argument
// checks, host<->device copies and the launches of the build, trace and
// render kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "vh_internal.h"

namespace vh {

static thread_local int g_last_status = VH_OK;
static thread_local std::string g_last_message;

int fail(int code, const std::string& message) {
    g_last_status = code;
    g_last_message = message;
    return code;
}
int hip_fail(hipError_t e, const char* what) {
    int code = (e == hipErrorOutOfMemory) ? VH_ERR_DEVICE : VH_ERR_DEVICE;
    return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

void device_retain(vh_device* d) { d->refs.fetch_add(1, std::memory_order_relaxed); }

void device_drop(vh_device* d) {
    if (d->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
    (void)hipSetDevice(d->ordinal);
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    d->lut_moments.release();
    d->lut_dielectric.release();
    d->streams.clear();
    if (d->static_host) (void)hipHostFree(d->static_host);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    // every object of the device has released its blocks into the cache
    pool_trim(d->ordinal);
    delete d;
}

// ------------------------------------------------------------ block cache
namespace {
struct BlockPool {
    std::mutex lock;
    std::multimap<size_t, void*> blocks;  // size -> block
    size_t cached = 0;
};
constexpr int kPoolDevices = 64;
constexpr size_t kPoolCap = size_t(16) << 30;
// Never destroyed: blocks may be released during static destruction.
BlockPool* pools() {
    static BlockPool* p = new BlockPool[kPoolDevices];
    return p;
}
// Sizes in 256-byte steps, and in 2 MiB steps from 2 MiB on.
size_t pool_round(size_t n) {
    const size_t g = n >= (size_t(2) << 20) ? (size_t(2) << 20) : size_t(256);
    return (n + g - 1) / g * g;
}
}  // namespace

hipError_t pool_alloc(void** p, size_t* bytes, int* device) {
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return e;
    const size_t n = pool_round(*bytes);
    if (d >= 0 && d < kPoolDevices) {
        BlockPool& bp = pools()[d];
        std::lock_guard<std::mutex> hold(bp.lock);
        auto it = bp.blocks.lower_bound(n);
        if (it != bp.blocks.end() && it->first <= n + n / 4) {
            *p = it->second;
            *bytes = it->first;
            *device = d;
            bp.cached -= it->first;
            bp.blocks.erase(it);
            return hipSuccess;
        }
    }
    e = hipMalloc(p, n);
    if (e == hipErrorOutOfMemory) {
        // the cached blocks first, then once more
        (void)hipGetLastError();
        pool_trim(d);
        e = hipMalloc(p, n);
    }
    if (e == hipSuccess) {
        *bytes = n;
        *device = d;
    }
    return e;
}

void pool_free(void* p, size_t bytes, int device) {
    if (device >= 0 && device < kPoolDevices) {
        BlockPool& bp = pools()[device];
        std::lock_guard<std::mutex> hold(bp.lock);
        if (bp.cached + bytes <= kPoolCap) {
            bp.blocks.emplace(bytes, p);
            bp.cached += bytes;
            return;
        }
    }
    (void)hipFree(p);
}

void pool_trim(int device) {
    if (device < 0 || device >= kPoolDevices) return;
    std::multimap<size_t, void*> blocks;
    {
        BlockPool& bp = pools()[device];
        std::lock_guard<std::mutex> hold(bp.lock);
        blocks.swap(bp.blocks);
        bp.cached = 0;
    }
    if (blocks.empty()) return;
    int cur = 0;
    const bool restore = hipGetDevice(&cur) == hipSuccess && cur != device;
    if (restore) (void)hipSetDevice(device);
    for (auto& b : blocks) (void)hipFree(b.second);
    if (restore) (void)hipSetDevice(cur);
}

StreamState* stream_state(vh_device* dev, hipStream_t s) {
    std::lock_guard<std::mutex> hold(dev->streams_lock);
    std::unique_ptr<StreamState>& st = dev->streams[s];
    if (!st) st.reset(new StreamState());
    return st.get();
}

namespace {
int clear() {
    g_last_status = VH_OK;
    g_last_message.clear();
    return VH_OK;
}
hipStream_t pick(vh_device* dev, void* stream) { return stream ? static_cast<hipStream_t>(stream) : dev->stream; }

// Copies `bytes` from a host or device source to a fresh device buffer.
int upload(DeviceBuffer& dst, const void* src, size_t bytes, bool device_src, hipStream_t s) {
    VH_HIP(dst.allocate(bytes));
    if (bytes)
        VH_HIP(hipMemcpyAsync(dst.ptr, src, bytes, device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    return VH_OK;
}
}  // namespace
}  // namespace vh

using namespace vh;

extern "C" {

uint32_t vh_abi_version(void) { return VH_ABI_VERSION; }

int vh_last_error(char* message, size_t length) {
    if (message && length) {
        size_t n = std::min(length - 1, g_last_message.size());
        std::memcpy(message, g_last_message.data(), n);
        message[n] = 0;
    }
    return g_last_status;
}

int vh_device_create(int ordinal, vh_device** out) {
    clear();
    if (!out) return fail(VH_ERR_USAGE, "vh_device_create: null output");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(VH_ERR_DEVICE, "no HIP device available");
    if (ordinal < 0 || ordinal >= count) return fail(VH_ERR_USAGE, "device ordinal out of range");
    VH_HIP(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    VH_HIP(hipGetDeviceProperties(&prop, ordinal));
    vh_device* d = new (std::nothrow) vh_device();
    if (!d) return fail(VH_ERR_DEVICE, "out of host memory");
    d->ordinal = ordinal;
    d->cu_count = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete d;
        return hip_fail(e, "hipStreamCreate");
    }
    int rc = bake_luts(d);
    // set-up of vh_trace_closest_static, with this device current: its
    // static scratch's address here and a pinned word for its error flags
    if (!rc) rc = static_scratch_address(&d->static_scratch);
    if (!rc) {
        e = hipHostMalloc(reinterpret_cast<void**>(&d->static_host), sizeof(uint32_t), hipHostMallocDefault);
        if (e != hipSuccess) {
            d->static_host = nullptr;
            rc = hip_fail(e, "hipHostMalloc");
        }
    }
    if (rc) {
        vh_device_release(d);
        return rc;
    }
    *out = d;
    return VH_OK;
}

void vh_device_release(vh_device* d) {
    if (d) device_drop(d);
}

int vh_device_synchronize(vh_device* d) {
    clear();
    if (!d) return fail(VH_ERR_USAGE, "null device");
    VH_HIP(hipSetDevice(d->ordinal));
    VH_HIP(hipStreamSynchronize(d->stream));
    return VH_OK;
}

// ------------------------------------------------------------------ BLAS
int vh_blas_build(vh_device* dev, const vh_geometry* g, uint32_t flags, void* stream, vh_blas** out) {
    clear();
    if (!dev || !g || !out) return fail(VH_ERR_USAGE, "vh_blas_build: null argument");
    *out = nullptr;
    if (!g->positions || !g->indices || g->triangle_count == 0 || g->vertex_count == 0)
        return fail(VH_ERR_SCENE, "vh_blas_build: empty geometry");
    if (g->position_stride < 12 || g->position_stride % 4)
        return fail(VH_ERR_USAGE, "vh_blas_build: position stride must be a multiple of 4 and >= 12");
    if (g->normals && (g->normal_stride < 12 || g->normal_stride % 4))
        return fail(VH_ERR_USAGE, "vh_blas_build: normal stride must be a multiple of 4 and >= 12");
    VH_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = pick(dev, stream);
    const bool on_device = (flags & VH_DEVICE_POINTERS) != 0;
    const uint32_t n = g->triangle_count, nv = g->vertex_count;
    vh_blas* b = new (std::nothrow) vh_blas();
    if (!b) return fail(VH_ERR_DEVICE, "out of host memory");
    b->device = dev;
    auto bail = [&](int rc) { delete b; return rc; };
    hipEvent_t e0, e1;
    VH_HIP(hipEventCreate(&e0));
    VH_HIP(hipEventCreate(&e1));
    VH_HIP(hipEventRecord(e0, s));
    int rc;
    // temporaries come from the stream's build scratch; device sources are read in place
    StreamState* st = stream_state(dev, s);
    std::lock_guard<std::mutex> hold(st->lock);
    BuildScratch& B = st->build;
    const size_t pos_bytes = (size_t)(nv - 1) * g->position_stride + 12;
    const float* pos_src = static_cast<const float*>(g->positions);
    if (!on_device) {
        VH_HIP(B.src.reserve(pos_bytes));
        VH_HIP(hipMemcpyAsync(B.src.ptr, g->positions, pos_bytes, hipMemcpyHostToDevice, s));
        pos_src = B.src.as<float>();
    }
    if ((rc = upload(b->indices, g->indices, sizeof(uint32_t) * 3 * (size_t)n, on_device, s))) return bail(rc);
    VH_HIP(b->positions.allocate(sizeof(float) * 3 * (size_t)nv));
    VH_HIP(B.lo.reserve(sizeof(float) * 3 * (size_t)n));
    VH_HIP(B.hi.reserve(sizeof(float) * 3 * (size_t)n));
    float* lo = B.lo.as<float>();
    float* hi = B.hi.as<float>();
    if ((rc = launch_triangle_bounds(s, B, pos_src, g->position_stride / 4, b->indices.as<uint32_t>(), n, lo, hi,
                                     b->positions.as<float>(), nv)))
        return bail(rc);
    if (g->normals) {
        const size_t nb = (size_t)(nv - 1) * g->normal_stride + 12;
        const float* nrm_src = static_cast<const float*>(g->normals);
        if (!on_device) {
            VH_HIP(B.src_nrm.reserve(nb));
            VH_HIP(hipMemcpyAsync(B.src_nrm.ptr, g->normals, nb, hipMemcpyHostToDevice, s));
            nrm_src = B.src_nrm.as<float>();
        }
        VH_HIP(b->normals.allocate(sizeof(float) * 3 * (size_t)nv));
        if ((rc = launch_compact3(s, nrm_src, g->normal_stride / 4, nv, b->normals.as<float>()))) return bail(rc);
    }
    if (g->triangle_materials)
        if ((rc = upload(b->materials, g->triangle_materials, sizeof(uint32_t) * (size_t)n, on_device, s)))
            return bail(rc);
    WideBuild wb;
    if ((rc = build_wide_bvh(dev, s, B, lo, hi, n, kLeafMax, &wb))) return bail(rc);
    VH_HIP(b->tris.allocate(sizeof(float) * 9 * (size_t)n));
    VH_HIP(b->tri_ids.allocate(sizeof(uint32_t) * (size_t)n));
    VH_HIP(b->surf.allocate(sizeof(float) * kSurfFloats * (size_t)n));
    if ((rc = launch_gather_triangles(s, b->positions.as<float>(), g->normals ? b->normals.as<float>() : nullptr,
                                      g->triangle_materials ? b->materials.as<uint32_t>() : nullptr,
                                      b->indices.as<uint32_t>(), wb.order, n, b->tris.as<float>(),
                                      b->surf.as<float>(), b->tri_ids.as<uint32_t>())))
        return bail(rc);
    // the node array leaves the scratch at its exact size
    VH_HIP(b->nodes.allocate(sizeof(WideNode) * (size_t)wb.node_count));
    VH_HIP(hipMemcpyAsync(b->nodes.ptr, wb.nodes, sizeof(WideNode) * (size_t)wb.node_count, hipMemcpyDeviceToDevice, s));
    float bounds[6];
    if ((rc = launch_bounds_reduce(s, B, lo, hi, n, bounds))) return bail(rc);
    VH_HIP(hipEventRecord(e1, s));
    VH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    BlasRecord& r = b->record;
    r.nodes = b->nodes.as<WideNode>();
    r.tris = b->tris.as<float>();
    r.tri_ids = b->tri_ids.as<uint32_t>();
    r.positions = b->positions.as<float>();
    r.normals = g->normals ? b->normals.as<float>() : nullptr;
    r.indices = b->indices.as<uint32_t>();
    r.materials = g->triangle_materials ? b->materials.as<uint32_t>() : nullptr;
    r.surf = b->surf.as<float>();
    r.surf_normals = g->normals ? 1u : 0u;
    for (int k = 0; k < 3; k++) { r.lo[k] = bounds[k]; r.hi[k] = bounds[3 + k]; }
    r.triangle_count = n;
    r.node_count = wb.node_count;
    b->info.triangle_count = n;
    b->info.node_count = wb.node_count;
    b->info.binary_node_count = wb.binary_nodes;
    b->info.max_depth = wb.max_depth;
    b->info.sah_cost = wb.sah_cost;
    b->info.build_ms = ms;
    b->info.device_bytes = b->nodes.bytes + b->tris.bytes + b->tri_ids.bytes + b->positions.bytes + b->normals.bytes +
                           b->indices.bytes + b->materials.bytes + b->surf.bytes;
    *out = b;
    return VH_OK;
}

int vh_blas_get_info(const vh_blas* b, vh_blas_info* info) {
    clear();
    if (!b || !info) return fail(VH_ERR_USAGE, "vh_blas_get_info: null argument");
    *info = b->info;
    return VH_OK;
}

void vh_blas_release(vh_blas* b) {
    if (!b) return;
    (void)hipSetDevice(b->device->ordinal);
    (void)hipStreamSynchronize(b->device->stream);
    delete b;
}

// ------------------------------------------------------------------ TLAS
int vh_tlas_build(vh_device* dev, vh_blas* const* blases, uint32_t blas_count, const vh_instance* instances,
                  uint32_t instance_count, uint32_t mode, uint32_t flags, void* stream, vh_tlas** inout) {
    clear();
    if (!dev || !inout || (blas_count && !blases) || (instance_count && !instances))
        return fail(VH_ERR_USAGE, "vh_tlas_build: null argument");
    if (mode != VH_TLAS_REBUILD && mode != VH_TLAS_REFIT) return fail(VH_ERR_USAGE, "vh_tlas_build: unknown mode");
    if (instance_count == 0) return fail(VH_ERR_SCENE, "vh_tlas_build: no instances");
    if (instance_count >= (1u << 30)) return fail(VH_ERR_USAGE, "vh_tlas_build: more than 2^30 instances");
    for (uint32_t i = 0; i < blas_count; i++)
        if (!blases[i] || blases[i]->device != dev) return fail(VH_ERR_USAGE, "vh_tlas_build: BLAS of another device");
    vh_tlas* t = *inout;
    if (mode == VH_TLAS_REFIT) {
        if (!t) return fail(VH_ERR_USAGE, "vh_tlas_build: refit needs an existing TLAS");
        if (t->instance_count != instance_count || t->blases.size() != blas_count)
            return fail(VH_ERR_USAGE, "vh_tlas_build: refit with a different instance or BLAS count");
    }
    VH_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = pick(dev, stream);
    bool fresh = (t == nullptr);
    if (fresh) {
        t = new (std::nothrow) vh_tlas();
        if (!t) return fail(VH_ERR_DEVICE, "out of host memory");
        t->device = dev;
    }
    auto bail = [&](int rc) { if (fresh) delete t; return rc; };
    int rc;
    // temporaries come from the stream's build scratch; a refit reuses the TLAS's buffers
    StreamState* st = stream_state(dev, s);
    std::lock_guard<std::mutex> hold(st->lock);
    BuildScratch& B = st->build;
    std::vector<BlasRecord> recs(blas_count);
    for (uint32_t i = 0; i < blas_count; i++) recs[i] = blases[i]->record;
    const size_t rec_bytes = sizeof(BlasRecord) * std::max<uint32_t>(blas_count, 1);
    VH_HIP(t->blas_records.reserve(rec_bytes));
    if (blas_count)
        VH_HIP(hipMemcpyAsync(t->blas_records.ptr, recs.data(), sizeof(BlasRecord) * blas_count,
                              hipMemcpyHostToDevice, s));
    const vh_instance* src = instances;
    if (!(flags & VH_DEVICE_POINTERS)) {
        VH_HIP(B.src.reserve(sizeof(vh_instance) * (size_t)instance_count));
        VH_HIP(hipMemcpyAsync(B.src.ptr, instances, sizeof(vh_instance) * (size_t)instance_count,
                              hipMemcpyHostToDevice, s));
        src = B.src.as<vh_instance>();
    }
    VH_HIP(t->instances.reserve(sizeof(InstanceRecord) * (size_t)instance_count));
    if ((rc = launch_instance_records(s, src, instance_count, blas_count, t->instances.as<InstanceRecord>())))
        return bail(rc);
    VH_HIP(B.lo.reserve(sizeof(float) * 3 * (size_t)instance_count));
    VH_HIP(B.hi.reserve(sizeof(float) * 3 * (size_t)instance_count));
    float* lo = B.lo.as<float>();
    float* hi = B.hi.as<float>();
    if ((rc = launch_instance_bounds(s, t->instances.as<InstanceRecord>(), t->blas_records.as<BlasRecord>(),
                                     instance_count, lo, hi)))
        return bail(rc);
    // A refit keeps the handle and instance order; the 8-wide hierarchy is
    // rebuilt from the new boxes (a clustering rebuild costs little more than a
    // quantised refit pass and keeps the tree quality of moving instances).
    WideBuild wb;
    if ((rc = build_wide_bvh(dev, s, B, lo, hi, instance_count, 1, &wb))) return bail(rc);
    VH_HIP(t->nodes.reserve(sizeof(WideNode) * (size_t)wb.node_count));
    VH_HIP(hipMemcpyAsync(t->nodes.ptr, wb.nodes, sizeof(WideNode) * (size_t)wb.node_count, hipMemcpyDeviceToDevice, s));
    VH_HIP(t->inst_ids.reserve(sizeof(uint32_t) * (size_t)instance_count));
    VH_HIP(hipMemcpyAsync(t->inst_ids.ptr, wb.order, sizeof(uint32_t) * (size_t)instance_count,
                          hipMemcpyDeviceToDevice, s));
    t->node_count = wb.node_count;
    t->instance_count = instance_count;
    t->blases.assign(blases, blases + blas_count);
    VH_HIP(hipStreamSynchronize(s));
    *inout = t;
    return VH_OK;
}

void vh_tlas_release(vh_tlas* t) {
    if (!t) return;
    (void)hipSetDevice(t->device->ordinal);
    (void)hipStreamSynchronize(t->device->stream);
    delete t;
}

// ------------------------------------------------------------------ trace
static int trace_common(const vh_tlas* t, const vh_ray* rays, uint32_t count, uint32_t mask, uint32_t flags,
                        void* stream, void* out, size_t out_elem, int mode) {
    clear();
    if (!t || (count && (!rays || !out))) return fail(VH_ERR_USAGE, "vh_trace: null argument");
    if (count == 0) return VH_OK;
    vh_device* dev = t->device;
    VH_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = pick(dev, stream);
    TlasView tv = t->view();
    if (flags & VH_DEVICE_POINTERS) {
        if (reinterpret_cast<uintptr_t>(rays) % 16u)
            return fail(VH_ERR_USAGE, "vh_trace: device ray array must be 16-byte aligned");
        return launch_trace(dev, s, tv, rays, count, mask, out, mode);
    }
    DeviceBuffer dr, dout;
    VH_HIP(dr.allocate(sizeof(vh_ray) * (size_t)count));
    VH_HIP(dout.allocate(out_elem * (size_t)count));
    VH_HIP(hipMemcpyAsync(dr.ptr, rays, sizeof(vh_ray) * (size_t)count, hipMemcpyHostToDevice, s));
    int rc = launch_trace(dev, s, tv, dr.ptr, count, mask, dout.ptr, mode);
    if (rc) return rc;
    VH_HIP(hipMemcpyAsync(out, dout.ptr, out_elem * (size_t)count, hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    return VH_OK;
}

int vh_trace_closest(const vh_tlas* t, const vh_ray* rays, uint32_t count, uint32_t mask, uint32_t flags, void* stream,
                     vh_hit* hits) {
    return trace_common(t, rays, count, mask, flags, stream, hits, sizeof(vh_hit), 0);
}
int vh_trace_any(const vh_tlas* t, const vh_ray* rays, uint32_t count, uint32_t mask, uint32_t flags, void* stream,
                 uint8_t* occluded) {
    return trace_common(t, rays, count, mask, flags, stream, occluded, 1, 1);
}
// Calls on one GPU share that GPU's static scratch; a lock per GPU ordinal
// keeps them from overlapping, whichever vh_device handle they come through.
static std::mutex g_static_trace_locks[64];

int vh_trace_closest_static(const vh_tlas* t, const vh_ray* rays, uint32_t count, uint32_t mask, void* stream,
                            vh_hit* hits) {
    clear();
    if (!t || (count && (!rays || !hits))) return fail(VH_ERR_USAGE, "vh_trace_closest_static: null argument");
    if (count > kStaticTraceMaxRays)
        return fail(VH_ERR_USAGE, "vh_trace_closest_static: more than 2^31 rays in one call");
    if (reinterpret_cast<uintptr_t>(rays) % 16u)
        return fail(VH_ERR_USAGE, "vh_trace_closest_static: the ray array must be 16-byte aligned device memory");
    if (reinterpret_cast<uintptr_t>(hits) % 4u)
        return fail(VH_ERR_USAGE, "vh_trace_closest_static: the hit array must be 4-byte aligned device memory");
    if (count == 0) return VH_OK;
    vh_device* dev = t->device;
    if (dev->ordinal < 0 || dev->ordinal >= 64 || !dev->static_scratch || !dev->static_host)
        return fail(VH_ERR_DEVICE, "vh_trace_closest_static: no static scratch on this device");
    std::lock_guard<std::mutex> hold(g_static_trace_locks[dev->ordinal]);
    VH_HIP(hipSetDevice(dev->ordinal));
    return launch_trace_static(dev, pick(dev, stream), t->view(), rays, count, mask, hits);
}

int vh_trace_closest_counts(const vh_tlas* t, const vh_ray* rays, uint32_t count, uint32_t mask, uint32_t flags,
                            void* stream, uint32_t* counts) {
    return trace_common(t, rays, count, mask, flags, stream, counts, 4 * sizeof(uint32_t), 2);
}

// ---------------------------------------------------------------- sampler
int vh_sampler_bits(vh_device* dev, const uint32_t* queries, uint32_t count, uint32_t samples, uint32_t seed,
                    uint32_t flags, void* stream, uint32_t* bits) {
    clear();
    if (!dev || (count && (!queries || !bits))) return fail(VH_ERR_USAGE, "vh_sampler_bits: null argument");
    if (count == 0) return VH_OK;
    VH_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = pick(dev, stream);
    const size_t bytes = 4 * sizeof(uint32_t) * (size_t)count;
    if (flags & VH_DEVICE_POINTERS) {
        if (reinterpret_cast<uintptr_t>(queries) % 16u || reinterpret_cast<uintptr_t>(bits) % 16u)
            return fail(VH_ERR_USAGE, "vh_sampler_bits: device arrays must be 16-byte aligned");
        int rc = launch_sampler_bits(s, queries, count, samples, seed, bits);
        if (rc) return rc;
        VH_HIP(hipStreamSynchronize(s));
        return VH_OK;
    }
    DeviceBuffer dq, dout;
    VH_HIP(dq.allocate(bytes));
    VH_HIP(dout.allocate(bytes));
    VH_HIP(hipMemcpyAsync(dq.ptr, queries, bytes, hipMemcpyHostToDevice, s));
    int rc = launch_sampler_bits(s, dq.as<uint32_t>(), count, samples, seed, dout.as<uint32_t>());
    if (rc) return rc;
    VH_HIP(hipMemcpyAsync(bits, dout.ptr, bytes, hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    return VH_OK;
}

// ------------------------------------------------------------------ scene
// Light tree over rectangle lights: crates/vork-render/src/lights/tree.rs:
// 43-69 (guides) and 106-170 (median spatial split).
static uint32_t tree_build(std::vector<vh_light_node>& nodes, const std::vector<double>& power,
                           const std::vector<float>& blo, const std::vector<float>& bhi, uint32_t* ids, uint32_t n,
                           uint32_t parent, double scale) {
    uint32_t index = (uint32_t)nodes.size();
    nodes.push_back(vh_light_node{});
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double p = 0;
    for (uint32_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::fmin(lo[k], blo[3 * ids[i] + k]);
            hi[k] = std::fmax(hi[k], bhi[3 * ids[i] + k]);
        }
        p += power[ids[i]] / scale;
    }
    float pf = (float)p;
    if (pf < FLT_MIN) pf = 0.f;
    vh_light_node node{};
    for (int k = 0; k < 3; k++) { node.lower_power[k] = lo[k]; node.upper_infinite[k] = hi[k]; }
    node.lower_power[3] = pf;
    node.upper_infinite[3] = 0.f;
    node.links[0] = node.links[1] = 0xffffffffu;
    node.links[2] = parent;
    node.links[3] = n;
    if (n == 1) {
        node.source[1] = ids[0];
        node.source[2] = ids[0];
    } else {
        double ex = (double)hi[0] - lo[0], ey = (double)hi[1] - lo[1], ez = (double)hi[2] - lo[2];
        int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
        std::stable_sort(ids, ids + n, [&](uint32_t a, uint32_t b) {
            double ca = (double)blo[3 * a + axis] + bhi[3 * a + axis], cb = (double)blo[3 * b + axis] + bhi[3 * b + axis];
            return ca < cb || (ca == cb && a < b);
        });
        uint32_t mid = n / 2;
        uint32_t l = tree_build(nodes, power, blo, bhi, ids, mid, index, scale);
        uint32_t r = tree_build(nodes, power, blo, bhi, ids + mid, n - mid, index, scale);
        node.links[0] = l;
        node.links[1] = r;
    }
    nodes[index] = node;
    return index;
}

// crates/vork-render/src/sampling.rs:149-175: inverse CDF of the 1.5 px
// Blackman-Harris film kernel, 1024 intervals.
static double filter_cdf(double x) {
    double t = M_PI * x / 1.5;
    return 0.5 + (0.35875 * x + 1.5 / M_PI * (0.48829 * std::sin(t) + 0.14128 * std::sin(2. * t) / 2. +
                                               0.01168 * std::sin(3. * t) / 3.)) /
                     (3. * 0.35875);
}
static void filter_table(float* table) {
    const int n = 1024;
    for (int i = 0; i <= n; i++) {
        if (i == 0) { table[i] = -1.5f; continue; }
        if (i == n) { table[i] = 1.5f; continue; }
        double u = (double)i / n, lo = -1.5, hi = 1.5;
        for (int k = 0; k < 48; k++) {
            double mid = (lo + hi) * 0.5;
            if (filter_cdf(mid) < u) lo = mid; else hi = mid;
        }
        table[i] = (float)((lo + hi) * 0.5);
    }
}

int vh_scene_create(vh_device* dev, const vh_tlas* tlas, const vh_scene_desc* desc, vh_scene** out) {
    clear();
    if (!dev || !tlas || !desc || !out) return fail(VH_ERR_USAGE, "vh_scene_create: null argument");
    *out = nullptr;
    if (desc->material_count == 0 || !desc->materials) return fail(VH_ERR_SCENE, "vh_scene_create: no materials");
    if (desc->light_count && !desc->lights) return fail(VH_ERR_USAGE, "vh_scene_create: null lights");
    if (tlas->device != dev) return fail(VH_ERR_USAGE, "vh_scene_create: TLAS of another device");
    for (uint32_t i = 0; i < desc->light_count; i++) {
        // the reference's light record: sides along the node's X / Y, normal -Z
        // (crates/vork-render/src/lights.rs:101-115), so the normal is
        // perpendicular to both sides
        const vh_rect_light& l = desc->lights[i];
        double uu = 0, vv = 0, nn = 0, nu = 0, nv = 0;
        bool finite = true;
        for (int k = 0; k < 3; k++) {
            uu += (double)l.u[k] * l.u[k]; vv += (double)l.v[k] * l.v[k]; nn += (double)l.normal[k] * l.normal[k];
            nu += (double)l.normal[k] * l.u[k]; nv += (double)l.normal[k] * l.v[k];
            finite = finite && std::isfinite(l.position[k]) && std::isfinite(l.u[k]) && std::isfinite(l.v[k]) &&
                     std::isfinite(l.normal[k]) && std::isfinite(l.radiance[k]);
        }
        if (!finite || !(nn > 0) || std::fabs(nu) > 1e-4 * std::sqrt(nn * uu) || std::fabs(nv) > 1e-4 * std::sqrt(nn * vv))
            return fail(VH_ERR_SCENE, "vh_scene_create: light " + std::to_string(i) +
                                          " is not finite or its normal is not perpendicular to its sides");
    }
    for (uint32_t i = 0; i < desc->material_count; i++) {
        const vh_material& m = desc->materials[i];
        bool ok = std::isfinite(m.base_metalness) && std::isfinite(m.specular_roughness) &&
                  std::isfinite(m.specular_ior) && std::isfinite(m.specular_weight) && m.specular_ior > 0.f;
        for (int k = 0; k < 3; k++) ok = ok && std::isfinite(m.base_color[k]);
        if (!ok) return fail(VH_ERR_SCENE, "vh_scene_create: non-finite material " + std::to_string(i));
    }
    VH_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = dev->stream;
    vh_scene* sc = new (std::nothrow) vh_scene();
    if (!sc) return fail(VH_ERR_DEVICE, "out of host memory");
    sc->device = dev;
    sc->tlas = tlas;
    auto bail = [&](int rc) { delete sc; return rc; };
    int rc;
    if ((rc = upload(sc->materials, desc->materials, sizeof(vh_material) * desc->material_count, false, s)))
        return bail(rc);
    sc->material_count = desc->material_count;
    sc->light_count = desc->light_count;
    if (desc->light_count) {
        if ((rc = upload(sc->lights, desc->lights, sizeof(vh_rect_light) * desc->light_count, false, s)))
            return bail(rc);
        const uint32_t n = desc->light_count;
        std::vector<double> power(n);
        std::vector<float> blo(3 * n), bhi(3 * n);
        std::vector<uint32_t> ids(n);
        const float lum[3] = {0.2126f, 0.7152f, 0.0722f};
        double scale = 0;
        for (uint32_t i = 0; i < n; i++) {
            const vh_rect_light& l = desc->lights[i];
            double pw = 0;
            for (int k = 0; k < 3; k++) pw += (double)l.radiance[k] * (double)lum[k];
            power[i] = pw > 0 ? pw : 0;
            scale = std::max(scale, power[i]);
            for (int k = 0; k < 3; k++) {
                float r = (std::fabs(l.u[k]) + std::fabs(l.v[k])) * 0.5f;
                blo[3 * i + k] = l.position[k] - r;
                bhi[3 * i + k] = l.position[k] + r;
            }
            ids[i] = i;
        }
        if (scale < DBL_MIN) scale = DBL_MIN;
        std::vector<vh_light_node> nodes;
        tree_build(nodes, power, blo, bhi, ids.data(), n, 0xffffffffu, scale);
        sc->light_nodes = (uint32_t)nodes.size();
        if ((rc = upload(sc->light_tree, nodes.data(), sizeof(vh_light_node) * nodes.size(), false, s))) return bail(rc);
    } else {
        VH_HIP(sc->lights.allocate(sizeof(vh_rect_light)));
        VH_HIP(sc->light_tree.allocate(sizeof(vh_light_node)));
    }
    // the same table for every scene: computed once (about 49,000 evaluations
    // of the filter's CDF)
    static const std::array<float, 1025> table = [] {
        std::array<float, 1025> t;
        filter_table(t.data());
        return t;
    }();
    if ((rc = upload(sc->filter, table.data(), sizeof(float) * table.size(), false, s))) return bail(rc);
    sc->camera = desc->camera;
    sc->sky = desc->sky;
    VH_HIP(hipStreamSynchronize(s));
    *out = sc;
    return VH_OK;
}

int vh_scene_set_camera(vh_scene* sc, const vh_camera* camera) {
    clear();
    if (!sc || !camera) return fail(VH_ERR_USAGE, "vh_scene_set_camera: null argument");
    sc->camera = *camera;
    return VH_OK;
}

void vh_scene_release(vh_scene* sc) {
    if (!sc) return;
    (void)hipSetDevice(sc->device->ordinal);
    (void)hipStreamSynchronize(sc->device->stream);
    delete sc;
}

void vh_render_params_default(vh_render_params* p, uint32_t width, uint32_t height, uint32_t samples) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->width = width;
    p->height = height;
    p->samples = samples;
    p->sample_begin = 0;
    p->sample_end = samples;
    p->seed = 0;
    // crates/vork-render/src/constants.rs:256-267
    p->max_bounces = 12;
    p->diffuse_bounces = 4;
    p->glossy_bounces = 4;
    p->caustics = 3;
    p->clamp_direct = 0.f;
    p->clamp_indirect = 10.f;
    p->light_threshold = 0.01f;
    p->batch_samples = 0;
    // constants.rs:264-266: adaptive sampling on, threshold 0.01, minimum from the sample count
    p->flags = VH_RENDER_ADAPTIVE;
    p->adaptive_min_samples = 0;
    p->adaptive_threshold = 0.f;
}

int vh_render(vh_scene* sc, const vh_render_params* p, uint32_t flags, void* stream, float* film) {
    return vh_render_guides(sc, p, flags, stream, film, nullptr);
}

int vh_render_shared(const vh_render_job* jobs, uint32_t count, void* stream) {
    clear();
    if (!jobs || count == 0 || count > 16) return fail(VH_ERR_USAGE, "vh_render_shared: 1 to 16 jobs");
    for (uint32_t j = 0; j < count; j++) {
        const vh_render_job& jb = jobs[j];
        const vh_render_params* p = &jb.params;
        if (!jb.scene || !jb.film) return fail(VH_ERR_USAGE, "vh_render_shared: null scene or film");
        if (jb.scene->device != jobs[0].scene->device)
            return fail(VH_ERR_USAGE, "vh_render_shared: the scenes must be on one device");
        if (p->width == 0 || p->height == 0 || p->samples == 0 || p->sample_end > p->samples ||
            p->sample_begin > p->sample_end)
            return fail(VH_ERR_USAGE, "vh_render_shared: job " + std::to_string(j) + " has an empty image or a "
                                      "sample range outside its exposure");
        if (p->flags & VH_RENDER_ADAPTIVE)
            return fail(VH_ERR_USAGE, "vh_render_shared: adaptive sampling is rendered by vh_render");
    }
    vh_device* dev = jobs[0].scene->device;
    VH_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = pick(dev, stream);
    return render_shared(jobs, count, s);
}

int vh_render_guides(vh_scene* sc, const vh_render_params* p, uint32_t flags, void* stream, float* film,
                     float* guides) {
    clear();
    if (!sc || !p || !film) return fail(VH_ERR_USAGE, "vh_render: null argument");
    if (p->width == 0 || p->height == 0) return fail(VH_ERR_USAGE, "vh_render: empty image");
    if (p->samples == 0 || p->sample_end > p->samples || p->sample_begin > p->sample_end)
        return fail(VH_ERR_USAGE, "vh_render: sample range outside the exposure");
    if ((uint64_t)p->width * p->height > (1ull << 31)) return fail(VH_ERR_USAGE, "vh_render: image too large");
    if (!(p->adaptive_threshold >= 0.f) || !std::isfinite(p->adaptive_threshold))
        return fail(VH_ERR_USAGE, "vh_render: adaptive threshold must be finite and >= 0");
    vh_device* dev = sc->device;
    VH_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = pick(dev, stream);
    if (flags & VH_DEVICE_POINTERS) return render_scene(sc, p, film, guides, s);
    size_t bytes = sizeof(float) * 4 * (size_t)p->width * p->height;
    DeviceBuffer f, g;
    VH_HIP(f.allocate(bytes));
    VH_HIP(hipMemcpyAsync(f.ptr, film, bytes, hipMemcpyHostToDevice, s));
    if (guides) {
        VH_HIP(g.allocate(2 * bytes));
        VH_HIP(hipMemcpyAsync(g.ptr, guides, 2 * bytes, hipMemcpyHostToDevice, s));
    }
    int rc = render_scene(sc, p, f.as<float>(), guides ? g.as<float>() : nullptr, s);
    if (rc) return rc;
    VH_HIP(hipMemcpyAsync(film, f.ptr, bytes, hipMemcpyDeviceToHost, s));
    if (guides) VH_HIP(hipMemcpyAsync(guides, g.ptr, 2 * bytes, hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    return VH_OK;
}

int vh_render_get_stats(const vh_scene* sc, vh_render_stats* stats) {
    clear();
    if (!sc || !stats) return fail(VH_ERR_USAGE, "vh_render_get_stats: null argument");
    std::lock_guard<std::mutex> hold(const_cast<vh_scene*>(sc)->state_lock);
    *stats = sc->stats;
    return VH_OK;
}

}  // extern "C"
