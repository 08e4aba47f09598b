// Ray-query kernels behind vh_trace_closest / vh_trace_any: persistent
// workgroups stride over the ray array, one ray per lane.
#include "vh_internal.h"
#include "vh_traverse.h"

namespace vh {
namespace {

constexpr int kTraceBlock = VH_BLOCK;

// Rays from a vh_ray array; results to vh_hit / occlusion bytes / visit counts
// (nodes, triangles, deepest stack entry, 0).
template <int kMode>
struct ArraySource {
    const vh_ray* rays;
    uint32_t n, mask;
    void* out;
    const InstanceRecord* instances;
    const BlasRecord* blases;
    __device__ uint32_t count() const { return n; }
    __device__ void fetch(uint32_t i, RayData& r) const {
        const float4* p = reinterpret_cast<const float4*>(rays + i);
        r.a = p[0];
        r.b = p[1];
    }
    // vh_ray: origin, tmin, direction, tmax
    __device__ void unpack(uint32_t i, const RayData& r, f3& o, f3& d, float& t0, float& t1, uint32_t& m,
                           uint32_t& key) const {
        o = mk(r.a.x, r.a.y, r.a.z);
        t0 = r.a.w;
        d = mk(r.b.x, r.b.y, r.b.z);
        t1 = r.b.w;
        m = mask;
        key = i;
    }
    __device__ void store(uint32_t i, const TraceHit& h, const TraceCounts& c) const {
        if (kMode == 0) {
            vh_hit hit;
            if (h.found) {
                // the hit's leaf-order triangle back to the mesh's primitive index
                hit.t = h.t; hit.instance = h.inst;
                hit.primitive = blases[instances[h.inst].blas].tri_ids[h.item];
                hit.custom_index = instances[h.inst].custom_index & 0xffffffu;
                hit.u = h.u; hit.v = h.v;
            } else {
                hit.t = -1.f; hit.instance = hit.primitive = hit.custom_index = 0xffffffffu; hit.u = hit.v = 0.f;
            }
            static_cast<vh_hit*>(out)[i] = hit;
        } else if (kMode == 1) {
            static_cast<uint8_t*>(out)[i] = h.found ? 1 : 0;
        } else {
            uint32_t* q = static_cast<uint32_t*>(out) + 4 * (size_t)i;  // 4-byte aligned caller arrays
            q[0] = c.nodes;
            q[1] = c.tris;
            q[2] = c.depth;
            q[3] = 0u;
        }
    }
};

template <int kMode>
VH_TRACE_KERNEL void k_trace(TlasView tv, ArraySource<kMode> src, uint32_t* overflow,
                                                        uint32_t* counters) {
    __shared__ uint64_t lds_stack[kLdsStack * kTraceBlock];
    Stack st = make_stack(lds_stack, reinterpret_cast<uint64_t*>(overflow));
    bool ok = trace_persistent<kMode == 1, kMode == 2>(tv, src, &counters[1], st);
    if (!ok) atomicOr(&counters[0], kStackOverflow);
}

// Zeroes `n` words (n <= blockDim.x): a fill kernel, so that the clear is an
// ordinary kernel node wherever the launch sequence is recorded.
__global__ void k_clear_words(uint32_t* w, uint32_t n) {
    if (threadIdx.x < n) w[threadIdx.x] = 0u;
}

// Scratch of vh_trace_closest_static, the closest-hit entry point that
// allocates nothing: 256 bytes of counters ([0] error flags, [1] ray claims),
// then the traversal overflow stacks of a grid of at most kStaticBlocks
// workgroups (kGlobalStack 8-byte entries per thread, the same depth as the
// allocating entry point, so both overflow on the same rays).  32768
// threads x 244 entries x 8 B = 64 MB per GPU, held by every
// process that loads the library until it exits.  Host code never uses this
// name as an address: each device's copy is looked up once with
// hipGetSymbolAddress and the kernels receive that pointer.
constexpr uint32_t kStaticBlocks = 128 * 256 / kTraceBlock;  // 32768 threads whatever the workgroup size
constexpr size_t kStaticCounterBytes = 256;
constexpr size_t kStaticWords =
    kStaticCounterBytes / sizeof(uint64_t) + (size_t)kStaticBlocks * kTraceBlock * kGlobalStack;
static __device__ __align__(256) uint64_t scratch[kStaticWords];

}  // namespace

int static_scratch_address(void** out) {
    void* p = nullptr;
    VH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(VH_ERR_DEVICE, "hipGetSymbolAddress(scratch) gave no address");
    *out = p;
    return VH_OK;
}

int launch_trace_static(vh_device* dev, hipStream_t s, const TlasView& tv, const vh_ray* rays, uint32_t count,
                        uint32_t mask, vh_hit* hits) {
    if (count == 0) return VH_OK;
    char* base = static_cast<char*>(dev->static_scratch);
    uint32_t* counters = reinterpret_cast<uint32_t*>(base);
    uint32_t* of = reinterpret_cast<uint32_t*>(base + kStaticCounterBytes);
    uint32_t grid = (count + kTraceBlock - 1) / kTraceBlock;
    if (grid > kStaticBlocks) grid = kStaticBlocks;
    // the counters keep the last call's values: clear them first
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, s, counters, 2u);
    hipLaunchKernelGGL(k_trace<0>, dim3(grid), dim3(kTraceBlock), 0, s, tv,
                       ArraySource<0>{rays, count, mask, hits, tv.instances, tv.blases}, of, counters);
    VH_HIP(hipGetLastError());
    VH_HIP(hipMemcpyAsync(dev->static_host, counters, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    if (*dev->static_host & kStackOverflow) return fail(VH_ERR_DEVICE, "traversal stack overflow");
    return VH_OK;
}

// Workgroups resident at once for the persistent trace kernels.
uint32_t trace_grid(vh_device* dev, uint32_t n) {
    uint32_t want = (n + kTraceBlock - 1) / kTraceBlock;
    uint32_t cap = (uint32_t)dev->cu_count * 8u;
    return want < cap ? (want ? want : 1u) : cap;
}

// Overflow stacks: kGlobalStack 8-byte entries per thread.
int ensure_overflow(DeviceBuffer& buf, uint32_t threads) {
    size_t need = (size_t)threads * kGlobalStack * 2 * sizeof(uint32_t);
    VH_HIP(buf.reserve(need));
    return VH_OK;
}

int launch_trace(vh_device* dev, hipStream_t s, const TlasView& tv, const void* rays, uint32_t count, uint32_t mask,
                 void* out, int mode) {
    if (count == 0) return VH_OK;
    uint32_t grid = trace_grid(dev, count);
    StreamState* st = stream_state(dev, s);
    std::lock_guard<std::mutex> hold(st->lock);
    int rc = ensure_overflow(st->overflow, grid * kTraceBlock);
    if (rc) return rc;
    VH_HIP(st->scratch_u32.reserve(2 * sizeof(uint32_t)));
    uint32_t* errors = st->scratch_u32.as<uint32_t>();  // [0] errors, [1] ray claims
    VH_HIP(hipMemsetAsync(errors, 0, 2 * sizeof(uint32_t), s));
    const vh_ray* r = static_cast<const vh_ray*>(rays);
    uint32_t* of = st->overflow.as<uint32_t>();
    if (mode == 0)
        hipLaunchKernelGGL(k_trace<0>, dim3(grid), dim3(kTraceBlock), 0, s, tv,
                           ArraySource<0>{r, count, mask, out, tv.instances, tv.blases}, of, errors);
    else if (mode == 1)
        hipLaunchKernelGGL(k_trace<1>, dim3(grid), dim3(kTraceBlock), 0, s, tv,
                           ArraySource<1>{r, count, mask, out, tv.instances, tv.blases}, of, errors);
    else
        hipLaunchKernelGGL(k_trace<2>, dim3(grid), dim3(kTraceBlock), 0, s, tv,
                           ArraySource<2>{r, count, mask, out, tv.instances, tv.blases}, of, errors);
    VH_HIP(hipGetLastError());
    uint32_t h = 0;
    VH_HIP(hipMemcpyAsync(&h, errors, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    if (h & kStackOverflow) return fail(VH_ERR_DEVICE, "traversal stack overflow");
    return VH_OK;
}

}  // namespace vh
