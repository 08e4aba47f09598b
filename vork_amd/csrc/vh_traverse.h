// Two-level traversal of the 8-wide BVH, one ray per lane, in persistent
// workgroups that keep every lane busy.
//
// Child order without sorting.  The builder places the children of a node in
// slots so that, for a ray whose direction has sign octant `oct` (bit a set
// when the direction is negative on axis a), visiting slot s in increasing
// order of s ^ oct is near-to-far in the common case (the ordering of
// compressed wide BVHs, Ylitie, Karras and Laine 2017).  A node visit then
// reduces its hit children to two masks:
//
//  * the node group: hit internal children, one bit per child at position
//    slot ^ oct, plus the node's internal-slot mask to turn a slot into the
//    child's index (child_base + number of internal slots below it);
//  * the item group: hit leaf children as a mask over the node's contiguous
//    items (triangles in a BLAS, instances in the TLAS).
//
// The lane keeps one group of each kind in registers; the stack holds the
// groups still to visit, one 8-byte entry per pushed group (word 0: base
// index, bit 31 set for an item group; word 1: mask).  Items are tested
// before the next node is visited, so a near hit shrinks the box tests of
// the rest of the traversal.
//
// Entering an instance pushes the lane's TLAS groups, moves the ray into
// object space and records the stack depth; once the BLAS groups are used up
// and the stack is back at that depth, the world ray is restored.
//
// The top kLdsStack entries of a lane live in LDS (interleaved across the
// workgroup), deeper ones in a global overflow array.
#pragma once

#include "vh_bvh.h"

namespace vh {

constexpr int kLdsStack = 12;         // entries per lane in LDS
constexpr int kGlobalStack = 244;     // overflow entries per lane
constexpr uint32_t kStackOverflow = 1u;
constexpr uint32_t kItemGroup = 0x80000000u;

// Idle lanes of a wave before it leaves its stepping loop to start new rays
// (starting costs the ray loads; lanes that start together also stay in step
// longer), for closest-hit (TRACE) and any-hit (ANY) rays.
#ifndef VH_TRACE_REFILL_MIN
#define VH_TRACE_REFILL_MIN 48
#endif
#ifndef VH_ANY_REFILL_MIN
#define VH_ANY_REFILL_MIN 48
#endif
// How a wave claims new rays from the device-wide counter of consecutive
// rays: 0, one atomic per refill for as many rays as it starts; 1, chunks of
// 64 into a wave-local reserve, an atomic only when the reserve runs out
// (cheap refills, for a low refill threshold).
#ifndef VH_TRACE_CLAIM_CHUNK
#define VH_TRACE_CLAIM_CHUNK 0
#endif
#ifndef VH_ANY_CLAIM_CHUNK
#define VH_ANY_CLAIM_CHUNK 0
#endif
// Items (triangles or instances) a lane tests per traversal step.
#ifndef VH_TRACE_ITEMS_PER_STEP
#define VH_TRACE_ITEMS_PER_STEP 1
#endif
// Waves per SIMD the trace kernels are compiled for (bounds their VGPRs:
// 512 / waves).  Any-hit rays keep less state (no barycentrics, ids or tie
// rule), so the shadow kernel fits 6 waves with few spills; closest hit is
// fastest at 4 (its spills grow faster than the latency hidden by more waves).
#ifndef VH_TRACE_WAVES_PER_EU
#define VH_TRACE_WAVES_PER_EU 4
#endif
#ifndef VH_ANY_WAVES_PER_EU
#define VH_ANY_WAVES_PER_EU 6
#endif
// Threads per workgroup of the trace and wavefront kernels.
#ifndef VH_BLOCK
#define VH_BLOCK 256
#endif
#define VH_TRACE_KERNEL __global__ __launch_bounds__(VH_BLOCK) __attribute__((amdgpu_waves_per_eu(VH_TRACE_WAVES_PER_EU)))
#define VH_ANY_KERNEL __global__ __launch_bounds__(VH_BLOCK) __attribute__((amdgpu_waves_per_eu(VH_ANY_WAVES_PER_EU)))
constexpr int kItemsPerStep = VH_TRACE_ITEMS_PER_STEP;

// Raw 32-byte ray record as a ray source stores it; the source decodes it.
struct RayData {
    float4 a, b;
};

// Entries are (word 0) | (word 1) << 32.
struct Stack {
    VH_LDS uint64_t* lds;       // &lds_entries[threadIdx.x]
    uint32_t lds_stride;        // blockDim.x
    VH_GLOBAL uint64_t* glob;   // &overflow_entries[global thread id]
    uint32_t glob_stride;       // threads in the grid
    int sp;
    __device__ __forceinline__ bool push(uint32_t a, uint32_t b) {
        const uint64_t e = (uint64_t)a | ((uint64_t)b << 32);
        if (sp < kLdsStack) {
            lds[sp * lds_stride] = e;
        } else if (sp < kLdsStack + kGlobalStack) {
            glob[(sp - kLdsStack) * glob_stride] = e;
        } else {
            return false;
        }
        sp++;
        return true;
    }
    __device__ __forceinline__ uint64_t pop() {
        sp--;
        if (sp < kLdsStack) return lds[sp * lds_stride];
        return glob[(sp - kLdsStack) * glob_stride];
    }
};

// Stack of the calling lane: `lds` holds kLdsStack * blockDim.x entries,
// `overflow` kGlobalStack entries per thread of the grid.
__device__ __forceinline__ Stack make_stack(uint64_t* lds, uint64_t* overflow) {
    Stack st;
    st.lds = (VH_LDS uint64_t*)lds + threadIdx.x;
    st.lds_stride = blockDim.x;
    st.glob = (VH_GLOBAL uint64_t*)overflow + blockIdx.x * blockDim.x + threadIdx.x;
    st.glob_stride = gridDim.x * blockDim.x;
    st.sp = 0;
    return st;
}

struct TraceHit {
    float t, u, v;
    uint32_t inst, item;  // instance and leaf-order triangle (the BLAS's tri_ids map it to the primitive)
    bool found;
};

struct TraceCounts {
    uint32_t nodes, tris;
    uint32_t depth;  // deepest stack entry the ray used (entries beyond kLdsStack live in global memory)
};

union NodeBits {
    vu4 v[5];
    WideNode n;
};

__device__ __forceinline__ void load_node(const WideNode* base, uint32_t i, NodeBits& nb) {
    const VH_GLOBAL vu4* p = (const VH_GLOBAL vu4*)(base + i);
#pragma unroll
    for (int k = 0; k < 5; k++) nb.v[k] = p[k];
}

__device__ __forceinline__ uint32_t octant(f3 inv) {
    return (inv.x < 0.f ? 1u : 0u) | (inv.y < 0.f ? 2u : 0u) | (inv.z < 0.f ? 4u : 0u);
}

// Relative size of the slab-distance error margin; see child_groups.
constexpr float kBoxMargin = 1.0f / 524288.0f;  // 2^-19
// Scale on the closest hit before it bounds a box test, so a triangle whose
// rounded t does not exceed the current hit is never culled by its box.
constexpr float kCullScale = 1.0f + 1.0f / 262144.0f;  // 1 + 2^-18

// Box tests of the eight children of one node.  Returns the hit internal
// children as bits at slot ^ oct and the hit leaf children's items as a mask
// over the node's items.
//
// Plane t of child c on axis a is q * (scale_a * inv_a) + (origin_a - o_a) *
// inv_a, one fused multiply-add per plane.  That rounds differently from the
// exact slab distance (decoded plane - o) / d, by at most a few units of
// 2^-24 * M_a with M_a = (|origin_a| + 255 scale_a + |o_a|) |inv_a|, the
// largest magnitude any term on the axis takes.  Near planes are shifted
// down and far planes up by 2^-19 M_a, which is larger than that error, so
// a child whose exact box meets [tmin, tmax] is always reported.  A zero
// direction component never reaches here (box_inv keeps inv_a finite); if an
// overflow still makes a plane infinite or NaN, NaN drops out of fmin / fmax,
// so the axis does not constrain the test and it stays conservative.
__device__ __forceinline__ void child_groups(const NodeBits& nb, f3 o, f3 inv, uint32_t oct, float tmin, float tmax,
                                             uint32_t& hits, uint32_t& items) {
    const WideNode& n = nb.n;
    const float sx = exp_scale(n.ex), sy = exp_scale(n.ey), sz = exp_scale(n.ez);
    const float ax = sx * inv.x, ay = sy * inv.y, az = sz * inv.z;
    const float bx = (n.origin[0] - o.x) * inv.x, by = (n.origin[1] - o.y) * inv.y, bz = (n.origin[2] - o.z) * inv.z;
    const float ex = ((255.f * sx + fabsf(n.origin[0])) + fabsf(o.x)) * fabsf(inv.x) * kBoxMargin;
    const float ey = ((255.f * sy + fabsf(n.origin[1])) + fabsf(o.y)) * fabsf(inv.y) * kBoxMargin;
    const float ez = ((255.f * sz + fabsf(n.origin[2])) + fabsf(o.z)) * fabsf(inv.z) * kBoxMargin;
    const float bnx = bx - ex, bfx = bx + ex, bny = by - ey, bfy = by + ey, bnz = bz - ez, bfz = bz + ez;
    // q planes: qlo x = v[2].xy, y = v[2].zw, z = v[3].xy; qhi x = v[3].zw, y = v[4].xy, z = v[4].zw.
    // A negative direction enters through the high plane.
    const bool px = !(oct & 1u), py = !(oct & 2u), pz = !(oct & 4u);
    const uint32_t nx0 = px ? nb.v[2].x : nb.v[3].z, nx1 = px ? nb.v[2].y : nb.v[3].w;
    const uint32_t fx0 = px ? nb.v[3].z : nb.v[2].x, fx1 = px ? nb.v[3].w : nb.v[2].y;
    const uint32_t ny0 = py ? nb.v[2].z : nb.v[4].x, ny1 = py ? nb.v[2].w : nb.v[4].y;
    const uint32_t fy0 = py ? nb.v[4].x : nb.v[2].z, fy1 = py ? nb.v[4].y : nb.v[2].w;
    const uint32_t nz0 = pz ? nb.v[3].x : nb.v[4].z, nz1 = pz ? nb.v[3].y : nb.v[4].w;
    const uint32_t fz0 = pz ? nb.v[4].z : nb.v[3].x, fz1 = pz ? nb.v[4].w : nb.v[3].y;
    const uint32_t imask = n.imask;
    uint32_t hm = 0;  // bit c: child c's box meets the ray
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t sh = (c & 3) * 8;
        const uint32_t nxw = c < 4 ? nx0 : nx1, fxw = c < 4 ? fx0 : fx1;
        const uint32_t nyw = c < 4 ? ny0 : ny1, fyw = c < 4 ? fy0 : fy1;
        const uint32_t nzw = c < 4 ? nz0 : nz1, fzw = c < 4 ? fz0 : fz1;
        const float tnx = __builtin_fmaf((float)((nxw >> sh) & 0xffu), ax, bnx);
        const float tfx = __builtin_fmaf((float)((fxw >> sh) & 0xffu), ax, bfx);
        const float tny = __builtin_fmaf((float)((nyw >> sh) & 0xffu), ay, bny);
        const float tfy = __builtin_fmaf((float)((fyw >> sh) & 0xffu), ay, bfy);
        const float tnz = __builtin_fmaf((float)((nzw >> sh) & 0xffu), az, bnz);
        const float tfz = __builtin_fmaf((float)((fzw >> sh) & 0xffu), az, bfz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        hm |= tn <= tf ? 1u << c : 0u;
    }
    // the hit internal children, bit c moved to c ^ oct (a swap of
    // neighbouring bits, pairs and nibbles for the octant's three bits)
    uint32_t h = hm & imask;
    if (oct & 1u) h = ((h & 0x55u) << 1) | ((h >> 1) & 0x55u);
    if (oct & 2u) h = ((h & 0x33u) << 2) | ((h >> 2) & 0x33u);
    if (oct & 4u) h = ((h & 0x0fu) << 4) | (h >> 4);
    // the hit leaf children's items (internal and empty slots hold no items)
    uint32_t it = 0, leaves = hm & ~imask;
    while (leaves) {
        const uint32_t c = (uint32_t)__builtin_ctz(leaves);
        leaves &= leaves - 1u;
        const uint32_t meta = ((c < 4 ? nb.v[1].z : nb.v[1].w) >> ((c & 3u) * 8u)) & 0xffu;
        if (!(meta & kMetaNoItems)) it |= ((1u << ((meta >> 5) + 1u)) - 1u) << (meta & 31u);
    }
    hits = h;
    items = it;
}

// One lane's traversal state.
template <bool kAny, bool kCount>
struct Tracer {
    f3 wo, wd;              // world ray
    f3 o, inv;              // ray of the current space (world or instance object space)
    uint32_t oct;
    f3 winv;                // the world ray's inv and oct, restored on leaving an instance
    uint32_t woct;
    f3 ld;                  // object-space direction of the triangle tests
    Shear lsh;
    float tmin, tmax;
    uint32_t mask;
    uint32_t cur_inst;
    bool cull, in_blas;
    int blas_sp;            // stack depth at the current instance's entry
    const WideNode* bnodes; // current BLAS
    const float* btris;
    const uint32_t* bids;
    uint32_t nbase, ngroup; // node group: child base, hits (bits 0-7) | internal-slot mask << 8
    uint32_t ibase, items;  // item group
    bool done;
    TraceHit hit;
    TraceCounts cnt;

    __device__ __forceinline__ void set_root() {
        nbase = 0;
        ngroup = (1u << oct) | (1u << 8);  // the root as slot 0 of a one-child group
        ibase = 0;
        items = 0;
    }

    __device__ __forceinline__ void begin(const TlasView& tv, f3 o_, f3 d_, float tmin_, float tmax_, uint32_t mask_,
                                          Stack& st) {
        wo = o_; wd = d_;
        o = o_; inv = box_inv(d_); oct = octant(inv);
        winv = inv; woct = oct;
        ld = d_;
        lsh = make_shear(d_);
        tmin = tmin_; tmax = tmax_; mask = mask_;
        cur_inst = 0; cull = false; in_blas = false; blas_sp = 0;
        bnodes = nullptr; btris = nullptr; bids = nullptr;
        set_root();
        hit.found = false; hit.t = tmax_; hit.u = hit.v = 0.f; hit.inst = hit.item = 0xffffffffu;
        cnt.nodes = cnt.tris = cnt.depth = 0;
        st.sp = 0;
        done = tv.node_count == 0;
    }

    __device__ __forceinline__ void leave_instance() {
        o = wo;
        inv = winv;
        oct = woct;
        in_blas = false;
    }

    // Enters instance `inst` unless it is masked out or invalid.
    __device__ __forceinline__ bool enter_instance(const TlasView& tv, uint32_t inst, Stack& st) {
        const VH_GLOBAL InstanceRecord& r = as_global(tv.instances)[inst];
        const uint32_t mf = r.mask_flags;
        if (!(mf & (1u << 9)) || !(mf & mask & 0xffu)) return true;
        if ((ngroup & 0xffu) && !st.push(nbase, ngroup)) return false;
        if (items && !st.push(ibase | kItemGroup, items)) return false;
        blas_sp = st.sp;
        if ((mf & kInstIdentity) && identity_exact(wo, wd)) {
            // the transformed ray is the world ray: no transform, no reciprocals
            o = wo;
            ld = wd;
            inv = winv;
            oct = woct;
        } else {
            float w2o[12];
#pragma unroll
            for (int k = 0; k < 12; k++) w2o[k] = r.w2o[k];
            o = xform_point(w2o, wo);
            ld = xform_dir(w2o, wd);
            inv = box_inv(ld);
            oct = octant(inv);
        }
        lsh = make_shear(ld);
        cull = (mf >> 8) & 1u;
        const VH_GLOBAL BlasRecord& b = as_global(tv.blases)[r.blas];
        bnodes = b.nodes;
        btris = b.tris;
        bids = b.tri_ids;
        cur_inst = inst;
        in_blas = true;
        set_root();
        return true;
    }

    // Tests one triangle of the current BLAS.
    __device__ __forceinline__ void test_triangle(uint32_t item) {
        const VH_GLOBAL float* tp = as_global(btris) + 9 * (size_t)item;
        const f3 a = mk(tp[0], tp[1], tp[2]), b = mk(tp[3], tp[4], tp[5]), c = mk(tp[6], tp[7], tp[8]);
        if (kCount) cnt.tris++;
        // admit equal t so that the tie rule can pick the lower id
        const float lim = (kAny || !hit.found) ? tmax : next_up(hit.t);
        float t, u, v;
        if (intersect_triangle(a, b, c, o, ld, lsh, tmin, lim, cull, &t, &u, &v)) {
            if (!(t < tmax)) return;
            if (kAny) {
                hit.found = true;
                hit.t = t;
                done = true;
                return;
            }
            // (t, instance, primitive) order (better_hit); t never exceeds the held hit here
            // (lim), and an instance is entered once per ray, so an equal t in the same
            // instance compares two primitives of the current BLAS
            bool take = !hit.found || t < hit.t;
            if (!take && t == hit.t)
                take = cur_inst != hit.inst ? cur_inst < hit.inst : as_global(bids)[item] < as_global(bids)[hit.item];
            if (take) {
                hit.found = true; hit.t = t; hit.u = u; hit.v = v; hit.inst = cur_inst; hit.item = item;
            }
        }
    }

    // Takes a group off the stack when the lane holds none; the ray is done
    // when the stack is empty.
    __device__ __forceinline__ void pop_phase(Stack& st) {
        if ((ngroup & 0xffu) || items) return;
        if (in_blas && st.sp == blas_sp) leave_instance();
        if (st.sp == 0) {
            done = true;
            return;
        }
        const uint64_t e = st.pop();
        const uint32_t e0 = (uint32_t)e, e1 = (uint32_t)(e >> 32);
        if (e0 & kItemGroup) {
            ibase = e0 & ~kItemGroup;
            items = e1;
        } else {
            nbase = e0;
            ngroup = e1;
        }
    }

    __device__ __forceinline__ bool wants_node() const { return (ngroup & 0xffu) && !items; }

    // Visits the next node of the node group.  Returns false when the stack
    // overflowed.
    __device__ __forceinline__ bool node_phase(const TlasView& tv, Stack& st) {
        const uint32_t hits = ngroup & 0xffu, imask = ngroup >> 8;
        const uint32_t slot = (uint32_t)__builtin_ctz(hits) ^ oct;
        const uint32_t rest = hits & (hits - 1u);
        const uint32_t node = nbase + (uint32_t)__popc(imask & ((1u << slot) - 1u));
        if (rest && !st.push(nbase, rest | (imask << 8))) return false;
        NodeBits nb;
        load_node(in_blas ? bnodes : tv.nodes, node, nb);
        if (kCount) cnt.nodes++;
        uint32_t h, it;
        child_groups(nb, o, inv, oct, tmin, kAny ? tmax : hit.t * kCullScale, h, it);
        nbase = nb.n.child_base;
        ngroup = h | ((uint32_t)nb.n.imask << 8);
        ibase = nb.n.item_base;
        items = it;
        return true;
    }

    // Tests up to kItemsPerStep items of the item group (an instance entry
    // ends the phase).  Returns false when the stack overflowed.
    __device__ __forceinline__ bool item_phase(const TlasView& tv, Stack& st) {
        for (int q = 0; q < kItemsPerStep && items && !done; q++) {
            const uint32_t item = ibase + (uint32_t)__builtin_ctz(items);
            items &= items - 1u;
            if (in_blas) {
                test_triangle(item);
            } else {
                if (!enter_instance(tv, as_global(tv.inst_ids)[item], st)) return false;
                break;
            }
        }
        return true;
    }
};

// Persistent trace loop.  `src` supplies rays and consumes results:
//   uint32_t count() const;
//   void fetch(uint32_t i, RayData& r) const;      // loads ray i
//   void unpack(uint32_t i, const RayData& r, f3& o, f3& d, float& tmin, float& tmax, uint32_t& mask,
//               uint32_t& key) const;
//   void store(uint32_t key, const TraceHit& h, const TraceCounts& c) const;
//   TlasView view(uint32_t key) const;            // with kPerRay: the ray's own TLAS
// `next` is a global counter of claimed ray indices, zero at launch.  With
// kPerRay every ray traverses the TLAS the source names for it (the rays of
// several scenes in one queue); otherwise all traverse `tv`.
//
// Lanes that finish wait until kRefillMin lanes of the wave are idle, then
// start consecutive rays claimed with one atomic per wave (with kChunk, from
// a reserve of 64 the wave claims at a time).  All waves claim from the one
// counter, so the rays in flight on the whole device are a narrow window of
// the queue (neighbouring pixels at the first bounce) and share the BVH data
// they touch in the caches; waves that each walked chunks of their own lost
// that (closest hit 6.9 -> 7.2-7.7 ms per bounce at 1920x1080 x 64 spp).  Each iteration of the inner loop every busy
// lane takes one traversal step.  Every lane leaves the outer loop once the
// queue is drained and its ray is stored.  Returns false when a stack
// overflowed.
template <bool kAny, bool kCount, class Src, bool kPerRay = false>
__device__ bool trace_persistent(const TlasView& tv0, const Src& src, uint32_t* next, Stack& st) {
    const uint32_t n = src.count();
    Tracer<kAny, kCount> tr;
    TlasView lane_tv;
    if (kPerRay) lane_tv = tv0;
    const TlasView& tv = kPerRay ? lane_tv : tv0;
    bool active = false, exhausted = false, ok = true;
    uint32_t key = 0;
    const uint64_t lt_mask = (1ull << __lane_id()) - 1ull;
    // Every thread's first ray is its global index, without an atomic (all
    // waves claiming one counter at once costs about 100 us per launch, and a
    // launch with fewer rays than threads then claims nothing); refills claim
    // rays from `first` on.
    const uint32_t first = gridDim.x * blockDim.x;
    constexpr uint32_t kRefillMin = kAny ? VH_ANY_REFILL_MIN : VH_TRACE_REFILL_MIN;
    constexpr bool kChunk = kAny ? VH_ANY_CLAIM_CHUNK : VH_TRACE_CLAIM_CHUNK;
    uint32_t res_next = 0, res_end = 0;  // with kChunk: the wave's claimed, unstarted rays
    bool fresh = true;
    while (true) {
        const uint64_t want = __ballot(!active && !exhausted);
        if (want) {
            uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
            if (!fresh) {
                if (kChunk) {
                    // the reserve first, then a fresh chunk (need <= 64)
                    const uint32_t need = (uint32_t)__popcll(want), j = (uint32_t)__popcll(want & lt_mask);
                    const uint32_t left = res_end - res_next;
                    uint32_t fresh_base = 0;
                    if (need > left) {
                        const int leader = __ffsll((unsigned long long)want) - 1;
                        if ((int)__lane_id() == leader) fresh_base = atomicAdd(next, 64u);
                        fresh_base = first + (uint32_t)__shfl(fresh_base, leader);
                    }
                    idx = j < left ? res_next + j : fresh_base + (j - left);
                    if (need > left) {
                        res_next = fresh_base + (need - left);
                        res_end = fresh_base + 64u;
                    } else {
                        res_next += need;
                    }
                } else {
                    const int leader = __ffsll((unsigned long long)want) - 1;
                    uint32_t base = 0;
                    if ((int)__lane_id() == leader) base = atomicAdd(next, (uint32_t)__popcll(want));
                    base = __shfl(base, leader);
                    idx = first + base + (uint32_t)__popcll(want & lt_mask);
                }
            }
            fresh = false;
            if (!active && !exhausted) {
                if (idx < n) {
                    RayData rd;
                    src.fetch(idx, rd);
                    f3 o, d;
                    float t0, t1;
                    uint32_t m;
                    src.unpack(idx, rd, o, d, t0, t1, m, key);
                    if constexpr (kPerRay) lane_tv = src.view(key);
                    tr.begin(tv, o, d, t0, t1, m, st);
                    active = !tr.done;
                    if (!active) src.store(key, tr.hit, tr.cnt);
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__ballot(active)) {
            if (!__ballot(!exhausted)) break;
            continue;
        }
        while (true) {
            if (active) {
                tr.pop_phase(st);
                if (!tr.done && tr.wants_node() && !tr.node_phase(tv, st)) {
                    ok = false;
                    tr.done = true;
                }
                if (!tr.done && tr.items && !tr.item_phase(tv, st)) {
                    ok = false;
                    tr.done = true;
                }
                // the phases only push (pops happen in pop_phase), so the
                // stack is at its deepest for this step here
                if (kCount) tr.cnt.depth = max(tr.cnt.depth, (uint32_t)st.sp);
                if (tr.done) {
                    src.store(key, tr.hit, tr.cnt);
                    active = false;
                }
            }
            if (!__ballot(active)) break;
            if ((uint32_t)__popcll(__ballot(!active && !exhausted)) >= kRefillMin) break;
        }
    }
    return ok;
}


// ---------------------------------------------------------------- packets
// Packet traversal (closest hit) for waves whose rays are coherent, such as
// the jittered camera samples of one pixel: the wave walks one stack, visits a
// node once for all its lanes, and each lane tests its own ray against it.
// A child is taken when any lane's box test hits it (each lane culls with its
// own closest hit), and every lane tests every triangle of the union of hit
// leaves.  The box test is conservative, so a triangle outside a lane's boxes
// cannot hit that lane's ray, and the closest hit by (t, instance, primitive)
// does not depend on the visiting order: each lane ends with the hit of the
// per-lane traversal.  Nodes and triangles, read at wave-uniform addresses,
// come in through the scalar unit into SGPRs.  The wave's stack lives in LDS and holds as many
// entries as a lane's stack in trace_persistent (kLdsStack + kGlobalStack), so
// both overflow at the same tree depth.
constexpr int kPacketStack = kLdsStack + kGlobalStack;

// Wave-uniform reads through the scalar unit: the address is made uniform
// with readfirstlane and read through the constant address space.
#define VH_CONST __attribute__((address_space(4)))
__device__ __forceinline__ uint64_t uniform_address(const void* p, uint64_t offset) {
    const uint64_t a = (uint64_t)p + offset;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32));
    return ((uint64_t)hi << 32) | lo;
}
// Lanes whose `pred` holds, straight from the compare mask: HIP's
// __ballot(int) widens the bool to an int and compares it again, two more
// vector instructions per vote.
__device__ __forceinline__ uint64_t wave_vote(bool pred) { return __builtin_amdgcn_ballot_w64(pred); }

// A value every lane holds alike (read at a wave-uniform address from LDS or
// through the vector unit), moved into an SGPR so that the compiler treats
// what depends on it as uniform: branches on it stay scalar.
__device__ __forceinline__ uint32_t wave_uniform(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
template <class T>
__device__ __forceinline__ const T* wave_uniform(const T* p) {
    const uint64_t a = (uint64_t)p;
    return (const T*)(((uint64_t)wave_uniform((uint32_t)(a >> 32)) << 32) | wave_uniform((uint32_t)a));
}
__device__ __forceinline__ void load_node_scalar(const WideNode* base, uint32_t i, NodeBits& nb) {
    const VH_CONST vu4* p = (const VH_CONST vu4*)uniform_address(base, (uint64_t)i * sizeof(WideNode));
#pragma unroll
    for (int k = 0; k < 5; k++) nb.v[k] = p[k];
}

struct PacketRay {
    f3 wo, wd;       // world ray
    f3 winv;         // the world ray's inv, restored on leaving an instance
    f3 o, inv, ld;   // ray of the current space
    Shear lsh;
    uint32_t oct, mask;
    float tmin, tmax;
    bool live, in_inst;  // holds a ray; takes part in the current instance
    TraceHit hit;
};

// Box tests of a node's children for every lane; the union over the wave of
// the hit internal children (bits at slot ^ order_oct) and of the hit leaf
// items.  With kUniform every lane taking part has order_oct as its own
// octant (the usual case for a packet): the near and far planes are then
// picked once for the wave, by scalar selects of the node's dwords, instead
// of per lane.
template <bool kUniform>
__device__ __forceinline__ void packet_children(const NodeBits& nb, const PacketRay& r, bool take_part,
                                                uint32_t order_oct, uint32_t& hits, uint32_t& items) {
    const WideNode& n = nb.n;
    const f3 o = r.o, inv = r.inv;
    const float tmax = r.hit.found ? r.hit.t * kCullScale : r.tmax;
    const float sx = exp_scale(n.ex), sy = exp_scale(n.ey), sz = exp_scale(n.ez);
    const float ax = sx * inv.x, ay = sy * inv.y, az = sz * inv.z;
    const float bx = (n.origin[0] - o.x) * inv.x, by = (n.origin[1] - o.y) * inv.y, bz = (n.origin[2] - o.z) * inv.z;
    const float ex = ((255.f * sx + fabsf(n.origin[0])) + fabsf(o.x)) * fabsf(inv.x) * kBoxMargin;
    const float ey = ((255.f * sy + fabsf(n.origin[1])) + fabsf(o.y)) * fabsf(inv.y) * kBoxMargin;
    const float ez = ((255.f * sz + fabsf(n.origin[2])) + fabsf(o.z)) * fabsf(inv.z) * kBoxMargin;
    const float bnx = bx - ex, bfx = bx + ex, bny = by - ey, bfy = by + ey, bnz = bz - ez, bfz = bz + ez;
    const uint32_t oct = kUniform ? order_oct : r.oct;
    const bool px = !(oct & 1u), py = !(oct & 2u), pz = !(oct & 4u);
    const uint32_t nx0 = px ? nb.v[2].x : nb.v[3].z, nx1 = px ? nb.v[2].y : nb.v[3].w;
    const uint32_t fx0 = px ? nb.v[3].z : nb.v[2].x, fx1 = px ? nb.v[3].w : nb.v[2].y;
    const uint32_t ny0 = py ? nb.v[2].z : nb.v[4].x, ny1 = py ? nb.v[2].w : nb.v[4].y;
    const uint32_t fy0 = py ? nb.v[4].x : nb.v[2].z, fy1 = py ? nb.v[4].y : nb.v[2].w;
    const uint32_t nz0 = pz ? nb.v[3].x : nb.v[4].z, nz1 = pz ? nb.v[3].y : nb.v[4].w;
    const uint32_t fz0 = pz ? nb.v[4].z : nb.v[3].x, fz1 = pz ? nb.v[4].w : nb.v[3].y;
    const uint32_t imask = n.imask;
    // the vote of a single compare is its lane mask; the lanes taking part are
    // masked in with one scalar AND per child
    const uint64_t part = wave_vote(take_part);
    uint32_t h = 0, it = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t sh = (c & 3) * 8;
        const uint32_t nxw = c < 4 ? nx0 : nx1, fxw = c < 4 ? fx0 : fx1;
        const uint32_t nyw = c < 4 ? ny0 : ny1, fyw = c < 4 ? fy0 : fy1;
        const uint32_t nzw = c < 4 ? nz0 : nz1, fzw = c < 4 ? fz0 : fz1;
        const float tnx = __builtin_fmaf((float)((nxw >> sh) & 0xffu), ax, bnx);
        const float tfx = __builtin_fmaf((float)((fxw >> sh) & 0xffu), ax, bfx);
        const float tny = __builtin_fmaf((float)((nyw >> sh) & 0xffu), ay, bny);
        const float tfy = __builtin_fmaf((float)((fyw >> sh) & 0xffu), ay, bfy);
        const float tnz = __builtin_fmaf((float)((nzw >> sh) & 0xffu), az, bnz);
        const float tfz = __builtin_fmaf((float)((fzw >> sh) & 0xffu), az, bfz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        if (wave_vote(tn <= tf) & part) {
            const uint32_t meta = ((c < 4 ? nb.v[1].z : nb.v[1].w) >> sh) & 0xffu;
            if ((imask >> c) & 1u) {
                h |= 1u << ((uint32_t)c ^ order_oct);
            } else if (!(meta & kMetaNoItems)) {
                it |= ((1u << ((meta >> 5) + 1u)) - 1u) << (meta & 31u);
            }
        }
    }
    hits = h;
    items = it;
}

// Traces the wave's rays (one per lane, `live` marks the lanes holding one).
// Returns false when the wave's stack overflowed (the rays' hits are then
// incomplete).
__device__ inline bool trace_packet(const TlasView& tv, PacketRay& r, uint64_t* stack) {
    const uint32_t lane = __lane_id();
    bool in_blas = false, cull = false;
    uint32_t cur_inst = 0;
    const WideNode* bnodes = nullptr;
    const float* btris = nullptr;
    const uint32_t* bids = nullptr;
    int sp = 0, blas_sp = 0;
    uint32_t nbase = 0, ibase = 0, items = 0;
    uint32_t order = (uint32_t)__builtin_amdgcn_readfirstlane((int)r.oct);
    uint32_t ngroup = tv.node_count ? (1u << order) | (1u << 8) : 0u;  // the root as slot 0
    r.in_inst = r.live;
    while (true) {
        if (items) {
            const uint32_t item = ibase + (uint32_t)__builtin_ctz(items);
            items &= items - 1u;
            if (in_blas) {
                const VH_CONST float* tp = (const VH_CONST float*)uniform_address(btris, 36ull * item);
                const f3 a = mk(tp[0], tp[1], tp[2]), b = mk(tp[3], tp[4], tp[5]), c = mk(tp[6], tp[7], tp[8]);
                if (r.in_inst) {
                    const float lim = r.hit.found ? next_up(r.hit.t) : r.tmax;
                    float t, u, v;
                    if (intersect_triangle(a, b, c, r.o, r.ld, r.lsh, r.tmin, lim, cull, &t, &u, &v) && t < r.tmax) {
                        bool take = !r.hit.found || t < r.hit.t;
                        if (!take && t == r.hit.t)
                            take = cur_inst != r.hit.inst ? cur_inst < r.hit.inst
                                                          : as_global(bids)[item] < as_global(bids)[r.hit.item];
                        if (take) {
                            r.hit.found = true; r.hit.t = t; r.hit.u = u; r.hit.v = v;
                            r.hit.inst = cur_inst; r.hit.item = item;
                        }
                    }
                }
                continue;
            }
            // an instance of the TLAS: enter it if some lane's mask takes it
            const uint32_t inst = wave_uniform(as_global(tv.inst_ids)[item]);
            const VH_GLOBAL InstanceRecord& rec = as_global(tv.instances)[inst];
            const uint32_t mf = wave_uniform(rec.mask_flags);
            const bool enter = r.live && (mf & (1u << 9)) && (mf & r.mask & 0xffu);
            if (!wave_vote(enter)) continue;
            if (sp + 2 > kPacketStack) return false;
            if (ngroup & 0xffu) {
                if (lane == 0) stack[sp] = (uint64_t)nbase | ((uint64_t)ngroup << 32);
                sp++;
            }
            if (items) {
                if (lane == 0) stack[sp] = (uint64_t)(ibase | kItemGroup) | ((uint64_t)items << 32);
                sp++;
            }
            blas_sp = sp;
            if ((mf & kInstIdentity) && identity_exact(r.wo, r.wd)) {
                r.o = r.wo;
                r.ld = r.wd;
                r.inv = r.winv;
            } else {
                float w2o[12];
#pragma unroll
                for (int k = 0; k < 12; k++) w2o[k] = rec.w2o[k];
                r.o = xform_point(w2o, r.wo);
                r.ld = xform_dir(w2o, r.wd);
                r.inv = box_inv(r.ld);
            }
            r.oct = octant(r.inv);
            r.lsh = make_shear(r.ld);
            r.in_inst = enter;
            cull = (mf >> 8) & 1u;
            const VH_GLOBAL BlasRecord& br = as_global(tv.blases)[rec.blas];
            bnodes = wave_uniform(br.nodes);
            btris = wave_uniform(br.tris);
            bids = wave_uniform(br.tri_ids);
            cur_inst = inst;
            in_blas = true;
            order = (uint32_t)__builtin_amdgcn_readfirstlane((int)r.oct);
            nbase = 0;
            ngroup = (1u << order) | (1u << 8);
            items = 0;
            continue;
        }
        if (ngroup & 0xffu) {
            const uint32_t hits = ngroup & 0xffu, imask = ngroup >> 8;
            const uint32_t slot = (uint32_t)__builtin_ctz(hits) ^ order;
            const uint32_t rest = hits & (hits - 1u);
            const uint32_t node = nbase + (uint32_t)__popc(imask & ((1u << slot) - 1u));
            if (rest) {
                if (sp >= kPacketStack) return false;
                if (lane == 0) stack[sp] = (uint64_t)nbase | ((uint64_t)(rest | (imask << 8)) << 32);
                sp++;
            }
            NodeBits nb;
            load_node_scalar(in_blas ? bnodes : tv.nodes, node, nb);
            uint32_t h, it;
            const bool take_part = in_blas ? r.in_inst : r.live;
            if (!wave_vote(take_part && r.oct != order))
                packet_children<true>(nb, r, take_part, order, h, it);
            else
                packet_children<false>(nb, r, take_part, order, h, it);
            nbase = nb.n.child_base;
            ngroup = h | ((uint32_t)nb.n.imask << 8);
            ibase = nb.n.item_base;
            items = it;
            continue;
        }
        // the group is used up: leave the instance, take the next group, or stop
        if (in_blas && sp == blas_sp) {
            r.o = r.wo;
            r.inv = r.winv;
            r.oct = octant(r.inv);
            in_blas = false;
            order = (uint32_t)__builtin_amdgcn_readfirstlane((int)r.oct);
        }
        if (sp == 0) break;
        sp--;
        const uint64_t e = stack[sp];
        const uint32_t e0 = wave_uniform((uint32_t)e), e1 = wave_uniform((uint32_t)(e >> 32));
        if (e0 & kItemGroup) {
            ibase = e0 & ~kItemGroup;
            items = e1;
        } else {
            nbase = e0;
            ngroup = e1;
        }
    }
    return true;
}

}  // namespace vh
