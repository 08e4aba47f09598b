// Two-level traversal of the 8-wide BVH, one ray per lane, in persistent
// workgroups that keep every lane busy:
//
//  * a lane whose ray finished takes the next ray from a queue (one atomic per
//    wave, shared out by ballot), so long rays do not idle the rest of the
//    wave;
//  * node steps and triangle steps alternate in phases: every lane visits
//    nodes until it holds a leaf (or finishes), then the wave tests leaves
//    together, so the two code paths do not split one wave;
//  * a stack entry carries the distance at which the ray enters it, and
//    entries farther than the current closest hit are dropped without a load.
//
// Stack entries (top two bits): 00 TLAS node, 01 instance, 10 BLAS node,
// 11 BLAS leaf (bits 28-29 item count - 1, bits 0-27 first item).  The top
// kLdsStack entries of a lane live in LDS (interleaved across the workgroup),
// deeper ones in a global overflow array.
#pragma once

#include "vh_bvh.h"

namespace vh {

constexpr int kLdsStack = 12;         // entries per lane in LDS
constexpr int kGlobalStack = 244;     // overflow entries per lane
constexpr uint32_t kStackOverflow = 1u;
constexpr uint32_t kTypeMask = 0xc0000000u;
constexpr uint32_t kTypeTlas = 0x00000000u, kTypeInst = 0x40000000u, kTypeBlas = 0x80000000u, kTypeLeaf = 0xc0000000u;
constexpr uint32_t kMaxLeafItems = 0x0fffffffu;

struct Stack {
    uint32_t* lds;        // &lds_entries[threadIdx.x]
    float* ldst;          // &lds_distances[threadIdx.x]
    uint32_t lds_stride;  // blockDim.x
    uint32_t* glob;       // &overflow_entries[global thread id]
    float* globt;         // &overflow_distances[global thread id]
    uint32_t glob_stride; // threads in the grid
    int sp;
    __device__ __forceinline__ bool push(uint32_t v, float t) {
        if (sp < kLdsStack) {
            lds[sp * lds_stride] = v;
            ldst[sp * lds_stride] = t;
        } else if (sp < kLdsStack + kGlobalStack) {
            glob[(sp - kLdsStack) * glob_stride] = v;
            globt[(sp - kLdsStack) * glob_stride] = t;
        } else {
            return false;
        }
        sp++;
        return true;
    }
    __device__ __forceinline__ uint32_t peek() const {
        const int i = sp - 1;
        return i < kLdsStack ? lds[i * lds_stride] : glob[(i - kLdsStack) * glob_stride];
    }
    __device__ __forceinline__ uint32_t pop(float& t) {
        sp--;
        if (sp < kLdsStack) {
            t = ldst[sp * lds_stride];
            return lds[sp * lds_stride];
        }
        t = globt[(sp - kLdsStack) * glob_stride];
        return glob[(sp - kLdsStack) * glob_stride];
    }
};

struct TraceHit {
    float t, u, v;
    uint32_t inst, prim;
    bool found;
};

struct TraceCounts {
    uint32_t nodes, tris;
};

// Compare-exchange of the 8-key sorting network (descending).
#define VH_CX(a, b)                     \
    {                                   \
        uint32_t hi_ = max(k[a], k[b]); \
        uint32_t lo_ = min(k[a], k[b]); \
        k[a] = hi_;                     \
        k[b] = lo_;                     \
    }

__device__ __forceinline__ void sort8_desc(uint32_t k[8]) {
    // Batcher's odd-even merge sort network, 19 compare-exchanges.
    VH_CX(0, 1) VH_CX(2, 3) VH_CX(4, 5) VH_CX(6, 7)
    VH_CX(0, 2) VH_CX(1, 3) VH_CX(4, 6) VH_CX(5, 7)
    VH_CX(1, 2) VH_CX(5, 6) VH_CX(0, 4) VH_CX(3, 7)
    VH_CX(1, 5) VH_CX(2, 6)
    VH_CX(1, 4) VH_CX(3, 6)
    VH_CX(2, 4) VH_CX(3, 5)
    VH_CX(3, 4)
}
#undef VH_CX

union NodeBits {
    uint4 v[5];
    WideNode n;
};

// meta byte of child `slot` without indexing a register array dynamically
__device__ __forceinline__ uint32_t node_meta(const NodeBits& nb, uint32_t slot) {
    uint32_t w = slot < 4u ? nb.v[1].z : nb.v[1].w;
    return (w >> ((slot & 3u) * 8u)) & 0xffu;
}

__device__ __forceinline__ void load_node(const WideNode* base, uint32_t i, NodeBits& nb) {
    const uint4* p = reinterpret_cast<const uint4*>(base + i);
#pragma unroll
    for (int k = 0; k < 5; k++) nb.v[k] = p[k];
}

constexpr uint32_t kKeyLive = 0x80000000u;

// Child entry keys: a live child's entry distance bits (non-negative, so they
// order like the values) with the low 3 bits replaced by its slot and the top
// bit set; 0 for a child that is missed or absent.
//
// The near plane of each axis is chosen once per node from the ray's
// direction signs, so every child costs six plane decodes, six slab
// distances and two three-way min/max.  Each plane is decoded to the float
// the builder verified, then (plane - o) * inv with two roundings: the far
// distance is scaled by (1 + 2 gamma(3)) so that the test stays conservative.
__device__ __forceinline__ void child_keys(const NodeBits& nb, f3 o, f3 inv, float tmin, float tmax, uint32_t k[8]) {
    const WideNode& n = nb.n;
    const float sx = exp_scale(n.ex), sy = exp_scale(n.ey), sz = exp_scale(n.ez);
    // q planes: qlo x = v[2].xy, y = v[2].zw, z = v[3].xy; qhi x = v[3].zw, y = v[4].xy, z = v[4].zw
    const bool px = inv.x >= 0.f, py = inv.y >= 0.f, pz = inv.z >= 0.f;
    const uint32_t nx0 = px ? nb.v[2].x : nb.v[3].z, nx1 = px ? nb.v[2].y : nb.v[3].w;
    const uint32_t fx0 = px ? nb.v[3].z : nb.v[2].x, fx1 = px ? nb.v[3].w : nb.v[2].y;
    const uint32_t ny0 = py ? nb.v[2].z : nb.v[4].x, ny1 = py ? nb.v[2].w : nb.v[4].y;
    const uint32_t fy0 = py ? nb.v[4].x : nb.v[2].z, fy1 = py ? nb.v[4].y : nb.v[2].w;
    const uint32_t nz0 = pz ? nb.v[3].x : nb.v[4].z, nz1 = pz ? nb.v[3].y : nb.v[4].w;
    const uint32_t fz0 = pz ? nb.v[4].z : nb.v[3].x, fz1 = pz ? nb.v[4].w : nb.v[3].y;
    const uint32_t count = n.count;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const uint32_t sh = (c & 3) * 8;
        const uint32_t nxw = c < 4 ? nx0 : nx1, fxw = c < 4 ? fx0 : fx1;
        const uint32_t nyw = c < 4 ? ny0 : ny1, fyw = c < 4 ? fy0 : fy1;
        const uint32_t nzw = c < 4 ? nz0 : nz1, fzw = c < 4 ? fz0 : fz1;
        float tnx, tny, tnz, tfx, tfy, tfz;
        {
            VH_EXACT
            tnx = (decode(n.origin[0], (uint8_t)(nxw >> sh), sx) - o.x) * inv.x;
            tfx = (decode(n.origin[0], (uint8_t)(fxw >> sh), sx) - o.x) * inv.x;
            tny = (decode(n.origin[1], (uint8_t)(nyw >> sh), sy) - o.y) * inv.y;
            tfy = (decode(n.origin[1], (uint8_t)(fyw >> sh), sy) - o.y) * inv.y;
            tnz = (decode(n.origin[2], (uint8_t)(nzw >> sh), sz) - o.z) * inv.z;
            tfz = (decode(n.origin[2], (uint8_t)(fzw >> sh), sz) - o.z) * inv.z;
        }
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        const bool live = (uint32_t)c < count && tn <= tf * (1.f + 2.f * kGamma3);
        k[c] = live ? ((f2u(tn) & 0x7ffffff8u) | (uint32_t)c | kKeyLive) : 0u;  // sign cleared: tn >= +-0
    }
}

// One lane's traversal state.
template <bool kAny, bool kCount>
struct Tracer {
    f3 o, d, inv;           // world ray
    f3 lo, ld, linv;        // ray in the current instance's object space
    Shear lsh;
    float tmin, tmax;
    uint32_t mask;
    uint32_t cur_inst;
    bool cull;
    const BlasRecord* blas;
    uint32_t leaf;          // pending leaf entry, 0 = none
    bool done;
    TraceHit hit;
    TraceCounts cnt;

    __device__ __forceinline__ void begin(const TlasView& tv, f3 o_, f3 d_, float tmin_, float tmax_, uint32_t mask_,
                                          Stack& st) {
        o = o_; d = d_; inv = safe_inv(d_);
        lo = o_; ld = d_; linv = inv;
        lsh = make_shear(d_);
        tmin = tmin_; tmax = tmax_; mask = mask_;
        cur_inst = 0; cull = false; blas = nullptr;
        leaf = 0;
        hit.found = false; hit.t = tmax_; hit.u = hit.v = 0.f; hit.inst = hit.prim = 0xffffffffu;
        cnt.nodes = cnt.tris = 0;
        st.sp = 0;
        done = tv.node_count == 0;
        if (!done) st.push(kTypeTlas, tmin_);
    }

    // Pops one entry: visits a node, enters an instance, or takes a leaf.
    // Returns false when the stack overflowed.
    __device__ __forceinline__ bool node_step(const TlasView& tv, Stack& st) {
        if (st.sp == 0) { done = true; return true; }
        float te;
        uint32_t e = st.pop(te);
        // entered beyond the closest hit so far: te is a rounded entry distance, so
        // compare with the same (1 + 2 gamma(3)) margin as the box test
        if (!kAny && te > hit.t * (1.f + 2.f * kGamma3)) return true;
        const uint32_t type = e & kTypeMask;
        if (type == kTypeLeaf) { leaf = e; return true; }
        if (type == kTypeInst) {
            const uint32_t inst = e & ~kTypeMask;
            const InstanceRecord& r = tv.instances[inst];
            const uint32_t mf = r.mask_flags;
            if (!(mf & (1u << 9)) || !(mf & mask & 0xffu)) return true;
            lo = xform_point(r.w2o, o);
            ld = xform_dir(r.w2o, d);
            linv = safe_inv(ld);
            lsh = make_shear(ld);
            cull = (mf >> 8) & 1u;
            blas = &tv.blases[r.blas];
            cur_inst = inst;
            return st.push(kTypeBlas, te);
        }
        const bool in_blas = type == kTypeBlas;
        NodeBits nb;
        load_node(in_blas ? blas->nodes : tv.nodes, e & ~kTypeMask, nb);
        if (kCount) cnt.nodes++;
        uint32_t k[8];
        child_keys(nb, in_blas ? lo : o, in_blas ? linv : inv, tmin, kAny ? tmax : hit.t, k);
        sort8_desc(k);
        // far to near, so the nearest entry is popped first; missed children sort last
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (k[j] == 0u) break;
            const uint32_t meta = node_meta(nb, k[j] & 7u);
            const float t = u2f(k[j] & ~(7u | kKeyLive));
            uint32_t entry;
            if (meta & kMetaInternal) {
                entry = (nb.n.child_base + (meta & 7u)) | type;
            } else if (in_blas) {
                entry = kTypeLeaf | (((meta >> 5) & 3u) << 28) | (nb.n.item_base + (meta & 31u));
            } else {
                entry = kTypeInst | tv.inst_ids[nb.n.item_base + (meta & 31u)];
            }
            if (!st.push(entry, t)) return false;
        }
        return true;
    }

    // Tests the pending leaf's triangles.
    __device__ __forceinline__ void leaf_step() {
        const uint32_t first = leaf & kMaxLeafItems, count = ((leaf >> 28) & 3u) + 1u;
        leaf = 0;
        for (uint32_t q = 0; q < count; q++) {
            const uint32_t item = first + q;
            const float* tp = blas->tris + 9 * (size_t)item;
            const f3 a = mk(tp[0], tp[1], tp[2]), b = mk(tp[3], tp[4], tp[5]), c = mk(tp[6], tp[7], tp[8]);
            float t, u, v;
            if (kCount) cnt.tris++;
            // admit equal t so that the tie rule can pick the lower id
            const float lim = (kAny || !hit.found) ? tmax : next_up(hit.t);
            if (intersect_triangle(a, b, c, lo, ld, lsh, tmin, lim, cull, &t, &u, &v)) {
                if (!(t < tmax)) continue;
                if (kAny) {
                    hit.found = true;
                    hit.t = t;
                    done = true;
                    return;
                }
                const uint32_t prim = blas->tri_ids[item];
                if (!hit.found || better_hit(t, cur_inst, prim, hit.t, hit.inst, hit.prim)) {
                    hit.found = true; hit.t = t; hit.u = u; hit.v = v; hit.inst = cur_inst; hit.prim = prim;
                }
            }
        }
    }
};

// Persistent trace loop.  `src` supplies rays and consumes results:
//   uint32_t count() const;
//   void load(uint32_t i, f3& o, f3& d, float& tmin, float& tmax, uint32_t& mask) const;
//   void store(uint32_t i, const TraceHit& h, const TraceCounts& c) const;
// `next` is a global counter of handed-out ray indices, zero at launch.
// Every lane leaves the loop once the queue is empty and its ray is stored.
// Returns false when a stack overflowed.
template <bool kAny, bool kCount, class Src>
__device__ bool trace_persistent(const TlasView& tv, const Src& src, uint32_t* next, Stack& st) {
    const uint32_t n = src.count();
    Tracer<kAny, kCount> tr;
    bool active = false, exhausted = false, ok = true;
    uint32_t ray = 0;
    const uint64_t lt_mask = (1ull << __lane_id()) - 1ull;
    while (true) {
        // lanes without a ray take new ones: one atomic per wave
        const uint64_t need = __ballot(!active && !exhausted);
        if (need) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(next, (uint32_t)__popcll(need));
            base = __shfl(base, leader);
            if (!active && !exhausted) {
                const uint32_t idx = base + (uint32_t)__popcll(need & lt_mask);
                if (idx < n) {
                    f3 o, d;
                    float t0, t1;
                    uint32_t m;
                    src.load(idx, o, d, t0, t1, m);
                    tr.begin(tv, o, d, t0, t1, m, st);
                    ray = idx;
                    active = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__ballot(active)) break;
        // node phase: visit nodes until every lane holds a leaf or is done.  A
        // lane that already holds a leaf keeps visiting node entries on top of
        // its stack meanwhile (speculative traversal, Aila and Laine 2009), so
        // it does not idle while the others search.
        while (true) {
            const bool searching = active && !tr.done && tr.leaf == 0;
            if (!__ballot(searching)) break;
            bool step = searching;
            if (active && !tr.done && tr.leaf != 0 && st.sp > 0) {
                // node entries only: a leaf or an instance entry would change
                // the object space the pending leaf is tested in
                const uint32_t top = st.peek() & kTypeMask;
                step = top == kTypeBlas || top == kTypeTlas;
            }
            if (step && !tr.node_step(tv, st)) {
                ok = false;
                tr.done = true;
            }
        }
        // leaf phase
        if (active && tr.leaf != 0 && !tr.done) tr.leaf_step();
        if (active && tr.done) {
            src.store(ray, tr.hit, tr.cnt);
            active = false;
        }
    }
    return ok;
}

}  // namespace vh
