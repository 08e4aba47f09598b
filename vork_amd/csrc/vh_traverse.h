// Two-level traversal of the 8-wide BVH: one ray per lane, stack entries in
// LDS (kLdsStack per lane, interleaved across the workgroup so a wave's
// pushes hit distinct banks) with an overflow tail in global memory.
#pragma once

#include "vh_bvh.h"

namespace vh {

constexpr int kLdsStack = 16;         // entries per lane in LDS
constexpr int kGlobalStack = 240;     // overflow entries per lane
constexpr uint32_t kExit = 0xffffffffu;
constexpr uint32_t kBlasBit = 0x80000000u;
constexpr uint32_t kInstBit = 0x40000000u;
constexpr uint32_t kStackOverflow = 1u;

struct Stack {
    uint32_t* lds;        // &lds[threadIdx.x]
    uint32_t lds_stride;  // blockDim.x
    uint32_t* glob;       // &overflow[global thread id]
    uint32_t glob_stride; // threads in the grid
    int sp;
    __device__ __forceinline__ bool push(uint32_t v) {
        if (sp < kLdsStack) lds[sp * lds_stride] = v;
        else if (sp < kLdsStack + kGlobalStack) glob[(sp - kLdsStack) * glob_stride] = v;
        else return false;
        sp++;
        return true;
    }
    __device__ __forceinline__ uint32_t pop() {
        sp--;
        return sp < kLdsStack ? lds[sp * lds_stride] : glob[(sp - kLdsStack) * glob_stride];
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

// Compare-exchange of the 8-key sorting network (ascending).
#define VH_CX(a, b)                        \
    {                                      \
        uint32_t lo_ = min(k[a], k[b]);    \
        uint32_t hi_ = max(k[a], k[b]);    \
        k[a] = lo_;                        \
        k[b] = hi_;                        \
    }

__device__ __forceinline__ void sort8(uint32_t k[8]) {
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

// Child entry keys: (entry distance bits with the low 3 bits replaced by the
// slot), or ~0 for a child that is missed or absent.
__device__ __forceinline__ void child_keys(const WideNode& n, f3 o, f3 inv, float tmin, float tmax, uint32_t k[8]) {
    float sx = exp_scale(n.ex), sy = exp_scale(n.ey), sz = exp_scale(n.ez);
#pragma unroll
    for (int c = 0; c < 8; c++) {
        float lox = decode(n.origin[0], n.qlo[0][c], sx), hix = decode(n.origin[0], n.qhi[0][c], sx);
        float loy = decode(n.origin[1], n.qlo[1][c], sy), hiy = decode(n.origin[1], n.qhi[1][c], sy);
        float loz = decode(n.origin[2], n.qlo[2][c], sz), hiz = decode(n.origin[2], n.qhi[2][c], sz);
        float t = box_enter(lox, loy, loz, hix, hiy, hiz, o, inv, tmin, tmax);
        bool live = c < n.count && t != __builtin_huge_valf();
        k[c] = live ? ((f2u(t) & ~7u) | (uint32_t)c) : 0xffffffffu;
    }
}

// Closest hit (kAny = false) or any hit (kAny = true) against the TLAS.
// Returns false when the stack overflowed (the result is then incomplete).
template <bool kAny, bool kCount>
__device__ bool traverse(const TlasView& tv, f3 o, f3 d, float tmin, float tmax, uint32_t mask, Stack& st,
                         TraceHit& hit, TraceCounts& cnt) {
    hit.found = false;
    hit.t = tmax;
    hit.inst = hit.prim = 0xffffffffu;
    hit.u = hit.v = 0.f;
    if (tv.node_count == 0) return true;
    const f3 inv = safe_inv(d);
    f3 lo = o, ld = d, linv = inv;
    Shear lsh = make_shear(d);
    bool cull = false;
    uint32_t cur_inst = 0;
    const BlasRecord* blas = nullptr;
    st.sp = 0;
    st.push(0u);
    bool ok = true;
    while (st.sp > 0) {
        uint32_t e = st.pop();
        if (e == kExit) continue;
        if ((e & (kBlasBit | kInstBit)) == kInstBit) {
            uint32_t inst = e & ~kInstBit;
            const InstanceRecord& r = tv.instances[inst];
            uint32_t mf = r.mask_flags;
            if (!(mf & (1u << 9)) || !(mf & mask & 0xffu)) continue;
            lo = xform_point(r.w2o, o);
            ld = xform_dir(r.w2o, d);
            linv = safe_inv(ld);
            lsh = make_shear(ld);
            cull = (mf >> 8) & 1u;
            blas = &tv.blases[r.blas];
            cur_inst = inst;
            if (!st.push(kExit) || !st.push(kBlasBit)) { ok = false; break; }
            continue;
        }
        const bool in_blas = (e & kBlasBit) != 0;
        NodeBits nb;
        load_node(in_blas ? blas->nodes : tv.nodes, e & ~kBlasBit, nb);
        if (kCount) cnt.nodes++;
        const f3 ro = in_blas ? lo : o;
        const f3 ri = in_blas ? linv : inv;
        uint32_t k[8];
        child_keys(nb.n, ro, ri, tmin, kAny ? tmax : hit.t, k);
        sort8(k);
        // leaves first, near to far
        if (in_blas) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (k[j] == 0xffffffffu) break;
                uint32_t meta = node_meta(nb, k[j] & 7u);
                if (meta & kMetaInternal) continue;
                uint32_t first = nb.n.item_base + (meta & 31u), count = ((meta >> 5) & 3u) + 1u;
                for (uint32_t q = 0; q < count; q++) {
                    uint32_t item = first + q;
                    const float* tp = blas->tris + 9 * (size_t)item;
                    f3 a = mk(tp[0], tp[1], tp[2]), b = mk(tp[3], tp[4], tp[5]), c = mk(tp[6], tp[7], tp[8]);
                    float t, u, v;
                    if (kCount) cnt.tris++;
                    // admit equal t so that the tie rule can pick the lower id
                    float lim = (kAny || !hit.found) ? tmax : next_up(hit.t);
                    if (intersect_triangle(a, b, c, lo, ld, lsh, tmin, lim, cull, &t, &u, &v)) {
                        if (!(t < tmax)) continue;
                        if (kAny) { hit.found = true; hit.t = t; return true; }
                        uint32_t prim = blas->tri_ids[item];
                        if (!hit.found || better_hit(t, cur_inst, prim, hit.t, hit.inst, hit.prim)) {
                            hit.found = true; hit.t = t; hit.u = u; hit.v = v; hit.inst = cur_inst; hit.prim = prim;
                        }
                    }
                }
            }
        }
        // internal children (and TLAS instances), far to near so the nearest is popped first
#pragma unroll
        for (int j = 7; j >= 0; j--) {
            if (k[j] == 0xffffffffu) continue;
            uint32_t slot = k[j] & 7u;
            uint32_t meta = node_meta(nb, slot);
            uint32_t entry;
            if (meta & kMetaInternal) {
                entry = (nb.n.child_base + (meta & 7u)) | (in_blas ? kBlasBit : 0u);
            } else if (!in_blas) {
                entry = kInstBit | tv.inst_ids[nb.n.item_base + (meta & 31u)];
            } else {
                continue;
            }
            if (!st.push(entry)) { ok = false; break; }
        }
        if (!ok) break;
    }
    return ok;
}

}  // namespace vh
