// Acceleration-structure layouts and the traversal of the hot path.
//
// A BLAS is a compressed 8-wide BVH over one mesh: 80-byte nodes holding an
// fp32 origin, per-axis power-of-two scales and 8-bit child boxes, plus a
// triangle array (three fp32 vertices, 36 bytes) in leaf order.  A TLAS is
// the same 8-wide layout whose leaves are instances; a leaf child of a TLAS
// node names exactly one instance.
#pragma once

#include "vh_math.h"

namespace vh {

// ------------------------------------------------------------ node layout
// 20 dwords = 80 bytes; loaded as five 16-byte vectors.
//  d0-2  origin xyz (fp32, the node's lower corner)
//  d3    exponents ex, ey, ez (biased by 127: scale = 2^(e-127)), imask
//        (bit s set when slot s holds an internal child)
//  d4    first internal child node index; the internal children are
//        contiguous, in slot order
//  d5    first item index of the leaf children (items are contiguous)
//  d6-7  meta[8]: leaf slot: bits 0-4 item offset from item_base, bits 5-6
//        item count - 1; internal or empty slot: kMetaNoItems
//  d8-19 qlo x[8], qlo y[8], qlo z[8], qhi x[8], qhi y[8], qhi z[8]
// Slots are assigned per node so that visiting slot s in increasing order of
// s ^ octant(ray direction) is roughly near-to-far.
struct alignas(16) WideNode {
    float origin[3];
    uint8_t ex, ey, ez, imask;
    uint32_t child_base;
    uint32_t item_base;
    uint8_t meta[8];
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
};
static_assert(sizeof(WideNode) == 80, "wide node is 80 bytes");

constexpr int kWidth = 8;
#ifndef VH_LEAF_MAX
#define VH_LEAF_MAX 2
#endif
constexpr int kLeafMax = VH_LEAF_MAX;  // triangles per BLAS leaf child (at most 4)
static_assert(kLeafMax >= 1 && kLeafMax <= 4, "leaf meta holds 1-4 items");
constexpr uint32_t kMetaNoItems = 0x80u;

VH_HD float exp_scale(uint8_t e) { return u2f(uint32_t(e) << 23); }

// Decoded child bound on one axis: origin + q * 2^e with one rounding (fused
// multiply-add).  The builder verifies these decoded values enclose the exact
// child bounds, and traversal decodes with the same expression.
VH_HD float decode(float origin, uint8_t q, float scale) { return __builtin_fmaf(float(q), scale, origin); }

// ------------------------------------------------------------ records
// One BLAS on the device: its node and triangle arrays plus the shading data
// of the mesh (original indexed vertices, normals and triangle materials).
constexpr uint32_t kSurfFloats = 32;

struct BlasRecord {
    const WideNode* nodes;
    const float* tris;        // 9 floats per triangle, leaf order
    const uint32_t* tri_ids;  // original primitive index per leaf-order triangle
    const float* positions;   // 3 per vertex
    const float* normals;     // 3 per vertex or null
    const uint32_t* indices;  // 3 per triangle
    const uint32_t* materials;  // per triangle or null
    // per leaf-order triangle one 128-byte record (kSurfFloats floats, one cache
    // line): vertex positions (9 floats), vertex normals (9, zero when the mesh
    // has none), the material (bits of float 18, zero when none), padding
    const float* surf;
    float lo[3], hi[3];       // local bounds
    uint32_t triangle_count, node_count;
    uint32_t surf_normals;    // 1 when the records hold vertex normals
};

// One TLAS instance.  w2o are the world-to-object rows computed on the host in
// double precision and rounded to float (the oracle's vo_invert_transform).
struct InstanceRecord {
    float o2w[12];
    float w2o[12];
    uint32_t blas;
    uint32_t custom_index;
    uint32_t material;
    uint32_t mask_flags;  // bits 0-7 mask, bit 8 cull back faces, bit 9 valid, bit 10 identity (kInstIdentity)
};
// w2o is exactly the identity: traversal enters the instance with the world
// ray itself whenever transforming it would leave it unchanged.
constexpr uint32_t kInstIdentity = 1u << 10;

// Whether the identity rows leave a ray bit for bit unchanged under
// xform_point / xform_dir: they do unless a component is -0.0 (the sums
// 1 x + 0 y + 0 z (+ 0) can turn a -0.0 into +0.0).
VH_HD bool neg_zero(float x) { return f2u(x) == 0x80000000u; }
VH_HD bool identity_exact(f3 o, f3 d) {
    return !(neg_zero(o.x) || neg_zero(o.y) || neg_zero(o.z) || neg_zero(d.x) || neg_zero(d.y) || neg_zero(d.z));
}

struct TlasView {
    const WideNode* nodes;
    const uint32_t* inst_ids;   // leaf item -> instance index
    const InstanceRecord* instances;
    const BlasRecord* blases;
    uint32_t node_count;
    uint32_t instance_count;
};

// ------------------------------------------------------------ primitives
// Watertight crossing of crates/vork-render/shaders/vork_medium_domain.slang:
// 40-66 (Woop, Benthin and Wald 2013 with a symbolic tie at exact zeros).
VH_HD float volume_edge(float px, float py, float qx, float qy) {
    VH_EXACT
    bool forward = px < qx || (px == qx && py < qy);
    float lox = forward ? px : qx, loy = forward ? py : qy;
    float hix = forward ? qx : px, hiy = forward ? qy : py;
    float area = lox * hiy - loy * hix;
    float tie = px != qx ? 1.f : (py != qy ? -1.f : 0.f);
    float canonical = area != 0.f ? area : tie;
    return forward ? canonical : -canonical;
}

// Per-ray shear of the watertight test: axis permutation and shear factors.
struct Shear {
    int kx, ky, kz;
    float sx, sy;
};
VH_HD Shear make_shear(f3 d) {
    VH_EXACT
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int z = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int x = (z + 1) % 3, y = (x + 1) % 3;
    if (comp(d, z) < 0.f) { int s = x; x = y; y = s; }
    Shear s;
    s.kx = x; s.ky = y; s.kz = z;
    s.sx = comp(d, x) / comp(d, z);
    s.sy = comp(d, y) / comp(d, z);
    return s;
}

// Closest-hit triangle test: determinant (|det| > 1e-20, culled when det < 0
// for culling instances), watertight crossing, then Moller-Trumbore t and
// clamped barycentrics (crates/vork-render/shaders/camera_ray.slang:24-35).
VH_HD bool intersect_triangle(f3 a, f3 b, f3 c, f3 o, f3 d, const Shear& sh, float tmin, float tmax, bool cull,
                              float* t_out, float* u_out, float* v_out) {
    VH_EXACT
    f3 e1 = b - a, e2 = c - a;
    f3 q = cross(d, e2);
    float det = dot(e1, q);
    if (!(fabsf(det) > 1e-20f)) return false;
    if (cull && det < 0.f) return false;
    f3 va = a - o, vb = b - o, vc = c - o;
    float pax = comp(va, sh.kx) - sh.sx * comp(va, sh.kz), pay = comp(va, sh.ky) - sh.sy * comp(va, sh.kz);
    float pbx = comp(vb, sh.kx) - sh.sx * comp(vb, sh.kz), pby = comp(vb, sh.ky) - sh.sy * comp(vb, sh.kz);
    float pcx = comp(vc, sh.kx) - sh.sx * comp(vc, sh.kz), pcy = comp(vc, sh.ky) - sh.sy * comp(vc, sh.kz);
    float eu = volume_edge(pbx, pby, pcx, pcy);
    float ev = volume_edge(pcx, pcy, pax, pay);
    float ew = volume_edge(pax, pay, pbx, pby);
    bool crosses = (eu > 0.f && ev > 0.f && ew > 0.f) || (eu < 0.f && ev < 0.f && ew < 0.f);
    if (!crosses) return false;
    f3 s = o - a;
    f3 r = cross(s, e1);
    float t = dot(e2, r) / det;
    if (!(t > tmin && t < tmax)) return false;
    float u = fmaxf(dot(s, q) / det, 0.f), v = fmaxf(dot(d, r) / det, 0.f);
    if (u + v > 1.f) { float k = u + v; u = u / k; v = v / k; }
    *t_out = t; *u_out = u; *v_out = v;
    return true;
}

// Reciprocal direction of the box tests.  Components below 1e-30 in
// magnitude (zero included) become +-1e-30 with their sign first, so an
// axis-parallel ray still prunes on that axis: infinite reciprocals would
// leave the axis unconstrained, and a horizontal ray through a dense instance
// field would then enter every box along its xz path.  The substitute ray
// drifts from the true one by at most 1e-30 * t on the axis, far inside the
// box test's margin (2^-19 of the axis' coordinate magnitude) for any t below
// 1e20.
VH_HD float box_dir(float x) { return fabsf(x) < 1e-30f ? copysignf(1e-30f, x) : x; }
VH_HD f3 box_inv(f3 d) { VH_EXACT return mk(1.f / box_dir(d.x), 1.f / box_dir(d.y), 1.f / box_dir(d.z)); }

// (t, instance, primitive) lexicographic order: the rule at equal t.
VH_HD bool better_hit(float t, uint32_t inst, uint32_t prim, float bt, uint32_t binst, uint32_t bprim) {
    if (t < bt) return true;
    if (t > bt) return false;
    return inst < binst || (inst == binst && prim < bprim);
}

}  // namespace vh
