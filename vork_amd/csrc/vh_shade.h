// Device restatement of the reference estimator for the closure subset the
// workloads use.  Each function names the reference file:line it follows;
// the CPU oracle (oracle/vork_oracle.c) restates the same lines.
#pragma once

#include "vh_math.h"

namespace vh {

// ------------------------------------------------------------- constants
constexpr uint32_t kSampleHashMix = 0x45d9f3bu;      // crates/vork-slang/src/registry.rs:389
constexpr uint32_t kStreamStep = 0x9e3779b9u;        // registry.rs:391
constexpr uint32_t kRouletteStart = 3u;              // renderer_registry.rs:98-100
constexpr float kRouletteMin = 0.05f, kRouletteMax = 0.95f;
constexpr float kLightTreeUniformMix = 0.25f;        // renderer_registry.rs:109-111
constexpr float kLightTreeMinDistanceSquared = 1e-12f;
constexpr float kGgxAlphaMin = 0.0001f;              // registry.rs:384
constexpr uint32_t kBrdfSize = 128u, kBrdfAlbedoSize = 64u, kBrdfIorSteps = 64u;  // registry.rs:375-378
constexpr uint32_t kEnvironmentSamples = 512u, kDielectricAlbedoSamples = 512u;
constexpr float kUintScale = 0.000000059604644775390625f;
constexpr float kRadicalInverseScale = 2.3283064365386963e-10f;
__device__ constexpr float kLuminance[3] = {0.2126f, 0.7152f, 0.0722f};

// --------------------------------------------------------------- sampler
// crates/vork-slang/src/vork_sampling.slang:3-15
VH_HD uint32_t pixel_hash(uint32_t seed) {
    uint32_t v = (seed ^ (seed >> 16)) * kSampleHashMix;
    v = (v ^ (v >> 16)) * kSampleHashMix;
    return v ^ (v >> 16);
}
VH_HD uint32_t screen_ray_hash(uint32_t px, uint32_t py, uint32_t index, uint32_t stream) {
    uint32_t spatial = pixel_hash((px * kSampleHashMix) ^ py);
    return pixel_hash(spatial ^ pixel_hash(index * kStreamStep + stream));
}
// crates/vork-render/shaders/path_sampler.slang:65-68
VH_HD uint32_t owen_reversed(uint32_t x, uint32_t seed) {
    seed = pixel_hash(seed);
    x ^= x * 0x3d20adeau;
    x += seed;
    x *= (seed >> 16) | 1u;
    x ^= x * 0x05526c56u;
    x ^= x * 0x53a22864u;
    return x;
}
VH_HD float sample_unit(uint32_t x) { return ((float)(x >> 9) + 0.5f) * (1.f / 8388608.f); }

__device__ __forceinline__ uint32_t rev(uint32_t x) { return __builtin_bitreverse32(x); }

struct Sampler {
    uint32_t index, seed, mask;
};
// path_sampler.slang:54-63
__device__ __forceinline__ Sampler make_sampler(uint32_t px, uint32_t py, uint32_t index, uint32_t count,
                                                uint32_t stream) {
    Sampler s;
    s.index = index;
    s.seed = screen_ray_hash(px, py, 0, stream);
    s.mask = count <= 1 ? 0u : (2u << (31 - __clz(count - 1))) - 1u;
    return s;
}
// path_sampler.slang:74-93: Sobol dimensions 1-3 from their generator
// recurrences (degree 1, 2, 3 primitive polynomials), one column per bit.
// The four scrambled 32-bit values of a group; group_values maps them to
// (0, 1) with 2^23 buckets.
__device__ __forceinline__ void group_bits(const Sampler& s, uint32_t group, uint32_t out[4]) {
    uint32_t seed = pixel_hash(s.seed ^ pixel_hash(group + 1u));
    uint32_t index = rev(owen_reversed(rev(s.index), seed ^ 0x6c8e9cf5u)) & s.mask;
    uint32_t x0 = 0, x1 = 0, x2 = 0;
    uint32_t p = 0x80000000u, qx = 0x80000000u, qy = 0xc0000000u;
    uint32_t rx = 0x80000000u, ry = 0xc0000000u, rz = 0x20000000u;
    for (uint32_t bits = index; bits != 0u; bits >>= 1) {
        if (bits & 1u) { x0 ^= p; x1 ^= qx; x2 ^= rx; }
        p ^= p >> 1;
        uint32_t nq = qy ^ qx ^ (qx >> 2);
        qx = qy; qy = nq;
        uint32_t nr = ry ^ rx ^ (rx >> 3);
        rx = ry; ry = rz; rz = nr;
    }
    out[0] = rev(owen_reversed(index, seed ^ 0xa511e9b3u));
    out[1] = rev(owen_reversed(rev(x0), seed ^ 0x63d83595u));
    out[2] = rev(owen_reversed(rev(x1), seed ^ 0x2f7c4a1bu));
    out[3] = rev(owen_reversed(rev(x2), seed ^ kStreamStep));
}
__device__ __forceinline__ void group_values(const Sampler& s, uint32_t group, float out[4]) {
    uint32_t b[4];
    group_bits(s, group, b);
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = sample_unit(b[k]);
}
// path_sampler.slang:72
__device__ __forceinline__ float path_random(uint32_t& state) {
    state = pixel_hash(state + kStreamStep);
    return sample_unit(state);
}

// ----------------------------------------------------------------- film
// crates/vork-render/shaders/path.slang:1712-1719
__device__ __forceinline__ float filter_offset(const float* table, float u) {
    float t = clampf(u, 0.f, 1.f) * 1024.f;
    uint32_t i = min((uint32_t)t, 1023u);
    return lerpf(table[i], table[i + 1], t - (float)i);
}

// --------------------------------------------------------------- BSDF
struct Closure {
    f3 base_color;
    float metal, spec_weight, roughness, ior;
    f3 n;  // geometry_normal
};
struct Prep {
    float p0, p1;  // diffuse, reflection probabilities (refraction is 0 here)
    f3 diffuse, metal_comp, smooth, interface_n;
    float bump_a2;
};
struct Luts {
    const float4* moments;   // [row alpha][column cosine]
    const float* dielectric; // [slice][r][nv]
};

VH_HD float specular_alpha(float r) { return fmaxf(r * r, kGgxAlphaMin); }  // openpbr_core.slang:5-7
// openpbr_core.slang:13-21
__device__ __forceinline__ float modulated_ior(float eta, float weight) {
    if (weight == 1.f || eta == 1.f) return eta;
    float contrast = div_a(eta - 1.f, eta + 1.f);
    float limit = u2f(0x3f7fffffu);
    float epsilon = clampf(sqrt_a(fmaxf(weight, 0.f)) * contrast, -limit, limit);
    return div_a(1.f + epsilon, 1.f - epsilon);
}
// openpbr_core.slang:22-27
VH_HD float fresnel(float c, float eta) {
    c = clampf(c, 0.f, 1.f);
    if (eta == 1.f) return 0.f;
    float si2 = div_a(1.f - c * c, eta * eta);
    if (si2 >= 1.f) return 1.f;
    float ct = sqrt_a(fmaxf(0.f, 1.f - si2));
    float rs = div_a(c - eta * ct, c + eta * ct), rp = div_a(eta * c - ct, eta * c + ct);
    return (rs * rs + rp * rp) * .5f;
}
// openpbr_core.slang:105-116
VH_HD float distribution(float nh, float alpha) {
    float x = (1.f - nh * nh) + nh * nh * alpha * alpha;
    return alpha * alpha * rcp_a(kPi * x * x);
}
VH_HD float smith_denominator(float no, float ni, float alpha) {
    float ci = fabsf(ni), a2 = alpha * alpha;
    float po = sqrt_a(a2 + (1.f - a2) * no * no), pi = sqrt_a(a2 + (1.f - a2) * ci * ci);
    return ci * po + no * pi;
}
VH_HD float g2(float no, float ni, float alpha) { return div_a(2.f * no * fabsf(ni), smith_denominator(no, ni, alpha)); }
// openpbr_core.slang:126-138 (Heitz 2018)
VH_HD f3 ggx_visible_normal(f3 view, float ax, float ay, float u0, float u1) {
    f3 v = normalize_a(mk(ax * view.x, ay * view.y, view.z));
    float ls = v.x * v.x + v.y * v.y;
    f3 t = ls > 0.f ? mk(-v.y, v.x, 0.f) * rsqrt_a(ls) : mk(1.f, 0.f, 0.f);
    f3 b = cross(v, t);
    float radius = sqrt_a(u0);
    float x = radius * cos_rev(u1);
    float y = lerpf(sqrt_a(fmaxf(1.f - x * x, 0.f)), radius * sin_rev(u1), .5f * (1.f + v.z));
    f3 p = t * x + b * y + v * sqrt_a(fmaxf(1.f - x * x - y * y, 0.f));
    return normalize_a(mk(ax * p.x, ay * p.y, fmaxf(p.z, 0.f)));
}
// openpbr_core.slang:141-152
VH_HD float ggx_visible_weight(float no, float ni, float alpha) {
    if (ni <= 0.f) return 0.f;
    float a2 = alpha * alpha;
    return div_a(ni * (no + sqrt_a(a2 + (1.f - a2) * no * no)), smith_denominator(no, ni, alpha));
}
VH_HD float ggx_visible_reflection(float no, float alpha, float u0, float u1, float* vh) {
    f3 view = mk(sqrt_a(fmaxf(1.f - no * no, 0.f)), 0.f, no);
    f3 h = ggx_visible_normal(view, alpha, alpha, u0, u1);
    f3 light = reflect3(-view, h);
    *vh = clampf(dot(view, h), 0.f, 1.f);
    return ggx_visible_weight(no, light.z, alpha);
}
// orthonormal_basis.slang:2-9 (Duff et al. 2017)
VH_HD f3 basis_direction(f3 n, f3 v) {
    float s = n.z >= 0.f ? 1.f : -1.f, a = -rcp_a(s + n.z), b = n.x * n.y * a;
    f3 r0 = mk(1.f + s * n.x * n.x * a, s * b, -s * n.x), r1 = mk(b, s + n.y * n.y * a, -n.y);
    return r0 * v.x + r1 * v.y + n * v.z;
}
// Bilinear fetch with clamp to edge on a grid with texel centres at integers.
__device__ __forceinline__ float bilinear1(const float* g, uint32_t w, uint32_t h, float x, float y) {
    x = clampf(x, 0.f, (float)(w - 1)); y = clampf(y, 0.f, (float)(h - 1));
    uint32_t x0 = (uint32_t)x, y0 = (uint32_t)y;
    uint32_t x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
    float fx = x - (float)x0, fy = y - (float)y0;
    return lerpf(lerpf(g[y0 * w + x0], g[y0 * w + x1], fx), lerpf(g[y1 * w + x0], g[y1 * w + x1], fx), fy);
}
__device__ __forceinline__ float4 bilinear4(const float4* g, uint32_t w, uint32_t h, float x, float y) {
    x = clampf(x, 0.f, (float)(w - 1)); y = clampf(y, 0.f, (float)(h - 1));
    uint32_t x0 = (uint32_t)x, y0 = (uint32_t)y;
    uint32_t x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
    float fx = x - (float)x0, fy = y - (float)y0;
    float4 a = g[y0 * w + x0], b = g[y0 * w + x1], c = g[y1 * w + x0], d = g[y1 * w + x1];
    float4 r;
    r.x = lerpf(lerpf(a.x, b.x, fx), lerpf(c.x, d.x, fx), fy);
    r.y = lerpf(lerpf(a.y, b.y, fx), lerpf(c.y, d.y, fx), fy);
    r.z = lerpf(lerpf(a.z, b.z, fx), lerpf(c.z, d.z, fx), fy);
    r.w = lerpf(lerpf(a.w, b.w, fx), lerpf(c.w, d.w, fx), fy);
    return r;
}
// crates/vork-render/shaders/frame.slang:41-51
__device__ __forceinline__ float4 brdf_moments(const Luts& L, float nv, float roughness) {
    float ux = sqrt_a(sqrt_a(clampf(nv, 0.f, 1.f)));
    float uy = log2_a(clampf(roughness * roughness, kGgxAlphaMin, 1.f) * (1.f / kGgxAlphaMin)) * (1.f / 13.2877123795494f);
    return bilinear4(L.moments, kBrdfSize, kBrdfSize, ux * (float)(kBrdfSize - 1), uy * (float)(kBrdfSize - 1));
}
// frame.slang:52-66
__device__ __forceinline__ float dielectric_albedo(const Luts& L, float nv, float roughness, float eta) {
    if (eta == 1.f) return 0.f;
    float at = div_a(eta, 1.f + eta) * (float)kBrdfIorSteps;
    uint32_t first = min((uint32_t)at, kBrdfIorSteps - 1);
    const uint32_t n = kBrdfAlbedoSize;
    float x = clampf(nv, 0.f, 1.f) * (float)(n - 1), y = clampf(roughness, 0.f, 1.f) * (float)(n - 1);
    float a = bilinear1(L.dielectric + first * n * n, n, n, x, y);
    float b = bilinear1(L.dielectric + (first + 1) * n * n, n, n, x, y);
    return lerpf(a, b, at - (float)first);
}
__device__ __forceinline__ float dielectric_ior(const Closure& c) { return modulated_ior(c.ior, c.spec_weight); }
// openpbr_core.slang:36-49 (white F82 tint: no edge correction)
__device__ __forceinline__ f3 reflection_fresnel_at_ior(f3 base, float metal, float sw, float cosine, float eta) {
    cosine = clampf(cosine, 0.f, 1.f);
    float fd = fresnel(cosine, eta);
    f3 dielectric = mk(fd, fd, fd);
    if (metal <= 0.f) return dielectric;
    const float x = 1.f - cosine, x2 = x * x;
    float f = x2 * x2 * x;
    f3 m = mk(fmaxf(base.x + (1.f - base.x) * f, 0.f), fmaxf(base.y + (1.f - base.y) * f, 0.f),
              fmaxf(base.z + (1.f - base.z) * f, 0.f));
    return lerp3(dielectric, m * sw, metal);
}
// crates/vork-render/shaders/bsdf_energy.slang:76-88 with bsdf_film.slang:110
__device__ __forceinline__ f3 compensated_surface_reflection(const Closure& c, float cosine, f3 comp_) {
    f3 f = reflection_fresnel_at_ior(c.base_color, c.metal, c.spec_weight, cosine, dielectric_ior(c));
    if (comp_.x > 0.f || comp_.y > 0.f || comp_.z > 0.f) {
        if (c.metal == 1.f) {
            f = f + comp_ * f;
        } else {
            f3 conductor = reflection_fresnel_at_ior(c.base_color, 1.f, c.spec_weight, cosine, dielectric_ior(c));
            f = f + (comp_ * conductor) * c.metal;
        }
    }
    return f;
}
// bsdf_energy.slang:44-74 (isotropic, white tint)
__device__ __forceinline__ f3 metal_compensation(const Luts& L, const Closure& c, f3 view) {
    if (c.metal <= 0.f || c.roughness == 0.f) return mk(0.f, 0.f, 0.f);
    float4 m = brdf_moments(L, fmaxf(dot(c.n, view), .0001f), c.roughness);
    float energy = m.x + m.y;
    float e_avg = brdf_moments(L, 0.f, c.roughness).w;
    float missing = rcp_a(fmaxf(energy, 1e-6f)) - 1.f;
    f3 f0 = c.base_color;
    const float k21 = 1.f / 21.f;
    f3 avg = mk(clampf(f0.x + (1.f - f0.x) * k21, 0.f, 1.f), clampf(f0.y + (1.f - f0.y) * k21, 0.f, 1.f),
                clampf(f0.z + (1.f - f0.z) * k21, 0.f, 1.f));
    return mk(missing * avg.x * e_avg * rcp_a(fmaxf(1.f - avg.x * (1.f - e_avg), 1e-6f)),
              missing * avg.y * e_avg * rcp_a(fmaxf(1.f - avg.y * (1.f - e_avg), 1e-6f)),
              missing * avg.z * e_avg * rcp_a(fmaxf(1.f - avg.z * (1.f - e_avg), 1e-6f)));
}
// crates/vork-render/shaders/bsdf_response.slang:98-110
__device__ __forceinline__ f3 direct_diffuse(const Luts& L, const Closure& c, f3 view) {
    if (c.metal < 1.f) {
        float albedo = dielectric_albedo(L, fmaxf(dot(c.n, view), .0001f), c.roughness, dielectric_ior(c));
        return c.base_color * ((1.f - c.metal) * (1.f - albedo) * kInvPi);
    }
    return mk(0.f, 0.f, 0.f);
}
// Denoiser guide albedo (path.slang:1553-1570): the closure's directional
// albedo toward the viewer, diffuse (1 - metal)(1 - E_spec) base plus the
// specular reflectance of the same GGX tables (dielectric: E_spec at the
// modulated IOR; metal: F0 * m.x + m.y with the multiple-scattering
// compensation, times the specular weight), clamped to [0, 1].
__device__ __forceinline__ f3 guide_albedo(const Luts& L, const Closure& c, f3 view) {
    const float nv = fmaxf(dot(c.n, view), .0001f);
    const float ed = c.metal < 1.f ? dielectric_albedo(L, nv, c.roughness, dielectric_ior(c)) : 0.f;
    f3 a = c.base_color * ((1.f - c.metal) * (1.f - ed)) + mk(ed, ed, ed) * (1.f - c.metal);
    if (c.metal > 0.f) {
        const float4 m = brdf_moments(L, nv, c.roughness);
        const f3 comp = metal_compensation(L, c, view);
        const f3 e = (c.base_color * m.x + mk(m.y, m.y, m.y)) * c.spec_weight;
        a = a + mk(e.x * (1.f + comp.x), e.y * (1.f + comp.y), e.z * (1.f + comp.z)) * c.metal;
    }
    return mk(clampf(a.x, 0.f, 1.f), clampf(a.y, 0.f, 1.f), clampf(a.z, 0.f, 1.f));
}
// crates/vork-render/shaders/bsdf.slang:197-212
__device__ __forceinline__ f3 valid_interface_normal(f3 normal, f3 geometric, f3 view) {
    float iz = dot(view, geometric);
    if (iz <= 0.f) return normal;
    if (dot(normal, geometric) < 0.f) normal = -normal;
    float threshold = fminf(.9f * iz, .01f);
    if (dot(normal, view) > 0.f && dot(reflect3(-view, normal), geometric) >= threshold) return normal;
    f3 tangent = normal - geometric * dot(normal, geometric);
    float ts = dot(tangent, tangent);
    if (ts <= 0.f) return geometric;
    f3 axis = tangent * rsqrt_a(ts);
    float ix = dot(view, axis), ls = ix * ix + iz * iz;
    float c2 = clampf(.5f * (1.f + div_a(iz * threshold - ix * sqrt_a(fmaxf(ls - threshold * threshold, 0.f)), ls)), 0.f, 1.f);
    return axis * sqrt_a(1.f - c2) + geometric * sqrt_a(c2);
}
// bsdf.slang:173-196
__device__ __forceinline__ void transport_probabilities(const Closure& c, f3 view, float& p0, float& p1) {
    float diffuse = fmaxf((c.base_color.x * kLuminance[0] + c.base_color.y * kLuminance[1] +
                           c.base_color.z * kLuminance[2]) * (1.f - c.metal), 0.f);
    float f = fresnel(fmaxf(dot(c.n, view), 0.f), dielectric_ior(c));
    float dielectric = f * (kLuminance[0] + kLuminance[1] + kLuminance[2]);
    float reflection = fmaxf(lerpf(dielectric, c.spec_weight, c.metal), 0.f);
    float total = diffuse + reflection;
    if (total > 0.f) {
        float t = rcp_a(fmaxf(total, 1e-20f));
        p0 = diffuse * t; p1 = reflection * t;
    } else {
        p0 = p1 = 0.f;
    }
}
// bsdf.slang:218-232
__device__ __forceinline__ Prep transport_prepare(const Luts& L, const Closure& c, f3 view, f3 smooth, f3 geometric) {
    Prep r;
    Closure base = c;
    base.n = valid_interface_normal(c.n, geometric, view);
    transport_probabilities(base, view, r.p0, r.p1);
    float deviation = fabsf(dot(smooth, c.n));
    r.bump_a2 = clampf(div_a(1.f - deviation * deviation, fmaxf(8.f * deviation * deviation, 1e-20f)), 0.f, 1.f);
    r.diffuse = direct_diffuse(L, base, view);
    r.metal_comp = metal_compensation(L, base, view);
    r.smooth = smooth;
    r.interface_n = base.n;
    return r;
}
// bsdf.slang:235-244
__device__ __forceinline__ float bump_shadow(const Closure& c, f3 direction, const Prep& pr) {
    float cosine = dot(pr.smooth, direction);
    if (cosine * dot(pr.smooth, c.n) * dot(c.n, direction) < 0.f) return 0.f;
    if (pr.bump_a2 == 0.f) return 1.f;
    float mu = fabsf(cosine);
    return mu > 0.f ? div_a(2.f * mu, mu + sqrt_a(pr.bump_a2 + (1.f - pr.bump_a2) * mu * mu)) : 0.f;
}
// crates/vork-render/shaders/bsdf_microfacet.slang:42-62
__device__ __forceinline__ float visible_scale(float roughness, f3 n, f3 view) {
    float nv = dot(n, view), a = specular_alpha(roughness);
    return 2.f * rcp_a(nv + sqrt_a(a * a + (1.f - a * a) * nv * nv));
}
__device__ __forceinline__ f3 visible_normal(float roughness, f3 n, f3 view, float u0, float u1) {
    f3 x = basis_direction(n, mk(1.f, 0.f, 0.f)), y = basis_direction(n, mk(0.f, 1.f, 0.f));
    float a = specular_alpha(roughness);
    f3 local = mk(dot(view, x), dot(view, y), dot(view, n));
    f3 s = ggx_visible_normal(local, a, a, u0, u1);
    return x * s.x + y * s.y + n * s.z;
}
// bsdf.slang:270-283
__device__ __forceinline__ float pdf_mask(const Closure& c, f3 view, f3 direction, uint32_t mask, const Prep& pr) {
    f3 n = pr.interface_n;
    float nl = dot(n, direction);
    float cd = dot(c.n, direction);
    float pdf = (mask & VH_RAY_DIFFUSE) ? pr.p0 * (cd > 0.f ? 1.f : 0.f) * fabsf(cd) * kInvPi : 0.f;
    if (nl <= 0.f) return pdf;
    f3 h = normalize_a(view + direction);
    float d = distribution(fmaxf(dot(n, h), 0.f), specular_alpha(c.roughness));
    return pdf + ((mask & VH_RAY_SPECULAR) ? pr.p1 * (d * visible_scale(c.roughness, n, view) / 4.f) : 0.f);
}
// bsdf_response.slang:111-124 (isotropic) and bsdf.slang:347-365
__device__ __forceinline__ f3 eval_mask(const Closure& c, f3 view, f3 direction, uint32_t mask, const Prep& pr) {
    f3 zero = mk(0.f, 0.f, 0.f);
    f3 spec = zero;
    if (mask & VH_RAY_SPECULAR) {
        f3 n = pr.interface_n;
        float nl = fmaxf(dot(n, direction), 0.f), nv = fmaxf(dot(n, view), .0001f);
        if (nl > 0.f) {
            f3 h = normalize_a(view + direction);
            float nh = fmaxf(dot(n, h), 0.f), vh = fmaxf(dot(view, h), 0.f);
            Closure base = c;
            base.n = n;
            f3 f = compensated_surface_reflection(base, vh, pr.metal_comp);
            float a = specular_alpha(c.roughness);
            spec = f * (distribution(nh, a) * g2(nv, nl, a) * rcp_a(4.f * nv * nl)) * nl;
        }
    }
    f3 diff = zero;
    if ((mask & VH_RAY_DIFFUSE) && dot(c.n, direction) > 0.f)
        diff = pr.diffuse * (bump_shadow(c, direction, pr) * dot(c.n, direction));
    return spec + diff;
}
struct BsdfSample {
    f3 direction, weight;
    float pdf;
    uint32_t event;
};
// bsdf.slang:446-510
__device__ __forceinline__ BsdfSample transport_sample(const Closure& c, f3 view, float r0, float r1, float r2,
                                                       const Prep& pr) {
    BsdfSample res;
    res.direction = mk(0.f, 0.f, 0.f);
    res.weight = mk(0.f, 0.f, 0.f);
    res.pdf = 0.f;
    res.event = 0;
    f3 n = pr.interface_n;
    if (r2 < pr.p0) {
        float radius = sqrt_a(r0);
        res.direction = basis_direction(c.n, mk(radius * cos_rev(r1), radius * sin_rev(r1), sqrt_a(1.f - r0)));
        res.event = VH_RAY_DIFFUSE;
    } else if (r2 < pr.p0 + pr.p1) {
        f3 h = visible_normal(c.roughness, n, view, r0, r1);
        res.direction = reflect3(-view, h);
        res.event = VH_RAY_SPECULAR;
    } else {
        return res;
    }
    if (res.event == VH_RAY_SPECULAR && dot(n, res.direction) <= 0.f) {
        res.direction = mk(0.f, 0.f, 0.f);
        res.event = 0;
        return res;
    }
    res.pdf = pdf_mask(c, view, res.direction, res.event, pr);
    if (res.pdf > 0.f) res.weight = eval_mask(c, view, res.direction, res.event, pr) * rcp_a(res.pdf);
    return res;
}
// path.slang:833-850
__device__ __forceinline__ float mis_power(float sampled, float other) {
    if (sampled <= 0.f) return 0.f;
    float ratio = div_a(other, sampled);
    return rcp_a(1.f + ratio * ratio);
}
__device__ __forceinline__ bool event_hemisphere(uint32_t event, f3 direction, f3 face, f3 smooth) {
    if (event == VH_RAY_TRANSMISSION) return dot(face, direction) < 0.f && dot(smooth, direction) < 0.f;
    return dot(smooth, direction) > 0.f;
}
__device__ __forceinline__ f3 analytic_mis_response(const Closure& c, f3 view, f3 direction, f3 geometric,
                                                    float light_pdf, const Prep& pr) {
    f3 value = mk(0.f, 0.f, 0.f);
    if (!event_hemisphere(VH_RAY_DIFFUSE, direction, geometric, pr.smooth)) return value;  // same test for both lobes
#pragma unroll
    for (int i = 0; i < 2; i++) {
        uint32_t e = i == 0 ? VH_RAY_DIFFUSE : VH_RAY_SPECULAR;
        float w = light_pdf > 0.f ? mis_power(light_pdf, pdf_mask(c, view, direction, e, pr)) : 1.f;
        value = value + eval_mask(c, view, direction, e, pr) * w;
    }
    return value;
}

// ------------------------------------------------------------ emitters
// crates/vork-render/shaders/emitter.slang:113-150 (rectangle, no window)
struct EmitEval {
    f3 radiance;
    float pdf, distance;
};
__device__ __forceinline__ EmitEval rect_eval(const vh_rect_light& l, f3 world, f3 direction) {
    EmitEval r;
    r.radiance = mk(0.f, 0.f, 0.f);
    r.pdf = 0.f;
    r.distance = 0.f;
    f3 center = load3(l.position) - world;
    if (dot(load3(l.normal), -center) <= 0.f) return r;
    f3 u = load3(l.u) * .5f, v = load3(l.v) * .5f;
    f3 normal = cross(u, v);
    float square = dot(normal, normal);
    if (square <= 0.f) return r;
    float denominator = dot(normal, direction);
    if (denominator == 0.f) return r;
    float distance = div_a(dot(normal, center), denominator);
    if (distance <= 0.f) return r;
    f3 q = direction * distance - center;
    float uu = dot(u, u), uv = dot(u, v), vv = dot(v, v), qu = dot(q, u), qv = dot(q, v);
    const float inv_square = rcp_a(square);
    float cx = (qu * vv - qv * uv) * inv_square, cy = (qv * uu - qu * uv) * inv_square;
    if (fabsf(cx) > 1.f || fabsf(cy) > 1.f) return r;
    const float root = sqrt_a(square);
    float area = 4.f * root;
    r.pdf = div_a(distance * distance, area * fabsf(denominator) * rcp_a(root));
    r.radiance = load3(l.radiance) * rcp_a(kPi * area);
    r.distance = distance;
    return r;
}
// emitter.slang:88-110 (area-uniform point) + 151-170
__device__ __forceinline__ EmitEval rect_sample(const vh_rect_light& l, f3 world, float u0, float u1, f3* direction) {
    f3 point = load3(l.position) + (load3(l.u) * (u0 - .5f) + load3(l.v) * (u1 - .5f));
    float distance = length_a(point - world);
    *direction = distance > 0.f ? (point - world) * rcp_a(distance) : mk(0.f, 0.f, 0.f);
    EmitEval e = rect_eval(l, world, *direction);
    e.distance = distance;
    return e;
}

// ---------------------------------------------------------- light tree
struct LightNode {
    float4 lower_power, upper_infinite;
    uint4 links, source;
};
// crates/vork-render/shaders/light_tree.slang:18-36
__device__ __forceinline__ float tree_log_importance(const LightNode& n, f3 point) {
    if (!(n.lower_power.w > 0.f)) return -FLT_MAX;
    float power = log2_a(n.lower_power.w);
    if (n.upper_infinite.w != 0.f) return power;
    f3 lo = mk(n.lower_power.x, n.lower_power.y, n.lower_power.z);
    f3 hi = mk(n.upper_infinite.x, n.upper_infinite.y, n.upper_infinite.z);
    f3 center = lo * 0.5f + hi * 0.5f;
    f3 radius = hi * 0.5f - lo * 0.5f;
    f3 delta = point - center;
    float distance = fmaxf(kLightTreeMinDistanceSquared, fmaxf(dot(delta, delta), dot(radius, radius)));
    if (!isfinite(distance)) return -FLT_MAX;
    return power - log2_a(distance);
}
__device__ __forceinline__ float tree_branch(const LightNode& a, const LightNode& b, f3 point) {
    float x = tree_log_importance(a, point), y = tree_log_importance(b, point);
    float scale = fmaxf(x, y);
    float uniform = (float)a.links.w / ((float)a.links.w + (float)b.links.w);
    if (scale == -FLT_MAX) return uniform;
    x = exp2_a(x - scale);
    y = exp2_a(y - scale);
    return div_a(x, x + y);
}
// light_tree.slang:58-84
__device__ __forceinline__ uint32_t light_pick(const LightNode* nodes, uint32_t node_count, f3 point, float ux, float uy,
                                               float* probability) {
    *probability = 0.f;
    if (node_count == 0) return 0xffffffffu;
    uint32_t node = 0;
    float prob = 1.f;
    const float one = 1.f - FLT_EPSILON * 0.5f;
    float choice = clampf(ux, 0.f, one);
    bool uniform = uy < kLightTreeUniformMix;
    uint32_t total = nodes[0].links.w;
    uint32_t rank = min((uint32_t)(choice * (float)total), total - 1);
    while (nodes[node].links.x != 0xffffffffu) {
        uint32_t cl = nodes[node].links.x, cr = nodes[node].links.y;
        float p = tree_branch(nodes[cl], nodes[cr], point);
        bool left;
        if (uniform) {
            uint32_t cnt = nodes[cl].links.w;
            left = rank < cnt;
            if (!left) rank -= cnt;
        } else {
            left = choice < p;
            choice = left ? div_a(choice, p) : div_a(choice - p, 1.f - p);
            choice = fminf(choice, one);
        }
        prob *= left ? p : 1.f - p;
        node = left ? cl : cr;
    }
    *probability = kLightTreeUniformMix / (float)total + (1.f - kLightTreeUniformMix) * prob;
    return nodes[node].source.y;
}

// ------------------------------------------------------ surface offsets
// crates/vork-render/shaders/vork_prelude.slang:155-181
__device__ __forceinline__ float next_float(float value, bool up) {
    if (value == 0.f) return u2f(up ? 1u : 0x80000001u);
    bool increasing = (value > 0.f) == up;
    return u2f(f2u(value) + (increasing ? 1u : 0xffffffffu));
}
__device__ __forceinline__ f3 surface_ray_origin(f3 world, f3 error, f3 n, f3 direction) {
    float scale = fmaxf(fabsf(n.x), fmaxf(fabsf(n.y), fabsf(n.z)));
    if (scale == 0.f) return world;
    n = n / scale;
    float amount = next_float(dot(vabs(n), error) / dot(n, n), true);
    f3 offset = n * (amount * (dot(n, direction) >= 0.f ? 1.f : -1.f));
    f3 o = world + offset;
    if (offset.x != 0.f) o.x = next_float(o.x, offset.x > 0.f);
    if (offset.y != 0.f) o.y = next_float(o.y, offset.y > 0.f);
    if (offset.z != 0.f) o.z = next_float(o.z, offset.z > 0.f);
    return o;
}
__device__ __forceinline__ void surface_ray_connection(f3 origin, f3 world, f3 direction, float distance, f3* dir_out,
                                                       float* dist_out) {
    if (distance == FLT_MAX || (origin.x == world.x && origin.y == world.y && origin.z == world.z)) {
        *dir_out = direction;
        *dist_out = distance;
        return;
    }
    f3 delta = direction * distance - (origin - world);
    float remaining = length_a(delta);
    *dir_out = remaining > 0.f ? delta * rcp_a(remaining) : direction;
    *dist_out = remaining;
}
// stream.slang:416-427
__device__ __forceinline__ f3 finite_unit_normal(f3 normal, f3 fallback) {
    float scale = fmaxf(fabsf(normal.x), fmaxf(fabsf(normal.y), fabsf(normal.z)));
    if (finite3(normal) && scale > 0.f) return normalize_a(normal * rcp_a(scale));
    float fs = fmaxf(fabsf(fallback.x), fmaxf(fabsf(fallback.y), fabsf(fallback.z)));
    if (finite3(fallback) && fs > 0.f) return normalize_a(fallback * rcp_a(fs));
    return mk(0.f, 0.f, 1.f);
}
// sample_clamp.slang:7-16
__device__ __forceinline__ f3 clamp_sample(f3 value, float limit, uint32_t& dropped) {
    float brightness = dot(vabs(value), mk(1.f / 3.f, 1.f / 3.f, 1.f / 3.f));
    if (!isfinite(brightness)) {
        dropped++;
        value = mk(0.f, 0.f, 0.f);
        brightness = 0.f;
    }
    return value * ((limit > 0.f && brightness > limit) ? div_a(limit, fmaxf(brightness, 1e-20f)) : 1.f);
}
// Escaping rays: constant or procedural sky.
__device__ __forceinline__ f3 sky_radiance(const vh_sky& k, f3 d) {
    if (k.kind == VH_SKY_CONSTANT) return load3(k.color);
    float y = d.y;
    f3 base = y >= 0.f ? lerp3(load3(k.color), load3(k.zenith), sqrt_a(y))
                       : lerp3(load3(k.color), load3(k.ground), fminf(1.f, -4.f * y));
    float cs = fmaxf(dot(d, load3(k.sun_direction)), 0.f);
    // cs^e as 2^(e log2 cs); 0^0 is 1 as in powf
    const float sun = k.sun_exponent == 0.f ? 1.f : exp2_a(k.sun_exponent * log2_a(cs));
    return base + load3(k.sun_radiance) * sun;
}

}  // namespace vh
