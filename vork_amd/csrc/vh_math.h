// Small vector math shared by the device kernels and the host side of the
// library.  Everything that must round exactly like the CPU oracle (ray
// transforms, the triangle test, the box test) lives in functions compiled
// with FP contraction off; see VH_EXACT.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#define VH_HD __host__ __device__ __forceinline__
// Expressions that must round as written (no FMA contraction): the oracle is
// compiled with -ffp-contract=off and this keeps the two bit-identical.
#define VH_EXACT _Pragma("clang fp contract(off)")
// Address-space qualifiers for device pointers whose space the compiler
// cannot infer (pointers loaded from memory, per-lane stacks that may live in
// LDS or global memory): they make it emit global_* / ds_* instead of flat_*
// accesses.
#define VH_GLOBAL __attribute__((address_space(1)))
#define VH_LDS __attribute__((address_space(3)))

namespace vh {

struct f3 { float x, y, z; };

VH_HD f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
VH_HD f3 operator+(f3 a, f3 b) { VH_EXACT return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
VH_HD f3 operator-(f3 a, f3 b) { VH_EXACT return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
VH_HD f3 operator*(f3 a, f3 b) { VH_EXACT return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
VH_HD f3 operator*(f3 a, float s) { VH_EXACT return mk(a.x * s, a.y * s, a.z * s); }
VH_HD f3 operator/(f3 a, float s) { VH_EXACT return mk(a.x / s, a.y / s, a.z / s); }
VH_HD f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
VH_HD float dot(f3 a, f3 b) { VH_EXACT return a.x * b.x + a.y * b.y + a.z * b.z; }
VH_HD f3 cross(f3 a, f3 b) {
    VH_EXACT return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
VH_HD float length(f3 a) { return sqrtf(dot(a, a)); }
VH_HD f3 normalize(f3 a) { return a / length(a); }
VH_HD f3 vabs(f3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
VH_HD float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
VH_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
VH_HD bool finite3(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
VH_HD bool nonzero(f3 a) { return a.x != 0.f || a.y != 0.f || a.z != 0.f; }
VH_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
// HLSL / SPIR-V lerp: a*(1-s)+b*s
VH_HD float lerpf(float a, float b, float s) { VH_EXACT return a * (1.f - s) + b * s; }
VH_HD f3 lerp3(f3 a, f3 b, float s) { return a * (1.f - s) + b * s; }
VH_HD f3 reflect3(f3 i, f3 n) { return i - n * (2.f * dot(n, i)); }
VH_HD f3 load3(const float* p) { return mk(p[0], p[1], p[2]); }

// Approximate arithmetic for shading, which is checked against the oracle
// image by image (relMSE far below the noise level), not bit by bit: the
// hardware reciprocal, square root, reciprocal square root, base-2
// exponential and logarithm, and sine/cosine of revolutions (about 1 ulp),
// in place of the correctly rounded sequences of 10-150 instructions each.
// Ray generation, traversal, surface offsets and the film keep the exact
// operators.  The host pass keeps the exact functions.
VH_HD float rcp_a(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.f / x;
#endif
}
VH_HD float sqrt_a(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
VH_HD float rsqrt_a(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.f / sqrtf(x);
#endif
}
VH_HD float exp2_a(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_exp2f(x);
#else
    return exp2f(x);
#endif
}
VH_HD float log2_a(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_logf(x);
#else
    return log2f(x);
#endif
}
// sin(2 pi r) and cos(2 pi r) for |r| below 256 revolutions.
VH_HD float sin_rev(float r) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sinf(r);
#else
    return sinf(6.28318530717958647692f * r);
#endif
}
VH_HD float cos_rev(float r) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_cosf(r);
#else
    return cosf(6.28318530717958647692f * r);
#endif
}
VH_HD float div_a(float a, float b) { return a * rcp_a(b); }
VH_HD f3 div_a(f3 a, float b) { return a * rcp_a(b); }
VH_HD float length_a(f3 a) { return sqrt_a(dot(a, a)); }
VH_HD f3 normalize_a(f3 a) { return a * rsqrt_a(dot(a, a)); }

// Plain 4 x u32 vector (HIP's uint4 is a class whose members only take
// generic-address-space objects).
typedef uint32_t vu4 __attribute__((ext_vector_type(4)));

template <class T>
__device__ __forceinline__ const VH_GLOBAL T* as_global(const T* p) {
    return (const VH_GLOBAL T*)p;
}

VH_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
VH_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// The next float above x (x finite); nextafterf(x, +inf) without libm.
VH_HD float next_up(float x) {
    if (x == 0.f) return u2f(1u);
    uint32_t u = f2u(x);
    return u2f(x > 0.f ? u + 1u : u - 1u);
}

// Object-to-world style 3x4 rows applied to a point / direction, left to
// right in float without contraction (matches the oracle's xform_point).
VH_HD f3 xform_point(const float* m, f3 p) {
    VH_EXACT return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
                       m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
VH_HD f3 xform_dir(const float* m, f3 d) {
    VH_EXACT return mk(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
                       m[8] * d.x + m[9] * d.y + m[10] * d.z);
}

// World-to-object rows of object-to-world rows, in double precision rounded
// to float, exactly as the oracle's vo_invert_transform evaluates them.
// Returns false for a singular or non-finite transform.
VH_HD bool invert_transform(const float* m, float* inv) {
    VH_EXACT
    for (int i = 0; i < 12; i++)
        if (!isfinite(m[i])) return false;
    double a00 = m[0], a01 = m[1], a02 = m[2], a10 = m[4], a11 = m[5], a12 = m[6], a20 = m[8], a21 = m[9], a22 = m[10];
    double t0 = m[3], t1 = m[7], t2 = m[11];
    double c00 = a11 * a22 - a12 * a21;
    double c01 = a12 * a20 - a10 * a22;
    double c02 = a10 * a21 - a11 * a20;
    double det = a00 * c00 + a01 * c01 + a02 * c02;
    if (det == 0.0 || !isfinite(det)) return false;
    double r[9];
    r[0] = c00 / det;
    r[1] = (a02 * a21 - a01 * a22) / det;
    r[2] = (a01 * a12 - a02 * a11) / det;
    r[3] = c01 / det;
    r[4] = (a00 * a22 - a02 * a20) / det;
    r[5] = (a02 * a10 - a00 * a12) / det;
    r[6] = c02 / det;
    r[7] = (a01 * a20 - a00 * a21) / det;
    r[8] = (a00 * a11 - a01 * a10) / det;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) inv[4 * i + j] = (float)r[3 * i + j];
        inv[4 * i + 3] = (float)(-(r[3 * i] * t0 + r[3 * i + 1] * t1 + r[3 * i + 2] * t2));
    }
    for (int i = 0; i < 12; i++)
        if (!isfinite(inv[i])) return false;
    return true;
}

// gamma(n) of f32 (PBRT), used by the conservative box test.
constexpr float kGamma3 = (3.f * (FLT_EPSILON * 0.5f)) / (1.f - 3.f * (FLT_EPSILON * 0.5f));
constexpr float kGamma7 = (7.f * (FLT_EPSILON * 0.5f)) / (1.f - 7.f * (FLT_EPSILON * 0.5f));
constexpr float kPi = 3.14159265358979323846f;
constexpr float kTau = 6.28318530717958647692f;
constexpr float kInvPi = 0.318309886183790671538f;

}  // namespace vh
