/*
 * vork_oracle.c - CPU oracle of the path-tracing hot path (test infrastructure).
 *
 * A plain-C restatement of the reference estimator for the closure subset the
 * workloads use: opaque triangles, the OpenPBR base lobe (Lambertian diffuse
 * under a GGX dielectric or a GGX F82 metal, no coat/fuzz/transmission),
 * rectangle area lights with power-heuristic MIS and a constant or procedural
 * sky.  It is the checker of the HIP library and never part of it.
 *
 * Built with -ffp-contract=off so that float expressions round exactly as
 * written; the library's intersection and transform code is compiled the same
 * way, which makes hit IDs and t comparable bit for bit.
 */
#include "vork_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_F 3.14159265358979323846f
#define TAU_F 6.28318530717958647692f
/* crates/vork-slang/src/registry.rs:389-391 */
#define SAMPLE_HASH_MIX 0x45d9f3bu
#define PIXEL_HASH_STREAM_STEP 0x9e3779b9u
/* crates/vork-slang/src/registry.rs:89 (gamma(7) of f32) */
#define FP32_GAMMA7 ((7.f * (FLT_EPSILON * 0.5f)) / (1.f - 7.f * (FLT_EPSILON * 0.5f)))
/* crates/vork-slang/src/renderer_registry.rs:98-100,109-113,127 */
#define ROULETTE_START 3u
#define ROULETTE_MIN 0.05f
#define ROULETTE_MAX 0.95f
#define LIGHT_TREE_UNIFORM_MIX 0.25f
#define LIGHT_TREE_MIN_DISTANCE_SQUARED 1e-12f
#define POSITIVE_EPSILON 1e-20f
/* crates/vork-slang/src/registry.rs:375-385 */
#define BRDF_SIZE 128u
#define BRDF_ALBEDO_SIZE 64u
#define BRDF_IOR_STEPS 64u
#define GGX_ALPHA_MIN 0.0001f
#define UINT_SCALE 0.000000059604644775390625f
#define RADICAL_INVERSE_SCALE 2.3283064365386963e-10f
#define ENVIRONMENT_SAMPLES 512u
#define DIELECTRIC_ALBEDO_SAMPLES 512u

static const float LUMINANCE[3] = {0.2126f, 0.7152f, 0.0722f};

typedef struct { float x, y, z; } v3;
static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld(const float* p) { return mk(p[0], p[1], p[2]); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 dvs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float len(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 nrm(v3 a) { return dvs(a, len(a)); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline v3 vabs(v3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
static inline float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline float maxc(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
static inline int finite3(v3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
/* HLSL/SPIR-V lerp: x*(1-s)+y*s */
static inline float lerpf(float a, float b, float s) { return a * (1.f - s) + b * s; }
static inline v3 lerp3(v3 a, v3 b, float s) { return add(scl(a, 1.f - s), scl(b, s)); }
static inline v3 reflect3(v3 i, v3 n) { return sub(i, scl(n, 2.f * dot(n, i))); }
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static inline uint32_t reverse_bits(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}

/* ======================================================================
 * Sampler: crates/vork-render/src/path_sampler.rs:1-80 (CPU reference),
 * crates/vork-render/shaders/path_sampler.slang:46-115,
 * crates/vork-slang/src/vork_sampling.slang:3-15.
 * ==================================================================== */

/* vork_sampling.slang:3-7, crates/vork-render/src/sampling.rs:113-118 */
uint32_t vo_pixel_hash(uint32_t seed) {
    uint32_t value = (seed ^ (seed >> 16)) * SAMPLE_HASH_MIX;
    value = (value ^ (value >> 16)) * SAMPLE_HASH_MIX;
    return value ^ (value >> 16);
}
void vo_pixel_hash_many(const uint32_t* in, uint32_t n, uint32_t* out) {
    for (uint32_t i = 0; i < n; i++) out[i] = vo_pixel_hash(in[i]);
}
/* vork_sampling.slang:8-15 */
uint32_t vo_screen_ray_hash(uint32_t px, uint32_t py, uint32_t index, uint32_t stream) {
    uint32_t spatial = vo_pixel_hash((px * SAMPLE_HASH_MIX) ^ py);
    return vo_pixel_hash(spatial ^ vo_pixel_hash(index * PIXEL_HASH_STREAM_STEP + stream));
}

/* path_sampler.rs:8-34: Joe-Kuo direction numbers of dimensions 2-4 from
 * their primitive polynomials (degree s, coefficients a, m_1..m_s). */
static uint32_t DIRECTIONS[3][32];
static int directions_ready = 0;
static void build_directions(void) {
    static const uint32_t S[3] = {1, 2, 3}, A[3] = {0, 1, 1};
    static const uint32_t M[3][3] = {{1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
    for (int d = 0; d < 3; d++) {
        uint32_t s = S[d], a = A[d];
        for (uint32_t k = 0; k < 32; k++) {
            if (k < s) {
                DIRECTIONS[d][k] = M[d][k] << (31 - k);
            } else {
                uint32_t value = DIRECTIONS[d][k - s] ^ (DIRECTIONS[d][k - s] >> s);
                for (uint32_t j = 1; j < s; j++)
                    if ((a >> (s - 1 - j)) & 1u) value ^= DIRECTIONS[d][k - j];
                DIRECTIONS[d][k] = value;
            }
        }
    }
    directions_ready = 1;
}
void vo_sobol_directions(uint32_t out[96]) {
    if (!directions_ready) build_directions();
    memcpy(out, DIRECTIONS, sizeof(DIRECTIONS));
}
/* path_sampler.rs:71-80: unscrambled Sobol fractions of dimensions 0-3 */
void vo_sobol(uint32_t index, uint32_t out[4]) {
    if (!directions_ready) build_directions();
    out[0] = reverse_bits(index);
    out[1] = out[2] = out[3] = 0;
    for (uint32_t k = 0; k < 32; k++)
        if ((index >> k) & 1u)
            for (int d = 0; d < 3; d++) out[d + 1] ^= DIRECTIONS[d][k];
}
/* path_sampler.rs:41-49, path_sampler.slang:65-68 */
static uint32_t owen_reversed(uint32_t x, uint32_t seed) {
    seed = vo_pixel_hash(seed);
    x ^= x * 0x3d20adeau;
    x += seed;
    x *= (seed >> 16) | 1u;
    x ^= x * 0x05526c56u;
    x ^= x * 0x53a22864u;
    return x;
}
/* path_sampler.slang:70 */
static inline float sample_unit(uint32_t x) { return ((float)(x >> 9) + 0.5f) * (1.f / 8388608.f); }

typedef struct { uint32_t index, seed, mask; } sampler_t;
/* path_sampler.slang:54-63 (sample_block = 2<<firstbithigh(count-1)) */
static sampler_t make_sampler(uint32_t px, uint32_t py, uint32_t index, uint32_t count, uint32_t stream) {
    sampler_t s;
    s.index = index;
    s.seed = vo_screen_ray_hash(px, py, 0, stream);
    if (count <= 1) s.mask = 0;
    else s.mask = (2u << (31 - __builtin_clz(count - 1))) - 1u;
    return s;
}
/* path_sampler.rs:55-64, path_sampler.slang:74-93 */
static void group_bits(sampler_t s, uint32_t group, uint32_t out[4]) {
    static const uint32_t SEEDS[4] = {0xa511e9b3u, 0x63d83595u, 0x2f7c4a1bu, 0x9e3779b9u};
    uint32_t seed = vo_pixel_hash(s.seed ^ vo_pixel_hash(group + 1u));
    uint32_t index = reverse_bits(owen_reversed(reverse_bits(s.index), seed ^ 0x6c8e9cf5u)) & s.mask;
    uint32_t x[4];
    vo_sobol(index, x);
    for (int c = 0; c < 4; c++) x[c] = reverse_bits(x[c]);
    x[0] = index;
    for (int c = 0; c < 4; c++) out[c] = reverse_bits(owen_reversed(x[c], seed ^ SEEDS[c]));
}
static void group_values(sampler_t s, uint32_t group, float out[4]) {
    uint32_t b[4];
    group_bits(s, group, b);
    for (int c = 0; c < 4; c++) out[c] = sample_unit(b[c]);
}
void vo_sample_group_bits(uint32_t px, uint32_t py, uint32_t index, uint32_t count, uint32_t stream,
                          uint32_t group, uint32_t out[4]) {
    group_bits(make_sampler(px, py, index, count, stream), group, out);
}
void vo_sample_group(uint32_t px, uint32_t py, uint32_t index, uint32_t count, uint32_t stream, uint32_t group,
                     float out[4]) {
    group_values(make_sampler(px, py, index, count, stream), group, out);
}
void vo_sample_group_bits_batch(uint32_t px, uint32_t py, uint32_t count, uint32_t stream, uint32_t group,
                                const uint32_t* indices, uint32_t n, uint32_t* out) {
    sampler_t s = make_sampler(px, py, 0, count, stream);
    for (uint32_t i = 0; i < n; i++) {
        s.index = indices ? indices[i] : i;
        group_bits(s, group, out + 4 * i);
    }
}
void vo_sample_group_bits_many(const uint32_t* px, const uint32_t* py, const uint32_t* index, uint32_t count,
                               uint32_t stream, const uint32_t* group, uint32_t n, uint32_t* out) {
    for (uint32_t i = 0; i < n; i++)
        group_bits(make_sampler(px[i], py[i], index[i], count, stream), group[i], out + 4 * i);
}
/* path_sampler.slang:72 */
float vo_path_random(uint32_t* state) {
    *state = vo_pixel_hash(*state + PIXEL_HASH_STREAM_STEP);
    return sample_unit(*state);
}

/* ======================================================================
 * Film filter: crates/vork-render/src/sampling.rs:149-175 (table),
 * path.slang:1712-1719 (lookup).
 * ==================================================================== */
double vo_filter_cdf(double x) {
    double t = M_PI * x / 1.5;
    return 0.5 + (0.35875 * x + 1.5 / M_PI * (0.48829 * sin(t) + 0.14128 * sin(2. * t) / 2. + 0.01168 * sin(3. * t) / 3.)) /
                     (3. * 0.35875);
}
void vo_filter_table(float table[VO_FILTER_INTERVALS + 1]) {
    const int n = VO_FILTER_INTERVALS;
    for (int i = 0; i <= n; i++) {
        if (i == 0) { table[i] = -1.5f; continue; }
        if (i == n) { table[i] = 1.5f; continue; }
        double u = (double)i / (double)n, lo = -1.5, hi = 1.5;
        for (int k = 0; k < 48; k++) {
            double mid = (lo + hi) * 0.5;
            if (vo_filter_cdf(mid) < u) lo = mid; else hi = mid;
        }
        table[i] = (float)((lo + hi) * 0.5);
    }
}
float vo_filter_offset(const float* table, float u) {
    float t = clampf(u, 0.f, 1.f) * (float)VO_FILTER_INTERVALS;
    uint32_t i = (uint32_t)t;
    if (i > VO_FILTER_INTERVALS - 1) i = VO_FILTER_INTERVALS - 1;
    return lerpf(table[i], table[i + 1], t - (float)i);
}

/* ======================================================================
 * Intersection: watertight crossing of
 * crates/vork-render/shaders/vork_medium_domain.slang:40-66 (Woop, Benthin
 * and Wald 2013 with a symbolic tie at exact zeros), t and determinant of
 * vork_medium_domain.slang:110-118, clamped barycentrics of
 * crates/vork-render/shaders/camera_ray.slang:24-35.
 * ==================================================================== */
/* vork_medium_domain.slang:44-52 */
static float volume_edge(float px, float py, float qx, float qy) {
    int forward = px < qx || (px == qx && py < qy);
    float lox = forward ? px : qx, loy = forward ? py : qy;
    float hix = forward ? qx : px, hiy = forward ? qy : py;
    float area = lox * hiy - loy * hix;
    float tie = px != qx ? 1.f : (py != qy ? -1.f : 0.f);
    float canonical = area != 0.f ? area : tie;
    return forward ? canonical : -canonical;
}
static int crosses(v3 a, v3 b, v3 c, v3 o, v3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int z = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int x = (z + 1) % 3, y = (x + 1) % 3;
    if (comp(d, z) < 0.f) { int s = x; x = y; y = s; }
    float sx = comp(d, x) / comp(d, z), sy = comp(d, y) / comp(d, z);
    v3 vs[3] = {sub(a, o), sub(b, o), sub(c, o)};
    float px[3], py[3];
    for (int i = 0; i < 3; i++) {
        px[i] = comp(vs[i], x) - sx * comp(vs[i], z);
        py[i] = comp(vs[i], y) - sy * comp(vs[i], z);
    }
    float u = volume_edge(px[1], py[1], px[2], py[2]);
    float v = volume_edge(px[2], py[2], px[0], py[0]);
    float w = volume_edge(px[0], py[0], px[1], py[1]);
    return (u > 0.f && v > 0.f && w > 0.f) || (u < 0.f && v < 0.f && w < 0.f);
}
int vo_volume_crosses(const float a[3], const float b[3], const float c[3], const float o[3], const float d[3]) {
    return crosses(ld(a), ld(b), ld(c), ld(o), ld(d));
}
/* Returns 1 on a hit with tmin < t < tmax.  The determinant test follows
 * vork_medium_domain.slang:112-113 (|det| > REG_POSITIVE_EPSILON); a back
 * face (det < 0: the ray and the winding normal point the same way) is
 * rejected only for culled instances. */
static int intersect(v3 a, v3 b, v3 c, v3 o, v3 d, float tmin, float tmax, int cull, float* t_out, float* u_out,
                     float* v_out) {
    v3 e1 = sub(b, a), e2 = sub(c, a);
    v3 q = cross(d, e2);
    float det = dot(e1, q);
    if (!(fabsf(det) > POSITIVE_EPSILON)) return 0;
    if (cull && det < 0.f) return 0;
    if (!crosses(a, b, c, o, d)) return 0;
    v3 s = sub(o, a);
    v3 r = cross(s, e1);
    float t = dot(e2, r) / det;
    if (!(t > tmin && t < tmax)) return 0;
    float u = fmaxf(dot(s, q) / det, 0.f), v = fmaxf(dot(d, r) / det, 0.f);
    if (u + v > 1.f) { float k = u + v; u = u / k; v = v / k; }
    *t_out = t; *u_out = u; *v_out = v;
    return 1;
}
int vo_intersect(const float a[3], const float b[3], const float c[3], const float o[3], const float d[3], float tmin,
                 float tmax, int cull_back, float* t, float* u, float* v) {
    return intersect(ld(a), ld(b), ld(c), ld(o), ld(d), tmin, tmax, cull_back, t, u, v);
}

/* ======================================================================
 * Scene: meshes (BLAS), instances (TLAS), and a median-split binary BVH of
 * the form crates/vork-render/src/medium_bounded.rs:251-287 builds.
 * ==================================================================== */
typedef struct { float lo[3], hi[3]; uint32_t left, right, count; } bnode; /* count>0: leaf of order[left..] */
typedef struct { bnode* nodes; uint32_t node_count; uint32_t* order; } bvh_t;

typedef struct {
    float* pos; float* nrm; uint32_t* idx; uint32_t* mat;
    uint32_t vcount, tcount;
    float lo[3], hi[3];
    bvh_t bvh;
} mesh_t;

struct vo_scene {
    mesh_t* meshes; uint32_t mesh_count;
    vh_instance* inst; uint32_t inst_count;
    float* w2o;      /* 12 per instance */
    float* n2w;      /* normal matrix rows (inverse transpose of O2W 3x3), 9 per instance */
    uint8_t* valid;
    bvh_t tlas;
    vh_material* materials; uint32_t material_count;
    vh_rect_light* lights; uint32_t light_count;
    /* light tree (light_tree.slang:3-8) */
    struct { float lower_power[4], upper_infinite[4]; uint32_t links[4], source[4]; } * tree;
    uint32_t tree_nodes;
    vh_camera camera;
    vh_sky sky;
    float filter[VO_FILTER_INTERVALS + 1];
    uint64_t stats[4];
};

typedef struct { float lo[3], hi[3], c[3]; } prim_box;

/* Centroid order on one axis, ties by primitive id. */
static inline int before(const prim_box* boxes, int axis, uint32_t a, uint32_t b) {
    float ca = boxes[a].c[axis], cb = boxes[b].c[axis];
    return ca < cb || (ca == cb && a < b);
}
/* Places the `k`-th smallest id at ids[k] with smaller ones before it (quickselect). */
static void select_nth(const prim_box* boxes, int axis, uint32_t* ids, uint32_t n, uint32_t k) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        /* median of three as pivot */
        uint32_t a = ids[lo], b = ids[mid], c = ids[hi], pv;
        if (before(boxes, axis, a, b)) pv = before(boxes, axis, b, c) ? b : (before(boxes, axis, a, c) ? c : a);
        else pv = before(boxes, axis, a, c) ? a : (before(boxes, axis, b, c) ? c : b);
        uint32_t i = lo, j = hi;
        while (i <= j) {
            while (before(boxes, axis, ids[i], pv)) i++;
            while (before(boxes, axis, pv, ids[j])) j--;
            if (i <= j) { uint32_t t = ids[i]; ids[i] = ids[j]; ids[j] = t; i++; if (j == 0) break; j--; }
        }
        if (k <= j) hi = j;
        else if (k >= i) lo = i;
        else return;
    }
}
static uint32_t bvh_build_range(bvh_t* b, prim_box* boxes, uint32_t* ids, uint32_t n, uint32_t* cursor) {
    uint32_t at = (*cursor)++;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; i++) {
        prim_box* p = &boxes[ids[i]];
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], p->lo[k]); hi[k] = fmaxf(hi[k], p->hi[k]);
            clo[k] = fminf(clo[k], p->c[k]); chi[k] = fmaxf(chi[k], p->c[k]);
        }
    }
    for (int k = 0; k < 3; k++) { b->nodes[at].lo[k] = lo[k]; b->nodes[at].hi[k] = hi[k]; }
    if (n <= 4) {
        b->nodes[at].left = (uint32_t)(ids - b->order);
        b->nodes[at].right = 0;
        b->nodes[at].count = n;
        return at;
    }
    /* medium_bounded.rs:272-282: split at the median centroid of the widest centroid axis */
    float ex = chi[0] - clo[0], ey = chi[1] - clo[1], ez = chi[2] - clo[2];
    int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
    uint32_t mid = n / 2;
    select_nth(boxes, axis, ids, n, mid);
    uint32_t l = bvh_build_range(b, boxes, ids, mid, cursor);
    uint32_t r = bvh_build_range(b, boxes, ids + mid, n - mid, cursor);
    b->nodes[at].left = l;
    b->nodes[at].right = r;
    b->nodes[at].count = 0;
    return at;
}
static void bvh_build(bvh_t* b, prim_box* boxes, uint32_t n) {
    b->nodes = (bnode*)calloc(n ? 2 * n : 1, sizeof(bnode));
    b->order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) b->order[i] = i;
    uint32_t cursor = 0;
    if (n) bvh_build_range(b, boxes, b->order, n, &cursor);
    b->node_count = cursor;
}

/* Ray/box slab test, conservative by (1 + 2 gamma(3)) on the far distance. */
static inline int box_hit(const float lo[3], const float hi[3], v3 o, v3 inv, float tmin, float tmax) {
    float t0x = (lo[0] - o.x) * inv.x, t1x = (hi[0] - o.x) * inv.x;
    float t0y = (lo[1] - o.y) * inv.y, t1y = (hi[1] - o.y) * inv.y;
    float t0z = (lo[2] - o.z) * inv.z, t1z = (hi[2] - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
    const float g = 1.f + 2.f * (3.f * FLT_EPSILON * 0.5f) / (1.f - 3.f * FLT_EPSILON * 0.5f);
    return tn <= tf * g;
}

/* World-to-object rows of object-to-world rows, in double precision. */
int vo_invert_transform(const float m[12], float inv[12]) {
    double a[3][3], t[3];
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) a[r][c] = m[4 * r + c];
        t[r] = m[4 * r + 3];
    }
    for (int i = 0; i < 12; i++)
        if (!isfinite(m[i])) return 0;
    double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    if (det == 0.0 || !isfinite(det)) return 0;
    double r[3][3];
    r[0][0] = c00 / det;
    r[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    r[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    r[1][0] = c01 / det;
    r[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    r[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    r[2][0] = c02 / det;
    r[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    r[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) inv[4 * i + j] = (float)r[i][j];
        inv[4 * i + 3] = (float)(-(r[i][0] * t[0] + r[i][1] * t[1] + r[i][2] * t[2]));
    }
    for (int i = 0; i < 12; i++)
        if (!isfinite(inv[i])) return 0;
    return 1;
}
/* o' = W2O (o, 1), d' = W2O (d, 0), evaluated left to right in float. */
static inline v3 xform_point(const float* m, v3 p) {
    return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
static inline v3 xform_dir(const float* m, v3 d) {
    return mk(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
              m[8] * d.x + m[9] * d.y + m[10] * d.z);
}
static inline v3 safe_inv(v3 d) { return mk(1.f / d.x, 1.f / d.y, 1.f / d.z); }

typedef struct { float t, u, v; uint32_t inst, prim; int found; } hitrec;

/* (t, instance, primitive) lexicographic: the tie rule at equal t. */
static inline int better(float t, uint32_t inst, uint32_t prim, const hitrec* h) {
    if (!h->found) return 1;
    if (t < h->t) return 1;
    if (t > h->t) return 0;
    return inst < h->inst || (inst == h->inst && prim < h->prim);
}

static inline void tri_verts(const mesh_t* m, uint32_t prim, v3* a, v3* b, v3* c) {
    const uint32_t* ix = m->idx + 3 * prim;
    *a = ld(m->pos + 3 * ix[0]); *b = ld(m->pos + 3 * ix[1]); *c = ld(m->pos + 3 * ix[2]);
}

/* One BLAS, closest or any; returns 1 when `any` found a hit. */
static int blas_trace(const vo_scene* s, uint32_t inst, v3 o, v3 d, float tmin, float tmax_world, hitrec* h, int any,
                      int brute) {
    const vh_instance* in = &s->inst[inst];
    const mesh_t* m = &s->meshes[in->blas];
    const float* w = s->w2o + 12 * inst;
    v3 lo = xform_point(w, o), ld_ = xform_dir(w, d);
    int cull = (in->flags & VH_INSTANCE_CULL_BACK_FACES) != 0;
    if (brute) {
        for (uint32_t p = 0; p < m->tcount; p++) {
            v3 a, b, c; float t, u, v;
            tri_verts(m, p, &a, &b, &c);
            float tmax = h->found ? h->t : tmax_world;
            /* include equal t so that the tie rule can apply */
            if (intersect(a, b, c, lo, ld_, tmin, h->found ? nextafterf(tmax, INFINITY) : tmax, cull, &t, &u, &v)) {
                if (!(t < tmax_world)) continue;
                if (any) return 1;
                if (better(t, inst, p, h)) { h->t = t; h->u = u; h->v = v; h->inst = inst; h->prim = p; h->found = 1; }
            }
        }
        return 0;
    }
    if (!m->bvh.node_count) return 0;
    v3 inv = safe_inv(ld_);
    uint32_t stack[128]; int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        uint32_t ni = stack[--sp];
        const bnode* n = &m->bvh.nodes[ni];
        float tmax = h->found ? h->t : tmax_world;
        if (!box_hit(n->lo, n->hi, lo, inv, tmin, tmax)) continue;
        if (n->count) {
            for (uint32_t k = 0; k < n->count; k++) {
                uint32_t p = m->bvh.order[n->left + k];
                v3 a, b, c; float t, u, v;
                tri_verts(m, p, &a, &b, &c);
                float lim = h->found ? nextafterf(h->t, INFINITY) : tmax_world;
                if (intersect(a, b, c, lo, ld_, tmin, lim, cull, &t, &u, &v)) {
                    if (!(t < tmax_world)) continue;
                    if (any) return 1;
                    if (better(t, inst, p, h)) { h->t = t; h->u = u; h->v = v; h->inst = inst; h->prim = p; h->found = 1; }
                }
            }
        } else {
            stack[sp++] = n->right;
            stack[sp++] = n->left;
        }
    }
    return 0;
}

static int scene_trace(const vo_scene* s, v3 o, v3 d, float tmin, float tmax, uint32_t mask, hitrec* h, int any,
                       int brute) {
    h->found = 0; h->t = tmax;
    if (brute || !s->tlas.node_count) {
        for (uint32_t i = 0; i < s->inst_count; i++) {
            if (!s->valid[i] || !(s->inst[i].mask & mask)) continue;
            if (blas_trace(s, i, o, d, tmin, tmax, h, any, brute) && any) return 1;
        }
        return h->found;
    }
    v3 inv = safe_inv(d);
    uint32_t stack[128]; int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        uint32_t ni = stack[--sp];
        const bnode* n = &s->tlas.nodes[ni];
        float lim = h->found ? h->t : tmax;
        if (!box_hit(n->lo, n->hi, o, inv, tmin, lim)) continue;
        if (n->count) {
            for (uint32_t k = 0; k < n->count; k++) {
                uint32_t i = s->tlas.order[n->left + k];
                if (!s->valid[i] || !(s->inst[i].mask & mask)) continue;
                if (blas_trace(s, i, o, d, tmin, tmax, h, any, 0) && any) return 1;
            }
        } else {
            stack[sp++] = n->right;
            stack[sp++] = n->left;
        }
    }
    return h->found;
}

void vo_trace_closest(const vo_scene* s, const vh_ray* rays, uint32_t count, uint32_t mask, vh_hit* hits,
                      int brute_force) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)count; i++) {
        hitrec h;
        const vh_ray* r = &rays[i];
        vh_hit* o = &hits[i];
        if (scene_trace(s, ld(r->origin), ld(r->direction), r->tmin, r->tmax, mask, &h, 0, brute_force)) {
            o->t = h.t; o->instance = h.inst; o->primitive = h.prim;
            o->custom_index = s->inst[h.inst].custom_index & 0xffffffu;
            o->u = h.u; o->v = h.v;
        } else {
            o->t = -1.f; o->instance = o->primitive = o->custom_index = 0xffffffffu; o->u = o->v = 0.f;
        }
    }
}
void vo_trace_any(const vo_scene* s, const vh_ray* rays, uint32_t count, uint32_t mask, uint8_t* occluded,
                  int brute_force) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)count; i++) {
        hitrec h;
        const vh_ray* r = &rays[i];
        occluded[i] = (uint8_t)scene_trace(s, ld(r->origin), ld(r->direction), r->tmin, r->tmax, mask, &h, 1,
                                           brute_force);
    }
}

/* ---------------------------------------------------------- light tree */
/* crates/vork-render/src/lights/tree.rs:43-69 (leaf guides) and 106-170
 * (median spatial split) */
static uint32_t tree_build(vo_scene* s, const double* power, const float (*blo)[3], const float (*bhi)[3],
                           uint32_t* ids, uint32_t n, uint32_t parent, double scale, uint32_t* cursor) {
    uint32_t index = (*cursor)++;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double p = 0.;
    for (uint32_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) { lo[k] = fminf(lo[k], blo[ids[i]][k]); hi[k] = fmaxf(hi[k], bhi[ids[i]][k]); }
        p += power[ids[i]] / scale;
    }
    float pf = (float)p;
    if (pf < FLT_MIN) pf = 0.f;
    for (int k = 0; k < 3; k++) { s->tree[index].lower_power[k] = lo[k]; s->tree[index].upper_infinite[k] = hi[k]; }
    s->tree[index].lower_power[3] = pf;
    s->tree[index].upper_infinite[3] = 0.f;
    s->tree[index].links[0] = s->tree[index].links[1] = 0xffffffffu;
    s->tree[index].links[2] = parent;
    s->tree[index].links[3] = n;
    memset(s->tree[index].source, 0, sizeof(s->tree[index].source));
    if (n == 1) {
        s->tree[index].source[0] = 0; s->tree[index].source[1] = ids[0]; s->tree[index].source[2] = ids[0];
    } else {
        double ex = (double)hi[0] - lo[0], ey = (double)hi[1] - lo[1], ez = (double)hi[2] - lo[2];
        int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
        /* order by (lo+hi) centre on the axis, ties by identity */
        for (uint32_t i = 1; i < n; i++) {
            uint32_t key = ids[i]; uint32_t j = i;
            double ck = (double)blo[key][axis] + (double)bhi[key][axis];
            while (j > 0) {
                uint32_t pj = ids[j - 1];
                double cp = (double)blo[pj][axis] + (double)bhi[pj][axis];
                if (cp > ck || (cp == ck && pj > key)) { ids[j] = pj; j--; } else break;
            }
            ids[j] = key;
        }
        uint32_t mid = n / 2;
        uint32_t l = tree_build(s, power, blo, bhi, ids, mid, index, scale, cursor);
        uint32_t r = tree_build(s, power, blo, bhi, ids + mid, n - mid, index, scale, cursor);
        s->tree[index].links[0] = l; s->tree[index].links[1] = r;
    }
    return index;
}
/* light_tree.slang:18-36 */
static float tree_log_importance(const vo_scene* s, uint32_t i, v3 point) {
    const float* lp = s->tree[i].lower_power;
    const float* ui = s->tree[i].upper_infinite;
    if (!(lp[3] > 0.f)) return -FLT_MAX;
    float power = log2f(lp[3]);
    if (ui[3] != 0.f) return power;
    v3 center = add(scl(ld(lp), 0.5f), scl(ld(ui), 0.5f));
    v3 radius = sub(scl(ld(ui), 0.5f), scl(ld(lp), 0.5f));
    v3 delta = sub(point, center);
    float distance = fmaxf(LIGHT_TREE_MIN_DISTANCE_SQUARED, fmaxf(dot(delta, delta), dot(radius, radius)));
    if (!isfinite(distance)) return -FLT_MAX;
    return power - log2f(distance);
}
static float tree_branch(const vo_scene* s, uint32_t a, uint32_t b, v3 point) {
    float x = tree_log_importance(s, a, point), y = tree_log_importance(s, b, point);
    float scale = fmaxf(x, y);
    float uniform = (float)s->tree[a].links[3] / ((float)s->tree[a].links[3] + (float)s->tree[b].links[3]);
    if (scale == -FLT_MAX) return uniform;
    x = exp2f(x - scale); y = exp2f(y - scale);
    return x / (x + y);
}
/* light_tree.slang:58-84 */
uint32_t vo_light_pick(const vo_scene* s, const float pt[3], float ux, float uy, float* probability) {
    *probability = 0.f;
    if (!s->tree_nodes) return 0xffffffffu;
    v3 point = ld(pt);
    uint32_t node = 0; float prob = 1.f;
    const float one = 1.f - FLT_EPSILON * 0.5f;
    float choice = clampf(ux, 0.f, one);
    int uniform = uy < LIGHT_TREE_UNIFORM_MIX;
    uint32_t total = s->tree[0].links[3];
    uint32_t rank = (uint32_t)(choice * (float)total);
    if (rank > total - 1) rank = total - 1;
    while (s->tree[node].links[0] != 0xffffffffu) {
        uint32_t cl = s->tree[node].links[0], cr = s->tree[node].links[1];
        float p = tree_branch(s, cl, cr, point);
        int left;
        if (uniform) {
            uint32_t cnt = s->tree[cl].links[3];
            left = rank < cnt;
            if (!left) rank -= cnt;
        } else {
            left = choice < p;
            choice = left ? choice / p : (choice - p) / (1.f - p);
            choice = fminf(choice, one);
        }
        prob *= left ? p : 1.f - p;
        node = left ? cl : cr;
    }
    *probability = LIGHT_TREE_UNIFORM_MIX / (float)total + (1.f - LIGHT_TREE_UNIFORM_MIX) * prob;
    return s->tree[node].source[1];
}
/* light_tree.slang:37-47 */
float vo_light_probability(const vo_scene* s, uint32_t light, const float pt[3]) {
    if (!s->tree_nodes) return 0.f;
    uint32_t child = 0;
    for (uint32_t i = 0; i < s->tree_nodes; i++)
        if (s->tree[i].links[0] == 0xffffffffu && s->tree[i].source[1] == light) child = i;
    float prob = 1.f;
    v3 point = ld(pt);
    while (s->tree[child].links[2] != 0xffffffffu) {
        uint32_t parent = s->tree[child].links[2];
        uint32_t cl = s->tree[parent].links[0], cr = s->tree[parent].links[1];
        float p = tree_branch(s, cl, cr, point);
        prob *= child == cl ? p : 1.f - p;
        child = parent;
    }
    return LIGHT_TREE_UNIFORM_MIX / (float)s->tree[0].links[3] + (1.f - LIGHT_TREE_UNIFORM_MIX) * prob;
}

vo_scene* vo_scene_create(const vo_mesh_desc* meshes, uint32_t mesh_count, const vh_instance* instances,
                          uint32_t instance_count, const vh_scene_desc* desc) {
    vo_scene* s = (vo_scene*)calloc(1, sizeof(vo_scene));
    s->mesh_count = mesh_count;
    s->meshes = (mesh_t*)calloc(mesh_count ? mesh_count : 1, sizeof(mesh_t));
    for (uint32_t mi = 0; mi < mesh_count; mi++) {
        const vo_mesh_desc* md = &meshes[mi];
        mesh_t* m = &s->meshes[mi];
        m->vcount = md->vertex_count; m->tcount = md->triangle_count;
        m->pos = (float*)malloc(sizeof(float) * 3 * (m->vcount ? m->vcount : 1));
        memcpy(m->pos, md->positions, sizeof(float) * 3 * m->vcount);
        if (md->normals) {
            m->nrm = (float*)malloc(sizeof(float) * 3 * (m->vcount ? m->vcount : 1));
            memcpy(m->nrm, md->normals, sizeof(float) * 3 * m->vcount);
        }
        m->idx = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (m->tcount ? m->tcount : 1));
        memcpy(m->idx, md->indices, sizeof(uint32_t) * 3 * m->tcount);
        if (md->triangle_materials) {
            m->mat = (uint32_t*)malloc(sizeof(uint32_t) * (m->tcount ? m->tcount : 1));
            memcpy(m->mat, md->triangle_materials, sizeof(uint32_t) * m->tcount);
        }
        prim_box* boxes = (prim_box*)malloc(sizeof(prim_box) * (m->tcount ? m->tcount : 1));
        for (int k = 0; k < 3; k++) { m->lo[k] = INFINITY; m->hi[k] = -INFINITY; }
        for (uint32_t t = 0; t < m->tcount; t++) {
            v3 a, b, c;
            tri_verts(m, t, &a, &b, &c);
            for (int k = 0; k < 3; k++) {
                float x = comp(a, k), y = comp(b, k), z = comp(c, k);
                boxes[t].lo[k] = fminf(x, fminf(y, z));
                boxes[t].hi[k] = fmaxf(x, fmaxf(y, z));
                boxes[t].c[k] = (x + y + z) / 3.f;
                m->lo[k] = fminf(m->lo[k], boxes[t].lo[k]);
                m->hi[k] = fmaxf(m->hi[k], boxes[t].hi[k]);
            }
        }
        bvh_build(&m->bvh, boxes, m->tcount);
        free(boxes);
    }
    s->inst_count = instance_count;
    s->inst = (vh_instance*)malloc(sizeof(vh_instance) * (instance_count ? instance_count : 1));
    memcpy(s->inst, instances, sizeof(vh_instance) * instance_count);
    s->w2o = (float*)calloc(12 * (instance_count ? instance_count : 1), sizeof(float));
    s->n2w = (float*)calloc(9 * (instance_count ? instance_count : 1), sizeof(float));
    s->valid = (uint8_t*)calloc(instance_count ? instance_count : 1, 1);
    prim_box* ib = (prim_box*)malloc(sizeof(prim_box) * (instance_count ? instance_count : 1));
    for (uint32_t i = 0; i < instance_count; i++) {
        const float* m = instances[i].transform;
        s->valid[i] = (uint8_t)(vo_invert_transform(m, s->w2o + 12 * i) && instances[i].blas < mesh_count);
        /* normal matrix = transpose of the inverse 3x3 */
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) s->n2w[9 * i + 3 * r + c] = s->w2o[12 * i + 4 * c + r];
        for (int k = 0; k < 3; k++) { ib[i].lo[k] = INFINITY; ib[i].hi[k] = -INFINITY; }
        if (s->valid[i]) {
            const mesh_t* me = &s->meshes[instances[i].blas];
            if (me->tcount == 0) { s->valid[i] = 0; }
            else {
                for (int corner = 0; corner < 8; corner++) {
                    v3 p = mk(corner & 1 ? me->hi[0] : me->lo[0], corner & 2 ? me->hi[1] : me->lo[1],
                              corner & 4 ? me->hi[2] : me->lo[2]);
                    v3 w = xform_point(m, p);
                    for (int k = 0; k < 3; k++) {
                        ib[i].lo[k] = fminf(ib[i].lo[k], comp(w, k));
                        ib[i].hi[k] = fmaxf(ib[i].hi[k], comp(w, k));
                    }
                }
                /* widen by a relative ulp budget: the box is a traversal guide */
                for (int k = 0; k < 3; k++) {
                    float e = (fabsf(ib[i].lo[k]) + fabsf(ib[i].hi[k])) * 4.f * FLT_EPSILON;
                    ib[i].lo[k] -= e; ib[i].hi[k] += e;
                }
            }
        }
        for (int k = 0; k < 3; k++) ib[i].c[k] = s->valid[i] ? 0.5f * (ib[i].lo[k] + ib[i].hi[k]) : 0.f;
        if (!s->valid[i])
            for (int k = 0; k < 3; k++) { ib[i].lo[k] = ib[i].hi[k] = 0.f; }
    }
    bvh_build(&s->tlas, ib, instance_count);
    free(ib);
    s->material_count = desc->material_count;
    s->materials = (vh_material*)malloc(sizeof(vh_material) * (desc->material_count ? desc->material_count : 1));
    memcpy(s->materials, desc->materials, sizeof(vh_material) * desc->material_count);
    s->light_count = desc->light_count;
    s->lights = (vh_rect_light*)malloc(sizeof(vh_rect_light) * (desc->light_count ? desc->light_count : 1));
    memcpy(s->lights, desc->lights, sizeof(vh_rect_light) * desc->light_count);
    if (desc->light_count) {
        uint32_t n = desc->light_count;
        s->tree = calloc(2 * n - 1, sizeof(*s->tree));
        double* power = (double*)malloc(sizeof(double) * n);
        float(*blo)[3] = malloc(sizeof(float) * 3 * n);
        float(*bhi)[3] = malloc(sizeof(float) * 3 * n);
        uint32_t* ids = (uint32_t*)malloc(sizeof(uint32_t) * n);
        double scale = 0.;
        for (uint32_t i = 0; i < n; i++) {
            const vh_rect_light* l = &desc->lights[i];
            double pw = 0.;
            for (int k = 0; k < 3; k++) pw += (double)l->radiance[k] * (double)LUMINANCE[k];
            power[i] = pw > 0. ? pw : 0.;
            if (power[i] > scale) scale = power[i];
            for (int k = 0; k < 3; k++) {
                float r = (fabsf(l->u[k]) + fabsf(l->v[k])) * 0.5f;
                blo[i][k] = l->position[k] - r; bhi[i][k] = l->position[k] + r;
            }
            ids[i] = i;
        }
        if (scale < DBL_MIN) scale = DBL_MIN;
        uint32_t cursor = 0;
        tree_build(s, power, (const float(*)[3])blo, (const float(*)[3])bhi, ids, n, 0xffffffffu, scale, &cursor);
        s->tree_nodes = cursor;
        free(power); free(blo); free(bhi); free(ids);
    }
    s->camera = desc->camera;
    s->sky = desc->sky;
    vo_filter_table(s->filter);
    return s;
}

void vo_scene_release(vo_scene* s) {
    if (!s) return;
    for (uint32_t i = 0; i < s->mesh_count; i++) {
        free(s->meshes[i].pos); free(s->meshes[i].nrm); free(s->meshes[i].idx); free(s->meshes[i].mat);
        free(s->meshes[i].bvh.nodes); free(s->meshes[i].bvh.order);
    }
    free(s->meshes); free(s->inst); free(s->w2o); free(s->n2w); free(s->valid);
    free(s->tlas.nodes); free(s->tlas.order);
    free(s->materials); free(s->lights); free(s->tree);
    free(s);
}

/* ======================================================================
 * GGX energy tables: the bake of crates/vork-render/shaders/environment.slang:
 * 14-120 (mode 2) and the lookups of crates/vork-render/shaders/frame.slang:
 * 41-66, with bilinear filtering on the baked grid.
 * ==================================================================== */
/* openpbr_core.slang:5-7 */
static inline float specular_alpha(float r) { return fmaxf(r * r, GGX_ALPHA_MIN); }
/* openpbr_core.slang:13-21 */
float vo_modulated_ior(float eta, float weight) {
    if (weight == 1.f || eta == 1.f) return eta;
    float contrast = (eta - 1.f) / (eta + 1.f);
    float limit = u2f(0x3f7fffffu);
    float epsilon = clampf(sqrtf(fmaxf(weight, 0.f)) * contrast, -limit, limit);
    return (1.f + epsilon) / (1.f - epsilon);
}
/* openpbr_core.slang:22-27 */
float vo_fresnel(float c, float eta) {
    c = clampf(c, 0.f, 1.f);
    if (eta == 1.f) return 0.f;
    float si2 = (1.f - c * c) / (eta * eta);
    if (si2 >= 1.f) return 1.f;
    float ct = sqrtf(fmaxf(0.f, 1.f - si2));
    float rs = (c - eta * ct) / (c + eta * ct), rp = (eta * c - ct) / (eta * c + ct);
    return (rs * rs + rp * rp) * .5f;
}
/* openpbr_core.slang:105-108 */
static inline float distribution(float nh, float alpha) {
    float x = (1.f - nh * nh) + nh * nh * alpha * alpha;
    return alpha * alpha / (PI_F * x * x);
}
/* openpbr_core.slang:111-116 */
static inline float smith_denominator(float no, float ni, float alpha) {
    float ci = fabsf(ni), a2 = alpha * alpha;
    float po = sqrtf(a2 + (1.f - a2) * no * no), pi = sqrtf(a2 + (1.f - a2) * ci * ci);
    return ci * po + no * pi;
}
static inline float g2(float no, float ni, float alpha) { return 2.f * no * fabsf(ni) / smith_denominator(no, ni, alpha); }
/* openpbr_core.slang:126-138 (Heitz 2018) */
static v3 ggx_visible_normal(v3 view, float ax, float ay, float u0, float u1) {
    v3 v = nrm(mk(ax * view.x, ay * view.y, view.z));
    float ls = v.x * v.x + v.y * v.y;
    v3 t = ls > 0.f ? dvs(mk(-v.y, v.x, 0.f), sqrtf(ls)) : mk(1.f, 0.f, 0.f);
    v3 b = cross(v, t);
    float radius = sqrtf(u0), phi = 2.f * PI_F * u1;
    float x = radius * cosf(phi);
    float y = lerpf(sqrtf(fmaxf(1.f - x * x, 0.f)), radius * sinf(phi), .5f * (1.f + v.z));
    v3 p = add(add(scl(t, x), scl(b, y)), scl(v, sqrtf(fmaxf(1.f - x * x - y * y, 0.f))));
    return nrm(mk(ax * p.x, ay * p.y, fmaxf(p.z, 0.f)));
}
/* openpbr_core.slang:141-145 */
static float ggx_visible_weight(float no, float ni, float alpha) {
    if (ni <= 0.f) return 0.f;
    float a2 = alpha * alpha;
    return ni * (no + sqrtf(a2 + (1.f - a2) * no * no)) / smith_denominator(no, ni, alpha);
}
/* openpbr_core.slang:147-152 */
static float ggx_visible_reflection(float no, float alpha, float u0, float u1, float* vh) {
    v3 view = mk(sqrtf(fmaxf(1.f - no * no, 0.f)), 0.f, no);
    v3 h = ggx_visible_normal(view, alpha, alpha, u0, u1);
    v3 light = reflect3(neg(view), h);
    *vh = clampf(dot(view, h), 0.f, 1.f);
    return ggx_visible_weight(no, light.z, alpha);
}
/* openpbr_core.slang:154-166 */
float vo_dielectric_directional_albedo(float no, float roughness, float eta, uint32_t samples) {
    if (eta == 1.f) return 0.f;
    if (roughness == 0.f) return eta <= 0.f ? 1.f : vo_fresnel(no, eta);
    float a = specular_alpha(roughness), result = 0.f;
    for (uint32_t i = 0; i < samples; i++) {
        float u0 = ((float)i + .5f) / (float)samples;
        float u1 = (float)(reverse_bits(i) >> 8) * UINT_SCALE;
        float vh;
        float weight = ggx_visible_reflection(no, a, u0, u1, &vh);
        result += (eta <= 0.f ? 1.f : vo_fresnel(vh, eta)) * weight;
    }
    return result / (float)samples;
}
/* environment.slang:19-27 */
static float brdf_cosine(uint32_t column) { float t = (float)column / (float)(BRDF_SIZE - 1); float t2 = t * t; return t2 * t2; }
static float brdf_alpha(uint32_t row) {
    if (row == 0) return GGX_ALPHA_MIN;
    if (row == BRDF_SIZE - 1) return 1.f;
    return exp2f(log2f(GGX_ALPHA_MIN) * (1.f - (float)row / (float)(BRDF_SIZE - 1)));
}
/* environment.slang:78-114: rgb moments of one texel, alpha = E_avg in column 0 */
void vo_bake_moments_texel(uint32_t column, uint32_t row, float out[4]) {
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    float nv = brdf_cosine(column), alpha = brdf_alpha(row);
    for (uint32_t i = 0; i < ENVIRONMENT_SAMPLES; i++) {
        float u0 = ((float)i + .5f) / (float)ENVIRONMENT_SAMPLES;
        float u1 = (float)reverse_bits(i) * RADICAL_INVERSE_SCALE;
        float vh;
        float g = ggx_visible_reflection(nv, alpha, u0, u1, &vh);
        float f = powf(1.f - vh, 5.f);
        s0 += (1.f - f) * g; s1 += f * g; s2 += vh * powf(1.f - vh, 6.f) * g;
    }
    out[0] = s0 / (float)ENVIRONMENT_SAMPLES;
    out[1] = s1 / (float)ENVIRONMENT_SAMPLES;
    out[2] = s2 / (float)ENVIRONMENT_SAMPLES;
    float average = 0.f;
    if (column == 0) {
        for (uint32_t j = 0; j < BRDF_SIZE; j++) {
            float nvj = brdf_cosine(j);
            float interval = .5f * (brdf_cosine(j + 1 < BRDF_SIZE ? j + 1 : BRDF_SIZE - 1) - brdf_cosine(j > 0 ? j - 1 : 0));
            for (uint32_t i = 0; i < ENVIRONMENT_SAMPLES; i++) {
                float u0 = ((float)i + .5f) / (float)ENVIRONMENT_SAMPLES;
                float u1 = (float)reverse_bits(i) * RADICAL_INVERSE_SCALE;
                float vh;
                average += interval * 2.f * nvj * ggx_visible_reflection(nvj, alpha, u0, u1, &vh);
            }
        }
        average /= (float)ENVIRONMENT_SAMPLES;
    }
    out[3] = average;
}
static float* MOMENTS = NULL;          /* [row][column][4] */
static float* DIELECTRIC[BRDF_IOR_STEPS + 1]; /* [slice][r][nv] */
static void ensure_moments(void) {
    if (MOMENTS) return;
#pragma omp critical(vo_lut)
    {
        if (!MOMENTS) {
            float* t = (float*)malloc(sizeof(float) * 4 * BRDF_SIZE * BRDF_SIZE);
#pragma omp parallel for schedule(dynamic, 1)
            for (int64_t k = 0; k < (int64_t)(BRDF_SIZE * BRDF_SIZE); k++)
                vo_bake_moments_texel((uint32_t)(k % BRDF_SIZE), (uint32_t)(k / BRDF_SIZE), t + 4 * k);
            MOMENTS = t;
        }
    }
}
static const float* dielectric_slice(uint32_t slice) {
    if (DIELECTRIC[slice]) return DIELECTRIC[slice];
#pragma omp critical(vo_lut)
    {
        if (!DIELECTRIC[slice]) {
            const uint32_t n = BRDF_ALBEDO_SIZE;
            float* t = (float*)malloc(sizeof(float) * n * n);
            float q = (float)slice / (float)BRDF_IOR_STEPS;
            float eta = slice == BRDF_IOR_STEPS ? 0.f : q / fmaxf(1.f - q, POSITIVE_EPSILON);
#pragma omp parallel for schedule(dynamic, 4)
            for (int64_t k = 0; k < (int64_t)(n * n); k++) {
                uint32_t cx = (uint32_t)(k % n), cy = (uint32_t)(k / n);
                float no = fmaxf((float)cx / (float)(n - 1), GGX_ALPHA_MIN);
                float r = (float)cy / (float)(n - 1);
                t[k] = vo_dielectric_directional_albedo(no, r, eta, DIELECTRIC_ALBEDO_SAMPLES);
            }
            DIELECTRIC[slice] = t;
        }
    }
    return DIELECTRIC[slice];
}
/* Bilinear fetch of a w x h grid with clamp to edge, texel centres at integers. */
static float bilinear(const float* grid, uint32_t w, uint32_t h, uint32_t stride, uint32_t channel, float x, float y) {
    x = clampf(x, 0.f, (float)(w - 1)); y = clampf(y, 0.f, (float)(h - 1));
    uint32_t x0 = (uint32_t)x, y0 = (uint32_t)y;
    uint32_t x1 = x0 + 1 < w ? x0 + 1 : w - 1, y1 = y0 + 1 < h ? y0 + 1 : h - 1;
    float fx = x - (float)x0, fy = y - (float)y0;
    float a = grid[(y0 * w + x0) * stride + channel], b = grid[(y0 * w + x1) * stride + channel];
    float c = grid[(y1 * w + x0) * stride + channel], d = grid[(y1 * w + x1) * stride + channel];
    return lerpf(lerpf(a, b, fx), lerpf(c, d, fx), fy);
}
/* frame.slang:41-51 */
void vo_brdf_moments(float nv, float roughness, float out[4]) {
    ensure_moments();
    float ux = sqrtf(sqrtf(clampf(nv, 0.f, 1.f)));
    float uy = log2f(clampf(roughness * roughness, GGX_ALPHA_MIN, 1.f) / GGX_ALPHA_MIN) / log2f(1.f / GGX_ALPHA_MIN);
    float x = ux * (float)(BRDF_SIZE - 1), y = uy * (float)(BRDF_SIZE - 1);
    for (uint32_t c = 0; c < 4; c++) out[c] = bilinear(MOMENTS, BRDF_SIZE, BRDF_SIZE, 4, c, x, y);
}
/* frame.slang:52-66 */
float vo_dielectric_albedo(float nv, float roughness, float eta) {
    if (eta == 1.f) return 0.f;
    float at = eta / (1.f + eta) * (float)BRDF_IOR_STEPS;
    uint32_t first = (uint32_t)at;
    if (first > BRDF_IOR_STEPS - 1) first = BRDF_IOR_STEPS - 1;
    const uint32_t n = BRDF_ALBEDO_SIZE;
    float x = clampf(nv, 0.f, 1.f) * (float)(n - 1), y = clampf(roughness, 0.f, 1.f) * (float)(n - 1);
    float a = bilinear(dielectric_slice(first), n, n, 1, 0, x, y);
    float b = bilinear(dielectric_slice(first + 1), n, n, 1, 0, x, y);
    return lerpf(a, b, at - (float)first);
}

/* ======================================================================
 * BSDF (closure subset).  References: bsdf.slang:184-510,
 * bsdf_response.slang:12-20,98-124, bsdf_energy.slang:44-88,
 * bsdf_film.slang:110-123, openpbr_core.slang:1-226,
 * orthonormal_basis.slang:1-9, dielectric_interface.slang:11-17.
 * ==================================================================== */
typedef struct {
    v3 base_color; float metal, spec_weight, roughness, ior;
    v3 n;        /* closure geometry_normal */
} closure_t;
typedef struct {
    float p[3];         /* diffuse, reflection, refraction */
    v3 diffuse;         /* direct_diffuse */
    v3 metal_comp;
    v3 smooth;
    float bump_a2;
    v3 interface_n;
} prep_t;

static inline float dielectric_ior(const closure_t* c) { return vo_modulated_ior(c->ior, c->spec_weight); }
/* orthonormal_basis.slang:2-9 (Duff et al. 2017) */
static inline v3 basis_direction(v3 n, v3 v) {
    float s = n.z >= 0.f ? 1.f : -1.f, a = -1.f / (s + n.z), b = n.x * n.y * a;
    v3 r0 = mk(1.f + s * n.x * n.x * a, s * b, -s * n.x), r1 = mk(b, s + n.y * n.y * a, -n.y);
    return add(add(scl(r0, v.x), scl(r1, v.y)), scl(n, v.z));
}
/* openpbr_core.slang:36-49 (F82 edge tint with white tint: no correction) */
static v3 reflection_fresnel_at_ior(v3 base_color, float metal, float spec_weight, float cosine, float eta) {
    cosine = clampf(cosine, 0.f, 1.f);
    float fd = vo_fresnel(cosine, eta);
    v3 dielectric = mk(fd, fd, fd);
    if (metal <= 0.f) return dielectric;
    v3 f0 = base_color;
    float f = powf(1.f - cosine, 5.f);
    v3 m = mk(fmaxf(f0.x + (1.f - f0.x) * f, 0.f), fmaxf(f0.y + (1.f - f0.y) * f, 0.f), fmaxf(f0.z + (1.f - f0.z) * f, 0.f));
    return lerp3(dielectric, scl(m, spec_weight), metal);
}
/* bsdf_film.slang:110-123 without film or thin walls */
static v3 surface_reflection(const closure_t* c, float cosine) {
    return reflection_fresnel_at_ior(c->base_color, c->metal, c->spec_weight, cosine, dielectric_ior(c));
}
/* bsdf_energy.slang:76-88 */
static v3 compensated_surface_reflection(const closure_t* c, float cosine, v3 comp_) {
    v3 f = surface_reflection(c, cosine);
    if (comp_.x > 0.f || comp_.y > 0.f || comp_.z > 0.f) {
        if (c->metal == 1.f) f = add(f, mul(comp_, f));
        else {
            v3 conductor = reflection_fresnel_at_ior(c->base_color, 1.f, c->spec_weight, cosine,
                                                     vo_modulated_ior(c->ior, c->spec_weight));
            f = add(f, scl(mul(comp_, conductor), c->metal));
        }
    }
    return f;
}
/* bsdf_energy.slang:44-74 (isotropic, white tint) */
static v3 metal_compensation(const closure_t* c, v3 view) {
    if (c->metal <= 0.f || c->roughness == 0.f) return mk(0.f, 0.f, 0.f);
    float m[4];
    vo_brdf_moments(fmaxf(dot(c->n, view), .0001f), c->roughness, m);
    float energy = m[0] + m[1];
    float e_avg_m[4];
    vo_brdf_moments(0.f, c->roughness, e_avg_m);
    float e_avg = e_avg_m[3];
    float missing = 1.f / fmaxf(energy, 1e-6f) - 1.f;
    v3 f0 = c->base_color;
    v3 avg = mk(clampf(f0.x + (1.f - f0.x) / 21.f, 0.f, 1.f), clampf(f0.y + (1.f - f0.y) / 21.f, 0.f, 1.f),
                clampf(f0.z + (1.f - f0.z) / 21.f, 0.f, 1.f));
    return mk(missing * avg.x * e_avg / fmaxf(1.f - avg.x * (1.f - e_avg), 1e-6f),
              missing * avg.y * e_avg / fmaxf(1.f - avg.y * (1.f - e_avg), 1e-6f),
              missing * avg.z * e_avg / fmaxf(1.f - avg.z * (1.f - e_avg), 1e-6f));
}
/* Denoiser guide albedo (path.slang:1553-1570) over this closure subset:
 * (1 - metal) ((1 - E_spec) base + E_spec) + metal (F0 m.x + m.y)(1 + comp)
 * spec_weight, clamped to [0, 1]; E_spec is the dielectric table at the
 * modulated IOR, m the GGX moments at the view cosine. */
static v3 guide_albedo(const closure_t* c, v3 view) {
    float nv = fmaxf(dot(c->n, view), .0001f);
    float ed = c->metal < 1.f ? vo_dielectric_albedo(nv, c->roughness, dielectric_ior(c)) : 0.f;
    v3 a = add(scl(c->base_color, (1.f - c->metal) * (1.f - ed)), scl(mk(ed, ed, ed), 1.f - c->metal));
    if (c->metal > 0.f) {
        float m[4];
        vo_brdf_moments(nv, c->roughness, m);
        v3 comp = metal_compensation(c, view);
        v3 e = scl(add(scl(c->base_color, m[0]), mk(m[1], m[1], m[1])), c->spec_weight);
        a = add(a, scl(mk(e.x * (1.f + comp.x), e.y * (1.f + comp.y), e.z * (1.f + comp.z)), c->metal));
    }
    return mk(clampf(a.x, 0.f, 1.f), clampf(a.y, 0.f, 1.f), clampf(a.z, 0.f, 1.f));
}

/* bsdf_response.slang:98-110 */
static v3 direct_diffuse(const closure_t* c, v3 view) {
    if (c->metal < 1.f) {
        float albedo = vo_dielectric_albedo(fmaxf(dot(c->n, view), .0001f), c->roughness, dielectric_ior(c));
        return scl(c->base_color, (1.f - c->metal) * (1.f - albedo) / PI_F);
    }
    return mk(0.f, 0.f, 0.f);
}
/* bsdf.slang:197-212 */
static void transport_valid_interface_normal(v3* out, v3 normal, v3 geometric, v3 view) {
    float iz = dot(view, geometric);
    if (iz <= 0.f) { *out = normal; return; }
    if (dot(normal, geometric) < 0.f) normal = neg(normal);
    float threshold = fminf(.9f * iz, .01f);
    if (dot(normal, view) > 0.f && dot(reflect3(neg(view), normal), geometric) >= threshold) { *out = normal; return; }
    v3 tangent = sub(normal, scl(geometric, dot(normal, geometric)));
    float ts = dot(tangent, tangent);
    if (ts <= 0.f) { *out = geometric; return; }
    v3 axis = scl(tangent, 1.f / sqrtf(ts));
    float ix = dot(view, axis), ls = ix * ix + iz * iz;
    float c2 = clampf(.5f * (1.f + (iz * threshold - ix * sqrtf(fmaxf(ls - threshold * threshold, 0.f))) / ls), 0.f, 1.f);
    *out = add(scl(axis, sqrtf(1.f - c2)), scl(geometric, sqrtf(c2)));
}
/* bsdf.slang:173-196 */
static void transport_probabilities(const closure_t* c, v3 view, float p[3]) {
    float diffuse = fmaxf((c->base_color.x * LUMINANCE[0] + c->base_color.y * LUMINANCE[1] + c->base_color.z * LUMINANCE[2]) *
                              (1.f - c->metal),
                          0.f);
    float f = vo_fresnel(fmaxf(dot(c->n, view), 0.f), dielectric_ior(c));
    float reflected = f;
    float dielectric = reflected * (LUMINANCE[0] + LUMINANCE[1] + LUMINANCE[2]);
    float reflection = fmaxf(lerpf(dielectric, c->spec_weight, c->metal), 0.f);
    float total = diffuse + reflection;
    if (total > 0.f) {
        float t = fmaxf(total, 1e-20f);
        p[0] = diffuse / t; p[1] = reflection / t; p[2] = 0.f;
    } else {
        p[0] = p[1] = p[2] = 0.f;
    }
}
/* bsdf.slang:218-232 */
static prep_t transport_prepare(const closure_t* c, v3 view, v3 smooth, v3 geometric) {
    prep_t r;
    closure_t base = *c;
    transport_valid_interface_normal(&base.n, c->n, geometric, view);
    transport_probabilities(&base, view, r.p);
    float deviation = fabsf(dot(smooth, c->n));
    r.bump_a2 = clampf((1.f - deviation * deviation) / fmaxf(8.f * deviation * deviation, 1e-20f), 0.f, 1.f);
    r.diffuse = direct_diffuse(&base, view);
    r.metal_comp = metal_compensation(&base, view);
    r.smooth = smooth;
    r.interface_n = base.n;
    return r;
}
/* bsdf.slang:235-244 */
static float transport_bump_shadow(const closure_t* c, v3 direction, const prep_t* pr) {
    float cosine = dot(pr->smooth, direction);
    if (cosine * dot(pr->smooth, c->n) * dot(c->n, direction) < 0.f) return 0.f;
    if (pr->bump_a2 == 0.f) return 1.f;
    float mu = fabsf(cosine);
    return mu > 0.f ? 2.f * mu / (mu + sqrtf(pr->bump_a2 + (1.f - pr->bump_a2) * mu * mu)) : 0.f;
}
/* bsdf_microfacet.slang:42-52 (isotropic) */
static inline float transport_visible_scale(const closure_t* c, v3 view) {
    float nv = dot(c->n, view), a = specular_alpha(c->roughness);
    return 2.f / (nv + sqrtf(a * a + (1.f - a * a) * nv * nv));
}
/* bsdf_microfacet.slang:53-62 */
static v3 transport_visible_normal(const closure_t* c, v3 view, float u0, float u1) {
    v3 n = c->n;
    v3 x = basis_direction(n, mk(1.f, 0.f, 0.f)), y = basis_direction(n, mk(0.f, 1.f, 0.f));
    float a = specular_alpha(c->roughness);
    v3 local = mk(dot(view, x), dot(view, y), dot(view, n));
    v3 s = ggx_visible_normal(local, a, a, u0, u1);
    return add(add(scl(x, s.x), scl(y, s.y)), scl(n, s.z));
}
/* bsdf.slang:270-283 */
static float transport_pdf_mask(const closure_t* c, v3 view, v3 direction, uint32_t mask, const prep_t* pr) {
    closure_t base = *c;
    base.n = pr->interface_n;
    float nl = dot(base.n, direction);
    float pdf = (mask & VH_RAY_DIFFUSE) ? pr->p[0] * (dot(c->n, direction) > 0.f ? 1.f : 0.f) * fabsf(dot(c->n, direction)) / PI_F : 0.f;
    if (nl <= 0.f) return pdf;
    v3 h = nrm(add(view, direction));
    float d = distribution(fmaxf(dot(base.n, h), 0.f), specular_alpha(base.roughness));
    return pdf + ((mask & VH_RAY_SPECULAR) ? pr->p[1] * (d * transport_visible_scale(&base, view) / 4.f) : 0.f);
}
/* bsdf_response.slang:111-124 (isotropic) */
static v3 direct_brdf_prepared(const closure_t* c, v3 view, v3 light, uint32_t mask, v3 diffuse, v3 comp_) {
    v3 n = c->n;
    float nl = fmaxf(dot(n, light), 0.f), nv = fmaxf(dot(n, view), .0001f);
    if (nl <= 0.f) return mk(0.f, 0.f, 0.f);
    v3 specular = mk(0.f, 0.f, 0.f);
    if (mask & VH_RAY_SPECULAR) {
        v3 h = nrm(add(view, light));
        float nh = fmaxf(dot(n, h), 0.f), vh = fmaxf(dot(view, h), 0.f);
        v3 f = compensated_surface_reflection(c, vh, comp_);
        float a = specular_alpha(c->roughness);
        specular = scl(f, distribution(nh, a) * g2(nv, nl, a) / (4.f * nv * nl));
    }
    v3 r = add((mask & VH_RAY_DIFFUSE) ? diffuse : mk(0.f, 0.f, 0.f), (mask & VH_RAY_SPECULAR) ? specular : mk(0.f, 0.f, 0.f));
    return scl(r, nl);
}
/* bsdf.slang:347-365 */
static v3 transport_eval_mask(const closure_t* c, v3 view, v3 direction, uint32_t mask, const prep_t* pr) {
    closure_t base = *c;
    base.n = pr->interface_n;
    float bump = transport_bump_shadow(c, direction, pr);
    v3 specular = direct_brdf_prepared(&base, view, direction, mask & VH_RAY_SPECULAR, mk(0.f, 0.f, 0.f), pr->metal_comp);
    int behind = dot(c->n, direction) <= 0.f;
    v3 diffuse = mk(0.f, 0.f, 0.f);
    if ((mask & VH_RAY_DIFFUSE) && !behind) diffuse = scl(pr->diffuse, bump * dot(c->n, direction));
    return add(specular, diffuse);
}
typedef struct { v3 direction, weight; float pdf; uint32_t event; } bsdf_sample;
/* bsdf.slang:446-510 */
static bsdf_sample transport_sample(const closure_t* c, v3 view, float r0, float r1, float r2, const prep_t* pr) {
    bsdf_sample res;
    memset(&res, 0, sizeof(res));
    v3 n = pr->interface_n;
    closure_t base = *c;
    base.n = n;
    if (r2 < pr->p[0]) {
        float radius = sqrtf(r0), phi = TAU_F * r1;
        res.direction = basis_direction(c->n, mk(radius * cosf(phi), radius * sinf(phi), sqrtf(1.f - r0)));
        res.event = VH_RAY_DIFFUSE;
    } else if (r2 < pr->p[0] + pr->p[1]) {
        v3 h = transport_visible_normal(&base, view, r0, r1);
        res.direction = reflect3(neg(view), h);
        res.event = VH_RAY_SPECULAR;
    } else {
        return res;
    }
    if (res.event == VH_RAY_SPECULAR && dot(n, res.direction) <= 0.f) { memset(&res, 0, sizeof(res)); return res; }
    res.pdf = transport_pdf_mask(c, view, res.direction, res.event, pr);
    if (res.pdf > 0.f) res.weight = dvs(transport_eval_mask(c, view, res.direction, res.event, pr), res.pdf);
    return res;
}
/* path.slang:833-836 */
static inline float mis_power(float sampled, float other) {
    if (sampled <= 0.f) return 0.f;
    float ratio = other / sampled;
    return 1.f / (1.f + ratio * ratio);
}
/* path.slang:837-840 (no thin walls) */
static inline int event_hemisphere(uint32_t event, v3 direction, v3 face, v3 smooth) {
    if (event == VH_RAY_TRANSMISSION) return dot(face, direction) < 0.f && dot(smooth, direction) < 0.f;
    return dot(smooth, direction) > 0.f;
}
/* path.slang:841-850 (transmission carries no response in this subset) */
static v3 analytic_mis_response(const closure_t* c, v3 view, v3 direction, v3 geometric, float light_pdf, const prep_t* pr) {
    v3 value = mk(0.f, 0.f, 0.f);
    const uint32_t events[2] = {VH_RAY_DIFFUSE, VH_RAY_SPECULAR};
    for (int i = 0; i < 2; i++) {
        uint32_t e = events[i];
        if (!event_hemisphere(e, direction, geometric, pr->smooth)) continue;
        float w = light_pdf > 0.f ? mis_power(light_pdf, transport_pdf_mask(c, view, direction, e, pr)) : 1.f;
        value = add(value, scl(transport_eval_mask(c, view, direction, e, pr), w));
    }
    return value;
}
static closure_t make_closure(const vh_material* m, v3 n) {
    closure_t c;
    c.base_color = ld(m->base_color);
    c.metal = m->base_metalness;
    c.spec_weight = m->specular_weight;
    c.roughness = m->specular_roughness;
    c.ior = m->specular_ior;
    c.n = n;
    return c;
}
void vo_bsdf_sample(const vh_material* m, const float n[3], const float face[3], const float view[3], const float random[3],
                    float out[8]) {
    closure_t c = make_closure(m, ld(n));
    prep_t pr = transport_prepare(&c, ld(view), ld(n), ld(face));
    bsdf_sample s = transport_sample(&c, ld(view), random[0], random[1], random[2], &pr);
    out[0] = s.direction.x; out[1] = s.direction.y; out[2] = s.direction.z;
    out[3] = s.weight.x; out[4] = s.weight.y; out[5] = s.weight.z;
    out[6] = s.pdf; out[7] = (float)s.event;
}
void vo_bsdf_eval(const vh_material* m, const float n[3], const float face[3], const float view[3], const float dir[3],
                  uint32_t mask, float out[4]) {
    closure_t c = make_closure(m, ld(n));
    prep_t pr = transport_prepare(&c, ld(view), ld(n), ld(face));
    v3 v = transport_eval_mask(&c, ld(view), ld(dir), mask, &pr);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
    out[3] = transport_pdf_mask(&c, ld(view), ld(dir), mask, &pr);
}

/* ======================================================================
 * Emitters: crates/vork-render/shaders/emitter.slang:78-150 (rectangle,
 * position.w == 3, not an ellipse, no window or profile).
 * ==================================================================== */
typedef struct { v3 radiance; float pdf, distance; } emit_eval;
static emit_eval rect_eval(const vh_rect_light* l, v3 world, v3 direction) {
    emit_eval r;
    memset(&r, 0, sizeof(r));
    v3 center = sub(ld(l->position), world);
    if (dot(ld(l->normal), neg(center)) <= 0.f) return r;
    v3 u = scl(ld(l->u), .5f), v = scl(ld(l->v), .5f);
    v3 normal = cross(u, v);
    float square = dot(normal, normal);
    if (square <= 0.f) return r;
    float denominator = dot(normal, direction);
    if (denominator == 0.f) return r;
    float distance = dot(normal, center) / denominator;
    if (distance <= 0.f) return r;
    v3 q = sub(scl(direction, distance), center);
    float uu = dot(u, u), uv = dot(u, v), vv = dot(v, v), qu = dot(q, u), qv = dot(q, v);
    float cx = (qu * vv - qv * uv) / square, cy = (qv * uu - qu * uv) / square;
    if (fabsf(cx) > 1.f || fabsf(cy) > 1.f) return r;
    float area = 4.f * sqrtf(square);
    r.pdf = distance * distance / (area * fabsf(denominator) / sqrtf(square));
    r.radiance = dvs(ld(l->radiance), PI_F * area);
    r.distance = distance;
    return r;
}
typedef struct { v3 direction; float distance; v3 radiance; float pdf; } emit_sample;
static emit_sample rect_sample(const vh_rect_light* l, v3 world, float u0, float u1) {
    emit_sample s;
    v3 point = ld(l->position);
    point = add(point, add(scl(ld(l->u), u0 - .5f), scl(ld(l->v), u1 - .5f)));
    s.distance = len(sub(point, world));
    s.direction = s.distance > 0.f ? dvs(sub(point, world), s.distance) : mk(0.f, 0.f, 0.f);
    emit_eval e = rect_eval(l, world, s.direction);
    s.radiance = e.radiance; s.pdf = e.pdf;
    return s;
}
void vo_rect_sample(const vh_rect_light* l, const float world[3], const float u[2], float out[8]) {
    emit_sample s = rect_sample(l, ld(world), u[0], u[1]);
    out[0] = s.direction.x; out[1] = s.direction.y; out[2] = s.direction.z; out[3] = s.distance;
    out[4] = s.radiance.x; out[5] = s.radiance.y; out[6] = s.radiance.z; out[7] = s.pdf;
}
void vo_rect_eval(const vh_rect_light* l, const float world[3], const float dir[3], float out[8]) {
    emit_eval e = rect_eval(l, ld(world), ld(dir));
    out[0] = dir[0]; out[1] = dir[1]; out[2] = dir[2]; out[3] = e.distance;
    out[4] = e.radiance.x; out[5] = e.radiance.y; out[6] = e.radiance.z; out[7] = e.pdf;
}

/* ======================================================================
 * Integrator: path_pixel (path.slang:1792-1969), path_camera_ray
 * (path.slang:1722-1775), integrate (path.slang:1313-1710), next_event
 * (path.slang:851-952), surface offsets (vork_prelude.slang:162-181),
 * material.slang:69-79, hit_geometry.slang:120-135, clamp_sample
 * (sample_clamp.slang:7-16).
 * ==================================================================== */
/* vork_prelude.slang:155-160 */
static inline float next_float(float value, int up) {
    if (value == 0.f) return u2f(up ? 1u : 0x80000001u);
    int increasing = (value > 0.f) == up;
    return u2f(f2u(value) + (increasing ? 1u : 0xffffffffu));
}
/* vork_prelude.slang:170-181 */
static v3 surface_ray_origin(v3 world, v3 error, v3 n, v3 direction) {
    float scale = fmaxf(fabsf(n.x), fmaxf(fabsf(n.y), fabsf(n.z)));
    if (scale == 0.f) return world;
    n = dvs(n, scale);
    float amount = next_float(dot(vabs(n), error) / dot(n, n), 1);
    v3 offset = scl(n, amount * (dot(n, direction) >= 0.f ? 1.f : -1.f));
    v3 origin = add(world, offset);
    float o[3] = {origin.x, origin.y, origin.z}, of[3] = {offset.x, offset.y, offset.z};
    for (int k = 0; k < 3; k++)
        if (of[k] != 0.f) o[k] = next_float(o[k], of[k] > 0.f);
    return mk(o[0], o[1], o[2]);
}
/* vork_prelude.slang:165-169: direction xyz and remaining distance */
static void surface_ray_connection(v3 origin, v3 world, v3 direction, float distance, v3* dir_out, float* dist_out) {
    if (distance == FLT_MAX || (origin.x == world.x && origin.y == world.y && origin.z == world.z)) {
        *dir_out = direction; *dist_out = distance; return;
    }
    v3 delta = sub(scl(direction, distance), sub(origin, world));
    float remaining = len(delta);
    *dir_out = remaining > 0.f ? dvs(delta, remaining) : direction;
    *dist_out = remaining;
}
/* material.slang:69-79 with the instance's local bounds */
static v3 surface_position_error(const vo_scene* s, uint32_t inst, v3 world) {
    const mesh_t* m = &s->meshes[s->inst[inst].blas];
    const float* model = s->inst[inst].transform;
    const float* inv = s->w2o + 12 * inst;
    v3 lo = ld(m->lo), hi = ld(m->hi);
    v3 magnitude = mk(fmaxf(fabsf(lo.x), fabsf(hi.x)), fmaxf(fabsf(lo.y), fabsf(hi.y)), fmaxf(fabsf(lo.z), fabsf(hi.z)));
    v3 span = vabs(sub(hi, lo));
    float extent = fmaxf(span.x, fmaxf(span.y, span.z));
    v3 aw = vabs(world);
    v3 li = mk(fabsf(inv[0]) * aw.x + fabsf(inv[1]) * aw.y + fabsf(inv[2]) * aw.z + fabsf(inv[3]),
               fabsf(inv[4]) * aw.x + fabsf(inv[5]) * aw.y + fabsf(inv[6]) * aw.z + fabsf(inv[7]),
               fabsf(inv[8]) * aw.x + fabsf(inv[9]) * aw.y + fabsf(inv[10]) * aw.z + fabsf(inv[11]));
    v3 local_error = scl(add(add(magnitude, mk(extent, extent, extent)), li), FP32_GAMMA7);
    v3 a = mk(fabsf(model[0]) * local_error.x + fabsf(model[1]) * local_error.y + fabsf(model[2]) * local_error.z,
              fabsf(model[4]) * local_error.x + fabsf(model[5]) * local_error.y + fabsf(model[6]) * local_error.z,
              fabsf(model[8]) * local_error.x + fabsf(model[9]) * local_error.y + fabsf(model[10]) * local_error.z);
    v3 b = mk(fabsf(model[0]) * magnitude.x + fabsf(model[1]) * magnitude.y + fabsf(model[2]) * magnitude.z + fabsf(model[3]),
              fabsf(model[4]) * magnitude.x + fabsf(model[5]) * magnitude.y + fabsf(model[6]) * magnitude.z + fabsf(model[7]),
              fabsf(model[8]) * magnitude.x + fabsf(model[9]) * magnitude.y + fabsf(model[10]) * magnitude.z + fabsf(model[11]));
    return add(a, scl(b, FP32_GAMMA7));
}
/* stream.slang:416-427 */
static v3 finite_unit_normal(v3 normal, v3 fallback) {
    float scale = fmaxf(fabsf(normal.x), fmaxf(fabsf(normal.y), fabsf(normal.z)));
    if (finite3(normal) && scale > 0.f) return nrm(dvs(normal, scale));
    float fs = fmaxf(fabsf(fallback.x), fmaxf(fabsf(fallback.y), fabsf(fallback.z)));
    if (finite3(fallback) && fs > 0.f) return nrm(dvs(fallback, fs));
    return mk(0.f, 0.f, 1.f);
}
/* path.slang:1727-1775 (pinhole, no shared jitter or lens) */
static void camera_ray(const vo_scene* s, const vh_render_params* p, uint32_t px, uint32_t py, const float cs[4], v3* origin,
                       v3* direction, float* tmax) {
    const vh_camera* c = &s->camera;
    float ndcx = ((float)px + .5f + vo_filter_offset(s->filter, cs[0])) / (float)p->width * 2.f + -1.f;
    float ndcy = ((float)py + .5f + vo_filter_offset(s->filter, cs[1])) / (float)p->height * -2.f + 1.f;
    v3 forward = ld(c->forward), right = ld(c->right), up = ld(c->up);
    v3 d = nrm(add(add(forward, scl(right, ndcx * c->tan_half_x)), scl(up, ndcy * c->tan_half_y)));
    float fd = dot(d, forward);
    float near_distance = c->clip_start / fd, far_distance = c->clip_end / fd;
    v3 eye = ld(c->eye);
    *origin = add(eye, scl(d, near_distance));
    *direction = d;
    *tmax = fmaxf(0.f, far_distance - len(sub(*origin, eye)));
}
void vo_camera_ray(const vo_scene* s, const vh_render_params* p, uint32_t px, uint32_t py, uint32_t index, float out[8]) {
    float cs[4];
    vo_sample_group(px, py, index, p->samples, p->seed, 0, cs);
    v3 o, d; float tmax;
    camera_ray(s, p, px, py, cs, &o, &d, &tmax);
    out[0] = o.x; out[1] = o.y; out[2] = o.z; out[3] = tmax;
    out[4] = d.x; out[5] = d.y; out[6] = d.z; out[7] = 0.f;
}
/* Escaping rays: the environment of a constant or procedural sky. */
static v3 sky_radiance(const vh_sky* k, v3 d) {
    if (k->kind == VH_SKY_CONSTANT) return ld(k->color);
    float y = d.y;
    v3 base;
    if (y >= 0.f) base = lerp3(ld(k->color), ld(k->zenith), sqrtf(y));
    else base = lerp3(ld(k->color), ld(k->ground), fminf(1.f, -4.f * y));
    float cs = fmaxf(dot(d, ld(k->sun_direction)), 0.f);
    return add(base, scl(ld(k->sun_radiance), powf(cs, k->sun_exponent)));
}
/* sample_clamp.slang:7-16 */
static v3 clamp_sample(v3 value, float limit, uint32_t* dropped) {
    float brightness = dot(vabs(value), mk(1.f / 3.f, 1.f / 3.f, 1.f / 3.f));
    if (!isfinite(brightness)) { (*dropped)++; value = mk(0.f, 0.f, 0.f); brightness = 0.f; }
    return scl(value, (limit > 0.f && brightness > limit) ? limit / fmaxf(brightness, 1e-20f) : 1.f);
}

typedef struct {
    v3 world, face, smooth; /* face and smooth normals turned to the incoming side */
    uint32_t material;
} surface_t;
/* hit_geometry.slang:120-135 and the double-sided facing of a hit */
static surface_t reconstruct(const vo_scene* s, const hitrec* h, v3 direction) {
    surface_t r;
    const vh_instance* in = &s->inst[h->inst];
    const mesh_t* m = &s->meshes[in->blas];
    const uint32_t* ix = m->idx + 3 * h->prim;
    v3 p0 = ld(m->pos + 3 * ix[0]), p1 = ld(m->pos + 3 * ix[1]), p2 = ld(m->pos + 3 * ix[2]);
    float b1 = h->u, b2 = h->v, b0 = 1.f - b1 - b2;
    v3 p = add(add(scl(p0, b0), scl(p1, b1)), scl(p2, b2));
    r.world = xform_point(in->transform, p);
    v3 face_o = cross(sub(p1, p0), sub(p2, p0));
    v3 n0 = m->nrm ? ld(m->nrm + 3 * ix[0]) : face_o;
    v3 n1 = m->nrm ? ld(m->nrm + 3 * ix[1]) : face_o;
    v3 n2 = m->nrm ? ld(m->nrm + 3 * ix[2]) : face_o;
    face_o = finite_unit_normal(face_o, n0);
    const float* nm = s->n2w + 9 * h->inst;
    v3 face = finite_unit_normal(mk(nm[0] * face_o.x + nm[1] * face_o.y + nm[2] * face_o.z,
                                    nm[3] * face_o.x + nm[4] * face_o.y + nm[5] * face_o.z,
                                    nm[6] * face_o.x + nm[7] * face_o.y + nm[8] * face_o.z),
                                 face_o);
    v3 ns = add(add(scl(n0, b0), scl(n1, b1)), scl(n2, b2));
    ns = mk(nm[0] * ns.x + nm[1] * ns.y + nm[2] * ns.z, nm[3] * ns.x + nm[4] * ns.y + nm[5] * ns.z,
            nm[6] * ns.x + nm[7] * ns.y + nm[8] * ns.z);
    ns = finite_unit_normal(ns, face);
    if (dot(ns, face) < 0.f) ns = neg(ns);
    if (dot(face, direction) > 0.f) { face = neg(face); ns = neg(ns); }
    r.face = face;
    r.smooth = ns;
    uint32_t local = m->mat ? m->mat[h->prim] : 0u;
    r.material = in->material + local;
    if (r.material >= s->material_count) r.material = s->material_count ? s->material_count - 1 : 0;
    return r;
}
/* path.slang:772-781 (no shadow-terminator offset, no hair) */
static v3 path_origin(const surface_t* sf, v3 error, v3 direction) {
    int reflection = dot(sf->smooth, direction) > 0.f;
    return surface_ray_origin(sf->world, error, reflection ? sf->smooth : sf->face, direction);
}

typedef struct { uint64_t ext, shadow; uint32_t dropped; } counters_t;

static int shadow_visible(const vo_scene* s, v3 origin, v3 dir, float dist, counters_t* cnt) {
    hitrec h;
    cnt->shadow++;
    return !scene_trace(s, origin, dir, 0.f, dist, VH_RAY_SHADOW, &h, 1, 0);
}
/* path.slang:851-952 (non-propagation branch, rectangle lights, no media,
 * no hair, no subsurface) */
static v3 next_event(const vo_scene* s, const vh_render_params* p, const closure_t* c, v3 view, const surface_t* sf, v3 error,
                     const prep_t* pr, const bsdf_sample* sample, sampler_t smp, uint32_t vertex, uint32_t* rng, counters_t* cnt) {
    v3 zero = mk(0.f, 0.f, 0.f);
    if (!s->light_count) return zero;
    float u[4];
    group_values(smp, 1u + vertex * 5u + 1u, u);
    float probability;
    float wp[3] = {sf->world.x, sf->world.y, sf->world.z};
    uint32_t i = vo_light_pick(s, wp, u[2], u[3], &probability);
    if (probability <= 0.f) return zero;
    const vh_rect_light* l = &s->lights[i];
    emit_sample src = rect_sample(l, sf->world, u[0], u[1]);
    float light_pdf = src.pdf;
    v3 power = src.pdf > 0.f ? dvs(src.radiance, src.pdf) : zero;
    v3 response = analytic_mis_response(c, view, src.direction, sf->face, light_pdf, pr);
    v3 value = dvs(mul(response, power), probability);
    float threshold = p->light_threshold;
    float importance = maxc(value);
    if (importance < threshold && threshold > 0.f) {
        float keep = importance / threshold;
        if (vo_path_random(rng) >= keep) value = zero; else value = dvs(value, keep);
    }
    if (value.x != 0.f || value.y != 0.f || value.z != 0.f) {
        v3 origin = path_origin(sf, error, src.direction);
        v3 cd; float cdist;
        surface_ray_connection(origin, sf->world, src.direction, src.distance, &cd, &cdist);
        if (!shadow_visible(s, origin, cd, cdist, cnt)) value = zero;
    }
    int continuation = sample->event != 0 && event_hemisphere(sample->event, sample->direction, sf->face, sf->smooth);
    if (continuation && dot(sample->direction, sample->direction) > 0.f) {
        emit_eval e = rect_eval(l, sf->world, sample->direction);
        if (e.pdf > 0.f) {
            float weight = sample->pdf > 0.f ? mis_power(sample->pdf, e.pdf) : 1.f;
            v3 second = scl(mul(sample->weight, e.radiance), weight / probability);
            float imp2 = maxc(second);
            if (imp2 < threshold && threshold > 0.f) {
                float keep = imp2 / threshold;
                if (vo_path_random(rng) >= keep) second = zero; else second = dvs(second, keep);
            }
            if (second.x != 0.f || second.y != 0.f || second.z != 0.f) {
                v3 origin = path_origin(sf, error, sample->direction);
                v3 cd; float cdist;
                surface_ray_connection(origin, sf->world, sample->direction, e.distance, &cd, &cdist);
                if (!shadow_visible(s, origin, cd, cdist, cnt)) second = zero;
            }
            value = add(value, second);
        }
    }
    return value;
}

/* One sample's RGBA value: (radiance, 1), or zero when non-finite
 * (path.slang:1948-1952).  With `guide` (8 floats) also the camera hit's
 * denoiser guides: albedo, 1, shading normal, depth t / (t + 1); a miss
 * keeps zeros and depth 1, a non-finite sample zeros them all. */
static v3 sample_radiance(const vo_scene* s, const vh_render_params* p, uint32_t px, uint32_t py, uint32_t index,
                          counters_t* cnt, float* alpha, float* guide) {
    if (guide) { memset(guide, 0, 8 * sizeof(float)); guide[7] = 1.f; }
    sampler_t smp = make_sampler(px, py, index, p->samples, p->seed);
    uint32_t rng = vo_screen_ray_hash(px, py, index, p->seed);
    float cs[4];
    group_values(smp, 0, cs);
    v3 origin, direction; float cam_tmax;
    camera_ray(s, p, px, py, cs, &origin, &direction, &cam_tmax);
    v3 throughput = mk(1.f, 1.f, 1.f), radiance = mk(0.f, 0.f, 0.f);
    uint32_t ray_event = VH_RAY_CAMERA, diffuse = 0, glossy = 0, dropped = 0;
    int nonfinite = 0;
    uint32_t vertex = 0;
    for (uint32_t bounce = 0; bounce <= p->max_bounces; vertex++) {
        hitrec h;
        float tmax = ray_event == VH_RAY_CAMERA ? cam_tmax : FLT_MAX;
        cnt->ext++;
        float limit = bounce == 0 ? p->clamp_direct : p->clamp_indirect;
        if (!scene_trace(s, origin, direction, 0.f, tmax, ray_event, &h, 0, 0)) {
            radiance = add(radiance, clamp_sample(mul(throughput, sky_radiance(&s->sky, direction)), limit, &dropped));
            break;
        }
        surface_t sf = reconstruct(s, &h, direction);
        const vh_material* mat = &s->materials[sf.material];
        closure_t c = make_closure(mat, sf.smooth);
        if (guide && vertex == 0) {
            v3 a = guide_albedo(&c, neg(direction));
            guide[0] = a.x; guide[1] = a.y; guide[2] = a.z; guide[3] = 1.f;
            guide[4] = sf.smooth.x; guide[5] = sf.smooth.y; guide[6] = sf.smooth.z; guide[7] = h.t / (h.t + 1.f);
        }
        v3 error = surface_position_error(s, h.inst, sf.world);
        float scatter[4];
        group_values(smp, 1u + vertex * 5u + 0u, scatter);
        v3 view = neg(direction);
        v3 interface_face = dot(sf.smooth, view) > 0.f ? sf.smooth : sf.face;
        prep_t pr = transport_prepare(&c, view, sf.smooth, interface_face);
        bsdf_sample bs = transport_sample(&c, view, scatter[0], scatter[1], scatter[2], &pr);
        if (!finite3(bs.direction) || !finite3(bs.weight) || !isfinite(bs.pdf)) { nonfinite = 1; break; }
        v3 ne = next_event(s, p, &c, view, &sf, error, &pr, &bs, smp, vertex, &rng, cnt);
        radiance = add(radiance, clamp_sample(mul(throughput, ne), limit, &dropped));
        if (bounce == p->max_bounces) break;
        uint32_t event = bs.event;
        ray_event = event;
        if (event == VH_RAY_DIFFUSE) { if (diffuse++ >= p->diffuse_bounces) break; }
        else if (event == VH_RAY_SPECULAR) { if (glossy++ >= p->glossy_bounces || (diffuse > 0 && (p->caustics & 1u) == 0)) break; }
        else break;
        if (!event_hemisphere(event, bs.direction, sf.face, sf.smooth)) break;
        throughput = mul(throughput, bs.weight);
        if (!finite3(throughput)) { nonfinite = 1; break; }
        if (maxc(throughput) <= 0.f) break;
        if (bounce >= ROULETTE_START) {
            float survive = clampf(maxc(throughput), ROULETTE_MIN, ROULETTE_MAX);
            float d[4];
            group_values(smp, 1u + vertex * 5u + 4u, d);
            if (d[2] > survive) break;
            throughput = dvs(throughput, survive);
        }
        direction = bs.direction;
        origin = path_origin(&sf, error, direction);
        bounce++;
    }
    cnt->dropped += dropped;
    if (!finite3(radiance)) nonfinite = 1;
    *alpha = nonfinite ? 0.f : 1.f;
    if (guide && nonfinite) memset(guide, 0, 8 * sizeof(float));
    return nonfinite ? mk(0.f, 0.f, 0.f) : radiance;
}

void vo_sample_radiance(const vo_scene* s, const vh_render_params* p, uint32_t px, uint32_t py, uint32_t index, float out[4]) {
    counters_t cnt = {0, 0, 0};
    float alpha;
    v3 r = sample_radiance(s, p, px, py, index, &cnt, &alpha, NULL);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = alpha;
}

void vo_render(const vo_scene* s, const vh_render_params* p, float* film32, double* film64, int threads) {
    vo_render_guides(s, p, film32, film64, NULL, threads);
}

/* vo_render that also sums every sample's guides in sample order into
 * `guides32` (8 floats per pixel) when it is not NULL. */
void vo_render_guides(const vo_scene* s, const vh_render_params* p, float* film32, double* film64, float* guides32,
                      int threads) {
    int64_t pixels = (int64_t)p->width * p->height;
    uint64_t ext = 0, shadow = 0, dropped = 0;
    vo_scene* ms = (vo_scene*)s;
    if (vo_dielectric_albedo(0.5f, 0.5f, 1.5f) < 0.f) return; /* warm the tables outside the loop */
    ensure_moments();
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : ext, shadow, dropped)
    for (int64_t i = 0; i < pixels; i++) {
        uint32_t px = (uint32_t)(i % p->width), py = (uint32_t)(i / p->width);
        counters_t cnt = {0, 0, 0};
        float acc[4];
        double acc64[4] = {0, 0, 0, 0};
        if (film32) memcpy(acc, film32 + 4 * i, sizeof(acc));
        for (uint32_t k = p->sample_begin; k < p->sample_end; k++) {
            float alpha, g[8];
            v3 r = sample_radiance(s, p, px, py, k, &cnt, &alpha, guides32 ? g : NULL);
            if (guides32)
                for (int q = 0; q < 8; q++) guides32[8 * i + q] += g[q];
            if (film32) { acc[0] += r.x; acc[1] += r.y; acc[2] += r.z; acc[3] += alpha; }
            acc64[0] += r.x; acc64[1] += r.y; acc64[2] += r.z; acc64[3] += alpha;
        }
        if (film32) memcpy(film32 + 4 * i, acc, sizeof(acc));
        if (film64)
            for (int c = 0; c < 4; c++) film64[4 * i + c] += acc64[c];
        ext += cnt.ext; shadow += cnt.shadow; dropped += cnt.dropped;
    }
    ms->stats[0] = (uint64_t)pixels * (p->sample_end - p->sample_begin);
    ms->stats[1] = ext; ms->stats[2] = shadow; ms->stats[3] = dropped;
}
/* vo_render over a list of pixels: out[4 i..4 i+3] = the sum of samples
 * [sample_begin, sample_end) of pixel pixels[i] (row-major index), each
 * sample (radiance, 1) or zero when non-finite, summed in sample order in
 * float as the film is. */
/* Camera rays as vh_ray records (origin, tmin 0, direction, tmax) for a list
 * of (row-major pixel, sample index) pairs. */
void vo_camera_rays(const vo_scene* s, const vh_render_params* p, const uint32_t* pixels, const uint32_t* indices,
                    uint32_t n, vh_ray* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        float r[8];
        vo_camera_ray(s, p, pixels[i] % p->width, pixels[i] / p->width, indices[i], r);
        out[i].origin[0] = r[0]; out[i].origin[1] = r[1]; out[i].origin[2] = r[2]; out[i].tmin = 0.f;
        out[i].direction[0] = r[4]; out[i].direction[1] = r[5]; out[i].direction[2] = r[6]; out[i].tmax = r[3];
    }
}

void vo_render_pixels(const vo_scene* s, const vh_render_params* p, const uint32_t* pixels, uint32_t n, float* out,
                      int threads) {
    if (vo_dielectric_albedo(0.5f, 0.5f, 1.5f) < 0.f) return; /* warm the tables outside the loop */
    ensure_moments();
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint32_t px = pixels[i] % p->width, py = pixels[i] / p->width;
        counters_t cnt = {0, 0, 0};
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        for (uint32_t k = p->sample_begin; k < p->sample_end; k++) {
            float alpha;
            v3 r = sample_radiance(s, p, px, py, k, &cnt, &alpha, NULL);
            acc[0] += r.x; acc[1] += r.y; acc[2] += r.z; acc[3] += alpha;
        }
        memcpy(out + 4 * i, acc, sizeof(acc));
    }
}

/* One sample value per (pixel, sample index) pair: out[4 i..4 i+3]. */
void vo_sample_radiance_many(const vo_scene* s, const vh_render_params* p, const uint32_t* pixels,
                             const uint32_t* indices, uint32_t n, float* out, int threads) {
    if (vo_dielectric_albedo(0.5f, 0.5f, 1.5f) < 0.f) return;
    ensure_moments();
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        counters_t cnt = {0, 0, 0};
        float alpha;
        v3 r = sample_radiance(s, p, pixels[i] % p->width, pixels[i] / p->width, indices[i], &cnt, &alpha, NULL);
        out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = alpha;
    }
}
void vo_render_stats(const vo_scene* s, uint64_t out[4]) { memcpy(out, s->stats, sizeof(s->stats)); }
