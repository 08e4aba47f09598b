// GPU construction of the 8-wide BVH used by both BLAS (triangles) and TLAS
// (instances).
//
//  1. primitive boxes and centroid bounds
//  2. 63-bit Morton codes of the centroids, radix-sorted with their ids
//  3. binary tree by parallel locally-ordered clustering (PLOC, Meister and
//     Bittner 2018): the clusters stay in Morton order; each round every
//     cluster finds the neighbour within +-kPlocRadius positions whose merged
//     box has the smallest surface area, mutual pairs merge into a new node
//     (its box is the union, so no refit pass is needed), and a prefix sum
//     compacts the survivors, until one cluster is left
//  4. top-down collapse into 8-wide nodes, one kernel per wide level: each
//     node opens its largest-area children until it holds 8; subtrees of at
//     most `leaf_max` primitives become leaf children whose items are written
//     contiguously per node; children are placed in octant-ordered slots
//
// Every pass, the radix sort included, is a kernel written here.
#include <cstring>

#include <algorithm>
#include <cmath>
#include <vector>

#include "vh_internal.h"

namespace vh {
namespace {


__device__ __forceinline__ uint32_t float_to_ordered(float f) {
    uint32_t u = f2u(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(uint32_t u) {
    return u2f((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_init_bounds(uint32_t* b) {
    if (threadIdx.x < 3) { b[threadIdx.x] = 0xffffffffu; b[3 + threadIdx.x] = 0u; }
}

// Centroid bounds: per-block reduction then ordered-uint atomics.
__global__ void k_centroid_bounds(const float* lo, const float* hi, uint32_t n, uint32_t* bounds) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int k = 0; k < 3; k++) {
            float c = 0.5f * (lo[3 * i + k] + hi[3 * i + k]);
            mn[k] = fminf(mn[k], c);
            mx[k] = fmaxf(mx[k], c);
        }
    }
    __shared__ float smn[3][256], smx[3][256];
    for (int k = 0; k < 3; k++) { smn[k][threadIdx.x] = mn[k]; smx[k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; k++) {
                smn[k][threadIdx.x] = fminf(smn[k][threadIdx.x], smn[k][threadIdx.x + s]);
                smx[k][threadIdx.x] = fmaxf(smx[k][threadIdx.x], smx[k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) {
            atomicMin(&bounds[k], float_to_ordered(smn[k][0]));
            atomicMax(&bounds[3 + k], float_to_ordered(smx[k][0]));
        }
}

__device__ __forceinline__ uint64_t spread21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(const float* lo, const float* hi, uint32_t n, const uint32_t* bounds, uint64_t* keys,
                         uint32_t* ids) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t q[3];
    for (int k = 0; k < 3; k++) {
        float bmin = ordered_to_float(bounds[k]), bmax = ordered_to_float(bounds[3 + k]);
        float ext = bmax - bmin;
        float c = 0.5f * (lo[3 * i + k] + hi[3 * i + k]);
        float f = ext > 0.f ? (c - bmin) / ext : 0.5f;
        f = fminf(fmaxf(f, 0.f), 1.f);
        q[k] = (uint64_t)fminf(f * 2097152.f, 2097151.f);
    }
    keys[i] = spread21(q[0]) | (spread21(q[1]) << 1) | (spread21(q[2]) << 2);
    ids[i] = i;
}

// ---------------------------------------------------------------- PLOC
// Node ids: 0..n-1 are the leaves (sorted primitive i), n..2n-2 the merged
// nodes in creation order; the root is 2n-2 (0 when n == 1).
#ifndef VH_PLOC_RADIUS
#define VH_PLOC_RADIUS 16
#endif
constexpr int kPlocRadius = VH_PLOC_RADIUS;
constexpr int kPlocBlock = 256;

__device__ __forceinline__ float half_area(const float lo[3], const float hi[3]) {
    float dx = fmaxf(hi[0] - lo[0], 0.f), dy = fmaxf(hi[1] - lo[1], 0.f), dz = fmaxf(hi[2] - lo[2], 0.f);
    return dx * dy + dy * dz + dz * dx;
}

__global__ void k_ploc_init(const float* plo, const float* phi, const uint32_t* sorted_ids, uint32_t n, float* nb,
                            uint32_t* size, uint32_t* clusters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t prim = sorted_ids[i];
    for (int k = 0; k < 3; k++) {
        nb[6 * i + k] = plo[3 * prim + k];
        nb[6 * i + 3 + k] = phi[3 * prim + k];
    }
    size[i] = 1u;
    clusters[i] = i;
}

// Nearest neighbour of each of the `count` clusters within +-kPlocRadius
// positions: smallest surface area of the merged box, ties to the lower
// position.  The block stages its clusters' boxes and the halo in LDS.
__global__ __launch_bounds__(kPlocBlock) void k_ploc_nn(const uint32_t* clusters, const uint32_t* state, const float* nb,
                                                        uint32_t* nn) {
    const uint32_t count = state[0];
    if (count <= 1) return;
    constexpr int kSpan = kPlocBlock + 2 * kPlocRadius;
    __shared__ float box[6][kSpan];
    const int64_t base = (int64_t)blockIdx.x * kPlocBlock - kPlocRadius;
    for (int k = threadIdx.x; k < kSpan; k += blockDim.x) {
        const int64_t j = base + k;
        if (j >= 0 && j < (int64_t)count) {
            const uint32_t c = clusters[j];
            for (int q = 0; q < 6; q++) box[q][k] = nb[6 * (size_t)c + q];
        }
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * kPlocBlock + threadIdx.x;
    if (i >= count) return;
    const int li = threadIdx.x + kPlocRadius;
    const float lo0 = box[0][li], lo1 = box[1][li], lo2 = box[2][li];
    const float hi0 = box[3][li], hi1 = box[4][li], hi2 = box[5][li];
    float best = INFINITY;
    uint32_t bj = i;
    for (int d = -kPlocRadius; d <= kPlocRadius; d++) {
        const int64_t j = (int64_t)i + d;
        if (d == 0 || j < 0 || j >= (int64_t)count) continue;
        const int lj = li + d;
        float lo[3] = {fminf(lo0, box[0][lj]), fminf(lo1, box[1][lj]), fminf(lo2, box[2][lj])};
        float hi[3] = {fmaxf(hi0, box[3][lj]), fmaxf(hi1, box[4][lj]), fmaxf(hi2, box[5][lj])};
        const float a = half_area(lo, hi);
        if (a < best) { best = a; bj = (uint32_t)j; }
    }
    nn[i] = bj;
}

// Per cluster position: bit 0 = the cluster survives into the next round
// (unpaired, or the lower position of a mutual pair, which becomes the
// merged node), bit 32 = it creates a merged node.
__global__ void k_ploc_flags(const uint32_t* nn, const uint32_t* state, uint64_t* flags) {
    const uint32_t count = state[0];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (count <= 1 || i >= count) return;
    const uint32_t j = nn[i];
    const bool mutual = j != i && nn[j] == i;
    const bool keep = !mutual || i < j;
    const bool owner = mutual && i < j;
    flags[i] = (uint64_t)(keep ? 1u : 0u) | ((uint64_t)(owner ? 1u : 0u) << 32);
}

// Exclusive scan of the round's flags with the count in device memory:
// k_scan_tiles scans each tile of kScanTile flags and stores its total,
// k_scan_sums turns the totals into tile offsets, k_ploc_merge adds them.
constexpr int kScanBlock = 256, kScanItems = 8, kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint64_t block_inclusive_scan(uint64_t v, uint64_t* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const uint64_t x = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += x;
        __syncthreads();
    }
    return sh[tid];
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const uint64_t* in, const uint32_t* state, uint64_t* out,
                                                          uint64_t* sums) {
    __shared__ uint64_t sh[kScanBlock];
    const uint32_t count = state[0], base = blockIdx.x * kScanTile;
    if (count <= 1 || base >= count) return;
    uint64_t v[kScanItems], t = 0;
    for (int k = 0; k < kScanItems; k++) {
        const uint32_t i = base + threadIdx.x * kScanItems + k;
        v[k] = i < count ? in[i] : 0;
        t += v[k];
    }
    const uint64_t incl = block_inclusive_scan(t, sh);
    uint64_t prefix = incl - t;
    for (int k = 0; k < kScanItems; k++) {
        const uint32_t i = base + threadIdx.x * kScanItems + k;
        if (i < count) out[i] = prefix;
        prefix += v[k];
    }
    if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = incl;
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint64_t* sums, const uint32_t* state) {
    __shared__ uint64_t sh[1024];
    const uint32_t count = state[0];
    if (count <= 1) return;
    const uint32_t tiles = (count + kScanTile - 1) / kScanTile;
    uint64_t carry = 0;
    for (uint32_t start = 0; start < tiles; start += 1024) {
        const uint32_t i = start + threadIdx.x;
        const uint64_t v = i < tiles ? sums[i] : 0;
        const uint64_t incl = block_inclusive_scan(v, sh);
        if (i < tiles) sums[i] = carry + incl - v;
        carry += sh[1023];
        __syncthreads();
    }
}

// ---- LSD radix sort of (63-bit Morton key, id) pairs, 8 bits per pass.
// Per pass: k_radix_hist counts each tile's digits; the counts, digit-major
// ([digit][tile]), go through the exclusive scan above, so digit d of tile t
// starts after every smaller digit and after digit d of the tiles before t;
// k_radix_scatter then ranks each tile's keys stably (in index order) within
// their digit and writes them there.
constexpr int kRadixBlock = 256, kRadixItems = 16, kRadixTile = kRadixBlock * kRadixItems;
constexpr int kRadixBins = 256;  // one per thread of a block
static_assert(kRadixBins == kRadixBlock, "a thread per digit");

__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const uint64_t* keys, uint32_t n, int shift,
                                                           uint64_t* counts, uint32_t tiles) {
    __shared__ uint32_t h[kRadixBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRadixTile;
    for (int r = 0; r < kRadixItems; r++) {
        const uint32_t i = base + r * kRadixBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    counts[(size_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// The tile's keys in rounds of one per thread (round r: keys base + 256 r +
// thread), so index order is (round, wave, lane).  Within a wave the lanes of
// one digit find each other by 8 ballots; a lane's rank is the count of its
// peers below it; the waves' counts per digit are laid out in wave order and
// a running slot per digit advances by each round's total.  That places every
// key at its stable position in the tile sorted by digit, in LDS; the sorted
// tile then goes out in index order, so the keys of one digit, consecutive
// in LDS, land in consecutive slots of the output (writes in runs).
__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(const uint64_t* keys, const uint32_t* vals, uint32_t n,
                                                              int shift, const uint64_t* counts,
                                                              const uint64_t* offsets, const uint64_t* tile_sums,
                                                              uint32_t tiles, uint64_t* keys_out, uint32_t* vals_out) {
    constexpr int kWaves = kRadixBlock / 64;
    __shared__ uint64_t tile_keys[kRadixTile];
    __shared__ uint32_t tile_vals[kRadixTile];
    __shared__ uint32_t out_base[kRadixBins];    // the digit's first output slot minus its first tile slot
    __shared__ uint32_t next_slot[kRadixBins];   // the digit's next slot in the sorted tile
    __shared__ uint32_t wave_slot[kWaves][kRadixBins];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t j = (size_t)tid * tiles + blockIdx.x;
    const uint32_t count = (uint32_t)counts[j];
    // exclusive scan of the tile's digit counts: each digit's first slot in the sorted tile
    wave_slot[0][tid] = count;
    __syncthreads();
    for (int off = 1; off < kRadixBins; off <<= 1) {
        const uint32_t x = tid >= off ? wave_slot[0][tid - off] : 0u;
        __syncthreads();
        wave_slot[0][tid] += x;
        __syncthreads();
    }
    const uint32_t first = wave_slot[0][tid] - count;
    next_slot[tid] = first;
    out_base[tid] = (uint32_t)(offsets[j] + tile_sums[j / kScanTile]) - first;
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * kRadixTile;
    for (int r = 0; r < kRadixItems; r++) {
        const uint32_t i = base + r * kRadixBlock + tid;
        const bool valid = i < n;
        const uint64_t key = valid ? keys[i] : 0ull;
        const uint32_t val = valid ? vals[i] : 0u;
        const uint32_t d = (uint32_t)(key >> shift) & 0xffu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t ones = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? ones : ~ones;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
#pragma unroll
        for (int w = 0; w < kWaves; w++) wave_slot[w][tid] = 0u;
        __syncthreads();
        if (valid && rank == 0) wave_slot[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t run = next_slot[tid];
#pragma unroll
            for (int w = 0; w < kWaves; w++) {
                const uint32_t c = wave_slot[w][tid];
                wave_slot[w][tid] = run;
                run += c;
            }
            next_slot[tid] = run;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = wave_slot[wave][d] + rank;
            tile_keys[pos] = key;
            tile_vals[pos] = val;
        }
        __syncthreads();  // wave_slot is cleared by the next round
    }
    const uint32_t in_tile = n - base < (uint32_t)kRadixTile ? n - base : (uint32_t)kRadixTile;
    for (uint32_t k = tid; k < in_tile; k += kRadixBlock) {
        const uint64_t key = tile_keys[k];
        const uint32_t pos = out_base[(uint32_t)(key >> shift) & 0xffu] + k;
        keys_out[pos] = key;
        vals_out[pos] = tile_vals[k];
    }
}

// Writes the merged nodes and the next round's cluster array; the last
// thread stores the round's totals (survivors, merged nodes) in state[2..3].
// ---- SAH-optimal collapse (Ylitie, Karras and Laine 2017, section 3): C(x, i)
// is the least SAH cost of the subtree at binary node x as a forest of at
// most i children of a wide node; a leaf of at most p_max primitives costs
// area * c_prim * primitives, a wide node area * c_node plus the best forest
// of 8 roots of its two subtrees.  Decisions per internal node, byte i - 1:
// kDpLeaf / kDpInternal for i = 1, else kDpFewer (the forest of i - 1 roots)
// or the number k of roots given to the left subtree; byte 8 is the split of
// the wide node's 8 roots.  A node's table is computed by the clustering
// round that creates it, from its children's tables of earlier rounds.
#ifndef VH_DP_CNODE
#define VH_DP_CNODE 1.0f
#endif
#ifndef VH_DP_CPRIM
#define VH_DP_CPRIM 0.3f
#endif
constexpr uint8_t kDpFewer = 0, kDpLeaf = 0xF0, kDpInternal = 0xF1;
constexpr uint32_t kDpStride = 9;

__device__ __forceinline__ float node_area(const float* nb, uint32_t r) {
    float lo[3] = {nb[6 * (size_t)r], nb[6 * (size_t)r + 1], nb[6 * (size_t)r + 2]};
    float hi[3] = {nb[6 * (size_t)r + 3], nb[6 * (size_t)r + 4], nb[6 * (size_t)r + 5]};
    return half_area(lo, hi);
}

// The table of internal node id = (l, r) of box area a over prims primitives.
__device__ void dp_table(const float* nb, const float* cost, uint32_t n, uint32_t l, uint32_t r, uint32_t id, float a,
                         uint32_t prims, uint32_t p_max, float* cost_out, uint8_t* dec_out) {
    float cl[8], cr[8];
    for (int s2 = 0; s2 < 2; s2++) {
        const uint32_t c = s2 ? r : l;
        float* out = s2 ? cr : cl;
        if (c < n) {
            const float v = node_area(nb, c) * VH_DP_CPRIM;
            for (int i = 0; i < 8; i++) out[i] = v;
        } else {
            const float4* src = reinterpret_cast<const float4*>(cost + (size_t)(c - n) * 8);
            const float4 u = src[0], v = src[1];
            out[0] = u.x; out[1] = u.y; out[2] = u.z; out[3] = u.w;
            out[4] = v.x; out[5] = v.y; out[6] = v.z; out[7] = v.w;
        }
    }
    float dist[9];
    uint8_t arg[9];
    for (int jj = 2; jj <= 8; jj++) {
        float best = INFINITY;
        uint8_t bk = 1;
        for (int k = 1; k < jj; k++) {
            const float v = cl[k - 1] + cr[jj - k - 1];
            if (v < best) { best = v; bk = (uint8_t)k; }
        }
        dist[jj] = best;
        arg[jj] = bk;
    }
    const float c_leaf = prims <= p_max ? a * VH_DP_CPRIM * (float)prims : INFINITY;
    const float c_int = dist[8] + a * VH_DP_CNODE;
    float co[8];
    uint8_t de[kDpStride];
    float prev = fminf(c_leaf, c_int);
    co[0] = prev;
    de[0] = c_leaf <= c_int ? kDpLeaf : kDpInternal;
    for (int i = 2; i <= 8; i++) {
        if (dist[i] < prev) {
            prev = dist[i];
            de[i - 1] = arg[i];
        } else {
            de[i - 1] = kDpFewer;
        }
        co[i - 1] = prev;
    }
    de[8] = arg[8];
    float4* dst = reinterpret_cast<float4*>(cost_out + (size_t)(id - n) * 8);
    dst[0] = make_float4(co[0], co[1], co[2], co[3]);
    dst[1] = make_float4(co[4], co[5], co[6], co[7]);
    for (int i = 0; i < (int)kDpStride; i++) dec_out[(size_t)(id - n) * kDpStride + i] = de[i];
}

__global__ void k_ploc_merge(const uint32_t* clusters, const uint32_t* nn, uint32_t* state, const uint64_t* flags,
                             const uint64_t* offsets, const uint64_t* tile_offsets, uint32_t n, float* nb,
                             uint32_t* left, uint32_t* right, uint32_t* size, uint32_t* next, float* dp_cost,
                             uint8_t* dp_dec, uint32_t p_max) {
    const uint32_t count = state[0], node_base = state[1];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (count <= 1 || i >= count) return;
    const uint64_t f = flags[i], o = offsets[i] + tile_offsets[i / kScanTile];
    if (f & 1u) {
        uint32_t id = clusters[i];
        if (f >> 32) {
            const uint32_t a = id, b = clusters[nn[i]];
            id = node_base + (uint32_t)(o >> 32);
            float lo[3], hi[3];
            for (int k = 0; k < 3; k++) {
                lo[k] = fminf(nb[6 * (size_t)a + k], nb[6 * (size_t)b + k]);
                hi[k] = fmaxf(nb[6 * (size_t)a + 3 + k], nb[6 * (size_t)b + 3 + k]);
                nb[6 * (size_t)id + k] = lo[k];
                nb[6 * (size_t)id + 3 + k] = hi[k];
            }
            left[id - n] = a;
            right[id - n] = b;
            const uint32_t prims = size[a] + size[b];
            size[id] = prims;
            dp_table(nb, dp_cost, n, a, b, id, half_area(lo, hi), prims, p_max, dp_cost, dp_dec);
        }
        next[(uint32_t)o] = id;
    }
    if (i == count - 1) {
        state[2] = (uint32_t)o + (uint32_t)(f & 1u);
        state[3] = (uint32_t)(o >> 32) + (uint32_t)(f >> 32);
    }
}

// The last rounds, once at most kTailMax clusters are left, in one block of
// kTailMax threads: the clusters' ids and boxes stay in LDS between rounds,
// and each round is the rounds above (k_ploc_nn, k_ploc_flags, the scan,
// k_ploc_merge, k_ploc_advance) with the same neighbours and node ids, with
// barriers for launches.  Merged nodes' boxes, sizes and tables go to global
// memory as k_ploc_merge writes them; later rounds of the block read them
// back after a barrier (one block shares its CU's caches).
constexpr uint32_t kTailMax = 1024;
__global__ __launch_bounds__(kTailMax) void k_ploc_tail(const uint32_t* clusters, uint32_t* state, uint32_t n,
                                                       float* nb, uint32_t* left, uint32_t* right, uint32_t* size,
                                                       float* dp_cost, uint8_t* dp_dec, uint32_t p_max,
                                                       uint32_t n_nodes) {
    __shared__ float box[6][kTailMax];
    __shared__ uint32_t cl[kTailMax], nnb[kTailMax];
    __shared__ uint64_t sh[kTailMax];
    const uint32_t tid = threadIdx.x;
    uint32_t count = state[0], node_base = state[1];
    if (count <= 1 || count > kTailMax) return;
    if (tid < count) {
        const uint32_t c = clusters[tid];
        cl[tid] = c;
        for (int q = 0; q < 6; q++) box[q][tid] = nb[6 * (size_t)c + q];
    }
    __syncthreads();
    bool failed = false;
    while (count > 1) {
        if (tid < count) {
            const float lo0 = box[0][tid], lo1 = box[1][tid], lo2 = box[2][tid];
            const float hi0 = box[3][tid], hi1 = box[4][tid], hi2 = box[5][tid];
            float best = INFINITY;
            uint32_t bj = tid;
            for (int d = -kPlocRadius; d <= kPlocRadius; d++) {
                const int64_t j = (int64_t)tid + d;
                if (d == 0 || j < 0 || j >= (int64_t)count) continue;
                float lo[3] = {fminf(lo0, box[0][j]), fminf(lo1, box[1][j]), fminf(lo2, box[2][j])};
                float hi[3] = {fmaxf(hi0, box[3][j]), fmaxf(hi1, box[4][j]), fmaxf(hi2, box[5][j])};
                const float a = half_area(lo, hi);
                if (a < best) { best = a; bj = (uint32_t)j; }
            }
            nnb[tid] = bj;
        }
        __syncthreads();
        uint64_t f = 0;
        if (tid < count) {
            const uint32_t j = nnb[tid];
            const bool mutual = j != tid && nnb[j] == tid;
            f = (uint64_t)((!mutual || tid < j) ? 1u : 0u) | ((uint64_t)((mutual && tid < j) ? 1u : 0u) << 32);
        }
        const uint64_t o = block_inclusive_scan(f, sh) - f;
        const uint64_t total = sh[kTailMax - 1];
        uint32_t id = 0;
        float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};
        if (f & 1u) {
            id = cl[tid];
            for (int k = 0; k < 3; k++) { lo[k] = box[k][tid]; hi[k] = box[3 + k][tid]; }
            if (f >> 32) {
                const uint32_t bp = nnb[tid], a = id, b = cl[bp];
                id = node_base + (uint32_t)(o >> 32);
                for (int k = 0; k < 3; k++) {
                    lo[k] = fminf(lo[k], box[k][bp]);
                    hi[k] = fmaxf(hi[k], box[3 + k][bp]);
                    nb[6 * (size_t)id + k] = lo[k];
                    nb[6 * (size_t)id + 3 + k] = hi[k];
                }
                left[id - n] = a;
                right[id - n] = b;
                const uint32_t prims = size[a] + size[b];
                size[id] = prims;
                dp_table(nb, dp_cost, n, a, b, id, half_area(lo, hi), prims, p_max, dp_cost, dp_dec);
            }
        }
        __syncthreads();  // every read of this round's LDS is done
        if (f & 1u) {
            cl[(uint32_t)o] = id;
            for (int k = 0; k < 3; k++) { box[k][(uint32_t)o] = lo[k]; box[3 + k][(uint32_t)o] = hi[k]; }
        }
        const uint32_t survivors = (uint32_t)total, merged = (uint32_t)(total >> 32);
        if (survivors >= count || merged == 0 || node_base + merged > n_nodes) {
            failed = true;
            break;
        }
        count = survivors;
        node_base += merged;
        __syncthreads();
    }
    if (tid == 0) {
        state[0] = failed ? 1u : count;
        state[1] = node_base;
        if (failed) state[4] = 1u;
    }
}

// Next round: survivors become the clusters, merged nodes advance the node
// base; a round that merged nothing sets the error flag and stops the rounds.
__global__ void k_ploc_advance(uint32_t* state, uint32_t n_nodes) {
    if (threadIdx.x != 0) return;
    const uint32_t count = state[0];
    if (count <= 1) return;
    const uint32_t survivors = state[2], merged = state[3];
    if (survivors >= count || merged == 0 || state[1] + merged > n_nodes) {
        state[4] = 1u;
        state[0] = 1u;
        return;
    }
    state[0] = survivors;
    state[1] += merged;
}

struct Tree {
    const uint32_t *left, *right, *size, *sorted_ids;
    const float* nb;
    uint32_t n;
    const uint8_t* dp_dec;     // per internal node, the decisions of dp_table
};

__device__ __forceinline__ bool is_leaf(const Tree& t, uint32_t r) { return r < t.n; }
__device__ __forceinline__ void ref_box(const Tree& t, uint32_t r, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; k++) { lo[k] = t.nb[6 * (size_t)r + k]; hi[k] = t.nb[6 * (size_t)r + 3 + k]; }
}
// Primitives under a subtree of at most 4 of them, left to right.
__device__ __forceinline__ uint32_t gather_items(const Tree& t, uint32_t r, uint32_t* out) {
    uint32_t st[4];
    int sp = 0;
    uint32_t c = 0;
    st[sp++] = r;
    while (sp > 0) {
        const uint32_t x = st[--sp];
        if (is_leaf(t, x)) {
            out[c++] = t.sorted_ids[x];
        } else if (sp + 2 <= 4) {
            st[sp++] = t.right[x - t.n];
            st[sp++] = t.left[x - t.n];
        }
    }
    return c;
}

// Quantise one axis of up to 8 child boxes against the node's lower corner,
// making sure the decoded boxes enclose the exact ones.
__device__ void quantise_axis(float origin, float node_hi, int m, const float* clo, const float* chi, uint8_t* ebyte,
                              uint8_t* qlo, uint8_t* qhi) {
    float extent = node_hi - origin;
    int e = -126;
    if (extent > 0.f) {
        int ex;
        float mant = frexpf(extent / 255.f, &ex);  // extent/255 = mant * 2^ex, mant in [0.5,1)
        e = (mant == 0.5f) ? ex - 1 : ex;
        if (e < -126) e = -126;
    }
    for (;; e++) {
        if (e > 127) e = 127;
        float scale = u2f(uint32_t(e + 127) << 23);
        bool ok = true;
        for (int i = 0; i < m; i++) {
            float a = floorf((clo[i] - origin) / scale), b = ceilf((chi[i] - origin) / scale);
            int ql = (int)fminf(fmaxf(a, 0.f), 255.f), qh = (int)fminf(fmaxf(b, 0.f), 255.f);
            while (ql > 0 && decode(origin, (uint8_t)ql, scale) > clo[i]) ql--;
            while (qh < 255 && decode(origin, (uint8_t)qh, scale) < chi[i]) qh++;
            if (decode(origin, (uint8_t)ql, scale) > clo[i] || decode(origin, (uint8_t)qh, scale) < chi[i]) {
                ok = false;
                break;
            }
            qlo[i] = (uint8_t)ql;
            qhi[i] = (uint8_t)qh;
        }
        if (ok || e == 127) { *ebyte = (uint8_t)(e + 127); return; }
    }
}

struct QueueItem { uint32_t binary; uint32_t wide; };

// A subtree becomes a wide node unless the collapse chose a leaf for it; a
// leaf never holds more than leaf_max primitives, whatever the table says.
__device__ __forceinline__ bool dp_internal(const Tree& t, uint32_t r, uint32_t leaf_max) {
    return !is_leaf(t, r) && (t.size[r] > leaf_max || t.dp_dec[(size_t)(r - t.n) * kDpStride] != kDpLeaf);
}

// The roots of the best forest of a wide node at internal binary node r.
__device__ __forceinline__ int dp_children(const Tree& t, uint32_t r, uint32_t* c) {
    uint32_t sn[16], si[16];
    int sp = 0, m = 0;
    const uint8_t* d0 = t.dp_dec + (size_t)(r - t.n) * kDpStride;
    const uint32_t k8 = d0[8] >= 1 && d0[8] <= 7 ? d0[8] : 4u;
    sn[sp] = t.right[r - t.n]; si[sp++] = 8 - k8;
    sn[sp] = t.left[r - t.n]; si[sp++] = k8;
    while (sp > 0) {
        const uint32_t x = sn[--sp];
        uint32_t i = si[sp];
        if (is_leaf(t, x)) { c[m++] = x; continue; }
        const uint8_t* d = t.dp_dec + (size_t)(x - t.n) * kDpStride;
        while (i > 1 && d[i - 1] == kDpFewer) i--;
        if (i == 1) { c[m++] = x; continue; }
        const uint32_t k = d[i - 1];
        if (k == 0 || k >= i) { c[m++] = x; continue; }
        sn[sp] = t.right[x - t.n]; si[sp++] = i - k;
        sn[sp] = t.left[x - t.n]; si[sp++] = k;
    }
    return m;
}

// The exclusive prefix of v (< 2^kBits) over the wave's lanes plus `counter`'s
// value before the wave's total was added to it: one atomic per wave, so the
// lanes of a level do not serialise on the counter.  Every lane of the wave
// takes part.
template <int kBits>
__device__ __forceinline__ uint32_t wave_alloc(uint32_t v, uint32_t* counter) {
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < kBits; b++) {
        const uint64_t m = __ballot((v >> b) & 1u);
        pre += (uint32_t)__popcll(m & lt) << b;
        tot += (uint32_t)__popcll(m) << b;
    }
    uint32_t base = 0;
    if (tot) {
        const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
        if ((int)__lane_id() == leader) base = atomicAdd(counter, tot);
        base = __shfl(base, leader);
    }
    return base + pre;
}

__global__ void k_collapse(Tree t, const QueueItem* in, const uint32_t* in_count, QueueItem* out, uint32_t* out_count,
                           uint32_t* node_counter, uint32_t* item_counter, WideNode* nodes, uint32_t* order,
                           uint32_t leaf_max, uint32_t n_prims, uint32_t* max_level, uint32_t level) {
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    // The grid follows an upper bound of the queue (the host reads the length
    // back every few levels), so whole workgroups and waves past the queue
    // leave at once (nothing in the kernel waits on a barrier); the other
    // lanes past it stay for the wave's allocations with no children.
    const uint32_t count = *in_count;
    if (blockIdx.x * blockDim.x >= count) return;
    const bool valid = idx < count;
    if (!__ballot(valid)) return;
    const QueueItem it = valid ? in[idx] : QueueItem{0u, 0u};
    uint32_t c[kWidth];
    int m = 0;
    if (!valid) {
        m = 0;
    } else if (is_leaf(t, it.binary)) {
        // a single-primitive tree: one leaf child
        c[m++] = it.binary;
    } else {
        m = dp_children(t, it.binary, c);
    }
    float clo[3][kWidth], chi[3][kWidth];
    float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < m; k++) {
        float lo[3], hi[3];
        ref_box(t, c[k], lo, hi);
        for (int a = 0; a < 3; a++) {
            clo[a][k] = lo[a]; chi[a][k] = hi[a];
            nlo[a] = fminf(nlo[a], lo[a]); nhi[a] = fmaxf(nhi[a], hi[a]);
        }
    }
    // Slot assignment: slot s should come first for rays whose direction has
    // sign octant s, i.e. hold the child that lies farthest towards -d_s with
    // d_s = (+-1, +-1, +-1) (minus where bit a of s is set).  Greedy on the
    // smallest centroid projection over all free (child, slot) pairs.
    int slot_child[kWidth];
    for (int q = 0; q < kWidth; q++) slot_child[q] = -1;
    {
        float proj[kWidth][3];
        for (int k = 0; k < m; k++)
            for (int a = 0; a < 3; a++) proj[k][a] = (clo[a][k] + chi[a][k]) - (nlo[a] + nhi[a]);
        uint32_t child_free = (1u << m) - 1u, slot_free = 0xffu;
        for (int round = 0; round < m; round++) {
            float best = INFINITY;
            int bk = -1, bq = -1;
            for (int k = 0; k < m; k++) {
                if (!((child_free >> k) & 1u)) continue;
                for (int q = 0; q < kWidth; q++) {
                    if (!((slot_free >> q) & 1u)) continue;
                    float cost = ((q & 1) ? -proj[k][0] : proj[k][0]) + ((q & 2) ? -proj[k][1] : proj[k][1]) +
                                 ((q & 4) ? -proj[k][2] : proj[k][2]);
                    if (bk < 0 || cost < best) { best = cost; bk = k; bq = q; }
                }
            }
            slot_child[bq] = bk;
            child_free &= ~(1u << bk);
            slot_free &= ~(1u << bq);
        }
    }
    WideNode node;
    for (int a = 0; a < 3; a++) node.origin[a] = nlo[a];
    uint8_t e[3];
    for (int a = 0; a < 3; a++) {
        uint8_t ql[kWidth], qh[kWidth];
        quantise_axis(nlo[a], nhi[a], m, clo[a], chi[a], &e[a], ql, qh);
        for (int q = 0; q < kWidth; q++) {
            const int k = slot_child[q];
            node.qlo[a][q] = k >= 0 ? ql[k] : 255;
            node.qhi[a][q] = k >= 0 ? qh[k] : 0;
        }
    }
    node.ex = e[0]; node.ey = e[1]; node.ez = e[2];
    uint32_t n_internal = 0, n_items = 0;
    if (valid) {
        for (int k = 0; k < m; k++) {
            if (dp_internal(t, c[k], leaf_max)) n_internal++; else n_items += t.size[c[k]];
        }
    }
    // the wave's wide children, primitives and queue entries in one atomic
    // each (a node's internal children take consecutive node ids and queue
    // slots, in slot order)
    const uint32_t child_base = wave_alloc<4>(n_internal, node_counter);
    const uint32_t item_base = wave_alloc<7>(n_items, item_counter);
    const uint32_t slot_base = wave_alloc<4>(n_internal, out_count);
    if (__ballot(valid && n_internal) && __lane_id() == (uint32_t)(__ffsll((unsigned long long)__ballot(1)) - 1))
        atomicMax(max_level, level + 1);
    if (!valid) return;
    node.child_base = n_internal ? child_base : 0u;
    node.item_base = n_items ? item_base : 0u;
    uint32_t ki = 0, off = 0, imask = 0;
    for (int q = 0; q < kWidth; q++) {
        node.meta[q] = (uint8_t)kMetaNoItems;
        const int k = slot_child[q];
        if (k < 0) continue;
        if (dp_internal(t, c[k], leaf_max)) {
            imask |= 1u << q;
            out[slot_base + ki].binary = c[k];
            out[slot_base + ki].wide = child_base + ki;
            ki++;
        } else {
            uint32_t prims[4];
            const uint32_t cnt = gather_items(t, c[k], prims);
            node.meta[q] = (uint8_t)(off | ((cnt - 1) << 5));
            for (uint32_t j = 0; j < cnt; j++) order[item_base + off + j] = prims[j];
            off += cnt;
        }
    }
    node.imask = (uint8_t)imask;
    nodes[it.wide] = node;
}

// SAH cost of the binary tree relative to its root, with a cost of 1 per
// node and per primitive: the sum over all 2n-1 nodes (leaves included) of
// area / root area.
__global__ void k_sah(const float* nb, uint32_t nodes, const float* rootb, double* out) {
    __shared__ double s[256];
    double acc = 0;
    float rlo[3] = {rootb[0], rootb[1], rootb[2]}, rhi[3] = {rootb[3], rootb[4], rootb[5]};
    double root = half_area(rlo, rhi);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nodes; i += gridDim.x * blockDim.x) {
        float lo[3] = {nb[6 * i], nb[6 * i + 1], nb[6 * i + 2]}, hi[3] = {nb[6 * i + 3], nb[6 * i + 4], nb[6 * i + 5]};
        acc += root > 0 ? half_area(lo, hi) / root : 0.0;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t k = blockDim.x / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, s[0]);
}

inline uint32_t blocks(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

}  // namespace

int build_wide_bvh(vh_device* dev, hipStream_t s, BuildScratch& B, const float* d_lo, const float* d_hi, uint32_t n,
                   uint32_t leaf_max, WideBuild* out) {
    if (n == 0) return fail(VH_ERR_USAGE, "cannot build a BVH over zero primitives");
    const uint32_t n_int = n > 1 ? n - 1 : 0;
    const size_t n_nodes = (size_t)n + n_int;
    DeviceBuffer &bounds = B.bounds, &keys = B.keys, &keys_sorted = B.keys_sorted, &ids = B.ids,
                 &ids_sorted = B.ids_sorted, &left = B.left, &right = B.right, &size = B.size, &nb = B.nb, &ca = B.ca,
                 &cb = B.cb, &nn = B.nn, &flags = B.flags, &offsets = B.offsets, &tmp = B.sort_tmp, &qa = B.qa,
                 &qb = B.qb, &counters = B.counters, &sah = B.sah, &tile_sums = B.tile_sums;
    VH_HIP(bounds.reserve(6 * sizeof(uint32_t)));
    VH_HIP(keys.reserve(sizeof(uint64_t) * n));
    VH_HIP(keys_sorted.reserve(sizeof(uint64_t) * n));
    VH_HIP(ids.reserve(sizeof(uint32_t) * n));
    VH_HIP(ids_sorted.reserve(sizeof(uint32_t) * n));
    VH_HIP(left.reserve(sizeof(uint32_t) * (n_int + 1)));
    VH_HIP(right.reserve(sizeof(uint32_t) * (n_int + 1)));
    VH_HIP(size.reserve(sizeof(uint32_t) * n_nodes));
    VH_HIP(nb.reserve(sizeof(float) * 6 * n_nodes));
    VH_HIP(ca.reserve(sizeof(uint32_t) * n));
    VH_HIP(cb.reserve(sizeof(uint32_t) * n));
    VH_HIP(nn.reserve(sizeof(uint32_t) * n));
    VH_HIP(flags.reserve(sizeof(uint64_t) * n));
    VH_HIP(offsets.reserve(sizeof(uint64_t) * n));
    VH_HIP(counters.reserve(sizeof(uint32_t) * 8));
    VH_HIP(sah.reserve(sizeof(double)));
    VH_HIP(B.dp_cost.reserve(sizeof(float) * 8 * (n_int + 1)));
    VH_HIP(B.dp_dec.reserve((size_t)kDpStride * (n_int + 1)));

    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, bounds.as<uint32_t>());
    uint32_t rb = std::min<uint32_t>(blocks(n, 256), 1024);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(rb), dim3(256), 0, s, d_lo, d_hi, n, bounds.as<uint32_t>());
    hipLaunchKernelGGL(k_morton, dim3(blocks(n, 256)), dim3(256), 0, s, d_lo, d_hi, n, bounds.as<uint32_t>(),
                       keys.as<uint64_t>(), ids.as<uint32_t>());
    const uint32_t* sorted_ids = nullptr;
    // 8 passes of 8 bits over the 63-bit keys, back and forth between the two
    // buffers: the sorted ids end where the keys started
    {
        const uint32_t tiles = (n + kRadixTile - 1) / kRadixTile;
        const size_t entries = (size_t)kRadixBins * tiles;
        const size_t scan_tiles = (entries + kScanTile - 1) / kScanTile;
        VH_HIP(tmp.reserve(sizeof(uint64_t) * (2 * entries + scan_tiles) + 64));
        uint64_t* counts = tmp.as<uint64_t>();
        uint64_t* offsets = counts + entries;
        uint64_t* sums = offsets + entries;
        uint32_t* sort_state = reinterpret_cast<uint32_t*>(sums + scan_tiles);
        const uint32_t sort_init[8] = {(uint32_t)entries, 0, 0, 0, 0, 0, 0, 0};
        VH_HIP(hipMemcpyAsync(sort_state, sort_init, sizeof(sort_init), hipMemcpyHostToDevice, s));
        uint64_t *k_in = keys.as<uint64_t>(), *k_out = keys_sorted.as<uint64_t>();
        uint32_t *v_in = ids.as<uint32_t>(), *v_out = ids_sorted.as<uint32_t>();
        for (int shift = 0; shift < 64; shift += 8) {
            hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kRadixBlock), 0, s, k_in, n, shift, counts, tiles);
            hipLaunchKernelGGL(k_scan_tiles, dim3((uint32_t)scan_tiles), dim3(kScanBlock), 0, s, counts, sort_state,
                               offsets, sums);
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, sort_state);
            hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kRadixBlock), 0, s, k_in, v_in, n, shift, counts,
                               offsets, sums, tiles, k_out, v_out);
            std::swap(k_in, k_out);
            std::swap(v_in, v_out);
        }
        sorted_ids = v_in;
    }
    hipLaunchKernelGGL(k_ploc_init, dim3(blocks(n, 256)), dim3(256), 0, s, d_lo, d_hi, sorted_ids, n,
                       nb.as<float>(), size.as<uint32_t>(), ca.as<uint32_t>());
    // Clustering rounds run from device-resident state ([0] clusters, [1]
    // next node id, [2..3] the round's totals, [4] no-progress flag); the host
    // reads it back every kRounds rounds and sizes the grids by the last count
    // it saw (later rounds have fewer clusters and exit early); once it sees at
    // most kTailMax clusters, one k_ploc_tail launch runs the remaining rounds.
    constexpr int kRounds = 6;
    VH_HIP(tile_sums.reserve(sizeof(uint64_t) * ((n + kScanTile - 1) / kScanTile + 1)));
    uint32_t* state = counters.as<uint32_t>();
    const uint32_t init_state[8] = {n, n, 0, 0, 0, 0, 0, 0};
    VH_HIP(hipMemcpyAsync(state, init_state, sizeof(init_state), hipMemcpyHostToDevice, s));
    uint32_t bound = n;
    uint32_t *cl_in = ca.as<uint32_t>(), *cl_out = cb.as<uint32_t>();
    while (bound > 1) {
        if (bound <= kTailMax) {
            hipLaunchKernelGGL(k_ploc_tail, dim3(1), dim3(kTailMax), 0, s, cl_in, state, n, nb.as<float>(),
                               left.as<uint32_t>(), right.as<uint32_t>(), size.as<uint32_t>(), B.dp_cost.as<float>(),
                               B.dp_dec.as<uint8_t>(), leaf_max, (uint32_t)n_nodes);
        } else for (int r = 0; r < kRounds; r++) {
            hipLaunchKernelGGL(k_ploc_nn, dim3(blocks(bound, kPlocBlock)), dim3(kPlocBlock), 0, s, cl_in, state,
                               nb.as<float>(), nn.as<uint32_t>());
            hipLaunchKernelGGL(k_ploc_flags, dim3(blocks(bound, 256)), dim3(256), 0, s, nn.as<uint32_t>(), state,
                               flags.as<uint64_t>());
            hipLaunchKernelGGL(k_scan_tiles, dim3(blocks(bound, kScanTile)), dim3(kScanBlock), 0, s,
                               flags.as<uint64_t>(), state, offsets.as<uint64_t>(), tile_sums.as<uint64_t>());
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, tile_sums.as<uint64_t>(), state);
            hipLaunchKernelGGL(k_ploc_merge, dim3(blocks(bound, 256)), dim3(256), 0, s, cl_in, nn.as<uint32_t>(), state,
                               flags.as<uint64_t>(), offsets.as<uint64_t>(), tile_sums.as<uint64_t>(), n,
                               nb.as<float>(), left.as<uint32_t>(), right.as<uint32_t>(), size.as<uint32_t>(), cl_out,
                               B.dp_cost.as<float>(), B.dp_dec.as<uint8_t>(), leaf_max);
            hipLaunchKernelGGL(k_ploc_advance, dim3(1), dim3(64), 0, s, state, (uint32_t)n_nodes);
            std::swap(cl_in, cl_out);
        }
        uint32_t host_state[8];
        VH_HIP(hipMemcpyAsync(host_state, state, sizeof(host_state), hipMemcpyDeviceToHost, s));
        VH_HIP(hipStreamSynchronize(s));
        if (host_state[4]) return fail(VH_ERR_DEVICE, "BVH clustering made no progress");
        bound = host_state[0];
    }
    const uint32_t root_id = (uint32_t)(n_nodes - 1);
    // Upper bound of wide nodes: every wide node but the root has > leaf_max
    // primitives under it except through the tree shape; n internal binary
    // nodes bound it.
    uint32_t max_nodes = std::max<uint32_t>(n_int, 1) + 1;
    VH_HIP(B.nodes.reserve(sizeof(WideNode) * max_nodes));
    VH_HIP(B.order.reserve(sizeof(uint32_t) * n));
    VH_HIP(qa.reserve(sizeof(QueueItem) * max_nodes));
    VH_HIP(qb.reserve(sizeof(QueueItem) * max_nodes));
    // counters: 0 node_counter, 1 item_counter, 2 count_a, 3 count_b, 4 max_level
    uint32_t init[8] = {1, 0, 1, 0, 0, 0, 0, 0};
    VH_HIP(hipMemcpyAsync(counters.ptr, init, sizeof(init), hipMemcpyHostToDevice, s));
    QueueItem root = {root_id, 0u};
    VH_HIP(hipMemcpyAsync(qa.ptr, &root, sizeof(root), hipMemcpyHostToDevice, s));
    Tree t{left.as<uint32_t>(), right.as<uint32_t>(), size.as<uint32_t>(), sorted_ids, nb.as<float>(),
           n, B.dp_dec.as<uint8_t>()};
    // Collapse level by level, reading the queue length back every kLevels
    // levels; a level's queue is at most kWidth times the last one.
    constexpr int kLevels = 4;
    uint32_t* cnt = counters.as<uint32_t>();
    uint32_t queue = 1, level = 0;
    QueueItem *qin = qa.as<QueueItem>(), *qout = qb.as<QueueItem>();
    uint32_t *cin = cnt + 2, *cout = cnt + 3;
    while (queue) {
        for (int k = 0; k < kLevels; k++) {
            VH_HIP(hipMemsetAsync(cout, 0, sizeof(uint32_t), s));
            hipLaunchKernelGGL(k_collapse, dim3(blocks(queue, 128)), dim3(128), 0, s, t, qin, cin, qout, cout, cnt,
                               cnt + 1, B.nodes.as<WideNode>(), B.order.as<uint32_t>(), leaf_max, n, cnt + 4,
                               level);
            std::swap(qin, qout);
            std::swap(cin, cout);
            level++;
            queue = (uint32_t)std::min<uint64_t>((uint64_t)queue * kWidth, max_nodes);
        }
        VH_HIP(hipMemcpyAsync(&queue, cin, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        VH_HIP(hipStreamSynchronize(s));
    }
    uint32_t host_counters[8];
    VH_HIP(hipMemcpyAsync(host_counters, cnt, sizeof(host_counters), hipMemcpyDeviceToHost, s));
    double sah_sum = 0;
    VH_HIP(hipMemsetAsync(sah.ptr, 0, sizeof(double), s));
    hipLaunchKernelGGL(k_sah, dim3(std::min<uint32_t>(blocks(n_nodes, 256), 512)), dim3(256), 0, s, nb.as<float>(),
                       (uint32_t)n_nodes, nb.as<float>() + 6 * (size_t)root_id, sah.as<double>());
    VH_HIP(hipMemcpyAsync(&sah_sum, sah.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    if (host_counters[1] != n) return fail(VH_ERR_DEVICE, "BVH collapse lost primitives");
    out->nodes = B.nodes.as<WideNode>();
    out->order = B.order.as<uint32_t>();
    out->node_count = host_counters[0];
    out->binary_nodes = n_int;
    out->max_depth = host_counters[4] + 1;  // the deepest level holding wide nodes, counted from 1
    out->sah_cost = sah_sum;
    return VH_OK;
}

// ------------------------------------------------------------ mesh kernels
namespace {
__global__ void k_compact_positions(const float* positions, uint32_t stride, uint32_t vertex_count, float* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= vertex_count) return;
    const float* p = positions + (size_t)i * stride;
    out[3 * i] = p[0]; out[3 * i + 1] = p[1]; out[3 * i + 2] = p[2];
}
__global__ void k_tri_bounds(const float* pos, const uint32_t* idx, uint32_t n, uint32_t vertex_count, float* lo,
                             float* hi, uint32_t* bad) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t a = idx[3 * i], b = idx[3 * i + 1], c = idx[3 * i + 2];
    if (a >= vertex_count || b >= vertex_count || c >= vertex_count) {
        atomicOr(bad, 1u);
        a = b = c = 0;
    }
    for (int k = 0; k < 3; k++) {
        float x = pos[3 * a + k], y = pos[3 * b + k], z = pos[3 * c + k];
        if (!isfinite(x) || !isfinite(y) || !isfinite(z)) atomicOr(bad, 2u);
        lo[3 * i + k] = fminf(x, fminf(y, z));
        hi[3 * i + k] = fmaxf(x, fmaxf(y, z));
    }
}
__global__ void k_gather_triangles(const float* pos, const float* nrm, const uint32_t* mat, const uint32_t* idx,
                                   const uint32_t* order, uint32_t n, float* tris, float* surf, uint32_t* tri_ids) {
    // 12-byte records move as single dwordx3 accesses
    struct alignas(4) V3 { float x, y, z; };
    struct alignas(4) I3 { uint32_t v[3]; };
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t prim = order[i];
    const I3 t = *reinterpret_cast<const I3*>(idx + 3 * (size_t)prim);
    V3* out = reinterpret_cast<V3*>(tris + 9 * (size_t)i);
    float r[kSurfFloats] = {};
    for (int v = 0; v < 3; v++) {
        const V3 p = *reinterpret_cast<const V3*>(pos + 3 * (size_t)t.v[v]);
        out[v] = p;
        r[3 * v] = p.x; r[3 * v + 1] = p.y; r[3 * v + 2] = p.z;
    }
    if (nrm) {
        for (int v = 0; v < 3; v++) {
            const V3 q = *reinterpret_cast<const V3*>(nrm + 3 * (size_t)t.v[v]);
            r[9 + 3 * v] = q.x; r[10 + 3 * v] = q.y; r[11 + 3 * v] = q.z;
        }
    }
    if (mat) r[18] = __uint_as_float(mat[prim]);
    // the whole line, padding included, so that no store is partial
    float4* os = reinterpret_cast<float4*>(surf + kSurfFloats * (size_t)i);
    for (int k = 0; k < (int)kSurfFloats / 4; k++) os[k] = make_float4(r[4 * k], r[4 * k + 1], r[4 * k + 2], r[4 * k + 3]);
    tri_ids[i] = prim;
}
__global__ void k_instance_bounds(const InstanceRecord* inst, const BlasRecord* blases, uint32_t n, float* lo, float* hi) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const InstanceRecord& r = inst[i];
    float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (r.mask_flags & (1u << 9)) {
        const BlasRecord& b = blases[r.blas];
        for (int corner = 0; corner < 8; corner++) {
            f3 p = mk(corner & 1 ? b.hi[0] : b.lo[0], corner & 2 ? b.hi[1] : b.lo[1], corner & 4 ? b.hi[2] : b.lo[2]);
            f3 w = xform_point(r.o2w, p);
            for (int k = 0; k < 3; k++) { l[k] = fminf(l[k], comp(w, k)); h[k] = fmaxf(h[k], comp(w, k)); }
        }
        // The box only guides traversal into the instance: widen it well past
        // the rounding of the world-to-object ray transform.
        for (int k = 0; k < 3; k++) {
            float e = (fabsf(l[k]) + fabsf(h[k])) * 64.f * FLT_EPSILON + 1e-30f;
            l[k] -= e; h[k] += e;
        }
    } else {
        for (int k = 0; k < 3; k++) { l[k] = h[k] = 0.f; }
    }
    for (int k = 0; k < 3; k++) { lo[3 * i + k] = l[k]; hi[3 * i + k] = h[k]; }
}
__global__ void k_compact3(const float* src, uint32_t stride, uint32_t count, float* dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float* p = src + (size_t)i * stride;
    dst[3 * i] = p[0]; dst[3 * i + 1] = p[1]; dst[3 * i + 2] = p[2];
}
__global__ void k_bounds_reduce(const float* lo, const float* hi, uint32_t n, uint32_t* out) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; k++) { mn[k] = fminf(mn[k], lo[3 * i + k]); mx[k] = fmaxf(mx[k], hi[3 * i + k]); }
    __shared__ float smn[3][256], smx[3][256];
    for (int k = 0; k < 3; k++) { smn[k][threadIdx.x] = mn[k]; smx[k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; k++) {
                smn[k][threadIdx.x] = fminf(smn[k][threadIdx.x], smn[k][threadIdx.x + s]);
                smx[k][threadIdx.x] = fmaxf(smx[k][threadIdx.x], smx[k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) {
            atomicMin(&out[k], float_to_ordered(smn[k][0]));
            atomicMax(&out[3 + k], float_to_ordered(smx[k][0]));
        }
}
__global__ void k_ordered_to_float(const uint32_t* in, float* out) {
    if (threadIdx.x < 6) out[threadIdx.x] = ordered_to_float(in[threadIdx.x]);
}
// Instance records: world-to-object rows (double-precision inverse), mask 0
// for a singular or non-finite transform or an unknown BLAS, as the placement
// kernel of crates/vork-render/src/gpu_tlas.rs:70-118 writes for such rows.
__global__ void k_instance_records(const vh_instance* inst, uint32_t n, uint32_t blas_count, InstanceRecord* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const vh_instance& s = inst[i];
    InstanceRecord r;
    for (int k = 0; k < 12; k++) r.o2w[k] = s.transform[k];
    bool valid = invert_transform(s.transform, r.w2o) && s.blas < blas_count;
    if (!valid)
        for (int k = 0; k < 12; k++) r.w2o[k] = 0.f;
    r.blas = valid ? s.blas : 0u;
    r.custom_index = s.custom_index & 0xffffffu;
    r.material = s.material;
    bool identity = valid;
    for (int k = 0; k < 12; k++) identity = identity && r.w2o[k] == ((k == 0 || k == 5 || k == 10) ? 1.f : 0.f);
    r.mask_flags = (valid ? (uint32_t)s.mask : 0u) | ((s.flags & VH_INSTANCE_CULL_BACK_FACES) ? (1u << 8) : 0u) |
                   (valid ? (1u << 9) : 0u) | (identity ? kInstIdentity : 0u);
    out[i] = r;
}
}  // namespace

int launch_compact3(hipStream_t s, const float* src, uint32_t stride_floats, uint32_t count, float* dst) {
    hipLaunchKernelGGL(k_compact3, dim3(blocks(count, 256)), dim3(256), 0, s, src, stride_floats, count, dst);
    VH_HIP(hipGetLastError());
    return VH_OK;
}

int launch_bounds_reduce(hipStream_t s, BuildScratch& B, const float* lo, const float* hi, uint32_t n, float out[6]) {
    VH_HIP(B.small.reserve(16 * sizeof(uint32_t)));
    uint32_t* b = B.small.as<uint32_t>();
    float* f = reinterpret_cast<float*>(b + 8);
    hipLaunchKernelGGL(k_init_bounds, dim3(1), dim3(64), 0, s, b);
    hipLaunchKernelGGL(k_bounds_reduce, dim3(std::min<uint32_t>(blocks(n, 256), 1024)), dim3(256), 0, s, lo, hi, n, b);
    hipLaunchKernelGGL(k_ordered_to_float, dim3(1), dim3(64), 0, s, b, f);
    VH_HIP(hipMemcpyAsync(out, f, 6 * sizeof(float), hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    return VH_OK;
}

int launch_instance_records(hipStream_t s, const vh_instance* inst, uint32_t n, uint32_t blas_count,
                            InstanceRecord* out) {
    hipLaunchKernelGGL(k_instance_records, dim3(blocks(n, 256)), dim3(256), 0, s, inst, n, blas_count, out);
    VH_HIP(hipGetLastError());
    return VH_OK;
}

int launch_triangle_bounds(hipStream_t s, BuildScratch& B, const float* positions, uint32_t stride_floats,
                           const uint32_t* indices, uint32_t n, float* lo, float* hi, float* compact_positions,
                           uint32_t vertex_count) {
    VH_HIP(B.small.reserve(16 * sizeof(uint32_t)));
    DeviceBuffer& bad = B.small;
    VH_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_compact_positions, dim3(blocks(vertex_count, 256)), dim3(256), 0, s, positions, stride_floats,
                       vertex_count, compact_positions);
    hipLaunchKernelGGL(k_tri_bounds, dim3(blocks(n, 256)), dim3(256), 0, s, compact_positions, indices, n, vertex_count,
                       lo, hi, bad.as<uint32_t>());
    uint32_t h = 0;
    VH_HIP(hipMemcpyAsync(&h, bad.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    if (h & 1u) return fail(VH_ERR_SCENE, "triangle index out of range of the vertex count");
    if (h & 2u) return fail(VH_ERR_SCENE, "non-finite vertex position");
    return VH_OK;
}

int launch_gather_triangles(hipStream_t s, const float* positions, const float* normals, const uint32_t* materials,
                            const uint32_t* indices, const uint32_t* order, uint32_t n, float* tris,
                            float* surf, uint32_t* tri_ids) {
    hipLaunchKernelGGL(k_gather_triangles, dim3(blocks(n, 256)), dim3(256), 0, s, positions, normals, materials,
                       indices, order, n, tris, surf, tri_ids);
    VH_HIP(hipGetLastError());
    return VH_OK;
}

int launch_instance_bounds(hipStream_t s, const InstanceRecord* inst, const BlasRecord* blases, uint32_t n, float* lo,
                           float* hi) {
    hipLaunchKernelGGL(k_instance_bounds, dim3(blocks(n, 256)), dim3(256), 0, s, inst, blases, n, lo, hi);
    VH_HIP(hipGetLastError());
    return VH_OK;
}

}  // namespace vh
