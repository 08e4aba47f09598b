// The 3x3 convolution of the denoiser's U-Net and its layer table, shared by
// the inference pass (vh_denoise.hip) and the training passes (vh_train.hip).
//
// Activations are NHWC bf16 with a multiple of 32 channels, so one K step of
// v_mfma_f32_16x16x32_bf16 is one tap over 32 input channels.  A workgroup
// (4 waves) owns an 8 x 16 pixel tile and its output channels (all of them,
// or a 32-channel slice per blockIdx.y on the coarse levels, whose tiles alone
// would leave most CUs idle).  For each block of 32 input channels it stages
// the (8+2) x (16+2) halo of the tile in LDS (fusing the nearest upsample and
// the skip concatenation of the decoder); each wave then walks the halo by
// rows, every A fragment feeding the up to three output rows it reaches, over
// its share of the tile's 8 pixel rows and of the output channels, with
// weights held as 16-byte fragments in registers.  Layers of up to 4 K blocks
// run in resident workgroups that keep their weights across tiles and
// prefetch the next K block's halo.  The epilogue adds the bias, applies
// ReLU and rounds to bf16; on the encoder it also writes the 2x2 average pool
// of the rounded output (the next level's input), whose 2x2 groups never
// straddle a tile or a wave's rows.  The first layer computes its input
// features from the film and guides while staging, and the output layer
// writes the radiance from its epilogue.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "vh_internal.h"

namespace vh {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kTileH = 8, kTileW = 16, kCB = 32;     // tile rows, columns, channels per K block
constexpr int kHaloH = kTileH + 2, kHaloW = kTileW + 2;
constexpr int kPix = kCB * 2 + 16;                   // LDS bytes per halo pixel (32 bf16 + pad)

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// Two values rounded to bf16 (nearest even) by one v_cvt_pk_bf16_f32, a in the
// low half: the same bits as f2bf for every non-NaN input.
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
    typedef __attribute__((ext_vector_type(2))) float f32x2;
    typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2{a, b}), bf16x2));
}

struct ConvArgs {
    const uint16_t* x1;  // channels [0, c1): at half resolution when up1
    const uint16_t* x2;  // channels [c1, c1 + c2) at full resolution (may be null when c2 == 0)
    const uint16_t* wt;  // [tap 9][cout][cin] bf16
    const float* bias;   // [cout]
    uint16_t* y;         // [h][w][cout]
    int h, w, c1, c2, up1, relu;
    int cout;            // output channels of the layer (the block computes kCout of them)
    uint16_t* pool;      // [(h + 1) / 2][(w + 1) / 2][cout] 2x2 average of y, or null
    // a batch of `images` (blockIdx.z; 0 means 1): each tensor holds its
    // images one after the other, derived strides
    int images;
    // the output layer of inference: when set, the epilogue writes the clean
    // radiance (expm1 of channels 0-2, alpha 1) here instead of y
    float4* radiance = nullptr;
    // the first layer of inference: when set, the halo is computed from the
    // film and guide sums of `feat_samples` samples (the features of
    // k_features) instead of read from x1
    const float4* film = nullptr;
    const float4* guides = nullptr;
    float feat_samples = 1.f;
};

// The tensors of image `img` of a batch: every pointer advanced by its
// image's stride (x1 at its own resolution).
__device__ __forceinline__ ConvArgs conv_image(ConvArgs a, int img) {
    if (img == 0) return a;
    const size_t px = (size_t)a.h * a.w;
    const size_t px1 = a.up1 ? (size_t)((a.h + 1) / 2) * ((a.w + 1) / 2) : px;
    const size_t pxp = (size_t)((a.h + 1) / 2) * ((a.w + 1) / 2);
    a.x1 += img * px1 * a.c1;
    if (a.x2) a.x2 += img * px * a.c2;
    a.y += img * px * a.cout;
    if (a.pool) a.pool += img * pxp * a.cout;
    if (a.radiance) a.radiance += img * px;
    if (a.film) {
        a.film += img * px;
        a.guides += img * px * 2;
    }
    return a;
}

// The epilogue of a tile through LDS, in two halves.  conv_tile_to_lds:
// bias, ReLU and bf16 rounding per D fragment (a lane holds output channel
// (block + col) of pixels 4 kg + j of a row) into an output tile of the
// block's kCout channels per pixel at `tile`.  conv_tile_store, after a
// barrier: whole 16-byte chunks (8 channels) to memory, and the 2x2 pool from
// the rounded tile (or, for the output layer, the radiance).
template <int kCout, int kMW>
__device__ __forceinline__ void conv_tile_to_lds(const ConvArgs& a, f32x4 (&acc)[kTileH / kMW][(kCout / 16) / (4 / kMW)],
                                                 uint16_t* tile, int co0) {
    constexpr int kNB = kCout / 16, kNW = 4 / kMW, kNBW = kNB / kNW, kRows = kTileH / kMW;
    constexpr int kOutPix = kCout + 8;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % kMW, wn = wave / kMW;
    const int kg = lane >> 4, col = lane & 15;
#pragma unroll
    for (int r = 0; r < kRows; r++) {
        const int ty = wm * kRows + r;
#pragma unroll
        for (int n = 0; n < kNBW; n++) {
            const int c = (wn * kNBW + n) * 16 + col;
            const float bias = a.bias[co0 + c];
#pragma unroll
            for (int j = 0; j < 4; j += 2) {
                float v0 = acc[r][n][j] + bias, v1 = acc[r][n][j + 1] + bias;
                if (a.relu) {
                    v0 = fmaxf(v0, 0.f);
                    v1 = fmaxf(v1, 0.f);
                }
                const uint32_t pk = f2bf2(v0, v1);
                tile[(ty * kTileW + kg * 4 + j) * kOutPix + c] = (uint16_t)pk;
                tile[(ty * kTileW + kg * 4 + j + 1) * kOutPix + c] = (uint16_t)(pk >> 16);
            }
        }
    }
}

template <int kCout>
__device__ __forceinline__ void conv_tile_store(const ConvArgs& a, const uint16_t* tile, int y0, int x0, int co0) {
    constexpr int kOutPix = kCout + 8;
    const int tid = threadIdx.x;
    if (a.radiance) {
        for (int px = tid; px < kTileH * kTileW; px += 256) {
            const int y = y0 + px / kTileW, x = x0 + px % kTileW;
            if (y >= a.h || x >= a.w) continue;
            const uint16_t* v = tile + px * kOutPix;
            a.radiance[(size_t)y * a.w + x] = make_float4(expm1f(fmaxf(bf2f(v[0]), 0.f)), expm1f(fmaxf(bf2f(v[1]), 0.f)),
                                                          expm1f(fmaxf(bf2f(v[2]), 0.f)), 1.f);
        }
        return;
    }
    constexpr int kChunks = kCout / 8;  // 16-byte chunks per pixel
    for (int i = tid; i < kTileH * kTileW * kChunks; i += 256) {
        const int px = i / kChunks, q = i % kChunks;
        const int y = y0 + px / kTileW, x = x0 + px % kTileW;
        if (y < a.h && x < a.w)
            *reinterpret_cast<uint4*>(a.y + ((size_t)y * a.w + x) * a.cout + co0 + q * 8) =
                *reinterpret_cast<const uint4*>(tile + px * kOutPix + q * 8);
    }
    if (a.pool) {
        const int ow = (a.w + 1) / 2;
        for (int i = tid; i < (kTileH / 2) * (kTileW / 2) * kChunks; i += 256) {
            const int g = i / kChunks, q = i % kChunks;
            const int gy = g / (kTileW / 2), gx = g % (kTileW / 2);
            const int y = y0 + 2 * gy, x = x0 + 2 * gx;
            if (y >= a.h || x >= a.w) continue;
            const bool right = x + 1 < a.w, down = y + 1 < a.h;
            const int p00 = (2 * gy) * kTileW + 2 * gx;
            const uint4 c00 = *reinterpret_cast<const uint4*>(tile + p00 * kOutPix + q * 8);
            const uint4 c01 = *reinterpret_cast<const uint4*>(tile + (p00 + 1) * kOutPix + q * 8);
            const uint4 c10 = *reinterpret_cast<const uint4*>(tile + (p00 + kTileW) * kOutPix + q * 8);
            const uint4 c11 = *reinterpret_cast<const uint4*>(tile + (p00 + kTileW + 1) * kOutPix + q * 8);
            // 1, 1/2 or 1/4: the product is the quotient, exactly
            const float inv_cnt = right && down ? 0.25f : right || down ? 0.5f : 1.f;
            uint32_t o[4];
            const uint32_t* w00 = &c00.x; const uint32_t* w01 = &c01.x;
            const uint32_t* w10 = &c10.x; const uint32_t* w11 = &c11.x;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float avg[2];
#pragma unroll
                for (int hlf = 0; hlf < 2; hlf++) {
                    const int sh = 16 * hlf;
                    // sum in the order (0,0), (0,1), (1,0), (1,1); the last row / column of
                    // an odd size averages what exists
                    float sum = bf2f((uint16_t)(w00[k] >> sh));
                    if (right) sum += bf2f((uint16_t)(w01[k] >> sh));
                    if (down) {
                        sum += bf2f((uint16_t)(w10[k] >> sh));
                        if (right) sum += bf2f((uint16_t)(w11[k] >> sh));
                    }
                    avg[hlf] = sum * inv_cnt;
                }
                o[k] = f2bf2(avg[0], avg[1]);
            }
            *reinterpret_cast<uint4*>(a.pool + ((size_t)(y / 2) * ow + x / 2) * a.cout + co0 + q * 8) =
                make_uint4(o[0], o[1], o[2], o[3]);
        }
    }

}

// Both halves of the epilogue.  The caller makes sure no wave still reads what
// `tile` overlays; the function ends with the tile read (no closing barrier).
template <int kCout, int kMW>
__device__ __forceinline__ void conv_epilogue(const ConvArgs& a, f32x4 (&acc)[kTileH / kMW][(kCout / 16) / (4 / kMW)],
                                              uint16_t* tile, int y0, int x0, int co0) {
    conv_tile_to_lds<kCout, kMW>(a, acc, tile, co0);
    __syncthreads();
    conv_tile_store<kCout>(a, tile, y0, x0, co0);
}

// The 9 taps of one K block for a wave's kRows output rows, walked by halo
// row: each A fragment (halo row h, column shift dx) is read from LDS once and
// feeds the up to three output rows h - dy it reaches, so a wave reads
// 3 (kRows + 2) fragments instead of 9 kRows.  Every accumulator still takes
// its taps in the order dy-major, dx-minor, so the sums are those of a
// tap-by-tap loop, bit for bit.
//
// With kPipe the next halo row's fragments are read before this row's MFMAs
// issue (three more fragments in registers), so a wave does not wait on each
// row's LDS latency: for the kernels that run few waves per SIMD.
template <int kRows, int kNBW, bool kPipe = false>
__device__ __forceinline__ void halo_rows_mfma(const uint8_t* halo, int row0, int col, int kg,
                                               const bf16x8 (&b)[9][kNBW], f32x4 (&acc)[kRows][kNBW]) {
    const uint8_t* base = halo + (row0 * kHaloW + col) * kPix + kg * 16;
    auto read_row = [&](int h, bf16x8 (&av)[3]) {
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
            av[dx] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(base + (h * kHaloW + dx) * kPix));
    };
    bf16x8 cur[3];
    if (kPipe) read_row(0, cur);
#pragma unroll
    for (int h = 0; h < kRows + 2; h++) {
        bf16x8 nxt[3];
        if (kPipe) {
            if (h + 1 < kRows + 2) read_row(h + 1, nxt);
        } else {
            read_row(h, cur);
        }
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
#pragma unroll
            for (int dy = 0; dy < 3; dy++) {
                const int rr = h - dy;
                if (rr < 0 || rr >= kRows) continue;
#pragma unroll
                for (int n = 0; n < kNBW; n++)
                    acc[rr][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[dx], b[dy * 3 + dx][n], acc[rr][n], 0, 0, 0);
            }
        }
        if (kPipe && h + 1 < kRows + 2) {
#pragma unroll
            for (int dx = 0; dx < 3; dx++) cur[dx] = nxt[dx];
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

constexpr int kHaloChunks = kHaloH * kHaloW * 4;              // 16-byte chunks of one K block's halo
constexpr int kChunksPerThread = (kHaloChunks + 255) / 256;

// Thread tid's chunks (i = tid + 256 j) of K block kb's halo for the tile at
// (y0, x0): zero outside the image.  The block comes from one input for
// every lane (c1 is a multiple of 32): x1, at half resolution when up1, or
// the skip x2.
__device__ __forceinline__ void halo_chunks(const ConvArgs& a, int y0, int x0, int kb, uint4 (&pre)[kChunksPerThread]) {
    const int tid = threadIdx.x;
    const bool first = kb * kCB < a.c1;
    const uint16_t* src = first ? a.x1 : a.x2;
    const int cs = first ? a.c1 : a.c2, cb = first ? kb * kCB : kb * kCB - a.c1;
    const int half = first && a.up1;
    const int ws = half ? (a.w + 1) / 2 : a.w;
#pragma unroll
    for (int j = 0; j < kChunksPerThread; j++) {
        const int i = tid + 256 * j;
        const int px = i >> 2, q = i & 3;
        const int y = y0 + px / kHaloW - 1, x = x0 + px % kHaloW - 1;
        pre[j] = make_uint4(0, 0, 0, 0);
        if (i < kHaloChunks && (unsigned)y < (unsigned)a.h && (unsigned)x < (unsigned)a.w) {
            const uint32_t p = (uint32_t)(y >> half) * (uint32_t)ws + (uint32_t)(x >> half);
            pre[j] = *reinterpret_cast<const uint4*>(src + (size_t)p * cs + cb + q * 8);
        }
    }
}

__device__ __forceinline__ void halo_store(uint8_t* halo, const uint4 (&pre)[kChunksPerThread]) {
#pragma unroll
    for (int j = 0; j < kChunksPerThread; j++) {
        const int i = threadIdx.x + 256 * j;
        if (i < kHaloChunks) *reinterpret_cast<uint4*>(halo + (i >> 2) * kPix + (i & 3) * 16) = pre[j];
    }
}

// kCout output channels from blockIdx.y * kCout, one 8x16 tile per workgroup
// (blockIdx.x) of image blockIdx.z; waves split as kMW (pixel rows) x kNW
// (channel blocks).  Per K block the halo and all 9 taps' weight fragments
// are loaded at once.
template <int kCout, int kMW>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs args) {
    const ConvArgs a = conv_image(args, blockIdx.z);
    constexpr int kNB = kCout / 16, kNW = 4 / kMW, kNBW = kNB / kNW, kRows = kTileH / kMW;
    static_assert(kNB % kNW == 0 && kTileH % kMW == 0, "wave split");
    constexpr int kOutPix = kCout + 8;  // bf16 per output-tile pixel (kCout + 16 B of pad)
    constexpr int kLds = kHaloH * kHaloW * kPix > kTileH * kTileW * kOutPix * 2 ? kHaloH * kHaloW * kPix
                                                                               : kTileH * kTileW * kOutPix * 2;
    __shared__ __attribute__((aligned(16))) uint8_t halo[kLds];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % kMW, wn = wave / kMW;
    const int tiles_x = (a.w + kTileW - 1) / kTileW;
    const int y0 = (blockIdx.x / tiles_x) * kTileH, x0 = (blockIdx.x % tiles_x) * kTileW;
    const int cin = a.c1 + a.c2;
    const int co0 = blockIdx.y * kCout;
    f32x4 acc[kRows][kNBW];
#pragma unroll
    for (int r = 0; r < kRows; r++)
#pragma unroll
        for (int n = 0; n < kNBW; n++) acc[r][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kg = lane >> 4, col = lane & 15;  // k group (8 channels) and row/column of the fragment
    for (int kb = 0; kb < cin / kCB; kb++) {
        bf16x8 b[9][kNBW];
#pragma unroll
        for (int t = 0; t < 9; t++)
#pragma unroll
            for (int n = 0; n < kNBW; n++) {
                const int co = co0 + (wn * kNBW + n) * 16 + col;
                b[t][n] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(a.wt + ((size_t)t * a.cout + co) * cin + kb * kCB + kg * 8));
            }
        uint4 pre[kChunksPerThread];
        halo_chunks(a, y0, x0, kb, pre);
        if (kb > 0) __syncthreads();  // the previous K block's MFMAs are done with the halo
        halo_store(halo, pre);
        __syncthreads();
        halo_rows_mfma<kRows, kNBW, true>(halo, wm * kRows, col, kg, b, acc);
    }
    __syncthreads();  // every wave is done with the halo
    conv_epilogue<kCout, kMW>(a, acc, reinterpret_cast<uint16_t*>(halo), y0, x0, co0);
}

// log(1 + x) for x >= 0 by Goldberg's method (D. Goldberg, "What every
// computer scientist should know about floating-point arithmetic", 1991,
// theorem 4): with u = fl(1 + x), log(u) x / (u - 1) cancels the rounding
// error of u.  log(u) comes from the hardware base-2 logarithm and the
// quotient from the hardware reciprocal: a few ulp from log1p (at most 3.6
// over 400,000 sampled x with an exact logarithm and reciprocal, each of
// which adds about one here), exact at 0 and +inf at +inf, against some 40
// instructions for log1pf; the feature channels round it to bf16.
__device__ __forceinline__ float log1p_nonneg(float x) {
    const float u = 1.f + x;
    if (u == 1.f || x == __builtin_huge_valf()) return x;
    return __log2f(u) * 0.69314718f * __fdividef(x, u - 1.f);
}

// The 32 input channels of one pixel from its film sum and guide sums (the
// log1p of the mean radiance, mean albedo, mean normal, mean depth, then
// zeros), as k_features computes them, in four 16-byte chunks.
__device__ __forceinline__ void feature_chunks(float4 f, float4 ga, float4 gn, float samples, uint4 (&c)[4]) {
    const float inv = 1.f / samples;
    float v[16];
    v[0] = log1p_nonneg(fmaxf(f.x * inv, 0.f));
    v[1] = log1p_nonneg(fmaxf(f.y * inv, 0.f));
    v[2] = log1p_nonneg(fmaxf(f.z * inv, 0.f));
    v[3] = ga.x * inv; v[4] = ga.y * inv; v[5] = ga.z * inv;
    v[6] = gn.x * inv; v[7] = gn.y * inv; v[8] = gn.z * inv;
    v[9] = gn.w * inv;
#pragma unroll
    for (int k = 10; k < 16; k++) v[k] = 0.f;
    uint32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = f2bf2(v[2 * k], v[2 * k + 1]);
    c[0] = make_uint4(w[0], w[1], w[2], w[3]);
    c[1] = make_uint4(w[4], w[5], w[6], w[7]);
    c[2] = c[3] = make_uint4(0, 0, 0, 0);
}

// Resident layers of at least this many K blocks read each halo row's
// fragments one row ahead of its MFMAs.
#ifndef VH_CONV_PIPE_KB
#define VH_CONV_PIPE_KB 1
#endif

// The same convolution with resident workgroups, for layers of at most 4
// input-channel blocks of 32 (kKB): a workgroup loads its waves' weight fragments for
// every tap and K block once, then walks tiles t = blockIdx.x + k gridDim.x
// over the batch's images; the weights are not fetched again per tile.  A
// layer of one K block reads the next tile's halo into registers one tile
// ahead (3 chunks of 16 bytes per thread), so its loads are in flight while
// the MFMAs of this one run; a layer of 2-4 K blocks stages all of a tile's
// K blocks at once after issuing their loads at the tile's start.  The halos
// and the output tile have LDS regions of their own: a tile's output waits in
// LDS and goes to memory during the next tile's staging.
template <int kCout, int kMW, int kKB, bool kFeat = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kKB == 1 ? 4 : kKB == 2 ? 3 : 2))) void k_conv3x3_resident(
    ConvArgs args, int tiles_per_image, int tiles_total) {
    constexpr int kNB = kCout / 16, kNW = 4 / kMW, kNBW = kNB / kNW, kRows = kTileH / kMW;
    static_assert(kNB % kNW == 0 && kTileH % kMW == 0, "wave split");
    constexpr int kOutPix = kCout + 8;
    // Layers of 2-4 K blocks stage all of a tile's K blocks at once, an LDS
    // halo per block: two barriers per tile instead of two per K block (at
    // 1920x1080 the decoder's first full-resolution layer, 3 K blocks, went
    // 142 -> 130 us and the 4-K-block layer 73 -> 70 us)
    constexpr bool kAll = kKB >= 2;
    constexpr int kHalos = kAll ? kKB : 1;             // halos staged per barrier
    constexpr int kHaloBytes = kHaloH * kHaloW * kPix;
    __shared__ __attribute__((aligned(16))) uint8_t halo[kHalos * kHaloBytes];
    __shared__ __attribute__((aligned(16))) uint16_t out_tile[kTileH * kTileW * kOutPix];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % kMW, wn = wave / kMW;
    const int kg = lane >> 4, col = lane & 15;
    const int cin = args.c1 + args.c2;
    const int tiles_x = (args.w + kTileW - 1) / kTileW;
    const int co0 = blockIdx.y * kCout;  // this workgroup's output-channel slice
    // weights of every K block and tap of the slice
    bf16x8 b[kKB][9][kNBW];
#pragma unroll
    for (int kb = 0; kb < kKB; kb++)
#pragma unroll
        for (int tap = 0; tap < 9; tap++)
#pragma unroll
            for (int n = 0; n < kNBW; n++) {
                const int co = co0 + (wn * kNBW + n) * 16 + col;
                b[kb][tap][n] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const uint4*>(args.wt + ((size_t)tap * args.cout + co) * cin + kb * kCB + kg * 8));
            }
    // one step = one K block of one tile (all of them when kAll); `pre` holds
    // the next step's halo chunks
    uint4 pre[kHalos][kChunksPerThread];
    auto fetch = [&](int t, int kb, uint4 (&dst)[kChunksPerThread]) {
        const int img = t / tiles_per_image, r = t - img * tiles_per_image;
        const ConvArgs a = conv_image(args, img);
        const int y0 = (r / tiles_x) * kTileH, x0 = (r % tiles_x) * kTileW;
        if constexpr (kFeat) {
            // thread tid < 180 holds halo pixel tid's film sum and guide sums
            static_assert(kKB == 1 && kChunksPerThread == 3, "features fill one K block, three float4 per pixel");
#pragma unroll
            for (int j = 0; j < 3; j++) dst[j] = make_uint4(0, 0, 0, 0);
            const int y = y0 + tid / kHaloW - 1, x = x0 + tid % kHaloW - 1;
            if (tid < kHaloH * kHaloW && y >= 0 && x >= 0 && y < a.h && x < a.w) {
                const size_t p = (size_t)y * a.w + x;
                dst[0] = __builtin_bit_cast(uint4, a.film[p]);
                dst[1] = __builtin_bit_cast(uint4, a.guides[2 * p]);
                dst[2] = __builtin_bit_cast(uint4, a.guides[2 * p + 1]);
            }
            return;
        }
        halo_chunks(a, y0, x0, kb, dst);
    };
    int t = blockIdx.x, prev = -1;
    if (!kAll && t < tiles_total) fetch(t, 0, pre[0]);
    // the global stores of tile `prev`, whose output waits in out_tile
    auto store_prev = [&]() {
        const int img = prev / tiles_per_image, r = prev - img * tiles_per_image;
        conv_tile_store<kCout>(conv_image(args, img), out_tile, (r / tiles_x) * kTileH, (r % tiles_x) * kTileW, co0);
    };
    for (; t < tiles_total; t += gridDim.x) {
        const int img = t / tiles_per_image;
        const ConvArgs a = conv_image(args, img);
        f32x4 acc[kRows][kNBW];
#pragma unroll
        for (int rr = 0; rr < kRows; rr++)
#pragma unroll
            for (int n = 0; n < kNBW; n++) acc[rr][n] = f32x4{0.f, 0.f, 0.f, 0.f};
        if constexpr (kAll) {
            // this tile's loads go out first and the previous tile's stores
            // run while they are in flight (holding them a tile ahead, as the
            // other layers do, measured no faster here: 133 vs 132 us)
#pragma unroll
            for (int kb = 0; kb < kKB; kb++) fetch(t, kb, pre[kb]);
            __syncthreads();  // the previous tile's MFMAs are done with the halos
            if (prev >= 0) store_prev();
#pragma unroll
            for (int kb = 0; kb < kKB; kb++) halo_store(halo + kb * kHaloBytes, pre[kb]);
            __syncthreads();
#pragma unroll
            for (int kb = 0; kb < kKB; kb++)
                halo_rows_mfma<kRows, kNBW, true>(halo + kb * kHaloBytes, wm * kRows, col, kg, b[kb], acc);
        } else {
#pragma unroll
            for (int kb = 0; kb < kKB; kb++) {
                __syncthreads();  // the previous step's MFMAs are done with the halo
                if constexpr (kFeat) {
                    // out-of-image pixels hold zero sums, whose features are zero
                    if (tid < kHaloH * kHaloW) {
                        uint4 c[4];
                        feature_chunks(__builtin_bit_cast(float4, pre[0][0]), __builtin_bit_cast(float4, pre[0][1]),
                                       __builtin_bit_cast(float4, pre[0][2]), a.feat_samples, c);
#pragma unroll
                        for (int q = 0; q < 4; q++) *reinterpret_cast<uint4*>(halo + tid * kPix + q * 16) = c[q];
                    }
                } else {
                    halo_store(halo, pre[0]);
                }
                // the next step's loads go out before this step's MFMAs, and the
                // previous tile's stores right after them: the wait for these
                // loads at the next step also waits for the stores, and both have
                // the MFMAs' time to complete
                if (kb + 1 < kKB) fetch(t, kb + 1, pre[0]);
                else if (t + (int)gridDim.x < tiles_total) fetch(t + gridDim.x, 0, pre[0]);
                if (kb == 0 && prev >= 0) store_prev();
                __syncthreads();
                halo_rows_mfma<kRows, kNBW, (kKB >= VH_CONV_PIPE_KB)>(halo, wm * kRows, col, kg, b[kb], acc);
            }
        }
        conv_tile_to_lds<kCout, kMW>(a, acc, out_tile, co0);
        prev = t;
    }
    if (prev >= 0) {
        __syncthreads();
        store_prev();
    }
}

// Input features from the film and guide sums of `samples` samples: log1p
// of the mean radiance, mean albedo, mean normal, mean depth; 32 channels.
__global__ void k_features(const float4* film, const float4* guides, float samples, int pixels, uint16_t* x) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    const float inv = 1.f / samples;
    const float4 f = film[p], ga = guides[2 * p], gn = guides[2 * p + 1];
    float v[32];
    for (int k = 0; k < 32; k++) v[k] = 0.f;
    v[0] = log1p_nonneg(fmaxf(f.x * inv, 0.f));
    v[1] = log1p_nonneg(fmaxf(f.y * inv, 0.f));
    v[2] = log1p_nonneg(fmaxf(f.z * inv, 0.f));
    v[3] = ga.x * inv; v[4] = ga.y * inv; v[5] = ga.z * inv;
    v[6] = gn.x * inv; v[7] = gn.y * inv; v[8] = gn.z * inv;
    v[9] = gn.w * inv;
    // 64 bytes per pixel as four 16-byte stores
    uint4* o = reinterpret_cast<uint4*>(x + (size_t)p * 32);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t w[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            w[k] = (uint32_t)f2bf(v[8 * q + 2 * k]) | ((uint32_t)f2bf(v[8 * q + 2 * k + 1]) << 16);
        o[q] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

inline uint32_t nblocks(size_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// The network: (input channels, output channels) of its 19 convolutions in
// execution order; levels 0-4 have 32, 64, 64, 96, 128 channels.
constexpr int kLayers = 19;
constexpr int kShape[kLayers][2] = {
    {32, 32},  {32, 32},                  // encoder level 0 (10 feature channels padded to 32)
    {32, 64},  {64, 64},                  // level 1
    {64, 64},  {64, 64},                  // level 2
    {64, 96},  {96, 96},                  // level 3
    {96, 128}, {128, 128},                // level 4
    {224, 96}, {96, 96},                  // decoder level 3: up(4) + skip 3
    {160, 64}, {64, 64},                  // level 2
    {128, 64}, {64, 64},                  // level 1
    {96, 32},  {32, 32},                  // level 0
    {32, 16},                             // output: log radiance in channels 0-2
};

// A resident launch: as many workgroups as the device holds at once (the
// occupancy the runtime reports for this kernel, asked once), since each
// walks a fixed share of the tiles and a workgroup that had to wait for a
// slot would start its share late.
template <int kCout, int kMW, int kKB, bool kFeat = false>
inline int launch_resident(const ConvArgs& a, int tiles_per_image, int total, uint32_t cu_count, hipStream_t s,
                           int slices = 1) {
    static const int per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_conv3x3_resident<kCout, kMW, kKB, kFeat>, 256, 0) !=
            hipSuccess)
            n = 1;
        return n > 0 ? n : 1;
    }();
    const dim3 grid((uint32_t)std::max(1, std::min<int>(total, (int)cu_count * per_cu / slices)), (uint32_t)slices);
    hipLaunchKernelGGL((k_conv3x3_resident<kCout, kMW, kKB, kFeat>), grid, dim3(256), 0, s, a, tiles_per_image, total);
    return VH_OK;
}

// Layers whose tiles give fewer than VH_DENOISE_SPLIT_BLOCKS blocks per CU
// split their output channels into 32-channel slices, one tile and slice per
// workgroup (more blocks, each staging the same halo); the others run in
// resident workgroups when they have at most 4 K blocks.
#ifndef VH_DENOISE_SPLIT_BLOCKS
#define VH_DENOISE_SPLIT_BLOCKS 3
#endif
// `force_split` takes the sliced kernel for any width that is a multiple of
// 32 (the backward pass's 160- and 224-channel outputs).
inline int launch_conv(const ConvArgs& a, uint32_t cu_count, hipStream_t s, bool force_split = false) {
    const uint32_t tiles = nblocks(a.w, kTileW) * nblocks(a.h, kTileH);
    const bool split = force_split || (a.cout > 32 && tiles < VH_DENOISE_SPLIT_BLOCKS * cu_count);
    const int cin = a.c1 + a.c2, images = a.images > 1 ? a.images : 1;
    if (a.film) {
        if (cin != kCB || a.cout != 32 || a.up1 || a.c2)
            return fail(VH_ERR_USAGE, "denoiser: features feed a 32-to-32-channel layer only");
        return launch_resident<32, 2, 1, true>(a, (int)tiles, (int)tiles * images, cu_count, s);
    }
    if (!split && cin <= 4 * kCB && (a.cout == 16 || a.cout == 32 || a.cout == 64)) {
        // resident workgroups: weights loaded once, the next halo prefetched
        const int total = (int)tiles * images;
        const int kb = cin / kCB;
        if (a.cout == 16 && kb == 1) return launch_resident<16, 4, 1>(a, (int)tiles, total, cu_count, s);
        if (a.cout == 16) return launch_resident<16, 4, 2>(a, (int)tiles, total, cu_count, s);
        if (a.cout == 32 && kb == 1) return launch_resident<32, 2, 1>(a, (int)tiles, total, cu_count, s);
        if (a.cout == 32 && kb == 2) return launch_resident<32, 2, 2>(a, (int)tiles, total, cu_count, s);
        if (a.cout == 32 && kb == 3) return launch_resident<32, 2, 3>(a, (int)tiles, total, cu_count, s);
        if (a.cout == 32) return launch_resident<32, 2, 4>(a, (int)tiles, total, cu_count, s);
        if (kb == 1) return launch_resident<64, 1, 1>(a, (int)tiles, total, cu_count, s);
        if (kb == 2) return launch_resident<64, 1, 2>(a, (int)tiles, total, cu_count, s);
        if (kb == 3) return launch_resident<64, 1, 3>(a, (int)tiles, total, cu_count, s);
        return launch_resident<64, 1, 4>(a, (int)tiles, total, cu_count, s);
    }
    const dim3 grid(tiles, split ? a.cout / 32 : 1, a.images > 1 ? a.images : 1);
    switch (split ? 32 : a.cout) {
        // wave splits and tap groups as measured fastest at 1920x1080: the
        // kernels are latency bound (a block waits on its halo and weight
        // loads), so fewer registers and more resident waves win
        case 16: hipLaunchKernelGGL((k_conv3x3<16, 4>), grid, dim3(256), 0, s, a); break;
        case 32: hipLaunchKernelGGL((k_conv3x3<32, 2>), grid, dim3(256), 0, s, a); break;
        case 64: hipLaunchKernelGGL((k_conv3x3<64, 1>), grid, dim3(256), 0, s, a); break;
        case 96: hipLaunchKernelGGL((k_conv3x3<96, 2>), grid, dim3(256), 0, s, a); break;
        case 128: hipLaunchKernelGGL((k_conv3x3<128, 1>), grid, dim3(256), 0, s, a); break;
        default: return fail(VH_ERR_USAGE, "denoiser: unsupported channel count");
    }
    return VH_OK;
}

}  // namespace
}  // namespace vh
