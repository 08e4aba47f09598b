// Learned denoiser of the path tracer's film (replaces the OIDN pass of
// crates/vork-render/src/denoise.rs): a 5-level U-Net of 3x3 convolutions in
// bf16 on MFMA with fp32 accumulation.
//
// Activations are NHWC bf16 with a multiple of 32 channels, so one K step of
// v_mfma_f32_16x16x32_bf16 is one tap over 32 input channels.  A workgroup
// (4 waves) owns an 8 x 16 pixel tile and its output channels (all of them,
// or a 32-channel slice per blockIdx.y on the coarse levels, whose tiles alone
// would leave most CUs idle); for each
// block of 32 input channels it stages the (8+2) x (16+2) halo of the tile in
// LDS (fusing the nearest upsample and the skip concatenation of the decoder),
// then each wave runs 9 taps of MFMAs over its share of the tile's 8 pixel
// rows and of the output channels, with weights read as 16-byte fragments.
// The epilogue adds the bias, applies ReLU and rounds to bf16; on the encoder
// it also writes the 2x2 average pool of the rounded output (the next level's
// input), whose 2x2 groups never straddle a tile or a wave's rows.
#include <cmath>
#include <cstring>
#include <vector>

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

struct ConvArgs {
    const uint16_t* x1;  // channels [0, c1): at half resolution when up1
    const uint16_t* x2;  // channels [c1, c1 + c2) at full resolution (may be null when c2 == 0)
    const uint16_t* wt;  // [tap 9][cout][cin] bf16
    const float* bias;   // [cout]
    uint16_t* y;         // [h][w][cout]
    int h, w, c1, c2, up1, relu;
    int cout;            // output channels of the layer (the block computes kCout of them)
    uint16_t* pool;      // [(h + 1) / 2][(w + 1) / 2][cout] 2x2 average of y, or null
};

// One 16-byte chunk of channels [ch, ch + 8) of halo pixel (hy, hx) of the tile
// at (y0, x0): zero outside the image.
__device__ __forceinline__ uint4 fetch_chunk(const ConvArgs& a, int y, int x, int ch) {
    if (y < 0 || x < 0 || y >= a.h || x >= a.w) return make_uint4(0, 0, 0, 0);
    if (ch < a.c1) {
        const int wh = a.up1 ? (a.w + 1) / 2 : a.w;
        const int yy = a.up1 ? y >> 1 : y, xx = a.up1 ? x >> 1 : x;
        return *reinterpret_cast<const uint4*>(a.x1 + ((size_t)yy * wh + xx) * a.c1 + ch);
    }
    return *reinterpret_cast<const uint4*>(a.x2 + ((size_t)y * a.w + x) * a.c2 + (ch - a.c1));
}

// kCout output channels from blockIdx.y * kCout; waves split as kMW (pixel
// rows) x kNW (channel blocks).
template <int kCout, int kMW>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a) {
    constexpr int kNB = kCout / 16, kNW = 4 / kMW, kNBW = kNB / kNW, kRows = kTileH / kMW;
    static_assert(kNB % kNW == 0 && kTileH % kMW == 0, "wave split");
    __shared__ __attribute__((aligned(16))) uint8_t halo[kHaloH * kHaloW * kPix];
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
    for (int cb = 0; cb < cin; cb += kCB) {
        __syncthreads();
        for (int i = tid; i < kHaloH * kHaloW * 4; i += 256) {
            const int px = i >> 2, q = i & 3;
            const int hy = px / kHaloW, hx = px % kHaloW;
            *reinterpret_cast<uint4*>(halo + px * kPix + q * 16) = fetch_chunk(a, y0 + hy - 1, x0 + hx - 1, cb + q * 8);
        }
        __syncthreads();
#pragma unroll 1
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;
            bf16x8 b[kNBW];
#pragma unroll
            for (int n = 0; n < kNBW; n++) {
                const int co = co0 + (wn * kNBW + n) * 16 + col;
                b[n] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(
                                                      a.wt + ((size_t)tap * a.cout + co) * cin + cb + kg * 8));
            }
#pragma unroll
            for (int r = 0; r < kRows; r++) {
                const int hy = wm * kRows + r + dy, hx = col + dx;
                const bf16x8 av = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(
                                                                 halo + (hy * kHaloW + hx) * kPix + kg * 16));
#pragma unroll
                for (int n = 0; n < kNBW; n++) acc[r][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b[n], acc[r][n], 0, 0, 0);
            }
        }
    }
    // D fragment: lane holds output channel (block + col) of pixels 4 kg + j of the row
    static_assert(kRows % 2 == 0, "pooling pairs rows within a wave");
#pragma unroll
    for (int r = 0; r < kRows; r++) {
        const int y = y0 + wm * kRows + r;
#pragma unroll
        for (int n = 0; n < kNBW; n++) {
            const int co = co0 + (wn * kNBW + n) * 16 + col;
            const float bias = a.bias[co];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int x = x0 + kg * 4 + j;
                float v = acc[r][n][j] + bias;
                if (a.relu) v = fmaxf(v, 0.f);
                const uint16_t h16 = f2bf(v);
                acc[r][n][j] = bf2f(h16);  // the stored value, for the pool below
                if (y < a.h && x < a.w) a.y[((size_t)y * a.w + x) * a.cout + co] = h16;
            }
        }
    }
    if (a.pool) {
        const int ow = (a.w + 1) / 2;
#pragma unroll
        for (int r = 0; r < kRows; r += 2) {
            const int y = y0 + wm * kRows + r;
            if (y >= a.h) continue;
#pragma unroll
            for (int n = 0; n < kNBW; n++) {
                const int co = co0 + (wn * kNBW + n) * 16 + col;
#pragma unroll
                for (int j = 0; j < 4; j += 2) {
                    const int x = x0 + kg * 4 + j;
                    if (x >= a.w) continue;
                    // sum in the order (0,0), (0,1), (1,0), (1,1); the last row / column of
                    // an odd size averages what exists
                    float s = acc[r][n][j];
                    int cnt = 1;
                    if (x + 1 < a.w) { s += acc[r][n][j + 1]; cnt++; }
                    if (y + 1 < a.h) {
                        s += acc[r + 1][n][j];
                        cnt++;
                        if (x + 1 < a.w) { s += acc[r + 1][n][j + 1]; cnt++; }
                    }
                    a.pool[((size_t)(y / 2) * ow + x / 2) * a.cout + co] = f2bf(s / (float)cnt);
                }
            }
        }
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
    v[0] = log1pf(fmaxf(f.x * inv, 0.f));
    v[1] = log1pf(fmaxf(f.y * inv, 0.f));
    v[2] = log1pf(fmaxf(f.z * inv, 0.f));
    v[3] = ga.x * inv; v[4] = ga.y * inv; v[5] = ga.z * inv;
    v[6] = gn.x * inv; v[7] = gn.y * inv; v[8] = gn.z * inv;
    v[9] = gn.w * inv;
    uint16_t* o = x + (size_t)p * 32;
    for (int k = 0; k < 32; k++) o[k] = f2bf(v[k]);
}

// Output: expm1 of the predicted log radiance (channels 0-2 of 16), alpha 1.
__global__ void k_output(const uint16_t* y, int pixels, float4* out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    const uint16_t* v = y + (size_t)p * 16;
    out[p] = make_float4(expm1f(fmaxf(bf2f(v[0]), 0.f)), expm1f(fmaxf(bf2f(v[1]), 0.f)),
                         expm1f(fmaxf(bf2f(v[2]), 0.f)), 1.f);
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

}  // namespace
}  // namespace vh

struct vh_denoiser {
    vh::DeviceRef device;  // first: dropped after the buffers below are freed
    vh::DeviceBuffer weights, biases, scratch;
    size_t w_off[vh::kLayers] = {}, b_off[vh::kLayers] = {};
    size_t scratch_bytes = 0;
    std::mutex lock;       // one denoise at a time per object (shared scratch)
};

namespace vh {
namespace {

// Layers whose tiles give fewer than kSplitBlocks blocks per CU split their
// output channels into 32-channel slices (more blocks, each staging the same
// halo).
#ifndef VH_DENOISE_SPLIT_BLOCKS
#define VH_DENOISE_SPLIT_BLOCKS 4
#endif
int launch_conv(const ConvArgs& a, uint32_t cu_count, hipStream_t s) {
    const uint32_t tiles = nblocks(a.w, kTileW) * nblocks(a.h, kTileH);
    const bool split = a.cout > 32 && tiles < VH_DENOISE_SPLIT_BLOCKS * cu_count;
    const dim3 grid(tiles, split ? a.cout / 32 : 1);
    switch (split ? 32 : a.cout) {
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

int denoise_run(vh_denoiser* dn, int W, int H, float samples, const float* film, const float* guides, float* out,
                hipStream_t s) {
    int h[5], w[5];
    h[0] = H; w[0] = W;
    for (int l = 1; l < 5; l++) { h[l] = (h[l - 1] + 1) / 2; w[l] = (w[l - 1] + 1) / 2; }
    const int C[5] = {32, 64, 64, 96, 128};
    // scratch: features, a temporary and a pooled input at the largest level
    // used, every encoder output, every decoder output, and the output layer
    auto px = [&](int l) { return (size_t)h[l] * w[l]; };
    size_t off = 0;
    auto take = [&](size_t elems) { size_t o = off; off += (elems * 2 + 255) & ~(size_t)255; return o; };
    const size_t o_feat = take(px(0) * 32), o_tmp = take(px(0) * 128), o_pool = take(px(1) * 96);
    size_t o_enc[5], o_dec[4];
    for (int l = 0; l < 5; l++) o_enc[l] = take(px(l) * C[l]);
    for (int l = 0; l < 4; l++) o_dec[l] = take(px(l) * C[l]);
    const size_t o_out = take(px(0) * 16);
    if (dn->scratch_bytes < off) {
        VH_HIP(dn->scratch.allocate(off));
        dn->scratch_bytes = off;
    }
    uint8_t* base = dn->scratch.as<uint8_t>();
    auto at = [&](size_t o) { return reinterpret_cast<uint16_t*>(base + o); };
    const uint16_t* wts = dn->weights.as<uint16_t>();
    const float* bs = dn->biases.as<float>();
    int layer = 0, rc;
    auto conv = [&](int l, const uint16_t* x1, int c1, int up1, const uint16_t* x2, int c2, uint16_t* y, bool relu,
                    uint16_t* pool = nullptr) {
        ConvArgs a;
        a.x1 = x1; a.x2 = x2; a.c1 = c1; a.c2 = c2; a.up1 = up1; a.pool = pool;
        a.wt = wts + dn->w_off[layer]; a.bias = bs + dn->b_off[layer];
        a.y = y; a.h = h[l]; a.w = w[l]; a.relu = relu ? 1 : 0;
        a.cout = kShape[layer][1];
        layer++;
        return launch_conv(a, (uint32_t)dn->device->cu_count, s);
    };
    hipLaunchKernelGGL(k_features, dim3(nblocks(px(0), 256)), dim3(256), 0, s, reinterpret_cast<const float4*>(film),
                       reinterpret_cast<const float4*>(guides), samples, (int)px(0), at(o_feat));
    // encoder
    const uint16_t* in = at(o_feat);
    int cin = 32;
    for (int l = 0; l < 5; l++) {
        if ((rc = conv(l, in, cin, 0, nullptr, 0, at(o_tmp), true))) return rc;
        // levels 0-3 also write the next level's input, the 2x2 pool of their output
        if ((rc = conv(l, at(o_tmp), C[l], 0, nullptr, 0, at(o_enc[l]), true, l < 4 ? at(o_pool) : nullptr))) return rc;
        in = at(o_pool);
        cin = C[l];
    }
    // decoder: nearest upsample of the coarser level fused with the skip
    const uint16_t* coarse = at(o_enc[4]);
    int cc = C[4];
    for (int l = 3; l >= 0; l--) {
        if ((rc = conv(l, coarse, cc, 1, at(o_enc[l]), C[l], at(o_tmp), true))) return rc;
        if ((rc = conv(l, at(o_tmp), C[l], 0, nullptr, 0, at(o_dec[l]), true))) return rc;
        coarse = at(o_dec[l]);
        cc = C[l];
    }
    if ((rc = conv(0, at(o_dec[0]), 32, 0, nullptr, 0, at(o_out), false))) return rc;
    hipLaunchKernelGGL(k_output, dim3(nblocks(px(0), 256)), dim3(256), 0, s, at(o_out), (int)px(0),
                       reinterpret_cast<float4*>(out));
    VH_HIP(hipGetLastError());
    return VH_OK;
}

}  // namespace vh

using namespace vh;

extern "C" {

int vh_denoiser_create(vh_device* dev, const void* blob, size_t bytes, vh_denoiser** out) {
    if (!dev || !blob || !out) return fail(VH_ERR_USAGE, "vh_denoiser_create: null argument");
    *out = nullptr;
    const uint8_t* p = static_cast<const uint8_t*>(blob);
    auto u32 = [&](size_t o) { uint32_t v; std::memcpy(&v, p + o, 4); return v; };
    if (bytes < 12 || std::memcmp(p, "VHDN", 4) != 0 || u32(4) != 1 || u32(8) != (uint32_t)kLayers)
        return fail(VH_ERR_USAGE, "vh_denoiser_create: not a version-1 weight blob of the 19-layer U-Net");
    std::vector<uint16_t> w;
    std::vector<float> b;
    size_t o = 12;
    size_t w_off[kLayers], b_off[kLayers];
    for (int l = 0; l < kLayers; l++) {
        if (o + 8 > bytes) return fail(VH_ERR_USAGE, "vh_denoiser_create: truncated weight blob");
        const uint32_t ci = u32(o), co = u32(o + 4);
        o += 8;
        if ((int)ci != kShape[l][0] || (int)co != kShape[l][1])
            return fail(VH_ERR_USAGE, "vh_denoiser_create: layer " + std::to_string(l) + " has shape " +
                                          std::to_string(ci) + "x" + std::to_string(co));
        const size_t nw = (size_t)9 * co * ci;
        if (o + nw * 2 + co * 4 > bytes) return fail(VH_ERR_USAGE, "vh_denoiser_create: truncated weight blob");
        w_off[l] = w.size();
        b_off[l] = b.size();
        w.resize(w.size() + nw);
        std::memcpy(w.data() + w_off[l], p + o, nw * 2);
        o += nw * 2;
        b.resize(b.size() + co);
        std::memcpy(b.data() + b_off[l], p + o, co * 4);
        o += co * 4;
    }
    if (o != bytes) return fail(VH_ERR_USAGE, "vh_denoiser_create: trailing bytes after the last layer");
    for (float v : b)
        if (!std::isfinite(v)) return fail(VH_ERR_USAGE, "vh_denoiser_create: non-finite bias");
    VH_HIP(hipSetDevice(dev->ordinal));
    vh_denoiser* dn = new (std::nothrow) vh_denoiser();
    if (!dn) return fail(VH_ERR_DEVICE, "out of host memory");
    dn->device = dev;
    for (int l = 0; l < kLayers; l++) { dn->w_off[l] = w_off[l]; dn->b_off[l] = b_off[l]; }
    hipError_t e = dn->weights.allocate(w.size() * 2);
    if (e == hipSuccess) e = dn->biases.allocate(b.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(dn->weights.ptr, w.data(), w.size() * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dn->biases.ptr, b.data(), b.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        delete dn;
        return hip_fail(e, "vh_denoiser_create");
    }
    *out = dn;
    return VH_OK;
}

int vh_denoise(vh_denoiser* dn, uint32_t width, uint32_t height, float samples, const float* film,
               const float* guides, uint32_t flags, void* stream, float* out) {
    if (!dn || !film || !guides || !out) return fail(VH_ERR_USAGE, "vh_denoise: null argument");
    if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 27))
        return fail(VH_ERR_USAGE, "vh_denoise: image size out of range");
    if (!(samples > 0.f) || !std::isfinite(samples)) return fail(VH_ERR_USAGE, "vh_denoise: samples must be > 0");
    vh_device* dev = dn->device;
    VH_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : dev->stream;
    std::lock_guard<std::mutex> hold(dn->lock);
    const size_t px = (size_t)width * height;
    if (flags & VH_DEVICE_POINTERS) {
        int rc = denoise_run(dn, (int)width, (int)height, samples, film, guides, out, s);
        if (rc) return rc;
        VH_HIP(hipStreamSynchronize(s));
        return VH_OK;
    }
    DeviceBuffer f, g, o;
    VH_HIP(f.allocate(px * 16));
    VH_HIP(g.allocate(px * 32));
    VH_HIP(o.allocate(px * 16));
    VH_HIP(hipMemcpyAsync(f.ptr, film, px * 16, hipMemcpyHostToDevice, s));
    VH_HIP(hipMemcpyAsync(g.ptr, guides, px * 32, hipMemcpyHostToDevice, s));
    int rc = denoise_run(dn, (int)width, (int)height, samples, f.as<float>(), g.as<float>(), o.as<float>(), s);
    if (rc) return rc;
    VH_HIP(hipMemcpyAsync(out, o.ptr, px * 16, hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    return VH_OK;
}

void vh_denoiser_release(vh_denoiser* dn) {
    if (!dn) return;
    (void)hipSetDevice(dn->device->ordinal);
    (void)hipStreamSynchronize(dn->device->stream);
    delete dn;
}

}  // extern "C"
