// Training of the denoiser's U-Net with hand-written kernels: the forward
// pass keeps every activation, then a backward pass per image computes the
// weight, bias and activation gradients, and an Adam step updates fp32 master
// weights and rewrites the bf16 weights the MFMA kernels read.
//
// Numerics follow the inference pass: activations and their gradients are
// bf16 on MFMA with fp32 accumulation; weight and bias gradients, the gradient
// buffers that sum several contributions, the loss and Adam are fp32.
//
//   forward   the 19 k_conv3x3 launches of the inference pass, each layer's
//             output kept (and the 2x2 pools of the encoder)
//   loss      k_loss: the relative-MSE of the radiance plus the L1 of its
//             log1p (the pair of losses the denoiser is trained with), and
//             its gradient with respect to the output layer
//   backward  per layer, last to first: k_relu_grad (the gradient through
//             the ReLU, rounded to bf16), k_wgrad (weight and bias gradients:
//             per tap a GEMM over pixels on MFMA), and the input gradient as a
//             k_conv3x3 over the output gradient with the weights flipped and
//             transposed; k_route then sends it where the input came from
//             (the previous layer, a 2x2 pool, or the decoder's upsample and
//             skip concatenation)
//   update    k_adam over every weight and bias
#include <cmath>
#include <cstring>
#include <vector>

#include "vh_conv.h"
#include "vh_internal.h"

namespace vh {
namespace {

constexpr int kLevelC[5] = {32, 64, 64, 96, 128};
inline int cpad32(int c) { return (c + 31) / 32 * 32; }

// Level (resolution) of each layer's input and output.
constexpr int kLevel[kLayers] = {0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 3, 3, 2, 2, 1, 1, 0, 0, 0};

// Where a layer's input comes from: x1 (layer output index, or -1 the
// features, or -2 - l the 2x2 pool of level l's encoder output), whether x1 is
// the coarser level upsampled, and x2 (a layer output concatenated after x1,
// or -1).
struct LayerInput {
    int x1, up1, x2;
};
constexpr LayerInput kInput[kLayers] = {
    {-1, 0, -1}, {0, 0, -1},                 // encoder level 0
    {-2, 0, -1}, {2, 0, -1},                 // level 1: pool of level 0's output (layer 1)
    {-3, 0, -1}, {4, 0, -1},                 // level 2
    {-4, 0, -1}, {6, 0, -1},                 // level 3
    {-5, 0, -1}, {8, 0, -1},                 // level 4
    {9, 1, 7}, {10, 0, -1},                  // decoder level 3: up(layer 9) ++ layer 7
    {11, 1, 5}, {12, 0, -1},                 // level 2: up(layer 11) ++ layer 5
    {13, 1, 3}, {14, 0, -1},                 // level 1: up(layer 13) ++ layer 3
    {15, 1, 1}, {16, 0, -1},                 // level 0: up(layer 15) ++ layer 1
    {17, 0, -1},                             // output
};

// ------------------------------------------------------------------ loss
// Per output pixel and radiance channel: pred = the layer's log radiance,
// d = expm1(max(pred, 0)), loss = (d - t)^2 / (t^2 + 0.01) + |pred - log1p(t)|,
// averaged over `count` values (the batch); writes d loss / d pred into the
// 16-channel fp32 gradient of the output layer (0 for channels 3-15) and
// adds the loss to `*sum`.
__global__ void k_loss(const uint16_t* y, const float* target, int pixels, float inv_count, float* grad, float* sum) {
    __shared__ float part[256];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (p < pixels) {
        for (int c = 0; c < 3; c++) {
            const float pred = bf2f(y[(size_t)p * 16 + c]), t = target[(size_t)p * 3 + c];
            const float e = expf(fmaxf(pred, 0.f));
            const float d = e - 1.f, w = 1.f / (t * t + 0.01f), lt = log1pf(t);
            const float r = pred - lt;
            l += (d - t) * (d - t) * w + fabsf(r);
            const float gd = pred >= 0.f ? 2.f * (d - t) * w * e : 0.f;
            grad[(size_t)p * 16 + c] = (gd + (r > 0.f ? 1.f : r < 0.f ? -1.f : 0.f)) * inv_count;
        }
        for (int c = 3; c < 16; c++) grad[(size_t)p * 16 + c] = 0.f;
    }
    part[threadIdx.x] = l;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) part[threadIdx.x] += part[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(sum, part[0] * inv_count);
}

// --------------------------------------------------------------- backward
// Gradient at the pre-activation, rounded to bf16, channels padded to cpad
// with zeros: g * (y > 0) when the layer has a ReLU, else g.
__global__ void k_relu_grad(const float* g, const uint16_t* y, int pixels, int c, int cp, int relu, uint16_t* dz) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)pixels * cp) return;
    const size_t p = i / cp;
    const int ch = (int)(i % cp);
    float v = 0.f;
    if (ch < c) {
        v = g[p * c + ch];
        if (relu && !(bf2f(y[p * c + ch]) > 0.f)) v = 0.f;
    }
    dz[i] = f2bf(v);
}

// Weight and bias gradients of one layer over the batch:
//   dW[tap][co][ci] = sum_p dz[p][co] * x[p + offset(tap)][ci],  db[co] = sum_p dz[p][co]
// as GEMMs with M = output channels, N = input channels and K = pixels.  A
// workgroup owns 64 output channels (blockIdx.y), 32 input channels
// (blockIdx.z) and every 8 x 16 pixel tile t = blockIdx.x + k gridDim.x of
// the batch; per tile it stages the output gradient [128 pixels][64] and the
// input's 10 x 18 halo [180 pixels][32] in LDS, and wave w computes taps
// w, w + 4, w + 8 (its A fragments, the output gradient, are shared by its
// taps; its B fragments are the input shifted by the tap), keeping all their
// accumulators across the tiles.  Each workgroup writes its partial sums once;
// k_wgrad_reduce adds the workgroups' partials to the gradients in a fixed
// order, so the result does not depend on scheduling.
struct WgradArgs {
    const uint16_t* dz;   // [images][h][w][cp] output gradient (bf16, padded)
    const uint16_t* x1;   // [images][h or h/2][w or w/2][c1]
    const uint16_t* x2;   // [images][h][w][c2] or null
    float* part;          // [gridDim.x][gridDim.y][gridDim.z][kWgradPart] partial sums
    int h, w, c1, c2, up1, cout, cp, images;
};
constexpr int kWgradW = 9 * 64 * 32;          // one workgroup's weight partials
constexpr int kWgradPart = kWgradW + 64;      // and its bias partials
constexpr int kDzStride = 64 + 8;             // bf16 per staged pixel of dz
constexpr int kXStride = 32 + 8;              // bf16 per staged halo pixel

// 8 input channels [ch, ch + 8) of pixel (y, x) of image `img`: zero outside.
__device__ __forceinline__ uint4 wgrad_fetch(const WgradArgs& a, int img, int y, int x, int ch) {
    if (y < 0 || x < 0 || y >= a.h || x >= a.w) return make_uint4(0, 0, 0, 0);
    if (ch < a.c1) {
        const int wh = a.up1 ? (a.w + 1) / 2 : a.w, hh = a.up1 ? (a.h + 1) / 2 : a.h;
        const int yy = a.up1 ? y >> 1 : y, xx = a.up1 ? x >> 1 : x;
        return *reinterpret_cast<const uint4*>(a.x1 + (((size_t)img * hh + yy) * wh + xx) * a.c1 + ch);
    }
    return *reinterpret_cast<const uint4*>(a.x2 + (((size_t)img * a.h + y) * a.w + x) * a.c2 + (ch - a.c1));
}

__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t sdz[kTileH * kTileW * kDzStride];
    __shared__ __attribute__((aligned(16))) uint16_t sx[kHaloH * kHaloW * kXStride];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, kg = lane >> 4;
    const int co0 = blockIdx.y * 64, cb = blockIdx.z * 32;
    const int tiles_x = (a.w + kTileW - 1) / kTileW, per_image = tiles_x * ((a.h + kTileH - 1) / kTileH);
    const int tiles = per_image * a.images;
    f32x4 acc[3][4][2];
#pragma unroll
    for (int t = 0; t < 3; t++)
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int n = 0; n < 2; n++) acc[t][m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bsum[4] = {0.f, 0.f, 0.f, 0.f};
    for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int img = t / per_image, r = t - img * per_image;
        const int y0 = (r / tiles_x) * kTileH, x0 = (r % tiles_x) * kTileW;
        __syncthreads();
        for (int i = tid; i < kTileH * kTileW * 8; i += 256) {
            const int px = i >> 3, q = i & 7;
            const int y = y0 + px / kTileW, x = x0 + px % kTileW, co = co0 + q * 8;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (y < a.h && x < a.w && co < a.cp)
                v = *reinterpret_cast<const uint4*>(a.dz + (((size_t)img * a.h + y) * a.w + x) * a.cp + co);
            *reinterpret_cast<uint4*>(sdz + px * kDzStride + q * 8) = v;
        }
        for (int i = tid; i < kHaloH * kHaloW * 4; i += 256) {
            const int px = i >> 2, q = i & 3;
            *reinterpret_cast<uint4*>(sx + px * kXStride + q * 8) =
                wgrad_fetch(a, img, y0 + px / kHaloW - 1, x0 + px % kHaloW - 1, cb + q * 8);
        }
        __syncthreads();
        // K step s: tile rows 2s and 2s + 1; lane group kg holds pixels
        // (2s + kg / 2, 8 (kg % 2) .. 8 (kg % 2) + 7)
#pragma unroll 1
        for (int st = 0; st < kTileH / 2; st++) {
            const int prow = 2 * st + (kg >> 1), pcol = (kg & 1) * 8;
            bf16x8 A[4];
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int j = 0; j < 8; j++)
                    A[m][j] = __builtin_bit_cast(__bf16, sdz[(prow * kTileW + pcol + j) * kDzStride + m * 16 + col]);
            if (wave == 0 && blockIdx.z == 0) {
#pragma unroll
                for (int m = 0; m < 4; m++)
#pragma unroll
                    for (int j = 0; j < 8; j++) bsum[m] += (float)A[m][j];
            }
#pragma unroll
            for (int ti = 0; ti < 3; ti++) {
                const int tap = wave + 4 * ti;
                if (tap >= 9) break;
                const int dy = tap / 3, dx = tap % 3;
#pragma unroll
                for (int n = 0; n < 2; n++) {
                    bf16x8 B;
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        B[j] = __builtin_bit_cast(
                            __bf16, sx[((prow + dy) * kHaloW + pcol + j + dx) * kXStride + n * 16 + col]);
#pragma unroll
                    for (int m = 0; m < 4; m++)
                        if (co0 + m * 16 < a.cout)
                            acc[ti][m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, acc[ti][m][n], 0, 0, 0);
                }
            }
        }
    }
    // D fragment: lane holds output channels m * 16 + 4 kg + i of input channel n * 16 + col
    float* part = a.part + ((size_t)(blockIdx.x * gridDim.y + blockIdx.y) * gridDim.z + blockIdx.z) * kWgradPart;
#pragma unroll
    for (int ti = 0; ti < 3; ti++) {
        const int tap = wave + 4 * ti;
        if (tap >= 9) break;
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int n = 0; n < 2; n++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    part[(tap * 64 + m * 16 + 4 * kg + i) * 32 + n * 16 + col] = acc[ti][m][n][i];
    }
    if (wave == 0) {
#pragma unroll
        for (int m = 0; m < 4; m++) {
            float v = bsum[m];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (kg == 0) part[kWgradW + m * 16 + col] = v;  // zero unless blockIdx.z == 0
        }
    }
}

// dW[tap][co][ci] += the workgroups' partials; db[co] likewise from the
// partials of input-channel block 0.  A block reduces 32 consecutive
// gradient entries: its 8 thread rows each sum every 8th workgroup's
// partials, then row 0 adds the 8 row sums, always in the same order.
constexpr int kReduceRows = 8;
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float* part, int chunks, int ny, int nz, int cout,
                                                      int cin, float* dw, float* db) {
    __shared__ float rows[kReduceRows][33];
    const int e = threadIdx.x & 31, g = threadIdx.x >> 5;
    const size_t i = (size_t)blockIdx.x * 32 + e;
    const size_t nw = (size_t)9 * cout * cin;
    size_t at = 0, first = 0;  // the entry's offset in a partial block, and its first block
    bool valid = true;
    if (i < nw) {
        const int ci = (int)(i % cin), co = (int)((i / cin) % cout), tap = (int)(i / ((size_t)cin * cout));
        at = (size_t)(tap * 64 + co % 64) * 32 + ci % 32;
        first = (size_t)(co / 64) * nz + ci / 32;
    } else if (i < nw + cout) {
        const int co = (int)(i - nw);
        at = (size_t)kWgradW + co % 64;
        first = (size_t)(co / 64) * nz;
    } else {
        valid = false;
    }
    float sum = 0.f;
    if (valid)
        for (int c = g; c < chunks; c += kReduceRows) sum += part[((size_t)c * ny * nz + first) * kWgradPart + at];
    rows[g][e] = sum;
    __syncthreads();
    if (g == 0 && valid) {
        float t = 0.f;
        for (int k = 0; k < kReduceRows; k++) t += rows[k][e];
        if (i < nw) dw[i] += t;
        else db[i - nw] += t;
    }
}

// Sends an input gradient dx (bf16, [h][w][cin] of the layer) to the
// gradients of the tensors the input was made of (fp32, accumulated):
//   mode 0: the layer output x1 at the same resolution: g1 += dx
//   mode 1: x1 was the 2x2 pool of g1's tensor (one level finer, fh x fw):
//           each fine pixel gets its coarse pixel's gradient / the pool count
//   mode 2: x1 upsampled (channels [0, c1)) ++ x2 (channels [c1, cin)):
//           g2 += the x2 part; each coarse pixel of g1 sums its children's
//           x1 part
//   Every tensor holds the batch's `images` one after the other.
__global__ void k_route(const uint16_t* dx, int h, int w, int c1, int c2, int mode, float* g1, float* g2, int fh,
                        int fw, int images) {
    const int cin = c1 + c2;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 0) {
        if (i >= (size_t)images * h * w * c1) return;
        g1[i] += bf2f(dx[i]);
    } else if (mode == 1) {
        // one thread per fine pixel and channel
        const size_t fine = (size_t)fh * fw * c1;
        if (i >= images * fine) return;
        const size_t img = i / fine, r = i - img * fine;
        const size_t q = r / c1;
        const int ch = (int)(r % c1);
        const int y = (int)(q / fw), x = (int)(q % fw);
        const int py = y >> 1, px = x >> 1;
        const int ny = min(2, fh - 2 * py), nx = min(2, fw - 2 * px);
        g1[i] += bf2f(dx[((img * h + py) * w + px) * c1 + ch]) / (float)(ny * nx);
    } else {
        // one thread per pixel and channel of the layer's input
        if (i >= (size_t)images * h * w * cin) return;
        const size_t p = i / cin;  // the pixel in the batch
        const int ch = (int)(i % cin);
        if (ch >= c1) {
            g2[p * c2 + (ch - c1)] += bf2f(dx[i]);
            return;
        }
        // the coarse pixel sums its (up to 4) children once, from its
        // top-left child's thread
        const size_t img = p / ((size_t)h * w), lp = p - img * h * w;
        const int y = (int)(lp / w), x = (int)(lp % w);
        if ((y & 1) || (x & 1)) return;
        const int cw = (w + 1) / 2, chh = (h + 1) / 2;
        float s = bf2f(dx[i]);
        if (x + 1 < w) s += bf2f(dx[(p + 1) * cin + ch]);
        if (y + 1 < h) {
            s += bf2f(dx[(p + w) * cin + ch]);
            if (x + 1 < w) s += bf2f(dx[(p + w + 1) * cin + ch]);
        }
        g1[((img * chh + (y >> 1)) * cw + (x >> 1)) * c1 + ch] += s;
    }
}

__global__ void k_fill(float* p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ------------------------------------------------------------------ update
// Adam (Kingma and Ba 2015) on the fp32 master weights and biases of every
// layer, then the bf16 copies the kernels read: W[tap][cout][cin] for the
// forward pass and W'[8 - tap][cin][cout padded to 32] for the input
// gradient.  `layer_of` maps a weight index to its layer (blob order).
struct AdamArgs {
    float* w;           // master weights, all layers, blob order
    float* m;
    float* v;
    const float* g;
    uint16_t* wb;       // bf16 forward layout
    uint16_t* wt;       // bf16 flipped and transposed layout
    size_t n;
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
};
struct LayerTable {
    uint32_t w_off[kLayers + 1];  // weight index ranges per layer
    uint32_t wt_off[kLayers];
    uint16_t cin[kLayers], cout[kLayers], coutp[kLayers];
};
__global__ void k_adam(AdamArgs a, LayerTable t, int update_bias_only) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (size_t)gridDim.x * blockDim.x) {
        const float g = a.g[i];
        const float m = a.b1 * a.m[i] + (1.f - a.b1) * g;
        const float v = a.b2 * a.v[i] + (1.f - a.b2) * g * g;
        a.m[i] = m;
        a.v[i] = v;
        const float w = a.w[i] - a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
        a.w[i] = w;
        if (update_bias_only) continue;
        const uint16_t h = f2bf(w);
        a.wb[i] = h;
        int l = 0;
        while (l + 1 < kLayers && i >= t.w_off[l + 1]) l++;
        const uint32_t k = (uint32_t)(i - t.w_off[l]);
        const uint32_t ci = k % t.cin[l], co = (k / t.cin[l]) % t.cout[l], tap = k / (t.cin[l] * t.cout[l]);
        a.wt[t.wt_off[l] + ((size_t)(8 - tap) * t.cin[l] + ci) * t.coutp[l] + co] = h;
    }
}

// Input features of one image (the inference pass's k_features).
inline uint32_t grid_for(size_t n) { return (uint32_t)std::min<size_t>((n + 255) / 256, 65535u * 8); }

}  // namespace
}  // namespace vh

struct vh_denoiser_trainer {
    vh::DeviceRef device;  // first: dropped after the buffers below are freed
    vh::LayerTable table{};
    size_t nw = 0, nb = 0, nwt = 0;     // weights, biases, transposed weights (elements)
    uint32_t b_off[vh::kLayers + 1] = {};
    vh::DeviceBuffer w, mw, vw, gw, wb, wt, b, mb, vb, gb, zero_bias, loss;
    vh::DeviceBuffer act, grad, tmp;    // activations, fp32 gradients, bf16 temporaries
    vh::DeviceBuffer wpart;             // weight-gradient partial sums of k_wgrad's workgroups
    int act_w = 0, act_h = 0, act_n = 0;  // size and images the buffers hold
    size_t y_off[vh::kLayers] = {}, pool_off[4] = {}, feat_off = 0;  // in `act` (bytes)
    size_t g_off[vh::kLayers] = {};                                  // in `grad` (bytes)
    size_t dz_off = 0, dx_off = 0;                                   // in `tmp` (bytes)
    uint32_t step = 0;
    std::mutex lock;
};

namespace vh {
namespace {

// Workgroups of a k_wgrad launch (4 per CU), hence partial-sum slots.
inline int wgrad_blocks(const vh_denoiser_trainer* tr) { return 4 * tr->device->cu_count; }

int trainer_buffers(vh_denoiser_trainer* tr, int W, int H, int N) {
    if (tr->act_w == W && tr->act_h == H && tr->act_n == N) return VH_OK;
    int h[5], w[5];
    h[0] = H; w[0] = W;
    for (int l = 1; l < 5; l++) { h[l] = (h[l - 1] + 1) / 2; w[l] = (w[l - 1] + 1) / 2; }
    auto px = [&](int l) { return (size_t)N * h[l] * w[l]; };  // pixels of the batch at level l
    size_t a = 0, g = 0, t_dz = 0, t_dx = 0;
    auto take = [](size_t& off, size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    tr->feat_off = take(a, px(0) * 32 * 2);
    for (int L = 0; L < kLayers; L++) {
        const int co = kShape[L][1], ci = kShape[L][0];
        tr->y_off[L] = take(a, px(kLevel[L]) * co * 2);
        tr->g_off[L] = take(g, px(kLevel[L]) * co * 4);
        t_dz = std::max(t_dz, px(kLevel[L]) * cpad32(co) * 2);
        t_dx = std::max(t_dx, px(kLevel[L]) * ci * 2);
    }
    for (int l = 0; l < 4; l++) tr->pool_off[l] = take(a, px(l + 1) * kLevelC[l] * 2);
    size_t t = 0;
    tr->dz_off = take(t, t_dz);
    tr->dx_off = take(t, t_dx);
    VH_HIP(tr->act.allocate(a));
    VH_HIP(tr->grad.allocate(g));
    VH_HIP(tr->tmp.allocate(t));
    tr->act_w = W;
    tr->act_h = H;
    tr->act_n = N;
    return VH_OK;
}

// Forward pass of the batch (every launch over all its images), every
// activation kept.
int trainer_forward(vh_denoiser_trainer* tr, const float* film, const float* guides, float samples, hipStream_t s) {
    const int W = tr->act_w, H = tr->act_h;
    int h[5], w[5];
    h[0] = H; w[0] = W;
    for (int l = 1; l < 5; l++) { h[l] = (h[l - 1] + 1) / 2; w[l] = (w[l - 1] + 1) / 2; }
    uint8_t* base = tr->act.as<uint8_t>();
    auto at = [&](size_t o) { return reinterpret_cast<uint16_t*>(base + o); };
    const size_t px0 = (size_t)tr->act_n * W * H;
    hipLaunchKernelGGL(k_features, dim3(nblocks(px0, 256)), dim3(256), 0, s, reinterpret_cast<const float4*>(film),
                       reinterpret_cast<const float4*>(guides), samples, (int)px0, at(tr->feat_off));
    for (int L = 0; L < kLayers; L++) {
        const LayerInput& in = kInput[L];
        const int l = kLevel[L];
        ConvArgs a;
        std::memset(&a, 0, sizeof(a));
        a.x1 = in.x1 == -1 ? at(tr->feat_off) : in.x1 <= -2 ? at(tr->pool_off[-2 - in.x1]) : at(tr->y_off[in.x1]);
        a.c1 = in.x1 == -1 ? 32 : in.x1 <= -2 ? kLevelC[-2 - in.x1] : kShape[in.x1][1];
        a.up1 = in.up1;
        a.x2 = in.x2 >= 0 ? at(tr->y_off[in.x2]) : nullptr;
        a.c2 = in.x2 >= 0 ? kShape[in.x2][1] : 0;
        a.wt = tr->wb.as<uint16_t>() + tr->table.w_off[L];
        a.bias = tr->b.as<float>() + tr->b_off[L];
        a.y = at(tr->y_off[L]);
        a.h = h[l];
        a.w = w[l];
        a.relu = L < kLayers - 1 ? 1 : 0;
        a.cout = kShape[L][1];
        // the encoder's second convolutions of levels 0-3 also write their pool
        a.pool = (L % 2 == 1 && L < 8) ? at(tr->pool_off[L / 2]) : nullptr;
        a.images = tr->act_n;
        int rc = launch_conv(a, (uint32_t)tr->device->cu_count, s);
        if (rc) return rc;
    }
    VH_HIP(hipGetLastError());
    return VH_OK;
}

// Backward pass of the batch (the output gradient is in g_off[18]); adds to
// the weight and bias gradients.
int trainer_backward(vh_denoiser_trainer* tr, hipStream_t s) {
    const int W = tr->act_w, H = tr->act_h;
    int h[5], w[5];
    h[0] = H; w[0] = W;
    for (int l = 1; l < 5; l++) { h[l] = (h[l - 1] + 1) / 2; w[l] = (w[l - 1] + 1) / 2; }
    uint8_t* act = tr->act.as<uint8_t>();
    uint8_t* grad = tr->grad.as<uint8_t>();
    auto y_at = [&](int L) { return reinterpret_cast<uint16_t*>(act + tr->y_off[L]); };
    auto g_at = [&](int L) { return reinterpret_cast<float*>(grad + tr->g_off[L]); };
    uint16_t* dz = reinterpret_cast<uint16_t*>(tr->tmp.as<uint8_t>() + tr->dz_off);
    uint16_t* dx = reinterpret_cast<uint16_t*>(tr->tmp.as<uint8_t>() + tr->dx_off);
    // the other layers' output gradients start at zero (they sum contributions)
    const size_t gbytes = tr->g_off[kLayers - 1];
    hipLaunchKernelGGL(k_fill, dim3(grid_for(gbytes / 4)), dim3(256), 0, s, g_at(0), gbytes / 4, 0.f);
    for (int L = kLayers - 1; L >= 0; L--) {
        const LayerInput& in = kInput[L];
        const int l = kLevel[L], co = kShape[L][1], ci = kShape[L][0], cp = cpad32(co);
        const int N = tr->act_n;
        const size_t P = (size_t)N * h[l] * w[l];  // pixels of the batch
        hipLaunchKernelGGL(k_relu_grad, dim3(nblocks(P * cp, 256)), dim3(256), 0, s, g_at(L), y_at(L), (int)P, co, cp,
                           L < kLayers - 1 ? 1 : 0, dz);
        WgradArgs wa;
        wa.dz = dz;
        wa.x1 = in.x1 == -1 ? reinterpret_cast<uint16_t*>(act + tr->feat_off)
                : in.x1 <= -2 ? reinterpret_cast<uint16_t*>(act + tr->pool_off[-2 - in.x1]) : y_at(in.x1);
        wa.c1 = in.x1 == -1 ? 32 : in.x1 <= -2 ? kLevelC[-2 - in.x1] : kShape[in.x1][1];
        wa.up1 = in.up1;
        wa.x2 = in.x2 >= 0 ? y_at(in.x2) : nullptr;
        wa.c2 = in.x2 >= 0 ? kShape[in.x2][1] : 0;
        wa.part = tr->wpart.as<float>();
        wa.h = h[l];
        wa.w = w[l];
        wa.cout = co;
        wa.cp = cp;
        wa.images = N;
        const int ny = (co + 63) / 64, nz = ci / 32;
        const int tiles = (int)(nblocks(w[l], kTileW) * nblocks(h[l], kTileH)) * N;
        const int chunks = std::max(1, std::min(tiles, wgrad_blocks(tr) / (ny * nz)));
        hipLaunchKernelGGL(k_wgrad, dim3(chunks, ny, nz), dim3(256), 0, s, wa);
        const size_t nred = (size_t)9 * co * ci + co;
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(nblocks(nred, 32)), dim3(256), 0, s, tr->wpart.as<float>(), chunks,
                           ny, nz, co, ci, tr->gw.as<float>() + tr->table.w_off[L], tr->gb.as<float>() + tr->b_off[L]);
        if (L == 0) break;  // the features need no gradient
        // input gradient: the convolution of dz with the flipped, transposed weights
        ConvArgs a;
        std::memset(&a, 0, sizeof(a));
        a.x1 = dz;
        a.c1 = cp;
        a.wt = tr->wt.as<uint16_t>() + tr->table.wt_off[L];
        a.bias = tr->zero_bias.as<float>();
        a.y = dx;
        a.h = h[l];
        a.w = w[l];
        a.relu = 0;
        a.cout = ci;
        a.images = N;
        const bool odd = ci != 16 && ci != 32 && ci != 64 && ci != 96 && ci != 128;
        int rc = launch_conv(a, (uint32_t)tr->device->cu_count, s, odd);
        if (rc) return rc;
        // route it to where the input came from
        if (in.x1 <= -2) {  // a pool of the finer level's encoder output (layer 2 (l - 1) + 1)
            const int src = 2 * (l - 1) + 1, fh = h[l - 1], fw = w[l - 1];
            const size_t n = (size_t)N * fh * fw * ci;
            hipLaunchKernelGGL(k_route, dim3(nblocks(n, 256)), dim3(256), 0, s, dx, h[l], w[l], ci, 0, 1, g_at(src),
                               (float*)nullptr, fh, fw, N);
        } else if (in.up1) {
            const int c1 = kShape[in.x1][1], c2 = kShape[in.x2][1];
            hipLaunchKernelGGL(k_route, dim3(nblocks(P * ci, 256)), dim3(256), 0, s, dx, h[l], w[l], c1, c2, 2,
                               g_at(in.x1), g_at(in.x2), 0, 0, N);
        } else {
            hipLaunchKernelGGL(k_route, dim3(nblocks(P * ci, 256)), dim3(256), 0, s, dx, h[l], w[l], ci, 0, 0,
                               g_at(in.x1), (float*)nullptr, 0, 0, N);
        }
    }
    VH_HIP(hipGetLastError());
    return VH_OK;
}

// Host-side initial weights: the default of a 3x3 convolution layer (weights
// and biases uniform in +-1 / sqrt(fan in), fan in = 9 x the layer's real
// input channels), padded channels zero, from a seeded splitmix64 stream.
uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace
}  // namespace vh

using namespace vh;

extern "C" {

int vh_trainer_create(vh_device* dev, const void* blob, size_t bytes, uint32_t seed, vh_denoiser_trainer** out) {
    if (!dev || !out) return fail(VH_ERR_USAGE, "vh_trainer_create: null argument");
    *out = nullptr;
    LayerTable t{};
    uint32_t b_off[kLayers + 1];
    size_t nw = 0, nb = 0, nwt = 0;
    for (int L = 0; L < kLayers; L++) {
        t.w_off[L] = (uint32_t)nw;
        t.wt_off[L] = (uint32_t)nwt;
        b_off[L] = (uint32_t)nb;
        t.cin[L] = (uint16_t)kShape[L][0];
        t.cout[L] = (uint16_t)kShape[L][1];
        t.coutp[L] = (uint16_t)cpad32(kShape[L][1]);
        nw += (size_t)9 * kShape[L][0] * kShape[L][1];
        nwt += (size_t)9 * kShape[L][0] * cpad32(kShape[L][1]);
        nb += kShape[L][1];
    }
    t.w_off[kLayers] = (uint32_t)nw;
    b_off[kLayers] = (uint32_t)nb;
    std::vector<float> w(nw, 0.f), b(nb, 0.f);
    if (blob) {
        // a version-1 blob (the layout vh_denoiser_create reads): bf16 weights, fp32 biases
        const uint8_t* p = static_cast<const uint8_t*>(blob);
        auto u32 = [&](size_t o) { uint32_t v; std::memcpy(&v, p + o, 4); return v; };
        if (bytes < 12 || std::memcmp(p, "VHDN", 4) != 0 || u32(4) != 1 || u32(8) != (uint32_t)kLayers)
            return fail(VH_ERR_USAGE, "vh_trainer_create: not a version-1 weight blob of the 19-layer U-Net");
        size_t o = 12;
        for (int L = 0; L < kLayers; L++) {
            if (o + 8 > bytes || (int)u32(o) != kShape[L][0] || (int)u32(o + 4) != kShape[L][1])
                return fail(VH_ERR_USAGE, "vh_trainer_create: layer " + std::to_string(L) + " does not match");
            o += 8;
            const size_t k = (size_t)9 * kShape[L][0] * kShape[L][1];
            if (o + k * 2 + kShape[L][1] * 4 > bytes) return fail(VH_ERR_USAGE, "vh_trainer_create: truncated blob");
            for (size_t i = 0; i < k; i++) {
                uint16_t h;
                std::memcpy(&h, p + o + 2 * i, 2);
                const uint32_t u = (uint32_t)h << 16;
                std::memcpy(&w[t.w_off[L] + i], &u, 4);
            }
            o += k * 2;
            std::memcpy(&b[b_off[L]], p + o, kShape[L][1] * 4);
            o += kShape[L][1] * 4;
        }
        if (o != bytes) return fail(VH_ERR_USAGE, "vh_trainer_create: trailing bytes after the last layer");
    } else {
        uint64_t st = 0x5eed0000ull + seed;
        for (int L = 0; L < kLayers; L++) {
            const int real_in = L == 0 ? 10 : kShape[L][0], real_out = L == kLayers - 1 ? 3 : kShape[L][1];
            const float bound = 1.f / std::sqrt(9.f * real_in);
            for (int tap = 0; tap < 9; tap++)
                for (int co = 0; co < kShape[L][1]; co++)
                    for (int ci = 0; ci < kShape[L][0]; ci++) {
                        const float u = (float)((splitmix(st) >> 11) * (1.0 / 9007199254740992.0));
                        if (co < real_out && ci < real_in)
                            w[t.w_off[L] + ((size_t)tap * kShape[L][1] + co) * kShape[L][0] + ci] =
                                (2.f * u - 1.f) * bound;
                    }
            for (int co = 0; co < real_out; co++)
                b[b_off[L] + co] = (2.f * (float)((splitmix(st) >> 11) * (1.0 / 9007199254740992.0)) - 1.f) * bound;
        }
    }
    VH_HIP(hipSetDevice(dev->ordinal));
    vh_denoiser_trainer* tr = new (std::nothrow) vh_denoiser_trainer();
    if (!tr) return fail(VH_ERR_DEVICE, "out of host memory");
    tr->device = dev;
    tr->table = t;
    tr->nw = nw;
    tr->nb = nb;
    tr->nwt = nwt;
    std::memcpy(tr->b_off, b_off, sizeof(b_off));
    auto bail = [&](hipError_t e) {
        delete tr;
        return hip_fail(e, "vh_trainer_create");
    };
    hipError_t e = hipSuccess;
    for (DeviceBuffer* d : {&tr->w, &tr->mw, &tr->vw, &tr->gw})
        if (e == hipSuccess) e = d->allocate(nw * 4);
    for (DeviceBuffer* d : {&tr->b, &tr->mb, &tr->vb, &tr->gb})
        if (e == hipSuccess) e = d->allocate(nb * 4);
    if (e == hipSuccess) e = tr->wb.allocate(nw * 2);
    if (e == hipSuccess) e = tr->wt.allocate(nwt * 2);
    if (e == hipSuccess) e = tr->zero_bias.allocate(256 * 4);
    if (e == hipSuccess) e = tr->loss.allocate(16);
    if (e == hipSuccess) e = tr->wpart.allocate((size_t)wgrad_blocks(tr) * kWgradPart * sizeof(float));
    if (e != hipSuccess) return bail(e);
    hipStream_t s = dev->stream;
    VH_HIP(hipMemcpyAsync(tr->w.ptr, w.data(), nw * 4, hipMemcpyHostToDevice, s));
    VH_HIP(hipMemcpyAsync(tr->b.ptr, b.data(), nb * 4, hipMemcpyHostToDevice, s));
    for (DeviceBuffer* d : {&tr->mw, &tr->vw, &tr->gw})
        hipLaunchKernelGGL(k_fill, dim3(grid_for(nw)), dim3(256), 0, s, d->as<float>(), nw, 0.f);
    for (DeviceBuffer* d : {&tr->mb, &tr->vb, &tr->gb})
        hipLaunchKernelGGL(k_fill, dim3(grid_for(nb)), dim3(256), 0, s, d->as<float>(), nb, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, tr->zero_bias.as<float>(), (size_t)256, 0.f);
    // the flipped copy's padded output channels stay zero
    VH_HIP(hipMemsetAsync(tr->wt.ptr, 0, nwt * 2, s));
    // bf16 copies of the initial weights: an Adam step with zero gradient and lr 0
    AdamArgs a{tr->w.as<float>(), tr->mw.as<float>(), tr->vw.as<float>(), tr->gw.as<float>(), tr->wb.as<uint16_t>(),
               tr->wt.as<uint16_t>(), nw, 0.f, 0.9f, 0.999f, 1e-8f, 1.f, 1.f};
    hipLaunchKernelGGL(k_adam, dim3(grid_for(nw)), dim3(256), 0, s, a, t, 0);
    VH_HIP(hipGetLastError());
    VH_HIP(hipStreamSynchronize(s));
    *out = tr;
    return VH_OK;
}

int vh_trainer_step(vh_denoiser_trainer* tr, uint32_t images, uint32_t width, uint32_t height, float samples,
                    const float* films, const float* guides, const float* targets, float learning_rate,
                    uint32_t flags, void* stream, float* loss) {
    if (!tr || !films || !guides || !targets || images == 0) return fail(VH_ERR_USAGE, "vh_trainer_step: null argument");
    if (width < 2 || height < 2 || (uint64_t)width * height > (1ull << 22))
        return fail(VH_ERR_USAGE, "vh_trainer_step: image size out of range (2x2 to 4 Mpixels)");
    if ((uint64_t)images * width * height > (1ull << 25))
        return fail(VH_ERR_USAGE, "vh_trainer_step: a batch holds at most 2^25 pixels");
    if (!(samples > 0.f) || !std::isfinite(learning_rate) || learning_rate < 0.f)
        return fail(VH_ERR_USAGE, "vh_trainer_step: samples must be > 0 and the learning rate finite and >= 0");
    vh_device* dev = tr->device;
    VH_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : dev->stream;
    std::lock_guard<std::mutex> hold(tr->lock);
    int rc = trainer_buffers(tr, (int)width, (int)height, (int)images);
    if (rc) return rc;
    const size_t P = (size_t)images * width * height;  // pixels of the batch
    hipLaunchKernelGGL(k_fill, dim3(grid_for(tr->nw)), dim3(256), 0, s, tr->gw.as<float>(), tr->nw, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(grid_for(tr->nb)), dim3(256), 0, s, tr->gb.as<float>(), tr->nb, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, tr->loss.as<float>(), (size_t)1, 0.f);
    const float inv_count = 1.f / (float)((double)P * 3.0);
    if ((rc = trainer_forward(tr, films, guides, samples, s))) return rc;
    const uint16_t* y18 = reinterpret_cast<const uint16_t*>(tr->act.as<uint8_t>() + tr->y_off[kLayers - 1]);
    float* g18 = reinterpret_cast<float*>(tr->grad.as<uint8_t>() + tr->g_off[kLayers - 1]);
    hipLaunchKernelGGL(k_loss, dim3(nblocks(P, 256)), dim3(256), 0, s, y18, targets, (int)P, inv_count, g18,
                       tr->loss.as<float>());
    if ((rc = trainer_backward(tr, s))) return rc;
    if (!(flags & VH_TRAIN_NO_UPDATE)) {
        tr->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, tr->step)));
        const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, tr->step)));
        AdamArgs aw{tr->w.as<float>(), tr->mw.as<float>(), tr->vw.as<float>(), tr->gw.as<float>(),
                    tr->wb.as<uint16_t>(), tr->wt.as<uint16_t>(), tr->nw, learning_rate, b1, b2, 1e-8f, c1, c2};
        hipLaunchKernelGGL(k_adam, dim3(grid_for(tr->nw)), dim3(256), 0, s, aw, tr->table, 0);
        AdamArgs ab{tr->b.as<float>(), tr->mb.as<float>(), tr->vb.as<float>(), tr->gb.as<float>(), nullptr, nullptr,
                    tr->nb, learning_rate, b1, b2, 1e-8f, c1, c2};
        hipLaunchKernelGGL(k_adam, dim3(grid_for(tr->nb)), dim3(256), 0, s, ab, tr->table, 1);
    }
    VH_HIP(hipGetLastError());
    float l = 0.f;
    VH_HIP(hipMemcpyAsync(&l, tr->loss.ptr, 4, hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l;
    if (!std::isfinite(l)) return fail(VH_ERR_DEVICE, "vh_trainer_step: non-finite loss");
    return VH_OK;
}

int vh_trainer_gradients(vh_denoiser_trainer* tr, float* weights, size_t weight_count, float* biases,
                         size_t bias_count) {
    if (!tr || !weights || !biases) return fail(VH_ERR_USAGE, "vh_trainer_gradients: null argument");
    if (weight_count != tr->nw || bias_count != tr->nb)
        return fail(VH_ERR_USAGE, "vh_trainer_gradients: counts must be " + std::to_string(tr->nw) + " and " +
                                      std::to_string(tr->nb));
    VH_HIP(hipSetDevice(tr->device->ordinal));
    std::lock_guard<std::mutex> hold(tr->lock);
    VH_HIP(hipMemcpy(weights, tr->gw.ptr, tr->nw * 4, hipMemcpyDeviceToHost));
    VH_HIP(hipMemcpy(biases, tr->gb.ptr, tr->nb * 4, hipMemcpyDeviceToHost));
    return VH_OK;
}

int vh_trainer_export(vh_denoiser_trainer* tr, void* blob, size_t capacity, size_t* bytes) {
    if (!tr || !bytes) return fail(VH_ERR_USAGE, "vh_trainer_export: null argument");
    size_t need = 12;
    for (int L = 0; L < kLayers; L++) need += 8 + (size_t)9 * kShape[L][0] * kShape[L][1] * 2 + kShape[L][1] * 4;
    *bytes = need;
    if (!blob) return VH_OK;
    if (capacity < need) return fail(VH_ERR_USAGE, "vh_trainer_export: blob buffer too small");
    VH_HIP(hipSetDevice(tr->device->ordinal));
    std::lock_guard<std::mutex> hold(tr->lock);
    std::vector<uint16_t> wb(tr->nw);
    std::vector<float> b(tr->nb);
    VH_HIP(hipMemcpy(wb.data(), tr->wb.ptr, tr->nw * 2, hipMemcpyDeviceToHost));
    VH_HIP(hipMemcpy(b.data(), tr->b.ptr, tr->nb * 4, hipMemcpyDeviceToHost));
    uint8_t* p = static_cast<uint8_t*>(blob);
    const uint32_t head[3] = {0, 1, (uint32_t)kLayers};
    std::memcpy(p, "VHDN", 4);
    std::memcpy(p + 4, &head[1], 8);
    size_t o = 12;
    for (int L = 0; L < kLayers; L++) {
        const uint32_t sh[2] = {(uint32_t)kShape[L][0], (uint32_t)kShape[L][1]};
        std::memcpy(p + o, sh, 8);
        o += 8;
        const size_t k = (size_t)9 * kShape[L][0] * kShape[L][1];
        std::memcpy(p + o, wb.data() + tr->table.w_off[L], k * 2);
        o += k * 2;
        std::memcpy(p + o, b.data() + tr->b_off[L], kShape[L][1] * 4);
        o += kShape[L][1] * 4;
    }
    return VH_OK;
}

void vh_trainer_release(vh_denoiser_trainer* tr) {
    if (!tr) return;
    (void)hipSetDevice(tr->device->ordinal);
    (void)hipStreamSynchronize(tr->device->stream);
    delete tr;
}

}  // extern "C"
