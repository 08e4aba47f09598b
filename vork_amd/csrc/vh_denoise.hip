// Learned denoiser of the path tracer's film (replaces the OIDN pass of
// crates/vork-render/src/denoise.rs): a 5-level U-Net of 3x3 convolutions in
// bf16 on MFMA with fp32 accumulation.  The convolution kernel and the layer
// table are in vh_conv.h; this file holds the inference pass and the
// denoiser object of the C ABI.
#include <cmath>
#include <cstring>
#include <vector>

#include "vh_conv.h"
#include "vh_internal.h"

struct vh_denoiser {
    vh::DeviceRef device;  // first: dropped after the buffers below are freed
    vh::DeviceBuffer weights, biases, scratch;
    size_t w_off[vh::kLayers] = {}, b_off[vh::kLayers] = {};
    size_t scratch_bytes = 0;
    std::mutex lock;       // one denoise at a time per object (shared scratch)
};

namespace vh {

int denoise_run(vh_denoiser* dn, int W, int H, float samples, const float* film, const float* guides, float* out,
                hipStream_t s) {
    int h[5], w[5];
    h[0] = H; w[0] = W;
    for (int l = 1; l < 5; l++) { h[l] = (h[l - 1] + 1) / 2; w[l] = (w[l - 1] + 1) / 2; }
    const int C[5] = {32, 64, 64, 96, 128};
    // scratch: a temporary and a pooled input at the largest level used,
    // every encoder output and every decoder output
    auto px = [&](int l) { return (size_t)h[l] * w[l]; };
    size_t off = 0;
    auto take = [&](size_t elems) { size_t o = off; off += (elems * 2 + 255) & ~(size_t)255; return o; };
    const size_t o_tmp = take(px(0) * 128), o_pool = take(px(1) * 96);
    size_t o_enc[5], o_dec[4];
    for (int l = 0; l < 5; l++) o_enc[l] = take(px(l) * C[l]);
    for (int l = 0; l < 4; l++) o_dec[l] = take(px(l) * C[l]);
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
                    uint16_t* pool = nullptr, float4* radiance = nullptr) {
        ConvArgs a;
        a.radiance = radiance;
        a.x1 = x1; a.x2 = x2; a.c1 = c1; a.c2 = c2; a.up1 = up1; a.pool = pool;
        a.wt = wts + dn->w_off[layer]; a.bias = bs + dn->b_off[layer];
        a.y = y; a.h = h[l]; a.w = w[l]; a.relu = relu ? 1 : 0;
        a.cout = kShape[layer][1];
        a.images = 1;
        layer++;
        return launch_conv(a, (uint32_t)dn->device->cu_count, s);
    };
    // encoder; the first layer computes its input features from the film and
    // guides while it stages them
    const uint16_t* in = nullptr;
    int cin = 32;
    for (int l = 0; l < 5; l++) {
        if (l == 0) {
            ConvArgs a;
            a.x1 = nullptr; a.x2 = nullptr; a.c1 = 32; a.c2 = 0; a.up1 = 0; a.pool = nullptr;
            a.wt = wts + dn->w_off[layer]; a.bias = bs + dn->b_off[layer];
            a.y = at(o_tmp); a.h = h[0]; a.w = w[0]; a.relu = 1;
            a.cout = kShape[layer][1];
            a.images = 1;
            a.film = reinterpret_cast<const float4*>(film);
            a.guides = reinterpret_cast<const float4*>(guides);
            a.feat_samples = samples;
            layer++;
            if ((rc = launch_conv(a, (uint32_t)dn->device->cu_count, s))) return rc;
        } else if ((rc = conv(l, in, cin, 0, nullptr, 0, at(o_tmp), true))) {
            return rc;
        }
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
    // the output layer writes the clean radiance from its epilogue
    if ((rc = conv(0, at(o_dec[0]), 32, 0, nullptr, 0, nullptr, false, nullptr, reinterpret_cast<float4*>(out))))
        return rc;
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
