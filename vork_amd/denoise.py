"""Learned denoiser of the path tracer's film: the U-Net the HIP kernels of
vork_amd/csrc/vh_denoise.hip run (replacing the OIDN pass of
crates/vork-render/src/denoise.rs, whose inputs are the colour and the
albedo/normal guides of path.slang:1553-1575).

Input: 10 channels per pixel, the log1p of the mean radiance of a few
samples, the mean guide albedo, normal and depth.  Network: 5 levels of
32, 64, 64, 96, 128 channels, two 3x3 convolutions with ReLU per level, 2x2
average pooling down, nearest upsampling and skip concatenation up, and a 3x3
output convolution predicting the log1p of the clean radiance.  Weights and
activations are bf16 on the device (fp32 accumulation).

`UNet` is the torch definition, used to train on pairs the path tracer renders
(`train`) and, with bf16 rounding between layers, as the fp32 reference of the
HIP kernels.  `export` writes the weight blob vh_denoiser_create reads;
`Denoiser` runs the HIP network through the C ABI.
"""
import ctypes as C
import os
import struct

import numpy as np

LEVELS = (32, 64, 64, 96, 128)
FEATURES = 10
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "denoise_weights.vhdn")


def _torch():
    import torch
    return torch


def make_unet():
    torch = _torch()
    nn = torch.nn

    class UNet(nn.Module):
        def __init__(self):
            super().__init__()
            c = LEVELS
            self.enc = nn.ModuleList()
            cin = FEATURES
            for l in range(5):
                self.enc.append(nn.ModuleList([nn.Conv2d(cin, c[l], 3, padding=1), nn.Conv2d(c[l], c[l], 3, padding=1)]))
                cin = c[l]
            self.dec = nn.ModuleList()
            for l in (3, 2, 1, 0):
                up = c[l + 1]
                self.dec.append(nn.ModuleList([nn.Conv2d(up + c[l], c[l], 3, padding=1),
                                               nn.Conv2d(c[l], c[l], 3, padding=1)]))
            self.out = nn.Conv2d(c[0], 3, 3, padding=1)

        def convs(self):
            """The convolutions in execution order."""
            out = []
            for a, b in self.enc:
                out += [a, b]
            for a, b in self.dec:
                out += [a, b]
            return out + [self.out]

        def forward(self, x, bf16=False):
            F = torch.nn.functional
            r = (lambda t: t.to(torch.bfloat16).to(torch.float32)) if bf16 else (lambda t: t)

            def conv(m, t, relu=True):
                w = r(m.weight)
                y = F.conv2d(t, w, m.bias, padding=1)
                return r(torch.relu(y) if relu else y)

            x = r(x)
            skips = []
            for l, (a, b) in enumerate(self.enc):
                if l > 0:
                    x = r(F.avg_pool2d(x, 2, 2, ceil_mode=True))
                x = conv(b, conv(a, x))
                skips.append(x)
            for k, (a, b) in enumerate(self.dec):
                skip = skips[3 - k]
                up = F.interpolate(x, scale_factor=2, mode="nearest")[..., :skip.shape[-2], :skip.shape[-1]]
                x = conv(b, conv(a, torch.cat([up, skip], 1)))
            return conv(self.out, x, relu=False)

    return UNet()


def features(film, guides, samples):
    """Network input (N, 10, H, W) from film sums (N, H, W, 4) and guide sums
    (N, H, W, 2, 4) of `samples` samples (torch tensors)."""
    torch = _torch()
    inv = 1.0 / samples
    rad = torch.log1p(torch.clamp(film[..., :3] * inv, min=0.0))
    g = guides * inv
    x = torch.cat([rad, g[..., 0, :3], g[..., 1, :3], g[..., 1, 3:4]], -1)
    return x.permute(0, 3, 1, 2).contiguous()


def radiance(pred):
    """Clean radiance (N, 3, H, W) from the network's log1p prediction."""
    torch = _torch()
    return torch.expm1(torch.clamp(pred, min=0.0))


def _bf16_bits(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def export(model):
    """The weight blob of vh_denoiser_create: "VHDN", version 1, 19 layers,
    then per layer cin, cout, bf16 weights [tap][cout][cin] (input channels
    padded to 32 on the first layer, outputs to 16 on the last) and fp32
    biases."""
    convs = model.convs()
    out = [b"VHDN", struct.pack("<II", 1, len(convs))]
    for k, m in enumerate(convs):
        w = m.weight.detach().float().cpu().numpy()           # (cout, cin, 3, 3)
        b = m.bias.detach().float().cpu().numpy()
        cout, cin = w.shape[:2]
        pc_in = 32 if k == 0 else cin
        pc_out = 16 if k == len(convs) - 1 else cout
        wp = np.zeros((pc_out, pc_in, 3, 3), np.float32)
        wp[:cout, :cin] = w
        bp = np.zeros(pc_out, np.float32)
        bp[:cout] = b
        taps = wp.transpose(2, 3, 0, 1).reshape(9, pc_out, pc_in)  # [tap = 3 dy + dx][cout][cin]
        out += [struct.pack("<II", pc_in, pc_out), _bf16_bits(taps).tobytes(), bp.astype("<f4").tobytes()]
    return b"".join(out)


def load_model(path=WEIGHTS):
    """The torch model with the weights of a blob written by `export`."""
    torch = _torch()
    model = make_unet()
    with open(path, "rb") as f:
        blob = f.read()
    assert blob[:4] == b"VHDN"
    o = 12
    for k, m in enumerate(model.convs()):
        ci, co = struct.unpack_from("<II", blob, o)
        o += 8
        w = np.frombuffer(blob, np.uint16, 9 * co * ci, o).astype(np.uint32) << 16
        o += 9 * co * ci * 2
        b = np.frombuffer(blob, "<f4", co, o)
        o += co * 4
        w = w.view(np.float32).reshape(3, 3, co, ci).transpose(2, 3, 0, 1)
        cout, cin = m.weight.shape[:2]
        with torch.no_grad():
            m.weight.copy_(torch.from_numpy(np.array(w[:cout, :cin])))
            m.bias.copy_(torch.from_numpy(np.array(b[:cout])))
    return model


class Denoiser:
    """The HIP U-Net on a device (vh_denoiser_*)."""

    def __init__(self, device, blob=None):
        from . import native
        self.native = native
        self.device = device
        if blob is None:
            with open(WEIGHTS, "rb") as f:
                blob = f.read()
        self._blob = C.create_string_buffer(blob, len(blob))
        self.handle = C.c_void_p()
        native.check(native.lib().vh_denoiser_create(device.handle, self._blob, len(blob), C.byref(self.handle)))

    def __call__(self, film, guides, samples, out=None, stream=None):
        """Denoised mean radiance (H, W, 4; alpha 1) of film sums (H, W, 4) and
        guide sums (H, W, 2, 4) of `samples` samples: numpy arrays, or torch
        CUDA tensors (then `out` is one too)."""
        native = self.native
        h, w = film.shape[:2]
        s = C.c_void_p(stream) if stream else None
        if isinstance(film, np.ndarray):
            film = np.ascontiguousarray(film, np.float32)
            guides = np.ascontiguousarray(guides, np.float32)
            out = np.zeros((h, w, 4), np.float32) if out is None else out
            native.check(native.lib().vh_denoise(self.handle, w, h, float(samples), native._ptr(film),
                                                 native._ptr(guides), 0, s, native._ptr(out)))
            return out
        if out is None:
            out = film.new_empty((h, w, 4))
        native.check(native.lib().vh_denoise(self.handle, w, h, float(samples), native._dev_ptr(film),
                                             native._dev_ptr(guides), native.abi.VH_DEVICE_POINTERS, s,
                                             native._dev_ptr(out)))
        return out

    def close(self):
        if self.handle:
            self.native.lib().vh_denoiser_release(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Trainer:
    """The denoiser's hand-written training kernels on a device
    (vh_trainer_*): forward pass, loss, backward pass and Adam in HIP."""

    def __init__(self, device, blob=None, seed=0):
        from . import native
        self.native = native
        self.device = device
        self._blob = C.create_string_buffer(blob, len(blob)) if blob is not None else None
        self.handle = C.c_void_p()
        native.check(native.lib().vh_trainer_create(device.handle, self._blob, len(blob) if blob else 0, seed,
                                                    C.byref(self.handle)))

    def step(self, films, guides, targets, samples, lr, update=True, stream=None):
        """One step over a batch: torch CUDA tensors films (N, H, W, 4) and
        guides (N, H, W, 2, 4) of `samples` samples, targets (N, H, W, 3) mean
        radiance.  Returns the loss."""
        native = self.native
        n, h, w = films.shape[:3]
        for t in (films, guides, targets):
            assert t.is_cuda and t.is_contiguous() and t.dtype == _torch().float32
        loss = C.c_float()
        native.check(native.lib().vh_trainer_step(
            self.handle, n, w, h, float(samples), native._dev_ptr(films), native._dev_ptr(guides),
            native._dev_ptr(targets), float(lr), 0 if update else 1, C.c_void_p(stream) if stream else None,
            C.byref(loss)))
        return loss.value

    def gradients(self):
        """(weights, biases) fp32 gradients of the last step, blob order."""
        convs = make_unet().convs()
        nw = sum(9 * (32 if k == 0 else m.in_channels) * (16 if k == len(convs) - 1 else m.out_channels)
                 for k, m in enumerate(convs))
        nb = sum(16 if k == len(convs) - 1 else m.out_channels for k, m in enumerate(convs))
        gw, gb = np.zeros(nw, np.float32), np.zeros(nb, np.float32)
        self.native.check(self.native.lib().vh_trainer_gradients(self.handle, gw.ctypes.data, nw, gb.ctypes.data,
                                                                 nb))
        return gw, gb

    def export(self):
        native = self.native
        n = C.c_size_t()
        native.check(native.lib().vh_trainer_export(self.handle, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        native.check(native.lib().vh_trainer_export(self.handle, buf, n.value, C.byref(n)))
        return buf.raw

    def close(self):
        if self.handle:
            self.native.lib().vh_trainer_release(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def split_gradients(gw, gb):
    """Per layer (weight gradient (cout, cin, 3, 3) at padded sizes, bias
    gradient) from the blob-order arrays of Trainer.gradients."""
    out, ow, ob = [], 0, 0
    convs = make_unet().convs()
    for k, m in enumerate(convs):
        ci = 32 if k == 0 else m.in_channels
        co = 16 if k == len(convs) - 1 else m.out_channels
        w = gw[ow:ow + 9 * ci * co].reshape(3, 3, co, ci).transpose(2, 3, 0, 1)
        out.append((w, gb[ob:ob + co]))
        ow += 9 * ci * co
        ob += co
    return out


def loss_fn(pred, target):
    """The training loss: relative MSE of the radiance plus L1 of log1p, the
    kernels' k_loss in torch (pred (N, 3, H, W) log radiance)."""
    torch = _torch()
    d = radiance(pred)
    return ((d - target) ** 2 / (target ** 2 + 1e-2)).mean() + (pred - torch.log1p(target)).abs().mean()


HBM_BYTES_PER_S = 8.0e12      # MI355X HBM3E
BF16_FLOPS_PER_S = 2.5e15     # MI355X dense bf16 MFMA


def layer_costs(width, height):
    """Per launch of one denoise at width x height: (name, flops, bytes) with
    the multiply-adds x 2 each convolution runs (padded channel counts) and
    the least HBM bytes it must move: its inputs read once (the decoder's
    coarse input at its own resolution), its bf16 output written once (and
    the encoder's 2x2 pool), its weights once.  The first layer reads film and
    guides; the output layer writes the fp32 radiance from its epilogue."""
    h, w = [height], [width]
    for _ in range(4):
        h.append((h[-1] + 1) // 2)
        w.append((w[-1] + 1) // 2)
    px = [h[l] * w[l] for l in range(5)]
    c = LEVELS
    out = []

    def conv(name, l, c1, c2, cout, up=False, pool=False):
        cin = c1 + c2
        fl = 2 * px[l] * 9 * cin * cout
        # the first layer reads the film and guide sums (48 B per pixel) and
        # computes its input features while staging them
        inp = px[0] * 48 if name == "enc0a" else (px[l + 1] if up else px[l]) * c1 * 2 + px[l] * c2 * 2
        by = inp + 9 * cin * cout * 2 + px[l] * cout * 2
        if pool:
            by += px[l + 1] * cout * 2
        out.append((name, fl, by))

    cin = 32
    for l in range(5):
        conv(f"enc{l}a", l, cin, 0, c[l])
        conv(f"enc{l}b", l, c[l], 0, c[l], pool=l < 4)
        cin = c[l]
    for l in (3, 2, 1, 0):
        conv(f"dec{l}a", l, c[l + 1], c[l], c[l], up=True)
        conv(f"dec{l}b", l, c[l], 0, c[l])
    # the output layer writes the fp32 radiance (16 B per pixel) from its epilogue
    cin = c[0]
    out.append(("out", 2 * px[0] * 9 * cin * 16, px[0] * cin * 2 + 9 * cin * 16 * 2 + px[0] * 16))
    return out


def speed_of_light_ms(width, height):
    """Least time of one denoise: every launch at the larger of its MFMA time
    and its HBM time."""
    return sum(max(f / BF16_FLOPS_PER_S, b / HBM_BYTES_PER_S) for _, f, b in layer_costs(width, height)) * 1e3


def flops(width, height):
    """Multiply-adds x 2 of one denoise (the padded channel counts the kernels run)."""
    h, w = [height], [width]
    for _ in range(4):
        h.append((h[-1] + 1) // 2)
        w.append((w[-1] + 1) // 2)
    c = LEVELS
    total = 0
    cin = 32
    for l in range(5):
        total += h[l] * w[l] * 9 * (cin * c[l] + c[l] * c[l])
        cin = c[l]
    for l in (3, 2, 1, 0):
        total += h[l] * w[l] * 9 * ((c[l + 1] + c[l]) * c[l] + c[l] * c[l])
    total += h[0] * w[0] * 9 * c[0] * 16
    return 2 * total
