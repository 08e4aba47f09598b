"""The denoiser's host side on the CPU: the exported weight blob has the layout
vh_denoiser_create reads (sizes, padding, bf16 rounding) and loads back into
the same network; the network runs on odd image sizes; the flop count
follows the layer table."""
import struct

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from vork_amd import denoise as D  # noqa: E402


def test_blob_layout_and_round_trip(tmp_path):
    torch.manual_seed(0)
    m = D.make_unet()
    blob = D.export(m)
    assert blob[:4] == b"VHDN" and struct.unpack_from("<II", blob, 4) == (1, 19)
    shapes, o = [], 12
    for _ in range(19):
        ci, co = struct.unpack_from("<II", blob, o)
        shapes.append((ci, co))
        o += 8 + 9 * ci * co * 2 + co * 4
    assert o == len(blob)
    assert shapes[0] == (32, 32) and shapes[-1] == (32, 16) and shapes[10] == (224, 96)
    p = tmp_path / "w.vhdn"
    p.write_bytes(blob)
    back = D.load_model(str(p))
    for a, b in zip(m.convs(), back.convs()):
        wa = a.weight.detach().to(torch.bfloat16).float()
        assert torch.equal(wa, b.weight.detach()) and torch.equal(a.bias.detach(), b.bias.detach())


def test_bf16_bits_round_to_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -2.5, 3.0e38, 1e-40], np.float32)
    bits = D._bf16_bits(x)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().astype(np.uint16)
    assert np.array_equal(bits, want)


def test_network_runs_on_odd_sizes_and_counts_flops():
    m = D.make_unet().eval()
    x = torch.randn(1, D.FEATURES, 37, 53)
    with torch.no_grad():
        y = m(x)
        yb = m(x, bf16=True)
    assert y.shape == (1, 3, 37, 53) and yb.shape == y.shape
    assert torch.allclose(y, yb, atol=0.1)
    # by hand at 1920x1080, multiply-adds per level (pixels x 9 x sum cin*cout):
    # 2073600 x 9 x 6656 + 518400 x 9 x 18432 + 129600 x 9 x 22528
    # + 32400 x 9 x 46080 + 8160 x 9 x 28672 = 252.03e9
    assert D.flops(1920, 1080) == 2 * (2073600 * 9 * 6656 + 518400 * 9 * 18432 + 129600 * 9 * 22528
                                       + 32400 * 9 * 46080 + 8160 * 9 * 28672)


def test_features_follow_the_guide_layout():
    film = torch.zeros(1, 2, 2, 4)
    film[..., :3] = 4.0
    g = torch.zeros(1, 2, 2, 2, 4)
    g[..., 0, :3] = 2.0
    g[..., 1, :3] = torch.tensor([0.0, 4.0, 0.0])
    g[..., 1, 3] = 2.0
    x = D.features(film, g, 4.0)
    assert x.shape == (1, 10, 2, 2)
    assert torch.allclose(x[0, :3], torch.log1p(torch.tensor(1.0)).expand(3, 2, 2))
    assert torch.allclose(x[0, 3:6], torch.full((3, 2, 2), 0.5))
    assert torch.allclose(x[0, 7], torch.ones(2, 2)) and torch.allclose(x[0, 9], torch.full((2, 2), 0.5))
