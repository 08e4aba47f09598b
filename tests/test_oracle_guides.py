"""Pins of the oracle's denoiser guides (path.slang:1553-1575 over this
closure subset) by closed forms: the guide albedo of a white material is 1
for a Lambertian without specular layer, for a white dielectric
((1 - E) + E) and for a white compensated metal (energy x 1 / energy); a
coloured Lambertian gives its base colour; coverage counts the camera hits;
one sample's shading normal has unit length; depth t / (t + 1) lies in
(0, 1) on a hit and is 1 on a miss."""
import numpy as np
import pytest

import oracle
from vork_amd import scene as S


def guides(sc, spp=1):
    p = sc.render_params(samples=spp, adaptive=False)
    return oracle.OracleScene(sc).render_guides(p)[1]


def with_materials(sc, **m):
    sc.materials = [dict(mat, **m) for mat in sc.materials]
    return sc


@pytest.mark.parametrize("material", [
    dict(base_color=[1.0, 1.0, 1.0], base_metalness=0.0, specular_weight=0.0),
    dict(base_color=[1.0, 1.0, 1.0], base_metalness=0.0, specular_weight=1.0, specular_roughness=0.3),
    dict(base_color=[1.0, 1.0, 1.0], base_metalness=1.0, specular_weight=1.0, specular_roughness=0.6),
])
def test_white_materials_have_unit_guide_albedo(material):
    g = guides(with_materials(S.config0(width=24, height=20), **material))
    hit = g[..., 0, 3] > 0
    assert hit.mean() > 0.5
    np.testing.assert_allclose(g[hit][:, 0, :3], 1.0, atol=2e-6)
    assert not g[~hit][:, 0, :].any()


def test_lambertian_guide_is_its_base_colour_and_geometry_is_consistent():
    sc = with_materials(S.config0(width=24, height=20), specular_weight=0.0, base_metalness=0.0)
    g = guides(sc)
    cols = {tuple(np.float32(m["base_color"])) for m in sc.materials}
    hit = g[..., 0, 3] == 1
    for a in g[hit][:, 0, :3]:
        assert tuple(a) in cols
    n = np.linalg.norm(g[hit][:, 1, :3], axis=-1)
    np.testing.assert_allclose(n, 1.0, atol=1e-5)
    d = g[..., 1, 3]
    assert ((d[hit] > 0) & (d[hit] < 1)).all() and (d[~hit] == 1).all()
    g4 = guides(sc, spp=4)
    assert (g4[..., 0, 3] <= 4).all() and (g4[..., 0, 3] >= 0).all()
