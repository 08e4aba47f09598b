"""scene.json loader: the reference's own fixture and round trips of the
workload generators (CPU only)."""
import copy
import json
import os

import numpy as np
import pytest

from vork_amd import abi, scene as S, scene_json as J

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "scene")


def _analytic():
    with open(os.path.join(GOLDEN, "analytic.json")) as f:
        return json.load(f)


def test_fixture_refuses_what_the_hot_path_does_not_render():
    # crates/vork-render/tests/fixtures/analytic.json: an orthographic camera
    # and a sun light, neither of which this build renders
    with pytest.raises(J.SceneError) as e:
        J.load(os.path.join(GOLDEN, "analytic.json"))
    assert e.value.status == abi.VH_ERR_SCENE
    assert "sun" in str(e.value) and e.value.where == "nodes.node.shape"
    doc = _analytic()
    doc["nodes"] = [n for n in doc["nodes"] if n.get("type") != "light"]
    with pytest.raises(J.SceneError) as e:
        J.load(doc, base_dir=GOLDEN)
    assert "projection" in e.value.where


def analytic_perspective():
    """The fixture with a perspective camera and a rectangle light for the sun."""
    doc = _analytic()
    doc["nodes"] = [n for n in doc["nodes"] if n.get("type") != "light"]
    doc["nodes"][0]["material"]["inputs"]["projection"] = "persp"
    doc["nodes"].append({"name": "key", "type": "light", "shape": "rectangle", "size": 2.0, "energy": 200.0,
                         "translation": [0, 4, 2], "target": [0, 0, 0]})
    return doc


def test_fixture_geometry_transforms_and_defaults():
    sc, warnings = J.load(analytic_perspective(), base_dir=GOLDEN)
    assert warnings == []
    assert (sc.width, sc.height, sc.spp, sc.seed) == (96, 64, 64, 0)
    assert [m.triangle_count for m in sc.meshes] == [12, 1472, 2]
    # the sphere's file buffers: unit positions (read from geometry.bin)
    r = np.linalg.norm(sc.meshes[1].positions, axis=1)
    assert np.allclose(r, 1.0, atol=1e-6)
    inst = sc.instances
    assert inst.shape == (3,)
    # plane: rotation [0.70710678,0,0,0.70710678] (90 degrees about +X), scale (4,1,2), at (0,0,-2)
    m = inst[2]["transform"].reshape(3, 4).astype(np.float64)
    assert np.allclose(m, [[4, 0, 0, 0], [0, 0, -2, 0], [0, 1, 0, -2]], atol=1e-6)
    # sphere: translation (1.5,0,-1), scale 0.5
    m = inst[1]["transform"].reshape(3, 4).astype(np.float64)
    assert np.allclose(m, [[0.5, 0, 0, 1.5], [0, 0.5, 0, 0], [0, 0, 0.5, -1]])
    # meshes without materials get std/openpbr defaults with roughness 0.3
    assert sc.materials[0]["specular_roughness"] == pytest.approx(0.3)
    assert sc.materials[0]["base_color"] == pytest.approx((0.8, 0.8, 0.8))
    # no environment light: RGB 0.05 world
    assert sc.sky["color"] == (0.05, 0.05, 0.05)
    # rectangle light aimed at the origin: normal toward the target, sides of 2
    l = sc.lights[0]
    n = np.asarray(l["normal"])
    assert np.allclose(n, -np.array([0, 4, 2]) / np.linalg.norm([0, 4, 2]), atol=1e-6)
    assert np.linalg.norm(l["u"]) == pytest.approx(2.0) and np.linalg.norm(l["v"]) == pytest.approx(2.0)
    assert np.allclose(l["radiance"], (200.0, 200.0, 200.0))
    # std/camera defaults: 50 mm lens, 36 mm sensor, clip 1..10 from the fixture
    c = sc.camera_struct()
    assert c.clip_start == pytest.approx(1.0) and c.clip_end == pytest.approx(10.0)
    assert c.tan_half_x == pytest.approx(36.0 / (2 * 50.0)) and np.allclose(list(c.eye), [0, 0, 5])
    assert np.allclose(list(c.forward), [0, 0, -1])


def test_render_call_fields_and_errors():
    sc, _ = J.load(analytic_perspective(), J.parse_set(["resolution=40x30", "samples=8", "seed=7"]), base_dir=GOLDEN)
    assert (sc.width, sc.height, sc.spp, sc.seed) == (40, 30, 8, 7)
    with pytest.raises(J.SceneError):
        J.parse_set(["sampls=8"])
    doc = analytic_perspective()
    doc["nodes"][1]["materials"] = [{"inputs": {"emission_luminance": 5.0}}]
    with pytest.raises(J.SceneError) as e:
        J.load(doc, base_dir=GOLDEN)
    assert "emission_luminance" in e.value.where
    doc = analytic_perspective()
    doc["nodes"][1]["materials"] = [{"inputs": {"base_color": [0.2, 0.4, 0.6], "base_weight": 0.5,
                                                "specular_roughness": 0.1, "base_metalness": 1.0}}]
    sc, _ = J.load(doc, base_dir=GOLDEN)
    assert sc.materials[0]["base_color"] == pytest.approx((0.1, 0.2, 0.3))
    assert sc.materials[0]["base_metalness"] == 1.0
    doc = analytic_perspective()
    doc["version"] = 1
    with pytest.raises(J.SceneError):
        J.load(doc, base_dir=GOLDEN)


def _same_lights(a, b):
    for la, lb in zip(a, b):
        assert np.allclose(la["position"], lb["position"], atol=1e-6)
        assert np.allclose(la["radiance"], lb["radiance"])
        na, nb = np.asarray(la["normal"]), np.asarray(lb["normal"])
        assert np.allclose(na / np.linalg.norm(na), nb / np.linalg.norm(nb), atol=1e-6)
        # same rectangle: the side vectors agree up to sign
        for k in ("u", "v"):
            assert np.allclose(np.abs(la[k]), np.abs(lb[k]), atol=1e-6)


@pytest.mark.parametrize("make", [lambda: S.config0(width=64, height=64),
                                  lambda: S.small_config1(width=48, height=27),
                                  lambda: S.config2(width=32, height=18, birds=500)])
def test_generators_round_trip_through_scene_json(tmp_path, make):
    sc = make()
    if sc.sky.get("kind") != abi.VH_SKY_CONSTANT:
        sc.sky = dict(kind=abi.VH_SKY_CONSTANT, color=(0.2, 0.25, 0.3))
    path = os.path.join(tmp_path, "scene.json")
    J.dump(sc, path, inline_limit=256)
    back, warnings = J.load(path, {"samples": sc.spp, "seed": sc.seed})
    assert warnings == []
    assert (back.width, back.height, back.spp) == (sc.width, sc.height, sc.spp)
    assert back.triangle_count == sc.triangle_count
    for a, b in zip(sc.meshes, back.meshes):
        assert np.array_equal(a.positions, b.positions) and np.array_equal(a.indices, b.indices)
        assert np.array_equal(a.normals, b.normals)
    src = np.asarray(sc.instance_array(), dtype=object) if not isinstance(sc.instances, np.ndarray) else None
    ia = sc.instances if isinstance(sc.instances, np.ndarray) else None
    t_src = ia["transform"] if ia is not None else np.array([list(i["transform"]) for i in sc.instances], np.float32)
    b_src = ia["blas"] if ia is not None else np.array([i["blas"] for i in sc.instances])
    order_a = np.lexsort((t_src[:, 11], t_src[:, 7], t_src[:, 3], b_src))
    order_b = np.lexsort((back.instances["transform"][:, 11], back.instances["transform"][:, 7],
                          back.instances["transform"][:, 3], back.instances["blas"]))
    assert np.array_equal(b_src[order_a], back.instances["blas"][order_b])
    assert np.allclose(t_src[order_a], back.instances["transform"][order_b], atol=2e-5)
    _same_lights(sc.lights, back.lights)
    ca, cb = sc.camera_struct(), back.camera_struct()
    for f in ("eye", "forward", "right", "up"):
        assert np.allclose(list(getattr(ca, f)), list(getattr(cb, f)), atol=1e-6)
    assert ca.tan_half_x == pytest.approx(cb.tan_half_x) and ca.tan_half_y == pytest.approx(cb.tan_half_y)
    del src
