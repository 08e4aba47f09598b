"""The session protocol on the CPU: the document diff rules
(crates/vork-render/src/shot/worker.rs:652-732, with the changed paths the
reference's session test expects), request validation and failure kinds
(crates/vork-cli/src/session.rs), and a loud device failure without a GPU."""
import io
import json
import os

import pytest

from vork_amd import scene as S
from vork_amd import scene_json as J
from vork_amd import session as W


def nodes(*items):
    return {"version": 2, "nodes": list(items)}


def test_diff_named_items_report_their_names():
    a = nodes({"name": "ball", "geometry": {"radius": 0.5}}, {"name": "key", "energy": 60})
    b = nodes({"name": "ball", "geometry": {"radius": 0.5}}, {"name": "key", "energy": 140})
    assert W.diff(a, b) == [["nodes", "key", "energy"]]
    c = nodes({"name": "ball", "geometry": {"radius": 0.7}}, {"name": "key", "energy": 140})
    assert W.diff(b, c) == [["nodes", "ball", "geometry", "radius"]]


def test_diff_numbers_by_value_and_scalar_lists_as_one_value():
    assert W.diff({"a": 1, "c": [1, 2]}, {"a": 1.0, "c": [1.0, 2.0]}) == []
    assert W.diff({"c": [1, 2, 3]}, {"c": [1, 2, 4]}) == [["c"]]
    assert W.diff({"c": [1, 2]}, {"c": [1, 2, 3]}) == [["c"]]
    assert W.diff({"a": True}, {"a": 1}) == [["a"]]
    assert W.diff({"a": 1}, {"b": 1}) == [["a"], ["b"]]


def test_diff_reorder_additions_and_removals():
    x, y = {"name": "x", "v": 1}, {"name": "y", "v": 2}
    assert W.diff(nodes(x, y), nodes(y, x)) == [["nodes"]]
    assert W.diff(nodes(x), nodes(x, y)) == [["nodes", "y"]]
    assert W.diff(nodes(x, y), nodes(y)) == [["nodes", "x"]]
    # unnamed items pair by position; a left-over is reported by index
    assert W.diff(nodes({"v": 1}, {"v": 2}), nodes({"v": 1})) == [["nodes", 1]]
    assert W.diff(nodes({"v": 1}), nodes({"v": 3})) == [["nodes", 0, "v"]]
    # duplicate names pair in order
    assert W.diff(nodes({"name": "d", "v": 1}, {"name": "d", "v": 2}),
                  nodes({"name": "d", "v": 1}, {"name": "d", "v": 5})) == [["nodes", "d", "v"]]


def serve(lines, worker=None):
    out = io.StringIO()
    W.serve(io.StringIO("".join(line + "\n" for line in lines)), out, worker)
    return [json.loads(x) for x in out.getvalue().splitlines()]


def test_requests_are_validated_with_usage_replies():
    replies = serve([
        "{not json",
        "[1, 2]",
        "",
        json.dumps({"id": 1, "op": "dance"}),
        json.dumps({"id": 2, "op": "open", "scene": "a.json", "document": {}}),
        json.dumps({"id": 3, "op": "open"}),
        json.dumps({"id": 4, "op": "open", "scene": "a.json", "colour": 1}),
        json.dumps({"id": 5, "op": "render", "out": "a.png"}),
        json.dumps({"id": 6, "op": "ask", "questions": []}),
        json.dumps({"id": 7, "op": "close"}),
        json.dumps({"id": 8, "op": "render", "out": "after-close.png"}),
    ])
    assert [r["id"] for r in replies] == [None, None, 1, 2, 3, 4, 5, 6, 7]
    assert all(r["ok"] is False and r["kind"] == "usage" for r in replies[:-1]), replies
    assert "requests are open" in replies[2]["error"]
    assert replies[-1] == {"id": 7, "ok": True}


def test_render_fields_are_checked_before_any_device_work():
    w = W.Worker()
    w.open_state, w.installed = {"revision": 1}, {}  # stand-ins: _options alone is exercised
    for req, text in [({"out": "a.png", "colour": 1}, "unknown field"),
                      ({"set": {}}, "missing field `out`"),
                      ({"out": "a.tiff"}, ".png or .exr"),
                      ({"out": "a.png", "set": {"bogus": 1}}, "unknown render-call field bogus"),
                      ({"out": "a.png", "profile": "std/draft"}, "std/path"),
                      ({"out": "a.png", "channels": ["albedo"]}, "colour"),
                      ({"out": "a.png", "at": 1.5}, "stills"),
                      ({"out": "a.png", "set": {"samples": 2.5}}, "samples")]:
        with pytest.raises(W.Failure) as e:
            w._options(req)
        assert e.value.kind == "usage" and text in str(e.value), (req, str(e.value))
    outs, ov, no_write = w._options({"out": ["a.png", "b.exr"], "set": {"resolution": "64x48", "samples": 8,
                                                                          "seed": "3"}})
    assert outs == ["a.png", "b.exr"] and ov == {"resolution": (64, 48), "samples": 8, "seed": 3} and not no_write
    assert w._options({"out": "a.png", "provisional": True})[1] == {"samples": 1, "denoise": False}


def test_open_failures_keep_their_kinds(tmp_path):
    sc = S.config0(width=24, height=16, spp=2)
    path = str(tmp_path / "s.json")
    J.dump(sc, path)
    doc = json.load(open(path))
    doc["nodes"][3]["energy"] = "bright"
    missing = str(tmp_path / "nope.json")
    replies = serve([json.dumps({"id": 1, "op": "open", "document": doc, "base": str(tmp_path)}),
                     json.dumps({"id": 2, "op": "open", "scene": missing}),
                     json.dumps({"id": 3, "op": "open", "document": [1], "base": str(tmp_path)})])
    assert replies[0]["kind"] == "scene" and "energy" in replies[0]["error"], replies[0]
    assert replies[1]["kind"] == "io" and "nope.json" in replies[1]["error"]
    assert replies[2]["kind"] == "scene"


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the failure without a GPU")
def test_open_without_a_gpu_is_a_device_failure(tmp_path):
    sc = S.config0(width=24, height=16, spp=2)
    path = str(tmp_path / "s.json")
    J.dump(sc, path)
    (r,) = serve([json.dumps({"id": 1, "op": "open", "scene": path})])
    assert r["ok"] is False and r["kind"] == "device", r


def test_scene_keys_separate_tables_from_geometry():
    a = S.config0(width=24, height=16, spp=2)
    b = S.config0(width=24, height=16, spp=2)
    assert W.scene_key(a) == W.scene_key(b) and W.geometry_key(a) == W.geometry_key(b)
    b.lights[0] = dict(b.lights[0], radiance=(1.0, 2.0, 3.0))
    assert W.scene_key(a) != W.scene_key(b) and W.geometry_key(a) == W.geometry_key(b)
