"""The session worker on the GPU, after crates/vork-cli/tests/session.rs:
a kept-alive worker keeps its device and installation across reopened
documents that only change values, reinstalls when geometry changes, keeps
the last good revision after a refused open, and its pictures are
byte-identical to one-shot renders of the same documents."""
import json
import os
import subprocess
import sys

import pytest

from vork_amd import render as R
from vork_amd import scene as S
from vork_amd import scene_json as J

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SET = {"samples": 4, "seed": 2}


def cold(path, out):
    """The one-shot render call of the same document and options."""
    assert R.main([path, "--out", out] + sum((["--set", f"{k}={v}"] for k, v in SET.items()), [])) == 0
    with open(out, "rb") as f:
        return f.read()


class Session:
    def __init__(self):
        self.p = subprocess.Popen([sys.executable, "-m", "vork_amd.session"], cwd=ROOT, stdin=subprocess.PIPE,
                                  stdout=subprocess.PIPE, text=True)

    def send(self, request):
        self.p.stdin.write(json.dumps(request) + "\n")
        self.p.stdin.flush()
        line = self.p.stdout.readline()
        assert line, f"worker closed stdout after {request}"
        reply = json.loads(line)
        assert reply["id"] == request["id"], reply
        return reply

    def render(self, rid, out, **extra):
        r = self.send(dict({"id": rid, "op": "render", "out": out, "set": dict(SET)}, **extra))
        assert r["ok"] is True, r
        assert r["result"]["frames"] == 1
        return r


def test_session_reuses_updates_reinstalls_and_matches_one_shot(tmp_path):
    d = str(tmp_path)
    dim = os.path.join(d, "dim.json")
    J.dump(S.config0(width=48, height=32, spp=4), dim)
    doc = json.load(open(dim))
    bright = json.loads(json.dumps(doc))
    bright["nodes"][3]["energy"] = 60.0
    moved = json.loads(json.dumps(bright))
    moved["nodes"][1]["translation"] = [0.3, 1.0, 0.0]
    paths = {}
    for name, document in (("bright", bright), ("moved", moved)):
        paths[name] = os.path.join(d, name + ".json")
        with open(paths[name], "w") as f:
            json.dump(document, f)
    want = {k: cold(p, os.path.join(d, f"cold-{k}.png")) for k, p in (("dim", dim), ("bright", paths["bright"]),
                                                                      ("moved", paths["moved"]))}
    assert want["dim"] != want["bright"] != want["moved"], "the edits must change the picture"

    s = Session()
    try:
        r = s.send({"id": 1, "op": "open", "scene": dim})
        assert r == {"id": 1, "ok": True, "revision": 1, "reuse": "reinstalled", "changed": []}
        r = s.render(2, os.path.join(d, "s0.png"))
        assert (r["reuse"], r["installs"], r["provisional"]) == ("reinstalled", 1, False)
        assert r["device_reused"] is True and r["revision"] == 1

        # the identical document again: nothing reinstalled
        r = s.send({"id": 3, "op": "open", "scene": dim})
        assert r == {"id": 3, "ok": True, "revision": 2, "reuse": "reused", "changed": []}
        r = s.render(4, os.path.join(d, "s0b.png"))
        assert (r["reuse"], r["installs"]) == ("reused", 1)

        # a light energy change keeps the acceleration structures
        r = s.send({"id": 5, "op": "open", "document": bright, "base": d})
        assert r == {"id": 5, "ok": True, "revision": 3, "reuse": "updated", "changed": [["nodes", "light0", "energy"]]}
        r = s.render(6, os.path.join(d, "s1.png"))
        assert (r["reuse"], r["installs"]) == ("updated", 1)

        # moving an instance is an installed fact: reinstalled on the kept device
        r = s.send({"id": 7, "op": "open", "document": moved, "base": d})
        assert r == {"id": 7, "ok": True, "revision": 4, "reuse": "reinstalled",
                     "changed": [["nodes", "instance1", "translation"]]}
        r = s.render(8, os.path.join(d, "s2.png"))
        assert (r["reuse"], r["installs"], r["device_reused"]) == ("reinstalled", 2, True)

        # a refused document keeps the last revision
        broken = json.loads(json.dumps(moved))
        broken["nodes"][3]["energy"] = "bright"
        r = s.send({"id": 9, "op": "open", "document": broken, "base": d})
        assert r["ok"] is False and r["kind"] == "scene" and "energy" in r["error"], r
        r = s.send({"id": 10, "op": "open", "scene": os.path.join(d, "nope.json")})
        assert r["ok"] is False and r["kind"] == "io", r
        r = s.render(11, os.path.join(d, "s2b.png"))
        assert (r["revision"], r["reuse"], r["installs"]) == (4, "reused", 2)

        # a provisional picture: one sample
        r = s.render(12, os.path.join(d, "p.png"), provisional=True)
        assert r["provisional"] is True and r["report"]["render_settings"]["samples"] == 1
        assert s.send({"id": 13, "op": "close"}) == {"id": 13, "ok": True}
        assert s.p.wait(timeout=60) == 0
    finally:
        if s.p.poll() is None:
            s.p.kill()
            s.p.wait()
    for name, out in (("dim", "s0.png"), ("dim", "s0b.png"), ("bright", "s1.png"), ("moved", "s2.png"),
                      ("moved", "s2b.png")):
        with open(os.path.join(d, out), "rb") as f:
            assert f.read() == want[name], f"session render {out} differs from the one-shot render"
