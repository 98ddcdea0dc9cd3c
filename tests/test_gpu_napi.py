"""The MI355X step driven from JavaScript through the N-API addon (three-avbd_amd/napi), as the
reference's World drives its GpuSolver3D: a World-style scene built and stepped in Node equals the
same scene through the C ABI from Python bit for bit (bodies and joints), and the reference's
error classes arrive in JavaScript."""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import avbd_amd
from avbd_amd import build, scenes
from helpers import new_box

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(build.NAPI_OUT),
                                                  reason="no node, or the addon is not built")]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "three-avbd_amd", "napi", "gpu-solver-3d.js")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scene():
    pyr = scenes.pyramid(4)[0].astype(np.float32)
    pyr[:, 11] = 1.0
    chain = np.vstack([new_box((6.0, 0.0, 4.0), (1.0, 1.0, 1.0), 0.0)] + [new_box((7.0 + k, 0.0, 4.0), (1.0, 0.6, 0.6), 2.0) for k in range(5)])
    a = len(pyr)
    specs = [{"a": a, "b": a + 1, "rA": [0.5, 0, 0], "rB": [-0.5, 0, 0], "yield": 1.0, "breakBend": 0.25},
             {"a": a + 1, "b": a + 2, "rA": [0.5, 0, 0], "rB": [-0.5, 0, 0], "angular": 0.0},
             {"a": a + 2, "b": a + 3, "rA": [0.5, 0, 0], "rB": [-0.5, 0, 0], "rest": [0, 0, 0.05, 1]},
             {"a": a + 3, "b": a + 4, "rA": [0.5, 0, 0], "rB": [-0.5, 0, 0], "fracture": 1e6, "linear": True},
             {"a": a + 4, "b": a + 5, "rA": [0.5, 0, 0], "rB": [-0.5, 0, 0]}]
    return {
        "params": {"dt": 1 / 60, "colorCap": 32, "flags": avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS},
        "pyramid": pyr.ravel().tolist(), "chain": chain.ravel().tolist(), "specs": specs,
        "springs": [{"a": a + 5, "b": 5, "rA": [0, 0, 0], "rB": [0, 0, 0.5], "stiffness": 50.0, "rest": 1.5}],
        "drag": {"b": 9, "world": [0.0, 0.5, 4.5], "body": [0.0, 0.0, 0.5], "moveTo": [0.3, 0.8, 4.6]},
        "aux": {"indices": [3, 4], "scales": [1.0, 2.0], "linear": [0.8, 0.0], "angular": [0.0, 1.5]},
        "disable": 8, "steps": 25, "gravity2": -5.0,
    }


def test_a_world_driven_from_javascript_steps_as_from_python(gpu, tmp_path):
    say("case: World-style scene built and stepped from Node through the N-API addon vs the same through ctypes")
    sc = scene()
    inp, out = tmp_path / "in.json", tmp_path / "out.bin"
    inp.write_text(json.dumps(sc))
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "napi", "world.js"), JS, str(inp), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    assert js["first"] == 0 and js["base"] == 11 and js["slots"] == [0, 1, 2, 3, 4] and js["spring"] == [5] and js["drag"] == 6
    assert js["refused"] == "StorageCapacityError" and js["afterDestroy"] == "Solver is destroyed"
    assert js["contacts"] > 0 and js["counters"]["overflow"] == 0
    # the same through the C ABI from Python
    p = avbd_amd.default_params(dt=1 / 60, color_cap=32, flags=sc["params"]["flags"])
    s = avbd_amd.GpuSolver3D(np.zeros((0, 40), np.float32), params=p, capacity={"bodies": 64, "joints": 2})
    s.add_bodies(np.asarray(sc["pyramid"], np.float32).reshape(-1, 40))
    s.add_bodies(np.asarray(sc["chain"], np.float32).reshape(-1, 40))
    s.append_joints(sc["specs"])
    s.append_springs(sc["springs"])
    d = sc["drag"]
    slot = s.append_world_joint(d["b"], d["world"], d["body"])
    s.set_world_anchor(slot, d["moveTo"])
    a = sc["aux"]
    s.set_body_aux(a["indices"], a["scales"], a["linear"], a["angular"])
    s.set_lifecycle(sc["disable"], 2)
    s.steps(sc["steps"])
    q = s.params
    q.gravity = sc["gravity2"]
    s.set_params(q)
    s.steps(sc["steps"])
    got = np.frombuffer(out.read_bytes(), np.float32)
    nb = s.body_count * 40
    bodies, joints = got[:nb].reshape(-1, 40), got[nb:].reshape(-1, 40)
    assert np.array_equal(bodies.view(np.uint32), s.read_bodies().view(np.uint32))
    assert np.array_equal(joints.view(np.uint32), s.read_joints().view(np.uint32))
    assert js["colorCap"] == s.params.color_cap  # the library's adapted cap, not the one JavaScript first sent
    s.destroy()
