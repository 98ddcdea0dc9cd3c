"""The N-API addon and the JavaScript GpuSolver3D over it (three-avbd_amd/napi), on the CPU: it
loads in Node, exports the C ABI's operations, packs avbd_params / avbd_joint_spec /
avbd_spring_spec byte for byte as the C header lays them out (compared with ctypes), and turns the
library's refusals into thrown Errors, all without a device call."""
import ctypes
import json
import math
import os
import shutil
import subprocess

import pytest

import avbd_amd
from avbd_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "three-avbd_amd", "napi", "gpu-solver-3d.js")

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(build.NAPI_OUT),
                                reason="no node, or the addon is not built")

REFERENCE_METHODS = {  # GpuSolver3D's operational surface (solver.ts) that World calls
    "step", "readBodies", "readJoints", "readCounters", "readContactList", "readPairs", "rewriteBodies", "addBodies",
    "addBody", "patchBody", "rewriteFixed", "appendJoints", "appendJoint", "appendWorldJoint", "appendSprings",
    "releaseJoints", "setWorldAnchor", "disableConstraint", "setBodyAux", "setLifecycle", "setFilters", "adapt",
    "jointInfo", "destroy"}


def _layout():
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "napi", "layout.js"), JS], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_the_addon_loads_and_mirrors_the_reference_class():
    out = _layout()
    assert REFERENCE_METHODS <= set(out["methods"])
    assert {"create", "destroy", "step", "steps", "addBodies", "appendJoints", "releaseJoints", "setBodyAux"} <= set(out["exports"])
    assert out["sizes"] == {"params": ctypes.sizeof(avbd_amd.solver.Params), "counters": ctypes.sizeof(avbd_amd.solver.Counters),
                            "jointSpec": ctypes.sizeof(avbd_amd.solver.JointSpec), "springSpec": ctypes.sizeof(avbd_amd.solver.SpringSpec)}


def test_refusals_become_thrown_errors():
    out = _layout()
    assert out["create"] == "Error: color_rounds must be even"
    assert out["handle"] == "TypeError: not a solver handle"


def test_packed_structs_match_the_c_layout():
    out = _layout()
    p = avbd_amd.default_params(dt=1 / 240, color_cap=24, flags=avbd_amd.FLAG_DEFAULTS | avbd_amd.solver.FLAG_ADAPT_COLORS)
    assert out["params"] == bytes(p).hex()
    j = avbd_amd.solver.JointSpec()
    j.a, j.b = 3, 7
    j.rA[:] = [0.5, 0, -1]
    j.rB[:] = [0, 0.25, 0]
    j.angular, j.fracture, j.linear = math.nan, math.nan, 1
    j.rest[:] = [0, 0, 0.1, 1]
    j.yield_, j.break_bend = 40.0, 0.5
    assert out["joint"] == bytes(j).hex()
    sp = avbd_amd.solver.SpringSpec()
    sp.a, sp.b = -1, 2
    sp.rA[:] = [1, 2, 3]
    sp.rB[:] = [4, 5, 6]
    sp.stiffness, sp.rest = 50.0, 1.5
    assert out["spring"] == bytes(sp).hex()
