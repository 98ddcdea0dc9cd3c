"""The C ABI library loads and exports every entry point include/avbd.h declares (no compute
calls: this runs without a GPU)."""
import os
import re

import avbd_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "avbd.h")).read()
    return sorted(set(re.findall(r"\b(avbd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for n in ("avbd_create", "avbd_step", "avbd_read_bodies", "avbd_seed", "avbd_destroy", "avbd_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = avbd_amd.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(avbd_amd.EXPORTS) == declared()


def test_create_rejects_invalid_params_before_touching_the_device():
    import numpy as np
    p = avbd_amd.default_params()
    p.color_rounds = 3
    try:
        avbd_amd.GpuSolver3D(np.zeros((1, 40), np.float32), params=p)
    except ValueError as e:
        assert "even" in str(e)
    else:
        raise AssertionError("odd colour rounds accepted")
