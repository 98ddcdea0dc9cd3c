"""Builds libavbd_mi355x.so in-tree with hipcc for gfx950.

-ffp-contract=off and IEEE division/sqrt (no fast-math) make every float operation round as
written; packed-FP32 instructions are disabled (the hardware's packed forms are avoided).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
OUT = os.path.join(HERE, "libavbd_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
    "-Wno-unused-value", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


NAPI_DIR = os.path.join(os.path.dirname(HERE), "napi")
NAPI_OUT = os.path.join(NAPI_DIR, "avbd_node.node")
NODE_INCLUDE = "/usr/include/node"
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")


def build_napi(verbose=False):
    """The N-API addon (napi/avbd_node.c) against Node's headers, linked to the library beside
    it; skipped when the image has no Node headers."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print("no Node headers: the N-API addon is not built")
        return None
    tmp = NAPI_OUT + ".tmp"
    cmd = ["cc", "-O2", "-fPIC", "-shared", "-Wall", "-I" + NODE_INCLUDE, "-I" + INCLUDE, os.path.join(NAPI_DIR, "avbd_node.c"),
           "-L" + HERE, "-lavbd_mi355x", "-Wl,-rpath,$ORIGIN/../avbd_amd", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cc (N-API addon) failed:\n" + r.stderr)
    os.replace(tmp, NAPI_OUT)
    if verbose:
        print("built", NAPI_OUT)
    return NAPI_OUT


def build(verbose=False):
    src = os.path.join(CSRC, "avbd.hip")
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_napi(verbose)
    return OUT
