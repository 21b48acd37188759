"""Build recipe for the product libraries (run by __graft_entry__.build()).

* ``rtmuq2d_amd/lib/librtm_hip.so`` -- the C-ABI library: HIP kernels for
  gfx950 plus the host C layer that mirrors rtm_iso2d's entry points,
  including this library's own codec for the ZFP 0.5.5 seismogram streams
  (``csrc/rtm_zfp.c``; no ZFP library is built or loaded).
* ``rtmuq2d_amd/lib/rtmuq2d_hip`` -- the process driver.
"""
from __future__ import annotations

import glob
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # every expression is evaluated as written (no fused multiply-add), which
    # keeps the FP32 step bit-identical to the reference's arithmetic
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the kernel arguments arrive in SGPRs (gfx950 kernarg preload) instead
    # of through the first dependent scalar load of every workgroup
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
C_FLAGS = ["-std=gnu99", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
           "-Wno-unused-function", "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[0])}")
    return r


RESOURCES = os.path.join(LIB, "kernel_resources.json")
_RES_KEYS = {"VGPRs": "vgprs", "AGPRs": "agprs", "TotalSGPRs": "sgprs",
             "ScratchSize [bytes/lane]": "scratch_bytes", "Occupancy [waves/SIMD]": "occupancy",
             "LDS Size [bytes/block]": "lds_bytes"}


def _resources(text: str) -> dict:
    """Per-kernel register, scratch, LDS and occupancy figures from the
    compiler's kernel-resource-usage remarks."""
    out, name = {}, None
    for line in text.splitlines():
        if "remark:" not in line:
            continue
        body = line.split("remark:", 1)[1].split(" [-Rpass", 1)[0].strip()
        if body.startswith("Function Name:"):
            name = body.split(":", 1)[1].strip()
            out[name] = {}
        elif name and ":" in body:
            k, v = body.rsplit(":", 1)
            if k.strip() in _RES_KEYS:
                out[name][_RES_KEYS[k.strip()]] = int(v.strip())
    return out


def build_hip() -> str:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    inc = ["-I", os.path.join(REPO, "include"), "-I", CSRC]
    objs = []
    res = {}
    for src in sorted(glob.glob(os.path.join(CSRC, "*.hip"))):
        o = os.path.join(BUILD, os.path.basename(src) + ".o")
        r = _run([HIPCC, *HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage", *inc, "-c", src, "-o", o])
        res.update(_resources(r.stderr))
        objs.append(o)
    with open(RESOURCES, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    for src in sorted(glob.glob(os.path.join(CSRC, "*.c"))):
        o = os.path.join(BUILD, os.path.basename(src) + ".o")
        _run(["gcc", *C_FLAGS, *inc, "-c", src, "-o", o])
        objs.append(o)
    out = os.path.join(LIB, "librtm_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-fopenmp", *objs, "-o", tmp, "-lm"])
    os.replace(tmp, out)
    return out


DRIVER_SRC = os.path.join(HERE, "driver", "rtm_main.c")
DRIVER = os.path.join(LIB, "rtmuq2d_hip")


def build_driver() -> str:
    """rtmuq2d_hip, the process driver, linked against librtm_hip.so in the
    same directory."""
    tmp = DRIVER + ".tmp"
    _run(["gcc", "-std=gnu99", "-O2", "-Wall", "-I", os.path.join(REPO, "include"), DRIVER_SRC, "-o", tmp,
          "-L", LIB, "-lrtm_hip", "-Wl,-rpath,$ORIGIN", "-lm"])
    os.replace(tmp, DRIVER)
    return DRIVER


def build() -> dict:
    return {"rtm": build_hip(), "driver": build_driver()}


if __name__ == "__main__":
    print(build())
