"""Write the golden fixtures under tests/golden/ from the reference build.

Runs in the build container only (it needs /root/reference and the recipes of
oracle/build.py).  For each configuration it generates the seeded inputs,
compresses the seismograms with ZFP 0.5.5 (the reference's field convention),
runs the reference's own rtm_routine (oracle/_ref, compiled in place) and
stores inputs and outputs:

  tiny.npz, asym.npz : every input and output array (small grids)
  c0.json            : SHA-256 of the c0 inputs and of the reference's c0
                       images (strict build), so a GPU run can show bit
                       equality without the reference present
  c0_native.npz      : the c0 simulation-1 image of the reference built with
                       -O3 -march=native (FMA contraction), for the relL2 gate
  c1s.json           : as c0.json for one full c1-size image (1701x351, CPML
                       50, nt=4000, TS=20), plus the relL2 between the strict
                       and the -march=native builds' images (an image
                       bit-identical to the strict one has that relL2 to the
                       native one; the 2.4 MB native image is not kept)

usage: python tests/golden/make_golden.py [tiny asym c0 c1s]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import common  # noqa: E402
import rtmuq2d_amd as R  # noqa: E402


def sha(a) -> str:
    if isinstance(a, (bytes, bytearray)):
        return hashlib.sha256(a).hexdigest()
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make(name: str) -> None:
    cfg = common.REF_CONFIGS[name]
    v_true, models, wav, seis = common.make_inputs(cfg)
    tol = cfg["zfp_tolerance"]
    zs = [R.zfp_compress(d, tol) for d in seis]
    with tempfile.TemporaryDirectory() as run:
        common.write_inputs(run, cfg, models, wav, zs)
        nsim = len(models)
        common.run_reference(run, name, "strict", 1, nsim)
        outs = [common.read_outputs(run, cfg, s) for s in range(1, nsim + 1)]
        native = None
        if name in ("c0", "c1s"):
            common.run_reference(run, name, "native", 1, 1)
            native = common.read_outputs(run, cfg, 1)[0]
    if name in ("c0", "c1s"):
        meta = {
            "v_true": sha(v_true), "wavelet": sha(wav),
            "models": [sha(m) for m in models], "zfp": [sha(z) for z in zs],
            "zfp_bytes": [len(z) for z in zs],
            "stackedBrute": [sha(b) for b, _ in outs],
            "stackedLaplacian": [sha(l) for _, l in outs],
            "brute_norm": [float(np.linalg.norm(b.astype(np.float64))) for b, _ in outs],
        }
        if name == "c1s":
            meta["strict_vs_native_relL2"] = common.rel_l2(outs[0][0], native)
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(meta, f, indent=1)
        if name == "c0":
            np.savez_compressed(os.path.join(HERE, "c0_native.npz"), brute=native)
    else:
        arrays = {
            "v_true": v_true, "wavelet": wav,
            "models": np.stack(models),
            "seismograms": np.stack(seis),
            "brute": np.stack([b for b, _ in outs]),
            "laplacian": np.stack([l for _, l in outs]),
        }
        for s, z in enumerate(zs):
            arrays[f"zfp_{s}"] = np.frombuffer(z, np.uint8)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
    print(name, "written")


if __name__ == "__main__":
    for n in sys.argv[1:] or ["tiny", "asym", "c0"]:
        make(n)
