"""Build recipe for the oracle (test infrastructure, never the product).

Two things are built here:

1. ``oracle/liboracle.so`` -- the C restatement in ``rtm_oracle_impl.c``,
   compiled twice (float and double) with ``-ffp-contract=off``.

2. ``oracle/_ref/<config>/`` -- the reference's own C, compiled *in place*
   from ``/root/reference/SRC`` (nothing is copied).  The reference takes all
   of its parameters from compile-time ``static const`` globals in
   ``SRC/fixed_parameters.h`` (lines 8-70) behind the include guard
   ``FIXED_PARAMETERS_H``.  A per-configuration header with the same guard and
   the same globals is force-included (``-include``), so the reference's own
   ``#include "fixed_parameters.h"`` becomes a no-op and the reference builds
   for that configuration.  ``TIME_SLICES`` stays the reference's 20
   (``SRC/forward_modeling.h:4``).  ``mpi_rtm_iso2d.c`` is replaced by
   ``oracle/ref_main.c`` (a driver that calls ``rtm_routine`` for a range of
   simulation ids, as ``SRC/mpi_rtm_iso2d.c:47-51`` does on one rank).

   Two flavours per configuration:
     * ``strict``: ``-O2 -fopenmp -ffp-contract=off`` -- evaluates every
       expression as written; the oracle must match it bit for bit;
     * ``native``: ``-O3 -fopenmp -march=native`` -- the reference Makefile's
       ``-O3 -fopenmp -xHost`` with gcc standing in for icc; used for the CPU
       baseline timing and the relL2 <= 1e-5 end-to-end check.

   The reference links ZFP 0.5.5; it is compiled here from the reference's
   vendored source (``ZFP_LIBRARY_0.5.5/src``, read in place, with the
   reference's settings: 64-bit stream words, OpenMP, ``-O3 -DNDEBUG``) into
   ``oracle/_ref/zfp/libzfp.so``.  That library is a checker too: the tests
   compare this build's own ZFP stream codec (``rtmuq2d_amd/csrc/rtm_zfp.c``)
   with it byte for byte.  The product neither builds nor loads it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_ROOT = "/root/reference"
REF_SRC = os.path.join(REF_ROOT, "SRC")
REF_ZFP = os.path.join(REF_ROOT, "ZFP_LIBRARY_0.5.5")
REF_ZFP_INC = os.path.join(REF_ZFP, "include")
REF_OUT = os.path.join(HERE, "_ref")
ZFP_LIB_DIR = os.path.join(REF_OUT, "zfp")
ZFP_LIB = os.path.join(ZFP_LIB_DIR, "libzfp.so")

# Reference translation units on the RTM path (SRC/, all but the MPI main).
REF_UNITS = [
    "stencilDFM.c", "absorbingBoundaryCondition.c", "forward_modeling.c",
    "backward_modeling.c", "image_conditions.c", "generic_functions.c",
    "math_functions.c", "pi_file.c", "rtm_iso2d.c", "compression_tools.c",
]

FLAVOURS = {
    "strict": ["-O2", "-fopenmp", "-ffp-contract=off"],
    "native": ["-O3", "-fopenmp", "-march=native"],
}

# Configurations the tests and the CPU baseline use.  Keys mirror the names
# in SRC/fixed_parameters.h.
REF_CONFIGS = {
    # small symmetric grid, 2 shots, 2 simulations
    "tiny": dict(Nx=64, Nz=48, thicknessLeft=16, thicknessRight=16, thicknessUp=16,
                 thicknessDown=16, spacingGrid=10.0, sampling=0.001, numberOfTimeStep=300,
                 sampligWavelet=300, firstLocationShot=20, depthShot=2, spacingShot=24,
                 numberOfShots=2, firstLocationReceiver=0, depthReceiver=2,
                 spacingReceiver=2, numberOfReceivers=64, numberOfSimulations=2,
                 zfp_tolerance=1e-4),
    # unequal CPML thicknesses on all four sides (exercises every strip index
    # and the right-pad row range of SRC/generic_functions.c:27)
    "asym": dict(Nx=72, Nz=40, thicknessLeft=14, thicknessRight=20, thicknessUp=18,
                 thicknessDown=13, spacingGrid=12.5, sampling=0.0012, numberOfTimeStep=240,
                 sampligWavelet=180, firstLocationShot=30, depthShot=3, spacingShot=10,
                 numberOfShots=1, firstLocationReceiver=1, depthReceiver=4,
                 spacingReceiver=3, numberOfReceivers=36, numberOfSimulations=1,
                 zfp_tolerance=1e-4),
    # config c0: 600x200, h=10, CPML 40, dt=1 ms, nt=2000, 2 simulations
    "c0": dict(Nx=600, Nz=200, thicknessLeft=40, thicknessRight=40, thicknessUp=40,
               thicknessDown=40, spacingGrid=10.0, sampling=0.001, numberOfTimeStep=2000,
               sampligWavelet=2000, firstLocationShot=300, depthShot=2, spacingShot=1,
               numberOfShots=1, firstLocationReceiver=0, depthReceiver=2, spacingReceiver=2,
               numberOfReceivers=600, numberOfSimulations=2, zfp_tolerance=1e-4),
    # config c1 (1701x351, CPML 50, dt=0.8 ms, nt=4000): one full image,
    # the CPU-baseline sample
    "c1s": dict(Nx=1701, Nz=351, thicknessLeft=50, thicknessRight=50, thicknessUp=50,
                thicknessDown=50, spacingGrid=10.0, sampling=0.0008, numberOfTimeStep=4000,
                sampligWavelet=4000, firstLocationShot=850, depthShot=2, spacingShot=1,
                numberOfShots=1, firstLocationReceiver=0, depthReceiver=2, spacingReceiver=2,
                numberOfReceivers=1701, numberOfSimulations=1, zfp_tolerance=1e-4),
}


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"command failed: {cmd[0]} (rc={r.returncode})")
    return r


def build_oracle() -> str:
    """Compile oracle/liboracle.so (float + double restatement)."""
    out = os.path.join(HERE, "liboracle.so")
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    src = os.path.join(HERE, "rtm_oracle_impl.c")
    common = ["gcc", "-std=gnu99", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC",
              "-Wall", "-Wno-unused-function", "-I", HERE]
    o32 = os.path.join(bdir, "oracle_f32.o")
    o64 = os.path.join(bdir, "oracle_f64.o")
    _run(common + ["-c", src, "-o", o32])
    _run(common + ["-DREAL=double", "-DSFX=_f64", "-DORACLE_DOUBLE", "-c", src, "-o", o64])
    _run(["gcc", "-shared", "-fopenmp", o32, o64, "-o", out, "-lm"])
    return out


def ref_available() -> bool:
    return os.path.isdir(REF_SRC)


def _config_header(cfg: dict) -> str:
    ints = ["numberOfSimulations", "Nx", "Nz", "firstLocationShot", "depthShot", "spacingShot",
            "numberOfShots", "firstLocationReceiver", "depthReceiver", "spacingReceiver",
            "numberOfReceivers", "numberOfTimeStep", "sampligWavelet", "thicknessUp",
            "thicknessDown", "thicknessLeft", "thicknessRight"]
    lines = [
        "/* generated by oracle/build.py: one configuration of the reference */",
        "#ifndef FIXED_PARAMETERS_H",
        "#define FIXED_PARAMETERS_H",
        'static char velocityFile[] = "./INPUTS/vp_%05d.bin";',
        'static char waveletFile[] = "./INPUTS/wavelet.bin";',
        'static char seismogramFile[] = "./INPUTS/seismogram_%05d.zfp";',
        'static char stackedFile[] = "./OUTPUTS/stackedBrute_%05d.bin";',
        'static char stackedLaplacianFile[] = "./OUTPUTS/stackedLaplacian_%05d.bin";',
        f"static const float zfp_tolerance = {float(cfg['zfp_tolerance'])!r};",
        f"static const float spacingGrid = {float(cfg['spacingGrid'])!r};",
        f"static const float sampling = {float(cfg['sampling'])!r};",
        "static const float cutFrequency = 50.0;",
    ]
    lines += [f"static const int {k} = {int(cfg[k])};" for k in ints]
    lines.append("#endif")
    return "\n".join(lines) + "\n"


def build_ref(name: str, flavour: str = "strict") -> str:
    """Compile the reference for configuration `name`; returns its output dir."""
    if not ref_available():
        raise FileNotFoundError(REF_SRC)
    if not os.path.exists(ZFP_LIB):
        build_ref_zfp()
    cfg = REF_CONFIGS[name]
    odir = os.path.join(REF_OUT, f"{name}_{flavour}")
    os.makedirs(odir, exist_ok=True)
    hdr = os.path.join(odir, "cfg.h")
    with open(hdr, "w") as f:
        f.write(_config_header(cfg))
    objs = []
    for unit in REF_UNITS:
        obj = os.path.join(odir, unit.replace(".c", ".o"))
        _run(["gcc", "-std=gnu99", "-fPIC", "-w", *FLAVOURS[flavour], "-include", hdr,
              "-I", REF_ZFP_INC, "-c", os.path.join(REF_SRC, unit), "-o", obj])
        objs.append(obj)
    lib = os.path.join(odir, "librtmref.so")
    _run(["gcc", "-shared", "-fopenmp", *objs, "-o", lib, "-L", ZFP_LIB_DIR, "-lzfp",
          f"-Wl,-rpath,{ZFP_LIB_DIR}", "-lm"])
    exe = os.path.join(odir, "ref_main")
    _run(["gcc", "-std=gnu99", "-O2", os.path.join(HERE, "ref_main.c"), "-o", exe,
          "-L", odir, "-lrtmref", f"-Wl,-rpath,{odir}", "-fopenmp"])
    with open(os.path.join(odir, "params.txt"), "w") as f:
        for k, v in cfg.items():
            f.write(f"{k} {v}\n")
    return odir


def build_ref_zfp() -> str:
    """ZFP 0.5.5 compiled from the reference's vendored source into
    oracle/_ref/zfp/libzfp.so (the prebuilt binaries the reference ships
    are never used)."""
    if not ref_available():
        raise FileNotFoundError(REF_SRC)
    src = os.path.join(REF_ZFP, "src")
    os.makedirs(ZFP_LIB_DIR, exist_ok=True)
    units = ["bitstream.c", "zfp.c"] + [f"{k}{d}{t}.c" for k in ("decode", "encode") for d in "1234" for t in "ilfd"]
    objs = []
    for u in units:
        o = os.path.join(ZFP_LIB_DIR, u.replace(".c", ".o"))
        _run(["gcc", "-std=gnu99", "-O3", "-DNDEBUG", "-fPIC", "-fopenmp", "-w", "-I", REF_ZFP_INC,
              "-c", os.path.join(src, u), "-o", o])
        objs.append(o)
    _run(["gcc", "-shared", "-fopenmp", *objs, "-o", ZFP_LIB, "-lm"])
    return ZFP_LIB


def build_all_refs() -> list:
    out = [build_ref_zfp()]
    for name in REF_CONFIGS:
        out.append(build_ref(name, "strict"))
    out.append(build_ref("c0", "native"))
    out.append(build_ref("c1s", "native"))
    return out


def main() -> None:
    print(build_oracle())
    if ref_available():
        for d in build_all_refs():
            print(d)
    else:
        print("reference sources absent: oracle/_ref not rebuilt")


if __name__ == "__main__":
    main()
