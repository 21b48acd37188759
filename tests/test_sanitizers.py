"""AddressSanitizer + UndefinedBehaviorSanitizer over the CPU builds of the
device kernel bodies and the host C layer (tests/sanitize_driver.cpp): the
LDS tiles, psi halos and clamped global indices of the tile, column and
mixed-tile phases on four grids (every tile class, partial tiles, a
grid smaller than one tile), FP32 and FP64, the 16-bit stores, Kahan
imaging; the CPML profiles, padding, Laplacian, raw file I/O and the ZFP
stream codec in every mode.  The CPU build of the kernels is the guard on
the indexing that replaced the reference's psi -> dpsi barrier
(SRC/absorbingBoundaryCondition.c:174-217).  Host code only: nothing here
builds for the GPU."""
from __future__ import annotations

import os
import subprocess

import common

CSRC = os.path.join(common.REPO, "rtmuq2d_amd", "csrc")
INC = os.path.join(common.REPO, "include")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def test_kernel_bodies_and_host_code_are_sanitizer_clean(tmp_path):
    objs = []
    for src in ("rtm_host.c", "rtm_zfp.c"):
        o = str(tmp_path / (src + ".o"))
        subprocess.run(["gcc", "-std=gnu99", "-O1", "-g", "-fopenmp", *SAN, "-I", INC, "-I", CSRC, "-c",
                        os.path.join(CSRC, src), "-o", o], check=True)
        objs.append(o)
    exe = str(tmp_path / "sanitize_driver")
    # -O0 with line tables only: the instrumented template bodies build in
    # about a minute (about five at -O1 -g), and no access is optimised away
    subprocess.run(["g++", "-std=c++17", "-O0", "-g1", "-fopenmp", "-ffp-contract=off", "-Wno-unknown-pragmas", *SAN,
                    "-I", INC, "-I", CSRC, "-I", os.path.join(common.REPO, "tests"),
                    os.path.join(common.REPO, "tests", "sanitize_driver.cpp"), *objs, "-o", exe, "-lm"], check=True)
    # the sanitizer runtime need not come first in the library list: the
    # environment may preload other libraries
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "sanitize driver: ok" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]
