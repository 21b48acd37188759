"""The C-ABI library: loads without a GPU, exports every symbol declared in
include/*.h, and its host-side pieces (params, profiles, padding, the
Laplacian post-filter, raw files, ZFP) match the oracle / reference."""
from __future__ import annotations

import ctypes as C
import glob
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import common
import oracle as O
import rtmuq2d_amd as R


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(common.REPO, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            if m.group(1) not in ("if", "while", "for", "return", "sizeof"):
                names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol():
    names = declared_functions()
    assert "rtm_routine" in names and "aiso_attenuation_cpml" in names and "rtm_plan_forward" in names
    out = subprocess.run(["nm", "-D", "--defined-only", R.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = R.lib()
    for n in names:
        getattr(L, n)


def test_no_kernel_spills_to_scratch():
    """The compiler's resource report from the last build: no kernel of the
    library uses per-lane scratch (spilled registers or a stack array cost a
    round trip to memory per access), and the FP32 step kernels keep at least
    three waves per SIMD."""
    path = os.path.join(common.REPO, "rtmuq2d_amd", "lib", "kernel_resources.json")
    if not os.path.exists(path):
        pytest.skip("no resource report: the library was not built by rtmuq2d_amd/build.py")
    import json
    res = json.load(open(path))
    assert any("k_tiles" in k for k in res) and any("k_columns" in k for k in res)
    scratch = {k: v for k, v in res.items() if v.get("scratch_bytes", 0)}
    assert not scratch, scratch
    for k, v in res.items():
        if ("k_tiles" in k or "k_columns" in k) and "IfL" in k and "Li8E" not in k:
            assert v["occupancy"] >= 3, (k, v)


def test_params_struct_layout_matches_the_header(tmp_path):
    """The ctypes mirror of rtm_params has the C struct's size and every
    field at the C offset (compiled against include/rtm_hip.h)."""
    fields = [f for f, _ in R.RtmParams._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"rtm_hip.h\"\nint main(void) {\n"
                   "    printf(\"%zu\\n\", sizeof(rtm_params));\n"
                   + "".join(f'    printf("%zu\\n", offsetof(rtm_params, {f}));\n' for f in fields)
                   + "    return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(common.REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == C.sizeof(R.RtmParams)
    for f, off in zip(fields, out[1:]):
        assert getattr(R.RtmParams, f).offset == off, f


def test_default_params_are_the_reference_configuration():
    p = R.default_params()
    assert (p.nx, p.nz, p.number_of_time_steps, p.sampling_wavelet) == (2721, 561, 14000, 800)
    assert (p.thickness_left, p.thickness_right, p.thickness_up, p.thickness_down) == (50, 50, 50, 50)
    assert (p.first_location_receiver, p.spacing_receiver, p.number_of_receivers) == (1, 3, 1361)
    assert abs(p.spacing_grid - 6.25) == 0 and abs(p.sampling - np.float32(0.0005)) == 0
    assert p.time_slices == 20
    assert p.velocity_file == b"./INPUTS/vp_marmousi2_%05d.bin"


def test_params_validation_errors():
    bad = R.make_params(number_of_time_steps=1001, time_slices=20)
    with pytest.raises(R.RtmError, match="multiple of time_slices"):
        R.set_params(bad)
    with pytest.raises(R.RtmError, match="thickness"):
        R.set_params(R.make_params(thickness_left=6))
    with pytest.raises(R.RtmError):
        R.Plan(bad)   # rejected on the host, before any device call


@pytest.mark.parametrize("name", ["tiny", "asym"])
def test_host_profiles_padding_laplacian(name):
    cfg = common.REF_CONFIGS[name]
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    R.set_params(p)
    prof = [np.zeros_like(a) for a in O.profiles(op)]
    R.lib().attenuation_profile_cpml(*[R.fptr(a) for a in prof])
    for a, b in zip(prof, O.profiles(op)):
        assert np.array_equal(a, b)
    model = np.random.default_rng(0).uniform(1500, 4500, (cfg["Nx"], cfg["Nz"])).astype(np.float32)
    assert np.array_equal(R.extend(p, model), O.extend(op, model))
    img = np.random.default_rng(1).standard_normal((cfg["Nx"], cfg["Nz"])).astype(np.float32)
    lp = np.zeros_like(img)
    R.lib().laplacian_filter2D(R.fptr(lp), R.fptr(img))
    assert np.array_equal(lp, O.laplacian(op, img))


def test_zfp_roundtrip_and_reference_streams():
    """The host codec reproduces the reference's committed seismogram
    streams byte for byte and decodes them within the stated tolerance."""
    g = np.load(os.path.join(common.GOLDEN, "tiny.npz"))
    cfg = common.REF_CONFIGS["tiny"]
    d = g["seismograms"][0]
    z = R.zfp_compress(d, cfg["zfp_tolerance"])
    assert z == g["zfp_0"].tobytes()
    with tempfile.NamedTemporaryFile(delete=False) as f:
        f.write(z)
    try:
        dd = R.zfp_decompress_file(f.name, *d.shape, cfg["zfp_tolerance"])
    finally:
        os.unlink(f.name)
    assert np.max(np.abs(dd - d)) <= cfg["zfp_tolerance"]
    assert not np.array_equal(dd, d)   # lossy: the decoded array is what the RTM consumes


def test_zfp_missing_file_reports_failure():
    a = np.zeros((4, 8), np.float32)
    rc = R.lib().zfp2d_lossyDecompression(b"/nonexistent/x.zfp", R.fptr(a), 4, 8, C.c_float(1e-4))
    assert rc == 1


def test_raw_files_roundtrip(tmp_path):
    a = np.random.default_rng(3).standard_normal((7, 5)).astype(np.float32)
    path = str(tmp_path / "m.bin").encode()
    assert R.lib().recording_file(path, R.fptr(a), 7, 5) == 0
    b = np.zeros_like(a)
    assert R.lib().reading_file(path, R.fptr(b), 7, 5) == 0
    assert np.array_equal(a, b)
    assert np.array_equal(np.fromfile(path.decode(), np.float32).reshape(7, 5), a)


def test_short_file_exits_like_the_reference(tmp_path):
    """reading_file on a short file exits the process with status 1
    (SRC/pi_file.c:25-28); run in a child process."""
    path = tmp_path / "short.bin"
    np.zeros(3, np.float32).tofile(path)
    code = ("import sys; sys.path.insert(0, %r); import numpy as np, rtmuq2d_amd as R; "
            "a = np.zeros(10, np.float32); R.lib().reading_file(%r, R.fptr(a), 2, 5)" %
            (common.REPO, str(path).encode()))
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True)
    assert r.returncode == 1 and "Error reading file" in r.stdout


DRIVER = os.path.join(common.REPO, "rtmuq2d_amd", "lib", "rtmuq2d_hip")


def _dry(*args):
    r = subprocess.run([DRIVER, *args, "--dry-run"], capture_output=True, text=True)
    return r.returncode, r.stdout, r.stderr


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_splits_simulations_like_the_reference(tmp_path):
    """rtmuq2d_hip gives rank r of N the simulation ids of the reference's
    MPI split, SRC/mpi_rtm_iso2d.c:39-43 (float arithmetic as there), for
    --ranks and for the per-device fan-out; FIRST LAST shift the range."""
    import math

    def ref(nsim, world, rank):
        per = np.float32(nsim) / np.float32(world)
        return (math.ceil(np.float32(per * np.float32(rank)) + np.float32(1)),
                math.ceil(np.float32(per * np.float32(rank + 1))))

    for nsim in (1, 2, 3, 7, 10, 64, 100):
        for world in (1, 2, 3, 4, 7, 8):
            f = tmp_path / "p.txt"
            f.write_text(f"numberOfSimulations = {nsim}\n")
            rc, out, _ = _dry("--params", str(f), "--devices", str(world))
            assert rc == 0
            got = [tuple(map(int, l.split()[-2:])) for l in out.splitlines() if l.startswith("rank ")]
            assert got == [ref(nsim, world, r) for r in range(world)], (nsim, world)
    rc, out, _ = _dry("--ranks", "4", "--rank", "3", "11", "20")
    assert rc == 0 and out.splitlines()[-1] == "rank 3 device 0 simulations %d %d" % tuple(
        10 + v for v in ref(10, 4, 3))


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_params_file(tmp_path):
    """Parameter files take rtm_params names or the reference's macro names,
    comments and quoted paths; unknown names and bad values are refused."""
    f = tmp_path / "p.txt"
    f.write_text("# c0 grid\nNx = 600\nnz=200  # inline comment\nthicknessLeft = 40\nspacing_grid = 10.0\n"
                 "stacked_file = \"./OUT/brute_%05d.bin\"\nnormalize_illumination = 1\n")
    rc, out, _ = _dry("--params", str(f))
    assert rc == 0
    kv = dict(l.split(" = ", 1) for l in out.splitlines() if " = " in l)
    assert kv["nx"] == "600" and kv["nz"] == "200" and kv["thickness_left"] == "40"
    assert kv["spacing_grid"] == "10" and kv["stacked_file"] == "./OUT/brute_%05d.bin"
    assert kv["normalize_illumination"] == "1" and kv["thickness_right"] == "50"   # default kept
    for bad in ("nonsense = 3\n", "Nx = 6x\n", "Nx 600\n"):
        f.write_text(bad)
        rc, _, err = _dry("--params", str(f))
        assert rc == 2 and err, bad
    assert _dry("--ranks", "2", "--rank", "2")[0] == 2


def _c_blocks(path):
    text = open(path).read()
    return re.findall(r"```c\n(.*?)```", text, flags=re.S)


def test_integration_guide_c_examples_compile(tmp_path):
    """The C examples of the integration guide compile against
    include/rtm_hip.h and link against the library: the device-resident API
    and the request server in a function body, and, where the reference's
    sources are present, the rtm_apply_fixed_parameters translation unit
    against the reference's own fixed_parameters.h."""
    blocks = _c_blocks(os.path.join(common.REPO, "INTEGRATION.md"))
    assert len(blocks) == 3
    fixed, api, server = blocks
    body = api.replace('#include "rtm_hip.h"\n', "")
    src = tmp_path / "api.c"
    src.write_text("#include <stdio.h>\n#include \"rtm_hip.h\"\n"
                   "int example(float *vpad, float *wavelet, float *seis, float *I, float *Is, float *image,\n"
                   "            int shot_x, int shot_z) {\n" + body + "\n" + server + "\nreturn 0;\n}\n")
    lib = os.path.dirname(R.LIB_PATH)
    subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Werror", "-Wno-unused-variable", "-shared", "-fPIC",
                    "-I", os.path.join(common.REPO, "include"), str(src), "-o", str(tmp_path / "api.so"),
                    "-L", lib, "-lrtm_hip"], check=True)
    ref_src = "/root/reference/SRC"
    if os.path.exists(os.path.join(ref_src, "fixed_parameters.h")):
        unit = tmp_path / "rtm_params_from_fixed.c"
        unit.write_text(fixed)
        subprocess.run(["gcc", "-std=gnu99", "-Wall", "-c", "-I", ref_src, "-I", os.path.join(common.REPO, "include"),
                        str(unit), "-o", str(tmp_path / "fixed.o")], check=True)
