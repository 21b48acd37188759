"""Pins the oracle against the reference itself.

* Against the reference's C compiled in place (oracle/_ref, strict build,
  -ffp-contract=off): every hot-path function bit for bit on random state,
  and rtm_routine end to end (these tests need /root/reference).
* Against the committed golden outputs of the reference (tests/golden/,
  written by tests/golden/make_golden.py): runs anywhere.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import tempfile

import numpy as np
import pytest

import common
import oracle as O
import rtmuq2d_amd as R

CFGS = ["tiny", "asym"]


def _ref_lib(name):
    if not common.ref_built(name):
        pytest.skip(f"reference build {name} absent (needs /root/reference)")
    return C.CDLL(os.path.join(common.ref_dir(name), "librtmref.so"))


def vp(a):
    return a.ctypes.data_as(C.c_void_p)


def _rand_state(cfg, seed):
    rng = np.random.default_rng(seed)
    X = cfg["thicknessLeft"] + cfg["Nx"] + cfg["thicknessRight"]
    Z = cfg["thicknessUp"] + cfg["Nz"] + cfg["thicknessDown"]
    f = lambda *s: rng.standard_normal(s).astype(np.float32)  # noqa: E731
    return X, Z, f


@pytest.mark.parametrize("name", CFGS)
def test_profiles_extend_laplacian_bitexact(name):
    ref = _ref_lib(name)
    cfg = common.REF_CONFIGS[name]
    op = common.orc_params(cfg)
    prof = O.profiles(op)
    prof_ref = [np.zeros_like(a) for a in prof]
    ref.attenuation_profile_cpml(*[vp(a) for a in prof_ref])
    for a, b in zip(prof, prof_ref):
        assert np.array_equal(a, b)
    model = np.random.default_rng(1).uniform(1500, 4000, (cfg["Nx"], cfg["Nz"])).astype(np.float32)
    e_ref = np.zeros((op.X, op.Z), np.float32)
    ref.extend_matrix(vp(e_ref), vp(model))
    assert np.array_equal(O.extend(op, model), e_ref)
    img = np.random.default_rng(2).standard_normal((cfg["Nx"], cfg["Nz"])).astype(np.float32)
    lp_ref = np.zeros_like(img)
    ref.laplacian_filter2D(vp(lp_ref), vp(img))
    assert np.array_equal(O.laplacian(op, img), lp_ref)


@pytest.mark.parametrize("name", CFGS)
def test_stencil_cpml_image_bitexact(name):
    ref = _ref_lib(name)
    cfg = common.REF_CONFIGS[name]
    op = common.orc_params(cfg)
    X, Z, f = _rand_state(cfg, 3)
    Cf = (np.abs(f(X, Z)) * 0.1).astype(np.float32)
    u, U = f(X, Z), f(X, Z)
    U_ref = U.copy()
    ref.iso_acoustic_wave_equation(vp(Cf), vp(u), vp(U_ref))
    assert np.array_equal(O.stencil(op, Cf, u, U), U_ref)

    TL, TR, TU, TD = cfg["thicknessLeft"], cfg["thicknessRight"], cfg["thicknessUp"], cfg["thicknessDown"]
    shapes = [(TL, Z), (TR, Z), (X, TU), (X, TD)]
    psi = [f(*s) for s in shapes]
    csi = [f(*s) for s in shapes]
    psi_r, csi_r = [a.copy() for a in psi], [a.copy() for a in csi]
    prof = O.profiles(op)
    aL, aR, aT, aB, bL, bR, bT, bB = prof
    U_ref = U.copy()
    ref.aiso_attenuation_cpml(vp(u), vp(U_ref), vp(Cf), vp(aL), vp(bL), vp(aR), vp(bR), vp(aT), vp(bT),
                              vp(aB), vp(bB), *[vp(a) for a in psi_r], *[vp(a) for a in csi_r])
    U_o = O.cpml(op, u, U, Cf, prof, psi, csi)
    assert np.array_equal(U_o, U_ref)
    for a, b in zip(psi + csi, psi_r + csi_r):
        assert np.array_equal(a, b)

    S, Rw = f(X, Z), f(X, Z)
    I, Is = f(cfg["Nx"], cfg["Nz"]), f(cfg["Nx"], cfg["Nz"])
    I_r, Is_r = I.copy(), Is.copy()
    ref.normalized_cross_correlation(vp(I_r), vp(Is_r), vp(S), vp(Rw))
    O.image(op, I, Is, S, Rw)
    assert np.array_equal(I, I_r) and np.array_equal(Is, Is_r)


@pytest.mark.parametrize("name", CFGS)
def test_rtm_routine_end_to_end_bitexact(name):
    """The reference's rtm_routine (files in, files out) against the oracle's
    shot loop on the same inputs: bit-identical stacked image and Laplacian.
    Also regenerates the inputs and checks them against the golden ones."""
    if not common.ref_built(name):
        pytest.skip("reference build absent")
    cfg = common.REF_CONFIGS[name]
    g = np.load(os.path.join(common.GOLDEN, f"{name}.npz"))
    v_true, models, wav, seis = common.make_inputs(cfg)
    assert np.array_equal(np.stack(models), g["models"]), "input generator drifted"
    assert np.array_equal(np.stack(seis), g["seismograms"])
    zs = [R.zfp_compress(d, cfg["zfp_tolerance"]) for d in seis]
    for s, z in enumerate(zs):
        assert z == g[f"zfp_{s}"].tobytes(), "ZFP stream drifted"
    with tempfile.TemporaryDirectory() as run:
        common.write_inputs(run, cfg, models, wav, zs)
        common.run_reference(run, name, "strict", 1, len(models))
        outs = [common.read_outputs(run, cfg, s + 1) for s in range(len(models))]
    op = common.orc_params(cfg)
    for s, m in enumerate(models):
        vpad = O.extend(op, m)
        stack = np.zeros((cfg["Nx"], cfg["Nz"]), np.float32)
        for si, (sx, sz) in enumerate(common.shot_positions(cfg)):
            with tempfile.NamedTemporaryFile(delete=False) as f:
                f.write(zs[si])
            d = R.zfp_decompress_file(f.name, cfg["numberOfReceivers"], cfg["numberOfTimeStep"],
                                      cfg["zfp_tolerance"])
            os.unlink(f.name)
            I, _ = O.rtm_shot(op, sx, sz, vpad, wav, d)
            stack = stack + I
        assert np.array_equal(stack, outs[s][0])
        assert np.array_equal(O.laplacian(op, stack), outs[s][1])


@pytest.mark.parametrize("name", CFGS)
def test_oracle_matches_golden(name):
    """Runs without the reference: the oracle's images equal the reference's
    committed outputs."""
    cfg = common.REF_CONFIGS[name]
    g = np.load(os.path.join(common.GOLDEN, f"{name}.npz"))
    op = common.orc_params(cfg)
    for s in range(g["models"].shape[0]):
        vpad = O.extend(op, g["models"][s])
        stack = np.zeros((cfg["Nx"], cfg["Nz"]), np.float32)
        for si, (sx, sz) in enumerate(common.shot_positions(cfg)):
            with tempfile.NamedTemporaryFile(delete=False) as f:
                f.write(g[f"zfp_{si}"].tobytes())
            d = R.zfp_decompress_file(f.name, cfg["numberOfReceivers"], cfg["numberOfTimeStep"],
                                      cfg["zfp_tolerance"])
            os.unlink(f.name)
            I, _ = O.rtm_shot(op, sx, sz, vpad, g["wavelet"], d)
            stack = stack + I
        assert np.array_equal(stack, g["brute"][s])
        assert np.array_equal(O.laplacian(op, stack), g["laplacian"][s])


def test_oracle_c0_matches_golden_checksums():
    """Config c0 (600x200, nt=2000, TS=20): the oracle's image has the
    checksum of the reference's image (strict build), and is within relL2
    1e-5 of the reference built -O3 -march=native (FMA contraction)."""
    cfg = common.REF_CONFIGS["c0"]
    meta = json.load(open(os.path.join(common.GOLDEN, "c0.json")))
    native = np.load(os.path.join(common.GOLDEN, "c0_native.npz"))["brute"]
    v_true, models, wav, seis = common.make_inputs(cfg)
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    assert sha(models[0]) == meta["models"][0]
    z = R.zfp_compress(seis[0], cfg["zfp_tolerance"])
    assert hashlib.sha256(z).hexdigest() == meta["zfp"][0]
    with tempfile.NamedTemporaryFile(delete=False) as f:
        f.write(z)
    d = R.zfp_decompress_file(f.name, cfg["numberOfReceivers"], cfg["numberOfTimeStep"], cfg["zfp_tolerance"])
    os.unlink(f.name)
    op = common.orc_params(cfg)
    I, _ = O.rtm_shot(op, *common.shot_positions(cfg)[0], O.extend(op, models[0]), wav, d)
    stack = np.zeros_like(I) + I
    assert sha(stack) == meta["stackedBrute"][0]
    assert common.rel_l2(stack, native) <= 1e-5
