"""GPU parity at the benchmark workloads' own grids and store layouts.

The small-grid parity tests (test_gpu_parity.py) cover every code path on
grids the oracle finishes in a second.  These run each benchmarked
configuration on its own grid, in the kernel shape the benchmark uses, and
compare with the oracle (or, for the 64-realization ensemble at full length,
with the plain plan) bit for bit:

* c1: the 1701x351 grid at TS = 1 with the FP32 store, 64x8 tiles, 2000
  steps, so that the store (4.78 GB) has snapshot offsets past 4 GiB;
* c2: one launch group of 13 realizations on the c1 grid (FP64 Welford
  against numpy), and all 64 realizations at nt = 4000, each batched image
  equal to its plain-plan image;
* c3: FP16 / BF16 stores and boundary saving on the c1 grid at TS = 4, with
  the Kahan image the benchmark uses;
* c4: one request of each serving grid class (3401x701 at h = 5 m,
  dt = 0.4 ms included), FP16 store at TS = 4, Kahan image, in one server;
* c0: both simulations of the reference's own run shape.

The time steps are shortened where the oracle would take minutes (noted per
test); the grids, thicknesses, store kinds and kernel shapes are the
benchmark's.  Snapshot store and imaging follow SRC/forward_modeling.c:111-124
and SRC/backward_modeling.c:112-120.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np
import pytest

import common
import oracle as O
import rtmuq2d_amd as R
from rtmuq2d_amd import synth

sys.path.insert(0, common.REPO)
import bench  # noqa: E402  (the benchmark's workload definitions)

pytestmark = pytest.mark.gpu

C1_SHAPE = (2, 100, 0, 0)        # 64x8 tiles: the autotuner's pick on the c1 grid


def say(*a):
    print(*a, flush=True)


def c1(nt, ts, **extra):
    return bench.c1_params().update(number_of_time_steps=nt, sampling_wavelet=nt, time_slices=ts, **extra)


_C1_INPUTS = {}


def c1_inputs(nt):
    """The benchmark's c1 models and wavelet, and the traces modelled by the
    oracle on the true model over nt steps (cached per nt)."""
    if nt not in _C1_INPUTS:
        p = c1(nt, 1)
        v_true, v_mig, w = bench.c1_inputs(p)
        op = O.from_rtm(p)
        _, d, _, _ = O.forward(op, p.first_location_shot, p.depth_shot, O.extend(op, v_true), w,
                               snaps=False, record=True)
        _C1_INPUTS[nt] = (v_true, v_mig, w, d)
    return _C1_INPUTS[nt]


_C1_REF = {}


@pytest.mark.parametrize("shape", [C1_SHAPE, (2, 100, R.INTERIOR_PAIRS, 0), (4, 100, R.INTERIOR_PAIRS, 64)])
def test_c1_ts1_fp32_store_past_4gib_bitexact(gpu_available, shape):
    """The plain c1 workload's path: c1 grid, FP32 snapshot every step, in
    64x8 tiles and with two-step pair launches on the inner image region.
    2000 steps make a 4.78 GB store, so the per-snapshot base pointers pass
    4 GiB; the image equals the oracle's over the same steps."""
    nt = 2000
    p = c1(nt, 1)
    nint = p.nx * p.nz
    assert nt * nint * 4 > 4 << 30
    say(f"[gpu] c1 TS=1 FP32 store, {nt} steps, store {nt * nint * 4 / 1e9:.2f} GB, shape {shape}")
    v_true, v_mig, w, d = c1_inputs(nt)
    op = O.from_rtm(p)
    vpad = O.extend(op, v_mig)
    sx, sz = p.first_location_shot, p.depth_shot
    if nt not in _C1_REF:
        _C1_REF[nt] = O.rtm_shot(op, sx, sz, vpad, w, d)
    I_ref, Is_ref = _C1_REF[nt]
    say("[gpu] c1 TS=1: oracle done")
    with R.Plan(p) as plan:
        plan.set_tiling(*shape)
        plan.set_velocity(vpad)
        plan.set_wavelet(w)
        plan.set_seismogram(d)
        plan.zero_image()
        plan.migrate(sx, sz)
        I, Is = plan.image()
    assert np.array_equal(I, I_ref), f"relL2 {common.rel_l2(I, I_ref)}"
    assert np.array_equal(Is, Is_ref)


_C2 = {}


def _c2_group_refs(nt, nsims):
    if (nt, nsims) not in _C2:
        p = c1(nt, 20)
        v_true, _, w, d = c1_inputs(nt)
        op = O.from_rtm(p)
        vpads = [O.extend(op, synth.ensemble_velocity(v_true, 100 + k)) for k in range(nsims)]
        refs = [O.rtm_shot(op, p.first_location_shot, p.depth_shot, v, w, d) for v in vpads]
        _C2[(nt, nsims)] = (p, vpads, w, d, refs)
    return _C2[(nt, nsims)]


@pytest.mark.parametrize("shape", [(4, 100, 0, 0), (4, 100, 1, 32), (4, 100, 4, 0), (2, 99, 4, 48)])
def test_c2_launch_group_of_13_on_c1_grid(gpu_available, shape):
    """One c2 launch group (13 realizations of the c1 grid, TS = 20, the
    size the group autotuner picks), steps shortened to 200: every image
    equals the oracle's for its realization, and the FP64 Welford mean and
    variance equal numpy's over those images."""
    nt, nsims = 200, 13
    say(f"[gpu] c2 launch group of {nsims} on the c1 grid, {nt} steps, shape {shape}")
    p, vpads, w, d, refs = _c2_group_refs(nt, nsims)
    say("[gpu] c2 group: oracle done")
    with R.Batch(p, nsims) as b:
        b.set_group(nsims)
        b.set_tiling(*shape)
        assert b.group() == (nsims, 1)
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(w)
        b.set_seismogram(d)
        b.zero_images()
        b.migrate(p.first_location_shot, p.depth_shot)
        for k in range(nsims):
            I, Is = b.image(k)
            assert np.array_equal(I, refs[k][0]), (k, common.rel_l2(I, refs[k][0]))
            assert np.array_equal(Is, refs[k][1]), k
        mean, var = b.welford()
    stack = np.stack([r[0].astype(np.float64) for r in refs])
    scale = np.abs(stack).max()
    assert np.allclose(mean, stack.mean(0), rtol=1e-13, atol=1e-13 * scale)
    assert np.allclose(var, stack.var(0, ddof=1), rtol=1e-10, atol=1e-12 * stack.var(0).max())


_C2_PLAIN = {}


@pytest.mark.parametrize("shape,group,lanes", [(None, 13, 1), ((4, 100, R.INTERIOR_PAIRS, 0), 13, 1), (None, 6, 2)])
def test_c2_full_ensemble_equals_plain_images(gpu_available, shape, group, lanes):
    """The whole c2 workload (64 realizations, nt = 4000, TS = 20): launch
    groups of 13 in the default shape and with two-step pair launches, and
    groups of at most 6 two at a time (12 groups on two lanes): every
    batched image equals the plain plan's image of the same realization
    (64x16 tiles), bit for bit (the plain path is bit-exact against the
    oracle and the reference at this size)."""
    nsims = 64
    p = c1(4000, 20)
    v_true, _, w = bench.c1_inputs(p)
    say(f"[gpu] c2 full ensemble: {nsims} x nt {p.number_of_time_steps}, batch vs plain, shape {shape}")
    sx, sz = p.first_location_shot, p.depth_shot
    vpads = [R.extend(p, synth.ensemble_velocity(v_true, 100 + k)) for k in range(nsims)]
    if not _C2_PLAIN:
        plain = []
        with R.Plan(p) as plan:
            plan.set_tiling(4, 100, 0, 0)
            plan.set_velocity(R.extend(p, v_true))
            plan.set_wavelet(w)
            plan.forward(sx, sz, record=True)
            d = plan.seismogram()
            plan.set_seismogram(d)
            for k, v in enumerate(vpads):
                plan.set_velocity(v)
                plan.zero_image()
                plan.migrate(sx, sz)
                plain.append(plan.image()[0])
        _C2_PLAIN["d"], _C2_PLAIN["plain"] = d, plain
        say("[gpu] c2 full ensemble: plain images done")
    d, plain = _C2_PLAIN["d"], _C2_PLAIN["plain"]
    with R.Batch(p, nsims) as b:
        b.set_group(group)
        b.set_lanes(lanes)
        if shape:
            b.set_tiling(*shape)
        assert b.group() == ((13, 5) if lanes == 1 else (6, 12))
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(w)
        b.set_seismogram(d)
        b.zero_images()
        b.migrate(sx, sz)
        for k in range(nsims):
            I, _ = b.image(k)
            assert np.array_equal(I, plain[k]), (k, common.rel_l2(I, plain[k]))


_C3 = {}


def _c3_oracle(nt):
    if nt not in _C3:
        p = c1(nt, 4, image_kahan=1)
        _, v_mig, w, d = c1_inputs(nt)
        op = O.from_rtm(p)
        vpad = O.extend(op, v_mig)
        S, _, _, _ = O.forward(op, p.first_location_shot, p.depth_shot, vpad, w, record=False)
        _C3[nt] = (p, op, vpad, w, d, S)
    return _C3[nt]


@pytest.mark.parametrize("mode", [R.SNAP_FP16, R.SNAP_BF16])
def test_c3_sixteen_bit_store_on_c1_grid(gpu_available, mode):
    """c3's 16-bit stores on the c1 grid at TS = 4 (the benchmark's shape and
    Kahan image), 400 steps: the image equals the oracle's Kahan backward pass
    over the oracle's FP32 snapshots put through the store's rule."""
    nt = 400
    say(f"[gpu] c3 16-bit store mode {mode} on the c1 grid, TS=4, {nt} steps, Kahan image")
    p, op, vpad, w, d, S = _c3_oracle(nt)
    deq, _, _ = common.quantize_snapshots(S, mode)
    I_ref, Is_ref, _, _ = O.backward(op, vpad, d, deq)
    q = p.copy().update(snapshot_mode=mode)
    with R.Plan(q) as plan:
        plan.set_tiling(*C1_SHAPE)
        plan.set_velocity(vpad)
        plan.set_wavelet(w)
        plan.set_seismogram(d)
        plan.zero_image()
        plan.migrate(p.first_location_shot, p.depth_shot)
        I, Is = plan.image()
    assert np.array_equal(I, I_ref), f"relL2 {common.rel_l2(I, I_ref)}"
    assert np.array_equal(Is, Is_ref)


def test_c3_int16_block_store_on_c1_grid(gpu_available):
    """The int16 block store on the c1 grid at TS = 4 (the benchmark's shape
    and Kahan image), 400 steps: the image equals the oracle's Kahan
    backward pass over the oracle's FP32 snapshots put through the store's
    rule."""
    nt = 400
    say(f"[gpu] c3 int16 block store on the c1 grid, TS=4, {nt} steps, Kahan image")
    p, op, vpad, w, d, S = _c3_oracle(nt)
    deq, _, _ = common.quantize_i16_blocks(S, op.t_left)
    I_ref, Is_ref, _, _ = O.backward(op, vpad, d, deq)
    q = p.copy().update(snapshot_mode=R.SNAP_I16)
    with R.Plan(q) as plan:
        plan.set_tiling(*C1_SHAPE)
        plan.set_velocity(vpad)
        plan.set_wavelet(w)
        plan.set_seismogram(d)
        plan.zero_image()
        plan.migrate(p.first_location_shot, p.depth_shot)
        I, Is = plan.image()
    assert np.array_equal(I, I_ref), f"relL2 {common.rel_l2(I, I_ref)}"
    assert np.array_equal(Is, Is_ref)


def test_c3_fp32_store_kahan_on_c1_grid(gpu_available):
    """c3's FP32 store with the Kahan image on the c1 grid at TS = 4: the
    image equals the oracle's compensated image bit for bit."""
    nt = 400
    say(f"[gpu] c3 FP32 store, Kahan image, c1 grid, TS=4, {nt} steps")
    p, op, vpad, w, d, S = _c3_oracle(nt)
    I_ref, Is_ref, _, _ = O.backward(op, vpad, d, S)
    with R.Plan(p) as plan:
        plan.set_tiling(*C1_SHAPE)
        plan.set_velocity(vpad)
        plan.set_wavelet(w)
        plan.set_seismogram(d)
        plan.zero_image()
        plan.migrate(p.first_location_shot, p.depth_shot)
        I, Is = plan.image()
    assert np.array_equal(I, I_ref), f"relL2 {common.rel_l2(I, I_ref)}"
    assert np.array_equal(Is, Is_ref)


def test_c3_boundary_saving_on_c1_grid(gpu_available):
    """c3's boundary saving on the c1 grid at TS = 4, Kahan image, 40 steps
    (the reconstruction's restatement runs one oracle forward pass per step):
    the image equals the oracle's backward pass over the reconstructed source
    field, bit for bit."""
    nt = 40
    say(f"[gpu] c3 boundary saving on the c1 grid, TS=4, {nt} steps")
    p = c1(nt, 4, image_kahan=1)
    _, v_mig, w, d = c1_inputs(nt)
    op = O.from_rtm(p)
    vpad = O.extend(op, v_mig)
    sx, sz = p.first_location_shot, p.depth_shot
    S_bs = common.boundary_saving_snapshots(op, sx, sz, vpad, w)
    I_ref, Is_ref, _, _ = O.backward(op, vpad, d, S_bs)
    with R.Plan(p.copy().update(snapshot_mode=R.SNAP_BOUNDARY)) as plan:
        plan.set_tiling(*C1_SHAPE)
        plan.set_velocity(vpad)
        plan.set_wavelet(w)
        plan.set_seismogram(d)
        plan.zero_image()
        plan.migrate(sx, sz)
        I, Is = plan.image()
    assert np.array_equal(I, I_ref), f"relL2 {common.rel_l2(I, I_ref)}"
    assert np.array_equal(Is, Is_ref)


# c4: steps per class in this test (the benchmark's nt is 3200 / 4000 / 4000
# / 8000; the two largest grids are shortened so that the oracle finishes in
# seconds)
C4_TEST_NT = {"600x200 h10": 3200, "1001x301 h10": 4000, "1701x351 h10": 1000, "3401x701 h5": 800}


def test_c4_one_request_of_each_grid_class(gpu_available):
    """One request of every c4 grid class in one server (FP16 store at
    TS = 4, Kahan image, the benchmark's kernel shape and launch-group
    budget), each with its own velocity realization, source position and
    traces: every image equals the oracle's for that request alone under the
    FP16 store rule, bit for bit."""
    reqs = []
    for k, cls in enumerate(bench.SERVE_CLASSES):
        nt = C4_TEST_NT[cls[0]]
        p = bench.serve_params(R, cls).update(number_of_time_steps=nt, sampling_wavelet=nt)
        assert p.image_kahan == 1 and p.time_slices == 4 and p.snapshot_mode == R.SNAP_FP16
        v = synth.folded_faulted(p.nx, p.nz, p.spacing_grid, 300 + k)
        wav = synth.ricker(cls[7], p.sampling, p.sampling_wavelet)
        g = np.random.default_rng(400 + k)
        seis = (g.standard_normal((p.number_of_receivers, nt)) * 1e-3).astype(np.float32)
        sx = int(g.integers(0, p.nx))
        op = O.from_rtm(p)
        vpad = O.extend(op, v)
        reqs.append((cls[0], p, op, vpad, wav, seis, sx))
    refs = []
    for label, p, op, vpad, wav, seis, sx in reqs:
        say(f"[gpu] c4 {label}: oracle, {p.number_of_time_steps} steps")
        S, _, _, _ = O.forward(op, sx, 2, vpad, wav, record=False)
        refs.append(O.backward(op, vpad, seis, common.quantize_snapshots(S, 1)[0])[0])
        del S
    say("[gpu] c4: serving the four requests")
    arena = sum(R.request_footprint(r[1]) for r in reqs) + (64 << 20)
    got = {}
    with R.Server(arena, max_slots=8, snapshot_mode=R.SNAP_FP16) as srv:
        srv.set_tiling(4, 100, 0, 0)                    # bench.py --serve-tiling default
        ids = [srv.submit(p, vpad, wav, seis, sx, 2) for _, p, _, vpad, wav, seis, sx in reqs]
        while srv.tick() > 0:
            for rid, I in srv.done():
                got[rid] = I
        for rid, I in srv.done():
            got[rid] = I
    assert sorted(got) == sorted(ids)
    for k, rid in enumerate(ids):
        assert np.array_equal(got[rid], refs[k]), (reqs[k][0], common.rel_l2(got[rid], refs[k]))


def test_c0_both_simulations_match_reference(gpu_available):
    """Config c0 as the reference runs it (run_rtm_simulations.local:6-10:
    two simulations): both images, through the drop-in rtm_routine, carry
    the reference strict build's checksums."""
    meta = json.load(open(os.path.join(common.GOLDEN, "c0.json")))
    cfg = common.REF_CONFIGS["c0"]
    v_true, models, wav, seis = common.make_inputs(cfg)
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    assert [sha(m) for m in models] == meta["models"][:2], "input generator drifted"
    zs = R.zfp_compress(seis[0], cfg["zfp_tolerance"])
    assert hashlib.sha256(zs).hexdigest() == meta["zfp"][0]
    p = common.rtm_params(cfg)
    with tempfile.TemporaryDirectory() as run:
        common.write_inputs(run, cfg, models, wav, [zs])       # one shot: files keyed by shot id
        cwd = os.getcwd()
        os.chdir(run)
        try:
            R.set_params(p)
            for s in (1, 2):
                say(f"[gpu] c0 simulation {s}")
                R.lib().rtm_routine(s)
        finally:
            os.chdir(cwd)
        for s in (1, 2):
            b, lap = common.read_outputs(run, cfg, s)
            assert sha(b) == meta["stackedBrute"][s - 1], f"simulation {s}"
            assert sha(lap) == meta["stackedLaplacian"][s - 1], f"simulation {s}"
