"""GPU parity of the HIP path against the oracle and the reference's golden
outputs (FP32 bit-exact; FP64 bit-exact against the FP64 oracle).

Everything goes through the C ABI of librtm_hip.so.
"""
from __future__ import annotations

import ctypes as C
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

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def golden(name):
    return np.load(os.path.join(common.GOLDEN, f"{name}.npz"))


@pytest.mark.parametrize("name", ["tiny", "asym"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_forward_backward_bitexact(gpu_available, name, dtype):
    say(f"[gpu] forward/backward {name} {np.dtype(dtype).name}")
    cfg = common.REF_CONFIGS[name]
    g = golden(name)
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    vpad = O.extend(op, g["models"][0])
    sx, sz = common.shot_positions(cfg)[0]
    S, d_orc, lu, lU = O.forward(op, sx, sz, vpad, g["wavelet"], dtype=dtype, record=True)
    real = R.REAL_F64 if dtype == np.float64 else R.REAL_F32
    with R.Plan(p, real) as plan:
        plan.set_velocity(vpad)
        plan.set_wavelet(g["wavelet"])
        plan.forward(sx, sz, record=True)
        a, b = plan.fields(f64=dtype == np.float64)
        assert np.array_equal(a, lu) and np.array_equal(b, lU), "forward fields differ"
        if dtype == np.float32:
            assert np.array_equal(plan.seismogram(), d_orc), "recorded seismogram differs"
        d = g["seismograms"][0]
        I, Is, lu2, lU2 = O.backward(op, vpad, d, S)
        plan.set_seismogram(d)
        plan.zero_image()
        plan.backward()
        a, b = plan.fields(f64=dtype == np.float64)
        assert np.array_equal(a, lu2) and np.array_equal(b, lU2), "backward fields differ"
        I2, Is2 = plan.image(f64=dtype == np.float64)
        assert np.array_equal(I2, I), f"image relL2={common.rel_l2(I2, I)}"
        assert np.array_equal(Is2, Is)


@pytest.mark.parametrize("rz", [1, 2, 4, 8])
@pytest.mark.parametrize("block", [0, 7, 100])
def test_tilings_and_graph_blocks_bitexact(gpu_available, rz, block):
    """Every tile height and graph-block size gives the same bits (tiny)."""
    say(f"[gpu] tiling rz={rz} graph_block={block}")
    cfg = common.REF_CONFIGS["tiny"]
    g = golden("tiny")
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    vpad = O.extend(op, g["models"][1])
    sx, sz = common.shot_positions(cfg)[1]
    d = g["seismograms"][1]
    I_orc, Is_orc = O.rtm_shot(op, sx, sz, vpad, g["wavelet"], d)
    with R.Plan(p) as plan:
        plan.set_tiling(rz, block)
        plan.set_velocity(vpad)
        plan.set_wavelet(g["wavelet"])
        plan.set_seismogram(d)
        plan.zero_image()
        plan.migrate(sx, sz)
        I, Is = plan.image()
    assert np.array_equal(I, I_orc) and np.array_equal(Is, Is_orc)


def _wide_case():
    """A grid whose interior spans three tile columns (one 128-column and one
    64-column item) and several column segments, with the source and the
    receivers in the interior."""
    cfg = dict(Nx=250, Nz=130, thicknessLeft=20, thicknessRight=22, thicknessUp=18, thicknessDown=21,
               spacingGrid=10.0, sampling=0.001, numberOfTimeStep=120, sampligWavelet=120,
               firstLocationShot=100, depthShot=30, spacingShot=1, numberOfShots=1,
               firstLocationReceiver=3, depthReceiver=35, spacingReceiver=3, numberOfReceivers=100,
               numberOfSimulations=1, zfp_tolerance=1e-4)
    rng = np.random.default_rng(21)
    model = rng.uniform(1800.0, 3200.0, (cfg["Nx"], cfg["Nz"])).astype(np.float32)
    from rtmuq2d_amd import synth
    wav = synth.ricker(25.0, 0.001, cfg["numberOfTimeStep"])
    seis = rng.standard_normal((cfg["numberOfReceivers"], cfg["numberOfTimeStep"])).astype(np.float32)
    return cfg, model, wav, seis


@pytest.mark.parametrize("rz,interior,seg", [(4, 1, 0), (2, 1, 12), (1, 1, 4), (8, 1, 32), (4, 1, 16), (4, 2, 0),
                                             (2, 2, 0), (2, 3, 0), (4, 3, 0)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_marched_interior_bitexact(gpu_available, rz, interior, seg, dtype):
    """Strip tiles plus the lane-shift column kernel on the interior
    (including a 64-column item), and mixed tile lists (64 x 4 tiles on the
    strips beside full interior tiles): fields, recorded traces and images
    bit-identical to the oracle."""
    say(f"[gpu] interior={interior} rz={rz} rows={seg} {np.dtype(dtype).name}")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=10)
    p = common.rtm_params(cfg, time_slices=10)
    vpad = O.extend(op, model)
    sx, sz = common.shot_positions(cfg)[0]
    S, d_orc, lu, lU = O.forward(op, sx, sz, vpad, wav, dtype=dtype, record=True)
    I, Is, lu2, lU2 = O.backward(op, vpad, seis, S)
    f64 = dtype == np.float64
    with R.Plan(p, R.REAL_F64 if f64 else R.REAL_F32) as plan:
        plan.set_tiling(rz, 100, interior, seg)
        plan.set_velocity(vpad)
        plan.set_wavelet(wav)
        plan.forward(sx, sz, record=True)
        a, b = plan.fields(f64=f64)
        assert np.array_equal(a, lu) and np.array_equal(b, lU)
        if not f64:
            assert np.array_equal(plan.seismogram(), d_orc)
        plan.set_seismogram(seis)
        plan.zero_image()
        plan.backward()
        a, b = plan.fields(f64=f64)
        assert np.array_equal(a, lu2) and np.array_equal(b, lU2)
        I2, Is2 = plan.image(f64=f64)
    assert np.array_equal(I2, I) and np.array_equal(Is2, Is)


def _band_case():
    """The wide grid with the receivers on padded row 18 and the source at
    padded column 66, both in the 4-point band around the two-step region
    (which starts at row 20 and column 68): tiles own those points and the
    pair items recompute them as margins.  55 steps, an odd count."""
    cfg, model, wav, seis = _wide_case()
    nt = 55
    cfg = dict(cfg, depthReceiver=0, firstLocationShot=46, depthShot=4, numberOfTimeStep=nt, sampligWavelet=nt,
               firstLocationReceiver=0, spacingReceiver=2, numberOfReceivers=250)
    from rtmuq2d_amd import synth
    seis = np.random.default_rng(5).standard_normal((250, nt)).astype(np.float32)
    return cfg, model, synth.ricker(25.0, 0.001, nt), seis


@pytest.mark.parametrize("case", ["wide", "band"])
@pytest.mark.parametrize("rz,block,seg", [(4, 100, 0), (1, 7, 8), (2, 5, 12), (8, 10, 32), (4, 1, 16), (2, 100, 0)])
def test_pair_launches_bitexact(gpu_available, case, rz, block, seg):
    """Two time steps per launch on the inner image region (k_pairs) beside
    one-step tiles (k_tiles over four u buffers), in graph blocks of `block`
    steps (odd blocks end on a launch whose second step is cut, 1 = eager):
    fields, recorded traces, snapshots and images bit-identical to the
    oracle."""
    say(f"[gpu] pair launches {case} rz={rz} block={block} seg={seg}")
    cfg, model, wav, seis = _band_case() if case == "band" else _wide_case()
    ts = 5 if case == "band" else 10
    op = common.orc_params(cfg, time_slices=ts)
    p = common.rtm_params(cfg, time_slices=ts)
    vpad = O.extend(op, model)
    sx, sz = cfg["firstLocationShot"], cfg["depthShot"]
    S, d_orc, lu, lU = O.forward(op, sx, sz, vpad, wav, record=True)
    I, Is, lu2, lU2 = O.backward(op, vpad, seis, S)
    with R.Plan(p) as plan:
        plan.set_tiling(rz, block, R.INTERIOR_PAIRS, seg)
        assert plan.tiling() == (rz, block, R.INTERIOR_PAIRS, seg)
        plan.set_velocity(vpad)
        plan.set_wavelet(wav)
        plan.forward(sx, sz, record=True)
        a, b = plan.fields()
        assert np.array_equal(a, lu) and np.array_equal(b, lU)
        assert np.array_equal(plan.seismogram(), d_orc)
        plan.set_seismogram(seis)
        plan.zero_image()
        plan.backward()
        a, b = plan.fields()
        assert np.array_equal(a, lu2) and np.array_equal(b, lU2)
        I2, Is2 = plan.image()
        assert np.array_equal(I2, I) and np.array_equal(Is2, Is)
        # back to one step per launch on the same plan: two buffers again
        plan.set_tiling(rz, 100, 0, 0)
        plan.zero_image()
        plan.migrate(sx, sz)
        I3, _ = plan.image()
    assert np.array_equal(I3, I)


def test_pair_launches_kahan_image_equals_tiles(gpu_available):
    """The compensated image through the pair items equals the tile
    kernel's compensated image bit for bit."""
    say("[gpu] pair launches, Kahan image")
    cfg, model, wav, seis = _wide_case()
    p = common.rtm_params(cfg, time_slices=4, image_kahan=1)
    vpad = O.extend(common.orc_params(cfg, time_slices=4), model)
    sx, sz = cfg["firstLocationShot"], cfg["depthShot"]
    out = []
    for shape in [(2, 100, 0, 0), (2, 100, R.INTERIOR_PAIRS, 16)]:
        with R.Plan(p) as plan:
            plan.set_tiling(*shape)
            plan.set_velocity(vpad)
            plan.set_wavelet(wav)
            plan.set_seismogram(seis)
            plan.zero_image()
            plan.migrate(sx, sz)
            out.append(plan.image())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("group,block", [(0, 100), (2, 9)])
def test_batch_pair_launches_bitexact(gpu_available, group, block):
    """A batch stepped with pair items (one launch group, or groups of 2
    with a ragged last group): every realization's image equals the
    oracle's."""
    say(f"[gpu] batch pair launches group={group} block={block}")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=10)
    p = common.rtm_params(cfg, time_slices=10)
    sx, sz = common.shot_positions(cfg)[0]
    vpads = [O.extend(op, model * np.float32(1 + 0.03 * k)) for k in range(3)]
    refs = [O.rtm_shot(op, sx, sz, v, wav, seis) for v in vpads]
    with R.Batch(p, len(vpads)) as b:
        b.set_group(group)
        b.set_tiling(4, block, R.INTERIOR_PAIRS, 24)
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(wav)
        b.set_seismogram(seis)
        b.zero_images()
        b.migrate(sx, sz)
        for k in range(len(vpads)):
            I, Is = b.image(k)
            assert np.array_equal(I, refs[k][0]) and np.array_equal(Is, refs[k][1]), k


def test_pair_shape_refusals(gpu_available):
    """Pair items step FP32 fields with the FP32 store: an FP64 plan, a
    16-bit store and the request server refuse the shape and keep theirs."""
    say("[gpu] pair shape refusals")
    cfg, _, _, _ = _wide_case()
    with R.Plan(common.rtm_params(cfg), R.REAL_F64) as plan:
        before = plan.tiling()
        with pytest.raises(R.RtmError):
            plan.set_tiling(4, 100, R.INTERIOR_PAIRS, 0)
        assert plan.tiling() == before
    with R.Plan(common.rtm_params(cfg, snapshot_mode=R.SNAP_FP16)) as plan:
        with pytest.raises(R.RtmError):
            plan.set_tiling(4, 100, R.INTERIOR_PAIRS, 0)
    with R.Server(1 << 28, max_slots=2, snapshot_mode=R.SNAP_FP32) as srv:
        with pytest.raises(R.RtmError):
            srv.set_tiling(4, 100, R.INTERIOR_PAIRS, 0)


@pytest.mark.parametrize("interior,seg", [(0, 0), (1, 16), (2, 0)])
def test_batch_marched_bitexact(gpu_available, interior, seg):
    """A batch with a streaming interior kernel: every realization's image
    equals the oracle's, in the batch as alone."""
    say(f"[gpu] batch with interior kernel {interior}")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=10)
    p = common.rtm_params(cfg, time_slices=10)
    sx, sz = common.shot_positions(cfg)[0]
    vpads = [O.extend(op, model * np.float32(1 + 0.03 * k)) for k in range(3)]
    refs = [O.rtm_shot(op, sx, sz, v, wav, seis) for v in vpads]
    with R.Batch(p, len(vpads)) as b:
        b.set_tiling(4, 100, interior, seg)
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(wav)
        b.set_seismogram(seis)
        b.zero_images()
        b.migrate(sx, sz)
        for k in range(len(vpads)):
            I, Is = b.image(k)
            assert np.array_equal(I, refs[k][0]) and np.array_equal(Is, refs[k][1]), k


@pytest.mark.parametrize("rz,group,interior,lanes", [(2, 0, 0, 1), (8, 0, 0, 1), (4, 2, 0, 1), (2, 4, 0, 1),
                                                     (2, 4, 3, 1), (4, 0, 3, 1), (4, 2, 0, 2), (4, 0, 0, 2),
                                                     (2, 1, 0, 3), (4, 2, 3, 4)])
def test_batch_ensemble_bitexact_and_welford(gpu_available, rz, group, interior, lanes):
    """A batch of realizations advanced in lockstep: each image bit-identical
    to the oracle's for that realization (alone or in the batch, in one
    launch group or in groups of at most 2 or 4: 5 realizations run as
    2, 2, 1 or as 3, 2; with 2 to 4 groups in flight at once on lanes of
    their own, the group count a multiple of the lanes), and the FP64
    Welford mean/variance equal to numpy's over the same images."""
    say(f"[gpu] batch ensemble rz={rz} group={group} interior={interior} lanes={lanes}")
    cfg = common.REF_CONFIGS["asym"]
    g = golden("asym")
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    sx, sz = common.shot_positions(cfg)[0]
    d = g["seismograms"][0]
    rng = np.random.default_rng(12)
    models = [g["models"][0] * np.float32(1 + 0.05 * k) + rng.uniform(-5, 5, g["models"][0].shape).astype(np.float32)
              for k in range(5)]
    vpads = [O.extend(op, m) for m in models]
    refs = [O.rtm_shot(op, sx, sz, v, g["wavelet"], d) for v in vpads]
    with R.Batch(p, len(models)) as b:
        b.set_tiling(rz, 100, interior)
        b.set_group(group)
        b.set_lanes(lanes)
        assert b.lanes() == lanes
        ng = (len(models) + group - 1) // group if group else 1
        if lanes > 1:
            ng = min(len(models), (ng + lanes - 1) // lanes * lanes)
        # balanced groups: the largest holds ceil(N / groups)
        assert b.group() == ((len(models) + ng - 1) // ng, ng)
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(g["wavelet"])
        b.set_seismogram(d)
        b.zero_images()
        b.migrate(sx, sz)
        for k in range(len(models)):
            I, Is = b.image(k)
            assert np.array_equal(I, refs[k][0]), k
            assert np.array_equal(Is, refs[k][1]), k
        mean, var = b.welford()
    stack = np.stack([r[0].astype(np.float64) for r in refs])
    assert np.allclose(mean, stack.mean(0), rtol=1e-13, atol=1e-13 * np.abs(stack).max())
    assert np.allclose(var, stack.var(0, ddof=1), rtol=1e-10, atol=1e-12 * stack.var(0).max())


def test_batch_group_autotune_keeps_results(gpu_available):
    """The group autotuner tries every candidate group size, keeps the smallest
    within 2% of the fastest with its best shape, and the images stay bit-identical to the oracle's;
    a new batch of the same grids starts from the tuned group."""
    say("[gpu] batch group autotune")
    cfg = common.REF_CONFIGS["asym"]
    g = golden("asym")
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    sx, sz = common.shot_positions(cfg)[0]
    d = g["seismograms"][0]
    vpads = [O.extend(op, g["models"][0] * np.float32(1 + 0.04 * k)) for k in range(9)]
    refs = [O.rtm_shot(op, sx, sz, v, g["wavelet"], d) for v in vpads[:2]]
    with R.Batch(p, len(vpads)) as b:
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(g["wavelet"])
        b.set_seismogram(d)
        rows = b.autotune_group(20)
        groups = [r["group"] for r in rows]
        assert groups[-1] == 9 and groups == sorted(set(groups)) and groups[0] >= 1
        assert all(r["us_per_sim_step"] > 0 for r in rows)
        fastest = min(r["us_per_sim_step"] for r in rows)
        best = min(r["group"] for r in rows if r["us_per_sim_step"] <= fastest * 1.02)
        gsize, ngroups = b.group()
        assert gsize == best and ngroups == (len(vpads) + best - 1) // best
        tiling = b.tiling()
        b.zero_images()
        b.migrate(sx, sz)
        for k in range(2):
            I, Is = b.image(k)
            assert np.array_equal(I, refs[k][0]) and np.array_equal(Is, refs[k][1]), k
    with R.Batch(p, len(vpads)) as b2:
        assert b2.group() == (gsize, ngroups) and b2.tiling() == tiling


def test_batch_lanes_autotune_keeps_results(gpu_available):
    """The lanes autotuner times one lane at the current group size, then two
    and three lanes; it keeps one lane while that is within 2% of the fastest
    row, else the fastest row, the images stay
    bit-identical to the oracle's, and a new batch of the same grids starts
    with the tuned lanes and group size."""
    say("[gpu] batch lanes autotune")
    cfg = common.REF_CONFIGS["asym"]
    g = golden("asym")
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    sx, sz = common.shot_positions(cfg)[0]
    d = g["seismograms"][0]
    vpads = [O.extend(op, g["models"][0] * np.float32(1 + 0.04 * k)) for k in range(9)]
    refs = [O.rtm_shot(op, sx, sz, v, g["wavelet"], d) for v in vpads[:3]]
    with R.Batch(p, len(vpads)) as b:
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(g["wavelet"])
        b.set_seismogram(d)
        rows = b.autotune_lanes(20)
        assert rows[0]["lanes"] == 1 and len(rows) >= 2 and all(r["lanes"] in (2, 3) for r in rows[1:])
        assert all(r["us_per_sim_step"] > 0 for r in rows)
        fastest = min(r["us_per_sim_step"] for r in rows)
        win = rows[0] if rows[0]["us_per_sim_step"] <= fastest * 1.02 else \
            next(r for r in rows if r["us_per_sim_step"] == fastest)
        assert b.lanes() == win["lanes"] and b.group()[0] == win["group"]
        state = (b.lanes(), b.group(), b.tiling())
        b.zero_images()
        b.migrate(sx, sz)
        for k in range(3):
            I, Is = b.image(k)
            assert np.array_equal(I, refs[k][0]) and np.array_equal(Is, refs[k][1]), k
    with R.Batch(p, len(vpads)) as b2:
        assert (b2.lanes(), b2.group(), b2.tiling()) == state


def test_refused_tiling_keeps_the_current_shape(gpu_available):
    """A shape no kernel exists for is refused before anything is released:
    the plan and the batch keep their current shape and still migrate to the
    oracle's image."""
    say("[gpu] refused tiling")
    cfg = common.REF_CONFIGS["asym"]
    g = golden("asym")
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    sx, sz = common.shot_positions(cfg)[0]
    d = g["seismograms"][0]
    vpad = O.extend(op, g["models"][0])
    ref = O.rtm_shot(op, sx, sz, vpad, g["wavelet"], d)
    with R.Plan(p) as plan, R.Batch(p, 2) as b:
        plan.set_velocity(vpad)
        plan.set_wavelet(g["wavelet"])
        plan.set_seismogram(d)
        for k in range(2):
            b.set_velocity(k, vpad)
        b.set_wavelet(g["wavelet"])
        b.set_seismogram(d)
        for obj in (plan, b):
            obj.set_tiling(2, 100, 0, 0)
            with pytest.raises(R.RtmError, match="mixed tiles"):
                obj.set_tiling(8, 100, 2, 0)
            with pytest.raises(R.RtmError, match="rows per thread"):
                obj.set_tiling(5, 100, 0, 0)
            assert obj.tiling() == (2, 100, 0, 0)
        plan.zero_image()
        plan.migrate(sx, sz)
        b.zero_images()
        b.migrate(sx, sz)
        images = [plan.image()] + [b.image(k) for k in range(2)]
    for I, Is in images:
        assert np.array_equal(I, ref[0]) and np.array_equal(Is, ref[1])


_BS_CACHE = {}


def _boundary_saving_expected(ts):
    """asym, shot 0: the oracle's forward fields and its backward pass over
    the reconstructed source field (numpy restatement of the reconstruction
    over oracle forward runs)."""
    if ts not in _BS_CACHE:
        cfg = common.REF_CONFIGS["asym"]
        g = golden("asym")
        op = common.orc_params(cfg, time_slices=ts)
        vpad = O.extend(op, g["models"][0])
        sx, sz = common.shot_positions(cfg)[0]
        _, _, lu, lU = O.forward(op, sx, sz, vpad, g["wavelet"], snaps=False, record=False)
        S_bs = common.boundary_saving_snapshots(op, sx, sz, vpad, g["wavelet"])
        I, Is, bu, bU = O.backward(op, vpad, g["seismograms"][0], S_bs)
        _BS_CACHE[ts] = (lu, lU, I, Is, bu, bU)
    return _BS_CACHE[ts]


@pytest.mark.parametrize("ts,interior", [(1, 0), (4, 1), (5, 0), (5, 1), (4, 2)])
def test_boundary_saving_bitexact(gpu_available, ts, interior):
    """Boundary-saving snapshot mode: the forward pass saves only the ring
    around the image region each step and the backward pass rebuilds the
    source field in time reverse.  Forward and receiver fields equal the
    oracle's, and the image equals the oracle's backward pass over the
    reconstructed source field (numpy restatement), bit for bit.  A second
    backward pass without a new forward pass is refused (the reconstruction
    consumed the saved state)."""
    say(f"[gpu] boundary saving TS={ts} interior={interior}")
    cfg = common.REF_CONFIGS["asym"]
    g = golden("asym")
    op = common.orc_params(cfg, time_slices=ts)
    p = common.rtm_params(cfg, time_slices=ts, snapshot_mode=R.SNAP_BOUNDARY)
    vpad = O.extend(op, g["models"][0])
    sx, sz = common.shot_positions(cfg)[0]
    lu, lU, I, Is, bu, bU = _boundary_saving_expected(ts)
    with R.Plan(p) as plan:
        plan.set_tiling(4, 100, interior, 16)
        plan.set_velocity(vpad)
        plan.set_wavelet(g["wavelet"])
        plan.forward(sx, sz)
        a, b = plan.fields()
        assert np.array_equal(a, lu) and np.array_equal(b, lU), "forward fields differ"
        plan.set_seismogram(g["seismograms"][0])
        plan.zero_image()
        plan.backward()
        a, b = plan.fields()
        assert np.array_equal(a, bu) and np.array_equal(b, bU), "receiver fields differ"
        I2, Is2 = plan.image()
        assert np.array_equal(I2, I), f"image relL2={common.rel_l2(I2, I)}"
        assert np.array_equal(Is2, Is)
        with pytest.raises(RuntimeError):
            plan.backward()


@pytest.mark.parametrize("spacing", [3, 1])
def test_boundary_saving_records_traces(gpu_available, spacing):
    """A boundary-saving forward pass records the same traces as the oracle:
    fused into the step (distinct receivers) and per step (spacing 1 puts
    every receiver on one point, the reference's spacing-1 indexing); with
    distinct receivers the image is still the reconstruction's, bit for
    bit."""
    say(f"[gpu] boundary saving recording spacing={spacing}")
    cfg = dict(common.REF_CONFIGS["asym"])
    cfg["spacingReceiver"] = spacing
    g = golden("asym")
    op = common.orc_params(cfg, time_slices=4)
    p = common.rtm_params(cfg, time_slices=4, snapshot_mode=R.SNAP_BOUNDARY)
    vpad = O.extend(op, g["models"][0])
    sx, sz = common.shot_positions(cfg)[0]
    _, d_orc, lu, lU = O.forward(op, sx, sz, vpad, g["wavelet"], snaps=False, record=True)
    with R.Plan(p) as plan:
        plan.set_tiling(4, 100, 1, 16)
        plan.set_velocity(vpad)
        plan.set_wavelet(g["wavelet"])
        plan.forward(sx, sz, record=True)
        assert np.array_equal(plan.seismogram(), d_orc), "recorded traces differ"
        a, b = plan.fields()
        assert np.array_equal(a, lu) and np.array_equal(b, lU), "forward fields differ"
        if spacing == 3:
            _, _, I, Is, _, _ = _boundary_saving_expected(4)
            plan.set_seismogram(g["seismograms"][0])
            plan.zero_image()
            plan.backward()
            I2, Is2 = plan.image()
            assert np.array_equal(I2, I) and np.array_equal(Is2, Is)


@pytest.mark.parametrize("rz,interior,seg", [(4, 1, 16), (2, 1, 0), (4, 2, 0), (2, 3, 0)])
def test_boundary_saving_with_column_interior_bitexact(gpu_available, rz, interior, seg):
    """Boundary saving on a grid wide enough for column items (a 128-column
    and a 64-column item): the column kernel images against the rebuilt
    source field as the tiles do, and the image equals the oracle's backward
    pass over the reconstructed source field, bit for bit."""
    say(f"[gpu] boundary saving, column interior rz={rz} interior={interior} seg={seg}")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=4)
    p = common.rtm_params(cfg, time_slices=4, snapshot_mode=R.SNAP_BOUNDARY)
    vpad = O.extend(op, model)
    sx, sz = common.shot_positions(cfg)[0]
    S_bs = common.boundary_saving_snapshots(op, sx, sz, vpad, wav)
    I, Is, _, _ = O.backward(op, vpad, seis, S_bs)
    with R.Plan(p) as plan:
        plan.set_tiling(rz, 100, interior, seg)
        plan.set_velocity(vpad)
        plan.set_wavelet(wav)
        plan.set_seismogram(seis)
        plan.zero_image()
        plan.migrate(sx, sz)
        I2, Is2 = plan.image()
    assert np.array_equal(I2, I), f"image relL2={common.rel_l2(I2, I)}"
    assert np.array_equal(Is2, Is)


@pytest.mark.parametrize("group", [0, 2])
def test_batch_boundary_saving_bitexact(gpu_available, group):
    """A batch in boundary-saving mode: each realization's image equals the
    oracle's backward pass over its own reconstructed source field (one
    launch group, or groups of 2 and 1)."""
    say(f"[gpu] batch boundary saving group={group}")
    cfg = common.REF_CONFIGS["asym"]
    g = golden("asym")
    op = common.orc_params(cfg, time_slices=4)
    p = common.rtm_params(cfg, time_slices=4, snapshot_mode=R.SNAP_BOUNDARY)
    sx, sz = common.shot_positions(cfg)[0]
    d = g["seismograms"][0]
    vpads = [O.extend(op, g["models"][0] * np.float32(1 + 0.05 * k)) for k in range(3)]
    with R.Batch(p, len(vpads)) as b:
        b.set_tiling(4, 100, 1, 16)
        b.set_group(group)
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(g["wavelet"])
        b.set_seismogram(d)
        b.zero_images()
        b.migrate(sx, sz)
        for k, v in enumerate(vpads):
            S_bs = common.boundary_saving_snapshots(op, sx, sz, v, g["wavelet"])
            I, Is, _, _ = O.backward(op, v, d, S_bs)
            I2, Is2 = b.image(k)
            assert np.array_equal(I2, I), (k, common.rel_l2(I2, I))
            assert np.array_equal(Is2, Is), k


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("ts,interior", [(1, 0), (4, 1), (5, 0), (4, 2)])
def test_sixteen_bit_snapshots_bitexact(gpu_available, mode, ts, interior):
    """FP16 / BF16 snapshot store with per-snapshot max-abs scale: the image
    equals the oracle's backward pass over the oracle's FP32 snapshots put
    through the store's rule (numpy restatement), bit for bit; TS = 1 covers
    the final-snapshot flush."""
    say(f"[gpu] 16-bit store mode={mode} TS={ts} interior={interior}")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=ts)
    p = common.rtm_params(cfg, time_slices=ts, snapshot_mode=mode)
    vpad = O.extend(op, model)
    sx, sz = common.shot_positions(cfg)[0]
    S, _, _, _ = O.forward(op, sx, sz, vpad, wav, record=False)
    deq, _, _ = common.quantize_snapshots(S, mode)
    I, Is, _, _ = O.backward(op, vpad, seis, deq)
    with R.Plan(p) as plan:
        plan.set_tiling(4, 100, interior, 16)
        plan.set_velocity(vpad)
        plan.set_wavelet(wav)
        plan.set_seismogram(seis)
        plan.zero_image()
        plan.migrate(sx, sz)
        I2, Is2 = plan.image()
    assert np.array_equal(I2, I) and np.array_equal(Is2, Is)


def test_batch_sixteen_bit_snapshots_bitexact(gpu_available):
    """A batch with the FP16 store: each realization's image equals the
    oracle's over its own quantized snapshots (each realization keeps its own
    scales)."""
    say("[gpu] batch FP16 store")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=4)
    p = common.rtm_params(cfg, time_slices=4, snapshot_mode=1)
    sx, sz = common.shot_positions(cfg)[0]
    vpads = [O.extend(op, model * np.float32(1 + 0.05 * k)) for k in range(3)]
    with R.Batch(p, len(vpads)) as b:
        b.set_tiling(4, 100, 1, 16)
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(wav)
        b.set_seismogram(seis)
        b.zero_images()
        b.migrate(sx, sz)
        for k, v in enumerate(vpads):
            S, _, _, _ = O.forward(op, sx, sz, v, wav, record=False)
            I, Is, _, _ = O.backward(op, v, seis, common.quantize_snapshots(S, 1)[0])
            I2, Is2 = b.image(k)
            assert np.array_equal(I2, I) and np.array_equal(Is2, Is), k


@pytest.mark.parametrize("ts,rz,interior", [(1, 2, 0), (4, 4, 0), (5, 1, 0), (4, 8, 0), (1, 2, 3), (4, 4, 3)])
def test_int16_block_snapshots_bitexact(gpu_available, ts, rz, interior):
    """int16 block store (a power-of-two scale per image row and 64-column
    block, stored in the step that computes the snapshot): the image equals
    the oracle's backward pass over the oracle's FP32 snapshots put through
    the store's rule (numpy restatement), bit for bit."""
    say(f"[gpu] int16 block store TS={ts} rz={rz} interior={interior}")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=ts)
    p = common.rtm_params(cfg, time_slices=ts, snapshot_mode=R.SNAP_I16)
    vpad = O.extend(op, model)
    sx, sz = common.shot_positions(cfg)[0]
    S, _, _, _ = O.forward(op, sx, sz, vpad, wav, record=False)
    deq, _, _ = common.quantize_i16_blocks(S, op.t_left)
    I, Is, _, _ = O.backward(op, vpad, seis, deq)
    with R.Plan(p) as plan:
        plan.set_tiling(rz, 100, interior, 0)
        plan.set_velocity(vpad)
        plan.set_wavelet(wav)
        plan.set_seismogram(seis)
        plan.zero_image()
        plan.migrate(sx, sz)
        I2, Is2 = plan.image()
    assert np.array_equal(I2, I) and np.array_equal(Is2, Is)


def test_int16_block_store_refuses_non_tile_shapes_and_tunes_tiles_only(gpu_available):
    """The int16 block store runs in tile shapes (of 4 or 8 wavefronts): a
    column or mixed interior is refused with the reason and the current
    shape kept; the autotuner times tile shapes only and keeps results
    exact."""
    say("[gpu] int16 block store: shapes")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=4)
    p = common.rtm_params(cfg, time_slices=4, snapshot_mode=R.SNAP_I16)
    vpad = O.extend(op, model)
    sx, sz = common.shot_positions(cfg)[0]
    S, _, _, _ = O.forward(op, sx, sz, vpad, wav, record=False)
    I, _, _, _ = O.backward(op, vpad, seis, common.quantize_i16_blocks(S, op.t_left)[0])
    with R.Plan(p) as plan:
        plan.set_tiling(2, 100, 0, 0)
        for interior in (1, 2):
            with pytest.raises(R.RtmError, match="tile shapes only"):
                plan.set_tiling(4, 100, interior, 16)
        assert plan.tiling()[2] == 0
        tuned = plan.autotune(20)
        assert tuned and all(t["interior"] in (0, 3) for t in tuned)   # tiles of 4 or 8 wavefronts
        plan.set_velocity(vpad)
        plan.set_wavelet(wav)
        plan.set_seismogram(seis)
        plan.zero_image()
        plan.migrate(sx, sz)
        assert np.array_equal(plan.image()[0], I)


def test_batch_int16_block_snapshots_bitexact(gpu_available):
    """A batch with the int16 block store, in launch groups of 2 (a ragged
    last group): each realization's image equals the oracle's over its own
    quantized snapshots."""
    say("[gpu] batch int16 block store")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=4)
    p = common.rtm_params(cfg, time_slices=4, snapshot_mode=R.SNAP_I16)
    sx, sz = common.shot_positions(cfg)[0]
    vpads = [O.extend(op, model * np.float32(1 + 0.05 * k)) for k in range(3)]
    with R.Batch(p, len(vpads)) as b:
        b.set_group(2)
        b.set_tiling(4, 100, 0, 0)
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(wav)
        b.set_seismogram(seis)
        b.zero_images()
        b.migrate(sx, sz)
        for k, v in enumerate(vpads):
            S, _, _, _ = O.forward(op, sx, sz, v, wav, record=False)
            I, Is, _, _ = O.backward(op, v, seis, common.quantize_i16_blocks(S, op.t_left)[0])
            I2, Is2 = b.image(k)
            assert np.array_equal(I2, I) and np.array_equal(Is2, Is), k


def test_injection_outside_stepped_region(gpu_available):
    """Source and receivers on row 2 of the padded grid (outside the stepped
    region 4 <= k < Z-4, inside the top CPML): the per-point path with the
    outer-column top pass and list injection, against the oracle."""
    say("[gpu] injection outside the stepped region")
    cfg = dict(common.REF_CONFIGS["tiny"])
    cfg["depthReceiver"] = 2 - cfg["thicknessUp"]
    cfg["firstLocationReceiver"] = -cfg["thicknessLeft"]      # first receiver at column 0
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    g = golden("tiny")
    vpad = O.extend(op, g["models"][0])
    sx, sz = 10, 2 - cfg["thicknessUp"]
    S, d, lu, lU = O.forward(op, sx, sz, vpad, g["wavelet"], record=True)
    I, Is = O.rtm_shot(op, sx, sz, vpad, g["wavelet"], d)
    with R.Plan(p) as plan:
        plan.set_velocity(vpad)
        plan.set_wavelet(g["wavelet"])
        plan.forward(sx, sz, record=True)
        a, b = plan.fields()
        assert np.array_equal(a, lu) and np.array_equal(b, lU)
        assert np.array_equal(plan.seismogram(), d)
        plan.set_seismogram(d)
        plan.zero_image()
        plan.backward()
        I2, Is2 = plan.image()
    assert np.array_equal(I2, I) and np.array_equal(Is2, Is)


@pytest.mark.parametrize("name", ["tiny", "asym"])
def test_rtm_routine_matches_reference_golden(gpu_available, name):
    """The drop-in rtm_routine on the reference's files reproduces the
    reference's own outputs bit for bit."""
    say(f"[gpu] rtm_routine {name}")
    cfg = common.REF_CONFIGS[name]
    g = golden(name)
    p = common.rtm_params(cfg)
    zs = [g[f"zfp_{s}"].tobytes() for s in range(cfg["numberOfShots"])]
    with tempfile.TemporaryDirectory() as run:
        common.write_inputs(run, cfg, list(g["models"]), g["wavelet"], zs)
        cwd = os.getcwd()
        os.chdir(run)
        try:
            R.set_params(p)
            for s in range(1, cfg["numberOfSimulations"] + 1):
                R.lib().rtm_routine(s)
        finally:
            os.chdir(cwd)
        for s in range(cfg["numberOfSimulations"]):
            b, lap = common.read_outputs(run, cfg, s + 1)
            assert np.array_equal(b, g["brute"][s]), f"sim {s + 1}: relL2={common.rel_l2(b, g['brute'][s])}"
            assert np.array_equal(lap, g["laplacian"][s])


def test_rtm_routine_illumination_normalization(gpu_available):
    """normalize_illumination = 1: rtm_routine writes the shot stack divided
    by the stacked self-correlation (the reference's full-survey line; a point
    with zero illumination would stay 0), and its Laplacian; the stacks are
    the oracle's per-shot images summed in shot order."""
    say("[gpu] rtm_routine illumination normalization")
    cfg = common.REF_CONFIGS["tiny"]
    g = golden("tiny")
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    p.normalize_illumination = 1
    zs = [g[f"zfp_{s}"].tobytes() for s in range(cfg["numberOfShots"])]
    vpad = O.extend(op, g["models"][0])
    with tempfile.TemporaryDirectory() as run:
        common.write_inputs(run, cfg, list(g["models"][:1]), g["wavelet"], zs)
        stack = np.zeros((op.nx, op.nz), np.float32)
        self_stack = np.zeros_like(stack)
        for s, (sx, sz) in enumerate(common.shot_positions(cfg)):
            seis = R.zfp_decompress_file(os.path.join(run, "INPUTS", f"seismogram_{s + 1:05d}.zfp"),
                                         cfg["numberOfReceivers"], cfg["numberOfTimeStep"], cfg["zfp_tolerance"])
            I, Is = O.rtm_shot(op, sx, sz, vpad, g["wavelet"], seis)
            stack = stack + I
            self_stack = self_stack + Is
        assert np.array_equal(stack, g["brute"][0])
        lit = self_stack != 0
        expect = np.zeros_like(stack)
        expect[lit] = stack[lit] / self_stack[lit]
        cwd = os.getcwd()
        os.chdir(run)
        try:
            R.set_params(p)
            R.lib().rtm_routine(1)
        finally:
            os.chdir(cwd)
            R.set_params(common.rtm_params(cfg))
        b, lap = common.read_outputs(run, cfg, 1)
    assert np.array_equal(b, expect)
    assert np.array_equal(lap, O.laplacian(op, expect))


def test_plan_stack_on_device(gpu_available):
    """Two shots stacked on the device: the stack is the FP32 sum of the shot
    images in shot order, the normalized stack divides by the self-image
    stack (0 where that is 0), and the Laplacian equals the host
    laplacian_filter2D of the same stack, bit for bit."""
    say("[gpu] device stack")
    cfg = common.REF_CONFIGS["tiny"]
    g = golden("tiny")
    p = common.rtm_params(cfg)
    op = common.orc_params(cfg)
    vpad = O.extend(op, g["models"][0])
    shots = common.shot_positions(cfg)
    R.set_params(p)
    with R.Plan(p) as plan:
        plan.set_velocity(vpad)
        plan.set_wavelet(g["wavelet"])
        zero, zero_lap = plan.stack()
        assert not zero.any() and not zero_lap.any()
        imgs = []
        for s, (sx, sz) in enumerate(shots):
            plan.set_seismogram(g["seismograms"][s])
            plan.zero_image()
            plan.migrate(sx, sz)
            imgs.append(plan.image())
            plan.stack_image()
        st, lp = plan.stack()
        stn, lpn = plan.stack(normalize=True)
    ref = np.zeros_like(imgs[0][0])
    ref_self = np.zeros_like(imgs[0][1])
    for I, Is in imgs:
        ref = (ref + I).astype(np.float32)
        ref_self = (ref_self + Is).astype(np.float32)
    assert np.array_equal(st, ref)
    with np.errstate(divide="ignore", invalid="ignore"):
        refn = np.where(ref_self != 0, ref / ref_self, 0).astype(np.float32)
    assert np.array_equal(stn, refn)
    for a, b in ((st, lp), (stn, lpn)):
        host = np.zeros_like(a)
        R.lib().laplacian_filter2D(R.fptr(host), R.fptr(np.ascontiguousarray(a)))
        assert np.array_equal(b, host)


@pytest.mark.parametrize("mode", ["ranks", "devices"])
def test_driver_runs_the_reference_simulations(gpu_available, mode):
    """rtmuq2d_hip, the MPI-free process driver, run on the reference's files
    with a parameter file: every simulation's outputs equal the reference's
    own, bit for bit, whether the ids run in one rank or in one child process
    per device."""
    import subprocess
    say(f"[gpu] driver {mode}")
    cfg = common.REF_CONFIGS["tiny"]
    g = golden("tiny")
    zs = [g[f"zfp_{s}"].tobytes() for s in range(cfg["numberOfShots"])]
    keys = {"Nx": "Nx", "Nz": "Nz", "thicknessLeft": "thicknessLeft", "thicknessRight": "thicknessRight",
            "thicknessUp": "thicknessUp", "thicknessDown": "thicknessDown", "spacingGrid": "spacingGrid",
            "sampling": "sampling", "numberOfTimeStep": "numberOfTimeStep", "sampligWavelet": "sampligWavelet",
            "firstLocationShot": "firstLocationShot", "depthShot": "depthShot", "spacingShot": "spacingShot",
            "numberOfShots": "numberOfShots", "firstLocationReceiver": "firstLocationReceiver",
            "depthReceiver": "depthReceiver", "spacingReceiver": "spacingReceiver",
            "numberOfReceivers": "numberOfReceivers", "numberOfSimulations": "numberOfSimulations",
            "zfp_tolerance": "zfp_tolerance"}
    lines = [f"{k} = {cfg[v]}" for k, v in keys.items()]
    lines += ["TIME_SLICES = 20", "velocity_file = ./INPUTS/vp_%05d.bin", "wavelet_file = ./INPUTS/wavelet.bin",
              "seismogram_file = ./INPUTS/seismogram_%05d.zfp", "stacked_file = ./OUTPUTS/stackedBrute_%05d.bin",
              "stacked_laplacian_file = ./OUTPUTS/stackedLaplacian_%05d.bin"]
    driver = os.path.join(common.REPO, "rtmuq2d_amd", "lib", "rtmuq2d_hip")
    with tempfile.TemporaryDirectory() as run:
        common.write_inputs(run, cfg, list(g["models"]), g["wavelet"], zs)
        with open(os.path.join(run, "params.txt"), "w") as f:
            f.write("\n".join(lines) + "\n")
        extra = ["--devices", "1"] if mode == "devices" else ["--ranks", "1", "--rank", "0"]
        r = subprocess.run([driver, "--params", "params.txt", *extra], cwd=run, capture_output=True, text=True,
                           timeout=100)
        assert r.returncode == 0, r.stderr[-2000:]
        assert "took" in r.stdout
        for s in range(cfg["numberOfSimulations"]):
            b, lap = common.read_outputs(run, cfg, s + 1)
            assert np.array_equal(b, g["brute"][s]), s
            assert np.array_equal(lap, g["laplacian"][s]), s


def test_legacy_modeling_symbols(gpu_available):
    """isotropicAcousticModeling/adjointModeling keep the reference's calling
    convention (images accumulated with +=) and match the oracle."""
    say("[gpu] legacy modeling symbols")
    cfg = common.REF_CONFIGS["tiny"]
    g = golden("tiny")
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    R.set_params(p)
    vpad = O.extend(op, g["models"][1])
    sx, sz = common.shot_positions(cfg)[1]
    d = g["seismograms"][1]
    I0 = np.random.default_rng(3).standard_normal((cfg["Nx"], cfg["Nz"])).astype(np.float32)
    Is0 = np.abs(I0)
    I_orc, Is_orc = O.rtm_shot(op, sx, sz, vpad, g["wavelet"], d)
    shot = np.array([sx, sz], np.int32)
    wav = np.ascontiguousarray(g["wavelet"], np.float32)
    R.lib().isotropicAcousticModeling(5, R.iptr(shot), R.fptr(vpad), R.fptr(wav))
    I, Is = I0.copy(), Is0.copy()
    rc = R.lib().adjointModeling(5, R.iptr(shot), R.fptr(vpad), R.fptr(wav),
                                 R.fptr(np.ascontiguousarray(d)), R.fptr(I), R.fptr(Is))
    assert rc == 1
    assert np.array_equal(I, I0 + I_orc)
    assert np.array_equal(Is, Is0 + Is_orc)


def _rand_fields(cfg, seed):
    rng = np.random.default_rng(seed)
    X = cfg["thicknessLeft"] + cfg["Nx"] + cfg["thicknessRight"]
    Z = cfg["thicknessUp"] + cfg["Nz"] + cfg["thicknessDown"]
    f = lambda: rng.standard_normal((X, Z)).astype(np.float32)  # noqa: E731
    return X, Z, f


@pytest.mark.parametrize("name", ["tiny", "asym"])
def test_legacy_step_symbols(gpu_available, name):
    """iso_acoustic_wave_equation, aiso_attenuation_cpml and the imaging
    conditions, host-pointer form, against the oracle on random state."""
    say(f"[gpu] legacy per-step symbols {name}")
    cfg = common.REF_CONFIGS[name]
    op = common.orc_params(cfg)
    R.set_params(common.rtm_params(cfg))
    X, Z, f = _rand_fields(cfg, 11)
    Cf = (np.abs(f()) * 0.1).astype(np.float32)
    u, U = f(), f()
    ref = O.stencil(op, Cf, u, U)
    mine = U.copy()
    R.lib().iso_acoustic_wave_equation(R.fptr(Cf), R.fptr(u), R.fptr(mine))
    assert np.array_equal(mine, ref)

    prof = O.profiles(op)
    TL, TR, TU, TD = cfg["thicknessLeft"], cfg["thicknessRight"], cfg["thicknessUp"], cfg["thicknessDown"]
    rng = np.random.default_rng(5)
    shapes = [(TL, Z), (TR, Z), (X, TU), (X, TD)]
    psi = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    csi = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    psi_o, csi_o = [a.copy() for a in psi], [a.copy() for a in csi]
    Uref = O.cpml(op, u, U, Cf, prof, psi_o, csi_o)
    Umine = U.copy()
    aL, aR, aT, aB, bL, bR, bT, bB = prof
    R.lib().aiso_attenuation_cpml(R.fptr(u), R.fptr(Umine), R.fptr(Cf), R.fptr(aL), R.fptr(bL),
                                  R.fptr(aR), R.fptr(bR), R.fptr(aT), R.fptr(bT), R.fptr(aB), R.fptr(bB),
                                  *[R.fptr(a) for a in psi], *[R.fptr(a) for a in csi])
    assert np.array_equal(Umine, Uref)
    for a, b in zip(psi, psi_o):
        assert np.array_equal(a, b)
    for a, b in zip(csi, csi_o):
        assert np.array_equal(a, b)

    S, Rw = f(), f()
    I0 = rng.standard_normal((cfg["Nx"], cfg["Nz"])).astype(np.float32)
    I_ref, Is_ref = I0.copy(), I0.copy()
    O.image(op, I_ref, Is_ref, S, Rw)
    I_m, Is_m = I0.copy(), I0.copy()
    R.lib().normalized_cross_correlation(R.fptr(I_m), R.fptr(Is_m), R.fptr(S), R.fptr(Rw))
    assert np.array_equal(I_m, I_ref) and np.array_equal(Is_m, Is_ref)
    I_m = I0.copy()
    R.lib().cross_correlation(R.fptr(I_m), R.fptr(S), R.fptr(Rw))
    assert np.array_equal(I_m, I_ref)


def test_source_receiver_delta(gpu_available):
    """One forward step from rest puts w[0]*C at the source and nothing else;
    one backward step puts d[r, nt-1] at each receiver, bit for bit."""
    say("[gpu] delta indexing")
    cfg = dict(common.REF_CONFIGS["asym"])
    cfg["numberOfTimeStep"] = 1
    p = common.rtm_params(cfg, time_slices=1)
    op = common.orc_params(cfg, time_slices=1)
    rng = np.random.default_rng(2)
    vpad = O.extend(op, rng.uniform(1500, 3000, (cfg["Nx"], cfg["Nz"])).astype(np.float32))
    Cf = O.coef(op, vpad)
    wav = rng.standard_normal(cfg["sampligWavelet"]).astype(np.float32)
    sx, sz = 17, 9
    with R.Plan(p) as plan:
        plan.set_velocity(vpad)
        plan.set_wavelet(wav)
        plan.forward(sx, sz)
        _, U = plan.fields()
        xs, zs = sx + cfg["thicknessLeft"], sz + cfg["thicknessUp"]
        expect = np.zeros_like(U)
        expect[xs, zs] = np.float32(0) + wav[0] * Cf[xs, zs]
        assert np.array_equal(U, expect)
        d = rng.standard_normal((cfg["numberOfReceivers"], 1)).astype(np.float32)
        plan.set_seismogram(d)
        plan.zero_image()
        plan.backward()
        _, U = plan.fields()
        expect = np.zeros_like(U)
        zr = cfg["thicknessUp"] + cfg["depthReceiver"]
        for r in range(cfg["numberOfReceivers"]):
            expect[cfg["thicknessLeft"] + cfg["firstLocationReceiver"] + (cfg["spacingReceiver"] - 1) * r, zr] = d[r, 0]
        assert np.array_equal(U, expect)


def _c0_inputs():
    cfg = common.REF_CONFIGS["c0"]
    v_true, models, wav, seis = common.make_inputs(cfg)
    return cfg, v_true, models, wav, seis


def test_c0_image_matches_reference(gpu_available):
    """Config c0 (600x200, CPML 40, nt=2000, TS=20): the GPU FP32 image is
    bit-identical to the reference's (strict build, checksums committed) and
    within relL2 <= 1e-5 of the reference built -O3 -march=native."""
    say("[gpu] c0 end-to-end")
    meta = json.load(open(os.path.join(common.GOLDEN, "c0.json")))
    native = np.load(os.path.join(common.GOLDEN, "c0_native.npz"))["brute"]
    cfg, v_true, models, wav, seis = _c0_inputs()
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    assert sha(models[0]) == meta["models"][0], "input generator drifted"
    zs = R.zfp_compress(seis[0], cfg["zfp_tolerance"])
    assert hashlib.sha256(zs).hexdigest() == meta["zfp"][0], "seismogram or ZFP stream drifted"
    p = common.rtm_params(cfg)
    with tempfile.TemporaryDirectory() as run:
        common.write_inputs(run, cfg, models[:1], wav, [zs])
        cwd = os.getcwd()
        os.chdir(run)
        try:
            R.set_params(p)
            R.lib().rtm_routine(1)
        finally:
            os.chdir(cwd)
        b, lap = common.read_outputs(run, cfg, 1)
    assert sha(b) == meta["stackedBrute"][0], f"relL2 vs native {common.rel_l2(b, native)}"
    assert sha(lap) == meta["stackedLaplacian"][0]
    assert common.rel_l2(b, native) <= 1e-5


def test_c1_full_size_image_matches_reference(gpu_available):
    """One full c1-size image (1701x351, CPML 50, nt=4000, TS=20, 1701
    receivers) through the drop-in rtm_routine: bit-identical to the
    reference's strict build (checksums committed), hence within the
    recorded relL2 of its -march=native build, which must be <= 1e-5."""
    say("[gpu] c1-size end-to-end")
    meta = json.load(open(os.path.join(common.GOLDEN, "c1s.json")))
    cfg = common.REF_CONFIGS["c1s"]
    v_true, models, wav, seis = common.make_inputs(cfg)
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    assert sha(models[0]) == meta["models"][0], "input generator drifted"
    zs = R.zfp_compress(seis[0], cfg["zfp_tolerance"])
    assert hashlib.sha256(zs).hexdigest() == meta["zfp"][0], "seismogram or ZFP stream drifted"
    p = common.rtm_params(cfg)
    with tempfile.TemporaryDirectory() as run:
        common.write_inputs(run, cfg, models, wav, [zs])
        cwd = os.getcwd()
        os.chdir(run)
        try:
            R.set_params(p)
            R.lib().rtm_routine(1)
        finally:
            os.chdir(cwd)
        b, lap = common.read_outputs(run, cfg, 1)
    assert sha(b) == meta["stackedBrute"][0], f"norm {np.linalg.norm(b)} vs {meta['brute_norm'][0]}"
    assert sha(lap) == meta["stackedLaplacian"][0]
    assert meta["strict_vs_native_relL2"] <= 1e-5


def _server_requests():
    """Requests on three grids (tiny, asym, wide), each with its own velocity,
    shot position and seismogram, FP16 snapshot store every 4 steps."""
    from rtmuq2d_amd import synth
    rng = np.random.default_rng(77)
    out = []
    for k, name in enumerate(["tiny", "asym", "wide", "tiny", "asym"]):
        if name == "wide":
            cfg, model, wav, seis = _wide_case()
        else:
            cfg = common.REF_CONFIGS[name]
            g = golden(name)
            model = (g["models"][0] * np.float32(1 + 0.04 * k)).astype(np.float32)
            wav = g["wavelet"]
            seis = rng.standard_normal((cfg["numberOfReceivers"], cfg["numberOfTimeStep"])).astype(np.float32)
        op = common.orc_params(cfg, time_slices=4)
        p = common.rtm_params(cfg, time_slices=4, snapshot_mode=1)
        sx = int(rng.integers(6, cfg["Nx"] - 6))
        sz = common.shot_positions(cfg)[0][1]
        out.append((cfg, op, p, O.extend(op, model), wav, seis, sx, sz))
    return out


@pytest.mark.parametrize("shape,group_bytes", [((4, 50, 1, 16), None), ((4, 50, 2, 0), None),
                                               ((2, 30, 0, 0), None), ((4, 50, 0, 0), 1), ((2, 50, 3, 0), None),
                                               ((2, 50, 3, 0), 1)])
def test_server_each_result_same_alone_and_in_the_mix(gpu_available, shape, group_bytes):
    """Continuous batching: requests of three grid sizes arrive while others
    are in flight, are admitted by HBM footprint (the arena holds about two at
    a time), share every launch, and each image equals the oracle's for that
    request alone (FP16 store rule applied), bit for bit; so it is also
    independent of the mix.  Column interior, mixed tiles and plain tiles;
    and a launch-group budget so small that each request is its own group."""
    say(f"[gpu] request server: alone vs in the mix, shape {shape} group_bytes {group_bytes}")
    reqs = _server_requests()
    refs = []
    for cfg, op, p, vpad, wav, seis, sx, sz in reqs:
        S, _, _, _ = O.forward(op, sx, sz, vpad, wav, record=False)
        refs.append(O.backward(op, vpad, seis, common.quantize_snapshots(S, 1)[0])[0])
    fp = [R.request_footprint(r[2]) for r in reqs]
    budget = int(sorted(fp)[-1] + sorted(fp)[-2] + (1 << 20))
    got = {}
    with R.Server(budget, max_slots=4, snapshot_mode=1) as srv:
        srv.set_tiling(*shape)
        if group_bytes is not None:
            srv.set_group_bytes(group_bytes)
        ids = []
        ids.append(srv.submit(*_sub(reqs[0])))
        ids.append(srv.submit(*_sub(reqs[1])))
        srv.tick()
        ids.append(srv.submit(*_sub(reqs[2])))          # arrives while two are in flight
        srv.tick()
        st = srv.stats()
        assert st["active"] <= 2 and st["bytes_in_use"] <= budget
        most_groups = srv.group_bytes()[1]
        ids.append(srv.submit(*_sub(reqs[3])))
        ids.append(srv.submit(*_sub(reqs[4])))
        while srv.tick() > 0:
            most_groups = max(most_groups, srv.group_bytes()[1])
            for rid, I in srv.done():
                got[rid] = I
        for rid, I in srv.done():
            got[rid] = I
    # a group per request (budget 1; the arena holds two), or at most one
    # group per pass (the default budget: a group holds one pass)
    assert most_groups == 2 if group_bytes == 1 else most_groups in (1, 2)
    assert sorted(got) == sorted(ids)
    for k, rid in enumerate(ids):
        assert np.array_equal(got[rid], refs[k]), k


def _sub(req):
    cfg, op, p, vpad, wav, seis, sx, sz = req
    return p, vpad, wav, seis, sx, sz


@pytest.mark.parametrize("group_bytes,lanes", [(None, 1), (1, 1), (1, 2), (1, 3)])
@pytest.mark.parametrize("block", [50, 100])
def test_server_int16_block_store_same_alone_and_in_the_mix(gpu_available, group_bytes, lanes, block):
    """The request server with the int16 block store: requests of three grid
    sizes arriving while others are in flight, each image equal to the
    oracle's for that request alone under the store's rule, in one launch
    group and with every request in a group of its own.  With blocks of 100
    steps (a multiple of the snapshot interval 4) the steps at which no
    member of a launch group stores or images run the light tile instance;
    with 50 every step runs the full one.  With each request in a group of
    its own, 2 or 3 lanes step the tick's groups at once."""
    say(f"[gpu] request server, int16 block store, group_bytes {group_bytes} lanes {lanes} block {block}")
    reqs = _server_requests()
    refs = []
    for cfg, op, p, vpad, wav, seis, sx, sz in reqs:
        S, _, _, _ = O.forward(op, sx, sz, vpad, wav, record=False)
        refs.append(O.backward(op, vpad, seis, common.quantize_i16_blocks(S, op.t_left)[0])[0])
    fp = [R.request_footprint(r[2]) for r in reqs]
    budget = int(sorted(fp)[-1] + sorted(fp)[-2] + (1 << 20))
    got = {}
    with R.Server(budget, max_slots=4, snapshot_mode=R.SNAP_I16) as srv:
        srv.set_tiling(4, block, 0, 0)
        srv.set_lanes(lanes)
        assert srv.lanes() == lanes
        with pytest.raises(R.RtmError, match="tile shapes only"):
            srv.set_tiling(4, 50, 1, 16)
        if group_bytes is not None:
            srv.set_group_bytes(group_bytes)
        ids = [srv.submit(*_sub(reqs[0])), srv.submit(*_sub(reqs[1]))]
        srv.tick()
        ids += [srv.submit(*_sub(r)) for r in reqs[2:]]
        while srv.tick() > 0:
            for rid, I in srv.done():
                got[rid] = I
        for rid, I in srv.done():
            got[rid] = I
    assert sorted(got) == sorted(ids)
    for k, rid in enumerate(ids):
        assert np.array_equal(got[rid], refs[k]), k


@pytest.mark.parametrize("group_bytes,block,lanes", [(1, 100, 1), (None, 100, 1), (None, 30, 1), (1, 100, 2)])
def test_server_fp32_store_matches_the_oracle(gpu_available, group_bytes, block, lanes):
    """A server with the FP32 snapshot store (the reference's scheme): every
    image equals the oracle's for that request alone, bit for bit, with all
    requests admitted at once, each in a launch group of its own or all in
    one (forward and backward members in the same launches).  Blocks of 100
    steps run the light tile instance where no member stores or images."""
    say(f"[gpu] request server, FP32 store, group_bytes {group_bytes} block {block}")
    reqs = _server_requests()[:3]
    refs = []
    for cfg, op, p, vpad, wav, seis, sx, sz in reqs:
        S, _, _, _ = O.forward(op, sx, sz, vpad, wav, record=False)
        refs.append(O.backward(op, vpad, seis, S)[0])
    got = {}
    with R.Server(sum(R.request_footprint(r[2].copy().update(snapshot_mode=0)) for r in reqs) + (8 << 20),
                  max_slots=4, snapshot_mode=0) as srv:
        srv.set_tiling(4, block, 0, 0)
        srv.set_lanes(lanes)
        if group_bytes is not None:
            srv.set_group_bytes(group_bytes)
        ids = [srv.submit(*_sub(r)) for r in reqs]
        while srv.tick() > 0:
            for rid, I in srv.done():
                got[rid] = I
        for rid, I in srv.done():
            got[rid] = I
    assert sorted(got) == sorted(ids)
    for k, rid in enumerate(ids):
        assert np.array_equal(got[rid], refs[k]), k


def test_server_rejects_requests_it_can_never_run(gpu_available):
    """Requests larger than the arena, with the source outside the stepped
    region, or with another snapshot store are refused at submission."""
    cfg, op, p, vpad, wav, seis, sx, sz = _server_requests()[2]
    assert R.request_footprint(p) > 2 << 20
    with R.Server(R.request_footprint(p) // 2, max_slots=2, snapshot_mode=1) as srv:
        with pytest.raises(R.RtmError, match="arena"):
            srv.submit(p, vpad, wav, seis, sx, sz)
    with R.Server(4 * R.request_footprint(p), max_slots=2, snapshot_mode=1) as srv:
        with pytest.raises(R.RtmError, match="outside"):
            srv.submit(p, vpad, wav, seis, sx, -cfg["thicknessUp"])
        srv.snapshot_mode = 0
        with pytest.raises(R.RtmError, match="snapshot mode"):
            srv.submit(p, vpad, wav, seis, sx, sz)


def test_autotune_keeps_results_and_caches_per_grid_class(gpu_available):
    """The autotuner times every candidate shape, keeps the fastest (the
    image stays bit-identical to the oracle's), and a later plan of the same
    grid class starts with that shape."""
    say("[gpu] autotune")
    cfg, model, wav, seis = _wide_case()
    op = common.orc_params(cfg, time_slices=10)
    p = common.rtm_params(cfg, time_slices=10)
    vpad = O.extend(op, model)
    sx, sz = common.shot_positions(cfg)[0]
    I_orc, Is_orc = O.rtm_shot(op, sx, sz, vpad, wav, seis)
    with R.Plan(p) as plan:
        plan.set_velocity(vpad)
        res = plan.autotune(50)
        assert len(res) >= 4 and all(r["ms_per_step"] > 0 for r in res)
        best = min(res, key=lambda r: r["ms_per_step"])
        chosen = plan.tiling()
        assert chosen == (best["rows_per_thread"], best["graph_block"], best["interior"], best["seg_rows"])
        plan.set_wavelet(wav)
        plan.set_seismogram(seis)
        plan.zero_image()
        plan.migrate(sx, sz)
        I, Is = plan.image()
    assert np.array_equal(I, I_orc) and np.array_equal(Is, Is_orc)
    with R.Plan(p) as plan2:
        assert plan2.tiling() == chosen
