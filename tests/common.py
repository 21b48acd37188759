"""Shared helpers for the test suite: configurations, seeded inputs in the
reference's file formats, and runners for the reference build in oracle/_ref.
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle as O  # noqa: E402  (test infrastructure)
import build as oracle_build  # noqa: E402
from rtmuq2d_amd import synth  # noqa: E402

GOLDEN = os.path.join(REPO, "tests", "golden")
REF_CONFIGS = oracle_build.REF_CONFIGS


def ref_dir(name: str, flavour: str = "strict") -> str:
    return os.path.join(oracle_build.REF_OUT, f"{name}_{flavour}")


def ref_built(name: str, flavour: str = "strict") -> bool:
    return os.path.exists(os.path.join(ref_dir(name, flavour), "librtmref.so"))


def orc_params(cfg: dict, time_slices: int = 20) -> O.OrcParams:
    return O.OrcParams(cfg["Nx"], cfg["Nz"], cfg["thicknessLeft"], cfg["thicknessRight"],
                       cfg["thicknessUp"], cfg["thicknessDown"], cfg["spacingGrid"], cfg["sampling"],
                       cfg["numberOfTimeStep"], cfg["sampligWavelet"], cfg["firstLocationReceiver"],
                       cfg["depthReceiver"], cfg["spacingReceiver"], cfg["numberOfReceivers"],
                       time_slices)


def rtm_params(cfg: dict, time_slices: int = 20, **extra):
    """rtmuq2d_amd.RtmParams for a reference configuration dict."""
    import rtmuq2d_amd as R
    p = R.make_params(
        nx=cfg["Nx"], nz=cfg["Nz"], thickness_left=cfg["thicknessLeft"],
        thickness_right=cfg["thicknessRight"], thickness_up=cfg["thicknessUp"],
        thickness_down=cfg["thicknessDown"], spacing_grid=cfg["spacingGrid"],
        sampling=cfg["sampling"], number_of_time_steps=cfg["numberOfTimeStep"],
        sampling_wavelet=cfg["sampligWavelet"], first_location_shot=cfg["firstLocationShot"],
        depth_shot=cfg["depthShot"], spacing_shot=cfg["spacingShot"],
        number_of_shots=cfg["numberOfShots"], first_location_receiver=cfg["firstLocationReceiver"],
        depth_receiver=cfg["depthReceiver"], spacing_receiver=cfg["spacingReceiver"],
        number_of_receivers=cfg["numberOfReceivers"],
        number_of_simulations=cfg["numberOfSimulations"], zfp_tolerance=cfg["zfp_tolerance"],
        time_slices=time_slices,
        velocity_file="./INPUTS/vp_%05d.bin", wavelet_file="./INPUTS/wavelet.bin",
        seismogram_file="./INPUTS/seismogram_%05d.zfp",
        stacked_file="./OUTPUTS/stackedBrute_%05d.bin",
        stacked_laplacian_file="./OUTPUTS/stackedLaplacian_%05d.bin")
    return p.update(**extra)


def shot_positions(cfg: dict):
    return [(cfg["firstLocationShot"] + s * cfg["spacingShot"], cfg["depthShot"])
            for s in range(cfg["numberOfShots"])]


def make_inputs(cfg: dict, seed: int = 7):
    """Seeded inputs of one configuration:
    true model, per-simulation migration models, wavelet, and per-shot
    seismograms forward-modelled (oracle, FP32) on the true model."""
    nx, nz = cfg["Nx"], cfg["Nz"]
    dt, nt, nw = cfg["sampling"], cfg["numberOfTimeStep"], cfg["sampligWavelet"]
    f0 = 15.0 if dt >= 0.001 else 20.0
    v_true = synth.layered_anomaly(nx, nz, seed, sigma=max(2.0, min(15.0, nx / 8)))
    sm = max(1.0, min(10.0, nx / 12))
    models = [synth.migration_velocity(v_true, seed + s, sigma_smooth=sm,
                                       sigma_field=max(2.0, min(20.0, nx / 6)))
              for s in range(1, cfg["numberOfSimulations"] + 1)]
    wav = synth.ricker(f0, dt, nw)
    op = orc_params(cfg)
    vpad = O.extend(op, v_true)
    seis = []
    for sx, sz in shot_positions(cfg):
        _, d, _, _ = O.forward(op, sx, sz, vpad, wav, snaps=False, record=True)
        seis.append(d)
    return v_true, models, wav, seis


def write_inputs(run: str, cfg: dict, models, wav, seis_bytes):
    os.makedirs(os.path.join(run, "INPUTS"), exist_ok=True)
    os.makedirs(os.path.join(run, "OUTPUTS"), exist_ok=True)
    os.makedirs(os.path.join(run, "TEMPORARY_FILES"), exist_ok=True)
    for s, m in enumerate(models, start=1):
        np.ascontiguousarray(m, np.float32).tofile(os.path.join(run, "INPUTS", f"vp_{s:05d}.bin"))
    np.ascontiguousarray(wav, np.float32).tofile(os.path.join(run, "INPUTS", "wavelet.bin"))
    for s, b in enumerate(seis_bytes, start=1):
        with open(os.path.join(run, "INPUTS", f"seismogram_{s:05d}.zfp"), "wb") as f:
            f.write(b)


def read_outputs(run: str, cfg: dict, sim: int):
    nx, nz = cfg["Nx"], cfg["Nz"]
    b = np.fromfile(os.path.join(run, "OUTPUTS", f"stackedBrute_{sim:05d}.bin"), np.float32)
    lap = np.fromfile(os.path.join(run, "OUTPUTS", f"stackedLaplacian_{sim:05d}.bin"), np.float32)
    return b.reshape(nx, nz), lap.reshape(nx, nz)


def run_reference(run: str, name: str, flavour: str = "strict", first: int = 1, last: int = 1,
                  threads: int | None = None) -> float:
    """Run the reference's rtm_routine for sims [first, last] in `run`;
    returns its own wall-clock seconds."""
    env = dict(os.environ)
    env["OMP_SCHEDULE"] = "static"
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    exe = os.path.join(ref_dir(name, flavour), "ref_main")
    r = subprocess.run([exe, str(first), str(last)], cwd=run, env=env, capture_output=True, text=True,
                       timeout=1800)
    assert r.returncode == 0, r.stdout + r.stderr
    for line in r.stdout.splitlines():
        if line.startswith("REF_SECONDS"):
            return float(line.split()[1])
    raise AssertionError("no REF_SECONDS line: " + r.stdout)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> nearest BF16 value (ties to even), computed in float64 from
    the definition: 8 significant bits, or multiples of 2^-133 below 2^-126."""
    x = np.asarray(x, np.float64)
    _, e = np.frexp(x)                      # x = m * 2^e, 0.5 <= |m| < 1
    quantum = np.exp2(np.maximum(e - 8, -133))   # 8 significant bits; BF16's subnormal grid 2^-133
    return (np.round(x / quantum) * quantum).astype(np.float32)   # numpy rounds half to even


def quantize_snapshots(S, mode: int):
    """The 16-bit snapshot store's rule, restated independently: snapshot c
    is multiplied by 2^(14-e), e the binary exponent of max |S_c| (clamped to
    [-100, 100]), rounded to FP16 (numpy's IEEE conversion, nearest-even) or
    BF16 (bf16_round), and read back times 2^(e-14).  Returns the read-back
    float32 snapshots, the 16-bit codes and the max-abs float bits."""
    S = np.asarray(S, np.float32)
    deq = np.empty_like(S)
    codes = np.empty(S.shape, np.uint16)
    mx = np.empty(len(S), np.uint32)
    for c in range(len(S)):
        m = np.float32(np.max(np.abs(S[c])))
        bits = int(np.array([m], np.float32).view(np.uint32)[0])
        e = min(max(((bits >> 23) & 0xFF) - 127, -100), 100)
        x = (S[c] * np.float32(2.0 ** (14 - e))).astype(np.float32)
        if mode == 1:
            h = x.astype(np.float16)
            codes[c] = h.view(np.uint16)
            back = h.astype(np.float32)
        else:
            back = bf16_round(x)
            codes[c] = (back.view(np.uint32) >> 16).astype(np.uint16)
        deq[c] = back * np.float32(2.0 ** (e - 14))
        mx[c] = bits
    return deq, codes, mx


def quantize_i16_blocks(S, TL: int):
    """The int16 block store's rule, restated independently: in every image
    row (z) of snapshot c, the image columns are grouped by the block
    (TL + ii - 4) // 64 of their padded column; a block with max |S| =
    f 2^E (0.5 <= f < 1, E from the float32 bits, clamped to [-100, 100])
    stores rint(S 2^(15-E)) (nearest even, clipped to +-32767) as int16 and
    E, and reads back code 2^(E-15).  S is the oracle's [c][nx][nz]; returns
    the read-back float32 snapshots [c][nx][nz], the codes [c][nz][nx]
    (uint16 bits) and the exponents [c][nz][blocks] (int8)."""
    S = np.asarray(S, np.float32)
    nc, nx, nz = S.shape
    blk = (TL + np.arange(nx) - 4) // 64
    b0 = blk[0]
    nb = blk[-1] - b0 + 1
    deq = np.empty_like(S)
    codes = np.empty((nc, nz, nx), np.int16)
    exps = np.empty((nc, nz, nb), np.int8)
    for b in range(nb):
        cols = np.nonzero(blk == b0 + b)[0]
        seg = S[:, cols, :]                                       # [c][cols][nz]
        m = np.abs(seg).max(axis=1).astype(np.float32)            # [c][nz]
        e = ((m.view(np.uint32) >> 23) & 0xFF).astype(np.int32) - 126
        e = np.clip(e, -100, 100)
        scale = np.exp2(15 - e).astype(np.float32)[:, None, :]
        q = np.clip(np.rint((seg * scale).astype(np.float32)), -32767, 32767).astype(np.int16)
        deq[:, cols, :] = (q.astype(np.float32) * np.exp2(e - 15).astype(np.float32)[:, None, :])
        codes[:, :, cols] = q.transpose(0, 2, 1)
        exps[:, :, b] = e
    return deq, codes.view(np.uint16), exps


# ---- boundary saving: the reconstructed source field, from the oracle ----

_LAP_W = np.float32([-2.97399944, 1.70507669, -0.25861812, 0.04577745, -0.00523630])


def ring_points(op):
    """(i, k) of the 4-wide ring around the image region in its saved order:
    top band (4 rows of nx+8 points, row by row), bottom band, then the left
    and right bands (4 columns of nz points, column by column)."""
    W = op.nx + 8
    xs = np.arange(W) + op.t_left - 4
    zs = np.arange(op.nz) + op.t_up
    ii, kk = [], []
    for r in range(4):
        ii.append(xs); kk.append(np.full(W, op.t_up - 4 + r))
    for r in range(4):
        ii.append(xs); kk.append(np.full(W, op.t_up + op.nz + r))
    for c in range(4):
        ii.append(np.full(op.nz, op.t_left - 4 + c)); kk.append(zs)
    for c in range(4):
        ii.append(np.full(op.nz, op.t_left + op.nx + c)); kk.append(zs)
    return np.concatenate(ii), np.concatenate(kk)


def stencil_old(C, old, cur, i0, i1, k0, k1):
    """(C lap(cur) + 2 cur) - old on [i0,i1) x [k0,k1) of [X][Z] float32
    fields, each operation rounded to float32 in the order of stencil_val
    (bracket operands x-, x+, z-, z+)."""
    f = np.float32
    c = cur[i0:i1, k0:k1]
    s = []
    for m in range(1, 5):
        s.append(((cur[i0 - m:i1 - m, k0:k1] + cur[i0 + m:i1 + m, k0:k1]) + cur[i0:i1, k0 - m:k1 - m])
                 + cur[i0:i1, k0 + m:k1 + m])
    lap = ((((f(2.0) * _LAP_W[0]) * c + _LAP_W[1] * s[0]) + _LAP_W[2] * s[1]) + _LAP_W[3] * s[2]) \
        + _LAP_W[4] * s[3]
    return (C[i0:i1, k0:k1] * lap + f(2.0) * c) - old[i0:i1, k0:k1]


def boundary_saving_snapshots(op, sx, sz, vpad, wav, plant: str = ""):
    """The source field the boundary-saving backward pass images against,
    as snapshots [nt/TS][nx][nz] of u^{tau+1} (the layout of the oracle's
    stored snapshots).  The ring saves U^{n+1} come from oracle forward runs
    of n+1 steps; the last two fields from the full run; inside the image
    region u^{m-1} = (C lap(u^m) + 2 u^m) - u^{m+1} + w[m] C at the source
    (forward step m solved for its old value), on the ring the save of
    forward step m-2.  ``plant`` puts a mistake in for the tests that show
    the check catches it: "no_source" drops the source term, "ring_late"
    restores the ring from the save one step later."""
    nt, TS = op.nt, op.time_slices
    ri, rk = ring_points(op)
    saves = []
    for n in range(max(nt - 2, 0)):
        q = O.OrcParams.from_buffer_copy(op)
        q.nt, q.time_slices = n + 1, 1    # no snapshots are taken; nt must be a multiple of TS
        _, _, _, lU = O.forward(q, sx, sz, vpad, wav, snaps=False, record=False)
        saves.append(lU[ri, rk].copy())
    _, _, lu, lU = O.forward(op, sx, sz, vpad, wav, snaps=False, record=False)
    C = O.coef(op, vpad, np.float32)
    i0, i1, k0, k1 = op.t_left, op.t_left + op.nx, op.t_up, op.t_up + op.nz
    si, sk = sx + op.t_left, sz + op.t_up
    fields = {nt: lU, nt - 1: lu}
    snaps = np.zeros((nt // TS, op.nx, op.nz), np.float32)
    for n in range(nt):
        m = nt - n
        fields.pop(m + 2, None)
        if (m - 1) % TS == 0:
            snaps[(m - 1) // TS] = fields[m][i0:i1, k0:k1]
        if 1 <= n <= nt - 2:
            new = fields[m + 1].copy()
            new[i0:i1, k0:k1] = stencil_old(C, fields[m + 1], fields[m], i0, i1, k0, k1)
            if m < op.n_wavelet and i0 <= si < i1 and k0 <= sk < k1 and plant != "no_source":
                new[si, sk] = np.float32(new[si, sk] + np.float32(wav[m]) * C[si, sk])
            new[ri, rk] = saves[min(m - 1, len(saves) - 1)] if plant == "ring_late" else saves[m - 2]
            fields[m - 1] = new
    return snaps
