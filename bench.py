"""Benchmark: RTM images/hour and G point-updates/s on one MI355X.

Workload (default, "c1"): plain single-simulation RTM, 1701x351 model,
h = 10 m, CPML 50 points on every side (padded 1801x451), dt = 0.8 ms,
nt = 4000, Ricker 20 Hz at the surface centre (850, 2), 1701 receivers at
iz = 2, FP32 stepping, the source wavefield stored every step (TS = 1) in
FP32 on HBM (9.55 GB).  The velocity is a seeded 12-layer folded and
faulted model with a 450 m water layer; the seismogram is forward-modelled
on it (FP32, receiver extraction on the GPU) and passed through ZFP 0.5.5
fixed-accuracy at tol 1e-4 on the host; the migration velocity is the true
model smoothed (sigma 10 points) times (1 + 0.02 g), g a seeded Gaussian
random field.

One step = one full image: forward pass (nt steps, snapshot store) then
backward pass (nt steps, receiver replacement, imaging), all inputs resident
in HBM.  images/hour = 3600 / seconds per step;
G updates/s = 2 nt X Z / seconds per step.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import contextlib
import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "RTM images/hour + Gpt-updates/s, batched vs plain, 1 GPU, image relL2<=1e-4"
HBM_PEAK_GBS = 8000.0
RELL2_GATE = 1e-4      # every image within this relative L2 of the FP64 run of the same scheme
BYTES_PER_UPDATE = 16  # read u^n, u^{n-1}, coef; write u^{n+1} (FP32)

# L2-to-fabric bytes per launch of the c1 plain step kernel (tiles 64x8),
# mean of a forward and a backward step at TS=1, from rocprofv3 passes over
# `bench.py --profile` counting the read requests by size (32, 64 and
# 128 B) and the write requests (32 and 64 B); a 1 GiB copy in the same run
# reads and writes 1 GiB on these counts.
C1_TRAFFIC_BYTES = 26377888


def full_step_bytes(p, ts):
    """Algorithmic bytes of one step, mean of forward and backward: 16 B per
    padded point, +16 B per active CPML point (psi, csi read and written),
    +4 B per interior point stored (forward) or +20 B per interior point
    imaged (snapshot read, I and Is read-modify-write), every ts steps, and
    +4 B per receiver trace sample (backward)."""
    X, Z = p.X, p.Z
    TL, TR, TU, TD = p.thickness_left, p.thickness_right, p.thickness_up, p.thickness_down
    xs = (TL - 8) + (TR - 8)
    zs = (TU - 8) + (TD - 8)
    strip = xs * (Z - 8) + zs * (X - 8)        # corner points carry both x and z variables
    nint = p.nx * p.nz
    fwd = 16 * X * Z + 16 * strip + 4 * nint / ts
    bwd = 16 * X * Z + 16 * strip + 20 * nint / ts + 4 * p.number_of_receivers
    return int((fwd + bwd) / 2)


def tiling_dict(shape):
    """(tile rows per thread, graph block, interior kernel 0 tiles / 1
    columns / 2 mixed tiles / 3 tiles of 8 wavefronts / 4 two-step pair
    items, rows per interior segment) as a dict."""
    rz, block, interior, seg = shape
    return {"rows_per_thread": rz, "graph_block": block,
            "interior": ("tiles", "columns", "mixed", "tiles of 8 wavefronts", "two-step pairs")[interior],
            "seg_rows": seg}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def c1_params():
    import rtmuq2d_amd as R
    nx, nz, T = 1701, 351, 50
    return R.make_params(
        nx=nx, nz=nz, thickness_left=T, thickness_right=T, thickness_up=T, thickness_down=T,
        spacing_grid=10.0, sampling=0.0008, number_of_time_steps=4000, sampling_wavelet=4000,
        first_location_shot=850, depth_shot=2, spacing_shot=1, number_of_shots=1,
        first_location_receiver=0, depth_receiver=2, spacing_receiver=2, number_of_receivers=nx,
        number_of_simulations=1, zfp_tolerance=1e-4, time_slices=1)


def c1_inputs(p, seed=2024):
    from rtmuq2d_amd import synth
    v_true = synth.folded_faulted(p.nx, p.nz, p.spacing_grid, seed)
    v_mig = synth.migration_velocity(v_true, seed + 1)
    w = synth.ricker(20.0, p.sampling, p.sampling_wavelet)
    return v_true, v_mig, w


@contextlib.contextmanager
def stdout_to_stderr():
    """Send what native code prints to fd 1 (rtm_routine's progress banner,
    as the reference prints it) to stderr, so that stdout carries only the
    result line."""
    libc = ctypes.CDLL(None)
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        yield
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)


def image_vs_reference(R, p, run, env):
    """The c1 image of the reference's rtm_routine (the timed -march=native
    build, and the build without contraction of a*b+c, oracle/_ref/c1s_strict)
    against this library's rtm_routine on the same input files in `run`
    (TS=20, the reference's compiled interval)."""
    nx, nz = p.nx, p.nz
    brute = os.path.join(run, "OUTPUTS", "stackedBrute_00001.bin")
    native = np.fromfile(brute, np.float32).reshape(nx, nz)
    strict = None
    ref_strict = os.path.join(REPO, "oracle", "_ref", "c1s_strict", "ref_main")
    if os.path.exists(ref_strict):
        r = subprocess.run([ref_strict, "1", "1"], cwd=run, env=env, capture_output=True, text=True, timeout=900)
        if r.returncode == 0:
            strict = np.fromfile(brute, np.float32).reshape(nx, nz)
    os.remove(brute)
    cwd = os.getcwd()
    os.chdir(run)
    try:
        R.set_params(p.copy().update(
            time_slices=20, velocity_file="./INPUTS/vp_%05d.bin", wavelet_file="./INPUTS/wavelet.bin",
            seismogram_file="./INPUTS/seismogram_%05d.zfp", stacked_file="./OUTPUTS/stackedBrute_%05d.bin",
            stacked_laplacian_file="./OUTPUTS/stackedLaplacian_%05d.bin"))
        with stdout_to_stderr():
            R.lib().rtm_routine(1)
    finally:
        os.chdir(cwd)
    gpu = np.fromfile(brute, np.float32).reshape(nx, nz)
    out = {"time_slices": 20, "relL2_vs_native_build": rel_l2(gpu, native)}
    if strict is not None:
        out["relL2_vs_strict_build"] = rel_l2(gpu, strict)
        out["bitexact_vs_strict_build"] = bool(np.array_equal(gpu, strict))
    log(f"[bench] c1 image of this library vs the reference's: {out}")
    return out


def host_cpu():
    """Model, logical CPUs and sockets of this host (/proc/cpuinfo)."""
    model, sockets, logical = "", set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and not model:
                    model = v
                elif k == "physical id":
                    sockets.add(v)
    except OSError:
        pass
    return {"model": model, "logical_cpus": logical or os.cpu_count(), "sockets": len(sockets) or None}


def cpu_baseline(p, v_mig, w, seis, runs=3):
    """The reference's own rtm_routine (oracle/_ref, -O3 -march=native
    OpenMP, its disk snapshots every 20 steps) on one full c1 image of the
    same inputs, on this host's cores (OMP_NUM_THREADS); the median of
    `runs` runs."""
    ref = os.path.join(REPO, "oracle", "_ref", "c1s_native", "ref_main")
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    nt_s = p.number_of_time_steps
    X, Z = p.X, p.Z
    import rtmuq2d_amd as R
    if not os.path.exists(ref):
        return None
    with tempfile.TemporaryDirectory() as run:
        for d in ("INPUTS", "OUTPUTS", "TEMPORARY_FILES"):
            os.makedirs(os.path.join(run, d))
        v_mig.astype(np.float32).tofile(os.path.join(run, "INPUTS", "vp_00001.bin"))
        w[:nt_s].astype(np.float32).tofile(os.path.join(run, "INPUTS", "wavelet.bin"))
        z = R.zfp_compress(np.ascontiguousarray(seis[:, :nt_s]), 1e-4)
        with open(os.path.join(run, "INPUTS", "seismogram_00001.zfp"), "wb") as f:
            f.write(z)
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_SCHEDULE="static",
                   OMP_PROC_BIND="close")
        times = []
        for _ in range(runs):
            r = subprocess.run([ref, "1", "1"], cwd=run, env=env, capture_output=True, text=True,
                               timeout=900)
            if r.returncode != 0:
                log("cpu baseline failed:", r.stdout[-500:], r.stderr[-500:])
                return None
            times.append([float(l.split()[1]) for l in r.stdout.splitlines() if l.startswith("REF_SECONDS")][0])
        secs = float(np.median(times))
        parity = image_vs_reference(R, p, run, env)
    upd = 2.0 * nt_s * X * Z
    sec_per_image = secs * (p.number_of_time_steps / nt_s)
    return {
        "value": round(3600.0 / sec_per_image, 3), "unit": "images/hour", "cores": cores,
        "kind": "reference",
        "gupdates_per_s": round(upd / secs / 1e9, 4),
        "sample": (f"reference rtm_routine (gcc -O3 -march=native, OpenMP, disk snapshots every "
                   f"20 steps) on the {X}x{Z} padded c1 grid, {nt_s} of {p.number_of_time_steps} "
                   f"steps; median of {runs} runs ({secs:.2f} s wall)"),
        "run_seconds": [round(t, 3) for t in times],
        "gpu_image_vs_reference": parity,
        "host": host_cpu(),
    }


def gate(rel):
    return bool(rel <= RELL2_GATE)


def headline_accuracy(R, p, v_mig, w, seis, I):
    """The headline image against the FP64 GPU run of the same scheme (FP64
    stepping, FP64 snapshot every step, FP64 image) on the same inputs."""
    t0 = time.perf_counter()
    with R.Plan(p, R.REAL_F64) as p64:
        p64.set_velocity(R.extend(p, v_mig))
        p64.set_wavelet(w)
        p64.set_seismogram(seis)
        p64.zero_image()
        p64.migrate(p.first_location_shot, p.depth_shot)
        I64, _ = p64.image(f64=True)
    rel = rel_l2(I, I64)
    err = np.abs(I.astype(np.float64) - I64)
    out = {"reference": "FP64 GPU plan of the same scheme (TS=1, FP64 snapshots)",
           "relL2_vs_fp64": rel, "max_abs_err_rel_to_max": float(err.max() / max(np.abs(I64).max(), 1e-300)),
           "gate": RELL2_GATE, "meets_gate": gate(rel), "fp64_seconds": round(time.perf_counter() - t0, 3)}
    log(f"[bench] c1 image vs FP64: relL2 {rel:.3e} (gate {RELL2_GATE:g}: {'met' if out['meets_gate'] else 'MISSED'})")
    return out


def c0_params(R):
    """Config c0 (run_rtm_simulations.local:6-10 runs it as two simulations):
    600x200, h 10 m, CPML 40, dt 1 ms, nt 2000, Ricker 15 Hz at (300, 2), 600
    receivers at iz 2 (spacing 2: the reference places receiver r at column
    first + (spacing - 1) r), the reference's snapshot interval 20."""
    T = 40
    return R.make_params(
        nx=600, nz=200, thickness_left=T, thickness_right=T, thickness_up=T, thickness_down=T,
        spacing_grid=10.0, sampling=0.001, number_of_time_steps=2000, sampling_wavelet=2000,
        first_location_shot=300, depth_shot=2, spacing_shot=1, number_of_shots=1,
        first_location_receiver=0, depth_receiver=2, spacing_receiver=2, number_of_receivers=600,
        number_of_simulations=2, zfp_tolerance=1e-4, time_slices=20,
        velocity_file="./INPUTS/vp_%05d.bin", wavelet_file="./INPUTS/wavelet.bin",
        seismogram_file="./INPUTS/seismogram_%05d.zfp", stacked_file="./OUTPUTS/stackedBrute_%05d.bin",
        stacked_laplacian_file="./OUTPUTS/stackedLaplacian_%05d.bin")


def run_c0(R, cpu=True, seed=7):
    """Config c0 end to end as the reference runs it: both simulations
    through this library's rtm_routine (velocity, wavelet and ZFP seismogram
    files in, image files out) on the GPU, then the reference's own
    rtm_routine (oracle/_ref/c0_native, -O3 -march=native OpenMP, its disk
    snapshots) on the same files on this host's cores; the images compared.
    Inputs: 5 flat layers with an anomaly (sigma 15 points), the seismogram
    forward-modelled on it on the GPU and ZFP-compressed at 1e-4, migration
    velocities smoothed (sigma 10) times (1 + 0.02 g_s), one g_s per
    simulation."""
    from rtmuq2d_amd import synth
    p = c0_params(R)
    v_true = synth.layered_anomaly(p.nx, p.nz, seed, sigma=15.0)
    models = [synth.migration_velocity(v_true, seed + s, sigma_smooth=10.0, sigma_field=20.0) for s in (1, 2)]
    w = synth.ricker(15.0, p.sampling, p.sampling_wavelet)
    with R.Plan(p) as plan:
        plan.set_velocity(R.extend(p, v_true))
        plan.set_wavelet(w)
        plan.forward(p.first_location_shot, p.depth_shot, record=True)
        seis = plan.seismogram()
    out = {"workload": "c0: two 600x200 simulations (CPML 40, nt 2000, TS 20) through rtm_routine, "
                       "files in and out, as run_rtm_simulations.local runs them"}
    with tempfile.TemporaryDirectory() as run:
        for d in ("INPUTS", "OUTPUTS", "TEMPORARY_FILES"):
            os.makedirs(os.path.join(run, d))
        for s, m in enumerate(models, start=1):
            m.astype(np.float32).tofile(os.path.join(run, "INPUTS", f"vp_{s:05d}.bin"))
        w.astype(np.float32).tofile(os.path.join(run, "INPUTS", "wavelet.bin"))
        with open(os.path.join(run, "INPUTS", "seismogram_00001.zfp"), "wb") as f:
            f.write(R.zfp_compress(seis, p.zfp_tolerance))
        brute = lambda s: os.path.join(run, "OUTPUTS", f"stackedBrute_{s:05d}.bin")  # noqa: E731
        cwd = os.getcwd()
        os.chdir(run)
        try:
            R.set_params(p)
            with stdout_to_stderr():
                R.lib().rtm_routine(1)          # warm: first plan creation, kernel loads
                t0 = time.perf_counter()
                for s in (1, 2):
                    R.lib().rtm_routine(s)
                gpu_s = time.perf_counter() - t0
        finally:
            os.chdir(cwd)
        gpu = [np.fromfile(brute(s), np.float32) for s in (1, 2)]
        out.update({"gpu_seconds": round(gpu_s, 4), "gpu_images_per_hour": round(2 * 3600.0 / gpu_s, 1)})
        ref = os.path.join(REPO, "oracle", "_ref", "c0_native", "ref_main")
        if cpu and os.path.exists(ref):
            cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
            env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_SCHEDULE="static", OMP_PROC_BIND="close")
            r = subprocess.run([ref, "1", "2"], cwd=run, env=env, capture_output=True, text=True, timeout=900)
            if r.returncode == 0:
                cpu_s = [float(l.split()[1]) for l in r.stdout.splitlines() if l.startswith("REF_SECONDS")][0]
                rels = [rel_l2(gpu[s], np.fromfile(brute(s + 1), np.float32)) for s in (0, 1)]
                out.update({"cpu_seconds": round(cpu_s, 3), "cpu_cores": cores, "cpu_kind": "reference",
                            "cpu_images_per_hour": round(2 * 3600.0 / cpu_s, 1),
                            "gpu_over_cpu": round(cpu_s / gpu_s, 2),
                            "relL2_vs_reference_native_build": rels,
                            "meets_1e-5": all(v <= 1e-5 for v in rels)})
            else:
                out["cpu_error"] = (r.stdout[-300:] + r.stderr[-300:])
    log(f"[bench] c0: {out}")
    return out


def reap_children():
    """Child processes of this one still alive (each is waited for, so none
    outlives the bench): [(pid, command)]."""
    kids = []
    try:
        for t in os.listdir("/proc/self/task"):
            with open(f"/proc/self/task/{t}/children") as f:
                kids += [int(v) for v in f.read().split()]
    except OSError:
        return None
    out = []
    for pid in sorted(set(kids)):
        try:
            with open(f"/proc/{pid}/cmdline", "rb") as f:
                cmd = f.read().replace(b"\0", b" ").decode(errors="replace").strip()
        except OSError:
            cmd = "?"
        out.append([pid, cmd])
        try:
            os.waitpid(pid, 0)
        except ChildProcessError:
            pass
    if out:
        log(f"[bench] child processes at exit: {out}")
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-precision", action="store_true", help="skip the snapshot-precision variants")
    ap.add_argument("--serve", type=int, default=1000, help="requests in the serving run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=1.0,
                    help="arrival rate as a multiple of the one-at-a-time plain rate")
    ap.add_argument("--serve-arena-gb", type=float, default=200.0, help="HBM arena of the request server")
    ap.add_argument("--serve-store", choices=("i16", "fp16"), default="i16",
                    help="request server's snapshot store: int16 blocks or FP16 with a per-snapshot scale")
    ap.add_argument("--serve-lanes", type=int, default=2, help="launch groups in flight at once in the server")
    ap.add_argument("--serve-group-mib", type=int, default=112,
                    help="launch-group budget of the server in MiB (0: the library's default, 224)")
    ap.add_argument("--serve-tiling", default="4,100,0,0",
                    help="request server kernel shape: rows_per_thread,graph_block,interior,seg_rows")
    ap.add_argument("--tiling", default=None,
                    help="fixed c1/c3 kernel shape rows_per_thread,graph_block,interior,seg_rows instead of "
                         "the autotuner's pick (for traced runs, whose per-dispatch overhead changes the pick)")
    ap.add_argument("--only-serve", action="store_true", help="run only the serving leg")
    ap.add_argument("--batch", type=int, default=64, help="realizations in the batched run (0: skip)")
    ap.add_argument("--batch-ts", type=int, default=20, help="snapshot interval of the batched run")
    ap.add_argument("--profile", action="store_true",
                    help="short fixed workload for counter profiles (see run_profile)")
    ap.add_argument("--profile-sims", type=int, default=13, help="realizations in the profile batch (a c2 launch group)")
    ap.add_argument("--profile-nt", type=int, default=400, help="time steps of the profile workload")
    ap.add_argument("--profile-block", type=int, default=100,
                    help="profile workload: steps per captured graph of the batch (1: eager launches)")
    ap.add_argument("--profile-batch-only", action="store_true",
                    help="profile workload: only the c2 launch group in 64x16 tiles (counter passes)")
    ap.add_argument("--profile-batch-shapes", default=None,
                    help="profile workload: the launch group's kernel shapes, 'rz,block,interior,seg;...'")
    ap.add_argument("--profile-lanes", type=int, default=1,
                    help="profile workload: the batch as that many launch groups in flight at once")
    ap.add_argument("--no-c0", dest="c0", action="store_false",
                    help="skip the c0 leg (two 600x200 simulations through rtm_routine, GPU vs CPU)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the RTM path shards by "
              "independent simulations, so N GPUs run N independent copies.", file=sys.stderr)
        sys.exit(2)

    import rtmuq2d_amd as R
    if R.lib().rtm_device_count() < 1:
        log("no HIP device")
        sys.exit(1)
    if args.profile:
        shapes = None
        if args.profile_batch_shapes:
            shapes = [tuple(int(v) for v in sh.split(",")) for sh in args.profile_batch_shapes.split(";")]
        run_profile(R, args.profile_nt, args.profile_sims, batch_only=args.profile_batch_only,
                    block=args.profile_block, shapes=shapes, lanes=args.profile_lanes)
        return
    if args.only_serve:
        print(json.dumps(run_serving(R, args.serve, args.serve_rate, args.serve_arena_gb,
                                     tuple(int(v) for v in args.serve_tiling.split(",")),
                                     store=R.SNAP_I16 if args.serve_store == "i16" else R.SNAP_FP16,
                                     lanes=args.serve_lanes, group_mib=args.serve_group_mib)), flush=True)
        return
    p = c1_params()
    X, Z, nt = p.X, p.Z, p.number_of_time_steps
    log(f"[bench] c1 plain RTM {p.nx}x{p.nz} padded {X}x{Z} nt={nt} TS=1 FP32; "
        f"plan footprint {R.footprint(p) / 2**30:.2f} GiB")
    v_true, v_mig, w = c1_inputs(p)
    plan = R.Plan(p)
    # seismogram: forward on the true model with receiver extraction, then ZFP
    plan.set_velocity(R.extend(p, v_true))
    plan.set_wavelet(w)
    t0 = time.time()
    plan.forward(p.first_location_shot, p.depth_shot, record=True)
    seis_true = plan.seismogram()
    t_model = time.time() - t0
    t0 = time.perf_counter()
    z = R.zfp_compress(seis_true, p.zfp_tolerance)
    t_enc = time.perf_counter() - t0
    with tempfile.NamedTemporaryFile(suffix=".zfp", delete=False) as f:
        f.write(z)
    t0 = time.perf_counter()
    seis = R.zfp_decompress_file(f.name, p.number_of_receivers, nt, p.zfp_tolerance)
    t_dec = time.perf_counter() - t0
    os.unlink(f.name)
    zfp_host = {"gather": [p.number_of_receivers, nt], "stream_bytes": len(z),
                "encode_seconds": round(t_enc, 4), "decode_seconds": round(t_dec, 4),
                "decode_MB_per_s": round(seis.nbytes / t_dec / 1e6, 1),
                "threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()}
    log(f"[bench] seismogram modelled in {t_model:.2f} s; ZFP {len(z) / 1e6:.1f} MB, host encode "
        f"{t_enc * 1e3:.1f} ms, decode (file read included) {t_dec * 1e3:.1f} ms")
    plan.set_velocity(R.extend(p, v_mig))
    plan.set_seismogram(seis)

    def one_image():
        plan.zero_image()
        plan.migrate(p.first_location_shot, p.depth_shot)

    # the library's autotuner picks the kernel shape (timed on this grid)
    tuning = []
    if args.tiling:
        plan.set_tiling(*(int(v) for v in args.tiling.split(",")))
    else:
        tuning = plan.autotune(400)
    for t in tuning:
        log(f"[bench] autotune {tiling_dict((t['rows_per_thread'], t['graph_block'], t['interior'], t['seg_rows']))}: "
            f"{t['ms_per_step'] * 1e3:.2f} us/step")
    best = (None, plan.tiling())
    log(f"[bench] using {tiling_dict(best[1])}")

    for i in range(args.warmup):
        one_image()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    plan.synchronize()
    fwd_ms, bwd_ms = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        one_image()
        f_ms, b_ms, _ = plan.last_timing()
        fwd_ms.append(f_ms)
        bwd_ms.append(b_ms)
        log(f"[bench] step {i + 1}/{args.steps}: fwd {f_ms * 1e3:.2f} us/step, bwd {b_ms * 1e3:.2f} us/step")
    plan.synchronize()
    secs = (time.perf_counter() - t0) / args.steps
    I, _ = plan.image()
    plan.close()
    accuracy = headline_accuracy(R, p, v_mig, w, seis, I)

    updates = 2.0 * nt * X * Z
    step_ms = (np.mean(fwd_ms) + np.mean(bwd_ms)) / 2.0
    bytes_per_launch = BYTES_PER_UPDATE * X * Z
    full_step = full_step_bytes(p, ts=1)
    achieved = bytes_per_launch / (step_ms * 1e-3) / 1e9
    plain = {
        "workload": "c1 plain single-simulation RTM: 1701x351, h=10 m, CPML 50, dt=0.8 ms, "
                    "nt=4000, 1701 receivers, FP32 snapshots every step on HBM",
        "images_per_hour": round(3600.0 / secs, 3),
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_image": round(secs * 1e3, 3),
        "grid_padded": [X, Z], "nt": nt, "time_slices": 1, "simulations": 1,
        "tiling": tiling_dict(best[1]),
        "autotune": [dict(tiling_dict((t["rows_per_thread"], t["graph_block"], t["interior"], t["seg_rows"])),
                          us_per_step=round(t["ms_per_step"] * 1e3, 3)) for t in tuning],
        "gupdates_per_s": round(updates / secs / 1e9, 3),
        "step_kernel_us": {"forward": round(float(np.mean(fwd_ms)) * 1e3, 3),
                           "backward": round(float(np.mean(bwd_ms)) * 1e3, 3)},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": C1_TRAFFIC_BYTES,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "full_step_bytes_per_launch": full_step,
            "full_step_achieved": round(full_step / (step_ms * 1e-3) / 1e9, 1),
        },
        "image_norm": float(np.linalg.norm(I.astype(np.float64))),
        "image_accuracy": accuracy,
        "meets_gate": accuracy["meets_gate"],
    }
    out = {
        "metric": METRIC,
        "value": plain["images_per_hour"],
        "unit": "images/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_image"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": plain["workload"], "grid_padded": [X, Z], "nt": nt, "time_slices": 1,
                   "simulations": 1, "tiling": plain["tiling"]},
        "roofline": plain["roofline"],
        "gupdates_per_s": plain["gupdates_per_s"],
        "image_accuracy": accuracy,
        "meets_gate": accuracy["meets_gate"],
        "plain": plain,
        "zfp_host": zfp_host,
    }
    if args.batch > 0:
        # the headline: the batched ensemble (c2), one step = one migration of
        # the whole batch; the plain path beside it at the same snapshot interval
        b = run_batch(R, p, v_true, w, seis, args.batch, args.batch_ts, steps=args.steps, warmup=args.warmup)
        out["batched"] = b
        out.update({
            "value": b["images_per_hour"],
            "ms_per_step": round(b["seconds"] * 1e3, 3),
            "config": {"workload": b["workload"], "grid_padded": [X, Z], "nt": nt,
                       "time_slices": b["time_slices"], "simulations": b["simulations"],
                       "launch_groups": b["launch_groups"], "tiling": b["tiling"]},
            "roofline": b["roofline"],
            "image_accuracy": {"reference": "FP64 GPU plan of the same scheme, realization 0",
                               "relL2_vs_fp64": b["image0_relL2_vs_fp64"], "gate": RELL2_GATE,
                               "meets_gate": b["meets_gate"]},
            "meets_gate": bool(b["meets_gate"] and accuracy["meets_gate"]),
            "gupdates_per_s": b["gupdates_per_s"],
        })
        out["batched_over_plain"] = {
            "ratio": b["batched_over_plain"],
            "batched": b["workload"] + f"; {b['images_per_hour']} images/h",
            "plain": f"the same realization alone on the plain path at TS={b['time_slices']}; "
                     f"{b['plain_images_per_hour_same_ts']} images/h",
            "plain_ts1_images_per_hour": plain["images_per_hour"],
            "target": 5.0,
        }
    if args.c0:
        out["c0"] = run_c0(R, cpu=not args.no_cpu)
    if not args.no_precision:
        out["precision"] = run_precision(R, p, v_mig, w, seis, best[1])
    if args.serve > 0:
        out["serving"] = run_serving(R, args.serve, args.serve_rate, args.serve_arena_gb,
                                     tuple(int(v) for v in args.serve_tiling.split(",")),
                                     store=R.SNAP_I16 if args.serve_store == "i16" else R.SNAP_FP16,
                                     lanes=args.serve_lanes, group_mib=args.serve_group_mib)
    if not args.no_cpu:
        log("[bench] timing the reference CPU path on a bounded sample ...")
        out["cpu_baseline"] = cpu_baseline(p, v_mig, w, seis)
    out["children_at_exit"] = reap_children()
    print(json.dumps(out), flush=True)


def run_profile(R, nt=400, nsims=13, batch_only=False, block=100, shapes=None, lanes=1):
    """Short, fixed workload for counter profiles (rocprofv3 --pmc runs one
    pass per counter group over the same dispatches):
      1. a 1 GiB device-to-device copy (torch), the calibration of the
         FETCH_SIZE / WRITE_SIZE counters (1 GiB read, 1 GiB written);
      2. the c1 plain workload for `nt` steps, tiles of 64x8, graph blocks;
      3. `nsims` realizations of c1 (a c2 launch group) for `nt` steps
         (TS=20), tiles of 64x16 only, then strip tiles beside the column
         kernel.
    batch_only: the copy, then step 3 in 64x16 tiles alone, so that every
    step-kernel dispatch of the run is the launch group's.  block: steps per
    captured graph of step 3 (1: every step launched on its own).  shapes:
    the batch's kernel shapes for step 3 instead (rows_per_thread,
    graph_block, interior, seg_rows each), one run per shape.  lanes: step
    3's realizations as that many launch groups in flight at once."""
    import torch
    x = torch.rand(1 << 28, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    p = c1_params().update(number_of_time_steps=nt, sampling_wavelet=nt)
    v_true, v_mig, w = c1_inputs(p)
    seis = (np.random.default_rng(3).standard_normal((p.number_of_receivers, nt)) * 1e-3).astype(np.float32)
    if not batch_only:
        with R.Plan(p) as plan:
            plan.set_tiling(2, 100, 0, 0)
            plan.set_velocity(R.extend(p, v_mig))
            plan.set_wavelet(w)
            plan.set_seismogram(seis)
            plan.zero_image()
            plan.migrate(p.first_location_shot, p.depth_shot)
            log("[profile] plain", plan.last_timing(), "image norm", float(np.linalg.norm(plan.image()[0])))
    pb = p.copy().update(time_slices=20)
    with R.Batch(pb, nsims) as b:
        b.set_group((nsims + lanes - 1) // lanes if lanes > 1 else 0)
        b.set_lanes(lanes)
        vpad = R.extend(pb, v_mig)
        for k in range(nsims):
            b.set_velocity(k, vpad)
        b.set_wavelet(w)
        b.set_seismogram(seis)
        if shapes is None:
            shapes = ((4, block, 0, 0),) if batch_only else ((4, block, 0, 0), (4, block, 1, 32))
        for shape in shapes:
            b.set_tiling(*shape)
            b.zero_images()
            b.migrate(pb.first_location_shot, pb.depth_shot)
            log(f"[profile] batch{nsims} lanes {lanes} {tiling_dict(shape)}", b.last_timing(),
                "image norms", [round(float(np.linalg.norm(b.image(k)[0])), 6) for k in (0, nsims - 1)])
    print(json.dumps({"profile": "done", "nt": nt}), flush=True)


def run_precision(R, p1, v_mig, w, seis, shape):
    """Config c3: the c1 workload with the source snapshot stored every 4
    steps in FP32, in FP16 or in BF16 (per-snapshot max-abs power-of-two
    scale), stepping FP32, image accumulated in FP32 with Kahan compensation
    (and, for reference, FP32 without it), plus boundary saving (no snapshot
    store: the 4-point ring around the image region saved every step in FP32
    and the source field rebuilt in time reverse during the backward pass);
    each image against the FP64 GPU run of the same scheme (FP64 stepping,
    FP64 snapshots, FP64 image)."""
    p = p1.copy().update(time_slices=4)
    sx, sz = p.first_location_shot, p.depth_shot
    vpad = R.extend(p, v_mig)

    def image(q, real=R.REAL_F32, shape=shape):
        with R.Plan(q, real) as plan:
            if real == R.REAL_F32:
                plan.set_tiling(*shape)
            plan.set_velocity(vpad)
            plan.set_wavelet(w)
            plan.set_seismogram(seis)
            plan.zero_image()
            plan.synchronize()
            t0 = time.perf_counter()
            plan.migrate(sx, sz)
            plan.synchronize()
            dt = time.perf_counter() - t0
            I, _ = plan.image(f64=real == R.REAL_F64)
        return I, dt

    I64, t64 = image(p, R.REAL_F64)
    nrm = float(np.max(np.abs(I64)))
    out = {"workload": "c3: c1 workload, snapshots every 4 steps (1000 stored), FP32 stepping, "
                       "errors against the FP64 GPU run of the same scheme",
           "fp64_seconds": round(t64, 4), "gate": RELL2_GATE, "variants": []}
    nint = p.nx * p.nz
    nt = p.number_of_time_steps
    ring = 8 * (p.nx + 8) + 8 * p.nz
    X = p.nx + p.thickness_left + p.thickness_right
    Z = p.nz + p.thickness_up + p.thickness_down
    for name, mode, kahan in (("fp32", R.SNAP_FP32, 1), ("fp16_scaled", R.SNAP_FP16, 1),
                              ("bf16_scaled", R.SNAP_BF16, 1), ("int16_block", R.SNAP_I16, 1),
                              ("fp32_no_kahan", R.SNAP_FP32, 0), ("boundary_saving", R.SNAP_BOUNDARY, 1)):
        q = p.copy().update(snapshot_mode=mode, image_kahan=kahan)
        I, dt = image(q, shape=shape if mode != R.SNAP_I16 or shape[2] == 0 else (shape[0], shape[1], 0, 0))
        err = np.abs(I.astype(np.float64) - I64)
        if mode == R.SNAP_BOUNDARY:     # ring saves + the two source-field buffers (pitch 64)
            store = nt * ring * 4 + 2 * Z * ((X + 63) // 64 * 64) * 4
        elif mode == R.SNAP_I16:        # codes + a byte per row and 64-column block
            nqb = ((p.thickness_left + p.nx - 5) >> 6) - ((p.thickness_left - 4) >> 6) + 1
            store = (nt // 4) * (nint * 2 + p.nz * nqb)
        else:
            store = (nt // 4) * nint * (2 if mode in (R.SNAP_FP16, R.SNAP_BF16) else 4)
        v = {"variant": name, "kahan_image": bool(kahan), "seconds_per_image": round(dt, 4),
             "snapshot_store_GB": round(store / 1e9, 3),
             "relL2_vs_fp64": rel_l2(I, I64), "max_abs_err": float(err.max()),
             "max_abs_err_rel_to_max": float(err.max() / nrm)}
        v["meets_gate"] = gate(v["relL2_vs_fp64"])
        out["variants"].append(v)
        log(f"[bench] c3 {name}: relL2 {v['relL2_vs_fp64']:.3e} ({'meets' if v['meets_gate'] else 'MISSES'} "
            f"{RELL2_GATE:g}) max-abs {v['max_abs_err']:.3e} "
            f"({v['max_abs_err_rel_to_max']:.2e} of max) {dt * 1e3:.1f} ms/image")
    return out


# c4 request classes: (label, nx, nz, h, dt, nt, CPML, Ricker f0); nt covers a
# 3.2 s record at each dt
SERVE_CLASSES = [
    ("600x200 h10", 600, 200, 10.0, 0.001, 3200, 40, 15.0),
    ("1001x301 h10", 1001, 301, 10.0, 0.0008, 4000, 50, 20.0),
    ("1701x351 h10", 1701, 351, 10.0, 0.0008, 4000, 50, 20.0),
    ("3401x701 h5", 3401, 701, 5.0, 0.0004, 8000, 50, 20.0),
]


def serve_params(R, cls, store=None):
    label, nx, nz, h, dt, nt, T, f0 = cls
    store = R.SNAP_FP16 if store is None else store
    return R.make_params(
        nx=nx, nz=nz, thickness_left=T, thickness_right=T, thickness_up=T, thickness_down=T,
        spacing_grid=h, sampling=dt, number_of_time_steps=nt, sampling_wavelet=nt,
        first_location_shot=nx // 2, depth_shot=2, spacing_shot=1, number_of_shots=1,
        first_location_receiver=0, depth_receiver=2, spacing_receiver=2, number_of_receivers=nx,
        number_of_simulations=1, zfp_tolerance=1e-4, time_slices=4, snapshot_mode=store, image_kahan=1)


def run_serving(R, nreq, rate_factor, arena_gb, shape=(4, 100, 0, 0), seed=11, store=None, lanes=1,
                group_mib=0):
    """Config c4: a Poisson stream of `nreq` migration requests, grid class
    drawn uniformly from SERVE_CLASSES, each with its own seeded folded /
    faulted velocity realization and its own source position along the
    surface, its own traces (a window of rows, at a per-request offset, of
    a seeded synthetic gather of its class with extra rows: distinct per
    request, and no copy), a 16-bit snapshot store every 4 steps (`store`:
    the int16 block store by default, or FP16), Kahan image.
    Two requests of each class, the first to arrive, are checked afterwards:
    the served image against the FP64 run of the same request (FP64
    stepping and snapshots) and against the same request run alone in a
    plan (equal bit for bit).  Arrivals come at `rate_factor` times the rate
    that one-at-a-time plain execution sustains; latency = completion -
    arrival on the wall clock; the plain comparison uses each class's
    measured one-at-a-time time (a plan autotuned on that class) in the same
    arrival order.  `shape` is the server's kernel shape: one shape serves
    every grid in a launch."""
    from rtmuq2d_amd import serve, synth
    store = R.SNAP_I16 if store is None else store
    rng = np.random.default_rng(seed)
    extra = 256                              # rows of offset a request's window can take
    ctx = []
    for cls in SERVE_CLASSES:
        p = serve_params(R, cls, store)
        wav = synth.ricker(cls[7], p.sampling, p.sampling_wavelet)
        g = np.random.default_rng(seed + cls[1])
        base = (g.standard_normal((p.number_of_receivers + extra, p.number_of_time_steps),
                                  dtype=np.float32) * np.float32(1e-3))
        ctx.append((p, wav, base))
    window = lambda base, p, o: base[o:o + p.number_of_receivers]  # noqa: E731
    # one-at-a-time plain service time per class
    service = []
    for k, cls in enumerate(SERVE_CLASSES):
        p, wav, base = ctx[k]
        seis = window(base, p, 0)
        v = synth.folded_faulted(p.nx, p.nz, p.spacing_grid, seed)
        with R.Plan(p) as plan:
            plan.set_velocity(R.extend(p, v))
            plan.autotune(100)                  # each class at its own best shape
            plan.set_wavelet(wav)
            plan.set_seismogram(seis)
            plan.zero_image()
            plan.migrate(p.nx // 2, 2)          # warm
            plan.synchronize()
            t0 = time.perf_counter()
            plan.zero_image()
            plan.migrate(p.nx // 2, 2)
            plan.synchronize()
            service.append(time.perf_counter() - t0)
        log(f"[bench] c4 plain one-at-a-time {cls[0]}: {service[-1] * 1e3:.1f} ms, "
            f"footprint {R.request_footprint(p) / 1e9:.2f} GB")
    classes = rng.integers(0, len(SERVE_CLASSES), nreq)
    plain_rate = nreq / float(np.sum([service[c] for c in classes]))
    rate = rate_factor * plain_rate
    arrivals = serve.poisson_arrivals(nreq, rate, seed)

    # every request's inputs exist before the clock starts (a client sends
    # its velocity model; making it is not part of serving)
    t0 = time.time()
    inputs = []
    for i in range(nreq):
        p, wav, base = ctx[int(classes[i])]
        v = synth.folded_faulted(p.nx, p.nz, p.spacing_grid, 1000 + i)
        g = np.random.default_rng(5000 + i)
        sx = int(g.integers(0, p.nx))
        seis = window(base, p, int(g.integers(0, extra + 1)))
        inputs.append((p, R.extend(p, v), wav, seis, sx, 2))
    log(f"[bench] c4 {nreq} request inputs made in {time.time() - t0:.1f} s")
    reqs = [serve.Request(float(arrivals[i]), (lambda i=i: inputs[i]), SERVE_CLASSES[classes[i]][0])
            for i in range(nreq)]
    order = np.argsort(arrivals, kind="stable")
    checked = []
    for k in range(len(SERVE_CLASSES)):
        checked += [int(i) for i in order if classes[i] == k][:2]
    served = {}

    def keep(i, img):
        if i in checked:
            served[i] = img

    with R.Server(int(arena_gb * 1e9), max_slots=48, snapshot_mode=store) as srv:
        srv.set_tiling(*shape)
        srv.set_lanes(lanes)
        if group_mib:
            srv.set_group_bytes(group_mib << 20)
        res = serve.run(srv, reqs, on_done=keep,
                        progress=lambda n, t: log(f"[bench] c4 {n} requests served at {t:.1f} s"))
        # capacity: the same requests all waiting at t = 0, in launch groups
        # (the default budget) and then with every active request in one group
        ncap = min(nreq, 100)
        cap = serve.run(srv, [serve.Request(0.0, r.make, r.label) for r in reqs[:ncap]])
        group_budget = srv.group_bytes()[0]
        srv.set_group_bytes(0)
        cap1 = serve.run(srv, [serve.Request(0.0, r.make, r.label) for r in reqs[:ncap]])
    cap_plain_s = float(np.sum([service[c] for c in classes[:ncap]]))
    seq = serve.sequential_latencies(arrivals, [service[c] for c in classes])
    seq_wall = max(a + l for a, l in zip(sorted(arrivals), seq))
    out = {
        "workload": (f"c4 serving: {nreq} Poisson requests at {rate:.2f}/s ({rate_factor}x the one-at-a-time "
                     "plain rate), classes " + ", ".join(c[0] for c in SERVE_CLASSES)
                     + f"; {STORE_NAMES[store]} snapshots every 4 steps, Kahan image"),
        "snapshot_store": STORE_NAMES[store],
        "requests": nreq, "arrival_rate_per_s": round(rate, 3), "server_tiling": tiling_dict(shape),
        "launch_groups_in_flight": lanes,
        "class_counts": [int((classes == k).sum()) for k in range(len(SERVE_CLASSES))],
        "latency_s": {k: round(v, 4) for k, v in res["latency_s"].items()},
        "images_per_hour": round(res["images_per_hour"], 1),
        "wall_seconds": round(res["wall_seconds"], 3),
        "max_active": res["max_active"], "max_queued": res["max_queued"],
        "plain_sequential": {
            "service_ms_per_class": {SERVE_CLASSES[k][0]: round(service[k] * 1e3, 2) for k in range(len(service))},
            "images_per_hour": round(nreq * 3600.0 / seq_wall, 1),
            "latency_s": {k: round(v, 4) for k, v in serve.percentiles(seq).items()},
        },
    }
    out["served_over_plain"] = round(out["images_per_hour"] / out["plain_sequential"]["images_per_hour"], 3)
    out["failed_requests"] = len(res.get("failed", {}))
    out["accuracy"] = serve_accuracy(R, [(i, SERVE_CLASSES[classes[i]][0], inputs[i], served.get(i)) for i in checked])
    out["meets_gate"] = all(a["meets_gate"] for a in out["accuracy"])
    out["capacity"] = {
        "requests": ncap, "served_seconds": round(cap["wall_seconds"], 3),
        "served_images_per_hour": round(cap["images_per_hour"], 1),
        "plain_sequential_seconds": round(cap_plain_s, 3),
        "plain_images_per_hour": round(ncap * 3600.0 / cap_plain_s, 1),
        "served_over_plain": round(cap_plain_s / cap["wall_seconds"], 3),
        "max_active": cap["max_active"],
        "launch_group_bytes": group_budget,
        "one_group_served_seconds": round(cap1["wall_seconds"], 3),
    }
    log(f"[bench] c4 capacity ({ncap} requests waiting): served {cap['wall_seconds']:.2f} s vs plain sequential "
        f"{cap_plain_s:.2f} s ({out['capacity']['served_over_plain']}x); one launch group "
        f"{cap1['wall_seconds']:.2f} s")
    log(f"[bench] c4 served {nreq} requests in {res['wall_seconds']:.1f} s: {out['images_per_hour']:.0f} "
        f"images/h, p50/p95/p99 {out['latency_s']['p50']:.2f}/{out['latency_s']['p95']:.2f}/"
        f"{out['latency_s']['p99']:.2f} s; plain sequential {out['plain_sequential']['images_per_hour']:.0f} "
        f"images/h, p50/p95/p99 {out['plain_sequential']['latency_s']['p50']:.2f}/"
        f"{out['plain_sequential']['latency_s']['p95']:.2f}/{out['plain_sequential']['latency_s']['p99']:.2f} s")
    return out


def serve_accuracy(R, items):
    """Each served image of `items` [(index, label, inputs, image)] against
    the FP64 plan of the same request (FP64 stepping, FP64 snapshots at the
    same interval, FP64 image) and against the FP32 plan of it alone with the
    server's store (the same arithmetic: equal bit for bit); and, beside it,
    the relL2 of the same request alone with the FP16 store."""
    out = []
    for i, label, (p, vpad, wav, seis, sx, sz), img in items:
        rec = {"request": i, "class": label}
        if img is None:
            rec.update(meets_gate=False, error="no image (request failed)")
            out.append(rec)
            continue
        with R.Plan(p.copy().update(snapshot_mode=R.SNAP_FP32), R.REAL_F64) as p64:
            p64.set_velocity(vpad)
            p64.set_wavelet(wav)
            p64.set_seismogram(seis)
            p64.zero_image()
            p64.migrate(sx, sz)
            I64, _ = p64.image(f64=True)
        alone = {}
        for mode in {p.snapshot_mode, R.SNAP_FP16}:
            with R.Plan(p.copy().update(snapshot_mode=mode)) as plan:
                plan.set_velocity(vpad)
                plan.set_wavelet(wav)
                plan.set_seismogram(seis)
                plan.zero_image()
                plan.migrate(sx, sz)
                alone[mode] = plan.image()[0]
        rel = rel_l2(img, I64)
        rec.update(relL2_vs_fp64=rel, meets_gate=gate(rel),
                   equals_plan_alone=bool(np.array_equal(img, alone[p.snapshot_mode])),
                   fp16_store_relL2_vs_fp64=rel_l2(alone[R.SNAP_FP16], I64))
        log(f"[bench] c4 request {i} ({label}): relL2 vs FP64 {rel:.3e}, equals the plan alone: "
            f"{rec['equals_plan_alone']}")
        out.append(rec)
    return out


STORE_NAMES = {0: "FP32", 1: "FP16 (per-snapshot scale)", 2: "BF16 (per-snapshot scale)", 3: "boundary saving",
               4: "int16 blocks (a power-of-two scale per 64-point run)"}


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_batch(R, p1, v_true, w, seis, nsims, ts, steps=3, warmup=1):
    """Config c2: `nsims` realizations of the c1 model, v_k = gauss(v_true, 10)
    * exp(0.03 g_k) with g_k a Gaussian random field of 20-point correlation
    length, sharing one seismogram; snapshots every `ts` steps (FP32);
    FP64 Welford mean/variance.  Compared with the plain path on the same
    per-realization workload (one simulation per pass), and with an FP64 run
    of realization 0 (the FP64 oracle of the same scheme)."""
    from rtmuq2d_amd import synth
    p = p1.copy().update(time_slices=ts)
    X, Z, nt = p.X, p.Z, p.number_of_time_steps
    sx, sz = p.first_location_shot, p.depth_shot
    t0 = time.time()
    vs = [R.extend(p, synth.ensemble_velocity(v_true, 100 + k)) for k in range(nsims)]
    log(f"[bench] {nsims} realizations generated in {time.time() - t0:.1f} s; batch footprint "
        f"{R.batch_footprint(p, nsims) / 2**30:.1f} GiB")

    # plain comparator: realization 0 alone, same snapshot interval
    plan = R.Plan(p)
    plan.set_velocity(vs[0])
    plan.set_wavelet(w)
    plan.set_seismogram(seis)
    plan.autotune(400)
    best = (None, plan.tiling())
    reps = 3
    plan.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.zero_image()
        plan.migrate(sx, sz)
    plan.synchronize()
    plain_secs = (time.perf_counter() - t0) / reps
    I_plain, _ = plan.image()
    plan.close()
    log(f"[bench] plain (1 sim, TS={ts}) {tiling_dict(best[1])}: {plain_secs * 1e3:.1f} ms/image")

    # FP64 oracle run of realization 0
    with R.Plan(p, R.REAL_F64) as p64:
        p64.set_velocity(vs[0])
        p64.set_wavelet(w)
        p64.set_seismogram(seis)
        p64.zero_image()
        p64.migrate(sx, sz)
        I64, _ = p64.image(f64=True)

    b = R.Batch(p, nsims)
    for k, v in enumerate(vs):
        b.set_velocity(k, v)
    b.set_wavelet(w)
    b.set_seismogram(seis)
    # launch groups: each candidate group size shape-tuned, the least time
    # per realization-step kept; then the shape table at that group size
    gsweep = b.autotune_group(200)
    for t in gsweep:
        log(f"[bench] batch group {t['group']}: {t['us_per_sim_step']:.2f} us per realization-step")
    sweep = b.autotune(200)
    for t in sweep:
        log(f"[bench] batch autotune {tiling_dict((t['rows_per_thread'], t['graph_block'], t['interior'], t['seg_rows']))}: "
            f"{t['ms_per_step'] * 1e3:.1f} us/step")
    # launch groups in flight at once: one lane at the tuned size, then two
    # lanes of smaller groups, each timed as a wave of groups at that shape
    lsweep = b.autotune_lanes(200)
    for t in lsweep:
        log(f"[bench] batch lanes {t['lanes']} x groups of {t['group']}: {t['us_per_sim_step']:.3f} us per "
            "realization-step")
    bbest = (None, b.tiling())
    gsize, ngroups = b.group()
    lanes = b.lanes()
    b_group = (gsize, ngroups)
    log(f"[bench] batch using {tiling_dict(bbest[1])}, {ngroups} launch groups of {gsize}, {lanes} in flight")
    for i in range(warmup):
        b.zero_images()
        b.migrate(sx, sz)
        log(f"[bench] batch warmup {i + 1}/{warmup} done")
    # the timed region: `steps` whole-batch migrations (each returns when its
    # last pass has ended, so the host clock brackets finished device work)
    runs, steps_fb = [], []
    t_all = time.perf_counter()
    for i in range(steps):
        b.zero_images()
        t0 = time.perf_counter()
        b.migrate(sx, sz)
        runs.append(time.perf_counter() - t0)
        steps_fb.append(b.last_timing()[:2])
        log(f"[bench] batch step {i + 1}/{steps}: {runs[-1]:.3f} s")
    batch_secs = (time.perf_counter() - t_all) / steps
    f_ms = float(np.mean([t[0] for t in steps_fb]))
    b_ms = float(np.mean([t[1] for t in steps_fb]))
    t0 = time.perf_counter()
    mean, var = b.welford()
    welford_s = time.perf_counter() - t0
    I0, _ = b.image(0)
    b.close()
    step_ms = (f_ms + b_ms) / 2.0          # one time step of every launch group
    achieved = BYTES_PER_UPDATE * X * Z * nsims / (step_ms * 1e-3) / 1e9
    gsize_, ngroups_ = b_group
    lanes_ = lanes
    plain_iph = 3600.0 / plain_secs
    batch_iph = nsims * 3600.0 / batch_secs
    out = {
        "workload": f"c2 ensemble: {nsims} realizations of the c1 model, TS={ts} FP32 snapshots, shared "
                    "seismogram, FP64 Welford mean/variance",
        "simulations": nsims, "time_slices": ts,
        "images_per_hour": round(batch_iph, 1),
        "gupdates_per_s": round(2.0 * nt * X * Z * nsims / batch_secs / 1e9, 2),
        "seconds": round(batch_secs, 4),
        "steps": steps, "warmup": warmup,
        "run_seconds": [round(t, 4) for t in runs],
        "plain_images_per_hour_same_ts": round(plain_iph, 1),
        "batched_over_plain": round(batch_iph / plain_iph, 3),
        "step_kernel_us": {"forward": round(f_ms * 1e3, 2), "backward": round(b_ms * 1e3, 2)},
        "tiling": tiling_dict(bbest[1]),
        "launch_groups": {"sims_per_group": gsize, "groups": ngroups, "in_flight": lanes,
                          "autotune_us_per_sim_step": {str(t["group"]): round(t["us_per_sim_step"], 3)
                                                       for t in gsweep},
                          "autotune_lanes_us_per_sim_step": {f"{t['lanes']}x{t['group']}": round(t["us_per_sim_step"], 3)
                                                             for t in lsweep}},
        "autotune": [dict(tiling_dict((t["rows_per_thread"], t["graph_block"], t["interior"], t["seg_rows"])),
                          us_per_step=round(t["ms_per_step"] * 1e3, 2)) for t in sweep],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "traffic_note": "not measured: under a rocprofv3 pass of L2 (TCC) counters the batched "
                                     "kernels skip their work (zero images), so no L2 reading describes them; "
                                     "SQ counter passes leave their work intact",
                     "algorithmic_bytes_per_launch": round(BYTES_PER_UPDATE * X * Z * nsims / ngroups_),
                     "launch_us": round(step_ms * 1e3 / ngroups_, 3),
                     "launch_note": f"{ngroups_} launches per time step, one per launch group of at most {gsize_} "
                                    f"realizations ({nsims} in all), {lanes_} in flight at once on streams of their "
                                    "own; bytes are the mean per launch and time the step's span over the number "
                                    "of launches"},
        "welford_seconds": round(welford_s, 4),
        "image0_relL2_vs_fp64": rel_l2(I0, I64),
        "meets_gate": gate(rel_l2(I0, I64)),
        "plain_relL2_vs_fp64": rel_l2(I_plain, I64),
        "image0_equals_plain": bool(np.array_equal(I0, I_plain)),
        "mean_image_norm": float(np.linalg.norm(mean)),
        "var_image_norm": float(np.linalg.norm(var)),
    }
    log(f"[bench] batched {batch_iph:.0f} images/h vs plain {plain_iph:.0f} ({out['batched_over_plain']}x); "
        f"relL2 vs FP64 {out['image0_relL2_vs_fp64']:.2e}")
    return out


if __name__ == "__main__":
    main()
