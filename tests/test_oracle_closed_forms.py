"""Pins of the oracle that do not depend on the reference: closed forms,
invariants and index identities the mathematics fixes.  Each is chosen so a
plausible slip (a dropped or transposed term, a wrong sign or index) fails."""
from __future__ import annotations

import math

import numpy as np
import pytest

import common
import oracle as O

A = np.array([-2.97399944, 1.70507669, -0.25861812, 0.04577745, -0.00523630], np.float64)


def _op(nx=40, nz=36, T=12, nt=20, ts=1, nrec=10, first=0, depth=3, spacing=3, h=10.0, dt=1e-3):
    return O.OrcParams(nx, nz, T, T + 2, T + 1, T + 3, h, dt, nt, nt, first, depth, spacing, nrec, ts)


def test_stencil_constant_field():
    """A0 + 2 sum A_m = 0: a constant field is advanced as U = 2u - U_old."""
    op = _op()
    rng = np.random.default_rng(0)
    Cf = rng.uniform(0.01, 0.2, (op.X, op.Z)).astype(np.float32)
    u = np.full((op.X, op.Z), 3.0, np.float64)
    Uold = rng.standard_normal((op.X, op.Z))
    U = O.stencil(op, Cf.astype(np.float64), u, Uold)
    inner = (slice(4, op.X - 4), slice(4, op.Z - 4))
    expect = 2 * u - Uold
    assert np.allclose(U[inner], expect[inner], atol=1e-6)
    # untouched border
    assert np.array_equal(U[:4], Uold[:4]) and np.array_equal(U[:, -4:], Uold[:, -4:])


@pytest.mark.parametrize("axis", [0, 1])
def test_stencil_quadratic_factor(axis):
    """u = i^2 (or k^2): lap = 2 sum m^2 A_m = 1.99764 (optimized, not 2),
    so a missing, swapped or mis-weighted neighbour term is caught."""
    op = _op()
    ii, kk = np.meshgrid(np.arange(op.X, dtype=np.float64), np.arange(op.Z, dtype=np.float64),
                         indexing="ij")
    u = (ii if axis == 0 else kk) ** 2
    Cf = np.ones_like(u)
    U = O.stencil(op, Cf, u, np.zeros_like(u))
    lap = (U - 2 * u)[4:-4, 4:-4]
    assert abs(2 * sum(m * m * A[m] for m in range(1, 5)) - 1.99764092) < 1e-7
    # the scheme's weights are the float32 values of the decimals; with them
    # lap(i^2) = i^2 (2 A0 + 4 sum A_m) + 2 sum m^2 A_m exactly (in FP64)
    Af = A.astype(np.float32).astype(np.float64)
    c2 = 2 * Af[0] + 4 * Af[1:].sum()
    factor = 2 * sum(m * m * Af[m] for m in range(1, 5))
    x = (ii if axis == 0 else kk)[4:-4, 4:-4]
    assert np.max(np.abs(lap - (c2 * x ** 2 + factor))) <= 1e-9


def test_profiles_closed_form():
    """a, b of each CPML side from the closed form of the profile
    (sigma_max from the LEFT thickness for every side, vmax 4000, 60 Hz)."""
    op = _op(T=14)
    sides = [("L", op.t_left, True), ("R", op.t_right, False), ("T", op.t_up, True), ("B", op.t_down, False)]
    got32 = O.profiles(op, np.float32)
    got64 = O.profiles(op, np.float64)
    h, dt = float(op.h), float(op.dt)
    Lleft = (op.t_left - 8) * h
    smax = 3 * math.pi * 3 * 4000.0 / Lleft
    amax = math.pi * 60.0 / (3 * math.sqrt(math.pi))
    for n, (name, T, outer_first) in enumerate(sides):
        Lm = (T - 8) * h
        a = np.zeros(T)
        b = np.zeros(T)
        for t in range(T - 8):
            d = (Lm - h * t) / Lm if outer_first else (h + h * t) / Lm
            s = smax * d * d
            al = amax * (1 - d) + float(np.float32(0.1)) * amax   # the reference's 0.1f
            b[t + 4] = math.exp(-(s + al) * dt)
            a[t + 4] = s * (b[t + 4] - 1) / (s + al)
        assert np.allclose(got64[n], a, rtol=1e-7, atol=0), name
        assert np.allclose(got64[4 + n], b, rtol=1e-12, atol=0), name
        assert np.allclose(got32[n], a, rtol=2e-5, atol=1e-7), name
        assert np.allclose(got32[4 + n], b, rtol=2e-6, atol=0), name
        assert np.all(got32[n][:4] == 0) and np.all(got32[n][T - 4:] == 0)


def test_cpml_zero_state_stays_zero_and_outside_strips_untouched():
    op = _op()
    rng = np.random.default_rng(1)
    Cf = rng.uniform(0.01, 0.2, (op.X, op.Z)).astype(np.float32)
    prof = O.profiles(op)
    shapes = [(op.t_left, op.Z), (op.t_right, op.Z), (op.X, op.t_up), (op.X, op.t_down)]
    psi = [np.zeros(s, np.float32) for s in shapes]
    csi = [np.zeros(s, np.float32) for s in shapes]
    z = np.zeros((op.X, op.Z), np.float32)
    U = O.cpml(op, z, z, Cf, prof, psi, csi)
    assert not U.any() and not any(a.any() for a in psi + csi)
    # random u, U: corrections only inside the four strips
    u = rng.standard_normal((op.X, op.Z)).astype(np.float32)
    U0 = rng.standard_normal((op.X, op.Z)).astype(np.float32)
    U = O.cpml(op, u, U0, Cf, prof, psi, csi)
    interior = (slice(op.t_left - 4, op.X - op.t_right + 4), slice(op.t_up - 4, op.Z - op.t_down + 4))
    assert np.array_equal(U[interior], U0[interior])
    assert not np.array_equal(U, U0)


def test_source_and_receiver_delta():
    """One forward step from rest: U = w[0]*C at the source, 0 elsewhere.
    One backward step: U = d[r, nt-1] at pos_r = first + (spacing-1) r."""
    op = _op(nt=1, ts=1, nrec=7, first=2, spacing=4, depth=5)
    rng = np.random.default_rng(2)
    model = rng.uniform(1500, 3000, (op.nx, op.nz)).astype(np.float32)
    vpad = O.extend(op, model)
    Cf = O.coef(op, vpad)
    w = rng.standard_normal(1).astype(np.float32)
    sx, sz = 11, 6
    S, d, _, U = O.forward(op, sx, sz, vpad, w)
    xs, zs = sx + op.t_left, sz + op.t_up
    expect = np.zeros_like(U)
    expect[xs, zs] = w[0] * Cf[xs, zs]
    assert np.array_equal(U, expect)
    zr = op.t_up + op.depth_receiver
    for r in range(op.n_receivers):   # recorded where the receivers sit
        assert d[r, 0] == U[op.t_left + op.first_receiver + (op.spacing_receiver - 1) * r, zr]
    seis = rng.standard_normal((op.n_receivers, 1)).astype(np.float32)
    _, _, _, Ub = O.backward(op, vpad, seis, np.zeros((1, op.nx, op.nz), np.float32))
    expect = np.zeros_like(Ub)
    for r in range(op.n_receivers):
        expect[op.t_left + 2 + 3 * r, zr] = seis[r, 0]
    assert np.array_equal(Ub, expect)


def test_image_closed_form():
    """Constant S and R: after n imaging calls I = n S R, Is = n S^2."""
    op = _op()
    S = np.full((op.X, op.Z), 0.5, np.float32)
    Rw = np.full((op.X, op.Z), -0.25, np.float32)
    I = np.zeros((op.nx, op.nz), np.float32)
    Is = np.zeros_like(I)
    for _ in range(12):
        O.image(op, I, Is, S, Rw)
    assert np.all(I == np.float32(12 * 0.5 * -0.25)) and np.all(Is == np.float32(12 * 0.25))


def test_snapshot_index_pairs_with_step():
    """Snapshot c holds the forward field after step c*TS (interior), so the
    backward pass at tau = c*TS images against the field of the same time."""
    op = _op(nt=40, ts=8)
    rng = np.random.default_rng(4)
    vpad = O.extend(op, rng.uniform(1500, 2500, (op.nx, op.nz)).astype(np.float32))
    w = rng.standard_normal(op.n_wavelet).astype(np.float32)
    S, _, _, _ = O.forward(op, 15, 10, vpad, w)
    for c in (0, 2, 4):
        q = _op(nt=c * 8 + 1, ts=1)
        q.n_wavelet = op.n_wavelet
        _, _, _, U = O.forward(q, 15, 10, vpad, w, snaps=False, record=False)
        assert np.array_equal(S[c], U[op.t_left:op.t_left + op.nx, op.t_up:op.t_up + op.nz])


def test_rtm_brute_force_tiny():
    """A direct double-loop restatement in numpy (float64) of one full shot on
    a tiny grid agrees with the FP64 oracle to rounding."""
    op = _op(nx=14, nz=12, T=10, nt=30, ts=3, nrec=6, first=1, depth=2, spacing=3, dt=1e-3)
    op.t_right = op.t_up = op.t_down = 10
    rng = np.random.default_rng(9)
    vpad = O.extend(op, rng.uniform(1500, 2500, (op.nx, op.nz)).astype(np.float32))
    w = rng.standard_normal(op.nt).astype(np.float32)
    seis = rng.standard_normal((op.n_receivers, op.nt)).astype(np.float32)
    I_o, Is_o = O.rtm_shot(op, 6, 3, vpad, w, seis, dtype=np.float64)
    I_b, Is_b = _brute_rtm(op, 6, 3, vpad.astype(np.float64), w.astype(np.float64), seis.astype(np.float64))
    assert np.allclose(I_o, I_b, rtol=1e-10, atol=1e-12 * np.abs(I_b).max())
    assert np.allclose(Is_o, Is_b, rtol=1e-10, atol=1e-12 * np.abs(Is_b).max())


def _brute_rtm(op, sx, sz, v, w, seis):
    X, Z, nt, TS = op.X, op.Z, op.nt, op.time_slices
    h, dt = float(op.h), float(op.dt)
    Cf = (v * dt / h) ** 2
    prof = O.profiles(op, np.float64)
    aL, aR, aT, aB, bL, bR, bT, bB = prof
    ax, bx, az, bz = np.zeros(X), np.zeros(X), np.zeros(Z), np.zeros(Z)
    ax[:op.t_left], bx[:op.t_left] = aL, bL
    ax[X - op.t_right:], bx[X - op.t_right:] = aR, bR
    az[:op.t_up], bz[:op.t_up] = aT, bT
    az[Z - op.t_down:], bz[Z - op.t_down:] = aB, bB
    B = [0.0, 0.8, -0.2, 0.025396825, -0.001785714]
    Bf = [float(np.float32(b)) for b in B]
    Af = [float(np.float32(a)) for a in A]
    inX = lambda i: 4 <= i < op.t_left - 4 or X - op.t_right + 4 <= i < X - 4  # noqa: E731
    inZ = lambda k: 4 <= k < op.t_up - 4 or Z - op.t_down + 4 <= k < Z - 4  # noqa: E731

    def run(direction, snaps=None):
        u, U = np.zeros((X, Z)), np.zeros((X, Z))
        pX, cX, pZ, cZ = np.zeros((X, Z)), np.zeros((X, Z)), np.zeros((X, Z)), np.zeros((X, Z))
        out_snaps = []
        I, Is = np.zeros((op.nx, op.nz)), np.zeros((op.nx, op.nz))
        for n in range(nt):
            new = U.copy()
            for i in range(4, X - 4):
                for k in range(4, Z - 4):
                    lap = 2 * Af[0] * u[i, k] + sum(
                        Af[m] * (u[i - m, k] + u[i + m, k] + u[i, k - m] + u[i, k + m]) for m in range(1, 5))
                    new[i, k] = Cf[i, k] * lap + 2 * u[i, k] - U[i, k]
            npX, npZ = pX.copy(), pZ.copy()
            for i in range(X):
                for k in range(4, Z - 4):
                    if inX(i):
                        npX[i, k] = bx[i] * pX[i, k] + ax[i] * sum(Bf[m] * (u[i + m, k] - u[i - m, k]) for m in range(1, 5)) / h
                for k in range(Z):
                    if inZ(k) and 4 <= i < X - 4:
                        npZ[i, k] = bz[k] * pZ[i, k] + az[k] * sum(Bf[m] * (u[i, k + m] - u[i, k - m]) for m in range(1, 5)) / h
            for i in range(4, X - 4):
                for k in range(4, Z - 4):
                    corr = []
                    if inX(i):
                        dp = sum(Bf[m] * (npX[i + m, k] - npX[i - m, k]) for m in range(1, 5)) / h
                        d2 = sum(Af[m] * (u[i + m, k] + u[i - m, k]) for m in range(1, 5)) + Af[0] * u[i, k]
                        cX[i, k] = bx[i] * cX[i, k] + ax[i] * (d2 / h ** 2 + dp)
                        corr.append(("x", Cf[i, k] * h * h * (dp + cX[i, k])))
                    if inZ(k):
                        dp = sum(Bf[m] * (npZ[i, k + m] - npZ[i, k - m]) for m in range(1, 5)) / h
                        d2 = sum(Af[m] * (u[i, k + m] + u[i, k - m]) for m in range(1, 5)) + Af[0] * u[i, k]
                        cZ[i, k] = bz[k] * cZ[i, k] + az[k] * (d2 / h ** 2 + dp)
                        corr.append(("z", Cf[i, k] * h * h * (dp + cZ[i, k])))
                    for _, c in corr:
                        new[i, k] += c
            pX, pZ = npX, npZ
            if direction == "fwd":
                if n < op.n_wavelet:
                    new[sx + op.t_left, sz + op.t_up] += w[n] * Cf[sx + op.t_left, sz + op.t_up]
                if n % TS == 0:
                    out_snaps.append(new[op.t_left:op.t_left + op.nx, op.t_up:op.t_up + op.nz].copy())
            else:
                tau = nt - 1 - n
                for r in range(op.n_receivers):
                    new[op.t_left + op.first_receiver + (op.spacing_receiver - 1) * r,
                        op.t_up + op.depth_receiver] = seis[r, tau]
                if tau % TS == 0:
                    s = snaps[tau // TS]
                    R = new[op.t_left:op.t_left + op.nx, op.t_up:op.t_up + op.nz]
                    I += s * R
                    Is += s * s
            u, U = new, u
        return out_snaps, I, Is

    snaps, _, _ = run("fwd")
    _, I, Is = run("bwd", snaps)
    return I, Is


def test_kahan_image_option_closed_form():
    """The oracle's Kahan image option (not in the reference; the device's
    compensated image is checked against it): with the snapshot of the first
    imaged step at 1 and every other snapshot at 2^-13, the exact self-image
    is 1 + (n-1) 2^-26.  Each 2^-26 is below half an FP32 ulp of 1, so plain
    summation stays at exactly 1, while compensated summation lands on the
    FP32 value nearest the exact sum (1 + 125 * 2^-23 for n = 1000).  A
    compensation dropped, applied with the wrong sign or updated in another
    order gives 1 or a value several ulps off."""
    n = 1000
    op = _op(nt=n, ts=1)
    rng = np.random.default_rng(9)
    vpad = O.extend(op, rng.uniform(1500, 2500, (op.nx, op.nz)).astype(np.float32))
    seis = np.zeros((op.n_receivers, op.nt), np.float32)
    snaps = np.full((n, op.nx, op.nz), np.float32(2.0 ** -13), np.float32)
    snaps[n - 1] = 1.0                       # imaged first: the backward pass starts at tau = nt-1
    _, Is_plain, _, _ = O.backward(op, vpad, seis, snaps)
    op.image_kahan = 1
    I_k, Is_k, _, _ = O.backward(op, vpad, seis, snaps)
    exact = 1.0 + (n - 1) * 2.0 ** -26
    nearest = np.float32(1.0 + round((exact - 1.0) * 2 ** 23) * 2.0 ** -23)
    assert np.all(Is_plain == np.float32(1.0))
    assert np.all(Is_k == nearest), (float(Is_k.min()), float(Is_k.max()), float(nearest))
    assert not I_k.any()                     # zero traces: the receiver field stays zero


def test_kahan_image_option_tracks_fp64():
    """On a modelled shot the compensated FP32 image is at least as close to
    the FP64 oracle's as the plain one, and differs from it only by rounding
    (the imaging products are the same)."""
    op = _op(nt=400, ts=1, nrec=20, spacing=2)
    rng = np.random.default_rng(10)
    vpad = O.extend(op, rng.uniform(1500, 2500, (op.nx, op.nz)).astype(np.float32))
    w = rng.standard_normal(op.n_wavelet).astype(np.float32) * 1e-2
    S, d, _, _ = O.forward(op, 20, 6, vpad, w)
    S64, _, _, _ = O.forward(op, 20, 6, vpad, w, dtype=np.float64)
    I, Is, _, _ = O.backward(op, vpad, d, S)
    I64, Is64, _, _ = O.backward(op, vpad, d, S64)
    op.image_kahan = 1
    Ik, Isk, _, _ = O.backward(op, vpad, d, S)
    assert common.rel_l2(Ik, I64) <= common.rel_l2(I, I64) * 1.001
    assert common.rel_l2(Isk, Is64) <= common.rel_l2(Is, Is64) * 1.001
    assert common.rel_l2(Ik, I) < 1e-5 and not np.array_equal(Ik, I)
