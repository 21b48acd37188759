"""Device-layout driver of the step arithmetic for the CPU harness (tests).

Mirrors, in numpy, what the HIP plan does between launches (state reset,
ping-pong of u and psi, per-step mode bits, snapshot indices) and calls the
per-point update of rtmuq2d_amd/csrc/rtm_step.h compiled for the CPU
(tests/cpu_step_harness.cpp).  Used to prove that the device arithmetic is
bit-identical to the oracle before any GPU run.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

import common
import oracle as O

M_STENCIL, M_CPML, M_SOURCE, M_INJECT, M_SNAP, M_IMAGE, M_KAHAN = 1, 2, 4, 8, 16, 32, 64
SNAP_WORDS = 32 * 64   # max-abs slots per snapshot (rtm::kSnapWords)
IN_PAIRS = 4           # rtm::IN_PAIRS, two-step launches over the inner image region
Q_I16 = 3              # rtm::Q_I16, the int16 block store


def i16_blocks(TL, nx):
    """Blocks of 64 padded columns (from column 4) holding image columns."""
    return ((TL + nx - 1 - 4) >> 6) - ((TL - 4) >> 6) + 1


def build_harness(outdir: str) -> C.CDLL:
    src = os.path.join(common.REPO, "tests", "cpu_step_harness.cpp")
    inc = os.path.join(common.REPO, "rtmuq2d_amd", "csrc")
    so = os.path.join(outdir, "libstep_harness.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I", inc,
                    src, "-o", so], check=True)
    return C.CDLL(so)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class DevSim:
    def __init__(self, hx, op: O.OrcParams, vpad, dtype=np.float32):
        self.hx, self.op, self.dt = hx, op, np.dtype(dtype)
        self.fn = hx.hx_step_f64 if self.dt == np.float64 else hx.hx_step_f32
        X, Z = op.X, op.Z
        self.X, self.Z, self.P = X, Z, (X + 63) // 64 * 64
        self.C = self.to_dev(O.coef(op, vpad, dtype))
        aL, aR, aT, aB, bL, bR, bT, bB = O.profiles(op, dtype)
        self.ax = np.zeros(X, dtype); self.bx = np.zeros(X, dtype)
        self.az = np.zeros(Z, dtype); self.bz = np.zeros(Z, dtype)
        self.ax[:op.t_left], self.bx[:op.t_left] = aL, bL
        self.ax[X - op.t_right:], self.bx[X - op.t_right:] = aR, bR
        self.az[:op.t_up], self.bz[:op.t_up] = aT, bT
        self.az[Z - op.t_down:], self.bz[Z - op.t_down:] = aB, bB
        if self.dt == np.float32:
            h = np.float32(op.h)
            self.hs = np.array([np.float32(1) / h, h * h, np.float32(1) / (h * h), 0], np.float32)
        else:
            h = float(op.h)
            self.hs = np.array([1.0 / h, h * h, 1.0 / (h * h), 0.0], np.float64)
        self.recmap = np.full(X, -1, np.int32)
        zr = op.t_up + op.depth_receiver
        for r in range(op.n_receivers):
            self.recmap[op.t_left + op.first_receiver + (op.spacing_receiver - 1) * r] = r
        self.rec_row = zr

    def to_dev(self, a):
        d = np.zeros((self.Z, self.P), self.dt)
        d[:, :self.X] = np.asarray(a).T
        return d

    def from_dev(self, d):
        return np.ascontiguousarray(d[:, :self.X].T)

    def reset(self):
        z = lambda: np.zeros((self.Z, self.P), self.dt)  # noqa: E731
        self.u = [z(), z()]
        self.psiX = [z(), z()]
        self.psiZ = [z(), z()]
        self.csiX, self.csiZ = z(), z()
        self.cur = self.pc = 0
        self.last = None

    def step(self, mode, src_j=-1, src_w=0.0, seis_row=None, snap=None, I=None, Is=None, Ic=None,
             Isc=None, edges=False):
        op = self.op
        g = np.array([self.X, self.Z, self.P, op.t_left, op.t_right, op.t_up, op.t_down, op.nx, op.nz,
                      mode, src_j, self.rec_row, op.n_receivers, int(edges)], np.int32)
        hs = self.hs.copy()
        hs[3] = src_w
        c, pc = self.cur, self.pc
        self.fn(_p(g), _p(hs), _p(self.u[c]), _p(self.u[c ^ 1]), _p(self.C), _p(self.ax), _p(self.bx),
                _p(self.az), _p(self.bz), _p(self.psiX[pc]), _p(self.psiX[pc ^ 1]), _p(self.csiX),
                _p(self.psiZ[pc]), _p(self.psiZ[pc ^ 1]), _p(self.csiZ), _p(self.recmap),
                _p(seis_row), _p(snap), _p(I), _p(Is), _p(Ic), _p(Isc))
        self.cur ^= 1
        self.pc ^= 1

    def forward(self, shot_x, shot_z, wav):
        op = self.op
        self.reset()
        TS = op.time_slices
        snaps = np.zeros((op.nt // TS, op.nz, op.nx), self.dt)
        src_j = (shot_z + op.t_up) * self.P + shot_x + op.t_left
        for n in range(op.nt):
            mode = M_STENCIL | M_CPML
            w = 0.0
            if n < op.n_wavelet:
                mode |= M_SOURCE
                w = wav[n]
            snap = None
            if n % TS == 0:
                mode |= M_SNAP
                snap = snaps[n // TS]
            self.step(mode, src_j, w, snap=snap)
        return snaps

    def backward(self, seis, snaps, kahan=False):
        op = self.op
        self.reset()
        TS = op.time_slices
        seis_t = np.ascontiguousarray(np.asarray(seis, np.float32).T)
        I = np.zeros((op.nz, op.nx), self.dt)
        Is, Ic, Isc = np.zeros_like(I), np.zeros_like(I), np.zeros_like(I)
        for n in range(op.nt):
            tau = op.nt - n - 1
            mode = M_STENCIL | M_CPML | M_INJECT
            snap = None
            if tau % TS == 0:
                mode |= M_IMAGE | (M_KAHAN if kahan else 0)
                snap = snaps[tau // TS]
            self.step(mode, seis_row=seis_t[tau], snap=snap, I=I, Is=Is, Ic=Ic, Isc=Isc)
        return np.ascontiguousarray(I.T), np.ascontiguousarray(Is.T)

    def last_fields(self):
        if self.last is not None:
            return self.from_dev(self.last[0]), self.from_dev(self.last[1])
        return self.from_dev(self.u[self.cur ^ 1]), self.from_dev(self.u[self.cur])

    # ---- the tiled kernel (rtm_tile.h), emulated phase by phase ----

    def _tiles_pass(self, phase, flags, rz, src_j=-1, wav=None, seis=None, snaps=None, record=False,
                    interior=0, seg=0, qmode=0, q16=None, smax=None, block=0):
        op = self.op
        self.reset()
        nint = op.nx * op.nz
        TS = op.time_slices
        fn = self.hx.hx_tiles_f64 if self.dt == np.float64 else self.hx.hx_tiles_f32
        wav = np.zeros(max(op.n_wavelet, 1), np.float32) if wav is None else np.ascontiguousarray(wav, np.float32)
        seis_t = (np.zeros((op.nt, op.n_receivers), np.float32) if seis is None
                  else np.ascontiguousarray(np.asarray(seis, np.float32).T))
        seis_rec = np.zeros((op.nt, op.n_receivers), np.float32)
        if snaps is None:
            snaps = np.zeros((op.nt // TS, op.nz, op.nx), self.dt)
        I = np.zeros((op.nz, op.nx), self.dt)
        Is, Ic, Isc = np.zeros_like(I), np.zeros_like(I), np.zeros_like(I)
        if q16 is None:
            q16 = np.zeros((op.nt // TS, op.nz, op.nx), np.uint16)
            smax = np.zeros((op.nt // TS, SNAP_WORDS), np.uint32)
        pairs = interior == IN_PAIRS
        if pairs:               # two-step launches step four u buffers (u^n in u[n & 3])
            self.u += [np.zeros((self.Z, self.P), self.dt), np.zeros((self.Z, self.P), self.dt)]
        arrs = [self.u[0], self.u[1], self.C, self.ax, self.bx, self.az, self.bz, self.psiX[0],
                self.psiX[1], self.psiZ[0], self.psiZ[1], self.csiX, self.csiZ, snaps, I, Is, Ic, Isc,
                q16, smax] + (self.u[2:4] if pairs else [])
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        hs = self.hs[:3].copy()
        # launches as the executor issues them: with two-step launches the
        # first of each pair steps the pair items at n and n+1, and `block`
        # (when given) splits the run into blocks (an odd block ends on a
        # launch whose second step is not part of it)
        run_len = block if block else op.nt
        for n in range(op.nt):
            t = n % run_len
            left = min(run_len, op.nt - (n - t)) - t
            g = np.array([self.X, self.Z, self.P, op.t_left, op.t_right, op.t_up, op.t_down, op.nx,
                          op.nz, op.nt, TS, op.n_wavelet, op.n_receivers, phase, flags, src_j,
                          self.rec_row, n, rz, seg, interior, qmode, int(t % 2 == 0),
                          int(left > 1)], np.int32)
            fn(_p(g), _p(hs), ptrs, _p(self.recmap), _p(wav), _p(seis_t), _p(seis_rec))
        self.cur = op.nt & 1
        if pairs:
            self.last = (self.u[(op.nt - 1) & 3], self.u[op.nt & 3])
        if qmode and qmode != Q_I16 and phase == 1 and (op.nt - 1) % TS == 0:
            c = (op.nt - 1) // TS
            qs = q16[c]
            self.hx.hx_quantize_interior(_p(self.u[op.nt & 1]), C.c_int(self.P), C.c_int(op.t_left),
                                         C.c_int(op.t_up), C.c_int(op.nx), C.c_int(op.nz), _p(qs),
                                         _p(smax[c]), C.c_int(qmode))
        return snaps, I, Is, seis_rec

    def forward_tiled(self, shot_x, shot_z, wav, rz=4, record=False, interior=0, seg=0, qmode=0, block=0):
        """Forward pass; with qmode (1 FP16, 2 BF16, 3 int16 blocks) the
        snapshots are the 16-bit store and its max-abs slots (for int16
        blocks, the block exponents): returns (q16, smax), rec."""
        op = self.op
        src_j = (shot_z + op.t_up) * self.P + shot_x + op.t_left
        flags = 4 | (2 if record else 0)
        TS = op.time_slices
        q16 = np.zeros((op.nt // TS, op.nz, op.nx), np.uint16)
        if qmode == Q_I16:      # the block exponents take the max-slot entry of the table
            smax = np.zeros((op.nt // TS, op.nz, i16_blocks(op.t_left, op.nx)), np.int8)
        else:
            smax = np.zeros((op.nt // TS, SNAP_WORDS), np.uint32)
        snaps, _, _, rec = self._tiles_pass(1, flags, rz, src_j=src_j, wav=wav, record=record,
                                            interior=interior, seg=seg, qmode=qmode, q16=q16, smax=smax,
                                            block=block)
        if qmode:
            return (q16, smax), np.ascontiguousarray(rec.T)
        return snaps, np.ascontiguousarray(rec.T)

    def backward_tiled(self, seis, snaps, rz=4, kahan=False, interior=0, seg=0, qmode=0, block=0):
        flags = 8 | (1 if kahan else 0)
        if qmode:
            q16, smax = snaps
            _, I, Is, _ = self._tiles_pass(2, flags, rz, seis=seis, interior=interior, seg=seg, qmode=qmode,
                                           q16=q16, smax=smax)
        else:
            _, I, Is, _ = self._tiles_pass(2, flags, rz, seis=seis, snaps=snaps, interior=interior, seg=seg,
                                           block=block)
        return np.ascontiguousarray(I.T), np.ascontiguousarray(Is.T)
