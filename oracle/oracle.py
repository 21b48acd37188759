"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU restatement of rtmUq2D's RTM hot path (rtm_oracle_impl.c).  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it,
and only as the checker, never as the thing measured or shipped.

All arrays use the reference's layouts: padded fields [X][Z] (z fastest),
images [Nx][Nz], seismograms [nrec][nt].
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("nz", C.c_int),
        ("t_left", C.c_int), ("t_right", C.c_int), ("t_up", C.c_int), ("t_down", C.c_int),
        ("h", C.c_float), ("dt", C.c_float),
        ("nt", C.c_int), ("n_wavelet", C.c_int),
        ("first_receiver", C.c_int), ("depth_receiver", C.c_int),
        ("spacing_receiver", C.c_int), ("n_receivers", C.c_int),
        ("time_slices", C.c_int),
        ("image_kahan", C.c_int),      # 1: Kahan-compensated image sums (an option of this build)
    ]

    @property
    def X(self):
        return self.t_left + self.nx + self.t_right

    @property
    def Z(self):
        return self.t_up + self.nz + self.t_down


def from_rtm(p) -> OrcParams:
    """OrcParams from an rtmuq2d_amd.RtmParams-like object."""
    return OrcParams(p.nx, p.nz, p.thickness_left, p.thickness_right, p.thickness_up,
                     p.thickness_down, p.spacing_grid, p.sampling, p.number_of_time_steps,
                     p.sampling_wavelet, p.first_location_receiver, p.depth_receiver,
                     p.spacing_receiver, p.number_of_receivers, p.time_slices,
                     int(getattr(p, "image_kahan", 0)))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run oracle/build.py")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _real(dtype):
    return "_f64" if np.dtype(dtype) == np.float64 else "_f32"


def coef(p: OrcParams, vpad, dtype=np.float32):
    out = np.zeros((p.X, p.Z), dtype)
    getattr(lib(), "orc_coef" + _real(dtype))(C.byref(p), _p(out), _p(np.ascontiguousarray(vpad, np.float32)))
    return out


def profiles(p: OrcParams, dtype=np.float32):
    """(aL, aR, aT, aB, bL, bR, bT, bB) as attenuation_profile_cpml."""
    arrs = [np.zeros(n, dtype) for n in (p.t_left, p.t_right, p.t_up, p.t_down) * 2]
    getattr(lib(), "orc_profiles" + _real(dtype))(C.byref(p), *[_p(a) for a in arrs])
    return arrs


def stencil(p: OrcParams, Cf, u, U):
    """One interior step in place on U (copy returned)."""
    U = np.array(U, copy=True)
    getattr(lib(), "orc_stencil" + _real(U.dtype))(C.byref(p), _p(Cf), _p(u), _p(U))
    return U


def cpml(p: OrcParams, u, U, Cf, prof, psi, csi):
    """One CPML update.  prof = (aL,aR,aT,aB,bL,bR,bT,bB); psi, csi =
    [L,R,T,B] lists (modified in place).  Returns the corrected U."""
    U = np.array(U, copy=True)
    aL, aR, aT, aB, bL, bR, bT, bB = prof
    getattr(lib(), "orc_cpml" + _real(U.dtype))(
        C.byref(p), _p(u), _p(U), _p(Cf), _p(aL), _p(bL), _p(aR), _p(bR), _p(aT), _p(bT), _p(aB),
        _p(bB), *[_p(a) for a in psi], *[_p(a) for a in csi])
    return U


def image(p: OrcParams, I, Is, S, R):
    getattr(lib(), "orc_image" + _real(I.dtype))(C.byref(p), _p(I), _p(Is), _p(S), _p(R))


def forward(p: OrcParams, shot_x, shot_z, vpad, wavelet, dtype=np.float32, snaps=True, record=True):
    """Returns (snapshots [nt/TS][Nx][Nz] or None, seismogram [nrec][nt] or
    None, u^{nt-1}, u^{nt})."""
    ns = p.nt // p.time_slices
    S = np.zeros((ns, p.nx, p.nz), dtype) if snaps else None
    d = np.zeros((p.n_receivers, p.nt), dtype) if record else None
    lu = np.zeros((p.X, p.Z), dtype)
    lU = np.zeros((p.X, p.Z), dtype)
    rc = getattr(lib(), "orc_forward" + _real(dtype))(
        C.byref(p), int(shot_x), int(shot_z), _p(np.ascontiguousarray(vpad, np.float32)),
        _p(np.ascontiguousarray(wavelet, np.float32)), _p(S), _p(d), _p(lu), _p(lU))
    assert rc == 0, rc
    return S, d, lu, lU


def backward(p: OrcParams, vpad, seis, snaps, I=None, Is=None):
    dtype = snaps.dtype
    I = np.zeros((p.nx, p.nz), dtype) if I is None else I
    Is = np.zeros((p.nx, p.nz), dtype) if Is is None else Is
    lu = np.zeros((p.X, p.Z), dtype)
    lU = np.zeros((p.X, p.Z), dtype)
    rc = getattr(lib(), "orc_backward" + _real(dtype))(
        C.byref(p), _p(np.ascontiguousarray(vpad, np.float32)),
        _p(np.ascontiguousarray(seis, np.float32)), _p(snaps), _p(I), _p(Is), _p(lu), _p(lU))
    assert rc == 0, rc
    return I, Is, lu, lU


def rtm_shot(p: OrcParams, shot_x, shot_z, vpad, wavelet, seis, dtype=np.float32):
    I = np.zeros((p.nx, p.nz), dtype)
    Is = np.zeros((p.nx, p.nz), dtype)
    rc = getattr(lib(), "orc_rtm_shot" + _real(dtype))(
        C.byref(p), int(shot_x), int(shot_z), _p(np.ascontiguousarray(vpad, np.float32)),
        _p(np.ascontiguousarray(wavelet, np.float32)), _p(np.ascontiguousarray(seis, np.float32)),
        _p(I), _p(Is))
    assert rc == 0, rc
    return I, Is


def extend(p: OrcParams, model):
    out = np.zeros((p.X, p.Z), np.float32)
    lib().orc_extend(C.byref(p), _p(out), _p(np.ascontiguousarray(model, np.float32)))
    return out


def laplacian(p: OrcParams, img):
    out = np.zeros((p.nx, p.nz), np.float32)
    lib().orc_laplacian(C.byref(p), _p(out), _p(np.ascontiguousarray(img, np.float32)))
    return out
