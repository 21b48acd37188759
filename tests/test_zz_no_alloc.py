"""rtm_batch_welford_noalloc (the Welford reduction with its partial
results in the library's fixed device scratch array) against
rtm_batch_welford, through the C ABI: bit-identical mean and variance on two
calls in a row, the second on a smaller batch (fewer realizations, a smaller
grid) whose results must not carry anything of the first; and a batch larger
than the scratch is refused with an error."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import common
import oracle as O
import rtmuq2d_amd as R

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _batch_images(name, nsims, seed, nt=None):
    cfg = dict(common.REF_CONFIGS[name])
    if nt:
        cfg["numberOfTimeStep"] = nt
        cfg["sampligWavelet"] = min(cfg["sampligWavelet"], nt)
    op = common.orc_params(cfg)
    p = common.rtm_params(cfg)
    sx, sz = common.shot_positions(cfg)[0]
    g = np.load(f"{common.GOLDEN}/{name}.npz")
    rng = np.random.default_rng(seed)
    base = g["models"][0]
    models = [base * np.float32(1 + 0.04 * k) + rng.uniform(-5, 5, base.shape).astype(np.float32)
              for k in range(nsims)]
    d = (rng.standard_normal((p.number_of_receivers, p.number_of_time_steps)) * 1e-2).astype(np.float32)
    return p, [O.extend(op, m) for m in models], g["wavelet"][:p.sampling_wavelet], d, sx, sz


def _both(name, nsims, seed):
    p, vpads, w, d, sx, sz = _batch_images(name, nsims, seed)
    with R.Batch(p, nsims) as b:
        for k, v in enumerate(vpads):
            b.set_velocity(k, v)
        b.set_wavelet(w)
        b.set_seismogram(d)
        b.zero_images()
        b.migrate(sx, sz)
        m0, v0 = b.welford()
        m1, v1 = b.welford_noalloc()
        images = np.stack([b.image(k)[0].astype(np.float64) for k in range(nsims)])
    return (m0, v0), (m1, v1), images


def test_welford_noalloc_equals_welford_on_two_calls_the_second_smaller(gpu_available):
    say("[gpu] welford_noalloc: tiny grid, 6 realizations")
    (m0, v0), (m1, v1), imgs = _both("tiny", 6, 21)
    assert np.abs(imgs).max() > 0 and np.abs(v0).max() > 0
    assert np.array_equal(m1, m0) and np.array_equal(v1, v0)
    say("[gpu] welford_noalloc: asym grid (smaller image), 3 realizations")
    (m0b, v0b), (m1b, v1b), imgs_b = _both("asym", 3, 22)
    assert m1b.shape != m1.shape and m1b.size < m1.size
    assert np.array_equal(m1b, m0b) and np.array_equal(v1b, v0b)
    assert np.allclose(m1b, imgs_b.mean(0), rtol=1e-13, atol=1e-13 * np.abs(imgs_b).max())


def test_welford_noalloc_refuses_a_batch_beyond_its_scratch(gpu_available):
    """1100 x 1000 = 1,100,000 image points exceed the 1,048,576 the scratch
    holds: the call fails with the limit in its message and writes nothing."""
    say("[gpu] welford_noalloc: a batch beyond the scratch")
    p = R.make_params(nx=1100, nz=1000, thickness_left=10, thickness_right=10, thickness_up=10, thickness_down=10,
                      spacing_grid=10.0, sampling=0.001, number_of_time_steps=4, sampling_wavelet=4,
                      first_location_shot=550, depth_shot=2, spacing_shot=1, number_of_shots=1,
                      first_location_receiver=0, depth_receiver=2, spacing_receiver=2, number_of_receivers=10,
                      number_of_simulations=1, zfp_tolerance=1e-4, time_slices=4)
    with R.Batch(p, 2) as b:
        m = np.full((p.nx, p.nz), 7.0)
        v = np.full((p.nx, p.nz), 7.0)
        rc = R.lib().rtm_batch_welford_noalloc(b.h, R.dptr(m), R.dptr(v))
        assert rc != 0
        assert "1048576" in R.lib().rtm_last_error().decode()
        assert (m == 7.0).all() and (v == 7.0).all()
