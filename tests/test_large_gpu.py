"""Lattices large enough to take the streaming-store path (both population
copies above 256 MB: non-temporal stores, alternating tile order) against
the oracle, FP32 and FP16 storage."""
import os
import sys

import numpy as np
import pytest

from latticeboltzmanncuda_amd import capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_streaming_store_d3q19_f32_channel():
    # 128 x 128 x 112 D3Q19 FP32: 2 x 1.84 M x 19 x 4 B = 279 MB
    dims, q, steps = (128, 128, 112), 19, 6
    say(f"case: streaming stores D3Q19 F32 {dims} x {steps}")
    solid, _ = capi.random_spheres(dims, 2, 0.2, 4.0, 8.0, True)
    kw = dict(tau=0.8, periodic=(1, 0, 0), force=(1e-5, 0, 0))
    with capi.Context(capi.make_config(*dims, lattice=q, **kw)) as c:
        assert 2 * c.device_bytes() > 0
        c.upload_solid(solid)
        c.init_equilibrium()
        c.step(steps)
        fg = c.populations(np.float64)
    mask = O.std_build_mask(q, dims, solid, (1, 0, 0))
    p = O.std_params(tau=0.8, force=(1e-5, 0, 0), periodic=(1, 0, 0))
    fo = O.std_run_f32(q, dims, O.std_init_eq(q, dims), mask, p, steps,
                       O.STORE_F32_SHIFTED)
    err = np.abs(fg - fo).max() / np.abs(fo).max()
    say(f"  max rel diff {err:.2e}")
    assert err < 1e-6


def test_streaming_store_d2q9_f16_cavity():
    # 2800 x 2800 D2Q9 FP16: 2 x 7.84 M x 9 x 2 B = 282 MB, which takes the
    # group step (two nodes per thread); the one-node step gives the same
    # bytes
    dims, q, steps = (2800, 2800, 1), 9, 6
    say(f"case: streaming stores D2Q9 F16 {dims} x {steps}")
    kw = dict(tau=0.6, lid_face=3, lid_u=(0.1, 0, 0))
    cfg = capi.make_config(*dims, lattice=q, storage=capi.STORE_F16, **kw)
    with capi.Context(cfg) as c:
        c.init_equilibrium()
        c.step(steps)
        _, ug = c.moments()
        fg = c.populations()
    os.environ["LBM_GROUP_STEP"] = "0"
    try:
        with capi.Context(cfg) as c:
            c.init_equilibrium()
            c.step(steps)
            f1 = c.populations()
    finally:
        del os.environ["LBM_GROUP_STEP"]
    assert np.array_equal(fg, f1)
    mask = O.std_build_mask(q, dims, None, (0, 0, 0), 3)
    p = O.std_params(tau=0.6, lid_face=3, lid_u=(0.1, 0, 0))
    fo = O.std_run_f32(q, dims, O.std_init_eq(q, dims), mask, p, steps,
                       O.STORE_F16_SHIFTED)
    _, uo = O.std_moments(q, dims, fo, mask)
    du = np.abs(ug[0] - uo[0]).max() / np.abs(uo[0]).max()
    say(f"  u rel diff {du:.2e}")
    assert du < 2e-3
