"""lbm_moments_noalloc -- the allocation-free variant of lbm_moments -- gives
the same bytes as lbm_moments through the C-ABI, on consecutive calls with
a smaller lattice second (the shared device array keeps the first call's
larger contents), and refuses lattices beyond its fixed scratch array."""
import ctypes as C
import sys

import numpy as np
import pytest

from latticeboltzmanncuda_amd import capi, workloads

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc(ctx, want_rho=True, want_u=True):
    rho = np.full(ctx.n, np.nan, np.float32) if want_rho else None
    u = np.full(3 * ctx.n, np.nan, np.float32) if want_u else None
    capi.check(capi.load().lbm_moments_noalloc(
        ctx.h, capi._ptr(rho), capi._ptr(u)), "lbm_moments_noalloc")
    return rho, None if u is None else u.reshape(3, ctx.n)


def test_noalloc_moments_equal_moments_on_consecutive_calls():
    # a standard channel (FP16 storage), then a smaller cavity (FP32), then
    # a reference-mode lattice: each against lbm_moments on the same state
    say("case: lbm_moments_noalloc vs lbm_moments, 3 lattices, shrinking")
    solid, _ = capi.random_spheres((96, 64, 48), 3, 0.15, 3.0, 6.0, True)
    big = capi.make_config(96, 64, 48, lattice=19, storage=capi.STORE_F16,
                           tau=0.8, periodic=(1, 0, 0), force=(1e-5, 0, 0))
    small = capi.make_config(40, 24, 1, lattice=9, tau=0.7, lid_face=3,
                             lid_u=(0.05, 0, 0))
    ref = capi.make_config(20, 16, 12, lattice=27, mode=capi.MODE_REFERENCE,
                           tau=0.8)
    rng = np.random.default_rng(4)
    for cfg, sol in ((big, solid), (small, None), (ref, None)):
        with capi.Context(cfg) as c:
            if cfg.mode == capi.MODE_REFERENCE:
                c.set_populations((0.001 + 0.001 * rng.random(27 * c.n))
                                  .astype(np.float32))
                c.step(3)
            else:
                c.upload_solid(sol)
                c.init_equilibrium()
                c.step(25)
            want_rho, want_u = c.moments()
            for wr, wu in ((True, True), (False, True), (True, False)):
                rho, u = noalloc(c, wr, wu)
                if wr:
                    assert np.array_equal(rho, want_rho)
                if wu:
                    assert np.array_equal(u, want_u)
            say(f"  {cfg.nx}x{cfg.ny}x{cfg.nz} Q{cfg.lattice}: equal")


def test_noalloc_moments_rejects_lattices_beyond_scratch():
    # 4,194,304 nodes fit rho + u (16 B per node); one more z-row does not,
    # but its u alone (12 B per node) still fits
    say("case: lbm_moments_noalloc size limit")
    fits = capi.make_config(256, 128, 128, lattice=19, storage=capi.STORE_F16,
                            tau=0.8, periodic=(1, 1, 1))
    over = capi.make_config(257, 128, 128, lattice=19, storage=capi.STORE_F16,
                            tau=0.8, periodic=(1, 1, 1))
    with capi.Context(fits) as c:
        c.init_equilibrium()
        assert 16 * c.n == SCRATCH_BYTES
        rho, u = noalloc(c)
        want_rho, want_u = c.moments()
        assert np.array_equal(rho, want_rho) and np.array_equal(u, want_u)
    with capi.Context(over) as c:
        c.init_equilibrium()
        with pytest.raises(capi.LbmError) as e:
            noalloc(c)
        assert e.value.code == capi.ERR_INVALID
        assert "maximum" in str(e.value)
        _, u = noalloc(c, want_rho=False)
        assert np.array_equal(u, c.moments()[1])
