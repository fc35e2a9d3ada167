"""Analytic flows on the GPU: body-force Poiseuille channel and moving-wall
Couette flow at steady state, D2Q9 and D3Q19, FP32 and FP64 storage.  The
closed forms hold for half-way bounce-back walls; Poiseuille is exact for
BGK at tau = 1/2 + sqrt(3/16) (the wall slip vanishes there)."""
import sys

import numpy as np
import pytest

from latticeboltzmanncuda_amd import capi

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dims(q, ny):
    # periodic in x (and z for D3Q19), walls half a link outside y = 0 and
    # y = ny - 1
    return (4, ny, 1) if q == 9 else (4, ny, 4)


def _profile(u, dims):
    # u_x averaged over x and z, as a function of y (node index
    # z + nz (y + ny x))
    nx, ny, nz = dims
    return u[0].reshape(nx, ny, nz).mean(axis=(0, 2))


@pytest.mark.parametrize("storage", [capi.STORE_F64, capi.STORE_F32])
@pytest.mark.parametrize("q", [9, 19])
def test_poiseuille_channel(q, storage):
    # u(y) = F / (2 rho nu) (y + 1/2) (ny - 1/2 - y)
    ny, F = 16, 1e-6
    tau = 0.5 + np.sqrt(3.0 / 16.0)
    nu = (tau - 0.5) / 3
    dims = _dims(q, ny)
    say(f"case: Poiseuille Q{q} {dims} storage {storage} x 6000")
    per = (1, 0, 0) if q == 9 else (1, 0, 1)
    cfg = capi.make_config(*dims, lattice=q, storage=storage, tau=float(tau),
                           periodic=per, force=(F, 0, 0))
    with capi.Context(cfg) as c:
        c.init_equilibrium()
        c.step(6000)
        rho, u = c.moments()
    y = np.arange(ny)
    ua = F / (2 * nu) * (y + 0.5) * (ny - 0.5 - y)
    prof = _profile(u, dims)
    err = np.abs(prof - ua).max() / ua.max()
    say(f"  max rel err {err:.2e}, max |u_y| {np.abs(u[1]).max():.2e}")
    if storage == capi.STORE_F64:
        # exact at this tau: only the convergence to steady state remains
        assert err < 1e-6
        assert np.abs(u[1]).max() < 1e-9
    else:
        # FP32 storage of f_i - w_i (a few 1e-4 here, collided from small
        # terms): 2-7e-5 of the profile after 6000 steps (storing f_i
        # itself, ~0.1 against increments of ~1e-9, left ~1.5e-3)
        assert err < 2e-4
        assert np.abs(u[1]).max() < 1e-7
    assert np.abs(rho - 1).max() < 1e-5


@pytest.mark.parametrize("storage", [capi.STORE_F64, capi.STORE_F32])
@pytest.mark.parametrize("q", [9, 19])
def test_couette_moving_wall(q, storage):
    # still wall at y = -1/2, wall moving at U at y = ny - 1/2:
    # u(y) = U (y + 1/2) / ny
    ny, U = 12, 0.01
    dims = _dims(q, ny)
    say(f"case: Couette Q{q} {dims} storage {storage} x 8000")
    per = (1, 0, 0) if q == 9 else (1, 0, 1)
    cfg = capi.make_config(*dims, lattice=q, storage=storage, tau=0.8,
                           periodic=per, lid_face=3, lid_u=(U, 0, 0))
    with capi.Context(cfg) as c:
        c.init_equilibrium()
        c.step(8000)
        _, u = c.moments()
    ua = U * (np.arange(ny) + 0.5) / ny
    prof = _profile(u, dims)
    err = np.abs(prof - ua).max() / U
    say(f"  max err / U {err:.2e}")
    # FP64: exact up to the convergence to steady state; FP32: rounding of
    # the stored deviations f_i - w_i (2-4e-6 measured)
    assert err < (1e-6 if storage == capi.STORE_F64 else 2e-5)
