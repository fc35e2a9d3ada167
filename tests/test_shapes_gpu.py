"""Standard-mode FP32 D3Q19 / D3Q27 steps on unusual shapes, against the
one-lattice batch (the persistent batch kernel: same bytes) and against the
oracle.  Shapes cover thin and long z-rows (nz = 4 to 260, tiles of 64
nodes spanning several rows or a part of one), periodic and walled z, a
moving wall, a body force, and runs of solid nodes beside fluid rows
(solid lanes load nothing; their fluid neighbours' bounced links must
still be right)."""
import sys

import numpy as np
import pytest

from latticeboltzmanncuda_amd import capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def block_solids(dims, pattern):
    # solid runs along z next to fluid rows, so that a fluid node's
    # z-neighbours are often solid while its x / y source rows are not
    nx, ny, nz = dims
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                          indexing="ij")
    if pattern == 0:
        s = ((z // 4) % 3 == 0) & ((x + y) % 2 == 0)
    else:
        s = ((z // 8) % 2 == 1) & (y % 3 != 1)
    return s.astype(np.uint8).reshape(-1)


CASES = [
    # q, dims, periodic, solids, lid, force
    (19, (8, 6, 4), (1, 1, 1), None, False, True),
    (19, (5, 3, 12), (0, 0, 0), None, True, False),
    (19, (3, 2, 260), (1, 1, 1), 0, False, True),
    (19, (7, 3, 20), (0, 1, 1), 1, False, True),
    (19, (4, 3, 256), (1, 0, 0), 1, True, False),
    (27, (6, 5, 8), (1, 0, 0), 0, False, True),
    (27, (3, 2, 260), (0, 1, 0), 1, False, True),
    (27, (5, 4, 12), (0, 0, 0), None, True, False),
    (19, (6, 5, 10), (1, 0, 1), 0, False, True),
]


@pytest.mark.parametrize("q,dims,per,pat,lid,force", CASES)
def test_step_matches_batch_and_oracle(q, dims, per, pat, lid, force):
    steps, tau = 23, 0.7
    say(f"case: step Q{q} {dims} periodic {per} solids {pat} "
        f"lid {lid} force {force} x {steps}")
    solid = None if pat is None else block_solids(dims, pat)
    kw = dict(tau=tau, periodic=per)
    okw = dict(tau=tau, periodic=per)
    lid_face = -1
    if lid:
        lid_face = 3
        kw.update(lid_face=3, lid_u=(0.05, 0, 0.02))
        okw.update(lid_face=3, lid_u=(0.05, 0, 0.02))
    if force:
        kw.update(force=(2e-5, 1e-6, -3e-6))
        okw.update(force=(2e-5, 1e-6, -3e-6))
    cfg = capi.make_config(*dims, lattice=q, **kw)
    with capi.Context(cfg) as c:
        c.upload_solid(solid)
        c.init_equilibrium()
        c.step(steps)
        alone = c.populations()
    with capi.Batch(capi.STORE_F32) as b:
        m = b.add(cfg, solid)
        b.step(steps, 0)
        b.synchronize()
        batched = m.populations()
    assert np.array_equal(alone, batched)
    mask = O.std_build_mask(q, dims, solid, per, lid_face)
    fo = O.std_run_f32(q, dims, O.std_init_eq(q, dims), mask,
                       O.std_params(**okw), steps, O.STORE_F32_SHIFTED)
    err = np.abs(alone - fo).max() / np.abs(fo).max()
    say(f"  identical to the batch kernel; max rel diff to oracle {err:.2e}")
    assert err < 1e-6
