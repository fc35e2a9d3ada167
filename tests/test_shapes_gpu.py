"""Standard-mode steps on unusual shapes -- FP32 D3Q19 / D3Q27, 16-bit
D3Q19 with and without solid nodes, FP32 D2Q9, and D2Q9
with 16-bit storage run by the group step (four nodes per thread where ny
is a multiple of 4, two where it is even; LBM_GROUP_STEP=1 selects it below
the size where it is chosen by itself) -- against the one-node step
(LBM_GROUP_STEP=0) and the one-lattice batch (same bytes), and against the
oracle.  Shapes cover thin and long z-rows (nz = 4 to 260, tiles of 64
nodes spanning several rows or a part of one), periodic and walled z, a
moving wall, a body force, periodic axes in every combination the wrap
code distinguishes, and runs of solid nodes beside fluid rows
(solid lanes load nothing with FP32 storage and gather like fluid lanes
with 16-bit storage; either way their fluid neighbours' bounced links
must still be right, and waves with and without solid nodes take
different code)."""
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


def block_solids(dims, pattern):
    # solid runs along the fastest axis next to fluid rows, so that a fluid
    # node's row neighbours are often solid while its other source rows are
    # not
    nx, ny, nz = dims
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                          indexing="ij")
    if pattern == 0:
        s = ((z // 4) % 3 == 0) & ((x + y) % 2 == 0)
    else:
        s = ((z // 8) % 2 == 1) & (y % 3 != 1)
    if nz == 1:  # D2Q9: rows along y
        s = ((y // 2) % 3 == 0) & (x % 2 == pattern)
    return s.astype(np.uint8).reshape(-1)


F32, F16, BF16 = capi.STORE_F32, capi.STORE_F16, capi.STORE_BF16
CASES = [
    # q, dims, periodic, solids, lid, force, storage
    (19, (8, 6, 4), (1, 1, 1), None, False, True, F32),
    (19, (5, 3, 12), (0, 0, 0), None, True, False, F32),
    (19, (3, 2, 260), (1, 1, 1), 0, False, True, F32),
    (19, (7, 3, 20), (0, 1, 1), 1, False, True, F32),
    (19, (4, 3, 256), (1, 0, 0), 1, True, False, F32),
    (27, (6, 5, 8), (1, 0, 0), 0, False, True, F32),
    (27, (3, 2, 260), (0, 1, 0), 1, False, True, F32),
    (27, (5, 4, 12), (0, 0, 0), None, True, False, F32),
    (19, (6, 5, 10), (1, 0, 1), 0, False, True, F32),
    (9, (64, 48, 1), (1, 0, 0), 0, False, True, F16),
    (9, (40, 130, 1), (0, 0, 0), None, True, False, F16),
    (9, (33, 66, 1), (1, 1, 0), 1, False, True, BF16),
    (9, (17, 258, 1), (1, 1, 0), 0, False, False, F16),
    (9, (24, 6, 1), (0, 0, 0), 1, True, True, F16),
    (9, (30, 31, 1), (1, 0, 0), 0, False, True, F16),  # odd ny
    # ny % 4 == 0: four nodes per thread
    (9, (16, 260, 1), (1, 1, 0), 1, False, True, F16),
    (9, (12, 8, 1), (0, 0, 0), None, True, True, F16),
    (9, (20, 132, 1), (0, 0, 0), 0, True, False, BF16),
    (9, (8, 64, 1), (1, 0, 0), 1, False, True, BF16),
    # periodic along y only (the wrap decodes y; walled x ends pull
    # unwrapped addresses that the bounce replaces), odd and even ny
    (9, (20, 34, 1), (0, 1, 0), 0, False, True, F16),
    (9, (18, 25, 1), (0, 1, 0), None, False, True, BF16),
    # FP32 D2Q9 (its step reads the 16-bit link masks too)
    (9, (22, 30, 1), (1, 0, 0), 1, True, True, F32),
    (9, (16, 20, 1), (0, 1, 0), 0, False, True, F32),
    # 3-D 16-bit lattices with solid nodes (every node reads its link
    # mask) and without (class map first)
    (19, (6, 5, 70), (1, 0, 0), 1, False, True, F16),
    (19, (7, 4, 66), (1, 0, 1), 0, True, False, BF16),
    (19, (5, 6, 40), (1, 1, 1), None, False, True, F16),
    # FP32 z-rows of whole tiles (a wavefront is one tile of one row): D3Q27
    # with a moving wall, periodic y, periodic z over three tiles, walls
    (27, (4, 3, 128), (1, 0, 0), 1, True, False, F32),
    (19, (3, 5, 128), (0, 1, 0), 0, False, True, F32),
    (19, (2, 3, 192), (0, 0, 1), None, False, True, F32),
    (19, (3, 4, 64), (0, 0, 0), 1, True, True, F32),
    # axes of one node: periodic (the node is its own neighbour) and
    # walled (every link along the axis bounces)
    (9, (1, 40, 1), (1, 1, 0), None, False, True, F16),
    (9, (7, 1, 1), (1, 0, 0), None, False, True, F16),
    (19, (1, 1, 70), (1, 1, 1), None, False, True, F32),
    (19, (2, 1, 9), (0, 1, 0), 0, False, True, BF16),
]


@pytest.mark.parametrize("q,dims,per,pat,lid,force,storage", CASES)
def test_step_matches_batch_and_oracle(q, dims, per, pat, lid, force,
                                       storage):
    steps, tau = 23, 0.7
    say(f"case: step Q{q} {dims} periodic {per} solids {pat} "
        f"lid {lid} force {force} storage {storage} x {steps}")
    solid = None if pat is None else block_solids(dims, pat)
    kw = dict(tau=tau, periodic=per)
    okw = dict(tau=tau, periodic=per)
    lid_face = -1
    uz = 0.0 if q == 9 else 1.0  # D2Q9: in-plane wall velocity and force
    if lid:
        lid_face = 3
        kw.update(lid_face=3, lid_u=(0.05, 0, 0.02 * uz))
        okw.update(lid_face=3, lid_u=(0.05, 0, 0.02 * uz))
    if force:
        kw.update(force=(2e-5, 1e-6, -3e-6 * uz))
        okw.update(force=(2e-5, 1e-6, -3e-6 * uz))
    cfg = capi.make_config(*dims, lattice=q, storage=storage, **kw)

    def run_alone():
        with capi.Context(cfg) as c:
            c.upload_solid(solid)
            c.init_equilibrium()
            c.step(steps)
            return c.populations()

    try:
        os.environ["LBM_GROUP_STEP"] = "1"  # the group step where it applies
        alone = run_alone()
        os.environ["LBM_GROUP_STEP"] = "0"  # one node per thread everywhere
        one_node = run_alone()
    finally:
        del os.environ["LBM_GROUP_STEP"]
    assert np.array_equal(alone, one_node)
    with capi.Batch(storage) as b:
        m = b.add(cfg, solid)
        b.step(steps, 0)
        b.synchronize()
        batched = m.populations()
    assert np.array_equal(alone, batched)
    mask = O.std_build_mask(q, dims, solid, per, lid_face)
    mode = {F32: O.STORE_F32_SHIFTED, F16: O.STORE_F16_SHIFTED,
            BF16: O.STORE_BF16_SHIFTED}[storage]
    fo = O.std_run_f32(q, dims, O.std_init_eq(q, dims), mask,
                       O.std_params(**okw), steps, mode)
    err = np.abs(alone - fo).max() / np.abs(fo).max()
    say(f"  same bytes one node per thread and in a batch; max rel diff to "
        f"oracle {err:.2e}")
    # FP32: rounding only; 16-bit: the oracle emulates the storage, a
    # result one 16-bit step apart where the FP32 collisions round
    # differently
    assert err < (1e-6 if storage == F32 else 2e-3)
