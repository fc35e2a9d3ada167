"""C4 batching: lattices stepped together by one persistent kernel (plain
launches and hipGraph chunks) give the same populations as each lattice
stepped alone; admission honours the HBM budget; residuals are right."""
import sys

import numpy as np
import pytest

from latticeboltzmanncuda_amd import capi, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mixed_specs():
    rng = np.random.default_rng(11)
    specs = []
    # D2Q9 cavities and channels with squares
    for (nx, ny), kind in [((40, 33), "cavity"), ((97, 64), "channel"),
                           ((32, 200), "cavity"), ((150, 41), "channel")]:
        solid = workloads.random_squares(nx, ny, 0.1, rng)
        if kind == "cavity":
            cfg = capi.make_config(nx, ny, 1, lattice=9, tau=0.7, lid_face=3,
                                   lid_u=(0.1, 0, 0))
        else:
            cfg = capi.make_config(nx, ny, 1, lattice=9, tau=0.8,
                                   periodic=(1, 0, 0), force=(1e-5, 0, 0))
        specs.append((cfg, solid))
    # D3Q19 channel with spheres and a cavity
    solid, _ = capi.random_spheres((48, 24, 20), 3, 0.15, 3.0, 5.0, True)
    specs.append((capi.make_config(48, 24, 20, lattice=19, tau=0.8,
                                   periodic=(1, 0, 0), force=(1e-5, 0, 0)),
                  solid))
    specs.append((capi.make_config(20, 20, 20, lattice=19, tau=0.6,
                                   lid_face=3, lid_u=(0.1, 0, 0)), None))
    return specs


def run_alone(cfg, solid, steps):
    with capi.Context(cfg) as c:
        c.upload_solid(solid)
        c.init_equilibrium()
        c.step(steps)
        return c.populations()


@pytest.mark.parametrize("chunk", [0, 10])
def test_batch_equals_alone(chunk):
    say(f"case: batch of 6 mixed lattices, chunk={chunk}")
    specs = mixed_specs()
    steps = 37  # odd: exercises the buffer parity bookkeeping
    with capi.Batch(capi.STORE_F32) as b:
        members = [b.add(cfg, solid) for cfg, solid in specs]
        b.set_launch(128, 64)
        b.step(steps, chunk)
        b.synchronize()
        got = [m.populations() for m in members]
    for (cfg, solid), g in zip(specs, got):
        ref = run_alone(cfg, solid, steps)
        assert np.array_equal(g, ref)


def test_batch_launch_shapes_and_autotune_agree():
    say("case: batch launch shapes + autotune")
    specs = mixed_specs()
    outs = []
    for shape in [(64, 0), (256, 16), (512, 1000), None]:
        with capi.Batch(capi.STORE_F16) as b:
            members = [b.add(capi.make_config(
                cfg.nx, cfg.ny, cfg.nz, lattice=cfg.lattice,
                storage=capi.STORE_F16, tau=cfg.tau,
                periodic=tuple(cfg.periodic), lid_face=cfg.lid_face,
                lid_u=tuple(cfg.lid_u), force=tuple(cfg.force)), solid)
                for cfg, solid in specs]
            if shape is None:
                block, grid = b.autotune(steps=4, use_cache=False)
                say(f"  autotuned block {block} grid {grid}")
                # autotuning stepped the members: start again from rest
                for m in members:
                    m.init_equilibrium()
            else:
                b.set_launch(*shape)
            b.step(24, 12)
            b.synchronize()
            outs.append([m.populations() for m in members])
    for o in outs[1:]:
        for a, c in zip(outs[0], o):
            assert np.array_equal(a, c)


def test_batch_add_remove_between_chunks():
    say("case: admit / evict between graph chunks")
    specs = mixed_specs()
    with capi.Batch(capi.STORE_F32) as b:
        m0 = b.add(*specs[0])
        m1 = b.add(*specs[1])
        b.step(20, 10)
        m2 = b.add(*specs[2])  # joins late
        b.step(20, 10)
        b.remove(m1)           # leaves early
        b.step(20, 10)
        b.synchronize()
        g0, g2 = m0.populations(), m2.populations()
    assert np.array_equal(g0, run_alone(*specs[0], 60))
    assert np.array_equal(g2, run_alone(*specs[2], 40))


def test_batch_budget_admission():
    say("case: HBM budget admission")
    cfg = capi.make_config(64, 64, 1, lattice=9, tau=0.7)
    need = capi.Batch.estimate_bytes(cfg)
    with capi.Batch(capi.STORE_F32, budget_bytes=3 * need + 4096) as b:
        b.add(cfg)
        b.add(cfg)
        with pytest.raises(capi.LbmError) as e:
            for _ in range(3):
                b.add(cfg)
        assert e.value.code == 6  # LBM_ERR_BUDGET
        nlat, _, used = b.info()
        assert nlat >= 2 and used <= 3 * need + 4096


def test_batch_residuals_match_host():
    say("case: batch residuals vs host computation")
    specs = mixed_specs()[:3]
    with capi.Batch(capi.STORE_F32) as b:
        members = [b.add(cfg, solid, track_residual=True)
                   for cfg, solid in specs]
        b.step(50)
        r0 = b.residuals()  # against the zero snapshot: 1 for moving flow
        u_prev = [m.moments()[1].astype(np.float64) for m in members]
        b.step(20)
        r1 = b.residuals()
        for m, up, (cfg, solid) in zip(members, u_prev, specs):
            _, u = m.moments()
            fl = solid == 0
            d = (u.astype(np.float64) - up)[:, fl]
            ref = np.sqrt((d ** 2).sum() / (u.astype(np.float64)[:, fl] ** 2).sum())
            assert abs(r0[m.slot] - 1.0) < 1e-9
            assert abs(r1[m.slot] - ref) <= 1e-6 * ref + 1e-12


def test_batch_member_parameter_change_between_chunks():
    say("case: tau / force of a member changed between chunks")
    cfg, solid = mixed_specs()[1]  # D2Q9 channel with squares
    with capi.Batch(capi.STORE_F32) as b:
        m = b.add(cfg, solid)
        b.step(20, 10)
        m.set_tau(0.65)
        m.set_body_force((2e-5, 0.0, 0.0))
        b.step(20, 10)
        b.synchronize()
        got = m.populations()
    with capi.Context(capi.make_config(
            cfg.nx, cfg.ny, cfg.nz, lattice=cfg.lattice, tau=cfg.tau,
            periodic=tuple(cfg.periodic), force=tuple(cfg.force))) as c:
        c.upload_solid(solid)
        c.init_equilibrium()
        c.step(20)
        c.set_tau(0.65)
        c.set_body_force((2e-5, 0.0, 0.0))
        c.step(20)
        ref = c.populations()
    assert np.array_equal(got, ref)
