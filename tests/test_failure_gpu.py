"""Failure handling on the GPU: a lattice that turns non-finite or
diverges is detected by the residual reduction and retired on its own --
the other lattices of the batch or server are unaffected, bit for bit --
and the admission gates reject what is known to be unstable.  Each case
forces the condition from the host (populations or parameters), so every
kernel still runs to its normal end."""
import sys

import numpy as np
import pytest

from latticeboltzmanncuda_amd import capi, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def stable_requests(n=5, seed=21):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        if k % 2:
            L = int(rng.integers(12, 20))
            solid, _ = capi.random_spheres((L, L, L), 50 + k, 0.08, 2.0, 3.0,
                                           True)
            cfg = capi.make_config(L, L, L, lattice=19, tau=0.8,
                                   periodic=(1, 0, 0), force=(1e-5, 0, 0))
        else:
            nx, ny = (int(v) for v in rng.integers(24, 64, size=2))
            solid = workloads.random_squares(nx, ny, 0.1, rng)
            cfg = capi.make_config(nx, ny, 1, lattice=9, tau=0.7,
                                   lid_face=3, lid_u=(0.05, 0, 0))
        out.append({"cfg": cfg, "solid": solid, "arrival_s": 0.0,
                    "max_steps": 300, "tol": 0.0})
    return out


def unstable_requests():
    # a cavity at U / nu = 2400 (tau 0.5005, lid 0.4): BGK blows up within
    # a few hundred steps; a channel driven by a body force of 3e38 per
    # node (finite, near the FP32 maximum), whose populations overflow to
    # Inf and then NaN within a few steps
    cav = capi.make_config(48, 48, 1, lattice=9, tau=0.5005, lid_face=3,
                           lid_u=(0.4, 0, 0))
    nan = capi.make_config(32, 24, 1, lattice=9, tau=0.8, periodic=(1, 0, 0),
                           force=(3e38, 0, 0))
    return [{"cfg": cav, "solid": None, "arrival_s": 0.0, "max_steps": 2000,
             "tol": 0.0},
            {"cfg": nan, "solid": None, "arrival_s": 0.0, "max_steps": 2000,
             "tol": 0.0}]


def plain_u(r, steps):
    with capi.Context(r["cfg"]) as c:
        c.upload_solid(r["solid"])
        c.init_equilibrium()
        c.step(steps)
        return c.moments()[1]


def test_serving_retires_diverged_requests_alone():
    say("case: serve 5 stable + 2 diverging requests (FP32)")
    good, bad = stable_requests(), unstable_requests()
    reqs = good[:2] + [bad[0]] + good[2:4] + [bad[1]] + good[4:]
    is_bad = [any(r is x for x in bad) for r in reqs]
    stats, res = capi.serve(reqs, storage=capi.STORE_F32, chunk=100,
                            keep_results=True)
    say(f"  stats {stats}")
    say(f"  status {[r['status'] for r in res]} steps "
        f"{[r['steps'] for r in res]}")
    for r, out, b in zip(reqs, res, is_bad):
        if b:
            assert out["status"] == capi.SERVE_DIVERGED
            assert out["steps"] < r["max_steps"]
            u = out["u"]
            assert (not np.isfinite(u).all()) or np.abs(u).max() > 0.5
        else:
            assert out["status"] == capi.SERVE_STEP_LIMIT
            assert np.array_equal(out["u"], plain_u(r, out["steps"]))
    assert stats["diverged"] == 2
    assert stats["completed"] == len(reqs)
    # the stability gate rejects both before they run; the stable requests
    # give the same bytes again
    stats2, res2 = capi.serve(reqs, storage=capi.STORE_F32, chunk=100,
                              keep_results=True,
                              max_cell_re=workloads.C4_MAX_CELL_RE)
    say(f"  gated stats {stats2}")
    assert stats2["unstable"] == 2 and stats2["diverged"] == 0
    for a, b, bd in zip(res, res2, is_bad):
        if bd:
            assert b["status"] == capi.SERVE_UNSTABLE and b["steps"] == 0
        else:
            assert b["status"] == a["status"]
            assert np.array_equal(a["u"], b["u"])


def test_batch_residual_flags_nan_member():
    say("case: batch of 3 with NaN populations in the middle member")
    cfgs = [capi.make_config(40, 30, 1, lattice=9, tau=0.7, lid_face=3,
                             lid_u=(0.05, 0, 0)),
            capi.make_config(24, 20, 16, lattice=19, tau=0.8,
                             periodic=(1, 0, 0), force=(1e-5, 0, 0)),
            capi.make_config(36, 28, 1, lattice=9, tau=0.9,
                             periodic=(1, 0, 0), force=(1e-5, 0, 0))]
    with capi.Batch(capi.STORE_F32) as b:
        ms = [b.add(c, None, track_residual=True) for c in cfgs]
        f = ms[1].populations()
        f[5 * ms[1].n + 17] = np.nan
        ms[1].set_populations(f)
        b.step(100, 50)
        res = b.residuals()
        got = [m.populations() for m in ms]
    say(f"  residuals {res}")
    assert np.isnan(res[ms[1].slot])
    for k in (0, 2):
        assert np.isfinite(res[ms[k].slot]) and res[ms[k].slot] >= 0
        with capi.Context(cfgs[k]) as c:
            c.init_equilibrium()
            c.step(100)
            assert np.array_equal(got[k], c.populations())


def test_batch_residual_flags_speed_above_limit():
    # finite but far too fast: a speed above LBM_DIVERGED_SPEED counts as
    # diverged
    say("case: batch member with |u| = 0.6 everywhere")
    cfg = capi.make_config(32, 32, 1, lattice=9, tau=0.8,
                           periodic=(1, 1, 1))
    with capi.Batch(capi.STORE_F32) as b:
        m = b.add(cfg, None, track_residual=True)
        u = np.zeros((3, m.n))
        u[0] = 0.6
        m.init_equilibrium(np.ones(m.n), u)
        b.step(2, 2)
        res = b.residuals()
    assert np.isnan(res[m.slot])


def test_batch_stability_limit_rejects_at_admission():
    say("case: batch admission with a stability limit")
    ok = capi.make_config(64, 64, 1, lattice=9, tau=0.6, lid_face=3,
                          lid_u=(0.1, 0, 0))          # U / nu = 3
    hot = capi.make_config(64, 64, 1, lattice=9, tau=0.51, lid_face=3,
                           lid_u=(0.1, 0, 0))         # U / nu = 30
    with capi.Batch(capi.STORE_F32) as b:
        b.set_stability_limit(15.0)
        b.add(ok)
        with pytest.raises(capi.LbmError) as e:
            b.add(hot)
        assert e.value.code == capi.ERR_UNSTABLE
        assert b.info()[0] == 1
        b.set_stability_limit(0.0)  # no limit: admitted
        b.add(hot)
        assert b.info()[0] == 2
