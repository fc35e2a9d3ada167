"""Snapshots (lbm_snapshot_save / lbm_snapshot_load): a run interrupted
by save + load into a fresh context continues bit for bit like the
uninterrupted run, in standard mode (FP32, FP16) and in reference mode
(flags, reflection records, flow injection, force register)."""
import os
import sys

import numpy as np
import pytest

from latticeboltzmanncuda_amd import capi, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "dataFiles")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("q,dims,storage", [
    (19, (40, 24, 20), capi.STORE_F32),
    (9, (96, 64, 1), capi.STORE_F16),
    (19, (24, 20, 16), capi.STORE_BF16),
])
def test_standard_snapshot_resumes_bit_exact(tmp_path, q, dims, storage):
    say(f"case: snapshot standard D{3 if q != 9 else 2}Q{q} {dims} storage {storage}")
    if q == 9:
        solid = workloads.random_squares(dims[0], dims[1], 0.1,
                                         np.random.default_rng(3))
        kw = dict(tau=0.7, lid_face=3, lid_u=(0.08, 0, 0))
    else:
        solid, _ = capi.random_spheres(dims, 5, 0.15, 2.0, 4.0, True)
        kw = dict(tau=0.8, periodic=(1, 0, 0), force=(2e-5, 0, 0))
    cfg = capi.make_config(*dims, lattice=q, storage=storage, **kw)
    path = tmp_path / "snap.bin"
    with capi.Context(cfg) as a:
        a.upload_solid(solid)
        a.init_equilibrium()
        a.step(31)  # odd: the current buffer is the second one
        a.save(path)
        a.step(20)
        want = a.populations()
    # a context created with other parameters takes the snapshot's
    other = capi.make_config(*dims, lattice=q, storage=storage, tau=1.3)
    with capi.Context(other) as b:
        b.load(path)
        assert abs(b.cfg.tau - kw["tau"]) < 1e-7
        b.step(20)
        got = b.populations()
    assert np.array_equal(got, want)


def test_reference_snapshot_resumes_bit_exact(tmp_path):
    say("case: snapshot reference mode 24^3 cube_angled, flow, force")
    dims, spacing, tau = (24, 24, 24), 0.05, 0.6
    mesh = O.Mesh(os.path.join(DATA, "cube_angled.obj"))
    ok, internal = mesh.insert(dims, spacing)
    assert ok
    flags, recs = mesh.preprocess(dims, spacing, internal * O.FLAG_INTERNAL)
    n = int(np.prod(dims))
    flow = np.full(27, 0.001, np.float32)
    flow[6] = 0.002
    f = np.tile(flow[:, None], (1, n)).reshape(-1).copy()
    cfg = capi.make_config(*dims, lattice=27, mode=capi.MODE_REFERENCE,
                           tau=tau)
    path = tmp_path / "ref.bin"
    with capi.Context(cfg) as a:
        a.set_populations(f)
        a.upload_flags(flags)
        a.upload_reflections(*recs)
        a.set_flow(capi.FLOW_PLANE, [0, 0], flow)
        a.set_force([0, 0, 0])
        a.step_reference(15, with_force=True)
        a.save(path)
        a.step_reference(10, with_force=True)
        want, want_F = a.populations(), a.force()
    with capi.Context(cfg) as b:
        b.load(path)
        assert np.array_equal(b.flags(), flags)
        b.step_reference(10, with_force=True)
        got, got_F = b.populations(), b.force()
    assert np.array_equal(got, want)
    assert np.array_equal(got_F, want_F)


def test_snapshot_rejects_other_lattices(tmp_path):
    say("case: snapshot into a different lattice")
    path = tmp_path / "s.bin"
    with capi.Context(capi.make_config(16, 16, 16, lattice=19)) as a:
        a.init_equilibrium()
        a.save(path)
    with capi.Context(capi.make_config(16, 16, 8, lattice=19)) as b:
        b.init_equilibrium()
        before = b.populations()
        with pytest.raises(capi.LbmError):
            b.load(path)
        assert np.array_equal(b.populations(), before)
    with open(path, "r+b") as fh:  # truncated file
        fh.truncate(200)
    with capi.Context(capi.make_config(16, 16, 16, lattice=19)) as c:
        with pytest.raises(capi.LbmError):
            c.load(path)


# Snapshot header (lbm_snapshot.hip SnapHeader, x86-64): magic 8 B, version
# and header size (u32 each), lbm_config (72 B) at 16, steps at 88, then the
# section sizes pop / mask / cls / flag (u64) at 96..127 and nrec (i32) at
# 128.
HDR_POP, HDR_NREC = 96, 128


def library_tau(c):
    """tau as the library holds it (lbm_get_config), not the Python copy."""
    import ctypes
    cfg = capi.Config()
    capi.check(c.lib.lbm_get_config(c.h, ctypes.byref(cfg)), "lbm_get_config")
    return cfg.tau


def test_snapshot_rejects_corrupt_headers_without_changing_context(tmp_path):
    say("case: snapshot with corrupt record count / flags")
    dims, spacing = (16, 16, 16), 0.08
    mesh = O.Mesh(os.path.join(DATA, "cube_angled.obj"))
    ok, internal = mesh.insert(dims, spacing)
    assert ok
    flags, recs = mesh.preprocess(dims, spacing, internal * O.FLAG_INTERNAL)
    assert len(recs[0]) > 0
    cfg = capi.make_config(*dims, lattice=27, mode=capi.MODE_REFERENCE,
                           tau=0.6)
    path = tmp_path / "ref.bin"
    with capi.Context(cfg) as a:
        a.set_populations(np.full(27 * a.n, 0.001, np.float32))
        a.upload_flags(flags)
        a.upload_reflections(*recs)
        a.step_reference(3, with_force=True)
        a.save(path)
    good = bytearray(open(path, "rb").read())
    assert int.from_bytes(good[HDR_NREC:HDR_NREC + 4], "little") == len(recs[0])
    hdr = int.from_bytes(good[12:16], "little")
    pop = int.from_bytes(good[HDR_POP:HDR_POP + 8], "little")

    def patched(fn):
        b = bytearray(good)
        fn(b)
        p = tmp_path / "bad.bin"
        p.write_bytes(bytes(b))
        return p

    bad = {
        # a record count near 2^31 (would size ~700 GB of host vectors)
        "huge nrec": patched(lambda b: b.__setitem__(
            slice(HDR_NREC, HDR_NREC + 4), (2 ** 31 - 1).to_bytes(4, "little"))),
        # one record more than the file holds
        "nrec + 1": patched(lambda b: b.__setitem__(
            slice(HDR_NREC, HDR_NREC + 4),
            (len(recs[0]) + 1).to_bytes(4, "little"))),
        # the first record's node flagged reflecting but not internal
        "flags": patched(lambda b: b.__setitem__(
            hdr + pop + int(recs[0][0]), capi.FLAG_REFLECTED)),
        # a trailing byte
        "size": patched(lambda b: b.extend(b"\0")),
    }
    other = capi.make_config(*dims, lattice=27, mode=capi.MODE_REFERENCE,
                             tau=1.1)
    with capi.Context(other) as c:
        f0 = np.full(27 * c.n, 0.002, np.float32)
        c.set_populations(f0)
        for name, p in bad.items():
            with pytest.raises(capi.LbmError) as e:
                c.load(p)
            say(f"  {name}: {e.value}")
            assert e.value.code == 1  # LBM_ERR_INVALID
            # nothing changed: parameters, populations, flags
            assert abs(library_tau(c) - 1.1) < 1e-7
            assert np.array_equal(c.populations(), f0)
            assert not c.flags().any()
        c.load(path)  # the intact file still loads
        assert abs(library_tau(c) - 0.6) < 1e-7
        assert np.array_equal(c.flags(), flags)


def test_batch_member_takes_snapshot_with_other_moving_wall(tmp_path):
    # a batch member loads snapshots whose moving wall differs from its own
    # (lattices with and without one share the batch kernel): after the
    # same steps it holds the same bytes as a plain context that loaded the
    # same file
    say("case: batch member loads snapshots with / without a moving wall")
    dims = (16, 16, 16)
    plain = capi.make_config(*dims, lattice=19, tau=0.7, periodic=(1, 0, 0),
                             force=(1e-5, 0, 0))
    walled = capi.make_config(*dims, lattice=19, tau=0.6, lid_face=3,
                              lid_u=(0.05, 0, 0))
    for saved, member in ((plain, walled), (walled, plain)):
        path = tmp_path / "m.bin"
        with capi.Context(saved) as a:
            a.init_equilibrium()
            a.step(7)
            a.save(path)
        with capi.Context(saved) as c:
            c.load(path)
            c.step(10)
            want = c.populations()
        with capi.Batch(capi.STORE_F32, 1 << 30) as bt:
            m = bt.add(member)
            bt.step(3, 0)  # tables built for the member's own parameters
            m.load(path)
            bt.step(10, 10)
            bt.synchronize()
            got = m.populations()
        assert np.array_equal(got, want)
