"""The command-line driver with the reference's protocol
(LatticeBoltzmannCuda.cpp:11-95): fused and staged loops print the same
force sequence, and the final force matches the oracle's main loop."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "latticeboltzmanncuda_amd", "bin", "LatticeBoltzmannHip")
DATA = os.path.join(ROOT, "tests", "golden", "dataFiles")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_cli(*extra, steps=50, tau=0.6):
    cmd = [CLI, os.path.join(DATA, "cube_angled.obj"), "8", "8", "8", "4", "4",
           "4", "1.0", str(tau), "0.05", str(steps), *extra]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return [l for l in r.stdout.splitlines() if l.strip()]


def test_cli_fused_and_staged_print_the_same_forces():
    say("case: CLI 32^3 cube_angled, fused vs staged, 50 steps")
    fused = run_cli()
    staged = run_cli("--staged")
    assert len(fused) == len(staged) == 50
    a = np.array([[float(v) for v in l.split()] for l in fused])
    b = np.array([[float(v) for v in l.split()] for l in staged])
    assert np.abs(a[0]).max() == 0.0  # the register starts at zero
    scale = max(np.abs(b).max(), 1e-30)
    say(f"  max force difference {np.abs(a - b).max() / scale:.2e}")
    assert np.abs(a - b).max() <= 1e-5 * scale


def test_cli_force_matches_oracle_main_loop():
    say("case: CLI 32^3 cube_angled vs oracle main loop, 40 steps")
    steps, tau, dims, spacing = 40, 0.6, (32, 32, 32), 0.05
    lines = run_cli("--quiet", steps=steps, tau=tau)
    F = np.array([float(v) for v in lines[-1].split()])
    mesh = O.Mesh(os.path.join(DATA, "cube_angled.obj"))
    ok, internal = mesh.insert(dims, spacing)
    assert ok
    flags, recs = mesh.preprocess(dims, spacing, internal * O.FLAG_INTERNAL)
    n = int(np.prod(dims))
    flow = np.full(27, 0.001, np.float32)
    flow[6] = 0.002
    f = np.tile(flow[:, None], (1, n)).reshape(-1).copy()
    Fo = np.zeros(3, np.float32)
    Fo64 = np.zeros(3, np.float64)
    O.ref_run_main(dims, f, flags, recs, tau, O.FLOW_PLANE, [0, 0], flow, Fo,
                   steps, Fo64)
    err = np.abs(F - Fo64).max() / np.abs(Fo64).max()
    say(f"  CLI {F} oracle exact {Fo64} rel {err:.2e}")
    assert err <= 1e-5


def read_record_file(path, dims):
    """The record file of Lattice::writeRecords: magic "LBMPTS1\\0", the
    three extents (int32), then one 144-byte LatticePoint per node in index
    order (the retrieve_data stream, Lattice.cpp:87-113)."""
    from latticeboltzmanncuda_amd import capi
    raw = open(path, "rb").read()
    assert raw[:8] == b"LBMPTS1\0"
    assert list(np.frombuffer(raw[8:20], np.int32)) == list(dims)
    pts = np.frombuffer(raw[20:], capi.POINT_DTYPE)
    assert len(pts) == int(np.prod(dims))
    return pts


def test_cli_record_files_match_oracle(tmp_path):
    # P1-style run (cube_angled at spacing 0.05 in 32^3, plane injection,
    # force on): the records written after k steps hold exactly the
    # oracle's populations after k steps, the node coordinates of the index
    # order and the model's flags; a second run that starts from that file
    # and takes m more steps writes exactly the oracle's state after k + m
    k, m, tau, dims, spacing = 30, 20, 0.6, (32, 32, 32), 0.05
    say(f"case: CLI record files, {k} + {m} steps vs oracle")
    first, second = tmp_path / "k.rec", tmp_path / "km.rec"
    run_cli("--quiet", "--write-records", str(first), steps=k, tau=tau)
    run_cli("--quiet", "--read-records", str(first), "--write-records",
            str(second), steps=m, tau=tau)
    mesh = O.Mesh(os.path.join(DATA, "cube_angled.obj"))
    ok, internal = mesh.insert(dims, spacing)
    assert ok
    flags, recs = mesh.preprocess(dims, spacing, internal * O.FLAG_INTERNAL)
    n = int(np.prod(dims))
    flow = np.full(27, 0.001, np.float32)
    flow[6] = 0.002
    f0 = np.tile(flow[:, None], (1, n)).reshape(-1).copy()
    x, y, z = np.meshgrid(np.arange(dims[0]), np.arange(dims[1]),
                          np.arange(dims[2]), indexing="ij")
    for path, steps in ((first, k), (second, k + m)):
        pts = read_record_file(path, dims)
        fo = O.ref_run_main(dims, f0, flags, recs, tau, O.FLOW_PLANE, [0, 0],
                            flow, np.zeros(3, np.float32), steps)
        assert np.array_equal(pts["f"].T.reshape(-1), fo), steps
        assert np.array_equal(pts["x"], x.reshape(-1))
        assert np.array_equal(pts["y"], y.reshape(-1))
        assert np.array_equal(pts["z"], z.reshape(-1))
        assert np.array_equal(pts["isInternal"],
                              (flags & O.FLAG_INTERNAL).astype(bool))
        assert np.array_equal(pts["isReflected"],
                              (flags & O.FLAG_REFLECTED).astype(bool))
        assert not pts["reflections"].any() and not pts["d_reflections"].any()
        say(f"  {steps} steps: populations, coordinates and flags exact")
