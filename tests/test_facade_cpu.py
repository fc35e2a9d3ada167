"""The C++ facade on the CPU: the GPU-free C++ suites (Geometry, Model,
compressInts) and the facade's geometry preprocessing against the oracle
(flags bit-exact, reflection records identical)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# LBM_FACADE_BIN selects another build of the suites, e.g. the sanitizer
# build (latticeboltzmanncuda_amd/bin_san, tests/run_sanitized.sh)
BIN = os.environ.get("LBM_FACADE_BIN") or os.path.join(
    ROOT, "latticeboltzmanncuda_amd", "bin")
DATA = os.path.join(ROOT, "tests", "golden", "dataFiles")


def run_suite(name, filt=None):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (python -m latticeboltzmanncuda_amd.build)")
    env = dict(os.environ, LBM_TEST_DATA=DATA)
    cmd = [exe] + ([filt] if filt else [])
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    return r.stdout


def test_geometry_suite():
    # reference test/GeometryTest.cpp
    run_suite("GeometryTest")


def test_model_suite():
    # reference test/ModelTest.cpp
    run_suite("ModelTest")


def test_compression_case():
    # reference test/CalculationsTest.cpp:279-284 (host only)
    out = run_suite("CalculationsTest", "CompressionTest")
    assert "1 tests ran, 0 failed" in out


def facade_geometry(tmp_path, mesh, dims, spacing):
    out = tmp_path / "geo.bin"
    r = subprocess.run([os.path.join(BIN, "GeometryDump"),
                        os.path.join(DATA, mesh), *map(str, dims),
                        repr(spacing), str(out)], timeout=120)
    assert r.returncode == 0
    raw = out.read_bytes()
    n = int(np.prod(dims))
    if len(raw) == 4 and np.frombuffer(raw, np.int32)[0] == -1:
        return None
    flags = np.frombuffer(raw[:n], np.uint8)
    nrec = int(np.frombuffer(raw[n:n + 4], np.int32)[0])
    off = n + 4
    nodes = np.frombuffer(raw[off:off + 4 * nrec], np.int32)
    off += 4 * nrec
    dirs = np.frombuffer(raw[off:off + 56 * nrec], np.uint32).reshape(nrec, 14)
    off += 56 * nrec
    w = np.frombuffer(raw[off:off + 324 * nrec], np.float32).reshape(nrec, 81)
    return flags, (nodes, dirs, w)


@pytest.mark.parametrize("mesh,dims,spacing", [
    ("cube_angled.obj", (3, 3, 3), 0.5),
    ("compat_cube.obj", (10, 10, 10), 0.1),
    ("compat_cube.obj", (32, 32, 32), 0.05),
    ("cube_angled.obj", (32, 32, 32), 0.05),
    ("cube_angled.obj", (24, 20, 28), 0.045),
    ("cube.obj", (8, 8, 8), 1.0),  # outside the box: ignored
])
def test_facade_geometry_matches_oracle(tmp_path, mesh, dims, spacing):
    got = facade_geometry(tmp_path, mesh, dims, spacing)
    m = O.Mesh(os.path.join(DATA, mesh))
    ok, internal = m.insert(dims, spacing)
    if not ok:
        assert got is None
        return
    flags_o, (node_o, dirs_o, w_o) = m.preprocess(dims, spacing,
                                                  internal * O.FLAG_INTERNAL)
    flags_f, (node_f, dirs_f, w_f) = got
    assert np.array_equal(flags_f, flags_o)
    assert np.array_equal(node_f, node_o)
    assert np.array_equal(dirs_f, dirs_o)
    assert np.abs(w_f - w_o).max() <= 1e-6
