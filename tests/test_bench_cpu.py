"""bench.py's contract pieces that run without a GPU: the command line,
the single-GPU guard and the CPU-baseline leg on a small sample."""
import importlib.util
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location(
        "bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_bench_defines_every_section():
    b = load_bench()
    for name in ("parse", "precision_section", "lattice2d_section",
                 "batched_section",
                 "serving_section", "reference_mode_section",
                 "cpu_baseline", "measured_traffic", "main"):
        assert callable(getattr(b, name)), name


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip())["error"]


def test_bench_traffic_file_is_committed():
    b = load_bench()
    t = b.measured_traffic()
    assert t is not None and t > 637534208  # above the algorithmic bytes


def test_cpu_baseline_leg_runs(monkeypatch):
    b = load_bench()
    from latticeboltzmanncuda_amd import capi
    # a small lattice in place of the 256x128x128 channel keeps this quick
    monkeypatch.setattr(b, "DIMS", (32, 16, 16))
    solid, _ = capi.random_spheres(b.DIMS, b.SEED, b.SOLID_FRACTION, 2.0,
                                   3.0, True)
    out = b.cpu_baseline(solid, 0.2)
    assert out["kind"] == "port" and out["unit"] == "MLUPS"
    assert out["value"] > 0 and out["cores"] >= 1
    assert np.isfinite(out["value"])
    assert out["single_thread_value"] > 0
    assert out["nproc"] >= 1
    # the reference's main loop on the CPU (reference-mode workload)
    ref = out["reference_main_loop"]
    assert ref["reflection_records"] > 0
    assert ref["all_threads"]["mlups"] > 0 and ref["single_thread"]["mlups"] > 0
