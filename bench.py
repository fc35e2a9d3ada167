#!/usr/bin/env python3
"""Headline benchmark: D3Q19 BGK body-force channel through random spheres,
256 x 128 x 128 nodes, FP32, tau = 0.8, periodic in x, half-way bounce-back
walls in y/z, spheres of radius U[4,8] from MT19937-64(seed 2) up to 20 %
solid, Guo force 1e-6 along x, rho = 1, u = 0.

A step is one fused pull collide+stream launch over the whole lattice with
the populations already resident in HBM.  Prints one JSON line.  Besides the
headline it carries five sections (skipped with --headline-only):
precision (Taylor-Green 256^3 in FP16/BF16/FP32 storage vs FP64),
lattice2d (one D2Q9 FP16 lattice of 4096^2, group vs one-node step),
batched (512 D2Q9 lattices, per-lattice launches vs the persistent batch
kernel, FP32 and FP16), serving (Poisson requests through the
continuous-batching server, batched vs per-lattice engine) and reference
(the reference's D3Q27 main loop, fused vs staged).

    python bench.py [--gpus 1] [--steps 5000] [--warmup 100]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("MLUPS (%HBM peak) plain vs batched fp16-store @≤1e-3 rel err; "
          "p99 latency")
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

DIMS = (256, 128, 128)
TAU = 0.8
FORCE = (1e-6, 0.0, 0.0)
SEED = 2
SOLID_FRACTION = 0.2


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="wall time budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the extra sections")
    ap.add_argument("--sections",
                    default="precision,lattice2d,batched,serving,reference",
                    help="comma list of the extra sections to run "
                         "(also: c4kernel, the batch kernel alone for "
                         "profilers)")
    ap.add_argument("--c3-steps", type=int, default=2000)
    ap.add_argument("--l2d-steps", type=int, default=1000)
    ap.add_argument("--ref-steps", type=int, default=1000)
    ap.add_argument("--c4-steps", type=int, default=500)
    ap.add_argument("--c4-launch", default="",
                    help="block,grid of the batch kernel instead of the "
                         "autotuner's choice")
    ap.add_argument("--c4-chunk", type=int, default=100,
                    help="steps per hipGraph replay of the batch (0: plain "
                         "launches, e.g. under a kernel tracer)")
    ap.add_argument("--c5-requests", type=int, default=64)
    ap.add_argument("--c5-lambda", default="10,200",
                    help="comma list of arrival rates (requests/s)")
    ap.add_argument("--c5-engines", default="batched,per_lattice",
                    help="batched, per_lattice, batched_capped (admission "
                         "also capped at --c5-node-cap resident nodes)")
    ap.add_argument("--c5-node-cap", type=int, default=32 << 20)
    ap.add_argument("--c5-max-steps", type=int, default=20000)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def precision_section(steps):
    """Taylor-Green 256^3, FP16 / BF16 / FP32 shifted storage against the
    FP64 GPU run and the analytic decay."""
    from latticeboltzmanncuda_amd import workloads
    res = workloads.precision_study(L=256, steps=steps, u0=0.05, tau=0.55)
    runs = {k: {kk: (round(vv, 9) if isinstance(vv, float) else vv)
                for kk, vv in v.items()} for k, v in res["runs"].items()}
    return {"workload": res["workload"],
            "analytic_ke_ratio": round(res["analytic_ke_ratio"], 9),
            "runs": runs}


def batched_section(steps, launch=None, chunk=100):
    """512 D2Q9 cavity/channel lattices (seed 4): plain per-lattice
    launches vs one persistent launch per step, replayed in hipGraph chunks
    of `chunk` steps (0: launched one by one), FP32 and FP16 storage; FP16
    velocity error against FP64 storage."""
    from latticeboltzmanncuda_amd import capi, workloads
    out = {"workload": "512 D2Q9 lattices (seed 4), sides log-uniform "
                       "32-512, cavity/channel, Re 100-1000, tau >= 0.51, "
                       "10% squares; admitted if cell Reynolds number "
                       f"U/nu <= {workloads.C4_MAX_CELL_RE:g}",
           "steps": steps, "chunk": chunk}
    drawn, clamped = workloads.c4_specs(512, storage=capi.STORE_F32)
    specs32, rejected = workloads.c4_gate(drawn)
    out["tau_clamped"] = clamped
    out["max_cell_re"] = workloads.C4_MAX_CELL_RE
    out["lattices"] = len(specs32)
    out["rejected_unstable"] = rejected
    out["nodes"] = int(sum(c.nx * c.ny for c, _, _ in specs32))
    out["nodes_drawn"] = int(sum(c.nx * c.ny for c, _, _ in drawn))
    for name, st, bpl in (("f32", capi.STORE_F32, 72),
                          ("f16", capi.STORE_F16, 36)):
        specs = [(workloads.with_storage(c, st), s, i)
                 for c, s, i in specs32]
        log(f"batched: {name} plain")
        plain = workloads.c4_plain_mlups(specs, steps)
        log(f"batched: {name} persistent")
        bat = workloads.c4_batched_mlups(specs, steps, chunk, True, st, launch)
        out[name] = {
            "plain_mlups": round(plain["mlups"], 1),
            "batched_mlups": round(bat["mlups"], 1),
            "speedup": round(bat["mlups"] / plain["mlups"], 3),
            "batched_hbm_frac": round(bat["mlups"] * 1e6 * bpl / 8e12, 4),
            "plain_hbm_frac": round(plain["mlups"] * 1e6 * bpl / 8e12, 4),
            "launch": {"block": bat["block"], "grid": bat["grid"]},
        }
    log("batched: f16 vs f64 velocity error")
    err = workloads.c4_velocity_error(specs32, capi.STORE_F16, 1000, chunk)
    out["f16"]["error_vs_f64"] = {
        k: (round(v, 8) if isinstance(v, float) else v)
        for k, v in err.items()}
    return out


def c4kernel_section(steps, launch=None):
    """The batch kernel alone, for profilers: the admitted lattices of the
    batching workload in FP16 stepped with plain launches (no graphs, so
    every dispatch is visible to a kernel tracer), the 4096^2 D2Q9 FP16
    channel with one node per thread as the single-lattice control, and a
    256 MiB copy of known size for calibrating byte counters."""
    from latticeboltzmanncuda_amd import capi, workloads
    out = {"steps": steps}
    out["copy_256MiB_gbs"] = round(capi.measure_copy_bandwidth(256 << 20, 5), 1)
    specs, rejected = workloads.c4_gate(
        workloads.c4_specs(512, 4, storage=capi.STORE_F16)[0])
    with capi.Batch(capi.STORE_F16) as b:
        for cfg, solid, _ in specs:
            b.add(cfg, solid)
        if launch:
            b.set_launch(*launch)
            blk, grd = launch
        else:
            blk, grd = b.autotune(steps=10)
        _, nodes, _ = b.info()
        b.step(10, 0)
        b.synchronize()
        us = b.time_steps(steps, 0) / steps * 1e3
    out["batch"] = {"lattices": len(specs), "rejected_unstable": rejected,
                    "nodes": nodes, "block": blk, "grid": grd,
                    "us_per_step": round(us, 2),
                    "glups": round(nodes / us / 1e3, 1),
                    "hbm_frac_36B": round(nodes * 36 / (us * 1e-6) / 8e12, 4)}
    n = 4096
    sq = workloads.random_squares(n, n, 0.1, np.random.default_rng(7))
    os.environ["LBM_GROUP_STEP"] = "0"
    try:
        cfg = capi.make_config(n, n, 1, lattice=capi.D2Q9,
                               storage=capi.STORE_F16, tau=0.8,
                               periodic=(1, 0, 0), force=(1e-6, 0, 0))
        with capi.Context(cfg) as c:
            c.upload_solid(sq)
            c.init_equilibrium()
            c.step(10)
            c.synchronize()
            us = c.time_steps(steps) / steps * 1e3
    finally:
        os.environ.pop("LBM_GROUP_STEP", None)
    out["single_4096_channel_one_node"] = {
        "us_per_step": round(us, 2), "glups": round(n * n / us / 1e3, 1),
        "hbm_frac_36B": round(n * n * 36 / (us * 1e-6) / 8e12, 4)}
    return out


def lattice2d_section(steps):
    """One D2Q9 FP16 lattice of 4096^2 nodes (two population copies of
    604 MB, streamed from HBM): lid-driven cavity and body-force channel
    with 10 % squares, run by the group step (two nodes per thread, chosen
    for lattices this size) and, for comparison, by the one-node step."""
    from latticeboltzmanncuda_amd import capi, workloads
    n = 4096
    sq = workloads.random_squares(n, n, 0.1, np.random.default_rng(7))
    out = {"workload": "D2Q9 FP16 4096^2 single lattice: cavity (lid 0.1, "
                       "tau 0.6) and channel (force 1e-6, tau 0.8, 10% "
                       "squares)", "steps": steps}
    for name, kw, solid in (
            ("cavity", dict(tau=0.6, lid_face=3, lid_u=(0.1, 0, 0)), None),
            ("channel", dict(tau=0.8, periodic=(1, 0, 0), force=(1e-6, 0, 0)),
             sq)):
        res = {}
        for step, env in (("group", None), ("one_node", "0")):
            if env is None:
                os.environ.pop("LBM_GROUP_STEP", None)
            else:
                os.environ["LBM_GROUP_STEP"] = env
            try:
                cfg = capi.make_config(n, n, 1, lattice=capi.D2Q9,
                                       storage=capi.STORE_F16, **kw)
                with capi.Context(cfg) as c:
                    c.upload_solid(solid)
                    c.init_equilibrium()
                    c.step(20)
                    c.synchronize()
                    us = c.time_steps(steps) / steps * 1e3
            finally:
                os.environ.pop("LBM_GROUP_STEP", None)
            res[step] = {"us_per_step": round(us, 2),
                         "mlups": round(n * n / us, 1),
                         "hbm_frac_36B": round(n * n * 36 / (us * 1e-6) / 8e12,
                                               4)}
        res["speedup"] = round(res["one_node"]["us_per_step"]
                               / res["group"]["us_per_step"], 3)
        out[name] = res
    return out


def serving_section(nreq, lams, max_steps, engines, node_cap):
    """Continuous-batching solve server, FP16 storage, Poisson arrivals:
    batched engine vs per-lattice launches on the same requests, for each
    arrival rate."""
    from latticeboltzmanncuda_amd import capi, workloads
    out = {"workload": f"{nreq} requests per rate (seed 5), Poisson "
                       "arrivals, 50% D2Q9 64-1024^2 squares, 50% D3Q19 "
                       "32-128^3 spheres, 5-25% solid, tau 0.55-1.0, "
                       "body-force channel, residual 1e-6 or "
                       f"{max_steps} steps, 200 GB budget, FP16 storage, "
                       "chunk 100", "rates": {}}
    eng = {"batched": (capi.ENGINE_BATCHED, 0),
           "batched_capped": (capi.ENGINE_BATCHED, node_cap),
           "per_lattice": (capi.ENGINE_PER_LATTICE, 0)}
    for lam in lams:
        reqs = workloads.c5_requests(lam, nreq, max_steps=max_steps)
        row = {}
        for name in engines:
            log(f"serving: lambda {lam} {name}")
            st, res = capi.serve(reqs, storage=capi.STORE_F16, chunk=100,
                                 engine=eng[name][0], progress_s=30.0,
                                 max_resident_nodes=eng[name][1])
            conv = sum(1 for r in res if r["status"] == 0)
            gbs = (st["population_bytes"] / st["wall_s"] / 1e9
                   if st["wall_s"] > 0 else 0.0)
            row[name] = {
                "mlups": round(st["mlups"], 1),
                # algorithmic population bytes (2 Q s per node and step,
                # Q 9 or 19, FP16) over the wall time, against 8 TB/s
                "population_gbs": round(gbs, 1),
                "hbm_frac": round(gbs / 8000.0, 4),
                "unstable": st["unstable"], "diverged": st["diverged"],
                "p50_latency_s": round(st["p50_latency_s"], 4),
                "p99_latency_s": round(st["p99_latency_s"], 4),
                "mean_latency_s": round(st["mean_latency_s"], 4),
                "wall_s": round(st["wall_s"], 3),
                "completed": st["completed"], "converged": conv,
                "rejected": st["rejected"],
                "max_resident": st["max_resident"],
                "peak_gb": round(st["peak_bytes"] / 1e9, 3),
            }
        out["rates"][str(lam)] = row
    log("serving: fp16 vs fp64 velocity error")
    err = workloads.c5_velocity_error(16, 2000)
    out["f16_error_vs_f64"] = {
        k: (round(v, 8) if isinstance(v, float) else v) for k, v in err.items()}
    return out


def reference_mode_section(steps):
    """The reference's own main loop (LatticeBoltzmannCuda.cpp:11-95) on a
    larger lattice: D3Q27 FP32 128^3, the angled cube of the reference's
    tests voxelised at spacing 0.0125, plane x = 0 injection, force on; the
    command-line driver runs it fused (one pass per step) and staged (the
    reference's four passes)."""
    import subprocess
    cli = os.path.join(ROOT, "latticeboltzmanncuda_amd", "bin",
                       "LatticeBoltzmannHip")
    mesh = os.path.join(ROOT, "tests", "golden", "dataFiles",
                        "cube_angled.obj")
    out = {"workload": "D3Q27 FP32 128^3 reference main loop, angled cube "
                       "(spacing 0.0125), plane injection, force on, tau 0.6",
           "steps": steps}
    n = 128 ** 3
    for name, extra in (("fused", []), ("staged", ["--staged"])):
        log(f"reference mode: {name}")
        r = subprocess.run([cli, mesh, "8", "8", "8", "16", "16", "16", "1.0",
                            "0.6", "0.0125", str(steps), "--quiet", "--time",
                            *extra], capture_output=True, text=True,
                           timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"reference-mode driver failed: {r.stderr}")
        t = json.loads(r.stdout.strip().splitlines()[-1])
        us = t["seconds"] / steps * 1e6
        out[name] = {"mlups": round(t["mlups"], 1),
                     "us_per_step": round(us, 2),
                     "reflection_records": t["reflection_records"],
                     "hbm_frac_216B": round(n * 216 / (us * 1e-6) / 8e12, 4)}
        if name == "fused":
            out["setup_s_gpu"] = round(t["setup_seconds"], 4)
    out["speedup"] = round(out["fused"]["mlups"] / out["staged"]["mlups"], 3)
    # the same voxelisation + reflection tables with the host builders
    # (one process, no GPU work), for the setup comparison
    dump = os.path.join(ROOT, "latticeboltzmanncuda_amd", "bin",
                        "GeometryDump")
    tmp = os.path.join("/tmp", f"lbm_geometry_{os.getpid()}.bin")
    t0 = time.perf_counter()
    r = subprocess.run([dump, mesh, "128", "128", "128", "0.0125", tmp],
                       capture_output=True, text=True, timeout=600)
    out["setup_s_host"] = round(time.perf_counter() - t0, 4)
    if os.path.exists(tmp):
        os.remove(tmp)
    if r.returncode != 0:
        raise RuntimeError(f"host geometry builder failed: {r.stderr}")
    return out


def cpu_baseline(solid: np.ndarray, seconds: float) -> dict:
    """The oracle's FP32 standard-mode step on the host cores, same
    workload, timed over a bounded number of steps."""
    from oracle import oracle as O

    q = 19
    mask = O.std_build_mask(q, DIMS, solid, (1, 0, 0))
    p = O.std_params(tau=TAU, force=FORCE, periodic=(1, 0, 0))
    f0 = O.std_init_eq(q, DIMS).astype(np.float32)
    n = int(np.prod(DIMS))
    threads = O.get_threads()
    chunk = 8

    def timed(budget, nthreads):
        O.set_threads(nthreads)
        # the GPU kernel's arithmetic: FP32 f_i - w_i, shifted collision
        st = O.STORE_F32_SHIFTED
        f = O.std_run_f32(q, DIMS, f0, mask, p, 1, st)  # warm-up step
        steps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget:
            f = O.std_run_f32(q, DIMS, f, mask, p, chunk, st)
            steps += chunk
        return steps, time.perf_counter() - t0

    steps, dt = timed(seconds, threads)
    # one thread, a quarter of the budget (a scalar-port reference point)
    steps1, dt1 = timed(seconds / 4, 1)
    O.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((l.split(":", 1)[1].strip() for l in fh
                          if l.startswith("model name")), "")
    except OSError:
        pass
    ref = reference_cpu_baseline(O, seconds / 4, threads)
    return {
        "value": round(n * steps / dt / 1e6, 2),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP32 (shifted) standard-mode step (C, OpenMP {threads} "
                   f"threads) on the same 256x128x128 sphere channel, "
                   f"{steps} steps in {dt:.1f} s; includes a per-call "
                   f"population copy-in/out every {chunk} steps"),
        "single_thread_value": round(n * steps1 / dt1 / 1e6, 2),
        "single_thread_sample": f"{steps1} steps in {dt1:.1f} s, 1 thread",
        "cpu_model": model,
        "nproc": os.cpu_count(),
        "reference_main_loop": ref,
    }


def reference_cpu_baseline(O, seconds, threads):
    """The oracle's restatement of the reference's main loop (stream,
    reflect with force, collide, plane injection; LatticeBoltzmannCuda.cpp
    :85-93) on the reference-mode bench workload: D3Q27 128^3 with the
    angled cube at spacing 0.0125, timed at all threads and at one."""
    dims, spacing = (128, 128, 128), 0.0125
    mesh = O.Mesh(os.path.join(ROOT, "tests", "golden", "dataFiles",
                               "cube_angled.obj"))
    ok, internal = mesh.insert(dims, spacing)
    flags, recs = mesh.preprocess(dims, spacing, internal * O.FLAG_INTERNAL)
    n = int(np.prod(dims))
    flow = np.full(27, 0.001, np.float32)
    flow[6] = 0.002
    f = np.tile(flow[:, None], (1, n)).reshape(-1).copy()
    force = np.zeros(3, np.float32)
    out = {"workload": "D3Q27 FP32 128^3 reference main loop, angled cube "
                       "(spacing 0.0125), plane injection, force on, tau 0.6",
           "reflection_records": int(len(recs[0]))}
    for key, nt, budget in (("all_threads", threads, seconds),
                            ("single_thread", 1, seconds)):
        O.set_threads(nt)
        g = O.ref_run_main(dims, f, flags, recs, 0.6, O.FLOW_PLANE, [0, 0],
                           flow, force, 1)  # warm-up step
        steps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget:
            g = O.ref_run_main(dims, g, flags, recs, 0.6, O.FLOW_PLANE,
                               [0, 0], flow, force, 2)
            steps += 2
        dt = time.perf_counter() - t0
        out[key] = {"mlups": round(n * steps / dt / 1e6, 2),
                    "threads": nt, "steps": steps,
                    "seconds": round(dt, 2)}
    O.set_threads(threads)
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles",
                            "k_std_step_d3q19_f32_channel_traffic.json")


def measured_traffic():
    """HBM bytes per launch of the step kernel from the committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs (calibrated on the copy
    kernel), or None when no such measurement is committed."""
    try:
        with open(TRAFFIC_FILE) as fh:
            return json.load(fh)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must "
                          "be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)
    from latticeboltzmanncuda_amd import capi

    solid, spheres = capi.random_spheres(DIMS, SEED, SOLID_FRACTION, 4.0, 8.0,
                                         True)
    n = int(np.prod(DIMS))
    fluid = int((solid == 0).sum())
    # practical HBM ceiling: a streaming copy of one population array
    copy_gbs = capi.measure_copy_bandwidth(n * 19 * 4, 20)
    cfg = capi.make_config(*DIMS, lattice=capi.D3Q19,
                           storage=capi.STORE_F32, tau=TAU,
                           periodic=(1, 0, 0), force=FORCE)
    ctx = capi.Context(cfg)
    ctx.upload_solid(solid)
    ctx.init_equilibrium()
    ctx.step(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ms = ctx.time_steps(args.steps)  # hipEvents on the context's stream
    wall = time.perf_counter() - t0
    ctx.synchronize()
    cons = ctx.conservation()
    if not np.isfinite(cons).all():
        raise RuntimeError("non-finite state after the timed steps")
    ms_per_step = ms / args.steps
    mlups = n * args.steps / (ms / 1e3) / 1e6
    # the step reads and writes all 19 FP32 populations of every node (solid
    # nodes included: their lanes store the rest state so that every store
    # writes whole cache lines)
    bytes_per_launch = n * 2 * 19 * 4
    achieved = bytes_per_launch / (ms_per_step / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "D3Q19 BGK channel 256x128x128, random spheres "
                        "(seed 2, r 4-8) 20% solid, tau 0.8, periodic x, "
                        "bounce-back y/z, Guo force 1e-6 x",
            "nodes": n,
            "fluid_nodes": fluid,
            "solid_fraction": round(1 - fluid / n, 4),
            "spheres": int(len(spheres)),
            "storage": "f32",
            "mflups": round(fluid * args.steps / (ms / 1e3) / 1e6, 1),
            "wall_s_timed_region": round(wall, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(),
            "kernel": "k_std_step<19, f32>",
            "bytes_per_launch": bytes_per_launch,
            "copy_gbs": round(copy_gbs, 1),
            "frac_of_copy": round(achieved / copy_gbs, 4),
        },
    }
    ctx.close()
    sections = set() if args.headline_only else set(args.sections.split(","))
    if "precision" in sections:
        log("precision")
        out["precision"] = precision_section(args.c3_steps)
    if "lattice2d" in sections:
        log("lattice2d")
        out["lattice2d"] = lattice2d_section(args.l2d_steps)
    launch = (tuple(int(v) for v in args.c4_launch.split(","))
              if args.c4_launch else None)
    if "batched" in sections:
        log("batched")
        out["batched"] = batched_section(args.c4_steps, launch, args.c4_chunk)
    if "c4kernel" in sections:
        log("c4kernel")
        out["c4kernel"] = c4kernel_section(args.c4_steps, launch)
    if "reference" in sections:
        log("reference mode")
        out["reference_mode"] = reference_mode_section(args.ref_steps)
    if "serving" in sections:
        log("serving")
        out["serving"] = serving_section(
            args.c5_requests, [float(v) for v in args.c5_lambda.split(",")],
            args.c5_max_steps, args.c5_engines.split(","), args.c5_node_cap)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(solid, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
