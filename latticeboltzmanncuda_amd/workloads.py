"""Synthetic workloads of the benchmark configurations, built on the C-ABI.

cavity          lid-driven cavity (D3Q19 n^3 or D2Q9 n^2), lid = y max face
channel         D3Q19 body-force channel through random spheres (periodic x)
taylor_green    z-extruded Taylor-Green vortex on a periodic box
precision_study FP16/BF16/FP32 shifted-storage runs against an FP64 run

All populations live on the GPU; only setup arrays and readbacks cross the
bus.
"""
from __future__ import annotations

import math

import numpy as np

from . import capi


def cavity(n: int, lattice: int = capi.D3Q19, tau: float = 0.6,
           lid_u: float = 0.1, storage: int = capi.STORE_F32
           ) -> capi.Context:
    nz = 1 if lattice == capi.D2Q9 else n
    ctx = capi.Context(capi.make_config(n, n, nz, lattice=lattice,
                                        storage=storage, tau=tau,
                                        lid_face=3, lid_u=(lid_u, 0, 0)))
    ctx.init_equilibrium()
    return ctx


def channel(dims=(256, 128, 128), tau=0.8, force=1e-6, seed=2,
            solid_fraction=0.2, storage=capi.STORE_F32):
    solid, spheres = capi.random_spheres(dims, seed, solid_fraction, 4.0,
                                         8.0, True)
    ctx = capi.Context(capi.make_config(*dims, lattice=capi.D3Q19,
                                        storage=storage, tau=tau,
                                        periodic=(1, 0, 0),
                                        force=(force, 0, 0)))
    ctx.upload_solid(solid)
    ctx.init_equilibrium()
    return ctx, solid, spheres


def taylor_green_fields(L: int, nz: int, u0: float):
    """rho (N) and u (3 x N) of the z-extruded vortex
    u = u0 sin(kx) cos(ky), v = -u0 cos(kx) sin(ky), w = 0, k = 2 pi / L,
    rho = 1 - (3 u0^2 / 4)(cos 2kx + cos 2ky)."""
    k = 2.0 * math.pi / L
    x = np.arange(L, dtype=np.float64)
    sx, cx = np.sin(k * x), np.cos(k * x)
    c2x = np.cos(2 * k * x)
    n = L * L * nz
    ux = (u0 * sx[:, None] * cx[None, :])[:, :, None]
    uy = (-u0 * cx[:, None] * sx[None, :])[:, :, None]
    rho = (1.0 - 0.75 * u0 * u0 * (c2x[:, None] + c2x[None, :]))[:, :, None]
    shape = (L, L, nz)
    u = np.empty((3, n), np.float64)
    u[0] = np.broadcast_to(ux, shape).reshape(-1)
    u[1] = np.broadcast_to(uy, shape).reshape(-1)
    u[2] = 0.0
    return np.broadcast_to(rho, shape).reshape(-1).copy(), u


def taylor_green(L: int = 256, nz: int | None = None, u0: float = 0.05,
                 tau: float = 0.55, lattice: int = capi.D3Q19,
                 storage: int = capi.STORE_F32) -> capi.Context:
    nz = L if nz is None else nz
    if lattice == capi.D2Q9:
        nz = 1
    ctx = capi.Context(capi.make_config(L, L, nz, lattice=lattice,
                                        storage=storage, tau=tau,
                                        periodic=(1, 1, 1)))
    rho, u = taylor_green_fields(L, nz, u0)
    ctx.init_equilibrium(rho, u)
    return ctx


def kinetic_energy(rho: np.ndarray, u: np.ndarray) -> float:
    return 0.5 * float(np.sum(rho.astype(np.float64) *
                              (u.astype(np.float64) ** 2).sum(0)))


def precision_study(L: int = 256, steps: int = 2000, u0: float = 0.05,
                    tau: float = 0.55,
                    storages=("f16", "bf16", "f32")) -> dict:
    """Runs the Taylor-Green vortex in FP64 storage (the reference run) and
    in each shifted 16-bit / FP32 storage, on the GPU, and reports velocity
    errors against the FP64 run, kinetic-energy decay against FP64 and the
    analytic rate exp(-4 nu k^2 t), and MLUPS of each run."""
    nu = (tau - 0.5) / 3.0
    k = 2.0 * math.pi / L
    analytic = math.exp(-4.0 * nu * k * k * steps)
    n = L * L * L
    out = {"workload": f"D3Q19 Taylor-Green {L}^3 periodic, u0={u0}, "
                       f"tau={tau}, {steps} steps",
           "analytic_ke_ratio": analytic, "runs": {}}

    def run(storage_name):
        st = capi.STORAGE_NAMES[storage_name]
        ctx = taylor_green(L, L, u0, tau, capi.D3Q19, st)
        rho0, u_init = ctx.moments()
        ke0 = kinetic_energy(rho0, u_init)
        ms = ctx.time_steps(steps)
        rho, u = ctx.moments()
        ctx.close()
        return rho, u, ke0, ms

    rho64, u64, ke0_64, ms64 = run("f64")
    ke64 = kinetic_energy(rho64, u64)
    norm2 = float(np.sqrt((u64.astype(np.float64) ** 2).sum()))
    umax = float(np.abs(u64).max())
    out["runs"]["f64"] = {
        "mlups": n * steps / (ms64 / 1e3) / 1e6,
        "ke_ratio": ke64 / ke0_64,
        "ke_ratio_vs_analytic": ke64 / ke0_64 / analytic - 1.0,
    }
    for name in storages:
        rho, u, ke0, ms = run(name)
        d = u.astype(np.float64) - u64.astype(np.float64)
        ke = kinetic_energy(rho, u)
        out["runs"][name] = {
            "mlups": n * steps / (ms / 1e3) / 1e6,
            "u_l2_rel_err": float(np.sqrt((d ** 2).sum())) / norm2,
            "u_linf_rel_err": float(np.abs(d).max()) / umax,
            "ke_ratio": ke / ke0,
            "ke_rel_err_vs_f64": (ke / ke0) / (ke64 / ke0_64) - 1.0,
            "ke_ratio_vs_analytic": ke / ke0 / analytic - 1.0,
        }
        del rho, u, d
    return out


# --------------------------------------------------------------- batching
def random_squares(nx: int, ny: int, frac: float, rng,
                   clear_top: int = 0) -> np.ndarray:
    """Solid mask (nx*ny bytes, index y + ny*x) of random axis-aligned
    squares, side U{2..max(2, min(nx, ny)//8)}, added until at least frac of
    the nodes are solid.  The top clear_top rows stay fluid (a square that
    touches a moving lid makes a corner pocket whose density the lid drains
    until BGK diverges)."""
    solid = np.zeros((nx, ny), np.uint8)
    smax = max(2, min(nx, ny) // 8)
    target = frac * nx * ny
    while solid.sum() < target:
        s = int(rng.integers(2, smax + 1))
        x0 = int(rng.integers(0, max(1, nx - s + 1)))
        y0 = int(rng.integers(0, max(1, ny - clear_top - s + 1)))
        solid[x0:x0 + s, y0:y0 + s] = 1
    return solid.reshape(-1)


def c4_specs(count: int = 512, seed: int = 4, smin: int = 32, smax: int = 512,
             U: float = 0.1, tau_min: float = 0.51, solid_frac: float = 0.1,
             storage: int = capi.STORE_F32):
    """The batching workload: count D2Q9 lattices, sides log-uniform in
    [smin, smax] per axis, half lid-driven cavities (lid = y max face at
    U) and half body-force channels (periodic x, walls y, force
    8 nu U / ny^2 for a centre-line speed U), Reynolds number U[100, 1000],
    tau = 1/2 + 3 U L / Re with L the lid length (cavity) or the height
    (channel), clamped to tau >= tau_min, random squares to solid_frac.
    Returns a list of (config, solid, info) and the number of clamped
    lattices."""
    rng = np.random.default_rng(seed)
    out, clamped = [], 0
    for _ in range(count):
        nx = int(round(math.exp(rng.uniform(math.log(smin), math.log(smax)))))
        ny = int(round(math.exp(rng.uniform(math.log(smin), math.log(smax)))))
        kind = "cavity" if rng.random() < 0.5 else "channel"
        re = float(rng.uniform(100.0, 1000.0))
        L = nx if kind == "cavity" else ny
        tau = 0.5 + 3.0 * U * L / re
        if tau < tau_min:
            tau = tau_min
            clamped += 1
        solid = random_squares(nx, ny, solid_frac, rng,
                               clear_top=2 if kind == "cavity" else 0)
        if kind == "cavity":
            cfg = capi.make_config(nx, ny, 1, lattice=capi.D2Q9,
                                   storage=storage, tau=tau, lid_face=3,
                                   lid_u=(U, 0, 0))
        else:
            nu = (tau - 0.5) / 3.0
            cfg = capi.make_config(nx, ny, 1, lattice=capi.D2Q9,
                                   storage=storage, tau=tau,
                                   periodic=(1, 0, 0),
                                   force=(8.0 * nu * U / (ny * ny), 0, 0))
        out.append((cfg, solid, {"kind": kind, "re": re, "tau": tau,
                                 "nx": nx, "ny": ny}))
    return out, clamped


# Admission bound of the batching workload on the BGK cell Reynolds number
# U / nu (lbm_cell_reynolds).  Measured on the 512 lattices of seed 4 after
# 1000 steps: every lattice whose FP64 run diverges (21 lid-driven
# cavities) has U / nu >= 17.4, and FP16 storage departs from FP64 by
# 1e-3 to 1.7e-1 on the near-unstable cavities at 14-19; below 15 the FP16
# aggregate error is 9.4e-4 (90th percentile 7.2e-4).
C4_MAX_CELL_RE = 15.0


def c4_gate(specs, max_cell_re: float = C4_MAX_CELL_RE):
    """Splits the batching workload at the stability bound: (admitted
    specs, number rejected).  The rule is the library's (lbm_cell_reynolds,
    the same one lbm_batch_add applies under lbm_batch_set_stability_limit),
    so the admitted set is what a batch with that limit accepts."""
    kept = [s for s in specs if capi.cell_reynolds(s[0]) <= max_cell_re]
    return kept, len(specs) - len(kept)


def with_storage(cfg: capi.Config, storage: int) -> capi.Config:
    c = capi.Config()
    C_fields = [f[0] for f in capi.Config._fields_]
    for name in C_fields:
        v = getattr(cfg, name)
        if hasattr(v, "__len__"):
            getattr(c, name)[:] = list(v)
        else:
            setattr(c, name, v)
    c.storage = storage
    return c


def c4_plain_mlups(specs, steps: int) -> dict:
    """Every lattice in its own context (own stream), one plain launch per
    lattice per step; wall-clock MLUPS over all lattices."""
    import time
    ctxs = []
    try:
        for cfg, solid, _ in specs:
            c = capi.Context(cfg)
            c.upload_solid(solid)
            c.init_equilibrium()
            ctxs.append(c)
        nodes = sum(c.cfg.nx * c.cfg.ny * c.cfg.nz for c in ctxs)
        capi.step_each(ctxs, 2)
        for c in ctxs:
            c.synchronize()
        t0 = time.perf_counter()
        capi.step_each(ctxs, steps)
        for c in ctxs:
            c.synchronize()
        dt = time.perf_counter() - t0
    finally:
        for c in ctxs:
            c.close()
    return {"mlups": nodes * steps / dt / 1e6, "ms_per_step": dt * 1e3 / steps,
            "nodes": nodes, "steps": steps}


def c4_batched_mlups(specs, steps: int, chunk: int = 100,
                     autotune: bool = True, storage: int = capi.STORE_F32,
                     launch=None) -> dict:
    """All lattices in one batch: one persistent launch per kernel class and
    step, replayed as hipGraph chunks of `chunk` steps, launch shape from
    the autotuner (or `launch` = (block, grid)); wall-clock MLUPS (hipEvents
    agree within launch jitter)."""
    import time
    with capi.Batch(storage) as b:
        for cfg, solid, _ in specs:
            b.add(cfg, solid)
        _, nodes, nbytes = b.info()
        if launch is not None:
            b.set_launch(*launch)
            block, grid = launch
        else:
            block, grid = b.autotune(steps=10) if autotune else b.launch()
        b.step(chunk or 10, chunk)  # warm-up (and capture, with chunks)
        b.synchronize()
        t0 = time.perf_counter()
        ms = b.time_steps(steps, chunk)
        dt = time.perf_counter() - t0
    return {"mlups": nodes * steps / dt / 1e6,
            "mlups_events": nodes * steps / (ms / 1e3) / 1e6,
            "ms_per_step": dt * 1e3 / steps, "nodes": nodes, "steps": steps,
            "block": block, "grid": grid, "device_bytes": nbytes}


def c4_velocity_error(specs, storage: int, steps: int, chunk: int = 100
                      ) -> dict:
    """Velocity of every lattice after `steps` batched steps in `storage`
    against the same batch in FP64 storage: aggregate relative L2 error
    (sum over lattices of |u - u64|^2 over sum |u64|^2), per-lattice
    median / 90th percentile / worst, over the lattices whose FP64 run stays
    finite and below twice the driving speed (the others are counted).

    The same batch in FP32 storage (the reference's own precision) sorts
    the compared lattices: "conditioned" ones, where FP32 and FP64 agree to
    1e-3, and ill-conditioned ones (near-unstable flows that amplify any
    rounding difference, or flows too slow for FP32 to resolve their
    increments), whose errors say little about the storage format.  Both
    aggregates are reported."""
    def run(st):
        out = []
        with capi.Batch(st) as b:
            members = [b.add(with_storage(cfg, st), solid)
                       for cfg, solid, _ in specs]
            b.set_launch(256, 0)
            b.step(steps, chunk)
            b.synchronize()
            for m in members:
                out.append(m.moments()[1].astype(np.float64))
        return out

    def rel(a, b):
        n2 = float((b ** 2).sum())
        d2 = float(((a - b) ** 2).sum())
        return d2, n2

    u64 = run(capi.STORE_F64)
    u32 = run(capi.STORE_F32) if storage != capi.STORE_F32 else None
    ux = run(storage)
    num = den = cnum = cden = 0.0
    per, unstable, diverged, cond, stalled = [], 0, 0, 0, 0
    for k, ((cfg, _, info), a, b) in enumerate(zip(specs, ux, u64)):
        # a lattice whose FP64 run is not finite or exceeds twice the
        # driving speed has no reference solution to compare with
        if not np.isfinite(b).all():
            unstable += 1
            continue
        if float(np.abs(b).max()) > 0.2:
            unstable += 1
            continue
        if not np.isfinite(a).all():
            diverged += 1
            continue
        d2, n2 = rel(a, b)
        num += d2
        den += n2
        per.append(math.sqrt(d2 / n2) if n2 > 0 else 0.0)
        # the storage's velocity stayed below 1 % of FP64's: the per-step
        # change of its populations is below half a unit in the last place
        if float(np.abs(a).max()) < 0.01 * float(np.abs(b).max()):
            stalled += 1
        ref32 = u32[k] if u32 is not None else a
        e2, _ = rel(ref32, b)
        if n2 > 0 and math.sqrt(e2 / n2) <= 1e-3:
            cond += 1
            cnum += d2
            cden += n2
    per = np.array(per) if per else np.zeros(1)
    return {"u_l2_rel_err": math.sqrt(num / den) if den > 0 else float("nan"),
            "u_l2_rel_err_conditioned":
                math.sqrt(cnum / cden) if cden > 0 else float("nan"),
            "conditioned": cond,
            "median_lattice": float(np.median(per)),
            "p90_lattice": float(np.percentile(per, 90)),
            "worst_lattice": float(per.max()),
            "compared": int(len(per)), "unstable_in_f64": unstable,
            "diverged_only_in_storage": diverged, "stalled_in_storage": stalled,
            "within_1e-3": int((per <= 1e-3).sum()), "steps": steps}


def c5_requests(lam: float, count: int, seed: int = 5,
                storage: int = capi.STORE_F16, max_steps: int = 20000,
                tol: float = 1e-6, u_target: float = 0.05) -> list:
    """The serving workload: `count` requests with Poisson arrivals at rate
    lam per second.  Half D2Q9 squares n^2 (n log-uniform 64..1024, random
    squares), half D3Q19 cubes n^3 (n log-uniform 32..128, random spheres of
    radius n/16..n/8, at least 2); solid fraction U[0.05, 0.25]; tau
    U[0.55, 1.0]; body-force channel (periodic x, bounce-back walls) with
    force 8 nu u_target / ny^2; run to residual tol or max_steps."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for _ in range(count):
        t += float(rng.exponential(1.0 / lam))
        frac = float(rng.uniform(0.05, 0.25))
        tau = float(rng.uniform(0.55, 1.0))
        nu = (tau - 0.5) / 3.0
        if rng.random() < 0.5:
            n = int(round(math.exp(rng.uniform(math.log(64), math.log(1024)))))
            solid = random_squares(n, n, frac, rng)
            cfg = capi.make_config(n, n, 1, lattice=capi.D2Q9,
                                   storage=storage, tau=tau,
                                   periodic=(1, 0, 0),
                                   force=(8.0 * nu * u_target / (n * n), 0, 0))
        else:
            n = int(round(math.exp(rng.uniform(math.log(32), math.log(128)))))
            rmin = max(2.0, n / 16.0)
            rmax = max(rmin + 1.0, n / 8.0)
            solid, _ = capi.random_spheres((n, n, n), int(rng.integers(1 << 31)),
                                           frac, rmin, rmax, True)
            cfg = capi.make_config(n, n, n, lattice=capi.D3Q19,
                                   storage=storage, tau=tau,
                                   periodic=(1, 0, 0),
                                   force=(8.0 * nu * u_target / (n * n), 0, 0))
        out.append({"cfg": cfg, "solid": solid, "arrival_s": t,
                    "max_steps": max_steps, "tol": tol})
    return out


def c5_velocity_error(count: int = 16, steps: int = 2000, seed: int = 5,
                      chunk: int = 100) -> dict:
    """FP16-store serving against the same requests served in FP64: the
    first `count` requests of the serving workload (seed), all arriving at
    once, each run for exactly `steps` steps (no early stop, so both
    storages compare the same time), batched engine.  Per request the
    relative L2 velocity error |u16 - u64| / |u64|; requests whose FP64 run
    is not finite are counted, not compared.  A request 'stalls' when its
    FP16 velocity stays below 1 % of the FP64 one: the per-step change of
    its stored populations (the body force's increment) is below half an
    FP16 unit in the last place, so rounding returns the old value."""
    out = {}
    reqs = {}
    for name, st in (("f16", capi.STORE_F16), ("f64", capi.STORE_F64)):
        rs = c5_requests(100.0, count, seed=seed, storage=st,
                         max_steps=steps, tol=0.0)
        for r in rs:
            r["arrival_s"] = 0.0
        _, res = capi.serve(rs, storage=st, chunk=chunk, keep_results=True)
        reqs[name] = res
    errs, d2, u2, unstable, stalled = [], 0.0, 0.0, 0, 0
    for a, b in zip(reqs["f16"], reqs["f64"]):
        u16 = a["u"].astype(np.float64)
        u64 = b["u"].astype(np.float64)
        if not np.isfinite(u64).all() or b["steps"] != steps:
            unstable += 1
            continue
        if np.abs(u16).max() < 0.01 * np.abs(u64).max():
            stalled += 1
        dd = float(((u16 - u64) ** 2).sum())
        uu = float((u64 ** 2).sum())
        d2 += dd
        u2 += uu
        errs.append(math.sqrt(dd / uu) if uu > 0 else 0.0)
    errs = np.array(errs)
    out.update({
        "requests": count, "steps": steps, "compared": int(len(errs)),
        "unstable_in_f64": unstable, "stalled_in_f16": stalled,
        "within_1e-3": int((errs <= 1e-3).sum()),
        "u_l2_rel_err": math.sqrt(d2 / u2) if u2 > 0 else 0.0,
        "median_request": float(np.median(errs)) if len(errs) else 0.0,
        "worst_request": float(errs.max()) if len(errs) else 0.0,
    })
    return out
