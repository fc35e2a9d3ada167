"""ctypes binding of the C-ABI in include/lbm.h (liblbm_hip.so).

This is the Python view of the product path: every call goes through the
HIP library.  There is no fallback: if the library was not built, or a GPU
call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "liblbm_hip.so")
HEADER = os.path.join(ROOT, "include", "lbm.h")

# enums of include/lbm.h
OK = 0
D2Q9, D3Q19, D3Q27 = 9, 19, 27
STORE_F32, STORE_F64, STORE_F16, STORE_BF16 = 0, 1, 2, 3
MODE_STANDARD, MODE_REFERENCE = 0, 1
ENGINE_BATCHED, ENGINE_PER_LATTICE = 0, 1
ERR_INVALID, ERR_BUDGET, ERR_UNSTABLE = 1, 6, 8
# serving request status
SERVE_CONVERGED, SERVE_STEP_LIMIT, SERVE_REJECTED, SERVE_DIVERGED, \
    SERVE_UNSTABLE = 0, 1, 2, 3, 4
FLOW_NONE, FLOW_POINT, FLOW_PLANE, FLOW_BOX, FLOW_MASK = 0, 1, 2, 3, 4
FLAG_INTERNAL, FLAG_REFLECTED = 1, 2
MASK_SOLID = 1 << 31
MASK_LID = 1 << 30
STORAGE_BYTES = {STORE_F32: 4, STORE_F64: 8, STORE_F16: 2, STORE_BF16: 2}
STORAGE_NAMES = {"f32": STORE_F32, "f64": STORE_F64, "f16": STORE_F16,
                 "bf16": STORE_BF16}


class LbmError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("lattice", C.c_int32), ("storage", C.c_int32), ("mode", C.c_int32),
        ("tau", C.c_float),
        ("periodic", C.c_int32 * 3),
        ("lid_face", C.c_int32),
        ("lid_u", C.c_float * 3),
        ("wall_rho", C.c_float),
        ("force", C.c_float * 3),
    ]


class LatticePoint(C.Structure):
    """The reference's 144-byte LatticePoint record."""
    _fields_ = [
        ("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32),
        ("particle_distribution", C.c_float * 27),
        ("isReflected", C.c_uint8), ("isInternal", C.c_uint8),
        ("pad_", C.c_uint8 * 6),
        ("reflections", C.c_void_p), ("d_reflections", C.c_void_p),
    ]


assert C.sizeof(LatticePoint) == 144

# numpy view of LatticePoint records
POINT_DTYPE = np.dtype({
    "names": ["x", "y", "z", "f", "isReflected", "isInternal", "pad",
              "reflections", "d_reflections"],
    "formats": [np.int32, np.int32, np.int32, (np.float32, 27), np.uint8,
                np.uint8, (np.uint8, 6), np.uint64, np.uint64],
    "offsets": [0, 4, 8, 12, 120, 121, 122, 128, 136],
    "itemsize": 144,
})

class Request(C.Structure):
    _fields_ = [
        ("cfg", Config),
        ("solid", C.c_void_p),
        ("arrival_s", C.c_double),
        ("max_steps", C.c_int32),
        ("tol", C.c_double),
        ("out_u", C.c_void_p),
        ("t_admit_s", C.c_double), ("t_done_s", C.c_double),
        ("steps", C.c_int32),
        ("residual", C.c_double),
        ("status", C.c_int32),
    ]


class ServeParams(C.Structure):
    _fields_ = [
        ("storage", C.c_int32),
        ("budget_bytes", C.c_int64),
        ("chunk", C.c_int32),
        ("max_resident", C.c_int32),
        ("autotune", C.c_int32),
        ("engine", C.c_int32),
        ("progress_s", C.c_double),
        ("max_resident_nodes", C.c_int64),
        ("max_cell_re", C.c_double),
    ]


class ServeStats(C.Structure):
    _fields_ = [
        ("wall_s", C.c_double),
        ("lattice_updates", C.c_int64),
        ("mlups", C.c_double),
        ("p50_latency_s", C.c_double), ("p99_latency_s", C.c_double),
        ("mean_latency_s", C.c_double),
        ("completed", C.c_int32), ("rejected", C.c_int32),
        ("max_resident", C.c_int32),
        ("peak_bytes", C.c_int64),
        ("unstable", C.c_int32),
        ("diverged", C.c_int32),
        ("population_bytes", C.c_double),
    ]


P = C.c_void_p
CTX = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64

# name -> (restype, argtypes)
_SIGNATURES = {
    "lbm_last_error": (C.c_char_p, []),
    "lbm_abi_version": (C.c_int, []),
    "lbm_device_info": (C.c_int, [P, C.c_char_p, C.c_size_t]),
    "lbm_create": (C.c_int, [C.POINTER(Config), P, C.POINTER(CTX)]),
    "lbm_destroy": (C.c_int, [CTX]),
    "lbm_get_config": (C.c_int, [CTX, C.POINTER(Config)]),
    "lbm_node_count": (i64, [CTX]),
    "lbm_device_bytes": (i64, [CTX]),
    "lbm_set_tau": (C.c_int, [CTX, C.c_float]),
    "lbm_set_body_force": (C.c_int, [CTX, P]),
    "lbm_synchronize": (C.c_int, [CTX]),
    "lbm_upload_points": (C.c_int, [CTX, P, i64]),
    "lbm_download_points": (C.c_int, [CTX, P, i64]),
    "lbm_set_populations_f32": (C.c_int, [CTX, P]),
    "lbm_get_populations_f32": (C.c_int, [CTX, P]),
    "lbm_set_populations_f64": (C.c_int, [CTX, P]),
    "lbm_get_populations_f64": (C.c_int, [CTX, P]),
    "lbm_init_equilibrium": (C.c_int, [CTX, P, P]),
    "lbm_upload_solid": (C.c_int, [CTX, P]),
    "lbm_get_mask": (C.c_int, [CTX, P]),
    "lbm_upload_flags": (C.c_int, [CTX, P]),
    "lbm_get_flags": (C.c_int, [CTX, P]),
    "lbm_upload_reflections": (C.c_int, [CTX, i32, P, P, P]),
    "lbm_set_flow": (C.c_int, [CTX, C.c_int, P, P, P]),
    "lbm_stream": (C.c_int, [CTX]),
    "lbm_collide": (C.c_int, [CTX, C.c_float]),
    "lbm_reflect": (C.c_int, [CTX, C.c_int]),
    "lbm_flow": (C.c_int, [CTX]),
    "lbm_prime_points": (C.c_int, [CTX]),
    "lbm_set_force": (C.c_int, [CTX, P]),
    "lbm_get_force": (C.c_int, [CTX, P]),
    "lbm_step": (C.c_int, [CTX, C.c_int]),
    "lbm_step_each": (C.c_int, [P, i32, C.c_int]),
    "lbm_step_reference": (C.c_int, [CTX, C.c_int, C.c_int]),
    "lbm_time_steps": (C.c_int, [CTX, C.c_int, C.POINTER(C.c_float)]),
    "lbm_time_steps_reference": (C.c_int, [CTX, C.c_int, C.c_int,
                                           C.POINTER(C.c_float)]),
    "lbm_moments": (C.c_int, [CTX, P, P]),
    "lbm_moments_noalloc": (C.c_int, [CTX, P, P]),
    "lbm_conservation": (C.c_int, [CTX, P]),
    "lbm_test_coordinates": (C.c_int, [CTX, C.POINTER(C.c_int)]),
    "lbm_batch_create": (C.c_int, [C.c_int, i64, P, C.POINTER(P)]),
    "lbm_batch_destroy": (C.c_int, [P]),
    "lbm_batch_estimate_bytes": (i64, [C.POINTER(Config), C.c_int]),
    "lbm_cell_reynolds": (C.c_double, [C.POINTER(Config)]),
    "lbm_batch_set_stability_limit": (C.c_int, [P, C.c_double]),
    "lbm_batch_add": (C.c_int, [P, C.POINTER(Config), P, C.c_int,
                                C.POINTER(i32)]),
    "lbm_batch_remove": (C.c_int, [P, i32]),
    "lbm_batch_member": (C.c_int, [P, i32, C.POINTER(CTX)]),
    "lbm_batch_info": (C.c_int, [P, C.POINTER(i32), C.POINTER(i64),
                                 C.POINTER(i64)]),
    "lbm_batch_set_launch": (C.c_int, [P, C.c_int, C.c_int]),
    "lbm_batch_get_launch": (C.c_int, [P, C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]),
    "lbm_batch_autotune": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
    "lbm_batch_step": (C.c_int, [P, C.c_int]),
    "lbm_batch_step_graph": (C.c_int, [P, C.c_int, C.c_int]),
    "lbm_batch_synchronize": (C.c_int, [P]),
    "lbm_batch_time_steps": (C.c_int, [P, C.c_int, C.c_int,
                                       C.POINTER(C.c_float)]),
    "lbm_batch_residuals": (C.c_int, [P, P, i32]),
    "lbm_serve_run": (C.c_int, [C.POINTER(ServeParams), C.POINTER(Request),
                                i32, C.POINTER(ServeStats)]),
    "lbm_build_reflections": (C.c_int, [CTX, P, i32, P, i32, C.c_float, i32,
                                         P, P, P, P]),
    "lbm_snapshot_save": (C.c_int, [CTX, C.c_char_p]),
    "lbm_snapshot_load": (C.c_int, [CTX, C.c_char_p]),
    "lbm_voxelize_mesh": (C.c_int, [CTX, P, i32, P, i32, C.c_float, P, P]),
    "lbm_measure_copy_bandwidth": (C.c_int, [i64, C.c_int,
                                             C.POINTER(C.c_double)]),
    "lbm_random_spheres": (i64, [i32, i32, i32, C.c_uint64, C.c_double,
                                 C.c_double, C.c_double, i32, P, i64, P]),
}

_lib = None


def header_symbols() -> list[str]:
    """Names of the functions declared in include/lbm.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_\w+)\s*\(", text)))


def load() -> C.CDLL:
    """Load liblbm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with "
            "`python -m latticeboltzmanncuda_amd.build` (there is no "
            "fallback path)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)


def check(code: int, where: str) -> None:
    if code != OK:
        msg = load().lbm_last_error()
        raise LbmError(code, where, msg.decode() if msg else "")


def device_info() -> tuple[int, str]:
    n = C.c_int(0)
    buf = C.create_string_buffer(256)
    check(load().lbm_device_info(C.byref(n), buf, 256), "lbm_device_info")
    return n.value, buf.value.decode()


def make_config(nx, ny, nz, lattice=D3Q19, storage=STORE_F32,
                mode=MODE_STANDARD, tau=0.6, periodic=(0, 0, 0), lid_face=-1,
                lid_u=(0.0, 0.0, 0.0), wall_rho=1.0,
                force=(0.0, 0.0, 0.0)) -> Config:
    c = Config()
    c.nx, c.ny, c.nz = nx, ny, nz
    c.lattice, c.storage, c.mode = lattice, storage, mode
    c.tau = tau
    c.periodic[:] = list(periodic)
    c.lid_face = lid_face
    c.lid_u[:] = list(lid_u)
    c.wall_rho = wall_rho
    c.force[:] = list(force)
    return c


def copy_config(cfg: Config) -> Config:
    c = Config()
    C.pointer(c)[0] = cfg
    return c


class Context:
    """One lattice resident on the GPU (an lbm_ctx handle).  Keeps its own
    copy of the configuration (set_tau / set_body_force update the copy, not
    the caller's object)."""

    def __init__(self, cfg: Config, stream: int | None = None):
        self.lib = load()
        cfg = copy_config(cfg)
        self.cfg = cfg
        h = CTX()
        check(self.lib.lbm_create(C.byref(cfg), stream, C.byref(h)),
              "lbm_create")
        self.h = h
        self.q = cfg.lattice
        self.dims = (cfg.nx, cfg.ny, cfg.nz)
        self.n = cfg.nx * cfg.ny * cfg.nz

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- exchange ---------------------------------------------------------
    def upload_points(self, pts: np.ndarray):
        assert pts.dtype == POINT_DTYPE and pts.size == self.n
        check(self.lib.lbm_upload_points(self.h, _ptr(pts), self.n),
              "lbm_upload_points")

    def download_points(self) -> np.ndarray:
        pts = np.zeros(self.n, POINT_DTYPE)
        check(self.lib.lbm_download_points(self.h, _ptr(pts), self.n),
              "lbm_download_points")
        return pts

    def set_populations(self, f: np.ndarray):
        f = np.ascontiguousarray(f)
        assert f.size == self.q * self.n
        if f.dtype == np.float64:
            check(self.lib.lbm_set_populations_f64(self.h, _ptr(f)),
                  "lbm_set_populations_f64")
        else:
            f = np.ascontiguousarray(f, dtype=np.float32)
            check(self.lib.lbm_set_populations_f32(self.h, _ptr(f)),
                  "lbm_set_populations_f32")

    def populations(self, dtype=np.float32) -> np.ndarray:
        out = np.empty(self.q * self.n, dtype)
        if dtype == np.float64:
            check(self.lib.lbm_get_populations_f64(self.h, _ptr(out)),
                  "lbm_get_populations_f64")
        else:
            check(self.lib.lbm_get_populations_f32(self.h, _ptr(out)),
                  "lbm_get_populations_f32")
        return out

    def init_equilibrium(self, rho=None, u=None):
        r = None if rho is None else np.ascontiguousarray(rho, np.float64)
        uu = None if u is None else np.ascontiguousarray(u, np.float64)
        check(self.lib.lbm_init_equilibrium(self.h, _ptr(r), _ptr(uu)),
              "lbm_init_equilibrium")

    # -- geometry ---------------------------------------------------------
    def upload_solid(self, solid):
        s = None if solid is None else np.ascontiguousarray(solid, np.uint8)
        check(self.lib.lbm_upload_solid(self.h, _ptr(s)), "lbm_upload_solid")

    def mask(self) -> np.ndarray:
        m = np.empty(self.n, np.uint32)
        check(self.lib.lbm_get_mask(self.h, _ptr(m)), "lbm_get_mask")
        return m

    def upload_flags(self, flags):
        f = np.ascontiguousarray(flags, np.uint8)
        check(self.lib.lbm_upload_flags(self.h, _ptr(f)), "lbm_upload_flags")

    def flags(self) -> np.ndarray:
        f = np.empty(self.n, np.uint8)
        check(self.lib.lbm_get_flags(self.h, _ptr(f)), "lbm_get_flags")
        return f

    def upload_reflections(self, node, dirs, weights):
        node = np.ascontiguousarray(node, np.int32)
        dirs = np.ascontiguousarray(dirs, np.uint32)
        weights = np.ascontiguousarray(weights, np.float32)
        check(self.lib.lbm_upload_reflections(self.h, len(node), _ptr(node),
                                              _ptr(dirs), _ptr(weights)),
              "lbm_upload_reflections")

    def set_flow(self, kind, params=None, induced=None, mask=None):
        p = None if params is None else np.ascontiguousarray(params, np.int32)
        ind = None if induced is None else np.ascontiguousarray(induced,
                                                                np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        check(self.lib.lbm_set_flow(self.h, kind, _ptr(p), _ptr(ind), _ptr(m)),
              "lbm_set_flow")

    # -- stages / steps ---------------------------------------------------
    def stream(self):
        check(self.lib.lbm_stream(self.h), "lbm_stream")

    def collide(self, tau):
        check(self.lib.lbm_collide(self.h, tau), "lbm_collide")

    def reflect(self, with_force=False):
        check(self.lib.lbm_reflect(self.h, int(with_force)), "lbm_reflect")

    def flow(self):
        check(self.lib.lbm_flow(self.h), "lbm_flow")

    def set_tau(self, tau: float):
        check(self.lib.lbm_set_tau(self.h, float(tau)), "lbm_set_tau")
        self.cfg.tau = float(tau)

    def set_body_force(self, force):
        a = np.ascontiguousarray(force, np.float32)
        check(self.lib.lbm_set_body_force(self.h, _ptr(a)),
              "lbm_set_body_force")
        self.cfg.force[:] = [float(v) for v in a]

    def set_force(self, f):
        a = np.ascontiguousarray(f, np.float32)
        check(self.lib.lbm_set_force(self.h, _ptr(a)), "lbm_set_force")

    def force(self) -> np.ndarray:
        a = np.zeros(3, np.float32)
        check(self.lib.lbm_get_force(self.h, _ptr(a)), "lbm_get_force")
        return a

    def step(self, n=1):
        check(self.lib.lbm_step(self.h, n), "lbm_step")

    def step_reference(self, n=1, with_force=False):
        check(self.lib.lbm_step_reference(self.h, n, int(with_force)),
              "lbm_step_reference")

    def time_steps(self, n) -> float:
        ms = C.c_float(0)
        check(self.lib.lbm_time_steps(self.h, n, C.byref(ms)),
              "lbm_time_steps")
        return ms.value

    def time_steps_reference(self, n, with_force=False) -> float:
        ms = C.c_float(0)
        check(self.lib.lbm_time_steps_reference(self.h, n, int(with_force),
                                                C.byref(ms)),
              "lbm_time_steps_reference")
        return ms.value

    def synchronize(self):
        check(self.lib.lbm_synchronize(self.h), "lbm_synchronize")

    # -- readback ---------------------------------------------------------
    def moments(self):
        rho = np.empty(self.n, np.float32)
        u = np.empty(3 * self.n, np.float32)
        check(self.lib.lbm_moments(self.h, _ptr(rho), _ptr(u)), "lbm_moments")
        return rho, u.reshape(3, self.n)

    def conservation(self) -> np.ndarray:
        out = np.zeros(4, np.float64)
        check(self.lib.lbm_conservation(self.h, _ptr(out)), "lbm_conservation")
        return out

    def test_coordinates(self) -> bool:
        ok = C.c_int(0)
        check(self.lib.lbm_test_coordinates(self.h, C.byref(ok)),
              "lbm_test_coordinates")
        return bool(ok.value)

    def save(self, path: str):
        check(self.lib.lbm_snapshot_save(self.h, str(path).encode()),
              "lbm_snapshot_save")

    def load(self, path: str):
        check(self.lib.lbm_snapshot_load(self.h, str(path).encode()),
              "lbm_snapshot_load")
        cfg = Config()
        check(self.lib.lbm_get_config(self.h, C.byref(cfg)), "lbm_get_config")
        self.cfg = cfg

    def device_bytes(self) -> int:
        return int(self.lib.lbm_device_bytes(self.h))


class Member(Context):
    """A lattice owned by a Batch: the per-lattice calls of Context work on
    it; it is freed by the batch, not by close()."""

    def __init__(self, batch, slot, handle, cfg):
        self.lib = load()
        self.cfg = cfg = copy_config(cfg)
        self.h = handle
        self.q = cfg.lattice
        self.dims = (cfg.nx, cfg.ny, cfg.nz)
        self.n = cfg.nx * cfg.ny * cfg.nz
        self.batch = batch
        self.slot = slot

    def close(self):
        self.h = None


class Batch:
    """Resident lattices stepped together by one persistent kernel."""

    def __init__(self, storage=STORE_F32, budget_bytes=200 * 10 ** 9,
                 stream=None):
        self.lib = load()
        h = P()
        check(self.lib.lbm_batch_create(storage, budget_bytes, stream,
                                        C.byref(h)), "lbm_batch_create")
        self.h = h
        self.storage = storage
        self.members: dict[int, Member] = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_batch_destroy(self.h)
            self.h = None
            self.members.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def estimate_bytes(cfg: Config, track_residual=False) -> int:
        return int(load().lbm_batch_estimate_bytes(C.byref(cfg),
                                                   int(track_residual)))

    def add(self, cfg: Config, solid=None, track_residual=False) -> Member:
        s = None if solid is None else np.ascontiguousarray(solid, np.uint8)
        slot = i32(-1)
        check(self.lib.lbm_batch_add(self.h, C.byref(cfg), _ptr(s),
                                     int(track_residual), C.byref(slot)),
              "lbm_batch_add")
        ctx = CTX()
        check(self.lib.lbm_batch_member(self.h, slot.value, C.byref(ctx)),
              "lbm_batch_member")
        m = Member(self, slot.value, ctx, cfg)
        self.members[slot.value] = m
        return m

    def remove(self, member: Member):
        check(self.lib.lbm_batch_remove(self.h, member.slot),
              "lbm_batch_remove")
        self.members.pop(member.slot, None)
        member.h = None

    def info(self):
        l, n, b = i32(0), i64(0), i64(0)
        check(self.lib.lbm_batch_info(self.h, C.byref(l), C.byref(n),
                                      C.byref(b)), "lbm_batch_info")
        return l.value, n.value, b.value

    def set_launch(self, block, grid):
        check(self.lib.lbm_batch_set_launch(self.h, block, grid),
              "lbm_batch_set_launch")

    def set_stability_limit(self, max_cell_re: float):
        """add() then raises LbmError (code ERR_UNSTABLE) for a lattice
        whose cell Reynolds number exceeds max_cell_re (0: no limit)."""
        check(self.lib.lbm_batch_set_stability_limit(self.h,
                                                      float(max_cell_re)),
              "lbm_batch_set_stability_limit")

    def launch(self):
        b, g = C.c_int(0), C.c_int(0)
        check(self.lib.lbm_batch_get_launch(self.h, C.byref(b), C.byref(g)),
              "lbm_batch_get_launch")
        return b.value, g.value

    def autotune(self, steps=10, use_cache=True):
        b, g = C.c_int(0), C.c_int(0)
        check(self.lib.lbm_batch_autotune(self.h, steps, int(use_cache),
                                          C.byref(b), C.byref(g)),
              "lbm_batch_autotune")
        return b.value, g.value

    def step(self, n=1, chunk=0):
        if chunk:
            check(self.lib.lbm_batch_step_graph(self.h, n, chunk),
                  "lbm_batch_step_graph")
        else:
            check(self.lib.lbm_batch_step(self.h, n), "lbm_batch_step")

    def synchronize(self):
        check(self.lib.lbm_batch_synchronize(self.h), "lbm_batch_synchronize")

    def time_steps(self, n, chunk=0) -> float:
        ms = C.c_float(0)
        check(self.lib.lbm_batch_time_steps(self.h, n, chunk, C.byref(ms)),
              "lbm_batch_time_steps")
        return ms.value

    def residuals(self) -> dict:
        nslots = max(self.members) + 1 if self.members else 1
        out = np.full(nslots, -1.0)
        check(self.lib.lbm_batch_residuals(self.h, _ptr(out), nslots),
              "lbm_batch_residuals")
        return {s: float(out[s]) for s in self.members}


def cell_reynolds(cfg: Config) -> float:
    """BGK cell Reynolds number U / nu of a standard-mode configuration
    (lbm_cell_reynolds: U is the moving wall's speed or the Poiseuille
    centre-line speed its body force drives across the narrowest walled
    extent)."""
    v = load().lbm_cell_reynolds(C.byref(cfg))
    if v < 0:
        check(int(-v), "lbm_cell_reynolds")
    return float(v)


def step_each(contexts, n: int) -> None:
    """n steps of every context, one single-lattice launch each per step."""
    arr = (CTX * len(contexts))(*[c.h for c in contexts])
    check(load().lbm_step_each(arr, len(contexts), n), "lbm_step_each")


def serve(requests, storage=STORE_F16, budget_bytes=200 * 10 ** 9,
          chunk=100, max_resident=0, autotune=True, keep_results=False,
          engine=ENGINE_BATCHED, progress_s=0.0, max_resident_nodes=0,
          max_cell_re=0.0):
    """Runs lbm_serve_run on a list of dicts with keys cfg (Config), solid
    (array or None), arrival_s, max_steps, tol.  engine: ENGINE_BATCHED
    (persistent kernel, hipGraph chunks) or ENGINE_PER_LATTICE (one plain
    launch per lattice per step).  max_cell_re > 0 rejects requests above
    that cell Reynolds number (status SERVE_UNSTABLE).  Returns (stats
    dict, list of per-request result dicts)."""
    n = len(requests)
    arr = (Request * n)()
    keep = []
    for k, r in enumerate(requests):
        q = arr[k]
        q.cfg = r["cfg"]
        s = r.get("solid")
        if s is not None:
            s = np.ascontiguousarray(s, np.uint8)
            keep.append(s)
            q.solid = s.ctypes.data
        q.arrival_s = r["arrival_s"]
        q.max_steps = r["max_steps"]
        q.tol = r["tol"]
        if keep_results:
            nn = r["cfg"].nx * r["cfg"].ny * r["cfg"].nz
            u = np.zeros(3 * nn, np.float32)
            keep.append(u)
            r["_u"] = u
            q.out_u = u.ctypes.data
    p = ServeParams()
    p.storage, p.budget_bytes, p.chunk = storage, budget_bytes, chunk
    p.max_resident, p.autotune = max_resident, int(autotune)
    p.engine = engine
    p.progress_s = progress_s
    p.max_resident_nodes = max_resident_nodes
    p.max_cell_re = max_cell_re
    st = ServeStats()
    check(load().lbm_serve_run(C.byref(p), arr, n, C.byref(st)),
          "lbm_serve_run")
    stats = {f[0]: getattr(st, f[0]) for f in ServeStats._fields_}
    results = []
    for k, r in enumerate(requests):
        q = arr[k]
        res = {"t_admit_s": q.t_admit_s, "t_done_s": q.t_done_s,
               "steps": q.steps, "residual": q.residual, "status": q.status,
               "latency_s": q.t_done_s - q.arrival_s}
        if keep_results:
            res["u"] = r.pop("_u").reshape(3, -1)
        results.append(res)
    return stats, results


def voxelize_mesh(ctx: "Context", vertices, faces, spacing: float):
    """Inside map (one byte per node) of a closed triangle mesh, computed on
    the GPU for the lattice of ctx."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
    f = np.ascontiguousarray(faces, np.int32).reshape(-1)
    out = np.zeros(ctx.n, np.uint8)
    cnt = C.c_int64(0)
    check(load().lbm_voxelize_mesh(ctx.h, _ptr(v), len(v) // 3, _ptr(f),
                                   len(f) // 3, float(spacing), _ptr(out),
                                   C.byref(cnt)), "lbm_voxelize_mesh")
    return out


def measure_copy_bandwidth(nbytes: int, iters: int = 20) -> float:
    """GB/s (read + write) of a plain device-to-device copy of nbytes."""
    g = C.c_double(0)
    check(load().lbm_measure_copy_bandwidth(nbytes, iters, C.byref(g)),
          "lbm_measure_copy_bandwidth")
    return g.value


def random_spheres(dims, seed, frac, rmin=4.0, rmax=8.0, periodic_x=True,
                   max_spheres=1 << 16):
    """Solid mask of the random-sphere channel (host setup in the library)."""
    n = int(np.prod(dims))
    solid = np.zeros(n, np.uint8)
    sph = np.zeros(4 * max_spheres, np.float64)
    ns = load().lbm_random_spheres(dims[0], dims[1], dims[2], seed, frac,
                                   rmin, rmax, 1 if periodic_x else 0,
                                   _ptr(solid), max_spheres, _ptr(sph))
    if ns < 0:
        check(int(-ns), "lbm_random_spheres")
    return solid, sph[: 4 * min(ns, max_spheres)].reshape(-1, 4)
