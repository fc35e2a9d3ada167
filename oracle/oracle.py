"""ctypes view of the CPU oracle (oracle/lib/liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker.  Imported by tests/, by
``__graft_entry__.smoke()`` and by ``bench.py``'s cpu_baseline leg; the
product never imports it.  The C sources (lbm_oracle.c / lbm_oracle.h) cite
the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LBM_ORACLE_LIB selects another build of the same sources, e.g. the
# sanitizer build of `make -C oracle sanitize` (oracle/lib_san/)
LIB_PATH = os.environ.get("LBM_ORACLE_LIB") or os.path.join(_HERE, "lib",
                                                            "liblbm_oracle.so")

FLAG_INTERNAL = 1
FLAG_REFLECTED = 2
FLOW_NONE, FLOW_POINT, FLOW_PLANE, FLOW_BOX = 0, 1, 2, 3
MASK_SOLID = 1 << 31
MASK_LID = 1 << 30
STORE_F32, STORE_F16_SHIFTED, STORE_BF16_SHIFTED, STORE_F32_SHIFTED = 0, 2, 3, 4


def build() -> None:
    """Compile the oracle with its own Makefile (gcc)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
I32 = C.c_int32
F32 = C.c_float


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


class StdParams(C.Structure):
    _fields_ = [
        ("tau", C.c_float),
        ("force", C.c_float * 3),
        ("lid_face", C.c_int32),
        ("lid_u", C.c_float * 3),
        ("wall_rho", C.c_float),
        ("periodic", C.c_int32 * 3),
    ]


def _declare(L: C.CDLL) -> None:
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_threads.restype = C.c_int
    L.orc_ref_stream.argtypes = [C.c_int, C.c_int, C.c_int, P, P]
    L.orc_ref_collide.argtypes = [C.c_int, C.c_int, C.c_int, P, C.c_float]
    L.orc_ref_reflect.argtypes = [C.c_int] * 3 + [P, P, P, C.c_int, P, P, P, P,
                                                  P]
    L.orc_ref_flow.argtypes = [C.c_int] * 3 + [P, C.c_int, P, P]
    L.orc_ref_run_main.argtypes = [C.c_int] * 3 + [P, P, C.c_int, P, P, P,
                                                  C.c_float, C.c_int, P, P, P,
                                                  P, C.c_int]
    L.orc_ref_run_lattice.argtypes = [C.c_int] * 3 + [P, C.c_float, C.c_int]
    L.orc_ref_moments.argtypes = [C.c_int] * 3 + [P, P, P]
    L.orc_mesh_load.argtypes = [C.c_char_p]
    L.orc_mesh_load.restype = P
    L.orc_mesh_free.argtypes = [P]
    for n in ("orc_mesh_nverts", "orc_mesh_nfaces", "orc_mesh_valid",
              "orc_mesh_closed"):
        getattr(L, n).argtypes = [P]
        getattr(L, n).restype = C.c_int
    L.orc_mesh_bbox.argtypes = [P, P]
    L.orc_mesh_inside.argtypes = [P, C.c_float, C.c_float, C.c_float]
    L.orc_mesh_inside.restype = C.c_int
    L.orc_mesh_reflection_vector.argtypes = [P, P, P, P]
    L.orc_rotate_vector.argtypes = [P, P, C.c_float, P]
    L.orc_distribute_vector.argtypes = [P, P, P]
    L.orc_insert_model.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_float, P]
    L.orc_insert_model.restype = C.c_int
    L.orc_preprocess_model.argtypes = [P, C.c_int, C.c_int, C.c_int,
                                       C.c_float, P, C.c_int, P, P, P]
    L.orc_preprocess_model.restype = C.c_int
    L.orc_compress_ints.argtypes = [P, C.c_int, C.c_int]
    L.orc_compress_ints.restype = C.c_uint
    L.orc_std_tables.argtypes = [C.c_int, P, P, P]
    L.orc_std_build_mask.argtypes = [C.c_int] * 4 + [P, P, C.c_int, P]
    L.orc_std_init_eq_f64.argtypes = [C.c_int] * 4 + [P, P, P]
    L.orc_std_run_f32.argtypes = [C.c_int] * 4 + [P, P, C.POINTER(StdParams),
                                                  C.c_int, C.c_int]
    L.orc_std_run_f64.argtypes = [C.c_int] * 4 + [P, P, C.POINTER(StdParams),
                                                  C.c_int]
    L.orc_std_moments_f64.argtypes = [C.c_int] * 4 + [P, P, P, P, P]
    L.orc_round_fp16.argtypes = [C.c_float]
    L.orc_round_fp16.restype = C.c_float
    L.orc_round_bf16.argtypes = [C.c_float]
    L.orc_round_bf16.restype = C.c_float
    L.orc_store16_sr.argtypes = [C.c_int, C.c_float, C.c_int, C.c_uint32]
    L.orc_store16_sr.restype = C.c_float
    L.orc_mt64_nth.argtypes = [C.c_uint64, C.c_int64]
    L.orc_mt64_nth.restype = C.c_uint64
    L.orc_random_spheres.argtypes = [C.c_int] * 3 + [C.c_uint64, C.c_double,
                                                    C.c_double, C.c_double,
                                                    C.c_int, P, C.c_int64, P]
    L.orc_random_spheres.restype = C.c_int64


def set_threads(n: int) -> None:
    lib().orc_set_threads(n)


def get_threads() -> int:
    return lib().orc_get_threads()


# ----------------------------------------------------------------- reference
def ref_stream(dims, f: np.ndarray) -> np.ndarray:
    out = np.empty_like(f)
    lib().orc_ref_stream(*dims, _ptr(f), _ptr(out))
    return out


def ref_collide(dims, f: np.ndarray, tau: float) -> np.ndarray:
    out = np.ascontiguousarray(f, dtype=np.float32).copy()
    lib().orc_ref_collide(*dims, _ptr(out), tau)
    return out


def ref_reflect(dims, f, flags, records, force=None, force64=None):
    """force: float accumulator, one addition per term (reference
    semantics); force64: FP64 sum of the pass added in."""
    out = np.empty_like(f)
    node, dirs, w = records
    lib().orc_ref_reflect(*dims, _ptr(f), _ptr(out), _ptr(flags), len(node),
                          _ptr(node), _ptr(dirs), _ptr(w), _ptr(force),
                          _ptr(force64))
    return out


def ref_flow(dims, f, kind, params, induced):
    out = f.copy()
    params = np.ascontiguousarray(params, dtype=np.int32)
    induced = np.ascontiguousarray(induced, dtype=np.float32)
    lib().orc_ref_flow(*dims, _ptr(out), kind, _ptr(params), _ptr(induced))
    return out


def ref_run_main(dims, f, flags, records, tau, flow_kind, flow_params, induced,
                 force, steps, force64=None):
    out = f.copy()
    node, dirs, w = records
    flow_params = np.ascontiguousarray(flow_params, dtype=np.int32)
    induced = np.ascontiguousarray(induced, dtype=np.float32)
    lib().orc_ref_run_main(*dims, _ptr(out), _ptr(flags), len(node), _ptr(node),
                           _ptr(dirs), _ptr(w), tau, flow_kind,
                           _ptr(flow_params), _ptr(induced), _ptr(force),
                           _ptr(force64), steps)
    return out


def ref_run_lattice(dims, f, tau, steps):
    out = f.copy()
    lib().orc_ref_run_lattice(*dims, _ptr(out), tau, steps)
    return out


def ref_moments(dims, f):
    n = int(np.prod(dims))
    rho = np.empty(n, np.float32)
    u = np.empty(3 * n, np.float32)
    lib().orc_ref_moments(*dims, _ptr(f), _ptr(rho), _ptr(u))
    return rho, u.reshape(3, n)


# ------------------------------------------------------------------ geometry
class Mesh:
    def __init__(self, path: str):
        self._h = lib().orc_mesh_load(path.encode())
        if not self._h:
            raise FileNotFoundError(path)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_mesh_free(self._h)
            self._h = None

    @property
    def nverts(self):
        return lib().orc_mesh_nverts(self._h)

    @property
    def nfaces(self):
        return lib().orc_mesh_nfaces(self._h)

    @property
    def valid(self):
        return bool(lib().orc_mesh_valid(self._h))

    @property
    def closed(self):
        return bool(lib().orc_mesh_closed(self._h))

    def bbox(self):
        out = np.zeros(6, np.float32)
        lib().orc_mesh_bbox(self._h, _ptr(out))
        return out

    def inside(self, x, y, z) -> bool:
        return bool(lib().orc_mesh_inside(self._h, x, y, z))

    def reflection_vector(self, p1, p2):
        a = np.asarray(p1, np.float32)
        b = np.asarray(p2, np.float32)
        out = np.zeros(3, np.float32)
        lib().orc_mesh_reflection_vector(self._h, _ptr(a), _ptr(b), _ptr(out))
        return out

    def insert(self, dims, spacing):
        internal = np.zeros(int(np.prod(dims)), np.uint8)
        ok = lib().orc_insert_model(self._h, *dims, spacing, _ptr(internal))
        return bool(ok), internal

    def preprocess(self, dims, spacing, flags):
        n = int(np.prod(dims))
        cap = max(16, n)
        node = np.zeros(cap, np.int32)
        dirs = np.zeros((cap, 14), np.uint32)
        w = np.zeros((cap, 81), np.float32)
        flags = flags.copy()
        nrec = lib().orc_preprocess_model(self._h, *dims, spacing, _ptr(flags),
                                          cap, _ptr(node), _ptr(dirs), _ptr(w))
        assert nrec >= 0
        return flags, (node[:nrec].copy(), dirs[:nrec].copy(), w[:nrec].copy())


def rotate_vector(axis, v, angle):
    a = np.asarray(axis, np.float32)
    b = np.asarray(v, np.float32)
    out = np.zeros(3, np.float32)
    lib().orc_rotate_vector(_ptr(a), _ptr(b), angle, _ptr(out))
    return out


def distribute_vector(v):
    a = np.asarray(v, np.float32)
    sc = np.zeros(27, np.float32)
    idx = np.zeros(27, np.int32)
    lib().orc_distribute_vector(_ptr(a), _ptr(sc), _ptr(idx))
    return sc, idx


def slot_dir(words, s: int) -> int:
    return int((int(words[s // 6]) >> (5 * (s % 6))) & 0x1F)


def compress_ints(values, bits=5) -> int:
    a = np.ascontiguousarray(values, dtype=np.int32)
    return int(lib().orc_compress_ints(_ptr(a), len(a), bits))


# ------------------------------------------------------------------ standard
def std_tables(q):
    e = np.zeros((q, 3), np.int32)
    inv = np.zeros(q, np.int32)
    w = np.zeros(q, np.float64)
    lib().orc_std_tables(q, _ptr(e), _ptr(inv), _ptr(w))
    return e, inv, w


def std_build_mask(q, dims, solid, periodic, lid_face=-1):
    n = int(np.prod(dims))
    mask = np.zeros(n, np.uint32)
    per = np.ascontiguousarray(periodic, dtype=np.int32)
    s = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8)
    lib().orc_std_build_mask(q, *dims, _ptr(s), _ptr(per), lid_face, _ptr(mask))
    return mask


def std_params(tau, force=(0, 0, 0), lid_face=-1, lid_u=(0, 0, 0),
               wall_rho=1.0, periodic=(0, 0, 0)) -> StdParams:
    p = StdParams()
    p.tau = tau
    p.force[:] = list(force)
    p.lid_face = lid_face
    p.lid_u[:] = list(lid_u)
    p.wall_rho = wall_rho
    p.periodic[:] = list(periodic)
    return p


def std_init_eq(q, dims, rho=None, u=None) -> np.ndarray:
    n = int(np.prod(dims))
    f = np.zeros(q * n, np.float64)
    r = None if rho is None else np.ascontiguousarray(rho, np.float64)
    uu = None if u is None else np.ascontiguousarray(u, np.float64)
    lib().orc_std_init_eq_f64(q, *dims, _ptr(r), _ptr(uu), _ptr(f))
    return f


def std_run_f32(q, dims, f, mask, params, steps, storage=STORE_F32):
    out = np.ascontiguousarray(f, dtype=np.float32).copy()
    lib().orc_std_run_f32(q, *dims, _ptr(out), _ptr(mask), C.byref(params),
                          steps, storage)
    return out


def std_run_f64(q, dims, f, mask, params, steps):
    out = np.ascontiguousarray(f, dtype=np.float64).copy()
    lib().orc_std_run_f64(q, *dims, _ptr(out), _ptr(mask), C.byref(params),
                          steps)
    return out


def std_moments(q, dims, f, mask, force=(0, 0, 0)):
    n = int(np.prod(dims))
    rho = np.zeros(n, np.float64)
    u = np.zeros(3 * n, np.float64)
    ff = np.ascontiguousarray(f, dtype=np.float64)
    F = np.asarray(force, np.float32)
    lib().orc_std_moments_f64(q, *dims, _ptr(ff), _ptr(mask), _ptr(F),
                              _ptr(rho), _ptr(u))
    return rho, u.reshape(3, n)


def round_fp16(x: float) -> float:
    return lib().orc_round_fp16(x)


def round_bf16(x: float) -> float:
    return lib().orc_round_bf16(x)


def store16_sr(storage: int, v: float, hash_: int, direction: int = 0) -> float:
    """The stored (unscaled) value of v, a population of `direction`, with
    stochastic rounding under the node hash (storage STORE_F16_SHIFTED or
    STORE_BF16_SHIFTED)."""
    return lib().orc_store16_sr(storage, v, direction, hash_ & 0xFFFFFFFF)


def mt64_nth(seed: int, n: int) -> int:
    return int(lib().orc_mt64_nth(seed, n))


def random_spheres(dims, seed, frac, rmin=4.0, rmax=8.0, periodic_x=True,
                   max_spheres=1 << 16):
    n = int(np.prod(dims))
    solid = np.zeros(n, np.uint8)
    sph = np.zeros(4 * max_spheres, np.float64)
    ns = lib().orc_random_spheres(*dims, seed, frac, rmin, rmax,
                                  1 if periodic_x else 0, _ptr(solid),
                                  max_spheres, _ptr(sph))
    return solid, sph[: 4 * min(ns, max_spheres)].reshape(-1, 4)
