"""The C-ABI library loads and exports every symbol include/lbm.h declares,
and its host-side setup code agrees with the oracle (CPU only: no GPU
compute call is made here)."""
import ctypes as C

import numpy as np
import pytest

from latticeboltzmanncuda_amd import capi
from oracle import oracle as O


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = capi.header_symbols()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the binding declares exactly the header's functions
    assert sorted(capi._SIGNATURES) == names


def test_abi_version_and_record_layout():
    lib = capi.load()
    assert lib.lbm_abi_version() == 1
    # LatticePoint is byte-compatible with the reference's 144-byte record
    assert C.sizeof(capi.LatticePoint) == 144
    assert capi.LatticePoint.particle_distribution.offset == 12
    assert capi.LatticePoint.isReflected.offset == 120
    assert capi.LatticePoint.isInternal.offset == 121
    assert capi.LatticePoint.reflections.offset == 128
    assert capi.LatticePoint.d_reflections.offset == 136
    assert capi.POINT_DTYPE.itemsize == 144


def test_error_reporting_without_gpu_work():
    lib = capi.load()
    h = capi.CTX()
    cfg = capi.make_config(0, 4, 4)
    rc = lib.lbm_create(C.byref(cfg), None, C.byref(h))
    assert rc == 1  # LBM_ERR_INVALID, rejected before any HIP call
    assert b"extents" in lib.lbm_last_error()
    cfg = capi.make_config(4, 4, 4, lattice=9)
    assert lib.lbm_create(C.byref(cfg), None, C.byref(h)) == 1
    cfg = capi.make_config(4, 4, 4, lattice=19, mode=capi.MODE_REFERENCE)
    assert lib.lbm_create(C.byref(cfg), None, C.byref(h)) == 5
    cfg = capi.make_config(4, 4, 4, tau=0.5)
    assert lib.lbm_create(C.byref(cfg), None, C.byref(h)) == 1
    cfg = capi.make_config(4, 4, 4, lid_face=3, periodic=(0, 1, 0))
    assert lib.lbm_create(C.byref(cfg), None, C.byref(h)) == 1
    # a 2-D lattice has no z axis, so no z force (its x / y force is fine)
    cfg = capi.make_config(4, 4, 1, lattice=9, force=(0, 0, 1e-6))
    assert lib.lbm_create(C.byref(cfg), None, C.byref(h)) == 1
    assert b"z component" in lib.lbm_last_error()
    with pytest.raises(capi.LbmError):
        capi.cell_reynolds(cfg)
    assert capi.cell_reynolds(capi.make_config(
        4, 4, 1, lattice=9, force=(1e-6, 0, 0), periodic=(1, 0, 0))) > 0
    assert lib.lbm_step(None, 1) == 1


def test_size_limits_and_null_arguments_without_gpu_work():
    lib = capi.load()
    h = capi.CTX()
    # every extent below 2^24 (24-bit coordinate multiplies)
    cfg = capi.make_config(1 << 24, 1, 1, lattice=19)
    assert lib.lbm_create(C.byref(cfg), None, C.byref(h)) == 5
    # population element indices below 2^31: 1024 x 1024 x 2048 x 19 is over
    cfg = capi.make_config(1024, 1024, 2048, lattice=19)
    assert lib.lbm_create(C.byref(cfg), None, C.byref(h)) == 5
    assert b"32-bit" in lib.lbm_last_error()
    # entry points of the snapshot and mesh paths reject null handles
    assert lib.lbm_snapshot_save(None, b"/tmp/x") == 1
    assert lib.lbm_snapshot_load(None, b"/tmp/x") == 1
    inside = np.zeros(8, np.uint8)
    cnt = C.c_int64()
    assert lib.lbm_voxelize_mesh(None, None, 0, None, 0, C.c_float(1.0),
                                 inside.ctypes.data_as(C.c_void_p),
                                 C.byref(cnt)) == 1
    n32 = C.c_int32()
    assert lib.lbm_build_reflections(None, None, 0, None, 0, C.c_float(1.0),
                                     0, None, None, None, C.byref(n32)) == 1


@pytest.mark.parametrize("dims,seed,px", [((64, 32, 32), 2, True),
                                          ((40, 24, 20), 7, False),
                                          ((96, 48, 40), 2, True)])
def test_random_spheres_bit_exact_vs_oracle(dims, seed, px):
    s1, sp1 = capi.random_spheres(dims, seed, 0.2, 4.0, 8.0, px)
    s2, sp2 = O.random_spheres(dims, seed, 0.2, 4.0, 8.0, px)
    assert (sp1 == sp2).all()
    assert (s1 == s2).all()
    assert s1.sum() >= 0.2 * np.prod(dims)


def test_cell_reynolds_rule_without_gpu_work():
    # U / nu with nu = (tau - 1/2) / 3: a lid at 0.1 and tau 0.51 gives 30
    cav = capi.make_config(64, 48, 1, lattice=capi.D2Q9, tau=0.51,
                           lid_face=3, lid_u=(0.1, 0, 0))
    assert capi.cell_reynolds(cav) == pytest.approx(30.0, rel=1e-6)
    # a body force F across walls W apart drives a Poiseuille channel at
    # F W^2 / (8 nu): the force 8 nu U / W^2 of the batching workload
    # gives U / nu back
    tau, W, U = 0.8, 40, 0.05
    nu = (tau - 0.5) / 3
    ch = capi.make_config(200, W, 1, lattice=capi.D2Q9, tau=tau,
                          periodic=(1, 0, 0), force=(8 * nu * U / W ** 2, 0, 0))
    assert capi.cell_reynolds(ch) == pytest.approx(U / nu, rel=1e-6)
    # 3-D channel: the narrowest walled extent (z = 20) sets W
    ch3 = capi.make_config(64, 40, 20, lattice=capi.D3Q19, tau=tau,
                           periodic=(1, 0, 0), force=(1e-6, 0, 0))
    assert capi.cell_reynolds(ch3) == pytest.approx(
        1e-6 * 20 ** 2 / (8 * nu) / nu, rel=1e-5)
    # no walls: the largest extent
    per = capi.make_config(64, 40, 20, lattice=capi.D3Q19, tau=tau,
                           periodic=(1, 1, 1), force=(1e-6, 0, 0))
    assert capi.cell_reynolds(per) == pytest.approx(
        1e-6 * 64 ** 2 / (8 * nu) / nu, rel=1e-5)
    # no drive: 0; invalid configuration: an error
    still = capi.make_config(16, 16, 16, tau=0.6)
    assert capi.cell_reynolds(still) == 0.0
    with pytest.raises(capi.LbmError):
        capi.cell_reynolds(capi.make_config(16, 16, 16, tau=0.5))
    with pytest.raises(capi.LbmError):
        capi.cell_reynolds(capi.make_config(4, 4, 4, lattice=27,
                                            mode=capi.MODE_REFERENCE))


def _c_layout(struct, fields, tmp_path):
    """offsetof of each field and sizeof of a struct of include/lbm.h, as
    the C compiler lays it out."""
    import os
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("%zu\\n", offsetof({struct}, {f}));\n'
                   for f in fields)
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n"
                   '#include "lbm.h"\nint main(void) {\n' + body +
                   f'  printf("%zu\\n", sizeof({struct}));\n  return 0;\n}}\n')
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(capi.ROOT, "include"), str(src),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True,
                         text=True).stdout.split()
    return [int(v) for v in out]


@pytest.mark.parametrize("cls,cname", [(capi.Config, "lbm_config"),
                                       (capi.Request, "lbm_request"),
                                       (capi.ServeParams, "lbm_serve_params"),
                                       (capi.ServeStats, "lbm_serve_stats")])
def test_struct_mirrors_match_header(cls, cname, tmp_path):
    # the ctypes mirrors have the header's fields, offsets and size
    fields = [f[0] for f in cls._fields_]
    got = _c_layout(cname, fields, tmp_path)
    assert got[:-1] == [getattr(cls, f).offset for f in fields]
    assert got[-1] == C.sizeof(cls)
