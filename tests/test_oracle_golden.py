"""Pins the CPU oracle against the reference's own known-answer tests,
closed forms and invariants (CPU only).

Each golden value below is the one the reference's gtest suites assert;
the test name cites the reference test (file:line) it re-expresses.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

E27, INV27, _ = O.std_tables(27)
SPLIT = np.array([0.0] + [1.0] * 6 + [0.5] * 12 + [1.0 / 3.0] * 8, np.float32)


def idx(x, y, z, dims):
    return z + dims[2] * (y + dims[1] * x)


# ------------------------------------------------------------------ tables
def test_tables_shell_order_and_inverse():
    # Constants.cpp:6-7: shells rest/axis/edge/corner, inverse is a mirror.
    norms = (E27 ** 2).sum(1)
    assert list(norms) == [0] + [1] * 6 + [2] * 12 + [3] * 8
    for i in range(27):
        assert (E27[INV27[i]] == -E27[i]).all()
    assert list(INV27[1:7]) == [7 - i for i in range(1, 7)]
    assert list(INV27[7:19]) == [25 - i for i in range(7, 19)]
    assert list(INV27[19:27]) == [45 - i for i in range(19, 27)]
    for q in (9, 19, 27):
        e, inv, w = O.std_tables(q)
        assert abs(w.sum() - 1.0) < 1e-15
        # second-moment isotropy: sum w e_a e_b = delta_ab / 3
        m2 = np.einsum("i,ia,ib->ab", w, e.astype(float), e.astype(float))
        dim = 2 if q == 9 else 3
        assert np.allclose(m2[:dim, :dim], np.eye(dim) / 3.0, atol=1e-15)
        for i in range(q):
            assert (e[inv[i]] == -e[i]).all()
    # D3Q19 is the prefix of the D3Q27 table; D2Q9 the z=0 subset in order
    e19, _, _ = O.std_tables(19)
    assert (e19 == E27[:19]).all()
    e9, _, _ = O.std_tables(9)
    z0 = [i for i in range(27) if E27[i][2] == 0]
    assert (e9 == E27[z0]).all()


# ------------------------------------------------------------ reference mode
def test_streaming_golden():
    # CalculationsTest.cpp:19-45
    dims = (1, 2, 1)
    f = np.zeros(27 * 2, np.float32)
    f[5 * 2 + 0] = 2.0
    t = O.ref_stream(dims, f)
    assert abs(t[5 * 2 + 1] - 2.0) < 1e-6
    assert t[5 * 2 + 0] == 0.0


def test_streaming_is_shift_with_zero_fill():
    # Brute force on a tiny box: a unit value at one node in direction j
    # lands at node + e_j, or vanishes when that node is outside.
    dims = (3, 4, 5)
    n = 60
    for j in range(27):
        for src in [(0, 0, 0), (1, 2, 3), (2, 3, 4), (1, 0, 4)]:
            f = np.zeros(27 * n, np.float32)
            f[j * n + idx(*src, dims)] = 1.0
            t = O.ref_stream(dims, f)
            dst = tuple(np.array(src) + E27[j])
            inside = all(0 <= dst[a] < dims[a] for a in range(3))
            assert t.sum() == (1.0 if inside else 0.0)
            if inside:
                assert t[j * n + idx(*dst, dims)] == 1.0


def test_streaming_mass_closed_form():
    rng = np.random.default_rng(0)
    dims = (5, 6, 7)
    n = 210
    f = rng.random(27 * n).astype(np.float32)
    t = O.ref_stream(dims, f).reshape(27, *dims)
    fr = f.reshape(27, *dims)
    for j in range(27):
        # surviving sources are the nodes y with y + e_j inside the box
        sl = tuple(slice(0, dims[a] - 1) if E27[j][a] == 1 else
                   (slice(1, dims[a]) if E27[j][a] == -1 else slice(None))
                   for a in range(3))
        assert np.isclose(t[j].sum(dtype=np.float64),
                          fr[j][sl].sum(dtype=np.float64), rtol=1e-6)


def test_collision_golden():
    # CalculationsTest.cpp:47-75
    f = np.zeros(27, np.float32)
    f[0] = 2.0
    out = O.ref_collide((1, 1, 1), f, 1.0)
    assert abs(out[0] - 0.592592593) < 1e-6
    assert abs(out[1] - 0.148148148) < 1e-6
    assert abs(out[2] - 0.148148148) < 1e-6


def test_full_golden():
    # CalculationsTest.cpp:77-105 (stream then collide; pins SPLIT u_y = 1)
    f = np.zeros(54, np.float32)
    f[5 * 2 + 0] = 2.0
    out = O.ref_run_lattice((1, 2, 1), f, 1.0, 1)
    assert abs(out[0 * 2 + 1] - -0.296296296) < 1e-6
    assert abs(out[1 * 2 + 1] - -0.074074074) < 1e-6
    assert abs(out[2 * 2 + 1] - 0.148148148) < 1e-6


def test_collision_mass_and_split_momentum():
    # sum eq = rho for any u (isotropy); with tau = 1 the post-collision
    # momentum is rho * u_split, u_split = sum f e SPLIT / rho.
    rng = np.random.default_rng(1)
    n = 64
    f = (0.001 + 0.001 * rng.random(27 * n)).astype(np.float32)
    out = O.ref_collide((4, 4, 4), f, 1.0).reshape(27, n).astype(np.float64)
    fr = f.reshape(27, n).astype(np.float64)
    assert np.allclose(out.sum(0), fr.sum(0), rtol=2e-6)
    mom_after = E27.T.astype(np.float64) @ out
    mom_split = (E27.T * SPLIT).astype(np.float64) @ fr
    assert np.allclose(mom_after, mom_split, rtol=1e-4, atol=1e-9)
    # relaxation with tau = 2 moves half way
    out2 = O.ref_collide((4, 4, 4), f, 2.0).reshape(27, n).astype(np.float64)
    assert np.allclose(out2, 0.5 * (fr + out), rtol=1e-5, atol=1e-10)


def _one_node(values):
    f = np.zeros(27, np.float32)
    for i, v in values.items():
        f[i] = v
    return f


@pytest.mark.parametrize("values,rho,u", [
    # f5 = 2, e5 = (0, 1, 0), |e|^2 = 1
    ({5: 2.0}, 2.0, (0.0, 1.0, 0.0)),
    # f7 = 2, e7 = (-1, -1, 0), SPLIT 1/2: u = 2 (-1, -1, 0) / 2 / 2
    ({7: 2.0}, 2.0, (-0.5, -0.5, 0.0)),
    # f19 = 3, e19 = (-1, -1, -1), SPLIT 1/3
    ({19: 3.0}, 3.0, (-1 / 3, -1 / 3, -1 / 3)),
    # all three shells at once, worked by hand: rho = 2,
    # sum f e SPLIT = 0.3 (1,0,0) + 0.2 (0,1,1)/2 + 0.6 (-1,1,1)/3
    #              = (0.1, 0.3, 0.3)  (without SPLIT it would be
    #                (-0.3, 0.8, 0.8))
    ({0: 0.9, 6: 0.3, 14: 0.2, 22: 0.6}, 2.0, (0.05, 0.15, 0.15)),
    # zero density: the velocity stays the undivided sum (density != 0
    # branch of SimCalcFuncs.cu:115-120), here 0
    ({}, 0.0, (0.0, 0.0, 0.0)),
    # zero density with nonzero populations: the sum is not divided
    ({1: 1.0, 6: -1.0}, 0.0, (-2.0, 0.0, 0.0)),
])
def test_reference_moments_hand_computed(values, rho, u):
    # SimCalcFuncs.cu:102-120: rho = sum f_i, u = sum f_i e_i SPLIT_i / rho
    r, v = O.ref_moments((1, 1, 1), _one_node(values))
    assert abs(r[0] - rho) < 1e-6
    assert np.allclose(v[:, 0], u, atol=1e-6), v[:, 0]


def test_reference_moments_node_order():
    # plane i of the population array holds direction i for every node in
    # index order n = z + nz (y + ny x): distinct nodes keep their values
    dims = (2, 3, 2)
    n = 12
    f = np.zeros(27 * n, np.float32)
    for k in range(n):
        f[6 * n + k] = 0.1 * (k + 1)      # e6 = (1, 0, 0)
        f[0 * n + k] = 1.0 - 0.1 * (k + 1)
    r, v = O.ref_moments(dims, f)
    assert np.allclose(r, 1.0, atol=1e-6)
    assert np.allclose(v[0], 0.1 * np.arange(1, n + 1), atol=1e-6)
    assert np.allclose(v[1:], 0.0)


def test_reference_moments_match_collision_momentum():
    # With tau = 1 the collision (pinned by CalculationsTest.cpp:47-105)
    # returns feq(rho, u) whose first moment is rho u exactly (D3Q27
    # isotropy), so the momentum after collision must equal rho * u from the
    # moments readback on the same state.
    rng = np.random.default_rng(11)
    dims = (3, 2, 4)
    n = 24
    f = (0.01 + 0.02 * rng.random(27 * n)).astype(np.float32)
    r, v = O.ref_moments(dims, f)
    out = O.ref_collide(dims, f, 1.0).reshape(27, n).astype(np.float64)
    mom = E27.T.astype(np.float64) @ out
    assert np.allclose(out.sum(0), r, rtol=2e-6)
    assert np.allclose(mom, r.astype(np.float64) * v, rtol=1e-4, atol=1e-8)


def test_collision_rest_state_fixed_point():
    w = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)
    f = np.repeat((w * 1.3).astype(np.float32), 8)
    out = O.ref_collide((2, 2, 2), f, 0.6)
    assert np.allclose(out, f, rtol=1e-6)


def _angled_setup(data_dir):
    dims = (3, 3, 3)
    mesh = O.Mesh(os.path.join(data_dir, "cube_angled.obj"))
    ok, internal = mesh.insert(dims, 0.5)
    assert ok
    flags, recs = mesh.preprocess(dims, 0.5, internal * O.FLAG_INTERNAL)
    return dims, mesh, flags, recs


def test_reflections_golden(data_dir):
    # CalculationsTest.cpp:107-146 and :148-197 (force)
    dims, mesh, flags, recs = _angled_setup(data_dir)
    f = np.zeros(27 * 27, np.float32)
    f[1 * 27 + idx(1, 1, 1, dims)] = 1.0
    force = np.zeros(3, np.float32)
    t = O.ref_reflect(dims, f, flags, recs, force)
    assert abs(t[17 * 27 + idx(2, 1, 2, dims)] - 0.367402) < 1e-6
    assert abs(t[4 * 27 + idx(1, 1, 2, dims)] - 0.341121) < 1e-6
    assert abs(t[26 * 27 + idx(2, 2, 2, dims)] - 0.291477) < 1e-6
    assert abs(t[1 * 27 + idx(1, 1, 1, dims)]) < 1e-6
    assert abs(force[0] - 0.28086) < 1e-6
    assert abs(force[1] - -0.194318) < 1e-6
    assert abs(force[2] - 0.648261) < 1e-6
    # without the force accumulator the populations are identical
    t2 = O.ref_reflect(dims, f, flags, recs, None)
    assert (t2 == t).all()


def test_reflection_force_closed_form(data_dir):
    # x = w17*1/2 + w26*1/3; y = w26*(1/3 - 1); z = w17*(1/2 + 1) + w26*1/3
    # (slots 3..5 pick DIRECTIONS[3..5] in the second term).
    dims, mesh, flags, recs = _angled_setup(data_dir)
    node, dirs, w = recs
    r = list(node).index(idx(1, 1, 1, dims))
    w17, w26 = w[r][3], w[r][5]
    f = np.zeros(27 * 27, np.float32)
    f[1 * 27 + idx(1, 1, 1, dims)] = 1.0
    force = np.zeros(3, np.float32)
    O.ref_reflect(dims, f, flags, recs, force)
    assert np.allclose(force, [w17 / 2 + w26 / 3, w26 * (1 / 3 - 1),
                               w17 * 1.5 + w26 / 3], atol=1e-6)


def test_flow_generation_golden():
    # CalculationsTest.cpp:199-234 and :236-277
    dims = (3, 3, 3)
    induced = np.zeros(27, np.float32)
    induced[0] = 1.0
    f = np.zeros(27 * 27, np.float32)
    out = O.ref_flow(dims, f, O.FLOW_POINT, [1, 1, 1], induced)
    assert abs(out[idx(1, 1, 1, dims)] - 1.0) < 1e-6
    assert abs(out[idx(0, 0, 0, dims)]) < 1e-6
    out = O.ref_flow(dims, f, O.FLOW_PLANE, [0, 0], induced)
    for y in range(3):
        for z in range(3):
            assert abs(out[idx(0, y, z, dims)] - 1.0) < 1e-6
            assert abs(out[idx(1, y, z, dims)]) < 1e-6


def test_compress_ints_golden():
    # CalculationsTest.cpp:279-284
    assert O.compress_ints([1, 2, 3, 4, 5, 6], 5) == 35754150


# ------------------------------------------------------------------ geometry
def test_model_import_and_closed(data_dir):
    # ModelTest.cpp:5-30
    tri = O.Mesh(os.path.join(data_dir, "triangle.obj"))
    assert tri.nverts == 3
    assert not tri.closed
    cube = O.Mesh(os.path.join(data_dir, "cube.obj"))
    assert cube.closed and cube.valid


def test_model_inside(data_dir):
    # ModelTest.cpp:32-45
    cube = O.Mesh(os.path.join(data_dir, "cube.obj"))
    assert cube.inside(0.0, 0.0, 0.0)
    assert not cube.inside(2.0, 0.0, 0.0)
    # boundary counts as inside; corners, edges and faces
    assert cube.inside(1.0, 0.0, 0.0)
    assert cube.inside(1.0, 1.0, 1.0)
    assert cube.inside(1.0, 1.0, 0.3)
    assert not cube.inside(1.0000001, 0.0, 0.0)


def test_model_check_error(data_dir):
    # ModelTest.cpp:47-54
    quad = O.Mesh(os.path.join(data_dir, "quad.obj"))
    assert not quad.valid


def test_model_intersection_normal(data_dir):
    # ModelTest.cpp:56-74
    m = O.Mesh(os.path.join(data_dir, "compat_cube.obj"))
    r = m.reflection_vector((0.5, 0.5, -1.0), (0.5, 0.5, 1.0))
    assert np.allclose(r, [0, 0, -1], atol=1e-6)
    r = m.reflection_vector((0.5, 0.5, -1.0), (0.5, 0.5, -2.0))
    assert (r == 0).all()


def test_model_inside_brute_force(data_dir):
    # the angled cube is convex: inside <=> on the inner side of all planes
    m = O.Mesh(os.path.join(data_dir, "cube_angled.obj"))
    rng = np.random.default_rng(3)
    pts = rng.random((400, 3)).astype(np.float32)
    for p in pts:
        x, y, z = (float(v) for v in p)
        xmax = 0.9 - 0.625 * (z - 0.1)
        expect = 0.1 <= x <= xmax and 0.1 <= y <= 0.9 and 0.1 <= z <= 0.9
        margin = min(abs(x - 0.1), abs(x - xmax), abs(y - 0.1), abs(y - 0.9),
                     abs(z - 0.1), abs(z - 0.9))
        if margin > 1e-5:
            assert m.inside(x, y, z) == expect, (x, y, z)


def test_coordinates_with_model(data_dir):
    # CoordinatesTest.cpp:137-169 (boundary inclusive: 1*0.1f == 0.1f)
    m = O.Mesh(os.path.join(data_dir, "compat_cube.obj"))
    ok, internal = m.insert((10, 10, 10), 0.1)
    assert ok
    for y in range(10):
        for x in range(10):
            assert not internal[10 * y + x]
    for y in range(1, 9):
        for x in range(1, 9):
            assert internal[500 + 10 * y + x]


def test_insert_model_outside_box_is_ignored(data_dir):
    # Lattice.cpp:236-244: a mesh not inside [0, N*h] leaves flags alone
    m = O.Mesh(os.path.join(data_dir, "cube.obj"))
    ok, _ = m.insert((4, 4, 4), 1.0)
    assert not ok


def test_distribute_vector_golden():
    # CoordinatesTest.cpp:171-198 (18 before 17 pins the tie-break)
    sc, ix = O.distribute_vector((1.0, 0.1, 0.1))
    top = sc[:3] / sc[:3].sum()
    assert list(ix[:3]) == [6, 18, 17]
    assert np.allclose(top, [0.4134132615337746, 0.2932933692331128,
                             0.2932933692331128], atol=1e-6)
    # scores are 2 - distance to each unit direction, all 27 present once
    assert sorted(ix) == list(range(27))
    assert (np.diff(sc) <= 0).all()


def test_preprocess_reflection_golden(data_dir):
    # CoordinatesTest.cpp:200-229
    dims, mesh, flags, recs = _angled_setup(data_dir)
    node, dirs, w = recs
    assert list(node) == [13]
    assert flags[13] & O.FLAG_REFLECTED
    assert np.allclose(w[0][3:6], [0.367402, 0.341121, 0.291477], atol=1e-6)
    assert [O.slot_dir(dirs[0], s) for s in (3, 4, 5)] == [17, 4, 26]
    # every used slot group is normalised and targets a fluid node
    for k in range(1, 27):
        base = O.INV = INV27[k] * 3
        ds = [O.slot_dir(dirs[0], base + r) for r in range(3)]
        assert all(d != 0 for d in ds)
        assert abs(w[0][base:base + 3].sum() - 1.0) < 1e-6


def test_quaternion_rotation_golden():
    # GeometryTest.cpp:11-32
    axis = (np.cos(np.pi * 0.25), np.sin(np.pi * 0.25), 0.0)
    r = O.rotate_vector(axis, (1.0, 0.0, 0.0), np.pi)
    assert np.allclose(r, [0, 1, 0], atol=1e-6)
    r = O.rotate_vector(axis, (1.0, 0.0, 0.0), np.pi / 2)
    assert np.allclose(r, [0.5, 0.5, -np.cos(np.pi * 0.25)], atol=1e-6)


# ------------------------------------------------------------ standard mode
def _box_walls_count(q, dims):
    e, _, _ = O.std_tables(q)
    n = int(np.prod(dims))
    total = 0
    for i in range(1, q):
        inner = 1
        for a in range(3):
            inner *= dims[a] - abs(int(e[i][a]))
        total += n - inner
    return total


@pytest.mark.parametrize("q,dims", [(19, (4, 5, 6)), (27, (3, 4, 5)),
                                    (9, (7, 5, 1))])
def test_mask_wall_link_count_closed_form(q, dims):
    mask = O.std_build_mask(q, dims, None, (0, 0, 0))
    bits = sum(bin(int(m)).count("1") for m in mask)
    assert bits == _box_walls_count(q, dims)
    # periodic everywhere: no bounced links
    mask = O.std_build_mask(q, dims, None, (1, 1, 1))
    assert (mask == 0).all()


def test_mask_obstacle_links():
    q, dims = 19, (3, 3, 3)
    solid = np.zeros(27, np.uint8)
    solid[13] = 1
    e, _, _ = O.std_tables(q)
    mask = O.std_build_mask(q, dims, solid, (1, 1, 1))
    assert mask[13] == O.MASK_SOLID
    for i in range(1, q):
        # the node at centre + e_i pulls direction i from the centre
        x = tuple((1 + e[i][a]) % 3 for a in range(3))
        assert mask[idx(*x, dims)] & (1 << i)


def test_mt19937_64_known_answer():
    # C++ [rand.predef]: 10000th output of default-seeded mt19937_64
    assert O.mt64_nth(5489, 10000) == 9981545732273789042


def test_random_spheres_voxelisation():
    dims = (40, 24, 20)
    solid, sph = O.random_spheres(dims, 2, 0.2, 4.0, 8.0, True)
    n = int(np.prod(dims))
    assert solid.sum() >= 0.2 * n
    # brute force over all nodes, periodic in x, from the sphere list
    x, y, z = np.meshgrid(np.arange(40), np.arange(24), np.arange(20),
                          indexing="ij")
    ref = np.zeros(dims, bool)
    for cx, cy, cz, r in sph:
        dx = np.abs(x - cx)
        dx = np.minimum(dx, 40 - dx)
        ref |= dx * dx + (y - cy) ** 2 + (z - cz) ** 2 <= r * r
    assert (ref.reshape(-1) == solid.astype(bool)).all()
    # the last sphere was needed: without it the fraction is below 20%
    ref2 = np.zeros(dims, bool)
    for cx, cy, cz, r in sph[:-1]:
        dx = np.abs(x - cx)
        dx = np.minimum(dx, 40 - dx)
        ref2 |= dx * dx + (y - cy) ** 2 + (z - cz) ** 2 <= r * r
    assert ref2.sum() < 0.2 * n


def test_storage_rounding():
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.normal(0, 1e-2, 2000),
                           rng.normal(0, 1e-6, 500),
                           rng.normal(0, 1e-4, 500)]).astype(np.float32)
    for v in vals:
        assert O.round_fp16(float(v)) == float(np.float32(v).astype(np.float16))
    # bf16: nearest of the two neighbouring bf16 values, ties to even
    for v in vals[:1000]:
        b = np.float32(v).view(np.uint32)
        lo = np.uint32(b & 0xFFFF0000)
        hi = np.uint32(lo + 0x10000)
        flo, fhi = lo.view(np.float32), hi.view(np.float32)
        r = np.float32(O.round_bf16(float(v)))
        dl, dh = abs(float(v) - float(flo)), abs(float(fhi) - float(v))
        if dl < dh:
            assert r == flo
        elif dh < dl:
            assert r == fhi
        else:
            assert r == (flo if (lo >> 16) % 2 == 0 else fhi)


def _periodic_run(q, dims, f, steps, **kw):
    mask = O.std_build_mask(q, dims, None, (1, 1, 1))
    p = O.std_params(periodic=(1, 1, 1), **kw)
    return mask, O.std_run_f64(q, dims, f, mask, p, steps)


@pytest.mark.parametrize("q", [9, 19, 27])
def test_std_equilibrium_fixed_point(q):
    dims = (6, 5, 1 if q == 9 else 4)
    f = O.std_init_eq(q, dims)
    _, out = _periodic_run(q, dims, f, 5, tau=0.7)
    assert np.allclose(out, f, rtol=1e-14, atol=1e-16)


@pytest.mark.parametrize("q", [9, 19, 27])
def test_std_conservation_and_guo_momentum(q):
    rng = np.random.default_rng(q)
    dims = (6, 5, 1 if q == 9 else 4)
    n = int(np.prod(dims))
    u = 0.02 * rng.standard_normal((3, n))
    if q == 9:
        u[2] = 0
    f = O.std_init_eq(q, dims, 1 + 0.01 * rng.standard_normal(n), u)
    f += 1e-4 * rng.standard_normal(f.shape)
    e, _, _ = O.std_tables(q)
    fr = f.reshape(q, n)
    m0, p0 = fr.sum(), e.T @ fr.sum(1)
    F = (1e-5, -2e-5, 0.0 if q == 9 else 3e-5)
    steps = 7
    _, out = _periodic_run(q, dims, f, steps, tau=0.8, force=F)
    o = out.reshape(q, n)
    assert abs(o.sum() - m0) < 1e-12 * m0
    # every step adds F to every node's momentum (Guo)
    F32 = np.array(F, np.float32).astype(np.float64)  # params hold floats
    assert np.allclose(e.T @ o.sum(1), p0 + steps * n * F32, rtol=0,
                       atol=1e-12)


def test_std_walls_conserve_mass():
    q, dims = 19, (6, 7, 5)
    rng = np.random.default_rng(7)
    solid = (rng.random(int(np.prod(dims))) < 0.1).astype(np.uint8)
    mask = O.std_build_mask(q, dims, solid, (1, 0, 0))
    f = O.std_init_eq(q, dims)
    f += 1e-3 * rng.random(f.shape)
    p = O.std_params(tau=0.9, periodic=(1, 0, 0))
    out = O.std_run_f64(q, dims, f, mask, p, 10)
    fl = solid == 0
    n = int(np.prod(dims))
    assert np.isclose(out.reshape(q, n)[:, fl].sum(),
                      f.reshape(q, n)[:, fl].sum(), rtol=1e-13)


def test_std_poiseuille_d2q9():
    # body-force channel between half-way walls at y = -1/2 and ny - 1/2:
    # u(y) = F/(2 rho nu) (y + 1/2)(ny - 1/2 - y); exact for BGK at
    # tau = 1/2 + sqrt(3/16).
    q, ny = 9, 16
    dims = (2, ny, 1)
    tau = 0.5 + np.sqrt(3.0 / 16.0)
    nu = (tau - 0.5) / 3
    F = 1e-6
    mask = O.std_build_mask(q, dims, None, (1, 0, 0))
    p = O.std_params(tau=tau, force=(F, 0, 0), periodic=(1, 0, 0))
    f = O.std_init_eq(q, dims)
    out = O.std_run_f64(q, dims, f, mask, p, 6000)
    rho, u = O.std_moments(q, dims, out, mask, (F, 0, 0))
    y = np.arange(ny)
    ua = F / (2 * nu) * (y + 0.5) * (ny - 0.5 - y)
    assert np.allclose(u[0][:ny], ua, rtol=2e-3)


def test_std_couette_moving_wall():
    # moving lid at y = ny - 1/2 with U, still wall at y = -1/2:
    # u(y) = U (y + 1/2) / ny at steady state.
    q, ny, U = 9, 12, 0.01
    dims = (2, ny, 1)
    mask = O.std_build_mask(q, dims, None, (1, 0, 0), lid_face=3)
    p = O.std_params(tau=0.8, lid_face=3, lid_u=(U, 0, 0),
                     periodic=(1, 0, 0))
    f = O.std_init_eq(q, dims)
    out = O.std_run_f64(q, dims, f, mask, p, 8000)
    _, u = O.std_moments(q, dims, out, mask)
    ua = U * (np.arange(ny) + 0.5) / ny
    assert np.allclose(u[0][:ny], ua, rtol=1e-3, atol=1e-7)


@pytest.mark.parametrize("q", [9, 19])
def test_std_taylor_green_decay(q):
    # u = u0 sin(kx) cos(ky) e^{-2 nu k^2 t}: kinetic energy decays as
    # e^{-4 nu k^2 t}.  The rate is measured after a 100-step transient;
    # BGK's error is second order in k (0.34% at L = 32, 0.07% at L = 64).
    L, u0, tau = 32, 0.02, 0.8
    nz = 1 if q == 9 else 2
    dims = (L, L, nz)
    nu = (tau - 0.5) / 3
    k = 2 * np.pi / L
    x, y, z = np.meshgrid(np.arange(L), np.arange(L), np.arange(nz),
                          indexing="ij")
    ux = u0 * np.sin(k * x) * np.cos(k * y)
    uy = -u0 * np.cos(k * x) * np.sin(k * y)
    rho = 1 - 3 * u0 ** 2 / 4 * (np.cos(2 * k * x) + np.cos(2 * k * y))
    u = np.stack([ux, uy, 0 * ux]).reshape(3, -1)
    f = O.std_init_eq(q, dims, rho.reshape(-1), u)
    mask, f1 = _periodic_run(q, dims, f, 100, tau=tau)
    _, f2 = _periodic_run(q, dims, f1, 200, tau=tau)
    _, u1 = O.std_moments(q, dims, f1, mask)
    _, u2 = O.std_moments(q, dims, f2, mask)
    ratio = (u2 ** 2).sum() / (u1 ** 2).sum()
    assert abs(ratio / np.exp(-4 * nu * k * k * 200) - 1) < 5e-3


def test_std_f32_tracks_f64():
    q, dims = 19, (8, 6, 5)
    rng = np.random.default_rng(11)
    solid = (rng.random(int(np.prod(dims))) < 0.15).astype(np.uint8)
    mask = O.std_build_mask(q, dims, solid, (1, 0, 0), lid_face=3)
    p = O.std_params(tau=0.8, force=(1e-5, 0, 0), lid_face=3,
                     lid_u=(0.05, 0, 0), periodic=(1, 0, 0))
    f = O.std_init_eq(q, dims)
    o64 = O.std_run_f64(q, dims, f, mask, p, 50)
    o32 = O.std_run_f32(q, dims, f, mask, p, 50)
    assert np.abs(o32 - o64).max() < 1e-6
    # the shifted form (f_i - w_i, collided from small terms) is closer still
    o32s = O.std_run_f32(q, dims, f, mask, p, 50, O.STORE_F32_SHIFTED)
    assert np.abs(o32s - o64).max() < 0.2 * np.abs(o32 - o64).max()
    # shifted fp16 storage stays close; bf16 is coarser
    o16 = O.std_run_f32(q, dims, f, mask, p, 50, O.STORE_F16_SHIFTED)
    ob16 = O.std_run_f32(q, dims, f, mask, p, 50, O.STORE_BF16_SHIFTED)
    assert np.abs(o16 - o64).max() < 1e-4
    assert np.abs(ob16 - o64).max() < 2e-3


def test_std_fp16_forced_channel_does_not_stall():
    # a channel spun up from rest by a small body force: each step moves the
    # stored deviations by ~3 w F (1 - omega/2) ~ 3e-9, far below half a
    # binary16 unit in the last place of values near 1e-5.  Rounding to
    # nearest would return the old value every step; the stochastic rounding
    # of forced lattices (with the 2^12 scale) keeps the velocity within
    # 1e-3 of the FP64 run.
    q, dims = 9, (8, 64, 1)
    mask = O.std_build_mask(q, dims, None, (1, 0, 0))
    p = O.std_params(tau=0.64, force=(5e-8, 0, 0), periodic=(1, 0, 0))
    f = O.std_init_eq(q, dims)
    o64 = O.std_run_f64(q, dims, f, mask, p, 2000)
    o16 = O.std_run_f32(q, dims, f, mask, p, 2000, O.STORE_F16_SHIFTED)
    _, u64 = O.std_moments(q, dims, o64, mask, (5e-8, 0, 0))
    _, u16 = O.std_moments(q, dims, o16, mask, (5e-8, 0, 0))
    err = np.sqrt(((u16 - u64) ** 2).sum() / (u64 ** 2).sum())
    assert err < 1e-3
    # BF16 (8-bit mantissa) also tracks the flow instead of stalling
    ob = O.std_run_f32(q, dims, f, mask, p, 2000, O.STORE_BF16_SHIFTED)
    _, ub = O.std_moments(q, dims, ob, mask, (5e-8, 0, 0))
    assert np.sqrt(((ub - u64) ** 2).sum() / (u64 ** 2).sum()) < 2e-2


def _f16_neighbours(y):
    # the two binary16 values around y (y itself twice when representable)
    h = np.float16(y)
    if float(h) == y:
        return float(h), float(h)
    lo = h if float(h) < y else np.nextafter(h, np.float16(-np.inf))
    hi = np.nextafter(lo, np.float16(np.inf))
    return float(lo), float(hi)


def _bf16_neighbours(v):
    b = int(np.float32(v).view(np.uint32))
    t = np.uint32(b & 0xFFFF0000).view(np.float32)
    u = np.uint32((b & 0xFFFF0000) + 0x10000).view(np.float32)
    return tuple(sorted((float(t), float(u))))


@pytest.mark.parametrize("storage", [O.STORE_F16_SHIFTED, O.STORE_BF16_SHIFTED])
def test_stochastic_rounding_is_unbiased(storage):
    # The 16-bit store of a forced lattice: every result is one of the two
    # neighbouring 16-bit values, and over the nodes of a lattice (the node
    # hash: low 24 bits of the element index in the blocked layout times
    # an odd constant) its mean is the stored value itself -- rounding to
    # nearest would be biased by up to half a unit for the same values.
    # A wrong shift, mask, multiplier or dropped hash fails one of the two.
    rng = np.random.default_rng(11)
    n = np.arange(3000, dtype=np.uint64) * 7 + 5  # nodes of a D3Q19 lattice
    own = (n >> np.uint64(6)) * np.uint64(19 * 64) + (n & np.uint64(63))
    hashes = ((own & np.uint64(0xFFFFFF)) * np.uint64(0x5BD1E9)) & np.uint64(0xFFFFFFFF)
    scale = 4096.0 if storage == O.STORE_F16_SHIFTED else 1.0
    # deviations of 1e-6 .. 1e-1, and (FP16: below 2^-14 once scaled, the
    # subnormal range, where the conversion rounds stochastically at the
    # subnormal unit) 1e-12 .. 1e-9
    vals = np.concatenate([rng.choice([-1, 1], 24) * 10 ** rng.uniform(-6, -1, 24),
                           rng.choice([-1, 1], 8) * 10 ** rng.uniform(-12, -9, 8)])
    vals = vals.astype(np.float32)
    for j, v in enumerate(vals):
        v = float(v)
        if storage == O.STORE_F16_SHIFTED:
            lo, hi = _f16_neighbours(np.float32(v * scale))
        else:
            lo, hi = _bf16_neighbours(v)
        if lo == hi:
            continue
        d = j % 19  # each direction's multiplier
        got = np.array([O.store16_sr(storage, v, int(h), d) * scale for h in hashes])
        assert np.all((got == lo) | (got == hi)), (v, set(got) - {lo, hi})
        p = (v * scale - lo) / (hi - lo)  # exact rounding-up probability
        sd = (hi - lo) * np.sqrt(p * (1 - p) / len(hashes))
        assert abs(got.mean() - v * scale) <= 4.5 * sd + 1e-12 * abs(v * scale), (v, got.mean(), sd)
