"""The oracle's FEM mesh fields (oracle/oracle_mesh.c), pinned three ways:

* bit for bit to the reference's TriLinearMesh3D / BiLinearMesh2D (golden vectors of
  tests/golden/make_golden.py: point location with the walk hint carried from point to
  point, barycentric coordinates, field and potential; grids, irregular meshes, the file
  loader) and to its engine (mesh-field runs in test_oracle.py);
* by the mathematics: a linear potential has the exact constant gradient in every cell,
  barycentric coordinates reconstruct the point, brute force over all cells finds the
  same containing cell, the axisymmetric field is radial in (x, y), potentials clamp;
* by the engine's wiring: a mesh whose field is exactly a uniform field (dyadic vertex
  coordinates and a power-of-two gradient) gives the uniform-field run bit for bit.
"""
import numpy as np
import pytest

KEYS = ("found", "bary", "hints", "E", "pot")


def _check(o, gold, tag):
    for k in KEYS:
        assert np.array_equal(o[k], gold[f"{tag}_{k}"]), (tag, k)


def test_grid3_queries_golden(oracle, gold):
    O = oracle
    m = O.mesh3_grid()
    _check(m.query(gold["mesh_pts3"], 0), gold, "mesh3_grid")
    (hx, hy, hz), (nx, ny, nz), _ = O.MESH3_GRID
    assert tuple(gold["mesh3_grid_counts"]) == ((nx + 1) * (ny + 1) * (nz + 1), 6 * nx * ny * nz)


def test_irregular3_queries_golden(oracle, gold):
    O = oracle
    xyz, V, tets = O.irregular_mesh3()
    m = O.Mesh.tet(xyz, V, tets, -30.0, 30.0)
    _check(m.query(gold["mesh_pts3i"], -1), gold, "mesh3_irr")


def test_file_loader_golden(oracle, gold, tmp_path):
    O = oracle
    cp, tp = O.write_mesh_files(str(tmp_path))
    m = O.Mesh.load3d(cp, tp, 1e-3, 1e-3)
    _check(m.query(gold["mesh_pts3i"], 7), gold, "mesh3_file")
    assert tuple(gold["mesh3_file_counts"]) == (120, 360)


def test_grid2_queries_golden(oracle, gold):
    O = oracle
    _check(O.mesh2_grid_rz().query(gold["mesh_pts3"], 0), gold, "mesh2_rz")
    (u0, u1), (v0, v1), (nu, nv), c = O.MESH2_GRID_XY
    _check(O.Mesh.grid2d(u0, u1, v0, v1, nu, nv, c, False).query(gold["mesh_pts3"], 5), gold, "mesh2_xy")
    uv, V, tris = O.irregular_mesh2()
    _check(O.Mesh.tri(uv, V, tris, False).query(gold["mesh_pts3i"], 2), gold, "mesh2_irr")


def test_linear_potential_has_exact_gradient(oracle):
    """V = a + b x + c y + d z is reproduced exactly by linear interpolation: the field of
    every cell is -(b, c, d) and the interpolated potential is V(p)."""
    O = oracle
    a, b, c, d = 0.7, 35.0, -12.0, 160.0
    m = O.Mesh.grid3d(0.03, 0.02, 0.012, 7, 5, 4, (a, b, c, d, 0.0, 0.0, 0.0))
    rng = np.random.default_rng(3)
    p = rng.uniform(-1, 1, (500, 3)) * [0.03, 0.02, 0.012]
    o = m.query(p, 0)
    assert o["found"].all()
    assert np.allclose(o["E"], [-b, -c, -d], rtol=1e-9, atol=0)
    assert np.allclose(o["pot"], a + b * p[:, 0] + c * p[:, 1] + d * p[:, 2], rtol=0, atol=1e-12)
    # the same in 2D, planar and axisymmetric (there V(r, z) = a + b r + d z)
    m2 = O.Mesh.grid2d(-0.03, 0.03, -0.02, 0.02, 6, 5, (a, b, c, 0.0, 0.0, 0.0), False)
    o = m2.query(p * [1, 1, 0], 1)
    assert o["found"].all() and np.allclose(o["E"], [-b, -c, 0.0], rtol=1e-9, atol=0)
    mr = O.Mesh.grid2d(0.0, 0.04, -0.012, 0.012, 8, 4, (a, b, d, 0.0, 0.0, 0.0), True)
    o = mr.query(p, 2)
    r = np.hypot(p[:, 0], p[:, 1])
    assert o["found"].all()
    assert np.allclose(o["E"], np.c_[-b * p[:, 0] / r, -b * p[:, 1] / r, np.full(len(p), -d)], rtol=1e-9, atol=1e-12)
    assert np.allclose(o["pot"], a + b * r + d * p[:, 2], rtol=0, atol=1e-12)


def _brute_force_cell(xyz, tets, p):
    """Cells containing p (all barycentric coordinates >= -1e-9), by solving every cell."""
    out = []
    for i, t in enumerate(tets):
        x = xyz[t]
        T = (x[:3] - x[3]).T
        lam = np.linalg.solve(T, p - x[3])
        if min(lam.min(), 1.0 - lam.sum()) >= -1e-9:
            out.append(i)
    return out


def test_located_cell_contains_point_brute_force(oracle):
    """The walk ends in a cell that brute force also finds to contain the point, its
    barycentric coordinates reconstruct the point and sum to 1, and points off the mesh
    are reported outside with a zero field."""
    O = oracle
    xyz, V, tets = O.irregular_mesh3()
    m = O.Mesh.tet(xyz, V, tets)
    st = np.sort(tets, axis=1)
    rng = np.random.default_rng(9)
    p = rng.uniform(-1.15, 1.15, (150, 3)) * [0.02, 0.015, 0.01]
    o = m.query(p, 0)
    for i in range(len(p)):
        inside_box = np.all(np.abs(p[i]) <= [0.02, 0.015, 0.01])
        cells = _brute_force_cell(xyz, st, p[i])
        if o["found"][i]:
            assert o["hints"][i] in cells
            b = o["bary"][i]
            assert abs(b.sum() - 1.0) < 1e-12 and b.min() >= -1e-10
            assert np.allclose(b @ xyz[st[o["hints"][i]]], p[i], rtol=0, atol=1e-14)
        else:
            assert not inside_box or not cells
            assert np.all(o["E"][i] == 0.0) and o["pot"][i] == 0.0


def test_hint_independence_for_interior_points(oracle):
    """For a point strictly inside one cell the result does not depend on where the walk
    starts (the reference keeps its start per thread; the device keeps it per carrier)."""
    O = oracle
    m = O.mesh3_grid()
    rng = np.random.default_rng(4)
    (hx, hy, hz), _, _ = O.MESH3_GRID
    p = rng.uniform(-1, 1, (300, 3)) * [hx, hy, hz]
    base = m.query(p, 0)
    ok = base["bary"].min(axis=1) > 1e-8
    assert ok.sum() > 250
    for h0 in (1, 431, 863, -5):
        o = m.query(p, h0)
        for i in np.flatnonzero(ok):
            # a fresh walk per point: the carried hint differs, the located cell does not
            o1 = m.query(p[i:i + 1], int(h0 if h0 >= 0 else 0) + i)
            assert o1["hints"][0] == base["hints"][i]
            assert np.array_equal(o1["E"][0], base["E"][i])
        assert np.array_equal(o["E"][ok], base["E"][ok])


def test_axisymmetric_field_is_radial(oracle):
    """mesh_field.hpp:412-417: E_r rotated back into (x, y) along (x, y) / r; on the axis
    only E_z remains."""
    O = oracle
    m = O.mesh2_grid_rz()
    rng = np.random.default_rng(5)
    p = rng.uniform(-0.02, 0.02, (200, 3)) * [1, 1, 0.5]
    p[:5, :2] = 0.0
    o = m.query(p, 0)
    r = np.hypot(p[:, 0], p[:, 1])
    on = r > 1e-15
    er = np.where(on, (o["E"][:, 0] * p[:, 0] + o["E"][:, 1] * p[:, 1]) / np.where(on, r, 1), 0.0)
    assert np.all(o["E"][~on, :2] == 0.0)
    assert np.allclose(o["E"][on, 0] * p[on, 1] - o["E"][on, 1] * p[on, 0], 0.0, atol=1e-12)
    # the same triangle and E_u whether the radius comes from (x, y) or is given directly
    flat = O.Mesh.grid2d(*O.MESH2_GRID_RZ[0], *O.MESH2_GRID_RZ[1], *O.MESH2_GRID_RZ[2], O.MESH2_GRID_RZ[3], False)
    q = flat.query(np.c_[r, p[:, 2], np.zeros(len(p))], 0)
    assert np.allclose(er[on], q["E"][on, 0], rtol=1e-12, atol=1e-12)
    assert np.array_equal(o["E"][:, 2], q["E"][:, 1])


def test_potential_clamp(oracle):
    """TriLinearMesh3D clamps vertex potentials to [V_min, V_max] at construction."""
    O = oracle
    xyz, V, tets = O.irregular_mesh3()
    m = O.Mesh.tet(xyz, V, tets, -30.0, 30.0)
    st = np.sort(tets, axis=1)
    for v in (0, 7, 33, 60, 119):
        o = m.query(xyz[v:v + 1] * (1 - 1e-9), 0)
        if o["found"][0]:
            cell = st[o["hints"][0]]
            w = o["bary"][0]
            assert abs(o["pot"][0] - w @ np.clip(V[cell], -30.0, 30.0)) < 1e-9


def test_loader_units_and_comments(oracle, tmp_path):
    """load_from_files: comment and blank lines skipped, coordinates and potentials scaled
    by the unit factors, the same mesh as constructing it from the arrays."""
    O = oracle
    cp, tp = O.write_mesh_files(str(tmp_path))
    xyz, V, tets = O.irregular_mesh3()
    a = O.Mesh.load3d(cp, tp, 1e-3, 1e-3)
    b = O.Mesh.tet(xyz, V, tets)
    p = O.mesh_query_points(100, (0.02, 0.015, 0.01), 3)
    qa, qb = a.query(p, 0), b.query(p, 0)
    ok = qb["bary"].min(axis=1) > 1e-6
    assert np.array_equal(qa["found"][ok], qb["found"][ok])
    assert np.allclose(qa["E"][ok], qb["E"][ok], rtol=1e-9, atol=1e-9)


def test_dyadic_mesh_field_equals_uniform_field(oracle):
    """A grid with dyadic vertex coordinates and V = -128 z has the field (0, 0, 128)
    exactly in every cell, so pairs drifting in it take the same steps, bit for bit, as
    in the uniform field: the mesh enters the charge step where the uniform field does."""
    O = oracle
    wl = O.workload(O.WL_PAIR)
    du = O.default_detector()
    du.efield[:] = (0.0, 0.0, 128.0)
    du.luke_sample_fraction = 0.0
    dm = O.default_detector()
    dm.luke_sample_fraction = 0.0
    O.set_mesh(dm, O.Mesh.grid3d(2.0 ** -4, 2.0 ** -4, 2.0 ** -6, 8, 8, 8, (0.0, 0.0, 0.0, -128.0, 0.0, 0.0, 0.0)))
    E = np.zeros(3)
    O.lib().orc_field_at(dm, O.dptr(np.array([0.011, -0.007, 0.003])), O.dptr(E))
    assert np.array_equal(E, [0.0, 0.0, 128.0])
    a = O.Engine(detector=du).run(wl, 6, 42, threads=2)
    b = O.Engine(detector=dm).run(wl, 6, 42, threads=2)
    assert a["stats_raw"] == b["stats_raw"]
    assert np.array_equal(a["hits"], b["hits"])


@pytest.mark.parametrize("dim", [3, 2])
def test_degenerate_cells_have_zero_field(oracle, dim):
    """Cells with |det| <= 1e-30 get a zero inverse and a zero field (mesh_field.hpp:305-309)."""
    O = oracle
    if dim == 3:
        xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.0]]) * 0.01
        m = O.Mesh.tet(xyz, [0.0, 1.0, 2.0, 3.0], [[0, 1, 2, 3]])
    else:
        uv = np.array([[0, 0], [1, 0], [2, 0.0]]) * 0.01
        m = O.Mesh.tri(uv, [0.0, 1.0, 2.0], [[0, 1, 2]], False)
    o = m.query(np.array([[0.0, 0.0, 0.0]]), 0)
    assert np.all(o["E"] == 0.0)
