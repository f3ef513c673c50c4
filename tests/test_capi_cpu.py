"""The C ABI library without a GPU: it loads, exports every symbol include/cmpmc_amd.h
declares, its struct sizes match the bindings, and its host-side pieces (presets, the
CrystalLattice table build, host map_k_to_vg, argument checks) equal the oracle's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cmpmc_amd.h")
GOLD = os.path.join(ROOT, "tests", "golden", "reference_vectors.npz")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(cmpmc_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol(lib_cmpmc):
    L = lib_cmpmc.capi.load()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/cmpmc_amd.h but not exported"
    assert set(lib_cmpmc.capi.EXPORTED_SYMBOLS) <= set(names)


def test_struct_sizes(lib_cmpmc):
    L = lib_cmpmc.capi.load()
    lib_cmpmc.capi.check_abi(L)
    # byte layouts of the reference's records on x86-64, which the C ABI mirrors:
    # PhononTrack and ChargeTrack (alignas(64), core/track.hpp:62,75), SensorHit
    # (core/track.hpp:99), SimStats (transport/engine.hpp:29)
    assert L.cmpmc_abi_sizeof(5) == 128 and L.cmpmc_abi_sizeof(6) == 128
    assert L.cmpmc_abi_sizeof(7) == 56 and L.cmpmc_abi_sizeof(8) == 208


def test_presets_equal_oracle(lib_cmpmc, oracle):
    a = lib_cmpmc.germanium()
    b = oracle.germanium()
    for name, _ in lib_cmpmc.capi.Crystal._fields_:
        va, vb = getattr(a, name), getattr(b, name if name != "lambda_" else "lambda_")
        if hasattr(va, "__len__"):
            assert np.array_equal(np.ctypeslib.as_array(va), np.ctypeslib.as_array(vb)), name
        else:
            assert va == vb, name
    da, db = lib_cmpmc.default_detector(), oracle.default_detector()
    assert da.radius == db.radius and da.half_z == db.half_z and da.max_bounces == db.max_bounces
    for s in range(3):
        assert da.surf[s].has_film == db.surf[s].has_film
        assert da.surf[s].prob_abs == db.surf[s].prob_abs
        assert da.surf[s].film.gap_energy == db.surf[s].film.gap_energy


def test_lattice_tables_bit_exact(lib_cmpmc, oracle):
    lat = lib_cmpmc.Lattice()
    assert np.array_equal(lat.oct, oracle.Engine().oct_table())
    g = np.load(GOLD)
    for m in range(3):
        d, mag = lat.map_k_to_vg(m, g["map_k"])
        assert np.array_equal(d, g["map_vg"][m][:, :3]) and np.array_equal(mag, g["map_vg"][m][:, 3])
    # decay tables through the reference's own draws (Rng(99), sample_*_energy_fraction)
    r = lib_cmpmc.Rng(99)
    lt, tt = [], []
    for _ in range(1000):
        for cdf, acc in ((lat.cdf_lt, lt), (lat.cdf_tt, tt)):
            u = r.uniform() * 1023
            i = min(int(u), 1022)
            f = u - i
            acc.append(cdf[i] * (1.0 - f) + cdf[i + 1] * f)
    assert np.array_equal(lt, g["decay_lt"]) and np.array_equal(tt, g["decay_tt"])
    info = g["lattice_info"]
    c = lat.c
    assert [c.e_cond_mass_eV, c.h_mass_eV, c.l0_e, c.l0_h, c.k_sound_e, c.k_sound_h] == list(info[:6])


def test_host_rng_streams(lib_cmpmc):
    g = np.load(GOLD)
    for i, s in enumerate(g["rng_seeds"]):
        for j, t in enumerate(g["rng_streams"]):
            r = lib_cmpmc.Rng.for_stream(int(s), int(t))
            assert [r.next_u64() for _ in range(16)] == [int(x) for x in g["rng_values"][i, j]]


def test_argument_errors(lib_cmpmc):
    cm = lib_cmpmc
    L = cm.capi.load()
    lat = cm.Lattice()
    k = np.zeros(3)
    assert L.cmpmc_map_k_to_vg(C.byref(lat.c), 7, 1, k.ctypes.data, k.ctypes.data, k.ctypes.data) == cm.capi.ERR_INVALID
    assert "bad argument" in L.cmpmc_last_error().decode()
    assert L.cmpmc_lattice_build(None, None, None, None, None) == cm.capi.ERR_INVALID
    bad = cm.germanium()
    bad.n_valleys = 99
    o = np.zeros((3, 256, 256, 4), dtype=np.float32)
    a, b = np.zeros(1024), np.zeros(1024)
    out = cm.capi.Lattice()
    assert L.cmpmc_lattice_build(C.byref(bad), o.ctypes.data, a.ctypes.data, b.ctypes.data, C.byref(out)) == 1


def test_model_without_device_reports_it(lib_cmpmc):
    cm = lib_cmpmc
    n = C.c_int(0)
    try:
        hip = C.CDLL("libamdhip64.so")
        hip.hipGetDeviceCount(C.byref(n))
    except OSError:
        pass
    if n.value > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(cm.CmpmcError) as e:
        cm.Engine()
    assert e.value.code == cm.capi.ERR_NO_DEVICE


def test_host_exact_kinematics_bit_exact(lib_cmpmc):
    """cmpmc_map_k_to_vg with KinematicsMode::Exact (crystal_lattice.hpp:77-85) on the
    host equals the reference build's values (golden, bitwise), degenerate axes included."""
    import os
    cm = lib_cmpmc
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz"))
    lat = cm.Lattice()
    lat.set_kinematics_mode(lat.KINEMATICS_EXACT)
    k = gold["map_exact_k"]
    for m in range(3):
        d, mag = lat.map_k_to_vg(m, k)
        assert np.array_equal(np.c_[d, mag], gold["map_exact_vg"][m])


def test_host_bilinear_kinematics_bit_exact(lib_cmpmc, oracle):
    """cmpmc_lattice_build_bilinear + cmpmc_map_k_to_vg with KinematicsMode::Bilinear
    (crystal_lattice.hpp:89-114, 350-368) on the host: the table equals the oracle's, and the
    lookups equal the reference build's values (golden, bitwise), poles and axes included."""
    import os
    cm = lib_cmpmc
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz"))
    lat = cm.Lattice()
    with pytest.raises(cm.CmpmcError):        # no table yet: mode 2 is refused
        lat.c.kinematics_mode = 2
        lat.map_k_to_vg(0, np.array([[0.0, 0.0, 1.0]]))
    lat.c.kinematics_mode = 0
    lat.set_kinematics_mode(lat.KINEMATICS_BILINEAR)
    k = gold["map_exact_k"]
    for m in range(3):
        d, mag = lat.map_k_to_vg(m, k)
        assert np.array_equal(np.c_[d, mag], gold["map_bil_vg"][m])
    assert np.array_equal(lat.bil, oracle.Engine().bil_table())


def test_host_symmetric_kinematics_equals_oracle(lib_cmpmc, oracle):
    """cmpmc_lattice_build_symmetric + cmpmc_map_k_to_vg in kinematics mode 3 (this build's
    cubic-symmetry table; no reference counterpart): the table and the lookups equal the
    oracle's bitwise, on random directions, the axes, the diagonals and a tie."""
    cm = lib_cmpmc
    lat = cm.Lattice()
    with pytest.raises(cm.CmpmcError):        # no table yet: mode 3 is refused
        lat.c.kinematics_mode = 3
        lat.map_k_to_vg(0, np.array([[0.0, 0.0, 1.0]]))
    lat.c.kinematics_mode = 0
    lat.set_kinematics_mode(lat.KINEMATICS_SYMMETRIC)
    eng = oracle.Engine()
    assert np.array_equal(lat.sym, eng.sym_table())
    g = np.random.default_rng(41)
    k = np.vstack([g.normal(size=(2000, 3)), [[0, 0, 1], [0, -1, 0], [1, 0, 0], [1, 1, 1], [-1, 1, -1], [1, 1, 0],
                                              [0.3, -0.3, 0.9]]])
    for m in range(3):
        d, mag = lat.map_k_to_vg(m, k)
        assert np.array_equal(np.c_[d, mag], eng.map_k_to_vg(m, k, symmetric=True))


def _gold():
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz"))


def test_host_film_cascade_equals_reference(lib_cmpmc):
    """KaplanFilm::absorb_phonon on the host (cmpmc_film_absorb_quantum) equals the reference
    build's cascades bit for bit: 200 cascades from Rng(seed) at 7.5, 15 and 31 meV (each
    edep, reflected count and sum, and the final stream state); the rate-competition loss
    probability and pair-breaking rate equal the reference's on the golden energies; a
    photon seed conserves energy exactly (absorbed + re-emitted = quantum, no thermal loss
    at T = 0)."""
    cm = lib_cmpmc
    gold = _gold()
    f = cm.aluminum_film()
    for tag, E, seed in (("k75", 7.5e-3, 11), ("k15", 15e-3, 12), ("k31", 31e-3, 13)):
        r = cm.Rng(seed)
        for i in range(200):
            edep, refl = cm.film_absorb(f, E, r)
            sm = 0.0
            for x in refl:
                sm += x
            assert edep == gold[f"{tag}_edep"][i] and len(refl) == gold[f"{tag}_nrefl"][i]
            assert sm == gold[f"{tag}_reflsum"][i]
        assert r.s == [int(x) for x in gold[f"{tag}_state"]]
    frc = cm.aluminum_film()
    frc.loss_model, frc.phonon_loss_time = 1, 4.0 * 600e-9 / (0.5 * 3260.0)
    for E, g_ref, p_ref in zip(gold["rc_E"], gold["rc_rate"], gold["rc_loss"]):
        p, g = cm.film_loss_probability(frc, float(E), 1.0)
        assert p == p_ref and g == g_ref
    f0 = cm.aluminum_film()
    f0.temperature = 0.0
    r = cm.Rng(3)
    for E in (0.5e-3, 2.0e-3, 13.9e-3):
        for _ in range(50):
            edep, refl = cm.film_absorb(f0, E, r, photon=True)
            assert abs(edep + refl.sum() - E) <= 1e-12 * E


def test_host_surrogate_equals_reference(lib_cmpmc):
    """KaplanSurrogateLUT (cmpmc_surrogate_*): the table built from Rng(1337) and
    absorb_phonon_fast equal the reference's on 3,000 calls from one stream (energies over
    [1.5, 45] Delta, so the exact cascade outside the table as well); each stored outcome
    conserves energy to FP32 rounding."""
    cm = lib_cmpmc
    gold = _gold()
    s = cm.KaplanSurrogate()
    r = cm.Rng(21)
    for i, E in enumerate(gold["surr_E"]):
        edep, refl = s.absorb(float(E), r)
        sm = 0.0
        for x in refl:
            sm += x
        assert edep == gold["surr_edep"][i] and len(refl) == gold["surr_nrefl"][i] and sm == gold["surr_reflsum"][i], i
    assert r.s == [int(x) for x in gold["surr_state"]]
    a, o, fr = s.tables()
    assert len(a) == 128 * 512 and o[0] == 0 and o[-1] == len(fr)
    tot = a.astype(np.float64) + np.add.reduceat(np.r_[fr, 0.0].astype(np.float64), o[:-1]) * (np.diff(o) > 0)
    # fractions of the incident energy: absorbed + re-emitted <= 1 (thermalised energy leaves)
    assert np.all(tot <= 1.0 + 1e-6) and np.mean(tot) > 0.5
    s.close()


@pytest.mark.parametrize("tag", ["amm_geal", "amm_siw"])
def test_host_amm_equals_reference(lib_cmpmc, tag):
    """The library's AMM model (the code the plain kernel runs, compiled for the host)
    equals the reference's P-SV coefficients, SH transmission and film-entry probability
    bit for bit over 300 angles (golden vectors written by the reference)."""
    cm = lib_cmpmc
    g = np.load(GOLD)
    m1, m2 = (tuple(row) for row in g[f"{tag}_media"])
    for i, th in enumerate(g[f"{tag}_theta"]):
        assert np.array_equal(cm.amm_psv(m1, m2, True, th), g[f"{tag}_psv"][i, :4], equal_nan=True)
        assert np.array_equal(cm.amm_psv(m1, m2, False, th), g[f"{tag}_psv"][i, 4:], equal_nan=True)
        L = cm.capi.load()
        a, b = cm.capi.Medium(*m1), cm.capi.Medium(*m2)
        got = [L.cmpmc_amm_sh(C.byref(a), C.byref(b), th), cm.amm_transmission(m1, m2, True, np.cos(th)),
               cm.amm_transmission(m1, m2, False, np.cos(th))]
        assert np.array_equal(got, [g[f"{tag}_{k}"][i] for k in ("sh", "tl", "tt")], equal_nan=True)


def test_host_masks_equal_reference(lib_cmpmc, oracle):
    """cmpmc_qet_mask equals ElectrodePatternMask::supercdms_qet_mask's pads (through the
    oracle's restatement, pinned to the reference), and cmpmc_find_channel with fill 0.8
    returns the reference's channels at 2000 points from the same stream; the grid preset
    centres nx x ny pads on the face with x-major channel ids."""
    cm = lib_cmpmc
    g = np.load(GOLD)
    pads = cm.qet_mask(0.0381, 0.8)
    for got, want in zip(pads, oracle.qet_pads(0.0381, 0.8)):
        assert (got.shape, got.channel_id) == want[:2]
        assert [got.cx, got.cy, got.r_min, got.r_max, got.phi_min, got.phi_max, got.fill_fraction] == \
            pytest.approx(list(want[2:]), rel=1e-15, abs=0.0)
    rng = cm.Rng(31)
    got = [cm.find_channel(pads, float(x), float(y), rng) for x, y in g["qet_xy"]]
    assert np.array_equal(np.array(got), g["qet_channels"])
    grid = cm.grid_mask(3, 2, 0.01, 0.02, 0.004, 0.005)
    assert [p.channel_id for p in grid] == list(range(6))
    assert [(p.cx, p.cy) for p in grid] == pytest.approx([(-0.01, -0.01), (-0.01, 0.01), (0.0, -0.01), (0.0, 0.01),
                                                           (0.01, -0.01), (0.01, 0.01)])
    assert cm.find_channel(grid, 0.0031, 0.0149) == 3 and cm.find_channel(grid, 0.0041, 0.01) == -1


def test_host_partition_equals_reference(lib_cmpmc, gold):
    """The library's energy partition on the host (cmpmc_ionization_yield,
    cmpmc_sample_fano_binomial, cmpmc_partition_deposit: the shared core the device
    generator runs) equals the reference's yields, Fano draws and partitioned tracks
    bit for bit (golden vectors written by the reference)."""
    cm = lib_cmpmc
    E = gold["yield_E"]
    si = cm.germanium()
    si.atomic_Z, si.atomic_A = 14.0, 28.085
    for ci, cr in enumerate((cm.germanium(), si)):
        for m in range(5):
            got = np.array([cm.ionization_yield(m, e, cr) for e in E])
            assert np.array_equal(got, gold["yield_values"][ci, m]), f"crystal {ci} model {m}"
    fixed = cm.partition_opts(empirical_energy_dependent_k=0).empirical
    assert np.array_equal(np.array([cm.ionization_yield(4, e, None, fixed) for e in E]), gold["yield_empirical_fixed_ge"])
    assert np.array_equal(np.array([cm.lindhard_yield(e, None, 14.0, 28.085) for e in E]), gold["yield_lindhard_si_in_ge"])
    for i, (mean, fano) in enumerate(gold["fano_cases"]):
        r = cm.Rng(100 + i)
        got = [cm.sample_fano_binomial(mean, fano, r) for _ in range(200)]
        assert np.array_equal(np.array(got, dtype=np.uint64), gold["fano_values"][i]), (mean, fano)
    lat = cm.Lattice()
    from test_oracle_partition import PARTITION_CASES, RESULT_KEYS
    for i, (e, nr, ch, seed, pos) in enumerate(PARTITION_CASES):
        o = cm.partition_opts(**{k.replace("empirical.", "empirical_"): v for k, v in ch.items()})
        r = cm.Rng(seed)
        out = cm.partition_deposit(lat, pos, e, bool(nr), r, o)
        res = out["result"]
        assert np.array_equal(np.array([float(res[k]) for k in RESULT_KEYS]), gold[f"part{i}_result"]), i
        q = out["charges"]
        got_q = np.c_[q["type"], q["pos"], q["p_eV"], q["ekin"], q["weight"], q["valley"], q["time"]] \
            if len(q) else np.zeros((0, 11))
        assert np.array_equal(got_q, gold[f"part{i}_charges"]), i
        p = out["phonons"]
        got_p = np.c_[p["mode"], p["pos"], p["k_dir"], p["vg_dir"], p["vg_mag"], p["energy"], p["weight"], p["time"]] \
            if len(p) else np.zeros((0, 14))
        assert np.array_equal(got_p, gold[f"part{i}_phonons"]), i
        assert np.array_equal(np.array(r.s, dtype=np.uint64), gold[f"part{i}_rng"]), i


def test_host_partition_capacity_error_keeps_stream(lib_cmpmc):
    cm = lib_cmpmc
    L = cm.capi.load()
    lat = cm.Lattice()
    o = cm.partition_opts()
    st = (C.c_uint64 * 4)(1, 2, 3, 4)
    nq, nph = C.c_uint64(), C.c_uint64()
    pos = (C.c_double * 3)(0.0, 0.0, 0.0)
    rc = L.cmpmc_partition_deposit(C.byref(lat.c), pos, 0.0, 1000.0, 0, C.byref(o), st, None, 0, C.byref(nq), None, 0,
                                   C.byref(nph), None)
    assert rc == cm.capi.ERR_CAPACITY and nq.value > 0 and nph.value > 0
    assert list(st) == [1, 2, 3, 4]


MESH_KEYS = ("found", "bary", "hints", "E", "pot")


def test_host_mesh_fields_equal_reference(lib_cmpmc, oracle, gold, tmp_path):
    """cmpmc_mesh*: construction (grids, explicit meshes with clamped potentials and
    shuffled cells, the two-file loader) and the walk / field / potential queries equal the
    reference's TriLinearMesh3D and BiLinearMesh2D bit for bit, walk hints included."""
    cm, O = lib_cmpmc, oracle
    (hx, hy, hz), (nx, ny, nz), c3 = O.MESH3_GRID
    cases = [("mesh3_grid", cm.Mesh.grid3d(hx, hy, hz, nx, ny, nz, lambda x, y, z: O.poly3(c3, x, y, z)),
              "mesh_pts3", 0)]
    xyz, V, tets = O.irregular_mesh3()
    cases.append(("mesh3_irr", cm.Mesh.tet(xyz, V, tets, -30.0, 30.0), "mesh_pts3i", -1))
    cp, tp = O.write_mesh_files(str(tmp_path))
    cases.append(("mesh3_file", cm.Mesh.load3d(cp, tp, 1e-3, 1e-3), "mesh_pts3i", 7))
    (u0, u1), (v0, v1), (nu, nv), c2 = O.MESH2_GRID_RZ
    cases.append(("mesh2_rz", cm.Mesh.grid2d(u0, u1, v0, v1, nu, nv, lambda u, v: O.poly2(c2, u, v), True),
                  "mesh_pts3", 0))
    (u0, u1), (v0, v1), (nu, nv), c2 = O.MESH2_GRID_XY
    cases.append(("mesh2_xy", cm.Mesh.grid2d(u0, u1, v0, v1, nu, nv, lambda u, v: O.poly2(c2, u, v), False),
                  "mesh_pts3", 5))
    uv, V2, tris = O.irregular_mesh2()
    cases.append(("mesh2_irr", cm.Mesh.tri(uv, V2, tris, False), "mesh_pts3i", 2))
    for tag, mesh, pts, h0 in cases:
        o = mesh.query(gold[pts], h0)
        for k in MESH_KEYS:
            assert np.array_equal(o[k], gold[f"{tag}_{k}"]), (tag, k)
    assert (cases[0][1].num_vertices, cases[0][1].num_cells) == tuple(gold["mesh3_grid_counts"])
    assert (cases[2][1].num_vertices, cases[2][1].num_cells) == tuple(gold["mesh3_file_counts"])
    assert (cases[3][1].num_vertices, cases[3][1].num_cells) == tuple(gold["mesh2_rz_counts"])


def test_host_mesh_argument_errors(lib_cmpmc, tmp_path):
    """A cell naming a vertex that does not exist, a non-finite coordinate and a missing
    file are errors (the reference reads out of bounds / builds an empty mesh)."""
    cm = lib_cmpmc
    xyz = np.eye(4, 3)
    with pytest.raises(cm.CmpmcError, match="names vertex"):
        cm.Mesh.tet(xyz, np.zeros(4), [[0, 1, 2, 4]])
    bad = xyz.copy()
    bad[2, 1] = np.nan
    with pytest.raises(cm.CmpmcError, match="not finite"):
        cm.Mesh.tet(bad, np.zeros(4), [[0, 1, 2, 3]])
    with pytest.raises(cm.CmpmcError, match="cannot open"):
        cm.Mesh.load3d(tmp_path / "none.txt", tmp_path / "none2.txt")
    empty = cm.Mesh.tet(np.zeros((0, 3)), np.zeros(0), np.zeros((0, 4)))
    E, _ = empty.field_at([0.0, 0.0, 0.0])
    assert empty.num_cells == 0 and np.all(E == 0.0)
