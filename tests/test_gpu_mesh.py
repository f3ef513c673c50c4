"""FEM mesh fields on the GPU (geometry/mesh_field.hpp; DetectorModel::mesh_field_3d /
mesh_field_2d, engine.hpp:151-155), against the CPU oracle whose mesh code is bit-exact
against the reference (tests/test_oracle_mesh.py, golden runs mesh3_pairs, mesh2_pairs and
mesh3_leman_ion in tests/test_oracle.py):

* the device's field evaluation (the walk and the cell field the charge step uses), FP64
  form: bit for bit the reference's find_tetrahedron / find_triangle / field_at on the
  golden query points, each walk started where the reference's carried hint started it;
  FP32 form: the same cells except within its tolerance of a cell boundary, the same
  fields to FP32 rounding;
* a mesh whose field is exactly uniform (dyadic grid, V = -128 z) gives the uniform-field
  plain run bit for bit: the mesh enters the charge step where the uniform field does;
* pairs drifting in a 3D grid field and in an axisymmetric (r, z) field, both forms, and
  Leman stepping with trap ionisation in the 3D field: per-run totals and collection
  distributions against the oracle on the same seeds (charge runs decorrelate from the CPU
  through device libm ULPs over ~10^4 steps per carrier, as the uniform-field charge tests
  describe), and clearly apart from the uniform-field distributions.
"""
import numpy as np
import pytest
from scipy import stats as sst

pytestmark = pytest.mark.gpu

KEYS = ("found", "hints", "E")


def _mesh_pair(cm, O, kind):
    """(product mesh, oracle mesh) of a named case."""
    if kind == "grid3":
        (hx, hy, hz), (nx, ny, nz), c = O.MESH3_GRID
        return cm.Mesh.grid3d(hx, hy, hz, nx, ny, nz, lambda x, y, z: O.poly3(c, x, y, z)), O.mesh3_grid()
    if kind == "rz":
        (u0, u1), (v0, v1), (nu, nv), c = O.MESH2_GRID_RZ
        return cm.Mesh.grid2d(u0, u1, v0, v1, nu, nv, lambda u, v: O.poly2(c, u, v), True), O.mesh2_grid_rz()
    if kind == "irr3":
        xyz, V, tets = O.irregular_mesh3()
        return cm.Mesh.tet(xyz, V, tets, -30.0, 30.0), O.Mesh.tet(xyz, V, tets, -30.0, 30.0)
    if kind == "xy":
        (u0, u1), (v0, v1), (nu, nv), c = O.MESH2_GRID_XY
        return (cm.Mesh.grid2d(u0, u1, v0, v1, nu, nv, lambda u, v: O.poly2(c, u, v), False),
                O.Mesh.grid2d(u0, u1, v0, v1, nu, nv, c, False))
    uv, V, tris = O.irregular_mesh2()
    return cm.Mesh.tri(uv, V, tris, False), O.Mesh.tri(uv, V, tris, False)


GOLD_SETS = [("mesh3_grid", "grid3", "mesh_pts3", 0), ("mesh3_irr", "irr3", "mesh_pts3i", -1),
             ("mesh2_rz", "rz", "mesh_pts3", 0), ("mesh2_xy", "xy", "mesh_pts3", 5),
             ("mesh2_irr", "irr2", "mesh_pts3i", 2)]


@pytest.mark.parametrize("tag,kind,pts,h0", GOLD_SETS)
def test_device_field_fp64_equals_reference(lib_cmpmc, oracle, gold, tag, kind, pts, h0):
    cm, O = lib_cmpmc, oracle
    print(f"case: device field FP64 {tag}", flush=True)
    mesh, _ = _mesh_pair(cm, O, kind)
    d = cm.default_detector()
    cm.set_mesh(d, mesh)
    eng = cm.Engine(detector=d)
    # each walk starts where the reference's walk for that point started: the hint the
    # previous point left (the first: the sequence's initial hint)
    start = np.r_[h0, gold[f"{tag}_hints"][:-1]].astype(np.int32)
    r = eng.field_at(gold[pts], cm.FORM_PLAIN, start=start)
    # (the axisymmetric projection's r = hypot(x, y) is the host library's to the last bit)
    assert np.array_equal(r["inside"], gold[f"{tag}_found"])
    assert np.array_equal(r["cell"], gold[f"{tag}_hints"])
    assert np.array_equal(r["E"], gold[f"{tag}_E"])


@pytest.mark.parametrize("kind", ["grid3", "rz", "irr3"])
def test_device_field_fp32_and_seed_starts(lib_cmpmc, oracle, kind):
    """FP32 walks land in the FP64 walk's cell for every point farther than 1e-4 (in
    barycentric terms) from a cell boundary, with the same field to FP32 rounding; walks
    from the seed boxes (a new carrier's start) land where walks from far cells do."""
    cm, O = lib_cmpmc, oracle
    print(f"case: device field FP32 {kind}", flush=True)
    mesh, omesh = _mesh_pair(cm, O, kind)
    d = cm.default_detector()
    cm.set_mesh(d, mesh)
    eng = cm.Engine(detector=d)
    half = (0.0382, 0.0382, 0.0128) if kind != "irr3" else (0.02, 0.015, 0.01)
    p = O.mesh_query_points(4000, half, 21)
    ref = omesh.query(p, 0)
    a = eng.field_at(p, cm.FORM_PLAIN)                 # seed-box starts, FP64
    b = eng.field_at(p, cm.FORM_BATCHED)               # seed-box starts, FP32
    c = eng.field_at(p, cm.FORM_PLAIN, start=np.arange(len(p)) * 7919 % max(mesh.num_cells, 1))
    clear = ref["bary"].min(axis=1) > 1e-4
    out = ref["found"] == 0
    assert clear.sum() > 0.5 * len(p)
    for r in (a, c):
        assert np.array_equal(r["inside"][clear], ref["found"][clear])
        assert np.array_equal(r["E"][clear], ref["E"][clear])
    assert np.array_equal(b["inside"][clear], ref["found"][clear])
    ins = clear & (ref["found"] == 1)
    assert np.allclose(b["E"][ins], ref["E"][ins], rtol=2e-5, atol=1e-4 * np.abs(ref["E"]).max())
    # outside the mesh (beyond the tolerance) both forms give zero
    far = out & (np.abs(p) > 1.02 * np.array(half)).any(axis=1)
    assert far.sum() > 100
    assert np.all(a["E"][far] == 0.0) and np.all(b["E"][far] == 0.0)


def _dyadic(cm):
    d = cm.default_detector()
    d.luke_sample_fraction = 0.0
    m = cm.Mesh.grid3d(2.0 ** -4, 2.0 ** -4, 2.0 ** -6, 8, 8, 8, lambda x, y, z: -128.0 * z)
    cm.set_mesh(d, m)
    u = cm.default_detector()
    u.luke_sample_fraction = 0.0
    u.efield[:] = (0.0, 0.0, 128.0)
    return cm.Engine(detector=d), cm.Engine(detector=u)


def test_dyadic_mesh_equals_uniform_field_plain(lib_cmpmc):
    cm = lib_cmpmc
    em, eu = _dyadic(cm)
    wl = cm.workload(cm.WL_PAIR)
    print("case: dyadic mesh vs uniform, plain, n=64", flush=True)
    a = em.run_workload(wl, 64, seed=42, form=cm.FORM_PLAIN)
    b = eu.run_workload(wl, 64, seed=42, form=cm.FORM_PLAIN)
    assert a["stats"]["charges_collected"] == 128
    assert a["stats_raw"] == b["stats_raw"]
    assert np.array_equal(a["hits"], b["hits"])


def test_dyadic_mesh_uniform_field_batched(lib_cmpmc):
    """The FP32 form in the same exact field: the same collection statistics as in the
    uniform field (its step bound |E| c / hbar c is formed per step, not once, so runs are
    equal in distribution rather than bit for bit)."""
    cm = lib_cmpmc
    em, eu = _dyadic(cm)
    wl = cm.workload(cm.WL_PAIR)
    print("case: dyadic mesh vs uniform, batched, n=2000", flush=True)
    a = em.run_workload(wl, 2000, seed=42, form=cm.FORM_BATCHED)
    b = eu.run_workload(wl, 2000, seed=42, form=cm.FORM_BATCHED)
    assert a["stats"]["charges_collected"] == b["stats"]["charges_collected"] == 4000
    assert a["stats"]["charge_steps"] == pytest.approx(b["stats"]["charge_steps"], rel=0.02)
    assert a["stats"]["total_luke_phonon_eV"] == pytest.approx(b["stats"]["total_luke_phonon_eV"], rel=0.01)
    assert sst.ks_2samp(a["hits"]["time"], b["hits"]["time"]).pvalue > 0.01


def _field_engines(cm, O, kind, leman=False, trap_ion=False, luke=0.0):
    mesh, omesh = _mesh_pair(cm, O, kind)
    d, od = cm.default_detector(), O.default_detector()
    d.luke_sample_fraction = od.luke_sample_fraction = luke
    if leman:
        d.charge_stepping = od.charge_stepping = 1
    cm.set_mesh(d, mesh)
    O.set_mesh(od, omesh)
    cr, ocr = cm.germanium(), O.germanium()
    if trap_ion:
        ocr = O.trap_ion_crystal()
        for f in ("e_trapping_mfp_m", "h_trapping_mfp_m", "e_trap_ion_d0_mfp_m", "e_trap_ion_a_mfp_m",
                  "h_trap_ion_d_mfp_m", "h_trap_ion_a0_mfp_m"):
            setattr(cr, f, getattr(ocr, f))
    return cm.Engine(cm.Lattice(cr), d), O.Engine(crystal=ocr, detector=od)


def _compare(g, o, n, name, uniform=None):
    gs, os_ = g["stats"], o["stats"]
    print(f"{name}: gpu steps {gs['charge_steps']} collected {gs['charges_collected']} Luke eV "
          f"{gs['total_luke_phonon_eV']:.4g}; oracle steps {os_['charge_steps']} collected "
          f"{os_['charges_collected']} Luke eV {os_['total_luke_phonon_eV']:.4g}", flush=True)
    assert gs["charge_steps"] == pytest.approx(os_["charge_steps"], rel=0.05)
    assert gs["total_luke_phonon_eV"] == pytest.approx(os_["total_luke_phonon_eV"], rel=0.025)
    assert gs["charge_recomb_edep_eV"] == pytest.approx(os_["charge_recomb_edep_eV"], rel=0.04)
    for f in ("time",):
        p = sst.ks_2samp(g["hits"][f], o["hits"][f]).pvalue
        assert p > 0.005, f"{name}: collection {f} KS p = {p:.3g}"
    # collection radii: the side-wall share (binomial) and the flat-face radii (KS); the
    # FP32 form's side-wall hits sit within FP32 rounding of R rather than on it, so the
    # two are compared apart
    R = 0.0381
    r_g = np.hypot(g["hits"]["pos"][:, 0], g["hits"]["pos"][:, 1])
    r_o = np.hypot(o["hits"]["pos"][:, 0], o["hits"]["pos"][:, 1])
    side_g, side_o = r_g > R * (1 - 1e-6), r_o > R * (1 - 1e-6)
    k1, n1, k2, n2 = side_g.sum(), len(r_g), side_o.sum(), len(r_o)
    pp = (k1 + k2) / (n1 + n2)
    z = abs(k1 / n1 - k2 / n2) / max(np.sqrt(pp * (1 - pp) * (1 / n1 + 1 / n2)), 1e-12)
    assert z < 4.0, f"{name}: side-wall collection {k1}/{n1} vs {k2}/{n2}"
    p = sst.ks_2samp(r_g[~side_g], r_o[~side_o]).pvalue
    assert p > 0.005, f"{name}: flat-face collection radius KS p = {p:.3g}"
    if uniform is not None:
        # the mesh field is what moved the carriers: the field's work (the Luke energy, for
        # the same start positions) differs from the uniform 4 V bias's by more than the
        # tolerance above (by 7 % for the 3D grid field, 57 % for the axisymmetric one)
        lu = uniform["stats"]["total_luke_phonon_eV"]
        print(f"{name}: Luke eV in the uniform field {lu:.4g}", flush=True)
        assert abs(gs["total_luke_phonon_eV"] / lu - 1.0) > 0.05


@pytest.mark.parametrize("form", ["plain", "batched"])
@pytest.mark.parametrize("kind", ["grid3", "rz"])
def test_pairs_in_mesh_field_vs_oracle(lib_cmpmc, oracle, kind, form):
    cm, O = lib_cmpmc, oracle
    eng, oeng = _field_engines(cm, O, kind)
    n, seed = 400, 42
    batched = form == "batched"
    print(f"case: pairs in the {kind} mesh field, {form}, n={n} seed={seed}", flush=True)
    wl, owl = cm.workload(cm.WL_PAIR), O.workload(O.WL_PAIR)
    g = eng.run_workload(wl, n, seed=seed, form=cm.FORM_BATCHED if batched else cm.FORM_PLAIN)
    o = oeng.run(owl, n, seed, batched=batched)
    assert g["stats"]["charges_collected"] == o["stats"]["charges_collected"] == 2 * n
    du = O.default_detector()
    O.set_bias(du, 4.0)
    du.luke_sample_fraction = 0.0
    uni = O.Engine(detector=du).run(owl, n, seed, batched=batched)
    _compare(g, o, n, f"{kind} {form}", uniform=uni)


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_leman_trap_ionisation_in_mesh_field_vs_oracle(lib_cmpmc, oracle, form):
    cm, O = lib_cmpmc, oracle
    eng, oeng = _field_engines(cm, O, "grid3", leman=True, trap_ion=True)
    batched = form == "batched"
    n, seed = (1500 if batched else 300), 44
    print(f"case: Leman + trap ionisation in the 3D mesh field, {form}, n={n}", flush=True)
    wl, owl = cm.workload(cm.WL_PAIR), O.workload(O.WL_PAIR)
    g = eng.run_workload(wl, n, seed=seed, form=cm.FORM_BATCHED if batched else cm.FORM_PLAIN)
    o = oeng.run(owl, n, seed, batched=batched)
    gs, os_ = g["stats"], o["stats"]
    print(f"trap ionisations {gs['trap_ionizations']} / {os_['trap_ionizations']}, trapped {gs['charges_trapped']} / "
          f"{os_['charges_trapped']}", flush=True)
    for k in ("trap_ionizations", "charges_trapped"):
        a, b = gs[k], os_[k]
        assert abs(a - b) <= 4.0 * np.sqrt(a + b + 1.0), f"{k}: {a} vs {b}"
    _compare(g, o, n, f"Leman trap-ion {form}")
