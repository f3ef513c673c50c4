"""GPU parity of the reference's optional physics modes, against the CPU oracle on the same
seeds, in both forms:

* KinematicsMode::Bilinear (crystal_lattice.hpp:89-114): the (cos theta, phi) table built
  on the host and read by the device in FP64 (plain) and FP32 (batched);
* FilmLossModel::KaplanRateCompetition (kaplan_film.hpp:91-99, 183-191): film phonons lost
  with probability tau_pb / (tau_pb + tau_l) from the exact Kaplan pair-breaking rate;
* IVRateModel::Matrix (charge_kernels.hpp:36-79; engine.hpp:197-199): the guarded
  deformation-potential + neutral-impurity intervalley rate, refused without
  allow_unvalidated and run with it;
* kinematics mode 3, this build's cubic-symmetry v_g table (no reference counterpart; its
  oracle is pinned by the node, invariance and accuracy checks of tests/test_oracle.py),
  which the batched kernel stages in LDS.

The oracle is pinned bit for bit to the reference on each (tests/test_oracle.py: golden
lookups and rates, runs c1_bil, film_rc, iv_matrix). Bars as for the other branch tests:
plain first-step branch codes equal to the oracle's, batched first-step branch counts equal
to the oracle's batched mode, ledger closure 1e-9, distributions (KS p > 0.01, means within
3 sigma), and each mode visibly changing the result it should change.
"""
import numpy as np
import pytest
from scipy import stats as sst

pytestmark = pytest.mark.gpu


def _closure(s):
    inj = s["total_injected_eV"]
    return abs(inj - s["total_deposited_eV"] - s["escaped_or_killed_eV"]) / inj


def _ks_ok(a, b, name, pmin=0.01):
    p = sst.ks_2samp(a, b).pvalue
    assert p > pmin, f"{name}: KS p = {p:.3g}"


def _z(a, b):
    sa = a.std(ddof=1) / np.sqrt(len(a))
    sb = b.std(ddof=1) / np.sqrt(len(b))
    return (a.mean() - b.mean()) / np.hypot(sa, sb)


def _phonon_parity(cm, O, eng, oeng, n, seed, tag, min_same=0.5):
    """Both forms against the oracle on n 7.5 meV L phonons from the origin."""
    wl = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    owl = O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    out = {}
    for form, batched in ((cm.FORM_PLAIN, False), (cm.FORM_BATCHED, True)):
        name = f"{tag} {'batched' if batched else 'plain'}"
        print(f"case: {name}, n={n} seed={seed}", flush=True)
        g = eng.run_workload(wl, n, seed=seed, form=form, first_branch=True)
        o = oeng.run(owl, n, seed, batched=batched, first_branch=True, threads=8)
        if batched:
            assert np.array_equal(np.bincount(g["first_branch"], minlength=16),
                                  np.bincount(o["first_branch"], minlength=16)), f"{name}: first-step branch counts"
        else:
            assert np.array_equal(g["first_branch"], o["first_branch"]), f"{name}: first-step branch codes"
            same = np.mean(np.diff(g["hit_offsets"].astype(np.int64)) == np.diff(o["hit_offsets"].astype(np.int64)))
            print(f"{name}: per-event hit counts identical for {same:.4f} of events", flush=True)
            assert same >= min_same
        s = g["stats"]
        assert _closure(s) <= 1e-9, f"{name}: ledger closure {_closure(s):.2e}"
        assert s["total_steps"] == pytest.approx(o["stats"]["total_steps"], rel=0.05)
        for f in ("time", "edep"):
            _ks_ok(g["hits"][f], o["hits"][f], f"{name} hit {f}")
            assert abs(_z(g["hits"][f], o["hits"][f])) < 3.0, f"{name} hit {f} mean"
        out[form] = g
    return out


def test_bilinear_device_queries(lib_cmpmc, oracle):
    """The device's bilinear lookups: FP64 (plain form) equal the host's (the reference's,
    bitwise) up to atan2's last bits in ocml; FP32 (batched form) within FP32 rounding of
    the interpolation."""
    cm = lib_cmpmc
    lat = cm.Lattice()
    lat.set_kinematics_mode(lat.KINEMATICS_BILINEAR)
    eng = cm.Engine(lat)
    g = np.random.default_rng(17)
    k = np.vstack([g.normal(size=(4000, 3)), [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [1, 1, 1]]])
    k /= np.linalg.norm(k, axis=1)[:, None]
    print(f"case: bilinear device queries, n={len(k)}", flush=True)
    for m in range(3):
        hd, hm = lat.map_k_to_vg(m, k)
        d64, m64 = eng.device_map_k_to_vg(m, k, form=cm.FORM_PLAIN)
        assert np.allclose(m64, hm, rtol=1e-12, atol=0) and np.allclose(d64, hd, rtol=0, atol=1e-12)
        d32, m32 = eng.device_map_k_to_vg(m, k, form=cm.FORM_BATCHED)
        assert np.allclose(m32, hm, rtol=2e-5, atol=0)
        # the angle from the cross product (arccos of a dot product near 1 would turn the
        # FP32 rounding of the vectors' norms, ~1e-7, into ~4e-4 rad)
        ang = np.arctan2(np.linalg.norm(np.cross(d32, hd), axis=1), np.sum(d32 * hd, axis=1))
        assert np.quantile(ang, 0.99) < 2e-5 and ang.max() < 1e-3, f"mode {m}: FP32 direction error {ang.max():.2e}"


def test_bilinear_transport_both_forms(lib_cmpmc, oracle):
    cm, O = lib_cmpmc, oracle
    lat = cm.Lattice()
    lat.set_kinematics_mode(lat.KINEMATICS_BILINEAR)
    eng = cm.Engine(lat)
    oeng = O.Engine()
    oeng.set_kinematics(2)
    r = _phonon_parity(cm, O, eng, oeng, 2000, 42, "bilinear kinematics")
    # the table's lookup error (polar step ~8 degrees near the z axis) moves the transport:
    # the engine really runs the bilinear lookups (an engine on the octahedral table with the
    # same seed takes other paths)
    t = cm.Engine().run_workload(cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 2000, seed=42,
                                 form=cm.FORM_PLAIN)
    assert r[cm.FORM_PLAIN]["stats_raw"] != t["stats_raw"]
    eng.set_kinematics(0)
    t2 = eng.run_workload(cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 2000, seed=42, form=cm.FORM_PLAIN)
    assert t2["stats_raw"] == t["stats_raw"], "switching the engine back to the table gives the table's run"


def test_symmetric_device_queries(lib_cmpmc, oracle):
    """Kinematics mode 3 (this build's cubic-symmetry table; no reference counterpart) on the
    device: FP64 lookups (plain form) equal the host's (the oracle's, bitwise) to 1e-12,
    FP32 lookups (batched form) within FP32 rounding of the interpolation."""
    cm = lib_cmpmc
    lat = cm.Lattice()
    lat.set_kinematics_mode(lat.KINEMATICS_SYMMETRIC)
    eng = cm.Engine(lat)
    g = np.random.default_rng(23)
    k = np.vstack([g.normal(size=(4000, 3)), [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [1, 1, 1],
                                              [1, -1, 0]]])
    k /= np.linalg.norm(k, axis=1)[:, None]
    print(f"case: symmetric device queries, n={len(k)}", flush=True)
    for m in range(3):
        hd, hm = lat.map_k_to_vg(m, k)
        d64, m64 = eng.device_map_k_to_vg(m, k, form=cm.FORM_PLAIN)
        assert np.allclose(m64, hm, rtol=1e-12, atol=0) and np.allclose(d64, hd, rtol=0, atol=1e-12)
        d32, m32 = eng.device_map_k_to_vg(m, k, form=cm.FORM_BATCHED)
        assert np.allclose(m32, hm, rtol=2e-5, atol=0)
        ang = np.arctan2(np.linalg.norm(np.cross(d32, hd), axis=1), np.sum(d32 * hd, axis=1))
        assert np.quantile(ang, 0.99) < 2e-5 and ang.max() < 1e-3, f"mode {m}: FP32 direction error {ang.max():.2e}"


def test_symmetric_transport_both_forms(lib_cmpmc, oracle):
    """Kinematics mode 3 in both forms against the oracle in mode 3 (the batched form stages
    the table in LDS); the engine really runs the table (the octahedral table's run on the
    same seed differs) and switching back gives the octahedral run."""
    cm, O = lib_cmpmc, oracle
    lat = cm.Lattice()
    lat.set_kinematics_mode(lat.KINEMATICS_SYMMETRIC)
    eng = cm.Engine(lat)
    oeng = O.Engine()
    oeng.set_kinematics(3)
    r = _phonon_parity(cm, O, eng, oeng, 2000, 42, "symmetric kinematics")
    wl = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    t = cm.Engine().run_workload(wl, 2000, seed=42, form=cm.FORM_PLAIN)
    assert r[cm.FORM_PLAIN]["stats_raw"] != t["stats_raw"]
    eng.set_kinematics(0)
    t2 = eng.run_workload(wl, 2000, seed=42, form=cm.FORM_PLAIN)
    assert t2["stats_raw"] == t["stats_raw"], "switching the engine back to the table gives the table's run"
    # a model created without a symmetry table builds its own: switching a default engine
    # to mode 3 gives the same batched run as the engine built on the mode-3 lattice
    e3 = cm.Engine()
    e3.set_kinematics(3)
    eng.set_kinematics(3)
    a = e3.run_workload(wl, 2000, seed=43, form=cm.FORM_BATCHED)
    b = eng.run_workload(wl, 2000, seed=43, form=cm.FORM_BATCHED)
    assert a["stats_raw"] == b["stats_raw"]


def test_rate_competition_both_forms(lib_cmpmc, oracle):
    """Both Al films on KaplanRateCompetition (tau_l = 4 d / (0.5 v_s) = 1.47 ns): both forms
    against the oracle, and the loss model visibly changes the collected energy per hit from
    the MeanFreePath default (by ~4 %, > 5 sigma at this size) in the oracle's direction."""
    cm, O = lib_cmpmc, oracle
    d = cm.default_detector()
    for i in (0, 1):
        d.surf[i].film.loss_model = 1
        d.surf[i].film.phonon_loss_time = O.RC_LOSS_TIME
    eng = cm.Engine(detector=d)
    oeng = O.Engine(detector=O.rate_competition_detector())
    n = 3000
    r = _phonon_parity(cm, O, eng, oeng, n, 42, "rate competition", min_same=0.8)
    wl = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    base = cm.Engine().run_workload(wl, n, seed=42, form=cm.FORM_BATCHED)
    zb = _z(r[cm.FORM_BATCHED]["hits"]["edep"], base["hits"]["edep"])
    print(f"rate competition vs mean free path, batched: edep per hit z = {zb:.1f}", flush=True)
    assert zb > 5.0
    bad = cm.default_detector()
    bad.surf[0].film.loss_model = 3
    with pytest.raises(cm.CmpmcError):
        cm.Engine(detector=bad)


def test_matrix_iv_guard_and_both_forms(lib_cmpmc, oracle):
    """IVRateModel::Matrix: refused without allow_unvalidated (UnvalidatedModuleError's code),
    run with it; pairs at 4 V against the oracle per carrier (intervalley scatters, Luke
    emissions, charge steps) and in collection-time distribution, in both forms. The Matrix
    rate scatters electrons between valleys ~1,000x more often than the default Quadratic
    model at this field."""
    cm, O = lib_cmpmc, oracle
    cr = cm.germanium()
    cr.iv_model = 2
    d = cm.default_detector()
    cm.set_bias(d, 4.0)
    d.luke_sample_fraction = 0.0
    wl = cm.workload(cm.WL_PAIR)
    guarded = cm.Engine(cm.Lattice(cr), d)
    with pytest.raises(cm.CmpmcError) as ei:
        guarded.run_workload(wl, 10, seed=1)
    assert ei.value.code == cm.capi.ERR_UNVALIDATED
    d.allow_unvalidated = 1
    eng = cm.Engine(cm.Lattice(cr), d)
    od = O.default_detector()
    O.set_bias(od, 4.0)
    od.luke_sample_fraction = 0.0
    oeng = O.Engine(crystal=O.matrix_iv_crystal(), detector=od)
    owl = O.workload(O.WL_PAIR)
    n, seed = 600, 42
    for form, batched in ((cm.FORM_PLAIN, False), (cm.FORM_BATCHED, True)):
        name = f"matrix IV {'batched' if batched else 'plain'}"
        print(f"case: {name}, n={n} seed={seed}", flush=True)
        g = eng.run_workload(wl, n, seed=seed, form=form)
        o = oeng.run(owl, n, seed, batched=batched, threads=8)
        gs, os_ = g["stats"], o["stats"]
        print(f"{name}: iv/carrier gpu {gs['intervalley_scatters'] / (2 * n):.2f} oracle "
              f"{os_['intervalley_scatters'] / (2 * n):.2f}", flush=True)
        for k in ("intervalley_scatters", "luke_phonons_emitted", "charge_steps"):
            assert gs[k] == pytest.approx(os_[k], rel=0.05), f"{name} {k}"
        assert gs["charges_collected"] == os_["charges_collected"] == 2 * n
        _ks_ok(g["hits"]["time"], o["hits"]["time"], f"{name} collection time")
        assert gs["intervalley_scatters"] > 100 * (2 * n)
    q = cm.Engine(cm.Lattice(), d).run_workload(wl, n, seed=seed, form=cm.FORM_BATCHED)
    assert q["stats"]["intervalley_scatters"] < 5 * (2 * n)


@pytest.mark.parametrize("mode", [1, 2], ids=["bulk_debye", "electrode_film"])
def test_charge_release_both_forms(lib_cmpmc, oracle, mode):
    """DetectorModel::ChargeRelease at carrier collection (engine.hpp:299-312, 888-927):
    BulkDebye releases E_rel as h nu_D phonons at the contact and records nothing in the
    carrier's hit; ElectrodeFilm sends E_rel into the face's Kaplan film as photon-seeded
    quanta (the hit keeps what the film absorbs, the escaping phonons are released). Pairs at
    4 V with Luke phonons counted, both forms against the oracle (pinned to the reference by
    the golden runs release_debye and release_film): every carrier collected, the carriers'
    recorded energy, the released phonons' transport (steps, film hits, deposited energy)
    and hit distributions."""
    cm, O = lib_cmpmc, oracle
    d = cm.default_detector()
    cm.set_bias(d, 4.0)
    d.luke_sample_fraction = 0.0
    d.charge_release = mode
    eng = cm.Engine(detector=d)
    od = O.default_detector()
    O.set_bias(od, 4.0)
    od.luke_sample_fraction = 0.0
    od.charge_release = mode
    oeng = O.Engine(detector=od)
    wl, owl = cm.workload(cm.WL_PAIR), O.workload(O.WL_PAIR)
    n, seed = 400, 42
    for form, batched in ((cm.FORM_PLAIN, False), (cm.FORM_BATCHED, True)):
        name = f"release {mode} {'batched' if batched else 'plain'}"
        print(f"case: {name}, n={n} seed={seed}", flush=True)
        g = eng.run_workload(wl, n, seed=seed, form=form)
        o = oeng.run(owl, n, seed, batched=batched, threads=8)
        gs, os_ = g["stats"], o["stats"]
        print(f"{name}: deposited {gs['total_deposited_eV']:.4f} vs {os_['total_deposited_eV']:.4f} eV, hits "
              f"{gs['sensor_hits']} vs {os_['sensor_hits']}, steps {gs['total_steps']} vs {os_['total_steps']}", flush=True)
        assert gs["charges_collected"] == os_["charges_collected"] == 2 * n
        assert gs["total_steps"] > 0, f"{name}: the released phonons are transported"
        for k in ("total_deposited_eV", "sensor_hits", "total_steps"):
            assert gs[k] == pytest.approx(os_[k], rel=0.05), f"{name} {k}"
        if mode == 1:
            assert gs["charge_recomb_edep_eV"] == 0.0 and os_["charge_recomb_edep_eV"] == 0.0
        else:
            assert gs["charge_recomb_edep_eV"] == pytest.approx(os_["charge_recomb_edep_eV"], rel=0.03)
        carrier = g["hits"]["excitation"] != 0
        assert carrier.sum() == 2 * n
        ph_g = g["hits"][~carrier]
        ph_o = o["hits"][o["hits"]["excitation"] == 0]
        for f in ("time", "edep"):
            _ks_ok(ph_g[f], ph_o[f], f"{name} phonon hit {f}")
    bad = cm.default_detector()
    bad.charge_release = 5
    with pytest.raises(cm.CmpmcError):
        cm.Engine(detector=bad)
