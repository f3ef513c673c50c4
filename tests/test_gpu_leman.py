"""Leman second-order charge stepping (DetectorModel::ChargeStepping::LemanSecondOrder,
engine.hpp:234-254, 337-342) on the GPU, both forms, against the CPU oracle (whose Leman
path is bit-exact against the reference: tests/test_oracle.py golden runs `leman` and
`leman_ion`; and meets the reference's own Leman physics checks:
tests/test_oracle_leman.py).

* plain form (FP64, the reference's streams and operation order) against the oracle and
  batched form (FP32) against the oracle's batched mode: Luke emissions and energy,
  collected carriers, charge steps, collection-time distributions;
* energy closure (tests/test_leman_stepping.cpp:40-42 in the reference, here over a
  whole run of pairs starting at the origin): the field's work on every carrier equals the
  Luke phonon energy it emitted plus its kinetic-energy change, to 1e-3;
* fewer steps than first-order stepping on the same pairs.
"""
import numpy as np
import pytest
from scipy import stats as sst

pytestmark = pytest.mark.gpu

GAP = 0.74
E_FIELD = 4.0 / 2.54e-2      # V/m at 4 V across 2.54 cm


def _engines(cm, O, leman, luke_fraction=0.0):
    d, od = cm.default_detector(), O.default_detector()
    cm.set_bias(d, 4.0)
    O.set_bias(od, 4.0)
    d.luke_sample_fraction = od.luke_sample_fraction = luke_fraction
    if leman:
        d.charge_stepping = od.charge_stepping = 1
    return cm.Engine(detector=d), O.Engine(detector=od)


def _closure(stats, hits):
    """|sum E_Luke + sum (E_kin final - E_kin start) - sum work| / sum work over a run of
    pairs that start at rest (1 ueV) at the origin and are all collected."""
    sg = np.where(hits["excitation"] == 1, -1.0, 1.0)
    work = sg * (-E_FIELD) * hits["pos"][:, 2]
    ekf = hits["incident_energy"] - 0.5 * GAP
    return abs(stats["total_luke_phonon_eV"] + np.sum(ekf - 1e-6) - work.sum()) / work.sum()


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_leman_against_oracle(lib_cmpmc, oracle, form):
    cm, O = lib_cmpmc, oracle
    eng, oeng = _engines(cm, O, True)
    wl, owl = cm.workload(cm.WL_PAIR), O.workload(O.WL_PAIR)
    n, seed = 300, 42
    batched = form == "batched"
    print(f"case: Leman stepping, {form}, n={n} seed={seed}", flush=True)
    g = eng.run_workload(wl, n, seed=seed, form=cm.FORM_BATCHED if batched else cm.FORM_PLAIN)
    o = oeng.run(owl, n, seed, batched=batched)
    gs, os_ = g["stats"], o["stats"]
    print(f"Leman {form}: gpu steps {gs['charge_steps']} Luke {gs['luke_phonons_emitted']}; oracle steps "
          f"{os_['charge_steps']} Luke {os_['luke_phonons_emitted']}", flush=True)
    assert gs["charges_collected"] == os_["charges_collected"] == 2 * n
    for k in ("luke_phonons_emitted", "total_luke_phonon_eV"):
        assert gs[k] == pytest.approx(os_[k], rel=0.03), k
    assert gs["charge_steps"] == pytest.approx(os_["charge_steps"], rel=0.05)
    p = sst.ks_2samp(g["hits"]["time"], o["hits"]["time"]).pvalue
    assert p > 0.01, f"collection time KS p = {p:.3g}"


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_leman_energy_closure_and_fewer_steps(lib_cmpmc, oracle, form):
    cm, O = lib_cmpmc, oracle
    wl = cm.workload(cm.WL_PAIR, pair_at_origin=1)
    n = 200
    f = cm.FORM_BATCHED if form == "batched" else cm.FORM_PLAIN
    print(f"case: Leman closure, {form}, n={n}", flush=True)
    lem = _engines(cm, O, True)[0].run_workload(wl, n, seed=9, form=f)
    first = _engines(cm, O, False)[0].run_workload(wl, n, seed=9, form=f)
    for r, what in ((lem, "Leman"), (first, "first order")):
        assert r["stats"]["charges_collected"] == 2 * n
        c = _closure(r["stats"], r["hits"])
        print(f"{form} {what}: closure {c:.2e}, steps {r['stats']['charge_steps']}", flush=True)
        assert c <= 1e-3, f"{what} energy closure {c:.2e}"
    assert lem["stats"]["charge_steps"] < first["stats"]["charge_steps"] / 1.5
    assert lem["stats"]["luke_phonons_emitted"] == pytest.approx(first["stats"]["luke_phonons_emitted"], rel=0.05)


def test_leman_luke_cascade_batched(lib_cmpmc, oracle):
    """Leman stepping with every Luke phonon enqueued and transported (C4 shape): the
    deposited energy per pair and phonon statistics against the oracle's batched mode."""
    cm, O = lib_cmpmc, oracle
    eng, oeng = _engines(cm, O, True, luke_fraction=1.0)
    wl, owl = cm.workload(cm.WL_PAIR), O.workload(O.WL_PAIR)
    n = 40
    print(f"case: Leman Luke cascade, batched, n={n}", flush=True)
    g = eng.run_workload(wl, n, seed=5, form=cm.FORM_BATCHED)
    o = oeng.run(owl, n, 5, batched=True)
    gs, os_ = g["stats"], o["stats"]
    for k in ("luke_phonons_emitted", "total_luke_phonon_eV", "total_deposited_eV", "sensor_hits"):
        assert gs[k] == pytest.approx(os_[k], rel=0.05), k
    assert gs["total_steps"] == pytest.approx(os_["total_steps"], rel=0.1)
    inj = gs["total_luke_phonon_eV"] + gs["charge_recomb_edep_eV"]
    assert abs(gs["total_deposited_eV"] + gs["escaped_or_killed_eV"] - inj) <= 1e-6 * inj
