"""Charge drift with Luke emission on the GPU, at the charge configurations:

* drift: e-/h+ pairs at 4 V across the 2.54 cm height, positions uniform in the volume,
  trapping on (electron and hole trapping lengths 0.5 m, which traps ~9 % of the
  carriers over the drift), Luke phonons counted but not tracked;
* Luke cascade: the same pairs without trapping, every Luke phonon (~3,600 per pair)
  enqueued and tracked to absorption.

The plain FP64 kernel follows the reference draw for draw, but ~20k steps per carrier
with cbrt/cos/sin on every Luke emission decorrelate from the CPU through libm ULPs. Per
run totals are therefore compared statistically with the oracle on the same seeds:
counts as binomial/Poisson within 4 sigma, continuous totals within stated relative
tolerances. The batched FP32 form is compared with the plain form at larger N.
"""
import numpy as np
import pytest
from scipy import stats as sst

pytestmark = pytest.mark.gpu

TRAP = 0.5


def _engines(cm, O, luke, trap):
    cr = cm.germanium()
    cr.e_trapping_mfp_m = cr.h_trapping_mfp_m = trap
    det = cm.default_detector()
    cm.set_bias(det, 4.0)
    det.luke_sample_fraction = luke
    ocr = O.germanium()
    ocr.e_trapping_mfp_m = ocr.h_trapping_mfp_m = trap
    odet = O.default_detector()
    O.set_bias(odet, 4.0)
    odet.luke_sample_fraction = luke
    return cm.Engine(cm.Lattice(cr), det), O.Engine(crystal=ocr, detector=odet)


def _binom_ok(k1, n1, k2, n2, name, nsig=4.0):
    p = (k1 + k2) / (n1 + n2)
    sd = np.sqrt(p * (1 - p) * (1 / n1 + 1 / n2))
    z = abs(k1 / n1 - k2 / n2) / max(sd, 1e-12)
    assert z < nsig, f"{name}: {k1}/{n1} vs {k2}/{n2} differ by {z:.2f} sigma"


def test_drift_with_trapping_plain_vs_oracle(lib_cmpmc, oracle):
    cm, O = lib_cmpmc, oracle
    eng, oeng = _engines(cm, O, 0.0, TRAP)
    n = 400
    print(f"case: charge drift with trapping, plain vs oracle, n={n}", flush=True)
    g = eng.run_workload(cm.workload(cm.WL_PAIR), n, seed=42, form=cm.FORM_PLAIN)
    o = oeng.run(O.workload(O.WL_PAIR), n, 42, threads=8)
    gs, os_ = g["stats"], o["stats"]
    assert gs["charges_collected"] + gs["charges_trapped"] == 2 * n
    assert os_["charges_collected"] + os_["charges_trapped"] == 2 * n
    _binom_ok(gs["charges_trapped"], 2 * n, os_["charges_trapped"], 2 * n, "trapped carriers")
    for k in ("charge_steps", "luke_phonons_emitted", "total_luke_phonon_eV"):
        assert gs[k] == pytest.approx(os_[k], rel=0.04), k
    assert gs["charge_recomb_edep_eV"] == pytest.approx(os_["charge_recomb_edep_eV"], rel=0.04)
    assert len(g["hits"]) == gs["charges_collected"]


def test_drift_with_trapping_batched_vs_plain(lib_cmpmc, oracle):
    cm, O = lib_cmpmc, oracle
    eng, _ = _engines(cm, O, 0.0, TRAP)
    n = 20000
    print(f"case: charge drift with trapping, batched vs plain, n={n}", flush=True)
    wl = cm.workload(cm.WL_PAIR)
    p = eng.run_workload(wl, n, seed=7, form=cm.FORM_PLAIN, hit_capacity=4 * n)
    b = eng.run_workload(wl, n, seed=7, form=cm.FORM_BATCHED, hit_capacity=4 * n)
    ps, bs = p["stats"], b["stats"]
    _binom_ok(bs["charges_trapped"], 2 * n, ps["charges_trapped"], 2 * n, "trapped carriers")
    for k in ("charge_steps", "luke_phonons_emitted", "total_luke_phonon_eV", "charge_recomb_edep_eV"):
        assert bs[k] == pytest.approx(ps[k], rel=0.01), k
    for f in ("time", "edep"):
        pv = sst.ks_2samp(p["hits"][f], b["hits"][f]).pvalue
        assert pv > 0.01, f"collected-carrier {f}: KS p = {pv:.3g}"
    # a collected carrier's hit carries its excitation type (1 electron, 2 hole)
    for r in (p, b):
        h = r["hits"]
        assert set(np.unique(h["excitation"])) <= {1, 2}


def test_luke_cascade_plain_vs_oracle(lib_cmpmc, oracle):
    """The plain kernel's Luke path (emitted phonons joining the event's stack): one lane
    runs an event's ~3,600 phonons in sequence, so the test tracks a 10 % sample of them
    (weight 10, the reference's luke_sample_fraction) to keep the lane time short."""
    cm, O = lib_cmpmc, oracle
    eng, oeng = _engines(cm, O, 0.1, 0.0)
    n = 40
    print(f"case: Luke phonons tracked (fraction 0.1), plain vs oracle, n={n}", flush=True)
    g = eng.run_workload(cm.workload(cm.WL_PAIR), n, seed=42, form=cm.FORM_PLAIN)
    o = oeng.run(O.workload(O.WL_PAIR), n, 42, threads=8)
    gs, os_ = g["stats"], o["stats"]
    assert gs["charges_collected"] == 2 * n and os_["charges_collected"] == 2 * n
    for k in ("luke_phonons_emitted", "total_luke_phonon_eV", "total_deposited_eV", "total_steps"):
        assert gs[k] == pytest.approx(os_[k], rel=0.05), k


def test_luke_cascade_batched_vs_oracle_batched(lib_cmpmc, oracle):
    cm, O = lib_cmpmc, oracle
    eng, oeng = _engines(cm, O, 1.0, 0.0)
    n = 200
    print(f"case: full Luke cascade, batched vs oracle batched, n={n}", flush=True)
    g = eng.run_workload(cm.workload(cm.WL_PAIR), n, seed=5, form=cm.FORM_BATCHED, hit_capacity=4000 * n)
    o = oeng.run(O.workload(O.WL_PAIR), n, 5, threads=16, batched=True)
    gs, os_ = g["stats"], o["stats"]
    assert gs["charges_collected"] == 2 * n
    for k in ("luke_phonons_emitted", "total_luke_phonon_eV", "total_deposited_eV", "total_steps", "sensor_hits"):
        assert gs[k] == pytest.approx(os_[k], rel=0.03), k
    # every emitted Luke phonon is tracked: phonon deposits plus losses account for the
    # Luke energy, and charge recombination adds the rest
    for s in (gs, os_):
        lhs = s["total_deposited_eV"] + s["escaped_or_killed_eV"]
        rhs = s["total_luke_phonon_eV"] + s["charge_recomb_edep_eV"]
        assert lhs == pytest.approx(rhs, rel=1e-9)
    ph_g = g["hits"][g["hits"]["excitation"] == 0]
    ph_o = o["hits"][o["hits"]["excitation"] == 0]
    for f in ("time", "edep"):
        pv = sst.ks_2samp(ph_g[f], ph_o[f]).pvalue
        assert pv > 0.01, f"Luke phonon hit {f}: KS p = {pv:.3g}"


def test_batched_charges_independent_of_carriers_per_wavefront(lib_cmpmc, oracle):
    """A small batch runs one carrier per wavefront on the full grid, and 38 per
    wavefront on a 4-block grid (16 wavefronts for 600 carriers); every carrier draws from
    its own stream, so statistics and hits must be identical."""
    cm, O = lib_cmpmc, oracle
    eng, _ = _engines(cm, O, 0.0, TRAP)
    n = 300
    print(f"case: charge drift, carriers per wavefront 1 vs 38, n={n}", flush=True)
    wl = cm.workload(cm.WL_PAIR)
    wide = eng.run_workload(wl, n, seed=11, form=cm.FORM_BATCHED, hit_capacity=4 * n)
    packed = eng.run_workload(wl, n, seed=11, form=cm.FORM_BATCHED, hit_capacity=4 * n, grid_blocks=4)
    assert wide["stats"]["charges_collected"] + wide["stats"]["charges_trapped"] == 2 * n
    assert wide["stats_raw"] == packed["stats_raw"], "charge statistics depend on the carriers per wavefront"
    assert np.array_equal(wide["hits"], packed["hits"])
