"""GPU parity of the transport branches the default detector never takes, against the CPU
oracle on the same seeds, in both forms:

* the reference benchmark's own phonon workload (benchmarks/cmpmc_bench.cpp:46-66): the
  demonstration polynomial surfaces (energy-dependent specular / diffuse / surface
  down-conversion split, surface_property.hpp:120-137), 100 bounces, the initial mode drawn
  by the benchmark's generator;
* Russian roulette every 32 bounces past 32 with survival 0.5 (engine.hpp:423-428,
  variance_reduction.hpp:17-29);
* half-specular surfaces (engine.hpp:543-579);
* charge drift at 4 V with trapping (C3 shape) and with trap impact ionisation
  (charge_kernels.hpp:101-126), batched form against the oracle's batched mode.

The oracle itself is pinned bit for bit to the reference on each of these detectors
(tests/test_oracle.py::test_run_events_golden: bench_dos, roulette, specular, trap_ion).

Bars: plain form (FP64, reference streams) - the first-step branch code of every event
equals the oracle's; batched form (FP32, per-track sub-streams) - first-step branch counts
equal the oracle's batched mode; both - ledger closure 1e-9, branch totals within Poisson
errors, hit time / energy / x / y distributions (two-sample KS p > 0.01) and means (3
sigma), top-face fraction (3 sigma).
"""
import numpy as np
import pytest
from scipy import stats as sst

pytestmark = pytest.mark.gpu

N_EV = 2000


def _closure(s):
    inj = s["total_injected_eV"]
    return abs(inj - s["total_deposited_eV"] - s["escaped_or_killed_eV"]) / inj


def _ks_ok(a, b, name, pmin=0.01):
    p = sst.ks_2samp(a, b).pvalue
    assert p > pmin, f"{name}: KS p = {p:.3g}"


def _mean_ok(a, b, name, nsig=3.0):
    sa = a.std(ddof=1) / np.sqrt(len(a))
    sb = b.std(ddof=1) / np.sqrt(len(b))
    z = abs(a.mean() - b.mean()) / np.hypot(sa, sb)
    assert z < nsig, f"{name}: means differ by {z:.2f} sigma"


def _count_ok(a, b, name, nsig=4.0, rel_floor=0.02):
    """Totals of one branch in two runs of the same events: within nsig Poisson errors, or
    rel_floor of the total (many counts in a run are correlated through shared tracks)."""
    tol = max(nsig * np.sqrt(a + b + 1.0), rel_floor * max(a, b))
    assert abs(a - b) <= tol, f"{name}: {a} vs {b} (tolerance {tol:.0f})"


def _hits_match(h, q, what):
    for f in ("time", "edep"):
        _ks_ok(h[f], q[f], f"{what} hit {f}")
        _mean_ok(h[f], q[f], f"{what} hit {f}")
    # batched positions are FP32: compare at FP32 resolution
    for i, ax in enumerate("xy"):
        _ks_ok(h["pos"][:, i].astype(np.float32), q["pos"][:, i].astype(np.float32), f"{what} hit {ax}")
    top_g, top_o = np.mean(h["pos"][:, 2] > 0), np.mean(q["pos"][:, 2] > 0)
    assert abs(top_g - top_o) < 3 * np.sqrt(0.25 / len(h) + 0.25 / len(q)), f"{what}: top/bottom split"


def _detectors(cm, O, which):
    d, od = cm.default_detector(), O.default_detector()
    if which == "bench":
        cm.demo_polynomial_surfaces(d)
        O.demo_polynomial_surfaces(od)
        mode = cm.MODE_BENCH_DOS
    elif which == "roulette":
        d.russian_roulette_bounces, d.russian_roulette_survival = 32, 0.5
        od.russian_roulette_bounces, od.russian_roulette_survival = 32, 0.5
        mode = 0
    elif which == "specular":
        cm.constant_split_surfaces(d, 0.5)
        O.constant_split_surfaces(od, 0.5)
        mode = 0
    else:
        raise ValueError(which)
    return d, od, mode


BRANCH_KEYS = ("total_steps", "isotope_scatters", "decays_L_to_TT", "decays_L_to_LT", "surface_refls",
               "surface_downconversions", "sensor_hits", "killed_tracks")


@pytest.mark.parametrize("which", ["bench", "roulette", "specular"])
def test_phonon_branches_both_forms(lib_cmpmc, oracle, which):
    cm, O = lib_cmpmc, oracle
    d, od, mode = _detectors(cm, O, which)
    eng = cm.Engine(detector=d)
    oeng = O.Engine(detector=od)
    wl = cm.workload(cm.WL_PHONON_POINT, mode=mode, energy_eV=7.5e-3)
    owl = O.workload(O.WL_PHONON_POINT, mode=mode, energy_eV=7.5e-3)
    seed = 12345
    print(f"case: {which}, plain and batched, n={N_EV} seed={seed}", flush=True)

    gp = eng.run_workload(wl, N_EV, seed=seed, form=cm.FORM_PLAIN, first_branch=True)
    op = oeng.run(owl, N_EV, seed, first_branch=True)
    assert np.array_equal(gp["first_branch"], op["first_branch"]), f"{which}: plain first-step branch codes differ"
    same = np.diff(gp["hit_offsets"].astype(np.int64)) == np.diff(op["hit_offsets"].astype(np.int64))
    print(f"{which}: plain per-event hit counts identical for {same.mean():.4f} of events", flush=True)
    # libm ULPs (ocml vs glibc) let a few percent of events diverge after their first steps
    assert same.mean() >= 0.8
    for k in BRANCH_KEYS:
        _count_ok(gp["stats"][k], op["stats"][k], f"{which} plain {k}")

    gb = eng.run_workload(wl, N_EV, seed=seed, form=cm.FORM_BATCHED, first_branch=True)
    ob = oeng.run(owl, N_EV, seed, batched=True, first_branch=True)
    bg = np.bincount(gb["first_branch"], minlength=16)
    bo = np.bincount(ob["first_branch"], minlength=16)
    assert np.array_equal(bg, bo), f"{which}: batched first-step branch counts gpu {bg} oracle {bo}"
    for k in BRANCH_KEYS:
        _count_ok(gb["stats"][k], ob["stats"][k], f"{which} batched {k}")

    for r, o, form in ((gp, op, "plain"), (gb, ob, "batched")):
        s = r["stats"]
        if which == "roulette":
            # a roulette kill books nothing (engine.hpp:423-428): the ledger closes only in
            # expectation, through the survivors' weights, in the reference as well (its
            # golden run of 200 events is 2.4 % short); compare with the oracle instead
            assert s["total_deposited_eV"] == pytest.approx(o["stats"]["total_deposited_eV"], rel=0.03)
            assert _closure(s) < 0.1
        else:
            assert _closure(s) <= 1e-9, f"{which} {form}: ledger closure {_closure(s):.2e}"
        assert s["total_injected_eV"] == pytest.approx(N_EV * 7.5e-3, rel=1e-12)
        _hits_match(r["hits"], o["hits"], f"{which} {form}")
        assert len(r["hits"]) == s["sensor_hits"] == int(r["hit_offsets"][-1])
    if which == "bench":
        assert gb["stats"]["surface_downconversions"] > 0 and gp["stats"]["surface_downconversions"] > 0
        # the benchmark generator draws ST / FT / L with weights 0.535 / 0.367 / 0.098
        assert gp["stats"]["decays_L_to_TT"] + gp["stats"]["decays_L_to_LT"] > 0
    if which == "roulette":
        base = cm.Engine().run_workload(wl, N_EV, seed=seed, form=cm.FORM_BATCHED, keep_hits=False)
        assert gb["stats"]["total_steps"] < base["stats"]["total_steps"], "roulette must shorten tracks"
        assert gb["stats"]["killed_tracks"] > base["stats"]["killed_tracks"]
        # survivors carry weight 1 / 0.5^k: some hit energies are multiples of their incident energy
        assert np.any(gb["hits"]["edep"] > gb["hits"]["incident_energy"] * 1.5)


# Trap impact-ionisation lengths of the four channels (e- on D0, e- on A-, h+ on D+,
# h+ on A0), all different so that a channel mapped to the wrong projectile or liberated
# carrier changes the mix of collected electrons and holes. Each carrier liberates about
# 0.2 more (a sub-critical cascade: ~530 ionisations per 1,000 pairs).
ION_MFP = (0.3, 0.5, 0.4, 0.8)


def _charge_engines(cm, O, trap_ion):
    cr, ocr = cm.germanium(), O.germanium()
    for c in (cr, ocr):
        c.e_trapping_mfp_m = c.h_trapping_mfp_m = 0.5
        if trap_ion:
            c.e_trap_ion_d0_mfp_m, c.e_trap_ion_a_mfp_m, c.h_trap_ion_d_mfp_m, c.h_trap_ion_a0_mfp_m = ION_MFP
    d, od = cm.default_detector(), O.default_detector()
    cm.set_bias(d, 4.0)
    O.set_bias(od, 4.0)
    d.luke_sample_fraction = od.luke_sample_fraction = 0.0
    return cm.Engine(cm.Lattice(cr), d), O.Engine(crystal=ocr, detector=od)


@pytest.mark.parametrize("trap_ion", [False, True], ids=["trapping", "trap_ionisation"])
def test_charge_drift_trapping_both_forms(lib_cmpmc, oracle, trap_ion):
    """C3 shape (e-/h+ pairs at 4 V, uniform positions, trapping 0.5 m), optionally with
    trap impact ionisation (four channels of different lengths): both forms against the
    oracle (plain mode and batched mode). Knock-on carriers make the totals a branching
    process (a sub-critical cascade has a large spread of total progeny), so quantities
    are compared per carrier: Luke emissions, trapping and ionisation per carrier, and
    the electron share of the collected carriers."""
    cm, O = lib_cmpmc, oracle
    eng, oeng = _charge_engines(cm, O, trap_ion)
    wl, owl = cm.workload(cm.WL_PAIR), O.workload(O.WL_PAIR)
    n, seed = 1000, 42
    print(f"case: charge drift trap_ion={trap_ion}, n={n} seed={seed}", flush=True)
    for form, batched in ((cm.FORM_PLAIN, False), (cm.FORM_BATCHED, True)):
        g = eng.run_workload(wl, n, seed=seed, form=form)
        o = oeng.run(owl, n, seed, batched=batched)
        gs, os_ = g["stats"], o["stats"]
        tag = f"{'trap-ion' if trap_ion else 'trapping'} {'batched' if batched else 'plain'}"
        carriers = []
        for s in (gs, os_):
            # every carrier ends collected or trapped
            assert s["charges_collected"] + s["charges_trapped"] == 2 * n + s["trap_ionizations"]
            carriers.append(2 * n + s["trap_ionizations"])
        print(f"{tag}: gpu {[gs[k] for k in ('trap_ionizations', 'charges_trapped', 'luke_phonons_emitted')]}, "
              f"oracle {[os_[k] for k in ('trap_ionizations', 'charges_trapped', 'luke_phonons_emitted')]}",
              flush=True)
        ng, no = carriers
        assert gs["luke_phonons_emitted"] / ng == pytest.approx(os_["luke_phonons_emitted"] / no, rel=0.05), \
            f"{tag} Luke emissions per carrier"
        for k in ("charges_trapped", "trap_ionizations"):
            pg, po = gs[k] / ng, os_[k] / no
            sig = np.sqrt(max(pg, 1e-3) * (1 - pg) / ng + max(po, 1e-3) * (1 - po) / no)
            assert abs(pg - po) < 4 * sig, f"{tag} {k} per carrier: {pg:.4f} vs {po:.4f}"
        if trap_ion:
            assert gs["trap_ionizations"] > 300
        else:
            assert gs["trap_ionizations"] == 0 and gs["charges_trapped"] > 20
        h, q = g["hits"], o["hits"]
        assert len(h) == gs["charges_collected"]
        eg, eo = np.mean(h["excitation"] == 1), np.mean(q["excitation"] == 1)
        assert abs(eg - eo) < 4 * np.sqrt(0.25 / len(h) + 0.25 / len(q)), f"{tag} electron share {eg:.3f} vs {eo:.3f}"
        _ks_ok(h["time"], q["time"], f"{tag} collection time")
        _mean_ok(h["time"], q["time"], f"{tag} collection time")
        assert abs(gs["total_deposited_eV"] - gs["charge_recomb_edep_eV"]) <= 1e-9 * gs["total_deposited_eV"]


def test_charge_stack_overflow_reruns(lib_cmpmc, oracle):
    """A knock-on carrier stack of one entry per lane of the batched charge kernel
    overflows under trap ionisation; the library reruns with the depth the kernel reports
    and the result equals a run with the default depth."""
    cm, O = lib_cmpmc, oracle
    eng, _ = _charge_engines(cm, O, True)
    wl = cm.workload(cm.WL_PAIR)
    print("case: charge stack overflow rerun, n=300", flush=True)
    a = eng.run_workload(wl, 300, seed=7, form=cm.FORM_BATCHED, headroom=1)
    b = eng.run_workload(wl, 300, seed=7, form=cm.FORM_BATCHED)
    assert a["passes"] >= 2 and b["passes"] == 1
    assert a["stats_raw"] == b["stats_raw"]
    assert np.array_equal(a["hits"], b["hits"])
    assert a["stats"]["trap_ionizations"] > 0


def test_time_budget_stops_a_long_run(lib_cmpmc):
    """A plain C1 run of 200,000 events (about 0.7 s of kernel time) with a 50 ms budget
    stops with CMPMC_ERR_CANCELLED and a message naming the budget; the model then runs
    normally, without a budget and with an ample one."""
    cm = lib_cmpmc
    eng = cm.Engine()
    wl = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    print("case: plain C1 n=200000 with a 0.05 s budget", flush=True)
    with pytest.raises(cm.CmpmcError) as ei:
        eng.run_workload(wl, 200_000, seed=3, form=cm.FORM_PLAIN, keep_hits=False, time_budget_s=0.05)
    assert ei.value.code == 6 and "time budget" in str(ei.value)
    for form in (cm.FORM_PLAIN, cm.FORM_BATCHED):
        with pytest.raises(cm.CmpmcError) as ei:
            eng.run_workload(wl, 200_000, seed=3, form=form, keep_hits=False, time_budget_s=1e-4)
        assert ei.value.code == 6
    a = eng.run_workload(wl, 500, seed=3, form=cm.FORM_PLAIN)
    b = eng.run_workload(wl, 500, seed=3, form=cm.FORM_PLAIN, time_budget_s=60.0)
    assert a["stats_raw"] == b["stats_raw"] and a["stats"]["total_steps"] > 0


def test_plain_form_refuses_c2_size_events(lib_cmpmc):
    """One lane runs a whole event in the plain form: events of the lognormal C2 workload
    (up to 20,000 phonons at up to 31 meV) are refused unless a time budget bounds the run;
    the batched form takes them."""
    cm = lib_cmpmc
    eng = cm.Engine()
    wl = cm.workload(cm.WL_PHONON_LOGN)
    with pytest.raises(cm.CmpmcError) as ei:
        eng.run_workload(wl, 2000, seed=1, form=cm.FORM_PLAIN, keep_hits=False)
    assert ei.value.code == cm.capi.ERR_INVALID and "batched form" in str(ei.value)
    small = cm.workload(cm.WL_PHONON_LOGN, logn_median=4.0, n_max=200)
    r = eng.run_workload(small, 50, seed=1, form=cm.FORM_PLAIN, keep_hits=False)
    assert r["stats"]["total_steps"] > 0
