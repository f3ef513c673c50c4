"""Graph-captured batched runs (cmpmc_plan_*) and the launch-option autotuner.

The batched form draws every track from its own sub-stream (event, track key), so its
results do not depend on how the tracks are scheduled: a plan replay must give exactly
what cmpmc_run_workload gives for the same seed and events (statistics bit for bit,
hits and offsets equal), whatever the recorded launch options, and so must options the
autotuner picks. Each graph is replayed several times with eager runs and
synchronisations in between, and every replay is checked.
"""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PLANS_OFF = ("graph-captured replays are switched off: in the first GPU run of this file the third replay of a "
             "plan did not complete (silent for 180 s) after two replays that matched eager runs bit for bit; "
             "no cause has been found, so cmpmc_plan_create returns an error")


def _same(a, b, what):
    assert a["stats_raw"] == b["stats_raw"], f"{what}: statistics differ"
    assert a["hits"].tobytes() == b["hits"].tobytes(), f"{what}: hits differ"
    assert np.array_equal(a["hit_offsets"], b["hit_offsets"]), f"{what}: hit offsets differ"


def _workloads(cm):
    return {
        "point 7.5 meV L": (cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 3000),
        "lognormal": (cm.workload(cm.WL_PHONON_LOGN, logn_median=4.0, logn_sigma=1.0, n_max=200), 600),
        "uniform count": (cm.workload(cm.WL_PHONON_UNIF, mode=0, energy_eV=7.5e-3, n_min=1, n_max=16), 800),
    }


@pytest.mark.skip(reason=PLANS_OFF)
@pytest.mark.parametrize("name", ["point 7.5 meV L", "lognormal", "uniform count"])
def test_plan_replays_equal_eager_runs(lib_cmpmc, name):
    cm = lib_cmpmc
    eng = cm.Engine()
    wl, n = _workloads(cm)[name]
    plan = eng.plan(wl, n)
    for seed in (3, 11, 3):
        print(f"case: plan {name}, n={n}, seed={seed}", flush=True)
        r = plan.run(seed)
        ref = eng.run_workload(wl, n, seed=seed, form=cm.FORM_BATCHED)
        _same(r, ref, f"{name} seed {seed}")
        s = r["stats"]
        assert abs(s["total_injected_eV"] - s["total_deposited_eV"] - s["escaped_or_killed_eV"]) \
            <= 1e-9 * s["total_injected_eV"]
    assert plan.recordings == 1
    plan.close()


@pytest.mark.skip(reason=PLANS_OFF)
def test_plan_charges_with_trapping_and_luke(lib_cmpmc):
    cm = lib_cmpmc
    cr = cm.germanium()
    cr.e_trapping_mfp_m = cr.h_trapping_mfp_m = 0.5
    det = cm.default_detector()
    cm.set_bias(det, 4.0)
    det.luke_sample_fraction = 0.0
    eng = cm.Engine(cm.Lattice(cr), det)
    wl = cm.workload(cm.WL_PAIR)
    plan = eng.plan(wl, 2000, hit_capacity=8000)
    for seed in (5, 6):
        print(f"case: plan charge drift with trapping, n=2000, seed={seed}", flush=True)
        _same(plan.run(seed), eng.run_workload(wl, 2000, seed=seed, form=cm.FORM_BATCHED, hit_capacity=8000),
              f"drift seed {seed}")
    plan.close()
    det2 = cm.default_detector()
    cm.set_bias(det2, 4.0)
    det2.luke_sample_fraction = 1.0
    eng2 = cm.Engine(cm.Lattice(), det2)
    plan2 = eng2.plan(wl, 20, hit_capacity=200000)
    for seed in (7, 8):
        print(f"case: plan full Luke cascade, n=20, seed={seed}", flush=True)
        r = plan2.run(seed)
        _same(r, eng2.run_workload(wl, 20, seed=seed, form=cm.FORM_BATCHED, hit_capacity=200000), f"Luke seed {seed}")
        assert r["stats"]["luke_phonons_emitted"] > 1000
    plan2.close()


@pytest.mark.skip(reason=PLANS_OFF)
def test_plan_first_event_id(lib_cmpmc):
    """A replay with first event id f runs events f..f+n-1 of the seed: their hits are
    those of the same events in one larger eager run."""
    cm = lib_cmpmc
    eng = cm.Engine()
    wl = cm.workload(cm.WL_PHONON_UNIF, mode=0, energy_eV=7.5e-3, n_min=1, n_max=8)
    n = 500
    print(f"case: plan first event id {n}, n={n}", flush=True)
    plan = eng.plan(wl, n)
    r = plan.run(9, first_event=n)
    big = eng.run_workload(wl, 2 * n, seed=9, form=cm.FORM_BATCHED)
    off = big["hit_offsets"].astype(np.int64)
    assert big["hits"][off[n]:off[2 * n]].tobytes() == r["hits"].tobytes()
    assert np.array_equal(np.diff(off[n:]), np.diff(r["hit_offsets"].astype(np.int64)))
    plan.close()


def _unif_total(cm, seed, n, n_min, n_max):
    return sum(n_min + int(cm.Rng.for_stream(seed, i).uniform() * (n_max - n_min + 1)) for i in range(n))


@pytest.mark.skip(reason=PLANS_OFF)
def test_plan_grows_and_records_again(lib_cmpmc):
    """Buffers that a replay outgrows are enlarged and the graph recorded again: a spawn
    pool of 2 entries per wave, and an initial queue sized from seed 1 that a later seed
    exceeds. Results equal eager runs with ample room."""
    cm = lib_cmpmc
    eng = cm.Engine()
    wl = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    print("case: plan with a 2-entry spawn pool, n=300", flush=True)
    plan = eng.plan(wl, 300, lane_depth=2)
    r = plan.run(5)
    assert plan.recordings >= 2 and r["passes"] >= 2
    _same(r, eng.run_workload(wl, 300, seed=5, form=cm.FORM_BATCHED), "small pool")
    plan.close()

    n, lo, hi = 2, 1, 3000
    wlu = cm.workload(cm.WL_PHONON_UNIF, mode=0, energy_eV=2.0e-3, n_min=lo, n_max=hi)
    t1 = _unif_total(cm, 1, n, lo, hi)
    cap = t1 + t1 // 16 + 1024
    seed = next(s for s in range(2, 500) if _unif_total(cm, s, n, lo, hi) > cap)
    print(f"case: plan queue growth, n={n}, seed 1 total {t1}, seed {seed} total "
          f"{_unif_total(cm, seed, n, lo, hi)} > {cap}", flush=True)
    plan = eng.plan(wlu, n)
    assert plan.recordings == 1
    r = plan.run(seed)
    assert plan.recordings == 2 and r["passes"] == 2
    _same(r, eng.run_workload(wlu, n, seed=seed, form=cm.FORM_BATCHED), "queue growth")
    plan.close()


TUNE_OFF = ("in two GPU runs of the whole suite this test went silent for 180 s right after its first tuning "
            "(20,000 events); an earlier run passed; no cause has been found in the code, so the test stays off "
            "the GPU until one is")


@pytest.mark.skip(reason=TUNE_OFF)
def test_autotune_cache_and_result_invariance(lib_cmpmc, tmp_path):
    cm = lib_cmpmc
    eng = cm.Engine()
    wl = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    cache = str(tmp_path / "tune.jsonl")
    print("case: autotune point 7.5 meV, n=20000", flush=True)
    t = eng.autotune(wl, 20000, seed=1, cache=cache)
    print(f"tuned: {t}", flush=True)
    assert t["block"] in (64, 128, 256) and t["waves"] in (2, 3, 4)
    assert t["refill_period"] in (1, 2, 4, 8) and t["lane_depth"] in (1024, 2048, 4096)
    assert not t["from_cache"] and t["trials"] >= 8
    assert 0 < t["best_ms"] <= t["default_ms"]
    lines = open(cache).read().splitlines()
    assert len(lines) == 1
    entry = json.loads(lines[0])
    assert entry["block_size"] == t["block"] and entry["waves_per_simd"] == t["waves"]
    again = eng.autotune(wl, 20000, seed=1, cache=cache)
    assert again["from_cache"] and all(again[k] == t[k] for k in ("block", "waves", "refill_period", "lane_depth"))
    # another workload gets its own line
    eng.autotune(cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 2000, seed=1, cache=cache)
    assert len(open(cache).read().splitlines()) == 2
    # the options change speed, never results
    opts = {k: t[k] for k in ("block", "waves", "refill_period", "lane_depth")}
    a = eng.run_workload(wl, 3000, seed=4, form=cm.FORM_BATCHED, **opts)
    b = eng.run_workload(wl, 3000, seed=4, form=cm.FORM_BATCHED)
    _same(a, b, "tuned vs default options")


def test_plan_create_reports_switched_off(lib_cmpmc):
    cm = lib_cmpmc
    eng = cm.Engine()
    with pytest.raises(cm.CmpmcError) as ei:
        eng.plan(cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 100)
    assert ei.value.code == 1 and "switched off" in str(ei.value)


def test_autotune_reports_switched_off(lib_cmpmc):
    cm = lib_cmpmc
    eng = cm.Engine()
    with pytest.raises(cm.CmpmcError) as ei:
        eng.autotune(cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 100)
    assert ei.value.code == 1 and "switched off" in str(ei.value)
