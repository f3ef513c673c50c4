"""cmpmc_run_workload_noalloc: the same runs as cmpmc_run_workload with every device
buffer carved from the library's fixed per-device scratch array. Bar: bit-identical
statistics and hits to cmpmc_run_workload on the same workload, options and seed, on two
calls in a row (the second smaller, so that it runs over what the first left in the
scratch), in both forms; no device memory taken by the call (hipMemGetInfo unchanged
across a call once the kernels are loaded); runs that do not fit the scratch refused with
CMPMC_ERR_CAPACITY; the ordinary entry point unaffected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    ok = a["stats_raw"] == b["stats_raw"]
    assert ok, f"{what}: statistics differ"
    ha, hb = a["hits"], b["hits"]
    assert len(ha) == len(hb), f"{what}: hit counts differ ({len(ha)} vs {len(hb)})"
    bad = [f for f in ha.dtype.names if ha[f].tobytes() != hb[f].tobytes()]
    assert not bad, f"{what}: hit fields {bad} differ"
    ok = np.array_equal(a["hit_offsets"], b["hit_offsets"])
    assert ok, f"{what}: hit offsets differ"


def _free_bytes():
    # the HIP runtime the library links (torch may carry another copy under libamdhip64.so)
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


CASES = [
    # (name, workload kwargs, form, extra options, first n, second n, hit capacity)
    ("phonons plain", dict(kind="point"), "plain", {}, 1500, 400, 20000),
    ("phonons batched", dict(kind="point"), "batched", dict(grid_blocks=8), 2000, 500, 20000),
    ("lognormal batched", dict(kind="logn"), "batched", dict(grid_blocks=8), 2, 1, 100000),
    ("pairs plain", dict(kind="pair"), "plain", {}, 300, 100, 4096),
    ("pairs batched", dict(kind="pair"), "batched", dict(grid_blocks=8), 300, 100, 4096),
]


def _workload(cm, kind):
    if kind == "point":
        return cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    if kind == "logn":
        return cm.workload(cm.WL_PHONON_LOGN, logn_median=200.0, n_max=600)
    return cm.workload(cm.WL_PAIR)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_run_workload(lib_cmpmc, case):
    cm = lib_cmpmc
    name, wkw, form, extra, n1, n2, cap = case
    d = cm.default_detector()
    if wkw["kind"] == "pair":
        cm.set_bias(d, 4.0)
        d.luke_sample_fraction = 0.0
    eng = cm.Engine(detector=d)
    wl = _workload(cm, wkw["kind"])
    f = cm.FORM_BATCHED if form == "batched" else cm.FORM_PLAIN
    for n, seed in ((n1, 11), (n2, 12)):
        print(f"case: no-alloc {name}, n={n} seed={seed}", flush=True)
        a = eng.run_workload(wl, n, seed=seed, form=f, hit_capacity=cap, no_alloc=True, **extra)
        b = eng.run_workload(wl, n, seed=seed, form=f, hit_capacity=cap, **extra)
        _same(a, b, f"{name} n={n}")
        assert a["stats"]["sensor_hits"] > 0 or a["stats"]["charges_collected"] > 0


def test_noalloc_takes_no_device_memory(lib_cmpmc):
    cm = lib_cmpmc
    eng = cm.Engine()
    wl = _workload(cm, "point")
    print("case: no-alloc memory check, plain 500 events", flush=True)
    eng.run_workload(wl, 500, seed=3, hit_capacity=40000, no_alloc=True)
    eng.run_workload(wl, 500, seed=3, form=cm.FORM_BATCHED, grid_blocks=8, hit_capacity=40000, no_alloc=True)
    eng.trim()
    before = _free_bytes()
    r1 = eng.run_workload(wl, 500, seed=4, hit_capacity=40000, no_alloc=True)
    r2 = eng.run_workload(wl, 500, seed=4, form=cm.FORM_BATCHED, grid_blocks=8, hit_capacity=40000, no_alloc=True)
    after = _free_bytes()
    assert after == before, f"device memory changed by {before - after} B across no-allocation calls"
    assert r1["stats"]["sensor_hits"] > 0 and r2["stats"]["sensor_hits"] > 0


def test_noalloc_refuses_runs_beyond_the_scratch(lib_cmpmc):
    cm = lib_cmpmc
    eng = cm.Engine()
    wl = _workload(cm, "point")
    print("case: no-alloc size limits", flush=True)
    with pytest.raises(cm.CmpmcError, match="scratch"):
        eng.run_workload(wl, 20000, seed=1, hit_capacity=1024, no_alloc=True)
    with pytest.raises(cm.CmpmcError, match="scratch"):
        eng.run_workload(wl, 1000, seed=1, form=cm.FORM_BATCHED, hit_capacity=1024, no_alloc=True)
    # a hit buffer beyond the scratch fails when it is carved
    with pytest.raises(cm.CmpmcError, match="scratch"):
        eng.run_workload(wl, 100, seed=1, hit_capacity=2_000_000, no_alloc=True)
    # the ordinary entry point and the no-allocation one still agree afterwards
    a = eng.run_workload(wl, 300, seed=2, hit_capacity=30000, no_alloc=True)
    b = eng.run_workload(wl, 300, seed=2, hit_capacity=30000)
    _same(a, b, "after refusals")
