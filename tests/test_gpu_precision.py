"""FP32 batched kinematics against FP64 plain on the same events (cmpmc_amd.precision).

The per-event observables (hits, steps, decays and reflections per event) are means over
independent events, so their z-scores are exact statements of agreement; the hit means and
KS p-values treat correlated hits of one event as independent and are held to looser
bounds. Both forms close the energy ledger to 1e-9.
"""
import math

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kin_mode", [0, 1], ids=["table", "exact"])
def test_fp32_agrees_with_fp64_per_observable(lib_cmpmc, kin_mode):
    cm = lib_cmpmc
    from cmpmc_amd import precision as P
    n = 20_000 if kin_mode == 0 else 6_000
    print(f"case: precision study, 7.5 meV L from the origin, n={n}, kinematics {kin_mode}", flush=True)
    wl = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    r = P.study(cm, wl, n, seed=5, kin_mode=kin_mode)
    c = r["fp32_vs_fp64"]
    print({k: round(v, 3) for k, v in c.items()}, flush=True)
    for form in ("fp64_plain", "fp32_batched"):
        assert r[form]["ledger_closure"] <= 1e-9
    assert abs(c["z_hits_per_event"]) < 4.5
    assert abs(c["z_hit_time_us"]) < 5.0 and abs(c["z_hit_edep_meV"]) < 5.0
    assert min(c["ks_p_time"], c["ks_p_edep"]) > 1e-4
    a, b = r["fp64_plain"], r["fp32_batched"]
    # steps, decays and reflections per event: Poisson-like counts, relative agreement
    assert abs(a["steps_per_event"] - b["steps_per_event"]) / a["steps_per_event"] < 0.03
    assert abs(a["decays_per_event"] - b["decays_per_event"]) < 5 * math.sqrt(a["decays_per_event"] / n) + 0.02
    assert abs(a["deposited_over_injected"] - b["deposited_over_injected"]) < 2e-4
