"""FP32-against-FP64 precision study of the transport forms.

The plain form steps every track in FP64 (the reference's operation order); the batched
form steps in FP32 with FP64 time-of-flight, energies, weights and ledger. Both are run
on the same workload and seed, and every observable a user reads from a run is compared:
per-event rates (hits, steps, decays, reflections), hit means (arrival time, deposited
energy, radius) with z-scores from their standard errors, the deposited fraction of the
injected energy, the top-face share of the hits, and two-sample Kolmogorov-Smirnov
p-values of the hit time, energy and radius distributions.

A hit's time, energy and radius are correlated with the other hits of its event, so the
KS p-values and hit-mean z-scores treat more samples as independent than there are; the
per-event z-scores (hits, steps, decays, reflections per event) do not.
"""
from __future__ import annotations

import math

import numpy as np


def _ks_2samp(a: np.ndarray, b: np.ndarray) -> float:
    """Two-sample Kolmogorov-Smirnov p-value (asymptotic Kolmogorov distribution)."""
    from scipy.stats import ks_2samp
    if len(a) == 0 or len(b) == 0:
        return float("nan")
    return float(ks_2samp(a, b).pvalue)


def _mean_se(x: np.ndarray) -> tuple[float, float]:
    if len(x) < 2:
        return float(np.mean(x)) if len(x) else float("nan"), float("nan")
    return float(np.mean(x)), float(np.std(x, ddof=1) / math.sqrt(len(x)))


def _z(a: tuple[float, float], b: tuple[float, float]) -> float:
    se = math.hypot(a[1], b[1])
    return (a[0] - b[0]) / se if se > 0 else float("nan")


def observables(run: dict, n_events: int) -> dict:
    """The observables of one run (run_workload's result, hits kept)."""
    s = run["stats"]
    h = run["hits"]
    off = np.asarray(run["hit_offsets"], dtype=np.int64)
    per_event = np.diff(off).astype(np.float64)
    r = np.hypot(h["pos"][:, 0], h["pos"][:, 1]) if len(h) else np.zeros(0)
    return {
        "events": n_events,
        "hits_per_event": _mean_se(per_event),
        "hit_time_us": _mean_se(h["time"] * 1e6),
        "hit_edep_meV": _mean_se(h["edep"] * 1e3),
        "hit_radius_mm": _mean_se(r * 1e3),
        "deposited_over_injected": s["total_deposited_eV"] / s["total_injected_eV"],
        "ledger_closure": abs(s["total_injected_eV"] - s["total_deposited_eV"] - s["escaped_or_killed_eV"])
        / s["total_injected_eV"],
        "steps_per_event": s["total_steps"] / n_events,
        "decays_per_event": (s["decays_L_to_TT"] + s["decays_L_to_LT"]) / n_events,
        "reflections_per_event": s["surface_refls"] / n_events,
        "top_face_fraction": float(np.mean(h["surface_id"] == 0)) if len(h) else float("nan"),
        "_samples": {"time": h["time"], "edep": h["edep"], "radius": r},
        "_counts": {"steps": s["total_steps"], "decays": s["decays_L_to_TT"] + s["decays_L_to_LT"],
                    "refls": s["surface_refls"]},
    }


def compare(ref: dict, other: dict) -> dict:
    """z-scores and KS p-values of `other` against `ref` (both from observables())."""
    out = {}
    for k in ("hits_per_event", "hit_time_us", "hit_edep_meV", "hit_radius_mm"):
        out["z_" + k] = _z(other[k], ref[k])
    for k in ("time", "edep", "radius"):
        out["ks_p_" + k] = _ks_2samp(ref["_samples"][k], other["_samples"][k])
    return out


def _public(o: dict) -> dict:
    return {k: (list(v) if isinstance(v, tuple) else v) for k, v in o.items() if not k.startswith("_")}


def study(cm, workload, n_events: int, seed: int, kin_mode: int = 0, crystal=None, detector=None,
          with_table_fp32: bool = False) -> dict:
    """Plain FP64 against batched FP32 on the same events (kinematics mode kin_mode: 0 the
    octahedral table, 1 the exact Christoffel solve). with_table_fp32 adds the batched form
    with the table when kin_mode is 1, for the table's own lookup error."""
    lat = cm.Lattice(crystal)
    lat.set_kinematics_mode(kin_mode)
    eng = cm.Engine(lat, detector)
    try:
        rp = eng.run_workload(workload, n_events, seed=seed, form=cm.FORM_PLAIN)
        rb = eng.run_workload(workload, n_events, seed=seed, form=cm.FORM_BATCHED)
        op, ob = observables(rp, n_events), observables(rb, n_events)
        res = {"kinematics": "exact" if kin_mode == 1 else "octahedral table", "events": n_events, "seed": seed,
               "fp64_plain": _public(op), "fp32_batched": _public(ob), "fp32_vs_fp64": compare(op, ob)}
        if with_table_fp32 and kin_mode == 1:
            eng.set_kinematics(0)
            rt = eng.run_workload(workload, n_events, seed=seed, form=cm.FORM_BATCHED)
            ot = observables(rt, n_events)
            res["fp32_batched_table"] = _public(ot)
            res["fp32_table_vs_fp64_exact"] = compare(op, ot)
        return res
    finally:
        eng.close()
