#!/usr/bin/env python3
"""Benchmark of the batched FP32 phonon transport on one MI355X.

Workload (the batched configuration): Ge cylinder R = 3.81 cm, half-height 1.27 cm, Al
Kaplan films on both faces, default surfaces; 1e6 events per step, each with a
lognormal phonon count (median 200, sigma 1, clipped to [1, 20000]), positions uniform
in the volume, modes by the Ge density of states, energies uniform in [1, 31] meV; FP32
kinematics with FP64 time and ledger, one persistent launch. Inputs are generated on
the device from the event seeds before the timed region; a step is one transport pass
over the whole batch.

Printed JSON line (one): value = events/s of the batched form. Beside it:
  plain_*        the plain per-event FP64 kernel on a subset of the same events
  c1_*           the 7.5 meV single-phonon workload: batched vs plain events/s
  roofline       200 B per phonon step (read + write of a 100 B FP32 track record) over
                 the transport kernels' device time, against 8 TB/s HBM
  cpu_baseline   the reference implementation compiled in place (oracle/_ref) or the
                 CPU oracle, on the host cores, over a bounded sample of the workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "phonon events/s (7.5 meV Ge), batched vs plain GPU; grad/fwd cost ratio"
BYTES_PER_STEP_FP32 = 200.0
HBM_PEAK = 8.0e12


def _stats_line(r):
    s = r["stats"]
    return s["total_steps"], s["total_injected_eV"], s["total_deposited_eV"], s["escaped_or_killed_eV"]


def cpu_baseline(n_events: int, seed: int) -> dict:
    """Times the reference (compiled in place) or the oracle on the host cores over a
    bounded sample of the same workload (the plain generator draw order)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes as O
    threads = min(os.cpu_count() or 1, 16)
    wl = O.workload(O.WL_PHONON_LOGN)
    kind = "reference" if O.ref_available() else "port"
    t = time.perf_counter()
    if kind == "reference":
        r = O.ref_run(wl, n_events, seed, threads=threads, keep_hits=False, hit_cap=1)
    else:
        r = O.Engine().run(wl, n_events, seed, threads=threads, keep_hits=False)
    wall = time.perf_counter() - t
    return {"value": n_events / wall, "unit": "events/s", "cores": threads, "kind": kind,
            "sample": f"first {n_events} events of the lognormal workload, seed {seed}, keep_hits=false, "
                      f"{r['stats']['total_steps']} phonon steps in {wall:.1f} s",
            "steps_per_s": r["stats"]["total_steps"] / wall}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=1_000_000)
    ap.add_argument("--plain-events", type=int, default=20_000)
    ap.add_argument("--c1-events", type=int, default=1_000_000)
    ap.add_argument("--cpu-events", type=int, default=2_000)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "multi-GPU sharding is described in DESIGN.md but not implemented", file=sys.stderr)
        sys.exit(2)

    import cmpmc_amd as cm
    info = cm.device_info()
    eng = cm.Engine()
    wl = cm.workload(cm.WL_PHONON_LOGN)
    kw = dict(keep_hits=False, form=cm.FORM_BATCHED)

    # warmup: the same workload on a 1% slice (allocations, code objects, caches)
    for _ in range(args.warmup):
        eng.run_workload(wl, max(1000, args.events // 100), seed=args.seed + 1, **kw)
        print(f"warmup done", file=sys.stderr, flush=True)

    times, kms, steps = [], [], 0
    for k in range(args.steps):
        t = time.perf_counter()
        r = eng.run_workload(wl, args.events, seed=args.seed + 1000 * k, **kw)
        times.append(time.perf_counter() - t)
        kms.append(r["kernel_ms"])
        s = r["stats"]
        steps += s["total_steps"]
        closure = abs(s["total_injected_eV"] - s["total_deposited_eV"] - s["escaped_or_killed_eV"]) / s["total_injected_eV"]
        print(f"step {k}: {times[-1]:.2f} s, kernels {kms[-1] / 1e3:.2f} s, {s['total_steps']:.3e} phonon steps, "
              f"ledger closure {closure:.1e}", file=sys.stderr, flush=True)
    wall = sum(times)
    value = args.events * args.steps / wall
    kernel_s = sum(kms) / 1e3
    steps_per_s_kernel = steps / kernel_s

    # plain per-event FP64 kernel on a subset of the same events
    t = time.perf_counter()
    rp = eng.run_workload(wl, args.plain_events, seed=args.seed, keep_hits=False, form=cm.FORM_PLAIN)
    plain_wall = time.perf_counter() - t
    plain_value = args.plain_events / plain_wall
    print(f"plain: {args.plain_events} events in {plain_wall:.2f} s", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (device generator, per-event seeds)",
        "config": {"workload": "batched FP32 phonon transport: Ge cylinder R=3.81 cm, h=1.27 cm, Al Kaplan films, "
                               "lognormal(200, 1) phonons/event clipped to [1, 20000], uniform volume, DOS modes, "
                               "E ~ U[1, 31] meV", "events_per_step": args.events, "form": "batched",
                   "device": info},
        "phonon_steps_per_s": steps / wall,
        "plain_events_per_s": plain_value,
        "plain_sample_events": args.plain_events,
        "batched_vs_plain": value / plain_value,
    }
    if not args.no_c1:
        c1 = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
        t = time.perf_counter()
        rb = eng.run_workload(c1, args.c1_events, seed=args.seed, keep_hits=False, form=cm.FORM_BATCHED)
        b_wall = time.perf_counter() - t
        t = time.perf_counter()
        rpl = eng.run_workload(c1, args.c1_events, seed=args.seed, keep_hits=False, form=cm.FORM_PLAIN)
        p_wall = time.perf_counter() - t
        out["c1_events"] = args.c1_events
        out["c1_batched_events_per_s"] = args.c1_events / b_wall
        out["c1_plain_events_per_s"] = args.c1_events / p_wall
        out["c1_batched_vs_plain"] = p_wall / b_wall
        out["c1_batched_steps_per_s"] = rb["stats"]["total_steps"] / b_wall
        print(f"c1: batched {b_wall:.2f} s, plain {p_wall:.2f} s", file=sys.stderr, flush=True)
    achieved = steps_per_s_kernel * BYTES_PER_STEP_FP32
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_batched_lognormal.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
    out["roofline"] = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK, "traffic": traffic,
                       "kernel_ms_per_step": sum(kms) / args.steps}
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_events, args.seed)
        except Exception as e:  # the CPU leg must not hide the GPU result
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
