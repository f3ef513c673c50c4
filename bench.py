#!/usr/bin/env python3
"""Benchmark of the track-transport path on one MI355X.

Headline workload (a step): 1,000,000 phonon events in one batched FP32 persistent
launch.  Detector: Ge cylinder R = 3.81 cm, half-height 1.27 cm, Al Kaplan films on both
flat faces, default Ge material and surfaces; each event is one L-mode 7.5 meV phonon
started at the origin with an isotropic direction, tracked to absorption.  Inputs are
generated on the device from the per-event seeds before the timed region.

value        events/s of the batched FP32 form over the timed steps, with the library's
             default launch options (which the autotuner, --tune, also picked for this
             workload) (wall clock around the library call, inputs generated on the device)
plain_*      the plain per-event FP64 kernel on the same 1e6 events (one run)
grad_*       the differentiated workload: 1e5 events of U{1..64} L-mode 7.5 meV phonons
             from uniform positions, theta = (Ge B, Ge A, f_diffuse 0.9, p_film 0.2);
             grad_fwd_ratio = wall time of (differentiable forward + backward kernel) /
             wall time of the batched forward of the same events at the same parameters
charge_*     e-/h+ pairs at 4 V across the 2.54 cm height, uniform in the volume, trapping
             on (electron and hole trapping lengths 0.5 m), Luke phonons counted: 1e5 pairs,
             batched FP32 and plain FP64 (pairs/s, charge steps/s)
             (charge.leman: the same pairs with Leman second-order stepping)
luke_*       the same pairs (trapping off) with every Luke phonon enqueued on the device
             and tracked to absorption: 2,000 events batched (the plain form on a slice
             when --luke-plain-events is given: its one-event-per-lane time is set by the
             slowest events, ~50-80 s)
deposit      the reference's end-to-end recoil demo: 100 eV nuclear recoils at the centre,
             partitioned on the device (EnergyPartition: Lindhard yield, Fano pair number,
             prompt phonons down-sampled to 1,000 tracks), drifted at 4 V with 5 % of the
             Luke phonons tracked: 2,000 events batched (events/s)
lognormal_*  the lognormal workload (phonon count per event lognormal(200, 1) clipped to
             [1, 20000], uniform positions, DOS modes, E ~ U[1, 31] meV) on a bounded
             slice of events, batched FP32 (and the plain kernel on a smaller slice when
             --lognormal-plain-events is given)
roofline     the binding resource of the batched phonon kernel, VALU issue: achieved =
             VALU wave-instructions per phonon step (SQ_INSTS_VALU over steps, from the
             committed rocprofv3 counters of the same kernel and workload) x this
             launch's steps / its device time (HIP events on the launch stream), against
             the issue peak (one wave64 VALU instruction per 4 cycles per SIMD, 4 SIMDs
             per CU, at the clock the counter run measured); beside it the measured HBM
             traffic per launch (rocprof FETCH_SIZE + WRITE_SIZE) against 8 TB/s, and the
             200 B/step record model of the path's bytes
cpu_baseline the reference implementation compiled in place (oracle/_ref) or else the
             CPU oracle, timed on the host: all cores the process may use (the box's CPU
             share) and one thread, on bounded samples of the headline workload; the
             same beside the lognormal, charge-drift, Luke-cascade and deposit lines
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
BYTES_PER_STEP_FP64 = 320.0
HBM_PEAK = 8.0e12
# the reference's end-to-end recoil demo (examples/phonon_benchmark.cpp:170-203): a 100 eV
# nuclear recoil partitioned by EnergyPartition (prompt phonons down-sampled to 1,000
# tracks), drifted at 4 V with 5 % of the Luke phonons tracked; here at the detector's centre
DEPOSIT_WL = dict(mode=1, energy_eV=100.0, pair_at_origin=1, partition=dict(max_sampled_phonons=1000))
# VALU issue of a CDNA compute unit: 4 SIMDs, one wave64 VALU instruction per 4 cycles each
VALU_WAVE_INSTS_PER_CU_CYCLE = 1.0


# FEM mesh fields of the mesh leg: a tetrahedral Kuhn grid over the crystal's bounding box
# (48 x 48 x 16 voxels, 221,184 tetrahedra) and an axisymmetric (r, z) grid (64 x 32
# cells, 4,096 triangles), potentials of about 4 V across the height with quadratic terms
# (V = 157.48 z + 2000 x^2 - 800 y z + 5000 z^2; V = 157.48 z + 1500 r^2 + 3000 r z - 4000 z^2);
# None: the uniform 4 V field on the same pairs.
MESH_CASES = {
    "uniform": None,
    "tet_grid": ("3d", (0.0382, 0.0382, 0.0128), (48, 48, 16), (0.0, 0.0, 0.0, 157.48, 2000.0, -800.0, 5000.0)),
    "rz_grid": ("rz", (0.0, 0.0382, -0.0128, 0.0128), (64, 32), (0.0, 0.0, 157.48, 1500.0, 3000.0, -4000.0)),
}


def mesh_from_spec(mesh_cls, spec):
    """A mesh of MESH_CASES with the library's or the oracle's Mesh class (the same vertex
    positions and potentials)."""
    kind, box, n, c = spec
    if kind == "3d":
        x = mesh_cls.grid3d_vertices(*box, *n)
        V = c[0] + c[1] * x[:, 0] + c[2] * x[:, 1] + c[3] * x[:, 2] + c[4] * x[:, 0] * x[:, 0] + \
            c[5] * x[:, 1] * x[:, 2] + c[6] * x[:, 2] * x[:, 2]
        if mesh_cls.__module__.startswith("oracle"):
            return mesh_cls.grid3d(*box, *n, c)
        return mesh_cls.grid3d(*box, *n, V)
    u = mesh_cls.grid2d_vertices(*box, *n)
    if mesh_cls.__module__.startswith("oracle"):
        return mesh_cls.grid2d(*box, *n, c, True)
    V = c[0] + c[1] * u[:, 0] + c[2] * u[:, 1] + c[3] * u[:, 0] * u[:, 0] + c[4] * u[:, 0] * u[:, 1] + \
        c[5] * u[:, 1] * u[:, 1]
    return mesh_cls.grid2d(*box, *n, V, True)


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _closure(s):
    return abs(s["total_injected_eV"] - s["total_deposited_eV"] - s["escaped_or_killed_eV"]) / s["total_injected_eV"]


def host_info() -> dict:
    """The host the CPU legs run on: logical CPUs, those this process may use, the
    threads the legs use (the CPU share the environment sets, else all usable CPUs) and
    the CPU model."""
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "threads": min(usable, share) if share > 0 else usable,
            "cpu_model": model}


def cpu_leg(wl_kind: str, n_events: int, seed: int, threads: int, crystal=None, detector=None) -> dict:
    """Wall time of the reference's run_events (compiled in place) or else the oracle, on
    `threads` host threads, generator and transport timed, lattice build excluded
    (parallel.hpp:79, 146), without hit storage."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes as O
    wl = {"c1": lambda: O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3),
          "lognormal": lambda: O.workload(O.WL_PHONON_LOGN),
          "pair": lambda: O.workload(O.WL_PAIR),
          "deposit": lambda: O.workload(O.WL_DEPOSIT, **DEPOSIT_WL)}[wl_kind]()
    kind = "reference" if O.ref_available() else "port"
    chunk = 16 if wl_kind == "c1" else 1
    t = time.perf_counter()
    if kind == "reference":
        r = O.ref_run(wl, n_events, seed, detector=detector, crystal=crystal, threads=threads, chunk=chunk,
                      keep_hits=False, hit_cap=1)
    else:
        r = O.Engine(crystal=crystal, detector=detector).run(wl, n_events, seed, threads=threads, chunk=chunk,
                                                             keep_hits=False)
    wall = time.perf_counter() - t
    s = r["stats"]
    return {"events": n_events, "threads": threads, "kind": kind, "wall_s": wall, "events_per_s": n_events / wall,
            "phonon_steps_per_s": s["total_steps"] / wall, "charge_steps_per_s": s["charge_steps"] / wall}


def cpu_pair(wl_kind: str, n_all: int, n_one: int, seed: int, host: dict, crystal=None, detector=None) -> dict:
    """The same workload on all the host threads and on one thread (the first events of
    the seed in each)."""
    a = cpu_leg(wl_kind, n_all, seed, host["threads"], crystal, detector)
    o = cpu_leg(wl_kind, n_one, seed, 1, crystal, detector)
    _log(f"cpu {wl_kind}: {a['events_per_s']:.4g} events/s on {a['threads']} threads, "
         f"{o['events_per_s']:.4g} on 1 thread")
    return {"all_cores": a, "one_thread": o}


def oracle_configs():
    """Crystal and detector of the charge-drift and Luke-cascade lines for the CPU legs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes as O
    c3 = O.germanium()
    c3.e_trapping_mfp_m = c3.h_trapping_mfp_m = 0.5
    d3 = O.default_detector()
    O.set_bias(d3, 4.0)
    d3.luke_sample_fraction = 0.0
    d4 = O.default_detector()
    O.set_bias(d4, 4.0)
    d4.luke_sample_fraction = 1.0
    return (c3, d3), (O.germanium(), d4)


def roofline(steps_per_launch: float, launch_s: float, device: str) -> dict:
    """The batched phonon kernel against its binding resource, VALU issue, from the
    committed counter summary of the same kernel and workload (VALU wave-instructions per
    phonon step, the clock of that run); HBM traffic per launch as the secondary ratio."""
    prof = os.path.join(ROOT, "profiles", "valu_batched_phonon_7p5meV.json")
    p = json.load(open(prof))
    per_step = p["valu_wave_insts_per_step"]
    n_cu = int(device.split(",")[1].split()[0])
    peak = VALU_WAVE_INSTS_PER_CU_CYCLE * n_cu * p["clock_hz"]
    achieved = per_step * steps_per_launch / launch_s
    hbm = p.get("hbm_bytes_per_launch")
    r = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "wave-instr/s", "frac": achieved / peak,
         "traffic": hbm, "kernel": "transport_batched_kernel", "kernel_ms_avg": launch_s * 1e3,
         "phonon_steps_per_launch": steps_per_launch, "valu_wave_insts_per_step": per_step,
         "clock_hz": p["clock_hz"], "source": "profiles/" + p["source"]}
    if hbm:
        r["hbm"] = {"achieved_GBps": hbm / launch_s / 1e9, "peak_GBps": HBM_PEAK / 1e9,
                    "frac": hbm / launch_s / HBM_PEAK}
    # the path's per-step record model (a 100 B FP32 track read and written every step);
    # tracks stay in registers here, so this is a byte count of the work, not traffic
    r["record_model"] = {"bytes_per_step": BYTES_PER_STEP_FP32,
                         "GBps": steps_per_launch * BYTES_PER_STEP_FP32 / launch_s / 1e9}
    return r


def precision_study(path: str, seed: int) -> None:
    """Plain FP64 against batched FP32 on the same events, per observable
    (cmpmc_amd.precision): the headline shape (1e5 events of one 7.5 meV L phonon from the
    origin) and the differentiated workload's shape (U{1..64} L phonons of 7.5 meV from
    uniform positions), each with the octahedral table and with the exact Christoffel
    solve (the batched table form beside the exact runs)."""
    import cmpmc_amd as cm
    from cmpmc_amd import precision as P
    c1 = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    c5 = cm.workload(cm.WL_PHONON_UNIF, mode=0, energy_eV=7.5e-3, n_min=1, n_max=64)
    out = {"device": cm.device_info(), "seed": seed, "cases": {}}
    for name, wl, n, kin in (("headline_table", c1, 100_000, 0), ("uniform_count_table", c5, 5_000, 0),
                             ("headline_exact", c1, 100_000, 1), ("uniform_count_exact", c5, 2_000, 1)):
        t = time.perf_counter()
        out["cases"][name] = P.study(cm, wl, n, seed, kin_mode=kin, with_table_fp32=True)
        c = out["cases"][name]["fp32_vs_fp64"]
        _log(f"precision {name}: {n} events in {time.perf_counter() - t:.1f} s; z(hits/event) "
             f"{c['z_hits_per_event']:.2f}, z(time) {c['z_hit_time_us']:.2f}, KS p time {c['ks_p_time']:.3f}")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({"precision_study": path, "cases": list(out["cases"])}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=1_000_000)
    ap.add_argument("--plain-events", type=int, default=1_000_000)
    ap.add_argument("--lognormal-events", type=int, default=1_000_000)
    ap.add_argument("--grad-events", type=int, default=100_000)
    ap.add_argument("--charge-events", type=int, default=100_000)
    ap.add_argument("--luke-events", type=int, default=2_000)
    ap.add_argument("--luke-plain-events", type=int, default=0)
    ap.add_argument("--deposit-events", type=int, default=2_000)
    ap.add_argument("--mesh-events", type=int, default=100_000, help="pairs drifting in FEM mesh fields (batched)")
    ap.add_argument("--mesh-plain-events", type=int, default=10_000, help="the same, plain form")
    ap.add_argument("--cpu-events", type=int, default=100_000, help="headline CPU leg, all cores")
    ap.add_argument("--cpu-events-1t", type=int, default=5_000, help="headline CPU leg, one thread")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--waves", type=int, default=0, help="register-budget variant (0: library default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lognormal", action="store_true")
    ap.add_argument("--precision", metavar="PATH", default="",
                    help="run only the FP32-against-FP64 precision study and write its tables to PATH (JSON)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 1 GiB device-to-device copy (a kernel of known traffic for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (--gpus {args.gpus} requested); the path's sharding over "
              "several GPUs is described, not implemented", file=sys.stderr)
        sys.exit(2)

    if args.precision:
        precision_study(args.precision, args.seed)
        return

    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        del x, y
        _log("copy calibration: 3 x 1 GiB device copies done")

    import cmpmc_amd as cm
    info = cm.device_info()
    eng = cm.Engine()
    wl = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    tune = None
    opts = {"waves": args.waves}
    kw = dict(keep_hits=False, form=cm.FORM_BATCHED, **opts)

    def step(seed):
        return eng.run_workload(wl, args.events, seed=seed, **kw)

    for w in range(args.warmup):
        step(args.seed + 7919 * (w + 1))
        _log(f"warmup {w} done")

    times, kms, steps = [], [], 0
    for k in range(args.steps):
        t = time.perf_counter()
        r = step(args.seed + 1000 * k)
        times.append(time.perf_counter() - t)
        kms.append(r["kernel_ms"])
        s = r["stats"]
        steps += s["total_steps"]
        _log(f"step {k}: {times[-1]:.3f} s, kernel {kms[-1] / 1e3:.3f} s, {s['total_steps']:.4e} phonon steps, "
             f"ledger closure {_closure(s):.1e}")
    wall = sum(times)
    value = args.events * args.steps / wall
    kernel_s = sum(kms) / 1e3

    t = time.perf_counter()
    rp = eng.run_workload(wl, args.plain_events, seed=args.seed, keep_hits=False, form=cm.FORM_PLAIN)
    plain_wall = time.perf_counter() - t
    plain_value = args.plain_events / plain_wall
    _log(f"plain: {args.plain_events} events in {plain_wall:.3f} s (kernel {rp['kernel_ms'] / 1e3:.3f} s)")

    out = {
        "metric": METRIC, "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (device generator from per-event seeds)",
        "config": {"workload": "7.5 meV Ge phonon events: Ge cylinder R=3.81 cm, half-height 1.27 cm, Al Kaplan "
                               "films on both faces, default surfaces; one L-mode 7.5 meV phonon per event from "
                               "the origin, isotropic, tracked to absorption",
                   "events_per_step": args.events, "form": "batched FP32 persistent (FP64 time and ledger)",
                   "options": opts, "device": info},
        "autotune": tune,
        "phonon_steps_per_s": steps / wall,
        "plain_events_per_s": plain_value,
        "plain_phonon_steps_per_s": rp["stats"]["total_steps"] / plain_wall,
        "plain_events": args.plain_events,
        "batched_vs_plain": value / plain_value,
        "grad_fwd_ratio": None,
    }
    out["roofline"] = roofline(steps / args.steps, kernel_s / args.steps, info)
    out["plain_roofline"] = {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK / 1e9,
                             "achieved": rp["stats"]["total_steps"] * BYTES_PER_STEP_FP64 / (rp["kernel_ms"] / 1e3) / 1e9}
    out["plain_roofline"]["frac"] = out["plain_roofline"]["achieved"] * 1e9 / HBM_PEAK

    if args.grad_events > 0:
        import numpy as np
        cr = cm.germanium()
        theta = np.array([cr.B_iso, cr.A_anh, 0.9, 0.2])
        det = cm.default_detector()
        cm.apply_diff_params(cr, det, theta)
        e5 = cm.Engine(cm.Lattice(cr), det)
        w5 = cm.workload(cm.WL_PHONON_UNIF, mode=0, energy_eV=7.5e-3, n_min=1, n_max=64)
        g = np.zeros(66)
        g[2:] = (np.arange(64) + 0.5) / 64.0
        e5.run_workload(w5, 2000, seed=1, keep_hits=False, form=cm.FORM_BATCHED)      # warm
        eng.diff_forward(w5, 2000, theta, seed=1)[1].backward(g)
        t = time.perf_counter()
        rf = e5.run_workload(w5, args.grad_events, seed=args.seed, keep_hits=False, form=cm.FORM_BATCHED)
        t_fwd = time.perf_counter() - t
        t = time.perf_counter()
        y, tape, info = eng.diff_forward(w5, args.grad_events, theta, seed=args.seed)
        t_dfwd = time.perf_counter() - t
        t = time.perf_counter()
        grad = tape.backward(g)
        t_bwd = time.perf_counter() - t
        out["grad_fwd_ratio"] = (t_dfwd + t_bwd) / t_fwd
        out["grad"] = {"events": args.grad_events, "forward_s": t_fwd, "diff_forward_s": t_dfwd, "backward_s": t_bwd,
                       "backward_kernel_ms": tape.last_kernel_ms, "tape_hits": len(tape),
                       "phonon_steps": rf["stats"]["total_steps"], "theta": theta.tolist(),
                       "grad_time_objective": grad.tolist()}
        _log(f"grad: forward {t_fwd:.3f} s, differentiable forward {t_dfwd:.3f} s, backward {t_bwd * 1e3:.2f} ms, "
             f"ratio {out['grad_fwd_ratio']:.3f}")
        del tape

    if args.charge_events > 0 or args.luke_events > 0:
        pair = cm.workload(cm.WL_PAIR)

        def pair_engine(luke_fraction, trap_mfp, leman=False):
            cr = cm.germanium()
            cr.e_trapping_mfp_m = cr.h_trapping_mfp_m = trap_mfp
            det = cm.default_detector()
            cm.set_bias(det, 4.0)
            det.luke_sample_fraction = luke_fraction
            det.charge_stepping = 1 if leman else 0
            return cm.Engine(cm.Lattice(cr), det)

        def timed(e, n, form):
            t = time.perf_counter()
            r = e.run_workload(pair, n, seed=args.seed, keep_hits=False, form=form)
            return time.perf_counter() - t, r["stats"]

        if args.charge_events > 0:
            e3 = pair_engine(0.0, 0.5)
            timed(e3, 1000, cm.FORM_BATCHED)
            tb, sb = timed(e3, args.charge_events, cm.FORM_BATCHED)
            tp, sp = timed(e3, args.charge_events, cm.FORM_PLAIN)
            out["charge"] = {"pairs": args.charge_events, "batched_pairs_per_s": args.charge_events / tb,
                             "plain_pairs_per_s": args.charge_events / tp, "batched_vs_plain": tp / tb,
                             "batched_charge_steps_per_s": sb["charge_steps"] / tb,
                             "trapped_fraction": sb["charges_trapped"] / (2.0 * args.charge_events),
                             "luke_per_pair": sb["luke_phonons_emitted"] / args.charge_events,
                             "trap_mfp_m": 0.5}
            _log(f"charge drift: batched {tb:.3f} s, plain {tp:.3f} s for {args.charge_events} pairs")
            # the same pairs with Leman second-order stepping (engine.hpp:234-254)
            e3l = pair_engine(0.0, 0.5, leman=True)
            timed(e3l, 1000, cm.FORM_BATCHED)
            tbl, sbl = timed(e3l, args.charge_events, cm.FORM_BATCHED)
            tpl, spl = timed(e3l, args.charge_events, cm.FORM_PLAIN)
            out["charge"]["leman"] = {
                "batched_pairs_per_s": args.charge_events / tbl, "plain_pairs_per_s": args.charge_events / tpl,
                "batched_vs_plain": tpl / tbl, "speedup_vs_first_order_batched": tb / tbl,
                "charge_steps_per_pair": sbl["charge_steps"] / args.charge_events,
                "first_order_charge_steps_per_pair": sb["charge_steps"] / args.charge_events,
                "luke_per_pair": sbl["luke_phonons_emitted"] / args.charge_events}
            _log(f"charge drift, Leman: batched {tbl:.3f} s, plain {tpl:.3f} s")
        if args.luke_events > 0:
            e4 = pair_engine(1.0, 0.0)
            timed(e4, 100, cm.FORM_BATCHED)
            tb, sb = timed(e4, args.luke_events, cm.FORM_BATCHED)
            out["luke"] = {"events": args.luke_events, "batched_events_per_s": args.luke_events / tb,
                           "luke_phonons_per_pair": sb["luke_phonons_emitted"] / args.luke_events,
                           "deposited_eV_per_event": sb["total_deposited_eV"] / args.luke_events,
                           "phonon_steps_per_s": sb["total_steps"] / tb}
            if args.luke_plain_events > 0:
                tp, sp = timed(e4, args.luke_plain_events, cm.FORM_PLAIN)
                out["luke"]["plain_events"] = args.luke_plain_events
                out["luke"]["plain_events_per_s"] = args.luke_plain_events / tp
            _log(f"Luke cascade: batched {tb:.3f} s for {args.luke_events} events")
            e4l = pair_engine(1.0, 0.0, leman=True)
            timed(e4l, 100, cm.FORM_BATCHED)
            tbl, sbl = timed(e4l, args.luke_events, cm.FORM_BATCHED)
            out["luke"]["leman"] = {"batched_events_per_s": args.luke_events / tbl,
                                    "speedup_vs_first_order": tb / tbl,
                                    "luke_phonons_per_pair": sbl["luke_phonons_emitted"] / args.luke_events,
                                    "deposited_eV_per_event": sbl["total_deposited_eV"] / args.luke_events}
            _log(f"Luke cascade, Leman: batched {tbl:.3f} s")

    if args.deposit_events > 0:
        dd = cm.default_detector()
        cm.set_bias(dd, 4.0)
        dd.luke_sample_fraction = 0.05
        ed = cm.Engine(detector=dd)
        dw = cm.workload(cm.WL_DEPOSIT, **DEPOSIT_WL)
        ed.run_workload(dw, 20, seed=1, keep_hits=False, form=cm.FORM_BATCHED)      # warm
        t = time.perf_counter()
        rd = ed.run_workload(dw, args.deposit_events, seed=args.seed, keep_hits=False, form=cm.FORM_BATCHED)
        tdw = time.perf_counter() - t
        sd = rd["stats"]
        out["deposit"] = {"events": args.deposit_events, "batched_events_per_s": args.deposit_events / tdw,
                          "phonon_steps_per_s": sd["total_steps"] / tdw, "charge_steps": sd["charge_steps"],
                          "pairs_per_event": sd["charges_collected"] / (2.0 * args.deposit_events),
                          "deposited_eV_per_event": sd["total_deposited_eV"] / args.deposit_events,
                          "kernel_s": rd["kernel_ms"] / 1e3}
        _log(f"deposit (100 eV NR at 4 V): {args.deposit_events} events in {tdw:.2f} s, "
             f"{sd['total_steps']:.3e} phonon steps")

    if args.mesh_events > 0:
        # pairs at 4 V-scale bias with trapping on, as the charge-drift leg, in FEM mesh
        # fields: a tetrahedral grid over the crystal's bounding box and an axisymmetric
        # (r, z) triangle grid (MESH_CASES), against the uniform 4 V field
        pair = cm.workload(cm.WL_PAIR)
        out["mesh"] = {"pairs": args.mesh_events, "plain_pairs": args.mesh_plain_events}
        for name, spec in MESH_CASES.items():
            det = cm.default_detector()
            det.luke_sample_fraction = 0.0
            if spec is None:
                cm.set_bias(det, 4.0)
                mesh = None
            else:
                mesh = mesh_from_spec(cm.Mesh, spec)
                cm.set_mesh(det, mesh)
            cr = cm.germanium()
            cr.e_trapping_mfp_m = cr.h_trapping_mfp_m = 0.5
            em = cm.Engine(cm.Lattice(cr), det)
            em.run_workload(pair, 1000, seed=1, keep_hits=False, form=cm.FORM_BATCHED)      # warm
            t = time.perf_counter()
            rb = em.run_workload(pair, args.mesh_events, seed=args.seed, keep_hits=False, form=cm.FORM_BATCHED)
            tb = time.perf_counter() - t
            row = {"cells": 0 if mesh is None else mesh.num_cells, "batched_pairs_per_s": args.mesh_events / tb,
                   "batched_charge_steps_per_s": rb["stats"]["charge_steps"] / tb,
                   "batched_kernel_s": rb["kernel_ms"] / 1e3,
                   "charge_steps_per_pair": rb["stats"]["charge_steps"] / args.mesh_events}
            if args.mesh_plain_events > 0:
                t = time.perf_counter()
                rp = em.run_workload(pair, args.mesh_plain_events, seed=args.seed, keep_hits=False, form=cm.FORM_PLAIN)
                tp = time.perf_counter() - t
                row["plain_pairs_per_s"] = args.mesh_plain_events / tp
                row["plain_charge_steps_per_s"] = rp["stats"]["charge_steps"] / tp
            out["mesh"][name] = row
            _log(f"mesh field {name}: batched {tb:.3f} s for {args.mesh_events} pairs"
                 + (f", plain {tp:.3f} s for {args.mesh_plain_events}" if args.mesh_plain_events > 0 else ""))
            em.close()

    if not args.no_lognormal and args.lognormal_events > 0:
        lw = cm.workload(cm.WL_PHONON_LOGN)
        t = time.perf_counter()
        rl = eng.run_workload(lw, args.lognormal_events, seed=args.seed, **kw)
        lwall = time.perf_counter() - t
        out["lognormal_events"] = args.lognormal_events
        out["lognormal_events_per_s"] = args.lognormal_events / lwall
        out["lognormal_phonon_steps_per_s"] = rl["stats"]["total_steps"] / lwall
        out["lognormal_kernel_s"] = rl["kernel_ms"] / 1e3
        _log(f"lognormal: {args.lognormal_events} events in {lwall:.2f} s, {rl['stats']['total_steps']:.3e} steps, "
             f"closure {_closure(rl['stats']):.1e}")

    if not args.no_cpu:
        try:
            host = host_info()
            c1 = cpu_pair("c1", args.cpu_events, args.cpu_events_1t, args.seed, host)
            a, o = c1["all_cores"], c1["one_thread"]
            out["cpu_baseline"] = {
                "value": a["events_per_s"], "unit": "events/s", "cores": a["threads"], "kind": a["kind"],
                "sample": f"first {a['events']} events (all cores) and first {o['events']} events (one thread) of "
                          f"the headline 7.5 meV workload, seed {args.seed}, FP64 run_events without hit storage, "
                          f"generator + transport timed",
                "phonon_steps_per_s": a["phonon_steps_per_s"], "one_thread": o, "host": host}
            (c3, d3), (c4, d4) = oracle_configs()
            if "lognormal_events" in out:
                out["lognormal_cpu"] = cpu_pair("lognormal", 64, 4, args.seed, host)
            if "charge" in out:
                out["charge"]["cpu"] = cpu_pair("pair", 4000, 400, args.seed, host, crystal=c3, detector=d3)
            if "luke" in out:
                out["luke"]["cpu"] = cpu_pair("pair", 64, 4, args.seed, host, crystal=c4, detector=d4)
            if "mesh" in out:
                sys.path.insert(0, os.path.join(ROOT, "oracle"))
                import oracle_ctypes as O
                for name, spec in MESH_CASES.items():
                    if spec is None:
                        continue
                    od = O.default_detector()
                    od.luke_sample_fraction = 0.0
                    O.set_mesh(od, mesh_from_spec(O.Mesh, spec))
                    ocr = O.germanium()
                    ocr.e_trapping_mfp_m = ocr.h_trapping_mfp_m = 0.5
                    out["mesh"][name]["cpu"] = cpu_pair("pair", 2000, 200, args.seed, host, crystal=ocr, detector=od)
            if "deposit" in out:
                sys.path.insert(0, os.path.join(ROOT, "oracle"))
                import oracle_ctypes as O
                od = O.default_detector()
                O.set_bias(od, 4.0)
                od.luke_sample_fraction = 0.05
                out["deposit"]["cpu"] = cpu_pair("deposit", 32, 2, args.seed, host, detector=od)
        except Exception as e:  # the CPU leg must not hide the GPU result
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
