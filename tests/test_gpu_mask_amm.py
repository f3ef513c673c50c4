"""Electrode masks and AMM film entry on the GPU, both forms, against the CPU oracle
(whose masked runs are bit-exact against the reference: tests/test_oracle.py golden runs
`mask_amm` and `mask_pairs`, and whose AMM tables match the reference's and its own
physics checks: test_amm_golden, test_amm_physics).

Detector: the QET layout (SurfaceProperty::electrode_mask = supercdms_qet_mask, fill 0.9)
and AMM film entry Ge -> Al (SurfaceProperty::amm_interface) on both flat faces.

* phonons (7.5 meV L at the centre): plain form (FP64, reference streams) - first-step
  branch codes equal the oracle's; batched form (FP32) - first-step branch counts equal
  the oracle's batched mode; both - ledger closure, branch totals within Poisson errors,
  the readout-channel distribution of the hits (chi-square), hit times and energies
  (KS), and every hit's channel equal to the pad geometrically under it;
* e-/h+ pairs at 4 V collected through a three-pad mask (core disk and two half-annuli,
  fill 0.9, gaps of a few micrometres): collected carriers per channel, carriers
  reflected from inactive metal and gaps. (A carrier reflected at a face at 4 V hops
  well under a micrometre before the field returns it, so one landing in the QET
  layout's millimetre-wide gaps random-walks for up to the reference's 20 M steps: that
  layout is exercised with phonons only.)
* the batched form's FP32 AMM entry table is covered by the phonon comparison: a wrong
  table moves `sensor_hits` and the branch totals away from the oracle's;
* invalid pads and media are refused at model creation.
"""
import numpy as np
import pytest
from scipy import stats as sst

pytestmark = pytest.mark.gpu

N_EV = 2000
FILL = 0.9
GE = (5320.0, 5310.0, 3250.0)


def _detectors(cm, O, bias=None):
    d = cm.default_detector()
    al = cm.medium_from_bulk_shear(2720.0, 76e9, 26e9)
    for i in (0, 1):
        cm.set_pads(d.surf[i], cm.qet_mask(d.radius, FILL))
        cm.set_amm(d.surf[i], GE, al)
    od = O.mask_amm_detector(FILL)
    if bias is not None:
        cm.set_bias(d, bias)
        O.set_bias(od, bias)
        d.luke_sample_fraction = od.luke_sample_fraction = 0.0
    return d, od


def _closure(s):
    inj = s["total_injected_eV"]
    return abs(inj - s["total_deposited_eV"] - s["escaped_or_killed_eV"]) / inj


def _count_ok(a, b, name, nsig=4.0, rel_floor=0.02):
    tol = max(nsig * np.sqrt(a + b + 1.0), rel_floor * max(a, b))
    assert abs(a - b) <= tol, f"{name}: {a} vs {b} (tolerance {tol:.0f})"


def _channels_match(g, o, name, nch=5):
    """Two samples of channel ids (0..nch-1) from the same distribution: chi-square
    contingency test p > 0.01."""
    cg = np.bincount(g, minlength=nch)[:nch]
    co = np.bincount(o, minlength=nch)[:nch]
    keep = (cg + co) > 0
    p = sst.chi2_contingency(np.vstack([cg[keep], co[keep]]))[1]
    print(f"{name}: channels gpu {cg.tolist()} oracle {co.tolist()} p = {p:.3g}", flush=True)
    assert p > 0.01, f"{name}: channel distributions differ (p = {p:.3g})"


def _geometric_channels(cm, radius, hits):
    pads = cm.qet_mask(radius, 1.0)
    return np.array([cm.find_channel(pads, float(h["pos"][0]), float(h["pos"][1])) for h in hits], dtype=np.int32)


BRANCH_KEYS = ("total_steps", "isotope_scatters", "decays_L_to_TT", "decays_L_to_LT", "surface_refls",
               "sensor_hits", "killed_tracks")


def test_masked_amm_phonons_both_forms(lib_cmpmc, oracle):
    cm, O = lib_cmpmc, oracle
    d, od = _detectors(cm, O)
    eng, oeng = cm.Engine(detector=d), O.Engine(detector=od)
    wl = cm.workload(cm.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    owl = O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    seed = 4242
    print(f"case: mask + AMM phonons, plain and batched, n={N_EV} seed={seed}", flush=True)
    gp = eng.run_workload(wl, N_EV, seed=seed, form=cm.FORM_PLAIN, first_branch=True)
    op = oeng.run(owl, N_EV, seed, first_branch=True)
    assert np.array_equal(gp["first_branch"], op["first_branch"]), "plain first-step branch codes differ"
    same = np.diff(gp["hit_offsets"].astype(np.int64)) == np.diff(op["hit_offsets"].astype(np.int64))
    print(f"plain per-event hit counts identical for {same.mean():.4f} of events", flush=True)
    assert same.mean() >= 0.8
    gb = eng.run_workload(wl, N_EV, seed=seed, form=cm.FORM_BATCHED, first_branch=True)
    ob = oeng.run(owl, N_EV, seed, batched=True, first_branch=True)
    assert np.array_equal(np.bincount(gb["first_branch"], minlength=16), np.bincount(ob["first_branch"], minlength=16))
    base = cm.Engine().run_workload(wl, N_EV, seed=seed, form=cm.FORM_BATCHED, keep_hits=False)
    for r, o, form in ((gp, op, "plain"), (gb, ob, "batched")):
        s, os_ = r["stats"], o["stats"]
        assert _closure(s) <= 1e-9, f"{form}: ledger closure {_closure(s):.2e}"
        for k in BRANCH_KEYS:
            _count_ok(s[k], os_[k], f"{form} {k}")
        h, q = r["hits"], o["hits"]
        assert len(h) == s["sensor_hits"] > 1000
        assert np.all((h["channel_id"] >= 0) & (h["channel_id"] <= 4)), f"{form}: hit off the QET channels"
        _channels_match(h["channel_id"], q["channel"], f"{form} phonon hits")
        for f in ("time", "edep"):
            p = sst.ks_2samp(h[f], q[f]).pvalue
            assert p > 0.01, f"{form} hit {f}: KS p = {p:.3g}"
        geo = _geometric_channels(cm, d.radius, h)
        mismatch = np.mean(geo != h["channel_id"])
        print(f"{form}: hit channel differs from the pad under it for {mismatch:.2e} of hits", flush=True)
        # FP32 positions may sit on the other side of a pad edge within rounding
        assert mismatch <= (0.0 if form == "plain" else 2e-3)
        # masked AMM faces change the transport relative to uniform films
        assert s["total_steps"] != base["stats"]["total_steps"]


def _pair_mask(radius):
    """Core disk (channel 0) and two half-annuli (1, 2), fill 0.9, gaps of 2-3 um."""
    r0, g = 0.5 * radius, 2e-6
    return [(0, 0, 0.0, 0.0, 0.0, r0, 0.0, 2 * np.pi, FILL),
            (1, 1, 0.0, 0.0, r0 + g, radius, 0.0, np.pi, FILL),
            (1, 2, 0.0, 0.0, r0 + g, radius, np.pi + 1e-4, 2 * np.pi - 1e-4, FILL)]


def test_masked_pairs_both_forms(lib_cmpmc, oracle):
    cm, O = lib_cmpmc, oracle
    d, od = cm.default_detector(), O.default_detector()
    for i in (0, 1):
        cm.set_pads(d.surf[i], _pair_mask(d.radius))
        O.set_pads(od.surf[i], _pair_mask(od.radius))
    cm.set_bias(d, 4.0)
    O.set_bias(od, 4.0)
    d.luke_sample_fraction = od.luke_sample_fraction = 0.0
    eng, oeng = cm.Engine(detector=d), O.Engine(detector=od)
    wl, owl = cm.workload(cm.WL_PAIR), O.workload(O.WL_PAIR)
    n, seed = 1000, 7
    print(f"case: masked pairs at 4 V, n={n} seed={seed}", flush=True)
    dd = cm.default_detector()
    cm.set_bias(dd, 4.0)
    dd.luke_sample_fraction = 0.0
    bare = cm.Engine(detector=dd).run_workload(wl, n, seed=seed, form=cm.FORM_PLAIN)
    for form, batched in ((cm.FORM_PLAIN, False), (cm.FORM_BATCHED, True)):
        g = eng.run_workload(wl, n, seed=seed, form=form)
        o = oeng.run(owl, n, seed, batched=batched)
        gs, os_ = g["stats"], o["stats"]
        tag = "batched" if batched else "plain"
        assert gs["charges_collected"] == os_["charges_collected"] == 2 * n
        _count_ok(gs["charge_surface_refls"], os_["charge_surface_refls"], f"{tag} charge_surface_refls")
        assert gs["charge_surface_refls"] > bare["stats"]["charge_surface_refls"] + 100, \
            "inactive metal and gaps must reflect carriers"
        h = g["hits"]
        face = h["surface_id"] < 2
        assert np.all((h["channel_id"][face] >= 0) & (h["channel_id"][face] <= 2))
        assert np.all(h["channel_id"][~face] == -1)
        oh = o["hits"]
        _channels_match(h["channel_id"][face], oh["channel"][oh["surface"] < 2], f"{tag} carriers", nch=3)
        pads = [list(r[:8]) + [1.0] for r in _pair_mask(d.radius)]
        geo = np.array([cm.find_channel(pads, float(x), float(y)) for x, y in h["pos"][face][:, :2]])
        assert np.mean(geo != h["channel_id"][face]) <= (0.0 if form == cm.FORM_PLAIN else 2e-3)


def test_mask_rejects_bad_pads(lib_cmpmc):
    """Invalid pad descriptions are refused at model creation (no device work)."""
    cm = lib_cmpmc
    d = cm.default_detector()
    cm.set_pads(d.surf[0], [(7, 0, 0.0, 0.0, 0.0, 0.01, 0.0, 6.28, 1.0)])
    with pytest.raises(cm.CmpmcError, match="unknown shape"):
        cm.Engine(detector=d)
    d = cm.default_detector()
    cm.set_amm(d.surf[0], (0.0, 5000.0, 3000.0), GE)
    with pytest.raises(cm.CmpmcError, match="AMM media"):
        cm.Engine(detector=d)
