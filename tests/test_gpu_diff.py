"""Differentiable batched transport (likelihood-ratio gradients) on the GPU.

Workload: the differentiated configuration - events of U{1..64} L-mode 7.5 meV phonons
from uniform positions in the default detector, parameters theta = (B, A, f_diffuse,
p_film) at (Ge B, Ge A, 0.9, 0.2). Observables: collected energy on the top and bottom
faces and the energy-weighted 64-bin arrival-time histogram over [0, 1 ms).

The gradient of g . y is checked against +-1 % common-random-number central finite
differences of the same forward pass: both are Monte-Carlo estimates, so the check is
statistical - 10 replicate seeds give means and standard errors, and the two must
agree within 4.5 combined standard errors (the score terms of B sum thousands of free
flights per track, so its replicate spread is wide and its standard error is itself
uncertain with 9 degrees of freedom), while the gradient of the parameters the arrival
time depends on strongly (B, p_film) must be significant (> 5 standard errors).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T_MAX = 1e-3
F0, P0 = 0.9, 0.2


@pytest.fixture(scope="module")
def diff(lib_cmpmc):
    cm = lib_cmpmc
    eng = cm.Engine()
    cr = cm.germanium()
    theta0 = np.array([cr.B_iso, cr.A_anh, F0, P0])
    wl = cm.workload(cm.WL_PHONON_UNIF, mode=0, energy_eV=7.5e-3, n_min=1, n_max=64)
    return cm, eng, wl, theta0


def _g_time():
    g = np.zeros(66)
    g[2:] = (np.arange(64) + 0.5) * (T_MAX / 64) * 1e3      # bin-centre time in ms
    return g


def test_diff_forward_observables_consistent(diff):
    cm, eng, wl, theta0 = diff
    print("case: diff forward n=2000 seed=11", flush=True)
    y, tape, info = eng.diff_forward(wl, 2000, theta0, seed=11, t_max_s=T_MAX)
    s = info["stats"]
    assert len(tape) == s["sensor_hits"]
    assert y[0] + y[1] == pytest.approx(s["total_deposited_eV"], rel=1e-12)
    assert y[2:].sum() <= y[0] + y[1] + 1e-12
    assert y[2:].sum() > 0.5 * (y[0] + y[1])
    closure = abs(s["total_injected_eV"] - s["total_deposited_eV"] - s["escaped_or_killed_eV"]) / s["total_injected_eV"]
    assert closure < 1e-9
    # same seed, same numbers (fixed-point sums of a schedule-independent run)
    y2, tape2, _ = eng.diff_forward(wl, 2000, theta0, seed=11, t_max_s=T_MAX, grid_blocks=97)
    assert np.array_equal(y, y2)
    g = _g_time()
    a, b = tape.backward(g), tape2.backward(g)
    assert np.allclose(a, b, rtol=1e-9, atol=0.0)


def test_diff_forward_equals_batched_run_with_same_parameters(diff):
    """The differentiable forward transports exactly like the batched form on a model
    whose crystal and detector carry the same parameter values."""
    cm, eng, wl, theta0 = diff
    cr = cm.germanium()
    det = cm.default_detector()
    cm.apply_diff_params(cr, det, theta0)
    e2 = cm.Engine(cm.Lattice(cr), det)
    y, tape, info = eng.diff_forward(wl, 1000, theta0, seed=5, t_max_s=T_MAX)
    r = e2.run_workload(wl, 1000, seed=5, form=cm.FORM_BATCHED, keep_hits=False)
    a, b = info["stats"], r["stats"]
    for k in ("total_steps", "isotope_scatters", "decays_L_to_TT", "decays_L_to_LT", "surface_refls", "sensor_hits",
              "killed_tracks"):
        assert a[k] == b[k], k
    assert a["total_deposited_eV"] == b["total_deposited_eV"]


def test_autograd_function(diff):
    torch = pytest.importorskip("torch")
    from cmpmc_amd.autograd import transport_observables
    cm, eng, wl, theta0 = diff
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    y = transport_observables(theta, eng, wl, 500, seed=3, t_max_s=T_MAX)
    g = torch.tensor(_g_time())
    (y * g).sum().backward()
    y_ref, tape, _ = eng.diff_forward(wl, 500, theta0, seed=3, t_max_s=T_MAX)
    assert np.array_equal(y.detach().numpy(), y_ref)
    assert np.allclose(theta.grad.numpy(), tape.backward(_g_time()), rtol=1e-9)


def test_gradient_matches_crn_finite_differences(diff):
    cm, eng, wl, theta0 = diff
    g = _g_time()
    R, N, H = 10, 20000, 0.01
    grads, fds = [], []
    for rep in range(R):
        seed = 1000 + rep
        print(f"case: gradient vs CRN FD replicate {rep} n={N} seed={seed}", flush=True)
        y, tape, _ = eng.diff_forward(wl, N, theta0, seed=seed, t_max_s=T_MAX)
        grads.append(tape.backward(g))
        fd = np.zeros(4)
        for j in range(4):
            tp, tm = theta0.copy(), theta0.copy()
            tp[j] *= 1 + H
            tm[j] *= 1 - H
            yp, _, _ = eng.diff_forward(wl, N, tp, seed=seed, t_max_s=T_MAX)
            ym, _, _ = eng.diff_forward(wl, N, tm, seed=seed, t_max_s=T_MAX)
            fd[j] = g @ (yp - ym) / (2 * H * theta0[j])
        fds.append(fd)
    grads, fds = np.array(grads), np.array(fds)
    mg, mf = grads.mean(0), fds.mean(0)
    sg, sf = grads.std(0, ddof=1) / np.sqrt(R), fds.std(0, ddof=1) / np.sqrt(R)
    z = np.abs(mg - mf) / np.hypot(sg, sf)
    rel = theta0 * mg / (g @ y)
    print("relative sensitivities theta d<t>/dtheta / <t>:", rel, flush=True)
    print("score:", mg, "+-", sg, flush=True)
    print("CRN FD:", mf, "+-", sf, flush=True)
    print("z:", z, flush=True)
    assert np.all(z < 4.5), f"score vs FD disagree: z = {z}"
    for j in (0, 3):
        assert abs(mg[j]) > 5 * sg[j], f"parameter {j}: gradient not resolved ({mg[j]} +- {sg[j]})"
