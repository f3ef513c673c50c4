"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

1. Golden vectors written by the reference's own implementation (tests/golden/
   reference_vectors.npz, see make_golden.py): the oracle must reproduce them bit for
   bit (RNG streams, samplers, v_g lookups, decay tables, Kaplan cascades, ray exits and
   whole run_events results, including hits).
2. Where the reference build itself is present (oracle/_ref), larger bit-exact runs.
3. What the mathematics fixes independently of both: cubic Christoffel closed forms,
   momentum closure of the decay angles, cascade energy closure, the 2D/Omega
   pair-breaking limit, sampler moments against the Kaplan densities, ledger closure.
"""
import ctypes as C
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_rng_streams_golden(oracle, gold):
    O = oracle
    L = O.lib()
    for i, s in enumerate(gold["rng_seeds"]):
        for j, t in enumerate(gold["rng_streams"]):
            r = O.Rng()
            L.orc_rng_for_stream(C.byref(r), int(s), int(t))
            got = [L.orc_next_u64(C.byref(r)) for _ in range(16)]
            assert got == [int(x) for x in gold["rng_values"][i, j]], (s, t)


def test_uniform_open_interval(oracle):
    O = oracle
    L = O.lib()
    r = O.Rng()
    L.orc_rng_seed(C.byref(r), 5)
    u = np.array([L.orc_uniform(C.byref(r)) for _ in range(20000)])
    assert (u > 0).all() and (u < 1).all()
    assert abs(u.mean() - 0.5) < 0.01


def test_transport_stream_xoshiro128(oracle):
    """The batched form's transport streams: xoshiro128** (Blackman & Vigna). From the
    state (1, 2, 3, 4) its first outputs are rotl(2 * 5, 7) * 9 = 11520, then 0 (s1 is
    2 ^ (3 ^ 1) = 0 after one step), then rotl(1029 * 5, 7) * 9 = 5927040, worked by hand
    from the published update; uniforms are ((u >> 8) | 1) * 2^-24 in (0, 1)."""
    O = oracle
    L = O.lib()
    r = O.Rng()
    r.s[0] = 1 | (2 << 32)
    r.s[1] = 3 | (4 << 32)
    r.w32 = 1
    assert [L.orc_next_u32(C.byref(r)) for _ in range(3)] == [11520, 0, 5927040]
    assert L.orc_next_bits(C.byref(r)) < 2 ** 32
    # a track stream: the first two splitmix64 outputs of its seed, as four 32-bit words
    a, b = O.Rng(), O.Rng()
    L.orc_rng_for_track(C.byref(a), 0x1234, 77)
    L.orc_rng32_for_track(C.byref(b), 0x1234, 77)
    assert b.w32 == 1 and (b.s[0], b.s[1]) == (a.s[0], a.s[1])
    u = np.array([L.orc_uniform(C.byref(b)) for _ in range(20000)])
    assert (u > 0).all() and (u < 1).all() and abs(u.mean() - 0.5) < 0.01
    assert np.all(np.round(u * 2 ** 24) == u * 2 ** 24)       # 24-bit uniforms


def test_samplers_golden(oracle, gold):
    O = oracle
    L = O.lib()
    r = O.Rng()
    L.orc_rng_seed(C.byref(r), 2024)
    out = np.zeros(7 * 256)
    d = np.zeros(3)
    for i in range(256):
        out[4 * i] = L.orc_normal(C.byref(r))
        L.orc_isotropic_dir(C.byref(r), O.dptr(d))
        out[4 * i + 1:4 * i + 4] = d
    n = np.array([0.0, 0.0, -1.0])
    for i in range(256):
        L.orc_lambertian_dir(C.byref(r), O.dptr(n), O.dptr(d))
        out[1024 + 3 * i:1024 + 3 * i + 3] = d
    assert np.array_equal(out, gold["sampler_values"])


def test_map_k_to_vg_golden(oracle, gold):
    eng = oracle.Engine()
    for m in range(3):
        assert np.array_equal(eng.map_k_to_vg(m, gold["map_k"]), gold["map_vg"][m])


def test_map_k_to_vg_exact_golden(oracle, gold):
    """KinematicsMode::Exact (crystal_lattice.hpp:77-85): Cardano roots with the Jacobi
    fallback near FT/ST degeneracies, bit-equal to the reference on axes, near-degenerate
    and random directions."""
    eng = oracle.Engine()
    for m in range(3):
        assert np.array_equal(eng.map_k_to_vg(m, gold["map_exact_k"], exact=True), gold["map_exact_vg"][m])


def test_map_k_to_vg_bilinear_golden(oracle, gold):
    """KinematicsMode::Bilinear (crystal_lattice.hpp:89-114, table 350-368): bit-equal to the
    reference on poles, axes, near-degenerate and random directions."""
    eng = oracle.Engine()
    for m in range(3):
        assert np.array_equal(eng.map_k_to_vg(m, gold["map_exact_k"], bilinear=True), gold["map_bil_vg"][m])


def test_bilinear_nodes_are_christoffel_solutions(oracle):
    """At a node of the (cos theta, phi) grid the bilinear weights select that node alone, so
    the lookup is the normalised Jacobi solve of the node's direction: it agrees with the
    closed-form (Cardano) solve there. A transposed or shifted table index fails this."""
    eng = oracle.Engine()
    g = np.random.default_rng(3)
    i = g.integers(1, 180, 40)
    j = g.integers(0, 360, 40)
    ct = -1.0 + 2.0 * i / 180.0
    phi = 2.0 * np.pi * j / 360.0
    st = np.sqrt(1.0 - ct * ct)
    k = np.c_[st * np.cos(phi), st * np.sin(phi), ct]
    for m in range(3):
        b = eng.map_k_to_vg(m, k, bilinear=True)
        e = eng.map_k_to_vg(m, k, exact=True)
        # L everywhere; the transverse modes away from their degeneracies (where the
        # mode labels of two solvers may differ)
        ok = np.ones(len(k), bool) if m == 0 else np.abs(e[:, 3] - eng.map_k_to_vg(3 - m, k, exact=True)[:, 3]) > 20.0
        assert ok.sum() >= 20
        assert np.allclose(b[ok, 3], e[ok, 3], rtol=1e-7, atol=0)
        assert np.allclose(b[ok, :3], e[ok, :3], atol=1e-6)


def _sym_nodes(n=33):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    s, t = i.ravel() / (n - 1), j.ravel() / (n - 1)
    k = np.c_[np.ones_like(s), s, s * t]
    return k / np.linalg.norm(k, axis=1)[:, None]


def test_symmetric_nodes_are_christoffel_solutions(oracle):
    """The cubic-symmetry table (kinematics mode 3, this build's; no reference counterpart)
    holds the Jacobi solve at k = (1, s, s t) / |.|: at every node the lookup agrees with the
    closed-form (Cardano) solve to FP32 rounding. A transposed (i, j) index, a wrong node
    spacing or a wrong component order fails this."""
    eng = oracle.Engine()
    k = _sym_nodes()
    for m in range(3):
        s = eng.map_k_to_vg(m, k, symmetric=True)
        e = eng.map_k_to_vg(m, k, exact=True)
        # L everywhere; the transverse modes away from their degeneracies
        ok = np.ones(len(k), bool) if m == 0 else np.abs(e[:, 3] - eng.map_k_to_vg(3 - m, k, exact=True)[:, 3]) > 20.0
        assert ok.sum() >= 500
        assert np.allclose(s[ok, 3], e[ok, 3], rtol=1e-6, atol=0)
        assert np.allclose(s[ok, :3], e[ok, :3], atol=2e-6)


def test_symmetric_lookup_is_cubic_invariant(oracle):
    """v_g(R k) = R v_g(k) for the 48 signed axis permutations R of the cubic group: the
    lookup of a permuted, sign-flipped direction is the permuted, sign-flipped lookup (up to
    the summation order of |v|). A wrong sign or axis in the mapping back fails this."""
    eng = oracle.Engine()
    g = np.random.default_rng(29)
    k = g.normal(size=(400, 3))
    k /= np.linalg.norm(k, axis=1)[:, None]
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    for m in range(3):
        base = eng.map_k_to_vg(m, k, symmetric=True)
        for p in perms:
            sg = g.choice([-1.0, 1.0], size=3)
            r = eng.map_k_to_vg(m, k[:, p] * sg, symmetric=True)
            assert np.allclose(r[:, :3], base[:, :3][:, p] * sg, rtol=0, atol=1e-15)
            assert np.allclose(r[:, 3], base[:, 3], rtol=1e-15, atol=0)


def test_symmetric_accuracy_against_exact(oracle):
    """On 3,000 random directions the 39 KiB cubic-symmetry table is as close to the exact
    solve as the 3 MiB octahedral table: direction error q99 <= 1.2e-3 rad, |v_g| relative
    error q99 <= 5e-4 in every mode (measured 7.0e-4 / 6.6e-4 / 1.2e-3 rad and 3.8e-4 /
    2.2e-4 / 4.2e-4 for L / ST / FT; the octahedral table's are 3.1e-4 / 1.9e-3 / 2.0e-3 rad)."""
    eng = oracle.Engine()
    g = np.random.default_rng(5)
    k = g.normal(size=(3000, 3))
    k /= np.linalg.norm(k, axis=1)[:, None]
    for m in range(3):
        s = eng.map_k_to_vg(m, k, symmetric=True)
        e = eng.map_k_to_vg(m, k, exact=True)
        ang = np.arctan2(np.linalg.norm(np.cross(s[:, :3], e[:, :3]), axis=1), np.sum(s[:, :3] * e[:, :3], axis=1))
        rel = np.abs(s[:, 3] / e[:, 3] - 1.0)
        assert np.quantile(ang, 0.99) <= 1.25e-3 and np.quantile(rel, 0.99) <= 5e-4, (m, np.quantile(ang, 0.99))


def test_iv_matrix_rate_golden(oracle, gold):
    """IVRateModel::Matrix (charge_kernels.hpp:36-79): bit-equal to the reference's rate of
    Germanium's defaults on both sides of each optical threshold; below the lowest threshold
    only the neutral-impurity term remains, which rises as sqrt(E) at small E."""
    O = oracle
    c = O.germanium()
    E = gold["iv_matrix_E"]
    r = np.array([O.lib().orc_intervalley_matrix_rate(C.byref(c), float(e)) for e in E])
    assert np.array_equal(r, gold["iv_matrix_rate"])
    small = E < 1e-4
    assert np.allclose(r[small] / np.sqrt(E[small]), r[small][0] / np.sqrt(E[small][0]), rtol=0.02)
    # the optical terms switch on above their thresholds: the rate jumps no more than the
    # impurity slope across them (e_func(0) = 0), and grows beyond
    assert np.all(np.diff(r[E < 0.2][np.argsort(E[E < 0.2])]) > -1e-6 * r.max())


def test_rate_competition_golden_and_pair_breaking_limit(oracle, gold):
    """FilmLossModel::KaplanRateCompetition (kaplan_film.hpp:69-99): the pair-breaking rate
    and loss probability are bit-equal to the reference's; 1/tau_pb(2 Delta) = 1/tau0
    exactly (the integral's limit); the 48-point quadrature agrees with an adaptive
    integration of the original singular integral; the loss probability is
    1 / (1 + tau_l / tau_pb) and 1 below the gap edge."""
    from scipy.integrate import quad
    O = oracle
    d = O.default_detector()
    f = d.surf[0].film
    f.loss_model, f.phonon_loss_time = 1, O.RC_LOSS_TIME
    L = O.lib()
    E = gold["rc_E"]
    rate = np.array([L.orc_kaplan_pair_breaking_rate(C.byref(f), float(e)) for e in E])
    loss = np.array([L.orc_film_loss_probability(C.byref(f), float(e), 1.0) for e in E])
    assert np.array_equal(rate, gold["rc_rate"]) and np.array_equal(loss, gold["rc_loss"])
    D, tau0 = f.gap_energy, f.phonon_lifetime
    assert abs(L.orc_kaplan_pair_breaking_rate(C.byref(f), 2.0 * D) * tau0 - 1.0) < 1e-12
    for W in (3.0, 10.0, 50.0):
        om = W * D
        num, _ = quad(lambda x: (x * (om - x) + D * D) / np.sqrt((x * x - D * D) * ((om - x) ** 2 - D * D)),
                      D, om - D, limit=400, epsabs=0, epsrel=1e-11, points=[0.5 * om])
        assert abs(L.orc_kaplan_pair_breaking_rate(C.byref(f), om) / (num / (np.pi * D * tau0)) - 1.0) < 1e-7
    above = E >= 2.0 * D
    assert np.allclose(loss[above], 1.0 / (1.0 + rate[above] * O.RC_LOSS_TIME), rtol=1e-15, atol=0)
    assert np.all(loss[~above] == 1.0)


def test_cardano_equals_jacobi(oracle):
    """tests/test_christoffel_cardano.cpp:52-53: the closed-form solve agrees with the
    Jacobi sweeps (group velocities and phase speeds) to 1e-9 relative."""
    O = oracle
    L = O.lib()
    c = O.germanium()
    vm = np.zeros((6, 6))                       # cubic Voigt matrix
    vm[:3, :3] = c.C12
    vm[[0, 1, 2], [0, 1, 2]] = c.C11
    vm[[3, 4, 5], [3, 4, 5]] = c.C44
    voigt = [[0, 5, 4], [5, 1, 3], [4, 3, 2]]
    Ct = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for l in range(3):
                for m in range(3):
                    Ct[i, j, l, m] = vm[voigt[i][j], voigt[l][m]]
    Ct = np.ascontiguousarray(Ct)
    Cp = Ct.ctypes.data
    rng = np.random.default_rng(5)
    for _ in range(500):
        k = rng.normal(size=3)
        k /= np.linalg.norm(k)
        a_vg, a_vp, b_vg, b_vp = np.zeros(9), np.zeros(3), np.zeros(9), np.zeros(3)
        L.orc_solve_christoffel_cardano(Cp, c.density, O.dptr(k), O.dptr(a_vg), O.dptr(a_vp))
        L.orc_solve_christoffel(Cp, c.density, O.dptr(k), O.dptr(b_vg), O.dptr(b_vp))
        assert np.allclose(a_vp, b_vp, rtol=1e-9, atol=0)
        assert np.allclose(a_vg, b_vg, rtol=1e-7, atol=1e-6 * np.abs(b_vg).max())


def test_decay_fractions_golden(oracle, gold):
    O = oracle
    L = O.lib()
    eng = O.Engine()
    r = O.Rng()
    L.orc_rng_seed(C.byref(r), 99)
    lt, tt = [], []
    for _ in range(1000):
        lt.append(L.orc_sample_lt_fraction(eng.lattice_ptr(), C.byref(r)))
        tt.append(L.orc_sample_tt_fraction(eng.lattice_ptr(), C.byref(r)))
    assert np.array_equal(lt, gold["decay_lt"]) and np.array_equal(tt, gold["decay_tt"])


@pytest.mark.parametrize("tag,E,seed", [("k75", 7.5e-3, 11), ("k15", 15e-3, 12), ("k31", 31e-3, 13)])
def test_kaplan_cascade_golden(oracle, gold, tag, E, seed):
    O = oracle
    L = O.lib()
    f = O.Film()
    L.orc_film_aluminum(C.byref(f))
    r = O.Rng()
    L.orc_rng_seed(C.byref(r), seed)
    refl = np.zeros(256)
    nr = C.c_int()
    for i in range(200):
        edep = L.orc_absorb_phonon(C.byref(f), E, O.dptr(refl), C.byref(nr), C.byref(r))
        assert edep == gold[f"{tag}_edep"][i]
        assert nr.value == gold[f"{tag}_nrefl"][i]
        s = 0.0
        for x in refl[:nr.value]:
            s += x
        assert s == gold[f"{tag}_reflsum"][i]
    assert [r.s[i] for i in range(4)] == [int(x) for x in gold[f"{tag}_state"]]


class _Surrogate:
    """The oracle's KaplanSurrogateLUT (test helper)."""

    def __init__(self, O, f, seed=1337):
        self.O, self.L = O, O.lib()
        self.buf = C.create_string_buffer(256)           # room for orc_surrogate
        if self.L.orc_surrogate_build(self.buf, C.byref(f), seed) != 0:
            raise MemoryError("orc_surrogate_build")

    def absorb(self, E, r, refl, nr):
        return self.L.orc_surrogate_absorb(self.buf, E, self.O.dptr(refl), C.byref(nr), C.byref(r))

    def close(self):
        self.L.orc_surrogate_free(self.buf)


def test_kaplan_surrogate_golden_and_reentry(oracle, gold):
    """KaplanSurrogateLUT (kaplan_film.hpp:325-383): the table built from Rng(1337) and its
    lookup are bit-equal to the reference's (3,000 absorb_phonon_fast calls over [1.5, 45]
    Delta from one stream: each edep, reflected count and sum, and the final state; energies
    outside [2, 40] Delta take the exact cascade). And tests/test_kaplan.cpp:95-107's own
    check: in a re-entry cavity (each re-emitted phonon returns with probability p_ret) the
    surrogate's collected fraction is the full cascade's to 3 %."""
    O = oracle
    L = O.lib()
    f = O.Film()
    L.orc_film_aluminum(C.byref(f))
    s = _Surrogate(O, f)
    try:
        r = O.Rng()
        L.orc_rng_seed(C.byref(r), 21)
        refl = np.zeros(256)
        nr = C.c_int()
        for i, E in enumerate(gold["surr_E"]):
            edep = s.absorb(float(E), r, refl, nr)
            assert edep == gold["surr_edep"][i] and nr.value == gold["surr_nrefl"][i], i
            sm = 0.0
            for x in refl[:nr.value]:
                sm += x
            assert sm == gold["surr_reflsum"][i], i
        assert [r.s[i] for i in range(4)] == [int(x) for x in gold["surr_state"]]
        D = f.gap_energy
        rng = O.Rng()
        L.orc_rng_seed(C.byref(rng), 31)

        def cavity(E, p_ret, lut):
            todo, edep = [E], 0.0
            while todo:
                e = todo.pop()
                edep += s.absorb(e, rng, refl, nr) if lut else L.orc_absorb_phonon(C.byref(f), e, O.dptr(refl),
                                                                                       C.byref(nr), C.byref(rng))
                for x in refl[:nr.value].copy():
                    if L.orc_uniform(C.byref(rng)) < p_ret:
                        todo.append(x)
            return edep / E

        for p_ret in (0.0, 0.9):
            for W in (3.0, 15.0, 35.0):
                a = np.mean([cavity(W * D, p_ret, False) for _ in range(4000)])
                b = np.mean([cavity(W * D, p_ret, True) for _ in range(4000)])
                assert b == pytest.approx(a, rel=0.03), (p_ret, W, a, b)
    finally:
        s.close()


def test_intersect_golden(oracle, gold):
    O = oracle
    L = O.lib()
    det = O.default_detector()
    n_in = np.zeros(3)
    sid = C.c_int()
    for i in range(len(gold["isect_p"])):
        p = np.ascontiguousarray(gold["isect_p"][i])
        v = np.ascontiguousarray(gold["isect_v"][i])
        t = L.orc_intersect_exit(C.byref(det), O.dptr(p), O.dptr(v), O.dptr(n_in), C.byref(sid))
        assert np.array_equal([t, *n_in, sid.value], gold["isect_out"][i])


def _gold_runs(O):
    det = O.default_detector()
    O.set_bias(det, 2.0)
    det.luke_sample_fraction = 0.02
    d3 = O.default_detector()
    O.set_bias(d3, 4.0)
    d3.luke_sample_fraction = 0.0
    dd = O.default_detector()
    O.demo_polynomial_surfaces(dd)
    dr = O.default_detector()
    dr.russian_roulette_bounces, dr.russian_roulette_survival = 32, 0.5
    ds = O.default_detector()
    O.constant_split_surfaces(ds, 0.5)
    dmx = O.default_detector()
    O.set_bias(dmx, 4.0)
    dmx.luke_sample_fraction = 0.0
    dt = O.default_detector()
    O.set_bias(dt, 4.0)
    dt.luke_sample_fraction = 0.0
    dl = O.default_detector()
    O.set_bias(dl, 4.0)
    dl.luke_sample_fraction = 0.0
    dl.charge_stepping = 1
    dl2 = O.default_detector()
    O.set_bias(dl2, 4.0)
    dl2.luke_sample_fraction = 0.02
    dl2.charge_stepping = 1
    dp = O.default_detector()
    O.set_bias(dp, 4.0)
    dp.luke_sample_fraction = 0.05
    return O.mesh_runs() | {
        "deposit_nr": (O.workload(O.WL_DEPOSIT, mode=1, energy_eV=100.0,
                                  partition=dict(max_sampled_phonons=12, max_sampled_pairs=8)), 4, 42, dp),
        "deposit_er": (O.workload(O.WL_DEPOSIT, mode=0, energy_eV=30.0, origin=(1e-3, -2e-3, 4e-3), pair_at_origin=1,
                                  partition=dict(max_sampled_phonons=10, max_sampled_pairs=4)), 3, 7, dp),
        "c1": (O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 400, 42, O.default_detector()),
        "lognormal": (O.workload(O.WL_PHONON_LOGN, logn_median=4.0), 40, 42, O.default_detector()),
        "mixed": (O.workload(O.WL_MIXED), 41, 77, det),
        "pairs": (O.workload(O.WL_PAIR), 6, 42, d3),
        "bench_demo": (O.workload(O.WL_PHONON_POINT, mode=2, energy_eV=7.5e-3), 100, 12345, dd),
        "bench_dos": (O.workload(O.WL_PHONON_POINT, mode=O.MODE_BENCH_DOS, energy_eV=7.5e-3), 300, 12345, dd),
        "roulette": (O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 200, 42, dr),
        "specular": (O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 200, 42, ds),
        "trap_ion": (O.workload(O.WL_PAIR), 6, 42, dt, O.trap_ion_crystal()),
        "leman": (O.workload(O.WL_PAIR), 6, 42, dl),
        "leman_ion": (O.workload(O.WL_PAIR), 6, 43, dl2, O.trap_ion_crystal()),
        "c1_exact": (O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 200, 42, O.default_detector()),
        "c1_bil": (O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 200, 42, O.default_detector()),
        "film_rc": (O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 200, 42, O.rate_competition_detector()),
        "iv_matrix": (O.workload(O.WL_PAIR), 6, 42, dmx, O.matrix_iv_crystal()),
        "release_debye": (O.workload(O.WL_PAIR), 6, 42, _release_detector(O, 1)),
        "release_film": (O.workload(O.WL_PAIR), 6, 42, _release_detector(O, 2)),
    }


def _release_detector(O, mode):
    d = O.default_detector()
    O.set_bias(d, 4.0)
    d.luke_sample_fraction = 0.0
    d.charge_release = mode
    return d


@pytest.mark.parametrize("tag", ["c1", "lognormal", "mixed", "pairs", "bench_demo", "bench_dos", "roulette",
                                 "specular", "trap_ion", "leman", "leman_ion", "c1_exact", "deposit_nr",
                                 "deposit_er", "mesh3_pairs", "mesh2_pairs", "mesh3_leman_ion", "c1_bil",
                                 "film_rc", "iv_matrix", "release_debye", "release_film"])
def test_run_events_golden(oracle, gold, tag):
    """Whole run_events results (SimStats bytes, hit_offsets, hits) equal the reference's
    (c1_exact: with KinematicsMode::Exact; c1_bil: KinematicsMode::Bilinear; film_rc: both
    films on FilmLossModel::KaplanRateCompetition; iv_matrix: the IVRateModel::Matrix
    intervalley rate for pairs at 4 V; release_debye, release_film: ChargeRelease BulkDebye and
    ElectrodeFilm at carrier collection; bench_dos: the reference benchmark's own phonon
    workload; roulette, specular, trap_ion: Russian roulette, half-specular surfaces, trap
    impact ionisation; leman, leman_ion: Leman second-order charge stepping; deposit_nr,
    deposit_er: generators that partition a nuclear / electron recoil with
    EnergyPartition::partition_deposit, transported at 4 V; mesh3_pairs, mesh2_pairs,
    mesh3_leman_ion: pairs drifting in FEM mesh fields, a tetrahedral grid and an
    axisymmetric triangle grid)."""
    O = oracle
    spec = _gold_runs(O)[tag]
    wl, n, seed, det = spec[:4]
    eng = O.Engine(crystal=spec[4] if len(spec) > 4 else None, detector=det)
    if tag.endswith("_exact"):
        eng.set_kinematics(1)
    elif tag.endswith("_bil"):
        eng.set_kinematics(2)
    r = eng.run(wl, n, seed, threads=4)
    assert r["stats_raw"] == gold[f"run_{tag}_stats"].tobytes()
    assert np.array_equal(r["hit_offsets"], gold[f"run_{tag}_offsets"])
    h = r["hits"][:600]
    got = np.c_[h["pos"], h["time"], h["edep"], h["incident"], h["surface"], h["mode"]]
    assert np.array_equal(got, gold[f"run_{tag}_hits"])


def test_thread_count_independence(oracle):
    """tests/test_parallel.cpp:70-77: bit-identical results for 1..4 threads."""
    O = oracle
    det = O.default_detector()
    O.set_bias(det, 2.0)
    det.luke_sample_fraction = 0.02
    eng = O.Engine(detector=det)
    wl = O.workload(O.WL_MIXED)
    base = eng.run(wl, 37, 77, threads=1)
    for t in (2, 3, 4):
        r = eng.run(wl, 37, 77, threads=t)
        assert r["stats_raw"] == base["stats_raw"]
        assert np.array_equal(r["hits"], base["hits"])
    assert base["hit_offsets"][-1] == len(base["hits"])


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                    "oracle", "_ref", "libcmpmc_ref.so")),
                    reason="the in-place reference build (oracle/_ref) is not present")
def test_against_reference_build(oracle):
    O = oracle
    eng = O.Engine()
    wl = O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3)
    a = O.ref_run(wl, 1500, 42, threads=8)
    b = eng.run(wl, 1500, 42, threads=8)
    assert a["stats_raw"] == b["stats_raw"]
    assert np.array_equal(a["hit_offsets"], b["hit_offsets"])
    for f in ("pos", "time", "edep", "incident", "surface", "mode"):
        assert np.array_equal(a["hits"][f], b["hits"][f])
    d4 = O.default_detector()
    O.set_bias(d4, 4.0)
    a = O.ref_run(O.workload(O.WL_PAIR), 1, 42, detector=d4, threads=1, chunk=1)
    b = O.Engine(detector=d4).run(O.workload(O.WL_PAIR), 1, 42, threads=1, chunk=1)
    assert a["stats_raw"] == b["stats_raw"]


# ------------------------------------------------------------------ closed forms
def test_christoffel_closed_forms(oracle):
    """tests/test_christoffel.cpp:29-46: cubic phase speeds along [100], [110], [111]."""
    O = oracle
    L = O.lib()
    c = O.germanium()
    eng = O.Engine(crystal=c)
    Cijkl = eng.lattice_ptr() + C.sizeof(O.Crystal)   # orc_lattice.C follows the crystal config
    r, c11, c12, c44 = c.density, c.C11, c.C12, c.C44
    vg = np.zeros((3, 3))
    vp = np.zeros(3)

    def solve(k):
        k = np.array(k, dtype=float)
        k /= np.linalg.norm(k)
        L.orc_solve_christoffel(Cijkl, r, O.dptr(k), O.dptr(vg), O.dptr(vp))
        return vp.copy(), vg.copy()

    p, _ = solve([1, 0, 0])
    assert p[0] == pytest.approx(np.sqrt(c11 / r), rel=1e-10)
    assert p[1] == pytest.approx(np.sqrt(c44 / r), rel=1e-10)
    p, _ = solve([1, 1, 0])
    assert p[0] == pytest.approx(np.sqrt((c11 + c12 + 2 * c44) / (2 * r)), rel=1e-10)
    a, b = np.sqrt(c44 / r), np.sqrt((c11 - c12) / (2 * r))
    assert p[1] == pytest.approx(max(a, b), rel=1e-10) and p[2] == pytest.approx(min(a, b), rel=1e-10)
    p, g = solve([1, 1, 1])
    assert p[0] == pytest.approx(np.sqrt((c11 + 2 * c12 + 4 * c44) / (3 * r)), rel=1e-10)
    assert p[1] == pytest.approx(np.sqrt((c11 - c12 + c44) / (3 * r)), rel=1e-10)
    # along [111] the L group velocity is parallel to k with |v_g| = v_p
    assert np.linalg.norm(g[0]) == pytest.approx(p[0], rel=1e-10)


def test_group_velocity_is_gradient(oracle):
    """tests/test_christoffel.cpp:52-72: v_g = grad_k omega by central differences."""
    O = oracle
    L = O.lib()
    c = O.germanium()
    eng = O.Engine(crystal=c)
    Cijkl = eng.lattice_ptr() + C.sizeof(O.Crystal)   # orc_lattice.C follows the crystal config
    rng = np.random.default_rng(4)
    vg = np.zeros((3, 3))
    vp = np.zeros(3)
    worst = 0.0
    for _ in range(200):
        k = rng.normal(size=3)
        k /= np.linalg.norm(k)
        L.orc_solve_christoffel(Cijkl, c.density, O.dptr(k), O.dptr(vg), O.dptr(vp))
        g0, p0 = vg.copy(), vp.copy()
        if abs(p0[1] - p0[2]) < 1e-3 * p0[1]:
            continue
        for m in range(3):
            grad = np.zeros(3)
            for d in range(3):
                kp, km = k.copy(), k.copy()
                kp[d] += 1e-6
                km[d] -= 1e-6
                kpu, kmu = kp / np.linalg.norm(kp), km / np.linalg.norm(km)
                L.orc_solve_christoffel(Cijkl, c.density, O.dptr(kpu), O.dptr(vg), O.dptr(vp))
                a = np.linalg.norm(kp) * vp[m]
                L.orc_solve_christoffel(Cijkl, c.density, O.dptr(kmu), O.dptr(vg), O.dptr(vp))
                b = np.linalg.norm(km) * vp[m]
                grad[d] = (a - b) / 2e-6
            worst = max(worst, np.linalg.norm(grad - g0[m]) / np.linalg.norm(g0[m]))
    assert worst < 1e-5


def test_decay_angles_momentum_closure(oracle):
    """tests/test_anharmonic_decay.cpp:44-70."""
    O = oracle
    L = O.lib()
    c = O.germanium()
    vL, vT = c.v_sound, c.v_trans
    d = vL / vT
    worst_lt = worst_tt = 0.0
    for i in range(1, 400):
        xl = (d - 1) / (d + 1) + ((1.0 - (d - 1) / (d + 1)) * i) / 400.0
        tL, tT = L.orc_theta_L_to_LT(C.byref(c), xl), L.orc_theta_T_from_LT(C.byref(c), xl)
        kL, kT, k0 = xl / vL, (1 - xl) / vT, 1 / vL
        worst_lt = max(worst_lt, np.hypot(kL * np.cos(tL) + kT * np.cos(tT) - k0, kL * np.sin(tL) - kT * np.sin(tT)) / k0)
        xt = 0.5 * (1 - 1 / d) + (1 / d) * i / 400.0
        t1, t2 = L.orc_theta_TT(C.byref(c), xt), L.orc_theta_TT(C.byref(c), 1 - xt)
        k1, k2 = xt / vT, (1 - xt) / vT
        worst_tt = max(worst_tt, np.hypot(k1 * np.cos(t1) + k2 * np.cos(t2) - k0, k1 * np.sin(t1) - k2 * np.sin(t2)) / k0)
    assert worst_lt < 1e-9 and worst_tt < 1e-9
    assert np.cos(L.orc_theta_TT(C.byref(c), 0.5)) == pytest.approx(1.0 / d, rel=1e-9)


def test_decay_fraction_moments(oracle):
    """tests/test_anharmonic_decay.cpp:72-89: sampled <x> vs the Leman (9)/(12) densities."""
    O = oracle
    L = O.lib()
    c = O.germanium()
    eng = O.Engine(crystal=c)
    r = O.Rng()
    L.orc_rng_seed(C.byref(r), 12345)
    lt = np.array([L.orc_sample_lt_fraction(eng.lattice_ptr(), C.byref(r)) for _ in range(200000)])
    tt = np.array([L.orc_sample_tt_fraction(eng.lattice_ptr(), C.byref(r)) for _ in range(200000)])
    d = c.v_sound / c.v_trans

    def moments(pdf, lo, hi):
        x = lo + (np.arange(200000) + 0.5) * (hi - lo) / 200000
        w = np.maximum(0.0, np.array([pdf(C.byref(c), xi) for xi in x[::50]]))
        xs = x[::50]
        return (w * xs).sum() / w.sum()

    mlt = moments(L.orc_tamura_lt_pdf, (d - 1) / (d + 1), 1.0)
    mtt = moments(L.orc_tamura_tt_pdf, 0.5 * (1 - 1 / d), 0.5 * (1 + 1 / d))
    assert lt.mean() == pytest.approx(mlt, rel=3e-3)
    assert tt.mean() == pytest.approx(mtt, rel=3e-3)


def test_kaplan_energy_closure_and_limit(oracle):
    """tests/test_pair_breaking.cpp:106-119 (closure) and tests/test_kaplan.cpp:60-70
    (eta = 2 Delta / Omega below 4 Delta in the no-loss limit)."""
    O = oracle
    L = O.lib()
    f = O.Film()
    L.orc_film_aluminum(C.byref(f))
    f.subgap_absorption = 0.0
    f.temperature = 0.0
    D = f.gap_energy
    r = O.Rng()
    L.orc_rng_seed(C.byref(r), 20261101)
    refl = np.zeros(256)
    nr = C.c_int()
    for W in (50.0, 150.0, 350.0):
        worst = 0.0
        for _ in range(300):
            e = L.orc_absorb_phonon(C.byref(f), W * D, O.dptr(refl), C.byref(nr), C.byref(r))
            e += refl[:nr.value].sum()
            worst = max(worst, abs(e / (W * D) - 1.0))
        assert worst < 1e-12
    thick = O.Film()
    L.orc_film_aluminum(C.byref(thick))
    thick.film_thickness, thick.subgap_absorption, thick.keep_all_phonons, thick.temperature = 1.0, 0.0, 0, 0.0
    for W in (2.5, 3.0, 3.9):
        s = sum(L.orc_absorb_phonon(C.byref(thick), W * D, O.dptr(refl), C.byref(nr), C.byref(r)) for _ in range(20000))
        assert s / 20000 / (W * D) == pytest.approx(2.0 / W, rel=5e-3)


def test_kaplan_sampler_moments(oracle):
    """tests/test_kaplan.cpp:50-58: pair-breaking and QP-emission samplers vs the Kaplan
    densities, integrated here with a sin^2 substitution."""
    O = oracle
    L = O.lib()
    f = O.Film()
    L.orc_film_aluminum(C.byref(f))
    D = f.gap_energy
    r = O.Rng()
    L.orc_rng_seed(C.byref(r), 31)

    def mean_pdf(p, lo, hi):
        u = (np.arange(200000) + 0.5) / 200000
        sn = np.sin(0.5 * np.pi * u)
        x = lo + (hi - lo) * sn * sn
        jac = (hi - lo) * np.pi * sn * np.cos(0.5 * np.pi * u)
        w = p(x) * jac
        return (w * x).sum() / w.sum()

    for W in (2.5, 5.0, 20.0):
        Om = W * D
        s = np.mean([L.orc_sample_qp_energy(C.byref(f), Om, C.byref(r)) for _ in range(100000)])
        ref = mean_pdf(lambda x: (x * (Om - x) + D * D) / np.sqrt((x * x - D * D) * ((Om - x) ** 2 - D * D)), D, Om - D)
        assert s / D == pytest.approx(ref / D, rel=4e-3)
        s2 = np.mean([L.orc_sample_phonon_energy(C.byref(f), Om, C.byref(r)) for _ in range(100000)])
        ref2 = mean_pdf(lambda x: (Om - x) ** 2 * (x - D * D / Om) / np.sqrt(x * x - D * D), D, Om)
        assert s2 / D == pytest.approx((Om - ref2) / D, rel=4e-3)


def test_ledger_closure(oracle):
    """tests/test_engine_energy_closure.cpp:19-21 / test_parallel.cpp:110-120."""
    O = oracle
    r = O.Engine().run(O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=4e-3), 500, 5, threads=4)
    s = r["stats"]
    assert abs(s["total_deposited_eV"] + s["escaped_or_killed_eV"] - s["total_injected_eV"]) <= 1e-9 * s["total_injected_eV"]


def test_batched_scheme_is_order_independent(oracle):
    """The batched form's per-track streams make every event a function of (seed, event)
    only: results do not depend on chunking or thread count."""
    O = oracle
    eng = O.Engine()
    wl = O.workload(O.WL_PHONON_LOGN, logn_median=3.0)
    a = eng.run(wl, 30, 9, threads=1, chunk=1, batched=True)
    b = eng.run(wl, 30, 9, threads=4, chunk=7, batched=True)
    assert a["stats"]["total_steps"] == b["stats"]["total_steps"]
    assert np.array_equal(a["hits"], b["hits"])


# ------------------------------------------------------------------ electrode masks and AMM film entry
# §8(f) rank 3: ElectrodePatternMask::find_channel (electrode_pattern.hpp:82-92) and
# IsotropicAMMInterface::transmission (interface_transmission.hpp:123-129), hooked into the
# film branch (engine.hpp:483-495) and charge collection (engine.hpp:287-293).
@pytest.mark.parametrize("tag", ["amm_geal", "amm_siw"])
def test_amm_golden(oracle, gold, tag):
    """P-SV flux coefficients, SH transmission and the engine's entry probability equal the
    reference's bit for bit over 300 angles."""
    O = oracle
    L = O.lib()
    m1, m2 = (O.Medium(*row) for row in gold[f"{tag}_media"])
    c = np.zeros(4)
    for i, th in enumerate(gold[f"{tag}_theta"]):
        # (exactly grazing incidence, theta = pi/2, is 0/0 in both: NaN)
        L.orc_amm_psv(C.byref(m1), C.byref(m2), 1, th, O.dptr(c))
        assert np.array_equal(c, gold[f"{tag}_psv"][i, :4], equal_nan=True)
        L.orc_amm_psv(C.byref(m1), C.byref(m2), 0, th, O.dptr(c))
        assert np.array_equal(c, gold[f"{tag}_psv"][i, 4:], equal_nan=True)
        got = [L.orc_amm_sh(C.byref(m1), C.byref(m2), th),
               L.orc_amm_transmission(C.byref(m1), C.byref(m2), 1, np.cos(th)),
               L.orc_amm_transmission(C.byref(m1), C.byref(m2), 0, np.cos(th))]
        want = [gold[f"{tag}_{k}"][i] for k in ("sh", "tl", "tt")]
        assert np.array_equal(got, want, equal_nan=True)


def _media(O):
    Si, Ge = (2330., 9000., 5400.), (5320., 5310., 3250.)
    Al2O3 = O.medium_from_lame(3866., 139.8e9, 166.24e9)
    Al = O.medium_from_bulk_shear(2720., 76e9, 26e9)
    W = O.medium_from_bulk_shear(18560., 297.672e9, 150.536e9)
    Ti = O.medium_from_bulk_shear(4265., 113.529e9, 63.5361e9)
    Ir = O.medium_from_bulk_shear(21980., 320e9, 210e9)
    # substrate, film, AMM L normal, T normal, SH flux average, DMM (tests/test_interface_amm.cpp:34-36)
    return [("Si/Al", Si, Al, 0.991, 0.960, 0.960, 0.743), ("Si/W", Si, W, 0.588, 0.621, 0.482, 0.778),
            ("Si/Ti", Si, Ti, 0.974, 0.982, 0.896, 0.658), ("Si/Ir", Si, Ir, 0.522, 0.527, 0.406, 0.752),
            ("Ge/Al", Ge, Al, 0.943, 0.881, 0.905, 0.510), ("Al2O3/Al", Al2O3, Al, 0.822, 0.748, 0.858, 0.810)], Si, Al


def test_amm_physics(oracle):
    """The reference's own AMM checks (tests/test_interface_amm.cpp:17-90) on the oracle:
    energy-flux conservation of the 4x4 P-SV solve at every angle (not imposed by the
    solver), the normal-incidence closed form 4 Z1 Z2 / (Z1 + Z2)^2 with no mode
    conversion, total transmission between identical media, SH and P-SV reciprocity, the
    published normal-incidence / flux-averaged / DMM table, DMM detailed balance."""
    O = oracle
    L = O.lib()
    pairs, Si, Al = _media(O)
    c = np.zeros(4)

    def psv(a, b, P, th):
        L.orc_amm_psv(C.byref(a), C.byref(b), P, th, O.dptr(c))
        return c.copy()

    for name, s, f, tL, tT, tSH, tD in pairs:
        ms, mf = O.Medium(*s), O.Medium(*f)
        worst = 0.0
        for th in (np.arange(400) + 0.5) / 400 * (np.pi / 2):
            for a, b in ((ms, mf), (mf, ms)):
                for P in (1, 0):
                    worst = max(worst, abs(psv(a, b, P, th).sum() - 1.0))
        assert worst <= 1e-10, f"{name}: flux conservation {worst:.2e}"
        p0, s0 = psv(ms, mf, 1, 0.0), psv(ms, mf, 0, 0.0)
        assert p0[2] == pytest.approx(L.orc_amm_normal(s[0] * s[1], f[0] * f[1]), rel=1e-12)
        assert s0[3] == pytest.approx(L.orc_amm_normal(s[0] * s[2], f[0] * f[2]), rel=1e-12)
        assert p0[1] + p0[3] + s0[0] + s0[2] <= 1e-14, f"{name}: mode conversion at normal incidence"
        n = 20000
        th = (np.arange(n) + 0.5) / n * (np.pi / 2)
        sh = sum(L.orc_amm_sh(C.byref(ms), C.byref(mf), t) * 2 * np.sin(t) * np.cos(t) for t in th) * (np.pi / 2) / n
        assert abs(p0[2] - tL) <= 6e-4 and abs(s0[3] - tT) <= 6e-4 and abs(sh - tSH) <= 6e-4, name
        assert abs(L.orc_dmm(C.byref(ms), C.byref(mf)) - tD) <= 6e-4, name
        assert L.orc_dmm(C.byref(ms), C.byref(mf)) + L.orc_dmm(C.byref(mf), C.byref(ms)) == pytest.approx(1.0, abs=1e-14)
    mSi = O.Medium(*Si)
    for th in (np.arange(100) + 0.5) / 100 * 1.5:
        assert abs(psv(mSi, mSi, 1, th)[2] - 1.0) <= 1e-12 and abs(psv(mSi, mSi, 0, th)[3] - 1.0) <= 1e-12
        assert abs(L.orc_amm_sh(C.byref(mSi), C.byref(mSi), th) - 1.0) <= 1e-12
    mAl = O.Medium(*Al)
    th1 = 0.6
    th2 = np.arcsin(Al[2] / Si[2] * np.sin(th1))
    assert L.orc_amm_sh(C.byref(mSi), C.byref(mAl), th1) == pytest.approx(L.orc_amm_sh(C.byref(mAl), C.byref(mSi), th2), rel=1e-12)
    thp2 = np.arcsin(Al[1] / Si[1] * np.sin(th1))
    assert psv(mSi, mAl, 1, th1)[2] == pytest.approx(psv(mAl, mSi, 1, thp2)[2], rel=1e-10)


def test_qet_mask_golden_and_geometry(oracle, gold):
    """find_channel of the QET layout (fill fraction 0.8) equals the reference's at 2000
    points; with fill 1 the channel follows from the radius and azimuth alone."""
    O = oracle
    L = O.lib()
    s = O.Surface()
    O.set_pads(s, O.qet_pads(0.0381, 0.8))
    rng = O.Rng()
    L.orc_rng_seed(C.byref(rng), 31)
    p = np.zeros(3)
    got = []
    for x, y in gold["qet_xy"]:
        p[0], p[1] = x, y
        got.append(L.orc_find_channel(C.byref(s), O.dptr(p), C.byref(rng)))
    assert np.array_equal(np.array(got), gold["qet_channels"])
    s1 = O.Surface()
    O.set_pads(s1, O.qet_pads(0.0381, 1.0))
    R = 0.0381
    for x, y in gold["qet_xy"]:
        p[0], p[1] = x, y
        r, phi = np.hypot(x, y), np.arctan2(y, x) % (2 * np.pi)
        sec = int(phi // (2 * np.pi / 3))
        off = phi - sec * 2 * np.pi / 3
        if r <= 0.55 * R:
            want = 0
        elif 0.57 * R <= r <= 0.90 * R and np.radians(1) <= off <= 2 * np.pi / 3 - np.radians(1):
            want = sec + 1
        elif 0.92 * R <= r <= 0.98 * R:
            want = 4
        else:
            want = -1
        if min(abs(r - k * R) for k in (0.55, 0.57, 0.90, 0.92, 0.98)) > 1e-9 and \
                min(abs(off - np.radians(1)), abs(off - 2 * np.pi / 3 + np.radians(1))) > 1e-9:
            assert L.orc_find_channel(C.byref(s1), O.dptr(p), C.byref(rng)) == want


@pytest.mark.parametrize("tag", ["mask_amm", "mask_pairs"])
def test_mask_runs_golden(oracle, gold, tag):
    """Whole runs with QET masks and AMM film entry on both faces (phonons; pairs at 4 V
    collected through the masks): statistics bytes, offsets, hits and channels equal the
    reference's."""
    O = oracle
    det = O.mask_amm_detector()
    if tag == "mask_pairs":
        O.set_bias(det, 4.0)
        det.luke_sample_fraction = 0.0
        wl, n = O.workload(O.WL_PAIR), 6
    else:
        wl, n = O.workload(O.WL_PHONON_POINT, mode=0, energy_eV=7.5e-3), 200
    r = O.Engine(detector=det).run(wl, n, 42, threads=4)
    assert r["stats_raw"] == gold[f"run_{tag}_stats"].tobytes()
    assert np.array_equal(r["hit_offsets"], gold[f"run_{tag}_offsets"])
    h = r["hits"][:600]
    got = np.c_[h["pos"], h["time"], h["edep"], h["incident"], h["surface"], h["mode"]]
    assert np.array_equal(got, gold[f"run_{tag}_hits"])
    assert np.array_equal(h["channel"], gold[f"run_{tag}_channels"])
