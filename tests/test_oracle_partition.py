"""The oracle's energy partition (oracle/oracle_partition.c) pinned to the reference:

* every ionisation-yield model (Lindhard, Lewin & Smith, Sarkis, IMPACT@TUNL, empirical
  k(E) with and without energy dependence) for Ge and Si targets over 68 energies from
  0.3 eV to 3 MeV (including the Sarkis knots' ends and the IMPACT and empirical
  boundaries), and the Lindhard yield of a Si projectile in Ge, bit for bit against the
  reference's golden vectors;
* the Fano binomial sampler: 200 draws in each of 11 (mean, F) cases covering every
  branch (mean <= 1, Poisson below and above 30, Gaussian limit, the exact binomial at an
  integral n_try, the interpolated rejection sampler), bit for bit;
* partition_deposit: result record, every charge and prompt phonon and the final stream
  state for 11 deposits (ER and NR, each NIEL model in use, without Fano fluctuations,
  without the charge cloud, with down-sampling, below the gap), bit for bit;
* the reference's own checks (tests/test_fano_partition.cpp): <N> = n_bar to 1 % and
  Var N / <N> = F to 5 % over 400,000 draws for F in {0.119, 0.15, 0.30} and n_bar in
  {2, 3.5, 7.3, 20, 100, 1e4}; and energy bookkeeping N E_g + sum E_kin w + sum E_ph w =
  omega to 1e-9 (here in Ge; the reference checks Si);
* mathematical checks that do not come from the reference: the Lindhard yield at a
  self-recoil equals the projectile form with (Z1, A1) = (Z, A); Lewin & Smith is
  Lindhard's k = 0.133 Z^2/3 A^-1/2 form; the sampler's mean <= 1 case returns 1; with
  Fano fluctuations off the pair count is floor(mean + u).
The run_events goldens `deposit_nr` / `deposit_er` (tests/test_oracle.py) pin the
deposit workload end to end.
"""
import ctypes as C

import numpy as np
import pytest


def _si(O):
    c = O.germanium()
    c.atomic_Z, c.atomic_A = 14.0, 28.085
    return c


def test_yields_golden(oracle, gold):
    O = oracle
    L = O.lib()
    E = gold["yield_E"]
    emp = O.partition_opts().empirical
    for ci, cr in enumerate((O.germanium(), _si(O))):
        for m in range(5):
            got = np.array([L.orc_ionization_yield(m, e, C.byref(cr), C.byref(emp)) for e in E])
            assert np.array_equal(got, gold["yield_values"][ci, m]), f"crystal {ci} model {m}"
    emp.energy_dependent_k = 0
    ge = O.germanium()
    got = np.array([L.orc_ionization_yield(4, e, C.byref(ge), C.byref(emp)) for e in E])
    assert np.array_equal(got, gold["yield_empirical_fixed_ge"])
    got = np.array([L.orc_lindhard_yield(e, C.byref(ge), 14.0, 28.085) for e in E])
    assert np.array_equal(got, gold["yield_lindhard_si_in_ge"])


def test_yield_identities(oracle):
    O = oracle
    L = O.lib()
    ge = O.germanium()
    for e in (1.0, 37.0, 1e3, 2.5e5):
        self_ = L.orc_ionization_yield(0, e, C.byref(ge), None)
        assert self_ == L.orc_lindhard_yield(e, C.byref(ge), 32.0, 72.63)
        # Lewin & Smith: k h / (1 + k h), h = 0.7 eps^0.6 + 3 eps^0.15 + eps
        eps = 11.5e-3 * e * 32.0 ** (-7.0 / 3.0)
        h = 0.7 * eps ** 0.6 + 3.0 * eps ** 0.15 + eps
        k = 0.133 * 32.0 ** (2.0 / 3.0) / np.sqrt(72.63)
        assert L.orc_ionization_yield(1, e, C.byref(ge), None) == pytest.approx(k * h / (1 + k * h), rel=1e-12)
    assert L.orc_ionization_yield(2, 100.0, C.byref(ge), None) == L.orc_ionization_yield(1, 100.0, C.byref(ge), None)
    assert L.orc_ionization_yield(3, 100.0, C.byref(ge), None) == L.orc_ionization_yield(1, 100.0, C.byref(ge), None)


def test_fano_golden(oracle, gold):
    O = oracle
    L = O.lib()
    for i, (mean, fano) in enumerate(gold["fano_cases"]):
        r = O.Rng()
        L.orc_rng_seed(C.byref(r), 100 + i)
        got = [L.orc_sample_fano_binomial(mean, fano, C.byref(r)) for _ in range(200)]
        assert np.array_equal(np.array(got, dtype=np.uint64), gold["fano_values"][i]), (mean, fano)
    assert np.all(gold["fano_values"][0] == 1)            # mean <= 1 returns 1


@pytest.mark.parametrize("F", [0.119, 0.15, 0.30])
def test_fano_moments(oracle, F):
    """tests/test_fano_partition.cpp: mean n_bar to 1 %, variance / mean = F to 5 %."""
    O = oracle
    r = O.Rng()
    O.lib().orc_rng_seed(C.byref(r), 77)
    n = 400000
    buf = np.zeros(n, dtype=np.uint64)
    for nb in (2.0, 3.5, 7.3, 20.0, 100.0, 1.0e4):
        O.lib().orc_fano_batch(nb, F, C.byref(r), n, buf.ctypes.data_as(C.POINTER(C.c_uint64)))
        x = buf.astype(np.float64)
        m, v = x.mean(), x.var()
        assert m == pytest.approx(nb, rel=0.01), f"F={F} nbar={nb}: <N> = {m}"
        assert v / m == pytest.approx(F, rel=0.05), f"F={F} nbar={nb}: Var/<N> = {v / m}"


def _partition_opts(O, changes):
    o = O.partition_opts()
    for k, v in changes.items():
        if k.startswith("empirical."):
            setattr(o.empirical, k.split(".", 1)[1], v)
        else:
            setattr(o, k, v)
    return o


PARTITION_CASES = [
    (50.0, 0, {}, 1, (0.0, 0.0, 0.0)),
    (1000.0, 0, {}, 2, (1e-3, -2e-3, 5e-3)),
    (100.0, 1, {}, 3, (0.0, 0.0, 0.0)),
    (1000.0, 1, {"niel_model": 1}, 4, (0.0, 0.0, 0.0)),
    (500.0, 1, {"niel_model": 4}, 5, (0.0, 0.0, 0.0)),
    (150.0, 1, {"niel_model": 4, "empirical.energy_dependent_k": 0}, 6, (0.0, 0.0, 0.0)),
    (200.0, 0, {"use_fano_fluctuations": 0}, 7, (0.0, 0.0, 0.0)),
    (200.0, 0, {"use_charge_cloud": 0}, 8, (2e-3, 0.0, 0.0)),
    (5000.0, 0, {"max_sampled_pairs": 7, "max_sampled_phonons": 11}, 9, (0.0, 0.0, 0.0)),
    (2.0, 0, {}, 10, (0.0, 0.0, 0.0)),
    (0.5, 1, {}, 11, (0.0, 0.0, 0.0)),
]
RESULT_KEYS = ("total_edep_eV", "ionizing_eV", "niel_eV", "lindhard_yield", "n_pairs_true", "n_pairs_sampled",
               "n_phonons_true", "n_phonons_sampled", "charge_weight", "phonon_weight", "prompt_phonon_eV")


@pytest.mark.parametrize("i", range(len(PARTITION_CASES)))
def test_partition_golden(oracle, gold, i):
    O = oracle
    e, nr, ch, seed, pos = PARTITION_CASES[i]
    r = O.Rng()
    O.lib().orc_rng_seed(C.byref(r), seed)
    out = O.Engine().partition(pos, e, bool(nr), _partition_opts(O, ch), r)
    res = out["result"]
    assert np.array_equal(np.array([float(res[k]) for k in RESULT_KEYS]), gold[f"part{i}_result"])
    q = gold[f"part{i}_charges"]
    got_q = np.c_[out["charge_type"], out["charge_pos"], out["charge_p"], out["charge_ekin"], out["charge_weight"],
                  out["charge_valley"], np.zeros(len(out["charge_type"]))] if len(out["charge_type"]) else np.zeros((0, 11))
    assert np.array_equal(got_q, q)
    p = gold[f"part{i}_phonons"]
    got_p = np.c_[out["phonon_mode"], np.tile(pos, (len(out["phonon_mode"]), 1)), out["phonon_k"], out["phonon_vg"],
                  out["phonon_vg_mag"], out["phonon_energy"], out["phonon_weight"],
                  np.zeros(len(out["phonon_mode"]))] if len(out["phonon_mode"]) else np.zeros((0, 14))
    assert np.array_equal(got_p, p)
    assert np.array_equal(np.array(list(r.s), dtype=np.uint64), gold[f"part{i}_rng"])


@pytest.mark.parametrize("E,nr", [(50.0, False), (1000.0, False), (300.0, True)])
def test_partition_energy_bookkeeping(oracle, E, nr):
    """tests/test_fano_partition.cpp: N E_g + sum E_kin w + sum E_ph w = omega (to 1e-9),
    without down-sampling."""
    O = oracle
    r = O.Rng()
    O.lib().orc_rng_seed(C.byref(r), 77)
    eng = O.Engine()
    o = O.partition_opts(max_sampled_pairs=1 << 30, max_sampled_phonons=1 << 30)
    out = eng.partition((0.0, 0.0, 0.0), E, nr, o, r)
    res = out["result"]
    total = res["n_pairs_true"] * eng.crystal.bandgap + np.sum(out["charge_ekin"] * out["charge_weight"]) + \
        np.sum(out["phonon_energy"] * out["phonon_weight"])
    assert total == pytest.approx(E, rel=1e-9)
    assert len(out["charge_type"]) == 2 * res["n_pairs_true"]
    assert len(out["phonon_energy"]) == res["n_phonons_true"]


def test_partition_without_fano_is_floor_mean_plus_u(oracle):
    O = oracle
    eng = O.Engine()
    o = O.partition_opts(use_fano_fluctuations=0, use_charge_cloud=0)
    for seed in range(20):
        r = O.Rng()
        O.lib().orc_rng_seed(C.byref(r), seed)
        u = O.Rng()
        O.lib().orc_rng_seed(C.byref(u), seed)
        n = eng.partition((0, 0, 0), 100.0, False, o, r)["result"]["n_pairs_true"]
        assert n == int(np.floor(100.0 / 2.96 + O.lib().orc_uniform(C.byref(u))))
