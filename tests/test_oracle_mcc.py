"""Pins the oracle's Philox stream, cross-section lookup, null-collision MCC
and swarm loop.  The reference's mt19937_64 stream cannot be reproduced, so
the reference comparisons are statistical; the closed forms below (rates,
conservation, threshold removal, zero-cross-section swarm) are exact."""
import math

import numpy as np
import pytest

import oracle_py
from aurora_pic import workloads
from conftest import golden


def test_philox_known_answer_vectors(oracle):
    # Random123 known-answer vectors for philox4x32, 10 rounds
    assert oracle.philox([0, 0, 0, 0], [0, 0]).tolist() == [
        0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2).tolist() == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                         [0xA4093822, 0x299F31D0]).tolist() == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_u01_is_open_and_float32_exact(oracle):
    for u in [0, 1, 511, 512, 0x7FFFFFFF, 0xFFFFFFFF]:
        x = oracle.u01(u)
        assert 0.0 < x < 1.0
        assert float(np.float32(x)) == x
    assert oracle.u01(0) == 2.0 ** -24 and oracle.u01(0xFFFFFFFF) == 1.0 - 2.0 ** -24


def test_stream_uniform_moments(oracle):
    s = oracle.stream(123, 4, 5, 6)
    x = np.array([oracle.stream_u01(s) for _ in range(40000)])
    assert abs(x.mean() - 0.5) < 0.005 and abs(x.var() - 1 / 12) < 0.002


@pytest.mark.parametrize("c", [0, 1, 2])
def test_xs_eval_matches_reference_golden(oracle, c):
    g = golden(f"xs_{c}.npz")
    out = np.array([oracle.xs_eval(g["e"], g["s"], float(q)) for q in g["q"]])
    assert np.max(np.abs(out - g["out"])) <= 1e-15 * np.max(np.abs(g["s"])) + 1e-300


def test_xs_known_answer(oracle):
    # test_core.cpp:300-321
    e, s = [0.0, 1.0, 2.0], [0.0, 0.5, 1.0]
    assert abs(oracle.xs_eval(e, s, 0.5) - 0.25) < 1e-15
    assert abs(oracle.xs_eval(e, s, 3.0) - 1.0) < 1e-15


def _unit_table_model(kind, threshold=0.0, max_frequency=2.0, sigma_scale=1.0,
                      max_candidates=64):
    e = [0.0, 1.0, 2.0]
    s = [0.0, 0.5 * sigma_scale, 1.0 * sigma_scale]
    return oracle_py.MccModel([(kind, threshold, e, s)], 1.0, max_frequency, 1.0,
                              max_candidates, 0.0)


def test_elastic_candidate_and_acceptance_rates(oracle):
    # test_core.cpp:353-367: candidates ~ nu_max dt N = 2000, accepted ~ 707
    m = _unit_table_model(workloads.ELASTIC)
    cand = acc = 0
    for p in range(10000):
        v, st = oracle.collide(m, [math.sqrt(2.0), 0.0, 0.0], 0.1, 8128, 0, p, 0)
        assert abs(np.linalg.norm(v) - math.sqrt(2.0)) < 1e-14
        cand += st.candidates
        acc += st.chan[0]
        assert st.status == 0
    assert 1800 < cand < 2200
    assert 620 < acc < 800


def test_excitation_removes_threshold(oracle):
    # test_core.cpp:596-634
    m = _unit_table_model(workloads.EXCITATION, threshold=0.5, sigma_scale=2.0,
                          max_frequency=4.0)
    accepted = 0
    for p in range(50):
        v, st = oracle.collide(m, [2.0, 0.0, 0.0], 2.0, 77, 0, p, 0)
        assert st.status == 0
        accepted += st.chan[0]
        e0 = 0.5 * 4.0
        e1 = 0.5 * float(np.dot(v, v))
        assert abs((e0 - e1) - 0.5 * st.chan[0]) < 1e-13
        assert abs(st.energy_change[0] - (e1 - e0)) < 1e-13
    assert accepted > 0


def test_ionization_energy_partition_closes(oracle):
    # test_core.cpp:636-684: primary + secondaries + thresholds = initial
    m = _unit_table_model(workloads.IONIZATION, threshold=0.5, max_frequency=10.0,
                          max_candidates=256)
    events = 0
    for p in range(40):
        v, st = oracle.collide(m, [4.0, 0.0, 0.0], 10.0, 314159, 0, p, 0)
        assert st.status == 0
        if st.nsec > 8:
            continue
        prod = 0.5 * float(np.dot(v, v))
        for s in range(st.nsec):
            prod += 0.5 * float(np.dot(st.sec_v[s], st.sec_v[s]))
        assert abs(prod + 0.5 * st.nsec - 8.0) < 1e-12
        assert st.nsec == st.chan[0]
        events += st.nsec
    assert events > 0


def test_frequency_bound_and_candidate_limit_errors(oracle):
    # rate above nu_max -> status 1 (Collision.cpp:617-623)
    m = _unit_table_model(workloads.ELASTIC, max_frequency=0.1)
    _, st = oracle.collide(m, [math.sqrt(2.0), 0, 0], 0.1, 1, 0, 0, 0)
    assert st.status == 1
    # candidate limit (Collision.cpp:864-870)
    m = _unit_table_model(workloads.ELASTIC, max_frequency=50.0, max_candidates=2)
    statuses = {oracle.collide(m, [0.1, 0, 0], 1.0, 2, 0, p, 0)[1].status for p in range(20)}
    assert 2 in statuses


def test_mcc_statistics_match_reference_golden(oracle):
    """Same inputs as the reference's collide() run on the synthetic gas:
    per-channel collision counts agree within 5 Poisson sigma."""
    g = golden("mcc_stats.npz")
    model = oracle_py.model_from_gas(workloads.synthetic_gas(rows=200))
    vin = g["vin"].astype(np.float64)
    dt = float(g["dt"])
    counts = np.zeros(2 + 3 + 1, dtype=np.int64)
    for p in range(vin.shape[0]):
        _, st = oracle.collide(model, vin[p], dt, 8128, 0, p, 0)
        assert st.status == 0
        counts[0] += st.candidates
        counts[1] += st.nulls
        for c in range(3):
            counts[2 + c] += st.chan[c]
        counts[5] += st.nsec
    ref = g["counts"]
    for k in range(6):
        sigma = math.sqrt(max(ref[k], 1))
        assert abs(int(counts[k]) - int(ref[k])) <= 5 * sigma + 2, (k, counts[k], ref[k])


def test_swarm_zero_cross_section_closed_form(oracle):
    # cf. test_core.cpp:6190-6230: with sigma = 0 the electron is only
    # accelerated: v_x after step s is -(s+1) a dt exactly.
    e = np.array([0.0, 1.0e-16])
    m = oracle_py.MccModel([(workloads.ELASTIC, 0.0, e, np.zeros(2))], 1e20, 1e7,
                           workloads.ME, 64, workloads.HELIUM_MASS)
    cfg = dict(neutral_density=1e20, max_frequency=1e7, timestep=1e-9, reduced_field_td=1.0,
               steps=40, burn_in=10, particles=5, blocks=5, initial_mean_energy_ev=0.0,
               max_energy_ev=1e3)
    r = oracle.swarm(m, cfg, np.zeros(15), seed=1)
    a_dt = workloads.EV * 1.0e-21 * 1e20 / workloads.ME * 1e-9
    s1 = np.arange(11, 41, dtype=float)
    assert abs(r.drift - a_dt * s1.mean()) < 1e-13 * a_dt * s1.mean()
    e_exp = 0.5 * workloads.ME * a_dt ** 2 * np.mean(s1 ** 2) / workloads.EV
    assert abs(r.mean_energy_ev - e_exp) < 1e-12 * e_exp
    assert r.collisions[0] == 0


def test_swarm_matches_reference_golden_statistically(oracle):
    g = golden("swarm_300td.npz")
    model = oracle_py.model_from_gas(workloads.synthetic_gas(rows=200))
    cfg = {k: (float(g[k]) if k not in ("steps", "burn_in", "particles", "blocks")
               else int(g[k])) for k in ["neutral_density", "max_frequency", "timestep",
                                          "reduced_field_td", "steps", "burn_in", "particles",
                                          "blocks", "initial_mean_energy_ev",
                                          "max_energy_ev"]}
    rng = np.random.default_rng(1)
    sigma = math.sqrt(2.0 * cfg["initial_mean_energy_ev"] * workloads.EV / (3.0 * workloads.ME))
    vel0 = rng.normal(0.0, sigma, size=(cfg["particles"], 3))
    r = oracle.swarm(model, cfg, vel0, seed=271828)
    ref = g["out"]
    assert r.status == 0
    # drift and mean energy within 5 combined standard errors (+2 % floor)
    for mine, se, theirs, tse in [(r.drift, r.drift_se, ref[0], ref[1]),
                                  (r.mean_energy_ev, r.mean_energy_se, ref[2], ref[3])]:
        tol = 5 * math.hypot(se, tse) + 0.02 * abs(theirs)
        assert abs(mine - theirs) <= tol, (mine, theirs, tol)
    # ionization rate within 5 Poisson sigma
    assert abs(r.ion_rate - ref[4]) <= 5 * math.hypot(r.ion_rate_se, ref[5]) + 1e-9 * ref[4]


def test_block_mean_error_known_answer(oracle):
    # Swarm.cpp:103-134 on a hand-checked case
    m, se = oracle.block_mean_error([1.0, 2.0, 3.0, 4.0, 5.0, 6.0], 3)
    assert m == 3.5
    assert abs(se - math.sqrt(((1.5 - 3.5) ** 2 + 0 + (5.5 - 3.5) ** 2) / (3 * 2))) < 1e-15


def _elastic_model(max_frequency=2.0, neutral_mass=0.0, options=None, kind=workloads.ELASTIC,
                   neutral_vstd=0.0, particle_mass=1.0):
    e, s = [0.0, 1.0, 2.0], [0.0, 0.5, 1.0]
    return oracle_py.MccModel([(kind, 0.0, e, s)], 1.0, max_frequency, particle_mass, 64,
                              neutral_mass, neutral_vstd, options=options)


def test_henyey_greenstein_mean_cosine(oracle):
    # test_core.cpp:397-440: mean cosine 0.8 table, single-collision events
    # keep |v| and scatter with <cos> = 0.8 +- 0.04
    mc = ([0.0, 2.0], [0.8, 0.8])
    m = _elastic_model(options={0: (oracle_py.HENYEY_GREENSTEIN, 0, mc[0], mc[1])})
    n = cos_sum = 0
    for p in range(20000):
        v, st = oracle.collide(m, [math.sqrt(2.0), 0.0, 0.0], 0.1, 161803, 0, p, 0)
        assert abs(np.linalg.norm(v) - math.sqrt(2.0)) < 1e-14
        if st.chan[0] != 1:
            continue
        cos_sum += v[0] / math.sqrt(2.0)
        n += 1
    assert n > 1200
    assert abs(cos_sum / n - 0.8) < 0.04


def test_center_of_mass_frame_selects_reduced_energy(oracle):
    # test_core.cpp:530-549: equal masses, CoM energy 0.5 -> sigma 0.25
    m = _elastic_model(neutral_mass=1.0, options={0: (0, oracle_py.FRAME_CENTER_OF_MASS,
                                                         None, None)})
    acc = 0
    for p in range(10000):
        _, st = oracle.collide(m, [math.sqrt(2.0), 0.0, 0.0], 0.1, 141421, 0, p, 0)
        acc += st.chan[0]
    assert 280 < acc < 440
    m2 = _elastic_model(neutral_mass=1.0)
    acc2 = sum(oracle.collide(m2, [math.sqrt(2.0), 0, 0], 0.1, 141421, 0, p, 0)[1].chan[0]
               for p in range(10000))
    assert acc2 > 1.6 * acc       # projectile frame: sigma 0.5


def test_backward_equal_mass_exchanges_velocity(oracle):
    # test_core.cpp:559-583: equal-mass backward scattering stops the projectile
    m = _elastic_model(neutral_mass=1.0,
                       options={0: (oracle_py.BACKWARD, oracle_py.FRAME_CENTER_OF_MASS, None,
                                    None)})
    seen = False
    for p in range(10000):
        v, st = oracle.collide(m, [math.sqrt(2.0), 0.0, 0.0], 0.1, 173205, 0, p, 0)
        if st.chan[0] != 1:
            continue
        assert np.linalg.norm(v) < 1e-14
        seen = True
        break
    assert seen


def test_charge_exchange_makes_slow_ion(oracle):
    # test_core.cpp:794-817 (cold neutral) and 818-852 (thermal neutral)
    m = _elastic_model(kind=oracle_py.CHARGE_EXCHANGE, neutral_mass=40.0, particle_mass=40.0)
    seen = False
    for p in range(10000):
        v, st = oracle.collide(m, [1.0, 0.0, 0.0], 0.01, 141421, 0, p, 0)
        if st.chan[0] == 0:
            continue
        assert st.chan[0] == 1 and np.linalg.norm(v) == 0.0 and st.nsec == 0
        seen = True
        break
    assert seen
    mt = _elastic_model(kind=oracle_py.CHARGE_EXCHANGE, neutral_mass=1.0, max_frequency=20.0,
                        neutral_vstd=1.0)
    seen = False
    for p in range(10000):
        v, st = oracle.collide(mt, [2.0, 0.0, 0.0], 0.01, 173205, 0, p, 0)
        if st.chan[0] == 0:
            continue
        assert 0.0 < np.linalg.norm(v) <= 8.0 and st.nsec == 0
        seen = True
        break
    assert seen
    unsafe = _elastic_model(kind=oracle_py.CHARGE_EXCHANGE, neutral_mass=1.0, max_frequency=1.0,
                            neutral_vstd=1.0)
    _, st = oracle.collide(unsafe, [2.0, 0.0, 0.0], 0.01, 1, 0, 0, 0)
    assert st.status == 1


# ------------------------------------------------------------ 1V collide
def test_1v_elastic_rates_and_speed(oracle):
    # test_core.cpp:336-367 (the scalar collide): candidates ~ nu_max dt N =
    # 2000, accepted ~ n sigma v dt N = 707, |v| kept to 1e-14; the sign
    # flips in about half of the accepted events (forward or backward)
    m = _unit_table_model(workloads.ELASTIC)
    cand = acc = flips = 0
    for p in range(10000):
        v, st = oracle.collide_1v(m, math.sqrt(2.0), 0.1, 8128, 0, p, 0)
        assert st.status == 0
        assert abs(abs(v) - math.sqrt(2.0)) < 1e-14
        cand += st.candidates
        acc += st.chan[0]
        flips += v < 0
    assert 1800 < cand < 2200
    assert 620 < acc < 800
    # a particle reversed by an odd number of its accepted events
    assert 0.3 * acc < flips < 0.7 * acc


def test_1v_excitation_removes_threshold(oracle):
    m = _unit_table_model(workloads.EXCITATION, threshold=0.5, sigma_scale=2.0,
                          max_frequency=4.0)
    accepted = 0
    for p in range(50):
        v, st = oracle.collide_1v(m, 2.0, 2.0, 77, 0, p, 0)
        assert st.status == 0
        accepted += st.chan[0]
        e0, e1 = 0.5 * 4.0, 0.5 * v * v
        assert abs((e0 - e1) - 0.5 * st.chan[0]) < 1e-13
        assert abs(st.energy_change[0] - (e1 - e0)) < 1e-13
    assert accepted > 0


def test_1v_heavy_neutral_elastic_and_charge_exchange(oracle):
    # elastic on a neutral at rest of mass M: v' = m v / (m + M) +/- M |v| / (m + M)
    e, s = [0.0, 10.0], [1.0, 1.0]
    m = oracle_py.MccModel([(workloads.ELASTIC, 0.0, e, s)], 1.0, 4.0, 1.0, 64, 3.0)
    for p in range(40):
        v, st = oracle.collide_1v(m, 1.5, 0.2, 5, 0, p, 0)
        if st.chan[0] == 1:
            assert min(abs(v - (1.5 + 4.5) / 4.0), abs(v - (1.5 - 4.5) / 4.0)) < 1e-14
    # charge exchange on cold neutrals stops the ion
    m = oracle_py.MccModel([(oracle_py.CHARGE_EXCHANGE, 0.0, e, s)], 1.0, 4.0, 1.0, 64, 1.0)
    stopped = 0
    for p in range(40):
        v, st = oracle.collide_1v(m, 1.5, 0.5, 6, 0, p, 0)
        if st.chan[0]:
            assert v == 0.0
            stopped += 1
    assert stopped > 0


def test_decision_margins(oracle):
    # the margin is the smallest relative distance to a decision boundary:
    # with a zero cross section the decisions are the waiting times against
    # the time left (|t_wait - left| / dt) and each candidate's selection
    # against the zero cumulative rate (selection / nu_max = its uniform)
    m = _unit_table_model(workloads.ELASTIC, sigma_scale=0.0)
    v, st = oracle.collide(m, [1.0, 0.0, 0.0], 0.1, 3, 0, 0, 0)
    s = oracle.stream(3, 0, 0, 0)
    left, margin = 0.1, 1.0
    while True:
        t = -math.log(oracle.stream_u01(s)) / 2.0
        margin = min(margin, abs(t - left) / 0.1)
        if t >= left:
            break
        left -= t
        margin = min(margin, oracle.stream_u01(s))   # the selection of a candidate
    assert st.margin == pytest.approx(margin, rel=1e-12)
    assert 0.0 < st.margin <= 1.0
