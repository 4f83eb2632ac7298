"""GPU parity of the null-collision MCC (FP32, Philox uniforms) against the
FP64 oracle fed the same uniform stream, the reference's closed forms, and
the reference's own statistics (tests/golden/mcc_stats.npz).

Bar: per particle, velocities agree to 1e-4 relative for >= 99 % of the
particles, and every particle that does not sits within FP32 rounding of a
decision boundary: the oracle's smallest relative distance of a branch
decision of that particle to its boundary (waiting time against the time
left, selection against the cumulative rates, energy against a threshold,
scatter sign) is at most MARGIN, where FP32 and FP64 may round to different
sides; aggregate counts agree within 5 Poisson sigma."""
import math
import os
import tempfile

import numpy as np
import pytest

import oracle_py
from aurora_pic import workloads
from conftest import golden

pytestmark = pytest.mark.gpu

# FP32 rounding of a decision: the device's uniforms equal the oracle's, its
# waiting times, rates, cumulative sums and energies carry a few FP32 ulps
# (2^-24 each), and its velocities after earlier collisions a few more
MARGIN = 1e-5


@pytest.fixture
def ctx():
    from aurora_pic import hip
    c = hip.Context(0)
    c.mesh([17], [1.0], 1, 1.0)
    yield c
    c.close()


def _load(ctx, channels, density, max_frequency, mass, max_candidates=64, neutral_mass=0.0,
          seed=8128):
    sid = ctx.add_species(-1.0, mass, 1.0, 64)
    extra = ctx.add_species(1.0, 1000.0, 1.0, 64)
    gas = dict(channels=channels, neutral_density=density, max_frequency=max_frequency,
               neutral_mass=neutral_mass, max_candidates=max_candidates)
    ctx.mcc_load(sid, gas, product_species=(sid, extra), seed=seed)
    return gas


def test_collide_matches_oracle_same_stream(ctx, oracle):
    g = golden("mcc_stats.npz")
    gas = workloads.synthetic_gas(rows=200)
    model = oracle_py.model_from_gas(workloads.synthetic_gas(rows=200))
    _load(ctx, gas["channels"], gas["neutral_density"], gas["max_frequency"], workloads.ME,
          neutral_mass=gas["neutral_mass"], seed=8128)
    vin = g["vin"].astype(np.float64)
    dt = float(g["dt"])
    vout, counts, ech = ctx.mcc_collide(vin, dt, step=0)
    ocounts = np.zeros(6, dtype=np.int64)
    close = 0
    far_margins = []
    for p in range(vin.shape[0]):
        ov, st = oracle.collide(model, vin[p], dt, 8128, 0, p, 0)
        ocounts += [st.candidates, st.nulls, st.chan[0], st.chan[1], st.chan[2], st.nsec]
        if np.max(np.abs(vout[p] - ov)) <= 1e-4 * np.linalg.norm(ov) + 1e-3:
            close += 1
        else:
            far_margins.append(st.margin)
    assert close >= 0.99 * vin.shape[0], close
    print(f"{len(far_margins)} of {vin.shape[0]} differ; their largest decision margin "
          f"{max(far_margins, default=0.0):.2e}", flush=True)
    assert max(far_margins, default=0.0) <= MARGIN
    for k in range(6):
        assert abs(int(counts[k]) - int(ocounts[k])) <= 5 * math.sqrt(max(ocounts[k], 1)) + 2
    # and against the reference's own mt19937_64 statistics on the same inputs
    ref = g["counts"]
    for k in range(6):
        assert abs(int(counts[k]) - int(ref[k])) <= 5 * math.sqrt(max(ref[k], 1)) + 2, k


def test_elastic_rates_and_energy_conservation(ctx):
    # test_core.cpp:353-367 closed forms on the device
    e, s = [0.0, 1.0, 2.0], [0.0, 0.5, 1.0]
    _load(ctx, [(workloads.ELASTIC, 0.0, e, s)], 1.0, 2.0, 1.0)
    v = np.tile([math.sqrt(2.0), 0.0, 0.0], (10000, 1))
    vout, counts, _ = ctx.mcc_collide(v, 0.1, step=3)
    assert 1800 < counts[0] < 2200
    assert 620 < counts[2] < 800
    assert np.max(np.abs(np.linalg.norm(vout, axis=1) - math.sqrt(2.0))) < 1e-6


def test_excitation_removes_threshold(ctx):
    e, s = [0.0, 1.0, 2.0], [0.0, 1.0, 2.0]
    _load(ctx, [(workloads.EXCITATION, 0.5, e, s)], 1.0, 4.0, 1.0)
    v = np.tile([2.0, 0.0, 0.0], (200, 1))
    vout, counts, ech = ctx.mcc_collide(v, 2.0, step=1)
    e1 = 0.5 * np.sum(vout ** 2, axis=1)
    loss = 2.0 - e1
    assert counts[2] > 0
    assert abs(loss.sum() - 0.5 * counts[2]) < 1e-4 * counts[2]
    assert abs(ech[0] + loss.sum()) < 1e-4 * counts[2]


def test_ionization_counts_secondaries(ctx):
    e, s = [0.0, 1.0, 2.0], [0.0, 0.5, 1.0]
    _load(ctx, [(workloads.IONIZATION, 0.5, e, s)], 1.0, 10.0, 1.0, max_candidates=256)
    v = np.tile([4.0, 0.0, 0.0], (500, 1))
    vout, counts, ech = ctx.mcc_collide(v, 1.0, step=0)
    assert counts[2] > 0 and counts[3] == counts[2]      # one secondary per event


def test_rate_above_bound_error(ctx):
    from aurora_pic import hip
    e, s = [0.0, 1.0, 2.0], [0.0, 0.5, 1.0]
    _load(ctx, [(workloads.ELASTIC, 0.0, e, s)], 1.0, 0.1, 1.0)
    with pytest.raises(hip.HipError, match="exceeds configured max_frequency"):
        ctx.mcc_collide(np.tile([math.sqrt(2.0), 0, 0], (10, 1)), 0.1, step=0)


def _oracle_step_1d(oracle, model, state, n, L, eps, dt, t1, phi_r, qm, w, seed, step):
    """One reference-semantics 1D step (Simulation.cpp:1008-1138) in FP64:
    push all species, deposit, Dirichlet solve at t+dt, then MCC on
    species 0 with products appended (secondary -> species 0, ion -> 1)."""
    for s, st in enumerate(state):
        st["pos"], st["vh"], st["alive"], _ = oracle.push([n], [L], 1, st["pos"], st["vh"],
                                                         st["alive"], st["E"], qm[s], dt)
    rho = np.zeros(n)
    for s, st in enumerate(state):
        rho += oracle.deposit_cic([n], [L], 1, st["pos"], qm[s] * 0 + st["q"], w[s], st["alive"])
    _, E = oracle.poisson([n], [L], 1, eps, rho, 0.0, phi_r)
    e0 = state[0]
    Ep = oracle.gather([n], [L], 1, e0["pos"], E[0])
    products = []
    for p in range(len(e0["alive"])):
        if not e0["alive"][p]:
            continue
        v = e0["vh"][p].copy()
        v[0] += 0.5 * qm[0] * Ep[p] * dt
        v, st = oracle.collide(model, v, dt, seed, 0, int(e0["uid"][p]), step)
        e0.setdefault("margin", np.ones(len(e0["alive"])))[p] = st.margin
        for k in range(st.nsec):
            products.append((p, np.array(st.sec_v[k]), np.array(st.ion_v[k])))
        v[0] -= 0.5 * qm[0] * Ep[p] * dt
        e0["vh"][p] = v
    return E, rho, products


def test_step_1d_with_mcc_matches_oracle(oracle):
    """Config-[1]-shaped 1D3V RF discharge step: push, deposit, Thomas solve
    with the RF electrode at t+dt, MCC with ionization products."""
    from aurora_pic import hip
    rng = np.random.default_rng(5)
    n, L = 65, 0.067
    eps = workloads.EPS0
    dt = 1.0 / (400 * 13.56e6)
    N = 3000
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    model = oracle_py.model_from_gas(workloads.synthetic_gas(rows=1000, neutral_density=1.0e22))
    species = [(-workloads.EV, workloads.ME, 8.0), (workloads.EV, 6.67e-27, 0.026)]
    w = [2.56e14 * L / N] * 2
    state = []
    c = hip.Context(0)
    c.mesh([n], [L], 1, eps)
    c.electrodes([0, 0, 0, 0], [0.0, 200.0, 13.56e6, 0.0])
    sids = []
    for q, m, te in species:
        sid = c.add_species(q, m, w[0], 4 * N)
        pos = rng.uniform(0.1 * L, 0.9 * L, size=N)
        vel = rng.normal(0, math.sqrt(te * workloads.EV / m), size=(N, 3))
        c.upload(sid, pos, vel)
        sids.append(sid)
    c.mcc_load(sids[0], gas, product_species=(sids[0], sids[1]), seed=77)
    c.deposit()
    c.solve(0.0, 0.0, True)
    for sid, (q, m, te) in zip(sids, species):
        p, v, a, u = c.download(sid)
        state.append(dict(pos=p[:, 0].copy(), vh=v.copy(), alive=a.copy(), uid=u.copy(), q=q))
    _, _, E0 = c.fields()
    for st in state:
        st["E"] = E0[0]
    c.step(dt, 1)
    qm = [q / m for q, m, _ in species]
    phi_r = 200.0 * math.sin(2 * math.pi * 13.56e6 * dt)
    E1, rho1, products = _oracle_step_1d(oracle, model, state, n, L, eps, dt, dt, phi_r, qm, w,
                                         77, 0)
    rg, _, Eg = c.fields()
    assert np.max(np.abs(Eg[0] - E1[0])) <= 1e-6 * np.max(np.abs(E1[0]))
    # electrons: the original particles keep their slots; products follow
    p, v, a, u = c.download(sids[0])
    e0 = state[0]
    keep = e0["alive"] == 1
    close = np.max(np.abs(v[:N][keep] - e0["vh"][keep]), axis=1) <= (
        1e-4 * np.linalg.norm(e0["vh"][keep], axis=1) + 1.0)
    assert close.mean() >= 0.99
    far = e0["margin"][keep][~close]
    print(f"{far.size} of {close.size} differ; largest decision margin "
          f"{far.max(initial=0.0):.2e}", flush=True)
    assert far.max(initial=0.0) <= MARGIN
    slots_e, _ = c.count(sids[0])
    slots_i, _ = c.count(sids[1])
    assert abs((slots_e - N) - len(products)) <= max(2, 0.05 * len(products))
    assert slots_i - N == slots_e - N
    counts, ech = c.mcc_stats()
    assert counts[2 + 2] == slots_e - N          # ionizations = products
    # product uids continue after the uploaded ones, in parent order
    assert np.array_equal(u[N:slots_e], np.arange(N, slots_e, dtype=np.uint32))
    c.close()


def _compare_same_stream(ctx, oracle, model, vin, dt, seed, tol_rel=1e-4, tol_abs=1e-3):
    vout, counts, _ = ctx.mcc_collide(vin, dt, step=0)
    nchan = len(model.channels)
    ocounts = np.zeros(3 + nchan, dtype=np.int64)
    close = 0
    far_margins = []
    for p in range(vin.shape[0]):
        ov, st = oracle.collide(model, vin[p], dt, seed, 0, p, 0)
        assert st.status == 0
        ocounts += [st.candidates, st.nulls] + [st.chan[c] for c in range(nchan)] + [st.nsec]
        if np.max(np.abs(vout[p] - ov)) <= tol_rel * np.linalg.norm(ov) + tol_abs:
            close += 1
        else:
            far_margins.append(st.margin)
    assert close >= 0.99 * vin.shape[0], close
    assert max(far_margins, default=0.0) <= MARGIN, max(far_margins)
    for k in range(3 + nchan):
        assert abs(int(counts[k]) - int(ocounts[k])) <= 5 * math.sqrt(max(ocounts[k], 1)) + 2, k
    return vout, counts


def test_anisotropic_com_electron_model_matches_oracle(ctx, oracle):
    """Henyey-Greenstein elastic scattering in the centre-of-mass frame with
    an energy-dependent mean cosine, plus excitation and ionization."""
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=200)
    e_mc = np.array([0.0, 10.0 * workloads.EV, 2000.0 * workloads.EV])
    v_mc = np.array([0.1, 0.6, 0.9])
    _load(ctx, gas["channels"], gas["neutral_density"], gas["max_frequency"], workloads.ME,
          neutral_mass=gas["neutral_mass"], seed=99)
    ctx.mcc_channel_options(0, hip.SCATTER_HENYEY_GREENSTEIN, hip.FRAME_CENTER_OF_MASS, e_mc, v_mc)
    model = oracle_py.MccModel(gas["channels"], gas["neutral_density"], gas["max_frequency"],
                               workloads.ME, 64, gas["neutral_mass"],
                               options={0: (oracle_py.HENYEY_GREENSTEIN,
                                            oracle_py.FRAME_CENTER_OF_MASS, e_mc, v_mc)})
    rng = np.random.default_rng(3)
    vin = rng.normal(0.0, 3.0e6, size=(4000, 3))
    _compare_same_stream(ctx, oracle, model, vin, 0.099 / gas["max_frequency"] * 5, 99)


def test_thermal_ion_charge_exchange_backward_matches_oracle(ctx, oracle):
    """Ion model: resonant charge exchange and backward elastic scattering,
    both in the centre-of-mass frame, against 300 K thermal neutrals."""
    from aurora_pic import hip
    mi = 40.0 * workloads.AMU
    e = np.geomspace(1e-3, 1e3, 300) * workloads.EV
    cx = 5e-19 * np.ones_like(e)
    el = 3e-19 / np.sqrt(1.0 + e / workloads.EV)
    chans = [(hip.CHARGE_EXCHANGE, 0.0, e, cx), (hip.ELASTIC, 0.0, e, el)]
    n0, T = 1e21, 300.0
    vmax = math.sqrt(2 * 1e3 * workloads.EV / mi)
    numax = 1.2 * n0 * (5e-19 + 3e-19) * (vmax + 8 * math.sqrt(workloads.KB * T / mi) + 2e4)
    sid = ctx.add_species(workloads.EV, mi, 1.0, 64)
    gas = dict(channels=chans, neutral_density=n0, max_frequency=numax, neutral_mass=mi)
    ctx.mcc_load(sid, gas, seed=4242, neutral_temperature=T)
    ctx.mcc_channel_options(0, hip.SCATTER_ISOTROPIC, hip.FRAME_CENTER_OF_MASS)
    ctx.mcc_channel_options(1, hip.SCATTER_BACKWARD, hip.FRAME_CENTER_OF_MASS)
    model = oracle_py.MccModel(chans, n0, numax, mi, 64, mi, math.sqrt(workloads.KB * T / mi),
                               options={0: (0, oracle_py.FRAME_CENTER_OF_MASS, None, None),
                                        1: (oracle_py.BACKWARD, oracle_py.FRAME_CENTER_OF_MASS,
                                            None, None)})
    rng = np.random.default_rng(4)
    vin = rng.normal(0.0, 2.0e3, size=(3000, 3))
    _, counts = _compare_same_stream(ctx, oracle, model, vin, 5.0 / numax, 4242, tol_abs=1e-1)
    assert counts[2] > 0 and counts[3] > 0


def test_reference_closed_forms_for_channel_options(ctx):
    """test_core.cpp:397-440 (HG mean cosine 0.8), 530-549 (centre-of-mass
    cross section), 559-583 (equal-mass backward), 794-817 (charge
    exchange) on the device."""
    from aurora_pic import hip
    e, s = [0.0, 1.0, 2.0], [0.0, 0.5, 1.0]
    v0 = np.tile([math.sqrt(2.0), 0.0, 0.0], (20000, 1))
    _load(ctx, [(hip.ELASTIC, 0.0, e, s)], 1.0, 2.0, 1.0, seed=161803)
    ctx.mcc_channel_options(0, hip.SCATTER_HENYEY_GREENSTEIN, 0, [0.0, 2.0], [0.8, 0.8])
    vout, counts, _ = ctx.mcc_collide(v0, 0.1, step=0)
    assert np.max(np.abs(np.linalg.norm(vout, axis=1) - math.sqrt(2.0))) < 1e-5
    s2 = float(np.float32(math.sqrt(2.0)))    # the uploaded FP32 velocity
    moved = vout[:, 0] != s2
    cosines = vout[moved, 0] / math.sqrt(2.0)
    assert moved.sum() > 1200
    assert abs(cosines.mean() - 0.8) < 0.04
    # centre-of-mass frame (equal masses): sigma(0.5) = 0.25 -> ~354 of 10000
    _load(ctx, [(hip.ELASTIC, 0.0, e, s)], 1.0, 2.0, 1.0, neutral_mass=1.0, seed=141421)
    ctx.mcc_channel_options(0, 0, hip.FRAME_CENTER_OF_MASS)
    _, counts, _ = ctx.mcc_collide(v0[:10000], 0.1, step=0)
    assert 280 < counts[2] < 440
    # backward, equal masses: the projectile stops
    ctx.mcc_channel_options(0, hip.SCATTER_BACKWARD, hip.FRAME_CENTER_OF_MASS)
    vout, counts, _ = ctx.mcc_collide(v0[:10000], 0.1, step=1)
    hit = vout[:, 0] != s2
    assert hit.sum() > 0 and np.max(np.linalg.norm(vout[hit], axis=1)) < 1e-6
    # resonant charge exchange with a cold neutral: the ion stops exactly
    _load(ctx, [(hip.CHARGE_EXCHANGE, 0.0, e, s)], 1.0, 2.0, 40.0, neutral_mass=40.0, seed=141421)
    vout, counts, _ = ctx.mcc_collide(np.tile([1.0, 0.0, 0.0], (10000, 1)), 0.01, step=0)
    hit = counts[2]
    assert hit > 0 and np.sum(np.linalg.norm(vout, axis=1) == 0.0) == hit
    assert counts[3] == 0


def test_channel_option_errors(ctx):
    from aurora_pic import hip
    e, s = [0.0, 1.0, 2.0], [0.0, 0.5, 1.0]
    _load(ctx, [(hip.EXCITATION, 0.5, e, s)], 1.0, 4.0, 1.0)
    with pytest.raises(hip.HipError, match="supported only for elastic channels"):
        ctx.mcc_channel_options(0, hip.SCATTER_BACKWARD, 0)
    with pytest.raises(hip.HipError, match="requires positive neutral_mass"):
        ctx.mcc_channel_options(0, 0, hip.FRAME_CENTER_OF_MASS)
    _load(ctx, [(hip.ELASTIC, 0.0, e, s)], 1.0, 2.0, 1.0)
    with pytest.raises(hip.HipError, match="must cover the cross-section table"):
        ctx.mcc_channel_options(0, hip.SCATTER_HENYEY_GREENSTEIN, 0, [0.5, 2.0], [0.2, 0.2])
    with pytest.raises(hip.HipError, match="projectile mass equal to neutral_mass"):
        _load(ctx, [(hip.CHARGE_EXCHANGE, 0.0, e, s)], 1.0, 2.0, 39.0, neutral_mass=40.0)


def test_1v_collide_matches_oracle_same_stream(ctx, oracle):
    """The scalar collide (Collision.cpp:831-905) of a 1D1V run: elastic on
    4 amu neutrals at 300 K (centre-of-mass kinematics, thermal sampling)
    and excitation, against the oracle's scalar collide on the same stream;
    the y and z components stay untouched."""
    e = np.geomspace(1e-3, 1e3, 300) * workloads.EV
    chans = [(workloads.ELASTIC, 0.0, e, np.full(300, 6e-20)),
             (workloads.EXCITATION, 19.8 * workloads.EV, e,
              np.where(e > 19.8 * workloads.EV, 3e-20, 0.0))]
    nu = workloads.max_collision_frequency(chans, 1e22, margin=1.5)
    ctx.velocity_dimensions(1)
    sid = ctx.add_species(-1.0, workloads.ME, 1.0, 64)
    gas = dict(channels=chans, neutral_density=1e22, max_frequency=nu,
               neutral_mass=workloads.HELIUM_MASS)
    ctx.mcc_load(sid, gas, seed=99, neutral_temperature=300.0)
    vstd = math.sqrt(workloads.KB * 300.0 / workloads.HELIUM_MASS)
    model = oracle_py.model_from_gas(dict(gas, particle_mass=workloads.ME), neutral_vstd=vstd)
    rng = np.random.default_rng(4)
    n = 8000
    vin = np.zeros((n, 3))
    vin[:, 0] = rng.normal(0.0, 2.5e6, n)
    vin[:, 1:] = rng.normal(0.0, 1e5, (n, 2))
    dt = 0.09 / nu * 4
    vout, counts, _ = ctx.mcc_collide(vin, dt, step=2)
    assert np.array_equal(vout[:, 1:], vin[:, 1:].astype(np.float32).astype(np.float64))
    ocounts = np.zeros(5, dtype=np.int64)
    close, far = 0, []
    for p in range(n):
        ov, st = oracle.collide_1v(model, vin[p, 0], dt, 99, 0, p, 2)
        assert st.status == 0
        ocounts += [st.candidates, st.nulls, st.chan[0], st.chan[1], 0]
        if abs(vout[p, 0] - ov) <= 1e-4 * abs(ov) + 1e-3:
            close += 1
        else:
            far.append(st.margin)
    assert close >= 0.99 * n, close
    assert max(far, default=0.0) <= MARGIN, max(far)
    assert ocounts[2] > 100 and ocounts[3] > 10        # both channels fired
    for k in range(4):
        assert abs(int(counts[k]) - int(ocounts[k])) <= 5 * math.sqrt(max(ocounts[k], 1)) + 2, k
    # ionization has no scalar form (Collision.cpp:836-846)
    from aurora_pic import hip
    ion = ctx.add_species(1.0, 1000.0, 1.0, 64)
    ctx.mcc_load(sid, workloads.synthetic_gas(rows=100), product_species=(sid, ion))
    with pytest.raises(hip.HipError, match="require the 3V collision interface"):
        ctx.mcc_collide(vin[:8], dt, step=0)


def test_second_model_targets_another_species(ctx):
    """Two named models on one context (Simulation.cpp:490-518): electrons
    and ions collide in the same step on their own streams, their
    statistics are kept apart, a third model on a targeted species is
    refused, and a checkpoint carries both."""
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=200, neutral_density=1e22)
    e = np.geomspace(1e-3, 1e3, 100) * workloads.EV
    ion_gas = dict(channels=[(workloads.ELASTIC, 0.0, e, np.full(100, 5e-19)),
                             (4, 0.0, e, np.full(100, 3e-19))],
                   neutral_density=1e22, max_frequency=1e22 * 8e-19 * 1.5e5,
                   neutral_mass=6.67e-27)          # resonant: the ion's own mass
    c = hip.Context(0)
    try:
        c.mesh([65], [0.067], hip.DIRICHLET, workloads.EPS0)
        c.electrodes([0, 0, 0, 0], [0.0, 200.0, 13.56e6, 0.0])
        se = c.add_species(-workloads.EV, workloads.ME, 1e9, 20000)
        si = c.add_species(workloads.EV, 6.67e-27, 1e9, 20000)
        # hot electrons: ionization fires and its ions join the species that
        # the second model collides
        c.init_maxwellian(se, 4000, math.sqrt(20 * workloads.EV / workloads.ME), seed=1)
        c.init_maxwellian(si, 4000, math.sqrt(0.026 * workloads.EV / 6.67e-27), seed=2)
        c.mcc_load(se, gas, product_species=(se, si), seed=5)
        k = c.mcc_add(si, ion_gas, seed=5, neutral_temperature=300.0)
        assert k == 1
        with pytest.raises(hip.HipError, match="multiple collision models target species"):
            c.mcc_add(si, ion_gas)
        dt = 1.0 / (400 * 13.56e6)
        c.initialize(dt)
        c.step(dt, 50)
        ce, _ = c.mcc_model_stats(0)
        ci, ei = c.mcc_model_stats(1)
        assert ce[2] > 0 and ce[4] > 0          # electron elastic and ionization
        assert ci[2] > 0 and ci[3] > 0          # ion elastic and charge exchange
        assert np.array_equal(c.mcc_stats()[0], ce)
        live_i = c.count(si)[1]
        path = os.path.join(tempfile.mkdtemp(), "two_models.ckpt")
        c.checkpoint_save(path)
        c.step(dt, 10)
        after = c.mcc_model_stats(1)[0]
        c.checkpoint_load(path)
        assert np.array_equal(c.mcc_model_stats(1)[0], ci)
        assert c.count(si)[1] == live_i
        c.step(dt, 10)
        assert np.array_equal(c.mcc_model_stats(1)[0], after)   # the restart repeats
    finally:
        c.close()


def test_1d_step_refuses_attachment():
    """The 1D step refuses an attaching channel with the reference's message
    (Simulation.cpp:437-439); the standalone collide keeps it."""
    from aurora_pic import hip
    e = np.geomspace(1e-3, 1e3, 50) * workloads.EV
    gas = dict(channels=[(workloads.ELASTIC, 0.0, e, np.full(50, 1e-20)),
                         (workloads.ATTACHMENT, 0.0, e, np.full(50, 1e-21))],
               neutral_density=1e21, max_frequency=workloads.max_collision_frequency(
                   [(0, 0.0, e, np.full(50, 1.1e-20))], 1e21, margin=1.5),
               neutral_mass=0.0)
    c = hip.Context(0)
    try:
        c.mesh([33], [0.01], hip.DIRICHLET, workloads.EPS0)
        se = c.add_species(-workloads.EV, workloads.ME, 1e6, 1000)
        c.init_maxwellian(se, 100, math.sqrt(2 * workloads.EV / workloads.ME), seed=3)
        c.mcc_load(se, gas, seed=4)
        c.initialize(1e-12)
        with pytest.raises(hip.HipError, match="1D MCC does not support attachment"):
            c.step(1e-12, 1)
    finally:
        c.close()


def test_1d_step_capacity_error_with_counts_on_device():
    """A 1D MCC step whose products overflow max_particles_per_species fails
    with the reference's message (Simulation.cpp:929-964) when the slot
    counts are kept on the device for the call: the device-side commit
    checks every model's products at once and places none of that step's,
    and the host's counts come back within the capacity.  With room, the
    same run places one secondary electron and one ion per ionization."""
    from aurora_pic import hip
    cfg = workloads.config_discharge()
    hot = math.sqrt(100.0 * workloads.EV / workloads.ME)   # ionizing electrons

    def run(extra, steps):
        c = hip.Context(0)
        try:
            c.mesh([65], cfg["L"], hip.DIRICHLET, cfg["eps"])
            n = 4000
            se = c.add_species(-workloads.EV, workloads.ME, 1e6, n + extra)
            si = c.add_species(workloads.EV, 6.67e-27, 1e6, n + extra)
            c.init_maxwellian(se, n, hot, seed=11)
            c.init_maxwellian(si, n, workloads.maxwellian_sigma(0.026, 6.67e-27), seed=12)
            c.mcc_load(se, cfg["gas"], product_species=(se, si), seed=13)
            c.initialize(cfg["dt"])
            err = None
            try:
                c.step(cfg["dt"], steps)
            except hip.HipError as e:
                err = e
            counts = [c.count(s) for s in (se, si)]
            stats = c.mcc_stats()
            return err, counts, stats, n
        finally:
            c.close()

    err, counts, stats, n = run(2, 400)
    assert err is not None and "reactive collisions exceeded max_particles_per_species" in str(err)
    assert all(slots <= n + 2 for slots, _ in counts)
    err, counts, stats, n = run(100000, 200)
    assert err is None
    ionizations = int(stats[0][2 + 2])
    assert ionizations > 0
    assert counts[0][0] == n + ionizations and counts[1][0] == n + ionizations


def test_1d_mcc_steps_with_sorts_inside_one_call():
    """Sorts inside one aurora_hip_step call (the slot counts leave the
    device for the sort and return after it) give the particles, counts and
    collision statistics of the same steps taken one call at a time."""
    from aurora_pic import hip
    cfg = workloads.config_discharge()
    hot = math.sqrt(30.0 * workloads.EV / workloads.ME)

    def run(calls):
        c = hip.Context(0)
        try:
            c.mesh([65], cfg["L"], hip.DIRICHLET, cfg["eps"])
            n = 6000
            se = c.add_species(-workloads.EV, workloads.ME, 1e6, 4 * n)
            si = c.add_species(workloads.EV, 6.67e-27, 1e6, 4 * n)
            c.init_maxwellian(se, n, hot, seed=21)
            c.init_maxwellian(si, n, workloads.maxwellian_sigma(0.026, 6.67e-27), seed=22)
            c.mcc_load(se, cfg["gas"], product_species=(se, si), seed=23)
            c.initialize(cfg["dt"])
            c.sort_interval(7)
            for k in calls:
                c.step(cfg["dt"], k)
            out = []
            for s in (se, si):
                p, v, a, u = c.download(s)
                o = np.argsort(u)
                out.append((c.count(s), p[o], v[o], a[o], u[o]))
            return out, c.mcc_stats()[0]
        finally:
            c.close()

    one, stats_one = run([30])
    many, stats_many = run([1] * 30)
    assert np.array_equal(stats_one, stats_many)
    assert stats_one[2 + 2] > 0          # ionizations happened
    for a, b in zip(one, many):
        assert a[0] == b[0]
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(x, y)
