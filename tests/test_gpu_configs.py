"""End-to-end runs of the measured 1D and 2D configurations on the GPU.

- Two-stream 1D1V (FP64 particles, 256-node periodic grid, 25,600
  electrons at +-3 v_th): steps against the FP64 oracle from the same
  state, then the full 2,000-step run with its energy invariant.
- Langmuir 2D3V with uniform B_z (FP32 particles, FFT Poisson): steps
  against the oracle on a small mesh, then the full 256 x 256 x 64-per-cell
  configuration for a bounded number of steps with its invariants.

Tolerances: FP64 particles agree with the FP64 oracle to 1e-9 relative
(the only difference is the fixed-point deposit, 2^-41 of a particle per
contribution); FP32 particles to 1e-5 over 10 steps."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _two_stream_state(N, L, rng):
    k = 2 * math.pi / L
    x = rng.uniform(0, L, N)
    x = (x + 1e-3 * np.sin(k * x)) % L
    v = np.where(np.arange(N) % 2 == 0, 3.0, -3.0) + rng.normal(0, 0.1, N)
    vel = np.zeros((N, 3))
    vel[:, 0] = v
    return x, vel


def test_two_stream_fp64_steps_match_oracle(oracle):
    from aurora_pic import hip
    n, L, N, dt = 256, 4 * math.pi, 25600, 0.1
    q, m, w = -1.0, 1.0, L / N
    x, vel = _two_stream_state(N, L, np.random.default_rng(12345))
    with hip.Context(0) as c:
        c.mesh([n], [L], hip.PERIODIC, 1.0)
        sid = c.add_species(q, m, w, N, precision=64)
        c.upload(sid, x, vel, kind=hip.VELOCITY_CENTRED)
        c.initialize(dt)
        pos = x.copy()
        alive = np.ones(N, dtype=np.uint8)
        rho = oracle.deposit_cic([n], [L], 0, pos, q, w, alive)
        _, E = oracle.poisson([n], [L], 0, 1.0, rho)
        vh = oracle.init_half([n], [L], 0, pos, vel, alive, E, q / m, dt)
        steps = 40
        c.step(dt, steps)
        for _ in range(steps):
            pos, vh, alive, _ = oracle.push([n], [L], 0, pos, vh, alive, E, q / m, dt)
            pos = pos.reshape(-1)
            rho = oracle.deposit_cic([n], [L], 0, pos, q, w, alive)
            _, E = oracle.poisson([n], [L], 0, 1.0, rho)
        _, _, Eg = c.fields()
        assert np.max(np.abs(Eg[0] - E)) <= 1e-9 * np.max(np.abs(E))
        p, v, a, u = c.download(sid)
        d = (p[:, 0] - pos + L / 2) % L - L / 2
        assert np.max(np.abs(d)) <= 1e-9 * L
        assert np.max(np.abs(v[:, 0] - vh[:, 0])) <= 1e-9 * 3.0


def test_two_stream_full_run_conserves_energy():
    """The 2,000-step two-stream run: total energy (kinetic at the
    synchronised velocity + field) stays within 1 %, no particle is lost."""
    from aurora_pic import hip
    n, L, N, dt = 256, 4 * math.pi, 25600, 0.1
    x, vel = _two_stream_state(N, L, np.random.default_rng(12345))
    with hip.Context(0) as c:
        c.mesh([n], [L], hip.PERIODIC, 1.0)
        sid = c.add_species(-1.0, 1.0, L / N, N, precision=64)
        c.upload(sid, x, vel, kind=hip.VELOCITY_CENTRED)
        c.initialize(dt)
        ke0, fe0, _, live0 = c.energy(dt)
        c.step(dt, 2000)
        ke1, fe1, _, live1 = c.energy(dt)
    e0, e1 = ke0.sum() + fe0, ke1.sum() + fe1
    assert live0 == live1 == N
    assert abs(e1 - e0) <= 0.01 * e0


def _langmuir_state(cells, ppc, L, vth, rng):
    N = cells * cells * ppc
    k = 2 * math.pi / L
    # 1 % cos(kx) density perturbation by rejection on x
    xs = []
    while sum(len(a) for a in xs) < N:
        cand = rng.uniform(0, L, 2 * N)
        keep = rng.uniform(0, 1.01, 2 * N) < 1.0 + 0.01 * np.cos(k * cand)
        xs.append(cand[keep])
    x = np.concatenate(xs)[:N]
    y = rng.uniform(0, L, N)
    vel = rng.normal(0, vth, (N, 3))
    return np.stack([x, y], axis=1), vel


def test_langmuir_2d_steps_match_oracle(oracle):
    from aurora_pic import hip
    cells, ppc, L, dt = 32, 16, 32.0, 0.05
    pos, vel = _langmuir_state(cells, ppc, L, 0.5, np.random.default_rng(4242))
    N = pos.shape[0]
    q, m, w = -1.0, 1.0, L * L / N
    B = [0.0, 0.0, 0.5]
    with hip.Context(0) as c:
        c.mesh([cells, cells], [L, L], hip.PERIODIC, 1.0)
        c.uniform_b(B)
        sid = c.add_species(q, m, w, N)
        c.upload(sid, pos, vel, kind=hip.VELOCITY_CENTRED)
        c.sort()
        c.initialize(dt)
        # the oracle continues from the device's own FP32 state after initialize
        p0, v0, a0, u0 = c.download(sid)
        o = np.argsort(u0)
        ps, vh, alive = p0[o].copy(), v0[o].copy(), a0[o].copy()
        rho = oracle.deposit_cic([cells, cells], [L, L], 0, ps, q, w, alive)
        _, E = oracle.poisson([cells, cells], [L, L], 0, 1.0, rho)
        c.step(dt, 10)
        for _ in range(10):
            ps, vh, alive, _ = oracle.push([cells, cells], [L, L], 0, ps, vh, alive, E, q / m,
                                           dt, B=B)
            rho = oracle.deposit_cic([cells, cells], [L, L], 0, ps, q, w, alive)
            _, E = oracle.poisson([cells, cells], [L, L], 0, 1.0, rho)
        _, _, Eg = c.fields()
        assert np.max(np.abs(Eg.reshape(E.shape) - E)) <= 1e-5 * np.max(np.abs(E))
        p, v, a, u = c.download(sid)
        o = np.argsort(u)
        d = (p[o] - ps + L / 2) % L - L / 2
        assert np.max(np.abs(d)) <= 1e-5 * L
        assert np.max(np.abs(v[o] - vh)) <= 1e-5 * np.max(np.abs(vh))


def test_langmuir_2d_full_size_invariants():
    """256 x 256 cells, 64 electrons per cell (4,194,304), uniform B_z,
    200 steps with the periodic cell sort: particle count, charge (fixed
    point, exact) and total energy within 1 %."""
    from aurora_pic import hip, workloads as wl
    cfg = wl.config_langmuir_2d()
    s = cfg["species"][0]
    with hip.Context(0) as c:
        c.mesh(cfg["n"], cfg["L"], hip.PERIODIC, cfg["eps"])
        c.uniform_b(cfg["B"])
        sid = c.add_species(s["charge"], s["mass"], s["weight"], s["particles"])
        c.init_maxwellian(sid, s["particles"], wl.maxwellian_sigma(s["temperature_ev"], s["mass"]),
                          cfg["seed"])
        c.sort()
        c.initialize(cfg["dt"])
        c.sort_interval(20)
        q0 = int(c.charge_fixed(sid).astype(np.uint64).sum())
        ke0, fe0, _, live0 = c.energy(cfg["dt"])
        c.step(cfg["dt"], 200)
        ke1, fe1, _, live1 = c.energy(cfg["dt"])
        q1 = int(c.charge_fixed(sid).astype(np.uint64).sum())
    assert live0 == live1 == s["particles"]
    assert q0 == q1
    assert abs((ke1.sum() + fe1) - (ke0.sum() + fe0)) <= 0.01 * (ke0.sum() + fe0)
