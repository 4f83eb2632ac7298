"""Branching / resampled electron swarms (Swarm.cpp:940-1250 with
population_model = branching_resampled) on the CPU.

- The oracle's systematic resample of M equal-weight electrons to N, in
  closed form, against the reference's sequential cumulative-weight walk
  (Swarm.cpp:616-643) run here in exact rational arithmetic, for every u on
  a grid and every (M, N) up to 24.
- The oracle's block slope (Swarm.cpp:136-197) against numpy's least-squares
  fit per block, and on an exact line.
- effective_population_limit (Swarm.cpp:240-262) on its branches.
- The oracle's branching run of tests/golden/swarm_branching/branching.swarm
  (a synthetic attaching and ionizing gas: 60 Td attachment only, 300 Td
  mixed, 1000 Td ionization dominated) against the reference's own runs over
  16 seeds (branching_golden.npz): drift, mean energy, growth rate,
  ionization and attachment rates, Townsend coefficients within 4 sigma.
- The swarm CSV writer on the reference's own branching CSV, cell for cell.
"""
import csv
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle_py
from aurora_pic import gasdata

HERE = os.path.dirname(os.path.abspath(__file__))
BR_DIR = os.path.join(HERE, "golden", "swarm_branching")
EV = 1.602176634e-19
KB = 1.380649e-23
ME = 9.1093837139e-31

COMPARED = ("drift", "mean_energy_ev", "growth_rate", "ionization_rate", "attachment_rate",
            "net_creation_rate", "townsend_growth", "townsend_rate_balance", "collisions0",
            "candidates")


def branching_setup():
    cfg = gasdata.load_swarm_config(os.path.join(BR_DIR, "branching.swarm"))
    ds = gasdata.load_gas_dataset(cfg.gas_data_file)
    gas = gasdata.swarm_gas(ds, cfg)
    opts = {k: (c.angular, c.energy_frame, c.mean_cosine_energy, c.mean_cosine)
            for k, c in enumerate(ds.channels)}
    vstd = math.sqrt(KB * cfg.neutral_temperature / ds.neutral_mass)
    return cfg, ds, gas, oracle_py.model_from_gas(gas, neutral_vstd=vstd, options=opts)


def check_branching_against_reference(results, tolerance_sigma=4.0):
    """results[f]: dict of COMPARED values for field f.  The reference's
    spread over 16 seeds sets the scale; the mean of 16 runs carries its own
    1/sqrt(16) error."""
    g = np.load(os.path.join(BR_DIR, "branching_golden.npz"))
    cols = list(g["columns"])
    mean, std, n = g["seed_mean"], g["seed_std"], int(g["seeds"])
    for f, got in enumerate(results):
        for name in COMPARED:
            c = cols.index(name)
            scale = std[f, c] * math.sqrt(1.0 + 1.0 / n)
            if scale == 0.0:                  # e.g. no ionization at 60 Td
                assert got[name] == mean[f, c], (f, name, got[name])
                continue
            assert abs(got[name] - mean[f, c]) <= tolerance_sigma * scale, \
                (f, name, got[name], mean[f, c], std[f, c])


def _sequential_source(j, M, N, u):
    """The reference's walk with weights 1: output weight M / N, threshold
    (u + j) M / N, first source whose cumulative weight reaches it."""
    w_out = Fraction(M, N)
    threshold = u * w_out + j * w_out
    src, cum = 0, Fraction(1)
    while src + 1 < M and cum < threshold:
        src += 1
        cum += 1
    return src


def test_systematic_source_matches_sequential_walk(oracle):
    for M in range(1, 25):
        for N in range(1, 25):
            for k in (1, 2 ** 22, 2 ** 23 - 1, 2 ** 23, 2 ** 23 + 1, 3 * 2 ** 22, 2 ** 24 - 1):
                u = Fraction(k, 2 ** 24)
                for j in range(N):
                    assert oracle.systematic_source(j, M, N, float(u)) == \
                        _sequential_source(j, M, N, u), (M, N, k, j)


def test_block_slope_error(oracle):
    t = np.arange(400) * 2.5e-3
    y = 3.0 - 7.5 * t
    s, se = oracle.block_slope_error(y, 2.5e-3, 8)
    assert abs(s + 7.5) <= 1e-9 and se <= 1e-9
    rng = np.random.default_rng(1)
    y = np.cumsum(rng.normal(0.3, 1.0, 600))
    s, se = oracle.block_slope_error(y, 0.5, 6)
    slopes = [np.polyfit(np.arange(100) * 0.5, y[b * 100:(b + 1) * 100], 1)[0]
              for b in range(6)]
    assert abs(s - np.mean(slopes)) <= 1e-9 * max(1.0, abs(s))
    assert abs(se - np.std(slopes, ddof=1) / math.sqrt(6)) <= 1e-9 * se
    with pytest.raises(ValueError):
        oracle.block_slope_error(y[:7], 0.5, 4)


def test_effective_population_limit():
    cfg = gasdata.SwarmConfig()
    for particles, want in ((512, 1536), (100, 1124), (4000, 8000),
                            (5_000_000, 5_000_000), (10_000_000, 10_000_000),
                            (4_999_000, 5_001_000)):
        cfg.particles = particles
        assert gasdata.effective_population_limit(cfg) == want, particles
    cfg.population_limit = 777
    assert gasdata.effective_population_limit(cfg) == 777


def oracle_branching_results(oracle, cfg, model, vel0_of):
    out = []
    limit = gasdata.effective_population_limit(cfg)
    for f, td in enumerate(cfg.reduced_fields_td):
        c = dict(neutral_density=cfg.neutral_density, max_frequency=cfg.max_frequency,
                 timestep=cfg.timestep, reduced_field_td=td, steps=cfg.steps,
                 burn_in=cfg.burn_in_steps, particles=cfg.particles,
                 blocks=cfg.uncertainty_blocks,
                 initial_mean_energy_ev=cfg.initial_mean_energy_ev,
                 max_energy_ev=cfg.max_energy_ev)
        r, g = oracle.swarm_branching(model, c, vel0_of(f), cfg.seed + f, limit)
        assert r.status == 0, r.status
        assert g.final_particles == cfg.particles
        out.append(dict(drift=r.drift, mean_energy_ev=r.mean_energy_ev,
                        growth_rate=g.growth_rate, ionization_rate=r.ion_rate,
                        attachment_rate=g.attachment_rate,
                        net_creation_rate=g.net_creation_rate,
                        townsend_growth=g.townsend_growth,
                        townsend_rate_balance=g.townsend_rate_balance,
                        collisions0=r.collisions[0], candidates=r.candidates))
    return out


def test_oracle_branching_matches_reference(oracle):
    cfg, ds, gas, model = branching_setup()
    sigma = math.sqrt(2.0 * cfg.initial_mean_energy_ev * EV / (3.0 * ME))
    results = oracle_branching_results(
        oracle, cfg, model,
        lambda f: np.random.default_rng(cfg.seed + f).normal(0.0, sigma, (cfg.particles, 3)))
    check_branching_against_reference(results)
    # the three regimes: attachment only, mixed, ionization dominated
    assert results[0]["growth_rate"] < 0 < results[1]["growth_rate"] < results[2]["growth_rate"]


def test_oracle_population_limit_and_extinction(oracle):
    cfg, ds, gas, model = branching_setup()
    sigma = math.sqrt(2.0 * cfg.initial_mean_energy_ev * EV / (3.0 * ME))
    vel0 = np.random.default_rng(3).normal(0.0, sigma, (cfg.particles, 3))
    c = dict(neutral_density=cfg.neutral_density, max_frequency=cfg.max_frequency,
             timestep=cfg.timestep, reduced_field_td=1000.0, steps=cfg.steps,
             burn_in=cfg.burn_in_steps, particles=cfg.particles, blocks=cfg.uncertainty_blocks,
             initial_mean_energy_ev=cfg.initial_mean_energy_ev, max_energy_ev=cfg.max_energy_ev)
    r, _ = oracle.swarm_branching(model, c, vel0, 5, cfg.particles + 1)
    assert r.status == 5                      # population limit
    # one electron and an attachment cross section 5000x the table's (the
    # total rate stays below max_frequency): the electron attaches, with a
    # probability near 7 % per step, and the swarm dies out
    strong = oracle_py.model_from_gas(strong_attachment_gas(gas),
                                      neutral_vstd=model_vstd(cfg, ds))
    c.update(reduced_field_td=0.0, particles=1, blocks=2, steps=1000, burn_in=500)
    r, _ = oracle.swarm_branching(strong, c, vel0[:1], 5, 10_000)
    assert r.status == 6                      # extinct


def strong_attachment_gas(gas):
    return dict(gas, channels=[ch if ch[0] != 3 else (ch[0], ch[1], ch[2], np.asarray(ch[3]) * 5e3)
                               for ch in gas["channels"]])


def model_vstd(cfg, ds):
    return math.sqrt(KB * cfg.neutral_temperature / ds.neutral_mass)


def _reference_rows():
    with open(os.path.join(BR_DIR, "branching_reference.csv")) as f:
        rows = list(csv.reader(f))
    return rows[0], rows[1:]


def test_csv_writer_reproduces_reference_branching_csv(tmp_path):
    from types import SimpleNamespace
    from aurora_pic import swarmio
    head, body = _reference_rows()
    cfg, ds, gas, model = branching_setup()
    assert swarmio.header(ds) == head
    chans = [c.name for c in ds.channels]
    results, growth = [], []
    for row in body:
        col = dict(zip(head, row))
        results.append(SimpleNamespace(
            drift=float(col["electron_drift_velocity_m_s"]),
            drift_se=float(col["mean_velocity_x_standard_error_m_s"]),
            mean_energy_ev=float(col["mean_energy_ev"]),
            mean_energy_se=float(col["mean_energy_standard_error_ev"]),
            max_energy_ev=float(col["maximum_observed_energy_ev"]),
            candidates=int(col["collision_candidates"]), nulls=int(col["null_collisions"]),
            collisions=[int(col[f"{c}_collisions"]) for c in chans],
            rate=[float(col[f"{c}_rate_per_electron_s"]) for c in chans],
            rate_se=[float(col[f"{c}_rate_standard_error_s"]) for c in chans],
            longitudinal_diffusion=0.0, transverse_diffusion=0.0))
        growth.append(SimpleNamespace(
            initial_weight=float(col["initial_total_electron_weight"]),
            final_weight=float(col["final_total_electron_weight"]),
            final_particles=int(col["final_computational_particles"]),
            growth_rate=float(col["temporal_growth_rate_s"]),
            growth_rate_se=float(col["temporal_growth_rate_standard_error_s"]),
            townsend_available=col["townsend_available"] == "yes",
            townsend_growth=float(col["growth_over_flux_drift_townsend_1_m"] or 0),
            townsend_growth_se=float(
                col["growth_over_flux_drift_townsend_standard_error_1_m"] or 0),
            townsend_rate_balance=float(col["rate_balance_effective_townsend_1_m"] or 0),
            townsend_rate_balance_se=float(
                col["rate_balance_effective_townsend_standard_error_1_m"] or 0)))
    out = str(tmp_path / "branching.csv")
    swarmio.write_swarm_benchmark_csv(out, cfg, ds, results, growth)
    with open(out) as f:
        mine = list(csv.reader(f))
    assert mine[0] == head
    strip = ("gas_data_file", "spatial_profile_file")
    for got, want in zip(mine[1:], body):
        for name, a, b in zip(head, got, want):
            if name in strip:
                a, b = os.path.basename(a), os.path.basename(b)
            assert a == b, (name, a, b)
