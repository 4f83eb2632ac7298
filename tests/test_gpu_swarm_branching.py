"""Branching / resampled electron swarms on the device
(aurora_hip_swarm_branching; Swarm.cpp:940-1250 with population_model =
branching_resampled).

- tests/golden/swarm_branching/branching.swarm through the readers and the
  device: drift, mean energy, growth rate, ionization and attachment rates,
  Townsend coefficients and counts within 4 sigma of the reference's own
  runs over 16 seeds (60 Td attachment only, 300 Td mixed, 1000 Td
  ionization dominated); the population stays at `particles`.
- The device result in the reference's CSV schema: every column describing
  the run equals the reference's own branching CSV.
- A run repeats to the bit.
- The reference's errors: the population limit, extinction (one electron, a
  strong attachment cross section), attachment refused by the fixed model,
  and the branching-model configuration checks."""
import csv
import os

import numpy as np
import pytest

from aurora_pic import gasdata

pytestmark = pytest.mark.gpu


@pytest.fixture
def bctx():
    from aurora_pic import hip
    from test_branching_swarm import branching_setup
    cfg, ds, gas, _ = branching_setup()
    c = hip.Context(0)
    c.swarm_gas(gas)
    gasdata.apply_channel_options(c, ds, model=1)
    yield c, cfg, ds, gas
    c.close()


def _run(c, cfg, limit=None):
    limit = gasdata.effective_population_limit(cfg) if limit is None else limit
    return [c.swarm_branching(q, limit) for q in gasdata.swarm_requests(cfg)]


def test_branching_example_matches_reference(bctx):
    from test_branching_swarm import check_branching_against_reference
    c, cfg, ds, gas = bctx
    out = _run(c, cfg)
    results = []
    for r, g in out:
        assert r.status == 0 and g.final_particles == cfg.particles
        assert g.initial_weight == cfg.particles
        results.append(dict(drift=r.drift, mean_energy_ev=r.mean_energy_ev,
                            growth_rate=g.growth_rate, ionization_rate=r.ionization_rate,
                            attachment_rate=g.attachment_rate,
                            net_creation_rate=g.net_creation_rate,
                            townsend_growth=g.townsend_growth,
                            townsend_rate_balance=g.townsend_rate_balance,
                            collisions0=r.collisions[0], candidates=r.candidates))
    check_branching_against_reference(results)
    assert results[0]["growth_rate"] < 0 < results[1]["growth_rate"] < results[2]["growth_rate"]
    # the total weight follows the growth: W grows at 1000 Td, shrinks at 60 Td
    assert out[2][1].final_weight > 4 * cfg.particles
    assert out[0][1].final_weight < cfg.particles


def test_branching_csv_matches_reference_columns(bctx, tmp_path):
    from aurora_pic import swarmio
    from test_branching_swarm import _reference_rows
    c, cfg, ds, gas = bctx
    out = _run(c, cfg)
    path = str(tmp_path / "branching.csv")
    swarmio.write_swarm_benchmark_csv(path, cfg, ds, [r for r, _ in out], [g for _, g in out])
    head, ref = _reference_rows()
    with open(path) as f:
        mine = list(csv.reader(f))
    assert mine[0] == head
    describe = head[:head.index("collision_model_signature") + 2] + [
        "diffusion_available", "longitudinal_diffusion_m2_s", "transverse_diffusion_m2_s",
        "initial_total_electron_weight", "final_computational_particles", "townsend_available",
        "spatial_townsend_available"]
    for got, want in zip(mine[1:], ref):
        for name in describe:
            i = head.index(name)
            a, b = got[i], want[i]
            if name in ("gas_data_file", "spatial_profile_file"):
                a, b = os.path.basename(a), os.path.basename(b)
            assert a == b, (name, a, b)


def test_branching_run_repeats_to_the_bit(bctx):
    c, cfg, ds, gas = bctx
    q = gasdata.swarm_requests(cfg)[1]
    lim = gasdata.effective_population_limit(cfg)
    (r1, g1), (r2, g2) = c.swarm_branching(q, lim), c.swarm_branching(q, lim)
    assert bytes(r1) == bytes(r2) and bytes(g1) == bytes(g2)


def test_branching_errors(bctx):
    from aurora_pic import hip
    from test_branching_swarm import strong_attachment_gas
    c, cfg, ds, gas = bctx
    qs = gasdata.swarm_requests(cfg)
    with pytest.raises(hip.HipError, match="exceeded population_limit before resampling"):
        c.swarm_branching(qs[2], cfg.particles + 1)
    with pytest.raises(hip.HipError, match="population_limit must exceed particles"):
        c.swarm_branching(qs[2], cfg.particles)
    bad = gasdata.swarm_requests(cfg)[1]
    bad.blocks = 2000                          # one sampling step per block
    with pytest.raises(hip.HipError, match="at least two sampling steps"):
        c.swarm_branching(bad, 10_000)
    with pytest.raises(hip.HipError, match="fixed-population swarm does not support attachment"):
        c.swarm_submit(qs[0])
    # one electron and a strong attachment cross section: extinction
    c.swarm_gas(strong_attachment_gas(gas))
    one = gasdata.swarm_requests(cfg)[0]
    one.particles, one.reduced_field_td = 1, 0.0
    one.steps, one.burn_in, one.blocks = 1000, 500, 2
    with pytest.raises(hip.HipError, match="electron ensemble became extinct"):
        c.swarm_branching(one, 10_000)


def test_branching_needs_a_creation_channel():
    from aurora_pic import hip
    from test_branching_swarm import branching_setup
    cfg, ds, gas, _ = branching_setup()
    c = hip.Context(0)
    c.swarm_gas(dict(gas, channels=[ch for ch in gas["channels"] if ch[0] in (0, 1)]))
    with pytest.raises(hip.HipError, match="requires an ionization or attachment channel"):
        c.swarm_branching(gasdata.swarm_requests(cfg)[0], 10_000)
    c.close()
