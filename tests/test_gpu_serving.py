"""The continuous-batching request loop on the GPU (serving.serve): Poisson
arrivals of swarm requests of 1e4-1e5 electrons admitted against an HBM
budget while others run, mixed with a running 1D3V discharge ensemble.
Every served result equals the same request run alone, bit for bit; the
ensemble's instances equal the same ensemble stepped with no swarms beside
it; admission refuses requests while the budget is full and admits them as
memory frees up; a request that can never fit is rejected, not retried."""
import numpy as np
import pytest

from aurora_pic import serving, workloads

pytestmark = pytest.mark.gpu

FIELDS = ("drift", "drift_se", "mean_energy_ev", "mean_energy_se", "ionization_rate",
          "ionization_rate_se", "candidates", "nulls", "max_energy_ev",
          "longitudinal_diffusion", "transverse_diffusion")


def _requests(count, steps, seed=5):
    cfg = workloads.config_swarm_requests(count, seed=seed)
    cfg = dict(cfg, particles=np.clip(cfg["particles"], 10_000, 100_000))
    reqs = serving.make_requests(cfg, count=count, steps=steps, burn_in=steps // 2)
    return cfg, reqs


def _same(a, b):
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        if not (x == y or (np.isnan(x) and np.isnan(y))):
            return False
    return list(a.collisions) == list(b.collisions) and a.status == b.status == 0


def test_serve_mixed_with_ensemble_equals_lone_runs():
    from aurora_pic import hip
    count, steps = 16, 600
    cfg, reqs = _requests(count, steps)
    arrivals = np.cumsum(np.random.default_rng(3).exponential(1.0 / 2000.0, size=count))
    big = hip.SwarmRequest(reduced_field_td=100.0, timestep=cfg["timestep"], particles=400_000,
                           steps=steps, burn_in=steps // 2, blocks=cfg["blocks"],
                           initial_mean_energy_ev=1.0, max_energy_ev=1000.0, seed=77)
    reqs.append(big)
    arrivals = np.append(arrivals, arrivals[3])
    ctx = hip.Context(0)
    ctx.swarm_gas(cfg["gas"])
    fps = sorted(ctx.swarm_footprint(r) for r in reqs[:count])
    budget = sum(fps[-3:])                     # the three largest at a time
    assert ctx.swarm_footprint(big) > budget
    ens_ctx, ens, dt = serving.discharge_ensemble(instances=3, particles=6000, graph_steps=10)
    ticks = [0]

    def tick():
        ens.step(dt, 10)
        ticks[0] += 1

    rep = serving.serve(ctx, reqs, arrivals, budget, steps_per_launch=150, on_launch=tick,
                        work_item_limit=10 ** 12)
    print("serving:", rep.summary(), flush=True)
    served = {s.index: s for s in rep.served}
    assert len(served) == count + 1
    assert served[count].rejected and "exceeds the HBM budget" in served[count].rejected
    assert rep.refused_for_budget > 0          # requests waited for memory ...
    assert all(s.result is not None for i, s in served.items() if i < count)   # ... and ran
    assert rep.max_active >= 2                 # requests ran together
    assert np.all(rep.latencies() > 0)
    lone, _, _ = serving.run_sequential(ctx, reqs[:count])
    for i in range(count):
        assert _same(served[i].result, lone[i]), i
    ctx.close()
    # the ensemble advanced beside the swarms exactly as it does alone
    steps_run = 10 * ticks[0]
    assert steps_run > 0
    ref_ctx, ref_ens, _ = serving.discharge_ensemble(instances=3, particles=6000, graph_steps=10)
    ref_ens.step(dt, steps_run)
    for k in range(3):
        for sid in range(2):
            a, b = ens.download(k, sid), ref_ens.download(k, sid)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
    ens_ctx.close()
    ref_ctx.close()


def test_serve_rejects_work_above_the_item_limit():
    from aurora_pic import hip
    cfg, reqs = _requests(3, 400)
    ctx = hip.Context(0)
    ctx.swarm_gas(cfg["gas"])
    limit = sorted(r.particles for r in reqs)[1] * 400    # the largest request is above it
    over = [i for i, r in enumerate(reqs) if r.particles * r.steps > limit]
    assert len(over) == 1
    rep = serving.serve(ctx, reqs, np.zeros(3), 0, steps_per_launch=400,
                        work_item_limit=limit)
    for s in rep.served:
        if s.index in over:
            assert "swarm scan exceeds work_item_limit" in s.rejected
        else:
            assert s.result is not None
    ctx.close()


def test_active_electron_cap_serves_first_come_first_served():
    """With a cap on the admitted electrons, requests wait in arrival order
    and never run more than the cap allows together; results still equal
    the lone runs."""
    from aurora_pic import hip
    count, steps = 8, 400
    cfg, reqs = _requests(count, steps, seed=9)
    ctx = hip.Context(0)
    ctx.swarm_gas(cfg["gas"])
    cap = max(r.particles for r in reqs) * 2
    rep = serving.serve(ctx, reqs, np.zeros(count), 0, steps_per_launch=100,
                        work_item_limit=10 ** 12, max_active_particles=cap)
    print("capped:", rep.summary(), flush=True)
    assert 1 <= rep.max_active < count
    served = sorted(rep.served, key=lambda s: s.index)
    admits = [s.admit_s for s in served]
    assert admits == sorted(admits)            # admitted in arrival order
    lone, _, _ = serving.run_sequential(ctx, reqs)
    for i in range(count):
        assert _same(served[i].result, lone[i]), i
    ctx.close()
