"""Continuous-batching request loop for electron-swarm solves.

Requests arrive over time (each with an arrival timestamp).  The loop admits
an arrived request as soon as its HBM footprint fits the budget
(``aurora_hip_swarm_submit`` refuses it with AURORA_ERR_CAPACITY otherwise,
and it stays queued), advances every admitted request together by a bounded
number of steps per launch (``aurora_hip_swarm_launch``: one persistent
kernel over the chunks of all active requests, queued without waiting, so
other device work and admissions overlap it), and collects results as
requests finish.  State stays on the device from admission to the result.

A request's result does not depend on what it is batched with (integer
statistics, per-request RNG streams), so the batched answer equals the one a
sequential single-request run gives, bit for bit; the loop reports
per-request latency (arrival -> result) as p50/p99 beside throughput.

The reference's swarm driver (run_swarm_benchmark, Swarm.cpp:1513-1716)
runs field points one after another on the CPU; seeds follow its rule
``seed = base + index`` (Swarm.cpp:1513-1515).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import hip


@dataclass
class Served:
    index: int
    arrival_s: float
    request: object
    admit_s: float = float("nan")
    done_s: float = float("nan")
    result: object = None
    refusals: int = 0          # admission attempts refused for lack of HBM budget
    rejected: str = ""         # why the request was never admitted (empty when served)

    @property
    def latency_s(self):
        return self.done_s - self.arrival_s


@dataclass
class ServingReport:
    served: list
    wall_s: float
    electron_steps: int
    launches: int
    max_active: int
    refused_for_budget: int
    timeline: list = field(default_factory=list)

    def latencies(self):
        return np.array([s.latency_s for s in self.served if not s.rejected])

    def rejected(self):
        return [s for s in self.served if s.rejected]

    def summary(self):
        lat = self.latencies()
        return {
            "requests": len(self.served),
            "wall_s": self.wall_s,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "requests_per_s": len(self.served) / self.wall_s,
            "electron_steps_per_s": self.electron_steps / self.wall_s,
            "launches": self.launches,
            "max_active": self.max_active,
            "budget_refusals": self.refused_for_budget,
            "rejected": len(self.rejected()),
        }


def make_requests(cfg, count=None, steps=None, burn_in=None, particle_scale=1.0):
    """SwarmRequest objects for a config_swarm_requests() workload
    (optionally truncated / shortened for a bounded run)."""
    n = len(cfg["arrival_s"]) if count is None else count
    steps = cfg["steps"] if steps is None else steps
    burn = cfg["burn_in"] if burn_in is None else burn_in
    out = []
    for i in range(n):
        parts = max(256, int(round(cfg["particles"][i] * particle_scale)))
        out.append(hip.SwarmRequest(
            reduced_field_td=float(cfg["reduced_field_td"][i]), timestep=cfg["timestep"],
            particles=parts, steps=steps, burn_in=burn, blocks=cfg["blocks"],
            initial_mean_energy_ev=cfg["initial_mean_energy_ev"],
            max_energy_ev=cfg["max_energy_ev"], seed=cfg["seed"] + i))
    return out


WORK_ITEM_LIMIT = 100_000_000      # SwarmBenchmarkConfig::work_item_limit default (Swarm.hpp:42)


def work_items_exceeded(request, work_item_limit=WORK_ITEM_LIMIT):
    """The reference's work-item gate for one field point (Swarm.cpp:340-359:
    fields x particles x steps, here 1 x particles x steps): the message it
    raises when the request exceeds work_item_limit, else ""."""
    if work_item_limit <= 0:
        return "swarm work_item_limit must be positive"
    if request.particles > 0 and request.steps > work_item_limit // request.particles:
        return ("swarm scan exceeds work_item_limit; raise it explicitly only on an "
                "appropriate compute host")
    return ""


def serve(ctx, requests, arrivals_s, hbm_budget, steps_per_launch=1000, clock=time.perf_counter,
          on_launch=None, work_item_limit=WORK_ITEM_LIMIT, progress=None, progress_every_s=30.0,
          max_active_particles=None):
    """Run the request loop until every request has a result.

    ctx: hip.Context with the swarm gas loaded.  arrivals_s: arrival time of
    each request relative to the start of the loop (non-decreasing).
    A request whose own footprint exceeds hbm_budget can never be admitted:
    it is rejected on arrival (Served.rejected says why) instead of waiting.
    Admission is first come, first served: a request waiting for memory
    holds back the requests that arrived after it.  A request above
    work_item_limit (particles x steps) is rejected with the reference's
    message; the default is the reference's 1e8, and a GPU deployment raises
    it explicitly.  max_active_particles, when given, caps the electrons of
    the admitted requests together (a request larger than the cap runs
    alone): once the GPU is full, admitting more only shares it more thinly,
    so a cap of a few GPU-fulls keeps service first come, first served and
    the latency of early arrivals short, at the same throughput.

    Swarm launches are queued without waiting (aurora_hip_swarm_launch), so
    while one runs the loop keeps admitting arrivals (they join the next
    launch) and calls on_launch, e.g. to step a 1D ensemble on its own
    context and stream beside the swarms; it then collects the launch and
    the requests it finished.  progress(served, total, seconds), when
    given, is called about every progress_every_s seconds.  Returns a
    ServingReport."""
    order = np.argsort(np.asarray(arrivals_s), kind="stable")
    pending = collections.deque(Served(int(i), float(arrivals_s[i]), requests[i]) for i in order)
    queued = collections.deque()
    active = {}
    served = []
    launches = max_active = refused = 0
    electron_steps = 0
    in_flight = False
    t0 = clock()
    next_report = progress_every_s

    def admit():
        nonlocal refused
        now = clock() - t0
        while pending and pending[0].arrival_s <= now:
            s = pending.popleft()
            fp = ctx.swarm_footprint(s.request)
            s.rejected = work_items_exceeded(s.request, work_item_limit)
            if not s.rejected and hbm_budget and fp > hbm_budget:
                s.rejected = f"footprint {fp} B exceeds the HBM budget of {hbm_budget} B"
            if s.rejected:
                s.done_s = now
                served.append(s)
                continue
            queued.append(s)
        while queued:
            s = queued[0]
            if max_active_particles is not None and active and \
                    sum(a.request.particles for a in active.values()) + s.request.particles > \
                    max_active_particles:
                break
            try:
                h = ctx.swarm_submit(s.request, hbm_budget=hbm_budget)
            except hip.HipError as e:
                if e.code != hip.ERR_CAPACITY:
                    raise
                if not active:
                    # nothing holds memory, yet the request does not fit
                    raise RuntimeError(f"request {s.index} refused with no request running: {e}")
                refused += 1
                s.refusals += 1
                break
            s.admit_s = clock() - t0
            active[h] = s
            queued.popleft()

    while pending or queued or active:
        admit()
        if not in_flight:
            runnable = [s for s in active.values() if getattr(s, "_steps", 0) < s.request.steps]
            if not runnable:
                if pending and not active:
                    wait = pending[0].arrival_s - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                continue
            max_active = max(max_active, len(runnable))
            for s in runnable:
                done_steps = getattr(s, "_steps", 0)
                k = min(steps_per_launch, s.request.steps - done_steps)
                s._steps = done_steps + k
                electron_steps += k * s.request.particles
            ctx.swarm_launch(steps_per_launch)
            launches += 1
            in_flight = True
        if on_launch is not None:
            on_launch()
        in_flight = ctx.swarm_collect(wait=on_launch is None)
        if in_flight:
            continue
        t_done = clock() - t0
        if progress is not None and t_done >= next_report:
            progress(len(served), len(requests), t_done)
            next_report = t_done + progress_every_s
        for h in list(active):
            done, r = ctx.swarm_poll(h)
            if done:
                s = active.pop(h)
                s.done_s = t_done
                s.result = r
                served.append(s)
    wall = clock() - t0
    served.sort(key=lambda s: s.index)
    return ServingReport(served, wall, electron_steps, launches, max_active, refused)


def run_sequential(ctx, requests, steps_per_launch=None):
    """Each request alone on the GPU, one after another (the single-instance
    baseline).  Returns (results, wall seconds, electron-steps)."""
    results = []
    steps_total = 0
    t0 = time.perf_counter()
    for r in requests:
        h = ctx.swarm_submit(r)
        while True:
            ctx.swarm_advance(steps_per_launch or r.steps)
            done, res = ctx.swarm_poll(h)
            if done:
                break
        results.append(res)
        steps_total += r.steps * r.particles
    return results, time.perf_counter() - t0, steps_total


def discharge_ensemble(instances=64, particles=100_000, capacity_factor=3, seed=2026,
                       graph_steps=50):
    """Distinct-seed instances of the 1D3V RF discharge (513 nodes, 0.067 m,
    200 V at 13.56 MHz, synthetic gas at 1e21 m^-3) with `particles` per
    instance (half electrons, half ions), as one batched ensemble.  Returns
    (context, ensemble, dt); every instance is initialised on the context
    (Maxwellian loading, t = 0 field, half step) and loaded."""
    from . import workloads as wl
    cfg = wl.config_discharge()
    n_each = particles // 2
    L = cfg["L"][0]
    weight = 2.56e14 * L / n_each
    ctx = hip.Context(0)
    ctx.mesh(cfg["n"], cfg["L"], hip.DIRICHLET, cfg["eps"])
    ctx.electrodes([0.0, 0.0, 0.0, 0.0], [0.0, cfg["phi_right_amplitude"], cfg["frequency"], 0.0])
    cap = capacity_factor * n_each
    sids = [ctx.add_species(s["charge"], s["mass"], weight, cap) for s in cfg["species"]]
    ctx.mcc_load(sids[0], cfg["gas"], product_species=(sids[0], sids[1]), seed=seed)
    ens = ctx.ensemble_create(instances, [cap, cap])
    ens.graph(graph_steps)
    for k in range(instances):
        for sid, s in zip(sids, cfg["species"]):
            ctx.init_maxwellian(sid, n_each, wl.maxwellian_sigma(s["temperature_ev"], s["mass"]),
                                seed=seed + 7919 * k + 2 * sid + 1)
        ctx.initialize(cfg["dt"])
        ens.load(k, seed + 1000 + k)
    return ctx, ens, cfg["dt"]
