"""Benchmark: particle-updates/s of the 3D3V periodic plasma with tabulated B
(128^3 cells, 16 electrons + 16 ions per cell = 67,108,864 particles,
adaptive cell sort) on one MI355X, with the reference CPU path timed on a
bounded sample beside it.

One step = gather + Boris push (tabulated B_y(x)) + drift + periodic
boundary + fixed-point deposit for every particle, FP64 FFT Poisson solve.
Inputs are generated on the device and resident in HBM before timing.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "aurora-pic_amd"))

# bytes per particle-update of the dominant kernel: read + write of the
# 28-byte SoA state (int32 cell, 3 x FP32 offset, 3 x FP32 velocity)
BYTES_PER_UPDATE = 56
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_threads():
    """Threads the CPU baseline may use: the process's CPU share
    (OMP_NUM_THREADS when set, else the affinity mask)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# Time of the restatement's step over the reference's own Simulation3D::step,
# measured in the build container (8 x Intel Xeon @ 2.70 GHz, one thread,
# 128^3 mesh, 2+2 particles per cell, same tabulated B): 0.895.
PORT_OVER_REFERENCE_TIME = 0.895


def cpu_baseline(cfg, ppc_sample: int):
    """The oracle's FP64 restatement of Simulation3D::step (or_step3d: push,
    deposit and FFT solve serial as in the reference, the synchronisation
    loop on OpenMP threads as the reference's runtime_parallel_for) on the
    same mesh and field with fewer particles per cell: one step on one
    thread, then one on every thread of the host share."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py
    from aurora_pic import workloads as wl

    orc = oracle_py.Oracle()
    cells = cfg["n"][0]
    sample = wl.config_plasma_3d(cells=cells, ppc=ppc_sample)
    sp = [(s["charge"], s["mass"], s["weight"], s["particles"],
           wl.maxwellian_sigma(s["temperature_ev"], s["mass"])) for s in sample["species"]]
    ax, bx, bv = sample["btable"]
    st = orc.plasma3d_init(sample["n"], sample["L"], sample["eps"], sample["dt"], sp,
                           btable=(ax, bx, bv), seed=sample["seed"])
    threads = host_threads()
    rates = {}
    for th in (1, threads):
        t0 = time.perf_counter()
        live = orc.step3d(st, th)
        rates[th] = live / (time.perf_counter() - t0)
        log(f"cpu baseline: {th} thread(s) {rates[th]:.4g} particle-updates/s")
    total = sum(s[3] for s in sp)
    return {"value": rates[threads], "unit": "particle-updates/s", "cores": threads,
            "kind": "port", "single_thread_value": rates[1],
            "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads": threads},
            "port_over_reference_time": PORT_OVER_REFERENCE_TIME,
            "sample": f"FP64 restatement of Simulation3D::step (oracle/pic_oracle.c or_step3d, "
                      f"-O3) on the {cells}^3 mesh with {ppc_sample}+{ppc_sample} particles per "
                      f"cell ({total} particles), tabulated B_y(x), one step on 1 thread and one "
                      f"on {threads} threads (OpenMP on the synchronisation loop, where the "
                      f"reference parallelises); host {cpu_model()}, nproc {os.cpu_count()}"}


def discharge_leg(args):
    """The 1D3V bounded RF discharge with MCC at its full size (513 nodes,
    5e5 electrons + 5e5 ions, FP32, synthetic gas at 1e21 m^-3, ionization
    products): particle-updates/s of the whole step (push, deposit, field
    solve at t + dt, MCC and product placement), after a warm-up."""
    from aurora_pic import hip
    from aurora_pic import workloads as wl

    cfg = wl.config_discharge()
    dt = cfg["dt"]
    ctx = hip.Context(0)
    ctx.mesh(cfg["n"], cfg["L"], hip.DIRICHLET, cfg["eps"])
    ctx.electrodes([0.0, 0.0, 0.0, 0.0], [0.0, cfg["phi_right_amplitude"], cfg["frequency"], 0.0])
    sids = []
    for k, s in enumerate(cfg["species"]):
        sid = ctx.add_species(s["charge"], s["mass"], s["weight"], 4 * s["particles"])
        ctx.init_maxwellian(sid, s["particles"],
                            wl.maxwellian_sigma(s["temperature_ev"], s["mass"]), cfg["seed"] + k)
        sids.append(sid)
    ctx.mcc_load(sids[0], cfg["gas"], product_species=(sids[0], sids[1]), seed=cfg["seed"])
    ctx.initialize(dt)
    ctx.step(dt, args.discharge_warmup)
    ctx.sync()
    live0 = sum(ctx.count(sid)[1] for sid in sids)
    t0 = time.perf_counter()
    ctx.step(dt, args.discharge_steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    live1 = sum(ctx.count(sid)[1] for sid in sids)
    ctx.timing(True)
    ctx.step(dt, 20)
    ctx.sync()
    kernels = {}
    for name in ["push_deposit", "push", "rho", "solve_thomas_1d", "ledger", "work", "mcc",
                 "append_products", "clear_charge"]:
        ms, n = ctx.kernel_time(name)
        if n:
            kernels[name] = {"avg_ms": ms / n, "launches": n}
    ctx.timing(False)
    counts, _ = ctx.mcc_stats()
    ctx.close()
    updates = 0.5 * (live0 + live1) * args.discharge_steps
    # the same step through the batched path with one instance: steps run in
    # HIP-graph-captured segments of 50, with no host round trip per step
    # (an instance equals the context's run: particles, fields and counts
    # bit for bit, ledgers to rounding)
    from aurora_pic import serving
    ectx, ens, edt = serving.discharge_ensemble(instances=1, particles=2 * cfg["species"][0]
                                                ["particles"], capacity_factor=4, graph_steps=50)
    ens.step(edt, args.discharge_warmup)
    ectx.sync()
    e_live0 = sum(ens.count(0, sid)[1] for sid in range(2))
    t0 = time.perf_counter()
    ens.step(edt, args.discharge_steps)
    ectx.sync()
    e_wall = time.perf_counter() - t0
    e_live1 = sum(ens.count(0, sid)[1] for sid in range(2))
    ectx.close()
    e_updates = 0.5 * (e_live0 + e_live1) * args.discharge_steps
    # the batched discharge alone: 64 distinct-seed instances of 1e5
    # particles advanced together in the same graph-captured segments
    mctx, mens, mdt = serving.discharge_ensemble(instances=64, particles=100_000,
                                                 capacity_factor=3, graph_steps=50)
    msteps = 200
    mens.step(mdt, 50)
    mctx.sync()
    m_live0 = sum(mens.count(k, sid)[1] for k in range(64) for sid in range(2))
    t0 = time.perf_counter()
    mens.step(mdt, msteps)
    mctx.sync()
    m_wall = time.perf_counter() - t0
    m_live1 = sum(mens.count(k, sid)[1] for k in range(64) for sid in range(2))
    mctx.close()
    m_updates = 0.5 * (m_live0 + m_live1) * msteps
    return {"workload": "1D3V RF discharge: 513 nodes, 5e5 electrons + 5e5 ions (FP32), 200 V "
                        "at 13.56 MHz, synthetic gas at 1e21 m^-3 with ionization products",
            "particle_updates_per_s": updates / wall,
            "ms_per_step": 1e3 * wall / args.discharge_steps,
            "steps": args.discharge_steps, "warmup": args.discharge_warmup,
            "live_particles": [live0, live1], "ionizations": int(counts[4]),
            "kernels_ms": kernels,
            "graph_path": {"particle_updates_per_s": e_updates / e_wall,
                           "ms_per_step": 1e3 * e_wall / args.discharge_steps,
                           "live_particles": [e_live0, e_live1],
                           "note": "one-instance ensemble: the context's step captured in "
                                   "HIP graphs of 50 steps"},
            "ensemble_64": {"instances": 64, "particles_per_instance": 100_000,
                            "particle_updates_per_s": m_updates / m_wall,
                            "ms_per_step": 1e3 * m_wall / msteps, "steps": msteps,
                            "note": "64 distinct-seed instances advanced together in "
                                    "HIP-graph-captured segments of 50 steps"}}


def other_configs_leg(args):
    """Particle-updates/s of the two smaller configurations, timed the same
    way (state resident, whole steps): [0] the 1D1V two-stream (25,600 FP64
    electrons, 256 periodic nodes, spectral solve, 2,000 steps) and [2] the
    2D3V Langmuir wave (4,194,304 FP32 electrons, 256^2 periodic nodes, FFT,
    uniform B_z, 1,000 steps; LDS-tiled push, adaptive sort)."""
    import numpy as np
    from aurora_pic import hip
    from aurora_pic import workloads as wl

    out = {}
    c0 = wl.config_two_stream()
    rng = np.random.default_rng(c0["seed"])
    s0 = c0["species"][0]
    N = s0["particles"]
    L = c0["L"][0]
    u, spread = c0["beams"]
    x = rng.uniform(0.0, L, N)
    x = np.mod(x + c0["perturbation"] * np.sin(2.0 * np.pi * x / L) * L / (2.0 * np.pi), L)
    v = np.zeros((N, 3))
    v[:, 0] = np.where(np.arange(N) % 2 == 0, u, -u) + rng.normal(0.0, spread, N)
    ctx = hip.Context(0)
    ctx.mesh(c0["n"], c0["L"], hip.PERIODIC, c0["eps"])
    sid = ctx.add_species(s0["charge"], s0["mass"], s0["weight"], N, 64)
    ctx.upload(sid, x[:, None], v, kind=1)
    ctx.initialize(c0["dt"])
    ctx.step(c0["dt"], 100)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(c0["dt"], c0["steps"])
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.close()
    out["two_stream_1d1v"] = {"particles": N, "steps": c0["steps"],
                              "particle_updates_per_s": N * c0["steps"] / wall,
                              "ms_per_step": 1e3 * wall / c0["steps"], "precision": "f64"}
    c2 = wl.config_langmuir_2d()
    s2 = c2["species"][0]
    N2 = s2["particles"]
    Lx, Ly = c2["L"]
    rng = np.random.default_rng(c2["seed"])
    pos = np.stack([rng.uniform(0.0, Lx, N2), rng.uniform(0.0, Ly, N2)], axis=1)
    k = 2.0 * np.pi / Lx
    pos[:, 0] = np.mod(pos[:, 0] + c2["perturbation"] / k * np.sin(k * pos[:, 0]), Lx)
    vel = rng.normal(0.0, wl.maxwellian_sigma(1.0, s2["mass"]), (N2, 3))
    ctx = hip.Context(0)
    ctx.mesh(c2["n"], c2["L"], hip.PERIODIC, c2["eps"], c2["depth"])
    ctx.uniform_b(c2["B"])
    # the LDS-tiled 2D push on 16 x 8 cell tiles, adaptive cell sort
    ctx.tiling(16, 8, 1, 2)
    ctx.push_variant(512, 1)
    ctx.sort_policy(hip.SORT_ADAPTIVE)
    sid = ctx.add_species(s2["charge"], s2["mass"], s2["weight"], N2)
    ctx.upload(sid, pos, vel, kind=1)
    ctx.sort()
    ctx.initialize(c2["dt"])
    ctx.step(c2["dt"], 50)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(c2["dt"], c2["steps"])
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.close()
    out["langmuir_2d3v"] = {"particles": N2, "steps": c2["steps"],
                            "particle_updates_per_s": N2 * c2["steps"] / wall,
                            "ms_per_step": 1e3 * wall / c2["steps"], "precision": "f32",
                            "push": "LDS-tiled 16x8 halo 2, adaptive sort"}
    return out


def sequential_estimate(ctx, reqs, strata, log_each=False):
    """The sequential single-instance baseline of a request stream, from a
    stratified sample: requests ordered by particle count, cut into `strata`
    bins of equal count; each bin's median request runs alone on the GPU
    (through the same kernel) and stands for every request of its bin.
    Returns (estimated seconds for all requests, sample results by index,
    per-bin detail)."""
    from aurora_pic import serving
    order = sorted(range(len(reqs)), key=lambda i: reqs[i].particles)
    bins = [order[k * len(order) // strata:(k + 1) * len(order) // strata]
            for k in range(strata)]
    total, sample, detail = 0.0, {}, []
    for b in bins:
        if not b:
            continue
        i = b[len(b) // 2]
        res, wall, _ = serving.run_sequential(ctx, [reqs[i]])
        sample[i] = res[0]
        total += wall * len(b)
        detail.append({"particles": int(reqs[i].particles), "requests": len(b), "alone_s": wall})
        if log_each:
            log(f"alone: request {i} ({reqs[i].particles} electrons) {wall:.4f} s x {len(b)}")
    return total, sample, detail


def serving_leg(args):
    """Batched swarm serving: Poisson arrivals at 200 req/s (E/N log-uniform
    1-1000 Td, log-normal particle counts, synthetic gas), admitted by HBM
    footprint and advanced together, mixed with a 1D3V discharge ensemble
    that steps on its own context and stream while the swarm launches run
    (latency); plus the capacity of the batched path (every request present
    at t = 0) against the same requests run one at a time, each alone on
    the GPU (every request, or a stratified sample for the full stream)."""
    from aurora_pic import hip, serving
    from aurora_pic import workloads as wl
    import numpy as np

    cfg = wl.config_swarm_requests(4096)
    R = len(cfg["arrival_s"]) if args.serving_full else args.serving_requests
    steps = cfg["steps"] if args.serving_full else args.serving_steps
    burn = steps // 2
    reqs = serving.make_requests(cfg, count=R, steps=steps, burn_in=burn)
    ens_ctx, ens, ens_dt = serving.discharge_ensemble(args.ensemble_instances,
                                                      args.ensemble_particles)
    ens_steps = [0]
    last_tick = [-1e30]

    def ensemble_tick():
        # a steady background load: one block of ensemble steps per
        # --ensemble-interval seconds, whatever the swarm launches' length
        now = time.perf_counter()
        if now - last_tick[0] < args.ensemble_interval:
            return
        last_tick[0] = now
        ens.step(ens_dt, args.ensemble_steps_per_launch)
        ens_steps[0] += args.ensemble_steps_per_launch

    ctx = hip.Context(0)
    ctx.swarm_gas(cfg["gas"])
    budget = args.serving_hbm_gib << 30
    # the reference's 1e8 work-item gate (Swarm.cpp:340-359), raised
    # explicitly for the GPU: a 1e6-electron request of 20,000 steps is 2e10
    work_items = 10 ** 12
    # capacity: every request present at t = 0, advanced together
    def progress(done, total, secs):
        log(f"  {done}/{total} requests served after {secs:.0f} s")

    sat = serving.serve(ctx, reqs, np.zeros(R), budget, steps_per_launch=args.serving_launch_steps,
                        work_item_limit=work_items, progress=progress)
    log(f"capacity: {R} requests in {sat.wall_s:.3f} s batched")
    # the same requests one at a time, each alone on the GPU
    if args.serving_full:
        seq_wall, seq_sample, strata = sequential_estimate(ctx, reqs, 16, log_each=True)
        seq_results = {i: r for i, r in seq_sample.items()}
        seq_kind = "stratified sample of 16 requests (median of each particle-count sixteenth)"
    else:
        res, seq_wall, _ = serving.run_sequential(ctx, reqs)
        seq_results = dict(enumerate(res))
        strata = None
        seq_kind = "every request"
    electron_steps = sum(r.steps * r.particles for r in reqs)
    log(f"sequential ({seq_kind}): {seq_wall:.3f} s")
    log(f"serving {R} swarm requests x {steps} steps (arrivals 200/s, budget "
        f"{args.serving_hbm_gib} GiB, at most {args.serving_active_electrons:.3g} electrons "
        f"admitted) + {args.ensemble_instances} discharge instances")
    t0 = time.perf_counter()
    cap = int(args.serving_active_electrons) or None
    rep = serving.serve(ctx, reqs, cfg["arrival_s"][:R], budget,
                        steps_per_launch=args.serving_launch_steps, on_launch=ensemble_tick,
                        work_item_limit=work_items, progress=progress, max_active_particles=cap)
    wall = time.perf_counter() - t0
    summ = rep.summary()
    ens_updates = ens_steps[0] * args.ensemble_instances * args.ensemble_particles
    log(f"serving: {json.dumps(summ)}; ensemble {ens_steps[0]} steps")
    # every served result against its sequential run (bit-identical by design)
    diffs = []
    for i, r in seq_results.items():
        for b in (rep.served[i].result, sat.served[i].result):
            for f in ("drift", "mean_energy_ev", "ionization_rate"):
                x, y = getattr(r, f), getattr(b, f)
                diffs.append(abs(x - y) / max(abs(x), 1e-300))
    ctx.close()
    ens_ctx.close()
    lat = rep.latencies()
    return {
        "workload": f"{R} electron-swarm requests, Poisson arrivals 200/s, E/N 1-1000 Td "
                    f"log-uniform, particles log-normal median 5e4 [1e4, 1e6], {steps} steps "
                    f"({burn} burn-in), synthetic elastic/excitation/ionization gas at 1e21 "
                    f"m^-3; mixed with {args.ensemble_instances} distinct-seed 1D3V discharge "
                    f"instances of {args.ensemble_particles} particles (HIP-graph ensemble on "
                    f"its own stream, {args.ensemble_steps_per_launch} steps every "
                    f"{args.ensemble_interval} s while swarm launches run)",
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "requests_per_s_offered_load": summ["requests_per_s"],
        "wall_s": wall,
        "max_active": summ["max_active"],
        "active_electron_cap": cap,
        "launches": summ["launches"],
        "budget_refusals": summ["budget_refusals"],
        "capacity_requests_per_s": R / sat.wall_s,
        "capacity_electron_steps_per_s": electron_steps / sat.wall_s,
        "sequential": seq_kind,
        "sequential_strata": strata,
        "sequential_requests_per_s": R / seq_wall,
        "sequential_electron_steps_per_s": electron_steps / seq_wall,
        "batched_over_sequential": seq_wall / sat.wall_s,
        "max_rel_diff_vs_sequential": max(diffs) if diffs else None,
        "ensemble_steps": ens_steps[0],
        "ensemble_particle_updates_per_s": ens_updates / wall,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=16)
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--cpu-ppc", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--sort", choices=("adaptive", "fixed"), default="adaptive",
                    help="adaptive: per-species renewal rule on timed pushes; fixed: every "
                         "20 steps")
    ap.add_argument("--tile", default=None,
                    help="fixed tile shape tx,ty,tz,halo (skips the autotuner)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--counter-pass", default=None, metavar="VARIANTS",
                    help="counter-collection driver for rocprofv3 --pmc: a flat device copy "
                         "(1 GiB read + 1 GiB written per launch, two launches), then two "
                         "steps of each push variant 'tx,ty,tz,halo,block,prefetch;...' right "
                         "after a sort; prints no JSON line")
    ap.add_argument("--swarm-counter-pass", action="store_true",
                    help="counter-collection driver for the swarm kernel: 64 requests of the "
                         "served stream's distribution, two launches of 500 steps with every "
                         "request active; prints no JSON line")
    ap.add_argument("--no-discharge", action="store_true")
    ap.add_argument("--discharge-steps", type=int, default=1000)
    ap.add_argument("--discharge-warmup", type=int, default=100)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-full", action="store_true",
                    help="all 4,096 requests x 20,000 steps (minutes)")
    ap.add_argument("--serving-requests", type=int, default=256)
    ap.add_argument("--serving-steps", type=int, default=4000)
    ap.add_argument("--serving-launch-steps", type=int, default=500)
    ap.add_argument("--serving-hbm-gib", type=int, default=8)
    ap.add_argument("--serving-active-electrons", type=float, default=2e6,
                    help="cap on the electrons of the admitted requests together (0: none)")
    ap.add_argument("--ensemble-instances", type=int, default=64)
    ap.add_argument("--ensemble-particles", type=int, default=100_000)
    ap.add_argument("--ensemble-steps-per-launch", type=int, default=50)
    ap.add_argument("--ensemble-interval", type=float, default=0.1,
                    help="seconds between blocks of ensemble steps while swarms are served")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the per-step loop of one 3D3V instance runs on a "
              "single MI355X; sharding across GPUs is not built.",
              file=sys.stderr)
        sys.exit(2)

    from aurora_pic import hip
    from aurora_pic import workloads as wl

    if args.swarm_counter_pass:
        from aurora_pic import serving
        scfg = wl.config_swarm_requests(4096)
        reqs = serving.make_requests(scfg, count=64, steps=1000, burn_in=500)
        sctx = hip.Context(0)
        sctx.swarm_gas(scfg["gas"])
        for r in reqs:
            sctx.swarm_submit(r, hbm_budget=0)
        for k in range(2):
            sctx.swarm_advance(500)
            sctx.sync()
            log(f"swarm counter pass: launch {k}, "
                f"{sum(r.particles for r in reqs) * 500} electron-steps")
        sctx.close()
        return

    if args.copy_calibration:
        import torch
        x = torch.empty(2 ** 28, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB write")
        del x, y
    cfg = wl.config_plasma_3d(cells=args.cells, ppc=args.ppc)
    dt = cfg["dt"]
    ctx = hip.Context(0)
    log("device:", ctx.info())
    ctx.mesh(cfg["n"], cfg["L"], 0, cfg["eps"])
    axis, bx, bv = cfg["btable"]
    ctx.b_profile(axis, bx, bv)
    sids = []
    for k, s in enumerate(cfg["species"]):
        sid = ctx.add_species(s["charge"], s["mass"], s["weight"], s["particles"])
        ctx.init_maxwellian(sid, s["particles"], wl.maxwellian_sigma(s["temperature_ev"],
                                                                     s["mass"]), cfg["seed"] + k)
        sids.append(sid)
    total = sum(s["particles"] for s in cfg["species"])
    ctx.sort()
    ctx.initialize(dt)
    if args.counter_pass:
        log("copy calibration:", ctx.stream_bandwidth(2, 2 << 30, 1), "GB/s")
        for v in args.counter_pass.split(";"):
            shape = [int(x) for x in v.split(",")]
            ctx.tiling(*shape[:4])
            ctx.push_variant(shape[4], shape[5])
            ctx.sort()
            ctx.step(dt, 2)
            ctx.sync()
            log("counter pass: variant", v)
        ctx.close()
        return
    ctx.sort_interval(cfg["sort_interval"])
    adaptive = args.sort == "adaptive"
    if adaptive:
        ctx.sort_policy(hip.SORT_ADAPTIVE)
    sort_desc = "adaptive" if adaptive else cfg["sort_interval"]
    tuned = None
    if args.tile:
        shape = [int(v) for v in args.tile.split(",")]
        ctx.tiling(*shape[:4])
        if len(shape) > 4:
            ctx.push_variant(shape[4], shape[5])
        ctx.sort()
        tuned = {"tile": shape[:3], "halo": shape[3], "block": shape[4] if len(shape) > 4 else 512,
                 "prefetch": shape[5] if len(shape) > 4 else 0, "sort": sort_desc, "fixed": True}
    elif not args.no_autotune:
        from aurora_pic import autotune
        best, table = autotune.tune(ctx, dt, sort_intervals=(cfg["sort_interval"],), log=log,
                                    adaptive=adaptive, steps=40 if adaptive else 20)
        tuned = {"tile": list(best[1][:3]), "halo": best[1][3], "block": best[1][4],
                 "prefetch": best[1][5], "sort": sort_desc,
                 "candidates_ms_per_step": {
                     f"{s[0]}x{s[1]}x{s[2]}h{s[3]}b{s[4]}p{s[5]}": round(ms, 4)
                     for ms, s, _ in table}}
    log(f"particles {total}, warmup {args.warmup} steps")
    ctx.step(dt, args.warmup)
    ctx.sync()

    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    ctx.step(dt, args.steps)
    ctx.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = total * args.steps / elapsed
    log(f"{ms_per_step:.3f} ms/step, {value:.4g} particle-updates/s")

    # per-kernel HIP-event timing on the context stream
    ctx.timing(True)
    ctx.step(dt, args.timing_steps)
    ctx.sync()
    kernels = {}
    for name in ["push_deposit", "rho", "fft_fwd_x", "fft_axis", "fft_last", "fft_inv_x",
                 "green", "fft_load", "fft_unpack", "clear_charge", "sort"]:
        ms, n = ctx.kernel_time(name)
        if n:
            kernels[name] = {"avg_ms": ms / n, "launches": n}
    ctx.timing(False)
    pd = kernels.get("push_deposit", {"avg_ms": float("nan")})
    per_launch_particles = total / len(sids)
    achieved = BYTES_PER_UPDATE * per_launch_particles / (pd["avg_ms"] * 1e-3) / 1e9
    log("kernels:", json.dumps(kernels))
    ke, fe, l1, live = ctx.energy(dt)
    log(f"live {live}, kinetic {ke.sum():.6e} J, field {fe:.6e} J")
    # the attainable HBM bandwidth of hand-written streaming kernels
    streams = {k: ctx.stream_bandwidth(k) for k in range(6)}
    stream_peak = {"copy_gbs": max(streams[k] for k in (0, 2, 3)),
                   "soa7_in_place_gbs": max(streams[k] for k in (1, 4, 5)),
                   "patterns_gbs": [round(streams[k], 1) for k in range(6)]}
    log("stream peak:", json.dumps(stream_peak))
    ctx.close()

    # HBM bytes per launch of the variant that ran, from the counter runs
    # recorded per variant (null when that variant was not measured)
    traffic = None
    variant = None
    if tuned is not None:
        variant = ",".join(str(v) for v in (*tuned["tile"], tuned["halo"],
                                            tuned.get("block", 512), tuned.get("prefetch", 0)))
    tfile = os.path.join(REPO, "profiles", "push_deposit_plasma3d_traffic.txt")
    if variant and os.path.exists(tfile):
        with open(tfile) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 4 and parts[0] == "variant" and parts[1] == variant and \
                        parts[2] == "bytes_per_launch":
                    traffic = int(parts[3])

    discharge_out = None
    if not args.no_discharge:
        try:
            discharge_out = discharge_leg(args)
            log("discharge:", json.dumps(discharge_out))
        except Exception as e:  # a secondary leg must not hide the headline result
            log("discharge leg failed:", repr(e))
            discharge_out = {"error": repr(e)}

    others = None
    if not args.no_discharge:
        try:
            others = other_configs_leg(args)
            log("other configs:", json.dumps(others))
        except Exception as e:  # a secondary leg must not hide the headline result
            log("other configs failed:", repr(e))
            others = {"error": repr(e)}

    serving_out = None
    if not args.no_serving:
        try:
            serving_out = serving_leg(args)
        except Exception as e:  # the serving leg must not hide the headline result
            log("serving leg failed:", repr(e))
            serving_out = {"error": repr(e)}

    base = None
    if not args.no_cpu_baseline:
        log("cpu baseline (FP64 restatement, bounded sample)...")
        try:
            base = cpu_baseline(cfg, args.cpu_ppc)
        except Exception as e:  # the baseline must not hide the GPU result
            log("cpu baseline failed:", e)
    out = {
        "metric": "particle-updates/s (gather+push+deposit+MCC) vs CPU; batched swarm "
                  "p50/p99 latency",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plasma_3d3v: 128^3 cells, 16 e + 16 i per cell "
                               "(67,108,864 particles), Te 5 eV, Ti 0.1 eV, 131 amu ions, "
                               "tabulated B_y(x) 512 rows, FFT Poisson (FP64), "
                               f"cell sort {'adaptive' if adaptive else 'every 20 steps'}" if args.cells == 128 and args.ppc == 16 else
                               f"plasma_3d3v {args.cells}^3 x {args.ppc}+{args.ppc} ppc",
                   "particles": total, "cells": args.cells ** 3, "dt_s": dt,
                   "autotune": tuned, "kernels_ms": kernels},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "push_deposit (gather+Boris+drift+wrap+int64 deposit)",
                     "variant": variant,
                     "bytes_per_update": BYTES_PER_UPDATE,
                     "attainable": stream_peak},
        "cpu_baseline": base,
        "discharge_1d3v": discharge_out,
        "other_configs": others,
        "serving": serving_out,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
