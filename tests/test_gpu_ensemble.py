"""Batched 1D ensembles on the GPU: every instance equals the same
discharge run alone on a context (bit for bit: particles, uids, fields and
collision counters), graph replays equal direct launches, and the
all-or-nothing capacity rule is reported per instance.  The single-context
1D step itself is checked against the oracle in test_gpu_mcc.py."""
import math

import numpy as np
import pytest

from aurora_pic import workloads

pytestmark = pytest.mark.gpu

N = 1500
NX, L = 65, 0.067
DT = 1.0 / (400 * 13.56e6)
SPECIES = [(-workloads.EV, workloads.ME, 8.0), (workloads.EV, 6.67e-27, 0.026)]


def _template(hip, gas, capacity=4 * N, mcc_seed=1, nx=NX):
    c = hip.Context(0)
    c.mesh([nx], [L], hip.DIRICHLET, workloads.EPS0)
    c.electrodes([0, 0, 0, 0], [0.0, 200.0, 13.56e6, 0.0])
    sids = [c.add_species(q, m, 2.56e14 * L / N, capacity) for q, m, _ in SPECIES]
    c.mcc_load(sids[0], gas, product_species=(sids[0], sids[1]), seed=mcc_seed)
    return c, sids


def _init_instance(c, sids, k):
    for sid, (q, m, te) in zip(sids, SPECIES):
        c.init_maxwellian(sid, N, math.sqrt(te * workloads.EV / m), seed=1000 * k + 7 * sid + 3)
    c.initialize(DT)


def _alone(hip, gas, k, mcc_seed, steps, nx=NX):
    c, sids = _template(hip, gas, mcc_seed=mcc_seed, nx=nx)
    _init_instance(c, sids, k)
    c.step(DT, steps)
    out = [c.download(s) for s in sids]
    fields = c.fields()
    counts, ech = c.mcc_stats()
    walls = [c.losses(s) for s in sids]
    c.close()
    return out, fields, counts, ech, walls


def _by_uid(pos, vel, alive, uid):
    m = alive == 1
    order = np.argsort(uid[m], kind="stable")
    return uid[m][order], pos[m][order], vel[m][order]


def _build_ensemble(hip, gas, instances, graph_steps, capacity=4 * N, nx=NX):
    c, sids = _template(hip, gas, capacity, nx=nx)
    ens = c.ensemble_create(instances, [capacity, capacity])
    ens.graph(graph_steps)
    for k in range(instances):
        _init_instance(c, sids, k)
        ens.load(k, 500 + k)
    return c, sids, ens


def test_instances_equal_single_context_runs():
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    steps = 120                     # two graphs of 50 + 20 direct steps
    c, sids, ens = _build_ensemble(hip, gas, 4, 50)
    ens.step(DT, steps)
    t, s = ens.time()
    assert s == steps
    for k in (0, 3):
        ref, ref_fields, ref_counts, ref_ech, ref_walls = _alone(hip, gas, k, 500 + k, steps)
        for sid in sids:
            a = _by_uid(*ens.download(k, sid))
            b = _by_uid(*ref[sid])
            assert a[0].shape == b[0].shape, (k, sid, a[0].shape, b[0].shape)
            assert np.array_equal(a[0], b[0])
            assert np.array_equal(a[1], b[1])
            assert np.array_equal(a[2], b[2])
        rho, phi, E = ens.fields(k)
        assert np.array_equal(rho, ref_fields[0])
        assert np.array_equal(phi, ref_fields[1])
        assert np.array_equal(E, ref_fields[2][0])
        counts, ech, absorbed, absorbed_e = ens.stats(k, 3)
        assert np.array_equal(counts, ref_counts)
        assert counts[4] > 0          # ionizations happened: products were placed
        # FP64 ledgers: the same terms summed in another fixed order (the lone
        # context keeps dead slots in place, the ensemble compacts them), so
        # they agree to rounding: 1e-12 relative
        assert np.allclose(ech, ref_ech, rtol=1e-12, atol=0.0), (ech, ref_ech)
        assert ech[1] < 0 and ech[2] < 0   # excitation and ionization take energy
        for sid in sids:
            ref_n, ref_e = ref_walls[sid]
            assert np.array_equal(absorbed[sid], ref_n[:2].astype(np.uint64))
            assert np.allclose(absorbed_e[sid], ref_e, rtol=1e-12, atol=0.0), (absorbed_e, ref_e)
        assert absorbed[0].sum() > 0 and absorbed_e[0].sum() > 0   # electrons reached the walls
    c.close()


def test_instances_equal_lone_runs_on_a_long_mesh():
    # 1500 nodes: the field step's Dirichlet solve on a mesh beyond 1024
    # nodes, in the ensemble and on a context
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    c, sids, ens = _build_ensemble(hip, gas, 2, 20, nx=1500)
    ens.step(DT, 40)
    for k in (0, 1):
        ref, ref_fields, ref_counts, _, _ = _alone(hip, gas, k, 500 + k, 40, nx=1500)
        for sid in sids:
            a = _by_uid(*ens.download(k, sid))
            b = _by_uid(*ref[sid])
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
        rho, phi, E = ens.fields(k)
        assert np.array_equal(phi, ref_fields[1])
        assert np.array_equal(E, ref_fields[2][0])
        assert np.array_equal(ens.stats(k, 3)[0], ref_counts)
    c.close()


def test_ensemble_ledgers_repeat_to_the_bit():
    # every FP64 tally of an instance is summed in a fixed order: two
    # ensembles of the same instances give identical ledgers
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    runs = []
    for _ in range(2):
        c, sids, ens = _build_ensemble(hip, gas, 3, 25)
        ens.step(DT, 60)
        runs.append([ens.stats(k, 3) for k in range(3)])
        c.close()
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert a[3][0].sum() > 0


def test_graph_replays_equal_direct_launches():
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    results = []
    for g in (20, 0):
        c, sids, ens = _build_ensemble(hip, gas, 3, g)
        for _ in range(3):           # three replays with host work between
            ens.step(DT, 20)
            _ = ens.count(1, 0)
        results.append([[_by_uid(*ens.download(k, s)) for s in sids] for k in range(3)])
        c.close()
    for k in range(3):
        for s in range(2):
            for a, b in zip(results[0][k][s], results[1][k][s]):
                assert np.array_equal(a, b)


def test_instances_differ_by_seed_and_compaction_keeps_live():
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    c, sids, ens = _build_ensemble(hip, gas, 2, 25)
    ens.step(DT, 50)
    for k in range(2):
        for sid in sids:
            slots, live = ens.count(k, sid)
            assert slots == live          # compacted after the graph
    a = _by_uid(*ens.download(0, 0))
    b = _by_uid(*ens.download(1, 0))
    assert a[0].shape != b[0].shape or not np.array_equal(a[1], b[1])
    c.close()


def test_capacity_error_names_instance():
    from aurora_pic import hip
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    c, sids, ens = _build_ensemble(hip, gas, 2, 0, capacity=N)   # no room for products
    with pytest.raises(hip.HipError, match="max_particles_per_species \\(ensemble instance"):
        ens.step(DT, 50)
    c.close()
