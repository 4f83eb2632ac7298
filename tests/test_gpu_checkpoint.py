"""Checkpoint / restart of a context (the reference's restart guarantee:
continuing from a restart file equals the uninterrupted run).

- 3D3V, tabulated B, LDS-tiled push under the adaptive sort: 12 steps, save,
  8 more; a fresh context loads the file and runs the same 8 steps.
  Particles (by uid), fixed-point charge and fields agree to the bit.
- 1D3V RF discharge with null-collision MCC and ionization products
  (absorbing walls, appended products with new uids, the RNG keyed by step):
  the same, plus the MCC counters and the wall ledger.
- A context configured differently refuses the file (physics signature).
- 2D3V with uniform B_z and an FP64 species beside an FP32 one: the same.
- A species with no particles goes through steps, sorts, energies, ledgers
  and a checkpoint round trip."""
import math

import numpy as np
import pytest

from aurora_pic import workloads

pytestmark = pytest.mark.gpu


def _ctx_3d():
    from aurora_pic import hip
    rng = np.random.default_rng(3)
    n, L = [32, 32, 32], [1.0, 1.0, 1.0]
    N = 32 ** 3 * 2
    x = np.linspace(0, 1.0, 65)
    bvals = np.stack([0.1 * np.ones(65), 3.0 * np.exp(-((x - 0.5) / 0.2) ** 2), np.zeros(65)],
                     axis=1)
    c = hip.Context(0)
    c.mesh(n, L, 0, 1.0)
    c.b_profile(0, x, bvals)
    c.tiling(8, 8, 4, 2)
    c.push_variant(512, 1)
    c.sort_policy(hip.SORT_ADAPTIVE, 5)
    sids = []
    for q, m in ((-1.0, 1.0), (1.0, 50.0)):
        sid = c.add_species(q, m, 1.0 / N, N)
        c.upload(sid, rng.uniform(0, 1, (N, 3)), rng.normal(0, 0.5 / math.sqrt(m), (N, 3)),
                 kind=1)
        sids.append(sid)
    c.sort()
    return c, sids


def _state(c, sids):
    out = []
    for sid in sids:
        p, v, a, u = c.download(sid)
        o = np.argsort(u)
        out.append((p[o], v[o], a[o], u[o], c.charge_fixed(sid)))
    rho, phi, E = c.fields()
    return out, (rho, phi, E)


def _assert_same(a, b):
    for (pa, va, aa, ua, qa), (pb, vb, ab, ub, qb) in zip(a[0], b[0]):
        assert np.array_equal(ua, ub)
        assert np.array_equal(pa, pb) and np.array_equal(va, vb) and np.array_equal(aa, ab)
        assert np.array_equal(qa, qb)
    for fa, fb in zip(a[1], b[1]):
        assert np.array_equal(fa, fb)


def test_checkpoint_restart_3d_is_bit_exact(tmp_path):
    path = str(tmp_path / "run3d.ckpt")
    dt = 0.01
    c, sids = _ctx_3d()
    c.initialize(dt)
    c.step(dt, 12)
    c.checkpoint_save(path)
    c.step(dt, 8)
    a = _state(c, sids)
    t_a = c.time()
    c.close()
    d, dsids = _ctx_3d()
    d.checkpoint_load(path)
    d.step(dt, 8)
    b = _state(d, dsids)
    assert d.time() == t_a
    _assert_same(a, b)
    d.close()


def _ctx_1d(seed=77, weight_scale=1.0):
    from aurora_pic import hip
    rng = np.random.default_rng(5)
    n, L = 65, 0.067
    N = 3000
    gas = workloads.synthetic_gas(rows=1000, neutral_density=1.0e22)
    c = hip.Context(0)
    c.mesh([n], [L], 1, workloads.EPS0)
    c.electrodes([0, 0, 0, 0], [0.0, 200.0, 13.56e6, 0.0])
    sids = []
    w = 2.56e14 * L / N * weight_scale
    for q, m, te in ((-workloads.EV, workloads.ME, 8.0), (workloads.EV, 6.67e-27, 0.026)):
        sid = c.add_species(q, m, w, 4 * N)
        c.upload(sid, rng.uniform(0.1 * L, 0.9 * L, size=N),
                 rng.normal(0, math.sqrt(te * workloads.EV / m), size=(N, 3)), kind=1)
        sids.append(sid)
    c.mcc_load(sids[0], gas, product_species=(sids[0], sids[1]), seed=seed)
    return c, sids


def test_checkpoint_restart_1d_mcc_is_bit_exact(tmp_path):
    path = str(tmp_path / "run1d.ckpt")
    dt = 1.0 / (400 * 13.56e6)
    c, sids = _ctx_1d()
    c.initialize(dt)
    c.step(dt, 40)
    c.checkpoint_save(path)
    c.step(dt, 30)
    a = _state(c, sids)
    stats_a = c.mcc_stats()
    walls_a = [c.losses(s) for s in sids]
    c.close()
    d, dsids = _ctx_1d()
    d.checkpoint_load(path)
    d.step(dt, 30)
    b = _state(d, dsids)
    _assert_same(a, b)
    stats_b = d.mcc_stats()
    assert np.array_equal(stats_a[0], stats_b[0]) and np.array_equal(stats_a[1], stats_b[1])
    for (ca, ea), (cb, eb) in zip(walls_a, [d.losses(s) for s in dsids]):
        assert np.array_equal(ca, cb) and np.array_equal(ea, eb)
    # more electrons than were loaded: products were appended before the save
    assert len(b[0][0][3]) > 3000
    d.close()


def test_checkpoint_refuses_other_physics(tmp_path):
    from aurora_pic import hip
    path = str(tmp_path / "run1d.ckpt")
    dt = 1.0 / (400 * 13.56e6)
    c, _ = _ctx_1d()
    c.initialize(dt)
    c.step(dt, 2)
    c.checkpoint_save(path)
    sig = c.physics_signature()
    c.close()
    for kwargs in (dict(seed=78), dict(weight_scale=2.0)):
        d, _ = _ctx_1d(**kwargs)
        assert d.physics_signature() != sig
        with pytest.raises(hip.HipError, match="physics signature mismatch"):
            d.checkpoint_load(path)
        d.close()


def test_checkpoint_restart_2d_fp64_is_bit_exact(tmp_path):
    """2D3V with a uniform B_z, an FP64 species beside an FP32 one: the
    restart continues to the bit, FP64 arrays included."""
    from aurora_pic import hip
    path = str(tmp_path / "run2d.ckpt")
    dt = 0.02

    def ctx():
        rng = np.random.default_rng(11)
        c = hip.Context(0)
        c.mesh([64, 32], [2.0, 1.0], 0, 1.0)
        c.uniform_b([0.0, 0.0, 0.4])
        sids = []
        for q, m, prec in ((-1.0, 1.0, 64), (1.0, 25.0, 32)):
            N = 5000
            sid = c.add_species(q, m, 2.0 / N, N, precision=prec)
            c.upload(sid, rng.uniform(0, 1, (N, 2)) * [2.0, 1.0],
                     rng.normal(0, 0.2 / math.sqrt(m), (N, 3)), kind=1)
            sids.append(sid)
        return c, sids

    c, sids = ctx()
    c.initialize(dt)
    c.step(dt, 15)
    c.checkpoint_save(path)
    c.step(dt, 10)
    a = _state(c, sids)
    c.close()
    d, dsids = ctx()
    d.checkpoint_load(path)
    d.step(dt, 10)
    _assert_same(a, _state(d, dsids))
    d.close()


def test_empty_species_through_step_sort_checkpoint(tmp_path):
    """A species with no particles (before and after an upload of zero)
    takes part in every call: steps, sorts, energies, ledgers, checkpoints."""
    from aurora_pic import hip
    path = str(tmp_path / "empty.ckpt")
    rng = np.random.default_rng(2)
    c = hip.Context(0)
    c.mesh([16, 16, 16], [1.0, 1.0, 1.0], 0, 1.0)
    full = c.add_species(-1.0, 1.0, 1.0 / 4096, 4096)
    empty = c.add_species(1.0, 100.0, 1.0, 16)
    c.upload(full, rng.uniform(0, 1, (4096, 3)), rng.normal(0, 0.1, (4096, 3)))
    c.upload(empty, np.zeros((0, 3)), np.zeros((0, 3)))
    c.sort()
    c.initialize(0.01)
    c.step(0.01, 5)
    c.sort()
    ke, fe, l1, live = c.energy(0.01)
    assert live == 4096 and ke[1] == 0.0 and ke[0] > 0.0
    assert c.count(empty) == (0, 0)
    counts, e2 = c.losses(empty)
    assert counts.sum() == 0
    c.checkpoint_save(path)
    c.step(0.01, 3)
    s = _state(c, [full, empty])
    c.close()
    d = hip.Context(0)
    d.mesh([16, 16, 16], [1.0, 1.0, 1.0], 0, 1.0)
    full2 = d.add_species(-1.0, 1.0, 1.0 / 4096, 4096)
    empty2 = d.add_species(1.0, 100.0, 1.0, 16)
    d.checkpoint_load(path)
    d.step(0.01, 3)
    _assert_same(s, _state(d, [full2, empty2]))
    d.close()
