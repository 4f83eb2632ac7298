"""Pins the FP64 oracle restatement of deposit / gather / push / Poisson.

Each check is one the reference itself holds (test_core.cpp known answers),
a golden vector written by the reference (tests/golden), the live reference
when oracle/_ref is built, or a closed form chosen so that a dropped term,
a wrong sign or a transposed index fails it."""
import math

import numpy as np
import pytest

from conftest import golden

CASES = ["1d_periodic", "1d_dirichlet", "2d_periodic", "2d_dirichlet", "3d_periodic",
         "3d_dirichlet"]


# ------------------------------------------------------------ deposit
@pytest.mark.parametrize("case", CASES)
def test_deposit_matches_reference_golden(oracle, case):
    g = golden(f"deposit_gather_{case}.npz")
    rho = oracle.deposit_cic(g["n"], g["L"], int(g["bc"]), g["pos"], float(g["charge"]),
                             float(g["weight"]), g["alive"])
    scale = np.max(np.abs(g["rho"]))
    assert np.max(np.abs(rho - g["rho"])) <= 1e-13 * scale


@pytest.mark.parametrize("case", CASES)
def test_gather_matches_reference_golden(oracle, case):
    g = golden(f"deposit_gather_{case}.npz")
    out = oracle.gather(g["n"], g["L"], int(g["bc"]), g["pos"], g["E"])
    assert np.max(np.abs(out - g["gathered"])) <= 1e-13 * np.max(np.abs(g["E"]))


def test_cic_wrapped_weights_2d_3d(oracle):
    # test_core.cpp:1874-1904: wrapped corner weights and charge conservation
    pos2 = [[0.125, 0.125], [1.125, -0.125], [0.5, 0.5]]
    rho = oracle.deposit_cic([4, 4], [1.0, 1.0], 0, pos2, 2.0, 0.5, alive=[1, 1, 0])
    assert abs(rho.sum() * 0.25 * 0.25 - 2.0) < 1e-12
    assert abs(rho[0] - 8.0) < 1e-12 and abs(rho[1] - 8.0) < 1e-12 and abs(rho[4] - 4.0) < 1e-12
    pos3 = [[0.125, 0.125, 0.125], [1.125, -0.125, 0.875], [0.5, 0.5, 0.5]]
    rho = oracle.deposit_cic([4, 4, 4], [1.0] * 3, 0, pos3, 2.0, 0.5, alive=[1, 1, 0])
    assert abs(rho.sum() * 0.25 ** 3 - 2.0) < 1e-12
    assert abs(rho[0] - 16.0) < 1e-12 and abs(rho[1] - 16.0) < 1e-12
    assert abs(rho[4] - 8.0) < 1e-12 and abs(rho[48] - 8.0) < 1e-12


def test_cic_exact_single_particle_3d(oracle):
    # test_core.cpp:1935-1976: exact trilinear shape, written independently
    n, L = [6, 5, 4], [1.2, 1.0, 0.8]
    d = [L[a] / n[a] for a in range(3)]
    p = [0.37, 0.46, 0.51]
    rho = oracle.deposit_cic(n, L, 0, [p], 1.25, 0.8)
    g = [p[a] / d[a] for a in range(3)]
    i = [int(math.floor(v)) for v in g]
    f = [g[a] - i[a] for a in range(3)]
    exp = np.zeros(120)
    for c in range(2):
        for b in range(2):
            for a in range(2):
                w = (f[0] if a else 1 - f[0]) * (f[1] if b else 1 - f[1]) * (f[2] if c else 1 - f[2])
                exp[((i[2] + c) * 5 + i[1] + b) * 6 + i[0] + a] += 1.0 * w / (d[0] * d[1] * d[2])
    assert np.max(np.abs(rho - exp)) < 1e-12


def test_dirichlet_half_volume_1d(oracle):
    # test_core.cpp:2131-2146: particle at the wall fills a half control volume
    rho = oracle.deposit_cic([5], [1.0], 1, [0.0], 1.0, 1.0)
    assert abs(rho[0] - 1.0 / (0.5 * 0.25)) < 1e-12 and rho[1] == 0.0


def test_gather_affine_exact_3d(oracle):
    # test_core.cpp:2051-2079
    n, L = [5, 4, 3], [1.0, 0.8, 0.6]
    x = np.arange(5) * 0.2
    y = np.arange(4) * 0.2
    z = np.arange(3) * 0.2
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    E = np.stack([1 + 2 * X - 3 * Y + 0.5 * Z, -0.25 + 0.75 * X + 1.5 * Y - 2 * Z,
                  3 - X + 0.25 * Y + 4 * Z]).reshape(3, -1)
    p = np.array([[0.37, 0.31, 0.26]])
    out = oracle.gather(n, L, 0, p, E)[0]
    ex = [1 + 2 * .37 - 3 * .31 + .5 * .26, -.25 + .75 * .37 + 1.5 * .31 - 2 * .26,
          3 - .37 + .25 * .31 + 4 * .26]
    assert np.max(np.abs(out - ex)) < 1e-12
    wrapped = oracle.gather(n, L, 0, [[0.37 + 1.0, 0.31 - 0.8, 0.26]], E)[0]
    assert np.max(np.abs(wrapped - out)) < 1e-12


# ------------------------------------------------------- fixed point
def _random_cells(rng, n, count, bc):
    dim = len(n)
    cells_per_axis = [k if bc == 0 else k - 1 for k in n]
    idx = [rng.integers(0, c, size=count) for c in cells_per_axis]
    cell = idx[0].copy()
    stride = 1
    for a in range(1, dim):
        stride *= n[a - 1]
        cell += idx[a] * stride
    off = rng.random((count, dim)).astype(np.float32).astype(np.float64)
    return cell.astype(np.int32), off


@pytest.mark.parametrize("n,bc", [([64], 0), ([33], 1), ([12, 10], 0), ([9, 7], 1),
                                  ([8, 6, 5], 0), ([6, 5, 4], 1)])
def test_fixed_point_deposit_invariants(oracle, n, bc):
    rng = np.random.default_rng(5)
    cell, off = _random_cells(rng, n, 3000, bc)
    acc = oracle.deposit_fixed(n, bc, cell, off)
    perm = rng.permutation(cell.size)
    # order independence: integer adds
    assert np.array_equal(acc, oracle.deposit_fixed(n, bc, cell[perm], off[perm]))
    # each particle deposits 2^40 up to one half-unit per corner
    corners = 2 ** len(n)
    assert abs(int(acc.sum()) - cell.size * 2 ** 40) <= cell.size * corners // 2
    # splitting the particle set gives the same total (chunked/batched runs)
    a1 = oracle.deposit_fixed(n, bc, cell[:1000], off[:1000])
    a2 = oracle.deposit_fixed(n, bc, cell[1000:], off[1000:])
    assert np.array_equal(acc, a1 + a2)


@pytest.mark.parametrize("n,L,bc", [([64], [2.0], 0), ([33], [0.067], 1),
                                    ([12, 10], [1.2, 1.0], 0), ([8, 6, 5], [1.0, .8, .6], 0),
                                    ([6, 5, 4], [1.2, 1.0, 0.8], 1)])
def test_fixed_point_rho_matches_reference_deposit(oracle, n, L, bc):
    """Fixed-point rho differs from the FP64 reference deposit of the same
    positions by at most the 2^-40 quantisation per contribution."""
    rng = np.random.default_rng(6)
    cell, off = _random_cells(rng, n, 2000, bc)
    dim = len(n)
    d = [L[a] / (n[a] if bc == 0 else n[a] - 1) for a in range(dim)]
    idx = []
    c = cell.astype(np.int64)
    for a in range(dim):
        idx.append(c % n[a])
        c //= n[a]
    pos = np.stack([(idx[a] + off[:, a]) * d[a] for a in range(dim)], axis=1)
    acc = oracle.deposit_fixed(n, bc, cell, off)
    rho_fx = oracle.fixed_to_rho(n, L, bc, [acc], [-1.5 * 0.75])
    rho_ref = oracle.deposit_cic(n, L, bc, pos, -1.5, 0.75)
    vol = np.prod(d)
    contributions = 2000 * 2 ** dim
    bound = contributions * (2.0 ** -40 + 1e-15) * 1.5 * 0.75 / (vol * 0.125)
    assert np.max(np.abs(rho_fx - rho_ref)) <= bound


# ------------------------------------------------------------ Poisson
@pytest.mark.parametrize("case", ["1d_periodic", "1d_dirichlet", "2d_pow2", "2d_mixed",
                                  "3d_pow2", "3d_mixed"])
def test_poisson_matches_reference_golden(oracle, case):
    g = golden(f"solve_{case}.npz")
    phi, E = oracle.poisson(g["n"], g["L"], int(g["bc"]), float(g["eps"]), g["rho"],
                            float(g["phi_left"]), float(g["phi_right"]))
    assert np.max(np.abs(phi - g["phi"])) <= 1e-11 * np.max(np.abs(g["phi"]))
    assert np.max(np.abs(E - g["E"])) <= 1e-11 * np.max(np.abs(g["E"]))


def test_poisson_1d_sine_known_answer(oracle):
    # test_core.cpp:253-266
    x = np.arange(64) / 64.0
    phi, E = oracle.poisson([64], [1.0], 0, 1.0, np.sin(2 * np.pi * x))
    assert np.max(np.abs(E[0] + np.cos(2 * np.pi * x) / (2 * np.pi))) < 1e-12
    _, E2 = oracle.poisson([64], [1.0], 0, 2.0, np.sin(2 * np.pi * x))
    assert np.max(np.abs(E2[0] - 0.5 * E[0])) < 1e-13


@pytest.mark.parametrize("n", [[16, 12, 10], [8, 16, 8]])
def test_poisson_3d_sine_cosine_known_answer(oracle, n):
    # test_core.cpp:1360-1438 (analytic mode, non power-of-two included)
    L = [1.0, 1.25, 1.5]
    kx, ky, kz = 2 * np.pi / L[0], 4 * np.pi / L[1], 6 * np.pi / L[2]
    k2 = kx * kx + ky * ky + kz * kz
    z = np.arange(n[2]) * L[2] / n[2]
    y = np.arange(n[1]) * L[1] / n[1]
    x = np.arange(n[0]) * L[0] / n[0]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = np.sin(kx * X) * np.cos(ky * Y) * np.sin(kz * Z)
    phi, E = oracle.poisson(n, L, 0, 1.0, rho.reshape(-1))
    assert np.max(np.abs(phi - (rho / k2).reshape(-1))) < 1e-12
    ex = -kx * np.cos(kx * X) * np.cos(ky * Y) * np.sin(kz * Z) / k2
    ey = ky * np.sin(kx * X) * np.sin(ky * Y) * np.sin(kz * Z) / k2
    ez = -kz * np.sin(kx * X) * np.cos(ky * Y) * np.cos(kz * Z) / k2
    assert np.max(np.abs(E[0] - ex.reshape(-1))) < 1e-12
    assert np.max(np.abs(E[1] - ey.reshape(-1))) < 1e-12
    assert np.max(np.abs(E[2] - ez.reshape(-1))) < 1e-12


def test_dirichlet_tridiagonal_closed_forms(oracle):
    n, L, eps = 17, 2.0, 1.5
    x = np.linspace(0, L, n)
    # rho = 0: linear potential between the electrodes, uniform E
    phi, E = oracle.poisson([n], [L], 1, eps, np.zeros(n), 4.0, -2.0)
    assert np.max(np.abs(phi - (4.0 - 3.0 * x))) < 1e-13
    assert np.max(np.abs(E[0] - 3.0)) < 1e-12
    # uniform rho: the discrete parabola is exact for -phi'' = rho/eps
    r = 0.7
    phi, _ = oracle.poisson([n], [L], 1, eps, np.full(n, r), 0.0, 0.0)
    assert np.max(np.abs(phi - r / (2 * eps) * x * (L - x))) < 1e-12
    # dense solve on a small random case
    rng = np.random.default_rng(3)
    rho = rng.normal(size=n)
    phi, _ = oracle.poisson([n], [L], 1, eps, rho, 1.0, 2.0)
    dx = L / (n - 1)
    A = 2 * np.eye(n - 2) - np.eye(n - 2, k=1) - np.eye(n - 2, k=-1)
    b = rho[1:-1] * dx * dx / eps
    b[0] += 1.0
    b[-1] += 2.0
    assert np.max(np.abs(phi[1:-1] - np.linalg.solve(A, b))) < 1e-12


# ---------------------------------------------------------------- push
def test_boris_matches_reference_golden(oracle):
    g = golden("boris.npz")
    for i in range(len(g["qm"])):
        out = oracle.boris(g["v"][i], g["E"][i], g["B"][i], float(g["qm"][i]), float(g["h"][i]))
        assert np.max(np.abs(out - g["out"][i])) <= 1e-15 * max(1.0, np.max(np.abs(out)))


def test_boris_cyclotron_angle(oracle):
    # test_core.cpp:1291-1342: rotation by 2 atan(|qB| dt / 2) per step
    qm, dt = -0.75, 0.03
    B = np.array([0.4, -0.8, 1.1])
    v0 = np.array([0.6, -0.25, 0.9])
    v = v0.copy()
    for _ in range(29):
        v = oracle.boris(v, np.zeros(3), B, qm, dt)
    bm = np.linalg.norm(B)
    axis = B / bm
    angle = 29 * -2.0 * math.atan(qm * bm * dt / 2.0)
    # Rodrigues rotation about axis by angle
    exp = (v0 * math.cos(angle) + np.cross(axis, v0) * math.sin(angle) +
           axis * np.dot(axis, v0) * (1 - math.cos(angle)))
    assert np.max(np.abs(v - exp)) < 1e-13
    assert abs(np.linalg.norm(v) - np.linalg.norm(v0)) < 1e-13


def test_table_eval_matches_reference_golden(oracle):
    g = golden("table.npz")
    for x, valid, out in zip(g["q"], g["valid"], g["out"]):
        r = oracle.table_eval(g["coords"], g["values"], float(x))
        assert (r is not None) == bool(valid)
        if valid:
            assert np.max(np.abs(r - out)) <= 1e-15 * max(1.0, np.max(np.abs(out)))


def test_table_eval_known_answer(oracle):
    # test_core.cpp:1108-1150
    r = oracle.table_eval([0.0, 0.5, 1.0], [[0, 1, 2], [1, 3, 4], [2, 5, 6]], 0.25)
    assert np.allclose(r, [0.5, 2.0, 3.0], atol=1e-15)
    assert oracle.table_eval([0.0, 0.5, 1.0], [[0, 1, 2], [1, 3, 4], [2, 5, 6]], -0.01) is None


def test_leapfrog_constant_acceleration(oracle):
    # test_core.cpp:1068-1107 through the species push with a uniform field
    n, L = [8, 8], [100.0, 100.0]
    acc = np.array([1.5, -0.75])
    E = np.stack([np.full(64, acc[0]), np.full(64, acc[1])])
    pos = np.array([[50.25, 50.75]])
    v0 = np.array([[-0.2, 0.3, 0.0]])
    alive = np.array([1], dtype=np.uint8)
    dt = 0.1
    vh = oracle.init_half(n, L, 0, pos, v0, alive, E, 1.0, dt)
    for _ in range(6):
        pos, vh, alive, _ = oracle.push(n, L, 0, pos, vh, alive, E, 1.0, dt)
    v = oracle.sync(n, L, 0, pos, vh, alive, E, 1.0, dt)
    t = 0.6
    assert np.max(np.abs(pos[0] - (np.array([50.25, 50.75]) + v0[0, :2] * t + 0.5 * acc * t * t))) < 1e-12
    assert np.max(np.abs(v[0, :2] - (v0[0, :2] + acc * t))) < 1e-13


def test_particle_boundaries_one_step(oracle):
    # test_core.cpp:2002-2050 analytic one-step outcomes (2D)
    n, L = [5, 5], [1.0, 1.0]
    E = np.zeros((2, 25))
    pos = np.array([[0.03, 0.5], [0.94, 0.25], [0.2, 0.03], [0.7, 0.92]])
    vh = np.array([[-0.4, 0, 0], [0.5, 0, 0], [0, -0.4, 0], [0, 0.6, 0]], dtype=float)
    alive = np.ones(4, dtype=np.uint8)
    pb = [1, 2, 3, 2, 3, 3]  # left absorb, right reflect, bottom periodic, top reflect
    pos, vh, alive, absorbed = oracle.push(n, L, 1, pos, vh, alive, E, 0.0, 0.2, pb=pb)
    assert absorbed[0] == 1 and absorbed[1:].sum() == 0 and alive.tolist() == [0, 1, 1, 1]
    assert abs(pos[1, 0] - 0.96) < 1e-12 and abs(vh[1, 0] + 0.5) < 1e-12
    assert abs(pos[2, 1] - 0.95) < 1e-12 and abs(vh[2, 1] + 0.4) < 1e-12
    assert abs(pos[3, 1] - 0.96) < 1e-12 and abs(vh[3, 1] + 0.6) < 1e-12


# ------------------------------------------------------ live reference
def test_live_reference_random_cases(oracle, reference):
    rng = np.random.default_rng(11)
    for n, L, bc in [([40], [3.0], 0), ([21], [0.5], 1), ([10, 14], [1.0, 1.4], 0),
                     ([7, 9, 6], [0.7, 0.9, 0.6], 0), ([7, 6, 5], [1, 1, 1], 1)]:
        dim = len(n)
        pos = rng.uniform(-0.2, 1.2, size=(500, dim)) * np.asarray(L)
        if bc == 1:
            pos = np.clip(pos, 0, np.asarray(L))
        a = oracle.deposit_cic(n, L, bc, pos, 2.0, 0.3)
        b = reference.deposit(n, L, bc, pos, 2.0, 0.3)
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
        E = rng.normal(size=(dim, int(np.prod(n))))
        assert np.max(np.abs(oracle.gather(n, L, bc, pos, E) - reference.gather(n, L, bc, pos, E))) < 1e-13


def test_fixed_point_3d_weights_exact(oracle):
    """The 3D fixed-point weight of every corner is wx * fl(wy * wz) * 2^40
    rounded once, half to even (brute force in exact rationals), including
    offsets small enough that fl(wy * wz) rounds and products on ties."""
    from fractions import Fraction
    rng = np.random.default_rng(11)
    offs = [rng.random(3).astype(np.float32).astype(np.float64) for _ in range(200)]
    tiny = np.float32(2.0 ** -20)
    offs += [np.array([0.5, tiny, np.float32(3.0 * 2.0 ** -21)], dtype=np.float64),
             np.array([0.25, 0.5, 0.75]), np.array([0.0, 0.0, 0.0]),
             np.array([np.float32(1.0 - 2.0 ** -24), tiny, tiny], dtype=np.float64)]
    # float32 offsets where rounding fl(fl(wx * wy) * wz) first and then to
    # the grid lands one unit away from the single rounding
    offs += [np.array([0.7479357719421387, 0.7795912623405457, 0.9397904276847839]),
             np.array([0.22637362778186798, 0.8544207811355591, 0.6317999362945557]),
             np.array([0.6504586935043335, 0.44892260432243347, 0.829235851764679])]
    n = [4, 4, 4]
    for o in offs:
        acc = oracle.deposit_fixed(n, 0, np.array([0], dtype=np.int32), o.reshape(1, 3))
        wx = [1.0 - o[0], o[0]]
        wy = [1.0 - o[1], o[1]]
        wz = [1.0 - o[2], o[2]]
        for c in range(2):
            for b in range(2):
                wyz = wy[b] * wz[c]                       # one FP64 rounding
                for a in range(2):
                    exact = Fraction(wx[a]) * Fraction(wyz) * 2 ** 40
                    assert acc[(c * 4 + b) * 4 + a] == round(exact), (o, a, b, c)


def test_step3d_is_the_pinned_pieces_in_order(oracle):
    """or_step3d (the CPU baseline's Simulation3D::step, Simulation3D.cpp:
    342-365) is the pinned push, deposit, periodic solve and gather in the
    reference's order: one step equals that composition bit for bit, and the
    OpenMP synchronisation loop gives the same state on 1 and 4 threads."""
    import copy
    n, L, eps, dt = [8, 6, 4], [1.0, 0.75, 0.5], 1.0, 0.01
    species = [(-1.0, 1.0, 1e-3, 600, 0.3), (1.0, 50.0, 1e-3, 400, 0.05)]
    st = oracle.plasma3d_init(n, L, eps, dt, species, seed=11)
    st0 = copy.deepcopy(st)
    st4 = copy.deepcopy(st)
    oracle.step3d(st, 1)
    oracle.step3d(st4, 4)
    for key in ("pos", "vh", "v", "alive"):
        for a, b in zip(st[key], st4[key]):
            assert np.array_equal(a, b)
    for key in ("rho", "phi", "E"):
        assert np.array_equal(st[key], st4[key])
    # the composition from the same start
    rho = np.zeros(int(np.prod(n)))
    E0 = st0["E"].reshape(3, -1)
    moved = []
    for k, (q, m, w, count, _) in enumerate(species):
        pos, vh, alive, _ = oracle.push(n, L, 0, st0["pos"][k].reshape(-1, 3),
                                        st0["vh"][k].reshape(-1, 3), st0["alive"][k], E0,
                                        q / m, dt)
        assert np.array_equal(pos.reshape(-1), st["pos"][k])
        assert np.array_equal(vh.reshape(-1), st["vh"][k])
        rho += oracle.deposit_cic(n, L, 0, pos, q, w)
        moved.append((pos, vh, q / m))
    # one accumulator over both species there, a sum of per-species arrays here
    assert np.max(np.abs(rho - st["rho"])) <= 1e-14 * np.max(np.abs(rho))
    _, E = oracle.poisson(n, L, 0, eps, st["rho"])
    assert np.array_equal(E.reshape(-1), st["E"])
    for k, (pos, vh, qm) in enumerate(moved):
        e = oracle.gather(n, L, 0, pos, E)
        assert np.array_equal(vh + 0.5 * qm * e * dt, st["v"][k].reshape(-1, 3))
