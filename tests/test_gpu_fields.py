"""GPU parity of the deposit, gather + push and field solves through the
C ABI, against the oracle (and the reference's golden vectors).

Bars: charge density bit-exact against the oracle's fixed-point definition
on the device's own (cell, offset) state; fields and one-step push within
1e-6 relative (max-norm over the array, relative to the array's scale) of
the FP64 oracle restatement."""
import math

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

CASES = ["1d_periodic", "1d_dirichlet", "2d_periodic", "2d_dirichlet", "3d_periodic",
         "3d_dirichlet"]


@pytest.fixture
def hip_ctx():
    from aurora_pic import hip
    c = hip.Context(0)
    yield c
    c.close()


def _setup_case(ctx, g, precision=32):
    n, L, bc = g["n"], g["L"], int(g["bc"])
    ctx.mesh(n, L, bc, 1.0)
    sid = ctx.add_species(float(g["charge"]), 1.0, float(g["weight"]), len(g["alive"]) + 16,
                          precision)
    vel = np.zeros((len(g["alive"]), 3))
    ctx.upload(sid, g["pos"], vel, g["alive"])
    return sid


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("precision", [32, 64])
def test_deposit_bit_exact_against_fixed_point_oracle(oracle, hip_ctx, case, precision):
    print(f"case={case} precision={precision}", flush=True)
    g = golden(f"deposit_gather_{case}.npz")
    sid = _setup_case(hip_ctx, g, precision)
    hip_ctx.deposit()
    acc = hip_ctx.charge_fixed(sid)
    cell, off, _ = hip_ctx.download_raw(sid)
    expect = oracle.deposit_fixed(g["n"], int(g["bc"]), cell, off)
    assert np.array_equal(acc, expect)
    rho = hip_ctx.rho()
    rho_fx = oracle.fixed_to_rho(g["n"], g["L"], int(g["bc"]), [expect],
                                 [float(g["charge"]) * float(g["weight"])])
    assert np.array_equal(rho, rho_fx)
    # against the reference's FP64 deposit of the same input positions: the
    # difference is the offset rounding (FP32: 2^-24 per weight) plus 2^-40
    ulp = 2.0 ** -24 if precision == 32 else 2.0 ** -52
    dim = len(g["n"])
    touch = oracle.deposit_fixed(g["n"], int(g["bc"]), cell, np.full_like(off, 0.5))
    count = touch / (2.0 ** 40 / 2 ** dim)
    d = [g["L"][a] / (g["n"][a] if int(g["bc"]) == 0 else g["n"][a] - 1) for a in range(dim)]
    vmin = np.prod(d) / 2 ** dim
    bound = count * (dim * 2 * ulp + 2.0 ** -40) * abs(float(g["charge"]) * float(g["weight"])) / vmin
    assert np.all(np.abs(rho - g["rho"]) <= bound + 1e-300)


def test_deposit_order_independent(oracle, hip_ctx):
    g = golden("deposit_gather_3d_periodic.npz")
    sid = _setup_case(hip_ctx, g)
    hip_ctx.deposit()
    a = hip_ctx.charge_fixed(sid)
    perm = np.random.default_rng(0).permutation(len(g["alive"]))
    hip_ctx.upload(sid, g["pos"][perm], np.zeros((len(perm), 3)), g["alive"][perm])
    hip_ctx.deposit()
    assert np.array_equal(a, hip_ctx.charge_fixed(sid))


@pytest.mark.parametrize("case", ["1d_periodic", "1d_dirichlet", "2d_pow2", "2d_mixed",
                                  "3d_pow2", "3d_mixed"])
def test_field_solve_matches_reference_golden(hip_ctx, case):
    print(f"case={case}", flush=True)
    g = golden(f"solve_{case}.npz")
    hip_ctx.mesh(g["n"], g["L"], int(g["bc"]), float(g["eps"]))
    hip_ctx.set_rho(g["rho"])
    hip_ctx.solve(float(g["phi_left"]), float(g["phi_right"]), True)
    _, phi, E = hip_ctx.fields()
    assert np.max(np.abs(E - g["E"])) <= 1e-10 * np.max(np.abs(g["E"]))
    assert np.max(np.abs(phi - g["phi"])) <= 1e-10 * np.max(np.abs(g["phi"]))


def _dirichlet_long_double(rho, dx, eps, left, right):
    """The reference's Thomas recurrence (FieldSolver.cpp:733-759) in numpy
    long double (the x87 80-bit format on this host): the reference's
    algorithm with a 64-bit mantissa, a yardstick for an FP64 solve."""
    ld = np.longdouble
    n = len(rho)
    m = n - 2
    d = rho[1:n - 1].astype(ld) * (ld(dx) * ld(dx)) / ld(eps)
    d[0] += ld(left)
    d[m - 1] += ld(right)
    cp = np.zeros(m, dtype=ld)
    dp = np.zeros(m, dtype=ld)
    cp[0], dp[0] = ld(-0.5), d[0] / 2
    for i in range(1, m):
        den = ld(2) + cp[i - 1]
        cp[i] = ld(0) if i == m - 1 else ld(-1) / den
        dp[i] = (d[i] + dp[i - 1]) / den
    phi = np.zeros(n, dtype=ld)
    phi[0], phi[n - 1], phi[n - 2] = left, right, dp[m - 1]
    for i in range(m - 2, -1, -1):
        phi[i + 1] = dp[i] - cp[i] * phi[i + 2]
    E = np.zeros(n, dtype=ld)
    E[1:n - 1] = -(phi[2:] - phi[:n - 2]) / (2 * ld(dx))
    E[0] = -(phi[1] - phi[0]) / ld(dx)
    E[n - 1] = -(phi[n - 1] - phi[n - 2]) / ld(dx)
    return phi, E


@pytest.mark.parametrize("n", [513, 1500, 4097])
def test_dirichlet_solve_large_meshes(oracle, hip_ctx, n):
    """The 1D Dirichlet solve (prefix sums in place of the reference's
    serial recurrence) on meshes up to 4097 nodes: within 1e-12 of the
    reference's recurrence in 80-bit arithmetic, and within 1e-10 of the
    oracle's FP64 recurrence (whose own rounding grows with n)."""
    print(f"n={n}", flush=True)
    rng = np.random.default_rng(n)
    L, eps = 0.067, 8.8541878128e-12
    rho = rng.normal(0.0, 1e-6, n) + 2e-7 * np.sin(np.linspace(0, 7, n))
    hip_ctx.mesh([n], [L], 1, eps)
    hip_ctx.set_rho(rho)
    hip_ctx.solve(0.0, 173.0, True)
    _, phi, E = hip_ctx.fields()
    ref_phi, ref_E = _dirichlet_long_double(rho, L / (n - 1), eps, 0.0, 173.0)
    assert np.max(np.abs(phi - ref_phi)) <= 1e-12 * float(np.max(np.abs(ref_phi)))
    assert np.max(np.abs(E[0] - ref_E)) <= 1e-12 * float(np.max(np.abs(ref_E)))
    o_phi, o_E = oracle.poisson([n], [L], 1, eps, rho, 0.0, 173.0)
    assert np.max(np.abs(phi - o_phi)) <= 1e-10 * np.max(np.abs(o_phi))
    assert np.max(np.abs(E[0] - o_E)) <= 1e-10 * np.max(np.abs(o_E))


def test_field_solve_3d_analytic_mode(hip_ctx):
    # test_core.cpp:1360-1438 on a 32^3 power-of-two mesh
    n, L = [32, 32, 32], [1.0, 1.25, 1.5]
    kx, ky, kz = 2 * np.pi / L[0], 4 * np.pi / L[1], 6 * np.pi / L[2]
    k2 = kx * kx + ky * ky + kz * kz
    z, y, x = [np.arange(n[a]) * L[a] / n[a] for a in (2, 1, 0)]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = np.sin(kx * X) * np.cos(ky * Y) * np.sin(kz * Z)
    hip_ctx.mesh(n, L, 0, 1.0)
    hip_ctx.set_rho(rho.reshape(-1))
    hip_ctx.solve(0, 0, True)
    _, phi, E = hip_ctx.fields()
    assert np.max(np.abs(phi - (rho / k2).reshape(-1))) < 1e-12
    ex = -kx * np.cos(kx * X) * np.cos(ky * Y) * np.sin(kz * Z) / k2
    ez = -kz * np.sin(kx * X) * np.cos(ky * Y) * np.cos(kz * Z) / k2
    assert np.max(np.abs(E[0] - ex.reshape(-1))) < 1e-12
    assert np.max(np.abs(E[2] - ez.reshape(-1))) < 1e-12


def _spectral_reference(rho, n, L, eps):
    """FieldSolver.cpp:442-499, 659-731 in numpy: full complex transform of
    rho - mean, phi_k = rho_k / (eps k^2) with continuous k (Nyquist index
    positive), E_k = -i k phi_k, real part of each inverse."""
    dim = len(n)
    r = rho.reshape(n[::-1])
    R = np.fft.fftn(r - r.mean())
    ks = []
    for a in range(dim):
        idx = np.arange(n[a])
        ks.append(2 * np.pi / L[a] * np.where(idx <= n[a] // 2, idx, idx - n[a]))
    K = np.meshgrid(*ks[::-1], indexing="ij")[::-1]
    k2 = sum(k * k for k in K)
    k2.flat[0] = 1.0
    phi_k = R / (eps * k2)
    phi_k.flat[0] = 0.0
    E = np.stack([np.real(np.fft.ifftn(-1j * K[a] * phi_k)).reshape(-1) for a in range(dim)])
    return np.real(np.fft.ifftn(phi_k)).reshape(-1), E


@pytest.mark.parametrize("n", [[4, 4], [4, 1024], [1024, 4], [64, 8], [4, 4, 4], [4, 8, 64],
                               [128, 4, 4], [8, 4, 16], [32, 4, 8], [256, 16], [512, 8]])
def test_half_spectrum_solve_edge_shapes(hip_ctx, n):
    """The power-of-two solve on half spectra at the edges of its blocking:
    4-point axes (64 lines per block, partial last blocks) and 1024-point
    axes (one line per block), and x axes of every power of two from 4 to
    1024 (the forward x transform's passes of two radix-2 stages end in a
    one-stage pass on odd powers), against the reference's formula
    evaluated with numpy's full complex FFT; E and phi within 1e-11
    relative."""
    print(f"n={n}", flush=True)
    rng = np.random.default_rng(len(n) * 100 + n[0])
    L = [1.0, 1.3, 0.7][:len(n)]
    rho = rng.normal(size=int(np.prod(n)))
    hip_ctx.mesh(n, L, 0, 2.0)
    hip_ctx.set_rho(rho)
    hip_ctx.solve(0.0, 0.0, True)
    _, phi, E = hip_ctx.fields()
    phi_ref, E_ref = _spectral_reference(rho, n, L, 2.0)
    assert np.max(np.abs(phi - phi_ref)) <= 1e-11 * np.max(np.abs(phi_ref))
    assert np.max(np.abs(E[:len(n)] - E_ref)) <= 1e-11 * np.max(np.abs(E_ref))


def _smooth_field(n, L, dim, rng):
    coords = [np.arange(n[a]) * L[a] / n[a] for a in range(dim)]
    grids = np.meshgrid(*coords[::-1], indexing="ij")[::-1]
    E = []
    for a in range(dim):
        f = 0.3 + np.zeros_like(grids[0])
        for b in range(dim):
            f = f + rng.uniform(-1, 1) * np.sin(2 * np.pi * grids[b] / L[b] + rng.uniform(0, 6))
        E.append(f.reshape(-1))
    return np.stack(E)


PUSH_CASES = [
    ("1d_periodic", [64], [2.0], 0, None, None),
    ("1d_dirichlet", [65], [1.0], 1, None, None),
    ("2d_uniform_bz", [32, 24], [1.0, 0.75], 0, [0.0, 0.0, 2.5], None),
    ("3d_leapfrog", [16, 12, 10], [1.0, 0.75, 0.6], 0, None, None),
    ("3d_table_by", [16, 12, 10], [1.0, 0.75, 0.6], 0, None, "table"),
]


@pytest.mark.parametrize("name,n,L,bc,B,btable", PUSH_CASES)
def test_one_step_push_matches_oracle(oracle, hip_ctx, name, n, L, bc, B, btable):
    print(f"case={name}", flush=True)
    rng = np.random.default_rng(7)
    dim = len(n)
    N = 4000
    pos = rng.uniform(0, 1, size=(N, dim)) * np.asarray(L)
    vel = rng.normal(0, 0.4, size=(N, 3))
    E = _smooth_field(n, L, dim, rng)
    qm, dt = -1.3, 0.05
    hip_ctx.mesh(n, L, bc, 1.0)
    if B is not None:
        hip_ctx.uniform_b(B)
    bt = None
    if btable == "table":
        x = np.linspace(0, L[0], 57)
        vals = np.stack([0.2 * np.cos(x), 1.5 + np.sin(3 * x), 0.4 * x], axis=1)
        hip_ctx.b_profile(0, x, vals)
        bt = (0, x, vals)
    sid = hip_ctx.add_species(qm, 1.0, 1.0, N)
    hip_ctx.upload(sid, pos, vel)
    hip_ctx.set_electric(E)
    pos0, vel0, alive0, _ = hip_ctx.download(sid)
    hip_ctx.push(dt)
    pos1, vel1, alive1, _ = hip_ctx.download(sid)
    opos, ovel, oalive, _ = oracle.push(n, L, bc, pos0, vel0, alive0, E, qm, dt, B=B, btable=bt)
    # absorption decisions agree except within rounding of a wall
    scale_x = np.asarray(L)
    mismatch = alive1 != oalive
    if mismatch.any():
        near = np.min(np.minimum(np.abs(opos[mismatch]), np.abs(opos[mismatch] - scale_x)), axis=1)
        assert np.all(near <= 1e-6 * np.max(scale_x))
    keep = (alive1 == 1) & (oalive == 1)
    dx = pos1[keep] - opos[keep]
    if bc == 0:
        dx = (dx + 0.5 * scale_x) % scale_x - 0.5 * scale_x
    assert np.max(np.abs(dx)) <= 1e-6 * np.max(scale_x)
    vscale = np.max(np.abs(ovel[keep]))
    assert np.max(np.abs(vel1[keep] - ovel[keep])) <= 1e-6 * vscale


BOUNDARY_CASES = [
    # (name, nodes, lengths, per-face policy left, right, bottom, top, back, front)
    ("2d_absorb_all", [17, 13], [1.0, 0.75], [1, 1, 1, 1, 1, 1]),
    ("2d_reflect_all", [17, 13], [1.0, 0.75], [2, 2, 2, 2, 2, 2]),
    ("2d_mixed", [17, 13], [1.0, 0.75], [1, 2, 2, 1, 1, 1]),
    ("3d_absorb_all", [9, 11, 7], [1.0, 0.75, 0.5], [1, 1, 1, 1, 1, 1]),
    ("3d_reflect_all", [9, 11, 7], [1.0, 0.75, 0.5], [2, 2, 2, 2, 2, 2]),
    ("3d_mixed", [9, 11, 7], [1.0, 0.75, 0.5], [2, 1, 1, 2, 2, 1]),
]


@pytest.mark.parametrize("name,n,L,pb", BOUNDARY_CASES)
@pytest.mark.parametrize("precision", [32, 64])
def test_particle_boundaries_match_oracle(oracle, hip_ctx, name, n, L, pb, precision):
    """Absorbing and reflecting walls of 2D/3D Dirichlet meshes
    (Simulation3D.cpp:30-84, Simulation2D.cpp:43-97): one push with a field
    and speeds that carry many particles through a wall (some through two
    reflections), against the oracle's restatement: survivors' positions and
    velocities within 1e-6, absorbed counts per face equal (a particle within
    rounding of a wall may go either way)."""
    print(f"case={name} precision={precision}", flush=True)
    rng = np.random.default_rng(21)
    dim = len(n)
    N = 6000
    pos = rng.uniform(0, 1, size=(N, dim)) * np.asarray(L)
    vel = rng.normal(0, 3.0, size=(N, 3))
    E = _smooth_field(n, L, dim, rng)
    qm, dt = -1.3, 0.1
    hip_ctx.mesh(n, L, 1, 1.0)
    hip_ctx.boundaries(pb)
    sid = hip_ctx.add_species(qm, 1.0, 1.0, N, precision)
    hip_ctx.upload(sid, pos, vel)
    hip_ctx.set_electric(E)
    pos0, vel0, alive0, _ = hip_ctx.download(sid)
    hip_ctx.push(dt)
    pos1, vel1, alive1, _ = hip_ctx.download(sid)
    absorbed, _ = hip_ctx.losses(sid)
    opos, ovel, oalive, oabs = oracle.push(n, L, 1, pos0, vel0, alive0, E, qm, dt, pb=pb)
    scale_x = np.asarray(L)
    mismatch = alive1 != oalive
    if mismatch.any():
        p = np.where(mismatch[:, None], opos, 0.0)[mismatch]
        near = np.min(np.minimum(np.abs(p), np.abs(p - scale_x)), axis=1)
        assert np.all(near <= 1e-5 * np.max(scale_x)), near
    assert mismatch.sum() <= 2
    assert np.abs(absorbed - oabs).sum() <= mismatch.sum()
    if 1 in pb:
        assert absorbed.sum() > 100            # walls absorbed particles
    keep = (alive1 == 1) & (oalive == 1)
    assert keep.sum() > N // 4
    tol = 1e-6 if precision == 32 else 1e-10
    assert np.max(np.abs(pos1[keep] - opos[keep])) <= tol * np.max(scale_x)
    assert np.all((pos1[keep] >= 0) & (pos1[keep] <= scale_x))
    vscale = np.max(np.abs(ovel[keep]))
    assert np.max(np.abs(vel1[keep] - ovel[keep])) <= tol * vscale


def test_step_3d_matches_oracle(oracle, hip_ctx):
    """Initialize + one full step (push, deposit, FFT solve) of a two-species
    3D plasma with a tabulated B against the FP64 oracle from the same state."""
    rng = np.random.default_rng(3)
    n, L = [16, 16, 16], [1.0, 1.0, 1.0]
    N = 16 * 16 * 16 * 4
    dt = 0.02
    species = [(-1.0, 1.0, 1.0 / N), (1.0, 50.0, 1.0 / N)]
    x = np.linspace(0, 1.0, 33)
    bvals = np.stack([np.zeros(33), 2.0 * np.exp(-((x - 0.5) / 0.2) ** 2), np.zeros(33)], axis=1)
    hip_ctx.mesh(n, L, 0, 1.0)
    hip_ctx.b_profile(0, x, bvals)
    sids = []
    states = []
    for q, m, w in species:
        pos = rng.uniform(0, 1, size=(N, 3))
        vel = rng.normal(0, 0.3 / math.sqrt(m), size=(N, 3))
        sid = hip_ctx.add_species(q, m, w, N)
        hip_ctx.upload(sid, pos, vel, kind=1)
        sids.append(sid)
    for sid in sids:
        p, v, a, _ = hip_ctx.download(sid)
        states.append([p, v, a])
    hip_ctx.initialize(dt)
    bt = (0, x, bvals)

    def oracle_fields(sts):
        rho = np.zeros(4096)
        for (q, m, w), (p, v, a) in zip(species, sts):
            rho += oracle.deposit_cic(n, L, 0, p, q, w, a)
        _, E = oracle.poisson(n, L, 0, 1.0, rho)
        return rho, E

    _, E0 = oracle_fields(states)
    _, _, Eg0 = hip_ctx.fields()
    assert np.max(np.abs(Eg0 - E0)) <= 1e-6 * np.max(np.abs(E0))
    for (q, m, w), st in zip(species, states):
        st[1] = oracle.init_half(n, L, 0, st[0], st[1], st[2], E0, q / m, dt, btable=bt)
    hip_ctx.step(dt, 1)
    for (q, m, w), st in zip(species, states):
        st[0], st[1], st[2], _ = oracle.push(n, L, 0, st[0], st[1], st[2], E0, q / m, dt,
                                             btable=bt)
    rho1, E1 = oracle_fields(states)
    rg, _, Eg1 = hip_ctx.fields()
    assert np.max(np.abs(Eg1 - E1)) <= 1e-6 * np.max(np.abs(E1))
    for sid, (q, m, w), st in zip(sids, species, states):
        p, v, a, _ = hip_ctx.download(sid)
        d = (p - st[0] + 0.5) % 1.0 - 0.5
        assert np.max(np.abs(d)) <= 1e-6
        assert np.max(np.abs(v - st[1])) <= 1e-6 * np.max(np.abs(st[1]))
        # fused deposit is bit-exact on the device's own post-drift state
        cell, off, _ = hip_ctx.download_raw(sid)
        assert np.array_equal(hip_ctx.charge_fixed(sid), oracle.deposit_fixed(n, 0, cell, off))


def test_sort_preserves_particles_and_charge(hip_ctx):
    rng = np.random.default_rng(9)
    n, L = [16, 8, 8], [2.0, 1.0, 1.0]
    N = 20000
    hip_ctx.mesh(n, L, 0, 1.0)
    sid = hip_ctx.add_species(-1.0, 1.0, 1.0, N)
    hip_ctx.upload(sid, rng.uniform(0, 1, (N, 3)) * L, rng.normal(0, 1, (N, 3)))
    hip_ctx.set_electric(np.zeros((3, 16 * 64)))
    hip_ctx.step(0.01, 3)
    acc = hip_ctx.charge_fixed(sid)
    p0, v0, _, u0 = hip_ctx.download(sid)
    hip_ctx.sort()
    hip_ctx.deposit()
    assert np.array_equal(acc, hip_ctx.charge_fixed(sid))
    p1, v1, _, u1 = hip_ctx.download(sid)
    o0, o1 = np.argsort(u0), np.argsort(u1)
    assert np.array_equal(u0[o0], u1[o1])
    assert np.array_equal(p0[o0], p1[o1]) and np.array_equal(v0[o0], v1[o1])
    # sorted by tile-major cell key: cells are non-decreasing within tiles order
    cell, _, _ = hip_ctx.download_raw(sid)
    assert len(cell) == N


def _tile_keys(cell, cells, tile):
    """Tile-major sort key of linear 3D cell indices (x fastest)."""
    i, j, k = cell % cells[0], (cell // cells[0]) % cells[1], cell // (cells[0] * cells[1])
    tx, ty, tz = [min(t, m) for t, m in zip(tile, cells)]
    ntx, nty = -(-cells[0] // tx), -(-cells[1] // ty)
    return ((k // tz) * nty + j // ty) * ntx + i // tx


@pytest.mark.parametrize("name,n,tile,spread,walls", [
    ("near", (32, 32, 32), (8, 8, 8, 2), 0.5, False),
    ("far", (32, 32, 32), (4, 4, 4, 2), 1.5, False),
    ("two_tiles", (16, 8, 8), (8, 8, 4, 2), 1.0, False),
    ("absorbing", (17, 17, 17), (4, 4, 4, 2), 1.5, True)])
def test_local_sort_matches_radix_sort(name, n, tile, spread, walls):
    """The counting sort of a tile-sorted species (particles that moved into
    a neighbouring tile counted by direction, those that moved further
    placed through a short radix sort, dead ones dropped) against the full
    radix sort, over three sort cycles: the same particles in every tile,
    tile keys non-decreasing in slot order, no dead slot kept, the same
    states and charge by uid; a second run gives the same slot order."""
    from aurora_pic import hip
    print(f"case={name}", flush=True)
    rng = np.random.default_rng(5)
    N = 40000
    L = [1.0, 1.0, 1.0]
    cells = list(n)          # cell indices run over the mesh extents (nodes when walled)
    pos = rng.uniform(0, 1, (N, 3))
    vel = rng.normal(0, spread, (N, 3))
    E = _smooth_field(list(n), L, 3, rng)
    runs = []
    for method in (hip.SORT_RADIX, hip.SORT_LOCAL, hip.SORT_LOCAL):
        with hip.Context(0) as c:
            c.mesh(list(n), L, 1 if walls else 0, 1.0)
            if walls:
                c.boundaries([1] * 6)
            c.tiling(*tile)
            c.sort_method(method)
            sid = c.add_species(-1.0, 1.0, 1.0 / N, N)
            c.upload(sid, pos, vel)
            c.set_electric(E)
            c.sort()
            for _ in range(3):
                for _ in range(4):
                    c.push(0.01)
                c.sort()
            cell, _, _ = c.download_raw(sid)
            p_, v_, a_, u_ = c.download(sid)
            c.deposit()
            runs.append(dict(cell=cell, uid=u_, pos=p_, vel=v_, alive=a_,
                             charge=c.charge_fixed(sid)))
    radix, local, again = runs
    for r in runs:
        assert np.all(r["cell"] >= 0) and np.all(r["alive"] == 1)
        assert np.all(np.diff(_tile_keys(r["cell"], cells, tile)) >= 0)
    assert len(local["uid"]) == len(radix["uid"])
    if walls:
        assert len(local["uid"]) < N            # walls absorbed particles
    keyed = [np.lexsort((r["uid"], _tile_keys(r["cell"], cells, tile))) for r in (radix, local)]
    assert np.array_equal(radix["uid"][keyed[0]], local["uid"][keyed[1]])
    o_r, o_l = np.argsort(radix["uid"]), np.argsort(local["uid"])
    assert np.array_equal(radix["pos"][o_r], local["pos"][o_l])
    assert np.array_equal(radix["vel"][o_r], local["vel"][o_l])
    assert np.array_equal(radix["charge"], local["charge"])
    assert np.array_equal(local["cell"], again["cell"]) and np.array_equal(local["uid"], again["uid"])


@pytest.mark.parametrize("shape", [(8, 8, 8, 2, 512, 0), (8, 8, 8, 1, 512, 0),
                                   (4, 4, 4, 2, 512, 0), (16, 8, 4, 2, 512, 0),
                                   (8, 8, 4, 2, 256, 1), (8, 8, 4, 2, 512, 1),
                                   (8, 8, 8, 2, 256, 1), (8, 8, 4, 2, 512, 3),
                                   (8, 8, 4, 1, 512, 3), (8, 8, 4, 2, 512, 5),
                                   (8, 8, 8, 2, 512, 5), (8, 8, 8, 1, 512, 1),
                                   (8, 8, 8, 1, 512, 5)])
def test_tiled_push_bit_identical_to_untiled(shape):
    """The LDS-tiled 3D push performs the same operations as the untiled
    kernel: particle state and fixed-point charge agree bit for bit, also
    after particles drift beyond the tile halo and across a re-sort."""
    from aurora_pic import hip
    rng = np.random.default_rng(21)
    n, L = [32, 32, 32], [1.0, 1.0, 1.0]
    N = 32 ** 3 * 2
    x = np.linspace(0, 1.0, 65)
    bvals = np.stack([0.1 * np.ones(65), 3.0 * np.exp(-((x - 0.5) / 0.2) ** 2), np.zeros(65)],
                     axis=1)
    pos = rng.uniform(0, 1, (N, 3))
    vel = rng.normal(0, 0.5, (N, 3))   # fast: many particles leave the halo
    results = []
    # tiled re-sorting every 7 steps, tiled never re-sorting (threshold 1:
    # particles stay where the first sort put them), tiled under the
    # adaptive policy (timed, and capped at 5 steps), untiled
    for tiled, threshold, policy in ((True, 0.0, None), (True, 1.0, None), (True, 0.0, 0),
                                     (True, 0.0, 5), (False, 0.0, None)):
        with hip.Context(0) as c:
            c.mesh(n, L, 0, 1.0)
            c.b_profile(0, x, bvals)
            if tiled:
                c.tiling(*shape[:4])
                c.push_variant(*shape[4:])
            else:
                c.tiling(0)
            c.sort_threshold(threshold)
            if policy is not None:
                c.sort_policy(hip.SORT_ADAPTIVE, policy)
            sid = c.add_species(-1.0, 1.0, 1.0 / N, N)
            c.upload(sid, pos, vel, kind=1)
            c.sort()
            c.initialize(0.01)
            c.sort_interval(7)
            c.step(0.01, 16)
            p_, v_, a_, u_ = c.download(sid)
            o = np.argsort(u_)
            results.append((c.charge_fixed(sid), p_[o], v_[o]))
    for other in results[1:]:
        assert np.array_equal(results[0][0], other[0])
        assert np.array_equal(results[0][1], other[1])
        assert np.array_equal(results[0][2], other[2])


@pytest.mark.parametrize("shape", [(8, 8, 1, 2, 512, 1), (8, 8, 1, 2, 512, 0),
                                   (8, 8, 1, 1, 512, 1), (8, 8, 1, 2, 256, 1),
                                   (16, 16, 1, 2, 512, 1), (16, 8, 1, 2, 512, 1)])
@pytest.mark.parametrize("bfield", ["none", "uniform", "table"])
def test_tiled_push_2d_bit_identical_to_untiled(shape, bfield):
    """The LDS-tiled 2D push performs the untiled k_push<2>'s operations:
    particle state and fixed-point charge agree bit for bit with no field,
    a uniform B and a tabulated B(x), across re-sorts (fixed interval and
    the adaptive policy) and with particles leaving the tile halo."""
    from aurora_pic import hip
    print(f"case={shape} {bfield}", flush=True)
    rng = np.random.default_rng(23)
    n, L = [64, 48], [1.0, 0.75]
    N = 64 * 48 * 8
    pos = rng.uniform(0, 1, (N, 2)) * np.asarray(L)
    vel = rng.normal(0, 0.5, (N, 3))
    x = np.linspace(0, 1.0, 65)
    bvals = np.stack([0.1 * np.ones(65), 3.0 * np.exp(-((x - 0.5) / 0.2) ** 2),
                      1.5 * np.ones(65)], axis=1)
    results = []
    for tiled, policy in ((True, None), (True, 5), (False, None)):
        with hip.Context(0) as c:
            c.mesh(n, L, 0, 1.0)
            if bfield == "uniform":
                c.uniform_b([0.0, 0.0, 2.0])
            elif bfield == "table":
                c.b_profile(0, x, bvals)
            if tiled:
                c.tiling(*shape[:4])
                c.push_variant(*shape[4:])
            else:
                c.tiling(0)
            if policy is not None:
                c.sort_policy(hip.SORT_ADAPTIVE, policy)
            sid = c.add_species(-1.0, 1.0, 1.0 / N, N)
            c.upload(sid, pos, vel, kind=1)
            c.sort()
            c.initialize(0.01)
            c.sort_interval(7)
            c.step(0.01, 16)
            p_, v_, a_, u_ = c.download(sid)
            o = np.argsort(u_)
            results.append((c.charge_fixed(sid), p_[o], v_[o]))
    for other in results[1:]:
        assert np.array_equal(results[0][0], other[0])
        assert np.array_equal(results[0][1], other[1])
        assert np.array_equal(results[0][2], other[2])


def test_full_size_3d_invariants(hip_ctx):
    """The 3D3V workload at full size (128^3 cells, 16 + 16 particles per
    cell): charge conservation of the fixed-point deposit, fused deposit equal
    to a separate deposit of the same state, live count."""
    import aurora_pic.workloads as wl
    cfg = wl.config_plasma_3d()
    hip_ctx.mesh(cfg["n"], cfg["L"], 0, cfg["eps"])
    axis, bx, bv = cfg["btable"]
    hip_ctx.b_profile(axis, bx, bv)
    sids = []
    for k, s in enumerate(cfg["species"]):
        sid = hip_ctx.add_species(s["charge"], s["mass"], s["weight"], s["particles"])
        hip_ctx.init_maxwellian(sid, s["particles"], wl.maxwellian_sigma(s["temperature_ev"],
                                                                         s["mass"]), 100 + k)
        sids.append(sid)
    hip_ctx.initialize(cfg["dt"])
    hip_ctx.step(cfg["dt"], 2)
    N = cfg["species"][0]["particles"]
    for sid in sids:
        acc = hip_ctx.charge_fixed(sid)
        # the node total exceeds int64: compare modulo 2^64
        total = int(acc.view(np.uint64).sum(dtype=np.uint64))
        diff = (total - N * 2 ** 40) % 2 ** 64
        diff = diff - 2 ** 64 if diff >= 2 ** 63 else diff
        assert abs(diff) <= N * 4
        slots, live = hip_ctx.count(sid)
        assert live == N
    fused = [hip_ctx.charge_fixed(s) for s in sids]
    hip_ctx.deposit()
    for s, a in zip(sids, fused):
        assert np.array_equal(a, hip_ctx.charge_fixed(s))
    ke, fe, l1, live = hip_ctx.energy(cfg["dt"])
    assert live == 2 * N and np.all(np.isfinite(ke)) and np.isfinite(fe)


def test_error_paths(hip_ctx):
    from aurora_pic import hip
    hip_ctx.mesh([8, 8, 8], [1, 1, 1], 0, 1.0)
    with pytest.raises(hip.HipError, match="does not cover the simulation domain"):
        hip_ctx.b_profile(0, [0.1, 0.5], [[0, 0, 1], [0, 0, 2]])
    with pytest.raises(hip.HipError, match="strictly increasing"):
        hip_ctx.b_profile(0, [0.0, 0.0, 1.0], [[0, 0, 1]] * 3)
    with pytest.raises(hip.HipError, match="permittivity"):
        hip_ctx.mesh([8], [1.0], 0, 0.0)


@pytest.mark.parametrize("n", [[16, 16, 16], [32, 16]])
def test_fields_after_steps_are_current(oracle, n):
    """Steps skip rho, phi and the FP64 field when no kernel reads them; the
    readers recompute them from the last deposit.  The fields read after a
    step are the solve of that step's charge, and reading them in the middle
    of a run does not change the run (bit for bit)."""
    from aurora_pic import hip
    rng = np.random.default_rng(8)
    dim = len(n)
    L = [1.0] * dim
    N = int(np.prod(n)) * 4
    pos = rng.uniform(0, 1, (N, dim))
    vel = rng.normal(0, 0.3, (N, 3))
    finals = []
    for peek in (False, True):
        with hip.Context(0) as c:
            c.mesh(n, L, 0, 1.0)
            sid = c.add_species(-1.0, 1.0, 1.0 / N, N)
            c.upload(sid, pos, vel, kind=1)
            c.sort()
            c.initialize(0.01)
            c.step(0.01, 5)
            if peek:
                rho, phi, E = c.fields()
                phi_o, E_o = oracle.poisson(n, L, 0, 1.0, rho)
                assert np.max(np.abs(E.reshape(-1) - E_o.reshape(-1))) <= \
                    1e-10 * np.max(np.abs(E_o))
                assert np.max(np.abs(phi - phi_o)) <= 1e-10 * np.max(np.abs(phi_o))
                ke, fe, l1, live = c.energy(0.01)
                assert live == N and fe > 0.0
            c.step(0.01, 5)
            p_, v_, a_, u_ = c.download(sid)
            o = np.argsort(u_)
            finals.append((c.charge_fixed(sid), p_[o], v_[o]))
    assert np.array_equal(finals[0][0], finals[1][0])
    assert np.array_equal(finals[0][1], finals[1][1])
    assert np.array_equal(finals[0][2], finals[1][2])


@pytest.mark.parametrize("dim", [2, 3])
def test_tiled_push_ragged_and_empty_tiles(dim):
    """Ragged inputs for the tiled pushes and the local sort: every particle
    in one tile (all other tiles empty), a species with one particle and an
    empty species beside it; results equal the untiled push bit for bit,
    across re-sorts."""
    from aurora_pic import hip
    print(f"case=dim{dim}", flush=True)
    rng = np.random.default_rng(31)
    n = [32, 32] if dim == 2 else [16, 16, 16]
    L = [1.0] * dim
    N = 5000
    pos = 0.1 + 0.05 * rng.uniform(0, 1, (N, dim))     # all inside the first tile
    vel = rng.normal(0, 0.05, (N, 3))
    shape = (8, 8, 1, 2, 512, 1) if dim == 2 else (8, 8, 4, 2, 512, 5)
    results = []
    for tiled in (True, False):
        with hip.Context(0) as c:
            c.mesh(n, L, 0, 1.0)
            c.uniform_b([0.0, 0.5, 1.0])
            if tiled:
                c.tiling(*shape[:4])
                c.push_variant(*shape[4:])
            else:
                c.tiling(0)
            s0 = c.add_species(-1.0, 1.0, 1.0 / N, N)
            s1 = c.add_species(1.0, 10.0, 1.0, 4)
            s2 = c.add_species(1.0, 10.0, 1.0, 4)
            c.upload(s0, pos, vel, kind=1)
            c.upload(s1, np.full((1, dim), 0.7), np.array([[0.3, -0.2, 0.1]]), kind=1)
            c.upload(s2, np.zeros((0, dim)), np.zeros((0, 3)), kind=1)
            c.sort()
            c.initialize(0.01)
            c.sort_interval(5)
            c.step(0.01, 12)
            out = []
            for sid in (s0, s1, s2):
                p_, v_, a_, u_ = c.download(sid)
                o = np.argsort(u_)
                out.append((c.charge_fixed(sid), p_[o], v_[o]))
            assert c.count(s2)[0] == 0 and c.count(s1)[1] == 1
            results.append(out)
    for a, b in zip(results[0], results[1]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_local_sort_when_every_particle_dies():
    """Absorbing walls take every particle: the local sort leaves an empty,
    tile-sorted species, and stepping on is well defined."""
    from aurora_pic import hip
    rng = np.random.default_rng(37)
    N = 3000
    with hip.Context(0) as c:
        c.mesh([17, 17, 17], [1.0, 1.0, 1.0], 1, 1.0)
        c.boundaries([1] * 6)
        c.tiling(4, 4, 4, 2)
        sid = c.add_species(-1.0, 1.0, 1.0 / N, N)
        c.upload(sid, rng.uniform(0, 1, (N, 3)), rng.normal(0, 1.0e5, (N, 3)))
        c.set_electric(np.zeros((3, 17 ** 3)))
        c.sort()
        c.push(0.01)
        absorbed, _ = c.losses(sid)
        assert absorbed.sum() == N
        c.sort()                                  # the local counting sort
        assert c.count(sid) == (0, 0)
        c.push(0.01)
        c.sort()
        assert c.count(sid) == (0, 0)


def test_autotune_picks_a_refined_candidate():
    """The autotuner times every candidate, re-times the fastest over a
    longer window and leaves the best of those configured; the tuned
    context's steps equal an untiled context's bit for bit."""
    from aurora_pic import hip
    from aurora_pic import autotune
    rng = np.random.default_rng(31)
    n, L = [32, 32, 32], [1.0, 1.0, 1.0]
    N = 32 ** 3 * 4
    pos = rng.uniform(0, 1, (N, 3))
    vel = rng.normal(0, 0.3, (N, 3))
    shapes = [(8, 8, 8, 2, 512, 0), (8, 8, 4, 2, 512, 1), (8, 8, 4, 2, 512, 5)]
    results = []
    for tiled in (True, False):
        with hip.Context(0) as c:
            c.mesh(n, L, 0, 1.0)
            sid = c.add_species(-1.0, 1.0, 1.0 / N, N)
            c.upload(sid, pos, vel, kind=1)
            c.initialize(0.01)
            if tiled:
                logs = []
                best, table = autotune.tune(c, 0.01, shapes=shapes, steps=4, refine=2,
                                            adaptive=True, log=logs.append)
                assert len(table) == 3 and best[1] in shapes
                assert sum("refined" in line for line in logs) == 2
                assert best in sorted(table)[:3]
                steps_done = 3 * 5 + 2 * 11
            else:
                c.tiling(0)
                c.step(0.01, 3 * 5 + 2 * 11)
                steps_done = 3 * 5 + 2 * 11
            c.step(0.01, 5)
            p_, v_, a_, u_ = c.download(sid)
            o = np.argsort(u_)
            results.append((steps_done, c.charge_fixed(sid), p_[o], v_[o]))
    assert results[0][0] == results[1][0]
    assert np.array_equal(results[0][1], results[1][1])
    assert np.array_equal(results[0][2], results[1][2])
    assert np.array_equal(results[0][3], results[1][3])


@pytest.mark.parametrize("n", [4, 96, 100, 256, 1024, 2048])
def test_spectral_1d_solve_sizes(hip_ctx, n):
    """The 1D periodic solve on power-of-two meshes (radix-2 FFTs in LDS,
    tables in LDS up to 1024 nodes) and on others (direct sums), against
    the reference's formula evaluated with numpy's FFT: phi and E within
    1e-11 relative."""
    print(f"n={n}", flush=True)
    rng = np.random.default_rng(n)
    L = [2.7]
    rho = rng.normal(size=n) + 0.3 * np.sin(np.linspace(0, 5, n))
    hip_ctx.mesh([n], L, 0, 1.3)
    hip_ctx.set_rho(rho)
    hip_ctx.solve(0.0, 0.0, True)
    _, phi, E = hip_ctx.fields()
    phi_ref, E_ref = _spectral_reference(rho, [n], L, 1.3)
    assert np.max(np.abs(phi - phi_ref)) <= 1e-11 * np.max(np.abs(phi_ref))
    assert np.max(np.abs(E[0] - E_ref[0])) <= 1e-11 * np.max(np.abs(E_ref))
