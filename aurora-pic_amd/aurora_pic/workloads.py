"""Synthetic inputs for the five measured configurations.

Every parameter the configuration descriptions leave open is fixed here and
documented; the same builders feed the parity tests and bench.py.

Physical constants are the CODATA values the reference uses
(include/pic/Types.hpp:7-10, src/Swarm.cpp:23-26).
"""
from __future__ import annotations

import math

import numpy as np

EV = 1.602176634e-19          # J per eV, elementary charge (C)
ME = 9.1093837139e-31         # electron mass (kg)
EPS0 = 8.8541878128e-12       # vacuum permittivity (F/m)
AMU = 1.66053906660e-27       # kg
KB = 1.380649e-23             # J/K
HELIUM_MASS = 4.002602 * AMU  # synthetic gas neutral mass (kg)

ELASTIC, EXCITATION, IONIZATION, ATTACHMENT = 0, 1, 2, 3


# --------------------------------------------------------------- gas tables
def synthetic_gas_tables(rows: int = 1000):
    """Elastic (flat 6e-20 m^2), excitation (19.8 eV) and ionization
    (24.6 eV) on ``rows`` log-spaced energies 1e-3..1e3 eV.  Inelastic
    channels use the Lotz form sigma = A ln(E/I) / (E I) above threshold
    (A = 3e-18 eV^2 m^2 excitation, 6e-18 ionization; peaks near 2.8e-21 and
    3.6e-21 m^2).  Energies are returned in joules, as the reference's
    tables after energy_scale (Collision.cpp:159-219)."""
    e_ev = np.geomspace(1.0e-3, 1.0e3, rows)
    elastic = np.full(rows, 6.0e-20)

    def lotz(a, thr):
        s = np.zeros(rows)
        m = e_ev > thr
        s[m] = a * np.log(e_ev[m] / thr) / (e_ev[m] * thr)
        return s

    e_j = e_ev * EV
    return [
        (ELASTIC, 0.0, e_j, elastic),
        (EXCITATION, 19.8 * EV, e_j.copy(), lotz(3.0e-18, 19.8)),
        (IONIZATION, 24.6 * EV, e_j.copy(), lotz(6.0e-18, 24.6)),
    ]


def max_collision_frequency(channels, neutral_density, mass=ME, margin=1.05):
    """nu_max = margin * max over table rows of n * sum(sigma) * v.  With
    sigma held constant past the last row (Collision.cpp:226-227) the bound
    stays valid up to margin^2 times the last tabulated energy."""
    e = channels[0][2]
    total = np.zeros_like(e)
    for _, _, ec, sc in channels:
        total += np.interp(e, ec, sc)
    v = np.sqrt(2.0 * e / mass)
    return margin * float(np.max(neutral_density * total * v))


def synthetic_gas(rows: int = 1000, neutral_density: float = 1.0e21):
    channels = synthetic_gas_tables(rows)
    return dict(channels=channels, neutral_density=neutral_density,
                neutral_mass=HELIUM_MASS, particle_mass=ME,
                max_frequency=max_collision_frequency(channels, neutral_density),
                max_candidates=64)


# ------------------------------------------------------------ B profile
def gaussian_by_profile(length_x: float, rows: int = 512, peak: float = 0.02,
                        width_fraction: float = 0.125):
    """B_y(x) = peak * exp(-((x - Lx/2) / (width_fraction Lx))^2) on ``rows``
    equally spaced x covering [0, Lx] (table format x Bx By Bz,
    PrescribedField.cpp:136-183)."""
    x = np.linspace(0.0, length_x, rows)
    by = peak * np.exp(-(((x - 0.5 * length_x) / (width_fraction * length_x)) ** 2))
    values = np.zeros((rows, 3))
    values[:, 1] = by
    return x, values


# -------------------------------------------------------------- configs
def debye_length(n, te_ev):
    return math.sqrt(EPS0 * te_ev / (n * EV))


def plasma_frequency(n, mass=ME):
    return math.sqrt(n * EV * EV / (EPS0 * mass))


def config_two_stream():
    """1D1V two-stream, normalised units, FP64 particles."""
    L = 2.0 * math.pi / 0.5
    N = 25600
    return dict(name="two_stream_1d1v", dim=1, n=[256], L=[L], bc=0, eps=1.0,
                dt=0.1, steps=2000, seed=12345, precision="f64",
                species=[dict(name="electrons", charge=-1.0, mass=1.0, weight=L / N,
                              particles=N)],
                beams=(3.0, 0.1), perturbation=1.0e-3)


def config_discharge():
    """1D3V bounded RF discharge with MCC (helium-like synthetic gas)."""
    L = 0.067
    freq = 13.56e6
    ne = 5 * 10 ** 5
    n0 = 2.56e14
    return dict(name="rf_discharge_1d3v", dim=1, n=[513], L=[L], bc=1, eps=EPS0,
                dt=1.0 / (400.0 * freq), steps=10000, seed=2026, precision="f32",
                phi_left=0.0, phi_right_amplitude=200.0, frequency=freq,
                species=[dict(name="electrons", charge=-EV, mass=ME, weight=n0 * L / ne,
                              particles=ne, temperature_ev=2.0),
                         dict(name="ions", charge=EV, mass=6.67e-27, weight=n0 * L / ne,
                              particles=ne, temperature_ev=0.026)],
                gas=synthetic_gas(1000, 1.0e21), max_particles_per_species=10 ** 7)


def config_langmuir_2d():
    """2D3V periodic Langmuir wave with uniform B_z."""
    n = 1.0e16
    lam = debye_length(n, 1.0)
    wp = plasma_frequency(n)
    cells = 256
    L = cells * lam
    N = 64 * cells * cells
    return dict(name="langmuir_2d3v", dim=2, n=[cells, cells], L=[L, L], bc=0, eps=EPS0,
                dt=0.05 / wp, steps=1000, seed=4242, precision="f32", depth=1.0,
                B=[0.0, 0.0, 0.01],
                species=[dict(name="electrons", charge=-EV, mass=ME,
                              weight=n * L * L * 1.0 / N, particles=N,
                              temperature_ev=1.0)],
                perturbation=0.01)


def config_plasma_3d(cells: int = 128, ppc: int = 16):
    """3D3V periodic electron-ion plasma with tabulated B_y(x)."""
    n = 1.0e16
    lam = debye_length(n, 5.0)
    wp = plasma_frequency(n)
    L = cells * lam
    N = ppc * cells ** 3
    w = n * L ** 3 / N
    bx, bv = gaussian_by_profile(L)
    return dict(name="plasma_3d3v", dim=3, n=[cells] * 3, L=[L] * 3, bc=0, eps=EPS0,
                dt=0.05 / wp, steps=500, sort_interval=20, seed=777, precision="f32",
                btable=(0, bx, bv),
                species=[dict(name="electrons", charge=-EV, mass=ME, weight=w, particles=N,
                              temperature_ev=5.0),
                         dict(name="ions", charge=EV, mass=131.0 * AMU, weight=w,
                              particles=N, temperature_ev=0.1)])


def config_swarm_requests(count: int = 4096, seed: int = 99):
    """Poisson-arriving swarm requests: E/N log-uniform 1-1000 Td, particle
    counts log-normal (median 5e4, sigma 0.8, clipped to [1e4, 1e6])."""
    rng = np.random.default_rng(seed)
    gas = synthetic_gas(1000, 1.0e21)
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, size=count))
    en_td = np.exp(rng.uniform(math.log(1.0), math.log(1000.0), size=count))
    particles = np.clip(np.exp(rng.normal(math.log(5.0e4), 0.8, size=count)), 1e4, 1e6)
    return dict(name="swarm_serving", gas=gas, arrival_s=arrivals, reduced_field_td=en_td,
                particles=particles.astype(np.int64), steps=20000, burn_in=10000,
                blocks=10, timestep=0.099 / gas["max_frequency"],
                initial_mean_energy_ev=1.0, max_energy_ev=1000.0, seed=seed)


def maxwellian_sigma(temperature_ev, mass):
    return math.sqrt(temperature_ev * EV / mass)
