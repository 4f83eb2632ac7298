"""Write tests/golden/run1d/<case>/ from the REAL reference.

Each case directory holds a case.txt (mesh, drive, species; read the same
way by tests/test_gpu_run1d.py and by the host test program's --run1d mode)
and the outputs of the reference's own 1D Simulation::run() on it, compiled
unmodified into oracle/_ref by oracle/Makefile: scalars.csv,
boundary_losses.csv, power_transfer.csv and the first and last
fields_<step>.csv.

  rf_1d1v_aps     1D1V, the initial state from a .aps v1 file written here
                  (state.aps): the identical-input channel of the reference.
  rf_1d3v_seeded  1D3V, both sides load the species from the config seed
                  (uniform positions in [init_x_min, init_x_max], Maxwellian
                  velocities from std::mt19937_64).

Both are collisionless RF discharges between absorbing walls, so every
output column is a deterministic function of the input.

    python tests/golden/make_run1d_golden.py
"""
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle_py  # noqa: E402

OUT = os.path.join(HERE, "run1d")
KEEP = ("scalars.csv", "boundary_losses.csv", "power_transfer.csv")

BASE = dict(nx=65, length=1.0, dt=0.02, steps=400, output_interval=100, dirichlet=1,
            electrodes=[0.0, 0.0, 0.0, 0.0, 0.0, 0.002, 0.05, 0.0])


def species(n, vth_e=0.05, vth_i=0.005):
    w = 1.0 / n
    return [dict(name="electrons", charge=-1.0, mass=1.0, weight=w, particles=n,
                 drift_velocity=0.0, thermal_velocity=vth_e, init_x_min=0.05, init_x_max=0.95),
            dict(name="ions", charge=1.0, mass=100.0, weight=w, particles=n,
                 drift_velocity=0.0, thermal_velocity=vth_i, init_x_min=0.05, init_x_max=0.95)]


def write_case(path, case, aps=None):
    with open(path, "w") as f:
        for k in ("nx", "length", "dt", "steps", "output_interval", "dirichlet",
                  "velocity_dimensions", "seed"):
            f.write(f"{k} {case[k]!r}\n")
        f.write("electrodes " + " ".join(repr(float(v)) for v in case["electrodes"]) + "\n")
        if aps:
            f.write(f"initial_state {aps}\n")
        for s in case["species"]:
            f.write("species " + s["name"] + " " + " ".join(
                repr(float(s[k])) if k != "particles" else str(s[k])
                for k in oracle_py.Reference.RUN1D_SPECIES_KEYS) + "\n")


def write_aps(path, case, rng):
    """An .aps v1 file (ParticleState.cpp:119-540 format) for a 1D1V case."""
    rows = []
    for s in case["species"]:
        x = rng.uniform(s["init_x_min"], s["init_x_max"], s["particles"])
        v = rng.normal(s["drift_velocity"], s["thermal_velocity"], s["particles"])
        rows += [f"particle {s['name']} {float(a)!r} 0 0 {float(b)!r} 0 0" for a, b in zip(x, v)]
    with open(path, "w") as f:
        f.write("AuroraPIC-particle-state-v1\ndimension 1\nunits normalized\n"
                "weighting species_constant\nvelocity_staggering time_centered\n"
                f"particle_count {len(rows)}\nrecords\n")
        f.write("\n".join(rows) + "\nend\n")


def run(ref, name, case, aps=None):
    d = os.path.join(OUT, name)
    os.makedirs(d, exist_ok=True)
    write_case(os.path.join(d, "case.txt"), case, "state.aps" if aps else None)
    tmp = tempfile.mkdtemp(prefix="run1d_golden_")
    try:
        ref.run1d(case, tmp, os.path.join(d, "state.aps") if aps else "")
        for f in KEEP + ("fields_0.csv", f"fields_{case['steps']}.csv"):
            shutil.copy(os.path.join(tmp, f), os.path.join(d, f))
    finally:
        shutil.rmtree(tmp)
    print("wrote", d)


def main():
    ref = oracle_py.Reference()
    case = dict(BASE, velocity_dimensions=1, seed=5, species=species(1500))
    d = os.path.join(OUT, "rf_1d1v_aps")
    os.makedirs(d, exist_ok=True)
    write_aps(os.path.join(d, "state.aps"), case, np.random.default_rng(20260))
    run(ref, "rf_1d1v_aps", case, aps=True)
    case = dict(BASE, velocity_dimensions=3, seed=11, species=species(1500))
    run(ref, "rf_1d3v_seeded", case)


if __name__ == "__main__":
    main()
