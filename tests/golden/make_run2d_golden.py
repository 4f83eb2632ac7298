"""Write tests/golden/run2d/<case>/ from the REAL reference.

A periodic 2D3V plasma in normalized units with a uniform B_z (the
reference's 2D3V Boris push): case.txt (read the same way by
tests/test_gpu_run3d.py and the host test program's --run2d mode) and the
scalars.csv the reference's own Simulation2D::run() writes for it (compiled
unmodified into oracle/_ref by oracle/Makefile).  Both sides load the
species from the config seed with std::mt19937_64 (Species2D.cpp:29-75).

    python tests/golden/make_run2d_golden.py
"""
import os
import shutil
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle_py  # noqa: E402

OUT = os.path.join(HERE, "run2d")


def main():
    N = 64 * 64 * 4
    w = 4.0 * 4.0 * 0.5 / N
    case = dict(n=[64, 64], length=[4.0, 4.0], depth=0.5, dt=0.05, steps=60, output_interval=20,
                seed=13, B=[0.0, 0.0, 0.3],
                species=[dict(name="electrons", charge=-1.0, mass=1.0, weight=w, particles=N,
                              thermal_velocity=0.2),
                         dict(name="ions", charge=1.0, mass=100.0, weight=w, particles=N,
                              thermal_velocity=0.02)])
    d = os.path.join(OUT, "plasma2d_bz")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "case.txt"), "w") as f:
        f.write("n " + " ".join(str(v) for v in case["n"]) + "\n")
        f.write("length " + " ".join(repr(float(v)) for v in case["length"]) + "\n")
        f.write("B " + " ".join(repr(float(v)) for v in case["B"]) + "\n")
        for k in ("depth", "dt", "steps", "output_interval", "seed"):
            f.write(f"{k} {case[k]!r}\n")
        for s in case["species"]:
            f.write(f"species {s['name']} {s['charge']!r} {s['mass']!r} {s['weight']!r} "
                    f"{s['particles']} {s['thermal_velocity']!r}\n")
    tmp = tempfile.mkdtemp(prefix="run2d_golden_")
    try:
        oracle_py.Reference().run2d(case, tmp)
        shutil.copy(os.path.join(tmp, "scalars.csv"), os.path.join(d, "scalars.csv"))
    finally:
        shutil.rmtree(tmp)
    print("wrote", d)
    print(open(os.path.join(d, "scalars.csv")).read())


if __name__ == "__main__":
    main()
