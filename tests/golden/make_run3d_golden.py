"""Write tests/golden/run3d/<case>/ from the REAL reference.

A periodic 3D3V plasma in normalized units with a tabulated B_y(x) (the
shape of the benchmark workload, scaled down): case.txt (read the same way
by tests/test_gpu_run3d.py and the host test program's --run3d mode) and
the scalars.csv the reference's own Simulation3D::run() writes for it
(compiled unmodified into oracle/_ref by oracle/Makefile).  Both sides load
the species from the config seed with std::mt19937_64 (Species3D.cpp:28-76),
so they start from the same particles.

    python tests/golden/make_run3d_golden.py
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

OUT = os.path.join(HERE, "run3d")


def case_plasma():
    n, L = [16, 16, 16], [4.0, 4.0, 4.0]
    N = 16 ** 3 * 4
    x = np.linspace(0.0, 4.0, 65)
    b = np.stack([0.05 * np.ones(65), 0.5 * np.exp(-((x - 2.0) / 0.8) ** 2), np.zeros(65)], 1)
    w = 4.0 ** 3 / N
    return dict(n=n, length=L, dt=0.05, steps=60, output_interval=20, seed=9, b_axis=0,
                bcoords=x, bvalues=b,
                species=[dict(name="electrons", charge=-1.0, mass=1.0, weight=w, particles=N,
                              thermal_velocity=0.2),
                         dict(name="ions", charge=1.0, mass=100.0, weight=w, particles=N,
                              thermal_velocity=0.02)])


def write_case(d, case):
    with open(os.path.join(d, "btable.txt"), "w") as f:
        for xx, row in zip(case["bcoords"], case["bvalues"]):
            f.write(" ".join(repr(float(v)) for v in [xx, *row]) + "\n")
    with open(os.path.join(d, "case.txt"), "w") as f:
        f.write("n " + " ".join(str(v) for v in case["n"]) + "\n")
        f.write("length " + " ".join(repr(float(v)) for v in case["length"]) + "\n")
        for k in ("dt", "steps", "output_interval", "seed"):
            f.write(f"{k} {case[k]!r}\n")
        f.write(f"btable {case['b_axis']} btable.txt\n")
        for s in case["species"]:
            f.write(f"species {s['name']} {s['charge']!r} {s['mass']!r} {s['weight']!r} "
                    f"{s['particles']} {s['thermal_velocity']!r}\n")


def main():
    ref = oracle_py.Reference()
    case = case_plasma()
    d = os.path.join(OUT, "plasma3d_btable")
    os.makedirs(d, exist_ok=True)
    write_case(d, case)
    tmp = tempfile.mkdtemp(prefix="run3d_golden_")
    try:
        ref.run3d(case, tmp)
        shutil.copy(os.path.join(tmp, "scalars.csv"), os.path.join(d, "scalars.csv"))
    finally:
        shutil.rmtree(tmp)
    print("wrote", d)
    print(open(os.path.join(d, "scalars.csv")).read())


if __name__ == "__main__":
    main()
