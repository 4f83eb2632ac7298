"""Write tests/golden/swarm/synthetic_swarm_golden.npz from the REAL reference.

Runs the reference's own ``load_swarm_benchmark_config`` and
``run_swarm_benchmark`` (compiled unmodified into oracle/_ref by
oracle/Makefile) on tests/golden/swarm/synthetic_swarm.swarm -- the
reference's examples/synthetic_swarm.swarm data files, kept here as
fixtures: 256 electrons, 1/5/10 Td, 300 K neutrals, Henyey-Greenstein
elastic scattering.  The reference's own result CSV of that run is kept
beside it (synthetic_swarm_reference.csv) for the column-by-column check of
the CSV writer.  Because the device's random stream differs from
std::mt19937_64, the fixture also records the spread of the reference
itself over independent seeds, which the statistical parity test uses.

    python tests/golden/make_swarm_golden.py
"""
import os
import re
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle_py  # noqa: E402

SWARM = os.path.join(HERE, "swarm", "synthetic_swarm.swarm")
SEEDS = 16


def with_seed(src_dir, seed):
    """A copy of the example directory whose .swarm file uses another seed."""
    tmp = tempfile.mkdtemp(prefix="swarm_golden_")
    for name in os.listdir(src_dir):
        shutil.copy(os.path.join(src_dir, name), tmp)
    path = os.path.join(tmp, "synthetic_swarm.swarm")
    text = open(path).read()
    text = re.sub(r"(?m)^seed\s*=.*$", f"seed = {seed}", text)
    open(path, "w").write(text)
    return tmp, path


def main():
    ref = oracle_py.Reference()
    csv = os.path.join(HERE, "swarm", "synthetic_swarm_reference.csv")
    example = ref.swarm_file(SWARM, csv_path=csv)
    runs = []
    for k in range(SEEDS):
        tmp, path = with_seed(os.path.dirname(SWARM), 1_000_003 * (k + 1))
        try:
            runs.append(ref.swarm_file(path))
        finally:
            shutil.rmtree(tmp)
    runs = np.stack(runs)                       # [seed, field, col]
    out = os.path.join(HERE, "swarm", "synthetic_swarm_golden.npz")
    np.savez_compressed(out, columns=np.asarray(oracle_py.Reference.SWARM_FILE_COLS),
                        example=example, seed_mean=runs.mean(axis=0),
                        seed_std=runs.std(axis=0, ddof=1), seeds=SEEDS)
    print("wrote", out)
    for row in example:
        print(" ".join(f"{v:.6g}" for v in row))


if __name__ == "__main__":
    main()
