"""Write the swarm golden fixtures from the REAL reference:
tests/golden/swarm/synthetic_swarm_golden.npz (fixed population) and
tests/golden/swarm_branching/branching_golden.npz (branching / resampled
population on a synthetic attaching and ionizing gas: 60 Td attachment only,
300 Td mixed, 1000 Td ionization dominated).

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

EXAMPLES = [("swarm", "synthetic_swarm.swarm", "synthetic_swarm_golden.npz",
             "synthetic_swarm_reference.csv"),
            ("swarm_branching", "branching.swarm", "branching_golden.npz",
             "branching_reference.csv")]
SEEDS = 16


def with_seed(src_dir, name, seed):
    """A copy of the example directory whose .swarm file uses another seed."""
    tmp = tempfile.mkdtemp(prefix="swarm_golden_")
    for f in os.listdir(src_dir):
        shutil.copy(os.path.join(src_dir, f), tmp)
    path = os.path.join(tmp, name)
    text = open(path).read()
    text = re.sub(r"(?m)^seed\s*=.*$", f"seed = {seed}", text)
    open(path, "w").write(text)
    return tmp, path


def main():
    ref = oracle_py.Reference()
    for sub, name, npz, csv_name in EXAMPLES:
        src = os.path.join(HERE, sub)
        example = ref.swarm_file(os.path.join(src, name), csv_path=os.path.join(src, csv_name))
        runs = []
        for k in range(SEEDS):
            tmp, path = with_seed(src, name, 1_000_003 * (k + 1))
            try:
                runs.append(ref.swarm_file(path))
            finally:
                shutil.rmtree(tmp)
        runs = np.stack(runs)                   # [seed, field, col]
        out = os.path.join(src, npz)
        np.savez_compressed(out, columns=np.asarray(oracle_py.Reference.SWARM_FILE_COLS),
                            example=example, seed_mean=runs.mean(axis=0),
                            seed_std=runs.std(axis=0, ddof=1), seeds=SEEDS)
        print("wrote", out)
        for row in example:
            print(" ".join(f"{v:.6g}" for v in row))


if __name__ == "__main__":
    main()
