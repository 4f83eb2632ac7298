"""External particle state (.aps version 1) in the C++ host library (CPU).

- The host-only cases of host/tests/host_tests.cpp: the reference's example
  file (examples/external_state_1d.aps, kept in tests/golden/aps) has the
  signature its example config pins (15689689587873937355); the reference's
  round-trip, validation and writer tests (test_initialization.cpp:786-930);
  malformed files fail with the reference's messages.
- Signatures and error texts against the reference's own reader
  (oracle/_ref, where built) on generated files: interleaved species,
  1D/2D/3D, values that exercise the binary64 text round trip."""
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(REPO, "aurora-pic_amd", "lib", "aurora_host_tests")
GOLDEN = os.path.join(REPO, "tests", "golden")


def _signature(path):
    out = subprocess.run([TESTS, "--aps-signature", path], capture_output=True, text=True,
                         timeout=60).stdout.strip()
    if out.startswith("error: "):
        return None, out[len("error: "):]
    return int(out), None


def test_host_state_file_cases():
    out = subprocess.run([TESTS, "--cpu", GOLDEN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 failed" in out.stdout and "PASS large_run_guard" in out.stdout


def test_reference_example_signature():
    sig, err = _signature(os.path.join(GOLDEN, "aps", "external_state_1d.aps"))
    assert err is None and sig == 15689689587873937355


def _write(path, dim, rows, units="normalized", head=None, tail="end\n"):
    with open(path, "w") as f:
        f.write(head if head is not None else
                f"AuroraPIC-particle-state-v1\ndimension {dim}\nunits {units}\n"
                f"weighting species_constant\nvelocity_staggering time_centered\n"
                f"particle_count {len(rows)}\nrecords\n")
        for r in rows:
            f.write(r + "\n")
        f.write(tail)


def _random_rows(rng, dim, n, names):
    rows = []
    for _ in range(n):
        x = rng.uniform(0, 1, 3)
        v = rng.normal(0, 1, 3)
        if dim < 3:
            x[2] = 0.0
        if dim == 1:
            x[1] = 0.0
            v[1] = v[2] = 0.0
        name = names[rng.integers(len(names))]
        vals = " ".join(repr(float(t)) for t in list(x) + list(v))
        rows.append(f"particle {name} {vals}")
    return rows


def test_signatures_match_reference_reader(tmp_path, reference):
    rng = np.random.default_rng(33)
    for k, dim in enumerate([1, 2, 3, 3, 2]):
        p = str(tmp_path / f"s{k}.aps")
        rows = _random_rows(rng, dim, 50 + 7 * k, ["electrons", "ions", "Xe+"][:1 + k % 3])
        _write(p, dim, rows, units="si" if k % 2 else "normalized")
        mine, err = _signature(p)
        theirs, rerr = reference.aps_signature(p)
        assert err is None and rerr is None, (err, rerr)
        assert mine == theirs, (dim, mine, theirs)


MALFORMED = [
    ("AuroraPIC-particle-state-v2\n", None),
    ("AuroraPIC-particle-state-v1\ndimension 5\n", None),
    ("AuroraPIC-particle-state-v1\ndimension 2\nunits cgs\n", None),
    ("AuroraPIC-particle-state-v1\ndimension 2\nunits si\nweighting per_particle\n", None),
    ("AuroraPIC-particle-state-v1\ndimension 2\nunits si\nweighting species_constant\n"
     "velocity_staggering half_step\n", None),
    (None, ["particle e 0.1 0.2 0.3 1 2 3"]),             # inactive z in 2D
    (None, ["particle e 0.1 0.2 0 1 2 inf"]),             # not finite
    (None, ["particle e 0.1 0.2 0 1 2"]),                 # short record
]


@pytest.mark.parametrize("head,rows", MALFORMED)
def test_malformed_files_fail_like_reference(tmp_path, reference, head, rows):
    p = str(tmp_path / "bad.aps")
    if rows is None:
        _write(p, 2, [], head=head, tail="")
    else:
        _write(p, 2, rows)
    mine, err = _signature(p)
    theirs, rerr = reference.aps_signature(p)
    assert mine is None and theirs is None
    assert err == rerr, (err, rerr)
