"""CPU checks of the C-ABI boundary: the library loads without a GPU and
exports every function include/aurora_hip.h declares (no compute calls)."""
import ctypes
import os
import subprocess

import pytest

from aurora_pic import hip


def test_header_declares_the_boundary():
    names = hip.declared_symbols()
    for required in ["aurora_hip_create", "aurora_hip_deposit", "aurora_hip_field_solve",
                     "aurora_hip_push", "aurora_hip_step", "aurora_hip_mcc_load",
                     "aurora_hip_swarm_submit", "aurora_hip_last_error"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(hip.LIB_PATH):
        pytest.skip("libaurora_hip.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in hip.declared_symbols() if s not in exported]
    assert not missing, missing


def test_library_loads_without_gpu():
    if not os.path.exists(hip.LIB_PATH):
        pytest.skip("libaurora_hip.so not built")
    lib = ctypes.CDLL(hip.LIB_PATH)
    for s in hip.declared_symbols():
        assert hasattr(lib, s)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(FileNotFoundError):
        hip._LIB = None
        hip.load(str(tmp_path / "absent.so"))
    hip._LIB = None
