"""The C++20 host classes (aurora-pic_amd/host): the library and its test
program build and list their cases without a GPU; on the GPU every case of
host/tests/host_tests.cpp passes (the reference-style tests of the
FieldSolver / deposit / NullCollisionModel / Simulation* / SwarmService
surface)."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "aurora-pic_amd", "lib")
TESTS = os.path.join(LIB, "aurora_host_tests")


def test_host_program_lists_cases_without_gpu():
    assert os.path.exists(TESTS), "build with aurora-pic_amd/build.py first"
    out = subprocess.run([TESTS, "--list"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    names = out.stdout.split()
    assert "simulation3d_tabulated_field_energy" in names and len(names) >= 12


def test_host_library_exports_the_classes():
    out = subprocess.run(["nm", "-DC", os.path.join(LIB, "libaurora_gpu.so")],
                         capture_output=True, text=True, timeout=60).stdout
    for sym in ["pic::gpu::FieldSolver::solve", "pic::gpu::NullCollisionModel::collide",
                "pic::gpu::Simulation3D::step", "pic::gpu::Simulation::step",
                "pic::gpu::Simulation2D::step", "pic::gpu::SwarmService::submit",
                "pic::gpu::deposit_charge_cic", "pic::gpu::load_external_particle_state",
                "pic::gpu::write_external_particle_state",
                "pic::gpu::external_particle_state_signature"]:
        assert sym in out, sym


@pytest.mark.gpu
def test_host_cases_pass_on_gpu():
    out = subprocess.run([TESTS], capture_output=True, text=True, timeout=300)
    print(out.stdout[-4000:])
    assert out.returncode == 0, out.stdout[-4000:] + out.stderr[-2000:]
    assert " 0 failed" in out.stdout
