"""The C++ host layer (include/chmc/*.hh: config, potential_energy, set_integrator on a
hamiltonian_field, identical_particles, LEAP, HMC_class host loop vs device loop) driven from a compiled
program over the C ABI."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "chemhmc_amd", "bin", "host_api_check")


def test_host_api_program():
    assert os.path.exists(BIN), "build first: make -C chemhmc_amd all"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
