"""CPU check of the GPU table-design code (csrc/hip/huff_design.hpp, compiled for the host)
against std::priority_queue / std::sort on tie-heavy histograms and adversarial sort inputs."""
import os
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.timeout(300)
def test_design_restatement_matches_std_library():
    src = os.path.join(HERE, "cpp", "design_check.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "design_check")
        subprocess.run(["g++", "-std=c++17", "-O2", "-o", exe, src], check=True, timeout=240)
        r = subprocess.run([exe, "300"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("OK"), r.stdout
