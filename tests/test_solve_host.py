"""The FP64 SE(3) update of the solver (kdh_solve.h), compiled for the host
and checked by tests/native/rodrigues_check.cpp: rodrigues_d against the
matrix exponential of [w]x as a long-double power series, across its
small-angle Taylor branch, the 0.25 rad switch and the closed form.  CPU only."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_rodrigues_against_matrix_exponential():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "rodrigues_check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I/opt/rocm/include",
                        "-D__HIP_PLATFORM_AMD__", os.path.join(HERE, "native", "rodrigues_check.cpp"), "-o", exe],
                       check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        print(r.stdout)
        assert r.returncode == 0, r.stdout + r.stderr
