"""CPU check of the adjoint's cell functions: builds tests/adjoint_cells.cpp
(the VJPs of muram_main_amd/csrc/mrm_adjoint.h and the forward cell functions
they differentiate, compiled for the host) and runs its finite-difference
comparisons against the committed EOS table."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_adjoint_cell_vjps_match_finite_differences(tmp_path):
    exe = str(tmp_path / "adjoint_cells")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-fno-fast-math", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
           "-I", os.path.join(ROOT, "muram_main_amd", "csrc"), "-o", exe, "-x", "hip",
           os.path.join(HERE, "adjoint_cells.cpp")]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([exe, os.path.join(HERE, "golden", "eos_small.dat")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
