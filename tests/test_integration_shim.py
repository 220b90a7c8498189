"""The drop-in shim (muram_main_amd/shim/mrm_dropin.C, the C++ file a
maintainer adds to the reference's build) compiles against the reference's
own headers and include/muram_step.h: g++ -fsyntax-only.  Host only; skipped
where the reference tree or MPICH's headers are absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "muram_main_amd", "shim", "mrm_dropin.C")
REF = "/root/reference"
MPI_INC = "/opt/conda/include"


@pytest.mark.skipif(not (os.path.isdir(os.path.join(REF, "include")) and os.path.exists(os.path.join(MPI_INC, "mpi.h"))
                         and shutil.which("g++")), reason="reference headers, mpi.h or g++ not present")
def test_shim_compiles_against_reference_headers():
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-w", "-DMURAM_FFTW", f"-I{MPI_INC}", f"-I{REF}/include",
           f"-I{REF}/src/rt", f"-I{ROOT}/include", SHIM]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[:4000]


def test_shim_calls_declared_entry_points():
    # every entry point the shim forwards to is declared in the header
    code = open(SHIM).read()
    header = open(os.path.join(ROOT, "include", "muram_step.h")).read()
    called = set(re.findall(r"\b(mrm_[a-z_A-Z0-9]+)\(", code))
    assert called
    for fn in called:
        assert re.search(rf"\b{fn}\(", header), fn
