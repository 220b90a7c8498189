"""The drop-in shim printed in INTEGRATION.md (the C++ file a maintainer adds
to the reference's build) compiles against the reference's own headers and
include/muram_step.h: g++ -fsyntax-only on the extracted code block.  Host
only; skipped where the reference tree or MPICH's headers are absent."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
MPI_INC = "/opt/conda/include"


@pytest.mark.skipif(not (os.path.isdir(os.path.join(REF, "include")) and os.path.exists(os.path.join(MPI_INC, "mpi.h"))
                         and shutil.which("g++")), reason="reference headers, mpi.h or g++ not present")
def test_shim_compiles_against_reference_headers(tmp_path):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    assert blocks, "no C++ block in INTEGRATION.md"
    src = tmp_path / "mrm_dropin.C"
    src.write_text(blocks[0])
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-w", "-DMURAM_FFTW", f"-I{MPI_INC}", f"-I{REF}/include",
           f"-I{REF}/src/rt", f"-I{ROOT}/include", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[:4000]
    # every entry point the shim forwards to is declared in the header
    header = open(os.path.join(ROOT, "include", "muram_step.h")).read()
    for fn in set(re.findall(r"\b(mrm_[a-z_A-Z0-9]+)\(", blocks[0])):
        assert re.search(rf"\b{fn}\(", header), fn
