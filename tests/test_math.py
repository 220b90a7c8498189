"""CPU check of the logarithms the step's hot loops use
(muram_main_amd/csrc/mrm_math.h), compiled for the host with g++: the FP64
mlog against the C library (glibc) over 2e6 arguments per sweep -- near 1,
the density and specific-energy ranges of the EOS lookup, every normal
double, denormals -- to within one ulp; the FP32 mlogf against the exact
logarithm (glibc's FP64 log) to within one FP32 ulp everywhere and 0.6 ulp
over the density and specific-energy ranges (correct rounding is 0.5); the
special values (0, negative, NaN, inf) of both."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_log_within_one_ulp(tmp_path):
    exe = str(tmp_path / "math_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "math_check.cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    sweeps = [l.split() for l in out if l.startswith("log ")]
    assert len(sweeps) == 5
    for fn, name, n, worst, at in sweeps:
        print(fn, name, n, "max ulp", worst, "at", at)
        assert float(worst) <= 1.0, (fn, name, worst, at)
    sweeps = [l.split() for l in out if l.startswith("logf ")]
    assert len(sweeps) == 4
    for fn, name, n, worst, at in sweeps:
        print(fn, name, n, "max ulp", worst, "at", at)
        assert float(worst) <= (0.6 if name in ("density", "specific_energy") else 1.0), (fn, name, worst, at)
    special = [l for l in out if l.startswith("special")][0]
    assert special.split()[1] == "0"
