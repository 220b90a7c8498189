"""CPU checks of the C ABI: the gfx950 library loads and exports exactly the
entry points include/muram_step.h declares (no GPU calls)."""
import ctypes as C
import os
import re

import pytest

from muram_main_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "muram_step.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mrm_\w+)\s*\(", src, re.M)))


def test_header_and_python_mirror_agree():
    assert declared() == sorted(abi.EXPORTS)


@pytest.mark.skipif(not os.path.exists(abi.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(abi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_c_layout():
    assert C.sizeof(abi.GridDesc) == 9 * 4 + 4 + 6 * 8   # 9 ints, pad to 8, 6 doubles
    assert C.sizeof(abi.PhysDesc) == (9 + 5 + 4 + 4 + 10 + 2 + 2 + 2 + 4 + 4 + 4 + 10 + 3) * 8 + 8
    assert C.sizeof(abi.RunDesc) == 5 * 8 + 4 * 4
