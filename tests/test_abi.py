"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/mpcd_gpu.h declares, and the Python mirrors of its structs have the C
sizes.  No compute call is made (no GPU here)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import gpu_util as U

ROOT = U.ROOT
HEADER = os.path.join(ROOT, "include", "mpcd_gpu.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mpcd_gpu_\w+)\s*\(", txt)))


def test_header_and_python_exports_agree():
    M = U.package()
    assert sorted(M.EXPORTS) == declared()


def test_library_exports_every_declared_symbol():
    M = U.package()
    if not os.path.exists(M.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    L = C.CDLL(M.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), f"{name} missing from {M.LIB_PATH}"
    out = subprocess.run(["nm", "-D", "--defined-only", M.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared():
        assert re.search(rf"\sT {name}$", out, re.M), f"{name} is not an exported text symbol"


def test_struct_sizes_match_c(tmp_path):
    """compile a tiny C program against the header and compare sizeof()"""
    M = U.package()
    src = str(tmp_path / "sizes.c")
    exe = src[:-2]
    with open(src, "w") as f:
        f.write('#include <stdio.h>\n#include "mpcd_gpu.h"\nint main(void){printf("%zu %zu %zu %zu\\n",'
                'sizeof(mpcd_gpu_species),sizeof(mpcd_gpu_wall),sizeof(mpcd_gpu_params),sizeof(mpcd_draws));return 0;}\n')
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    sizes = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert sizes == [C.sizeof(M.Species), C.sizeof(M.Wall), C.sizeof(M.Params), C.sizeof(M.DrawsC)]


def test_bin_noalloc_rejects_sizes_above_its_maximum():
    """the size checks of the entry point that allocates nothing are host
    logic: they fail before any device call (none is made here)"""
    M = U.package()
    if not os.path.exists(M.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    sys.path.insert(0, os.path.dirname(__file__))
    from test_gpu_parity import _random_box
    import numpy as np
    c, _, _, _ = _random_box(3, (4, 4, 4), 2, 0)
    cfg = U.gpu_cfg(c, 2.0, np.zeros(3))
    with pytest.raises(M.MPCDError, match="MPCD_NOALLOC_MAX_PARTICLES"):
        M.bin_noalloc(cfg, 2097153, None, None, None, None, None, None, None)
    with pytest.raises(M.MPCDError, match="MPCD_NOALLOC_MAX_CELLS"):
        M.bin_noalloc(dict(cfg, L=[128, 128, 64]), 100, None, None, None, None, None, None, None)
