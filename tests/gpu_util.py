"""Shared helpers for the GPU tests: load the package, build keyed draws with
the oracle, and run the GPU library through its C-ABI."""
import importlib.util
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def package():
    if "niamh_mpcd_amd" in sys.modules:
        return sys.modules["niamh_mpcd_amd"]
    spec = importlib.util.spec_from_file_location(
        "niamh_mpcd_amd", os.path.join(ROOT, "niamh-mpcd_amd", "__init__.py"),
        submodule_search_locations=[os.path.join(ROOT, "niamh-mpcd_amd")])
    m = importlib.util.module_from_spec(spec)
    sys.modules["niamh_mpcd_amd"] = m
    spec.loader.exec_module(m)
    return m


def load_golden(name):
    z = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    g = {k: z[k] for k in z.files}
    g["case_d"] = json.loads(str(g["case"]))
    return g


def gpu_cfg(c, gndnst, avv):
    cfg = dict(dim=c["dim"], L=c["L"], dt=c["dt"], kbt=c["kbt"], rotang=c["rotang"], tau=c["tau"], grav=c["grav"],
               avv=list(avv), gndnst=gndnst, rtech=c["rtech"], lc=c["lc"], tstech=c["tstech"], galinv=c["galinv"],
               species=c["species"], walls=c["walls"])
    return cfg


class KeyedRecorder:
    """Runs the oracle in sequential mode (== the reference's draws) and
    records every consumed variate into the keyed layout of the C-ABI."""

    def __init__(self, oracle_sys, seed, ncells, n):
        import pyoracle
        self.po = pyoracle
        self.s = oracle_sys
        self.d = pyoracle.Draws(pyoracle.DRAW_SEQ, seed)
        self.nc, self.n = ncells, n
        self.stride = oracle_sys.lc_stride()

    def step(self):
        shift = np.zeros(3)
        cell = np.zeros(self.nc * 4)
        ghost = np.zeros(self.nc * 3)
        part = np.zeros(self.n * 3)
        lc = np.zeros(self.n * self.stride)
        escape = np.zeros(2 * self.n * self.po.ESC_STRIDE)
        retherm = np.zeros(self.n * self.po.RETH_STRIDE)
        self.d.set_record(shift, cell, ghost, part, lc, self.stride)
        self.d.set_rare(escape, retherm, self.n)
        self.s.step(self.d)
        assert not self.d.rare_overflow(), "a step needed more escape tries or pushes than a keyed entry holds"
        return dict(shift=shift, cell=cell, ghost=ghost, part=part, lc=lc, lc_stride=self.stride, escape=escape,
                    retherm=retherm)


def hip_runtime():
    """the HIP runtime the library itself runs on (the libamdhip64 it mapped
    into this process), for tests that hand it device memory of their own:
    a second runtime, such as the one bundled with torch, would hold its own
    device state"""
    import ctypes as C
    M = package()
    M.lib()
    paths = []
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line and line.split()[-1] not in paths:
                paths.append(line.split()[-1])
    # with torch loaded first, the library binds to torch's copy (same soname)
    # and that is the only runtime mapped; otherwise the library's is the other
    own = [p for p in paths if "/torch/" not in p]
    path = own[0] if own else (paths[0] if len(paths) == 1 else None)
    if path is None:
        raise RuntimeError(f"cannot tell the library's HIP runtime among {paths}")
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DevBuf:
    """device memory from the library's HIP runtime, filled from and read
    into numpy arrays"""

    def __init__(self, hip, a):
        import ctypes as C
        self.hip, self.a = hip, np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(1, self.a.nbytes)) == 0, "hipMalloc"
        assert hip.hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, 1) == 0, "hipMemcpy H2D"

    @property
    def addr(self):
        return self.ptr.value

    def read(self):
        out = np.empty_like(self.a)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0, "hipMemcpy D2H"
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = None
