"""ctypes wrapper around oracle/_build/liboracle.so (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only as the checker.  It builds the oracle on first use if the
shared library is missing and `make` + `gcc` are available.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

MAXSPEC = 10
MAXWALL = 16

DRAW_SEQ, DRAW_KEYED, DRAW_PHILOX = 0, 1, 2
P_SHIFT, P_CELL, P_GHOST, P_NORMAL, P_LC = 1, 2, 3, 4, 5
ESC_STRIDE, RETH_STRIDE = 32, 24


class Species(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("mass", "rfc", "tumble", "chihi", "act", "mfpot", "damp",
                                          "sigwidth", "sigpos", "minactratio")] + [("vdist", C.c_int)]


class Wall(C.Structure):
    _fields_ = [("A", C.c_double * 3), ("Q", C.c_double * 3), ("R", C.c_double), ("DN", C.c_double),
                ("DT", C.c_double), ("DVN", C.c_double), ("DVT", C.c_double), ("DVxyz", C.c_double * 3),
                ("MVN", C.c_double), ("MVT", C.c_double), ("MUN", C.c_double), ("MUT", C.c_double),
                ("MUxyz", C.c_double * 3), ("DUxyz", C.c_double * 3), ("V", C.c_double * 3),
                ("L", C.c_double * 3), ("INV", C.c_int), ("PHANTOM", C.c_int), ("COLL_TYPE", C.c_int),
                ("inter", C.c_int * MAXSPEC), ("P", C.c_double * 4), ("B", C.c_double * 3), ("ABS", C.c_int)]


class Params(C.Structure):
    _fields_ = [("dim", C.c_int), ("L", C.c_int * 3), ("dt", C.c_double), ("kbt", C.c_double),
                ("rotang", C.c_double), ("tau", C.c_double), ("grav", C.c_double * 3),
                ("avv", C.c_double * 3), ("gndnst", C.c_double), ("rtech", C.c_int), ("lc", C.c_int),
                ("tstech", C.c_int), ("galinv", C.c_int), ("true_random_sign", C.c_int),
                ("nspec", C.c_int), ("nwall", C.c_int), ("sp", Species * MAXSPEC),
                ("wall", Wall * MAXWALL)]


class XCfg(C.Structure):
    """one replica of the exact-sum FP32 form (exact_step.c)"""
    _fields_ = [("dim", C.c_int), ("L", C.c_int * 3), ("rtech", C.c_int), ("true_sign", C.c_int),
                ("galinv", C.c_int), ("nspec", C.c_int), ("cell_off", C.c_int), ("dt", C.c_double),
                ("kbt", C.c_double), ("cosA", C.c_double), ("sinA", C.c_double), ("grav", C.c_double * 3),
                ("avv", C.c_double * 3), ("mass", C.c_double * 10), ("damp", C.c_double * 10)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def build_ref():
    subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(Params), C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_state.argtypes = [vp, dp, dp, dp, ip]
        L.orc_get_state.argtypes = [vp, dp, dp, dp]
        L.orc_ncells.argtypes = [vp]
        L.orc_get_order.argtypes = [vp, ip, ip]
        L.orc_get_cells.argtypes = [vp, ip, dp, dp, dp, dp, dp, dp]
        L.orc_get_walls.argtypes = [vp, dp]
        L.orc_kbtnow.restype = C.c_double
        L.orc_kbtnow.argtypes = [vp]
        L.orc_avs.restype = C.c_double
        L.orc_avs.argtypes = [vp]
        L.orc_counters.argtypes = [vp, C.POINTER(C.c_long)]
        L.orc_lc_stride.argtypes = [vp]
        L.orc_zero_net_momentum.argtypes = [vp]
        for f in ("orc_step", "orc_phase_bin", "orc_phase_collide", "orc_phase_stream"):
            getattr(L, f).argtypes = [vp, vp]
        L.orc_last_shift.argtypes = [vp, dp]
        L.orc_flow.argtypes = [vp, C.c_int]
        L.orc_get_flow.argtypes = [vp, dp, C.POINTER(C.c_long), C.c_int]
        L.orc_draws_new.restype = vp
        L.orc_draws_new.argtypes = [C.c_int, C.c_uint64, C.c_uint32]
        L.orc_draws_free.argtypes = [vp]
        L.orc_draws_set_keyed.argtypes = [vp, dp, dp, dp, dp, dp, C.c_int]
        L.orc_draws_set_record.argtypes = [vp, dp, dp, dp, dp, dp, C.c_int]
        L.orc_draws_counts.argtypes = [vp, C.POINTER(C.c_uint64)]
        L.orc_draws_set_rare.argtypes = [vp, dp, dp, C.c_int]
        L.orc_lc_decisions.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_draws_rare_overflow.argtypes = [vp]
        L.orc_eigen.argtypes = [dp, C.c_int, dp]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_exact_step.restype = C.c_int
        L.orc_exact_step.argtypes = [C.POINTER(XCfg), C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), ip, ip,
                                     C.c_uint64, C.c_uint32, C.c_uint32, dp]
        L.orc_philox_uniform.restype = C.c_double
        L.orc_philox_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32]
        L.orc_philox_normal.restype = C.c_double
        L.orc_philox_normal.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32]
        for f in ("orc_philox_unitvec", "orc_philox_unitvec_turn"):
            getattr(L, f).argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, dp]
        L.orc_sincos_turn.argtypes = [C.c_double, dp, dp]
        L.orc_log_unit.restype = C.c_double
        L.orc_log_unit.argtypes = [C.c_double]
        L.orc_philox_normal_turn.restype = C.c_double
        L.orc_philox_normal_turn.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32]
        L.orc_rodrigues.argtypes = [dp, dp, C.c_double]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


def make_params(cfg: dict) -> Params:
    """cfg uses the reference's names: dim, L, dt, kbt, rotang, tau, grav, avv,
    gndnst, rtech, lc, tstech, galinv, species (list of dicts), walls (list of dicts)."""
    p = Params()
    p.dim = cfg["dim"]
    for i in range(3):
        p.L[i] = int(cfg["L"][i])
        p.grav[i] = float(cfg.get("grav", (0, 0, 0))[i])
        p.avv[i] = float(cfg.get("avv", (0, 0, 0))[i])
    p.dt = cfg.get("dt", 0.1)
    p.kbt = cfg.get("kbt", 1.0)
    p.rotang = cfg.get("rotang", 1.570796)
    p.tau = cfg.get("tau", 0.5)
    p.gndnst = cfg.get("gndnst", 0.0)
    p.rtech = cfg.get("rtech", 2)
    p.lc = cfg.get("lc", 0)
    p.tstech = cfg.get("tstech", 0)
    p.galinv = cfg.get("galinv", 1)
    p.true_random_sign = cfg.get("true_random_sign", 0)
    sps = cfg["species"]
    p.nspec = len(sps)
    for i, sp in enumerate(sps):
        for k in ("mass", "rfc", "tumble", "chihi", "act", "mfpot", "damp", "sigwidth", "sigpos",
                  "minactratio"):
            setattr(p.sp[i], k, float(sp[k]))
        p.sp[i].vdist = int(sp.get("vdist", 4))
    ws = cfg["walls"]
    p.nwall = len(ws)
    for i, w in enumerate(ws):
        W = p.wall[i]
        for k in ("A", "Q", "DVxyz", "MUxyz", "DUxyz", "V", "L"):
            v = w.get(k, (1, 1, 1) if k == "MUxyz" else (0, 0, 0))
            for j in range(3):
                getattr(W, k)[j] = float(v[j])
        for k, dv in (("R", 0), ("DN", 0), ("DT", 0), ("DVN", 0), ("DVT", 0), ("MVN", 1), ("MVT", 1),
                      ("MUN", 1), ("MUT", 1)):
            setattr(W, k, float(w.get(k, dv)))
        W.INV = int(w.get("INV", 0))
        W.PHANTOM = int(w.get("PHANTOM", 0))
        W.COLL_TYPE = int(w.get("COLL_TYPE", 1))
        for j, x in enumerate(w.get("P", (1, 1, 1, 1))):
            W.P[j] = float(x)
        for j, x in enumerate(w.get("B", (0, 0, 0))):
            W.B[j] = float(x)
        W.ABS = int(w.get("ABS", 0))
        for j in range(MAXSPEC):
            W.inter[j] = 1
    return p


class Draws:
    def __init__(self, mode: int, seed: int, replica: int = 0):
        self.L = lib()
        self.h = self.L.orc_draws_new(mode, seed, replica)
        self._keep = []

    def set_keyed(self, shift, cell, ghost, part, lc, lc_stride):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                for a in (shift, cell, ghost, part, lc)]
        self._keep = arrs
        self.L.orc_draws_set_keyed(self.h, *[_dp(a) for a in arrs], int(lc_stride))

    def set_record(self, shift, cell, ghost, part, lc, lc_stride):
        self._keep = [shift, cell, ghost, part, lc]
        self.L.orc_draws_set_record(self.h, *[_dp(a) for a in (shift, cell, ghost, part, lc)], int(lc_stride))

    def set_rare(self, escape, retherm, n):
        """escape [2 * n * ESC_STRIDE] / retherm [n * RETH_STRIDE] arrays: read in
        keyed mode (None: Philox), written in sequential mode"""
        self._rare = [escape, retherm]
        self.L.orc_draws_set_rare(self.h, _dp(escape), _dp(retherm), int(n))

    def rare_overflow(self):
        return bool(self.L.orc_draws_rare_overflow(self.h))

    def counts(self):
        out = (C.c_uint64 * 3)()
        self.L.orc_draws_counts(self.h, out)
        return tuple(out)

    def __del__(self):
        try:
            self.L.orc_draws_free(self.h)
        except Exception:
            pass


class System:
    def __init__(self, cfg: dict, Q, V, U=None, spid=None):
        self.L = lib()
        self.cfg = cfg
        self.params = make_params(cfg)
        self.N = int(Q.shape[0])
        self.h = self.L.orc_create(C.byref(self.params), self.N)
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        U = np.ascontiguousarray(U if U is not None else np.zeros_like(Q), dtype=np.float64)
        spid = np.ascontiguousarray(spid if spid is not None else np.zeros(self.N), dtype=np.int32)
        rc = self.L.orc_set_state(self.h, _dp(Q), _dp(V), _dp(U), _ip(spid))
        if rc:
            raise ValueError("initial positions outside the box")
        self.nc = self.L.orc_ncells(self.h)

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def step(self, draws: Draws):
        rc = self.L.orc_step(self.h, draws.h)
        if rc:
            raise RuntimeError(f"oracle step failed ({rc})")

    def phase(self, name: str, draws: Draws):
        rc = getattr(self.L, "orc_phase_" + name)(self.h, draws.h)
        if rc:
            raise RuntimeError(f"oracle phase {name} failed ({rc})")

    def state(self):
        Q = np.zeros((self.N, 3))
        V = np.zeros((self.N, 3))
        U = np.zeros((self.N, 3))
        self.L.orc_get_state(self.h, _dp(Q), _dp(V), _dp(U))
        return Q, V, U

    def order(self):
        o = np.zeros(self.N, np.int32)
        cs = np.zeros(self.nc + 1, np.int32)
        self.L.orc_get_order(self.h, _ip(o), _ip(cs))
        return o, cs

    def cells(self):
        nc = self.nc
        out = dict(pop=np.zeros(nc, np.int32), mass=np.zeros(nc), vcm=np.zeros((nc, 3)), cm=np.zeros((nc, 3)),
                   S=np.zeros(nc), dir=np.zeros((nc, 3)), E=np.zeros((nc, 9)))
        self.L.orc_get_cells(self.h, _ip(out["pop"]), _dp(out["mass"]), _dp(out["vcm"]), _dp(out["cm"]),
                             _dp(out["S"]), _dp(out["dir"]), _dp(out["E"]))
        return out

    def walls(self):
        w = np.zeros((max(1, self.params.nwall), 3))
        self.L.orc_get_walls(self.h, _dp(w))
        return w[: self.params.nwall]

    def kbtnow(self):
        return self.L.orc_kbtnow(self.h)

    def avs(self):
        return self.L.orc_avs(self.h)

    def zero_net_momentum(self):
        """galileantrans + zeroExtraDims (mtools.c:673-744)"""
        self.L.orc_zero_net_momentum(self.h)

    def counters(self):
        out = (C.c_long * 4)()
        self.L.orc_counters(self.h, out)
        return tuple(out)

    def lc_decisions(self):
        """(particle words [N], cell words [ncell]) of the last collision phase"""
        part, cell = np.zeros(self.N, np.uint32), np.zeros(self.nc, np.uint32)
        u32 = C.POINTER(C.c_uint32)
        self.L.orc_lc_decisions(self.h, part.ctypes.data_as(u32), cell.ctypes.data_as(u32))
        return part, cell

    def lc_stride(self):
        return self.L.orc_lc_stride(self.h)

    def last_shift(self):
        s = np.zeros(3)
        self.L.orc_last_shift(self.h, _dp(s))
        return s

    def flow(self, on=True):
        """with on, every later step ends with localFLOW + sumFLOW (mpc.c:226-246)"""
        self.L.orc_flow(self.h, int(bool(on)))

    def get_flow(self, reset=False):
        """(FLOW [ncell, 3], steps summed); reset clears both afterwards"""
        f, n = np.zeros((self.L.orc_ncells(self.h), 3)), C.c_long(0)
        self.L.orc_get_flow(self.h, _dp(f), C.byref(n), int(bool(reset)))
        return f, int(n.value)


def eigen(m: np.ndarray):
    m = np.ascontiguousarray(m, dtype=np.float64).copy()
    dim = m.shape[0]
    ev = np.zeros(3)
    lib().orc_eigen(_dp(m), dim, _dp(ev))
    return ev[:dim], m


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return tuple(o)


def rodrigues(vec, axis, theta):
    v = np.ascontiguousarray(vec, dtype=np.float64).copy()
    a = np.ascontiguousarray(axis, dtype=np.float64).copy()
    lib().orc_rodrigues(_dp(v), _dp(a), float(theta))
    return v


def exact_step(cfg: dict, x, v, sp=None, seed=1, replica=0, step=0, cell_off=0, ids=None):
    """One step of the exact-sum FP32 form on a box of periodic faces
    (exact_step.c).  cfg: the keys of the library's mpcd_gpu_params dicts
    (dim, L, dt, kbt, rotang, grav, avv, rtech, galinv, true_random_sign,
    species).  x, v: (n, 3) float32 positions and velocities; ids: the keys of
    the particles' draws (default: their index).  Returns the new x, v
    (float32) and KBTNOW."""
    import math
    p = XCfg()
    p.dim = int(cfg["dim"])
    for i in range(3):
        p.L[i] = int(cfg["L"][i]) if i < p.dim else 1
        p.grav[i] = float(cfg.get("grav", (0, 0, 0))[i])
        p.avv[i] = float(cfg.get("avv", (0, 0, 0))[i])
    p.rtech = int(cfg.get("rtech", 0))
    p.true_sign = int(cfg.get("true_random_sign", 0))
    p.galinv = int(cfg.get("galinv", 1))
    sps = cfg["species"]
    p.nspec = len(sps)
    p.cell_off = int(cell_off)
    p.dt = float(cfg["dt"])
    p.kbt = float(cfg["kbt"])
    p.cosA = math.cos(float(cfg.get("rotang", 0.0)))
    p.sinA = math.sin(float(cfg.get("rotang", 0.0)))
    for k, s_ in enumerate(sps):
        p.mass[k] = float(s_["mass"])
        p.damp[k] = float(s_.get("damp", 0.0))
    x = np.ascontiguousarray(x, dtype=np.float32).copy()
    v = np.ascontiguousarray(v, dtype=np.float32).copy()
    spa = None if sp is None else np.ascontiguousarray(sp, dtype=np.int32)
    ida = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
    kbt = C.c_double()
    fp = C.POINTER(C.c_float)
    bad = lib().orc_exact_step(C.byref(p), int(x.shape[0]), x.ctypes.data_as(fp), v.ctypes.data_as(fp),
                               None if spa is None else spa.ctypes.data_as(C.POINTER(C.c_int)),
                               None if ida is None else ida.ctypes.data_as(C.POINTER(C.c_int)), int(seed),
                               int(replica), int(step), C.byref(kbt))
    if bad:
        raise RuntimeError(f"exact_step: {bad} particles outside the box")
    return x, v, kbt.value
