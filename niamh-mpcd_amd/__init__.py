"""niamh-mpcd_amd -- MI355X (gfx950) MPCD solvent library, Python side.

The product is the C-ABI shared library built from csrc/ (libmpcd_gpu.so,
entry points declared in include/mpcd_gpu.h).  This module builds it with
hipcc for gfx950 and binds it with ctypes so Python callers (tests, bench.py,
the calibration harness) drive exactly the entry points a C host would.

There is no CPU fallback: every compute call goes through the HIP library,
and loading fails loudly if it has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_PATH = os.path.join(HERE, "_build", "libmpcd_gpu.so")
INCLUDE = os.path.join(ROOT, "include")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # FP64 parity with the reference needs the reference's (unfused) operation order
    "-ffp-contract=off",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-Wno-pass-failed",
]

MAXSPEC = 10
MAXWALL = 16

# collision operators, LC modes, thermostats, precisions (include/mpcd_gpu.h)
ARBAXIS, ORTHAXIS, MPCAT, RAT = 0, 1, 2, 3
DIPOLE_DIR_SUM, DIPOLE_DIR_AV, DIPOLE_DIR_SIG, DIPOLE_DIR_SIG_SUM = 16, 17, 20, 21
ISOF, LCL, LCG = 0, 1, 2
NOTHERM, VSC, BEREND = 0, 1, 2
FP64, FP32, FP32_EXACT = 0, 1, 2

STATUS = {0: "MPCD_OK", 1: "MPCD_EINVAL", 2: "MPCD_ENOMEM", 3: "MPCD_EDEVICE", 4: "MPCD_EUNSUPPORTED",
          5: "MPCD_EESCAPE", 6: "MPCD_ECAPACITY"}


def sources():
    return [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]


def build(verbose: bool = False) -> str:
    """Compile csrc/*.hip into _build/libmpcd_gpu.so for gfx950 (always recompiles)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, *sources(), "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    build_host(verbose)
    return LIB_PATH


HOST_PATH = os.path.join(HERE, "_build", "mpcd_host")


def build_host(verbose: bool = False) -> str:
    """Compile the C host (host/*.c) against the built library with gcc."""
    cc = shutil.which("gcc") or "cc"
    host = os.path.join(HERE, "host")
    cmd = [cc, "-O2", "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-Wall", "-I", INCLUDE,
           os.path.join(host, "mpcd_host.c"), os.path.join(host, "json.c"), os.path.join(host, "checkpoint.c"),
           "-L", os.path.dirname(LIB_PATH), "-lmpcd_gpu", "-Wl,-rpath,$ORIGIN", "-lm", "-o", HOST_PATH + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(HOST_PATH + ".tmp", HOST_PATH)
    return HOST_PATH


class Species(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("mass", "rfc", "tumble", "chihi", "act", "mfpot", "damp",
                                          "sigwidth", "sigpos", "minactratio")] + [("vdist", C.c_int),
                                                                                    ("pad_", C.c_int)]


class Wall(C.Structure):
    _fields_ = [("A", C.c_double * 3), ("Q", C.c_double * 3), ("R", C.c_double), ("DN", C.c_double),
                ("DT", C.c_double), ("DVN", C.c_double), ("DVT", C.c_double), ("DVxyz", C.c_double * 3),
                ("MVN", C.c_double), ("MVT", C.c_double), ("MUN", C.c_double), ("MUT", C.c_double),
                ("MUxyz", C.c_double * 3), ("DUxyz", C.c_double * 3), ("V", C.c_double * 3),
                ("L", C.c_double * 3), ("INV", C.c_int), ("PHANTOM", C.c_int), ("COLL_TYPE", C.c_int),
                ("inter_mask", C.c_int), ("P", C.c_double * 4), ("B", C.c_double * 3), ("ABS", C.c_int),
                ("pad_", C.c_int)]


class Params(C.Structure):
    _fields_ = [("dim", C.c_int), ("L", C.c_int * 3), ("dt", C.c_double), ("kbt", C.c_double),
                ("rotang", C.c_double), ("tau", C.c_double), ("grav", C.c_double * 3), ("avv", C.c_double * 3),
                ("gndnst", C.c_double), ("rtech", C.c_int), ("lc", C.c_int), ("tstech", C.c_int),
                ("galinv", C.c_int), ("true_random_sign", C.c_int), ("precision", C.c_int),
                ("nspec", C.c_int), ("nwall", C.c_int), ("sp", Species * MAXSPEC), ("wall", Wall * MAXWALL)]


class DrawsC(C.Structure):
    _fields_ = [("shift", C.POINTER(C.c_double)), ("cell", C.POINTER(C.c_double)),
                ("ghost", C.POINTER(C.c_double)), ("part", C.POINTER(C.c_double)),
                ("lc", C.POINTER(C.c_double)), ("lc_stride", C.c_int),
                ("escape", C.POINTER(C.c_double)), ("retherm", C.POINTER(C.c_double))]


ESC_STRIDE, RETHERM_STRIDE = 32, 24  # MPCD_ESC_STRIDE, MPCD_RETHERM_STRIDE


EXPORTS = ["mpcd_gpu_create", "mpcd_gpu_destroy", "mpcd_gpu_last_error", "mpcd_gpu_upload", "mpcd_gpu_download",
           "mpcd_gpu_bin", "mpcd_gpu_collide", "mpcd_gpu_stream", "mpcd_gpu_step", "mpcd_gpu_run",
           "mpcd_gpu_synchronize", "mpcd_gpu_get_cells", "mpcd_gpu_get_order", "mpcd_gpu_ncells",
           "mpcd_gpu_lc_stride", "mpcd_gpu_profile", "mpcd_gpu_timing", "mpcd_gpu_autotune",
           "mpcd_gpu_tile_fill", "mpcd_gpu_checkpoint", "mpcd_gpu_restore", "mpcd_gpu_record",
           "mpcd_gpu_adjoint_reset", "mpcd_gpu_step_backward", "mpcd_gpu_set_layout",
           "mpcd_gpu_layout_stats", "mpcd_gpu_tangent", "mpcd_gpu_tangent_download", "mpcd_gpu_upload_aos",
           "mpcd_gpu_download_aos", "mpcd_gpu_flow", "mpcd_gpu_flow_read", "mpcd_gpu_lc_record",
           "mpcd_gpu_lc_decisions", "mpcd_gpu_bin_noalloc", "mpcd_gpu_kbtnow", "mpcd_gpu_get_step",
           "mpcd_gpu_set_step",
           "mpcd_gpu_zero_net_momentum", "mpcd_gpu_step_backward_state", "mpcd_gpu_observable"]

NTHETA = 4 + MAXSPEC  # MPCD_NTHETA: parameter adjoints per replica
OBS_BAND_VELOCITY, OBS_MEAN_S = 0, 1

_lib = None


def lib():
    """Load the built library; raises if it is missing (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    L.mpcd_gpu_create.argtypes = [C.POINTER(Params), C.c_int, i64p, C.c_uint64, C.POINTER(vp)]
    L.mpcd_gpu_destroy.argtypes = [vp]
    L.mpcd_gpu_last_error.restype = C.c_char_p
    L.mpcd_gpu_last_error.argtypes = [vp]
    L.mpcd_gpu_upload.argtypes = [vp, C.c_int, dp, dp, dp, ip]
    L.mpcd_gpu_download.argtypes = [vp, C.c_int, dp, dp, dp]
    L.mpcd_gpu_bin.argtypes = [vp, C.POINTER(DrawsC), C.c_int, ip, ip, ip]
    L.mpcd_gpu_collide.argtypes = [vp, C.POINTER(DrawsC), dp]
    L.mpcd_gpu_stream.argtypes = [vp, i64p]
    L.mpcd_gpu_step.argtypes = [vp, C.POINTER(DrawsC), dp]
    L.mpcd_gpu_run.argtypes = [vp, C.c_int, C.c_int]
    L.mpcd_gpu_synchronize.argtypes = [vp]
    L.mpcd_gpu_get_step.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.mpcd_gpu_set_step.argtypes = [vp, C.c_uint32]
    L.mpcd_gpu_get_cells.argtypes = [vp, C.c_int, ip, dp, dp, dp]
    L.mpcd_gpu_lc_record.argtypes = [vp, C.c_int]
    L.mpcd_gpu_kbtnow.argtypes = [vp, dp]
    L.mpcd_gpu_zero_net_momentum.argtypes = [vp]
    L.mpcd_gpu_bin_noalloc.argtypes = [C.POINTER(Params), C.c_int64, vp, dp, vp, vp, vp, vp, vp, vp]
    L.mpcd_gpu_lc_decisions.argtypes = [vp, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.mpcd_gpu_get_order.argtypes = [vp, C.c_int, ip, ip]
    L.mpcd_gpu_ncells.argtypes = [vp, C.c_int]
    L.mpcd_gpu_lc_stride.argtypes = [vp]
    L.mpcd_gpu_profile.argtypes = [vp, C.c_int]
    L.mpcd_gpu_timing.argtypes = [vp, dp, C.c_int]
    L.mpcd_gpu_autotune.argtypes = [vp, C.c_int, dp]
    L.mpcd_gpu_tile_fill.argtypes = [vp]
    L.mpcd_gpu_tile_fill.restype = C.c_double
    L.mpcd_gpu_set_layout.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    L.mpcd_gpu_layout_stats.argtypes = [vp, C.POINTER(C.c_int64)]
    L.mpcd_gpu_checkpoint.argtypes = [vp, C.c_int]
    L.mpcd_gpu_restore.argtypes = [vp, C.c_int]
    L.mpcd_gpu_record.argtypes = [vp, C.c_int]
    L.mpcd_gpu_adjoint_reset.argtypes = [vp]
    L.mpcd_gpu_step_backward.argtypes = [vp, dp, dp, dp]
    L.mpcd_gpu_step_backward_state.argtypes = [vp, dp, dp, dp]
    L.mpcd_gpu_observable.argtypes = [vp, C.c_int, dp, dp, dp]
    L.mpcd_gpu_tangent.argtypes = [vp, C.c_int, dp, dp]
    L.mpcd_gpu_upload_aos.argtypes = [vp, C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                      C.c_int64]
    L.mpcd_gpu_download_aos.argtypes = [vp, C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64]
    L.mpcd_gpu_tangent_download.argtypes = [vp, C.c_int, dp, dp, dp]
    L.mpcd_gpu_flow.argtypes = [vp, C.c_int]
    L.mpcd_gpu_flow_read.argtypes = [vp, C.c_int, dp, i64p, C.c_int]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


def make_params(cfg: dict, precision: int = FP64) -> Params:
    """Parameters from a dict using the reference's names (inputList / spec / bc)."""
    p = Params()
    p.dim = int(cfg["dim"])
    for i in range(3):
        p.L[i] = int(cfg["L"][i]) if i < len(cfg["L"]) else 1
        p.grav[i] = float(cfg.get("grav", (0, 0, 0))[i])
        p.avv[i] = float(cfg.get("avv", (0, 0, 0))[i])
    p.dt = float(cfg.get("dt", 0.1))
    p.kbt = float(cfg.get("kbt", 1.0))
    p.rotang = float(cfg.get("rotang", 1.570796))
    p.tau = float(cfg.get("tau", 0.5))
    p.gndnst = float(cfg.get("gndnst", 0.0))
    p.rtech = int(cfg.get("rtech", MPCAT))
    p.lc = int(cfg.get("lc", ISOF))
    p.tstech = int(cfg.get("tstech", NOTHERM))
    p.galinv = int(cfg.get("galinv", 1))
    p.true_random_sign = int(cfg.get("true_random_sign", 0))
    p.precision = int(precision)
    sps = cfg["species"]
    p.nspec = len(sps)
    for i, sp in enumerate(sps):
        for k in ("mass", "rfc", "tumble", "chihi", "act", "mfpot", "damp", "sigwidth", "sigpos", "minactratio"):
            setattr(p.sp[i], k, float(sp.get(k, 0.0)))
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
        W.inter_mask = int(w.get("inter_mask", (1 << MAXSPEC) - 1))
        for j, x in enumerate(w.get("P", (1, 1, 1, 1))):
            W.P[j] = float(x)
        for j, x in enumerate(w.get("B", (0, 0, 0))):
            W.B[j] = float(x)
        W.ABS = int(w.get("ABS", 0))
    return p


def bin_noalloc(cfg, n, q_dev, shift=None, prev_perm_dev=None, prev_cell_dev=None, cell_dev=None, count_dev=None,
                perm_dev=None, stream=None, precision=FP64):
    """mpcd_gpu_bin_noalloc: binning without a context or allocation; every
    *_dev argument is a device address (int, e.g. a tensor's data_ptr()),
    stream a hipStream_t address or None (null stream)"""
    L = lib()
    p = make_params(cfg, precision)
    sh = None if shift is None else (C.c_double * 3)(*[float(x) for x in shift])
    st = L.mpcd_gpu_bin_noalloc(C.byref(p), int(n), q_dev, sh, prev_perm_dev, prev_cell_dev, cell_dev, count_dev,
                                perm_dev, stream)
    if st != 0:
        L.mpcd_gpu_last_error.restype = C.c_char_p
        msg = L.mpcd_gpu_last_error(None)
        raise MPCDError(f"bin_noalloc failed: {STATUS.get(st, st)}: {msg.decode() if msg else ''}")


class MPCDError(RuntimeError):
    pass


class Context:
    """One GPU context holding one or more replicas (include/mpcd_gpu.h)."""

    def __init__(self, cfgs, n_particles, seed: int = 1, precision: int = FP64):
        if isinstance(cfgs, dict):
            cfgs = [cfgs]
            n_particles = [n_particles]
        self.L = lib()
        self.cfgs = list(cfgs)
        self.nrep = len(self.cfgs)
        self.params = (Params * self.nrep)(*[make_params(c, precision) for c in self.cfgs])
        self.n = np.asarray(n_particles, dtype=np.int64)
        h = C.c_void_p()
        st = self.L.mpcd_gpu_create(self.params, self.nrep, self.n.ctypes.data_as(C.POINTER(C.c_int64)),
                                    C.c_uint64(seed), C.byref(h))
        if st != 0:
            raise MPCDError(f"mpcd_gpu_create failed: {STATUS.get(st, st)}")
        self.h = h
        self.ncells = [self.L.mpcd_gpu_ncells(self.h, r) for r in range(self.nrep)]
        self.lc_stride = self.L.mpcd_gpu_lc_stride(self.h)

    def _check(self, st, what):
        if st not in (0,):
            msg = self.L.mpcd_gpu_last_error(self.h).decode()
            raise MPCDError(f"{what}: {STATUS.get(st, st)} {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.L.mpcd_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, r, Q, V, U=None, spid=None):
        Q = np.ascontiguousarray(Q, np.float64)
        V = np.ascontiguousarray(V, np.float64)
        U = None if U is None else np.ascontiguousarray(U, np.float64)
        spid = None if spid is None else np.ascontiguousarray(spid, np.int32)
        self._check(self.L.mpcd_gpu_upload(self.h, r, _dp(Q), _dp(V), _dp(U), _ip(spid)), "upload")

    def upload_aos(self, r, rec, q="Q", v="V", u=None, spid=None):
        """upload from a numpy structured array of particle records (e.g. the
        reference's particleMPC layout); q, v, u name double[3] fields, spid an int32 field"""
        rec = np.ascontiguousarray(rec)
        f = rec.dtype.fields
        off = lambda name: -1 if name is None else f[name][1]
        self._check(self.L.mpcd_gpu_upload_aos(self.h, r, rec.ctypes.data, rec.dtype.itemsize, off(q), off(v), off(u),
                                               off(spid)), "upload_aos")

    def download_aos(self, r, rec, q="Q", v="V", u=None):
        """write Q, V (and U) into a contiguous numpy structured array in place"""
        assert rec.flags["C_CONTIGUOUS"]
        f = rec.dtype.fields
        off = lambda name: -1 if name is None else f[name][1]
        self._check(self.L.mpcd_gpu_download_aos(self.h, r, rec.ctypes.data, rec.dtype.itemsize, off(q), off(v),
                                                 off(u)), "download_aos")

    def download(self, r):
        n = int(self.n[r])
        Q, V, U = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self.L.mpcd_gpu_download(self.h, r, _dp(Q), _dp(V), _dp(U)), "download")
        return Q, V, U

    @staticmethod
    def _draws(d):
        if d is None:
            return None, None
        keep = {k: (None if d.get(k) is None else np.ascontiguousarray(d[k], np.float64))
                for k in ("shift", "cell", "ghost", "part", "lc", "escape", "retherm")}
        s = DrawsC(_dp(keep["shift"]), _dp(keep["cell"]), _dp(keep["ghost"]), _dp(keep["part"]), _dp(keep["lc"]),
                   int(d.get("lc_stride", 0)), _dp(keep["escape"]), _dp(keep["retherm"]))
        return s, keep

    def bin(self, draws=None, shifted=1, outputs=False):
        s, keep = self._draws(draws)
        ntot, nct = int(self.n.sum()), int(sum(self.ncells))
        cell = np.zeros(ntot, np.int32) if outputs else None
        cnt = np.zeros(nct, np.int32) if outputs else None
        perm = np.zeros(ntot, np.int32) if outputs else None
        self._check(self.L.mpcd_gpu_bin(self.h, C.byref(s) if s else None, shifted, _ip(cell), _ip(cnt), _ip(perm)),
                    "bin")
        return (cell, cnt, perm) if outputs else None

    def collide(self, draws=None):
        s, keep = self._draws(draws)
        kbt = np.zeros(self.nrep)
        self._check(self.L.mpcd_gpu_collide(self.h, C.byref(s) if s else None, _dp(kbt)), "collide")
        return kbt

    def stream(self):
        ev = np.zeros(4, np.int64)
        st = self.L.mpcd_gpu_stream(self.h, ev.ctypes.data_as(C.POINTER(C.c_int64)))
        if st not in (0, 5):
            self._check(st, "stream")
        return ev

    def step(self, draws=None):
        s, keep = self._draws(draws)
        kbt = np.zeros(self.nrep)
        self._check(self.L.mpcd_gpu_step(self.h, C.byref(s) if s else None, _dp(kbt)), "step")
        return kbt

    def run(self, nsteps, use_graph=True):
        self._check(self.L.mpcd_gpu_run(self.h, int(nsteps), int(bool(use_graph))), "run")

    def get_step(self):
        """the draw counter (steps taken since the context began, or since set_step)"""
        out = C.c_uint32(0)
        self._check(self.L.mpcd_gpu_get_step(self.h, C.byref(out)), "get_step")
        return int(out.value)

    def set_step(self, step):
        """continue a saved run: the draws of the next step are those of `step`"""
        self._check(self.L.mpcd_gpu_set_step(self.h, int(step)), "set_step")

    def synchronize(self):
        self._check(self.L.mpcd_gpu_synchronize(self.h), "synchronize")

    def autotune(self, steps_per_trial=5):
        """pick the fill of run()'s layout by timing real steps; returns ms per step"""
        ms = C.c_double(0.0)
        self._check(self.L.mpcd_gpu_autotune(self.h, int(steps_per_trial), C.byref(ms)), "autotune")
        return ms.value

    def tile_fill(self):
        return self.L.mpcd_gpu_tile_fill(self.h)

    def set_layout(self, bucketed=True, fill=0.0, capacity=0):
        """layout of run(): per-tile buckets (default) or the cell-sorted arrays;
        fill > 0 sets the mean bucket population / bucket capacity, capacity > 0
        lowers the usable slots per bucket"""
        self._check(self.L.mpcd_gpu_set_layout(self.h, 1 if bucketed else 0, float(fill), int(capacity)),
                    "set_layout")

    def layout_stats(self):
        """(bucketed steps, bucket overflows, usable slots per bucket, buckets)"""
        out = (C.c_int64 * 4)()
        self._check(self.L.mpcd_gpu_layout_stats(self.h, out), "layout_stats")
        return tuple(int(x) for x in out)

    # --- checkpoints and the adjoint -------------------------------------------------------------------
    def checkpoint(self, slot):
        self._check(self.L.mpcd_gpu_checkpoint(self.h, int(slot)), "checkpoint")

    def restore(self, slot):
        self._check(self.L.mpcd_gpu_restore(self.h, int(slot)), "restore")

    def record(self, max_steps):
        self._check(self.L.mpcd_gpu_record(self.h, int(max_steps)), "record")

    def adjoint_reset(self):
        self._check(self.L.mpcd_gpu_adjoint_reset(self.h), "adjoint_reset")

    def step_backward_state(self, adj_in=None, want_lambda=False):
        """pops one taped step (every operator); adj_in [n, 9] = dJ/d(Q, V, U) at
        the end of that step; returns (the adjoint [n, 9] at its start or None,
        dJ/dtheta [nrep, NTHETA]: GRAV[0..2], 0, ACT of each species)"""
        n = int(self.n.sum())
        a = None if adj_in is None else np.ascontiguousarray(adj_in, np.float64).reshape(-1)
        if a is not None and a.size != 9 * n:
            raise ValueError("adj_in must hold 9 values per particle")
        lam = np.zeros(9 * n) if want_lambda else None
        th = np.zeros(NTHETA * self.nrep)
        self._check(self.L.mpcd_gpu_step_backward_state(self.h, _dp(a), _dp(lam), _dp(th)), "step_backward_state")
        return (None if lam is None else lam.reshape(n, 9)), th.reshape(self.nrep, NTHETA)

    def observable(self, kind, arg=None, grad=False):
        """(value [nrep], gradient [n, 9] or None) of OBS_BAND_VELOCITY (arg =
        component, axis, lo, hi) or OBS_MEAN_S for the current state"""
        n = int(self.n.sum())
        a = None if arg is None else np.ascontiguousarray(arg, np.float64)
        val = np.zeros(self.nrep)
        g = np.zeros(9 * n) if grad else None
        self._check(self.L.mpcd_gpu_observable(self.h, int(kind), _dp(a), _dp(val), _dp(g)), "observable")
        return val, (None if g is None else g.reshape(n, 9))

    def step_backward(self, adj_in=None, want_lambda=False):
        """pops one taped step; returns (lambda or None, dJ/dtheta [nrep, 4])"""
        n = int(self.n.sum())
        a = None if adj_in is None else np.ascontiguousarray(adj_in, np.float64).reshape(-1)
        if a is not None and a.size != 3 * n:
            raise ValueError("adj_in must hold 3 values per particle")
        lam = np.zeros(3 * n) if want_lambda else None
        th = np.zeros(4 * self.nrep)
        self._check(self.L.mpcd_gpu_step_backward(self.h, _dp(a), _dp(lam), _dp(th)), "step_backward")
        return (None if lam is None else lam.reshape(n, 3)), th.reshape(self.nrep, 4)

    def tangent(self, on=True, dgrav=None, dact=None):
        """forward-mode tangent along (d GRAV, d ACT per species)/d theta; on=False stops it"""
        g = None if dgrav is None else np.ascontiguousarray(dgrav, np.float64).reshape(3)
        a = None if dact is None else np.ascontiguousarray(dact, np.float64).reshape(-1)
        self._check(self.L.mpcd_gpu_tangent(self.h, 1 if on else 0, _dp(g), _dp(a)), "tangent")

    def tangent_download(self, r=0):
        """d(Q, V, U)/d theta of replica r in particle-index order"""
        n = int(self.n[r])
        dQ, dV, dU = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self.L.mpcd_gpu_tangent_download(self.h, r, _dp(dQ), _dp(dV), _dp(dU)), "tangent_download")
        return dQ, dV, dU

    def kbtnow(self):
        """KBTNOW of the last completed step, per replica"""
        k = np.zeros(self.nrep)
        self._check(self.L.mpcd_gpu_kbtnow(self.h, _dp(k)), "kbtnow")
        return k

    def zero_net_momentum(self):
        """the rest-frame transformation of an output step (galileantrans,
        mtools.c:673-718, + zeroExtraDims), on every replica"""
        self._check(self.L.mpcd_gpu_zero_net_momentum(self.h), "zero_net_momentum")

    def lc_record(self, on=True):
        """record the nematic path's discrete decisions of every later step"""
        self._check(self.L.mpcd_gpu_lc_record(self.h, int(bool(on))), "lc_record")

    def lc_decisions(self, r=0):
        """(particle words [n], cell words [ncell]) of the last step (uint32)"""
        part, cell = np.zeros(int(self.n[r]), np.uint32), np.zeros(self.ncells[r], np.uint32)
        u32 = C.POINTER(C.c_uint32)
        self._check(self.L.mpcd_gpu_lc_decisions(self.h, r, part.ctypes.data_as(u32), cell.ctypes.data_as(u32)),
                    "lc_decisions")
        return part, cell

    def cells(self, r=0):
        nc = self.ncells[r]
        pop, vcm, S, d = np.zeros(nc, np.int32), np.zeros((nc, 3)), np.zeros(nc), np.zeros((nc, 3))
        self._check(self.L.mpcd_gpu_get_cells(self.h, r, _ip(pop), _dp(vcm), _dp(S), _dp(d)), "get_cells")
        return dict(pop=pop, vcm=vcm, S=S, dir=d)

    def flow(self, on=True):
        """start (or stop) the running flow field: every later step adds each
        unshifted cell's centre-of-mass velocity (localFLOW + sumFLOW,
        mpc.c:226-246)"""
        self._check(self.L.mpcd_gpu_flow(self.h, int(bool(on))), "flow")

    def flow_read(self, r=0, reset=False):
        """(sums [ncell, 3], steps summed) of replica r; reset clears both
        afterwards, as flowout does (pout.c:1696)"""
        out, n = np.zeros((self.ncells[r], 3)), np.zeros(1, np.int64)
        self._check(self.L.mpcd_gpu_flow_read(self.h, r, _dp(out), n.ctypes.data_as(C.POINTER(C.c_int64)),
                                              int(bool(reset))), "flow_read")
        return out, int(n[0])

    FAMILIES = ["count", "scan", "scatter", "sort", "gather", "props", "ghost", "lc", "collide", "temp", "stream",
                "fused", "bucket", "flow"]

    def profile(self, enable=True):
        self._check(self.L.mpcd_gpu_profile(self.h, int(bool(enable))), "profile")

    def timing(self):
        """{family: (total_ms, launches)} of the kernels timed since profile(True)"""
        F = len(self.FAMILIES)
        out = np.zeros(2 * F)
        self._check(self.L.mpcd_gpu_timing(self.h, _dp(out), 2 * F), "timing")
        return {f: (out[i], int(out[F + i])) for i, f in enumerate(self.FAMILIES)}

    def order(self, r=0):
        o = np.zeros(int(self.n[r]), np.int32)
        cs = np.zeros(self.ncells[r] + 1, np.int32)
        self._check(self.L.mpcd_gpu_get_order(self.h, r, _ip(o), _ip(cs)), "get_order")
        return o, cs
