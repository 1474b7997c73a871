"""Benchmark: MPCD particle-steps/s on one MI355X, FP32-state batched and
plain FP64, with the HBM roofline of the dominant kernel and the reference's
CPU step timed on this host.

Workload (both forms): 3D SRD, periodic 64^3 cells x 10 particles per cell
(2 621 440 particles per replica), ARBAXIS rotation 130 degrees, dt 0.1,
kT 1, random grid shift every step, no thermostat.
  * headline `value`: the exact-sum FP32 form (FP32 particle state and
    arithmetic, cell sums of quantised terms added exactly in FP64, so no
    list order), B independent replicas (default 8) advancing together in
    one context;
  * `fp32_list_order`: the same replicas with the FP32 state stepped in the
    reference's list order (FP64 cell sums in list order);
  * `fp64_plain`: FP64 state, one replica.
A step is one full MPCD step (shift, bin, cell sums, collision, shift back,
stream, walls) over every particle.  Inputs are resident in HBM before the
timed region; the timed region is bracketed by device synchronisation.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--replicas B]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "MPCD particle-steps/s on 1 MI355X (plain FP64 vs FP32 batched) + % HBM peak"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle-step of the whole step: one read and one
# write of the persistent particle state, r and v (3D: 6 words), r and v in
# 2D (4 words), r, v and u for 2D nematics (6 words)
B_ALG = {"fp64": 96.0, "fp32": 48.0, "fp32x": 48.0}
B_ALG_2D = {"fp64": 64.0, "fp32": 32.0}
B_ALG_LC2D = {"fp64": 96.0, "fp32": 48.0}
PREC = {"fp64": 8, "fp32": 4, "fp32x": 4}


def fam_bytes(fam, form, dim=3, lc=False):
    """algorithmic bytes per particle that one launch of a kernel family must
    move (each input read once, each output written once); p = bytes per
    real, c = components (dim), u = orientation words (nematics)"""
    p, c = PREC[form], dim
    u = c if lc else 0
    return {"count": c * p + 24,               # read r, slot cell, unshifted cell; write cell, rank, 64-bit key
            "scatter": 24,                     # read cell, rank, key; write key
            "sort": 20,                        # read + write keys; write slot cells
            "gather": 2 * (2 * c + u) * p + 24,  # read key, state, id, species; write state, id, species
            "props": (2 * c + u) * p + 4,      # read r, v (u), species
            "collide": (2 * c + u + 2 * c) * p + 4,  # read r, v (u), species; write r, v
            "lc": 2 * u * p + 4,               # read and write u, read species
            "stream": 2 * (2 * c + u) * p + 8,  # read state, species; write state, unshifted cell
            # fused: read key, r, v, id; write r, v, id, slot cell, next cell, rank, key
            "fused": 4 * c * p + 44,
            # bucketed step: read r, v, id (and the list-order key, 8 B, but in the exact-sum form) from the
            # bucket; write them to the next bucket
            "bucket": 4 * c * p + (8 if form == "fp32x" else 24),
            "ghost": 0, "temp": 0, "scan": 0, "flow": 0}[fam]


ROTANG_130 = 2.2689280275926285


def package():
    spec = importlib.util.spec_from_file_location(
        "niamh_mpcd_amd", os.path.join(ROOT, "niamh-mpcd_amd", "__init__.py"),
        submodule_search_locations=[os.path.join(ROOT, "niamh-mpcd_amd")])
    m = importlib.util.module_from_spec(spec)
    sys.modules["niamh_mpcd_amd"] = m
    spec.loader.exec_module(m)
    return m


def periodic_walls(L):
    ws = []
    for ax in range(len(L)):
        a1 = [0.0, 0.0, 0.0]
        a1[ax] = 1.0
        a2 = [0.0, 0.0, 0.0]
        a2[ax] = -1.0
        ws += [dict(A=a1, R=0.0, DN=float(L[ax])), dict(A=a2, R=-float(L[ax]), DN=float(L[ax]))]
    return ws


def srd_cfg(L, ppc, dim=3, rtech=0, rotang=ROTANG_130):
    sp = dict(mass=1.0, rfc=0.01, tumble=2.0, chihi=0.5, act=0.05, mfpot=10.0, damp=0.0, sigwidth=1.0, sigpos=1.0,
              minactratio=0.0, vdist=4)
    return dict(dim=dim, L=list(L) + [1] * (3 - len(L)), dt=0.1, kbt=1.0, rotang=rotang, tau=0.5, grav=[0, 0, 0],
                avv=[0, 0, 0], gndnst=float(ppc), rtech=rtech, lc=0, tstech=0, galinv=1, species=[sp],
                walls=periodic_walls(L))


def initial_state(L, ppc, seed, dim=3):
    rng = np.random.default_rng(seed)
    N = int(np.prod(L[:dim])) * ppc
    Q = np.zeros((N, 3))
    for d in range(dim):
        Q[:, d] = rng.uniform(0.0, L[d], N)
    V = np.zeros((N, 3))
    V[:, :dim] = rng.normal(size=(N, dim))
    V -= V.mean(axis=0)
    return Q, V


def small_box(M, steps=1000):
    """the reference's CPU-runnable configuration on the GPU: 2D SRD 30x30
    cells x 20 ppc (18 000 particles), ORTHAXIS 90 degrees, dt 0.1, periodic,
    FP64, 1000 steps -- eager launches against hipGraph replays, where the
    launch overhead of a step dominates -- beside the reference's own step
    on one host core"""
    L, ppc = (30, 30), 20
    cfg = srd_cfg(L, ppc, dim=2, rtech=1, rotang=math.pi / 2)
    Q, V = initial_state(L, ppc, seed=3, dim=2)
    N = Q.shape[0]
    out = {}
    for name, graph in (("eager", False), ("hipgraph", True)):
        ctx = M.Context(cfg, N, seed=17, precision=M.FP64)
        ctx.upload(0, Q, V)
        ctx.run(10, use_graph=graph)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.run(steps, use_graph=graph)
        ctx.synchronize()
        sec = time.perf_counter() - t0
        ctx.close()
        out[name] = dict(value=N * steps / sec, ms_per_step=1e3 * sec / steps)
    out["workload"] = (f"2D SRD {L[0]}x{L[1]} cells x {ppc} ppc ({N} particles), ORTHAXIS 90 deg, dt 0.1, kT 1, "
                       f"periodic, FP64, {steps} steps")
    out["unit"] = "particle-steps/s"
    return out


def graph_run(ctx, steps, n_total):
    """the same steps replayed as hipGraphs (groups of two or six steps), no profiling"""
    ctx.run(12, use_graph=True)  # records the graph
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(steps, use_graph=True)
    ctx.synchronize()
    sec = time.perf_counter() - t0
    return dict(value=n_total * steps / sec, ms_per_step=1e3 * sec / steps)


def timed_run(M, ctx, steps, warmup, n_total, prof_steps=20):
    """warmup untimed steps, then exactly `steps` timed eager steps with
    nothing else on the stream; then a separate run of prof_steps steps with
    HIP events around every kernel family for the kernel times (the events
    add a few microseconds between kernels, so they stay out of the timed
    run)"""
    ctx.run(warmup, use_graph=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(steps, use_graph=False)
    ctx.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    ctx.profile(True)
    ctx.run(prof_steps, use_graph=False)
    ctx.synchronize()
    timing = ctx.timing()
    ctx.profile(False)
    return dict(sec=sec, ms_per_step=1e3 * sec / steps, value=n_total * steps / sec, timing=timing,
                prof_steps=prof_steps)


def roofline(timing, n_per_launch, form, b_alg=None, families=None, traffic_tag=None, dim=3, lc=False):
    """the dominant kernel against HBM.  `achieved` counts the algorithmic
    bytes of a particle-step (one read and one write of the particle state,
    b_alg per particle; 48 / 96 B in 3D) for every particle of one launch,
    over the kernel's mean launch time (HIP events on the library stream).
    What the kernel itself must move -- its records, ids and binning keys --
    is kernel_bytes, with kernel_frac beside it; `traffic` is the HBM bytes
    per launch that the committed rocprofv3 counter summary measured."""
    items = [kv for kv in timing.items() if families is None or kv[0] in families]
    fam, (ms, cnt) = max(items, key=lambda kv: kv[1][0])
    avg_s = (ms / max(cnt, 1)) * 1e-3
    b = B_ALG[form] if b_alg is None else b_alg
    bytes_alg = b * n_per_launch
    kbytes = fam_bytes(fam, form, dim, lc) * n_per_launch
    achieved = bytes_alg / avg_s / 1e9
    return dict(bound="hbm", kernel_family=fam, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=achieved / HBM_PEAK_GBS, traffic=measured_traffic(fam, traffic_tag or form),
                avg_launch_ms=ms / max(cnt, 1), alg_bytes_per_particle=b, alg_bytes_per_launch=bytes_alg,
                kernel_bytes_per_particle=fam_bytes(fam, form, dim, lc), kernel_bytes=kbytes,
                kernel_frac=kbytes / avg_s / 1e9 / HBM_PEAK_GBS)


def measured_traffic(fam, tag):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any"""
    path = os.path.join(ROOT, "profiles", f"traffic_{fam}_{tag}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(L=(64, 64, 64), ppc=10, steps=4, dim=3, rtech=0, rotang=ROTANG_130):
    """the reference's own C step (oracle/_ref/ref_harness, built from the
    reference's sources) timed on one host core on a bounded sample of the
    same workload; falls back to the oracle port when the build is absent"""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    Q, V = initial_state(L, ppc, seed=99, dim=dim)
    N = Q.shape[0]
    if os.path.exists(harness):
        with tempfile.TemporaryDirectory() as d:
            case = os.path.join(d, "case.txt")
            state = os.path.join(d, "state.bin")
            Lc = list(L) + [1] * (3 - len(L))
            lines = [f"dim {dim}", "L {} {} {}".format(*Lc), f"N {N}", f"steps {steps}", f"rtech {rtech}", "lc 0",
                     "tstech 0", "galinv 1", "dt 0.1", "kbt 1.0", f"rotang {rotang!r}", "tau 0.5", "grav 0 0 0",
                     "avv 0 0 0", "seed 5", "species 1 0.01 2 0.5 0.05 10 0 1 1 0 4 0.007"]
            for w in periodic_walls(L):
                v = list(w["A"]) + [0, 0, 0] + [w["R"], w["DN"], 0, 0, 0] + [0, 0, 0] + [1, 1, 1, 1] + \
                    [1, 1, 1] + [0, 0, 0] + [0, 0, 0] + [0, 0, 1]
                lines.append("wall " + " ".join(repr(float(x)) for x in v))
            open(case, "w").write("\n".join(lines) + "\n")
            with open(state, "wb") as f:
                f.write(np.hstack([Q, V, np.zeros_like(Q)]).astype("<f8").tobytes())
                f.write(np.zeros(N, "<i4").tobytes())
            out = subprocess.run([harness, case, state, os.path.join(d, "o.bin"), "timing"], capture_output=True,
                                 text=True, check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
            r = json.loads(out.stdout.strip().splitlines()[-1])
            sec = r["seconds"]
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
        import cases as CS
        import pyoracle
        cfg = srd_cfg(L, ppc, dim=dim, rtech=rtech, rotang=rotang)
        cfg["walls"] = CS.periodic(dim, L)
        s = pyoracle.System(cfg, Q, V, np.zeros_like(Q), None)
        d = pyoracle.Draws(pyoracle.DRAW_PHILOX, 5)
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step(d)
        sec = time.perf_counter() - t0
        kind = "port"
    return dict(value=N * steps / sec, unit="particle-steps/s", cores=1, kind=kind,
                sample=f"{dim}D SRD {'x'.join(str(x) for x in L)} cells x {ppc} ppc ({N} particles), {steps} steps, "
                       f"one core")


def nematic_gradients(M, Lx=128, Ly=64, ppc=20, K=200, ckpt=20, g=5e-4, act=0.03, hg=(1e-11,), ha=(1e-12,)):
    """gradients of two time-averaged observables of a 2D active nematic
    channel (lc 1, dipole activity DIPOLE_DIR_SUM, U = 10, tumbling 0.5,
    rfc 0.01; bounce-back phantom walls in y with homeotropic strong
    anchoring; body force g along x) with respect to g and the activity:
    the x-velocity of the centreline band and the mean scalar order
    parameter of the cells.  The reverse mode of the whole step
    (mpcd_gpu_step_backward_state, checkpoint every `ckpt` steps) gives both
    derivatives of one observable in one backward pass; the forward-mode
    tangent (one run per direction) gives the same pathwise derivative
    independently; central finite differences with common random numbers
    at the steps hg, ha are reported beside them.
    The pathwise derivative holds every discrete decision fixed.  A difference
    step perturbs the orientations by h times their tangents, which grow to
    O(100) over 200 steps, and Metropolis acceptances (about 2e8 comparisons at
    128 x 64 x 20) flip once h exceeds about 1e-12; below about 1e-13 the
    FP64 round-off of the observables dominates.  At that size finite
    differences therefore resolve the derivative to about 1 %, and the exact
    check is reverse against tangent; a smaller box has a clean window and
    checks it to 1e-3."""
    import importlib
    H = importlib.import_module("niamh_mpcd_amd.adjoint")
    rng = np.random.default_rng(4)
    N = Lx * Ly * ppc
    Q = np.zeros((N, 3))
    Q[:, 0] = rng.uniform(0, Lx, N)
    Q[:, 1] = rng.uniform(0, Ly, N)
    V = np.zeros((N, 3))
    V[:, :2] = rng.normal(size=(N, 2))
    V -= V.mean(axis=0)
    th = rng.uniform(0, 2 * np.pi, N)
    U = np.zeros((N, 3))
    U[:, 0], U[:, 1] = np.cos(th), np.sin(th)

    def cfg(gg, aa):
        sp = dict(mass=1.0, rfc=0.01, tumble=0.5, chihi=0.5, act=aa, mfpot=10.0, damp=0.0, sigwidth=1.0,
                  sigpos=1.0, minactratio=0.0, vdist=4)
        anch = dict(DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1, MUN=1.0, MUT=0.0)
        ws = [dict(A=[1.0, 0, 0], R=0.0, DN=float(Lx)), dict(A=[-1.0, 0, 0], R=-float(Lx), DN=float(Lx)),
              dict(A=[0, 1.0, 0], R=0.0, **anch), dict(A=[0, -1.0, 0], R=-float(Ly), **anch)]
        return dict(dim=2, L=[Lx, Ly, 1], dt=0.1, kbt=1.0, rotang=1.570796, tau=0.5, grav=[gg, 0.0, 0.0],
                    avv=[0.0, 0.0, 0.0], gndnst=float(ppc), rtech=16, lc=1, tstech=0, galinv=1, species=[sp],
                    walls=ws)

    band = (Ly // 2 - 1, Ly // 2 + 1)
    OBS = {"centreline_velocity": (M.OBS_BAND_VELOCITY, (0, 1, band[0], band[1])), "mean_S": (M.OBS_MEAN_S, None)}
    seed = 13
    out = {}
    for name, obs in OBS.items():
        t0 = time.perf_counter()
        run = H.ReverseRun(M, cfg(g, act), Q, V, K, seed, obs, U0=U, ckpt=ckpt)
        try:
            J = run.forward()
            t1 = time.perf_counter()
            dg, dact = run.backward()
            t2 = time.perf_counter()
        finally:
            run.close()
        out[name] = dict(J=J, dJ_dg=float(dg[0]), dJ_dact=float(dact[0]), forward_s=t1 - t0, backward_s=t2 - t1)

    def tangent(direction):
        """both observables and their tangents along one direction, with the
        reverse runs' call sequence (an observable after every step)"""
        ctx = M.Context(cfg(g, act), N, seed=seed, precision=M.FP64)
        try:
            ctx.upload(0, Q, V, U=U)
            ctx.tangent(True, **direction)
            d = {k: 0.0 for k in OBS}
            for _ in range(K):
                ctx.step()
                ctx.observable(*OBS["centreline_velocity"])
                Qs, Vs, Us = ctx.download(0)
                _, dV, dU = ctx.tangent_download(0)
                y = np.trunc(Qs[:, 1])
                m = (y >= band[0]) & (y < band[1])
                d["centreline_velocity"] += dV[m, 0].mean() / K
                # dS = v0^T dQ v0 of each counted cell's 2x2 order tensor
                cell = np.trunc(Qs[:, 0]).astype(np.int64) * (Ly + 1) + y.astype(np.int64)
                nc = (Lx + 1) * (Ly + 1)
                pop = np.bincount(cell, minlength=nc)
                sxx = np.bincount(cell, Us[:, 0] ** 2, nc) * 2 / np.maximum(pop, 1) - 1
                sxy = np.bincount(cell, Us[:, 0] * Us[:, 1], nc) * 2 / np.maximum(pop, 1)
                tr2 = np.sqrt(sxx ** 2 + sxy ** 2)  # the eigenvalue of the traceless tensor
                v0 = np.stack([sxx + tr2, sxy], 1)
                v0 /= np.maximum(np.linalg.norm(v0, axis=1), 1e-300)[:, None]
                w = v0[cell]
                term = (dU[:, 0] * w[:, 0] + dU[:, 1] * w[:, 1]) * (Us[:, 0] * w[:, 0] + Us[:, 1] * w[:, 1])
                dS = 4.0 * np.bincount(cell, term, nc) / np.maximum(pop, 1)
                d["mean_S"] += dS[pop > 1].mean() / K
            return d
        finally:
            ctx.close()

    t0 = time.perf_counter()
    tg = tangent(dict(dgrav=[1.0, 0.0, 0.0]))
    ta = tangent(dict(dact=[1.0]))
    t_tan = (time.perf_counter() - t0) / 2

    def J_of(gg, aa):
        ctx = M.Context(cfg(gg, aa), N, seed=seed, precision=M.FP64)
        ctx.upload(0, Q, V, U=U)
        j = {k: 0.0 for k in OBS}
        for _ in range(K):
            ctx.step()
            for k, o in OBS.items():
                j[k] += ctx.observable(*o)[0][0] / K
        ctx.close()
        return j

    rel = lambda a, b: abs(a - b) / max(abs(b), 1e-300)  # noqa: E731
    fd_g = {h: (J_of(g + h, act), J_of(g - h, act)) for h in hg}
    fd_a = {h: (J_of(g, act + h), J_of(g, act - h)) for h in ha}
    for name in OBS:
        o = out[name]
        o["tangent_dJ_dg"], o["tangent_dJ_dact"] = tg[name], ta[name]
        o["rel_reverse_vs_tangent"] = max(rel(o["dJ_dg"], tg[name]), rel(o["dJ_dact"], ta[name]))
        o["fd_dJ_dg"] = {str(h): (p[name] - m[name]) / (2 * h) for h, (p, m) in fd_g.items()}
        o["fd_dJ_dact"] = {str(h): (p[name] - m[name]) / (2 * h) for h, (p, m) in fd_a.items()}
        o["rel_err_fd_g"] = {k: rel(o["dJ_dg"], v) for k, v in o["fd_dJ_dg"].items()}
        o["rel_err_fd_act"] = {k: rel(o["dJ_dact"], v) for k, v in o["fd_dJ_dact"].items()}
    return dict(
        workload=f"2D active nematic channel {Lx}x{Ly} cells x {ppc} ppc ({N} particles), lc 1, DIPOLE_DIR_SUM, "
                 f"U 10, tumbling 0.5, rfc 0.01, activity {act}, g {g}, homeotropic strong anchoring at bounce-back "
                 f"phantom walls, K {K} steps, checkpoint every {ckpt}; J = time average of the x-velocity of the "
                 f"centreline band y in [{band[0]}, {band[1]}) or of the mean S of the cells",
        tangent_s_per_direction=t_tan, **out)


def adjoint_check(M, Lx=128, Ly=64, ppc=20, K=200, g=5e-4, angle=2.0, h=1e-11):
    """gradient of the time-averaged centreline velocity of a 2D SRD channel
    (bounce-back phantom walls in y, body force g along x) with respect to g
    and the rotation angle: adjoint (checkpoint every 20 steps) against central
    finite differences with common random numbers.  The pathwise gradient
    holds cell assignments fixed, so h must be small enough that no particle
    changes cell between the two runs (at this size h = 1e-9 already moves a
    few dozen particles across cell boundaries over 200 steps); 1e-11 is
    below that and above FP64 round-off of the observable."""
    import importlib
    H = importlib.import_module("niamh_mpcd_amd.adjoint")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    rng = np.random.default_rng(3)
    N = Lx * Ly * ppc
    Q = np.zeros((N, 3))
    Q[:, 0] = rng.uniform(0, Lx, N)
    Q[:, 1] = rng.uniform(0, Ly, N)
    V = np.zeros((N, 3))
    V[:, :2] = rng.normal(size=(N, 2))
    V -= V.mean(axis=0)

    def cfg(gg, aa):
        sp = dict(mass=1.0, vdist=4)
        ws = [dict(A=[1.0, 0, 0], R=0.0, DN=float(Lx)), dict(A=[-1.0, 0, 0], R=-float(Lx), DN=float(Lx)),
              dict(A=[0, 1.0, 0], R=0.0, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1),
              dict(A=[0, -1.0, 0], R=-float(Ly), DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1)]
        return dict(dim=2, L=[Lx, Ly, 1], dt=0.1, kbt=1.0, rotang=aa, tau=0.5, grav=[gg, 0.0, 0.0],
                    avv=[0.0, 0.0, 0.0], gndnst=float(ppc), rtech=0, lc=0, tstech=0, galinv=1, species=[sp],
                    walls=ws)

    band = (Ly // 2 - 1, Ly // 2 + 1)
    t0 = time.perf_counter()
    run = H.AdjointRun(M, cfg(g, angle), Q, V, K, 11, band, ckpt=20)
    J = run.forward()
    t1 = time.perf_counter()
    dg, da = run.backward()
    t2 = time.perf_counter()
    run.close()

    def J_of(gg, aa):
        r = H.AdjointRun(M, cfg(gg, aa), Q, V, K, 11, band, ckpt=20)
        j = r.forward()
        r.close()
        return j

    fd_g = (J_of(g + h, angle) - J_of(g - h, angle)) / (2 * h)
    fd_a = (J_of(g, angle + h) - J_of(g, angle - h)) / (2 * h)
    return dict(workload=f"2D SRD channel {Lx}x{Ly} x {ppc} ppc, phantom bounce-back walls in y, g={g} along x, "
                         f"rotation angle {angle}, K={K} steps, checkpoint every 20, FP64; "
                         f"J = time-averaged mean v_x of the cells {band[0]}..{band[1] - 1} across the channel",
                J=J, dJ_dg=float(dg[0]), dJ_dangle=da, fd_dJ_dg=fd_g, fd_dJ_dangle=fd_a, fd_h=h,
                rel_err_g=abs(dg[0] - fd_g) / abs(fd_g), rel_err_angle=abs(da - fd_a) / abs(fd_a),
                forward_s=t1 - t0, backward_s=t2 - t1)


def lc_throughput(M, L=256, ppc=20, steps=40, warmup=5):
    """the 2D active-nematic workload (256^2 x 20 ppc,
    lc = 1, DIPOLE_DIR_SUM, U = 10, tumble 0.5, rfc 0.01, activity 0.05,
    random isotropic orientations): particle-steps/s of the split path in FP64
    and in the FP32-state form, with the per-kernel times"""
    import importlib
    V_ = importlib.import_module("niamh_mpcd_amd.validate")
    out = {}
    for name, prec in (("fp64", M.FP64), ("fp32", M.FP32)):
        cfg = dict(dim=2, L=[L, L, 1], dt=0.1, kbt=1.0, rotang=ROTANG_130, tau=0.5, grav=[0, 0, 0], avv=[0, 0, 0],
                   gndnst=float(ppc), rtech=16, lc=1, tstech=0, galinv=1,
                   species=[V_.species(mfpot=10.0, tumble=0.5, rfc=0.01, act=0.05)], walls=V_.periodic_walls(2, (L, L)))
        N = L * L * ppc
        ctx = M.Context(cfg, N, seed=31, precision=prec)
        rng = np.random.default_rng(31)
        Q = np.zeros((N, 3))
        Q[:, 0] = rng.uniform(0, L, N)
        Q[:, 1] = rng.uniform(0, L, N)
        V = np.zeros((N, 3))
        V[:, :2] = rng.normal(size=(N, 2))
        V -= V.mean(axis=0)
        th = rng.uniform(0, 2 * math.pi, N)
        U = np.zeros((N, 3))
        U[:, 0] = np.cos(th)
        U[:, 1] = np.sin(th)
        ctx.upload(0, Q, V, U)
        r = timed_run(M, ctx, steps, warmup, N)
        ctx.close()
        # the dominant kernel against HBM with the 2D nematic step's algorithmic bytes
        # (r, v, u read and written once: 96 / 48 B per particle-step)
        rl = roofline(r["timing"], N, name, b_alg=B_ALG_LC2D[name], traffic_tag=f"lc2d_{name}", dim=2, lc=True)
        out[name] = dict(value=r["value"], ms_per_step=r["ms_per_step"],
                         kernels_ms_per_step={k: v[0] / r["prof_steps"] for k, v in r["timing"].items() if v[1]},
                         roofline=rl, step_hbm_frac=r["value"] * B_ALG_LC2D[name] / 1e9 / HBM_PEAK_GBS)
    out["workload"] = (f"2D active nematic {L}x{L} x {ppc} ppc (lc=1, collision 16 DIPOLE_DIR_SUM, U=10, tumble 0.5, "
                       f"rfc 0.01, activity 0.05), {steps} timed steps after {warmup}; split path")
    out["unit"] = "particle-steps/s"
    return out


def copy_probe(mib):
    """a plain copy of known size: what FETCH_SIZE / WRITE_SIZE read for it
    calibrates the counters of the step kernels"""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--replicas", type=int, default=8)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--ppc", type=int, default=10)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-validate", action="store_true", help="skip the FP32-vs-FP64 observable report")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the pathwise gradient checks (adjoint of a 2D SRD channel, tangent of an active "
                         "nematic channel; gradients vs finite differences)")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-replica MPC-AT+a channel ensemble (throughput + Poiseuille viscosities)")
    ap.add_argument("--no-lc", action="store_true", help="skip the 2D active-nematic throughput line")
    ap.add_argument("--only", choices=["fp32", "fp64", "both", "lc", "ensemble", "gradients"], default="both",
                    help="lc / ensemble: only the 2D active nematic or the 64-replica MPC-AT+a ensemble throughput "
                         "(--steps timed steps; its JSON instead of the bench line); gradients: only the reverse "
                         "mode, tangent and finite differences of the active nematic channel")
    ap.add_argument("--fill", type=float, default=0.0,
                    help="fix the bucket fill of the run layout instead of autotuning it (reproducible profiles)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "replicas would shard across GPUs with no communication", file=sys.stderr)
        sys.exit(2)
    M = package()
    if args.only == "lc":
        print(json.dumps(lc_throughput(M, steps=args.steps, warmup=args.warmup)))
        return
    if args.only == "gradients":
        print(json.dumps(dict(channel_128x64=nematic_gradients(M, hg=(1e-11, 1e-12), ha=(1e-12, 1e-13)),
                              box_32x16=nematic_gradients(M, Lx=32, Ly=16, K=100, hg=(1e-10,), ha=(1e-10,)))))
        return
    if args.only == "ensemble":
        import importlib
        e = importlib.import_module("niamh_mpcd_amd.batched").run(M, M.FP32_EXACT, steps=0, time_steps=args.steps)
        e["roofline"] = roofline({"bucket": (e["bucket_avg_launch_ms"], 1)}, e["particles"], "fp32x",
                                 traffic_tag="ensemble_fp32x")
        e.pop("per_replica")
        print(json.dumps(e))
        return
    L = (args.L,) * 3
    cfg = srd_cfg(L, args.ppc)
    result = {}
    if args.only in ("fp32", "both"):
        B = args.replicas
        n1 = int(np.prod(L)) * args.ppc

        def batched_ctx(prec, seed):
            x = M.Context([cfg] * B, [n1] * B, seed=seed, precision=prec)
            for r in range(B):
                Q, V = initial_state(L, args.ppc, seed=100 + r)
                x.upload(r, Q, V)
            if args.fill > 0:
                x.set_layout(True, fill=args.fill)
                return x, 0.0
            return x, x.autotune(3)

        # the headline: the exact-sum FP32 form (order-free cell sums, 28-byte records)
        ctx, tuned_ms = batched_ctx(M.FP32_EXACT, 2024)
        print(f"[bench] fp32 exact-sum batched: {B} x {n1} particles, tile fill {ctx.tile_fill():.2f} "
              f"({tuned_ms:.3f} ms/step in the autotuner), warmup {args.warmup}, steps {args.steps}",
              file=sys.stderr, flush=True)
        r32x = timed_run(M, ctx, args.steps, args.warmup, B * n1)
        r32x["graph"] = graph_run(ctx, args.steps, B * n1)
        r32x["tile_fill"] = ctx.tile_fill()
        result["fp32x"] = (r32x, B * n1, B)
        ctx.close()
        # the FP32 state stepped in the reference's list order (FP64 cell sums in list order)
        ctx, tuned_ms = batched_ctx(M.FP32, 2024)
        print(f"[bench] fp32 list-order batched: tile fill {ctx.tile_fill():.2f}", file=sys.stderr, flush=True)
        r32 = timed_run(M, ctx, max(20, args.steps // 2), args.warmup, B * n1)
        r32["graph"] = graph_run(ctx, max(20, args.steps // 2), B * n1)
        r32["tile_fill"] = ctx.tile_fill()
        # the same steps with the running flow field on (the reference's flow
        # output, localFLOW + sumFLOW every step): what it adds to a step
        kf = max(10, args.steps // 4)
        ctx.flow(True)
        rf = timed_run(M, ctx, kf, 2, B * n1)
        ctx.flow(False)
        r32["flow_field"] = dict(steps=kf, ms_per_step=rf["ms_per_step"],
                                 flow_kernels_ms_per_step=rf["timing"]["flow"][0] / rf["prof_steps"],
                                 summed_steps=ctx.flow_read(0)[1])
        # the same on the cell-sorted layout (the path of parity steps,
        # thermostats and the LC split path): FP64 atomics per run of equal
        # cells in a wavefront instead of tile partials
        ctx.set_layout(False)
        ctx.flow(True)
        rd = timed_run(M, ctx, kf, 2, B * n1)
        ctx.flow(False)
        r32["flow_field"]["cell_sorted"] = dict(ms_per_step=rd["ms_per_step"],
                                                flow_kernels_ms_per_step=rd["timing"]["flow"][0] / rd["prof_steps"])
        result["fp32"] = (r32, B * n1, B)
        ctx.close()
    if args.only in ("fp64", "both"):
        n1 = int(np.prod(L)) * args.ppc
        ctx = M.Context(cfg, n1, seed=2025, precision=M.FP64)
        Q, V = initial_state(L, args.ppc, seed=7)
        ctx.upload(0, Q, V)
        k64 = args.fp64_steps or args.steps
        if args.fill > 0:
            ctx.set_layout(True, fill=args.fill)
        else:
            ctx.autotune(3)
        print(f"[bench] fp64 plain: {n1} particles, tile fill {ctx.tile_fill():.2f}, warmup {args.warmup}, "
              f"steps {k64}", file=sys.stderr, flush=True)
        r64 = timed_run(M, ctx, k64, args.warmup, n1)
        r64["graph"] = graph_run(ctx, k64, n1)
        r64["tile_fill"] = ctx.tile_fill()
        result["fp64"] = (r64, n1, 1)
        ctx.close()

    validation = None
    if not args.no_validate and args.only == "both":
        import importlib
        print("[bench] FP32-state observables vs FP64", file=sys.stderr, flush=True)
        validation = importlib.import_module("niamh_mpcd_amd.validate").fp32_report(M, quick=True)
    ensemble = None
    if not args.no_ensemble and args.only == "both":
        import importlib
        print("[bench] 64-replica MPC-AT+a channel ensemble", file=sys.stderr, flush=True)
        e = importlib.import_module("niamh_mpcd_amd.batched").run(M, M.FP32_EXACT)
        by_ppc = {}
        for x in e["per_replica"]:
            by_ppc.setdefault(x["ppc"], []).append(x["nu"])
        th = {x["ppc"]: x["nu_theory"] for x in e["per_replica"]}
        ensemble = {k: e[k] for k in ("particles", "replicas", "value", "ms_per_step", "steps", "samples",
                                      "kernels_ms_per_step")}
        # the bucketed MPC-AT+a kernel against HBM (48 B per particle-step, 3D FP32 state)
        ensemble["roofline"] = roofline({"bucket": (e["bucket_avg_launch_ms"], 1)}, e["particles"], "fp32x",
                                        traffic_tag="ensemble_fp32x")
        ensemble["step_hbm_frac"] = e["value"] * B_ALG["fp32x"] / 1e9 / HBM_PEAK_GBS
        ensemble.update(
            workload="64 replicas of 3D MPC-AT+a channels: L^3 with L ~ U{16..48}, ppc ~ U{5..20}, dt 0.05, kT 1, "
                     "bounce-back phantom walls in z, g = 1e-3 along x, exact-sum FP32 form, autotuned tiles, "
                     "hipGraph replays",
            unit="particle-steps/s",
            viscosity_by_ppc={str(k): dict(mean=float(np.mean(v)), n=len(v), theory=th[k],
                                           ratio=float(np.mean(v)) / th[k]) for k, v in sorted(by_ppc.items())},
            viscosity_theory="MPC-AT+a, Noguchi & Gompper 2008 (init.c:849-872), Poisson cell populations of mean ppc",
            per_replica=[[x["L"], x["ppc"], round(x["nu"], 4), round(x["nu_theory"], 4)] for x in e["per_replica"]])
    lc = None
    if not args.no_lc and args.only == "both":
        print("[bench] 2D active nematic (LC) throughput", file=sys.stderr, flush=True)
        lc = lc_throughput(M)
    adjoint = None
    if not args.no_adjoint and args.only == "both":
        print("[bench] pathwise adjoint vs finite differences", file=sys.stderr, flush=True)
        adjoint = adjoint_check(M)
    tangent = None
    if not args.no_adjoint and args.only == "both":
        print("[bench] reverse mode and tangent of the active nematic channel vs finite differences",
              file=sys.stderr, flush=True)
        tangent = dict(channel_128x64=nematic_gradients(M, hg=(1e-11, 1e-12), ha=(1e-12, 1e-13)),
                       box_32x16=nematic_gradients(M, Lx=32, Ly=16, K=100, hg=(1e-10,), ha=(1e-10,)))
    cpu = None if args.no_cpu else cpu_baseline()
    small = None
    if args.only == "both":
        print("[bench] 2D SRD 30x30 x 20 ppc, eager and hipGraph", file=sys.stderr, flush=True)
        small = small_box(M)
        if not args.no_cpu:
            small["cpu_reference"] = cpu_baseline(L=(30, 30), ppc=20, steps=1000, dim=2, rtech=1,
                                                  rotang=math.pi / 2)
    head_key = "fp32x" if "fp32x" in result else "fp64"
    r, n, B = result[head_key]
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps if head_key == "fp32x" else (args.fp64_steps or args.steps),
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state / exact f64 cell sums" if head_key == "fp32x" else "f64",
        "data": "synthetic (uniform positions, Maxwell-Boltzmann velocities)",
        "config": {
            "workload": f"3D SRD {L[0]}^3 cells x {args.ppc} ppc, ARBAXIS 130 deg, dt 0.1, kT 1, random grid shift, "
                        f"periodic; {'FP32 state / exact-sum FP64 cell sums, ' + str(B) + ' replicas' if head_key == 'fp32x' else 'FP64, 1 replica'}",
            "replicas": B, "particles": n, "collision": "SRD ARBAXIS",
            "precision": "fp32_exact" if head_key == "fp32x" else head_key,
        },
        "roofline": roofline(r["timing"], n, head_key),
        "cpu_baseline": cpu,
        "kernels_ms_per_step": {k: v[0] / r["prof_steps"] for k, v in r["timing"].items() if v[1]},
        # the same workload replayed as hipGraphs of two fused steps (no per-kernel event timing)
        "hipgraph": r["graph"],
        "tile_fill": r["tile_fill"],
        "step_hbm_frac": r["value"] * B_ALG[head_key] / 1e9 / HBM_PEAK_GBS,
        "fp32_validation": validation,
        "ensemble_mpcat_a": ensemble,
        "adjoint": adjoint,
        "gradients_active_nematic": tangent,
        "lc_active_nematic": lc,
        "flow_field": result["fp32"][0].get("flow_field") if "fp32" in result else None,
        "small_box_2d": small,
    }
    if "fp32" in result:
        r32, n32, _ = result["fp32"]
        line["fp32_list_order"] = {"value": r32["value"], "ms_per_step": r32["ms_per_step"],
                                   "hipgraph": r32["graph"], "tile_fill": r32["tile_fill"],
                                   "roofline": roofline(r32["timing"], n32, "fp32"),
                                   "kernels_ms_per_step": {k: v[0] / r32["prof_steps"]
                                                           for k, v in r32["timing"].items() if v[1]},
                                   "step_hbm_frac": r32["value"] * B_ALG["fp32"] / 1e9 / HBM_PEAK_GBS}
    if "fp64" in result and head_key != "fp64":
        r64, n64, _ = result["fp64"]
        line["fp64_plain"] = {"value": r64["value"], "ms_per_step": r64["ms_per_step"], "hipgraph": r64["graph"],
                              "roofline": roofline(r64["timing"], n64, "fp64"),
                              "step_hbm_frac": r64["value"] * B_ALG["fp64"] / 1e9 / HBM_PEAK_GBS}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
