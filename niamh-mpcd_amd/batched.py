"""The batched ensemble: many independent MPC-AT+a channel replicas of
different box sizes, densities and seeds advancing together in one context.

Each replica is a planar channel: periodic in x and y, no-slip bounce-back
walls with phantom particles at z = 0 and z = H (read.c:1514-1520,
domainWalls == 2), body force g along x, started from rest.  The flow is
the start-up of Poiseuille flow,

    v_x(z, t) = g z (H - z) / (2 nu)
                - sum_{n odd} 4 g H^2 / (nu pi^3 n^3) sin(n pi z / H) exp(-n^2 pi^2 nu t / H^2),

so every replica yields a kinematic viscosity from its sampled,
population-weighted cell velocities (localPROP VCM, mpc.c:162-165) by a
one-parameter least-squares fit of nu, whether or not the largest channels
have reached the steady parabola.
"""
from __future__ import annotations

import math
import time

import numpy as np

from .validate import mpcat_a_viscosity


def channel_walls(L):
    ws = []
    for ax in range(2):
        a1 = [0.0, 0.0, 0.0]
        a1[ax] = 1.0
        a2 = [0.0, 0.0, 0.0]
        a2[ax] = -1.0
        ws += [dict(A=a1, R=0.0, DN=float(L[ax])), dict(A=a2, R=-float(L[ax]), DN=float(L[ax]))]
    wall = dict(DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1)
    ws += [dict(A=[0.0, 0.0, 1.0], R=0.0, **wall), dict(A=[0.0, 0.0, -1.0], R=-float(L[2]), **wall)]
    return ws


def ensemble(nrep=64, lmin=16, lmax=48, ppc_min=5, ppc_max=20, seed=2024):
    """replica boxes (same L on all axes) and densities drawn uniformly"""
    rng = np.random.default_rng(seed)
    Ls = rng.integers(lmin, lmax + 1, nrep)
    ppcs = rng.integers(ppc_min, ppc_max + 1, nrep)
    return [(int(L), int(p)) for L, p in zip(Ls, ppcs)]


def channel_cfg(L, ppc, g=1e-3, dt=0.05, kbt=1.0):
    sp = dict(mass=1.0, rfc=0.01, tumble=0.5, chihi=0.5, act=0.0, mfpot=0.0, damp=0.0, sigwidth=1.0, sigpos=1.0,
              minactratio=0.0, vdist=4)
    return dict(dim=3, L=[L, L, L], dt=dt, kbt=kbt, rotang=math.pi / 2, tau=0.5, grav=[g, 0.0, 0.0],
                avv=[0.0, 0.0, 0.0], gndnst=float(ppc), rtech=3, lc=0, tstech=0, galinv=1, species=[sp],
                walls=channel_walls((L, L, L)))


def startup_profile(z, t, H, g, nu, nmax=199):
    """start-up Poiseuille profile at heights z and time t"""
    v = g * z * (H - z) / (2.0 * nu)
    for n in range(1, nmax + 1, 2):
        v = v - 4.0 * g * H * H / (nu * math.pi ** 3 * n ** 3) * np.sin(n * math.pi * z / H) * \
            math.exp(-n * n * math.pi ** 2 * nu * t / (H * H))
    return v


def fit_startup(profiles, times, H, g):
    """least-squares nu of sampled profiles (cells at heights c + 1/2)"""
    from scipy.optimize import minimize_scalar
    z = np.arange(H) + 0.5

    def cost(lognu):
        nu = math.exp(lognu)
        return sum(float(np.sum((p - startup_profile(z, t, H, g, nu)) ** 2)) for p, t in zip(profiles, times))

    res = minimize_scalar(cost, bounds=(math.log(0.02), math.log(20.0)), method="bounded",
                          options=dict(xatol=1e-6))
    return float(math.exp(res.x))


def run(M, precision, reps=None, steps=3000, sample_from=600, every=100, g=1e-3, dt=0.05, seed=7, time_steps=400):
    """runs the ensemble; returns the throughput of a plain stretch of graph
    replays and each replica's Poiseuille viscosity"""
    reps = reps or ensemble()
    cfgs = [channel_cfg(L, p, g=g, dt=dt) for L, p in reps]
    ns = [L ** 3 * p for L, p in reps]
    ctx = M.Context(cfgs, ns, seed=seed, precision=precision)
    for r, (L, p) in enumerate(reps):
        rng = np.random.default_rng(seed * 1000 + r)
        n = ns[r]
        Q = rng.uniform(0.0, L, (n, 3))
        V = rng.normal(size=(n, 3))
        V -= V.mean(axis=0)
        ctx.upload(r, Q, V)
    trial = 2
    tile_ms = ctx.autotune(trial)  # 6 bucket fills x (1 + trial) real steps
    # throughput: graph replays with nothing else in the timed region (the
    # warm-up records the graph: groups of six steps in the exact-sum form)
    warm = 12
    ctx.run(warm, use_graph=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(time_steps, use_graph=True)
    ctx.synchronize()
    sec = time.perf_counter() - t0
    # kernel times of a few eager steps (HIP events on the library stream)
    ctx.profile(True)
    ctx.run(4, use_graph=False)
    ctx.synchronize()
    kern = ctx.timing()
    ctx.profile(False)
    done = 6 * (1 + trial) + warm + time_steps + 4  # autotune trials, graph warm-up, timed stretch, profiled steps
    prof = [[] for _ in reps]
    times = []
    nz = [L + 1 for L, _ in reps]
    samples = 0
    while done < steps:
        k = min(every, steps - done)
        ctx.run(k, use_graph=True)
        done += k
        if done < sample_from:
            continue
        samples += 1
        times.append(done)
        for r, (L, _) in enumerate(reps):
            c = ctx.cells(r)
            z = np.arange(ctx.ncells[r]) % nz[r]
            ok = z < L
            pop = c["pop"][ok].astype(np.float64)
            mom = np.bincount(z[ok], weights=pop * c["vcm"][ok, 0], minlength=L)[:L]
            w = np.bincount(z[ok], weights=pop, minlength=L)[:L]
            prof[r].append(mom / np.maximum(w, 1.0))
    out = []
    for r, (L, p) in enumerate(reps):
        nu_th = mpcat_a_viscosity(float(p), dt)["nu"]
        if not times:
            out.append(dict(L=L, ppc=p, nu=None, nu_theory=nu_th))
            continue
        prof_r = np.asarray([pr for pr in prof[r]])
        nu = fit_startup(prof_r, np.asarray(times) * dt, L, g)
        tau = L * L / (math.pi ** 2 * max(nu, 1e-9)) / dt  # slowest mode, in steps
        out.append(dict(L=L, ppc=p, nu=nu, nu_theory=nu_th, vmax=float(prof_r[-1].max()),
                        steps_over_tau=float(steps / tau)))
    ctx.close()
    N = int(sum(ns))
    return dict(particles=N, replicas=len(reps), value=N * time_steps / sec, ms_per_step=1e3 * sec / time_steps,
                tile_ms=tile_ms, steps=steps, samples=samples, per_replica=out,
                kernels_ms_per_step={k: v[0] / 4 for k, v in kern.items() if v[1]},
                bucket_avg_launch_ms=kern["bucket"][0] / max(kern["bucket"][1], 1))
