"""Physics checks of the FP32-state forms (MPCD_FP32 and the exact-sum
MPCD_FP32_EXACT) against the FP64 path.

Both precisions run the same replicas from the same initial states with the
same counter-based draws (the generator is keyed by seed, replica, step,
purpose and cell or slot, not by the precision), and each observable is
reported with its relative error against the FP64 value:

  * temperature after the run (TEMP, therm.c:209-235, about the mean velocity),
  * total-momentum drift per particle,
  * the kinematic shear viscosity from the decay of a transverse shear wave
    v_x(z) = A sin(2 pi z / L_z), next to the analytic kinetic-plus-collisional
    SRD value with a random grid shift (Kikuchi et al. 2003, 3D, cell size 1:
    nu_kin = (kT dt / 2m) [5n / ((n - 1 + e^-n)(2 - cos a - cos 2a)) - 1],
    nu_coll = (1 / (18 dt)) ((n - 1 + e^-n) / n)(1 - cos a)),
  * the mean nematic order S of an active nematic (per-cell Q-tensor
    eigenvalue, mpc.c:189-190).

The FP64 GPU path is the one pinned to the CPU oracle (and through it to the
reference) at 1e-12 relative on the committed fixtures.
"""
from __future__ import annotations

import math

import numpy as np

ROT130 = 2.2689280275926285


def srd_viscosity_3d(n: float, alpha: float, dt: float, kbt: float = 1.0, m: float = 1.0) -> dict:
    """kinematic viscosity of 3D SRD with a random grid shift, cell size 1"""
    f = n - 1.0 + math.exp(-n)
    nu_kin = kbt * dt / (2.0 * m) * (5.0 * n / (f * (2.0 - math.cos(alpha) - math.cos(2.0 * alpha))) - 1.0)
    nu_coll = 1.0 / (18.0 * dt) * (f / n) * (1.0 - math.cos(alpha))
    return dict(nu=nu_kin + nu_coll, nu_kin=nu_kin, nu_coll=nu_coll)


def mpcat_a_viscosity(n: float, dt: float, kbt: float = 1.0, m: float = 1.0, dim: int = 3) -> dict:
    """kinematic viscosity of MPC-AT+a (Andersen thermostat with angular-momentum
    conservation), cell size 1, Poisson-distributed cell populations of mean n:
    the kinetic and collisional parts of Noguchi & Gompper, Phys. Rev. E 78,
    016706 (2008), as theory_trans evaluates them with A = B = 1
    (init.c:849-872; the smrtPow quirk of init.c:812 does not enter here)"""
    e = math.exp(-n)
    c = (1.0 - e * (1.0 + n)) + (1.0 + 1.0 / dim) * n * e / (dim + 2.0)
    c += 0.5 * (dim - 0.5 * (3.0 * dim + 2.0)) * (1.0 - e * (1.0 + n + 0.5 * n * n)) / n
    eta_kin = n * kbt * dt * (1.0 / c - 0.5)
    eta_coll = m / (24.0 * dt) * (n - 7.0 / 5.0 + e * (7.0 / 5.0 + 2.0 * n / 5.0 + (1.0 / dim - 0.3) * n * n))
    return dict(nu=(eta_kin + eta_coll) / (n * m), nu_kin=eta_kin / (n * m), nu_coll=eta_coll / (n * m))


def periodic_walls(dim, L):
    ws = []
    for ax in range(dim):
        a1 = [0.0, 0.0, 0.0]
        a1[ax] = 1.0
        a2 = [0.0, 0.0, 0.0]
        a2[ax] = -1.0
        ws += [dict(A=a1, R=0.0, DN=float(L[ax])), dict(A=a2, R=-float(L[ax]), DN=float(L[ax]))]
    return ws


def species(**kw):
    sp = dict(mass=1.0, rfc=0.01, tumble=0.5, chihi=0.5, act=0.05, mfpot=10.0, damp=0.0, sigwidth=1.0,
              sigpos=1.0, minactratio=0.0, vdist=4)
    sp.update(kw)
    return sp


def srd_cfg(L, ppc, dt=0.1, rotang=ROT130, kbt=1.0, rtech=0):
    return dict(dim=3, L=list(L), dt=dt, kbt=kbt, rotang=rotang, tau=0.5, grav=[0, 0, 0], avv=[0, 0, 0],
                gndnst=float(ppc), rtech=rtech, lc=0, tstech=0, galinv=1, species=[species()],
                walls=periodic_walls(3, L))


def _cell_z(ctx, r, L):
    nz = L[2] + 1
    idx = np.arange(ctx.ncells[r])
    return (idx % nz).astype(np.float64)


def shear_wave(M, precision, L=(64, 64, 32), ppc=10, nrep=8, steps=400, every=10, amp=0.2, seed=5,
               dt=0.1, rotang=ROT130, rtech=0):
    """decay of a transverse shear wave in nrep replicas (collision operator
    rtech: SRD ARBAXIS by default); returns the fitted kinematic viscosity
    (mean and standard error over replicas), the final temperature and the
    total-momentum drift per particle"""
    cfg = srd_cfg(L, ppc, dt=dt, rotang=rotang, rtech=rtech)
    N = int(np.prod(L)) * ppc
    ctx = M.Context([cfg] * nrep, [N] * nrep, seed=seed, precision=precision)
    k = 2.0 * math.pi / L[2]
    p0 = []
    for r in range(nrep):
        rng = np.random.default_rng(1000 * seed + r)
        Q = np.empty((N, 3))
        for d in range(3):
            Q[:, d] = rng.uniform(0.0, L[d], N)
        V = rng.normal(size=(N, 3))
        V -= V.mean(axis=0)
        V[:, 0] += amp * np.sin(k * Q[:, 2])
        V -= V.mean(axis=0)
        ctx.upload(r, Q, V)
        p0.append(V.sum(axis=0))
    zc = [_cell_z(ctx, r, L) for r in range(nrep)]
    sink = [np.sin(k * (z + 0.5)) for z in zc]
    t, A = [], [[] for _ in range(nrep)]
    for s in range(0, steps + 1, every):
        if s:
            ctx.run(every, use_graph=True)
        t.append(s * dt)
        for r in range(nrep):
            c = ctx.cells(r)
            pop = c["pop"].astype(np.float64)
            A[r].append(2.0 * np.sum(pop * c["vcm"][:, 0] * sink[r]) / np.sum(pop))
    t = np.asarray(t)
    A = np.asarray(A)

    def fit(a):
        # log-linear fit while the mode is well above the thermal noise
        use = a > 0.25 * a[0]
        return -np.polyfit(t[use], np.log(a[use]), 1)[0] / (k * k)

    nus = [fit(A[r]) for r in range(nrep)]
    # the replica-mean decay, with a jackknife error over replicas
    nu_mean = fit(A.mean(axis=0))
    jk = np.array([fit(np.delete(A, r, axis=0).mean(axis=0)) for r in range(nrep)])
    nu_jk_err = float(np.sqrt((nrep - 1) / nrep * np.sum((jk - jk.mean()) ** 2)))
    temps, drift = [], []
    for r in range(nrep):
        Q, V, _ = ctx.download(r)
        Vm = V - V.mean(axis=0)
        temps.append(float(np.sum(Vm * Vm) / (3.0 * N)))
        drift.append(float(np.abs(V.sum(axis=0) - p0[r]).max() / N))
    ctx.close()
    nus = np.asarray(nus)
    return dict(nu=float(nu_mean), nu_err=nu_jk_err, nu_rep=nus.tolist(),
                T=float(np.mean(temps)), momentum_drift=float(max(drift)), amplitude=A.mean(axis=0).tolist(),
                t=t.tolist(), particles=N * nrep, A=A, fit=fit)


def paired_rel_err(w32: dict, w64: dict) -> tuple:
    """(nu32 - nu64) / nu64 from the replica-mean decays and its jackknife
    error with the same replica left out of both precisions: the two runs
    share initial states and draws, so their noise is partly common"""
    A32, A64, fit = w32["A"], w64["A"], w64["fit"]
    nrep = A64.shape[0]
    d = (fit(A32.mean(axis=0)) - fit(A64.mean(axis=0))) / fit(A64.mean(axis=0))
    jk = []
    for r in range(nrep):
        a64 = np.delete(A64, r, axis=0).mean(axis=0)
        a32 = np.delete(A32, r, axis=0).mean(axis=0)
        jk.append((fit(a32) - fit(a64)) / fit(a64))
    jk = np.asarray(jk)
    return float(d), float(np.sqrt((nrep - 1) / nrep * np.sum((jk - jk.mean()) ** 2)))


def nematic_order(M, precision, L=(64, 64), ppc=20, steps=300, every=20, start=100, seed=9):
    """mean per-cell scalar order S of a 2D active nematic (collision
    operator 16, lc = 1), averaged over cells with POP > 1 and samples"""
    cfg = dict(dim=2, L=list(L) + [1], dt=0.1, kbt=1.0, rotang=ROT130, tau=0.5, grav=[0, 0, 0], avv=[0, 0, 0],
               gndnst=float(ppc), rtech=16, lc=1, tstech=0, galinv=1,
               species=[species(mfpot=10.0, tumble=0.5, rfc=0.01, act=0.05)], walls=periodic_walls(2, L))
    N = int(np.prod(L)) * ppc
    ctx = M.Context(cfg, N, seed=seed, precision=precision)
    rng = np.random.default_rng(seed)
    Q = np.zeros((N, 3))
    Q[:, 0] = rng.uniform(0, L[0], N)
    Q[:, 1] = rng.uniform(0, L[1], N)
    V = np.zeros((N, 3))
    V[:, :2] = rng.normal(size=(N, 2))
    V -= V.mean(axis=0)
    th = rng.uniform(0, 2 * math.pi, N)
    U = np.zeros((N, 3))
    U[:, 0] = np.cos(th)
    U[:, 1] = np.sin(th)
    ctx.upload(0, Q, V, U)
    S = []
    for s in range(every, steps + 1, every):
        ctx.run(every, use_graph=False)
        if s >= start:
            c = ctx.cells(0)
            ok = c["pop"] > 1
            S.append(float(c["S"][ok].mean()))
    ctx.close()
    return dict(S=float(np.mean(S)), S_series=S, particles=N)


def fp32_report(M, quick: bool = True) -> dict:
    """the FP32-state observables with their relative errors against FP64"""
    # the headline's box (64^3, 10 per cell); the wave's amplitude and the run's
    # length set the statistical error of the decay rate
    kw = dict(L=(64, 64, 64), nrep=32, steps=1200, every=40, amp=0.5) if quick else \
        dict(L=(64, 64, 64), nrep=96, steps=1200, every=40, amp=0.5)
    w64 = shear_wave(M, M.FP64, **kw)
    w32 = shear_wave(M, M.FP32, **kw)
    w32x = shear_wave(M, M.FP32_EXACT, **kw)
    th = srd_viscosity_3d(10.0, ROT130, 0.1)
    n64 = nematic_order(M, M.FP64)
    n32 = nematic_order(M, M.FP32)

    def rel(a, b):
        return abs(a - b) / abs(b)

    paired = paired_rel_err(w32, w64)
    paired_x = paired_rel_err(w32x, w64)

    return dict(
        workload=f"3D SRD {kw['L'][0]}x{kw['L'][1]}x{kw['L'][2]} x 10 ppc x {kw['nrep']} replicas, shear wave "
                 f"A={kw['amp']} along z, {kw['steps']} steps; "
                 "2D active nematic 64^2 x 20 ppc (collision 16, U=10), 300 steps",
        temperature_fp64=w64["T"], temperature_fp32=w32["T"], temperature_rel_err=rel(w32["T"], w64["T"]),
        momentum_drift_per_particle_fp64=w64["momentum_drift"],
        momentum_drift_per_particle_fp32=w32["momentum_drift"],
        viscosity_fp64=w64["nu"], viscosity_fp32=w32["nu"], viscosity_stderr_fp64=w64["nu_err"],
        viscosity_stderr_fp32=w32["nu_err"], viscosity_rel_err=rel(w32["nu"], w64["nu"]),
        viscosity_rel_err_stderr=paired[1], viscosity_signed_rel_err=paired[0],
        viscosity_rel_err_stderr_unpaired=math.hypot(w64["nu_err"], w32["nu_err"]) / w64["nu"],
        viscosity_theory=th["nu"], viscosity_fp64_vs_theory=rel(w64["nu"], th["nu"]),
        S_fp64=n64["S"], S_fp32=n32["S"], S_rel_err=rel(n32["S"], n64["S"]),
        # the exact-sum FP32 form (the batched headline form; SRD, no nematic mode)
        fp32_exact=dict(temperature=w32x["T"], temperature_rel_err=rel(w32x["T"], w64["T"]),
                        momentum_drift_per_particle=w32x["momentum_drift"], viscosity=w32x["nu"],
                        viscosity_stderr=w32x["nu_err"], viscosity_rel_err=rel(w32x["nu"], w64["nu"]),
                        viscosity_signed_rel_err=paired_x[0], viscosity_rel_err_stderr=paired_x[1],
                        viscosity_vs_theory=rel(w32x["nu"], th["nu"])))
