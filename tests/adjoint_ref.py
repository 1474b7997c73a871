"""FP64 CPU adjoint on the oracle (test infrastructure).

The forward pass runs the oracle phase by phase with Philox draws (the same
keyed draws as the GPU's internal generator) and keeps, per step, what the
pathwise adjoint holds fixed: the list order and cell offsets of the shifted
bin, the pre-collision velocities, the cells' (ghost-corrected) VCM and mass,
and each particle's wall factors.  The backward pass is a plain numpy
restatement of the transposed maps, written independently of the HIP kernels:

  lambda' = wmul * (lambda + dJ_t/dv)
  SRD (2D):  lambda1_i = R^T lambda'_i + (m_i / M)(I - R)^T sum_cell lambda'
  dJ/dg += dt sum_i lambda1_i,  dJ/dangle += sum_i lambda'_i . (dR/dangle)(v1_i - VCM)

The observable is the time average over K steps of the mean x-velocity of the
particles in a band of cells across the channel (y cells [lo, hi)).
"""
import math

import numpy as np

import pyoracle as P


def band_mask(Q, lo, hi):
    y = np.trunc(Q[:, 1])
    return (y >= lo) & (y < hi)


def forward(cfg, Q0, V0, K, seed, band, tape=True):
    s = P.System(cfg, Q0, V0)
    d = P.Draws(P.DRAW_PHILOX, seed)
    g = np.asarray(cfg["grav"], dtype=np.float64)
    dim, dt = cfg["dim"], cfg["dt"]
    J = 0.0
    steps = []
    for _ in range(K):
        s.phase("bin", d)
        if tape:
            order, cstart = s.order()
            _, V, _ = s.state()
            v1 = V.copy()
            for j in range(dim):
                v1[:, j] = V[:, j] + dt * g[j]  # acc_all, same rounding as the step
        s.phase("collide", d)
        if tape:
            cells = s.cells()
            _, Vc, _ = s.state()
        s.phase("stream", d)
        Q, Vs, _ = s.state()
        m = band_mask(Q, *band)
        J += Vs[m, 0].mean() / K
        if tape:
            with np.errstate(divide="ignore", invalid="ignore"):
                mult = np.where(Vc != 0.0, np.round(Vs / np.where(Vc != 0.0, Vc, 1.0)), 1.0)
            steps.append(dict(order=order.copy(), cstart=cstart.copy(), v1=v1, vcm=cells["vcm"].copy(),
                              mass=cells["mass"].copy(), mult=mult, mask=m))
    return J, steps


def backward(cfg, steps, K):
    """dJ/dg (3) and dJ/dangle for a 2D SRD run with the reference's
    always-negative rotation sign and no phantom walls"""
    assert cfg["dim"] == 2 and cfg["rtech"] in (0, 1)
    dt = cfg["dt"]
    a = cfg["rotang"]
    sA, cA = math.sin(a), math.cos(a)
    sgn = -1.0  # (long)u > 0.5 never holds for u in [0,1) (mpc.c:1291)
    Sn = sgn * sA
    R = np.array([[cA, -Sn], [Sn, cA]])
    dR = np.array([[-sA, -sgn * cA], [sgn * cA, -sA]])
    N = steps[0]["v1"].shape[0]
    lam = np.zeros((N, 3))
    dg = np.zeros(3)
    da = 0.0
    for st in reversed(steps):
        adj = np.zeros((N, 3))
        mk = st["mask"]
        adj[mk, 0] = 1.0 / (K * mk.sum())
        lp = st["mult"] * (lam + adj)
        new = lp.copy()
        order, cstart = st["order"], st["cstart"]
        for c in range(len(cstart) - 1):
            idx = order[cstart[c]:cstart[c + 1]]
            if len(idx) < 2:
                continue
            L = lp[idx, :2]
            Lc = L.sum(axis=0)
            G = Lc - R.T @ Lc
            w = 1.0 / st["mass"][c]  # unit masses
            new[idx, :2] = L @ R + w * G[None, :]
            dv = st["v1"][idx, :2] - st["vcm"][c, :2]
            da += float(np.sum(L * (dv @ dR.T)))
        dg[:2] += dt * new[:, :2].sum(axis=0)
        lam = new
    return dg, da


def channel_cfg(Lx, Ly, ppc, g, rotang, dt=0.1, phantom=0):
    sp = dict(mass=1.0, rfc=0.01, tumble=2.0, chihi=0.5, act=0.0, mfpot=0.0, damp=0.0, sigwidth=1.0, sigpos=1.0,
              minactratio=0.0, vdist=4, len=0.007)
    ws = [dict(A=[1.0, 0, 0], R=0.0, DN=float(Lx)), dict(A=[-1.0, 0, 0], R=-float(Lx), DN=float(Lx)),
          dict(A=[0, 1.0, 0], R=0.0, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=phantom),
          dict(A=[0, -1.0, 0], R=-float(Ly), DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=phantom)]
    return dict(dim=2, L=[Lx, Ly, 1], dt=dt, kbt=1.0, rotang=rotang, tau=0.5, grav=[g, 0.0, 0.0],
                avv=[0.0, 0.0, 0.0], gndnst=float(ppc), rtech=0, lc=0, tstech=0, galinv=1, species=[sp], walls=ws)


def initial(Lx, Ly, ppc, seed):
    rng = np.random.default_rng(seed)
    N = Lx * Ly * ppc
    Q = np.zeros((N, 3))
    Q[:, 0] = rng.uniform(0, Lx, N)
    Q[:, 1] = rng.uniform(0, Ly, N)
    V = np.zeros((N, 3))
    V[:, :2] = rng.normal(size=(N, 2))
    V -= V.mean(axis=0)
    return Q, V
