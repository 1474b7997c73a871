"""Golden-fixture cases: small systems that exercise every row of the solvent
hot path (grid shift, binning order, cell sums, ghost particles, SRD ARBAXIS /
ORTHAXIS, MPC-AT, MPC-AT+a, nematic LC with and without activity, thermostats,
periodic and bounce-back planar walls, wall velocity, anchoring).

Each case is recorded from the reference's own timestep() (oracle/_ref) by
make_golden.py with injected sequential draws; tests replay the same initial
state through the oracle and the GPU library.

Wall dictionaries use the reference's bc field names (SRDclss.h:128-262);
`domainWalls`-style periodic planes follow read.c:1441-1583.
"""
from __future__ import annotations

import math

import numpy as np


def species(**kw):
    sp = dict(mass=1.0, rfc=0.01, tumble=2.0, chihi=0.5, act=0.05, mfpot=10.0, damp=0.0, sigwidth=1.0,
              sigpos=1.0, minactratio=0.0, vdist=4, len=0.007)
    sp.update(kw)
    return sp


def pbc_walls(ax, L):
    """the two periodic planes of axis `ax` (read.c:1459-1504, domainWalls == 1)"""
    A1 = [0.0, 0.0, 0.0]
    A1[ax] = 1.0
    A2 = [0.0, 0.0, 0.0]
    A2[ax] = -1.0
    return [dict(A=A1, R=0.0, DN=float(L)), dict(A=A2, R=-float(L), DN=float(L))]


def solid_walls(ax, L, **kw):
    """the two no-slip bounce-back planes of axis `ax` (read.c:1514-1520, domainWalls == 2)"""
    A1 = [0.0, 0.0, 0.0]
    A1[ax] = 1.0
    A2 = [0.0, 0.0, 0.0]
    A2[ax] = -1.0
    base = dict(DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1)
    base.update(kw)
    w1 = dict(A=A1, R=0.0, **base)
    w2 = dict(A=A2, R=-float(L), **base)
    return [w1, w2]


def periodic(dim, L):
    ws = []
    for ax in range(dim):
        ws += pbc_walls(ax, L[ax])
    return ws


def CS_pbc_x(L):
    return pbc_walls(0, L)


def case(name, dim, L, ppc, rtech, steps, seed, lc=0, tstech=0, rotang=math.pi / 2, dt=0.1, kbt=1.0,
         grav=(0, 0, 0), walls=None, sp=None, galinv=1, tau=0.5, qbox=None, kick=None, zeronetmom=0, avv=None):
    c = dict(name=name, dim=dim, L=list(L) + [1] * (3 - len(L)), ppc=ppc, rtech=rtech, steps=steps,
             seed=seed, lc=lc, tstech=tstech, rotang=rotang, dt=dt, kbt=kbt, grav=list(grav),
             walls=walls if walls is not None else periodic(dim, L), species=[sp or species()],
             galinv=galinv, tau=tau)
    if qbox is not None:
        c["qbox"] = [list(map(float, b)) for b in qbox]  # initial positions: [lo, hi) per axis
    if kick is not None:
        c["kick"] = [[int(i), [float(x) for x in v]] for i, v in kick]  # particle i starts with velocity v
    if zeronetmom:
        # after every step s with s % zeronetmom == 0: the rest-frame transformation of outputResults
        # (galileantrans + zeroExtraDims, pout.c:2557-2562)
        c["zeronetmom"] = int(zeronetmom)
    if avv is not None:
        c["avv"] = [float(x) for x in avv]  # the AVV given to the reference instead of the initial mean velocity
    return c


CASES = [
    # 2D SRD, ORTHAXIS, 90 degrees (the reference's CPU-runnable configuration, scaled down)
    case("srd2d_orth", 2, (8, 8), 10, rtech=1, steps=4, seed=11),
    # 2D SRD, ARBAXIS, gravity (acc_all before the cell sums), thermostat VSC
    case("srd2d_arb_grav_vsc", 2, (10, 6), 8, rtech=0, steps=4, seed=12, grav=(0.01, -0.02, 0), tstech=1),
    # 3D SRD, ARBAXIS, 130 degrees
    case("srd3d_arb", 3, (6, 5, 4), 8, rtech=0, steps=3, seed=13, rotang=2.2689280275926285),
    # 3D SRD, ORTHAXIS (keeps the reference's x-axis matrix entry -0.1*S)
    case("srd3d_orth", 3, (5, 5, 5), 6, rtech=1, steps=3, seed=14),
    # 3D MPC-AT with damping
    case("at3d_damp", 3, (5, 4, 6), 7, rtech=2, steps=3, seed=15, sp=species(damp=0.1)),
    # 3D MPC-AT+a channel: z bounce-back walls with ghost particles, gravity in x, Berendsen
    case("rat3d_channel", 3, (6, 5, 6), 8, rtech=3, steps=4, seed=16, grav=(1e-3, 0, 0), dt=0.05, tstech=2,
         walls=pbc_walls(0, 6) + pbc_walls(1, 5) + solid_walls(2, 6)),
    # 2D MPC-AT+a periodic, two species of different mass
    case("rat2d_2sp", 2, (9, 7), 9, rtech=3, steps=4, seed=17),
    # 2D active nematic (DIPOLE_DIR_SUM), local S
    case("lc2d_active", 2, (8, 8), 12, rtech=16, lc=1, steps=3, seed=18,
         sp=species(mfpot=10.0, tumble=0.5, rfc=0.01, act=0.05)),
    # 2D active nematic channel: y walls, homeotropic strong anchoring, gravity; DIPOLE_DIR_SIG_SUM
    case("lc2d_channel", 2, (10, 6), 12, rtech=21, lc=1, steps=3, seed=19, grav=(5e-4, 0, 0),
         sp=species(mfpot=10.0, tumble=0.5, rfc=0.01, act=0.03),
         walls=pbc_walls(0, 10) + solid_walls(1, 6, MUN=1.0, MUT=0.0)),
    # 2D passive nematic with global S (lc = 2)
    case("lcg2d_passive", 2, (7, 7), 10, rtech=3, lc=2, steps=3, seed=20, sp=species(mfpot=5.0, tumble=1.5)),
    # 3D nematic, strong field: Gaussian Maier-Saupe sampler
    case("lc3d_gauss", 3, (4, 4, 4), 10, rtech=3, lc=1, steps=3, seed=21, sp=species(mfpot=12.0)),
    # 3D nematic, weak field: exponential and Metropolis samplers; DIPOLE_DIR_AV
    case("lc3d_weak", 3, (4, 4, 3), 10, rtech=17, lc=1, steps=3, seed=22, sp=species(mfpot=0.6)),
    # 2D SRD Couette: bounce-back walls in y, the top one moving (wall velocity in velBC)
    case("srd2d_couette", 2, (8, 6), 10, rtech=0, steps=4, seed=23,
         walls=pbc_walls(0, 8) + [dict(A=[0, 1, 0], R=0.0, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1),
                                  dict(A=[0, -1, 0], R=-6.0, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1,
                                       V=[0.2, 0, 0])]),
]

# bin()'s escape repair (mpc.c:970-982) and the wall re-thermalisation after
# NBOUNCE bounces (bc.c:1461-1488), which the periodic and channel cases above
# never reach:
CASES += [
    # 2D SRD with an open y axis (no wall there): particles drift out and are
    # re-placed, in the shifted bin and in the unshifted bin after streaming
    case("srd2d_open", 2, (8, 8), 10, rtech=0, steps=4, seed=24, dt=1.0, walls=pbc_walls(0, 8)),
    # the same for a 2D active nematic: escapes also redraw the orientation
    case("lc2d_open", 2, (8, 8), 12, rtech=16, lc=1, steps=3, seed=25, dt=1.0, walls=pbc_walls(0, 8),
         sp=species(mfpot=10.0, tumble=0.5, rfc=0.01, act=0.05)),
    # 3D MPC-AT, open z, y walls inside the box (at y = 1 and y = 5 of 6): a
    # re-placed position beyond them is redrawn (replacePos_WithCheck)
    case("at3d_open_inset", 3, (6, 6, 6), 6, rtech=2, steps=3, seed=26, dt=1.0,
         walls=pbc_walls(0, 6) + [dict(A=[0, 1, 0], R=1.0, DN=0.0, MVN=-1.0, MVT=-1.0),
                                  dict(A=[0, -1, 0], R=-5.0, DN=0.0, MVN=-1.0, MVT=-1.0)],
         qbox=[(0, 6), (1.2, 4.8), (0, 6)]),
    # 3D nematic with an open z axis: the 3D orientation redraw
    case("lc3d_open", 3, (4, 4, 4), 8, rtech=3, lc=1, steps=3, seed=27, dt=1.0,
         walls=pbc_walls(0, 4) + pbc_walls(1, 4), sp=species(mfpot=0.6)),
    # 2D MPC-AT channel, four species (vDist 0, 2, 3, 4): four particles start
    # so fast that their cells bounce more than 2 NBOUNCE times in one step,
    # which rewinds and then re-thermalises them with push()
    case("at2d_nbounce", 2, (8, 4), 10, rtech=2, steps=3, seed=28,
         walls=pbc_walls(0, 8) + solid_walls(1, 4), kick=[(0, (0.3, 2e5)), (1, (-0.2, -2e5)), (2, (0.1, 1.5e5)),
                                                          (3, (0.0, -1.8e5))]),
]

# the periodic rest-frame transformation (galileantrans, mtools.c:673-718,
# with zeroExtraDims) of the output step, every zeronetmom steps; the walls
# exchange momentum with the fluid, so the net momentum drifts between resets
CASES += [
    # 2D SRD Couette, two species of different mass, a given AVV: reset after steps 0, 2 and 4
    case("srd2d_couette_znm", 2, (8, 6), 10, rtech=0, steps=5, seed=29, zeronetmom=2, avv=(0.05, -0.02, 0.0),
         walls=pbc_walls(0, 8) + [dict(A=[0, 1, 0], R=0.0, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1),
                                  dict(A=[0, -1, 0], R=-6.0, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1,
                                       V=[0.3, 0, 0])]),
    # 3D MPC-AT channel with bounce-back walls in z, two species: reset after every step
    case("at3d_channel_znm", 3, (5, 4, 5), 7, rtech=2, steps=3, seed=30, zeronetmom=1, avv=(0.0, 0.01, 0.0),
         walls=pbc_walls(0, 5) + pbc_walls(1, 4) + solid_walls(2, 5)),
]

# the general wall engine (bc.c:55-150, 251-405, 1252-1556): curved surfaces
# (ellipses and circles with the quadratic crossing time, superellipses and
# pipes with the secant method), their periodic images (shiftBC), inverted
# surfaces, wavy planes, and strong anchoring at a curved wall
def _disc(Q, R, P=(2, 2, 2, 2), A=(1.0, 1.0, 0.0), **kw):
    return dict(A=list(A), Q=list(Q), R=R, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1, P=list(P), **kw)


CASES += [
    # 2D MPC-AT, a disc obstacle in a periodic box
    case("circle2d_at", 2, (12, 10), 10, rtech=2, steps=4, seed=31, dt=0.5,
         walls=periodic(2, (12, 10)) + [_disc((6.0, 5.0, 0.0), 3.0)]),
    # 2D SRD, an elliptic obstacle across the periodic x boundary (its image is the wall
    # shiftBC moves for the particles on the far side)
    case("ellipse2d_pbc", 2, (12, 10), 10, rtech=0, steps=4, seed=32, dt=0.5,
         walls=periodic(2, (12, 10)) + [_disc((0.7, 5.0, 0.0), 1.0, A=(1 / 3.0, 1 / 2.5, 0.0))]),
    # 2D SRD, a square-ish superellipse obstacle (P = 8: the secant crossing time)
    case("square2d_srd", 2, (12, 10), 10, rtech=0, steps=4, seed=33, dt=0.5,
         walls=periodic(2, (12, 10)) + [_disc((6.0, 5.0, 0.0), 1.0, P=(8, 8, 1, 8), A=(1 / 3.0, 1 / 3.0, 0.0))]),
    # 3D MPC-AT+a in a pipe along x (inverted cylinder, P = 1 2 2 2: the secant crossing time)
    case("pipe3d_rat", 3, (8, 8, 8), 8, rtech=3, steps=3, seed=34, dt=0.3, grav=(2e-3, 0, 0),
         walls=periodic(3, (8, 8, 8)) + [_disc((0.0, 4.0, 4.0), 3.4, P=(1, 2, 2, 2), A=(0.0, 1.0, 1.0), INV=1)]),
    # 3D MPC-AT in a cylinder along x with P = 2 2 2 2 (the quadratic crossing time)
    case("cyl3d_at", 3, (6, 8, 8), 8, rtech=2, steps=3, seed=35, dt=0.3,
         walls=periodic(3, (6, 8, 8)) + [_disc((0.0, 4.0, 4.0), 3.4, A=(0.0, 1.0, 1.0), INV=1)]),
    # 2D SRD channel between wavy walls (cos modulation along x) inside the flat domain walls
    case("wavy2d_srd", 2, (20, 12), 10, rtech=0, steps=4, seed=36, dt=0.5,
         walls=CS_pbc_x(20) + solid_walls(1, 12) +
         [dict(A=[0, 1, 0], Q=[0, 2.5, 0], R=0.0, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1, P=[1, 1, 1, 1],
               B=[0.8, 2 * math.pi / 10.0, 0.0]),
          dict(A=[0, -1, 0], Q=[0, 9.5, 0], R=0.0, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1, P=[1, 1, 1, 1],
               B=[0.8, 2 * math.pi / 10.0, 0.0])],
         qbox=[(0, 20), (3.4, 8.6)]),
    # 2D active nematic around a disc with homeotropic strong anchoring
    case("lc2d_disc", 2, (12, 10), 12, rtech=16, lc=1, steps=3, seed=37, dt=0.3,
         sp=species(mfpot=10.0, tumble=0.5, rfc=0.01, act=0.05),
         walls=periodic(2, (12, 10)) + [_disc((6.0, 5.0, 0.0), 2.6, MUN=1.0, MUT=0.0)]),
]

# rat2d_2sp uses two species
CASES[6]["species"] = [species(mass=1.0), species(mass=2.5)]
next(c for c in CASES if c["name"] == "at2d_nbounce")["species"] = [species(vdist=v) for v in (0, 2, 3, 4)]
for _n in ("srd2d_couette_znm", "at3d_channel_znm"):
    next(c for c in CASES if c["name"] == _n)["species"] = [species(mass=1.0), species(mass=1.7)]


def wall_w(w: dict, Q: np.ndarray, dim: int) -> np.ndarray:
    """calcW (bc.c:55-150) of rows of Q for the wall dict w (four-fold symmetric
    surfaces; the planar and the ellipsoidal wavy forms), the wall at its Q"""
    A = np.asarray(w["A"], float)
    q0 = np.asarray(w.get("Q", (0, 0, 0)), float)
    P = list(w.get("P", (1, 1, 1, 1)))
    B = list(w.get("B", (0, 0, 0)))
    W = np.zeros(Q.shape[0])
    for i in range(dim):
        t = A[i] * (Q[:, i] - q0[i])
        if w.get("ABS", 0):
            t = np.abs(t)
        W += t ** P[i]
    r = abs(w.get("R", 0.0)) if w.get("ABS", 0) else w.get("R", 0.0)
    W -= r ** P[3]
    if B[0] != 0.0:
        if all(abs(x - 1.0) < 1e-8 for x in P[:dim]):
            W1 = (A[1] * Q[:, 0] - A[0] * Q[:, 1]) / np.sqrt(A[0] ** 2 + A[1] ** 2)
        else:
            W1 = np.arctan2(A[1] * (Q[:, 1] - q0[1]), A[0] * (Q[:, 0] - q0[0]))
        W += B[0] * np.cos(B[1] * W1)
    if w.get("INV", 0):
        W = -W
    return W


def rest_frame_due(c: dict, k: int) -> bool:
    """whether the output step after step k (runtime k) applies the rest-frame transformation"""
    z = c.get("zeronetmom", 0)
    return z > 0 and k % z == 0


def initial_state(c: dict):
    """uniform positions, Maxwell-Boltzmann velocities, random orientations"""
    rng = np.random.default_rng(1000 + c["seed"])
    dim = c["dim"]
    L = c["L"]
    vol = int(np.prod(L[:dim]))
    N = vol * c["ppc"]
    Q = np.zeros((N, 3))
    V = np.zeros((N, 3))
    U = np.zeros((N, 3))
    nsp = len(c["species"])
    spid = (np.arange(N) % nsp).astype(np.int32)
    qbox = c.get("qbox") or [(0.0, L[d]) for d in range(dim)]
    for d in range(dim):
        Q[:, d] = rng.uniform(qbox[d][0], qbox[d][1], N)
    # curved and wavy walls: no particle starts inside one (replace, init.c:1240-1243, with
    # replacePos_WithCheck's test W <= TOL)
    curved = [w for w in c["walls"] if "P" in w or "B" in w]
    for _ in range(1000):
        bad = np.zeros(N, bool)
        for w in curved:
            bad |= wall_w(w, Q, dim) <= 0.05
        if not bad.any():
            break
        for d in range(dim):
            Q[bad, d] = rng.uniform(qbox[d][0], qbox[d][1], int(bad.sum()))
    masses = np.array([s["mass"] for s in c["species"]])[spid]
    V[:, :dim] = rng.normal(size=(N, dim)) * np.sqrt(c["kbt"] / masses)[:, None]
    if dim == 3:
        u = rng.normal(size=(N, 3))
        U = u / np.linalg.norm(u, axis=1)[:, None]
    else:
        th = rng.uniform(0, 2 * np.pi, N)
        U[:, 0] = np.cos(th)
        U[:, 1] = np.sin(th)
    avv = np.zeros(3)
    avv[:dim] = V[:, :dim].mean(axis=0)
    if "avv" in c:
        avv = np.array(c["avv"], dtype=float)
    for i, v in c.get("kick", []):
        V[i, :dim] = v[:dim]
    return Q, V, U, spid, avv


def oracle_cfg(c: dict, gndnst: float, avv) -> dict:
    return dict(dim=c["dim"], L=c["L"], dt=c["dt"], kbt=c["kbt"], rotang=c["rotang"], tau=c["tau"],
                grav=c["grav"], avv=list(avv), gndnst=gndnst, rtech=c["rtech"], lc=c["lc"], tstech=c["tstech"],
                galinv=c["galinv"], species=c["species"], walls=c["walls"])
