"""GPU parity of the general wall engine (curved, inverted, superelliptic and
wavy walls; bc.c:55-150 calcW, 251-405 crossing times and periodic images,
1252-1556 chooseBC and the surface normals; ghost cells and strong anchoring
at curved walls, mpc.c:351-567).

The reference-recorded fixtures of these walls (circle2d_at, ellipse2d_pbc,
square2d_srd, pipe3d_rat, cyl3d_at, wavy2d_srd, lc2d_disc) are replayed by
test_gpu_parity.test_gpu_matches_reference through the whole-step entry;
here the same kinds of walls run at larger sizes with the library's own
generator against the oracle's Philox mode through the three-phase API, for
several steps: list order and populations bit-exact, FP64 r, v (and u)
within 1e-12 relative (plus 1e-12 absolute), and the event counters
(escapes, bounces past NBOUNCE, rewinds, re-thermalisations) equal to the
oracle's; the whole-step entry and mpcd_gpu_run give the same bits.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
import gpu_util as U  # noqa: E402
import cases as CS  # noqa: E402
from test_gpu_parity import _assert_close, assert_lc_close, assert_same_decisions  # noqa: E402


def _surface(w, Q, dim):
    """calcW of an integer-exponent, non-wavy surface, for placing the initial
    particles outside every wall (test set-up only)"""
    A = np.asarray(w["A"], float)
    q0 = np.asarray(w.get("Q", (0, 0, 0)), float)
    P = np.asarray(w.get("P", (1, 1, 1, 1)), float)
    W = np.zeros(Q.shape[0])
    for i in range(dim):
        W += (A[i] * (Q[:, i] - q0[i])) ** P[i]
    W -= w["R"] ** P[3]
    return -W if w.get("INV", 0) else W


def _box_outside(dim, L, ppc, walls, seed, wavy_band=None):
    rng = np.random.default_rng(seed)
    N = int(np.prod(L[:dim])) * ppc
    Q = np.zeros((N, 3))
    filled = 0
    while filled < N:
        cand = np.zeros((4 * N, 3))
        for d in range(dim):
            cand[:, d] = rng.uniform(0, L[d], 4 * N)
        ok = np.ones(4 * N, bool)
        for w in walls:
            if "B" in w:
                continue
            ok &= _surface(w, cand, dim) > 1e-6
        if wavy_band is not None:
            ax, lo, hi = wavy_band
            ok &= (cand[:, ax] > lo) & (cand[:, ax] < hi)
        cand = cand[ok][:N - filled]
        Q[filled:filled + cand.shape[0]] = cand
        filled += cand.shape[0]
    V = np.zeros((N, 3))
    V[:, :dim] = rng.normal(size=(N, dim))
    Uo = np.zeros((N, 3))
    if dim == 3:
        u = rng.normal(size=(N, 3))
        Uo = u / np.linalg.norm(u, axis=1)[:, None]
    else:
        th = rng.uniform(0, 2 * np.pi, N)
        Uo[:, 0], Uo[:, 1] = np.cos(th), np.sin(th)
    return Q, V, Uo


def _disc(Q, R, P=(2, 2, 2, 2), A=(1.0, 1.0, 0.0), **kw):
    return dict(A=list(A), Q=list(Q), R=R, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1, P=list(P), **kw)


def _wavy(y, sign, amp, k):
    return dict(A=[0, sign, 0], Q=[0, y, 0], R=0.0, DN=0.0, MVN=-1.0, MVT=-1.0, PHANTOM=1, P=[1, 1, 1, 1],
                B=[amp, k, 0.0])


# name: dim, L, ppc, rtech, lc, walls, dt, grav, extra
GEOMS = {
    # MPC-AT around a disc obstacle
    "disc2d_at": (2, (40, 30), 12, 2, 0, CS.periodic(2, (40, 30)) + [_disc((20.0, 15.0, 0.0), 6.0)], 0.3,
                  (2e-3, 0, 0), None),
    # SRD around an ellipse across the periodic x boundary (its image acts on the far side)
    "ellipse2d_pbc": (2, (32, 24), 10, 0, 0,
                      CS.periodic(2, (32, 24)) + [_disc((1.0, 12.0, 0.0), 1.0, A=(1 / 4.0, 1 / 3.0, 0.0))], 0.4,
                      (0, 0, 0), None),
    # SRD around a superellipse (P = 8: the secant crossing time)
    "square2d_srd": (2, (32, 24), 10, 1, 0,
                     CS.periodic(2, (32, 24)) + [_disc((16.0, 12.0, 0.0), 1.0, P=(8, 8, 1, 8),
                                                       A=(1 / 5.0, 1 / 5.0, 0.0))], 0.4, (0, 0, 0), None),
    # SRD between wavy walls inside the flat domain walls
    "wavy2d_srd": (2, (40, 16), 10, 0, 0,
                   CS.CS_pbc_x(40) + CS.solid_walls(1, 16) + [_wavy(3.0, 1, 0.8, 2 * np.pi / 10.0),
                                                               _wavy(13.0, -1, 0.8, 2 * np.pi / 10.0)], 0.4,
                   (1e-3, 0, 0), (1, 4.0, 12.0)),
    # MPC-AT+a in a pipe along x (inverted cylinder, P = 1 2 2 2: the secant crossing time)
    "pipe3d_rat": (3, (12, 12, 12), 8, 3, 0,
                   CS.periodic(3, (12, 12, 12)) + [_disc((0.0, 6.0, 6.0), 5.3, P=(1, 2, 2, 2), A=(0.0, 1.0, 1.0),
                                                        INV=1)], 0.2, (2e-3, 0, 0), None),
    # MPC-AT in a cylinder along z with P = 2 2 2 2 (the quadratic crossing time)
    "cyl3d_at": (3, (12, 12, 8), 8, 2, 0,
                 CS.periodic(3, (12, 12, 8)) + [_disc((6.0, 6.0, 0.0), 5.3, A=(1.0, 1.0, 0.0), INV=1)], 0.2,
                 (0, 0, 0), None),
    # an active nematic around a disc with homeotropic strong anchoring
    "lc2d_disc": (2, (24, 20), 12, 16, 1,
                  CS.periodic(2, (24, 20)) + [_disc((12.0, 10.0, 0.0), 4.3, MUN=1.0, MUT=0.0)], 0.2, (0, 0, 0),
                  None),
    # a passive nematic around a disc with planar strong anchoring, MPC-AT+a
    "lc2d_disc_planar": (2, (24, 20), 12, 3, 1,
                         CS.periodic(2, (24, 20)) + [_disc((12.0, 10.0, 0.0), 4.3, MUN=0.0, MUT=1.0)], 0.2,
                         (0, 0, 0), None),
}


def _setup(name, seed):
    import pyoracle
    M = U.package()
    dim, L, ppc, rtech, lc, walls, dt, grav, band = GEOMS[name]
    Q, V, Uo = _box_outside(dim, L, ppc, walls, seed, band)
    c = dict(name=name, dim=dim, L=list(L) + [1] * (3 - len(L)), rtech=rtech, lc=lc, tstech=0,
             rotang=2.2689280275926285, dt=dt, kbt=1.0, grav=list(grav), walls=walls,
             species=[CS.species(mfpot=10.0, tumble=0.5, act=0.05)], galinv=1, tau=0.5)
    gnd = float(ppc)
    osys = pyoracle.System(CS.oracle_cfg(c, gnd, np.zeros(3)), Q, V, Uo, None)
    ctx = M.Context(U.gpu_cfg(c, gnd, np.zeros(3)), Q.shape[0], seed=seed)
    ctx.upload(0, Q, V, Uo if lc else None, None)
    if lc:
        ctx.lc_record(True)
    return M, pyoracle, c, osys, ctx


def _compare(c, osys, ctx, what):
    Qo, Vo, Uor = osys.state()
    Qg, Vg, Ug = ctx.download(0)
    np.testing.assert_array_equal(ctx.order(0)[0], osys.order()[0], err_msg=f"{what}: list order")
    close = assert_lc_close if c["lc"] else _assert_close
    close(Qg, Qo, f"{what}: Q")
    close(Vg, Vo, f"{what}: V")
    if c["lc"]:
        close(Ug, Uor, f"{what}: U")


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GEOMS))
def test_walls_philox_vs_oracle(name):
    M, pyoracle, c, osys, ctx = _setup(name, 41)
    d = pyoracle.Draws(pyoracle.DRAW_PHILOX, 41, 0)
    print(f"[gpu walls] {name} N={osys.state()[0].shape[0]}", flush=True)
    for k in range(4):
        osys.step(d)
        ctx.bin(None, shifted=1)
        ctx.collide(None)
        events = ctx.stream()  # running counts since the context was made
        if c["lc"]:
            assert_same_decisions(ctx, osys, f"{name} step {k}")
        _compare(c, osys, ctx, f"{name} step {k}")
        assert tuple(events) == tuple(osys.counters()), (tuple(events), osys.counters())
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["disc2d_at", "pipe3d_rat", "lc2d_disc"])
def test_walls_phases_and_run(name):
    """the three-phase API and mpcd_gpu_run give the whole step's result bit for
    bit with curved walls (every call takes the unfused kernels)"""
    M, pyoracle, c, osys, a = _setup(name, 43)
    _, _, _, _, b = _setup(name, 43)
    _, _, _, _, r = _setup(name, 43)
    print(f"[gpu walls phases] {name}", flush=True)
    for k in range(3):
        a.step(None)
        b.bin(None, shifted=1)
        b.collide(None)
        b.stream()
    r.run(3)
    for x, y, z in zip(a.download(0), b.download(0), r.download(0)):
        np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(x, z)
    np.testing.assert_array_equal(a.order(0)[0], b.order(0)[0])
    for x in (a, b, r):
        x.close()


@pytest.mark.gpu
def test_walls_refuse_derivatives():
    """the tangent and the adjoint tape cover planar walls only"""
    M, pyoracle, c, osys, ctx = _setup("disc2d_at", 47)
    with pytest.raises(M.MPCDError):
        ctx.tangent(True, dgrav=(1.0, 0.0, 0.0))
    with pytest.raises(M.MPCDError):
        ctx.record(4)
    ctx.close()
