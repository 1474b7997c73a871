"""CPU: the oracle's restatement of the exact-sum FP32 form (oracle/exact_step.c).

What pins it:
  * order independence, the property the form is built on: the particles
    shuffled, stepped and put back give the same bits (positions,
    velocities, KBTNOW) for SRD, MPC-AT and MPC-AT+a with two species;
  * one SRD step agrees with the FP64 oracle's step (pinned to the
    reference's own fixtures) from the same state with the same counter-based
    draws, to FP32 precision;
  * SRD conserves the momentum of every cell and its kinetic energy to FP32
    rounding; MPC-AT+a conserves momentum;
  * with the rotation angle 0 an SRD step is a pure drift with the periodic
    wrap, computed independently here in float32 numpy;
  * the SRD axis's sine and cosine of 2 pi u (octant reduction and Taylor
    polynomials, no library sine) agree with numpy's to 5e-14 at random
    points, every octant boundary and the ends of [0, 1), and the axis agrees
    with the FP64 oracle's library-sine axis from the same draws.
"""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as P  # noqa: E402


def box_walls(L):
    ws = []
    for ax in range(len(L)):
        a1 = [0.0, 0.0, 0.0]
        a1[ax] = 1.0
        a2 = [0.0, 0.0, 0.0]
        a2[ax] = -1.0
        ws += [dict(A=a1, R=0.0, DN=float(L[ax])), dict(A=a2, R=-float(L[ax]), DN=float(L[ax]))]
    return ws


def cfg(dim=3, L=(6, 5, 4), rtech=0, nspec=1, grav=(0.0, 0.0, 0.0), rotang=2.268928, dt=0.1):
    sp = [dict(mass=1.0 + 0.5 * k, rfc=0.01, tumble=2.0, chihi=0.5, act=0.0, mfpot=10.0, damp=0.0, sigwidth=1.0,
               sigpos=1.0, minactratio=0.0, vdist=4) for k in range(nspec)]
    L = list(L[:dim]) + [1] * (3 - dim)
    return dict(dim=dim, L=L, dt=dt, kbt=1.0, rotang=rotang, tau=0.5, grav=list(grav), avv=[0.0, 0.0, 0.0],
                gndnst=10.0, rtech=rtech, lc=0, tstech=0, galinv=1, species=sp, walls=box_walls(L[:dim]))


def state(c, ppc=10, seed=3):
    rng = np.random.default_rng(seed)
    dim = c["dim"]
    n = int(np.prod(c["L"][:dim])) * ppc
    x = np.zeros((n, 3), np.float32)
    v = np.zeros((n, 3), np.float32)
    for d in range(dim):
        x[:, d] = rng.uniform(0.0, c["L"][d], n).astype(np.float32)
        v[:, d] = rng.normal(0.0, 1.0, n).astype(np.float32)
    sp = rng.integers(0, len(c["species"]), n).astype(np.int32)
    return x, v, sp


CASES = {
    "srd3d_arb": dict(),
    "srd3d_orth": dict(rtech=1),
    "srd2d": dict(dim=2, L=(9, 7)),
    "at3d_grav": dict(rtech=2, grav=(1e-3, 0.0, 0.0)),
    "rat3d_2sp": dict(rtech=3, nspec=2),
    "rat2d": dict(dim=2, L=(9, 7), rtech=3),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_exact_step_is_order_independent(name):
    c = cfg(**CASES[name])
    x, v, sp = state(c)
    ids = np.arange(x.shape[0], dtype=np.int32)
    x1, v1, k1 = P.exact_step(c, x, v, sp, seed=11, step=5, ids=ids)
    perm = np.random.default_rng(1).permutation(x.shape[0])
    x2, v2, k2 = P.exact_step(c, x[perm], v[perm], sp[perm], seed=11, step=5, ids=ids[perm])
    inv = np.argsort(perm)
    assert np.array_equal(x2[inv], x1) and np.array_equal(v2[inv], v1), name
    assert k1 == k2
    assert abs(k1 - 1.0) < 0.2


def test_srd_step_matches_fp64_oracle():
    """One ARBAXIS SRD step: the exact FP32 form vs the FP64 oracle's Philox
    mode (same shift and cell axes), from the same FP32-representable state"""
    c = cfg(L=(8, 7, 6))
    x, v, _ = state(c, seed=5)
    n = x.shape[0]
    sysd = P.System(c, x.astype(np.float64), v.astype(np.float64), np.zeros((n, 3)), np.zeros(n, np.int32))
    sysd.step(P.Draws(P.DRAW_PHILOX, 21, 0))
    Q, V, _ = sysd.state()
    x1, v1, kbt = P.exact_step(c, x, v, seed=21, step=0)
    dq = np.abs(x1 - Q)
    dq = np.minimum(dq, np.abs(dq - np.array(c["L"], np.float32)))  # the same point across the periodic wrap
    dv = np.abs(v1 - V)
    # a particle within FP32 rounding of a cell face can bin differently
    ok = (dq.max(axis=1) < 2e-5) & (dv.max(axis=1) < 2e-5)
    assert ok.mean() > 0.995, (ok.mean(), dq.max(), dv.max())
    assert abs(kbt - sysd.kbtnow()) < 1e-5 * max(1.0, kbt)


def test_srd_conserves_cell_momentum_and_energy():
    c = cfg(L=(6, 6, 6), rotang=1.3)
    x, v, _ = state(c, seed=8)
    x1, v1, _ = P.exact_step(c, x, v, seed=2, step=1)
    # total momentum and kinetic energy (SRD rotates relative velocities)
    p0, p1 = v.astype(np.float64).sum(axis=0), v1.astype(np.float64).sum(axis=0)
    e0, e1 = (v.astype(np.float64) ** 2).sum(), (v1.astype(np.float64) ** 2).sum()
    assert np.abs(p1 - p0).max() < 1e-6 * x.shape[0] ** 0.5 * 10
    assert abs(e1 - e0) / e0 < 1e-5
    # the velocities changed (collisions happened)
    assert np.abs(v1 - v).max() > 0.1


def test_rat_conserves_momentum():
    c = cfg(rtech=3, nspec=2)
    x, v, sp = state(c, seed=9)
    m = np.array([s["mass"] for s in c["species"]])[sp][:, None]
    x1, v1, _ = P.exact_step(c, x, v, sp, seed=4, step=2)
    p0 = (m * v.astype(np.float64)).sum(axis=0)
    p1 = (m * v1.astype(np.float64)).sum(axis=0)
    assert np.abs(p1 - p0).max() < 1e-4


def test_zero_angle_is_a_pure_drift():
    """rotation angle 0 (cos 1, sin 0): SRD leaves v to rounding, and the
    positions are the grid shift, the shift back and the drift with the
    periodic wrap, in float32"""
    c = cfg(dim=2, L=(9, 7), rotang=0.0)
    x, v, _ = state(c, seed=12)
    x1, v1, _ = P.exact_step(c, x, v, seed=6, step=3)
    s = np.array([P.lib().orc_philox_uniform(6, 0, 3, P.P_SHIFT, 0, j) for j in range(2)]).astype(np.float32)
    L = np.array(c["L"][:2], np.float32)
    q = x[:, :2].copy()
    t = q + s
    q = np.where((t > L) & (t <= L + s), t - L, t).astype(np.float32)
    t = q + (-s)
    q = np.where((t < 0) & (t >= -s), t + L, t).astype(np.float32)
    t = (q + np.float32(0.1) * v1[:, :2]).astype(np.float32)
    q = np.where(t < 0, t + L, np.where(t >= L, t - L, t)).astype(np.float32)
    assert np.abs(v1 - v).max() <= 4e-7 * max(1.0, float(np.abs(v).max()))
    assert np.array_equal(x1[:, :2], q)
    assert math.isfinite(float(x1.sum()))


def _sincos_turn(u):
    sn, cs = C.c_double(0.0), C.c_double(0.0)
    P.lib().orc_sincos_turn(float(u), C.byref(sn), C.byref(cs))
    return sn.value, cs.value


def test_sincos_turn_matches_library_sine():
    rng = np.random.default_rng(17)
    ulp = 2.0 ** -53
    us = list(rng.integers(0, 2 ** 53, 4000, dtype=np.int64).astype(np.float64) * ulp)
    for o in range(8):  # both sides of every octant boundary
        for d in (-2, -1, 0, 1, 2):
            n = o * 2 ** 50 + d
            if 0 <= n < 2 ** 53:
                us.append(n * ulp)
    us.append(1.0 - ulp)
    worst = 0.0
    for u in us:
        sn, cs = _sincos_turn(u)
        worst = max(worst, abs(sn - math.sin(2.0 * math.pi * u)), abs(cs - math.cos(2.0 * math.pi * u)))
        assert abs(sn * sn + cs * cs - 1.0) < 1e-13, u
    assert worst < 5e-14, worst
    # exact special values: the octant reduction makes these exact
    assert _sincos_turn(0.0) == (0.0, 1.0)
    assert _sincos_turn(0.25) == (1.0, 0.0)
    assert _sincos_turn(0.5) == (0.0, -1.0)
    assert _sincos_turn(0.75) == (-1.0, 0.0)


def test_unitvec_turn_matches_library_axis():
    a = (C.c_double * 3)()
    b = (C.c_double * 3)()
    for key in range(500):
        P.lib().orc_philox_unitvec(9, 1, 4, P.P_CELL, key, 0, a)
        P.lib().orc_philox_unitvec_turn(9, 1, 4, P.P_CELL, key, 0, b)
        assert max(abs(a[d] - b[d]) for d in range(3)) < 1e-13, key
        assert a[2] == b[2]  # z is drawn the same way
        assert abs(b[0] ** 2 + b[1] ** 2 + b[2] ** 2 - 1.0) < 1e-13


def test_log_unit_matches_library_log():
    L = P.lib()
    rng = np.random.default_rng(23)
    ulp = 2.0 ** -53
    us = list(1.0 - rng.integers(0, 2 ** 53, 4000, dtype=np.int64).astype(np.float64) * ulp)
    us += [1.0, 1.0 - ulp, ulp, 2.0 * ulp, 0.5, 0.25, math.sqrt(0.5), np.nextafter(math.sqrt(0.5), 0.0),
           np.nextafter(math.sqrt(0.5), 1.0)]
    for u in us:
        got, ref = L.orc_log_unit(float(u)), math.log(u)
        assert abs(got - ref) <= 4e-16 * max(1.0, abs(ref)), (u, got, ref)
    assert L.orc_log_unit(1.0) == 0.0
    assert L.orc_log_unit(0.5) == -0.69314718055994530942


def test_normal_turn_matches_library_normal():
    L = P.lib()
    for key in range(300):
        for k in range(3):
            a = L.orc_philox_normal(5, 0, 7, P.P_NORMAL, key, k)
            b = L.orc_philox_normal_turn(5, 0, 7, P.P_NORMAL, key, k)
            assert abs(a - b) < 1e-13 * max(1.0, abs(a)), (key, k, a, b)


def test_body_force_sets_the_momentum():
    """in a periodic box the collisions conserve momentum, so after k steps
    the body force has added N m g dt k, up to the unbiased FP32 rounding of
    the collisions (a random walk: about 5e-4 after 240 steps here, the same
    along y, where no force acts).  The kick enters once per cell in FP64;
    added per particle in FP32 (a fixed small increment) its rounding would
    depend only on the binade of v and fall 1.1e-2 short here."""
    c = cfg(dim=2, L=(30, 20), grav=(0.0005, 0.0, 0.0))
    x, v, _ = state(c, ppc=10, seed=3)
    p0 = v.astype(np.float64).sum(axis=0)
    ids = np.arange(x.shape[0], dtype=np.int32)
    k = 240
    for s in range(k):
        x, v, _ = P.exact_step(c, x, v, seed=5, step=s, ids=ids)
    dp = v.astype(np.float64).sum(axis=0) - p0
    expect = x.shape[0] * 0.0005 * 0.1 * k
    assert abs(dp[0] - expect) < 2.5e-3, (dp[0], expect)
    assert abs(dp[1]) < 2.5e-3
