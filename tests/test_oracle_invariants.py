"""Pins of the oracle that the recorded fixtures leave open, by what the
mathematics fixes (CPU only).

  * Philox-4x32-10 against the published Random123 known-answer vectors;
  * the Q-tensor eigen-solver (mtools.c:1230-1476) against numpy.linalg.eigh;
  * the Rodrigues rotation (mtools.c:1572-1584) against the matrix exponential;
  * SRD ARBAXIS: |v - VCM| per particle and the total kinetic energy and
    momentum are conserved; ORTHAXIS 3D keeps momentum but not energy (the
    reference's R[2][1] = -.1 S, mpc.c:1215);
  * MPC-AT: momentum conserved; MPC-AT+a additionally conserves every cell's
    angular momentum about its centre of mass when POP >= 3 (mpc.c:1540);
  * bin() is the stable partition of the previous list order by
    (new cell, arrived) (mpc.c:948-992, addlink 1069-1087);
  * escape repair (mpc.c:970-982): particles that leave an unwalled axis are
    re-placed inside the box and counted.
"""
import numpy as np
import pytest
import scipy.linalg

import cases as CS
import pyoracle as P


def make(name, dim, L, ppc, rtech, seed=3, galinv=1, walls=None, sp=None, rotang=2.2689280275926285, grav=(0, 0, 0)):
    c = CS.case(name, dim, L, ppc, rtech, steps=1, seed=seed, rotang=rotang, walls=walls, sp=sp, galinv=galinv,
                grav=grav)
    Q, V, U, spid, avv = CS.initial_state(c)
    cfg = CS.oracle_cfg(c, float(ppc), avv)
    return P.System(cfg, Q, V, U, spid), c


def cells_of(order, cstart):
    cell = np.empty(len(order), np.int64)
    for k in range(len(cstart) - 1):
        cell[order[cstart[k]:cstart[k + 1]]] = k
    return cell


# --- Philox-4x32-10 known answers (Random123, kat_vectors) -----------------------------------------------
@pytest.mark.parametrize("ctr,key,out", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, out):
    assert P.philox(ctr, key) == out


# --- eigen-solver ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_eigen_matches_numpy(dim):
    rng = np.random.default_rng(11 + dim)
    for _ in range(200):
        a = rng.normal(size=(dim, dim))
        m = 0.5 * (a + a.T)
        m -= np.trace(m) / dim * np.eye(dim)  # traceless, like the Q tensor
        ev, vec = P.eigen(m)
        ref = np.linalg.eigvalsh(m)
        assert ev[0] == pytest.approx(ref[-1], rel=1e-10, abs=1e-12)
        if dim == 3:
            assert ev[0] >= ev[1] - 1e-12 and ev[1] >= ev[2] - 1e-12
            assert sorted(ev) == pytest.approx(list(ref), rel=1e-9, abs=1e-11)
        for k in range(dim):
            v = vec[k]
            assert np.linalg.norm(v) == pytest.approx(1.0, rel=1e-12)
            assert np.linalg.norm(m @ v - ev[k] * v) < 1e-8


# --- Rodrigues rotation ------------------------------------------------------------------------------------
def test_rodrigues_is_the_matrix_exponential():
    rng = np.random.default_rng(5)
    for _ in range(100):
        v = rng.normal(size=3)
        ax = rng.normal(size=3)
        th = rng.uniform(-np.pi, np.pi)
        n = ax / np.linalg.norm(ax)
        K = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
        ref = scipy.linalg.expm(th * K) @ v
        got = P.rodrigues(v, ax, th)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


# --- collision invariants ----------------------------------------------------------------------------------
def momentum_energy(s):
    Q, V, U = s.state()
    return V.sum(axis=0), 0.5 * (V * V).sum()


@pytest.mark.parametrize("rtech", [0, 2, 3])
def test_momentum_conserved(rtech):
    s, _ = make("inv", 3, (6, 6, 6), 8, rtech)
    d = P.Draws(P.DRAW_PHILOX, 21)
    p0, _ = momentum_energy(s)
    for _ in range(4):
        s.step(d)
    p1, _ = momentum_energy(s)
    assert np.allclose(p1, p0, atol=1e-10)


def test_srd_arbaxis_conserves_energy_and_relative_speed():
    s, c = make("inv", 3, (6, 6, 6), 8, 0, galinv=0)
    d = P.Draws(P.DRAW_PHILOX, 22)
    s.phase("bin", d)
    Q, V0, _ = s.state()
    order, cstart = s.order()
    cell = cells_of(order, cstart)
    nc = len(cstart) - 1
    pop = np.bincount(cell, minlength=nc)
    vcm = np.zeros((nc, 3))
    np.add.at(vcm, cell, V0)
    vcm[pop > 0] /= pop[pop > 0, None]
    s.phase("collide", d)
    _, V1, _ = s.state()
    many = pop[cell] > 1
    r0 = np.linalg.norm(V0 - vcm[cell], axis=1)
    r1 = np.linalg.norm(V1 - vcm[cell], axis=1)
    assert np.allclose(r1[many], r0[many], rtol=1e-12, atol=1e-12)
    assert 0.5 * (V1 * V1).sum() == pytest.approx(0.5 * (V0 * V0).sum(), rel=1e-12)
    assert not np.allclose(V1[many], V0[many])


def test_srd_orthaxis_3d_keeps_momentum_not_energy():
    s, _ = make("inv", 3, (6, 6, 6), 8, 1, galinv=0)
    d = P.Draws(P.DRAW_PHILOX, 23)
    s.phase("bin", d)
    _, V0, _ = s.state()
    s.phase("collide", d)
    _, V1, _ = s.state()
    assert np.allclose(V1.sum(axis=0), V0.sum(axis=0), atol=1e-10)
    e0, e1 = (V0 * V0).sum(), (V1 * V1).sum()
    assert abs(e1 - e0) > 1e-6 * e0  # the x-axis matrix is not a rotation


@pytest.mark.parametrize("rtech,conserves", [(3, True), (2, False)])
def test_cell_angular_momentum(rtech, conserves):
    s, _ = make("inv", 3, (5, 5, 5), 8, rtech, galinv=0)
    d = P.Draws(P.DRAW_PHILOX, 24)
    s.phase("bin", d)
    Q, V0, _ = s.state()
    order, cstart = s.order()
    s.phase("collide", d)
    _, V1, _ = s.state()
    worst, moved = 0.0, 0.0
    for k in range(len(cstart) - 1):
        idx = order[cstart[k]:cstart[k + 1]]
        if len(idx) < 3:
            continue
        cm = Q[idx].mean(axis=0)
        r = Q[idx] - cm
        L0 = np.cross(r, V0[idx]).sum(axis=0)
        L1 = np.cross(r, V1[idx]).sum(axis=0)
        scale = np.abs(np.cross(r, V0[idx])).sum() + 1e-300
        worst = max(worst, np.abs(L1 - L0).max() / scale)
        moved = max(moved, np.abs(L1 - L0).max() / scale)
    if conserves:
        assert worst < 1e-12
    else:
        assert moved > 1e-3


# --- bin() order rule -------------------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_bin_is_a_stable_partition(dim):
    L = (8, 8) if dim == 2 else (5, 5, 5)
    s, _ = make("inv", dim, L, 6, 0)
    d = P.Draws(P.DRAW_PHILOX, 25)
    for _ in range(2):
        s.step(d)
    order0, cstart0 = s.order()
    cell0 = cells_of(order0, cstart0)
    s.phase("bin", d)
    order1, cstart1 = s.order()
    cell1 = cells_of(order1, cstart1)
    assert (cell1 != cell0).any()
    arrived = (cell1[order0] != cell0[order0]).astype(np.int64)
    expect = order0[np.lexsort((arrived, cell1[order0]))]  # stable: ties keep the previous order
    assert np.array_equal(order1, expect)
    assert cstart1[-1] == len(order0)


# --- escape repair ----------------------------------------------------------------------------------------
def test_escape_repair_replaces_and_counts():
    L = (6, 6, 6)
    walls = CS.pbc_walls(0, L[0]) + CS.pbc_walls(1, L[1])  # nothing on z
    s, _ = make("inv", 3, L, 6, 0, walls=walls)
    d = P.Draws(P.DRAW_PHILOX, 26)
    for _ in range(6):
        s.step(d)
    esc = s.counters()[0]
    Q, V, _ = s.state()
    assert esc > 0
    # every particle truncates into a cell 0..L (the cell index is (int)Q,
    # so (-1, 0) is cell 0, mpc.c:958-960)
    c = np.trunc(Q[:, 2]).astype(int)
    assert ((c >= 0) & (c <= L[2])).all()
    assert (Q[:, 2] < -0.5).any() or esc > 0
    assert np.isfinite(V).all()
