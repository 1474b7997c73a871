"""GPU: the exact-sum FP32 form (precision MPCD_FP32_EXACT) through the C-ABI.

  * parity: on boxes of periodic faces, every step of the bucketed kernels
    and of the dense kernels equals the oracle's restatement
    (oracle/exact_step.c) bit for bit -- positions, velocities and KBTNOW --
    for SRD (ARBAXIS, ORTHAXIS, 2D), MPC-AT with gravity and MPC-AT+a with two
    species, on a context of two replicas of different sizes;
  * layout independence: with walls (bounce-back phantom planes, a moving
    wall, an open axis that makes particles escape, curved walls), the
    bucketed run, the dense run, a run whose buckets overflow and hipGraph
    replays all hold the same bits after every compared step;
  * the entry points outside the form fail with MPCD_EUNSUPPORTED.
"""
import os
import sys

import numpy as np
import pytest

import cases as CS
import gpu_util as U

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
ROT3 = 2.2689280275926285


def _cases(name):
    if name == "srd3d_arb":
        return [CS.case("x", 3, (7, 6, 5), 9, 0, 1, 1, rotang=ROT3), CS.case("x", 3, (5, 8, 6), 7, 0, 1, 2,
                                                                            rotang=ROT3)]
    if name == "srd3d_orth":
        return [CS.case("x", 3, (6, 6, 7), 10, 1, 1, 3, rotang=ROT3)]
    if name == "srd2d":
        return [CS.case("x", 2, (21, 17), 12, 0, 1, 4), CS.case("x", 2, (9, 13), 8, 0, 1, 5)]
    if name == "at3d_grav":
        return [CS.case("x", 3, (8, 7, 6), 8, 2, 1, 6, grav=(2e-3, 0, 1e-3), sp=CS.species(damp=0.1))]
    if name == "rat3d_2sp":
        out = []
        for L, seed in (((7, 7, 6), 7), ((5, 6, 8), 8)):
            c = CS.case("x", 3, L, 9, 3, 1, seed, dt=0.05)
            c["species"] = [CS.species(mass=1.0), CS.species(mass=2.5)]
            out.append(c)
        return out
    if name == "rat2d":
        return [CS.case("x", 2, (15, 11), 9, 3, 1, 9)]
    raise ValueError(name)


def _ctx(M, cases, seed, **kw):
    cfgs, ns, states = [], [], []
    for c in cases:
        cfgs.append(U.gpu_cfg(c, float(c["ppc"]), [0.0, 0.0, 0.0]))
        ns.append(int(np.prod(c["L"][:c["dim"]])) * c["ppc"])
    ctx = M.Context(cfgs, ns, seed=seed, precision=M.FP32_EXACT)
    for r, c in enumerate(cases):
        Q, V, _, spid, _ = CS.initial_state(c)
        Q = Q.astype(np.float32).astype(np.float64)
        V = V.astype(np.float32).astype(np.float64)
        ctx.upload(r, Q, V, spid=spid if len(c["species"]) > 1 else None)
        states.append((Q.astype(np.float32), V.astype(np.float32), spid))
    return ctx, cfgs, states


@pytest.mark.parametrize("layout", ["bucketed", "dense"])
@pytest.mark.parametrize("name", ["srd3d_arb", "srd3d_orth", "srd2d", "at3d_grav", "rat3d_2sp", "rat2d"])
def test_exact_matches_oracle(name, layout):
    import pyoracle

    M = U.package()
    cases = _cases(name)
    seed = 314
    print(f"[exact] {name} {layout}", flush=True)
    ctx, cfgs, states = _ctx(M, cases, seed)
    ctx.set_layout(layout == "bucketed")
    offs = np.cumsum([0] + [ctx.ncells[r] for r in range(len(cases))])
    xs = [s[0].copy() for s in states]
    vs = [s[1].copy() for s in states]
    step = 0
    for chunk in (1, 1, 3, 2):
        ctx.run(chunk)
        kb = ctx.kbtnow()
        for _ in range(chunk):
            for r, c in enumerate(cases):
                sp = states[r][2] if len(c["species"]) > 1 else None
                xs[r], vs[r], kbt = pyoracle.exact_step(cfgs[r], xs[r], vs[r], sp, seed=seed, replica=r, step=step,
                                                        cell_off=int(offs[r]))
            step += 1
        for r in range(len(cases)):
            Q, V, _ = ctx.download(r)
            dq = np.abs(Q - xs[r]).max()
            dv = np.abs(V - vs[r]).max()
            assert np.array_equal(Q, xs[r].astype(np.float64)), f"{name} {layout} step {step} replica {r}: |dQ| {dq}"
            assert np.array_equal(V, vs[r].astype(np.float64)), f"{name} {layout} step {step} replica {r}: |dV| {dv}"
        # KBTNOW of the last step, replica by replica
        assert kb[len(cases) - 1] == kbt, f"{name} {layout} step {step}: KBTNOW {kb} vs {kbt}"
    if layout == "bucketed":
        st = ctx.layout_stats()
        assert st[0] == step and st[1] == 0, st
    ctx.close()


def _wall_cases(kind):
    if kind == "channel_rat":
        walls = CS.pbc_walls(0, 9) + CS.pbc_walls(1, 7) + CS.solid_walls(2, 8)
        c = CS.case("w", 3, (9, 7, 8), 9, 3, 1, 21, walls=walls, grav=(2e-3, 0, 0), dt=0.05)
        c["species"] = [CS.species(mass=1.0), CS.species(mass=1.5)]
        return [c, dict(c, seed=22)]
    if kind == "couette_srd2d":
        walls = CS.pbc_walls(0, 16) + CS.solid_walls(1, 12, PHANTOM=0)
        walls[3]["V"] = [0.3, 0.0, 0.0]
        walls[3]["DVxyz"] = [0.0, 0.0, 0.0]
        return [CS.case("w", 2, (16, 12), 10, 0, 1, 23, walls=walls)]
    if kind == "open_at":
        # no walls on y: particles leave the box and are re-placed
        walls = CS.pbc_walls(0, 8) + CS.pbc_walls(2, 7)
        return [CS.case("w", 3, (8, 6, 7), 8, 2, 1, 24, walls=walls)]
    if kind == "cylinder_srd":
        walls = CS.pbc_walls(0, 10) + CS.pbc_walls(1, 10) + CS.pbc_walls(2, 6)
        walls.append(dict(A=[1.0, 1.0, 0.0], Q=[5.0, 5.0, 0.0], R=2.0, P=[2, 2, 1, 2], DN=0.0, MVN=-1.0, MVT=-1.0,
                          INV=0))
        return [CS.case("w", 3, (10, 10, 6), 8, 0, 1, 25, rotang=ROT3, walls=walls, qbox=None)]
    raise ValueError(kind)


def _state(ctx, n):
    return [ctx.download(r)[:2] for r in range(n)]


@pytest.mark.parametrize("kind", ["channel_rat", "couette_srd2d", "open_at", "cylinder_srd"])
def test_exact_layouts_agree(kind):
    M = U.package()
    cases = _wall_cases(kind)
    print(f"[exact] layouts {kind}", flush=True)
    runs = {}
    for mode in ("dense", "bucketed", "overflow", "graph"):
        ctx, _, _ = _ctx(M, cases, 77)
        if kind == "cylinder_srd":
            # start outside the cylinder
            Q, V, _ = ctx.download(0)
            rq = np.hypot(Q[:, 0] - 5.0, Q[:, 1] - 5.0)
            Q[rq < 2.05, 0] = np.mod(Q[rq < 2.05, 0] + 5.0, 10.0)
            Q = Q.astype(np.float32).astype(np.float64)
            ctx.upload(0, Q, V)
        ctx.set_layout(mode != "dense", capacity=(40 if mode == "overflow" else 0))
        snaps = []
        for chunk in (2, 5, 8):
            ctx.run(chunk, use_graph=(mode == "graph"))
            snaps.append((_state(ctx, len(cases)), ctx.kbtnow().copy()))
        runs[mode] = (snaps, ctx.layout_stats())
        ctx.close()
    base = runs["dense"][0]
    for mode in ("bucketed", "overflow", "graph"):
        for k, (a, b) in enumerate(zip(base, runs[mode][0])):
            for r in range(len(cases)):
                assert np.array_equal(a[0][r][0], b[0][r][0]), f"{kind} {mode} snapshot {k} replica {r}: Q"
                assert np.array_equal(a[0][r][1], b[0][r][1]), f"{kind} {mode} snapshot {k} replica {r}: V"
            assert np.array_equal(a[1], b[1]), f"{kind} {mode} snapshot {k}: KBTNOW"
    assert runs["overflow"][1][1] > 0, "the lowered capacity did not overflow"
    assert runs["bucketed"][1][1] == 0 and runs["bucketed"][1][0] > 0


def test_exact_refuses_other_entry_points():
    M = U.package()
    cases = _cases("srd3d_orth")
    ctx, _, _ = _ctx(M, cases, 5)
    for call in (lambda: ctx.bin(), lambda: ctx.collide(), lambda: ctx.stream(), lambda: ctx.flow(True)):
        with pytest.raises(RuntimeError, match="MPCD_EUNSUPPORTED"):
            call()
    ctx.step()
    ctx.close()
    c = CS.case("x", 3, (6, 6, 6), 10, 0, 1, 3, tstech=1)
    with pytest.raises(RuntimeError):
        M.Context([U.gpu_cfg(c, 10.0, [0, 0, 0])], [2160], seed=1, precision=M.FP32_EXACT)
