"""GPU: the bucketed layout of mpcd_gpu_run gives the cell-sorted layout's
results bit for bit -- every particle's position and velocity, the list order
of every cell, and the state left for the phase-by-phase entry points (whose
binning reads the cell of every slot).  A lowered bucket capacity forces
overflows: the run then finishes on the cell-sorted path from the last
complete step, and must still agree."""
import numpy as np
import pytest

import cases as CS
import gpu_util as U

pytestmark = pytest.mark.gpu

ROT3 = 2.2689280275926285


def kind_cfgs(kind):
    """list of (case, ppc) replicas"""
    if kind == "srd3d":
        return [CS.case("b", 3, (8, 8, 8), 8, 0, 1, 1, rotang=ROT3), CS.case("b", 3, (6, 10, 7), 8, 0, 1, 1,
                                                                                  rotang=ROT3)]
    if kind == "srd3d_orth":
        return [CS.case("b", 3, (7, 6, 9), 10, 1, 1, 1, rotang=ROT3)]
    if kind == "srd2d":
        return [CS.case("b", 2, (24, 20), 12, 1, 1, 1), CS.case("b", 2, (9, 31), 7, 1, 1, 1)]
    if kind == "at_walls_grav":
        walls = CS.pbc_walls(0, 8) + CS.pbc_walls(1, 8) + CS.solid_walls(2, 8)
        return [CS.case("b", 3, (8, 8, 8), 8, 2, 1, 1, walls=walls, grav=(1e-3, 0, 0), sp=CS.species(damp=0.1)),
                CS.case("b", 3, (8, 8, 8), 6, 2, 1, 1, walls=walls, grav=(1e-3, 0, 0), sp=CS.species(damp=0.1))]
    if kind == "rat_channel":
        walls = CS.pbc_walls(0, 9) + CS.pbc_walls(1, 7) + CS.solid_walls(2, 8)
        return [CS.case("b", 3, (9, 7, 8), 9, 3, 1, 1, walls=walls, grav=(2e-3, 0, 0), dt=0.05)]
    if kind == "rat2d_2sp":
        c = CS.case("b", 2, (15, 11), 9, 3, 1, 1)
        c["species"] = [CS.species(mass=1.0), CS.species(mass=2.5)]
        return [c]
    raise ValueError(kind)


def make_ctx(M, cases, prec, seed=91):
    cfgs, ns = [], []
    for c in cases:
        cfgs.append(U.gpu_cfg(c, float(c["ppc"]), [0.0, 0.0, 0.0]))
        ns.append(int(np.prod(c["L"][:c["dim"]])) * c["ppc"])
    ctx = M.Context(cfgs, ns, seed=seed, precision=prec)
    for r, c in enumerate(cases):
        cc = dict(c)
        cc["seed"] = 40 + r
        Q, V, _, spid, _ = CS.initial_state(cc)
        ctx.upload(r, Q, V, spid=spid if len(c["species"]) > 1 else None)
    return ctx


def snapshot(ctx, nrep):
    out = []
    for r in range(nrep):
        Q, V, _ = ctx.download(r)
        order, cstart = ctx.order(r)
        out.append((Q.copy(), V.copy(), order.copy(), cstart.copy()))
    return out


def assert_same(a, b, what):
    for r, (x, y) in enumerate(zip(a, b)):
        for k, name in enumerate(("Q", "V", "order", "cstart")):
            assert np.array_equal(x[k], y[k]), f"{what}: replica {r} {name} differs"


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["srd3d", "srd3d_orth", "srd2d", "at_walls_grav", "rat_channel", "rat2d_2sp"])
def test_bucketed_equals_cell_sorted(prec, kind):
    M = U.package()
    P = M.FP64 if prec == "fp64" else M.FP32
    cases = kind_cfgs(kind)
    print(f"[bucket] {kind} {prec}", flush=True)
    a = make_ctx(M, cases, P)
    b = make_ctx(M, cases, P)
    a.set_layout(False)
    for n, g in ((7, False), (6, True), (1, True), (5, False)):
        a.run(n, use_graph=g)
        b.run(n, use_graph=g)
        assert_same(snapshot(a, len(cases)), snapshot(b, len(cases)), f"run({n}, graph={g})")
    st = b.layout_stats()
    assert st[0] >= 12 and st[1] == 0, st  # the first step of the first run builds the bin
    # the phase-by-phase path after a bucketed run (binning reads the slot cells)
    for ctx in (a, b):
        ctx.bin()
        ctx.collide()
        ctx.stream()
    assert_same(snapshot(a, len(cases)), snapshot(b, len(cases)), "phases after the runs")
    a.run(4)
    b.run(4)
    assert_same(snapshot(a, len(cases)), snapshot(b, len(cases)), "run after the phases")
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_bucket_overflow_falls_back(prec):
    M = U.package()
    P = M.FP64 if prec == "fp64" else M.FP32
    cases = [CS.case("b", 3, (8, 8, 8), 8, 0, 1, 1, rotang=ROT3)]
    print(f"[bucket overflow] {prec}", flush=True)
    # tiles of 0.5 x (bucket capacity) / 8 cells at fill 0.5: the usable
    # slots are set to the median over steps of the fullest tile, so about
    # every other step overflows -- at the move into buckets or inside a run
    probe = make_ctx(M, cases, P)
    probe.set_layout(True, fill=0.5)
    full = probe.layout_stats()[2]
    tc = int(0.5 * full / 8)
    probe.set_layout(False)
    maxes = []
    for _ in range(8):
        probe.run(1)
        _, cstart = probe.order(0)
        pop = np.diff(cstart)
        maxes.append(int(np.add.reduceat(pop, np.arange(0, len(pop), tc)).max()))
    probe.close()
    cap = int(np.median(maxes))
    print(f"[bucket overflow] fullest tiles {maxes}, capacity {cap}", flush=True)
    a = make_ctx(M, cases, P)
    b = make_ctx(M, cases, P)
    a.set_layout(False)
    b.set_layout(True, fill=0.5, capacity=cap)
    for n in (6, 9, 5, 8, 12):
        a.run(n, use_graph=True)
        b.run(n, use_graph=True)
        assert_same(snapshot(a, len(cases)), snapshot(b, len(cases)), f"run({n})")
    st = b.layout_stats()
    print("[bucket overflow] stats", st, flush=True)
    assert st[2] == cap and st[3] == (9 ** 3 + tc - 1) // tc, st
    assert st[1] > 0, "no overflow happened: the fallback was not exercised"
    assert st[0] > 0, "no bucketed step completed"
    a.close()
    b.close()
