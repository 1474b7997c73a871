"""GPU: the batched form's hipGraph replays and the tile autotuner change no
result.  A graph replays two fused steps with the step number read from device
memory; eager steps and replays of the same run must agree bit for bit, and
so must runs with different tile tables (without a thermostat the tiling only
regroups whole cells)."""
import numpy as np
import pytest

import cases as CS
import gpu_util as U

pytestmark = pytest.mark.gpu


def cfg3(L, rtech=0, tstech=0, walls=None, grav=(0, 0, 0)):
    c = CS.case("g", 3, L, 8, rtech, steps=1, seed=1, rotang=2.2689280275926285, tstech=tstech, walls=walls,
                grav=grav)
    return c, U.gpu_cfg(c, 8.0, [0.0, 0.0, 0.0])


def make_ctx(M, cfgs, prec, seed=77):
    n = [int(np.prod(c["L"])) * 8 for c, _ in cfgs]
    ctx = M.Context([g for _, g in cfgs], n, seed=seed, precision=prec)
    states = []
    for r, (c, _) in enumerate(cfgs):
        cc = dict(c)
        cc["seed"] = 10 + r
        Q, V, _, _, _ = CS.initial_state(cc)
        ctx.upload(r, Q, V)
        states.append((Q, V))
    return ctx


def snapshot(ctx, nrep):
    out = []
    for r in range(nrep):
        Q, V, _ = ctx.download(r)
        out.append((Q.copy(), V.copy()))
    return out


def same(a, b):
    for (qa, va), (qb, vb) in zip(a, b):
        if not (np.array_equal(qa, qb) and np.array_equal(va, vb)):
            return False
    return True


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["srd", "at_walls_grav", "srd_vsc"])
def test_graph_replay_matches_eager(prec, kind):
    M = U.package()
    P = M.FP64 if prec == "fp64" else M.FP32
    if kind == "srd":
        cfgs = [cfg3((8, 8, 8)), cfg3((6, 10, 7))]
    elif kind == "srd_vsc":
        cfgs = [cfg3((8, 8, 8), tstech=1), cfg3((6, 10, 7), tstech=1)]
    else:
        walls = CS.pbc_walls(0, 8) + CS.pbc_walls(1, 8) + CS.solid_walls(2, 8)
        cfgs = [cfg3((8, 8, 8), rtech=3, walls=walls, grav=(1e-3, 0, 0)),
                cfg3((8, 8, 8), rtech=3, walls=walls, grav=(1e-3, 0, 0))]
    print(f"[graph] {kind} {prec}", flush=True)
    a = make_ctx(M, cfgs, P)
    b = make_ctx(M, cfgs, P)
    a.run(9, use_graph=False)
    b.run(9, use_graph=True)  # one eager step to reach an even step, then 4 replays
    a.synchronize()
    b.synchronize()
    sa, sb = snapshot(a, 2), snapshot(b, 2)
    assert same(sa, sb)
    a.run(4, use_graph=False)
    b.run(4, use_graph=True)  # the recorded graph is reused
    assert same(snapshot(a, 2), snapshot(b, 2))
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_autotune_changes_no_result(prec):
    M = U.package()
    P = M.FP64 if prec == "fp64" else M.FP32
    cfgs = [cfg3((8, 8, 8)), cfg3((6, 10, 7))]
    print(f"[autotune] {prec}", flush=True)
    a = make_ctx(M, cfgs, P)
    b = make_ctx(M, cfgs, P)
    trial = 2
    ms = a.autotune(trial)  # 6 bucket fills x (1 + trial) real steps
    assert ms > 0.0
    assert 0.2 <= a.tile_fill() <= 0.9
    assert a.layout_stats()[0] > 0, "the contexts take the bucketed layout"
    b.run(6 * (1 + trial), use_graph=False)
    a.run(6, use_graph=True)
    b.run(6, use_graph=True)
    assert same(snapshot(a, 2), snapshot(b, 2))
    a.close()
    b.close()
