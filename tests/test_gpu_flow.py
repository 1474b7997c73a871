"""GPU: the running flow field (mpcd_gpu_flow / mpcd_gpu_flow_read), the
reference's localFLOW + sumFLOW (mpc.c:226-246) at the end of every step.

* test_flow_matches_reference: golden cases replayed with the reference's
  recorded draws.  After the case's steps the GPU's sums equal the FLOW the
  reference itself accumulated (tests/golden/flow.npz, recorded with its flow
  output on) and the oracle's, and the step count is the case's.  The GPU adds
  the terms of a cell with atomics in no fixed order, so the bar is FP64
  round-off: 1e-12 absolute and relative on sums of O(1) velocities.
* test_flow_runs: whole runs with the internal generator.  The bucketed
  layout, the cell-sorted layout and graph replays agree with the oracle in
  its Philox mode to the same bar, steps taken with the flow off add nothing,
  and a read with reset starts the sums and the count again from zero (as
  flowout clears FLOW, pout.c:1696).
"""
import os

import numpy as np
import pytest

import cases as CS
import gpu_util as U

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12
ROT3 = 2.2689280275926285


def _close(a, b, what):
    err = np.abs(np.asarray(a) - np.asarray(b))
    lim = TOL + TOL * np.abs(b)
    assert (err <= lim).all(), f"{what}: worst |diff| {err.max():.3e} (limit {lim.max():.3e})"


@pytest.mark.parametrize("name", ["srd2d_arb_grav_vsc", "srd3d_arb", "at3d_damp", "rat3d_channel", "rat2d_2sp",
                                  "lc2d_channel", "srd2d_couette"])
def test_flow_matches_reference(name):
    import pyoracle
    M = U.package()
    g = U.load_golden(name)
    c = g["case_d"]
    ref = np.load(os.path.join(HERE, "golden", "flow.npz"))["FLOW_" + name]
    N = g["Q0"].shape[0]
    print(f"[gpu flow] case={name} N={N} steps={c['steps']}", flush=True)
    ctx = M.Context(U.gpu_cfg(c, float(g["gndnst"]), g["avv"]), N, seed=c["seed"], precision=M.FP64)
    ctx.upload(0, g["Q0"], g["V0"], g["U0"] if c["lc"] else None, g["spid"])
    osys = pyoracle.System(CS.oracle_cfg(c, float(g["gndnst"]), g["avv"]), g["Q0"], g["V0"], g["U0"], g["spid"])
    osys.flow(True)
    ctx.flow(True)
    rec = U.KeyedRecorder(osys, c["seed"], ctx.ncells[0], N)
    for k in range(c["steps"]):
        ctx.step(rec.step())
        if CS.rest_frame_due(c, k):
            ctx.zero_net_momentum()
            osys.zero_net_momentum()
    flow, n = ctx.flow_read(0)
    oflow, on = osys.get_flow()
    assert n == on == c["steps"]
    np.testing.assert_array_equal(oflow, ref)
    _close(flow, ref, f"{name}: GPU FLOW vs the reference's")
    assert np.all(flow[:, c["dim"]:] == 0.0)
    ctx.close()


def _box(kind):
    if kind == "srd3d":
        return CS.case("f", 3, (8, 8, 8), 8, 0, 1, 1, rotang=ROT3)
    if kind == "at_channel":
        walls = CS.pbc_walls(0, 8) + CS.pbc_walls(1, 8) + CS.solid_walls(2, 8)
        return CS.case("f", 3, (8, 8, 8), 8, 2, 1, 1, walls=walls, grav=(1e-3, 0, 0), sp=CS.species(damp=0.1))
    if kind == "srd2d_2sp":
        c = CS.case("f", 2, (20, 16), 9, 1, 1, 1)
        c["species"] = [CS.species(mass=1.0), CS.species(mass=2.5)]
        return c
    raise ValueError(kind)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["srd3d", "at_channel", "srd2d_2sp"])
def test_flow_runs(prec, kind):
    import pyoracle
    M = U.package()
    P = M.FP64 if prec == "fp64" else M.FP32
    c = _box(kind)
    cc = dict(c)
    cc["seed"] = 40
    Q, V, _, spid, _ = CS.initial_state(cc)
    sp = spid if len(c["species"]) > 1 else None
    cfg = U.gpu_cfg(c, float(c["ppc"]), [0.0, 0.0, 0.0])
    print(f"[gpu flow] runs {kind} {prec} N={Q.shape[0]}", flush=True)
    ctxs = []
    for layout in (False, True):
        x = M.Context(cfg, Q.shape[0], seed=77, precision=P)
        x.upload(0, Q, V, spid=sp)
        x.set_layout(layout)
        ctxs.append(x)
    osys = pyoracle.System(CS.oracle_cfg(c, float(c["ppc"]), np.zeros(3)), Q, V, np.zeros_like(Q), spid)
    d = pyoracle.Draws(pyoracle.DRAW_PHILOX, 77, 0)

    def advance(n, graph, on):
        for x in ctxs:
            x.flow(on)
            x.run(n, use_graph=graph)
        osys.flow(on)
        for _ in range(n):
            osys.step(d)

    def check(what, reset=False):
        """FP64: both layouts against the oracle.  FP32: the layouts hold the
        same particles bit for bit, so their sums agree to round-off (the
        FP32 state itself drifts from the FP64 oracle)"""
        of, on = osys.get_flow(reset=reset)
        got = [x.flow_read(0, reset=reset) for x in ctxs]
        for (f, n), lay in zip(got, ("cell-sorted", "bucketed")):
            assert n == on, f"{what} {lay}: {n} steps summed, the oracle {on}"
            if prec == "fp64":
                _close(f, of, f"{what} {lay} vs oracle")
        _close(got[1][0], got[0][0], f"{what}: bucketed vs cell-sorted")

    advance(2, False, False)
    assert ctxs[0].flow_read(0)[1] == 0
    advance(3, False, True)
    check("3 eager steps")
    advance(6, True, True)
    check("+6 steps in graph pairs")
    advance(2, True, False)
    check("+2 steps with the flow off")
    check("read with reset", reset=True)
    advance(5, True, True)
    check("+5 steps after the reset")
    st = ctxs[1].layout_stats()
    assert st[0] > 0, "no bucketed step ran"
    for x in ctxs:
        x.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_flow_bucket_overflow(prec):
    """A bucket overflow discards the overflowing step and finishes the run on
    the cell-sorted path: the discarded steps add nothing, the steps run
    again add once, so the sums and the count match a cell-sorted run."""
    M = U.package()
    P = M.FP64 if prec == "fp64" else M.FP32
    c = _box("srd3d")
    cc = dict(c)
    cc["seed"] = 40
    Q, V, _, _, _ = CS.initial_state(cc)
    cfg = U.gpu_cfg(c, float(c["ppc"]), [0.0, 0.0, 0.0])
    print(f"[gpu flow] bucket overflow {prec}", flush=True)
    # usable slots per bucket: the median over steps of the fullest tile at
    # fill 0.5, so that about every other step overflows
    probe = M.Context(cfg, Q.shape[0], seed=77, precision=P)
    probe.upload(0, Q, V)
    probe.set_layout(True, fill=0.5)
    tc = int(0.5 * probe.layout_stats()[2] / 8)
    probe.set_layout(False)
    maxes = []
    for _ in range(8):
        probe.run(1)
        pop = np.diff(probe.order(0)[1])
        maxes.append(int(np.add.reduceat(pop, np.arange(0, len(pop), tc)).max()))
    probe.close()
    cap = int(np.median(maxes))
    a = M.Context(cfg, Q.shape[0], seed=77, precision=P)
    b = M.Context(cfg, Q.shape[0], seed=77, precision=P)
    for x in (a, b):
        x.upload(0, Q, V)
        x.flow(True)
    a.set_layout(False)
    b.set_layout(True, fill=0.5, capacity=cap)
    for n in (6, 9, 5):
        a.run(n, use_graph=True)
        b.run(n, use_graph=True)
        fa, na = a.flow_read(0)
        fb, nb = b.flow_read(0)
        assert na == nb, f"run({n}): {nb} steps summed, the cell-sorted run {na}"
        _close(fb, fa, f"run({n}): bucketed with overflows vs cell-sorted")
    st = b.layout_stats()
    assert st[1] > 0, "no overflow happened: the fallback was not exercised"
    a.close()
    b.close()
