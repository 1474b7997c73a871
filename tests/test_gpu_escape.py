"""GPU parity of the rare events of the step: bin()'s escape repair
(mpc.c:970-982 with replacePos_WithCheck, init.c:1462-1483) and the wall
re-thermalisation after NBOUNCE bounces (bc.c:1461-1488, push() init.c:1261).

* test_rare_fixture_phases: the reference-recorded cases with escapes
  (open axes, walls inside the box, nematics) and with re-thermalisations
  (four velocity distributions) replayed through the three-phase API, which
  runs the split kernels, against the reference's own state after every step;
  test_gpu_parity.test_gpu_matches_reference replays the same fixtures
  through the whole-step entry point (the fused kernel for the isotropic
  operators).
* test_rare_philox_*: the same kinds of systems at larger sizes with the
  library's own generator, against the oracle's Philox mode (the keys of the
  escape and re-thermalisation draws are shared), through mpcd_gpu_step and
  through mpcd_gpu_run, whose bucketed kernels re-place escapes in the
  shifted bin inline and finish streaming escapes in the wall-list kernel.
FP64 throughout: order and populations bit-exact, r, v, u within 1e-12
relative (plus 1e-12 absolute near zero).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
import gpu_util as U  # noqa: E402
import cases as CS  # noqa: E402
from test_gpu_parity import _assert_close, _random_box, assert_lc_close, assert_same_decisions  # noqa: E402

RARE = ["srd2d_open", "lc2d_open", "at3d_open_inset", "lc3d_open", "at2d_nbounce"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", RARE)
def test_rare_fixture_phases(name):
    import pyoracle
    M = U.package()
    g = U.load_golden(name)
    c = g["case_d"]
    N = g["Q0"].shape[0]
    print(f"[gpu rare phases] case={name} N={N} steps={c['steps']}", flush=True)
    ctx = M.Context(U.gpu_cfg(c, float(g["gndnst"]), g["avv"]), N, seed=c["seed"])
    ctx.upload(0, g["Q0"], g["V0"], g["U0"] if c["lc"] else None, g["spid"])
    if c["lc"]:
        ctx.lc_record(True)
    osys = pyoracle.System(CS.oracle_cfg(c, float(g["gndnst"]), g["avv"]), g["Q0"], g["V0"], g["U0"], g["spid"])
    rec = U.KeyedRecorder(osys, c["seed"], ctx.ncells[0], N)
    for k in range(c["steps"]):
        d = rec.step()
        ctx.bin(d, shifted=1)
        ctx.collide(d)
        events = ctx.stream()  # running counts since the context was made
        order, cstart = ctx.order(0)
        np.testing.assert_array_equal(np.diff(cstart), g["pop"][k], err_msg=f"{name} step {k} populations")
        np.testing.assert_array_equal(order, g["order"][k], err_msg=f"{name} step {k} list order")
        Q, V, Uo = ctx.download(0)
        close = _assert_close
        if c["lc"]:
            assert_same_decisions(ctx, osys, f"{name} step {k}")
            close = assert_lc_close
        close(Q, g["Q"][k], f"{name} step {k} Q")
        close(V, g["V"][k], f"{name} step {k} V")
        if c["lc"]:
            close(Uo, g["U"][k], f"{name} step {k} U")
    # the oracle counted the same events: escapes (both binnings), particles
    # past NBOUNCE, rewinds, re-thermalisations
    assert tuple(events) == tuple(osys.counters()), (tuple(events), osys.counters())
    ctx.close()


def _open_box(kind, seed):
    """systems whose particles leave the box (open axes, walls inside it) or
    bounce more than 2 NBOUNCE times in a step"""
    if kind == "srd2d_open":
        L = (24, 20)
        c, Q, V, Uo = _random_box(2, L, 10, 0, seed=seed, walls=CS.pbc_walls(0, L[0]), dt=1.0)
    elif kind == "srd3d_open":
        L = (16, 12, 14)
        c, Q, V, Uo = _random_box(3, L, 10, 0, seed=seed, walls=CS.pbc_walls(0, L[0]) + CS.pbc_walls(1, L[1]),
                                  dt=1.0)
    elif kind in ("at3d_inset", "rat3d_inset"):
        L = (12, 12, 10)
        walls = CS.pbc_walls(0, L[0]) + [dict(A=[0, 1, 0], R=2.0, DN=0.0, MVN=-1.0, MVT=-1.0),
                                         dict(A=[0, -1, 0], R=-10.0, DN=0.0, MVN=-1.0, MVT=-1.0)]
        c, Q, V, Uo = _random_box(3, L, 8, 2 if kind == "at3d_inset" else 3, seed=seed, walls=walls, dt=1.0)
        Q[:, 1] = 2.2 + (Q[:, 1] / L[1]) * 7.6
    elif kind == "at2d_nbounce":
        L = (20, 6)
        c, Q, V, Uo = _random_box(2, L, 10, 2, seed=seed, walls=CS.pbc_walls(0, L[0]) + CS.solid_walls(1, L[1]))
        c["species"] = [CS.species(vdist=v) for v in (0, 2, 3, 4)]
        V[:8, 1] = 2e5 * np.where(np.arange(8) % 2 == 0, 1.0, -1.0)
    elif kind == "lc2d_open":
        L = (20, 16)
        c, Q, V, Uo = _random_box(2, L, 12, 16, lc=1, seed=seed, walls=CS.pbc_walls(0, L[0]), dt=1.0,
                                  sp=dict(mfpot=10.0, tumble=0.5, act=0.05))
    else:
        raise ValueError(kind)
    spid = (np.arange(Q.shape[0]) % len(c["species"])).astype(np.int32)
    return c, Q, V, Uo, spid


def _philox_pair(kind, seed):
    import pyoracle
    M = U.package()
    c, Q, V, Uo, spid = _open_box(kind, seed)
    gnd = 10.0
    osys = pyoracle.System(CS.oracle_cfg(c, gnd, np.zeros(3)), Q, V, Uo, spid)
    d = pyoracle.Draws(pyoracle.DRAW_PHILOX, seed, 0)
    ctx = M.Context(U.gpu_cfg(c, gnd, np.zeros(3)), Q.shape[0], seed=seed)
    ctx.upload(0, Q, V, Uo if c["lc"] else None, spid)
    return c, osys, d, ctx


def _compare(ctx, osys, lc, what):
    Qo, Vo, Uor = osys.state()
    Qg, Vg, Ug = ctx.download(0)
    o_ord, o_cs = osys.order()
    g_ord, g_cs = ctx.order(0)
    np.testing.assert_array_equal(g_cs, o_cs, err_msg=f"{what}: populations")
    np.testing.assert_array_equal(g_ord, o_ord, err_msg=f"{what}: list order")
    close = assert_lc_close if lc else _assert_close
    close(Qg, Qo, f"{what}: Q")
    close(Vg, Vo, f"{what}: V")
    if lc:
        close(Ug, Uor, f"{what}: U")


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["srd2d_open", "srd3d_open", "at3d_inset", "at2d_nbounce", "lc2d_open"])
def test_rare_philox_steps(kind):
    """mpcd_gpu_step (fused kernel, or the split kernels for nematics) vs the oracle, step by step"""
    c, osys, d, ctx = _philox_pair(kind, 41)
    print(f"[gpu rare philox step] {kind} N={ctx.n[0]}", flush=True)
    if c["lc"]:
        ctx.lc_record(True)
    for k in range(4):
        osys.step(d)
        ctx.step(None)
        if c["lc"]:
            assert_same_decisions(ctx, osys, f"{kind} step {k}")
        _compare(ctx, osys, c["lc"], f"{kind} step {k}")
    ev = osys.counters()
    assert ev[0] > 0 or ev[3] > 0, f"{kind}: no rare event happened ({ev})"
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["srd2d_open", "srd3d_open", "at3d_inset", "at2d_nbounce"])
@pytest.mark.parametrize("graph", [False, True])
def test_rare_philox_run(kind, graph):
    """mpcd_gpu_run (bucketed kernels after the first step) vs the oracle"""
    c, osys, d, ctx = _philox_pair(kind, 43)
    print(f"[gpu rare philox run] {kind} graph={graph} N={ctx.n[0]}", flush=True)
    for k in range(6):
        osys.step(d)
    ctx.run(6, use_graph=graph)
    stats = ctx.layout_stats()
    assert stats[0] >= 4 and stats[1] == 0, f"{kind}: the run did not stay bucketed ({stats})"
    _compare(ctx, osys, c["lc"], f"{kind} run")
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["srd2d_open", "srd3d_open", "at3d_inset", "rat3d_inset", "at2d_nbounce"])
def test_rare_run_equals_steps(kind):
    """mpcd_gpu_run's bucketed kernels (shifted-bin escapes re-placed by the
    preparation kernel, streaming escapes in the wall-list kernel) give the
    same bits as the cell-sorted steps of mpcd_gpu_step.  This also covers
    MPC-AT+a, whose comparison with the oracle over many steps is left to the
    keyed-draw fixtures: the generator's normals go through the device's and
    the host's log/sin/cos, which may differ in the last bit, and the inverse
    inertia tensor of a cell holding a far re-placed particle amplifies that."""
    M = U.package()
    c, Q, V, Uo, spid = _open_box(kind, 47)
    cfg = U.gpu_cfg(c, 10.0, np.zeros(3))
    a = M.Context(cfg, Q.shape[0], seed=47)
    b = M.Context(cfg, Q.shape[0], seed=47)
    for x in (a, b):
        x.upload(0, Q, V, None, spid)
    print(f"[gpu rare run=steps] {kind} N={a.n[0]}", flush=True)
    a.run(8, use_graph=False)
    for _ in range(8):
        b.step(None)
    assert a.layout_stats()[0] >= 6
    for x, y in zip(a.download(0), b.download(0)):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a.order(0)[0], b.order(0)[0])
    a.close()
    b.close()
