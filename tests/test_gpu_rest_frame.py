"""GPU parity of the periodic rest-frame transformation of the output step
(galileantrans, mtools.c:673-718, with zeroExtraDims, mtools.c:727-744;
called by outputResults, pout.c:2557-2562, when runtime % zeroNetMom == 0).

The reference-recorded fixtures srd2d_couette_znm and at3d_channel_znm run
through test_gpu_parity.test_gpu_matches_reference (whole steps with
injected draws, mpcd_gpu_zero_net_momentum after the due steps).  Here:
* mpcd_gpu_run's bucketed layout between resets, against the oracle's
  Philox mode (FP64: order and populations bit-exact, r and v within 1e-12);
* the FP32 state: after a reset the mass-weighted mean velocity is AVV to
  FP32 rounding, and the cell values are those of the step (computed before
  the reset, as outputResults computes localPROP before galileantrans).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
import gpu_util as U  # noqa: E402
import cases as CS  # noqa: E402
from test_gpu_parity import _assert_close, _random_box  # noqa: E402


def _channel(two_species=True):
    L = (12, 10, 8)
    walls = CS.pbc_walls(0, L[0]) + CS.pbc_walls(1, L[1]) + CS.solid_walls(2, L[2])
    c, Q, V, Uo = _random_box(3, L, 10, 2, seed=61, walls=walls)
    if two_species:
        c["species"] = [CS.species(mass=1.0), CS.species(mass=1.6)]
    spid = (np.arange(Q.shape[0]) % len(c["species"])).astype(np.int32)
    return c, Q, V, spid


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph", [False, True])
def test_rest_frame_run_vs_oracle(use_graph):
    import pyoracle
    M = U.package()
    c, Q, V, spid = _channel()
    avv = np.array([0.02, -0.01, 0.0])
    osys = pyoracle.System(CS.oracle_cfg(c, 10.0, avv), Q, V, None, spid)
    d = pyoracle.Draws(pyoracle.DRAW_PHILOX, 19, 0)
    ctx = M.Context(U.gpu_cfg(c, 10.0, avv), Q.shape[0], seed=19)
    ctx.upload(0, Q, V, None, spid)
    print(f"[gpu rest frame] MPC-AT channel N={Q.shape[0]} graph={use_graph}", flush=True)
    m = np.where(spid == 0, 1.0, 1.6)
    for blk in range(3):
        for _ in range(4):
            osys.step(d)
        osys.zero_net_momentum()
        ctx.run(4, use_graph=use_graph)
        ctx.zero_net_momentum()
        Qg, Vg, _ = ctx.download(0)
        Qo, Vo, _ = osys.state()
        o_ord, o_cs = osys.order()
        g_ord, g_cs = ctx.order(0)
        np.testing.assert_array_equal(g_cs, o_cs, err_msg=f"block {blk}: populations")
        np.testing.assert_array_equal(g_ord, o_ord, err_msg=f"block {blk}: list order")
        _assert_close(Qg, Qo, f"block {blk}: Q")
        _assert_close(Vg, Vo, f"block {blk}: V")
        mean = (m[:, None] * Vg).sum(axis=0) / m.sum()
        np.testing.assert_allclose(mean, avv, atol=1e-13, err_msg=f"block {blk}: mean velocity after the reset")
    assert ctx.layout_stats()[0] >= 6, "the runs did not use the bucketed layout"
    ctx.close()


@pytest.mark.gpu
def test_rest_frame_fp32_and_cells():
    M = U.package()
    c, Q, V, spid = _channel()
    avv = np.array([0.0, 0.03, 0.0])
    a = M.Context(U.gpu_cfg(c, 10.0, avv), Q.shape[0], seed=23, precision=M.FP32)
    b = M.Context(U.gpu_cfg(c, 10.0, avv), Q.shape[0], seed=23, precision=M.FP32)
    for x in (a, b):
        x.upload(0, Q, V, None, spid)
        x.run(5, use_graph=False)
    print(f"[gpu rest frame fp32] MPC-AT channel N={Q.shape[0]}", flush=True)
    m = np.where(spid == 0, 1.0, 1.6)
    _, V0, _ = a.download(0)
    drift = (m[:, None] * V0).sum(axis=0) / m.sum() - avv
    assert np.abs(drift).max() > 1e-4, "the walls left the net momentum unchanged"
    # b resets before its cells are read: they are still the step's
    b.zero_net_momentum()
    ca, cb = a.cells(0), b.cells(0)
    np.testing.assert_array_equal(cb["pop"], ca["pop"])
    np.testing.assert_array_equal(cb["vcm"], ca["vcm"])
    _, Vg, _ = b.download(0)
    mean = (m[:, None] * Vg).sum(axis=0) / m.sum()
    np.testing.assert_allclose(mean, avv, atol=2e-7)
    # every particle moved by the same vector: the mean drift less AVV
    np.testing.assert_allclose(V0 - Vg, np.broadcast_to(drift, V0.shape), atol=2e-6)
    a.close()
    b.close()
