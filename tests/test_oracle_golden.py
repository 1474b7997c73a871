"""Pins the oracle against the reference's own C step.

Every fixture in tests/golden/ was recorded by running the reference's
timestep() (compiled from its sources by oracle/Makefile into oracle/_ref)
on a known initial state with injected sequential draws.  The oracle, run in
its sequential draw mode on the same initial state, must reproduce:

  * the linked-list order of every cell after every step (bit-exact),
  * per-cell populations (bit-exact),
  * the number of draws of each kind consumed (bit-exact),
  * positions, velocities and orientations (bit-exact; the test tolerates
    1e-13 relative only to survive a different libm on another host),
  * KBTNOW and the per-cell S, director and centre-of-mass velocity,
  * with the flow output on, the running flow sum FLOW after the case's steps
    (localFLOW + sumFLOW, mpc.c:226-246; tests/golden/flow.npz, recorded from
    a second reference run with outFlags.FLOWOUT set).

Parts of the oracle that no fixture exercises, and how they are pinned
instead (tests/test_oracle_invariants.py):
  * escape repair (bin(), mpc.c:970-982)      -- forced from the host, counted
  * wall re-thermalisation after NBOUNCE       -- not reachable with the
    planar walls in scope; parity unpinned beyond the counter
  * keyed and Philox draw modes                -- checked against the
    sequential mode via recorded keyed arrays and against published
    Philox-4x32-10 known-answer vectors.
"""
import json
import os

import numpy as np
import pytest

import cases as C
import pyoracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")
NAMES = [c["name"] for c in C.CASES]


def load(name):
    z = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def oracle_for(g):
    c = json.loads(str(g["case"]))
    cfg = C.oracle_cfg(c, float(g["gndnst"]), g["avv"])
    sys_ = pyoracle.System(cfg, g["Q0"], g["V0"], g["U0"], g["spid"])
    return c, sys_


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_reference(name):
    g = load(name)
    c, s = oracle_for(g)
    order0, _ = s.order()
    np.testing.assert_array_equal(order0, g["order0"])
    d = pyoracle.Draws(pyoracle.DRAW_SEQ, c["seed"])
    for k in range(c["steps"]):
        s.step(d)
        if C.rest_frame_due(c, k):
            s.zero_net_momentum()
        order, cstart = s.order()
        pop = np.diff(cstart)
        np.testing.assert_array_equal(pop, g["pop"][k], err_msg=f"{name} step {k} populations")
        np.testing.assert_array_equal(order, g["order"][k], err_msg=f"{name} step {k} list order")
        assert tuple(d.counts()) == tuple(int(x) for x in g["draws"][k]), f"{name} step {k} draw counts"
        Q, V, U = s.state()
        for lab, a, b in (("Q", Q, g["Q"][k]), ("V", V, g["V"][k]), ("U", U, g["U"][k])):
            np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13, err_msg=f"{name} step {k} {lab}")
        np.testing.assert_allclose(s.kbtnow(), g["kbt"][k], rtol=1e-13, err_msg=f"{name} step {k} KBTNOW")
        cells = s.cells()
        occ = pop > 1
        np.testing.assert_allclose(cells["vcm"], g["VCM"][k], rtol=1e-12, atol=1e-13)
        if c["lc"]:
            np.testing.assert_allclose(cells["S"][occ], g["S"][k][occ], rtol=1e-11, atol=1e-12)
            np.testing.assert_allclose(cells["dir"][occ], g["DIR"][k][occ], rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_bitexact_on_this_host(name):
    """With the glibc the fixtures were recorded against (the same image runs every test) the match is exact."""
    g = load(name)
    c, s = oracle_for(g)
    d = pyoracle.Draws(pyoracle.DRAW_SEQ, c["seed"])
    for k in range(c["steps"]):
        s.step(d)
        if C.rest_frame_due(c, k):
            s.zero_net_momentum()
    Q, V, U = s.state()
    diff = max(np.abs(Q - g["Q"][-1]).max(), np.abs(V - g["V"][-1]).max(), np.abs(U - g["U"][-1]).max())
    assert diff == 0.0, f"{name}: oracle differs from the reference by {diff:.3e} (expected bit-exact)"


@pytest.mark.parametrize("name", NAMES)
def test_oracle_flow_matches_reference(name):
    g = load(name)
    c, s = oracle_for(g)
    ref = np.load(os.path.join(GOLD, "flow.npz"), allow_pickle=False)["FLOW_" + name]
    s.flow(True)
    d = pyoracle.Draws(pyoracle.DRAW_SEQ, c["seed"])
    for k in range(c["steps"]):
        s.step(d)
        if C.rest_frame_due(c, k):
            s.zero_net_momentum()
    f, n = s.get_flow()
    assert n == c["steps"]
    np.testing.assert_allclose(f, ref, rtol=1e-13, atol=1e-13, err_msg=f"{name} FLOW")
    assert np.all(f[:, c["dim"]:] == 0.0)
    # a read with reset clears the sums and the count, as flowout clears FLOW (pout.c:1696)
    f2, n2 = s.get_flow(reset=True)
    assert n2 == n and np.array_equal(f2, f)
    f3, n3 = s.get_flow()
    assert n3 == 0 and not f3.any()
