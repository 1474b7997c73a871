"""GPU parity: the HIP library (through its C-ABI) against the reference's own
step and against the oracle.

* test_gpu_matches_reference: every golden case recorded from the reference
  (tests/golden) is replayed on the GPU with the reference's draws injected
  in the keyed layout (recorded by the oracle's sequential mode, which equals
  the reference draw for draw).  List order and cell populations must match
  bit-exactly; FP64 positions, velocities and orientations within 1e-12
  relative (plus 1e-12 absolute near zero), and so are KBTNOW after the
  thermostat and the cell values the library returns (populations, VCM, and
  for nematics S and the director) against the reference's own after every
  step.  The nematic path calls cos, sin, acos, atan and exp, whose device
  (ocml) and host (glibc) results may differ in the last bit; those
  differences stay below 1e-12 on every case here (the largest seen is
  about 3e-13, on the 3D director of a near-isotropic cell), and a flipped
  discrete decision (a Metropolis acceptance, the S <= 1 clamp, a branch of
  the eigenvector or atan2 forms) would show as an O(1) difference.
* test_gpu_phases: the three-phase API (bin / collide / stream) gives the
  same result as the whole-step entry point.
* test_gpu_philox_vs_oracle: internal counter-based RNG mode against the
  oracle's Philox mode on larger boxes.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
import gpu_util as U  # noqa: E402
import cases as CS  # noqa: E402

NAMES = [c["name"] for c in CS.CASES]
RTOL = 1e-12
ATOL = 1e-12


def _assert_close(a, b, what, rtol=RTOL, atol=ATOL):
    a = np.asarray(a)
    b = np.asarray(b)
    err = np.abs(a - b)
    lim = atol + rtol * np.abs(b)
    bad = err > lim
    if bad.any():
        i = np.unravel_index(np.argmax(err - lim), err.shape)
        raise AssertionError(f"{what}: {bad.sum()} entries out of tolerance; worst at {i}: {a[i]!r} vs {b[i]!r}")


LC_ROUND = 1e-9  # bound of a rounding difference once every decision is shown equal


def assert_same_decisions(ctx, osys, what):
    """the nematic path's discrete decisions of the last step (Metropolis
    acceptances, samplers, redraws and signs per particle, the side of the
    dipole force; eigen-solver branches and the S <= 1 clamp per cell) are the
    oracle's, word for word"""
    gp, gc = ctx.lc_decisions(0)
    op, oc = osys.lc_decisions()
    bp, bc = np.nonzero(gp != op)[0], np.nonzero(gc != oc)[0]
    assert bp.size == 0 and bc.size == 0, (
        f"{what}: flipped decisions at particles {bp[:6].tolist()} ({[hex(x) for x in gp[bp[:3]]]} vs "
        f"{[hex(x) for x in op[bp[:3]]]}) and cells {bc[:6].tolist()} ({[hex(x) for x in gc[bc[:3]]]} vs "
        f"{[hex(x) for x in oc[bc[:3]]]})")
    assert (gp & 1).any(), f"{what}: no particle was recorded"


def assert_lc_close(a, b, what):
    """a value of a nematic run: within 1e-12 relative (plus 1e-12 absolute),
    or -- the run's decisions having been checked equal at every step by
    assert_same_decisions -- within LC_ROUND, the device's and the host's
    cos/sin/acos/exp rounding carried through the steps; the count of such
    entries is printed"""
    a = np.asarray(a)
    b = np.asarray(b)
    err = np.abs(a - b)
    over = err > ATOL + RTOL * np.abs(b)
    if over.any():
        print(f"  {what}: {over.sum()} entries above 1e-12 with every decision equal, largest {err.max():.2e}",
              flush=True)
    _assert_close(a, b, what, rtol=LC_ROUND, atol=LC_ROUND)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_reference(name):
    import pyoracle
    M = U.package()
    g = U.load_golden(name)
    c = g["case_d"]
    cfg = U.gpu_cfg(c, float(g["gndnst"]), g["avv"])
    print(f"[gpu parity] case={name} N={g['Q0'].shape[0]} steps={c['steps']}", flush=True)
    ctx = M.Context(cfg, g["Q0"].shape[0], seed=c["seed"], precision=M.FP64)
    ctx.upload(0, g["Q0"], g["V0"], g["U0"] if c["lc"] else None, g["spid"])
    if c["lc"]:
        ctx.lc_record(True)
    o, cs = ctx.order(0)
    np.testing.assert_array_equal(o, g["order0"], err_msg=f"{name}: initial binning order")
    ocfg = CS.oracle_cfg(c, float(g["gndnst"]), g["avv"])
    osys = pyoracle.System(ocfg, g["Q0"], g["V0"], g["U0"], g["spid"])
    rec = U.KeyedRecorder(osys, c["seed"], ctx.ncells[0], g["Q0"].shape[0])
    for k in range(c["steps"]):
        draws = rec.step()
        kbt = ctx.step(draws)
        if CS.rest_frame_due(c, k):
            # the rest-frame transformation of the output step (galileantrans)
            ctx.zero_net_momentum()
            osys.zero_net_momentum()
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
        # KBTNOW and the cell values of the step's last localPROP (mpc.c:4942),
        # which the C host's outputs read through mpcd_gpu_get_cells
        close(kbt[0], g["kbt"][k], f"{name} step {k} KBTNOW")
        close(ctx.kbtnow()[0], g["kbt"][k], f"{name} step {k} mpcd_gpu_kbtnow")
        cells = ctx.cells(0)
        np.testing.assert_array_equal(cells["pop"], g["pop"][k], err_msg=f"{name} step {k} get_cells pop")
        close(cells["vcm"], g["VCM"][k], f"{name} step {k} VCM")
        if c["lc"]:
            close(cells["S"], g["S"][k], f"{name} step {k} S")
            close(cells["dir"], g["DIR"][k], f"{name} step {k} DIR")
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["srd3d_arb", "rat3d_channel", "lc2d_channel"])
def test_gpu_phases(name):
    """bin + collide + stream phases == one step (same injected draws)."""
    import pyoracle
    M = U.package()
    g = U.load_golden(name)
    c = g["case_d"]
    cfg = U.gpu_cfg(c, float(g["gndnst"]), g["avv"])
    N = g["Q0"].shape[0]
    a = M.Context(cfg, N, seed=3)
    b = M.Context(cfg, N, seed=3)
    for x in (a, b):
        x.upload(0, g["Q0"], g["V0"], g["U0"] if c["lc"] else None, g["spid"])
    osys = pyoracle.System(CS.oracle_cfg(c, float(g["gndnst"]), g["avv"]), g["Q0"], g["V0"], g["U0"], g["spid"])
    rec = U.KeyedRecorder(osys, c["seed"], a.ncells[0], N)
    for k in range(2):
        d = rec.step()
        a.step(d)
        cell, cnt, perm = b.bin(d, shifted=1, outputs=True)
        assert cnt.sum() == N
        assert sorted(perm.tolist()) == list(range(N))
        b.collide(d)
        b.stream()
    for x, y in zip(a.download(0), b.download(0)):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a.order(0)[0], b.order(0)[0])


def _random_box(dim, L, ppc, rtech, lc=0, seed=5, walls=None, **kw):
    rng = np.random.default_rng(seed)
    N = int(np.prod(L[:dim])) * ppc
    Q = np.zeros((N, 3))
    V = np.zeros((N, 3))
    Uo = np.zeros((N, 3))
    for d in range(dim):
        Q[:, d] = rng.uniform(0, L[d], N)
    V[:, :dim] = rng.normal(size=(N, dim))
    if dim == 3:
        u = rng.normal(size=(N, 3))
        Uo = u / np.linalg.norm(u, axis=1)[:, None]
    else:
        th = rng.uniform(0, 2 * np.pi, N)
        Uo[:, 0], Uo[:, 1] = np.cos(th), np.sin(th)
    c = dict(name="rand", dim=dim, L=list(L) + [1] * (3 - len(L)), rtech=rtech, lc=lc, tstech=kw.get("tstech", 0),
             rotang=kw.get("rotang", 2.2689280275926285), dt=kw.get("dt", 0.1), kbt=1.0,
             grav=kw.get("grav", [0, 0, 0]), walls=walls if walls is not None else CS.periodic(dim, L),
             species=[CS.species(**kw.get("sp", {}))], galinv=1, tau=0.5)
    return c, Q, V, Uo


@pytest.mark.gpu
@pytest.mark.parametrize("spec", [
    (3, (16, 16, 16), 10, 0, 0),
    (3, (12, 10, 14), 10, 3, 0),
    (2, (40, 40), 20, 16, 1),
    # 600 particles per cell: every cell tile of the unfused kernels is above
    # their LDS capacity, so the lane-per-cell fallbacks run
    (2, (3, 3), 600, 16, 1),
    (3, (3, 3, 3), 600, 3, 0),
])
def test_gpu_philox_vs_oracle(spec):
    import pyoracle
    M = U.package()
    dim, L, ppc, rtech, lc = spec
    c, Q, V, Uo = _random_box(dim, L, ppc, rtech, lc, sp=dict(mfpot=10.0, tumble=0.5, act=0.05))
    gnd = float(ppc)
    ocfg = CS.oracle_cfg(c, gnd, np.zeros(3))
    osys = pyoracle.System(ocfg, Q, V, Uo, None)
    d = pyoracle.Draws(pyoracle.DRAW_PHILOX, 77, 0)
    ctx = M.Context(U.gpu_cfg(c, gnd, np.zeros(3)), Q.shape[0], seed=77)
    ctx.upload(0, Q, V, Uo if lc else None, None)
    if lc:
        ctx.lc_record(True)
    print(f"[gpu philox] dim={dim} L={L} ppc={ppc} rtech={rtech} lc={lc}", flush=True)
    for k in range(3):
        osys.step(d)
        ctx.step(None)
        if lc:
            assert_same_decisions(ctx, osys, f"step {k}")
    Qo, Vo, Uor = osys.state()
    Qg, Vg, Ug = ctx.download(0)
    o_ord, _ = osys.order()
    g_ord, _ = ctx.order(0)
    np.testing.assert_array_equal(g_ord, o_ord)
    close = assert_lc_close if lc else _assert_close
    close(Qg, Qo, "Q")
    close(Vg, Vo, "V")
    if lc:
        close(Ug, Uor, "U")


@pytest.mark.gpu
@pytest.mark.parametrize("rtech", [20, 21, 17])
def test_gpu_active_species_vs_oracle(rtech):
    """two species, one active with a MINACTRATIO near its mean share (about
    half of the cells count enough of it to be active), in a channel with homeotropic anchoring
    at bounce-back phantom walls: the GPU's internal draws against the oracle's
    (the same Philox variates)"""
    import pyoracle
    M = U.package()
    L, ppc = (14, 10), 12
    walls = CS.pbc_walls(0, L[0]) + CS.solid_walls(1, L[1], MUN=1.0, MUT=0.0)
    c, Q, V, Uo = _random_box(2, L, ppc, rtech, 1, seed=17, walls=walls, grav=[5e-4, 0, 0])
    c["species"] = [CS.species(mfpot=10.0, tumble=0.5, act=0.05, minactratio=0.65, sigwidth=0.8, sigpos=0.9),
                    CS.species(mass=1.5, mfpot=6.0, tumble=0.8, act=0.02)]
    spid = (np.arange(Q.shape[0]) % 3 == 0).astype(np.int32)  # one particle in three of species 1
    gnd = float(ppc)
    osys = pyoracle.System(CS.oracle_cfg(c, gnd, np.zeros(3)), Q, V, Uo, spid)
    d = pyoracle.Draws(pyoracle.DRAW_PHILOX, 91, 0)
    ctx = M.Context(U.gpu_cfg(c, gnd, np.zeros(3)), Q.shape[0], seed=91)
    ctx.upload(0, Q, V, Uo, spid)
    print(f"[gpu philox] active species rtech={rtech}", flush=True)
    ctx.lc_record(True)
    for k in range(3):
        osys.step(d)
        ctx.step(None)
        assert_same_decisions(ctx, osys, f"step {k}")
    Qo, Vo, Uor = osys.state()
    Qg, Vg, Ug = ctx.download(0)
    o_ord, _ = osys.order()
    g_ord, _ = ctx.order(0)
    np.testing.assert_array_equal(g_ord, o_ord)
    assert_lc_close(Qg, Qo, "Q")
    assert_lc_close(Vg, Vo, "V")
    assert_lc_close(Ug, Uor, "U")


@pytest.mark.gpu
@pytest.mark.parametrize("rtech,mfpot", [(3, 12.0), (21, 0.6), (17, 3.0)])
def test_gpu_nematic_3d_channel_vs_oracle(rtech, mfpot):
    """3D nematic channel: periodic x and y, bounce-back phantom walls in z
    with planar strong anchoring (MUN 0, MUT 1), gravity; the Gaussian,
    cubic-root and Metropolis Maier-Saupe samplers by the mean-field strength"""
    import pyoracle
    M = U.package()
    L, ppc = (6, 5, 6), 10
    walls = CS.pbc_walls(0, L[0]) + CS.pbc_walls(1, L[1]) + CS.solid_walls(2, L[2], MUN=0.0, MUT=1.0)
    c, Q, V, Uo = _random_box(3, L, ppc, rtech, 1, seed=29, walls=walls, grav=[1e-3, 0, 0],
                              sp=dict(mfpot=mfpot, tumble=0.7, act=0.04))
    gnd = float(ppc)
    osys = pyoracle.System(CS.oracle_cfg(c, gnd, np.zeros(3)), Q, V, Uo, None)
    d = pyoracle.Draws(pyoracle.DRAW_PHILOX, 5, 0)
    ctx = M.Context(U.gpu_cfg(c, gnd, np.zeros(3)), Q.shape[0], seed=5)
    ctx.upload(0, Q, V, Uo, None)
    print(f"[gpu philox] 3D nematic channel rtech={rtech} U={mfpot}", flush=True)
    ctx.lc_record(True)
    for k in range(3):
        osys.step(d)
        ctx.step(None)
        assert_same_decisions(ctx, osys, f"step {k}")
    Qo, Vo, Uor = osys.state()
    Qg, Vg, Ug = ctx.download(0)
    o_ord, _ = osys.order()
    g_ord, _ = ctx.order(0)
    np.testing.assert_array_equal(g_ord, o_ord)
    assert_lc_close(Qg, Qo, "Q")
    assert_lc_close(Vg, Vo, "V")
    assert_lc_close(Ug, Uor, "U")
