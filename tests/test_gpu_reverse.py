"""GPU: the reverse mode of the whole step (mpcd_gpu_step_backward_state) and
the observables it differentiates (mpcd_gpu_observable).

* For every operator family the tangent covers (2D active nematic channels
  with anchoring walls and both dipole rules, a 3D active nematic channel,
  MPC-AT+a, MPC-AT, SRD between phantom walls), dJ/dg and dJ/d(activity) of
  the time-averaged band velocity and, for the 2D nematics, of the mean
  scalar order parameter, computed backwards with checkpoints every few
  steps, equal the forward-mode tangent's to 1e-9 relative.  The tangent of
  the mean S comes from the tangent orientations through the first-order
  eigenvalue perturbation of the 2x2 order tensor (dS = v0^T dQ v0), written
  here in numpy.
* The observables' values are the numpy values of the downloaded state.
* A backward pass leaves the context's state as it was.
* The reverse gradient agrees with central finite differences (common
  random numbers) on a 2D active nematic channel.
* On a 16 x 8 active nematic channel it equals the FP64 CPU adjoint
  (tests/adjoint_lc_ref.py: torch autograd over a restatement of the step
  driven by the oracle with the same counter-based draws) to 1e-9.
"""
import importlib

import numpy as np
import pytest

import cases as CS
import gpu_util as U
from test_gpu_tangent import TAN_CASES, lc_channel, problem

pytestmark = pytest.mark.gpu

REL = 1e-9


def harness():
    U.package()
    return importlib.import_module("niamh_mpcd_amd.adjoint")


def band_obs(cfg):
    """the x velocity in a band of cells across the walls (or the middle of a periodic box)"""
    axis = 1 if cfg["dim"] == 2 else 2
    L = cfg["L"][axis]
    return (0, axis, L // 2 - 1, L // 2 + 1)


def mean_S_2d(Q, Uo, L):
    """mean S over the unshifted cells with more than one particle, and for
    each counted cell (members, top eigenvector, population)"""
    cell = np.trunc(Q[:, 0]).astype(int) * (L[1] + 1) + np.trunc(Q[:, 1]).astype(int)
    order = np.argsort(cell, kind="stable")
    vals, starts, counts = np.unique(cell[order], return_index=True, return_counts=True)
    S, info = [], []
    for s0, n in zip(starts, counts):
        if n < 2:
            continue
        idx = order[s0:s0 + n]
        u = Uo[idx, :2]
        Qt = 2.0 * (u.T @ u) / n - np.eye(2)
        w, v = np.linalg.eigh(Qt)
        S.append(w[-1])
        info.append((idx, v[:, -1], n))
    return float(np.mean(S)), info


def tangent_reference(M, cfg, Q, V, Uo, spid, K, seed, obs, direction):
    """J and dJ/dtheta along one direction by the tangent, with the same call
    sequence as the reverse harness (an observable after every step)"""
    ctx = M.Context(cfg, Q.shape[0], seed=seed, precision=M.FP64)
    try:
        ctx.upload(0, Q, V, U=Uo if cfg["lc"] else None, spid=spid if len(cfg["species"]) > 1 else None)
        ctx.tangent(True, **direction)
        J = dJ = 0.0
        for _ in range(K):
            ctx.step()
            val, _ = ctx.observable(*obs)
            Qs, Vs, Us = ctx.download(0)
            _, dV, dU = ctx.tangent_download(0)
            if obs[0] == M.OBS_BAND_VELOCITY:
                comp, axis, lo, hi = obs[1]
                y = np.trunc(Qs[:, axis])
                m = (y >= lo) & (y < hi)
                assert val[0] == pytest.approx(Vs[m, comp].mean(), rel=1e-13, abs=1e-15)
                J += Vs[m, comp].mean() / K
                dJ += dV[m, comp].mean() / K
            else:
                S, info = mean_S_2d(Qs, Us, cfg["L"])
                assert val[0] == pytest.approx(S, rel=1e-12)
                dS = [4.0 / n * np.sum((dU[idx, :2] @ v0) * (Us[idx, :2] @ v0)) for idx, v0, n in info]
                J += S / K
                dJ += float(np.mean(dS)) / K
        return J, dJ
    finally:
        ctx.close()


CASES = [
    ("lc2d_channel", "band"), ("lc2d_channel", "S"), ("lc2d_channel_sig", "band"), ("lc2d_channel_sig", "S"),
    ("lc3d_weak", "band"), ("rat3d_channel", "band"), ("at3d_damp", "band"), ("srd2d_phantom", "band"),
]


@pytest.mark.parametrize("name,kind", CASES)
def test_reverse_equals_tangent(name, kind):
    M = U.package()
    H = harness()
    c = TAN_CASES[name]()
    cfg, Q, V, Uo, spid = problem(c)
    obs = (M.OBS_MEAN_S, None) if kind == "S" else (M.OBS_BAND_VELOCITY, band_obs(cfg))
    K, seed = 10, 3
    sp = spid if len(cfg["species"]) > 1 else None
    print(f"[reverse] {name} {kind} N={Q.shape[0]} K={K}", flush=True)
    J, dg, dact = H.reverse_value_and_grad(M, cfg, Q, V, K, seed, obs, U0=Uo if cfg["lc"] else None, spid=sp, ckpt=4)
    nsp = len(cfg["species"])
    Jt, tg = tangent_reference(M, cfg, Q, V, Uo, spid, K, seed, obs, dict(dgrav=[1.0, 0.0, 0.0]))
    print(f"[reverse] {name} {kind} J={J} dJ/dg: reverse {dg[0]!r} tangent {tg!r}", flush=True)
    assert J == pytest.approx(Jt, rel=1e-13)
    assert dg[0] == pytest.approx(tg, rel=REL, abs=1e-12 * abs(tg) + 1e-300)
    if cfg["lc"] and cfg["rtech"] >= 16:
        _, ta = tangent_reference(M, cfg, Q, V, Uo, spid, K, seed, obs, dict(dact=[1.0] * nsp))
        da = float(np.sum(dact[:nsp]))
        print(f"[reverse] {name} {kind} dJ/dact: reverse {da!r} tangent {ta!r}", flush=True)
        assert ta != 0.0
        assert da == pytest.approx(ta, rel=REL)


def test_backward_leaves_the_state():
    M = U.package()
    c = TAN_CASES["lc2d_channel"]()
    cfg, Q, V, Uo, spid = problem(c)
    obs = (M.OBS_MEAN_S, None)
    print("[reverse] state after backward", flush=True)
    a = M.Context(cfg, Q.shape[0], seed=7, precision=M.FP64)
    b = M.Context(cfg, Q.shape[0], seed=7, precision=M.FP64)
    for x in (a, b):
        x.upload(0, Q, V, U=Uo)
    a.record(3)
    for _ in range(3):
        a.step()
        b.step()
    before = a.download(0)
    a.adjoint_reset()
    for _ in range(3):
        _, g = a.observable(*obs, grad=True)
        a.step_backward_state(g)
    for x, y in zip(a.download(0), before):
        np.testing.assert_array_equal(x, y)
    a.record(0)
    for _ in range(2):
        a.step()
        b.step()
    for x, y in zip(a.download(0), b.download(0)):
        np.testing.assert_array_equal(x, y)
    a.close()
    b.close()


def test_reverse_matches_finite_differences_active_nematic():
    """central differences (common random numbers) of the mean S and of the
    centreline velocity of a 2D active nematic channel"""
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed = 16, 8, 12, 20, 8
    g, act = 5e-4, 0.03

    def make(gg, aa):
        return problem(lc_channel(Lx, Ly, ppc, gg, aa))

    cfg, Q, V, Uo, _ = make(g, act)
    for obs in ((M.OBS_BAND_VELOCITY, (0, 1, 3, 5)), (M.OBS_MEAN_S, None)):
        print(f"[reverse] FD active nematic obs={obs[0]}", flush=True)
        J, dg, dact = H.reverse_value_and_grad(M, cfg, Q, V, K, seed, obs, U0=Uo, ckpt=5)

        def J_of(gg, aa):
            cf, q, v, u, _ = make(gg, aa)
            run = H.ReverseRun(M, cf, q, v, K, seed, obs, U0=u, ckpt=5)
            try:
                return run.forward()
            finally:
                run.close()

        h = 1e-9
        fg = (J_of(g + h, act) - J_of(g - h, act)) / (2 * h)
        fa = (J_of(g, act + h) - J_of(g, act - h)) / (2 * h)
        print(f"[reverse] dJ/dg {dg[0]!r} FD {fg!r}; dJ/dact {dact[0]!r} FD {fa!r}", flush=True)
        assert dg[0] == pytest.approx(fg, rel=1e-3)
        assert dact[0] == pytest.approx(fa, rel=1e-3)


def test_reverse_torch_wrapper():
    torch = pytest.importorskip("torch")
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed = 10, 6, 10, 6, 2
    obs = (M.OBS_MEAN_S, None)

    def prob(g, a):
        cfg, Q, V, Uo, _ = problem(lc_channel(Lx, Ly, ppc, g, a))
        return M, cfg, Q, V, Uo, K, seed, obs

    theta = torch.tensor([5e-4, 0.03], dtype=torch.float64, requires_grad=True)
    J = H.active_channel_reverse(theta, prob)
    J.backward()
    _, cfg, Q, V, Uo, _, _, _ = prob(5e-4, 0.03)
    J2, dg, dact = H.reverse_value_and_grad(M, cfg, Q, V, K, seed, obs, U0=Uo)
    assert J.detach().item() == J2
    # the adjoints of shared cell values are summed with atomics: two backward
    # passes agree to rounding, not bit for bit
    assert theta.grad[0].item() == pytest.approx(dg[0], rel=1e-12)
    assert theta.grad[1].item() == pytest.approx(dact[0], rel=1e-12)


@pytest.mark.parametrize("rtech", [16, 21])
@pytest.mark.parametrize("kind", ["band", "S"])
def test_reverse_equals_cpu_adjoint(rtech, kind):
    import adjoint_lc_ref as R
    torch = pytest.importorskip("torch")
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed = 16, 8, 20, 10, 21
    cfg, Q, V, Uo, _ = problem(lc_channel(Lx, Ly, ppc, 5e-4, 0.03, rtech=rtech))
    obs = (M.OBS_MEAN_S, None) if kind == "S" else (M.OBS_BAND_VELOCITY, (0, 1, 3, 5))
    print(f"[reverse] vs CPU adjoint 16x8x20 rtech={rtech} {kind} K={K}", flush=True)
    J, dg, dact = H.reverse_value_and_grad(M, cfg, Q, V, K, seed, obs, U0=Uo, ckpt=5)
    ocfg = dict(cfg)
    Jc, tg, ta, worst = R.run(ocfg, Q, V, Uo, K, seed, ("S", None) if kind == "S" else ("band", (3, 5)))
    assert worst <= 1.0, f"the CPU restatement left the oracle ({worst:.2f} x 1e-12)"
    Jc.backward()
    print(f"[reverse] J {J!r} cpu {float(Jc.detach())!r}; dJ/dg {dg[0]!r} cpu {float(tg.grad)!r}; "
          f"dJ/dact {dact[0]!r} cpu {float(ta.grad)!r}", flush=True)
    assert J == pytest.approx(float(Jc.detach()), rel=1e-12)
    assert dg[0] == pytest.approx(float(tg.grad), rel=REL)
    assert dact[0] == pytest.approx(float(ta.grad), rel=REL)
