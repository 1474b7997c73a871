"""GPU: the forward-mode tangent behind mpcd_gpu_tangent.

  * a run with the tangent on computes the same values bit for bit as the
    plain FP64 unfused path (LC active channel with anchoring, 3D nematic,
    MPC-AT+a channel, MPC-AT, SRD with phantom walls);
  * on a 2D SRD channel its dJ/dg equals the reverse-mode adjoint's;
  * on a 2D active nematic channel (bounce-back walls with homeotropic strong
    anchoring, dipole activity, body force) dJ/dg and dJ/d(activity) agree
    with central finite differences of the GPU's own observable with common
    random numbers, and with finite differences of the oracle on the CPU;
  * on a 3D MPC-AT+a channel dJ/dg agrees with finite differences;
  * the torch.autograd.Function wrapper returns the tangent's gradient.
"""
import importlib

import numpy as np
import pytest

import adjoint_ref as A
import cases as CS
import gpu_util as U

pytestmark = pytest.mark.gpu


def harness():
    U.package()
    return importlib.import_module("niamh_mpcd_amd.adjoint")


def lc_channel(Lx, Ly, ppc, g, act, rtech=16):
    """2D active nematic channel: periodic x, bounce-back phantom walls in y
    with homeotropic strong anchoring (MUN 1, MUT 0)"""
    return CS.case("tan", 2, (Lx, Ly), ppc, rtech=rtech, lc=1, steps=0, seed=0, grav=(g, 0, 0),
                   sp=CS.species(mfpot=10.0, tumble=0.5, rfc=0.01, act=act),
                   walls=CS.pbc_walls(0, Lx) + CS.solid_walls(1, Ly, MUN=1.0, MUT=0.0))


def problem(c, state_seed=5):
    cc = dict(c)
    cc["seed"] = state_seed
    Q, V, Uo, spid, avv = CS.initial_state(cc)
    cfg = U.gpu_cfg(c, float(c["ppc"]), avv)
    return cfg, Q, V, Uo, spid


def run(M, cfg, Q, V, Uo, spid, K, seed, tangent):
    """K steps on the unfused path: phase by phase (plain) or whole steps with
    the tangent on (which always run unfused)"""
    ctx = M.Context(cfg, Q.shape[0], seed=seed, precision=M.FP64)
    ctx.upload(0, Q, V, U=Uo if cfg["lc"] else None, spid=spid if len(cfg["species"]) > 1 else None)
    if tangent:
        ctx.tangent(True, dgrav=[1.0, 0.5, 0.0], dact=[1.0] * len(cfg["species"]))
    for _ in range(K):
        if tangent:
            ctx.step()
        else:
            ctx.bin()
            ctx.collide()
            ctx.stream()
    out = ctx.download(0)
    cells = ctx.cells(0)
    ctx.close()
    return out, cells


TAN_CASES = {
    "lc2d_channel": lambda: lc_channel(12, 8, 12, 5e-4, 0.03),
    "lc2d_channel_sig": lambda: lc_channel(12, 8, 12, 5e-4, 0.03, rtech=21),
    "lc3d_weak": lambda: CS.case("t", 3, (4, 4, 3), 10, rtech=17, lc=1, steps=0, seed=0, grav=(1e-3, 0, 0),
                                 sp=CS.species(mfpot=0.6)),
    "lc3d_gauss": lambda: CS.case("t", 3, (4, 4, 4), 10, rtech=3, lc=1, steps=0, seed=0, sp=CS.species(mfpot=12.0)),
    "rat3d_channel": lambda: CS.case("t", 3, (6, 5, 6), 8, rtech=3, steps=0, seed=0, grav=(1e-3, 0, 0), dt=0.05,
                                     walls=CS.pbc_walls(0, 6) + CS.pbc_walls(1, 5) + CS.solid_walls(2, 6)),
    "at3d_damp": lambda: CS.case("t", 3, (5, 4, 6), 7, rtech=2, steps=0, seed=0, sp=CS.species(damp=0.1),
                                 grav=(1e-3, 0, 0)),
    "srd2d_phantom": lambda: CS.case("t", 2, (10, 8), 10, rtech=0, steps=0, seed=0, grav=(1e-3, 0, 0),
                                     walls=CS.pbc_walls(0, 10) + CS.solid_walls(1, 8)),
}


@pytest.mark.parametrize("name", sorted(TAN_CASES))
def test_tangent_values_are_the_fp64_values(name):
    M = U.package()
    c = TAN_CASES[name]()
    cfg, Q, V, Uo, spid = problem(c)
    print(f"[tangent] values {name}", flush=True)
    plain, pcells = run(M, cfg, Q, V, Uo, spid, 6, 3, tangent=False)
    tan, tcells = run(M, cfg, Q, V, Uo, spid, 6, 3, tangent=True)
    for a, b, what in zip(plain, tan, ("Q", "V", "U")):
        assert np.array_equal(a, b), f"{name}: {what} differs with the tangent on"
    for k in ("pop", "vcm", "S", "dir"):
        assert np.array_equal(pcells[k], tcells[k]), f"{name}: cell {k} differs with the tangent on"


def test_tangent_equals_reverse_adjoint_srd():
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed = 12, 8, 10, 30, 41
    g, ang, band = 5e-4, 2.0, (3, 5)
    Q, V = A.initial(Lx, Ly, ppc, 7)
    cfg = A.channel_cfg(Lx, Ly, ppc, g, ang)
    print("[tangent] vs reverse adjoint", flush=True)
    J, dg, _ = H.value_and_grad(M, cfg, Q, V, K, seed, band, ckpt=10)
    Jt, (tx, ty) = H.tangent_value_and_grad(M, cfg, Q, V, K, seed, band,
                                            [dict(dgrav=[1.0, 0.0, 0.0]), dict(dgrav=[0.0, 1.0, 0.0])])
    print("[tangent]", J, Jt, dg[0], tx, dg[1], ty, flush=True)
    assert Jt == J
    assert tx == pytest.approx(dg[0], rel=1e-9)
    assert ty == pytest.approx(dg[1], rel=1e-6, abs=1e-12)


def fd_gpu(H, M, make, g, act, h, which, K, seed, band):
    def J_of(gg, aa):
        cfg, Q, V, Uo, _ = make(gg, aa)
        ctx = M.Context(cfg, Q.shape[0], seed=seed, precision=M.FP64)
        ctx.upload(0, Q, V, U=Uo if cfg["lc"] else None)
        j = 0.0
        for _ in range(K):
            ctx.step()
            Qs, Vs, _ = ctx.download(0)
            y = np.trunc(Qs[:, 1 if cfg["dim"] == 2 else 2])
            m = (y >= band[0]) & (y < band[1])
            j += Vs[m, 0].mean() / K
        ctx.close()
        return j

    if which == "g":
        return (J_of(g + h, act) - J_of(g - h, act)) / (2 * h)
    return (J_of(g, act + h) - J_of(g, act - h)) / (2 * h)


def test_tangent_matches_finite_differences_active_nematic():
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed, band = 24, 12, 12, 30, 8, (5, 7)
    g, act = 5e-4, 0.03

    def make(gg, aa):
        return problem(lc_channel(Lx, Ly, ppc, gg, aa))

    cfg, Q, V, Uo, _ = make(g, act)
    print("[tangent] active nematic channel", flush=True)
    J, (dg, da) = H.tangent_value_and_grad(M, cfg, Q, V, K, seed, band,
                                           [dict(dgrav=[1.0, 0.0, 0.0]), dict(dact=[1.0])], U0=Uo)
    fg = fd_gpu(H, M, make, g, act, 1e-9, "g", K, seed, band)
    fa = fd_gpu(H, M, make, g, act, 1e-9, "act", K, seed, band)
    print("[tangent]", J, dg, fg, da, fa, flush=True)
    assert da != 0.0
    assert dg == pytest.approx(fg, rel=1e-3)
    assert da == pytest.approx(fa, rel=1e-3)


def test_tangent_matches_cpu_oracle_finite_differences():
    """the tiny box of the FP64 CPU check: central differences of the
    oracle's own run (Philox draws, as the GPU's) against the GPU tangent"""
    import pyoracle as P
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed, band = 10, 6, 10, 8, 12, (2, 4)
    g, act = 5e-4, 0.03

    def make(gg, aa):
        return problem(lc_channel(Lx, Ly, ppc, gg, aa), state_seed=9)

    def J_cpu(gg, aa):
        cfg, Q, V, Uo, _ = make(gg, aa)
        s = P.System(cfg, Q, V, Uo)
        d = P.Draws(P.DRAW_PHILOX, seed)
        j = 0.0
        for _ in range(K):
            s.step(d)
            Qs, Vs, _ = s.state()
            m = (np.trunc(Qs[:, 1]) >= band[0]) & (np.trunc(Qs[:, 1]) < band[1])
            j += Vs[m, 0].mean() / K
        return j

    cfg, Q, V, Uo, _ = make(g, act)
    print("[tangent] cpu oracle", flush=True)
    J, (dg, da) = H.tangent_value_and_grad(M, cfg, Q, V, K, seed, band,
                                           [dict(dgrav=[1.0, 0.0, 0.0]), dict(dact=[1.0])], U0=Uo)
    Jc = J_cpu(g, act)
    h = 1e-7
    fg = (J_cpu(g + h, act) - J_cpu(g - h, act)) / (2 * h)
    fa = (J_cpu(g, act + h) - J_cpu(g, act - h)) / (2 * h)
    print("[tangent]", J, Jc, dg, fg, da, fa, flush=True)
    assert J == pytest.approx(Jc, rel=1e-10)
    assert dg == pytest.approx(fg, rel=1e-4)
    assert da == pytest.approx(fa, rel=1e-4)


def test_tangent_matches_finite_differences_mpcat_a():
    M = U.package()
    H = harness()
    K, seed, band = 30, 4, (3, 5)
    g = 1e-3

    def make(gg, _aa=None):
        c = CS.case("t", 3, (10, 8, 8), 8, rtech=3, steps=0, seed=0, grav=(gg, 0, 0), dt=0.05,
                    walls=CS.pbc_walls(0, 10) + CS.pbc_walls(1, 8) + CS.solid_walls(2, 8))
        return problem(c)

    cfg, Q, V, Uo, _ = make(g)
    print("[tangent] mpc-at+a channel", flush=True)
    J, (dg,) = H.tangent_value_and_grad(M, cfg, Q, V, K, seed, band, [dict(dgrav=[1.0, 0.0, 0.0])], axis=2)
    fg = fd_gpu(H, M, make, g, None, 1e-9, "g", K, seed, band)
    print("[tangent]", J, dg, fg, flush=True)
    assert dg == pytest.approx(fg, rel=1e-3)


def test_active_torch_wrapper():
    torch = pytest.importorskip("torch")
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed, band = 10, 6, 10, 6, 2, (2, 4)

    def prob(g, a):
        cfg, Q, V, Uo, _ = problem(lc_channel(Lx, Ly, ppc, g, a))
        return M, cfg, Q, V, Uo, K, seed, band

    theta = torch.tensor([5e-4, 0.03], dtype=torch.float64, requires_grad=True)
    J = H.active_channel_velocity(theta, prob)
    J.backward()
    _, cfg, Q, V, Uo, _, _, _ = prob(5e-4, 0.03)
    J2, (dg, da) = H.tangent_value_and_grad(M, cfg, Q, V, K, seed, band,
                                            [dict(dgrav=[1.0, 0.0, 0.0]), dict(dact=[1.0])], U0=Uo)
    assert J.detach().item() == J2
    assert theta.grad[0].item() == dg and theta.grad[1].item() == da


@pytest.mark.parametrize("mfpot", [12.0, 3.0])
def test_tangent_matches_finite_differences_3d_nematic(mfpot):
    """3D active nematic channel (eig3 director, Gaussian or Metropolis
    Maier-Saupe sampling, planar anchoring at phantom walls in z)"""
    M = U.package()
    H = harness()
    K, seed, band = 12, 6, (2, 4)
    g, act = 1e-3, 0.04

    def make(gg, aa):
        c = CS.case("t", 3, (6, 5, 6), 10, rtech=16, lc=1, steps=0, seed=0, grav=(gg, 0, 0),
                    sp=CS.species(mfpot=mfpot, tumble=0.7, rfc=0.01, act=aa),
                    walls=CS.pbc_walls(0, 6) + CS.pbc_walls(1, 5) + CS.solid_walls(2, 6, MUN=0.0, MUT=1.0))
        return problem(c, state_seed=3)

    cfg, Q, V, Uo, _ = make(g, act)
    print(f"[tangent] 3D nematic U={mfpot}", flush=True)
    J, (dg, da) = H.tangent_value_and_grad(M, cfg, Q, V, K, seed, band,
                                           [dict(dgrav=[1.0, 0.0, 0.0]), dict(dact=[1.0])], U0=Uo, axis=2)
    fg = fd_gpu(H, M, make, g, act, 1e-9, "g", K, seed, band)
    fa = fd_gpu(H, M, make, g, act, 1e-9, "act", K, seed, band)
    print("[tangent]", J, dg, fg, da, fa, flush=True)
    assert dg == pytest.approx(fg, rel=1e-3)
    assert da == pytest.approx(fa, rel=1e-3, abs=1e-6 * abs(dg))
