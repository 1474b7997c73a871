"""CPU: the FP64 CPU adjoint of the 2D active nematic step
(tests/adjoint_lc_ref.py, torch autograd over a restatement of the step).

Its forward run equals the oracle's state after every step (1e-12), and its
gradients of the time-averaged band velocity and mean order parameter with
respect to the body force and the activity agree with central finite
differences of the oracle's own observable taken with common random numbers
(the same counter-based draws)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import adjoint_lc_ref as R  # noqa: E402
import cases as CS  # noqa: E402
import pyoracle as P  # noqa: E402


def channel(Lx, Ly, ppc, g, act, rtech=16, state_seed=5):
    c = CS.case("cpu", 2, (Lx, Ly), ppc, rtech=rtech, lc=1, steps=0, seed=state_seed, grav=(g, 0, 0),
                sp=CS.species(mfpot=10.0, tumble=0.5, rfc=0.01, act=act),
                walls=CS.pbc_walls(0, Lx) + CS.solid_walls(1, Ly, MUN=1.0, MUT=0.0))
    Q, V, U, spid, avv = CS.initial_state(c)
    return CS.oracle_cfg(c, float(ppc), avv), Q, V, U


def J_oracle(cfg, Q, V, U, K, seed, obs):
    pb = R.Problem(cfg)
    s = P.System(cfg, Q, V, U)
    d = P.Draws(P.DRAW_PHILOX, seed)
    j = 0.0
    for _ in range(K):
        s.step(d)
        Qs, Vs, Us = s.state()
        if obs[0] == "band":
            j += float(pb.band_velocity(R._f(Qs), R._f(Vs), *obs[1])) / K
        else:
            j += float(pb.mean_S(R._f(Qs), R._f(Us))) / K
    return j


@pytest.mark.parametrize("rtech", [16, 21])
@pytest.mark.parametrize("obs", [("band", (2, 4)), ("S", None)])
def test_cpu_lc_adjoint_matches_oracle_fd(rtech, obs):
    Lx, Ly, ppc, K, seed = 10, 6, 10, 5, 12
    g, act = 5e-4, 0.03
    cfg, Q, V, U = channel(Lx, Ly, ppc, g, act, rtech)
    J, tg, ta, worst = R.run(cfg, Q, V, U, K, seed, obs)
    assert worst <= 1.0, f"the torch step differs from the oracle's ({worst:.2f} x 1e-12)"
    J.backward()
    assert float(J) == pytest.approx(J_oracle(cfg, Q, V, U, K, seed, obs), rel=1e-11)
    h = 1e-7
    cp, _, _, _ = channel(Lx, Ly, ppc, g + h, act, rtech)
    cm, _, _, _ = channel(Lx, Ly, ppc, g - h, act, rtech)
    fg = (J_oracle(cp, Q, V, U, K, seed, obs) - J_oracle(cm, Q, V, U, K, seed, obs)) / (2 * h)
    cp, _, _, _ = channel(Lx, Ly, ppc, g, act + h, rtech)
    cm, _, _, _ = channel(Lx, Ly, ppc, g, act - h, rtech)
    fa = (J_oracle(cp, Q, V, U, K, seed, obs) - J_oracle(cm, Q, V, U, K, seed, obs)) / (2 * h)
    print(obs[0], rtech, float(J), float(tg.grad), fg, float(ta.grad), fa)
    assert float(tg.grad) == pytest.approx(fg, rel=1e-4)
    assert float(ta.grad) == pytest.approx(fa, rel=1e-4)
