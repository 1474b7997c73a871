"""GPU: the pathwise adjoint behind mpcd_gpu_step_backward.

  * on a 12 x 8 2D SRD channel it equals the FP64 CPU adjoint of the oracle
    (tests/adjoint_ref.py) on the same trajectory (the GPU's Philox draws are
    the oracle's), for dJ/dg and dJ/d(rotation angle);
  * on a 48 x 24 channel with phantom walls (ghost-corrected VCM), 60 steps and
    a checkpoint every 20 steps it agrees with central finite differences of
    the GPU's own observable with common random numbers;
  * the torch.autograd.Function wrapper returns the same gradient.
"""
import importlib

import numpy as np
import pytest

import adjoint_ref as A
import gpu_util as U

pytestmark = pytest.mark.gpu


def harness():
    U.package()
    return importlib.import_module("niamh_mpcd_amd.adjoint")


def test_gpu_adjoint_equals_cpu_adjoint():
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed = 12, 8, 10, 30, 41
    g, ang, band = 5e-4, 2.0, (3, 5)
    Q, V = A.initial(Lx, Ly, ppc, 7)
    cfg = A.channel_cfg(Lx, Ly, ppc, g, ang)
    Jc, steps = A.forward(cfg, Q, V, K, seed, band)
    dgc, dac = A.backward(cfg, steps, K)
    print("[adjoint] gpu vs cpu", flush=True)
    Jg, dgg, dag = H.value_and_grad(M, cfg, Q, V, K, seed, band, ckpt=10)
    print("[adjoint]", Jc, Jg, dgc[0], dgg[0], dac, dag, flush=True)
    assert Jg == pytest.approx(Jc, rel=1e-10)
    assert dgg[0] == pytest.approx(dgc[0], rel=1e-8)
    assert dgg[1] == pytest.approx(dgc[1], rel=1e-6, abs=1e-12)
    assert dag == pytest.approx(dac, rel=1e-8)


def test_gpu_adjoint_matches_finite_differences():
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed = 48, 24, 20, 60, 5
    g, ang, band = 5e-4, 2.0, (11, 13)
    Q, V = A.initial(Lx, Ly, ppc, 3)

    def cfg(gg, aa):
        return A.channel_cfg(Lx, Ly, ppc, gg, aa, phantom=1)

    print("[adjoint] fd", flush=True)
    J, dg, da = H.value_and_grad(M, cfg(g, ang), Q, V, K, seed, band, ckpt=20)
    h = 1e-9

    def J_of(gg, aa):
        run = H.AdjointRun(M, cfg(gg, aa), Q, V, K, seed, band, ckpt=20)
        j = run.forward()
        run.close()
        return j

    fd_g = (J_of(g + h, ang) - J_of(g - h, ang)) / (2 * h)
    fd_a = (J_of(g, ang + h) - J_of(g, ang - h)) / (2 * h)
    print("[adjoint]", J, dg[0], fd_g, da, fd_a, flush=True)
    assert dg[0] == pytest.approx(fd_g, rel=1e-3)
    assert da == pytest.approx(fd_a, rel=1e-3)


def test_torch_wrapper():
    torch = pytest.importorskip("torch")
    M = U.package()
    H = harness()
    Lx, Ly, ppc, K, seed = 12, 8, 10, 20, 9
    band = (3, 5)
    Q, V = A.initial(Lx, Ly, ppc, 2)

    def problem(g, a):
        return H.AdjointRun(M, A.channel_cfg(Lx, Ly, ppc, g, a), Q, V, K, seed, band, ckpt=10)

    theta = torch.tensor([5e-4, 2.0], dtype=torch.float64, requires_grad=True)
    J = H.centreline_velocity(theta, problem)
    J.backward()
    J2, dg, da = H.value_and_grad(M, A.channel_cfg(Lx, Ly, ppc, 5e-4, 2.0), Q, V, K, seed, band, ckpt=10)
    assert J.detach().item() == pytest.approx(J2, rel=1e-12)
    assert theta.grad[0].item() == pytest.approx(dg[0], rel=1e-12)
    assert theta.grad[1].item() == pytest.approx(da, rel=1e-12)
