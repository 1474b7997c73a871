"""Transport check of the collision operators the batched form runs: the
kinematic viscosity measured from the decay of a transverse shear wave in
periodic boxes against the analytic kinetic-plus-collisional value with a
random grid shift -- SRD ARBAXIS (Kikuchi et al. 2003) and MPC-AT+a
(Noguchi & Gompper 2008, the formulas of theory_trans, init.c:849-872) at
three densities.  These are statistical checks of the GPU kernels' physics,
not parity (the kernels are pinned to the reference bit for bit elsewhere):
the fit's noise is about 1 % here.

SRD agrees with its theory to about 1 %.  MPC-AT+a comes out below the
mean-field formula by an amount that shrinks with the density: on MI355X
0.878 of it at 5 particles per cell, 0.928 at 10 and 0.975 at 20, and the
64-replica Poiseuille channels of the batched ensemble (an independent
measurement, with walls) give the same deficit.  The kernels match the
reference's own MPC-AT+a step bit for bit on the recorded fixtures, so the
deficit lies between the reference's operator and its formula, not in this
port.  The test holds the measured ratio to that band, so a kernel error of a
few percent (a wrong inertia term, a dropped thermal draw) fails it."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(__file__))
import gpu_util as U  # noqa: E402


# measured / theory ratios of MPC-AT+a and the band each must fall in
MPCAT_A_BAND = {5: (0.85, 0.91), 10: (0.90, 0.96), 20: (0.95, 1.00)}


@pytest.mark.gpu
@pytest.mark.parametrize("ppc", [5, 10, 20])
def test_mpcat_a_viscosity_matches_theory(ppc):
    import importlib
    M = U.package()
    V = importlib.import_module("niamh_mpcd_amd.validate")
    print(f"[gpu viscosity] MPC-AT+a shear wave ppc={ppc}", flush=True)
    w = V.shear_wave(M, M.FP32, L=(32, 32, 16), ppc=ppc, nrep=8, steps=240, every=10, amp=0.3, dt=0.05, rtech=3)
    th = V.mpcat_a_viscosity(float(ppc), 0.05)["nu"]
    print(f"  nu = {w['nu']:.4f} +- {w['nu_err']:.4f}, theory {th:.4f}, ratio {w['nu'] / th:.4f}", flush=True)
    lo, hi = MPCAT_A_BAND[ppc]
    assert lo < w["nu"] / th < hi


@pytest.mark.gpu
def test_srd_viscosity_matches_theory():
    import importlib
    M = U.package()
    V = importlib.import_module("niamh_mpcd_amd.validate")
    print("[gpu viscosity] SRD ARBAXIS 130 deg shear wave ppc=10", flush=True)
    w = V.shear_wave(M, M.FP32, L=(32, 32, 16), ppc=10, nrep=8, steps=200, every=10, amp=0.3)
    th = V.srd_viscosity_3d(10.0, V.ROT130, 0.1)["nu"]
    print(f"  nu = {w['nu']:.4f} +- {w['nu_err']:.4f}, theory {th:.4f}, ratio {w['nu'] / th:.4f}", flush=True)
    assert abs(w["nu"] / th - 1.0) < 0.03
