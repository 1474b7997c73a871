"""GPU: the FP32-state form against the FP64 path (which the parity tests pin
to the oracle at 1e-12).  Same replicas, initial states and keyed draws in
both precisions; the trajectories separate (FP32 rounding moves particles
across cell boundaries), so the comparison is of observables:

  * temperature after 400 steps: relative error < 5e-3 (the 0.5% target);
  * total-momentum drift per particle: FP64 at round-off (< 1e-13), FP32 < 1e-7;
  * shear-wave viscosity: each precision within 4% of the analytic SRD value
    (Kikuchi et al. 2003) and FP32 within 0.5% + 3 standard errors of FP64;
  * active-nematic mean order S: relative error < 5e-3.
"""
import importlib

import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def V():
    U.package()
    return importlib.import_module("niamh_mpcd_amd.validate")


def test_fp32_shear_wave(V):
    M = U.package()
    print("[fp32] shear wave fp64", flush=True)
    w64 = V.shear_wave(M, M.FP64, nrep=8)
    print("[fp32] shear wave fp32", flush=True)
    w32 = V.shear_wave(M, M.FP32, nrep=8)
    th = V.srd_viscosity_3d(10.0, V.ROT130, 0.1)["nu"]
    print("[fp32]", w64["nu"], w64["nu_err"], w32["nu"], w32["nu_err"], th, w64["T"], w32["T"],
          w64["momentum_drift"], w32["momentum_drift"], flush=True)
    assert abs(w32["T"] - w64["T"]) / w64["T"] < 5e-3
    assert w64["momentum_drift"] < 1e-13
    assert w32["momentum_drift"] < 1e-7
    assert abs(w64["nu"] - th) / th < 0.04
    assert abs(w32["nu"] - th) / th < 0.04
    sig = (w64["nu_err"] ** 2 + w32["nu_err"] ** 2) ** 0.5
    assert abs(w32["nu"] - w64["nu"]) < 0.005 * w64["nu"] + 3.0 * sig


def test_fp32_nematic_order(V):
    M = U.package()
    print("[fp32] nematic", flush=True)
    n64 = V.nematic_order(M, M.FP64)
    n32 = V.nematic_order(M, M.FP32)
    print("[fp32]", n64["S"], n32["S"], flush=True)
    assert 0.5 < n64["S"] < 1.0
    assert abs(n32["S"] - n64["S"]) / n64["S"] < 5e-3
