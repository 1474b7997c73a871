"""GPU: a small ragged ensemble of MPC-AT+a channels (different L, ppc and
seeds in one context) develops start-up Poiseuille flow; the fitted
viscosities are positive, finite and agree across box sizes within the
statistical scatter expected at this size."""
import importlib

import numpy as np
import pytest

import gpu_util as U

pytestmark = pytest.mark.gpu


def test_small_channel_ensemble():
    M = U.package()
    B = importlib.import_module("niamh_mpcd_amd.batched")
    reps = [(10, 10), (12, 10), (14, 10), (12, 10), (10, 10), (14, 10)]
    print("[ensemble] 6 channels", flush=True)
    r = B.run(M, M.FP32, reps=reps, steps=1500, sample_from=300, every=50, time_steps=100, g=2e-3)
    nus = np.array([x["nu"] for x in r["per_replica"]])
    print("[ensemble]", nus, r["value"], flush=True)
    assert np.isfinite(nus).all() and (nus > 0.3).all() and (nus < 1.5).all()
    assert nus.std() / nus.mean() < 0.2
    assert r["value"] > 0
