"""mpcd_gpu_bin_noalloc, the binning entry point that allocates nothing,
against mpcd_gpu_bin on the same positions: binin (no shift) and the shifted
relinking of that binning, on two calls in a row, the second with fewer
particles and cells (the shared scratch array keeps the first call's
contents).  Integers, so bit-exact."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
import gpu_util as U  # noqa: E402
from test_gpu_parity import _random_box  # noqa: E402


@pytest.mark.gpu
def test_bin_noalloc_matches_bin():
    M = U.package()
    hip = U.hip_runtime()
    for dim, L, ppc, seed in [(3, (12, 10, 8), 10, 3), (2, (9, 7), 6, 4)]:
        c, Q, V, _ = _random_box(dim, L, ppc, 0, seed=seed)
        N = Q.shape[0]
        cfg = U.gpu_cfg(c, float(ppc), np.zeros(3))
        print(f"[gpu bin_noalloc] dim={dim} L={L} N={N}", flush=True)
        ctx = M.Context(cfg, N, seed=seed)
        ctx.upload(0, Q, V)
        nc = ctx.ncells[0]
        cell0, cnt0, perm0 = ctx.bin(None, shifted=0, outputs=True)  # binin after the upload
        shift = np.random.default_rng(seed).uniform(size=3)
        cell1, cnt1, perm1 = ctx.bin(dict(shift=shift, lc_stride=0), shifted=1, outputs=True)
        ctx.close()

        q = U.DevBuf(hip, np.ascontiguousarray(Q, np.float64))
        cell = U.DevBuf(hip, np.full(N, -7, np.int32))
        count = U.DevBuf(hip, np.full(nc, -7, np.int32))
        perm = U.DevBuf(hip, np.full(N, -7, np.int32))
        M.bin_noalloc(cfg, N, q.addr, None, None, None, cell.addr, count.addr, perm.addr)
        np.testing.assert_array_equal(cell.read(), cell0, err_msg="binin cells")
        np.testing.assert_array_equal(count.read(), cnt0, err_msg="binin populations")
        np.testing.assert_array_equal(perm.read(), perm0, err_msg="binin order")

        pperm, pcell = U.DevBuf(hip, perm.read()), U.DevBuf(hip, cell.read())
        M.bin_noalloc(cfg, N, q.addr, shift[:dim], pperm.addr, pcell.addr, cell.addr, count.addr, perm.addr)
        np.testing.assert_array_equal(cell.read(), cell1, err_msg="shifted cells")
        np.testing.assert_array_equal(count.read(), cnt1, err_msg="shifted populations")
        np.testing.assert_array_equal(perm.read(), perm1, err_msg="shifted list order")
        for b in (q, cell, count, perm, pperm, pcell):
            b.free()
