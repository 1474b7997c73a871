"""GPU: upload and download through an array of structures laid out like the
reference's particleMPC (SRDclss.h:34-44: Q, V, U, T double[3], S_flag and
SPID int, q double, next and previous pointers; 128 B).  The state the
library gets from the records is the packed upload's bit for bit, a download
writes Q, V and U only, and offsets outside a record are rejected."""
import numpy as np
import pytest

import cases as CS
import gpu_util as U

pytestmark = pytest.mark.gpu

PARTICLE_MPC = np.dtype([("Q", "<f8", 3), ("V", "<f8", 3), ("U", "<f8", 3), ("T", "<f8", 3), ("S_flag", "<i4"),
                         ("SPID", "<i4"), ("q", "<f8"), ("next", "<u8"), ("previous", "<u8")])


def test_particle_mpc_records():
    assert PARTICLE_MPC.itemsize == 128
    M = U.package()
    c = CS.case("aos", 2, (10, 8), 10, rtech=16, lc=1, steps=0, seed=4, grav=(5e-4, 0, 0),
                sp=CS.species(mfpot=10.0, tumble=0.5, act=0.03),
                walls=CS.pbc_walls(0, 10) + CS.solid_walls(1, 8, MUN=1.0, MUT=0.0))
    c["species"] = [CS.species(mfpot=10.0, tumble=0.5, act=0.03), CS.species(mass=1.5, mfpot=8.0, act=0.0)]
    cc = dict(c)
    cc["seed"] = 21
    Q, V, Uo, spid, avv = CS.initial_state(cc)
    cfg = U.gpu_cfg(c, 10.0, avv)
    n = Q.shape[0]
    rec = np.zeros(n, PARTICLE_MPC)
    rec["Q"], rec["V"], rec["U"], rec["SPID"] = Q, V, Uo, spid
    rec["T"] = 7.0
    rec["S_flag"] = 1
    rec["q"] = -3.5
    rec["next"] = np.arange(n, dtype=np.uint64) * 128
    before = rec.copy()
    print("[aos] packed vs records", flush=True)
    a = M.Context(cfg, n, seed=8, precision=M.FP64)
    b = M.Context(cfg, n, seed=8, precision=M.FP64)
    a.upload(0, Q, V, U=Uo, spid=spid)
    b.upload_aos(0, rec, q="Q", v="V", u="U", spid="SPID")
    for _ in range(4):
        a.step()
        b.step()
    Qa, Va, Ua = a.download(0)
    b.download_aos(0, rec, q="Q", v="V", u="U")
    assert np.array_equal(rec["Q"], Qa) and np.array_equal(rec["V"], Va) and np.array_equal(rec["U"], Ua)
    for f in ("T", "S_flag", "SPID", "q", "next", "previous"):
        assert np.array_equal(rec[f], before[f]), f
    # a field that would run past the record is refused before any transfer
    with pytest.raises(M.MPCDError):
        b.upload_aos(0, rec, q="previous", v="V")
    a.close()
    b.close()
