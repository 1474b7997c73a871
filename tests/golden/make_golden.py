"""Record golden fixtures from the reference's own C step.

Builds oracle/_ref/ref_harness (the reference's sources compiled in place by
oracle/Makefile, linked with the draw-injecting oracle/ref_rand.c), runs every
case of cases.py through it, and stores what the reference computed:

  tests/golden/<case>.npz:
    Q0, V0, U0, spid, avv    initial state given to the reference
    gndnst, pbc              the reference's GnDNST (init.c:1002) and XYZPBC
    Q, V, U   [steps, N, 3]  state after every step (particle-index order)
    pop       [steps, nc]    per-cell populations (compact cell layout)
    order     [steps, N]     particles in list order (cells x-major)
    kbt       [steps]        KBTNOW after every step
    draws     [steps, 3]     cumulative (uniform, gauss, unit-vector) draws consumed
    S, DIR, VCM              per-cell values after every step

  tests/golden/ckpt_<case>.dat: the reference's own checkpoint file
    (checkpoint(), pout.c:2325-2404) after the case's steps, for CKPT_CASES,
    with simSteps 50 and warmupSteps 5
  tests/golden/flow.npz: FLOW_<case> [nc, 3], the reference's running flow
    sum (localFLOW + sumFLOW, mpc.c:226-246) after the case's steps, from a
    second run of the same case with flow output on (compact cell layout)

The compact cell layout drops the reference's always-empty z = 1 slab in 2D.
Run:  python tests/golden/make_golden.py   (needs /root/reference; the
fixtures it writes are committed, the reference is not needed afterwards).
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import cases as C  # noqa: E402
import pyoracle  # noqa: E402


def wall_line(w):
    g = lambda k, d: w.get(k, d)  # noqa: E731
    v = (list(g("A", (0, 0, 0))) + list(g("Q", (0, 0, 0))) +
         [g("R", 0.0), g("DN", 0.0), g("DT", 0.0), g("DVN", 0.0), g("DVT", 0.0)] + list(g("DVxyz", (0, 0, 0))) +
         [g("MVN", 1.0), g("MVT", 1.0), g("MUN", 1.0), g("MUT", 1.0)] + list(g("MUxyz", (1, 1, 1))) +
         list(g("DUxyz", (0, 0, 0))) + list(g("V", (0, 0, 0))) +
         [g("INV", 0), g("PHANTOM", 0), g("COLL_TYPE", 1)])
    assert len(v) == 30
    return "wall " + " ".join(repr(float(x)) for x in v)


def write_case(c, Q, V, U, spid, avv, d):
    N = Q.shape[0]
    lines = [f"dim {c['dim']}", "L {} {} {}".format(*c["L"]), f"N {N}", f"steps {c['steps']}",
             f"rtech {c['rtech']}", f"lc {c['lc']}", f"tstech {c['tstech']}", f"galinv {c['galinv']}",
             f"dt {c['dt']!r}", f"kbt {c['kbt']!r}", f"rotang {c['rotang']!r}", f"tau {c['tau']!r}",
             "grav {!r} {!r} {!r}".format(*[float(x) for x in c["grav"]]),
             "avv {!r} {!r} {!r}".format(*[float(x) for x in avv]), f"seed {c['seed']}"]
    if c.get("zeronetmom"):
        lines.append(f"zeronetmom {c['zeronetmom']}")
    for sp in c["species"]:
        lines.append("species " + " ".join(repr(float(sp[k])) for k in (
            "mass", "rfc", "tumble", "chihi", "act", "mfpot", "damp", "sigwidth", "sigpos", "minactratio",
            "vdist", "len")))
    for w in c["walls"]:
        lines.append(wall_line(w))
        if "P" in w or "B" in w or "ABS" in w:
            # a curved or wavy surface (bc.c:55-150): exponents, wavy amplitude and wavenumbers, |.| terms
            v = list(w.get("P", (1, 1, 1, 1))) + list(w.get("B", (0, 0, 0))) + [w.get("ABS", 0)]
            lines.append("shape " + " ".join(repr(float(x)) for x in v))
    cp = os.path.join(d, "case.txt")
    sp_ = os.path.join(d, "state.bin")
    with open(cp, "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(sp_, "wb") as f:
        f.write(np.hstack([Q, V, U]).astype("<f8").tobytes())
        f.write(np.asarray(spid, "<i4").tobytes())
    return cp, sp_


def compact_cells(arr, c):
    """reference layout over XYZ_P1 -> compact layout (drop the z = 1 slab in 2D)"""
    if c["dim"] == 3:
        return arr
    nx, ny = c["L"][0] + 1, c["L"][1] + 1
    a = arr.reshape((nx, ny, 2) + arr.shape[1:])
    assert np.all(a[:, :, 1] == 0) or a.dtype != np.int32
    return a[:, :, 0].reshape((nx * ny,) + arr.shape[1:])


def parse(path, c, N):
    dim = c["dim"]
    nref = (c["L"][0] + 1) * (c["L"][1] + 1) * ((c["L"][2] + 1) if dim == 3 else 2)
    b = open(path, "rb").read()
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(b, dtype=dt, count=n, offset=off)
        off += a.nbytes
        return a

    gnd = take("<f8", 1)[0]
    pbc = take("<i4", 3).copy()
    recs = []
    while off < len(b):
        tag = take("<i4", 1)[0]
        if tag == -2:
            recs.append(dict(tag=tag, FLOW=compact_cells(take("<f8", nref * 3).reshape(nref, 3).copy(), c)))
            break
        kbt = take("<f8", 1)[0]
        cnt = take("<u8", 3).copy()
        st = take("<f8", N * 9).reshape(N, 9)
        pop = take("<i4", nref).copy()
        order = take("<i4", N).copy()
        cellv = take("<f8", nref * 7).reshape(nref, 7)
        recs.append(dict(tag=tag, kbt=kbt, cnt=cnt, Q=st[:, 0:3].copy(), V=st[:, 3:6].copy(), U=st[:, 6:9].copy(),
                         pop=compact_cells(pop, c), order=order, S=compact_cells(cellv[:, 0].copy(), c),
                         DIR=compact_cells(cellv[:, 1:4].copy(), c), VCM=compact_cells(cellv[:, 4:7].copy(), c)))
    return gnd, pbc, recs


def record(c, outdir):
    Q, V, U, spid, avv = C.initial_state(c)
    N = Q.shape[0]
    with tempfile.TemporaryDirectory() as d:
        cp, sp_ = write_case(c, Q, V, U, spid, avv, d)
        out = os.path.join(d, "out.bin")
        subprocess.run([pyoracle.REF_HARNESS, cp, sp_, out], check=True)
        gnd, pbc, recs = parse(out, c, N)
    steps = recs[1:]
    assert recs[0]["tag"] == -1 and len(steps) == c["steps"]
    np.savez_compressed(
        os.path.join(outdir, c["name"] + ".npz"),
        case=json.dumps(c), Q0=Q, V0=V, U0=U, spid=spid, avv=avv, gndnst=gnd, pbc=pbc,
        order0=recs[0]["order"], pop0=recs[0]["pop"],
        Q=np.stack([r["Q"] for r in steps]), V=np.stack([r["V"] for r in steps]),
        U=np.stack([r["U"] for r in steps]), pop=np.stack([r["pop"] for r in steps]),
        order=np.stack([r["order"] for r in steps]), kbt=np.array([r["kbt"] for r in steps]),
        draws=np.stack([r["cnt"] for r in steps]), S=np.stack([r["S"] for r in steps]),
        DIR=np.stack([r["DIR"] for r in steps]), VCM=np.stack([r["VCM"] for r in steps]))
    return N


def record_flow(c):
    """FLOW after the case's steps, from a run with the flow output on"""
    Q, V, U, spid, avv = C.initial_state(c)
    with tempfile.TemporaryDirectory() as d:
        cp, sp_ = write_case(c, Q, V, U, spid, avv, d)
        out = os.path.join(d, "out.bin")
        subprocess.run([pyoracle.REF_HARNESS, cp, sp_, out, "flow"], check=True)
        _, _, recs = parse(out, c, Q.shape[0])
    assert recs[-1]["tag"] == -2 and len(recs) == c["steps"] + 2
    return recs[-1]["FLOW"]


CKPT_CASES = ["rat3d_channel", "lc2d_channel", "rat2d_2sp"]


def record_ckpt(c):
    """the reference's checkpoint file after the case's steps"""
    Q, V, U, spid, avv = C.initial_state(c)
    with tempfile.TemporaryDirectory() as d:
        cp, sp_ = write_case(c, Q, V, U, spid, avv, d)
        with open(cp, "a") as f:
            f.write("simsteps 50\nwarmup 5\n")
        out = os.path.join(d, "out.bin")
        subprocess.run([pyoracle.REF_HARNESS, cp, sp_, out, "ckpt"], check=True)
        return open(out + ".ckpt").read()


def main():
    pyoracle.build_ref()
    names = sys.argv[1:]
    for c in C.CASES:
        if names and c["name"] not in names:
            continue
        N = record(c, HERE)
        print(f"{c['name']}: N={N} steps={c['steps']}")
    if not names or "ckpt" in names:
        for name in CKPT_CASES:
            c = next(x for x in C.CASES if x["name"] == name)
            with open(os.path.join(HERE, f"ckpt_{name}.dat"), "w") as f:
                f.write(record_ckpt(c))
        print("ckpt: " + ", ".join(CKPT_CASES))
    if not names or "flow" in names:
        np.savez_compressed(os.path.join(HERE, "flow.npz"), **{"FLOW_" + c["name"]: record_flow(c) for c in C.CASES})
        print("flow: FLOW of every case")


if __name__ == "__main__":
    main()
