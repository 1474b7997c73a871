"""The C host (niamh-mpcd_amd/host): reading of the reference's input.json.

CPU: `mpcd_host --dump-params` prints the parameters it read without
touching the GPU.  The checks cover the reference's defaults
(read.c:1058-1114, 1153-1440, 1609-1680), the aInv -> A inversion, the
domainWalls override (read.c:1441-1583), the anchoring overrides, species
interaction masks, densities, repeated "comment" keys and the options this
path rejects.  GPU: the host runs a periodic box, whose momentum the body force sets in
closed form, and a channel between bounce-back walls."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import gpu_util as U


def host():
    M = U.package()
    path = os.path.join(os.path.dirname(M.LIB_PATH), "mpcd_host")
    if not os.path.exists(path):
        pytest.skip("C host not built (run __graft_entry__.build())")
    return path


def run_dump(tmp_path, cfg, text=None):
    f = tmp_path / "input.json"
    f.write_text(text if text is not None else json.dumps(cfg))
    r = subprocess.run([host(), str(f), "--dump-params"], capture_output=True, text=True)
    return r


def dump(tmp_path, cfg=None, text=None):
    r = run_dump(tmp_path, cfg, text)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_defaults_and_duplicate_keys(tmp_path):
    text = '''{
        "comment": "repeated comment keys, as in the reference's sample inputs",
        "comment": "the first of each key is used",
        "domain": [16, 12, 8],
        "species": [{"comment": "x", "pop": 1000}]
    }'''
    d = dump(tmp_path, text=text)
    assert d["dim"] == 3 and d["L"] == [16, 12, 8]
    assert (d["kbt"], d["dt"], d["tau"], d["rotang"]) == (1.0, 0.1, 0.5, 1.570796)
    assert (d["rtech"], d["lc"], d["tstech"], d["galinv"], d["simSteps"], d["warmUp"]) == (2, 0, 0, 1, 2000, 0)
    assert (d["flowOut"], d["velOut"], d["dirSOut"], d["densOut"]) == (0, 0, 0, 0)  # read.c:1732-1738
    sp = d["species"][0]
    assert sp["pop"] == 1000 and sp["mass"] == 1.0 and sp["vdist"] == 0
    assert (sp["rfc"], sp["tumble"], sp["chihi"], sp["act"], sp["mfpot"]) == (0.01, 2.0, 0.5, 0.05, 10.0)
    assert (sp["sigwidth"], sp["sigpos"], sp["minactratio"], sp["damp"]) == (1.0, 1.0, 0.0, 0.0)
    # no BC list: periodic box walls (domainWalls override)
    assert len(d["walls"]) == 6
    for k, w in enumerate(d["walls"]):
        ax, side = divmod(k, 2)
        assert w["A"][ax] == (-1.0 if side else 1.0) and w["DN"] == [16, 12, 8][ax]
        assert w["MVN"] == 1.0 and w["MVT"] == 1.0 and w["PHANTOM"] == 0


def test_density_solid_domain_walls_and_collop(tmp_path):
    d = dump(tmp_path, dict(domain=[10, 6], rTech=0, collOp=3, domainWalls=2, seed=5,
                            species=[dict(dens=7.0, mass=2.0, vDist=4)]))
    assert d["dim"] == 2 and d["rtech"] == 3  # both set: the later tag, collOp, wins (cJson.c:265-279)
    assert d["seed"] == 5 and d["species"][0]["pop"] == 420 and d["species"][0]["mass"] == 2.0
    assert len(d["walls"]) == 4
    for w in d["walls"]:
        assert w["DN"] == 0.0 and w["MVN"] == -1.0 and w["MVT"] == -1.0 and w["PHANTOM"] == 1
        assert w["MUxyz"] == [1.0, 1.0, 1.0]  # read.c:1558-1559


def test_bc_entries(tmp_path):
    bc = [dict(aInv=[0, 0.5, 0], Q=[0, 0, 0], P=[1, 1, 1, 1], R=0, DN=0, MVN=-1, MVT=-1, phantom=1,
               homeotropic=1, V=[0.2, 0, 0], interSRD=[1, 0]),
          dict(aInv=[0, -1, 0], Q=[0, 9, 0], P=[1, 1, 1, 1], planar=1, inv=1)]
    d = dump(tmp_path, dict(domain=[9, 9], BC=bc, species=[dict(pop=10), dict(pop=20)]))
    w0, w1 = d["walls"]
    assert w0["A"] == [0.0, 2.0, 0.0] and w0["MUN"] == 1.0 and w0["MUT"] == 0.0 and w0["V"] == [0.2, 0.0, 0.0]
    assert w0["PHANTOM"] == 1 and w0["inter_mask"] == 1
    # absent keys take the reference's defaults: R 2, DN 1, MVN = MVT = 1
    assert w1["A"] == [0.0, -1.0, 0.0] and w1["R"] == 2.0 and w1["DN"] == 1.0 and w1["MVN"] == 1.0
    assert w1["MUN"] == 0.0 and w1["MUT"] == 1.0 and w1["INV"] == 1 and w1["inter_mask"] == 3
    # an absent MUxyz is 1 on every axis (read.c:1344-1348), a given one is taken as it is
    assert w0["MUxyz"] == [1.0, 1.0, 1.0] and w1["MUxyz"] == [1.0, 1.0, 1.0]
    bc[1]["MUxyz"] = [0.5, -1.0, 1.0]
    assert dump(tmp_path, dict(domain=[9, 9], BC=bc))["walls"][1]["MUxyz"] == [0.5, -1.0, 1.0]


def test_domain_walls_follow_the_bc_list(tmp_path):
    """domainWalls = 1 appends the periodic domain walls after the BC entries
    (read.c:1441-1520): a solid wall pair in y plus four periodic walls"""
    bc = [dict(aInv=[0, 1, 0], Q=[0, 0, 0], P=[1, 1, 1, 1], R=0, DN=0, MVN=-1, MVT=-1),
          dict(aInv=[0, -1, 0], Q=[0, 6, 0], P=[1, 1, 1, 1], R=0, DN=0, MVN=-1, MVT=-1)]
    d = dump(tmp_path, dict(domain=[10, 6], BC=bc, domainWalls=1))
    assert len(d["walls"]) == 6
    assert d["walls"][0]["A"] == [0.0, 1.0, 0.0] and d["walls"][0]["MVN"] == -1.0
    assert [w["DN"] for w in d["walls"][2:]] == [10.0, 10.0, 6.0, 6.0]
    # a BC list alone adds no domain walls; no BC list gives periodic ones
    assert len(dump(tmp_path, dict(domain=[10, 6], BC=bc))["walls"]) == 2
    assert len(dump(tmp_path, dict(domain=[10, 6]))["walls"]) == 4


def test_rest_frame_flag(tmp_path):
    """rFrame (default 1) stays on for the periodic reset only with
    zeroNetMom > 0 and no body force (init.c:2155-2156)"""
    d = dump(tmp_path, dict(domain=[8, 8], zeroNetMom=100))
    assert (d["rFrame"], d["zeroNetMom"]) == (1, 100)
    assert dump(tmp_path, dict(domain=[8, 8]))["rFrame"] == 0
    assert dump(tmp_path, dict(domain=[8, 8], zeroNetMom=100, rFrame=0))["rFrame"] == 0
    assert dump(tmp_path, dict(domain=[8, 8], zeroNetMom=100, grav=[1e-3, 0, 0]))["rFrame"] == 0


@pytest.mark.gpu
def test_host_zero_net_momentum(tmp_path):
    """A channel between bounce-back walls: the walls change the total
    momentum every step.  With zeroNetMom = 3 the output step of every
    runtime divisible by 3 takes it back to zero (galileantrans,
    pout.c:2557-2562), before the solvent.dat line of that runtime."""
    cfg = {
        "domain": [24, 16], "simSteps": 99, "collOp": 2, "seed": 5, "zeroNetMom": 3,
        "species": [{"dens": 10, "vDist": 4, "mass": 1.0}, {"dens": 4, "vDist": 4, "mass": 2.5}],
        "BC": [
            {"aInv": [1, 0, 0], "Q": [0, 0, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 24},
            {"aInv": [-1, 0, 0], "Q": [24, 0, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 24},
            {"aInv": [0, 1, 0], "Q": [0, 0, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 0, "MVN": -1, "MVT": -1,
             "phantom": 1},
            {"aInv": [0, -1, 0], "Q": [0, 16, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 0, "MVN": -1, "MVT": -1,
             "phantom": 1},
        ],
    }
    print("[host] 2D channel 24x16, two species, zeroNetMom 3, 100 steps", flush=True)
    info, rows = run_host(tmp_path, cfg, "--every", "10")
    assert [int(x[0]) for x in rows] == list(range(9, 100, 10))
    reset = [r for r in rows if int(r[0]) % 3 == 0]
    other = [r for r in rows if int(r[0]) % 3 != 0]
    assert len(reset) == 4
    for r in reset:
        assert abs(r[3]) < 1e-9 and abs(r[4]) < 1e-9, r
    assert max(abs(r[3]) + abs(r[4]) for r in other) > 1e-3


def test_rtech_alone(tmp_path):
    assert dump(tmp_path, dict(domain=[8, 8], rTech=1))["rtech"] == 1
    assert dump(tmp_path, dict(domain=[8, 8], collOp=0))["rtech"] == 0


@pytest.mark.parametrize("cfg,msg", [
    (dict(domain=[8, 8], mdIn="md.inp"), "mdIn"),
    (dict(domain=[8, 8], checkpointIn="c.dat"), "cannot open the checkpoint"),
    (dict(domain=[8, 8], nSwim=2), "nSwim"),
    (dict(domain=[8, 8], BC=[dict(aInv=[0.25, 0.25, 0], Q=[4, 4, 0], P=[2, 2, 1, 1], R=4, dsplc=1)]),
     "displaceable"),
    (dict(domain=[8, 8], BC=[dict(aInv=[1, 0, 0], Q=[0, 0, 0], P=[1, 1, 1, 1], rotSym=[2, 2])]), "rotSym"),
    (dict(domain=[8]), "domain"),
])
def test_rejected_options(tmp_path, cfg, msg):
    r = run_dump(tmp_path, cfg)
    assert r.returncode != 0 and msg in r.stderr


def test_malformed_json(tmp_path):
    r = run_dump(tmp_path, None, text='{"domain": [8, 8], "kbt": }')
    assert r.returncode != 0 and "JSON line 1" in r.stderr


def test_text_after_the_first_object_is_ignored(tmp_path):
    """cJSON_Parse (the reference's readJson, read.c:993) reads the first
    value and ignores what follows it; so does the host: a second object
    after the first changes nothing."""
    text = '{"domain": [8, 6], "kbt": 2.0}\n\n{"checkpointIn": "./checkpoint.dat", "kbt": 3.0}\n'
    d = dump(tmp_path, text=text)
    assert d["L"][:2] == [8, 6] and d["kbt"] == 2.0 and d["runtime"] == 0


def run_host(tmp_path, cfg, *args):
    f = tmp_path / "input.json"
    f.write_text(json.dumps(cfg))
    out = tmp_path / "out"
    r = subprocess.run([host(), str(f), "-o", str(out), *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    rows = [list(map(float, l.split())) for l in open(out / "solvent.dat") if not l.startswith("#")]
    return info, rows


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [(), ("--replicas", "2", "--no-graph"), ("--fp32x",)])
def test_host_periodic_box_momentum(tmp_path, extra):
    """SRD in a periodic box conserves momentum, so the body force alone sets
    it: P_x = N m g dt (steps done) from the zero initial momentum (in the
    exact-sum FP32 form to the FP32 rounding of each collision)."""
    rel, ab = (1e-4, 2e-3) if "--fp32x" in extra else (2e-6, 1e-6)
    cfg = {"domain": [30, 20], "simSteps": 199, "warmUp": 9, "rTech": 0, "grav": [0.0005, 0, 0], "seed": 3,
           "species": [{"dens": 10, "vDist": 4}]}
    print(f"[host] periodic 2D box 30x20x10, 10 + 200 steps {extra}", flush=True)
    info, rows = run_host(tmp_path, cfg, "--every", "50", *extra)
    assert info["steps"] == 200 and info["particles_per_replica"] == 6000
    assert [int(x[0]) for x in rows] == [49, 99, 149, 199]
    for step, kbt, ke, px, py, pz in rows:
        assert 0.9 < kbt < 1.1
        expect = 6000 * 1.0 * 0.0005 * 0.1 * (10 + step + 1)  # warm-up: 9 + 1 steps
        assert abs(px - expect) < rel * expect, (step, px, expect)
        assert abs(py) < ab and pz == 0.0


@pytest.mark.gpu
def test_host_runs_a_channel(tmp_path):
    """A channel between bounce-back phantom walls: every wall hit reverses a
    particle's velocity, so the total momentum random-walks by far more than
    the body force adds in 200 steps; the checks are the step count, the
    temperature (SRD and bounce-back keep the kinetic energy) and the plane."""
    cfg = {
        "domain": [30, 20], "simSteps": 199, "collOp": 0, "grav": [0.0005, 0, 0], "seed": 3,
        "species": [{"dens": 10, "vDist": 4}],
        "BC": [
            {"aInv": [1, 0, 0], "Q": [0, 0, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 30},
            {"aInv": [-1, 0, 0], "Q": [30, 0, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 30},
            {"aInv": [0, 1, 0], "Q": [0, 0, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 0, "MVN": -1, "MVT": -1,
             "phantom": 1},
            {"aInv": [0, -1, 0], "Q": [0, 20, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 0, "MVN": -1, "MVT": -1,
             "phantom": 1},
        ],
    }
    print("[host] 2D channel 30x20x10, 200 steps", flush=True)
    info, rows = run_host(tmp_path, cfg, "--every", "50")
    assert info["steps"] == 200 and info["particles_per_replica"] == 6000
    assert [int(x[0]) for x in rows] == [49, 99, 149, 199]
    for step, kbt, ke, px, py, pz in rows:
        assert 0.9 < kbt < 1.1
        assert pz == 0.0


@pytest.mark.gpu
def test_host_runs_an_active_nematic(tmp_path):
    """the host on a 2D active nematic input (lc 1, collOp 16, random
    orientations): the run completes and the temperature stays near kT"""
    cfg = {"domain": [24, 24], "simSteps": 99, "lc": 1, "collOp": 16, "seed": 8,
           "species": [{"dens": 20, "vDist": 4, "mfpot": 10, "tumble": 0.5, "rfc": 0.01, "act": 0.05}]}
    print("[host] 2D active nematic 24x24x20, 100 steps", flush=True)
    info, rows = run_host(tmp_path, cfg, "--every", "25")
    assert info["steps"] == 100 and info["particles_per_replica"] == 24 * 24 * 20
    assert [int(x[0]) for x in rows] == [24, 49, 74, 99]
    for step, kbt, ke, px, py, pz in rows:
        assert 0.8 < kbt < 1.2


@pytest.mark.gpu
def test_host_writes_the_flow_field(tmp_path):
    """flowOut: 13 header lines, then one block of L0 x L1 cell lines at every
    runtime divisible by flowOut (pout.c:2690), each holding the sum over the
    interval divided by flowOut (pout.c:1691).  In a periodic box the body
    force gives every particle the same drift g dt s after s steps, so the
    mean over cells of a block is g dt times the mean step count of its
    interval, up to thermal noise (about 2e-3 here; a missing division or a
    sum that is not restarted after a block is off by 0.05 or more)."""
    g, dt, F = 0.02, 0.1, 50
    cfg = {"domain": [30, 20], "simSteps": 199, "rTech": 0, "grav": [g, 0, 0], "seed": 5, "dt": dt,
           "flowOut": F, "species": [{"dens": 50, "vDist": 4}]}
    print("[host] periodic 2D box 30x20x50 with flowOut 50, 200 steps", flush=True)
    info, rows = run_host(tmp_path, cfg, "--every", "100")
    assert info["steps"] == 200
    lines = open(tmp_path / "out" / "flowfield.dat").read().splitlines()
    assert lines[13 - 1].split() == ["t", "QX", "QY", "QZ", "VcmX", "VcmY", "VcmZ"]
    data = np.array([[float(x) for x in l.split()] for l in lines[13:]])
    assert data.shape == (4 * 600, 7)
    blocks = data.reshape(4, 600, 7)
    for b, r in enumerate((0, 50, 100, 150)):
        blk = blocks[b]
        assert np.allclose(blk[:, 0], r * dt)
        assert sorted(set(blk[:, 1].astype(int))) == list(range(30)) and sorted(set(blk[:, 2].astype(int))) == list(range(20))
        assert np.all(blk[:, 3] == 0) and np.all(blk[:, 6] == 0.0)
        # steps done at runtime s: s + 1; the block at r sums runtimes r - F + 1 .. r (just r = 0 for the first)
        expect = g * dt * 1.0 / F if r == 0 else g * dt * (r - (F - 1) / 2.0 + 1.0)
        assert abs(blk[:, 4].mean() - expect) < 1e-2, (r, blk[:, 4].mean(), expect)
        assert abs(blk[:, 5].mean()) < 1e-2


@pytest.mark.gpu
def test_host_flow_field_resolves_a_channel(tmp_path):
    """The cell ordering of flowfield.dat: between bounce-back walls in y a
    body force along x starts a flow that the walls hold back, so the
    averaged vx of the middle rows is well above that of the wall rows (about
    0.25 against under 0.15 after 150 steps; row means carry about 0.01 of
    noise), and the two wall rows agree."""
    g, F = 0.02, 50
    cfg = {
        "domain": [30, 20], "simSteps": 199, "collOp": 0, "grav": [g, 0, 0], "seed": 9, "flowOut": F,
        "species": [{"dens": 20, "vDist": 4}],
        "BC": [
            {"aInv": [1, 0, 0], "Q": [0, 0, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 30},
            {"aInv": [-1, 0, 0], "Q": [30, 0, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 30},
            {"aInv": [0, 1, 0], "Q": [0, 0, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 0, "MVN": -1, "MVT": -1,
             "phantom": 1},
            {"aInv": [0, -1, 0], "Q": [0, 20, 0], "P": [1, 1, 1, 1], "R": 0, "DN": 0, "MVN": -1, "MVT": -1,
             "phantom": 1},
        ],
    }
    print("[host] 2D channel 30x20x20 with flowOut 50, 200 steps", flush=True)
    run_host(tmp_path, cfg, "--every", "200")
    lines = open(tmp_path / "out" / "flowfield.dat").read().splitlines()
    blk = np.array([[float(x) for x in l.split()] for l in lines[13:]]).reshape(4, 600, 7)[-1]
    vx = np.zeros((30, 20))
    vx[blk[:, 1].astype(int), blk[:, 2].astype(int)] = blk[:, 4]
    rows = vx.mean(axis=0)
    middle, walls = rows[8:12].mean(), 0.5 * (rows[0] + rows[19])
    print(f"[host] middle rows {middle:.4f}, wall rows {rows[0]:.4f} {rows[19]:.4f}", flush=True)
    assert middle - walls > 0.08, rows
    assert abs(rows[0] - rows[19]) < 0.05, rows
    assert np.all(np.diff(rows[:10]) > -0.03) and np.all(np.diff(rows[10:]) < 0.03), rows


@pytest.mark.gpu
def test_host_velocity_and_director_fields(tmp_path):
    """velOut and dirSOut on a 2D active nematic.  With flowOut = 1 the flow
    field at a runtime holds that one step's cell velocities, summed by the
    flow kernels, and the velocity field holds the same velocities from the
    list-order cell sums (localPROP): the two files agree cell by cell to the
    printed precision.  The director field holds a unit director and S in
    [0, 1] in every occupied cell (orderout, pout.c:2021-2033)."""
    cfg = {"domain": [16, 12], "simSteps": 9, "lc": 1, "collOp": 16, "seed": 4, "flowOut": 1, "velOut": 1,
           "dirSOut": 5, "species": [{"dens": 20, "vDist": 4, "mfpot": 10, "tumble": 0.5, "act": 0.05}]}
    print("[host] 2D active nematic 16x12x20, 10 steps, flowOut 1, velOut 1, dirSOut 5", flush=True)
    run_host(tmp_path, cfg, "--every", "10")
    out = tmp_path / "out"

    def table(name):
        lines = open(out / name).read().splitlines()
        return lines[12], np.array([[float(x) for x in l.split()] for l in lines[13:]])

    hf, flow = table("flowfield.dat")
    hv, vel = table("velfield.dat")
    hd, dirs = table("directorfield.dat")
    assert hf.split() == hv.split() == ["t", "QX", "QY", "QZ", "VcmX", "VcmY", "VcmZ"]
    assert hd.split() == ["t", "QX", "QY", "QZ", "NX", "NY", "NZ", "S"]
    assert flow.shape == vel.shape == (10 * 16 * 12, 7)
    np.testing.assert_array_equal(flow[:, :4], vel[:, :4])
    np.testing.assert_allclose(flow[:, 4:], vel[:, 4:], rtol=2e-5, atol=2e-6)
    assert np.abs(vel[:, 4:6]).max() > 0.05 and np.all(vel[:, 6] == 0.0)
    assert dirs.shape == (2 * 16 * 12, 8)
    assert sorted(set(np.round(dirs[:, 0], 2))) == [0.0, 0.5]
    occ = np.abs(dirs[:, 4:7]).sum(axis=1) > 0
    assert occ.mean() > 0.95
    np.testing.assert_allclose(np.linalg.norm(dirs[occ, 4:7], axis=1), 1.0, atol=1e-4)
    assert np.all(dirs[:, 6] == 0.0)
    assert np.all((dirs[occ, 7] >= 0.0) & (dirs[occ, 7] <= 1.0))


# --- checkpoints (checkpoint(), pout.c:2325-2404; readchckpnt(), read.c:669) ---------------------------------------

CKPT_CASES = ["rat3d_channel", "lc2d_channel", "rat2d_2sp"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ckpt_fields(text):
    """(name, token) pairs of a checkpoint file, in the order checkpoint() writes them"""
    tok = iter(text.split())
    out = []

    def take(*names):
        for nm in names:
            out.append((nm, next(tok)))

    take("simSteps", "warmupSteps", "dt", "seed", "DIM", "X", "Y", "Z", "KBT", "KBTNOW", "RFRAME", "zeroNetMom",
         "GALINV", "TSTECH", "RTECH", "LC", "TAU", "RA", "FRICCO", "noHI", "inCOMP", "MULTIPHASE", "MFPLAYERH",
         "G0", "G1", "G2", "MAG0", "MAG1", "MAG2", "MD_mode", "stepsMD", "GPOP", "NSPECI", "runtime", "warmtime", "C",
         "S", "GRAV_FLAG", "MAG_FLAG", "tolD", "AVVEL", "AVS", "avDIR0", "avDIR1", "avDIR2", "S4", "stdN", "VOL",
         "AVV0", "AVV1", "AVV2", "AVNOW0", "AVNOW1", "AVNOW2", *[f"theory{k}" for k in range(6)], "chckpntIn",
         "MDmode", *[f"flag{k}" for k in range(32)], *[f"more{k}" for k in range(17)])
    ns = int(dict(out)["NSPECI"])
    for s in range(ns):
        take(*[f"sp{s}.{k}" for k in ("MASS", "POP", "QDIST", "VDIST", "ODIST", "RFC", "LEN", "TUMBLE", "CHIHI", "CHIA",
                                       "ACT", "MFPOT", "SIGWIDTH", "SIGPOS", "DAMP", "VOL", "nDNST", "mDNST",
                                       "MINACTRATIO", "BS")], *[f"sp{s}.M{j}" for j in range(ns)],
             *[f"sp{s}.theory{k}" for k in range(6)])
    take("GnDNST", "GmDNST", "GMASS", "maxXYZ", "NBC")
    for w in range(int(dict(out)["NBC"])):
        take(*[f"w{w}.{k}" for k in ("COLL_TYPE", "PHANTOM", "E", "Q0", "Q1", "Q2", "V0", "V1", "V2", "O0", "O1", "O2",
                                       "L0", "L1", "L2", "G0", "G1", "G2", "A0", "A1", "A2", "AINV0", "AINV1", "AINV2",
                                       "P0", "P1", "P2", "P3", "R", "B0", "B1", "B2", "DN", "DT", "DVN", "DVT", "DVx",
                                       "DVy", "DVz", "MVN", "MVT", "MUN", "MUT", "MUx", "MUy", "MUz", "DUx", "DUy",
                                       "DUz", "KBT", "DSPLC", "INV", "MASS")], *[f"w{w}.hist{k}" for k in range(28)],
             *[f"w{w}.INTER{j}" for j in range(12)])
    for i in range(int(dict(out)["GPOP"])):
        take(*[f"p{i}.{k}" for k in ("S_flag", "SPID", "q", "Q0", "Q1", "Q2", "V0", "V1", "V2", "U0", "U1", "U2",
                                       "T0", "T1", "T2")])
    take("NS", *[f"swimmer{k}" for k in range(9 + 19)])
    assert next(tok, None) is None, "tokens after the swimmer records"
    return out


# fields this path does not carry (the reference fills them from state it alone has)
NOT_CARRIED = ("warmupSteps", "KBTNOW", "AVNOW", "VOL", "AVVEL", "AVS", "avDIR", "S4", "stdN", "theory", "flag",
               ".POP", ".LEN", ".CHIA", ".QDIST", ".ODIST", ".VOL", ".nDNST", ".mDNST", ".BS", ".M", "GmDNST", "GMASS",
               "maxXYZ", ".hist", ".S_flag", ".q", ".T0", ".T1", ".T2")


@pytest.mark.parametrize("name", CKPT_CASES)
def test_checkpoint_layout_and_values(tmp_path, name):
    """The reference's own checkpoint file (tests/golden/ckpt_<case>.dat) is
    read, and written back by --write-checkpoint: the two files hold the same
    fields in the same order, and every field this path carries (inputs,
    species, boundaries, particles) is the same token."""
    ref = open(os.path.join(GOLDEN, f"ckpt_{name}.dat")).read()
    f = tmp_path / "input.json"
    f.write_text(json.dumps({"checkpointIn": os.path.join(GOLDEN, f"ckpt_{name}.dat")}))
    out = tmp_path / "mine.dat"
    r = subprocess.run([host(), str(f), "--write-checkpoint", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    a, b = ckpt_fields(ref), ckpt_fields(open(out).read())
    assert [k for k, _ in a] == [k for k, _ in b]
    diff = [(k, x, y) for (k, x), (_, y) in zip(a, b) if x != y and not any(s in k for s in NOT_CARRIED)]
    assert not diff, diff[:10]


@pytest.mark.parametrize("name", CKPT_CASES)
def test_checkpoint_read(tmp_path, name):
    """The parameters and particles the host takes from the reference's
    checkpoint are the recorded case's (to the file's six decimals)."""
    g = U.load_golden(name)
    c = g["case_d"]
    d = dump(tmp_path, dict(checkpointIn=os.path.join(GOLDEN, f"ckpt_{name}.dat")))
    assert (d["dim"], d["rtech"], d["lc"], d["tstech"], d["galinv"]) == (c["dim"], c["rtech"], c["lc"], c["tstech"],
                                                                         c["galinv"])
    assert d["L"][:c["dim"]] == c["L"][:c["dim"]]
    assert d["simSteps"] == 50 and d["warmUp"] == 0 and d["runtime"] == c["steps"] - 1
    for k in ("dt", "kbt", "rotang", "tau"):
        assert abs(d[k] - c[k]) < 1e-6, k
    assert np.allclose(d["grav"], c["grav"], atol=1e-6) and np.allclose(d["avv"], g["avv"], atol=1e-6)
    assert abs(d["gndnst"] - float(g["gndnst"])) < 1e-6
    pops = np.bincount(g["spid"], minlength=len(c["species"]))
    assert [s["pop"] for s in d["species"]] == list(pops)
    for s, sp in zip(d["species"], c["species"]):
        for k in ("mass", "rfc", "tumble", "chihi", "act", "mfpot", "damp", "sigwidth", "sigpos", "minactratio"):
            assert abs(s[k] - sp[k]) < 1e-6, k
    assert len(d["walls"]) == len(c["walls"])
    for w, cw in zip(d["walls"], c["walls"]):
        assert np.allclose(w["A"], cw["A"]) and np.allclose(w["Q"], cw.get("Q", (0, 0, 0)))
        for k, dflt in (("R", 0.0), ("DN", 0.0), ("MVN", 1.0), ("MVT", 1.0), ("MUN", 1.0), ("MUT", 1.0)):
            assert abs(w[k] - cw.get(k, dflt)) < 1e-6, k
        assert w["PHANTOM"] == cw.get("PHANTOM", 0) and w["INV"] == cw.get("INV", 0)
        assert w["inter_mask"] == (1 << len(c["species"])) - 1


@pytest.mark.gpu
def test_host_checkpoint_and_resume(tmp_path):
    """checkpointOut writes DIR/checkpoint.dat after the outputs of every
    runtime divisible by it (mpcd.c:310-312).  A run recovered from it skips
    the warm-up and starts its temporal loop at the checkpoint's runtime
    (mpcd.c:257, 290), so that runtime's step runs again.  In a periodic box
    SRD keeps the momentum the body force gives, N m g dt per step from zero:
    after the resume it counts 5 warm-up steps, 51 steps to runtime 50 and
    the resumed ones (to the file's six decimals per velocity)."""
    g, dt, N = 0.001, 0.1, 20 * 12 * 10
    cfg = {"domain": [20, 12], "simSteps": 59, "warmUp": 4, "collOp": 0, "seed": 3, "checkpointOut": 25, "dt": dt,
           "grav": [g, 0, 0], "species": [{"dens": 10, "vDist": 4}]}
    print("[host] periodic 2D box 20x12x10, checkpointOut 25, then a resume", flush=True)
    run_host(tmp_path, cfg, "--every", "20")
    ck = tmp_path / "out" / "checkpoint.dat"
    f = dict(ckpt_fields(open(ck).read()))
    assert (int(f["runtime"]), int(f["warmtime"]), int(f["simSteps"]), int(f["warmupSteps"])) == (50, 5, 59, 0)
    assert int(f["flag29"]) == 25 and int(f["GPOP"]) == N
    res = tmp_path / "resume"
    res.mkdir()
    info, rows = run_host(res, {"checkpointIn": str(ck)}, "--every", "5")
    assert info["steps"] == 10
    assert [int(x[0]) for x in rows] == [54, 59]
    for rt, kbt, ke, px, py, pz in rows:
        expect = N * g * dt * (5 + 51 + (rt - 50 + 1))
        assert abs(px - expect) < 2e-3, (rt, px, expect)
        assert 0.9 < kbt < 1.1
    # the resumed run wrote its own checkpoint at runtime 50 again (the flag came from the file)
    f2 = dict(ckpt_fields(open(res / "out" / "checkpoint.dat").read()))
    assert int(f2["runtime"]) == 50


@pytest.mark.gpu
def test_host_density_field(tmp_path):
    """densOut: per cell the population, the mass and the population of each
    species (densityout, pout.c:1744-1757); every block holds each species'
    whole population, the mass column is the species-weighted sum and the
    MPCD population is the population (no MD, no swimmers)."""
    cfg = {"domain": [12, 10], "simSteps": 19, "rTech": 1, "seed": 6, "densOut": 10,
           "species": [{"pop": 700, "mass": 1.0, "vDist": 4}, {"pop": 500, "mass": 2.5, "vDist": 4}]}
    print("[host] 2D box 12x10, two species, densOut 10", flush=True)
    run_host(tmp_path, cfg, "--every", "20")
    lines = open(tmp_path / "out" / "densityfield.dat").read().splitlines()
    assert lines[12].split() == ["t", "QX", "QY", "QZ", "pop", "mass", "popSRD", "popMD", "popSW", "SP0", "SP1"]
    data = np.array([[float(x) for x in l.split()] for l in lines[13:]])
    assert data.shape == (2 * 120, 11)
    for blk, t in zip(data.reshape(2, 120, 11), (0.0, 1.0)):
        assert np.allclose(blk[:, 0], t)
        assert blk[:, 9].sum() == 700 and blk[:, 10].sum() == 500
        assert np.array_equal(blk[:, 4], blk[:, 9] + blk[:, 10]) and np.array_equal(blk[:, 6], blk[:, 4])
        assert np.allclose(blk[:, 5], blk[:, 9] * 1.0 + blk[:, 10] * 2.5, atol=1e-4)
        assert not blk[:, 7].any() and not blk[:, 8].any()


@pytest.mark.gpu
@pytest.mark.parametrize("lc", [0, 1])
def test_host_energy_output(tmp_path, lc):
    """energyOut (enout, pout.c:1389-1422): in a periodic box SRD keeps the
    kinetic energy exactly, so every line shows the same MPCD_kin, equal to
    solvent.dat's kinetic energy (the mean velocity AVV is zero); without LC
    the nematic term is 0 and Total = MPCD_kin.  With LC the nematic
    mean-field energy (calcE_LC) is positive and Total is the sum."""
    cfg = {"domain": [16, 12], "simSteps": 19, "seed": 2, "energyOut": 5, "rTech": 16 if lc else 0, "lc": lc,
           "species": [{"dens": 20, "vDist": 4, "mfpot": 10, "tumble": 0.5, "act": 0.0}]}
    print(f"[host] energy output, lc={lc}", flush=True)
    info, rows = run_host(tmp_path, cfg, "--every", "5")
    lines = open(tmp_path / "out" / "energy.dat").read().splitlines()
    assert lines[12].split() == ["t", "MPCD_kin", "MPCD_nem", "BC_kin", "BC_rot", "Total", "KBT"]
    e = np.array([[float(x) for x in l.split()] for l in lines[13:]])
    assert e.shape == (4, 7) and np.allclose(e[:, 0], [0.0, 0.5, 1.0, 1.5])
    assert not e[:, 3].any() and not e[:, 4].any()
    assert np.allclose(e[:, 5], e[:, 1] + e[:, 2], atol=2e-6)
    assert np.all((0.8 < e[:, 6]) & (e[:, 6] < 1.2))
    if lc:
        assert np.all(e[:, 2] > 0)
    else:
        assert not e[:, 2].any()
        assert np.allclose(e[:, 1], e[0, 1], rtol=1e-9)
        ke = [r[2] for r in rows]
        assert np.allclose(e[1:, 1], ke[:3], rtol=1e-6)


def test_checkpoint_outside_the_path_is_refused(tmp_path):
    """A checkpoint with MD coupling, swimmers, a displaceable or turned wall,
    or a truncated particle list stops the host with a message naming the
    cause; a wavy wall is read with its wavy terms."""
    text = open(os.path.join(GOLDEN, "ckpt_lc2d_channel.dat")).read()
    fields = ckpt_fields(text)
    names = [k for k, _ in fields]

    def variant(edits, cut=None):
        tok = [v for _, v in fields]
        for k, v in edits.items():
            tok[names.index(k)] = v
        if cut is not None:
            tok = tok[:cut]
        return " ".join(tok)

    cases = [(variant({"MD_mode": "1"}), "MD coupling"),
             (variant({"NS": "2"}), "swimmers"),
             (variant({"w0.DSPLC": "1"}), "displaceable"),
             (variant({"w0.O1": "0.500000"}), "turned"),
             (variant({}, cut=names.index("p5.Q0")), "ends or breaks")]
    for k, (body, msg) in enumerate(cases):
        ck = tmp_path / f"ck{k}.dat"
        ck.write_text(body)
        r = run_dump(tmp_path, {"checkpointIn": str(ck)})
        assert r.returncode != 0 and msg in r.stderr, (msg, r.stderr)
    ck = tmp_path / "wavy.dat"
    ck.write_text(variant({"w0.B0": "0.500000", "w0.B1": "0.250000", "w0.P2": "2.000000"}))
    r = run_dump(tmp_path, {"checkpointIn": str(ck)})
    assert r.returncode == 0, r.stderr
    w0 = json.loads(r.stdout)["walls"][0]
    assert w0["B"] == [0.5, 0.25, 0.0] and w0["P"][2] == 2.0


def test_curved_walls_pass_through(tmp_path):
    """a cylinder and a wavy plane reach the library with their exponents,
    wavy terms and |.| flag; a density fills the accessible volume
    (accessibleVolume, init.c:954-990): the box minus the disc"""
    cyl = dict(aInv=[1, 1, 0], Q=[8, 8, 0], P=[2, 2, 1, 2], R=4, inv=1, abs=0)
    wavy = dict(aInv=[0, 1, 0], Q=[0, 2, 0], P=[1, 1, 1, 1], R=0, wavy=[0.5, 0.25, 0])
    d = dump(tmp_path, dict(domain=[16, 16], BC=[cyl, wavy], domainWalls=1, species=[dict(dens=10)]))
    w0, w1 = d["walls"][0], d["walls"][1]
    assert w0["P"] == [2, 2, 1, 2] and w0["INV"] == 1 and w0["ABS"] == 0 and w0["R"] == 4
    assert w1["B"] == [0.5, 0.25, 0] and w1["P"] == [1, 1, 1, 1]
    # inside the inverted cylinder of radius 4 only, so the accessible area is
    # about pi 4^2 = 50.3 of 256 minus the part below the wavy plane near y = 2
    pop = d["species"][0]["pop"]
    assert 0 < pop < 10 * 16 * 16
    ref = 10 * math.pi * 16.0
    assert abs(pop - ref) < 0.05 * ref, (pop, ref)


def _ckpt_particles(path):
    f = ckpt_fields(open(path).read())
    d = dict(f)
    n = int(d["GPOP"])
    return np.array([[float(d[f"p{i}.Q{k}"]) for k in range(3)] for i in range(n)])


@pytest.mark.gpu
def test_host_runs_curved_walls(tmp_path):
    """pipeFlow- and wavyWalls-shaped inputs through the host and the
    general wall engine: the particles are placed in the space the walls leave
    (a density fills the accessible volume), stay there after 20 steps (every
    checkpointed position inside the pipe, or between the wavy walls, to the
    file's six decimals) and the body force drives a flow along the pipe."""
    pipe = {"domain": [24, 12, 12], "domainWalls": 1, "simSteps": 20, "collOp": 2, "seed": 8, "checkpointOut": 20,
            "grav": [0.01, 0, 0], "species": [{"dens": 10}],
            "BC": [{"Q": [0, 6, 6], "aInv": [0, 1, 1], "P": [1, 2, 2, 2], "R": 5, "DN": 0, "phantom": 1,
                    "MVN": -1, "MVT": -1, "inv": 1}]}
    print("[host] pipe 24x12x12, R 5, MPC-AT, checkpoint at 20", flush=True)
    (tmp_path / "pipe").mkdir()
    info, rows = run_host(tmp_path / "pipe", pipe, "--every", "10")
    Q = _ckpt_particles(tmp_path / "pipe" / "out" / "checkpoint.dat")
    n = Q.shape[0]
    assert abs(n - 10 * math.pi * 25 * 24) < 0.03 * 10 * math.pi * 25 * 24, n
    r2 = (Q[:, 1] - 6) ** 2 + (Q[:, 2] - 6) ** 2
    assert r2.max() <= 25 + 1e-4, r2.max()
    assert rows[-1][3] > 0.0  # momentum along the pipe
    k = 2 * math.pi / 20
    wavy = {"domain": [40, 16], "domainWalls": 1, "simSteps": 20, "collOp": 0, "seed": 9, "checkpointOut": 20,
            "species": [{"dens": 10}], "grav": [0.005, 0, 0],
            "BC": [{"Q": [0, 2, 0], "aInv": [0, 1, 0], "P": [1, 1, 1, 1], "R": 0, "wavy": [0.5, k, 0], "DN": 0,
                    "MVN": -1, "MVT": -1},
                   {"Q": [0, 14, 0], "aInv": [0, -1, 0], "P": [1, 1, 1, 1], "R": 0, "wavy": [0.5, k, 0], "DN": 0,
                    "MVN": -1, "MVT": -1}]}
    print("[host] wavy channel 40x16, SRD, checkpoint at 20", flush=True)
    (tmp_path / "wavy").mkdir()
    run_host(tmp_path / "wavy", wavy, "--every", "10")
    Q = _ckpt_particles(tmp_path / "wavy" / "out" / "checkpoint.dat")
    lo = 2 - 0.5 * np.cos(k * Q[:, 0])
    hi = 14 + 0.5 * np.cos(k * Q[:, 0])
    assert (Q[:, 1] >= lo - 1e-4).all() and (Q[:, 1] <= hi + 1e-4).all()


def _bin_particles(path, n):
    """the particle arrays at the end of a binary checkpoint: Q, V, U (FP64) and the species"""
    raw = open(path, "rb").read()
    tail = n * (9 * 8 + 4)
    return raw[-tail:]


def test_binary_checkpoint_round_trip(tmp_path):
    """--binary-checkpoint writes the lossless layout (magic, parameters, loop
    state, draw counter, FP64 particles); reading it back and writing it again
    gives the same particle bytes and the same parameters"""
    cfg = {"domain": [10, 8], "collOp": 3, "seed": 4, "species": [{"pop": 400, "mass": 1.0}, {"pop": 300, "mass": 2.0}],
           "BC": [{"Q": [5, 4, 0], "aInv": [1, 1, 0], "P": [2, 2, 1, 2], "R": 2, "DN": 0, "MVN": -1, "MVT": -1}]}
    f = tmp_path / "in.json"
    f.write_text(json.dumps(cfg))
    a, b = tmp_path / "a.bin", tmp_path / "b.bin"
    r = subprocess.run([host(), str(f), "--write-checkpoint", str(a), "--binary-checkpoint"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert open(a, "rb").read(8) == b"MPCDCKB1"
    g = tmp_path / "in2.json"
    g.write_text(json.dumps({"checkpointIn": str(a)}))
    r = subprocess.run([host(), str(g), "--write-checkpoint", str(b), "--binary-checkpoint"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert _bin_particles(a, 700) == _bin_particles(b, 700)
    d1 = dump(tmp_path, cfg)
    d2 = json.loads(subprocess.run([host(), str(g), "--dump-params"], capture_output=True, text=True,
                                   timeout=60).stdout)
    assert d1["walls"] == d2["walls"] and [s["mass"] for s in d1["species"]] == [s["mass"] for s in d2["species"]]
    assert [s["pop"] for s in d2["species"]] == [400, 300]
    # a binary file cut short is refused with a message
    c = tmp_path / "c.bin"
    c.write_bytes(open(a, "rb").read()[:-100])
    g.write_text(json.dumps({"checkpointIn": str(c)}))
    r = subprocess.run([host(), str(g), "--dump-params"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cut short" in r.stderr


@pytest.mark.gpu
def test_binary_checkpoint_resume_is_bit_exact(tmp_path):
    """in the exact-sum form a run resumed from a binary checkpoint (state and
    draw counter) ends on the same bits as the run that went straight
    through: 21 steps, against 11 steps, a checkpoint and 10 more"""
    cfg = {"domain": [12, 10, 8], "collOp": 2, "seed": 21, "grav": [0.002, 0, 0], "simSteps": 20, "checkpointOut": 10,
           "species": [{"dens": 8}]}
    print("[host] binary checkpoint resume, MPC-AT 12x10x8, exact-sum form", flush=True)
    full, half, res = tmp_path / "full", tmp_path / "half", tmp_path / "res"
    for d in (full, half, res):
        d.mkdir()
    run_host(full, cfg, "--fp32x", "--binary-checkpoint", "--every", "10")
    run_host(half, dict(cfg, simSteps=10), "--fp32x", "--binary-checkpoint", "--every", "10")
    run_host(res, {"checkpointIn": str(half / "out" / "checkpoint.bin")}, "--fp32x", "--binary-checkpoint",
             "--steps", "20", "--every", "10")
    n = 12 * 10 * 8 * 8
    a = _bin_particles(full / "out" / "checkpoint.bin", n)
    b = _bin_particles(res / "out" / "checkpoint.bin", n)
    assert a == b
