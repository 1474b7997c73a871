"""FP64 CPU adjoint of the 2D active nematic step (test infrastructure).

A torch (float64, CPU) restatement of one step of the 2D active nematic on
the split path, written from the reference routines and independent of the
HIP kernels, differentiated by torch.autograd (reverse mode):

  gridShift_all(+) (mpc.c:728-736) and the shifted bin -- the list order
    and the cells come from the oracle's run (the pathwise derivative holds
    the cell assignment fixed);
  acc_all (mpc.c:697); localPROP (mpc.c:64-216: mass, VCM, centre of mass,
    moment of inertia, the order tensor, its largest eigenvalue and
    eigenvector by eigenvalues2x2/eigenvectors2x2, mtools.c:1230-1305);
  localVelGrad (velGradD2Q9 with the edge copies, lc.c:2070-2162);
  ghostPart with strong anchoring at phantom walls (mpc.c:351-567,
    oriBC lc.c:1290-1393);
  LCcollision with the 2D Metropolis Maier-Saupe sampler (lc.c:341-411,
    427-544) and the torque it leaves, then jefferysTorque (lc.c:559-718);
  dipoleAndersenROT_LC, DIPOLE_DIR_SUM (lc.c:1749-2043);
  gridShift_all(-), stream_all (mpc.c:680) and MPC_BCcollision with planar
    walls (bc.c:1363-1496, including the orientation boundary condition);
  the observables: the mean x velocity of a band of cells and the mean
    order parameter of the cells with more than one particle.

The draws are those the oracle consumed in the same step with the
counter-based generator (recorded in the keyed layout), so the CPU run and a
GPU run with the same seed see the same variates; the discrete decisions
(Metropolis acceptances, dipole side, wall events) are taken from the
torch values, which equal the oracle's to rounding: forward() checks the
torch state against the oracle's after every step.
"""
import math

import numpy as np
import torch

import pyoracle as P

TOL = 1e-8  # definitions.h:57
PI = 3.14159265358979323846
NBOUNCE = 50


def _f(x):
    return torch.as_tensor(x, dtype=torch.float64)


def _val(x):
    """the value of a float or a tensor"""
    return float(x.detach()) if torch.is_tensor(x) else float(x)


def _norm2(x):
    """normalise rows in place of a copy (normn: only when the length is not ~0)"""
    lng = torch.sqrt((x * x).sum(dim=-1, keepdim=True))
    ok = (lng.detach() - 0.0).abs() >= TOL
    return torch.where(ok, x / torch.where(ok, lng, torch.ones_like(lng)), x)


def _rodrigues(vec, ax, theta):
    """rodriguesRotation (mtools.c:1572-1584) for rows of vec (3 columns)"""
    ax = _norm2(ax)
    oldl = torch.sqrt((vec * vec).sum(-1, keepdim=True))
    cp = torch.cross(ax, vec, dim=-1)
    dp = (ax * vec).sum(-1, keepdim=True)
    th = theta.unsqueeze(-1)
    out = vec * torch.cos(th) + cp * torch.sin(th) + ax * dp * (1.0 - torch.cos(th))
    newl = torch.sqrt((out * out).sum(-1, keepdim=True))
    return out * (oldl / newl)


def _ref_atan2(y, x):
    """the reference's atan2 (mtools.c:342-356), y >= 0 here"""
    xv, yv = x.detach(), y.detach()
    safe = torch.where(xv != 0.0, x, torch.ones_like(x))
    at = torch.atan(y / safe)
    out = torch.where(xv > 0.0, at, torch.zeros_like(at))
    out = torch.where((xv < 0.0) & (yv >= 0.0), at + PI, out)
    out = torch.where((xv < 0.0) & (yv < 0.0), at - PI, out)
    out = torch.where((xv.abs() <= TOL) & (xv == 0.0) & (yv > 0.0), torch.full_like(at, 0.5 * PI), out)
    return out


def _ori_bc(u, W, n):
    """oriBC (lc.c:1290-1393) on rows of u (3 columns), wall normal n (3)"""
    u = _norm2(u)
    sgn = torch.where((u * n).sum(-1, keepdim=True).detach() >= 0.0, 1.0, -1.0)
    u = u * sgn
    un = (u * n).sum(-1, keepdim=True) * n
    ut = u - un
    if abs(W["MUN"]) <= TOL and abs(W["MUT"]) <= TOL:
        u = _f(np.asarray(W["DUxyz"], dtype=float)).expand_as(u).clone()
    else:
        u = un * W["MUN"] + ut * W["MUT"]
        u = u * _f(np.asarray(W["MUxyz"], dtype=float))
    return _norm2(u)


class Problem:
    """one 2D active nematic replica: cfg in the oracle's format"""

    def __init__(self, cfg):
        assert cfg["dim"] == 2 and cfg["lc"] == 1 and cfg["rtech"] in (16, 21) and cfg["tstech"] == 0
        assert len(cfg["species"]) == 1
        self.cfg = cfg
        self.L = [int(x) for x in cfg["L"]]
        self.nx, self.ny = self.L[0] + 1, self.L[1] + 1
        self.nc = self.nx * self.ny
        self.dt, self.kbt, self.gnd = cfg["dt"], cfg["kbt"], cfg["gndnst"]
        sp = cfg["species"][0]
        self.sp = sp
        self.walls = []
        for w in cfg["walls"]:
            d = dict(A=np.asarray(w["A"], float), R=float(w.get("R", 0.0)), DN=float(w.get("DN", 0.0)),
                     DT=float(w.get("DT", 0.0)), DVN=float(w.get("DVN", 0.0)), DVT=float(w.get("DVT", 0.0)),
                     MVN=float(w.get("MVN", 1.0)), MVT=float(w.get("MVT", 1.0)), MUN=float(w.get("MUN", 1.0)),
                     MUT=float(w.get("MUT", 1.0)), MUxyz=np.asarray(w.get("MUxyz", (1, 1, 1)), float),
                     DUxyz=np.asarray(w.get("DUxyz", (0, 0, 0)), float), DVxyz=np.asarray(w.get("DVxyz", (0, 0, 0)), float),
                     V=np.asarray(w.get("V", (0, 0, 0)), float), INV=int(w.get("INV", 0)),
                     PHANTOM=int(w.get("PHANTOM", 0)))
            assert not d["V"].any() and not d["INV"]
            d["anch"] = abs(d["MUN"]) <= TOL or abs(d["MUT"]) <= TOL
            d["axis"] = int(np.nonzero(d["A"])[0][0])
            self.walls.append(d)
        self.pbc = [any(w["axis"] == j and w["DN"] > 0 for w in self.walls) for j in range(2)]

    # -- the step -----------------------------------------------------------------------------------------------------
    def step(self, Q, V, U, g, act, rec):
        """one step from particle-order tensors (N x 3); rec holds what the
        oracle's step fixed: shift, list order and cell offsets of the shifted
        bin, wall positions, draws"""
        dt, kbt, L = self.dt, self.kbt, self.L
        sp = self.sp
        m = float(sp["mass"])
        s = rec["shift"]
        # gridShift_all(+)
        cols = []
        for j in range(3):
            q = Q[:, j]
            if j < 2:
                q = q + s[j]
                wrap = (q.detach() > L[j]) & (q.detach() <= L[j] + s[j])
                q = torch.where(wrap, q - L[j], q)
            cols.append(q)
        Qs = torch.stack(cols, dim=1)
        order, cst = rec["order"], rec["cstart"]
        perm = torch.as_tensor(order, dtype=torch.long)
        N = perm.numel()
        pop = np.diff(cst)
        cell = torch.as_tensor(np.repeat(np.arange(self.nc), pop), dtype=torch.long)
        q, v, u = Qs[perm], V[perm], U[perm]
        # acc_all
        v = torch.cat([(v[:, 0] + dt * g).unsqueeze(1), v[:, 1:]], dim=1)
        # localPROP
        popt = torch.as_tensor(pop, dtype=torch.float64)
        mass = popt * m
        msafe = torch.where(mass > 0, mass, torch.ones_like(mass))
        zc = torch.zeros(self.nc, 3, dtype=torch.float64)
        vcm = zc.index_add(0, cell, m * v) / msafe.unsqueeze(1)
        vcm = torch.cat([vcm[:, :2], torch.zeros(self.nc, 1, dtype=torch.float64)], dim=1)
        cm = zc.index_add(0, cell, m * q) / msafe.unsqueeze(1)
        cm = torch.cat([cm[:, :2], torch.zeros(self.nc, 1, dtype=torch.float64)], dim=1)
        r = q - cm[cell]
        I2 = torch.zeros(self.nc, dtype=torch.float64).index_add(0, cell, m * (r[:, 0] ** 2 + r[:, 1] ** 2))
        n_gt1 = torch.as_tensor(pop > 1)
        psafe = torch.as_tensor(np.maximum(pop, 1), dtype=torch.float64)
        sxx = torch.zeros(self.nc, dtype=torch.float64).index_add(0, cell, u[:, 0] * u[:, 0]) * (2.0 / psafe) - 1.0
        sxy = torch.zeros(self.nc, dtype=torch.float64).index_add(0, cell, u[:, 0] * u[:, 1]) * (2.0 / psafe)
        syy = torch.zeros(self.nc, dtype=torch.float64).index_add(0, cell, u[:, 1] * u[:, 1]) * (2.0 / psafe) - 1.0
        S, dirv = self._eig2(sxx, sxy, syy, n_gt1)
        S = torch.where(S.detach() > 1.0, torch.ones_like(S), S)
        S = torch.where(n_gt1, S, torch.zeros_like(S))
        dirv = torch.where(n_gt1.unsqueeze(1), dirv, torch.zeros_like(dirv))
        # localVelGrad from the cells' VCM before the ghost particles
        E = self._velgrad(vcm)
        # ghostPart + strong anchoring at phantom walls
        vcm, S, dirv, u = self._ghost(vcm, S, dirv, u, pop, cst, rec)
        # LCcollision + jefferysTorque
        u, Tq = self._lc(u, cell, pop, S, dirv, E, rec)
        # dipoleAndersenROT_LC
        v = self._collide(q, v, u, Tq, cell, pop, cst, vcm, cm, I2, dirv, act, rec)
        # gridShift_all(-)
        cols = []
        for j in range(3):
            qq = q[:, j]
            if j < 2:
                sg = -1.0 * s[j]
                qq = qq + sg
                wrap = (qq.detach() < 0.0) & (qq.detach() >= sg)
                qq = torch.where(wrap, qq + L[j], qq)
            cols.append(qq)
        q = torch.stack(cols, dim=1)
        # back to particle order
        inv = torch.empty_like(perm)
        inv[perm] = torch.arange(N)
        Q1, V1, U1 = q[inv], v[inv], u[inv]
        # stream_all + MPC_BCcollision
        return self._stream(Q1, V1, U1, rec["wq_b"])

    def _eig2(self, a, b, d, counted):
        """largest eigenvalue and its eigenvector (eigenvalues2x2, eigenvectors2x2) of [[a, b], [b, d]]
        (the branch m10 != 0, which every counted cell takes here)"""
        det = a * d - b * b
        tr = a + d
        sq = torch.sqrt(torch.clamp(tr * tr * 0.25 - det, min=0.0))
        e0 = tr * 0.5 + sq
        assert bool((b.detach().abs() >= TOL)[counted].all()), "an axis-aligned order tensor"
        v0 = torch.stack([e0 - d, b], dim=1)
        lng = torch.sqrt((v0 * v0).sum(1, keepdim=True))
        ok = lng.detach() >= TOL
        v0 = torch.where(ok, v0 / torch.where(ok, lng, torch.ones_like(lng)), v0)
        return e0, torch.cat([v0, torch.zeros(v0.shape[0], 1, dtype=torch.float64)], dim=1)

    def _idx(self, a, b):
        return a * self.ny + b

    def _velgrad(self, vcm):
        X, Y = self.L[0], self.L[1]
        uv = [(1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]
        w = [1. / 9.] * 4 + [1. / 36.] * 4
        A = np.arange(self.nx)[:, None].repeat(self.ny, 1)
        B = np.arange(self.ny)[None, :].repeat(self.nx, 0)
        inner = (A >= 1) & (A < X) & (B >= 1) & (B < Y)
        ia, ib = A[inner], B[inner]
        ci = torch.as_tensor(self._idx(ia, ib), dtype=torch.long)
        E = torch.zeros(self.nc, 3, 3, dtype=torch.float64)
        Ein = torch.zeros(len(ia), 3, 3, dtype=torch.float64)
        for i in range(2):
            for j in range(2):
                acc = torch.zeros(len(ia), dtype=torch.float64)
                for k in range(8):
                    nb = torch.as_tensor(self._idx(ia + uv[k][0], ib + uv[k][1]), dtype=torch.long)
                    acc = acc + w[k] * uv[k][j] * vcm[nb, i]
                Ein[:, i, j] = acc * 3.
        E = E.index_put((ci,), Ein)
        # edge copies: x faces, then y faces
        for b in range(Y):
            E = E.index_put((torch.tensor([self._idx(0, b)]),), E[self._idx(1, b)].unsqueeze(0))
            E = E.index_put((torch.tensor([self._idx(X, b)]),), E[self._idx(X - 1, b)].unsqueeze(0))
        for a in range(X):
            E = E.index_put((torch.tensor([self._idx(a, 0)]),), E[self._idx(a, 1)].unsqueeze(0))
            E = E.index_put((torch.tensor([self._idx(a, Y)]),), E[self._idx(a, Y - 1)].unsqueeze(0))
        return E

    def _wall_W(self, W, wq, pts):
        """calcW_PLANE (bc.c:306) of rows of pts"""
        Wv = torch.zeros(pts.shape[0], dtype=torch.float64)
        for i in range(2):
            Wv = Wv + W["A"][i] * (pts[:, i] - wq[i])
        return Wv - W["R"]

    def _ghost(self, vcm, S, dirv, u, pop, cst, rec):
        wq_s = rec["wq_s"]
        gdraw = rec["ghost"]
        X, Y = self.L[0], self.L[1]
        vcm_rows, S_rows, dir_rows = list(vcm), list(S), list(dirv)
        u_new = u
        for c in range(self.nc):
            if pop[c] <= 0:
                continue
            a, b = divmod(c, self.ny)
            corners = np.array([[a + (j & 1), b + ((j >> 1) & 1)] for j in range(4)], dtype=float)
            numBC, wallindex = 0, -1
            for wi, W in enumerate(self.walls):
                if not (W["PHANTOM"] and W["anch"]):
                    continue
                if any((W["A"][0] * (p[0] - wq_s[wi][0]) + W["A"][1] * (p[1] - wq_s[wi][1])) - W["R"] < 0.0
                       for p in corners):
                    numBC += 1
                    wallindex = wi
            flagW = 0
            b0, e0 = cst[c], cst[c + 1]
            for wi, W in enumerate(self.walls):
                if not W["PHANTOM"]:
                    continue
                if any((W["A"][0] * (p[0] - wq_s[wi][0]) + W["A"][1] * (p[1] - wq_s[wi][1])) - W["R"] < 0.0
                       for p in corners):
                    flagW = 1
                if flagW and numBC == 1 and wallindex == wi and W["anch"]:
                    n = _f(np.concatenate([W["A"][:2] / np.linalg.norm(W["A"][:2]), [0.0]]))
                    idx = torch.arange(b0, e0)
                    u_new = u_new.index_put((idx,), _ori_bc(u_new[b0:e0], W, n))
                    S_rows[c] = _f(1.0)
                    d3 = _norm2(torch.cat([u_new[b0, :2], _f([0.0])]).unsqueeze(0))[0]
                    dir_rows[c] = d3
            if flagW and pop[c] < self.gnd and pop[c] > 0:
                invN = 1.0 / (self.gnd - pop[c])
                sigma = math.sqrt(self.kbt / invN)
                row = vcm_rows[c]
                comps = []
                for k in range(3):
                    if k < 2:
                        comps.append((row[k] + sigma * gdraw[c * 3 + k]) / self.gnd)
                    else:
                        comps.append(row[k])
                vcm_rows[c] = torch.stack(comps)
        return torch.stack(vcm_rows), torch.stack(S_rows), torch.stack(dir_rows), u_new

    def _lc(self, u, cell, pop, S, dirv, E, rec):
        sp, dt = self.sp, self.dt
        lcd, stride = rec["lc"], rec["lc_stride"]
        N = u.shape[0]
        coll = torch.as_tensor(pop > 1)[cell]
        DIR, Sv, Ec = dirv[cell], S[cell], E[cell]
        xaxis = _f([1.0, 0.0, 0.0]).expand(N, 3)
        rotAx = _norm2(torch.cross(DIR, xaxis, dim=-1))
        cosang = (DIR[:, :2] * xaxis[:, :2]).sum(1) / torch.sqrt((DIR[:, :2] ** 2).sum(1)) / 1.0
        rotAngle = -1.0 * torch.acos(cosang)
        effM = float(sp["mfpot"]) * 0.5 * 2.0
        KBT = self.kbt
        # Maier-Saupe, 2D Metropolis (lc.c:341-411)
        c3 = DIR[:, 0] * u[:, 1] - DIR[:, 1] * u[:, 0]
        ang0 = _ref_atan2(torch.abs(c3), DIR[:, 0] * u[:, 0] + DIR[:, 1] * u[:, 1])
        redraw = ((ang0.detach() - 0.5 * PI).abs() <= TOL) | ((ang0.detach() + 0.5 * PI).abs() <= TOL)
        lc = lambda t, j: float(lcd[t * stride + j])  # noqa: E731
        ang_fin = np.zeros(N)
        uz_fin = np.zeros(N)
        cnt = np.zeros(N, int)
        uz0_val = (DIR[:, 0] * u[:, 0] + DIR[:, 1] * u[:, 1]).detach().numpy()
        ang0_val = ang0.detach().numpy().copy()
        Sval = Sv.detach().numpy()
        for t in range(N):
            if not bool(coll[t]):
                continue
            a0 = PI * (0.5 - lc(t, 0)) if bool(redraw[t]) else ang0_val[t]
            uz = uz0_val[t]
            nann = int(5 + 0.1 * effM * Sval[t] / KBT)
            k = 0
            for i in range(nann):
                w0 = effM * Sval[t] * uz * uz / KBT
                ang1 = PI * (0.5 - lc(t, 1 + 2 * i))
                uz1 = math.cos(ang1)
                w1 = effM * Sval[t] * uz1 * uz1 / KBT
                dw = w1 - w0
                if dw >= 0.0 or lc(t, 2 + 2 * i) <= math.exp(dw):
                    k += 1
                    a0, uz = ang1, uz1
            ang_fin[t], uz_fin[t], cnt[t] = a0, uz, k
        moved = torch.as_tensor(cnt > 0)
        newU = torch.stack([_f(uz_fin), torch.sin(_f(ang_fin)), torch.zeros(N, dtype=torch.float64)], dim=1)
        newU = _rodrigues(newU, rotAx, rotAngle)
        U1 = torch.where((moved & coll).unsqueeze(1), newU, u)
        dU = (U1 - u) / dt
        Tq = torch.cross(U1, dU, dim=-1) * float(sp["rfc"])
        # jefferysTorque (lc.c:559-718) with the velocity gradient
        D = 0.5 * (Ec + Ec.transpose(1, 2))   # D[j][i] = (E_ij + E_ji) / 2
        Wm = 0.5 * (Ec - Ec.transpose(1, 2))  # W[j][i] = (E_ij - E_ji) / 2, so W_ab = (E_ba - E_ab) / 2
        Wm = Wm.transpose(1, 2)
        uDu = torch.einsum("ni,nik,nk->n", U1, D, U1)
        uD = torch.einsum("ni,nij->nj", U1, D)
        uw = torch.einsum("ni,nij->nj", U1, Wm)
        dudt = uw + float(sp["tumble"]) * (uD - U1 * uDu.unsqueeze(1))
        w = torch.cross(U1, dudt, dim=-1) * float(sp["chihi"]) * dt
        theta = torch.sqrt((w * w).sum(1))
        w = _norm2(w)
        U2 = _rodrigues(U1, w, theta)
        U2 = torch.where(coll.unsqueeze(1), U2, u)
        Tq = torch.where(coll.unsqueeze(1), Tq, torch.zeros_like(Tq))
        return U2, Tq

    def _collide(self, q, v, u, Tq, cell, pop, cst, vcm, cm, I2, dirv, act, rec):
        sp, dt, kbt = self.sp, self.dt, self.kbt
        m = float(sp["mass"])
        sig = math.sqrt(kbt / m)
        nrm = torch.as_tensor(np.asarray(rec["part"]).reshape(-1, 3), dtype=torch.float64)
        N = q.shape[0]
        coll = torch.as_tensor(pop > 1)[cell]
        n_c = torch.as_tensor(pop, dtype=torch.float64)
        # the cell's activity (lc.c:1811-1841): ACT / m summed over its particles, times dt
        ACT = (n_c * act / m)
        if self.cfg["rtech"] in (20, 21):
            sigW = float(sp["sigwidth"])
            sigP = float(sp["sigpos"])
            fall = 1.0 - torch.tanh((n_c - self.gnd * (1 + sigP)) / (self.gnd * sigW))
            ACT = ACT * fall / 2
        ACT = ACT * dt
        PA = dirv
        rq = q - cm[cell]
        pW = (PA[cell][:, :2] * rq[:, :2]).sum(1)
        pm = torch.where(pW.detach() <= 0.0, -1.0, 1.0).to(torch.float64)
        Mc = torch.where(n_c > 0, n_c * m, torch.ones_like(n_c))
        rv = sig * nrm
        rv = torch.cat([rv[:, :2], torch.zeros(N, 1, dtype=torch.float64)], dim=1)
        zc = torch.zeros(self.nc, 3, dtype=torch.float64)
        RS = zc.index_add(0, cell, m * rv) / Mc.unsqueeze(1)
        ACTs = ACT[cell]
        AS = zc.index_add(0, cell, m * (pm * ACTs).unsqueeze(1) * PA[cell]) / Mc.unsqueeze(1)
        dv = m * (v - rv - (pm * ACTs).unsqueeze(1) * PA[cell])
        dv = torch.cat([dv[:, :2], torch.zeros(N, 1, dtype=torch.float64)], dim=1)
        am = torch.cross(torch.cat([rq[:, :2], torch.zeros(N, 1, dtype=torch.float64)], 1), dv, dim=-1)
        Lz = torch.zeros(self.nc, dtype=torch.float64).index_add(0, cell, am[:, 2] + Tq[:, 2] * dt)
        Wz = Lz / torch.where(I2 != 0, I2, torch.ones_like(I2))
        at = torch.stack([-Wz[cell] * rq[:, 1], Wz[cell] * rq[:, 0], torch.zeros(N, dtype=torch.float64)], 1)
        vnew = vcm[cell] + rv - RS[cell] + at + (pm * ACTs).unsqueeze(1) * PA[cell] - AS[cell] - float(sp["damp"]) * vcm[cell]
        vnew = torch.cat([vnew[:, :2], v[:, 2:]], dim=1)
        return torch.where(coll.unsqueeze(1), vnew, v)

    def _stream(self, Q, V, U, wq_b):
        dt, m = self.dt, float(self.sp["mass"])
        Q = Q + torch.cat([dt * V[:, :2], torch.zeros(Q.shape[0], 1, dtype=torch.float64)], dim=1)
        hit = np.zeros(Q.shape[0], bool)
        for wi, W in enumerate(self.walls):
            hit |= (self._wall_W(W, wq_b[wi], Q).detach().numpy() < 0.0)
        rows_q, rows_v, rows_u = list(Q), list(V), list(U)
        for i in np.nonzero(hit)[0]:
            q, vv, uu = Q[i], V[i], U[i]
            t_left, t_coll = dt, 0.0
            cnt = 0
            while True:
                tmin, chW, chosen = t_left, 1.0, 0
                for wi, W in enumerate(self.walls):
                    tw = float(self._wall_W(W, wq_b[wi], q.unsqueeze(0))[0].detach())
                    if tw < 0.0:
                        tc = W["R"] + sum(W["A"][j] * (wq_b[wi][j] - q[j]) for j in range(2))
                        tc = tc / (W["A"][0] * (-vv[0]) + W["A"][1] * (-vv[1]))
                        t1 = t_left - tc
                        if _val(t1) < _val(tmin):
                            tmin, chosen, chW = t1, wi, tw
                t_coll = tmin
                if chW > -TOL:
                    break
                cnt += 1
                assert cnt <= NBOUNCE, "a particle bounced more than NBOUNCE times"
                W = self.walls[chosen]
                if _val(t_coll) > -TOL:
                    q = q - (t_left - t_coll) * vv
                    n = _f(np.concatenate([W["A"][:2] / np.linalg.norm(W["A"][:2]), [0.0]]))
                    Vr = vv
                    VN = (Vr * n).sum() * n
                    VT = Vr - VN
                    VN = VN * W["MVN"] + W["DVN"] * n
                    VT = VT * W["MVT"] + W["DVT"]
                    Vr = VN + VT + _f(W["DVxyz"])
                    Vr = torch.cat([Vr[:2], vv[2:]])
                    Vr = (Vr - vv) * m
                    vv = vv + Vr * 1.0 * (1.0 / m)
                    PN = (q * n).sum() * n
                    PT = q - PN
                    PN = PN + W["DN"] * n * (1. - 2. * TOL)
                    PT = PT + W["DT"]
                    q = torch.cat([(PN + PT)[:2], q[2:]])
                    uu = _ori_bc(uu.unsqueeze(0), W, n)[0]
                    t_left = t_left - t_coll
                    q = q + t_left * torch.cat([vv[:2], _f([0.0])])
                    if _val(self._wall_W(W, wq_b[chosen], q.unsqueeze(0))[0]) < -TOL:
                        q = q - t_left * torch.cat([vv[:2], _f([0.0])])
                else:
                    q = q - (t_left - t_coll) * torch.cat([vv[:2], _f([0.0])])
            rows_q[i], rows_v[i], rows_u[i] = q, vv, uu
        return torch.stack(rows_q), torch.stack(rows_v), torch.stack(rows_u)

    # -- observables --------------------------------------------------------------------------------------------------
    def band_velocity(self, Q, V, lo, hi):
        y = np.trunc(Q[:, 1].detach().numpy())
        mk = torch.as_tensor((y >= lo) & (y < hi))
        return V[mk, 0].mean()

    def mean_S(self, Q, U):
        a = np.trunc(Q[:, 0].detach().numpy()).astype(int)
        b = np.trunc(Q[:, 1].detach().numpy()).astype(int)
        cell = torch.as_tensor(a * self.ny + b, dtype=torch.long)
        pop = np.bincount(cell.numpy(), minlength=self.nc)
        psafe = torch.as_tensor(np.maximum(pop, 1), dtype=torch.float64)
        z = torch.zeros(self.nc, dtype=torch.float64)
        sxx = z.index_add(0, cell, U[:, 0] * U[:, 0]) * (2.0 / psafe) - 1.0
        sxy = z.index_add(0, cell, U[:, 0] * U[:, 1]) * (2.0 / psafe)
        syy = z.index_add(0, cell, U[:, 1] * U[:, 1]) * (2.0 / psafe) - 1.0
        tr = sxx + syy
        det = sxx * syy - sxy * sxy
        S = tr * 0.5 + torch.sqrt(tr * tr * 0.25 - det)
        S = torch.where(S.detach() > 1.0, torch.ones_like(S), S)
        keep = torch.as_tensor(pop > 1)
        return S[keep].mean()


def run(cfg, Q0, V0, U0, K, seed, obs, check=True):
    """J = (1/K) sum_t obs_t and the torch leaves (g, act): J.backward() gives
    the FP64 CPU adjoint.  The oracle runs beside it (Philox draws, recorded)
    and fixes the step's list order; with check, the torch state must equal
    the oracle's after every step (1e-12)."""
    pb = Problem(cfg)
    g = torch.tensor(float(cfg["grav"][0]), dtype=torch.float64, requires_grad=True)
    act = torch.tensor(float(cfg["species"][0]["act"]), dtype=torch.float64, requires_grad=True)
    s = P.System(cfg, Q0, V0, U0)
    d = P.Draws(P.DRAW_PHILOX, seed)
    stride = s.lc_stride()
    N = Q0.shape[0]
    Q, V, U = _f(Q0), _f(V0), _f(U0)
    J = _f(0.0)
    worst = 0.0
    for _ in range(K):
        rec = dict(shift=np.zeros(3), cell=np.zeros(pb.nc * 4), ghost=np.zeros(pb.nc * 3), part=np.zeros(N * 3),
                   lc=np.zeros(N * stride), lc_stride=stride)
        d.set_record(rec["shift"], rec["cell"], rec["ghost"], rec["part"], rec["lc"], stride)
        rec["wq_prev"] = s.walls().copy()
        s.phase("bin", d)
        rec["shift"] = s.last_shift()
        rec["order"], rec["cstart"] = s.order()
        rec["wq_s"] = s.walls().copy()
        s.phase("collide", d)
        rec["wq_b"] = s.walls().copy()
        s.phase("stream", d)
        Q, V, U = pb.step(Q, V, U, g, act, rec)
        if check:
            Qo, Vo, Uo = s.state()
            for a, b in ((Q, Qo), (V, Vo), (U, Uo)):
                err = np.abs(a.detach().numpy() - b)
                lim = 1e-12 + 1e-12 * np.abs(b)
                worst = max(worst, float((err / lim).max()))
        if obs[0] == "band":
            J = J + pb.band_velocity(Q, V, *obs[1]) / K
        else:
            J = J + pb.mean_S(Q, U) / K
    return J, g, act, worst
