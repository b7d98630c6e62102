"""Pure-PyTorch FP64 transcription of the DG right-hand side and the SSP-RK
stage, written for autograd (test infrastructure: the checker of the HIP
adjoint).

It follows the same reference routines as the C oracle -- flux_calc_pnts_r
and flux_hllc (src/flux.f90:23-193,466-647), calc_flux_x/y/z with the
boundary ghosts (src/flux.f90:197-450, src/boundary.f90:40-338),
innerintegral2 and glflux (src/flux.f90:751-1022), source_calc
(src/sources.f90:12-84), the limiter (src/flux.f90:1077-1164) and
advance_time_level (src/integrator.f90:125-252) -- vectorised over cells,
with the basis tables taken from the oracle.  Its forward values are pinned
against the oracle in tests/test_oracle_pins.py; autograd through it is then
the reference for the hand-written vector-Jacobian products.

The limiter is applied with its factors computed under no_grad (the active
set frozen, as the adjoint does).  State layout: (nbasis, nQ, nz, ny, nx).
"""
from __future__ import annotations

import numpy as np
import torch

RH, MX, MY, MZ, EN = 0, 1, 2, 3, 4
PXX, PYY, PZZ, PXY, PXZ, PYZ = 5, 6, 7, 8, 9, 10
# 13-moment model: traceless stress without pzz (= -pxx - pyy), heat flux
T_PXX, T_PYY, T_PXY, T_PXZ, T_PYZ, T_QX = 5, 6, 7, 8, 9, 10
BC = {"periodic": 0, "outflow": 1, "wall": 2, "noslip": 3, "inflow": 4}


def degrees(nb):
    """Legendre degrees (x, y, z) of the modes 1,x,y,z,yz,zx,xy,(xyz | P2x,P2y,P2z)."""
    d = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0)]
    if nb == 8:
        d.append((1, 1, 1))
    elif nb == 10:
        d += [(2, 0, 0), (0, 2, 0), (0, 0, 2)]
    return d[:nb]


def _L(a, x):
    return np.ones_like(x) if a == 0 else (x if a == 1 else 1.5 * x * x - 0.5)


def _dL(a, x):
    return np.zeros_like(x) if a == 0 else (np.ones_like(x) if a == 1 else 3.0 * x)


class DG:
    def __init__(self, cfg, tables, dtype=torch.float64, device="cpu", inflow=None):
        """inflow: x-low inflow state (nQ, nface, nz, ny) when cfg.bc[0] is "inflow"."""
        self.cfg = cfg
        self.inflow = None if inflow is None else torch.as_tensor(np.asarray(inflow), dtype=dtype, device=device)
        self.nb, self.nQ = cfg.nbasis, cfg.nQ
        q = cfg.iquad
        t = tables
        T = lambda a: torch.as_tensor(np.asarray(a), dtype=dtype, device=device)  # noqa: E731
        self.bint = T(t["bint"])                    # (npg, nb)
        self.w2, self.w3, self.cb = T(t["w2"]), T(t["w3"]), T(t["cbasis"])
        self.bf = {("x", 1): T(t["bxp"]), ("x", 0): T(t["bxm"]), ("y", 1): T(t["byp"]), ("y", 0): T(t["bym"]),
                   ("z", 1): T(t["bzp"]), ("z", 0): T(t["bzm"])}
        # derivative tables d(phi_r)/d(xi_d) at the volume points (x slowest, z fastest)
        xq = np.asarray(t["xq"])
        X, Y, Z = np.meshgrid(xq, xq, xq, indexing="ij")
        X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
        deg = degrees(self.nb)
        D = np.zeros((3, q ** 3, self.nb))
        for r, (a, b, c) in enumerate(deg):
            D[0, :, r] = _dL(a, X) * _L(b, Y) * _L(c, Z)
            D[1, :, r] = _L(a, X) * _dL(b, Y) * _L(c, Z)
            D[2, :, r] = _L(a, X) * _L(b, Y) * _dL(c, Z)
        self.dphi = T(D)
        self.dxi = (cfg.nx / cfg.lx, cfg.ny / cfg.ly, cfg.nz / cfg.lz)
        g = cfg.gamma
        self.gamma, self.gm1 = g, g - 1.0
        self.P_floor = cfg.T_floor * cfg.rh_floor
        self.rh_floor = cfg.rh_floor
        self.colvis = cfg.colvis
        self.colq = getattr(cfg, "col_q", 0.0)
        self.bc = [BC[b] for b in cfg.bc]

    # ----------------------------------------------------------- points ----
    def prim(self, Q):
        ri = 1.0 / Q[RH]
        u = [Q[MX] * ri, Q[MY] * ri, Q[MZ] * ri]
        P = self.gm1 * (Q[EN] - 0.5 * ri * (Q[MX] ** 2 + Q[MY] ** 2 + Q[MZ] ** 2))
        P = torch.where(P < self.P_floor, torch.full_like(P, self.P_floor), P)
        return ri, u, P

    def flux13(self, Q, d):
        """13-moment flux in direction d: the relaxation model's stress rows for
        the five stored components (pzz = -pxx - pyy) and the Maxwell-Cattaneo
        heat flux (flux col_q*T of q_d; q_d in the energy flux)."""
        ri, u, P = self.prim(Q)
        un = u[d]
        tzz = -(Q[T_PXX] + Q[T_PYY])
        rows = [(Q[T_PXX], Q[T_PXY], Q[T_PXZ]), (Q[T_PXY], Q[T_PYY], Q[T_PYZ]), (Q[T_PXZ], Q[T_PYZ], tzz)][d]
        F = [None] * 13
        F[RH] = Q[MX + d]
        for a in range(3):
            F[MX + a] = Q[MX + a] * un + rows[a] + (P if a == d else 0.0)
        F[EN] = (Q[EN] + P) * un + rows[0] * u[0] + rows[1] * u[1] + rows[2] * u[2] + Q[T_QX + d]
        c4, c2, cv = 4.0 / 3.0 * self.colvis, 2.0 / 3.0 * self.colvis, self.colvis
        z = torch.zeros_like(un)
        F[T_PXX] = c4 * un if d == 0 else -c2 * un
        F[T_PYY] = c4 * un if d == 1 else -c2 * un
        F[T_PXY] = cv * u[1] if d == 0 else (cv * u[0] if d == 1 else z)
        F[T_PXZ] = cv * u[2] if d == 0 else (cv * u[0] if d == 2 else z)
        F[T_PYZ] = cv * u[2] if d == 1 else (cv * u[1] if d == 2 else z)
        for a in range(3):
            F[T_QX + a] = self.colq * P * ri if a == d else z
        return torch.stack(F)

    def flux(self, Q, d):
        """Q: (nQ, ...) -> F (nQ, ...) in direction d (flux_calc_pnts_r, ivis 0/1)."""
        if self.nQ == 13:
            return self.flux13(Q, d)
        ri, u, P = self.prim(Q)
        un = u[d]
        F = [None] * self.nQ
        if self.nQ >= 11:
            rows = [(PXX, PXY, PXZ), (PXY, PYY, PYZ), (PXZ, PYZ, PZZ)][d]
            Pn = [Q[rows[0]], Q[rows[1]], Q[rows[2]]]
        else:
            Pn = [torch.zeros_like(un)] * 3
        F[RH] = Q[MX + d]
        for a in range(3):
            F[MX + a] = Q[MX + a] * un + Pn[a] + (P if a == d else 0.0)
        F[EN] = (Q[EN] + P) * un + Pn[0] * u[0] + Pn[1] * u[1] + Pn[2] * u[2]
        if self.nQ >= 11:
            c4, c2, cv = 4.0 / 3.0 * self.colvis, 2.0 / 3.0 * self.colvis, self.colvis
            z = torch.zeros_like(un)
            F[PXX] = c4 * un if d == 0 else -c2 * un
            F[PYY] = c4 * un if d == 1 else -c2 * un
            F[PZZ] = c4 * un if d == 2 else -c2 * un
            F[PXY] = cv * u[1] if d == 0 else (cv * u[0] if d == 1 else z)
            F[PXZ] = cv * u[2] if d == 0 else (cv * u[0] if d == 2 else z)
            F[PYZ] = cv * u[2] if d == 1 else (cv * u[1] if d == 2 else z)
        return torch.stack(F)

    def riemann(self, QL, QR, d):
        FL, FR = self.flux(QL, d), self.flux(QR, d)
        out = 0.5 * (FL + FR)
        if self.cfg.riemann == "central":
            return out
        if self.cfg.riemann == "roe":
            return torch.cat([self.roe(QL, QR, FL, FR, d), out[5:]], dim=0)
        if self.cfg.riemann != "hllc":
            raise NotImplementedError("the transcription has central, HLLC and Roe fluxes")
        g, gm1 = self.gamma, self.gm1
        IP, I1, I2 = MX + d, MX + (d + 1) % 3, MX + (d + 2) % 3
        rL, rR = QL[RH], QR[RH]
        riL, riR = 1.0 / rL, 1.0 / rR
        uLv = [QL[MX + a] * riL for a in range(3)]
        uRv = [QR[MX + a] * riR for a in range(3)]
        uL, uR = uLv[d], uRv[d]
        pL = gm1 * (QL[EN] - 0.5 * rL * (uLv[0] ** 2 + uLv[1] ** 2 + uLv[2] ** 2))
        pR = gm1 * (QR[EN] - 0.5 * rR * (uRv[0] ** 2 + uRv[1] ** 2 + uRv[2] ** 2))
        cL, cR = torch.sqrt(g * pL * riL), torch.sqrt(g * pR * riR)
        SL = torch.minimum(uL - cL, uR - cR)
        SR = torch.maximum(uR + cR, uL + cL)
        aR, aL = rR * (SR - uR), rL * (SL - uL)
        den = aR - aL
        den = torch.where(den == 0, torch.full_like(den, self.rh_floor), den)
        SM = (aR * uR - aL * uL + pL - pR) / den
        sel = torch.full_like(SL, -1, dtype=torch.int64)
        sel = torch.where(SL > 0, 0, sel)
        sel = torch.where((SL <= 0) & (0 < SM), 1, sel)
        sel = torch.where((SM <= 0) & (0 <= SR), 2, sel)
        sel = torch.where(SR < 0, 3, sel)
        pstar = rL * (uL - SL) * (uL - SM) + pL

        def star(Qk, Fk, Sk, uk, pk):
            dd = Sk - SM
            dd = torch.where(dd == 0, torch.full_like(dd, self.rh_floor), dd)
            rid = 1.0 / dd
            sq = Sk - uk
            s = {RH: Qk[RH] * sq * rid, IP: (sq * Qk[IP] + pstar - pk) * rid, I1: sq * Qk[I1] * rid,
                 I2: sq * Qk[I2] * rid, EN: (sq * Qk[EN] - pk * uk + pstar * SM) * rid}
            return [Fk[f] + Sk * (s[f] - Qk[f]) for f in range(5)]

        sL = star(QL, FL, SL, uL, pL)
        sR = star(QR, FR, SR, uR, pR)
        rows = []
        for f in range(5):
            v = out[f]
            v = torch.where(sel == 0, FL[f], v)
            v = torch.where(sel == 1, sL[f], v)
            v = torch.where(sel == 2, sR[f], v)
            v = torch.where(sel == 3, FR[f], v)
            rows.append(v)
        return torch.cat([torch.stack(rows), out[5:]], dim=0)

    def roe(self, QL, QR, FL, FR, d):
        """Roe flux of rh..en (legacy Eulderink-Mellema form with half the
        thermal enthalpy, src/ces/10_moment/dg_3D_hydro_lqc_SEAN.f90:1614-1789);
        local Lax-Friedrichs where the Roe sound speed squared is not positive."""
        gm1, cp = self.gm1, self.gamma / self.gm1
        IP, I1, I2 = MX + d, MX + (d + 1) % 3, MX + (d + 2) % 3
        sk, sw = [], []
        for Q in (QL, QR):
            a, b, c = Q[IP] / Q[RH], Q[I1] / Q[RH], Q[I2] / Q[RH]
            vsq = a * a + b * b + c * c
            Pri = gm1 * (Q[EN] - 0.5 * Q[RH] * vsq)
            k = torch.sqrt(Q[RH])
            sk.append(k)
            sw.append([k * a, k * b, k * c, 0.5 * k * (vsq + cp * Pri / Q[RH])])
        Qri = 1.0 / QL[RH]
        delrho = QR[RH] * Qri - 1.0
        dele = (QR[EN] - QL[EN]) * Qri
        dm1, dm2, dm3 = ((QR[f] - QL[f]) * Qri for f in (IP, I1, I2))
        dsk = 0.5 * (sk[1] + sk[0])
        vc0, vc1, vc2, hc = (0.5 * (sw[1][j] + sw[0][j]) / dsk for j in range(4))
        vcsq = vc0 * vc0 + vc1 * vc1 + vc2 * vc2
        asq = gm1 * (hc - 0.5 * vcsq)
        ok = asq > 0
        asq_s = torch.where(ok, asq, torch.ones_like(asq))
        sc = torch.sqrt(asq_s)
        l0, l1, l2 = (vc0 - sc).abs(), (vc0 + sc).abs(), vc0.abs()
        ea2i = gm1 / asq_s
        vcdel = vc0 * dm1 + vc1 * dm2 + vc2 * dm3 - dele
        a9e = 0.5 * ea2i * (0.5 * vcsq * delrho - vcdel)
        sk9 = 0.5 * (dm1 - vc0 * delrho) / sc
        a0, a1 = a9e - sk9, a9e + sk9
        a2 = ea2i * ((hc - vcsq) * delrho + vcdel)
        a3, a4 = dm2 - vc1 * delrho, dm3 - vc2 * delrho
        t0, t1, t2, t3, t4 = a0 * l0, a1 * l1, a2 * l2, a3 * l2, a4 * l2
        dd = {RH: t0 + t1 + t2,
              EN: t0 * (hc - sc * vc0) + t1 * (hc + sc * vc0) + t2 * (0.5 * vcsq) + t3 * vc1 + t4 * vc2,
              IP: t0 * (vc0 - sc) + t1 * (vc0 + sc) + t2 * vc0,
              I1: (t0 + t1 + t2) * vc1 + t3, I2: (t0 + t1 + t2) * vc2 + t4}

        def cfcal(Q):
            ri = 1.0 / Q[RH]
            v = [Q[MX + a] * ri for a in range(3)]
            P = gm1 * (Q[EN] - 0.5 * Q[RH] * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
            return v[d].abs() + torch.sqrt(self.gamma * P * ri)
        cw = torch.fmax(cfcal(QL), cfcal(QR))
        rows = []
        for f in range(5):
            roe = 0.5 * (FL[f] + FR[f] - dd[f] * QL[RH])
            llf = 0.5 * (FL[f] + FR[f]) - 0.5 * cw * (QR[f] - QL[f])
            rows.append(torch.where(ok, roe, llf))
        return torch.stack(rows)

    # ------------------------------------------------------------- faces ----
    def _trace(self, q, d, side):
        """(nface, nQ, nz, ny, nx) trace of every cell on its minus (0) or plus (1) face."""
        return torch.einsum("pr,rfzyx->pfzyx", self.bf[("xyz"[d], side)], q)

    def _ghost(self, own, d, side):
        kind = self.bc[2 * d + side]
        ex = own.clone()
        if kind == BC["inflow"]:
            # x_bc_inflow (src/boundary.f90:571-592): own is (nface, nQ, nz, ny, 1)
            g = self.inflow.permute(1, 0, 2, 3).unsqueeze(-1)
            ex[:, RH] = g[:, RH]
            ex[:, MX] = g[:, MX]
            ex[:, MY] = 0.0
            ex[:, MZ] = 0.0
            ex[:, EN] = g[:, EN]
        elif kind == BC["noslip"]:
            ex[:, MX:MZ + 1] = 0.0
        elif kind == BC["wall"]:
            ex[:, MX + d] = 0.0
        elif kind == BC["outflow"]:
            n = own[:, MX + d]
            if d == 0 and side == 1:
                ex[:, MX] = torch.where(n > 0, -own[:, MX], own[:, MX])
            else:
                v = n.amax(dim=0, keepdim=True) if side == 0 else n.amin(dim=0, keepdim=True)
                cond = (v > 0) if side == 0 else (v < 0)
                ex[:, MX + d] = torch.where(cond.expand_as(n), torch.zeros_like(n), n)
        return ex

    def _face_fluxes(self, q, d):
        """Fluxes on the n+1 faces along axis d: (nface, nQ, ...) with that axis n+1 long."""
        ax = 4 - d  # tensor axis of direction d in (p, f, z, y, x)
        plus = self._trace(q, d, 1)    # each cell's + face
        minus = self._trace(q, d, 0)   # each cell's - face
        n = q.shape[2 + (2 - d)]
        if self.bc[2 * d] == BC["periodic"]:
            left = torch.cat([plus.narrow(ax, n - 1, 1), plus], dim=ax)
            right = torch.cat([minus, minus.narrow(ax, 0, 1)], dim=ax)
        else:
            lo_own = minus.narrow(ax, 0, 1)
            hi_own = plus.narrow(ax, n - 1, 1)
            left = torch.cat([self._ghost(lo_own, d, 0), plus], dim=ax)
            right = torch.cat([minus, self._ghost(hi_own, d, 1)], dim=ax)
        P = left.shape[0]
        shp = left.shape
        QL = left.permute(1, 0, 2, 3, 4).reshape(self.nQ, -1)
        QR = right.permute(1, 0, 2, 3, 4).reshape(self.nQ, -1)
        Fh = self.riemann(QL, QR, d)
        return Fh.reshape(self.nQ, P, *shp[2:]).permute(1, 0, 2, 3, 4)

    # --------------------------------------------------------------- rhs ----
    def rhs(self, q, coll=None):
        nb = self.nb
        cfg = self.cfg
        # volume integral
        Qg = torch.einsum("gr,rfzyx->gfzyx", self.bint, q)
        npg = Qg.shape[0]
        Qf = Qg.permute(1, 0, 2, 3, 4).reshape(self.nQ, -1)
        integ = 0.0
        for d in range(3):
            F = self.flux(Qf, d).reshape(self.nQ, npg, *q.shape[2:])
            integ = integ + self.dxi[d] * torch.einsum("g,gr,fgzyx->rfzyx", self.w3, self.dphi[d], F)
        integ = 0.25 * self.cb.view(nb, 1, 1, 1, 1) * integ
        # surface terms
        glf = 0.0
        for d in range(3):
            Fh = self._face_fluxes(q, d)
            ax = 4 - d
            n = q.shape[2 + (2 - d)]
            Fp = Fh.narrow(ax, 1, n)      # face i+1/2 of cell i
            Fm = Fh.narrow(ax, 0, n)      # face i-1/2
            bp, bm = self.bf[("xyz"[d], 1)], self.bf[("xyz"[d], 0)]
            s = torch.einsum("p,pr,pfzyx->rfzyx", self.w2, bp, Fp) - torch.einsum("p,pr,pfzyx->rfzyx", self.w2, bm, Fm)
            glf = glf + self.dxi[d] * s
        glf = 0.25 * self.cb.view(nb, 1, 1, 1, 1) * glf
        glf = torch.cat([glf[:1], glf[1:] - integ[1:]], dim=0)
        rhs = -glf
        # sources: explicit relaxation (ivis 0), body force
        src = torch.zeros_like(q)
        c = cfg.coll if coll is None else coll
        if self.nQ == 11 and cfg.ivis == 0 and cfg.coll != 0.0:
            src = torch.cat([src[:, :PXX], -c * q[:, PXX:]], dim=1)
        if self.nQ == 13 and cfg.ivis == 0:
            src = torch.cat([src[:, :PXX], -c * q[:, PXX:T_QX], -cfg.coll_q * q[:, T_QX:]], dim=1)
        bf = cfg.body_force
        if any(b != 0.0 for b in bf):
            rows = [src[:, f] for f in range(self.nQ)]
            for a in range(3):
                rows[MX + a] = rows[MX + a] + bf[a] * q[:, RH]
            rows[EN] = rows[EN] + bf[0] * q[:, MX] + bf[1] * q[:, MY] + bf[2] * q[:, MZ]
            src = torch.stack(rows, dim=1)
        return rhs + src

    # ----------------------------------------------------------- limiter ----
    def limit_factors(self, q):
        """Per-cell (floor, zero, th_rho, th_p) of the limiter, computed without gradients."""
        with torch.no_grad():
            nb = self.nb
            faces = torch.cat([self.bf[("x", 0)], self.bf[("x", 1)], self.bf[("y", 0)], self.bf[("y", 1)],
                               self.bf[("z", 0)], self.bf[("z", 1)]], dim=0)
            rho = q[:, RH].clone()
            floor = rho[0] < self.rh_floor
            rmin = torch.einsum("pr,rzyx->pzyx", faces, rho).amin(dim=0)
            th_rho = torch.ones_like(rmin)
            lim = rmin < self.rh_floor
            th = ((self.rh_floor - rho[0]) / (rmin - rho[0])).clamp(0.0, 1.0).abs()
            th_rho = torch.where(lim, th, th_rho)
            rho_l = torch.cat([rho[:1], th_rho * rho[1:]], dim=0)
            Pave = self.gm1 * (q[0, EN] - 0.5 * (q[0, MX] ** 2 + q[0, MY] ** 2 + q[0, MZ] ** 2) / rho_l[0])
            epsP = self.rh_floor * self.cfg.T_floor
            zero = floor | (Pave < epsP)
            e = [torch.einsum("pr,rzyx->pzyx", faces, rho_l)] + \
                [torch.einsum("pr,rzyx->pzyx", faces, q[:, f]) for f in (MX, MY, MZ, EN)]
            P = self.gm1 * (e[4] - 0.5 / e[0] * (e[1] ** 2 + e[2] ** 2 + e[3] ** 2))
            pmin = P.amin(dim=0)
            thp = ((Pave - epsP) / (Pave - pmin)).clamp(0.0, 1.0)
            th_p = torch.where((pmin < epsP) & (Pave != pmin), thp, torch.ones_like(thp))
            return floor, zero, th_rho, th_p

    def limit(self, q):
        """The limiter with frozen factors (differentiable in q, not in the factors)."""
        if not self.cfg.limiter:
            return q
        floor, zero, th_rho, th_p = self.limit_factors(q)
        out = q.clone()
        keep = (~zero).to(q.dtype)
        out[1:, RH] = q[1:, RH] * th_rho * th_p * keep
        for f in (MX, MY, MZ, EN):
            out[1:, f] = q[1:, f] * th_p * keep
        out[0, RH] = torch.where(floor, torch.full_like(q[0, RH], self.rh_floor), q[0, RH])
        return out

    # ------------------------------------------------------------- stage ----
    def euler(self, m, dt, coll=None):
        """v = m + dt*R(m), stress fields divided by 1 + dt*coll when ivis = 1."""
        v = m + dt * self.rhs(m, coll=coll)
        if self.cfg.ivis == 1 and self.nQ >= 11:
            c = self.cfg.coll if coll is None else coll
            faci = 1.0 / (1.0 + dt * c)
            if self.nQ == 13:
                faciq = 1.0 / (1.0 + dt * self.cfg.coll_q)
                v = torch.cat([v[:, :PXX], faci * v[:, PXX:T_QX], faciq * v[:, T_QX:]], dim=1)
            else:
                v = torch.cat([v[:, :PXX], faci * v[:, PXX:]], dim=1)
        return v
