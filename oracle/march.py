"""CPU restatement of the WInDS free-wake time march for a rigid rotor.

TEST INFRASTRUCTURE ONLY (the checker of the device march); the engine never
imports it.  Arrays keep MATLAB's shapes and every step follows the reference
files literally:

  setup            core/initials.m:44-152 (two ages at t1, bound circulation
                   0.5*|U|*chord*cl, r0 from T0 = 2 pi R / (12 TSR |U|), vcore,
                   induced velocity, forward Euler)
  time loop        WInDS.m:127-183 (Kelvin update, vcore, shed + trail
                   induced velocity at every node, integrator, Kutta-Joukowski)
  Kelvin           WInDS.m:131-140 / pcc.m:43-49 / KuttaJoukowski.m:285-288
  vcore            core/vcore.m:30-51 (global time, not filament age)
  filamentmod      core/filamentmod.m:27-41 (sqrt of the already-square-rooted
                   lengths, strain between equal age indices of two times)
  fe / pcc / ab4   numerical/fe.m:30-31, pcc.m:31-75, ab4.m:34-42
  Kutta-Joukowski  core/KuttaJoukowski.m:42-49 (fixed point) and kj() :173-298
  perform          postproc/perform.m:42-62

Rigid-rotor kinematics (core/kinematics.m, core/velocity.m) are restated
analytically: blade b at azimuth 2 pi b/nb + Omega t, quarter chord r e_r,
trailing edge r e_r + 0.75 c e_c, bxn = e_c, bzn = e_r, byn = bzn x bxn, and
the blade-station inflow U - Omega r e_t (the reference differentiates the
positions numerically, ctdiff.m).  When the wake reaches max_ages the oldest
age is dropped; filamentmod then compares equal age indices 2..end-1 of both
times (the reference never truncates).  Biot-Savart sums come from the C
oracle (oracle/bs_oracle.c) in the MEX or BiotSavart.m semantics.
"""
from __future__ import annotations

import math

import numpy as np

from . import bs as O

ALPHA_RL = 1.25643   # core/constants.m:22
NU = 1e-5            # :23
DELTA = 100.0        # :24
A1 = 6.5e-5          # :25
RHO = 1.23           # :28


class OracleMarch:
    def __init__(self, rotor, max_ages: int, d=1e-4, co: float = math.inf,
                 semantics: str = "mex", integ: str = "pcc", relax: float = 0.25,
                 tol: float = 1e-8, maxiter: int = 30, threads: int | None = None):
        self.r = rotor
        self.max_ages = max_ages
        self.d, self.co, self.sem, self.integ = d, co, semantics, integ
        self.relax, self.tol, self.maxiter = relax, tol, maxiter
        self.threads = threads
        self.nb, self.nst, self.ns = rotor.nb, rotor.nst, rotor.ns
        self.dt, self.om = rotor.dt, rotor.omega
        self.U = rotor.uinf
        self.history = []          # (p, T, Q, P, kj iterations)
        self._setup()

    # ------------------------------------------------------------------ kinematics
    def _frames(self, t):
        psi = 2 * math.pi * np.arange(self.nb) / self.nb + self.om * t
        er = np.stack([np.zeros(self.nb), np.cos(psi), np.sin(psi)], axis=1)    # [nb,3]
        et = np.stack([np.zeros(self.nb), -np.sin(psi), np.cos(psi)], axis=1)
        ex = np.array([1.0, 0.0, 0.0])
        return er, et, ex

    def _ec(self, et, ex, theta_deg):
        th = np.radians(theta_deg + self.r.pitch_deg)                              # [n]
        return (-np.cos(th)[:, None, None] * et[None, :, :]
                + np.sin(th)[:, None, None] * ex[None, None, :])                   # [n,nb,3]

    def quarter_trail(self, t):
        """pos.quarter and pos.trail, [nst,3,nb] each (kinematics.m:49-50)."""
        er, et, ex = self._frames(t)
        r = self.r.r_trail
        q = r[:, None, None] * er[None, :, :]                                     # [nst,nb,3]
        tr = q + 0.75 * self.r.chord_trail[:, None, None] * self._ec(et, ex, self.r.twist_trail)
        return np.transpose(q, (0, 2, 1)), np.transpose(tr, (0, 2, 1))

    def lifting_line(self, t):
        """pos.bound [ns,3,nb], frames bxn/byn/bzn [ns,3,nb], vel.blade [ns,3,nb]."""
        er, et, ex = self._frames(t)
        rn = self.r.r_node
        bound = rn[:, None, None] * er[None, :, :]
        bx = self._ec(et, ex, self.r.twist_node)                                  # [ns,nb,3]
        bz = np.broadcast_to(er[None, :, :], bx.shape)
        by = np.cross(bz, bx)
        vb = self.U[None, None, :] - (self.om * rn)[:, None, None] * et[None, :, :]
        vblade = np.stack([np.sum(bx * vb, -1), np.sum(by * vb, -1), np.sum(bz * vb, -1)], -1)
        tp = lambda a: np.transpose(a, (0, 2, 1))
        return tp(bound), tp(bx), tp(by), tp(bz), tp(vblade), tp(vb)

    # ------------------------------------------------------------------ pieces
    def _bs(self, F1, F2, P, g, rc):
        if self.sem == "mex":
            return O.bs_mex(F1, F2, P, g, self.d, threads=self.threads)
        return O.bs_matlab(F1, F2, P, g, rc, self.d, co=self.co, threads=self.threads)

    def induced(self, dom, gs, gt, rcs, rct, P):
        """Shed + trail calls of WInDS.m:150-154 (or KJ :216-218 with P = bound)."""
        us = self._bs(dom[:-1], dom[1:], P, gs, rcs)
        ut = self._bs(dom[:, :, 1:], dom[:, :, :-1], P, gt, rct)
        return us + ut

    @staticmethod
    def kelvin(gbound, gtrail_prev, A):
        """WInDS.m:131-140: new trail (A-1 ages) and shed (A ages) circulation."""
        ns, nb = gbound.shape
        z = np.zeros((1, nb))
        t0 = np.diff(np.concatenate([z, gbound, z], axis=0), axis=0)               # [nst,nb]
        gt = np.empty((ns + 1, 1, A - 1, nb))
        gt[:, 0, 0, :] = t0
        if A - 1 > 1:
            gt[:, 0, 1:, :] = gtrail_prev[:, 0, :A - 2, :]
        gs = np.empty((ns, 1, A, nb))
        gs[:, 0, 0, :] = gbound
        C = np.cumsum(gt[:-1, 0, :, :], axis=0)                                    # [ns,A-1,nb]
        Cz = np.concatenate([C, np.zeros((ns, 1, nb))], axis=1)
        gs[:, 0, 1:, :] = np.diff(Cz, axis=1)
        return gs, gt

    def vcore(self, p, dom, gs, gt, len_prev):
        """vcore.m:30-51 with filamentmod.m:27-41; returns rc_eff and lengths."""
        T = (p - 1) * self.dt
        res = np.abs(gs / NU)
        ret = np.abs(gt / NU)
        rcs = self.r0 ** 2 + 4 * ALPHA_RL * NU * (1 + A1 * res) * T
        rct = self.r0 ** 2 + 4 * ALPHA_RL * NU * (1 + A1 * ret) * T
        ls = O.filament_length(dom[:-1], dom[1:])
        lt = O.filament_length(dom[:, :, 1:], dom[:, :, :-1])
        rce_s, rce_t = rcs.copy(), rct.copy()
        if p > 3:
            At = dom.shape[2]
            it = slice(1, At - 2)   # trail ages 2:end-1
            ish = slice(1, At - 1)  # shed ages 2:end-1
            tnew = np.sqrt(lt[:, :, it, :])
            told = np.sqrt(len_prev[1][:, :, it, :])
            snew = np.sqrt(ls[:, :, ish, :])
            sold = np.sqrt(len_prev[0][:, :, ish, :])
            st = (tnew - told) / told
            ss = (snew - sold) / sold
            rce_t[:, :, it, :] = rct[:, :, it, :] * (1.0 / (1 + st))
            rce_s[:, :, ish, :] = rcs[:, :, ish, :] * (1.0 / (1 + ss))
        return rce_s, rce_t, (ls, lt)

    def _advect(self, dom, vel, t_next, vel_next=None, hist=None):
        A_new = min(dom.shape[2] + 1, self.max_ages)
        q, tr = self.quarter_trail(t_next)
        out = np.empty((self.nst, 3, A_new, self.nb))
        out[:, :, 0, :] = q
        out[:, :, 1, :] = tr
        n = A_new - 2
        src = dom[:, :, 1:1 + n, :]
        if vel_next is not None:       # pcc corrector
            out[:, :, 2:, :] = src + self.dt / 2 * (vel_next[:, :, 2:, :] + vel[:, :, 1:1 + n, :])
        elif hist is not None:         # ab4
            v0, v1, v2, v3 = hist
            acc = 55 * vel[:, :, 1:1 + n, :]
            for coef, vk, k in ((-59, v1, 1), (37, v2, 2), (-9, v3, 3)):
                sh = np.zeros_like(acc)
                if n > k:
                    sh[:, :, k:, :] = vk[:, :, 1:1 + n - k, :]
                acc = acc + coef * sh
            out[:, :, 2:, :] = src + self.dt / 24 * acc
        else:                          # forward Euler
            out[:, :, 2:, :] = src + self.dt * vel[:, :, 1:1 + n, :]
        return np.asfortranarray(out)

    def interp(self, table, aoa):
        return np.interp(aoa, self.r.alpha, table, left=np.nan, right=np.nan)

    # ------------------------------------------------------------------ setup
    def _setup(self):
        r = self.r
        self.r0 = math.sqrt(4 * ALPHA_RL * NU * DELTA * (2 * math.pi * r.R / (12 * r.tsr * np.linalg.norm(self.U))))
        q, tr = self.quarter_trail(0.0)
        dom = np.empty((self.nst, 3, 2, self.nb))
        dom[:, :, 0, :], dom[:, :, 1, :] = q, tr
        self.dom = np.asfortranarray(dom)
        _, bx, by, bz, vblade, _ = self.lifting_line(0.0)
        aoa = np.degrees(np.arctan2(-vblade[:, 1, :], vblade[:, 0, :]))           # [ns,nb]
        cl0 = self.interp(r.cl, aoa)
        self.gbound = 0.5 * np.linalg.norm(self.U) * r.chord_node[:, None] * cl0   # initials.m:124
        self.gs, self.gt = self.kelvin(self.gbound, np.zeros((self.nst, 1, 0, self.nb)), 2)
        rce_s, rce_t, lens = self.vcore(1, self.dom, self.gs, self.gt, None)
        u = self.induced(self.dom, self.gs, self.gt, rce_s, rce_t, self.dom)
        vel = self.U[None, :, None, None] + u
        self.lens = lens
        self.vel_hist = [vel]
        self.dom = self._advect(self.dom, vel, self.dt)
        self.p = 2

    # ------------------------------------------------------------------ one step
    def step(self):
        p = self.p
        t = (p - 1) * self.dt
        A = self.dom.shape[2]
        self.gs, self.gt = self.kelvin(self.gbound, self.gt, A)
        rce_s, rce_t, lens = self.vcore(p, self.dom, self.gs, self.gt, self.lens)
        u = self.induced(self.dom, self.gs, self.gt, rce_s, rce_t, self.dom)
        vel = self.U[None, :, None, None] + u
        # integrate to p+1
        if self.integ == "pcc":
            pred = self._advect(self.dom, vel, t + self.dt)
            if p > 2:
                gs1, gt1 = self.kelvin(self.gbound, self.gt, pred.shape[2])
                rs1, rt1, _ = self.vcore(p + 1, pred, gs1, gt1, lens)
                u1 = self.induced(pred, gs1, gt1, rs1, rt1, pred)
                vel1 = self.U[None, :, None, None] + u1
                pred = self._advect(self.dom, vel, t + self.dt, vel_next=vel1)
            dom_next = pred
        elif self.integ == "ab4" and p >= 4:
            h = self.vel_hist
            dom_next = self._advect(self.dom, vel, t + self.dt, hist=(vel, h[-1], h[-2], h[-3]))
        else:
            dom_next = self._advect(self.dom, vel, t + self.dt)
        # Kutta-Joukowski fixed point at time p
        iters, cl, cd, aoa = self.kj(t, rce_s, rce_t)
        T, Q, Pw = self.perform(cl, cd, aoa, t)
        self.history.append((p, T, Q, Pw, iters))
        self.last = dict(cl=cl, cd=cd, aoa=aoa, vel=vel, rce_s=rce_s, rce_t=rce_t)
        self.lens = lens
        self.vel_hist = (self.vel_hist + [vel])[-3:]
        self.dom = dom_next
        self.p = p + 1

    def kj(self, t, rce_s, rce_t):
        bound, bx, by, bz, vblade, _ = self.lifting_line(t)
        P = np.asfortranarray(bound.reshape(self.ns, 3, 1, self.nb))
        Vinf = np.linalg.norm(vblade, axis=1)                                      # [ns,nb]
        chord = self.r.chord_node[:, None]
        dg_max, j = 1.0, 0
        A = self.dom.shape[2]
        while dg_max > self.tol and j < self.maxiter:
            ub = self.induced(self.dom, self.gs, self.gt, rce_s, rce_t, P)[:, :, 0, :]  # [ns,3,nb]
            rot = np.stack([np.sum(bx * ub, 1), np.sum(by * ub, 1), np.sum(bz * ub, 1)], 1)
            tot = vblade + rot
            aoa = np.arctan2(-tot[:, 1, :], tot[:, 0, :]) * (180 / np.pi)
            cl = self.interp(self.r.cl, aoa)
            cd = self.interp(self.r.cd, aoa)
            if np.any(np.isnan(cl)):
                raise FloatingPointError("Diverging soln!!!")
            gnew = 0.5 * Vinf * chord * cl
            dg = gnew - self.gs[:, 0, 0, :]
            self.gs[:, 0, 0, :] = self.gs[:, 0, 0, :] + self.relax * dg
            self.gbound = self.gs[:, 0, 0, :].copy()
            # Kelvin update of trail age 1 and shed ages 2:end (KuttaJoukowski.m:285-288)
            z = np.zeros((1, self.nb))
            self.gt[:, 0, 0, :] = np.diff(np.concatenate([z, self.gbound, z], 0), axis=0)
            C = np.cumsum(self.gt[:-1, 0, :, :], axis=0)
            Cz = np.concatenate([C, np.zeros((self.ns, 1, self.nb))], axis=1)
            self.gs[:, 0, 1:, :] = np.diff(Cz, axis=1)
            dg = dg / (np.abs(gnew) + 1)
            dg_max = np.max(np.abs(dg))
            j += 1
        return j, cl, cd, aoa

    def perform(self, cl, cd, aoa, t):
        """perform.m:42-62 with velX = (axial inflow, in-plane speed) of the
        rigid rotor, phi = aoa - (pitch + twist)."""
        _, et_, ex = self._frames(t)
        rn = self.r.r_node
        vb = self.U[None, None, :] - (self.om * rn)[:, None, None] * et_[None, :, :]   # [ns,nb,3]
        v1 = vb[..., 0]
        v2 = -np.sum(vb * et_[None], -1)
        tsr = v2 / v1
        phi = aoa - (self.r.pitch_deg + self.r.twist_node[:, None])
        tphi = np.tan(np.radians(phi))
        aap = tsr / tphi
        a = aap * (v1 - tphi * v2) / (aap * v1 + tphi * v2)
        ap = a / aap
        cph, sph = np.cos(np.radians(phi)), np.sin(np.radians(phi))
        dCZ = cl * cph + cd * sph
        dCX = cl * sph - cd * cph
        W2 = (v1 * (1 - a)) ** 2 + (v2 * (1 + ap)) ** 2
        c = self.r.chord_node[:, None]
        dr = self.r.dr_node[:, None]
        dT = 0.5 * RHO * c * dCZ * W2 * dr
        dQ = 0.5 * RHO * c * dCX * W2 * rn[:, None] * dr
        dP = self.om * dQ          # (pi/30) * RotSpeed[rpm] = Omega
        return float(dT.sum()), float(dQ.sum()), float(dP.sum())
