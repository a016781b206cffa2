"""CPU restatement of the WInDS free-wake time march.

TEST INFRASTRUCTURE ONLY (the checker of the device march); the engine never
imports it.  Arrays keep MATLAB's shapes and every step follows the reference
files literally:

  setup            core/initials.m:44-152 (two ages at time index 1, bound
                   circulation 0.5*|U(1)|*chord*cl_init, r0(t) from
                   T0 = 2 pi R / (12 TSR(t) |U(t)|), vcore, induced velocity,
                   forward Euler)
  time loop        WInDS.m:127-183 (Kelvin update, vcore, shed + trail
                   induced velocity at every node, integrator, Kutta-Joukowski)
  Kelvin           WInDS.m:131-140 / pcc.m:43-49 / KuttaJoukowski.m:285-288
  vcore            core/vcore.m:30-51 (global time fastout.Time, not age)
  filamentmod      core/filamentmod.m:27-41 (sqrt of the already-square-rooted
                   lengths, strain between equal age indices of two times)
  fe / pcc / ab2 / ab4  numerical/fe.m:30-31, pcc.m:31-75, ab2.m:34-40,
                   ab4.m:34-42 (ab2.m's forward-Euler start passes dt where
                   fe.m takes user, an error in MATLAB; here the start is
                   fe.m's, as ab4.m's is)
  Kutta-Joukowski  core/KuttaJoukowski.m:42-49 (fixed point) and kj() :173-298,
                   each station interpolating its own airfoil's table
                   (blade.NFoil, :262-269), other sources at the lifting line
                   added to the induced velocity (:232-238)
  perform          postproc/perform.m:16-62 (velX = -DCMRot(vel.blade,
                   [-90 AeroTwst BldPitch], 'zzz'))

The rotor's motion comes from the same kinematic tables the device march
takes (winds.cases.tables: pos.quarter/trail/bound, pos.nodes.bxn/byn/bzn,
vel.blade, fastout, wind.infty per time index).  When the wake reaches
max_ages the oldest age is dropped; filamentmod then compares equal age
indices 2..end-1 of both times (the reference never truncates).  Biot-Savart
sums come from the C oracle (oracle/bs_oracle.c) in the MEX or BiotSavart.m
semantics.
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
    def __init__(self, tables, max_ages: int, d=1e-4, co: float = math.inf,
                 semantics: str = "mex", integ: str = "pcc", relax: float = 0.25,
                 tol: float = 1e-8, maxiter: int = 30, threads: int | None = None,
                 setup: bool = True):
        self.tb = tables
        self.max_ages = max_ages
        self.d, self.co, self.sem, self.integ = d, co, semantics, integ
        self.relax, self.tol, self.maxiter = relax, tol, maxiter
        self.threads = threads
        self.nb, self.nst, self.ns = tables.nb, tables.nst, tables.ns
        self.dt = tables.dt
        self.history = []          # (p, T, Q, P, kj iterations)
        # offsets of each airfoil table in the concatenated arrays
        self.foil_off = np.concatenate([[0], np.cumsum(tables.foil_len)[:-1]]).astype(int)
        if setup:
            self._setup()

    # ------------------------------------------------------------------ tables
    def wind(self, p):
        return np.asarray(self.tb.wind)[p - 1]

    def r0(self, p):
        """initials.m:134-135 at time index p."""
        T0 = 2 * math.pi * self.tb.tip_radius / (12 * self.tb.tsr[p - 1] * np.linalg.norm(self.wind(p)))
        return math.sqrt(4 * ALPHA_RL * NU * DELTA * T0)

    def quarter_trail(self, p):
        """pos.quarter and pos.trail at time index p, [nst,3,nb] each."""
        return self.tb.quarter[:, :, p - 1, :], self.tb.trail[:, :, p - 1, :]

    def lifting_line(self, p):
        """pos.bound, bxn, byn, bzn, vel.blade at time index p, [ns,3,nb] each."""
        t = self.tb
        return (t.bound[:, :, p - 1, :], t.bxn[:, :, p - 1, :], t.byn[:, :, p - 1, :],
                t.bzn[:, :, p - 1, :], t.vblade[:, :, p - 1, :])

    def polar(self, table, aoa):
        """cl or cd of every station from its airfoil (KuttaJoukowski.m:262-269);
        aoa [ns, nb]."""
        t = self.tb
        out = np.empty_like(aoa)
        for m in range(self.ns):
            k = int(t.foil[m]) - 1
            o, n = self.foil_off[k], int(t.foil_len[k])
            out[m] = np.interp(aoa[m], t.alpha[o:o + n], table[o:o + n], left=np.nan, right=np.nan)
        return out

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
        T = float(self.tb.time[p - 1])
        r0 = self.r0(p)
        res = np.abs(gs / NU)
        ret = np.abs(gt / NU)
        rcs = r0 ** 2 + 4 * ALPHA_RL * NU * (1 + A1 * res) * T
        rct = r0 ** 2 + 4 * ALPHA_RL * NU * (1 + A1 * ret) * T
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

    def _advect(self, dom, vel, p_next, vel_next=None, hist=None):
        """Nodes of time index p_next: ages 0 and 1 from the tables, older ages
        moved by fe.m, pcc.m's corrector (vel_next), ab4.m (hist of four
        velocity fields) or ab2.m (hist of two)."""
        A_new = min(dom.shape[2] + 1, self.max_ages)
        q, tr = self.quarter_trail(p_next)
        out = np.empty((self.nst, 3, A_new, self.nb))
        out[:, :, 0, :] = q
        out[:, :, 1, :] = tr
        n = A_new - 2
        src = dom[:, :, 1:1 + n, :]
        if vel_next is not None:       # pcc corrector
            out[:, :, 2:, :] = src + self.dt / 2 * (vel_next[:, :, 2:, :] + vel[:, :, 1:1 + n, :])
        elif hist is not None and len(hist) == 2:   # ab2
            v1 = hist[1]
            sh = np.zeros_like(vel[:, :, 1:1 + n, :])
            if n > 1:
                sh[:, :, 1:, :] = v1[:, :, 1:n, :]
            out[:, :, 2:, :] = src + self.dt / 2 * (3 * vel[:, :, 1:1 + n, :] - sh)
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

    # ------------------------------------------------------------------ setup
    def _setup(self):
        q, tr = self.quarter_trail(1)
        dom = np.empty((self.nst, 3, 2, self.nb))
        dom[:, :, 0, :], dom[:, :, 1, :] = q, tr
        self.dom = np.asfortranarray(dom)
        U1 = np.linalg.norm(self.wind(1))
        cl0 = np.asarray(self.tb.cl_init)
        self.gbound = np.repeat((0.5 * U1 * np.asarray(self.tb.chord_node) * cl0)[:, None],
                                self.nb, axis=1)                                 # initials.m:124
        self.gs, self.gt = self.kelvin(self.gbound, np.zeros((self.nst, 1, 0, self.nb)), 2)
        rce_s, rce_t, lens = self.vcore(1, self.dom, self.gs, self.gt, None)
        u = self.induced(self.dom, self.gs, self.gt, rce_s, rce_t, self.dom)
        vel = self.wind(1)[None, :, None, None] + u
        self.lens = lens
        self.vel_hist = [vel]
        self.dom = self._advect(self.dom, vel, 2)
        self.p = 2

    # ------------------------------------------------------------------ one step
    def step(self):
        p = self.p
        A = self.dom.shape[2]
        self.gs, self.gt = self.kelvin(self.gbound, self.gt, A)
        rce_s, rce_t, lens = self.vcore(p, self.dom, self.gs, self.gt, self.lens)
        u = self.induced(self.dom, self.gs, self.gt, rce_s, rce_t, self.dom)
        vel = self.wind(p)[None, :, None, None] + u
        # integrate to p+1
        if self.integ == "pcc":
            pred = self._advect(self.dom, vel, p + 1)
            if p > 2:
                gs1, gt1 = self.kelvin(self.gbound, self.gt, pred.shape[2])
                rs1, rt1, _ = self.vcore(p + 1, pred, gs1, gt1, lens)
                u1 = self.induced(pred, gs1, gt1, rs1, rt1, pred)
                vel1 = self.wind(p + 1)[None, :, None, None] + u1
                pred = self._advect(self.dom, vel, p + 1, vel_next=vel1)
            dom_next = pred
        elif self.integ == "ab4" and p >= 4:
            h = self.vel_hist
            dom_next = self._advect(self.dom, vel, p + 1, hist=(vel, h[-1], h[-2], h[-3]))
        elif self.integ == "ab2" and p >= 2:
            dom_next = self._advect(self.dom, vel, p + 1, hist=(vel, self.vel_hist[-1]))
        else:
            dom_next = self._advect(self.dom, vel, p + 1)
        # Kutta-Joukowski fixed point at time p
        iters, cl, cd, aoa = self.kj(p, rce_s, rce_t)
        T, Q, Pw = self.perform(cl, cd, aoa, p)
        self.history.append((p, T, Q, Pw, iters))
        self.last = dict(cl=cl, cd=cd, aoa=aoa, vel=vel, rce_s=rce_s, rce_t=rce_t)
        self.lens = lens
        self.vel_hist = (self.vel_hist + [vel])[-3:]
        self.dom = dom_next
        self.p = p + 1

    def kj(self, p, rce_s, rce_t):
        bound, bx, by, bz, vblade = self.lifting_line(p)
        P = np.asfortranarray(bound.reshape(self.ns, 3, 1, self.nb))
        extra = None
        if getattr(self.tb, "ll_extra", None) is not None:
            extra = self.tb.ll_extra[:, :, p - 1, :]
        Vinf = np.linalg.norm(vblade, axis=1)                                      # [ns,nb]
        chord = np.asarray(self.tb.chord_node)[:, None]
        dg_max, j = 1.0, 0
        while dg_max > self.tol and j < self.maxiter:
            ub = self.induced(self.dom, self.gs, self.gt, rce_s, rce_t, P)[:, :, 0, :]  # [ns,3,nb]
            if extra is not None:
                ub = ub + extra
            rot = np.stack([np.sum(bx * ub, 1), np.sum(by * ub, 1), np.sum(bz * ub, 1)], 1)
            tot = vblade + rot
            aoa = np.arctan2(-tot[:, 1, :], tot[:, 0, :]) * (180 / np.pi)
            cl = self.polar(self.tb.cl, aoa)
            cd = self.polar(self.tb.cd, aoa)
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

    def velX(self, p):
        """perform.m:19-24: vel.blade rotated by -90 deg, twist and pitch about
        z (DCMRot 'zzz', which composes to one rotation) and negated; [ns,3,nb]."""
        _, _, _, _, vb = self.lifting_line(p)
        tw = np.asarray(self.tb.twist_node)[:, None]
        pitch = np.asarray(self.tb.pitch).reshape(self.tb.n_times, -1)[p - 1][None, :]
        th = np.radians(-90.0 + tw + pitch)                                          # [ns,nb]
        c, s = np.cos(th), np.sin(th)
        out = np.empty_like(vb)
        out[:, 0, :] = -(c * vb[:, 0, :] - s * vb[:, 1, :])
        out[:, 1, :] = -(s * vb[:, 0, :] + c * vb[:, 1, :])
        out[:, 2, :] = -vb[:, 2, :]
        return out

    def perform(self, cl, cd, aoa, p):
        """perform.m:16-62 for time index p."""
        vX = self.velX(p)
        v1, v2 = vX[:, 0, :], vX[:, 1, :]
        tsr = v2 / v1
        tw = np.asarray(self.tb.twist_node)[:, None]
        pitch = np.asarray(self.tb.pitch).reshape(self.tb.n_times, -1)[p - 1][None, :]
        phi = aoa - (pitch + tw)
        tphi = np.tan(np.radians(phi))
        aap = tsr / tphi
        a = aap * (v1 - tphi * v2) / (aap * v1 + tphi * v2)
        ap = a / aap
        cph, sph = np.cos(np.radians(phi)), np.sin(np.radians(phi))
        dCZ = cl * cph + cd * sph
        dCX = cl * sph - cd * cph
        W2 = (v1 * (1 - a)) ** 2 + (v2 * (1 + ap)) ** 2
        c = np.asarray(self.tb.chord_node)[:, None]
        dr = np.asarray(self.tb.dr_node)[:, None]
        rn = np.asarray(self.tb.r_node)[:, None]
        dT = 0.5 * RHO * c * dCZ * W2 * dr
        dQ = 0.5 * RHO * c * dCX * W2 * rn * dr
        dP = (math.pi / 30) * float(self.tb.rot_speed[p - 1]) * dQ
        return float(dT.sum()), float(dQ.sum()), float(dP.sum())
