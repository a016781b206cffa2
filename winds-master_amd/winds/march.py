"""Device-resident free-wake time march (include/winds_march.h) from Python."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

INTEG = {"pcc": 0, "fe": 1, "ab4": 2, "ab2": 3}


class MarchCfg(C.Structure):
    _fields_ = [
        ("nb", C.c_int32), ("nst", C.c_int32), ("max_ages", C.c_int32), ("n_times", C.c_int32),
        ("r_node", C.c_void_p), ("chord_node", C.c_void_p), ("twist_node", C.c_void_p),
        ("dr_node", C.c_void_p), ("foil", C.c_void_p), ("cl_init", C.c_void_p),
        ("tip_radius", C.c_double), ("n_foils", C.c_int32), ("n_alpha", C.c_int32),
        ("foil_len", C.c_void_p), ("alpha", C.c_void_p), ("cl", C.c_void_p), ("cd", C.c_void_p),
        ("time", C.c_void_p), ("wind", C.c_void_p), ("tsr", C.c_void_p),
        ("rot_speed", C.c_void_p), ("azimuth", C.c_void_p), ("pitch", C.c_void_p),
        ("hub", C.c_void_p), ("quarter", C.c_void_p), ("trail", C.c_void_p),
        ("bound", C.c_void_p), ("bxn", C.c_void_p), ("byn", C.c_void_p), ("bzn", C.c_void_p),
        ("vblade", C.c_void_p), ("ll_extra", C.c_void_p),
        ("dt", C.c_double), ("relax", C.c_double), ("tol", C.c_double),
        ("maxiter", C.c_int32), ("integrator", C.c_int32),
        ("wake", _lib.BsOpts), ("ll", _lib.BsOpts),
    ]


_TABLE_FIELDS = ("r_node", "chord_node", "twist_node", "dr_node", "foil", "cl_init", "foil_len",
                 "alpha", "cl", "cd", "time", "wind", "tsr", "rot_speed", "azimuth", "pitch",
                 "hub", "quarter", "trail", "bound", "bxn", "byn", "bzn", "vblade", "ll_extra")


class MarchSurr(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("split", C.c_int32), ("target_ages", C.c_int32),
        ("check_every", C.c_int32), ("train_iters", C.c_int32), ("reserved", C.c_int32),
        ("tol", C.c_double), ("lr", C.c_double), ("in_norm", C.c_void_p),
        ("out_norm", C.c_void_p),
    ]


SURR_STATS = ("checks", "failed_checks", "network_steps", "direct_far_steps", "last_err_nodes",
              "last_err_ll", "max_err", "adam_steps", "last_err_nodes_of_far",
              "last_err_ll_of_far")


def _bind(L):
    if getattr(L, "_winds_march_bound", False):
        return L
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    L.winds_march_create.argtypes = [C.POINTER(MarchCfg), C.POINTER(C.c_void_p)]
    L.winds_march_destroy.argtypes = [P]
    L.winds_march_destroy.restype = None
    L.winds_march_step.argtypes = [P, C.c_int32]
    L.winds_march_sync.argtypes = [P]
    L.winds_march_info.argtypes = [P, C.POINTER(C.c_int64)]
    L.winds_march_get_state.argtypes = [P, D, D, D]
    L.winds_march_set_state.argtypes = [P, C.c_int32, C.c_int32, D, D, D]
    L.winds_march_get_perf.argtypes = [P, C.c_int32, C.c_int32, D]
    L.winds_march_get_kj.argtypes = [P, D, D, D]
    L.winds_march_uind_dev.argtypes = [P]
    L.winds_march_uind_dev.restype = C.c_void_p
    L.winds_march_set_surrogate.argtypes = [P, P, C.POINTER(MarchSurr)]
    L.winds_march_surr_stats.argtypes = [P, D]
    L.winds_march_set_graphs.argtypes = [P, C.c_int32]
    for f in ("create", "step", "sync", "info", "get_state", "set_state", "get_perf", "get_kj",
              "set_surrogate", "surr_stats", "set_graphs"):
        getattr(L, "winds_march_" + f).restype = C.c_int
    L._winds_march_bound = True
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class March:
    """Free-wake march on one GPU, driven by the kinematic tables of a case
    (winds.cases.tables: NREL 5-MW, a synthetic rotor, the elliptical wing,
    or any motion the caller states)."""

    def __init__(self, tables, max_ages: int, wake: _lib.BsOpts, ll: _lib.BsOpts | None = None,
                 integrator: str = "pcc", relax: float = 0.25, tol: float = 1e-8,
                 maxiter: int = 30):
        import torch  # the engine binds to torch's HIP runtime when torch is present
        torch.cuda.init()
        self.L = _bind(_lib.lib())
        self.tables = tables
        c = MarchCfg()
        keep = {}
        for name in _TABLE_FIELDS:
            a = getattr(tables, name)
            if a is None:
                setattr(c, name, None)
                continue
            dt = np.int32 if name in ("foil", "foil_len") else np.float64
            a = np.asfortranarray(a, dtype=dt)
            keep[name] = a
            setattr(c, name, a.ctypes.data)
        self._keep = keep
        c.nb, c.nst, c.max_ages, c.n_times = tables.nb, tables.nst, max_ages, tables.n_times
        c.tip_radius = tables.tip_radius
        c.n_foils, c.n_alpha = int(tables.foil_len.size), int(tables.alpha.size)
        c.dt, c.relax, c.tol, c.maxiter = tables.dt, relax, tol, maxiter
        c.integrator = INTEG[integrator]
        c.wake = wake
        c.ll = ll if ll is not None else _lib.opts(
            core=wake.core, vat_n=wake.vat_n, semantics=wake.semantics,
            precision=_lib.PREC_FP64, d=wake.d, co=wake.co)
        self.cfg = c
        self._m = C.c_void_p()
        _lib.check(self.L.winds_march_create(C.byref(c), C.byref(self._m)))

    def __del__(self):
        m = getattr(self, "_m", None)
        if m is not None and m.value:
            try:
                self.L.winds_march_destroy(m)
                m.value = None
            except Exception:       # interpreter shutdown
                pass

    def step(self, n: int = 1):
        _lib.check(self.L.winds_march_step(self._m, int(n)))

    def sync(self):
        _lib.check(self.L.winds_march_sync(self._m))

    def set_graphs(self, on: bool = True):
        """Replay recorded HIP graphs of the steps once the march is steady
        (wake at max_ages, no surrogate); the results do not change."""
        _lib.check(self.L.winds_march_set_graphs(self._m, 1 if on else 0))

    def info(self) -> dict:
        a = (C.c_int64 * 6)()
        _lib.check(self.L.winds_march_info(self._m, a))
        return {"p": a[0], "ages": a[1], "nodes": a[2], "segments": a[3], "ages_solved": a[4],
                "steps": a[5]}

    def state(self):
        """(nodes [nst,3,A,nb] of the next time index, shed [ns,1,A',nb] and
        trailing [nst,1,A'-1,nb] circulation of the last solved time index)."""
        r = self.tables
        inf = self.info()
        A = inf["ages"]
        nodes = np.zeros((r.nst, 3, A, r.nb), order="F")
        # circulation arrays are sized by the last solved time index
        gs = np.zeros(r.ns * self.cfg.max_ages * r.nb)
        gt = np.zeros(r.nst * self.cfg.max_ages * r.nb)
        _lib.check(self.L.winds_march_get_state(self._m, _dp(nodes), _dp(gs), _dp(gt)))
        return nodes, gs, gt

    def state_circulation(self, ages_solved: int):
        r = self.tables
        _, gs, gt = self.state()
        A = ages_solved
        gs = gs[: r.ns * A * r.nb].reshape((r.ns, 1, A, r.nb), order="F")
        gt = gt[: r.nst * (A - 1) * r.nb].reshape((r.nst, 1, A - 1, r.nb), order="F")
        return gs, gt

    def set_state(self, p: int, nodes, gshed, gtrail):
        nodes = np.asfortranarray(nodes, dtype=np.float64)
        gshed = np.asfortranarray(gshed, dtype=np.float64)
        gtrail = np.asfortranarray(gtrail, dtype=np.float64)
        _lib.check(self.L.winds_march_set_state(self._m, int(p), int(nodes.shape[2]), _dp(nodes),
                                                _dp(gshed), _dp(gtrail)))

    def perf(self, first: int = 0, n: int | None = None) -> np.ndarray:
        """Rows (p, T, Q, P, KJ iterations) of the stepped time indices."""
        done = self.info()["steps"]
        if n is None:
            n = done - first
        out = np.zeros((n, 5))
        _lib.check(self.L.winds_march_get_perf(self._m, int(first), int(n), _dp(out)))
        return out

    def kj(self):
        r = self.tables
        cl, cd, aoa = (np.zeros(r.ns * r.nb) for _ in range(3))
        _lib.check(self.L.winds_march_get_kj(self._m, _dp(cl), _dp(cd), _dp(aoa)))
        f = lambda a: a.reshape((r.ns, r.nb), order="F")
        return f(cl), f(cd), f(aoa)

    def set_surrogate(self, cfg=None, net=None, in_norm=None, out_norm=None):
        """Turn the far-wake surrogate on (cfg: winds.surrogate.MarchSurrogate)
        or off (cfg None).  The network must outlive its use by this march."""
        if cfg is None:
            _lib.check(self.L.winds_march_set_surrogate(self._m, None, None))
            self._surr_keep = None
            return
        s = MarchSurr()
        s.mode, s.split, s.target_ages = cfg.mode, cfg.split, cfg.target_ages
        s.check_every, s.train_iters = cfg.check_every, cfg.train_iters
        s.tol, s.lr = cfg.tol, cfg.lr
        keep = []
        if in_norm is not None:
            a = np.ascontiguousarray(in_norm, dtype=np.float32)
            b = np.ascontiguousarray(out_norm, dtype=np.float32)
            keep = [a, b]
            s.in_norm, s.out_norm = a.ctypes.data, b.ctypes.data
        self._surr_keep = (net, keep)
        _lib.check(self.L.winds_march_set_surrogate(self._m, net.handle if net is not None else None,
                                                    C.byref(s)))

    def surr_stats(self) -> dict:
        a = np.zeros(10)
        _lib.check(self.L.winds_march_surr_stats(self._m, _dp(a)))
        return dict(zip(SURR_STATS, a.tolist()))
