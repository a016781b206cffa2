"""Ground-effect vortex-panel plane on the device (include/winds_ground.h).

The reference's ground model (ground/Main_Driver.m) covers the plane z = 0
with Np x Np square vortex rings, finds their circulation so that the
normal velocity vanishes at every panel centre
(ground/VP_Influence_Coefficient_Matrix_v7.m), and adds the rings' induced
velocity at the wake and blade points (ground/VP_Induced_Velocity_LM.m).

    g = Ground(50, -500, 1500, -375, 375, co=0.05)   # Main_Driver.m:7-23
    gamma = g.solve(v_ext)          # v_ext [N, 3]: total velocity at centres
    v = g.induced(gamma, P)         # P [n, 3] -> v [n, 3]

Arrays cross this interface as torch tensors on the GPU (or numpy arrays,
copied in); rows are points as in the reference (N x 3), the C ABI takes
SoA [3][n].  Panels are numbered j = m + n Np (m along x, n along y).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _bind(L):
    if getattr(L, "_winds_ground_bound", False):
        return L
    P, D = C.c_void_p, C.c_double
    L.winds_ground_create.argtypes = [C.c_int32, D, D, D, D, D, C.POINTER(C.c_void_p)]
    L.winds_ground_create.restype = C.c_int
    L.winds_ground_destroy.argtypes = [P]
    L.winds_ground_destroy.restype = None
    L.winds_ground_panels.argtypes = [P]
    L.winds_ground_panels.restype = C.c_int64
    L.winds_ground_geometry_dev.argtypes = [P, P, P, P]
    L.winds_ground_geometry_dev.restype = C.c_int
    L.winds_ground_solve_dev.argtypes = [P, P, P, D, C.c_int32, P, C.POINTER(C.c_int32),
                                         C.POINTER(D)]
    L.winds_ground_solve_dev.restype = C.c_int
    L.winds_ground_induced_dev.argtypes = [P, P, P, C.c_int64, P, P]
    L.winds_ground_induced_dev.restype = C.c_int
    L._winds_ground_bound = True
    return L


def _soa(a, torch):
    """[n, 3] rows (numpy or torch) -> contiguous [3, n] float64 on the GPU."""
    t = torch.as_tensor(a, dtype=torch.float64)
    if t.ndim != 2 or t.shape[1] != 3:
        raise ValueError("points must be an [n, 3] array")
    return t.to("cuda").t().contiguous()


class Ground:
    """Np x Np vortex-ring panels over [x_min, x_max] x [y_min, y_max] on
    z = 0, with cut-off fraction co (absolute cut-off co (dx + dy) / 2)."""

    def __init__(self, np_side: int, x_min: float, x_max: float, y_min: float, y_max: float,
                 co: float = 0.05):
        import torch
        torch.cuda.init()
        self.torch = torch
        self.L = _bind(_lib.lib())
        self.np_side = int(np_side)
        self._g = C.c_void_p()
        _lib.check(self.L.winds_ground_create(self.np_side, float(x_min), float(x_max),
                                              float(y_min), float(y_max), float(co),
                                              C.byref(self._g)))
        self.n = int(self.L.winds_ground_panels(self._g))
        self.last_iterations = 0
        self.last_residual = 0.0

    def __del__(self):
        g = getattr(self, "_g", None)
        if g is not None and g.value:
            try:
                self.torch.cuda.synchronize()
                self.L.winds_ground_destroy(g)
                g.value = None
            except Exception:       # interpreter shutdown
                pass

    def _stream(self):
        return self.torch.cuda.current_stream().cuda_stream

    def geometry(self):
        """(centres [N, 3], influence matrix A [N, N]) as float64 tensors."""
        t = self.torch
        c = t.empty((3, self.n), dtype=t.float64, device="cuda")
        A = t.empty((self.n, self.n), dtype=t.float64, device="cuda")
        _lib.check(self.L.winds_ground_geometry_dev(self._g, c.data_ptr(), A.data_ptr(),
                                                    self._stream()))
        return c.t(), A

    def solve(self, v_ext, tol: float = 1e-13, max_iter: int = 0):
        """Circulation [N] with A gamma = -v_ext . n at the centres (the
        reference's inv(A) * V_n); v_ext is [N, 3]."""
        t = self.torch
        v = _soa(v_ext, t)
        if v.shape[1] != self.n:
            raise ValueError(f"v_ext must hold {self.n} rows, one per panel")
        gam = t.empty(self.n, dtype=t.float64, device="cuda")
        it = C.c_int32()
        res = C.c_double()
        _lib.check(self.L.winds_ground_solve_dev(self._g, v.data_ptr(), gam.data_ptr(),
                                                 float(tol), int(max_iter), self._stream(),
                                                 C.byref(it), C.byref(res)))
        self.last_iterations, self.last_residual = it.value, res.value
        return gam

    def induced(self, gamma, P):
        """Induced velocity [n, 3] of the panels with circulation gamma at
        points P [n, 3]."""
        t = self.torch
        p = _soa(P, t)
        g = t.as_tensor(gamma, dtype=t.float64).to("cuda").contiguous()
        if g.numel() != self.n:
            raise ValueError(f"gamma must hold {self.n} values, one per panel")
        n = p.shape[1]
        u = t.empty((3, n), dtype=t.float64, device="cuda")
        _lib.check(self.L.winds_ground_induced_dev(self._g, g.data_ptr(), p.data_ptr(), n,
                                                   u.data_ptr(), self._stream()))
        return u.t()


def grid_nodes(np_side: int, lo: float, hi: float) -> np.ndarray:
    """Node coordinates lo + k (hi - lo) / Np, k = 0..Np: the device's grid
    (Main_Driver.m:11-12 builds the same with the colon operator)."""
    d = (hi - lo) / np_side
    return lo + np.arange(np_side + 1) * d
