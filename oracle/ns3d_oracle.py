"""ctypes wrapper of oracle/libns3d_oracle.so -- TEST INFRASTRUCTURE ONLY.

``OracleField`` holds the reference's Field3D arrays on the host and runs the
reference's hot-path functions (restated in ns3d_oracle.cpp) on them.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libns3d_oracle.so")
NG = 3

_lib = None


def build():
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    from ns3d_wcns_amd.params import CParams

    lib = ctypes.CDLL(LIB_PATH)
    VP, DP = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    PP = ctypes.POINTER(CParams)
    lib.orc_create.restype = VP
    lib.orc_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [PP]
    lib.orc_destroy.argtypes = [VP]
    lib.orc_set_params.argtypes = [VP, PP]
    lib.orc_field.restype = DP
    lib.orc_field.argtypes = [VP, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_set_threads.restype = ctypes.c_int
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    for n in ("conserved_to_primitive", "primitive_to_conserved", "apply_boundary", "record_conserved_to0",
              "compute_invis_flux", "apply_boundary_halfnode_flux", "compute_invis_dflux", "compute_gradients",
              "compute_viscous_flux", "exchange_halos_viscous_flux", "compute_vis_flux", "compute_rhs"):
        getattr(lib, "orc_" + n).argtypes = [VP]
    lib.orc_runge_kutta_3.argtypes = [VP, ctypes.c_double]
    lib.orc_compute_timestep.restype = ctypes.c_double
    lib.orc_compute_timestep.argtypes = [VP]
    lib.orc_total_kinetic.restype = ctypes.c_double
    lib.orc_total_kinetic.argtypes = [VP]
    lib.orc_time_advance.restype = ctypes.c_int
    lib.orc_time_advance.argtypes = [VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, DP]
    lib.orc_interpolate_batch.argtypes = [PP, DP, ctypes.c_int, ctypes.c_double, DP]
    lib.orc_riemann_batch.argtypes = [PP, DP, DP, ctypes.c_int, ctypes.c_int, DP]
    lib.orc_face_flux_batch.argtypes = [PP, DP, DP, ctypes.c_int, ctypes.c_int, DP]
    lib.orc_eigen.argtypes = [DP, DP, ctypes.c_int, ctypes.c_double, DP, DP, DP]
    lib.orc_sutherland.restype = ctypes.c_double
    lib.orc_sutherland.argtypes = [PP, ctypes.c_double]
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n: int) -> int:
    return load().orc_set_threads(n)


class OracleField:
    """Reference-layout Field3D on the host with the restated hot path."""

    def __init__(self, grid, params):
        self.lib = load()
        self.grid, self.params = grid, params
        self.nx, self.ny, self.nz = grid.global_nx, grid.global_ny, grid.global_nz
        self.sx, self.sy, self.sz = self.nx + 2 * NG, self.ny + 2 * NG, self.nz + 2 * NG
        self.shape = (self.sz, self.sy, self.sx)
        cp = params.to_c()
        self.h = self.lib.orc_create(self.nx, self.ny, self.nz, grid.dx, grid.dy, grid.dz, ctypes.byref(cp))

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        cp = params.to_c()
        self.lib.orc_set_params(self.h, ctypes.byref(cp))
        self.params = params

    def view(self, name: str) -> np.ndarray:
        """Zero-copy numpy view of a Field3D array (cell arrays shaped
        (sz, sy, sx); face arrays flat)."""
        n = ctypes.c_size_t()
        p = self.lib.orc_field(self.h, name.encode(), ctypes.byref(n))
        if not p:
            raise KeyError(name)
        a = np.ctypeslib.as_array(p, shape=(n.value,))
        if n.value == self.sx * self.sy * self.sz:
            return a.reshape(self.shape)
        return a

    def interior(self, a):
        return a[NG:NG + self.nz, NG:NG + self.ny, NG:NG + self.nx]

    def initialize_primitive(self, rho, u, v, w, p):
        for name, a in (("rho", rho), ("u", u), ("v", v), ("w", w), ("p", p)):
            self.interior(self.view(name))[...] = a
        self.lib.orc_apply_boundary(self.h)
        self.lib.orc_primitive_to_conserved(self.h)

    def __getattr__(self, name):
        fn = getattr(load(), "orc_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda *a: fn(self.h, *a)

    def time_advance(self, max_steps, dt_fixed=-1.0, total_time=1e30, output_timefreq=1e30):
        dts = np.zeros(max(max_steps, 1))
        n = self.lib.orc_time_advance(self.h, max_steps, dt_fixed, total_time, output_timefreq, _dp(dts))
        return dts[:n]


def interpolate_batch(params, stencils: np.ndarray, flag: float) -> np.ndarray:
    st = np.ascontiguousarray(stencils, dtype=np.float64)
    out = np.empty(st.shape[0])
    cp = params.to_c()
    load().orc_interpolate_batch(ctypes.byref(cp), _dp(st), st.shape[0], flag, _dp(out))
    return out


def riemann_batch(params, UL: np.ndarray, UR: np.ndarray, dim: int) -> np.ndarray:
    a = np.ascontiguousarray(UL, dtype=np.float64)
    b = np.ascontiguousarray(UR, dtype=np.float64)
    out = np.empty_like(a)
    cp = params.to_c()
    load().orc_riemann_batch(ctypes.byref(cp), _dp(a), _dp(b), a.shape[0], dim, _dp(out))
    return out


def face_flux_batch(params, Ut: np.ndarray, ut: np.ndarray, dim: int) -> np.ndarray:
    """Ut, ut: (n, 5, 6) conserved / primitive stencils of cells i-2..i+3."""
    a = np.ascontiguousarray(Ut, dtype=np.float64)
    b = np.ascontiguousarray(ut, dtype=np.float64)
    out = np.empty((a.shape[0], 5))
    cp = params.to_c()
    load().orc_face_flux_batch(ctypes.byref(cp), _dp(a), _dp(b), a.shape[0], dim, _dp(out))
    return out


def eigen(prim_l, prim_r, dim, gamma):
    ul = np.ascontiguousarray(prim_l, dtype=np.float64)
    ur = np.ascontiguousarray(prim_r, dtype=np.float64)
    L = np.empty((5, 5)); R = np.empty((5, 5)); lam = np.empty(5)
    load().orc_eigen(_dp(ul), _dp(ur), dim, gamma, _dp(L), _dp(R), _dp(lam))
    return L, R, lam


def sutherland(params, T: float) -> float:
    cp = params.to_c()
    return load().orc_sutherland(ctypes.byref(cp), T)
