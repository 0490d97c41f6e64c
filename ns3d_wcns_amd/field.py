"""Device-resident ``Field3D`` of one periodic box and the reference's hot-path
entry points on it.

``DeviceField`` mirrors ``Field3D`` (field_structures.h:196-509) together with
the free functions of ns3d_func.h that act on it; each method forwards to one
C-ABI entry point of libns3d_hip (include/ns3d_hip.h).  Host arrays use the
reference's layout: shape (sz, sy, sx), x fastest, three ghost layers.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import CONSERVED, FIELDS, PRIMITIVE, RHS, check, load_library
from .params import GridDesc, SolverParams

NG = 3


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class DeviceField:
    """A context of libns3d_hip: the five conserved fields, their recorded
    copy, the primitives and the RHS on the device."""

    def __init__(self, grid: GridDesc, params: SolverParams, precision: int = _lib.PRECISION_FP64):
        self._lib = load_library()
        self.grid = grid
        self.params = params
        self.precision = precision
        self.nx, self.ny, self.nz = grid.global_nx, grid.global_ny, grid.global_nz
        self.sx, self.sy, self.sz = self.nx + 2 * NG, self.ny + 2 * NG, self.nz + 2 * NG
        self.shape = (self.sz, self.sy, self.sx)
        h = ctypes.c_void_p()
        cg, cp = grid.to_c(NG), params.to_c()
        check(self._lib.ns3d_create(ctypes.byref(h), ctypes.byref(cg), ctypes.byref(cp), precision))
        self._h = h

    # ---------------------------------------------------------------- lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.ns3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr: int):
        check(self._lib.ns3d_set_stream(self._h, ctypes.c_void_p(stream_ptr)))

    def set_params(self, params: SolverParams):
        cp = params.to_c()
        check(self._lib.ns3d_set_params(self._h, ctypes.byref(cp)))
        self.params = params

    def synchronize(self):
        check(self._lib.ns3d_synchronize(self._h))

    # ---------------------------------------------------------------- data
    def upload(self, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"{name}: expected shape {self.shape}, got {a.shape}")
        check(self._lib.ns3d_upload(self._h, FIELDS[name], _dptr(a)))

    def download(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        check(self._lib.ns3d_download(self._h, FIELDS[name], _dptr(out)))
        return out

    def device_ptr(self, name: str) -> int:
        p = ctypes.c_void_p()
        check(self._lib.ns3d_field_device_ptr(self._h, FIELDS[name], ctypes.byref(p)))
        return p.value

    def conserved(self) -> dict:
        return {n: self.download(n) for n in CONSERVED}

    def primitives(self) -> dict:
        return {n: self.download(n) for n in PRIMITIVE}

    def rhs(self) -> dict:
        return {n: self.download(n) for n in RHS}

    def interior(self, a: np.ndarray) -> np.ndarray:
        return a[NG:NG + self.nz, NG:NG + self.ny, NG:NG + self.nx]

    def download_face_flux(self, direction: int, comp: int) -> np.ndarray:
        n = self._lib.ns3d_face_count(self._h, direction)
        out = np.empty(n, dtype=np.float64)
        check(self._lib.ns3d_download_face_flux(self._h, direction, comp, _dptr(out)))
        return out

    def download_gradient(self, q: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        check(self._lib.ns3d_download_gradient(self._h, q, _dptr(out)))
        return out

    def download_viscous_flux(self, direction: int, comp: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        check(self._lib.ns3d_download_viscous_flux(self._h, direction, comp, _dptr(out)))
        return out

    # ---------------------------------------------------------------- Field3D members
    def conservedToPrimitive(self):
        check(self._lib.ns3d_conserved_to_primitive(self._h))

    def primitiveToConserved(self):
        check(self._lib.ns3d_primitive_to_conserved(self._h))

    def recordConservedTo0(self):
        check(self._lib.ns3d_record_conserved_to0(self._h))

    # ---------------------------------------------------------------- ns3d_func.h
    def apply_boundary(self):
        check(self._lib.ns3d_apply_boundary(self._h))

    def apply_boundary_halfnode_flux(self):
        check(self._lib.ns3d_apply_boundary_halfnode_flux(self._h))

    def exchange_halos_viscous_flux(self):
        check(self._lib.ns3d_exchange_halos_viscous_flux(self._h))

    def compute_invis_flux(self):
        check(self._lib.ns3d_compute_invis_flux(self._h))

    def compute_invis_dflux(self):
        check(self._lib.ns3d_compute_invis_dflux(self._h))

    def compute_gradients(self):
        check(self._lib.ns3d_compute_gradients(self._h))

    def compute_viscous_flux(self):
        check(self._lib.ns3d_compute_viscous_flux(self._h))

    def compute_vis_flux(self):
        check(self._lib.ns3d_compute_vis_flux(self._h))

    def compute_rhs(self):
        check(self._lib.ns3d_compute_rhs(self._h))

    def runge_kutta_3(self, dt: float):
        check(self._lib.ns3d_runge_kutta_3(self._h, float(dt)))

    def runge_kutta_3_dev(self, d_dt: int):
        check(self._lib.ns3d_runge_kutta_3_dev(self._h, ctypes.c_void_p(d_dt)))

    def compute_timestep(self) -> float:
        dt = ctypes.c_double()
        check(self._lib.ns3d_compute_timestep(self._h, ctypes.byref(dt)))
        return dt.value

    def compute_timestep_dev(self, d_dt: int):
        check(self._lib.ns3d_compute_timestep_dev(self._h, ctypes.c_void_p(d_dt)))

    def time_advance(self, max_steps: int, dt_fixed: float = -1.0, total_time: float = 1e30,
                     output_timefreq: float = 1e30) -> np.ndarray:
        dts = np.zeros(max(max_steps, 1), dtype=np.float64)
        taken = ctypes.c_int()
        check(self._lib.ns3d_time_advance(self._h, max_steps, dt_fixed, total_time, output_timefreq, _dptr(dts),
                                          ctypes.byref(taken)))
        return dts[:taken.value]

    def advance_steps(self, nsteps: int, d_dt_hist: int = 0):
        check(self._lib.ns3d_advance_steps(self._h, nsteps, ctypes.c_void_p(d_dt_hist or None)))

    def set_kernel_timing(self, enable: bool):
        check(self._lib.ns3d_set_kernel_timing(self._h, int(bool(enable))))

    def kernel_times(self) -> dict:
        """Summed HIP-event time (ms) and launch count per kernel class since
        the last call (see ns3d_set_kernel_timing)."""
        n = len(_lib.KERNEL_CLASSES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        check(self._lib.ns3d_kernel_times(self._h, ms, cnt))
        return {k: (ms[q], cnt[q]) for q, k in enumerate(_lib.KERNEL_CLASSES)}

    def total_kinetic_energy(self) -> float:
        out = ctypes.c_double()
        check(self._lib.ns3d_total_kinetic_energy(self._h, ctypes.byref(out)))
        return out.value

    # ------------------------------------------------ device buffers, derivatives
    # Device arrays are addresses (ints) of sx*sy*sz doubles in the ghosted
    # layout, five of them (rho, rho u, rho v, rho w, E) for a state.
    def copy_field_to(self, name: str, dst: int):
        check(self._lib.ns3d_copy_field_to_device(self._h, FIELDS[name], ctypes.c_void_p(dst)))

    def copy_field_from(self, name: str, src: int):
        check(self._lib.ns3d_copy_field_from_device(self._h, FIELDS[name], ctypes.c_void_p(src)))

    def set_state(self, ptrs):
        """Conserved state from five device arrays; primitives and ghost
        layers are made consistent with it."""
        check(self._lib.ns3d_set_state(self._h, _lib.ptr5(ptrs)))

    def get_state(self, ptrs):
        check(self._lib.ns3d_get_state(self._h, _lib.ptr5(ptrs)))

    def tangent_rhs(self, dU, out):
        """out = (dRHS/dU) dU at the current state."""
        check(self._lib.ns3d_tangent_rhs(self._h, _lib.ptr5(dU), _lib.ptr5(out)))

    def adjoint_rhs(self, lam, out):
        """out = (dRHS/dU)^T lam at the current state."""
        check(self._lib.ns3d_adjoint_rhs(self._h, _lib.ptr5(lam), _lib.ptr5(out)))

    def tangent_step(self, dt: float, dU):
        """One RK3 step of the state with fixed dt; dU becomes its tangent."""
        check(self._lib.ns3d_tangent_step(self._h, dt, _lib.ptr5(dU)))

    def adjoint_step(self, dt: float, adj):
        """adj <- (dU^{n+1}/dU^n)^T adj for the step from the current state
        U^n with fixed dt; the state stays U^n."""
        check(self._lib.ns3d_adjoint_step(self._h, dt, _lib.ptr5(adj)))

    def step_stages(self, dt: float, U1, U2):
        """step(dt), storing the RK3 stage states U^(1), U^(2) (five device
        pointers each, ghosted layout)."""
        check(self._lib.ns3d_step_stages(self._h, dt, _lib.ptr5(U1), _lib.ptr5(U2)))

    def adjoint_step_stages(self, dt: float, adj, U1, U2):
        """adjoint_step with the stage states stored by step_stages (no
        recomputation of stages 1 and 2)."""
        check(self._lib.ns3d_adjoint_step_stages(self._h, dt, _lib.ptr5(adj), _lib.ptr5(U1), _lib.ptr5(U2)))

    def step(self, dt: float):
        """recordConservedTo0 + runge_kutta_3 with a fixed dt."""
        self.recordConservedTo0()
        self.runge_kutta_3(dt)

    # ---------------------------------------------------------------- observables
    def energy_spectrum(self, nshell: int | None = None) -> np.ndarray:
        """E(k), k = 0..nshell-1 (default 2 nx + 1), computed on the device
        from the context's velocity fields (EnergySpectrum.cpp:63-156)."""
        n = nshell or 2 * self.nx + 1
        out = np.zeros(n)
        check(self._lib.ns3d_energy_spectrum(self._h, _dptr(out), n))
        return out

    def turbulence_statistics(self) -> dict:
        """Kinetic energy, TKE, u_rms, nu and the spectral dissipation,
        computed on the device (ana_isotropic.cpp:238-305)."""
        t = _lib.CTurbulence()
        check(self._lib.ns3d_turbulence_statistics(self._h, ctypes.byref(t)))
        return {n: getattr(t, n) for n, _ in _lib.CTurbulence._fields_}

    def snapshot(self) -> dict:
        """Interior copies of the conserved fields and the primitives."""
        return {name: self.interior(self.download(name)).copy() for name in CONSERVED + PRIMITIVE}

    # ---------------------------------------------------------------- set-up
    def initialize_primitive(self, rho, u, v, w, p):
        """Start from interior primitives (nz, ny, nx): fill the ghost layers
        (apply_boundary) and form the conserved fields (primitiveToConserved),
        as main.cpp:98-99 does after the initial condition."""
        for name, a in (("rho", rho), ("u", u), ("v", v), ("w", w), ("p", p)):
            # cells outside the interior keep Field3D::allocate's defaults (rho = 1, else 0)
            g = np.ones(self.shape) if name == "rho" else np.zeros(self.shape)
            self.interior(g)[...] = a
            self.upload(name, g)
        self.apply_boundary()
        self.primitiveToConserved()
        self.synchronize()
