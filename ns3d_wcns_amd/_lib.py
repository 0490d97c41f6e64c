"""ctypes binding of libns3d_hip.so (include/ns3d_hip.h).

The library is built in place by ``__graft_entry__.build()`` into
``ns3d_wcns_amd/lib/``.  Loading fails loudly when it is missing: there is no
CPU fallback for the device path.
"""
from __future__ import annotations

import ctypes
import os

from .params import CGrid, CParams

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libns3d_hip.so")

# enum ns3d_field
FIELDS = {
    "rho": 0, "rhou": 1, "rhov": 2, "rhow": 3, "E": 4,
    "rho0": 5, "rhou0": 6, "rhov0": 7, "rhow0": 8, "E0": 9,
    "u": 10, "v": 11, "w": 12, "p": 13, "T": 14,
    "rhs_rho": 15, "rhs_rhou": 16, "rhs_rhov": 17, "rhs_rhow": 18, "rhs_E": 19,
}
CONSERVED = ("rho", "rhou", "rhov", "rhow", "E")
PRIMITIVE = ("u", "v", "w", "p", "T")
RHS = ("rhs_rho", "rhs_rhou", "rhs_rhov", "rhs_rhow", "rhs_E")

PRECISION_FP64 = 0
PRECISION_MIXED = 1


KERNEL_CLASSES = ("viscous", "pencil_z", "pencil_y", "pencil_x", "halo", "cfl")


def ptr5(ptrs) -> "ctypes.Array":
    """Five device addresses (ints) as the C array ``double *[5]``."""
    ptrs = list(ptrs)
    if len(ptrs) != 5:
        raise ValueError("five field pointers expected")
    return _VP5(*[ctypes.c_void_p(int(p)) for p in ptrs])


class CFlops(ctypes.Structure):
    """``ns3d_flops`` of include/ns3d_hip.h."""

    _fields_ = [(n, ctypes.c_double) for n in ("face", "diff", "visc_flux", "visc_div", "update", "cfl", "per_stage",
                                                "per_step")]


class CTurbulence(ctypes.Structure):
    """``ns3d_turbulence`` of include/ns3d_hip.h."""

    _fields_ = [(n, ctypes.c_double) for n in ("kinetic_energy", "tke", "u_rms", "nu", "dissipation")]


class NS3DError(RuntimeError):
    """A nonzero status from the device library, with its message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"ns3d status {code}: {msg}")
        self.code = code


_lib = None

_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_VP5 = ctypes.c_void_p * 5
_SIGS = {
    "ns3d_create": [ctypes.POINTER(_VP), ctypes.POINTER(CGrid), ctypes.POINTER(CParams), ctypes.c_int],
    "ns3d_destroy": [_VP],
    "ns3d_set_stream": [_VP, _VP],
    "ns3d_set_params": [_VP, ctypes.POINTER(CParams)],
    "ns3d_synchronize": [_VP],
    "ns3d_upload": [_VP, ctypes.c_int, _DP],
    "ns3d_download": [_VP, ctypes.c_int, _DP],
    "ns3d_field_device_ptr": [_VP, ctypes.c_int, ctypes.POINTER(_VP)],
    "ns3d_conserved_to_primitive": [_VP],
    "ns3d_primitive_to_conserved": [_VP],
    "ns3d_record_conserved_to0": [_VP],
    "ns3d_apply_boundary": [_VP],
    "ns3d_apply_boundary_halfnode_flux": [_VP],
    "ns3d_exchange_halos_viscous_flux": [_VP],
    "ns3d_compute_invis_flux": [_VP],
    "ns3d_compute_invis_dflux": [_VP],
    "ns3d_download_face_flux": [_VP, ctypes.c_int, ctypes.c_int, _DP],
    "ns3d_compute_gradients": [_VP],
    "ns3d_compute_viscous_flux": [_VP],
    "ns3d_compute_vis_flux": [_VP],
    "ns3d_download_gradient": [_VP, ctypes.c_int, _DP],
    "ns3d_download_viscous_flux": [_VP, ctypes.c_int, ctypes.c_int, _DP],
    "ns3d_compute_rhs": [_VP],
    "ns3d_runge_kutta_3": [_VP, ctypes.c_double],
    "ns3d_runge_kutta_3_dev": [_VP, _VP],
    "ns3d_compute_timestep": [_VP, _DP],
    "ns3d_compute_timestep_dev": [_VP, _VP],
    "ns3d_compute_timestep_noalloc": [ctypes.POINTER(CGrid), ctypes.POINTER(CParams), _VP, _VP, _VP, _VP, _VP],
    "ns3d_time_advance": [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _DP,
                          ctypes.POINTER(ctypes.c_int)],
    "ns3d_advance_steps": [_VP, ctypes.c_int, _VP],
    "ns3d_total_kinetic_energy": [_VP, _DP],
    "ns3d_set_kernel_timing": [_VP, ctypes.c_int],
    "ns3d_kernel_times": [_VP, _DP, ctypes.POINTER(ctypes.c_int)],
    "ns3d_energy_spectrum": [_VP, _DP, ctypes.c_int],
    "ns3d_turbulence_statistics": [_VP, ctypes.POINTER(CTurbulence)],
    "ns3d_copy_field_to_device": [_VP, ctypes.c_int, _VP],
    "ns3d_copy_field_from_device": [_VP, ctypes.c_int, _VP],
    "ns3d_set_state": [_VP, _VP5],
    "ns3d_get_state": [_VP, _VP5],
    "ns3d_tangent_rhs": [_VP, _VP5, _VP5],
    "ns3d_adjoint_rhs": [_VP, _VP5, _VP5],
    "ns3d_tangent_step": [_VP, ctypes.c_double, _VP5],
    "ns3d_adjoint_step": [_VP, ctypes.c_double, _VP5],
    "ns3d_step_stages": [_VP, ctypes.c_double, _VP5, _VP5],
    "ns3d_adjoint_step_stages": [_VP, ctypes.c_double, _VP5, _VP5, _VP5],
    "ns3d_check_face_tangent": [ctypes.POINTER(CParams), _DP, _DP, _DP, _DP],
    "ns3d_check_face_adjoint": [ctypes.POINTER(CParams), _DP, _DP, _DP],
    "ns3d_check_visc_tangent": [ctypes.POINTER(CParams), _DP, _DP, _DP, _DP, _DP, _DP],
    "ns3d_check_visc_adjoint": [ctypes.POINTER(CParams), _DP, _DP, _DP, _DP, _DP],
    "ns3d_check_c2p_adjoint": [ctypes.POINTER(CParams), _DP, _DP, _DP],
    "ns3d_check_weno_pair_adjoint": [ctypes.POINTER(CParams), _DP, ctypes.c_double, ctypes.c_double, _DP, _DP],
    "ns3d_interpolate_batch": [ctypes.POINTER(CParams), _VP, ctypes.c_int, ctypes.c_double, _VP, _VP],
    "ns3d_riemann_batch": [ctypes.POINTER(CParams), _VP, _VP, ctypes.c_int, ctypes.c_int, _VP, _VP],
    "ns3d_scalar_interpolate": [ctypes.POINTER(CParams), _DP, ctypes.c_int, ctypes.c_double, _DP],
    "ns3d_scalar_riemann": [ctypes.POINTER(CParams), _DP, _DP, ctypes.c_int, _DP],
    "ns3d_scalar_face_flux": [ctypes.POINTER(CParams), _DP, _DP, ctypes.c_int, _DP],
}


def load_library(path: str | None = None):
    """Load libns3d_hip.so (once) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is not built: run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(p)
    for name, argtypes in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.ns3d_algorithmic_flops.restype = ctypes.c_int
    lib.ns3d_algorithmic_flops.argtypes = [ctypes.POINTER(CParams), ctypes.POINTER(CFlops)]
    lib.ns3d_last_error.restype = ctypes.c_char_p
    lib.ns3d_last_error.argtypes = []
    lib.ns3d_cell_count.restype = ctypes.c_size_t
    lib.ns3d_cell_count.argtypes = [_VP]
    lib.ns3d_face_count.restype = ctypes.c_size_t
    lib.ns3d_face_count.argtypes = [_VP, ctypes.c_int]
    _lib = lib
    return lib


def check(status: int):
    """Raise NS3DError for a nonzero status."""
    if status != 0:
        raise NS3DError(status, load_library().ns3d_last_error().decode())


def exported_symbols() -> list[str]:
    """Names of the entry points declared in include/ns3d_hip.h."""
    return sorted(list(_SIGS) + ["ns3d_last_error", "ns3d_cell_count", "ns3d_face_count", "ns3d_algorithmic_flops"])


def algorithmic_flops(params) -> dict:
    """Algorithmic FP operations per cell of the configured scheme (host-side
    count on the device numerics; needs no GPU)."""
    out = CFlops()
    cp = params.to_c()
    check(load_library().ns3d_algorithmic_flops(ctypes.byref(cp), ctypes.byref(out)))
    return {n: getattr(out, n) for n, _ in CFlops._fields_}
