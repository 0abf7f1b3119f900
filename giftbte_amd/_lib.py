"""ctypes binding of include/giftbte.h (libgiftbte_hip.so)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgiftbte_hip.so")

GBTE_OK, GBTE_BAD_INPUT, GBTE_DIVERGED, GBTE_NAN, GBTE_HIP_ERROR, GBTE_UNSUPPORTED = range(6)
GBTE_FP64, GBTE_FP32_SWEEP = 0, 1
GBTE_LINSOLVE_DIRECT, GBTE_LINSOLVE_PCG = 0, 1
GBTE_ORDER_1, GBTE_ORDER_2, GBTE_ORDER_2LIMIT = 1, 2, 3
GBTE_SCHEME_SYNTHETIC, GBTE_SCHEME_SEQUENTIAL = 0, 1
GBTE_MATRIX_LU, GBTE_MATRIX_BICGSTAB = 0, 1

_pi, _pd = C.POINTER(C.c_int), C.POINTER(C.c_double)


class Tables(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("n_cells", C.c_int), ("n_nodes", C.c_int), ("n_bounds", C.c_int), ("n_faces", C.c_int),
        ("n_bands", C.c_int), ("n_dirs", C.c_int), ("n_matters", C.c_int), ("material_dim", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("face_size", _pi), ("face_neighbor", _pi), ("face_bound", _pi), ("face_index", _pi),
        ("face_area", _pd), ("face_nx", _pd), ("face_ny", _pd), ("face_nz", _pd),
        ("face_cx", _pd), ("face_cy", _pd), ("face_cz", _pd),
        ("face_nvert", _pi), ("face_vert", _pi), ("cell_nvert", _pi), ("cell_vert", _pi),
        ("volume", _pd), ("cx", _pd), ("cy", _pd), ("cz", _pd), ("heat_source", _pd),
        ("matter", _pi), ("nbr_ptr", _pi), ("nbr_idx", _pi),
        ("node_x", _pd), ("node_y", _pd), ("node_z", _pd),
        ("bound_cell", _pi), ("bound_face", _pi), ("bound_type", _pi), ("bound_temp", _pd), ("bound_connect", _pi),
        ("fface_cell", _pi), ("fface_face", _pi), ("fface_bound", _pi),
        ("kappa_bulk", _pd), ("capacity_bulk", _pd),
        ("vg", _pd), ("tau", _pd), ("cap", _pd), ("lr", _pd), ("hr", _pd),
        ("dir_x", _pd), ("dir_y", _pd), ("dir_z", _pd), ("weight", _pd), ("initial_temp", _pd),
        ("max_iter", C.c_int), ("tol_temp", C.c_double), ("tol_flux", C.c_double), ("tref", C.c_double),
        ("order", C.c_int), ("scheme", C.c_int), ("matrix_solver", C.c_int),
    ]

    def shape_of(self, name: str):
        nc, nb, nf, nn = self.n_cells, self.n_bounds, self.n_faces, self.n_nodes
        nmb = self.n_matters * self.n_bands
        table = {
            "face_size": (nc,), "face_neighbor": (nc, 6), "face_bound": (nc, 6), "face_index": (nc, 6),
            "face_area": (nc, 6), "face_nx": (nc, 6), "face_ny": (nc, 6), "face_nz": (nc, 6),
            "face_cx": (nc, 6), "face_cy": (nc, 6), "face_cz": (nc, 6), "face_nvert": (nc, 6),
            "face_vert": (nc, 6, 4), "cell_nvert": (nc,), "cell_vert": (nc, 8),
            "volume": (nc,), "cx": (nc,), "cy": (nc,), "cz": (nc,), "heat_source": (nc,), "matter": (nc,),
            "nbr_ptr": (nc + 1,), "node_x": (nn,), "node_y": (nn,), "node_z": (nn,),
            "bound_cell": (nb, 2), "bound_face": (nb, 2), "bound_type": (nb,), "bound_temp": (nb,),
            "bound_connect": (nb,), "fface_cell": (nf, 2), "fface_face": (nf, 2), "fface_bound": (nf,),
            "kappa_bulk": (self.n_matters,), "capacity_bulk": (self.n_matters,),
            "vg": (self.n_matters, self.n_bands), "tau": (self.n_matters, self.n_bands),
            "cap": (self.n_matters, self.n_bands), "lr": (self.n_matters, self.n_bands),
            "hr": (self.n_matters, self.n_bands),
            "dir_x": (self.n_dirs,), "dir_y": (self.n_dirs,), "dir_z": (self.n_dirs,), "weight": (self.n_dirs,),
            "initial_temp": (nc,),
        }
        if name == "nbr_idx":
            return (int(self.array("nbr_ptr")[-1]),)
        del nmb
        return table[name]

    def array(self, name: str) -> np.ndarray:
        """Copy of one table as a numpy array."""
        shape = self.shape_of(name)
        n = int(np.prod(shape))
        ptr = getattr(self, name)
        if n == 0:
            dt = np.int32 if isinstance(ptr, _pi) else np.float64
            return np.zeros(shape, dt)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy().reshape(shape)


class Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("device", C.c_int), ("verbose", C.c_int), ("macro_rtol", C.c_double),
                ("linear_solver", C.c_int), ("fallback_kernels", C.c_int), ("sweep_kernel", C.c_int),
                ("gradient_kernel", C.c_int), ("source_layout", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("macro_passes", C.c_int), ("fourier_passes", C.c_int),
                ("converged", C.c_int), ("err_temp", C.c_double), ("err_flux", C.c_double),
                ("solve_ms", C.c_double), ("sweep_ms", C.c_double),
                ("transport_ms", C.c_double), ("macro_ms", C.c_double), ("sweep_levels", C.c_int),
                ("macro_cg_iters", C.c_int), ("source_ms", C.c_double), ("face_ms", C.c_double),
                ("grad_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AdjointStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("change", C.c_double), ("solve_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def load(path: str | None = None) -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    vp = C.c_void_p
    lib.gbte_problem_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.gbte_problem_tables.argtypes = [vp, C.POINTER(Tables)]
    lib.gbte_problem_free.argtypes = [vp]
    lib.gbte_problem_free.restype = None
    lib.gbte_problem_error.restype = C.c_char_p
    if hasattr(lib, "gbte_create"):
        lib.gbte_create.argtypes = [vp, C.POINTER(Options), C.POINTER(vp)]
        lib.gbte_fourier_solve.argtypes = [vp, C.POINTER(Stats)]
        lib.gbte_set_initial.argtypes = [vp, _pd, _pd, _pd, _pd]
        lib.gbte_solve.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.POINTER(Stats)]
        lib.gbte_get_fields.argtypes = [vp, _pd, _pd, _pd, _pd, _pd, _pd]
        lib.gbte_keff.argtypes = [vp, C.c_int, C.c_double, _pd]
        lib.gbte_last_error.argtypes = [vp]
        lib.gbte_last_error.restype = C.c_char_p
        lib.gbte_adjoint.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_double, _pd, _pd,
                                      C.POINTER(AdjointStats)]
        lib.gbte_tuning_log.argtypes = [vp]
        lib.gbte_tuning_log.restype = C.c_char_p
        lib.gbte_destroy.argtypes = [vp]
        lib.gbte_destroy.restype = None
        lib.gbte_write_outputs.argtypes = [vp, vp, C.c_char_p]
        lib.gbte_solve_batch.argtypes = [C.POINTER(vp), C.c_int, _pi, _pd, _pd, C.POINTER(Stats), _pi]
    lib.gbte_write_fields.argtypes = [vp, _pd, _pd, _pd, _pd, _pd, _pd, C.c_char_p]
    for f in ("gbte_flux_keff", "gbte_flux_keff_noalloc"):
        if hasattr(lib, f):
            getattr(lib, f).argtypes = [vp, vp, C.c_int, C.c_double, C.c_double, vp, vp]
    if path is None:
        _lib = lib
    return lib


EXPORTED = [
    "gbte_problem_load", "gbte_problem_tables", "gbte_problem_free", "gbte_create", "gbte_fourier_solve",
    "gbte_set_initial", "gbte_solve", "gbte_get_fields", "gbte_keff", "gbte_last_error", "gbte_problem_error",
    "gbte_destroy", "gbte_write_outputs", "gbte_write_fields", "gbte_solve_batch", "gbte_tuning_log",
    "gbte_flux_keff", "gbte_flux_keff_noalloc", "gbte_adjoint",
]


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_pd)
