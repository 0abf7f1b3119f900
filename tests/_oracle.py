"""ctypes access to the CPU oracle (oracle/liboracle.so) - test-side only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_pd = C.POINTER(C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle/liboracle.so is missing: run __graft_entry__.build()")
        L = C.CDLL(ORACLE_LIB)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_fourier.argtypes = [C.c_void_p, _pd, _pd, _pd, _pd]
        L.oracle_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _pd, _pd, _pd, _pd, _pd, _pd,
                                 C.POINTER(C.c_int), _pd]
        L.oracle_stages.argtypes = [C.c_void_p, C.c_int, C.c_int, _pd, _pd, _pd, _pd, _pd, _pd, _pd]
        L.oracle_lsq_gradient.argtypes = [C.c_void_p, _pd, _pd]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_pd)


def fourier(tables):
    nc = tables.n_cells
    T, qx, qy, qz = (np.zeros(nc) for _ in range(4))
    p = lib().oracle_fourier(C.byref(tables), _p(T), _p(qx), _p(qy), _p(qz))
    if p < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return p, {"T": T, "qx": qx, "qy": qy, "qz": qz}


def run(tables, max_iter, tol_temp, tol_flux):
    """Full steady run; returns (fields, info) with info = iterations, macro
    passes, Fourier passes, status, err_temp, err_flux."""
    nc, nb = tables.n_cells, tables.n_bands
    f = {k: np.zeros(nc) for k in ("T_lattice", "T_energy", "qx", "qy", "qz")}
    tb = np.zeros(nb * nc)
    info = (C.c_int * 4)()
    err = (C.c_double * 2)()
    rc = lib().oracle_run(C.byref(tables), int(max_iter), float(tol_temp), float(tol_flux), _p(f["T_lattice"]),
                          _p(f["T_energy"]), _p(f["qx"]), _p(f["qy"]), _p(f["qz"]), _p(tb), info, err)
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    f["T_band"] = tb.reshape(nb, nc)
    return f, {"iterations": info[0], "macro_passes": info[1], "fourier_passes": info[2], "status": info[3],
               "err_temp": err[0], "err_flux": err[1]}


def stages(tables, band, direction):
    t = tables
    nc, nb, nf, nn = t.n_cells, t.n_bounds, t.n_faces, t.n_nodes
    out = {"grad": np.zeros(3 * nc), "e_new": np.zeros(nc), "ebound_sys": np.zeros(nb * 2),
           "e_vertex": np.zeros(nn * t.n_matters), "gface": np.zeros(3 * nf * 2), "remacro": np.zeros(nc),
           "ebound_init": np.zeros(t.n_bands * t.n_dirs * nb * 2)}
    rc = lib().oracle_stages(C.byref(t), band, direction, *[_p(out[k]) for k in
                                                           ("grad", "e_new", "ebound_sys", "e_vertex", "gface",
                                                            "remacro", "ebound_init")])
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return out


def lsq_gradient(tables, field):
    nc = tables.n_cells
    f = np.ascontiguousarray(field, np.float64)
    g = np.zeros(3 * nc)
    lib().oracle_lsq_gradient(C.byref(tables), _p(f), _p(g))
    return g.reshape(3, nc)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
