"""giftbte_amd - MI355X solver for GiftBTE's steady StaticBTESynthetic path.

Python mirror of the reference's driver flow for this path
(src/BTEMain.cpp:327-393 -> SolutionAll::_Fourier_Solver / _BTE_Solver,
src/Solution/Solution.cpp:236-414) over the C ABI of libgiftbte_hip.so
(include/giftbte.h).  Everything numerical runs in the HIP library; this
module only moves arguments and results.  There is no CPU fallback: if the
library is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib

__all__ = ["Problem", "BTESolver", "SolveStats", "run_deck", "GBTEError"]


class GBTEError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[gbte {code}] {msg}")
        self.code = code


@dataclass
class SolveStats:
    iterations: int
    macro_passes: int
    fourier_passes: int
    converged: bool
    err_temp: float
    err_flux: float
    solve_ms: float
    sweep_ms: float
    transport_ms: float
    macro_ms: float
    sweep_levels: int
    macro_cg_iters: int
    source_ms: float
    face_ms: float
    grad_ms: float = 0.0


class Problem:
    """Deck + mesh + tables (replaces BTEMain.cpp:87-362)."""

    def __init__(self, deck_dir: str):
        self.lib = _lib.load()
        self._h = C.c_void_p()
        rc = self.lib.gbte_problem_load(deck_dir.encode(), C.byref(self._h))
        if rc != 0:
            raise GBTEError(rc, self.lib.gbte_problem_error().decode())
        self.tables = _lib.Tables()
        self.lib.gbte_problem_tables(self._h, C.byref(self.tables))

    @property
    def handle(self):
        return self._h

    def write_fields(self, out_dir: str, T_lattice, T_energy, qx, qy, qz, T_band):
        """The reference's output files from given fields (deck order), no
        device needed."""
        arrs = [np.ascontiguousarray(x, np.float64).ravel() for x in (T_lattice, T_energy, qx, qy, qz, T_band)]
        nc, nb = self.tables.n_cells, self.tables.n_bands
        if any(a.size != nc for a in arrs[:5]) or arrs[5].size != nb * nc:
            raise ValueError(f"write_fields needs {nc} values per field and {nb * nc} band temperatures")
        rc = self.lib.gbte_write_fields(self.handle, *[_lib.dptr(x) for x in arrs], out_dir.encode())
        if rc != 0:
            raise GBTEError(rc, "cannot write the output files")

    def array(self, name: str) -> np.ndarray:
        return self.tables.array(name)

    # the flattened mesh tables of a tetrahedral (or 2D) problem, as raw
    # little-endian arrays: what a solver that takes a built mesh reads
    # (the reference's BTEMesh public vectors, includes/BTEMesh/BTEMesh.h:27-37)
    TABLE_BINS = ["face_size", "face_neighbor", "face_bound", "face_index", "face_nvert", "face_vert", "cell_nvert",
                  "cell_vert", "matter", "face_area", "face_nx", "face_ny", "face_nz", "face_cx", "face_cy",
                  "face_cz", "volume", "cx", "cy", "cz", "heat_source", "initial_temp", "node_x", "node_y", "node_z",
                  "bound_cell", "bound_face", "bound_type", "bound_connect", "bound_temp", "fface_cell",
                  "fface_face", "fface_bound"]

    def write_tables(self, out_dir: str) -> None:
        """Write the flattened tables to out_dir/<name>.bin (int32 or
        float64), with sizes.bin (cells, nodes, boundaries, faces, dim) and
        scale.bin (L_x, L_y, L_z)."""
        os.makedirs(out_dir, exist_ok=True)
        t = self.tables
        for name in self.TABLE_BINS:
            a = self.array(name)
            a = a.astype("<i4" if a.dtype.kind == "i" else "<f8")
            a.tofile(os.path.join(out_dir, name + ".bin"))
        np.array([t.n_cells, t.n_nodes, t.n_bounds, t.n_faces, t.dim], "<i4").tofile(os.path.join(out_dir, "sizes.bin"))
        np.array([t.Lx, t.Ly, t.Lz], "<f8").tofile(os.path.join(out_dir, "scale.bin"))

    def close(self):
        if self._h:
            self.lib.gbte_problem_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BTESolver:
    """StaticFourier + StaticBTESynthetic on one GPU.

    ``fourier()`` is StaticFourier::_solve (StaticFourier.cpp:934-976);
    ``solve()`` is StaticBTESynthetic::solve (StaticBTESyntheticSolver.cpp:12-247);
    ``fields()`` returns what StaticBTESynthetic::_print_out writes
    (StaticBTESynthetic.cpp:2613-2653).
    """

    def __init__(self, problem: Problem, precision: str = "fp64", device: int = 0, verbose: bool = False,
                 macro_rtol: float = 0.0, linear_solver: str = "direct", fallback_kernels: bool = False,
                 sweep_kernel: str = "auto", gradient_kernel: str = "auto", source_layout: str = "auto"):
        self.problem = problem
        self.lib = problem.lib
        t = problem.tables
        self.nc, self.nb = t.n_cells, t.n_bands
        if linear_solver not in ("direct", "pcg"):
            raise ValueError("linear_solver must be 'direct' or 'pcg'")
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32' (FP32 storage of the per-system fields)")
        sweeps = {"auto": 0, "wavespec": 1, "level": 2, "chunked": 3, "lanes": 4}
        if sweep_kernel not in sweeps:
            raise ValueError(f"sweep_kernel must be one of {sorted(sweeps)}")
        grads = {"auto": 0, "gather": 1, "matrix": 2, "vertex": 3}
        if gradient_kernel not in grads:
            raise ValueError(f"gradient_kernel must be one of {sorted(grads)}")
        layouts = {"auto": 0, "chunk": 1, "position": 2, "cell": 3}
        if source_layout not in layouts:
            raise ValueError(f"source_layout must be one of {sorted(layouts)}")
        opts = _lib.Options(_lib.GBTE_FP64 if precision == "fp64" else _lib.GBTE_FP32_SWEEP, device, int(verbose),
                            macro_rtol,
                            _lib.GBTE_LINSOLVE_DIRECT if linear_solver == "direct" else _lib.GBTE_LINSOLVE_PCG,
                            int(bool(fallback_kernels)), sweeps[sweep_kernel], grads[gradient_kernel],
                            layouts[source_layout])
        self._h = C.c_void_p()
        rc = self.lib.gbte_create(problem.handle, C.byref(opts), C.byref(self._h))
        if rc != 0:
            msg = self.lib.gbte_last_error(self._h).decode() if self._h else "create failed"
            self.close()
            raise GBTEError(rc, msg)

    def _check(self, rc: int):
        if rc != 0:
            raise GBTEError(rc, self.lib.gbte_last_error(self._h).decode())

    def fourier(self) -> int:
        st = _lib.Stats()
        self._check(self.lib.gbte_fourier_solve(self._h, C.byref(st)))
        return st.fourier_passes

    def set_initial(self, T=None, qx=None, qy=None, qz=None):
        conv = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (T, qx, qy, qz)]
        self._check(self.lib.gbte_set_initial(self._h, *[_lib.dptr(a) for a in conv]))

    def solve(self, max_iter: int, tol_temp: float, tol_flux: float) -> SolveStats:
        st = _lib.Stats()
        rc = self.lib.gbte_solve(self._h, int(max_iter), float(tol_temp), float(tol_flux), C.byref(st))
        self._check(rc)
        return SolveStats(st.iterations, st.macro_passes, st.fourier_passes, bool(st.converged), st.err_temp,
                          st.err_flux, st.solve_ms, st.sweep_ms, st.transport_ms, st.macro_ms,
                          st.sweep_levels, st.macro_cg_iters, st.source_ms, st.face_ms, st.grad_ms)

    def fields(self) -> dict:
        nc, nb = self.nc, self.nb
        out = {k: np.zeros(nc) for k in ("T_lattice", "T_energy", "qx", "qy", "qz")}
        out["T_band"] = np.zeros(nb * nc)
        self._check(self.lib.gbte_get_fields(self._h, *[_lib.dptr(out[k]) for k in
                                                        ("T_lattice", "T_energy", "qx", "qy", "qz", "T_band")]))
        out["T_band"] = out["T_band"].reshape(nb, nc)
        return out

    def keff(self, axis: int, delta_t: float) -> float:
        v = C.c_double()
        self._check(self.lib.gbte_keff(self._h, axis, delta_t, C.byref(v)))
        return v.value

    def adjoint(self, axis: int = 0, delta_t: float = 2.0, max_iter: int = 100000, tol: float = 1e-13):
        """Sensitivities of keff(axis, delta_t) at the converged state of the
        last solve (gbte_adjoint): (dk/ds per band for tau_b -> s_b tau_b,
        dk/dT per boundary face (isothermal faces; 0 elsewhere), stats)."""
        ds = np.zeros(self.nb)
        dtb = np.zeros(self.problem.tables.n_bounds)
        st = _lib.AdjointStats()
        self._check(self.lib.gbte_adjoint(self._h, int(axis), float(delta_t), int(max_iter), float(tol),
                                          _lib.dptr(ds), _lib.dptr(dtb), C.byref(st)))
        return ds, dtb, st.as_dict()

    def tuning_log(self) -> str:
        """The kernel variants timed at set-up, their milliseconds per launch
        and the choices made (empty where nothing was tuned)."""
        return self.lib.gbte_tuning_log(self._h).decode()

    def write_outputs(self, out_dir: str = "."):
        """The reference's output files (TempLattice.dat, Tempband.dat,
        Temperature.dat, HeatFlux.dat, Result.vtk) in out_dir."""
        self._check(self.lib.gbte_write_outputs(self._h, self.problem.handle, out_dir.encode()))

    def close(self):
        if getattr(self, "_h", None):
            self.lib.gbte_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batch(solvers, max_iter, tol_temp, tol_flux):
    """BTESolver.solve on several independent solvers at once (gbte_solve_batch:
    one host thread per solver, the device runs them side by side).
    max_iter / tol_temp / tol_flux: one value or one per solver.  Returns the
    SolveStats list; raises GBTEError for the first failing solver."""
    n = len(solvers)
    if n == 0:
        return []
    lib = solvers[0].lib

    def per(x, typ):
        v = list(x) if hasattr(x, "__len__") else [x] * n
        return (typ * n)(*v)

    hs = (C.c_void_p * n)(*[s._h for s in solvers])
    st = (_lib.Stats * n)()
    status = (C.c_int * n)()
    rc = lib.gbte_solve_batch(hs, n, per(max_iter, C.c_int), per(tol_temp, C.c_double), per(tol_flux, C.c_double),
                              st, status)
    if rc != 0:
        bad = next((i for i in range(n) if status[i] != 0), None)
        if bad is None:
            raise GBTEError(rc, "solve_batch failed")
        raise GBTEError(status[bad], solvers[bad].lib.gbte_last_error(solvers[bad]._h).decode())
    return [SolveStats(x.iterations, x.macro_passes, x.fourier_passes, bool(x.converged), x.err_temp, x.err_flux,
                       x.solve_ms, x.sweep_ms, x.transport_ms, x.macro_ms, x.sweep_levels, x.macro_cg_iters,
                       x.source_ms, x.face_ms, x.grad_ms) for x in st]


def run_deck(deck_dir: str, max_iter: int | None = None, tol_temp: float | None = None,
             tol_flux: float | None = None, precision: str = "fp64", verbose: bool = False):
    """The reference's steady flow on one deck: Fourier warm start, then the
    synthetic BTE solve (BTEMain.cpp:372-393).  Returns (problem, solver, stats)."""
    p = Problem(deck_dir)
    s = BTESolver(p, precision=precision, verbose=verbose)
    s.fourier()
    t = p.tables
    st = s.solve(t.max_iter if max_iter is None else max_iter, t.tol_temp if tol_temp is None else tol_temp,
                 t.tol_flux if tol_flux is None else tol_flux)
    return p, s, st
