"""Benchmark: steady StaticBTESynthetic outer iterations on one MI355X.

Workload (the reference-runnable headline case): 2D nanoporous film, 500 nm
square minus a centred r = 150 nm diffuse pore, seeded Delaunay mesh of about
50k triangles, 40 Debye-like bands, 64 in-plane directions (CADOM, Nphi 16),
left/right isothermal 301/299 K, top/bottom periodic.  Synthetic data from the
in-repo seeded generator (giftbte_amd/decks.py).

A "step" is one outer iteration of StaticBTESynthetic::solve: every
band x direction transport sweep, the moment reductions, the boundary update
and the FP64 synthetic (macro) solve.  value = cell*band*direction updates/s
over the timed steps.  The time to a converged k_eff is reported beside it.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--converge]
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def make_deck(td, n_target, nbands, nphi, seed):
    from giftbte_amd import decks
    spec = decks.porous_spec(n_target=n_target, nbands=nbands, nphi=nphi, pore_radius=0.3, seed=seed, scale=5e-7)
    return decks.write_deck(os.path.join(td, "porous"), spec)


def cpu_baseline(problem, T0, target_s=15.0):
    """The CPU oracle's per-system transport loop on the same mesh and
    modes, timed on a bounded number of systems (single core)."""
    import ctypes as C
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    L = _oracle.lib()
    L.oracle_transport_sample.restype = C.c_double
    L.oracle_transport_sample.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
    t = problem.tables
    T0 = np.ascontiguousarray(T0, np.float64)
    ptr = T0.ctypes.data_as(C.POINTER(C.c_double))
    n = 4
    sec = L.oracle_transport_sample(C.byref(t), ptr, n)
    if sec <= 0:
        return None
    n = int(max(4, min(t.n_bands * t.n_dirs, n * target_s / max(sec, 1e-3))))
    sec = L.oracle_transport_sample(C.byref(t), ptr, n)
    upd = float(n) * t.n_cells
    return {"value": upd / sec, "unit": "cell*band*dir updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle transport loop (gradient, residual + exact sweep, boundary values, vertex "
                      f"IDW, face gradients, non-Fourier residual, moments) over {n} of "
                      f"{t.n_bands * t.n_dirs} band x direction systems of one outer iteration, same mesh; "
                      f"{sec:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=50000)
    ap.add_argument("--bands", type=int, default=40)
    ap.add_argument("--nphi", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--converge", action="store_true", help="also time a full solve to the deck tolerance")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device once (calibrates HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build runs the whole problem on a single MI355X "
              "(no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    from giftbte_amd import Problem, BTESolver

    if a.copy_probe > 0:
        import torch
        n = a.copy_probe * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(f"[bench] copy probe: {a.copy_probe} MiB read + {a.copy_probe} MiB written", file=sys.stderr, flush=True)
        del x, y

    with tempfile.TemporaryDirectory() as td:
        t_gen = time.time()
        deck = make_deck(td, a.cells, a.bands, a.nphi, a.seed)
        p = Problem(deck)
        t = p.tables
        nc, nB, nD = t.n_cells, t.n_bands, t.n_dirs
        print(f"[bench] deck {nc} cells x {nB} bands x {nD} directions ({time.time() - t_gen:.1f} s set-up)",
              file=sys.stderr, flush=True)
        s = BTESolver(p)
        t0 = time.time()
        fp = s.fourier()
        t_fourier = time.time() - t0
        print(f"[bench] fourier {fp} passes {t_fourier:.2f} s", file=sys.stderr, flush=True)
        T0 = s.fields()["T_lattice"]
        st_w = s.solve(a.warmup, 0.0, 0.0)
        print(f"[bench] warmup {st_w.iterations} it {st_w.solve_ms:.1f} ms", file=sys.stderr, flush=True)
        t1 = time.time()
        st = s.solve(a.steps, 0.0, 0.0)
        wall = time.time() - t1
        U = nc * nB * nD
        ms_step = wall * 1e3 / a.steps
        value = U * a.steps / wall
        # roofline of the transport step (sweep + vertex + face/moment kernels):
        # algorithmic bytes per update (2 + 2 dim) * 8 (read+write of e, write+read of grad)
        bytes_upd = (2 + 2 * t.dim) * 8
        tr_s = st.transport_ms / 1e3
        achieved = bytes_upd * U * a.steps / tr_s / 1e9 if tr_s > 0 else 0.0
        out = {
            "metric": "cell*band*dir updates/s (steady StaticBTESynthetic outer iterations, FP64)",
            "value": value, "unit": "cell*band*dir updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic (seeded porous Delaunay deck)",
            "config": {"workload": "2D porous film 500 nm, r=150 nm diffuse pore, seeded Delaunay",
                       "cells": nc, "bands": nB, "directions": nD, "systems": nB * nD,
                       "updates_per_step": U, "parallelism": "single GPU"},
            "breakdown_ms_per_step": {"transport": st.transport_ms / a.steps, "sweep": st.sweep_ms / a.steps,
                                      "macro": st.macro_ms / a.steps,
                                      "macro_passes_per_step": st.macro_passes / a.steps,
                                      "cg_iters_per_step": st.macro_cg_iters / a.steps,
                                      "sweep_levels": st.sweep_levels},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "kernel": "transport step (k_sweep levels + k_vertex_e + k_face_moments)",
                         "bytes_per_update": bytes_upd},
        }
        if a.converge:
            s2 = BTESolver(p)
            tc0 = time.time()
            s2.fourier()
            stc = s2.solve(t.max_iter, t.tol_temp, t.tol_flux)
            tc = time.time() - tc0
            out["time_to_converged_s"] = tc
            out["converged_iterations"] = stc.iterations
            out["keff"] = s2.keff(0, 2.0)
        if not a.no_cpu:
            out["cpu_baseline"] = cpu_baseline(p, T0)
        print(json.dumps(out))


if __name__ == "__main__":
    main()
