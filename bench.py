"""Benchmark: steady StaticBTESynthetic outer iterations on one MI355X.

Default workload (the 3D film): a 200 nm cube with a centred 80 nm cubic
void, seeded Delaunay tetrahedra from a jittered 37^3 lattice (about 295k
cells), 40 Debye-like bands, 96 directions (Gauss-Legendre, 4 polar x 24
azimuthal), x = 0 / x = 1 faces isothermal at 301 / 299 K, the other faces
and the void diffuse.  --workload 2d runs the 2D non-gray nanoporous film
instead: 500 nm square minus a centred r = 150 nm diffuse pore, seeded
Delaunay mesh of about 50k triangles, 40 bands, 64 in-plane directions,
left/right isothermal, top/bottom periodic.  Synthetic data from the in-repo
seeded generator (giftbte_amd/decks.py).

A "step" is one outer iteration of StaticBTESynthetic::solve: every
band x direction transport sweep, the moment reductions, the boundary update
and the FP64 synthetic (macro) solve.  Updates/s = cell*band*direction
updates per second over the timed steps (inputs resident in HBM).  Two
solvers run on the same deck: FP64 storage, and "mixed" (FP32 storage of
energy densities, gradients and sources; FP64 sweep arithmetic, temperature
and flux recovery, synthetic solve and reductions).  On the 3D workload the
headline value is the mixed solver's rate (the precision that workload is
defined in) with FP64 beside it; on 2D it is FP64 with mixed beside it.  For
both, the time to a converged k_eff (the deck's tolerances, from the Fourier
warm start) and the end-to-end time including gbte_create are reported, with
the mixed k_eff's error against FP64.

The CPU baseline is the reference itself, compiled from its sources, run with
MPI ranks on this host on a bounded sample: on 2D its BTE_CPU on the same mesh
with a subset of the directions; on 3D, where the reference cannot build the
mesh (its 3D mesh construction is removed and BTEMain.cpp:218-225 exits), its
StaticBTESynthetic solve() driven by oracle/ref_dump on our tables of a
smaller deck of the same generator.  Where those binaries are absent, the
oracle's single-core transport loop on the same mesh is the baseline.  The
oracle's figure is reported beside the reference's as cpu_port.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 3d|2d]
                       [--no-converge] [--no-cpu] [--batch N] [--adjoint]
"""
import argparse
import dataclasses
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILES = {"3d": os.path.join(ROOT, "profiles", "traffic_cube3d_300k.json"),
                 "2d": os.path.join(ROOT, "profiles", "traffic_porous50k.json")}


def make_deck(td, a):
    from giftbte_amd import decks
    if a.workload == "3d":
        spec = decks.cube_void_spec(m=a.lattice, nbands=a.bands, ntheta=2, nphi=6, seed=a.seed)
        return decks.write_deck(os.path.join(td, "cube3d"), spec)
    spec = decks.porous_spec(n_target=a.cells, nbands=a.bands, nphi=a.nphi, pore_radius=0.3, seed=a.seed, scale=5e-7)
    return decks.write_deck(os.path.join(td, "porous"), spec)


def cpu_baseline(problem, T0, target_s=12.0):
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
                      f"{sec:.1f} s on 1 of {host_cores()} host cores"}


def reference_baseline_3d(td, nbands, seed, lattice=14, iters=2, timeout=600):
    """The reference's own StaticBTESynthetic (compiled from its sources with
    oracle/ref_dump) on tetrahedra: the reference cannot build 3D meshes
    (BTEMain.cpp:218-225 exits), so ref_dump --tables --time hands it our
    tables of a cube-with-void deck of the same generator and runs solve()
    with the band x direction systems over MPI ranks on this host's cores.
    Bounded sample: a lattice of `lattice` cells per edge (about 17k
    tetrahedra at 14; the reference's own set-up, per rank, grows faster than
    its iterations), the same bands, 32 directions.  Rate = updates of one
    outer iteration / the reference's per-iteration inner-loop timer; rank
    count as for 2D."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
    mpi = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpi)):
        return None
    from giftbte_amd import decks, Problem
    spec = decks.cube_void_spec(m=lattice, nbands=nbands, ntheta=2, nphi=2, seed=seed)
    deck = decks.write_deck(os.path.join(td, "refdeck3d"), spec)
    p = Problem(deck)
    p.write_tables(os.path.join(td, "reftables3d"))
    t = p.tables
    nd = t.n_dirs
    cores = min(16, host_cores())
    ranks = max(r for r in range(1, nd + 1) if nd % r == 0 and r <= cores)
    t0 = time.time()
    try:
        r = subprocess.run(["timeout", str(timeout), mpi, "-np", str(ranks), exe, deck, os.path.join(td, "refout3d"),
                            str(iters), "0", "0", "--tables", os.path.join(td, "reftables3d"), "--time"],
                           cwd=deck, capture_output=True, text=True)
    except OSError:
        return None
    wall = time.time() - t0
    ms = [float(x) for x in re.findall(r"Time taken by inner loop: ([0-9.eE+-]+) milliseconds", r.stdout)]
    if r.returncode != 0 or not ms:
        print(f"[bench] reference 3D baseline failed (rc {r.returncode})", file=sys.stderr, flush=True)
        return None
    upd = float(t.n_cells) * t.n_bands * nd
    sec = sum(ms) / len(ms) / 1e3
    return {"value": upd / sec, "unit": "cell*band*dir updates/s", "cores": ranks, "kind": "reference",
            "sample": f"the reference's StaticBTESynthetic::solve (compiled from its sources; oracle/ref_dump "
                      f"--tables --time hands it our tetrahedral tables, its own 3D mesh reader being removed), "
                      f"mpiexec -np {ranks}, on a {t.n_cells}-tet cube-with-void deck of the same generator, "
                      f"{t.n_bands} bands, {nd} directions ({t.n_bands * nd} systems); mean of its per-iteration "
                      f"inner-loop timer over {len(ms)} outer iterations, {sec * 1e3:.0f} ms; {wall:.0f} s wall "
                      f"with set-up"}


def host_cores():
    """Cores this process may run on (the host's share on a shared box)."""
    return len(os.sched_getaffinity(0))


def reference_baseline(td, n_target, nbands, seed, nphi=4, iters=2, timeout=900):
    """The compiled reference (oracle/_ref/BTE_CPU, built from the reference
    sources by oracle/ref.mk) run with MPI ranks on this host's cores, on a
    bounded sample of the workload: the same mesh and bands, 4 * nphi of the
    directions.  Rate = updates of one outer iteration / the reference's own
    per-iteration "inner loop" timer (StaticBTESyntheticSolver.cpp:77-190),
    which leaves out its per-iteration file output.  The rank count is the
    largest divisor of the direction count within 16 cores (its per-system
    collectives need the divisor)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "BTE_CPU")
    mpi = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpi)):
        return None
    from giftbte_amd import decks
    nd = 4 * nphi
    cores = min(16, host_cores())
    ranks = max(r for r in range(1, nd + 1) if nd % r == 0 and r <= cores)
    spec = decks.porous_spec(n_target=n_target, nbands=nbands, nphi=nphi, pore_radius=0.3, seed=seed, scale=5e-7)
    spec = dataclasses.replace(spec, max_iter=iters, tol_temp=0.0, tol_flux=0.0)
    deck = decks.write_deck(os.path.join(td, "refdeck"), spec)
    t0 = time.time()
    try:
        r = subprocess.run(["timeout", str(timeout), mpi, "-np", str(ranks), exe], cwd=deck, capture_output=True,
                           text=True)
    except OSError:
        return None
    wall = time.time() - t0
    ms = [float(x) for x in re.findall(r"Time taken by inner loop: ([0-9.eE+-]+) milliseconds", r.stdout)]
    if r.returncode != 0 or not ms:
        print(f"[bench] reference baseline failed (rc {r.returncode})", file=sys.stderr, flush=True)
        return None
    from giftbte_amd import Problem
    t = Problem(deck).tables
    upd = float(t.n_cells) * t.n_bands * t.n_dirs
    sec = sum(ms) / len(ms) / 1e3
    return {"value": upd / sec, "unit": "cell*band*dir updates/s", "cores": ranks, "kind": "reference",
            "sample": f"the reference's BTE_CPU (compiled from its sources), mpiexec -np {ranks}, on the same "
                      f"{t.n_cells}-cell mesh and {t.n_bands} bands with {t.n_dirs} of the directions "
                      f"({t.n_bands * t.n_dirs} systems); mean of its per-iteration inner-loop timer over {len(ms)} "
                      f"outer iterations, {sec * 1e3:.0f} ms; {wall:.0f} s wall with set-up"}


def kernel_bytes_per_update(dim, word=8):
    """Algorithmic HBM bytes per cell*band*dir update of each transport kernel
    (word = 8 for FP64 storage, 4 for FP32), tables shared by all
    band x direction systems excluded:
      source     2D: reads the LSQ gradient (2 values), writes the sweep source;
                 3D: reads one face projection of the gradient per face (4
                 values: the cell's own for out-faces, the neighbour's for
                 in-faces), writes the sweep source
      sweep      reads the source, writes e
      2D face    k_tile_face reads e, writes the gradient (2 values)
      3D grad    reads e, writes the gradient's four face projections (the
                 vertex-sum variant's sums are intermediate traffic, counted
                 in the measured traffic, not in the algorithmic bytes)
      3D moments the moment kernels read e"""
    if dim == 3:
        return {"source": 5 * word, "sweep": 2 * word, "grad": 5 * word, "moments": word}
    return {"source": (dim + 1) * word, "sweep": 2 * word, "face": (1 + dim) * word}


def kernel_name(phase, dim, prec, tuning):
    """The kernel a timed phase launches (the tuner's choice on 3D)."""
    if dim == 2:
        return {"source": "k_source", "sweep": "k_sweep_pc", "face": "k_tile_face"}[phase]
    if phase == "sweep":
        if "sweep chosen: chunked" in tuning:
            return "k_sweep_lv"
        return "k_sweep_dir" if "sweep chosen: level-synchronous" in tuning else "k_sweep_pc"
    if phase == "grad":
        if "gradient chosen: k_grad3m" in tuning:
            return "k_grad3m"
        return "k_vsum3 + k_grad3v" if "gradient chosen: k_vsum3 + k_grad3v" in tuning else "k_grad3g"
    return {"source": "k_source3", "moments": "k_mom3 + k_vert3 + k_fin3"}[phase]


def load_traffic(kernel, cfg, prec, workload):
    """HBM bytes per launch of `kernel` at storage precision `prec` from the
    committed rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE scaled by a
    copy-kernel calibration), if it was measured on this workload."""
    try:
        with open(TRAFFIC_FILES[workload]) as f:
            tr = json.load(f)
    except (OSError, ValueError):
        return None
    if tr.get("workload") != {k: cfg[k] for k in ("cells", "bands", "directions")}:
        return None
    k = tr.get("kernels", {}).get(prec, {}).get(kernel)
    return None if k is None else k["hbm_bytes_per_launch"]


def run_batch(a):
    """--batch N: N independent porous instances of uneven size, 500 nm
    square with a pore of radius uniform in [50, 200] nm, log-uniform in
    [2k, 40k] triangles, 20 bands, 32 directions (seeded), each from its
    Fourier warm start to its deck tolerance, all side by side
    (gbte_solve_batch), in FP64 and then FP32 storage.  Prints one JSON line:
    whole-batch updates/s and wall time to every converged k_eff, and the
    largest per-instance k_eff error of FP32 against FP64."""
    import numpy as np
    from giftbte_amd import BTESolver, Problem, decks, solve_batch
    rng = np.random.default_rng(a.seed)
    with tempfile.TemporaryDirectory() as td:
        t_gen = time.time()
        probs = []
        for i in range(a.batch):
            n = int(np.exp(rng.uniform(np.log(2000), np.log(40000))))
            r = rng.uniform(50e-9, 200e-9) / 500e-9
            spec = decks.porous_spec(n_target=n, nbands=20, nphi=8, pore_radius=r, seed=a.seed + 100 + i, scale=5e-7)
            probs.append(Problem(decks.write_deck(os.path.join(td, f"inst{i}"), spec)))
        cells = [p.tables.n_cells for p in probs]
        print(f"[bench] batch of {a.batch}: {sum(cells)} cells in all ({min(cells)}..{max(cells)}), "
              f"{time.time() - t_gen:.1f} s set-up", file=sys.stderr, flush=True)
        res = {}
        for prec in ("fp64", "fp32"):
            sols = [BTESolver(p, precision=prec) for p in probs]
            for s in sols:
                s.fourier()
            t = [p.tables for p in probs]
            t0 = time.time()
            sts = solve_batch(sols, [x.max_iter for x in t], [x.tol_temp for x in t], [x.tol_flux for x in t])
            wall = time.time() - t0
            upd = sum(st.iterations * x.n_cells * x.n_bands * x.n_dirs for st, x in zip(sts, t))
            res[prec] = {"wall_s": wall, "updates_per_s": upd / wall, "iterations": [st.iterations for st in sts],
                         "converged": all(st.converged for st in sts), "keff": [s.keff(0, 2.0) for s in sols]}
            for s in sols:
                s.close()
            print(f"[bench] batch {prec}: {wall:.2f} s, {upd / wall:.3e} updates/s", file=sys.stderr, flush=True)
        k64, k32 = np.array(res["fp64"]["keff"]), np.array(res["fp32"]["keff"])
        out = {"metric": "batched time-to-converged k_eff & cell\u00b7band\u00b7dir updates/s, mixed vs FP64",
               "value": res["fp64"]["updates_per_s"], "unit": "cell*band*dir updates/s", "n_gpus": 1,
               "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded porous instances)",
               "config": {"workload": f"batch of {a.batch} 2D porous instances, 500 nm square, pore radius 50-200 nm, "
                                      f"2k-40k triangles, 20 bands, 32 dirs", "instances": a.batch,
                          "cells_total": int(sum(cells)), "parallelism": "single GPU, one stream per instance"},
               "fp64": {k: v for k, v in res["fp64"].items() if k != "keff"},
               "mixed": dict({k: v for k, v in res["fp32"].items() if k != "keff"},
                             keff_rel_err_vs_fp64_max=float(np.max(np.abs(k32 - k64) / np.abs(k64))),
                             speedup_time_to_converged_vs_fp64=res["fp64"]["wall_s"] / res["fp32"]["wall_s"])}
        print(json.dumps(out))


def run_adjoint(a):
    """--adjoint: the 2D porous film (as --workload 2d) solved in FP64 to
    margins of 1e-15 / 1e-13 from the Fourier warm start, then gbte_adjoint
    for dk_eff/d(tau scale) of every band and dk_eff/dT_hot.  Checked on the
    host against central differences (step 1e-4) of the GPU solve for the
    --adjoint-fd bands with the largest sensitivities, and against the closed
    form dk/dT_hot = k / (T_hot - T_cold).  Prints one JSON line."""
    import numpy as np
    from giftbte_amd import BTESolver, Problem, decks
    spec0 = decks.porous_spec(n_target=a.cells, nbands=a.bands, nphi=a.nphi, pore_radius=0.3, seed=a.seed,
                              scale=5e-7)
    with tempfile.TemporaryDirectory() as td:
        def solve(spec, name):
            p = Problem(decks.write_deck(os.path.join(td, name), spec))
            t0 = time.time()
            s = BTESolver(p)
            t1 = time.time()
            s.fourier()
            st = s.solve(100000, 1e-15, 1e-13)
            t2 = time.time()
            return p, s, st, t1 - t0, t2 - t1
        p, s, st, t_create, t_solve = solve(spec0, "base")
        t = p.tables
        k0 = s.keff(0, 2.0)
        print(f"[bench] adjoint: {t.n_cells} cells x {t.n_bands} bands x {t.n_dirs} dirs, forward {st.iterations} it "
              f"{t_solve:.2f} s (create {t_create:.2f} s), k {k0:.12g}", file=sys.stderr, flush=True)
        t0 = time.time()
        g, gtb, ast = s.adjoint(0, 2.0, 100000, 1e-13)
        t_adj = time.time() - t0
        bt, btemp = p.array("bound_type"), p.array("bound_temp")
        g_hot = float(gtb[(bt == 1) & (btemp == 301.0)].sum())
        s.close()
        print(f"[bench] adjoint: {ast['iterations']} it {t_adj:.2f} s, dk/dT_hot {g_hot:.10g} (k/2 {k0 / 2:.10g})",
              file=sys.stderr, flush=True)
        fd_bands = [int(b) for b in np.argsort(-np.abs(g))[:a.adjoint_fd]]
        fd, t_fd = {}, time.time()
        for b in fd_bands:
            ks = []
            for sgn in (1, -1):
                sc = np.ones(t.n_bands)
                sc[b] = 1 + sgn * 1e-4
                _, s2, _, _, _ = solve(dataclasses.replace(spec0, bands=decks.scale_tau(spec0.bands, sc)), f"b{b}{sgn}")
                ks.append(s2.keff(0, 2.0))
                s2.close()
            fd[b] = (ks[0] - ks[1]) / 2e-4
            print(f"[bench] band {b}: adjoint {g[b]:.10g} fd {fd[b]:.10g}", file=sys.stderr, flush=True)
        ga, gf = np.array([g[b] for b in fd_bands]), np.array([fd[b] for b in fd_bands])
        U = t.n_cells * t.n_bands * t.n_dirs
        out = {"metric": "adjoint dk_eff/d(tau scale) and dk_eff/dT_hot, time and agreement with finite differences",
               "value": t_adj, "unit": "s", "higher_is_better": False, "n_gpus": 1, "dtype": "f64",
               "data": "synthetic (seeded porous Delaunay deck, seeded Debye-like bands)",
               "config": {"workload": "2D porous film 500 nm, r=150 nm diffuse pore, seeded Delaunay, FP64 forward "
                                      "and adjoint", "cells": t.n_cells, "bands": t.n_bands, "directions": t.n_dirs,
                          "updates_per_iteration": U, "parallelism": "single GPU"},
               "forward": {"iterations": st.iterations, "solve_s": t_solve, "create_s": t_create, "keff": k0},
               "adjoint": {"iterations": ast["iterations"], "converged": bool(ast["converged"]), "time_s": t_adj,
                           "ms_per_iteration": t_adj * 1e3 / max(1, ast["iterations"]),
                           "dk_dtau_scale": g.tolist(), "dk_dT_hot": g_hot},
               "checks": {"fd_bands": fd_bands, "fd": [fd[b] for b in fd_bands],
                          "fd_rel_l2": float(np.linalg.norm(ga - gf) / np.linalg.norm(gf)),
                          "dT_hot_rel_err_vs_k_over_dT": abs(g_hot - k0 / 2.0) / abs(k0 / 2.0),
                          "fd_time_s": time.time() - t_fd}}
        print(json.dumps(out))


def run_precision(p, a, prec, cfg, U, T0_out):
    """Set-up, warm-up, K timed steps and (unless --no-converge) a solve to
    the deck tolerances from the Fourier warm start, on one handle."""
    from giftbte_amd import BTESolver
    t = p.tables
    K = a.steps
    t_c = time.time()
    s = BTESolver(p, precision=prec)
    create_s = time.time() - t_c
    tuning = s.tuning_log()
    fp = s.fourier()
    if not T0_out:
        T0_out.append(s.fields()["T_lattice"])
    st_w = s.solve(a.warmup, 0.0, 0.0)
    print(f"[bench] {prec}: gbte_create {create_s:.2f} s, fourier ({fp} passes), warmup {st_w.iterations} it "
          f"{st_w.solve_ms:.1f} ms", file=sys.stderr, flush=True)
    t1 = time.time()
    st = s.solve(K, 0.0, 0.0)  # returns after the device has finished
    wall = time.time() - t1
    bpu = kernel_bytes_per_update(t.dim, 8 if prec == "fp64" else 4)
    if t.dim == 3:
        ktime = {"source": st.source_ms, "sweep": st.sweep_ms, "grad": st.grad_ms,
                 "moments": st.face_ms - st.grad_ms}
    else:
        ktime = {"source": st.source_ms, "sweep": st.sweep_ms, "face": st.face_ms}
    kernels = {}
    for ph, ms in ktime.items():
        per = ms / K
        gbs = bpu[ph] * U / (per * 1e-3) / 1e9 if per > 0 else 0.0
        kernels[kernel_name(ph, t.dim, prec, tuning)] = {
            "ms_per_launch": per, "bytes_per_update": bpu[ph], "achieved_GBs": gbs,
            "frac_hbm_peak": gbs / HBM_PEAK_GBS}
    dom_ph = max(ktime, key=ktime.get)
    dom = kernel_name(dom_ph, t.dim, prec, tuning)
    r = {"value": U * K / wall, "ms_per_step": wall * 1e3 / K,
         "breakdown_ms_per_step": {"source": st.source_ms / K, "sweep": st.sweep_ms / K,
                                   "face_moments": st.face_ms / K, "gradient": st.grad_ms / K,
                                   "transport": st.transport_ms / K, "macro": st.macro_ms / K,
                                   "macro_passes_per_step": st.macro_passes / K,
                                   "sweep_levels": st.sweep_levels},
         "roofline": {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["achieved_GBs"],
                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kernels[dom]["frac_hbm_peak"],
                      "traffic": load_traffic(dom, cfg, prec, a.workload),
                      "bytes_per_update": bpu[dom_ph], "updates_per_launch": U},
         "kernels": kernels, "gbte_create_s": create_s,
         "tuning": [ln for ln in tuning.splitlines() if ln]}
    if not a.no_converge:
        tc0 = time.time()
        s.fourier()
        stc = s.solve(t.max_iter, t.tol_temp, t.tol_flux)
        r["time_to_converged_s"] = time.time() - tc0
        r["end_to_end_s"] = create_s + r["time_to_converged_s"]
        r["converged"] = bool(stc.converged)
        r["converged_iterations"] = stc.iterations
        r["keff_W_per_mK"] = s.keff(0, 2.0)
        print(f"[bench] {prec}: converged={r['converged']} in {stc.iterations} it, "
              f"{r['time_to_converged_s']:.3f} s, k_eff {r['keff_W_per_mK']:.9g}", file=sys.stderr, flush=True)
    s.close()
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("3d", "2d"), default="3d")
    ap.add_argument("--lattice", type=int, default=36, help="3d: lattice cells per cube edge (36: about 295k tets)")
    ap.add_argument("--cells", type=int, default=50000, help="2d: target triangle count")
    ap.add_argument("--bands", type=int, default=40)
    ap.add_argument("--nphi", type=int, default=16, help="2d: directions / 4")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-converge", action="store_true", help="skip the timed solve to the deck tolerance")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--precisions", default="fp64,fp32", help="storage precisions to run (default both)")
    ap.add_argument("--batch", type=int, default=0, metavar="N",
                    help="instead of the default line: N independent porous instances solved side by side")
    ap.add_argument("--adjoint", action="store_true",
                    help="instead of the default line: the adjoint sensitivities on the 2D porous workload")
    ap.add_argument("--adjoint-fd", type=int, default=2, help="bands checked by finite differences (--adjoint)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device once (calibrates HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build runs the whole problem on a single MI355X "
              "(no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    from giftbte_amd import Problem

    if a.batch > 0:
        run_batch(a)
        return
    if a.adjoint:
        run_adjoint(a)
        return
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
        deck = make_deck(td, a)
        p = Problem(deck)
        t = p.tables
        nc, nB, nD = t.n_cells, t.n_bands, t.n_dirs
        print(f"[bench] deck {nc} cells x {nB} bands x {nD} directions ({time.time() - t_gen:.1f} s set-up)",
              file=sys.stderr, flush=True)
        U = nc * nB * nD
        K = a.steps
        if a.workload == "3d":
            wl = (f"3D film: 200 nm cube, centred 80 nm cubic void, seeded Delaunay tetrahedra, {nB} bands, "
                  f"{nD} dirs (4 polar x 24 azimuthal)")
        else:
            wl = f"2D porous film 500 nm, r=150 nm diffuse pore, seeded Delaunay, {nB} bands, {nD} dirs"
        cfg = {"workload": wl, "cells": nc, "bands": nB, "directions": nD, "systems": nB * nD,
               "updates_per_step": U, "parallelism": "single GPU"}
        runs = {}
        T0 = []
        for prec in a.precisions.split(","):
            runs[prec] = run_precision(p, a, prec, cfg, U, T0)
        mixed_head = a.workload == "3d" and "fp32" in runs
        head = runs["fp32"] if mixed_head else runs[next(iter(runs))]
        out = {
            "metric": "time-to-converged k_eff (1e-6 rel) & cell\u00b7band\u00b7dir updates/s, mixed vs FP64",
            "value": head["value"], "unit": "cell*band*dir updates/s", "n_gpus": 1, "steps": K, "warmup": a.warmup,
            "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32" if mixed_head else "f64",
            "dtype_detail": ("f32 storage of e, gradients and sources, f32 gradient arithmetic; f64 sweep "
                             "arithmetic, recovery, synthetic solve and reductions") if mixed_head else "f64 throughout",
            "data": f"synthetic (seeded {'tetrahedral cube-with-void' if a.workload == '3d' else 'porous Delaunay'} "
                    f"deck, seeded Debye-like bands)",
            "config": cfg,
        }
        out.update({k: v for k, v in head.items() if k not in ("value", "ms_per_step")})
        if "fp64" in runs and "fp32" in runs:
            f64, f32 = runs["fp64"], runs["fp32"]
            other_key, other = ("fp64", f64) if mixed_head else ("mixed", f32)
            o = {"value": other["value"], "ms_per_step": other["ms_per_step"]}
            o.update({k: v for k, v in other.items() if k not in ("value", "ms_per_step")})
            out[other_key] = o
            out["mixed_vs_fp64"] = {"speedup_updates": f32["value"] / f64["value"]}
            if "keff_W_per_mK" in f64:
                out["mixed_vs_fp64"].update({
                    "keff_rel_err": abs(f32["keff_W_per_mK"] - f64["keff_W_per_mK"]) / abs(f64["keff_W_per_mK"]),
                    "speedup_time_to_converged": f64["time_to_converged_s"] / f32["time_to_converged_s"],
                    "speedup_end_to_end": f64["end_to_end_s"] / f32["end_to_end_s"]})
        if not a.no_cpu:
            port = cpu_baseline(p, T0[0])
            ref = (reference_baseline(td, a.cells, a.bands, a.seed) if a.workload == "2d"
                   else reference_baseline_3d(td, a.bands, a.seed))
            out["cpu_baseline"] = ref if ref is not None else port
            if ref is not None:
                out["cpu_port"] = port
        print(json.dumps(out))


if __name__ == "__main__":
    main()
