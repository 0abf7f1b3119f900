"""The batched workload at full size: 64 independent 2D porous instances of
uneven size (500 nm square, pore radius uniform in [50, 200] nm, 2k-40k
seeded Delaunay triangles, log-uniform), 20 bands of the Debye-like
dispersion and 32 directions, each solved from its Fourier warm start to its
deck tolerances side by side (gbte_solve_batch).

Checks:
- the batch equals the same instances solved one at a time: identical
  outer-iteration and synthetic-pass counts, fields within 1e-10 relative
  (max abs difference over max abs value; a batched solve forms each source in
  one pass, a single solve adds the C T_old term in the sweep, a rounding-level
  difference in summation order);
- FP32 storage against FP64, per instance: k_eff within 1e-6 relative, the
  outer-iteration count within one;
- the two smallest instances against the CPU oracle over five outer iterations
  from the Fourier warm start: fields within 1e-9 relative L2, the same
  synthetic-pass and Fourier-pass counts.
Measured figures are printed."""
import os
import time

import numpy as np
import pytest

import _oracle

pytestmark = pytest.mark.gpu

N_INST = 64


def _instances(root):
    from giftbte_amd import Problem, decks
    rng = np.random.default_rng(1)
    probs = []
    for i in range(N_INST):
        n = int(np.exp(rng.uniform(np.log(2000), np.log(40000))))
        r = rng.uniform(50e-9, 200e-9) / 500e-9
        spec = decks.porous_spec(n_target=n, nbands=20, nphi=8, pore_radius=r, seed=101 + i, scale=5e-7)
        probs.append(Problem(decks.write_deck(os.path.join(str(root), f"inst{i}"), spec)))
    return probs


def _keff(p, f, dT=2.0):
    """sum q_x V / (L_y dT), L_y the extent of the nodes across the heat flow."""
    return float((f["qx"] * p.array("volume")).sum() / (np.ptp(p.array("node_y")) * dT))


def _batch(probs, precision):
    from giftbte_amd import BTESolver, solve_batch
    sols = [BTESolver(p, precision=precision) for p in probs]
    for s in sols:
        s.fourier()
    t = [p.tables for p in probs]
    t0 = time.time()
    sts = solve_batch(sols, [x.max_iter for x in t], [x.tol_temp for x in t], [x.tol_flux for x in t])
    wall = time.time() - t0
    upd = sum(st.iterations * x.n_cells * x.n_bands * x.n_dirs for st, x in zip(sts, t))
    return sols, sts, wall, upd / wall


def test_batch64_matches_solo_and_fp32(tmp_path):
    from giftbte_amd import BTESolver
    t0 = time.time()
    probs = _instances(tmp_path)
    cells = [p.tables.n_cells for p in probs]
    print(f"\n[gpu batch64] {N_INST} instances, {sum(cells)} cells ({min(cells)}..{max(cells)}), "
          f"decks {time.time() - t0:.1f} s", flush=True)
    sols, sts, wall, rate = _batch(probs, "fp64")
    assert all(st.converged for st in sts)
    print(f"[gpu batch64] fp64 batch {wall:.2f} s, {rate:.3e} updates/s, iterations "
          f"{min(st.iterations for st in sts)}..{max(st.iterations for st in sts)}", flush=True)
    f64 = [s.fields() for s in sols]
    k64 = [_keff(p, f) for p, f in zip(probs, f64)]
    # the same handles, one at a time
    worst = 0.0
    t1 = time.time()
    for i, (p, s, st) in enumerate(zip(probs, sols, sts)):
        t = p.tables
        s.fourier()
        st1 = s.solve(t.max_iter, t.tol_temp, t.tol_flux)
        assert (st1.iterations, st1.macro_passes) == (st.iterations, st.macro_passes), i
        f = s.fields()
        for k, v in f.items():
            scale = max(np.abs(f64[i][k]).max(), 1e-300)
            d = np.abs(v - f64[i][k]).max() / scale
            worst = max(worst, d)
            assert d <= 1e-10, (i, k, d)
        s.close()
    print(f"[gpu batch64] one at a time {time.time() - t1:.2f} s; largest relative difference to the batch "
          f"{worst:.1e}", flush=True)
    # FP32 storage, batched
    sols32, sts32, wall32, rate32 = _batch(probs, "fp32")
    assert all(st.converged for st in sts32)
    errs = []
    for i, (p, s, st, st64) in enumerate(zip(probs, sols32, sts32, sts)):
        k = _keff(p, s.fields())
        errs.append(abs(k - k64[i]) / abs(k64[i]))
        assert abs(st.iterations - st64.iterations) <= 1, (i, st.iterations, st64.iterations)
        s.close()
    print(f"[gpu batch64] fp32 batch {wall32:.2f} s, {rate32:.3e} updates/s; k_eff vs fp64: largest rel. error "
          f"{max(errs):.1e}", flush=True)
    assert max(errs) <= 1e-6


def test_batch64_smallest_match_oracle(tmp_path):
    from giftbte_amd import BTESolver, solve_batch
    probs = _instances(tmp_path)
    order = np.argsort([p.tables.n_cells for p in probs])[:2]
    small = [probs[i] for i in order]
    sols = [BTESolver(p) for p in small]
    fps = [s.fourier() for s in sols]
    sts = solve_batch(sols, [5, 5], [0.0, 0.0], [0.0, 0.0])
    for p, s, fp, st in zip(small, sols, fps, sts):
        ref, info = _oracle.run(p.tables, 5, 0.0, 0.0)
        f = s.fields()
        errs = {k: _oracle.rel_l2(f[k], ref[k]) for k in ("T_lattice", "T_energy", "T_band")}
        errs["q"] = _oracle.rel_l2(np.concatenate([f["qx"], f["qy"]]), np.concatenate([ref["qx"], ref["qy"]]))
        print(f"\n[gpu batch64] {p.tables.n_cells} cells vs oracle: " +
              ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
        assert fp == info["fourier_passes"]
        assert (st.iterations, st.macro_passes) == (info["iterations"], info["macro_passes"])
        for k, v in errs.items():
            assert v < 1e-9, k
        s.close()
