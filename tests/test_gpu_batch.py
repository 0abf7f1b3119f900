"""Batched solves (gbte_solve_batch): independent porous instances of uneven
size solved side by side match the same instances solved one at a time.

Each handle keeps its own stream, graphs and buffers and the kernels are
deterministic.  A batched solve forms each source in one pass; a single solve
forms the C T_old term in the sweep instead, a rounding-level difference in
summation order.  The bar is 1e-10 relative (max abs difference over max abs
value), ten times tighter than the oracle parity bar, and the counts must
match.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (cells, bands, nphi, pore radius fraction, seed, max_iter, tolT, tolQ)
INSTANCES = [(2000, 4, 4, 0.2, 11, 12, 0.0, 0.0), (6000, 6, 4, 0.35, 12, 12, 0.0, 0.0),
             (1000, 4, 2, 0.1, 13, 200, 1e-6, 1e-4), (12000, 8, 4, 0.3, 14, 8, 0.0, 0.0),
             (3000, 5, 3, 0.25, 15, 10, 0.0, 0.0)]


def _problems(root):
    from giftbte_amd import Problem, decks
    out = []
    for i, (n, nb, nphi, r, seed, *_run) in enumerate(INSTANCES):
        spec = decks.porous_spec(n_target=n, nbands=nb, nphi=nphi, pore_radius=r, seed=seed, scale=5e-7)
        out.append(Problem(decks.write_deck(str(root / f"inst{i}"), spec)))
    return out


def test_batch_matches_one_at_a_time(tmp_path):
    from giftbte_amd import BTESolver, solve_batch
    probs = _problems(tmp_path)
    runs = [inst[5:] for inst in INSTANCES]
    print(f"\n[gpu batch] {len(probs)} instances, cells {[p.tables.n_cells for p in probs]}", flush=True)
    t0 = time.time()
    alone = []
    for p, (it, tT, tQ) in zip(probs, runs):
        s = BTESolver(p)
        s.fourier()
        st = s.solve(it, tT, tQ)
        alone.append((st, s.fields()))
        s.close()
    t_alone = time.time() - t0
    sols = [BTESolver(p) for p in probs]
    for s in sols:
        s.fourier()
    t1 = time.time()
    sts = solve_batch(sols, [r[0] for r in runs], [r[1] for r in runs], [r[2] for r in runs])
    t_batch = time.time() - t1
    print(f"[gpu batch] one at a time {t_alone:.2f} s (with set-up), batch solve {t_batch:.2f} s", flush=True)
    worst = 0.0
    for i, (s, st) in enumerate(zip(sols, sts)):
        st0, f0 = alone[i]
        assert (st.iterations, st.macro_passes) == (st0.iterations, st0.macro_passes), i
        f = s.fields()
        for k, v in f.items():
            scale = max(np.abs(f0[k]).max(), 1e-300)
            worst = max(worst, np.abs(v - f0[k]).max() / scale)
            assert np.abs(v - f0[k]).max() <= 1e-10 * scale, (i, k)
        s.close()
    print(f"[gpu batch] largest relative difference to one at a time: {worst:.1e}", flush=True)
