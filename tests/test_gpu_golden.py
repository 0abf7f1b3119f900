"""HIP path vs the compiled reference's own full-precision fixtures
(tests/golden/*.npz, written by tests/golden/make_golden.py from the reference
at np=1), directly: no oracle in the chain.

Bar: temperature, energy temperature, band temperatures and heat flux within
1e-9 relative L2 of the reference, the k_eff numerator sum q V along the heat-flow axis within 1e-8 of sum |q| V,
and the reference's outer-iteration, synthetic-pass and Fourier-pass counts.
The measured errors are printed per deck."""
import os
import sys

import numpy as np
import pytest

import _oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
import make_golden  # noqa: E402

FIELD_TOL = 1e-9
KEFF_TOL = 1e-8


@pytest.mark.parametrize("name", list(make_golden.golden_decks()))
def test_gpu_matches_reference_fixture(name, tmp_path):
    from giftbte_amd import Problem, BTESolver, decks
    factory, fmt, *_ = make_golden.golden_decks()[name]
    g = dict(np.load(os.path.join(GOLD, name + ".npz")))
    max_iter, tolT, tolQ = g["run"]
    print(f"\n[gpu golden] deck={name} run={tuple(g['run'])}", flush=True)
    d = decks.write_deck(str(tmp_path / name), factory(), mesh_format=fmt)
    p = Problem(d)
    sol = BTESolver(p)
    fp = sol.fourier()
    st = sol.solve(int(max_iter), float(tolT), float(tolQ))
    f = sol.fields()
    errs = {k: _oracle.rel_l2(f[k], g[k].reshape(f[k].shape)) for k in ("T_lattice", "T_energy", "T_band")}
    errs["q"] = _oracle.rel_l2(np.concatenate([f["qx"], f["qy"]]), np.concatenate([g["qx"], g["qy"]]))
    # the k_eff numerator along the axis the heat flows (x, or y for the
    # cross-plane film)
    vol = p.array("volume")
    ax = max(("qx", "qy"), key=lambda k: abs(float((g[k] * vol).sum())))
    kn, kr = float((f[ax] * vol).sum()), float((g[ax] * vol).sum())
    # relative to sum |q| V: a deck without a net flow (the heated film with
    # cold walls all round) has a numerator that is rounding noise
    errs["keff"] = abs(kn - kr) / float((np.abs(g[ax]) * vol).sum())
    print("  counts gpu (fourier, iterations, macro) =", (fp, st.iterations, st.macro_passes),
          "reference =", (int(g["count_fourier_passes"]), int(g["count_iterations"]), int(g["count_macro_passes"])))
    print("  rel errors vs reference: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert fp == int(g["count_fourier_passes"])
    assert st.iterations == int(g["count_iterations"])
    assert st.macro_passes == int(g["count_macro_passes"])
    for k in ("T_lattice", "T_energy", "T_band", "q"):
        assert errs[k] < FIELD_TOL, (k, errs[k])
    assert errs["keff"] < KEFF_TOL
