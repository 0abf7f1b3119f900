"""The tetrahedral (3D) path on the GPU against the CPU oracle.

The oracle's 3D branches are pinned on the reference's own 3D solver
branches (tests/golden/cube3d_*.npz, ref_dump --tables).  Bar, as for 2D:
FP64 fields within 1e-9 relative L2 of the oracle, k_eff within 1e-8, the
oracle's iteration / synthetic-pass / Fourier-pass counts; FP32 storage
within 1e-6 (fields and k_eff) with the same counts.  Measured errors are
printed."""
import numpy as np
import pytest

import _oracle
import decklib

pytestmark = pytest.mark.gpu

DECKS3D = ["cube3d_b8", "cube3d_gray_seq"]


def _keff_x(p, q, dT=2.0):
    """sum q_x V / (A_x dT), A_x the y-z extent of the nodes."""
    vol = p.array("volume")
    ys, zs = p.array("node_y"), p.array("node_z")
    return float((q * vol).sum() / ((ys.max() - ys.min()) * (zs.max() - zs.min()) * dT))


# every sweep kernel family the set-up tuner chooses from, forced, on the
# 8-band deck (the level-synchronous one-workgroup-per-direction kernel is
# FP64 only); on the 20-band deck the FP32 lane sweep can only run chunks of
# 16 and 4 bands (lanes per position set per chunk)
SWEEPS = [("cube3d_b8", "fp64", "wavespec", "auto"), ("cube3d_b8", "fp64", "level", "auto"),
          ("cube3d_b8", "fp64", "chunked", "auto"), ("cube3d_b8", "fp32", "wavespec", "auto"),
          ("cube3d_b8", "fp32", "chunked", "auto"), ("cube3d_b8", "fp64", "lanes", "auto"),
          ("cube3d_b8", "fp32", "lanes", "auto"), ("cube3d_b20", "fp32", "lanes", "auto"),
          ("cube3d_b20", "fp64", "auto", "auto")]
# each gradient kernel, forced, on both decks and precisions
GRADS = [(n, pr, "auto", g) for n in DECKS3D for pr in ("fp64", "fp32") for g in ("gather", "matrix", "vertex")]


# each layout of the sweep source, forced, with the chunked sweep
LAYOUTS = [("cube3d_b8", pr, "chunked", "auto", lay) for pr in ("fp64", "fp32") for lay in ("chunk", "position", "cell")]


@pytest.mark.parametrize("name,precision,sweep,grad,layout",
                         [(n, pr, "auto", "auto", "auto") for n in DECKS3D for pr in ("fp64", "fp32")]
                         + [c + ("auto",) for c in SWEEPS + GRADS] + LAYOUTS)
def test_3d_matches_oracle(name, precision, sweep, grad, layout, deck_root):
    from giftbte_amd import Problem, BTESolver
    d = decklib.make(name, deck_root)
    s = decklib.settings(name)
    print(f"\n[gpu 3d] deck={name} precision={precision} sweep={sweep} gradient={grad} S layout={layout} "
          f"settings={s}", flush=True)
    p = Problem(d)
    sol = BTESolver(p, precision=precision, sweep_kernel=sweep, gradient_kernel=grad, source_layout=layout)
    log = sol.tuning_log()
    if layout != "auto":
        names = {"chunk": "per band chunk", "position": "per position", "cell": "per cell"}
        assert "S layout chosen: " + names[layout] in log
    print("  " + log.replace("\n", "; "), flush=True)
    if (name, precision, sweep) == ("cube3d_b20", "fp32", "lanes"):
        assert "sweep chosen: chunked up to 16 bands" in log  # the only lane shape 20 FP32 bands admit
    fp = sol.fourier()
    st = sol.solve(s["max_iter"], s["tol_temp"], s["tol_flux"])
    ref, info = _oracle.run(p.tables, s["max_iter"], s["tol_temp"], s["tol_flux"])
    f = sol.fields()
    errs = {k: _oracle.rel_l2(f[k], ref[k]) for k in ("T_lattice", "T_energy", "T_band")}
    errs["q"] = _oracle.rel_l2(np.concatenate([f["qx"], f["qy"], f["qz"]]),
                               np.concatenate([ref["qx"], ref["qy"], ref["qz"]]))
    kg, kr = sol.keff(0, 2.0), _keff_x(p, ref["qx"])
    errs["keff"] = abs(kg - kr) / abs(kr)
    print(f"  counts gpu {(fp, st.iterations, st.macro_passes)} oracle "
          f"{(info['fourier_passes'], info['iterations'], info['macro_passes'])}; k_eff {kg:.10g}")
    print("  rel errors vs oracle: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    tol, ktol = (1e-9, 1e-8) if precision == "fp64" else (1e-6, 1e-6)
    assert fp == info["fourier_passes"]
    assert st.iterations == info["iterations"]
    assert st.macro_passes == info["macro_passes"]
    for k in ("T_lattice", "T_energy", "T_band", "q"):
        assert errs[k] < tol, k
    assert errs["keff"] < ktol


def test_3d_fallback_sweep_matches(deck_root):
    """The level-synchronous sweep on the same 3D deck gives the same fields."""
    from giftbte_amd import Problem, BTESolver
    d = decklib.make("cube3d_b8", deck_root)
    p = Problem(d)
    a = BTESolver(p)
    b = BTESolver(p, fallback_kernels=True)
    for sol in (a, b):
        sol.fourier()
        sol.solve(4, 0.0, 0.0)
    fa, fb = a.fields(), b.fields()
    e = _oracle.rel_l2(fb["qx"], fa["qx"])
    print(f"\n[gpu 3d] fallback sweep vs wave-specialised: q_x rel {e:.2e}", flush=True)
    assert e < 1e-12
