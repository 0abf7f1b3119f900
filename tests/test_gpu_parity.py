"""HIP path vs the CPU oracle on the same decks (FP64).

Bar (taken from the north-star contract): temperature and heat-flux fields
within 1e-9 relative L2 of the oracle, k_eff within 1e-8, and identical
outer-iteration, synthetic-pass and Fourier-pass counts.
"""
import numpy as np
import pytest

import _oracle
import decklib

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-9
KEFF_TOL = 1e-8


def _gpu_run(deck_dir, s, variant="direct"):
    from giftbte_amd import Problem, BTESolver
    p = Problem(deck_dir)
    if variant == "fallback":
        sol = BTESolver(p, fallback_kernels=True)
    else:
        sol = BTESolver(p, linear_solver=variant)
    fp = sol.fourier()
    st = sol.solve(s["max_iter"], s["tol_temp"], s["tol_flux"])
    return p, sol, fp, st


# default kernels on every deck; PCG linear solver and the fallback kernels
# (level-synchronous sweep, untiled face kernels) on two decks
CASES = ([(n, "direct") for n in decklib.DECKS] + [("porous_b4", "pcg")] +
         [("porous_b4", "fallback"), ("square_b20_diffuse", "fallback"), ("porous_b4_o1", "fallback"),
          ("porous_b4_2limit", "fallback")])


@pytest.mark.parametrize("name,variant", CASES)
def test_fields_match_oracle(name, variant, deck_root):
    d = decklib.make(name, deck_root)
    s = decklib.settings(name)
    print(f"\n[gpu parity] deck={name} variant={variant} settings={s}", flush=True)
    p, sol, fp, st = _gpu_run(d, s, variant)
    ref, info = _oracle.run(p.tables, s["max_iter"], s["tol_temp"], s["tol_flux"])
    assert st.fourier_passes == info["fourier_passes"]
    assert st.iterations == info["iterations"]
    assert st.macro_passes == info["macro_passes"]
    f = sol.fields()
    errs = {k: _oracle.rel_l2(f[k], ref[k]) for k in ("T_lattice", "T_energy", "T_band")}
    q = np.concatenate([f["qx"], f["qy"], f["qz"]])
    qr = np.concatenate([ref["qx"], ref["qy"], ref["qz"]])
    errs["q"] = _oracle.rel_l2(q, qr)
    print("  rel errors vs oracle: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < FIELD_TOL, k
    # k_eff along x with the deck's 2 K drop where it has one
    t = p.tables
    bt = p.array("bound_type")
    if (bt == 1).sum() and np.ptp(p.array("bound_temp")[bt == 1]) > 0:
        dT = 2.0
        k_gpu = sol.keff(0, dT)
        vol = p.array("volume")
        area = np.ptp(p.array("node_y"))
        if t.dim == 3:
            area *= np.ptp(p.array("node_z"))
        k_ref = float((ref["qx"] * vol).sum() / (area * dT))
        print(f"  k_eff gpu {k_gpu:.12g} oracle {k_ref:.12g} rel {abs(k_gpu - k_ref) / abs(k_ref):.2e}", flush=True)
        assert abs(k_gpu - k_ref) <= KEFF_TOL * abs(k_ref)
    del t


def test_fourier_matches_oracle(deck_root):
    d = decklib.make("porous_b4", deck_root)
    from giftbte_amd import Problem, BTESolver
    p = Problem(d)
    sol = BTESolver(p)
    fp = sol.fourier()
    pr, ref = _oracle.fourier(p.tables)
    assert fp == pr
    f = sol.fields()
    assert _oracle.rel_l2(f["T_lattice"], ref["T"]) < 1e-12
    assert _oracle.rel_l2(np.concatenate([f["qx"], f["qy"]]), np.concatenate([ref["qx"], ref["qy"]])) < 1e-9
