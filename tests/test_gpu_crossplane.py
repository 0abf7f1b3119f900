"""Known answer: cross-plane thermal conductivity of a silicon film.

The reference validates its solver on bin/examples/cross-plane (100 x 100
quads, Si 15-band DATABASE file, 3D Gauss directions Ntheta 4 / Nphi 4,
bottom 301 K / top 300 K isothermal, specular sides) by comparing
k = <q_y> L / dT at L = 10 nm, 100 nm and 1 um with the analytical model in
bin/examples/cross-plane/analytic_201806/TC.dat (verify_kappa_crossplane.m).
tests/golden/crossplane_TC.dat and Si_15.dat are those two data files; the
deck is rebuilt by the in-repo generator (same mesh, bands, angles and
boundaries).  The bar is the model's: the discrete-ordinate angular error
grows in the ballistic limit, so 3 % at 10 nm and 1 % at 100 nm and 1 um.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [(1e-8, 0.03), (1e-7, 0.01), (1e-6, 0.01)]


def _analytic(L):
    tc = np.loadtxt(os.path.join(GOLDEN, "crossplane_TC.dat"))
    order = np.argsort(tc[:, 0])
    return float(np.interp(np.log(L), np.log(tc[order, 0]), tc[order, 1]))


@pytest.mark.parametrize("length,tol", CASES)
def test_crossplane_keff_matches_analytic(length, tol, deck_root):
    from giftbte_amd import BTESolver, Problem, decks
    spec = decks.crossplane_spec(os.path.join(GOLDEN, "Si_15.dat"), length=length, n=100)
    d = decks.write_deck(os.path.join(str(deck_root), f"crossplane_{length:g}"), spec)
    print(f"\n[gpu crossplane] L={length:g} m", flush=True)
    p = Problem(d)
    t = p.tables
    s = BTESolver(p)
    s.fourier()
    st = s.solve(t.max_iter, t.tol_temp, t.tol_flux)
    assert st.converged
    k = s.keff(1, 1.0)  # sum q_y V / (L_x dT) = <q_y> L / dT on the L x L square
    ka = _analytic(length)
    print(f"[gpu crossplane] iterations {st.iterations}, k = {k:.4f} W/mK, model {ka:.4f}", flush=True)
    assert abs(k - ka) / ka < tol
