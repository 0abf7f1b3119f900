"""gbte_adjoint: dk_eff / d(tau scale per band) and dk_eff / dT_hot by the
discrete adjoint of the converged outer iteration, against central finite
differences (step 1e-4 relative) of the GPU forward solve and of the CPU
oracle, against the closed form dk/dT_hot = k / (T_hot - T_cold) (q is linear
in the wall temperatures and vanishes when they are equal), and through the
torch.autograd.Function.  Bar: ||g_adj - g_fd|| / ||g_fd|| <= 1e-5."""
import dataclasses
import os

import numpy as np
import pytest

import _oracle

pytestmark = pytest.mark.gpu

ADJ_TOL = 1e-5
H = 1e-4


def _spec(n_target=2000, nbands=4, nphi=4, **kw):
    from giftbte_amd import decks
    return decks.porous_spec(n_target=n_target, nbands=nbands, nphi=nphi, seed=3, **kw)


def _gpu_keff(spec, root, name, tol=(1e-15, 1e-13)):
    from giftbte_amd import decks, Problem, BTESolver
    p = Problem(decks.write_deck(os.path.join(root, name), spec))
    s = BTESolver(p)
    s.fourier()
    st = s.solve(100000, *tol)
    assert st.converged
    k = s.keff(0, 2.0)
    return p, s, k


def _fd(keff_of, spec, nb, h=H):
    from giftbte_amd import decks
    g = np.zeros(nb)
    for b in range(nb):
        sp, sm = np.ones(nb), np.ones(nb)
        sp[b], sm[b] = 1 + h, 1 - h
        kp = keff_of(dataclasses.replace(spec, bands=decks.scale_tau(spec.bands, sp)), f"p{b}")
        km = keff_of(dataclasses.replace(spec, bands=decks.scale_tau(spec.bands, sm)), f"m{b}")
        g[b] = (kp - km) / (2 * h)
    return g


@pytest.mark.parametrize("order,method", [("2", "Synthetic"), ("1", "Synthetic"), ("1", "Sequential")])
def test_adjoint_matches_gpu_finite_differences(order, method, tmp_path):
    spec = _spec(order=order, method=method)
    p, s, k0 = _gpu_keff(spec, str(tmp_path), "base")
    g, gtb, ast = s.adjoint(0, 2.0, 100000, 1e-13)
    s.close()
    assert ast["converged"] == 1

    def keff_of(sp, name):
        _, s2, k = _gpu_keff(sp, str(tmp_path), name)
        s2.close()
        return k

    fd = _fd(keff_of, spec, 4)
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    print(f"\n[adjoint] order {order} {method}: {ast['iterations']} adjoint iterations; adjoint {g} fd {fd} "
          f"rel L2 {err:.2e}", flush=True)
    assert err <= ADJ_TOL
    # the hot wall: closed form k / (T_hot - T_cold) and the finite difference
    bt, btemp = p.array("bound_type"), p.array("bound_temp")
    g_hot = gtb[(bt == 1) & (btemp == 301.0)].sum()
    assert np.all(gtb[bt != 1] == 0)
    assert abs(g_hot - k0 / 2.0) <= 1e-8 * abs(k0)


def test_adjoint_matches_oracle_finite_differences(tmp_path):
    """The CPU oracle (the reference's algorithm) differentiated by central
    differences on a small deck (846 cells, 3 bands, 8 directions), against
    the GPU adjoint.  Both stop at margins of 1e-14 / 1e-12 (tighter ones
    reach rounding level, where ten rises in a row count as divergence), so
    k is converged to about 1e-10 relative and the oracle's differences take
    a step of 1e-3 (truncation error about 1e-6 relative)."""
    from giftbte_amd import decks, Problem
    spec = _spec(n_target=1000, nbands=3, nphi=2)
    tol = (1e-14, 1e-12)
    p, s, k0 = _gpu_keff(spec, str(tmp_path), "base", tol)
    g, _, ast = s.adjoint(0, 2.0, 100000, 1e-13)
    s.close()

    def keff_of(sp, name):
        q = Problem(decks.write_deck(os.path.join(str(tmp_path), "o" + name), sp))
        ref, info = _oracle.run(q.tables, 100000, *tol)
        assert info["status"] == 0
        vol = q.array("volume")
        return float((ref["qx"] * vol).sum() / (np.ptp(q.array("node_y")) * 2.0))

    k_or = keff_of(spec, "base")
    fd = _fd(keff_of, spec, 3, h=1e-3)
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    print(f"\n[adjoint] oracle k {k_or:.12g} gpu k {k0:.12g}; adjoint {g} oracle fd {fd} rel L2 {err:.2e}", flush=True)
    assert abs(k_or - k0) <= 1e-9 * abs(k0)
    assert err <= ADJ_TOL


def test_autograd_function(tmp_path):
    """KeffFunction: backward gives the adjoint gradient; a central
    difference of the function itself along its largest component agrees."""
    import torch
    from giftbte_amd import autograd, decks
    spec = _spec(n_target=600, nbands=3, nphi=2)
    s = torch.tensor([1.0, 1.2, 0.9], dtype=torch.float64, requires_grad=True)
    th = torch.tensor(301.0, dtype=torch.float64, requires_grad=True)
    k = autograd.keff(s, th, spec, decks.LEFT + 1)
    k.backward()
    b = int(torch.argmax(torch.abs(s.grad)))  # the band the difference resolves best
    sp, sm = s.detach().clone(), s.detach().clone()
    sp[b] *= 1 + H
    sm[b] *= 1 - H
    kp = autograd.keff(sp, th.detach(), spec, decks.LEFT + 1)
    km = autograd.keff(sm, th.detach(), spec, decks.LEFT + 1)
    fd = float(kp - km) / (2 * H * float(s[b].detach()))
    print(f"\n[adjoint] autograd k {float(k):.12g} grad {s.grad.numpy()} dT {float(th.grad):.10g}; "
          f"fd band {b} {fd:.10g}", flush=True)
    assert abs(float(s.grad[b]) - fd) <= 1e-5 * abs(fd)
    assert abs(float(th.grad) - float(k) / 2.0) <= 1e-8 * abs(float(k))


def test_adjoint_refuses_what_it_does_not_cover(tmp_path):
    from giftbte_amd import decks, Problem, BTESolver, GBTEError, _lib
    p = Problem(decks.write_deck(str(tmp_path / "d"), _spec(n_target=300, nbands=2, nphi=2)))
    s = BTESolver(p)
    with pytest.raises(GBTEError) as e:  # no solve yet
        s.adjoint()
    assert e.value.code == _lib.GBTE_BAD_INPUT
    s.close()
    s = BTESolver(p, precision="fp32")
    s.fourier()
    s.solve(3, 0.0, 0.0)
    with pytest.raises(GBTEError) as e:
        s.adjoint()
    assert e.value.code == _lib.GBTE_UNSUPPORTED
    s.close()
