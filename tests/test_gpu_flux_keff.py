"""gbte_flux_keff: k_eff = sum q V / (A dT) reduced on the device from
device arrays, against a float64 numpy sum and against gbte_keff on a solved
deck (the cross-plane check of bin/examples/cross-plane/
verify_kappa_crossplane.m:8 in its sum q V / (A dT) form)."""
import numpy as np
import pytest

from _hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def _keff_dev(q, v, area, dT, n=None):
    from giftbte_amd import _lib
    lib = _lib.load()
    qd, vd, out = DeviceArray(q), DeviceArray(v), DeviceArray(n=1)
    rc = lib.gbte_flux_keff(qd.ptr, vd.ptr, qd.n if n is None else n, area, dT, out.ptr, None)
    k = out.get()[0]
    for a in (qd, vd, out):
        a.free()
    return rc, k


@pytest.mark.parametrize("n", [1, 255, 4096, 4097, 1_000_003])
def test_flux_keff_matches_numpy(n):
    from giftbte_amd import _lib
    rng = np.random.default_rng(n)
    q = rng.standard_normal(n) * 1e7 + 3e6
    v = rng.uniform(0.5, 2.0, n) * 1e-24
    print(f"\n[flux keff] n={n}", flush=True)
    rc, k = _keff_dev(q, v, 4e-14, 2.0)
    assert rc == _lib.GBTE_OK, _lib.load().gbte_problem_error().decode()
    ref = float(np.sum(q * v)) / (4e-14 * 2.0)
    scale = float(np.sum(np.abs(q * v))) / (4e-14 * 2.0)
    assert abs(k - ref) <= 1e-13 * scale


def test_flux_keff_rejects_bad_sizes():
    from giftbte_amd import _lib
    lib = _lib.load()
    z = np.zeros(8)
    rc, _ = _keff_dev(z, z, 1.0, 1.0, n=0)
    assert rc == _lib.GBTE_BAD_INPUT
    rc, _ = _keff_dev(z, z, 1.0, 1.0, n=16_777_217)  # rejected on the host before any launch
    assert rc == _lib.GBTE_BAD_INPUT and b"limit" in lib.gbte_problem_error()
    rc, _ = _keff_dev(z, z, 1.0, 0.0)
    assert rc == _lib.GBTE_BAD_INPUT


def test_flux_keff_matches_handle_keff(deck_root):
    """The porous test deck solved for 6 iterations: gbte_keff (host sum in
    cell order) and gbte_flux_keff on the same q and volumes."""
    import decklib
    from giftbte_amd import Problem, BTESolver, _lib
    d = decklib.make("porous_b4", deck_root)
    p = Problem(d)
    s = BTESolver(p)
    s.fourier()
    s.solve(6, 0.0, 0.0)
    k_h = s.keff(0, 2.0)
    q = s.fields()["qx"]
    s.close()
    rc, k_d = _keff_dev(q, p.array("volume"), float(np.ptp(p.array("node_y"))), 2.0)
    print(f"\n[flux keff] handle {k_h:.15g} device {k_d:.15g}", flush=True)
    assert rc == _lib.GBTE_OK
    assert abs(k_d - k_h) <= 1e-12 * abs(k_h)
