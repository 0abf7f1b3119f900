"""gbte_flux_keff: k_eff = sum q V / (A dT) reduced on the device from
device arrays, against a float64 numpy sum and against gbte_keff on a solved
deck (the cross-plane check of bin/examples/cross-plane/
verify_kappa_crossplane.m:8 in its sum q V / (A dT) form)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _keff_dev(lib, fn, q, v, area, dT):
    import torch
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc = getattr(lib, fn)(C.c_void_p(q.data_ptr()), C.c_void_p(v.data_ptr()), int(q.numel()), area, dT,
                          C.c_void_p(out.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, float(out.cpu()[0])


@pytest.mark.parametrize("n", [1, 255, 4096, 4097, 1_000_003])
def test_flux_keff_matches_numpy(n):
    import torch
    from giftbte_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(n)
    q = rng.standard_normal(n) * 1e7 + 3e6
    v = rng.uniform(0.5, 2.0, n) * 1e-24
    qd = torch.from_numpy(q).to("cuda:0")
    vd = torch.from_numpy(v).to("cuda:0")
    print(f"\n[flux keff] n={n}", flush=True)
    rc, k = _keff_dev(lib, "gbte_flux_keff", qd, vd, 4e-14, 2.0)
    assert rc == _lib.GBTE_OK, lib.gbte_problem_error().decode()
    ref = float(np.sum(q * v)) / (4e-14 * 2.0)
    scale = float(np.sum(np.abs(q * v))) / (4e-14 * 2.0)
    assert abs(k - ref) <= 1e-13 * scale


def test_flux_keff_rejects_bad_sizes():
    import torch
    from giftbte_amd import _lib
    lib = _lib.load()
    qd = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    rc, _ = _keff_dev(lib, "gbte_flux_keff", qd[:0], qd[:0], 1.0, 1.0)
    assert rc == _lib.GBTE_BAD_INPUT
    rc = lib.gbte_flux_keff(C.c_void_p(qd.data_ptr()), C.c_void_p(qd.data_ptr()), 16_777_217, 1.0, 1.0,
                            C.c_void_p(qd.data_ptr()), None)
    assert rc == _lib.GBTE_BAD_INPUT and b"limit" in lib.gbte_problem_error()
    rc, _ = _keff_dev(lib, "gbte_flux_keff", qd, qd, 1.0, 0.0)
    assert rc == _lib.GBTE_BAD_INPUT


def test_flux_keff_matches_handle_keff(deck_root):
    """The porous test deck solved for 6 iterations: gbte_keff (host sum in
    cell order) and gbte_flux_keff on the same q and volumes."""
    import torch
    import decklib
    from giftbte_amd import Problem, BTESolver, _lib
    d = decklib.make("porous_b4", deck_root)
    p = Problem(d)
    s = BTESolver(p)
    s.fourier()
    s.solve(6, 0.0, 0.0)
    k_h = s.keff(0, 2.0)
    q = s.fields()["qx"]
    vol = p.array("volume")
    area = float(np.ptp(p.array("node_y")))
    rc, k_d = _keff_dev(_lib.load(), "gbte_flux_keff", torch.from_numpy(q).to("cuda:0"),
                        torch.from_numpy(vol).to("cuda:0"), area, 2.0)
    s.close()
    print(f"\n[flux keff] handle {k_h:.15g} device {k_d:.15g}", flush=True)
    assert rc == _lib.GBTE_OK
    assert abs(k_d - k_h) <= 1e-12 * abs(k_h)
