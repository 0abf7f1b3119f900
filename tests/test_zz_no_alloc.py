"""gbte_flux_keff_noalloc (partial sums in the library's static device
array) against gbte_flux_keff (partial sums in a buffer it allocates): two
calls in a row, the second on fewer cells, each bit-identical to the
allocating entry point, then a size above the static array's limit."""
import numpy as np
import pytest

from _hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def test_noalloc_matches_allocating_entry_point():
    from giftbte_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(7)
    for n in (3_000_001, 70_000):  # the second call reuses the array with fewer partial sums
        qh, vh = rng.standard_normal(n) * 1e7, rng.uniform(0.5, 2.0, n) * 1e-24
        q, v, outs = DeviceArray(qh), DeviceArray(vh), DeviceArray(n=2)
        print(f"\n[no alloc] n={n}", flush=True)
        rc = lib.gbte_flux_keff(q.ptr, v.ptr, n, 3e-14, 2.0, outs.at(0), None)
        assert rc == _lib.GBTE_OK, lib.gbte_problem_error().decode()
        rc = lib.gbte_flux_keff_noalloc(q.ptr, v.ptr, n, 3e-14, 2.0, outs.at(1), None)
        assert rc == _lib.GBTE_OK, lib.gbte_problem_error().decode()
        a, b = outs.get()
        qv = qh * vh
        ref, scale = float(np.sum(qv)) / (3e-14 * 2.0), float(np.sum(np.abs(qv))) / (3e-14 * 2.0)
        print(f"[no alloc] allocating {a:.17g} static {b:.17g} numpy {ref:.17g}", flush=True)
        assert np.isfinite(a) and a == b
        assert abs(a - ref) <= 1e-13 * scale
        rc = lib.gbte_flux_keff_noalloc(q.ptr, v.ptr, 16_777_217, 1.0, 1.0, outs.at(0), None)
        assert rc == _lib.GBTE_BAD_INPUT and b"limit" in lib.gbte_problem_error()
        for x in (q, v, outs):
            x.free()
