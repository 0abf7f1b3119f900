"""Failure statuses forced from the host: the kernels run normally on data
that makes the iteration fail, and the library reports it as a status
(the reference prints and calls exit, StaticBTESynthetic.cpp:2719-2723,
StaticBTESyntheticSolver.cpp:206-221)."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

import _oracle

pytestmark = pytest.mark.gpu


def test_negative_relaxation_time_diverges(tmp_path):
    """One band of four with tau < 0 (the bulk conductivity stays positive,
    so the synthetic matrix is still definite): its sweeps amplify and the
    synthetic loop's change grows, 'cannot congerve' after ten growing
    passes, GBTE_DIVERGED, as the oracle reports."""
    from giftbte_amd import Problem, BTESolver, GBTEError, _lib, decks
    bands = decks.debye_bands(4, 1)
    bands.tau = bands.tau.copy()
    bands.tau[2] *= -1.0
    spec = dataclasses.replace(decks.gray_square_spec(6), bands=bands)
    d = decks.write_deck(str(tmp_path / "neg_tau"), spec)
    p = Problem(d)
    _, info = _oracle.run(p.tables, 200, 1e-12, 1e-12)
    assert info["status"] == _lib.GBTE_DIVERGED
    s = BTESolver(p)
    s.fourier()
    with pytest.raises(GBTEError) as e:
        s.solve(200, 1e-12, 1e-12)
    print(f"\n[gpu failures] negative tau: oracle status {info['status']} after {info['iterations']} iterations;"
          f" gpu {e.value}", flush=True)
    assert e.value.code == _lib.GBTE_DIVERGED


def test_nan_temperature_is_reported(tmp_path):
    """A NaN in the starting temperature reaches the iteration errors:
    GBTE_NAN, not a hang or a silent result."""
    from giftbte_amd import Problem, BTESolver, GBTEError, _lib, decks
    d = decks.write_deck(str(tmp_path / "nan_T"), decks.gray_square_spec(6))
    p = Problem(d)
    s = BTESolver(p)
    s.fourier()
    T = np.full(p.tables.n_cells, 300.0)
    T[3] = np.nan
    s.set_initial(T=T)
    with pytest.raises(GBTEError) as e:
        s.solve(20, 0.0, 0.0)
    print(f"\n[gpu failures] NaN start: {e.value}", flush=True)
    assert e.value.code == _lib.GBTE_NAN


def test_write_outputs_rejects_another_problem(tmp_path):
    """The output writer sizes its buffers from the problem argument: one that
    is not the handle's is refused before anything is written."""
    from giftbte_amd import Problem, BTESolver, GBTEError, _lib, decks
    p_small = Problem(decks.write_deck(str(tmp_path / "small"), decks.gray_square_spec(3)))
    p_big = Problem(decks.write_deck(str(tmp_path / "big"), decks.gray_square_spec(6)))
    s = BTESolver(p_big)
    s.fourier()
    s.solve(1, 0.0, 0.0)
    rc = s.lib.gbte_write_outputs(s._h, p_small.handle, str(tmp_path).encode())
    assert rc == _lib.GBTE_BAD_INPUT
    assert b"not the one the handle was created from" in s.lib.gbte_last_error(s._h)
