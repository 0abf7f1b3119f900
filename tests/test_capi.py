"""The C ABI: every entry point of include/giftbte.h is exported, problem
construction reports errors as status codes, and solver calls fail loudly
(never fall back to the CPU) when no GPU is present."""
import ctypes as C
import os
import re

import pytest

from giftbte_amd import _lib, decks, Problem, BTESolver, GBTEError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "giftbte.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gbte_[a-z_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED) == names


def test_library_does_not_link_the_oracle():
    with open(_lib.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"oracle_run" not in blob and b"liboracle" not in blob


def test_missing_deck_is_bad_input(tmp_path):
    with pytest.raises(GBTEError) as e:
        Problem(str(tmp_path / "nowhere"))
    assert e.value.code == _lib.GBTE_BAD_INPUT
    assert "CONTROL" in str(e.value)


def test_unsupported_features_are_reported(tmp_path):
    d = decks.write_deck(str(tmp_path / "t"), decks.gray_square_spec(3))
    with open(os.path.join(d, "input", "CONTROL"), "a") as f:
        f.write("State=\nTransient\n")
    with pytest.raises(GBTEError) as e:
        Problem(d)
    assert e.value.code == _lib.GBTE_UNSUPPORTED


def test_unclosed_boundary_is_bad_input(tmp_path):
    spec = decks.gray_square_spec(3, bcs=[(1, 3, 0.0), (2, 1, 299.0), (4, 1, 301.0)])
    d = decks.write_deck(str(tmp_path / "open"), spec)
    with pytest.raises(GBTEError) as e:
        Problem(d)
    assert e.value.code == _lib.GBTE_BAD_INPUT
    assert "neither boundary nor inner face" in str(e.value)


def test_tables_view(tmp_path):
    p = Problem(decks.write_deck(str(tmp_path / "v"), decks.gray_square_spec(5, nphi=2)))
    t = p.tables
    assert (t.n_cells, t.n_dirs, t.n_bands, t.dim) == (50, 8, 1, 2)
    assert t.max_iter == 10000 and t.tol_temp == 1e-8


def test_solver_without_gpu_fails_loudly(tmp_path):
    """On a host without a HIP device gbte_create returns GBTE_HIP_ERROR; there
    is no CPU path to fall back to."""
    lib = _lib.load()
    p = Problem(decks.write_deck(str(tmp_path / "g"), decks.gray_square_spec(3, nphi=2)))
    h = C.c_void_p()
    opts = _lib.Options(0, 0, 0, 0.0)
    rc = lib.gbte_create(p.handle, C.byref(opts), C.byref(h))
    msg = lib.gbte_last_error(h).decode()
    lib.gbte_destroy(h)
    if rc == 0:
        pytest.skip("a HIP device is present; covered by the gpu tests")
    assert rc == _lib.GBTE_HIP_ERROR and "HIP device" in msg
    with pytest.raises(GBTEError):
        BTESolver(p)


def test_option_errors_are_bad_input(tmp_path):
    """Option checks come before any device work, so they hold without a GPU."""
    lib = _lib.load()
    p = Problem(decks.write_deck(str(tmp_path / "o"), decks.gray_square_spec(3, nphi=2)))
    for opts, word in ((_lib.Options(7, 0, 0, 0.0), "precision"),
                       (_lib.Options(_lib.GBTE_FP32_SWEEP, 0, 0, 0.0, 0, 1), "fallback_kernels")):
        h = C.c_void_p()
        rc = lib.gbte_create(p.handle, C.byref(opts), C.byref(h))
        msg = lib.gbte_last_error(h).decode()
        lib.gbte_destroy(h)
        assert rc == _lib.GBTE_BAD_INPUT and word in msg
    with pytest.raises(ValueError):
        BTESolver(p, precision="fp16")


def test_scheme_in_tables_and_validated(tmp_path):
    """Order / IterativeScheme / MatrixSolver reach the tables; values the
    reference's dispatch exits on are bad input (Solution.cpp:287-406)."""
    import dataclasses
    base = decks.gray_square_spec(3, nphi=2)
    p = Problem(decks.write_deck(str(tmp_path / "syn"), base))
    t = p.tables
    assert (t.order, t.scheme, t.matrix_solver) == (_lib.GBTE_ORDER_2, _lib.GBTE_SCHEME_SYNTHETIC, _lib.GBTE_MATRIX_LU)
    p = Problem(decks.write_deck(str(tmp_path / "seq"), dataclasses.replace(base, method="Sequential",
                                                                            matrix="BICGSTAB", order="2limit")))
    t = p.tables
    assert (t.order, t.scheme, t.matrix_solver) == (_lib.GBTE_ORDER_2LIMIT, _lib.GBTE_SCHEME_SEQUENTIAL,
                                                    _lib.GBTE_MATRIX_BICGSTAB)
    for field, bad in (("method", "Jacobi"), ("order", "3"), ("matrix", "QR")):
        with pytest.raises(GBTEError) as e:
            Problem(decks.write_deck(str(tmp_path / f"bad_{field}"), dataclasses.replace(base, **{field: bad})))
        assert e.value.code == _lib.GBTE_BAD_INPUT



def test_batch_argument_errors():
    """gbte_solve_batch refuses empty batches and repeated handles before any
    device work."""
    lib = _lib.load()
    one = (C.c_int * 2)(1, 1)
    tol = (C.c_double * 2)(0.0, 0.0)
    hs = (C.c_void_p * 2)(None, None)
    assert lib.gbte_solve_batch(hs, 0, one, tol, tol, None, None) == _lib.GBTE_BAD_INPUT
    assert lib.gbte_solve_batch(hs, 2, one, tol, tol, None, None) == _lib.GBTE_BAD_INPUT


def test_solve_batch_checks_every_handle_first():
    """A null handle in a batch: nothing runs and the status entry is filled
    (repeated handles are refused as a bad argument before that)."""
    lib = _lib.load()
    hs = (C.c_void_p * 1)(None)
    mi = (C.c_int * 1)(1)
    tt = (C.c_double * 1)(0.0)
    st = (_lib.Stats * 1)()
    status = (C.c_int * 1)(-7)
    rc = lib.gbte_solve_batch(hs, 1, mi, tt, tt, st, status)
    assert rc == _lib.GBTE_HIP_ERROR
    assert list(status) == [_lib.GBTE_HIP_ERROR]
    hs2 = (C.c_void_p * 2)(None, None)
    assert lib.gbte_solve_batch(hs2, 2, (C.c_int * 2)(1, 1), (C.c_double * 2)(0, 0), (C.c_double * 2)(0, 0),
                                (_lib.Stats * 2)(), (C.c_int * 2)()) == _lib.GBTE_BAD_INPUT
