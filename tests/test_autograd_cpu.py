"""The torch.autograd.Function around the adjoint, wired to a stand-in for
the solve (no GPU): backward scales the stored gradient by the incoming one
and gives no gradient to the non-tensor arguments."""
import numpy as np
import pytest


def test_keff_function_backward(monkeypatch):
    torch = pytest.importorskip("torch")
    from giftbte_amd import autograd

    def fake(spec, s, t_hot, hot_index, axis, delta_t):
        return 2.0 + float(np.dot(s, [1.0, 2.0, 3.0])) + 0.5 * t_hot, np.array([1.0, 2.0, 3.0]), 0.5, None, None

    monkeypatch.setattr(autograd, "keff_and_grad", fake)
    s = torch.ones(3, dtype=torch.float64, requires_grad=True)
    th = torch.tensor(301.0, dtype=torch.float64, requires_grad=True)
    k = autograd.keff(s, th, None, 4)
    assert float(k) == pytest.approx(2.0 + 6.0 + 150.5)
    (3 * k).backward()
    assert s.grad.tolist() == [3.0, 6.0, 9.0]
    assert float(th.grad) == 1.5
