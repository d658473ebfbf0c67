"""Argument checks of the Python binding (no GPU). The C entry points take bare device pointers
without lengths, so a wrong size, dtype or step range must be refused before any call reaches
the library; these tests use meta tensors and stand-in objects, so nothing is launched."""
import pytest
import torch

import exadg_amd


def meta(*shape, dtype=torch.float64):
    return torch.empty(*shape, dtype=dtype, device="meta")


class FakeOp:
    def __init__(self, n):
        self.n = n
        self.handle = None


def test_check_vector_rejects_size_dtype_layout_and_host_memory():
    with pytest.raises(ValueError, match="entries"):
        exadg_amd.check_vector(meta(10), 11, torch.float64)
    with pytest.raises(ValueError, match="dtype"):
        exadg_amd.check_vector(meta(10, dtype=torch.float32), 10, torch.float64)
    with pytest.raises(ValueError, match="contiguous"):
        exadg_amd.check_vector(meta(4, 5).t(), 20, torch.float64)
    with pytest.raises(ValueError, match="device"):
        exadg_amd.check_vector(torch.zeros(10, dtype=torch.float64), 10, torch.float64)


def test_sequence_step_range_is_checked():
    P, B = meta(6, 100), meta(6, 100)
    with pytest.raises(ValueError, match="outside"):
        exadg_amd.check_sequence(P, B, 0, 7, history=3)
    with pytest.raises(ValueError, match="outside"):
        exadg_amd.check_sequence(P, B, -1, 6, history=3)
    with pytest.raises(ValueError, match="shorter"):
        exadg_amd.check_sequence(P, B, 2, 5, history=3)
    with pytest.raises(ValueError, match="same shape"):
        exadg_amd.check_sequence(P, meta(5, 100), 0, 5, history=3)
    with pytest.raises(ValueError, match="2-D"):
        exadg_amd.check_sequence(meta(600), meta(600), 0, 5, history=3)


def test_guess_train_and_evaluate_refuse_rows_past_the_end():
    net = object.__new__(exadg_amd.GuessNet)  # no handle: a call that got through would fail
    net.history, net.n_outputs, net.n_features = 3, 27, 108
    P, B = meta(6, 27 * 8), meta(6, 27 * 8)
    with pytest.raises(ValueError, match="outside"):
        net.train(P, B, 0, 10, 1)
    with pytest.raises(ValueError, match="outside"):
        net.evaluate(P, B, 2, 8)
    with pytest.raises(ValueError, match="history"):
        net.predict([meta(216)] * 2, meta(216), meta(216))
    with pytest.raises(ValueError, match="whole number"):
        net.predict([meta(216)] * 3, meta(216), meta(215))


def test_solver_refuses_mismatched_preconditioner_and_vectors():
    op = FakeOp(1000)
    prec = object.__new__(exadg_amd.Preconditioner)
    prec.n = 999
    with pytest.raises(ValueError, match="sizes differ"):
        exadg_amd.cg_solve(op, meta(1000), meta(1000), prec)
    with pytest.raises(ValueError, match="entries"):
        exadg_amd.cg_solve(op, meta(999), meta(1000), None)
    with pytest.raises(ValueError, match="dtype"):
        exadg_amd.cg_solve(op, meta(1000, dtype=torch.float32), meta(1000), None)
    prec.n = 1000
    with pytest.raises(ValueError, match="entries"):
        prec.vmult(meta(10), meta(1000))
    with pytest.raises(ValueError, match="sizes differ"):
        prec.smooth(FakeOp(10), meta(10), meta(10))


def test_preconditioner_keeps_its_operator_alive():
    """Point Jacobi refers to its operator: the Python object holds a reference to it."""
    import inspect
    src = inspect.getsource(exadg_amd.Preconditioner.__init__)
    assert "self._op = op" in src


def test_extrapolate_checks_history():
    with pytest.raises(ValueError, match="order"):
        exadg_amd.extrapolate([meta(10)] * 5, meta(10))
    with pytest.raises(ValueError, match="entries"):
        exadg_amd.extrapolate([meta(9), meta(10)], meta(10))
