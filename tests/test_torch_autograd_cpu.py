"""Host-side checks of the torch.autograd wrappers (no GPU needed): the module
imports, and tensors that are not on a GPU are refused before any library
call (there is no CPU path)."""
import pytest

torch = pytest.importorskip("torch")

from vulkpy_amd import torch_autograd as ta  # noqa: E402


@pytest.mark.parametrize("fn", [ta.relu, ta.sigmoid, ta.softmax])
def test_cpu_tensor_refused(fn):
    with pytest.raises(ValueError, match="GPU"):
        fn(torch.zeros((2, 3)))


def test_cpu_tensor_refused_dense_and_losses():
    with pytest.raises(ValueError, match="GPU"):
        ta.dense(torch.zeros((2, 3)), torch.zeros((4, 3)), torch.zeros(4))
    for loss in (ta.softmax_cross_entropy, ta.mse, ta.huber):
        with pytest.raises(ValueError, match="GPU"):
            loss(torch.zeros((2, 3)), torch.zeros((2, 3)))


def test_dtype_and_arguments_checked_first():
    with pytest.raises(TypeError):
        ta.relu(torch.zeros((2, 3), dtype=torch.float64))
    with pytest.raises(ValueError):
        ta.dense(torch.zeros((2, 3)), torch.zeros((4, 3)), precision="fp16")
    with pytest.raises(ValueError):
        ta.mse(torch.zeros((2, 3)), torch.zeros((2, 3)), reduce="max")
