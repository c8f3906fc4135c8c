"""The gradcheck wrapper imports without a GPU and is a torch.autograd.Function."""
import torch


def test_autograd_wrapper_is_a_function():
    from meamzilla_amd.autograd import ObjectiveFunction, objective
    assert issubclass(ObjectiveFunction, torch.autograd.Function)
    assert callable(objective)
