"""sparse_solve's value-type contract, checked before any device work (CPU).

The library stores FP32 matrices: float64 values that FP32 cannot hold exactly
are refused (they would be rounded, and x and the gradients would belong to
another matrix); float64 values that FP32 holds exactly are accepted.
"""
import pytest
import torch

from mpsparse_amd.autograd import _fp32_values, sparse_solve


def test_float64_values_not_in_fp32_are_refused():
    z = torch.zeros(3, dtype=torch.int64)
    v = torch.tensor([1.0, 0.1, 2.0], dtype=torch.float64)  # 0.1 is not an FP32 value
    with pytest.raises(ValueError, match="FP32"):
        sparse_solve(z, z, v, torch.ones(3, dtype=torch.float64))


def test_fp32_representable_float64_values_pass_through_exactly():
    v = torch.tensor([1.0, 0.1, -3.25e-7], dtype=torch.float32).double()
    out = _fp32_values(v)
    assert out.dtype == torch.float32 and torch.equal(out.double(), v)


def test_other_value_types_are_refused():
    with pytest.raises(TypeError):
        _fp32_values(torch.ones(3, dtype=torch.float16))
