"""The 16-bit rounding and decoding helpers the f16 / bf16 gemm parity tests
use, pinned on known bit patterns (ties to even, subnormals, the largest
finite f16) and against torch's own CPU conversions on random values."""
import numpy as np
import pytest

from tests import helpers as H


def test_bf16_known_patterns():
    x = np.array([1.0, -2.0, 3.140625, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20, 0.0, -0.0,
                  2.0 ** -130], np.float32)
    want = [0x3F80, 0xC000, 0x4049, 0x3F80, 0x3F82, 0x3F81, 0x0000, 0x8000, 0x0008]
    assert H.round_bf16(x).tolist() == want
    assert H.bf16_to_f64(np.array(want[:3], np.uint16)).tolist() == [1.0, -2.0, 3.140625]


def test_f16_known_patterns():
    x = np.array([1.0, 65504.0, 2.0 ** -24, -0.5, 1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11], np.float32)
    want = [0x3C00, 0x7BFF, 0x0001, 0xB800, 0x3C00, 0x3C02]
    assert H.round_f16(x).tolist() == want
    assert H.f16_to_f64(np.array(want[:4], np.uint16)).tolist() == [1.0, 65504.0, 2.0 ** -24, -0.5]


def test_rounding_matches_torch_cpu_conversions():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(100000) * np.exp2(rng.integers(-20, 20, 100000))).astype(np.float32)
    tb = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    th = torch.from_numpy(x).to(torch.float16).view(torch.int16).numpy().view(np.uint16)
    assert (H.round_bf16(x) == tb).all()
    assert (H.round_f16(x) == th).all()
