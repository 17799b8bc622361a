"""Host-side pieces of the surrogate's network files (no GPU)."""
import numpy as np

from ripple_amd.surrogate import _bf16_bits


def test_bf16_bits_round_to_nearest_even_like_torch():
    """save_network stores weights as the bf16 bit patterns the device's
    conversion produces (round to nearest, ties to even); torch's bfloat16
    cast rounds the same way."""
    import torch  # imported here only: the GPU test session keeps torch's HIP runtime out
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=100000).astype(np.float32) * np.float32(1e3),
                        # ties: 1 + 2^-8 rounds down to 1, 1 + 3 * 2^-8 up to 1 + 2^-6
                        np.array([0.0, -0.0, np.inf, -np.inf, 1.00390625, 1.01171875, 3.4e38, -1e-40],
                                 np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(_bf16_bits(x), ref)
    assert _bf16_bits(np.array([1.00390625], np.float32))[0] == 0x3F80
    assert _bf16_bits(np.array([1.01171875], np.float32))[0] == 0x3F82
    nan = _bf16_bits(np.array([np.nan], np.float32))[0]
    assert (nan & 0x7F80) == 0x7F80 and (nan & 0x7F) != 0


def test_source_value_field_forms():
    """source_value (Sources/Sources.jl:1450-1467): a number, an (Nx, Ny)
    array, or a function of (x, y, t), (x, y) or (t), tried in that order."""
    import numpy as np
    import pytest
    import ripple_amd as ra
    from ripple_amd.sources import source_value_field, _arity
    xn, yn = np.array([0.5, 1.5, 2.5]), np.array([10.0, 20.0])
    np.testing.assert_array_equal(source_value_field(0.3, xn, yn, 7.0), np.full((3, 2), 0.3))
    a = np.arange(6.0).reshape(3, 2)
    np.testing.assert_array_equal(source_value_field(a, xn, yn, 7.0), a)
    np.testing.assert_array_equal(source_value_field(lambda x, y, t: x + y + t, xn, yn, 7.0),
                                  xn[:, None] + yn[None, :] + 7.0)
    np.testing.assert_array_equal(source_value_field(lambda x, y: x * y, xn, yn, 7.0), xn[:, None] * yn[None, :])
    np.testing.assert_array_equal(source_value_field(lambda t: 2 * t, xn, yn, 7.0), np.full((3, 2), 14.0))
    import math
    # a function of scalars only is evaluated column by column
    np.testing.assert_allclose(source_value_field(lambda x, y: math.hypot(x, y), xn, yn, 0.0),
                               np.hypot(xn[:, None], yn[None, :]), rtol=1e-15)
    assert _arity(lambda x, y, t=0.0: 0) == 3 and _arity(lambda *a: 0) == 3
    with pytest.raises(ra.RippleArgumentError):
        _arity(lambda a, b, c, d: 0)
    with pytest.raises(ValueError):
        source_value_field(np.zeros((2, 2)), xn, yn, 0.0)
    with pytest.raises(ValueError):
        ra.QuadrupletTransferInteraction((1, 1), (1, 2), (1, 3), (1, 4), rate=np.zeros((2, 2, 2)))
