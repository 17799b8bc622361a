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
