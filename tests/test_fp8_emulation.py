"""tests/fp8_e4m3.py, the numpy restatement of OCP FP8 E4M3 that the FP8 Conv1d tests use as their reference,
pinned on the CPU against torch.float8_e4m3fn (the same format): every code decodes to the same value, and
encoding matches torch's round-to-nearest-even conversion on the grid, on the exact ties between neighbours and on
random values within the finite range."""
import numpy as np
import pytest

import fp8_e4m3 as F8

torch = pytest.importorskip("torch")


def test_decode_matches_torch_for_every_code():
    codes = torch.arange(256, dtype=torch.int32).to(torch.uint8)
    tv = codes.view(torch.float8_e4m3fn).to(torch.float64).numpy()
    ours = F8.DECODE
    nan = np.isnan(tv)
    assert np.array_equal(nan, np.isnan(ours))
    assert nan.sum() == 2  # S.1111.111 only
    assert np.array_equal(tv[~nan], ours[~nan])
    assert ours[0x7E] == 448.0 and ours[0x01] == 2.0 ** -9


def test_encode_matches_torch_within_range():
    rng = np.random.default_rng(3)
    grid = np.sort(F8.DECODE[np.isfinite(F8.DECODE)])
    mids = (grid[1:] + grid[:-1]) / 2
    x = np.concatenate([grid, mids, rng.uniform(-448, 448, 20000), rng.standard_normal(20000),
                        rng.standard_normal(5000) * 2.0 ** -8, [0.0, -0.0, 448.0, -448.0]])
    x = x.astype(np.float32).astype(np.float64)  # the kernels convert FP32 values
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    ours = F8.encode(x)
    bad = np.nonzero(ref != ours)[0]
    assert bad.size == 0, [(x[i], int(ref[i]), int(ours[i])) for i in bad[:8]]


def test_saturation_differs_from_torch_only_beyond_448():
    # torch turns values that round beyond 448 into NaN; the kernels' conversion saturates them to +-448.
    x = np.array([449.0, 470.0, 479.0, 480.0, 1e5, np.inf, -500.0, -np.inf])
    sat = F8.decode(F8.encode_saturating(x))
    assert np.array_equal(sat, np.copysign(448.0, x))
    tv = torch.from_numpy(x).to(torch.float8_e4m3fn).to(torch.float64).numpy()
    assert np.all(np.isnan(tv[[3, 4, 5, 6, 7]]))
