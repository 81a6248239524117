"""Pins the CPU oracle (oracle/bkma_oracle.c) before it is trusted as the
checker of the GPU kernels.  None of these checks restates the formula under
test; each one comes from the reference's own tests, a fixture, a closed form
or an invariant that a plausible transcription mistake breaks.

Pinning map (oracle function -> checks):
  orc_lag_basis        exact-rational fixture at p=1/3, nodal values,
                       reproduction of degree <= 4 polynomials, reflection
                       symmetry, NOT textbook Lagrange (degree-5 defect)
  orc_adv_foot/_cell   integer-shift known answer (bit-exact circular shift),
  orc_advect_f64       dt = 0 identity, analytic L1 < 1e-6 criterion of the
                       reference tests (impl_unittests.cpp:8-85,
                       validation_unittests.cpp:11-94), linearity
  orc_advect_f32_mixed distance to orc_advect_f64 within the FP32 budget,
                       integer-shift bit-exactness
  orc_advect_f32_blocked one pass of n steps within one FP32 rounding of
                       orc_advect_f64 (independent FP64 path), integer-shift
                       bit-exactness, error shrinking with the pass length,
                       remainder-pass composition
  orc_l1_error_adv     exactly 0 at t = 0 (ValidateNoIteration)
  orc_conv_literal_f64 constant-input known answer 1 + C*K*1.5*7.3 (the
                       reference's fill_buffer_conv1d values, init.hpp:50-72),
                       numpy.convolve of channel 0 with the ic-summed kernel,
                       batch invariance
  orc_conv_torch_f64   torch.nn.functional.conv1d in FP64
  orc_uniform          range, moments, seed/index independence
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------
# lag_basis
# ---------------------------------------------------------------------------
def test_lag_basis_exact_rational_fixture():
    fx = json.load(open(os.path.join(GOLDEN, "lag_basis_p1_3_exact.json")))
    p = float(Fraction(fx["p"]))
    c = O.lag_basis(2.0 + p)
    np.testing.assert_allclose(c * 729.0, fx["coef_times_729"], rtol=0, atol=1e-12)
    assert not np.allclose(c * 729.0, fx["textbook_lagrange_times_729"], atol=1e-3)


def test_lag_basis_nodal_values_exact():
    assert list(O.lag_basis(2.0)) == [0, 0, 1, 0, 0, 0]
    assert list(O.lag_basis(3.0)) == [0, 0, 0, 1, 0, 0]


@pytest.mark.parametrize("p", [0.0, 0.1, 1 / 3, 0.5, 0.77, 0.9, 0.999])
def test_lag_basis_reproduces_degree_4(p):
    # nodes at offsets -2..3 from leftNode, evaluation point p in [0, 1)
    c = O.lag_basis(2.0 + p)
    nodes = np.arange(-2, 4, dtype=np.float64)
    for m in range(5):
        assert abs(np.dot(c, nodes ** m) - p ** m) < 1e-13 * max(1, 3 ** m)


def test_lag_basis_reflection_symmetry():
    for p in np.linspace(0, 1, 17):
        np.testing.assert_allclose(O.lag_basis(2.0 + p), O.lag_basis(3.0 - p)[::-1], atol=1e-15)


def test_lag_basis_is_not_textbook_lagrange():
    # degree 5 is not reproduced except at p in {0, 1/2, 1}
    c = O.lag_basis(2.0 + 1 / 3)
    nodes = np.arange(-2, 4, dtype=np.float64)
    assert abs(np.dot(c, nodes ** 5) - (1 / 3) ** 5) > 1e-3
    c = O.lag_basis(2.5)
    assert abs(np.dot(c, nodes ** 5) - 0.5 ** 5) < 1e-13


# ---------------------------------------------------------------------------
# Advection step
# ---------------------------------------------------------------------------
def test_integer_shift_known_answer_bit_exact():
    # dt = 1/8, v = -1 + i0/128 => v*dt = (i0 - 128)/1024 = integer cells:
    # out[i1] = in[(i1 + 128 - i0) mod n1], exactly.
    p = O.make_params(256, 1024, 1, dt=1 / 8, minRealVx=-1, maxRealVx=1)
    d = O.uniform(256 * 1024, 0).reshape(256, 1024, 1)
    d0 = d.copy()
    O.advect_f64(d, p, 1)
    for i0 in range(256):
        assert np.array_equal(d[i0, :, 0], np.roll(d0[i0, :, 0], -(128 - i0)))
    f = d0.astype(np.float32)
    O.advect_f32_mixed(f, p, 1)
    assert np.array_equal(f, np.stack([np.roll(d0[i, :, 0], -(128 - i)) for i in range(256)])[:, :, None]
                          .astype(np.float32))


def test_integer_shift_strided_lines():
    p = O.make_params(8, 64, 5, dt=1 / 8, minRealVx=-1, maxRealVx=1)  # dvx = 1/4, v*dt*64 = 2*i0 - 8
    d = O.uniform(8 * 64 * 5, 7).reshape(8, 64, 5)
    d0 = d.copy()
    O.advect_f64(d, p, 1)
    for i0 in range(8):
        s = int(round((-1 + i0 / 4) * (1 / 8) * 64))
        assert np.array_equal(d[i0], np.roll(d0[i0], s, axis=0))


def test_zero_dt_is_identity():
    # bit-exact when dx is a power of two (every foot lands on its node)
    p = O.make_params(3, 128, 4, dt=0.0, minRealVx=-1, maxRealVx=1)
    d = O.uniform(3 * 128 * 4, 3).reshape(3, 128, 4)
    d0 = d.copy()
    O.advect_f64(d, p, 5)
    assert np.array_equal(d, d0)
    # otherwise (x*inv_dx rounds below the node) identity up to rounding
    p = O.make_params(3, 100, 4, dt=0.0, minRealVx=-1, maxRealVx=1)
    d = O.uniform(3 * 100 * 4, 3).reshape(3, 100, 4)
    d0 = d.copy()
    O.advect_f64(d, p, 5)
    np.testing.assert_allclose(d, d0, rtol=0, atol=1e-13)


def test_l1_zero_at_t0():
    rng = np.random.default_rng(1)
    for _ in range(3):
        p = O.make_params(int(rng.integers(1, 128)), int(rng.integers(1, 1024)), int(rng.integers(1, 64)))
        p.maxIter = 0
        f = O.fill_adv_sin4(p)
        hi, lo = O.l1_error_adv(f, p)
        assert hi == 0.0 and lo == 0.0


def test_reference_impl_criterion_default_params():
    # Impl.* tests: default ADVParams (n0=32, n1=1024, n2=32, dt=1e-4,
    # v in [0, 1), 100 steps) and L1 < EPS = 1e-6 (impl_unittests.cpp:8-28).
    p = O.make_params(32, 1024, 32, dt=1e-4, maxIter=100)
    f = O.fill_adv_sin4(p)
    O.advect_f64(f, p, 100)
    hi, _ = O.l1_error_adv(f, p)
    assert hi < 1e-6


def test_reference_validate_each_iteration():
    # Validation.ValidateEachIterFor10Iterations (validation_unittests.cpp:34-65)
    p = O.make_params(17, 512, 9, dt=1e-4)
    f = O.fill_adv_sin4(p)
    for it in range(1, 11):
        O.advect_f64(f, p, 1)
        p.maxIter = it
        assert O.l1_error_adv(f, p)[0] < 1e-6


def test_wide_velocity_range_accuracy():
    # v in [-1, 1), dt = 1e-3 (the benchmark configuration), 50 steps
    p = O.make_params(64, 1024, 2, dt=1e-3, minRealVx=-1, maxRealVx=1, maxIter=50)
    f = O.fill_adv_sin4(p)
    O.advect_f64(f, p, 50)
    assert O.l1_error_adv(f, p)[0] < 1e-6


def test_linearity():
    p = O.make_params(4, 96, 3, dt=1e-3, minRealVx=-1, maxRealVx=1)
    a = O.uniform(4 * 96 * 3, 11).reshape(4, 96, 3)
    b = O.uniform(4 * 96 * 3, 12).reshape(4, 96, 3)
    s = 0.5 * a + b
    for x in (a, b, s):
        O.advect_f64(x, p, 3)
    np.testing.assert_allclose(s, 0.5 * a + b, atol=1e-14)


def test_f32_mixed_close_to_f64():
    p = O.make_params(16, 1024, 8, dt=1e-3, minRealVx=-1, maxRealVx=1)
    d = O.uniform(16 * 1024 * 8, 1).reshape(16, 1024, 8)
    f = d.astype(np.float32)
    O.advect_f64(d, p, 50)
    O.advect_f32_mixed(f, p, 50)
    err = np.abs(f.astype(np.float64) - d).max() / np.abs(d).max()
    assert 0 < err < 1e-6


def test_f32_blocked_single_pass_is_one_rounding_of_f64():
    # n steps in one pass: FP64 throughout (fused instead of separate
    # multiply-add), then one rounding -- within one FP32 ulp of the rounded
    # FP64 path, and equal to it almost everywhere.
    p = O.make_params(8, 512, 4, dt=1e-3, minRealVx=-1, maxRealVx=1)
    d = O.uniform(8 * 512 * 4, 11).reshape(8, 512, 4)
    f = d.astype(np.float32)
    d = f.astype(np.float64)
    O.advect_f64(d, p, 6)
    O.advect_f32_blocked(f, p, 6, 6)
    ref = d.astype(np.float32)
    ulp = np.spacing(np.abs(ref))
    assert np.all(np.abs(f - ref) <= ulp)
    assert np.mean(f == ref) > 0.999


@pytest.mark.parametrize("ns", [1, 2, 3, 5])
def test_f32_blocked_integer_shift_bit_exact(ns):
    n0, n1, n2 = 4, 1024, 3
    p = O.make_params(n0, n1, n2, dt=1 / 512, minRealVx=-1, maxRealVx=1)
    f = O.uniform(n0 * n1 * n2, 12, np.float32).reshape(n0, n1, n2)
    g = f.copy()
    O.advect_f32_blocked(g, p, 7, ns)
    for i0 in range(n0):
        s = int(round((-1 + i0 / 2) * 2))
        assert np.array_equal(g[i0], np.roll(f[i0], 7 * s, axis=0))


def test_f32_blocked_error_shrinks_with_pass_length():
    p = O.make_params(16, 1024, 8, dt=1e-3, minRealVx=-1, maxRealVx=1)
    d = O.uniform(16 * 1024 * 8, 1).reshape(16, 1024, 8)
    f0 = d.astype(np.float32)
    O.advect_f64(d, p, 50)
    errs = []
    for ns in (1, 2, 5):
        f = f0.copy()
        O.advect_f32_blocked(f, p, 50, ns)
        errs.append(np.abs(f.astype(np.float64) - d).max() / np.abs(d).max())
    assert errs[0] > errs[1] > errs[2] > 0
    assert errs[2] < 1e-6


def test_f32_blocked_remainder_pass():
    p = O.make_params(3, 256, 2, dt=2e-3, minRealVx=-1, maxRealVx=1)
    f = O.uniform(3 * 256 * 2, 13, np.float32).reshape(3, 256, 2)
    a, b = f.copy(), f.copy()
    O.advect_f32_blocked(a, p, 7, 3)
    for k in (3, 3, 1):
        O.advect_f32_blocked(b, p, k, k)
    assert np.array_equal(a, b)
    c = f.copy()
    O.advect_f32_blocked(c, p, 4, 1)
    d = f.copy()
    O.advect_f32_mixed(d, p, 4)
    assert np.array_equal(c, d)


# ---------------------------------------------------------------------------
# Generator
# ---------------------------------------------------------------------------
def test_uniform_generator():
    u = O.uniform(1 << 20, 1)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 2e-3 and abs(u.var() - 1 / 12) < 2e-3
    assert not np.array_equal(u[:100], O.uniform(100, 2))
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 5e-3
    assert np.array_equal(O.uniform(10, 1), u[:10])
    assert np.array_equal(O.uniform(10, 1, np.float32), u[:10].astype(np.float32))


# ---------------------------------------------------------------------------
# Conv1d
# ---------------------------------------------------------------------------
def _literal_case(n0, n2, C, L, K, seed):
    rng = np.random.default_rng(seed)
    data = rng.standard_normal((n0, C * L, n2))
    w = rng.standard_normal((K, C, C))
    b = rng.standard_normal(C)
    return data, w, b


def test_conv_literal_constant_known_answer():
    # fill_buffer_conv1d values: data 7.3, weights 1.5, bias 1.0; C = K = 3.
    C, K, L = 3, 3, 64
    data = np.full((2, C * L, 3), 7.3)
    w = np.full((K, C, C), 1.5)
    b = np.ones(C)
    O.conv_literal_f64(data, w, b, K, C, L)
    nw = L - K + 1
    for oc in range(C):
        np.testing.assert_allclose(data[:, oc * L: oc * L + nw, :], 1 + C * K * 1.5 * 7.3, rtol=1e-15)
    assert abs(1 + C * K * 1.5 * 7.3 - 99.55) < 1e-12


def test_conv_literal_matches_numpy_convolve():
    C, K, L = 4, 5, 37
    data, w, b = _literal_case(3, 2, C, L, K, 5)
    ref = data.copy()
    O.conv_literal_f64(data, w, b, K, C, L)
    n1 = C * L
    nw = n1 - K + 1
    wsum = w.sum(axis=1)  # (K, C_out): the solver adds every ic with the same x
    for i0 in range(3):
        for i2 in range(2):
            x0 = ref[i0, :L, i2]
            out = np.empty(n1)
            for oc in range(C):
                out[oc * L:(oc + 1) * L] = b[oc] + np.convolve(x0, wsum[:, oc])[:L]
            exp = ref[i0, :, i2].copy()
            exp[:nw] = out[K - 1:]
            np.testing.assert_allclose(data[i0, :, i2], exp, rtol=1e-12, atol=1e-12)


def test_conv_literal_batch_invariance():
    C, K, L = 3, 3, 50
    rng = np.random.default_rng(0)
    line = rng.standard_normal(C * L)
    data = np.broadcast_to(line[None, :, None], (4, C * L, 3)).copy()
    w = rng.standard_normal((K, C, C))
    b = rng.standard_normal(C)
    O.conv_literal_f64(data, w, b, K, C, L)
    assert np.all(data == data[0:1, :, 0:1])


def test_conv_torch_matches_torch_conv1d():
    torch = pytest.importorskip("torch")
    B, C, L, K = 3, 8, 40, 5
    rng = np.random.default_rng(2)
    x = rng.standard_normal((B, C, L))
    w = rng.standard_normal((C, C, K)) / np.sqrt(C * K)
    b = rng.standard_normal(C)
    y = torch.nn.functional.conv1d(torch.from_numpy(x), torch.from_numpy(w), torch.from_numpy(b)).numpy()
    xx = x.copy()
    O.conv_torch_f64(xx, w, b)
    np.testing.assert_allclose(xx[:, :, :L - K + 1], y, rtol=1e-13, atol=1e-13)
    assert np.array_equal(xx[:, :, L - K + 1:], x[:, :, L - K + 1:])
