"""Pins of the learned indicator's CPU restatement (oracle/learned.py): gradient against central differences,
upwind-5 exactness, the WENO5-Z departure and the MUSCL clip at their closed-form cases, the limiters' values,
feature invariances, bfloat16 rounding."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import learned as L  # noqa: E402


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.4, L.NPARAM)
    X = rng.normal(0, 0.5, (40, 15))
    y = (rng.random(40) < 0.3).astype(float)
    loss, g = L.loss_and_grad(X, y, p, pos_weight=3.0)
    h = 1e-6
    for k in rng.choice(L.NPARAM, 60, replace=False).tolist() + [L.NPARAM - 1]:
        e = np.zeros(L.NPARAM)
        e[k] = h
        fd = (L.loss_and_grad(X, y, p + e, 3.0)[0] - L.loss_and_grad(X, y, p - e, 3.0)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    assert np.isclose(loss, L.loss_and_grad(X, y, L.pack(*L.unpack(p)).astype(np.float64), 3.0)[0], rtol=1e-5)


def test_upwind5_exact_for_quartic_cell_averages():
    # cell averages of x^4 on unit cells [j, j+1]: ((j+1)^5 - j^5) / 5; the face value at x = 3 (right of c2)
    avg = lambda j: ((j + 1.0) ** 5 - j ** 5) / 5.0   # noqa: E731
    c = [avg(j) for j in range(5)]
    assert abs(L.upwind5(*c) - 3.0 ** 4) < 1e-12
    assert abs(L.upwind5(*[avg(j) + 2.0 for j in range(5)]) - (81.0 + 2.0)) < 1e-12


def test_departure_closed_forms():
    j = np.arange(6, dtype=float)
    lin = (1.3 + 0.07 * j)[:, None, None] * np.ones((6, 5, 1))
    assert L.departure(lin, 1e-6, 3).max() < 1e-12            # linear data: tau = 0, weights = linear weights
    step = np.where(j < 3, 1.0, 0.125)[:, None, None] * np.ones((6, 5, 1))
    assert L.departure(step, 1e-6, 3).max() > 0.25              # a jump concentrates the weights
    flat = np.full((6, 5, 1), 2.0)
    assert L.departure(flat, 1e-6, 3).max() == 0.0


def test_limiters_closed_forms():
    # phi(1, 3) of the four limiters (Schemes.hpp:82-110): minmod 1, van Leer 2*3/4, MC min(2, 2, 6), Koren min(2, 5/3, 6);
    # opposite signs give 0, phi(a, a) = a, and phi is odd
    want = {0: 1.0, 1: 1.5, 2: 2.0, 3: 5.0 / 3.0}
    for lim, w in want.items():
        assert abs(float(L.limiter(lim, 1.0, 3.0)) - w) < 1e-15, lim
        assert float(L.limiter(lim, 1.0, -3.0)) == 0.0 and float(L.limiter(lim, 0.0, 2.0)) == 0.0
        assert float(L.limiter(lim, 0.5, 0.5)) == 0.5
        assert float(L.limiter(lim, -1.0, -3.0)) == -w


def test_clip_closed_forms():
    # the MUSCL clip label: zero on linear data for every limiter; a jump: the limited slope of the face value next to
    # it is far from the central one; a symmetric extremum (a = -b): phi = 0 = (a + b)/2, no clip
    j = np.arange(6, dtype=float)
    lin = (1.3 + 0.07 * j)[:, None, None] * np.ones((6, 5, 1))
    step = np.where(j < 3, 1.0, 0.125)[:, None, None] * np.ones((6, 5, 1))
    for lim in range(4):
        assert L.clip(lin, 1e-6, 3, lim).max() < 1e-12, lim
        assert L.clip(step, 1e-6, 3, lim).max() > 0.25, lim
    peak = np.array([1.0, 1.0, 2.0, 1.0, 1.0, 1.0])[:, None, None] * np.ones((6, 5, 1))   # c2 - c1 = -(c3 - c2)
    c = peak[:, 0, 0]
    a, b = c[2] - c[1], c[3] - c[2]
    assert a == -b and L.clip(peak[:, :, :], 1e-6, 3, 2)[0] > 0.0   # the right value's pair (c4 - c3, b) does clip
    only_left = L.limiter(2, a, b) - 0.5 * (a + b)
    assert only_left == 0.0
    # "both" is the larger of the two sources
    rng = np.random.default_rng(5)
    st = 1.0 + 0.3 * rng.random((6, 5, 50))
    both = L.label_value(st, 1e-6, 3, "both", 2)
    assert np.array_equal(both, np.maximum(L.departure(st, 1e-6, 3), L.clip(st, 1e-6, 3, 2)))


def test_departure_ignores_rounding_noise_in_a_near_zero_velocity():
    # a velocity component that is zero up to noise of 1e-7 beside a pressure of 71 (Mach 0.1): WENO5-Z regularised by
    # the component's own mean square sees jumps; the label, regularised by |v|^2 + p / rho, does not; a real jump of
    # that component (0.3, a tenth of the sound speed) still counts
    rng = np.random.default_rng(3)
    st = np.zeros((6, 5, 200))
    st[:, 0], st[:, 3] = 1.0, 71.0
    st[:, 4] = 1e-7 * rng.standard_normal((6, 200))
    own = np.array([L.weno5z_weights(*[st[s, 4] for s in range(5)], 1e-6) for _ in range(1)])[0]
    assert np.abs(own - L.D_LIN[:, None]).max() > 0.25
    assert L.departure(st, 1e-6, 3).max() < 1e-3
    st[3:, 4] += 0.3
    assert L.departure(st, 1e-6, 3).min() > 0.25


def test_feature_invariances():
    rng = np.random.default_rng(1)
    st = 1.0 + 0.1 * rng.random((6, 5, 50))
    f = L.features(st, 1e-6, 3)
    assert f.shape == (50, 15) and np.all(np.abs(f) <= np.sqrt(2.0) + 1e-12)   # |d1 - d0| <= sqrt(2) N
    st2 = st.copy()
    st2[:, 0] *= 7.0   # density scale: its five features do not change
    assert np.allclose(L.features(st2, 1e-6, 3)[:, :5], f[:, :5], atol=1e-12)
    ramp = (0.5 + 0.01 * np.arange(6.0))[:, None, None] * np.ones((6, 5, 1))
    fr = L.features(ramp, 0.0, 3)
    assert np.allclose(fr[0, :4], 0.0, atol=1e-12)           # a linear ramp: no second differences
    assert np.isclose(fr[0, 4], 1.0 / np.sqrt(5.0))          # five equal first differences
    step = np.where(np.arange(6) < 3, 1.0, 0.125)[:, None, None] * np.ones((6, 5, 1))
    fs = L.features(step, 0.0, 3)
    assert np.isclose(fs[0, 4], 1.0) and np.isclose(fs[0, 1], -1.0) and np.isclose(fs[0, 2], 1.0)   # one jump


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1.0 + 2.0 ** -9], dtype=np.float32)
    assert L.bf16(x).tolist() == [1.0, 1.0, 1.015625, -2.5, 1.0]
