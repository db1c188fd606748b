"""Pins the oracle's Euler/RK2/RK4 steps (oracle/integrators.py) and the
tracking driver (oracle/tracker.py) against closed forms."""
import numpy as np
import pytest

from oracle import integrators as I
from oracle import tracker as TR

F32, F64 = np.float32, np.float64


@pytest.mark.parametrize("name", ["euler", "rk2", "rk4"])
def test_constant_field_is_exact_translation(name):
    c = np.array([0.3, -1.25, 2.0], F32)
    x = np.random.default_rng(0).random((64, 3)).astype(F32)
    got = I.STEPPERS[name](x, 0.1, 0.01, lambda p, t: np.broadcast_to(c, p.shape))
    np.testing.assert_allclose(got, x + F32(0.01) * c, rtol=0, atol=2e-7)


def _amplification(A, h, order):
    z = h * A
    P = np.eye(3)
    term = np.eye(3)
    for k in range(1, order + 1):
        term = term @ z / k
        P = P + term
    return P


@pytest.mark.parametrize("name,order", [("euler", 1), ("rk2", 2), ("rk4", 4)])
def test_linear_field_amplification_polynomial(name, order):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((3, 3))
    x0 = rng.standard_normal((10, 3))
    h = 0.05
    got = I.STEPPERS[name](x0, 0.0, h, lambda p, t: p @ A.T, dtype=F64)
    want = x0 @ _amplification(A, h, order).T
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def test_rk4_exact_for_cubic_time_dependence():
    # dx/dt = (t^3, t^2, 1 - t): x(t+h) - x(t) is a polynomial integral RK4 reproduces
    f = lambda p, t: np.broadcast_to(np.array([t ** 3, t ** 2, 1 - t]), p.shape)
    t, h = 0.7, 0.3
    got = I.rk4_step(np.zeros((1, 3)), t, h, f, dtype=F64)[0]
    want = [((t + h) ** 4 - t ** 4) / 4, ((t + h) ** 3 - t ** 3) / 3, h - ((t + h) ** 2 - t ** 2) / 2]
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)


def test_rk2_midpoint_time_is_float64_sum():
    seen = []
    f = lambda p, t: (seen.append(t), np.zeros_like(p))[1]
    I.rk2_step(np.zeros((1, 3), F32), F32(0.1), F32(0.3), f)
    assert seen[1] == np.float32(float(np.float32(0.1)) + float(np.float32(0.3) * np.float32(0.5)))


@pytest.mark.parametrize("name,order", [("euler", 1), ("rk2", 2), ("rk4", 4)])
def test_convergence_order_on_rotation(name, order):
    f = lambda p, t: np.stack([-p[:, 1], p[:, 0], np.zeros(len(p))], 1)
    x0 = np.array([[1.0, 0.0, 0.0]])
    errs = []
    for n in (40, 80):
        x = x0.copy()
        h = 1.0 / n
        for i in range(n):
            x = I.STEPPERS[name](x, i * h, h, f, dtype=F64)
        errs.append(np.abs(x[0] - [np.cos(1.0), np.sin(1.0), 0.0]).max())
    assert abs(np.log2(errs[0] / errs[1]) - order) < 0.15


def test_time_grid_and_recording():
    times, dt = TR.time_grid((0.0, 1.0), 101)
    assert times.dtype == F32 and len(times) == 101
    assert dt == float(np.float32(times[1] - times[0]))
    x0 = np.zeros((3, 2), F32)
    t, pos = TR.track(x0, (0, 1), 11, lambda p, t: np.ones_like(p), lambda p: p, "euler",
                      record_every=5)
    assert pos.shape == (3, 3, 3)                  # x0 + steps 5 and 10
    np.testing.assert_allclose(pos[-1, :, 0], 1.0, atol=1e-6)
    assert (pos[:, :, 2] == pos[:, :, 2]).all()    # 2-D positions padded with z = 0
