"""Pins the oracle's dense output (sol(t), sol(t, Val{1}) of MIRKInterpolation,
MIRK/interpolation.jl:17-206) with the reference's own interpolation tests
(MIRKt/Core/mirk_basic_tests.jl:173-313: the two-point problem u'' = u with its
exponential solution, value and derivative at t = 0.001 and 0.04, atol 1e-6; the
multipoint problem with sin/cos solution at pi/6 and pi/3) and with what the
mathematics fixes: the extension reproduces the left node exactly (w(0) = 0), ends
at y_{i+1} - Phi_i (w(1) = b), its derivative is the derivative of its value, and its
error decreases at the method's order.
"""
import math

import numpy as np
import pytest

import pyoracle as O
import cases


def prob_tp(dt=0.001):
    """u'' = u / lam (lam = 1), u(0) = 1, u(1) = 0 (STANDARD BCs at the end points)."""
    return cases.linear([[0.0, 1.0], [1.0, 0.0]], c=0.0, tspan=(0.0, 1.0),
                        bc=[(0.0, 0, 1.0), (1.0, 0, 0.0)], dt=dt, u0=np.array([1.0, 0.0]), standard=True)


def tp_exact(t):
    a = 1.0
    d = 1 - math.exp(-2 * a)
    u = (math.exp(-a * t) - math.exp((t - 2) * a)) / d
    up = (-a * math.exp(-t * a) - a * math.exp((t - 2) * a)) / d
    return np.array([u, up]), np.array([up, u])


def prob_mp(dt=0.001):
    """u1' = u2, u2' = -u1 on [0, pi/2], sol(pi/6)[1] = 0.5, sol(pi/3)[2] = 0.5."""
    return cases.linear([[0.0, 1.0], [-1.0, 0.0]], c=0.0, tspan=(0.0, math.pi / 2),
                        bc=[(math.pi / 6, 0, 0.5), (math.pi / 3, 1, 0.5)], dt=dt, u0=np.array([0.0, 1.0]),
                        standard=True)


def _solved(spec, method, **kw):
    sol = O.solve(O.opts(method, **kw), spec)
    assert sol.retcode == 0
    return sol, spec.with_state(sol.t, sol.u.reshape(-1))


@pytest.mark.parametrize("method", [4, 6])
def test_reference_two_point_interpolation(method):
    sol, s = _solved(prob_tp(), method)
    u, du = O.interpolate(method, s, [0.001, 0.04, 0.0375, 0.5123])
    for k, t in enumerate([0.001, 0.04, 0.0375, 0.5123]):
        ex, dex = tp_exact(t)
        assert np.max(np.abs(u[k] - ex)) < 1e-6
        assert np.max(np.abs(du[k] - dex)) < 1e-6


@pytest.mark.parametrize("method", [4, 6])
def test_reference_multipoint_interpolation(method):
    for adaptive in (1, 0):
        sol, s = _solved(prob_mp(), method, adaptive=adaptive)
        u, _ = O.interpolate(method, s, [math.pi / 6, math.pi / 3])
        assert abs(u[0, 0] - 0.5) < 1e-6 and abs(u[0, 1] - math.cos(math.pi / 6)) < 1e-6
        assert abs(u[1, 0] - math.sin(math.pi / 3)) < 1e-6 and abs(u[1, 1] - 0.5) < 1e-6


@pytest.mark.parametrize("method", [4, 6])
def test_node_values_and_interval_ends(method):
    rng = np.random.default_rng(4)
    s = cases.random_state(cases.random_mesh(cases.bratu(2.0, N=12), rng), rng, 0.4)
    m, y = s.full_mesh(), s.full_guess().reshape(-1, 2)
    # the first node: interval 0 at tau = 0 gives y_0 exactly
    u, _ = O.interpolate(method, s, [m[0]])
    assert np.array_equal(u[0], y[0])
    # interior node t_i lies in interval i-1 at tau = 1: y_{i-1} + h sum b K = y_i - Phi_{i-1}
    phi = O.residual(method, s)[1:-1].reshape(-1, 2)  # [bca; Phi; bcb]
    u, _ = O.interpolate(method, s, m[1:])
    assert np.max(np.abs(u - (y[1:] - phi))) < 1e-13
    # a point just right of a node belongs to the next interval: continuity up to Phi
    u2, _ = O.interpolate(method, s, m[1:-1] + 1e-12)
    assert np.max(np.abs(u2 - y[1:-1])) < 1e-9


@pytest.mark.parametrize("method", [4, 6])
def test_derivative_is_the_derivative_of_the_value(method):
    rng = np.random.default_rng(5)
    s = cases.random_state(cases.random_mesh(cases.singpert(0.2, N=9), rng), rng, 0.5)
    m = s.full_mesh()
    t = (m[:-1] + m[1:]) / 2 + 0.1 * np.diff(m)
    eps = 1e-6 * np.diff(m)
    up, _ = O.interpolate(method, s, t + eps)
    um, _ = O.interpolate(method, s, t - eps)
    _, du = O.interpolate(method, s, t)
    fd = (up - um) / (2 * eps[:, None])
    assert np.max(np.abs(fd - du)) < 1e-6 * max(1.0, np.max(np.abs(du)))


@pytest.mark.parametrize("method,order", [(4, 4), (6, 6)])
def test_interpolant_converges_at_the_method_order(method, order):
    """Error of the extension at off-node points of the u'' = u problem under mesh halving."""
    errs = []
    for N in (10, 20):
        sol, s = _solved(prob_tp(1.0 / N), method, adaptive=0, abstol=1e-13)
        t = (np.arange(N) + 0.37) / N
        u, _ = O.interpolate(method, s, t)
        errs.append(max(np.max(np.abs(u[k] - tp_exact(tk)[0])) for k, tk in enumerate(t)))
    rate = math.log2(errs[0] / errs[1])
    assert rate > order - 0.6, (errs, rate)
