"""Pins the oracle's nonlinear solver: the default polyalgorithm
NewtonRaphson -> NewtonRaphson + BackTracking -> TrustRegion
(Core/default_internal_solve.jl:31-45), restated from the published algorithms
of its external stages (LineSearches.jl-style BackTracking, dogleg TrustRegion).

Each check is fixed by the mathematics rather than by the code: step lengths of
the line search computed by hand on polynomial merit functions (quadratic
interpolation, the [0.1, 0.5] clamps, the cubic model that is exact on a cubic),
dogleg steps whose length, direction and predicted reduction are checked against
their definitions, and Bratu starts from which undamped Newton fails where the
later stages must land on a closed-form root of the problem.
"""
import json
import math
import pathlib

import numpy as np
import pytest

import pyoracle as O
import cases

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "closed_forms.json").read_text())


# --------------------------------------------------------------- BackTracking
def test_backtracking_full_step_when_armijo_holds():
    # phi = 1 - 2a + 0.5 a^2: phi(1) = -0.5 <= 1 + 1e-4 * (-2): no reduction
    rc, a, trace = O.backtracking_poly([1.0, -2.0, 0.5])
    assert rc == 0 and a == 1.0 and list(trace) == [1.0]


def test_backtracking_quadratic_step_is_the_quadratic_minimiser():
    # phi = 1 - 2a + 4a^2: phi(1) = 3 fails Armijo; the quadratic through
    # phi(0), phi'(0), phi(1) is phi itself, minimised at a = 2/8 = 0.25
    rc, a, trace = O.backtracking_poly([1.0, -2.0, 4.0])
    assert rc == 0 and list(trace) == [1.0, 0.25] and a == 0.25


def test_backtracking_lower_clamp():
    # phi = 1 - 2a + 10a^2: minimiser 0.1 (exactly the lower clamp 0.1 * a);
    # phi = 1 - 2a + 20a^3: quadratic estimate 2 / (2 * 20) = 0.05 < 0.1 -> 0.1
    assert O.backtracking_poly([1.0, -2.0, 10.0])[2][1] == pytest.approx(0.1, abs=1e-15)
    rc, a, trace = O.backtracking_poly([1.0, -2.0, 0.0, 20.0])
    assert rc == 0 and list(trace) == [1.0, 0.1] and a == 0.1


def test_backtracking_upper_clamp():
    # phi(1) = 0.95 fails Armijo (phi(1) > 1 + 1e-4 * (-1000) = 0.9); the quadratic
    # estimate 1000 / (2 * 999.95) exceeds 0.5 and is cut to 0.5 * a
    rc, a, trace = O.backtracking_poly([1.0, -1000.0, 999.95])
    assert rc == 0 and list(trace) == [1.0, 0.5] and a == 0.5


def test_backtracking_cubic_step_is_exact_on_a_cubic():
    # phi = 1 - 2a + 5000 a^3: first reduction clamped to 0.1, phi(0.1) = 5.8 still
    # fails; the cubic through phi(0), phi'(0), phi(1), phi(0.1) is phi itself, whose
    # local minimiser is sqrt(2 / 15000)
    rc, a, trace = O.backtracking_poly([1.0, -2.0, 0.0, 5000.0])
    assert rc == 0 and len(trace) == 3
    assert trace[:2].tolist() == [1.0, 0.1]
    assert a == pytest.approx(math.sqrt(2.0 / 15000.0), rel=1e-12)


def test_backtracking_cubic_with_quadratic_part():
    # phi = 1 - a + 3a^2 + 400a^3 (a general cubic): phi(1) = 403 gives the quadratic
    # estimate 1/806, clamped to 0.1; phi(0.1) = 1.33 still fails, and the cubic model
    # through the four values is phi itself, minimised where 1200 a^2 + 6 a - 1 = 0
    rc, a, trace = O.backtracking_poly([1.0, -1.0, 3.0, 400.0])
    exact = (-6 + math.sqrt(36 + 4800)) / 2400
    assert rc == 0 and len(trace) == 3 and trace[1] == 0.1
    assert a == pytest.approx(exact, rel=1e-12)


def test_backtracking_nonfinite_trial_is_halved():
    # 1e308 (a^39 + a^40) overflows to inf at a = 1 and is finite at a = 0.5: the
    # non-finite trial is halved before any interpolation
    coef = np.zeros(41)
    coef[0], coef[1], coef[39], coef[40] = 1.0, -2.0, 1e308, 1e308
    rc, a, trace = O.backtracking_poly(coef)
    assert rc == 0
    assert trace[0] == 1.0 and trace[1] == 0.5
    assert 1.0 - 2.0 * a + 1e308 * (a ** 39 + a ** 40) <= 1.0 - 2e-4 * a


# --------------------------------------------------------------- dogleg
def test_dogleg_newton_step_inside_region():
    J = np.array([[2.0, 1.0], [0.0, 3.0]])
    F = np.array([1.0, -2.0])
    d, pred = O.dogleg(J, F, 10.0)
    assert np.allclose(d, np.linalg.solve(J, -F), rtol=1e-14)
    assert pred == pytest.approx(F @ F, rel=1e-14)


def test_dogleg_cut_steepest_descent():
    # J = I, F = (3, 4): Newton step (-3, -4) of length 5, g = F, Cauchy length
    # |g|^3 / |J g|^2 = 5 >= radius 2, so d = -(2/5) g; model reduction 25 - |F + d|^2 = 16
    d, pred = O.dogleg(np.eye(2), np.array([3.0, 4.0]), 2.0)
    assert np.allclose(d, [-1.2, -1.6], rtol=1e-15)
    assert pred == pytest.approx(16.0, rel=1e-14)


def test_dogleg_segment_between_cauchy_point_and_newton_step():
    J = np.diag([1.0, 10.0])
    F = np.array([1.0, 10.0])
    radius = 1.2
    d, pred = O.dogleg(J, F, radius)
    dn = np.linalg.solve(J, -F)
    g = J.T @ F
    dc = -(g @ g) / np.sum((J @ g) ** 2) * g
    assert np.linalg.norm(dc) < radius < np.linalg.norm(dn)
    assert np.linalg.norm(d) == pytest.approx(radius, rel=1e-13)
    # d lies on the segment from the Cauchy point to the Newton step
    tau = np.dot(d - dc, dn - dc) / np.dot(dn - dc, dn - dc)
    assert 0.0 < tau < 1.0
    assert np.allclose(d, dc + tau * (dn - dc), rtol=1e-12, atol=1e-14)
    # the predicted reduction is that of the linear model
    assert pred == pytest.approx(F @ F - np.sum((F + J @ d) ** 2), rel=1e-12)


def test_dogleg_predicted_reduction_random():
    rng = np.random.default_rng(7)
    for _ in range(20):
        n = 5
        J = rng.standard_normal((n, n)) + 3 * np.eye(n)
        F = rng.standard_normal(n)
        dn = np.linalg.solve(J, -F)
        for radius in (0.05, 0.3 * np.linalg.norm(dn), 0.8 * np.linalg.norm(dn)):
            d, pred = O.dogleg(J, F, radius)
            assert np.linalg.norm(d) == pytest.approx(radius, rel=1e-12)
            assert pred == pytest.approx(F @ F - np.sum((F + J @ d) ** 2), rel=1e-10, abs=1e-12)
            assert pred > 0


# --------------------------------------------------------------- stages end to end
def _upper(lam):
    return [e for e in GOLDEN["bratu_upper"] if e["lam"] == lam][0]


def _lower(lam):
    return [e for e in GOLDEN["bratu"] if e["lam"] == lam][0]


def _solve(spec, stages, method=4, **kw):
    return O.solve(O.opts(method, adaptive=0, abstol=1e-10, nlsolve_stages=stages, **kw), spec)


@pytest.mark.parametrize("lam,c", [(1.0, 4.0), (3.4, 3.0)])
def test_linesearch_rescues_a_diverging_newton(lam, c):
    """From u = c undamped Newton runs into a non-finite residual; the BackTracking
    stage lands on the upper Bratu branch, within the MIRK4 discretisation error of
    the closed form (h = 1/64)."""
    spec = cases.bratu(lam, N=64, u0=(c, 0.0))
    assert _solve(spec, 1).retcode == 3  # NewtonRaphson alone: Unstable
    sol = _solve(spec, 7)
    assert sol.retcode == 0 and sol.fallback_used & 1 and not sol.fallback_used & 16
    up = _upper(lam)
    assert abs(sol.u[32, 0] - up["u_half"]) < 1e-4 * up["u_half"]
    assert abs(sol.u[16, 0] - up["u_quarter"]) < 1e-4 * up["u_quarter"]


@pytest.mark.parametrize("lam,c", [(1.0, 6.0), (3.4, 4.0), (3.4, 6.0)])
def test_trust_region_rescues_when_linesearch_stalls(lam, c):
    """Undamped Newton diverges and the line search stalls (MaxIters); the
    TrustRegion stage, from the same start, reaches the upper branch."""
    spec = cases.bratu(lam, N=64, u0=(c, 0.0))
    assert _solve(spec, 1).retcode == 3
    assert _solve(spec, 2).retcode == 2
    sol = _solve(spec, 7)
    assert sol.retcode == 0 and sol.fallback_used & 16
    up = _upper(lam)
    assert abs(sol.u[32, 0] - up["u_half"]) < 1e-4 * up["u_half"]
    # and the same discrete root as plain Newton from a start near it
    near = cases.bratu(lam, N=64, guess=np.repeat(sol.u, 1, axis=0).reshape(-1) * 1.01)
    ref = _solve(near, 1)
    assert ref.retcode == 0
    assert np.max(np.abs(ref.u - sol.u)) < 1e-9 * np.max(np.abs(ref.u))


@pytest.mark.parametrize("entry", GOLDEN["bratu"], ids=lambda e: f"lam{e['lam']}")
def test_each_stage_alone_reaches_the_lower_branch(entry):
    """From the zero guess every stage run alone converges to the lower branch."""
    lam = entry["lam"]
    spec = cases.bratu(lam, N=64)
    sols = [_solve(spec, st) for st in (1, 2, 4)]
    for s in sols:
        assert s.retcode == 0
        assert abs(s.u[32, 0] - entry["u_half"]) < 2e-6 * max(1.0, entry["u_half"])
    for s in sols[1:]:
        assert np.max(np.abs(s.u - sols[0].u)) < 1e-9


def test_trust_region_fails_where_no_root_is_near():
    """lam above the fold (no solution): every stage fails, none claims success."""
    spec = cases.bratu(3.6, N=32)
    sol = _solve(spec, 7, max_newton_iters=200)
    assert sol.retcode != 0


def test_stage_mask_zero_iterations_returns_guess():
    """maxiters = 0 returns the guess untouched in every stage (MIRKt/Core/mirk_basic_tests.jl:678-720)."""
    spec = cases.bratu(1.0, N=16, u0=(0.3, -0.1))
    for st in (1, 2, 4, 7):
        sol = _solve(spec, st, max_newton_iters=0)
        assert sol.newton_iters == 0
        assert np.all(sol.u == np.array([0.3, -0.1]))
