"""GPU parity of the FIRK nested form (nested_nlsolve = true, FIRK/firk.jl:114-288) against
the oracle: residual, Newton step, defect, fixed-mesh and adaptive solves and dense output,
each through the C ABI.  The unknowns are the node values only; each interval's stages come
from an inner NewtonRaphson from k0 (u0, else the mean of the guess) stopped at
max|G| <= eps^(4/5).  Where an inner solve stops one iteration apart on the two sides the
stages differ by the last step's size (below 1e-13), hence the 1e-11 bounds on quantities
the expanded tests hold to 1e-12."""
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib
import pyoracle as O
import cases
from test_gpu_firk import ALL, _name, _specs, _by_m, _opts, _oopts

pytestmark = pytest.mark.gpu

NESTED = _abi.FIRK_NESTED


def _nname(code):
    return _name(code) + "_nested"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


@pytest.mark.parametrize("code", ALL, ids=_nname)
def test_nested_components_match_oracle(ctx, code):
    """Residual [bc; Phi] in the reference's layout, the Newton correction of the node
    values (the Jacobian through the implicit-function derivative of the inner solves)
    and the defect with the inner solves' stages, on random states of four problem kinds."""
    _say("firk nested components", _nname(code))
    nc = code + NESTED
    rng = np.random.default_rng(400 + code)
    for group in _by_m(_specs(rng)):
        res = _lib.firk_residual(ctx, nc, group)
        step = _lib.firk_newton_step(ctx, nc, group)
        dfc = _lib.firk_defect(ctx, nc, group)
        for s, g, d, (err, dn) in zip(group, res, step, dfc):
            assert g.shape == (s.M * (s.N + 1),)
            ref = O.firk_residual(nc, s)
            assert np.max(np.abs(g - ref)) <= 1e-11 * max(1.0, np.max(np.abs(ref))), (s.M, np.max(np.abs(g - ref)))
            rd = O.firk_newton_step(nc, s)
            assert np.max(np.abs(d - rd)) <= 1e-9 * max(1.0, np.max(np.abs(rd))), (s.M, np.max(np.abs(d - rd)))
            er, dr = O.firk_defect(nc, s)
            assert np.max(np.abs(err - er)) <= 1e-9 * max(1.0, np.max(np.abs(er)))
            assert abs(dn - dr) <= 1e-9 * max(1.0, dr)


def _compare(nc, got, ref, specs, tol=1e-9, abstol=None):
    """Same return code, mesh size and outer iterations; nodes within tol of the oracle's
    where the final meshes agree to 1e-12, else the oracle's nested Newton correction at
    the GPU's state within tol (and its defect there within abstol)."""
    for s, g, r in zip(specs, got, ref):
        assert g.retcode == r.retcode, (g.retcode, r.retcode)
        assert g.N == r.N and g.outer_iters == r.outer_iters, (g.N, r.N, g.outer_iters, r.outer_iters)
        if g.retcode != 0:
            continue
        if float(np.max(np.abs(g.t - r.t))) <= 1e-12:
            err = float(np.max(np.abs(g.u - r.u) / np.maximum(1.0, np.abs(r.u))))
            ek = float(np.max(np.abs(g.stages - r.stages) / np.maximum(1.0, np.abs(r.stages))))
            assert ek <= tol, ek
        else:
            st = s.with_state(g.t, g.u.reshape(-1))
            kst = None if nc >= NESTED else g.stages
            d = O.firk_newton_step(nc, st, kst)
            err = float(np.max(np.abs(d)) / max(1.0, np.max(np.abs(g.u))))
            if abstol is not None:
                assert O.firk_defect(nc, st, kst)[1] <= abstol
        assert err <= tol, err


@pytest.mark.parametrize("code", ALL, ids=_nname)
def test_nested_fixed_mesh_solve_matches_oracle(ctx, code):
    """adaptive = false on Bratu instances and a linear problem whose collocation solution
    is exact (FIRK test/nested/firk_basic_tests.jl's affineness case)."""
    _say("firk nested fixed-mesh solve", _nname(code))
    nc = code + NESTED
    rng = np.random.default_rng(500 + code)
    specs = [cases.bratu(lam, N=int(n)) for lam, n in zip(rng.uniform(0.5, 3.0, 5), rng.integers(4, 40, 5))]
    specs.append(cases.linear([[0, 1], [0, 0]], bc=([(0, 5.0)], [(0, 0.0)]), dt=0.2, u0=[5.0, -3.5]))
    o = _opts(nc, adaptive=0, newton_abstol=1e-12)
    got = _lib.solve_batch(ctx, o, specs)
    ref = [O.solve(_oopts(nc, adaptive=0, newton_abstol=1e-12), s) for s in specs]
    _compare(nc, got, ref, specs)
    assert all(g.newton_iters == r.newton_iters for g, r in zip(got, ref))
    u1 = got[-1].u[:, 0]
    assert np.max(np.abs(np.diff(u1) + 0.2)) + abs(u1[0] - 5) < 1e-10
    # the nested and expanded forms solve the same collocation equations
    ex = _lib.solve_batch(ctx, _opts(code, adaptive=0, newton_abstol=1e-12), specs)
    for g, e in zip(got, ex):
        assert g.retcode == e.retcode == 0
        assert np.max(np.abs(g.u - e.u)) <= 1e-9


@pytest.mark.parametrize("code", [102, 103, 105, 203, 204, 304, 403, 405], ids=_nname)
def test_nested_adaptive_solve_matches_oracle(ctx, code):
    """Adaptive nested solves (defect with inner-solved stages, mesh_selector!, the
    transfer with interp_eval!'s scaled step) of a Bratu batch and a singular-perturbation
    instance, in one launch."""
    _say("firk nested adaptive solve", _nname(code))
    nc = code + NESTED
    specs = [cases.bratu(lam, N=8) for lam in (0.5, 1.5, 3.0)] + [cases.singpert(0.02, N=16)]
    tol = 1e-6 if code % 100 == 2 else 1e-7
    o = _opts(nc, abstol=tol, newton_abstol=1e-12, max_num_subintervals=4000)
    got = _lib.solve_batch(ctx, o, specs)
    ref = [O.solve(_oopts(nc, abstol=tol, newton_abstol=1e-12, max_num_subintervals=4000), s) for s in specs]
    _compare(nc, got, ref, specs, abstol=tol)
    assert got[-1].retcode == 0 and got[-1].outer_iters > 1
    for g in got:
        assert g.defect_norm <= tol


@pytest.mark.parametrize("code", [103, 105], ids=_nname)
def test_nested_interpolation_reference_case(ctx, code):
    """FIRK test/nested/firk_basic_tests.jl: u'' = u, u(0) = 1, u(1) = 0 at dt = 0.001,
    adaptive; sol(0.001) = [0.998687464, -1.312035941] (atol 1e-6); the dense output
    (inner-solved stages) within 1e-11 of the oracle's at the GPU solution."""
    _say("firk nested interpolation", _nname(code))
    nc = code + NESTED
    spec = cases.linear([[0, 1], [1, 0]], tspan=(0.0, 1.0), bc=[(0.0, 0, 1.0), (1.0, 0, 0.0)], dt=0.001,
                        u0=[1.0, 0.0], standard=True)
    (g,) = _lib.solve_batch(ctx, _opts(nc), [spec])
    assert g.retcode == 0
    st = spec.with_state(g.t, g.u.reshape(-1))
    tq = np.array([0.001, 0.04, 0.5, 0.99999])
    ((u, du),) = _lib.firk_interpolate(ctx, nc, [st], [tq])
    assert np.max(np.abs(u[0] - [0.998687464, -1.312035941])) < 1e-6
    ur, dur = O.firk_interpolate(nc, st, tq)
    assert np.max(np.abs(u - ur)) <= 1e-11 and np.max(np.abs(du - dur)) <= 1e-10


def test_nested_user_api(ctx):
    """solve() with RadauIIa5(nested_nlsolve=True) and an inner tolerance, and sol(t) (FIRK
    test/nested/firk_basic_tests.jl's u'' = u case): the expanded solve's values."""
    _say("firk nested user api")
    import boundaryvaluediffeq_amd as bvp
    f = bvp.LinearTanhFunction(2)
    prob = bvp.BVProblem(f, [(0.0, 0, 1.0), (1.0, 0, 0.0)], [1.0, 0.0], (0.0, 1.0), [0.0, 0, 1, 1, 0])
    alg = bvp.RadauIIa5(nested_nlsolve=True, nested_nlsolve_kwargs={"abstol": 1e-12})
    sol = bvp.solve(prob, alg, dt=0.01)
    ex = bvp.solve(prob, bvp.RadauIIa5(), dt=0.01)
    assert bvp.successful_retcode(sol) and bvp.successful_retcode(ex)
    assert np.max(np.abs(sol(0.001) - [0.998687464, -1.312035941])) < 1e-6
    for t in (0.04, 0.37, 0.9):
        assert np.max(np.abs(sol(t) - ex(t))) <= 1e-8
    assert sol.stages.shape == (len(sol.t) - 1, 5, 2)


# --------------------------------------------------------------- interior STANDARD rows
@pytest.mark.parametrize("code", [102, 103, 105, 204, 304, 404], ids=_name)
@pytest.mark.parametrize("nested", [False, True], ids=["expanded", "nested"])
def test_interior_row_components_match_oracle(ctx, code, nested):
    """A row at pi/4 (inside an interval, dt = 0.07; on a node, dt = pi/64) and one at
    pi/2 (the README pendulum, FIRK test/expanded/firk_basic_tests.jl:5-22): residual,
    Newton step (the row's linearisation through EvalSol, the stage corrections eliminated
    through the interval record) and defect against the oracle."""
    import math
    _say("firk interior rows components", _name(code), nested)
    c = code + (NESTED if nested else 0)
    rng = np.random.default_rng(600 + code)
    for dt in (0.07, math.pi / 64):
        spec = cases.random_state(cases.pendulum(dt), rng, scale=0.1)
        K = None if nested else [rng.standard_normal((spec.N, code % 100, 2)) * 0.1]
        (g,) = _lib.firk_residual(ctx, c, [spec], K)
        ref = O.firk_residual(c, spec, None if nested else K[0])
        assert np.max(np.abs(g - ref)) <= 1e-11 * max(1.0, np.max(np.abs(ref))), np.max(np.abs(g - ref))
        (d,) = _lib.firk_newton_step(ctx, c, [spec], K)
        rd = O.firk_newton_step(c, spec, None if nested else K[0])
        assert np.max(np.abs(d - rd)) <= 1e-9 * max(1.0, np.max(np.abs(rd))), np.max(np.abs(d - rd))


@pytest.mark.parametrize("code", [102, 103, 105, 107, 203, 204, 304, 305, 403, 404], ids=_name)
@pytest.mark.parametrize("nested", [False, True], ids=["expanded", "nested"])
def test_reference_pendulum_solves_match_oracle(ctx, code, nested):
    """The reference's FIRK pendulum runs (test/expanded/firk_basic_tests.jl:5-51,
    test/nested/firk_pendulum_tests.jl): NewtonRaphson, dt = 0.005 (0.05 for Radau IIA 5
    and 7), adaptive, Newton tolerance 1e-12; same return code, mesh size, outer and
    Newton iterations as the oracle, nodes within 1e-9."""
    _say("firk pendulum solve", _name(code), nested)
    c = code + (NESTED if nested else 0)
    dt = 0.05 if code in (105, 107) else 0.005
    spec = cases.pendulum(dt)
    o = _opts(c, nlsolve_stages=_abi.STAGE_NEWTON, newton_abstol=1e-12)
    (g,) = _lib.solve_batch(ctx, o, [spec])
    r = O.solve(_oopts(c, nlsolve_stages=1, newton_abstol=1e-12), spec)
    _compare(c, [g], [r], [spec])
    assert g.retcode == 0 and g.newton_iters == r.newton_iters


# --------------------------------------------------------------- state dimension 6
@pytest.mark.parametrize("code", [103, 105, 204, 105 + NESTED], ids=lambda c: _nname(c - NESTED) if c > NESTED
                         else _name(c))
def test_swirling_flow_matches_oracle(ctx, code):
    """FIRK test/expanded/firk_basic_tests.jl:159-188: Swirling Flow III (M = 6, eps = 0.01,
    zero guess, dt = 0.01), adaptive at abstol 1e-6; same return code, mesh size and
    iterations as the oracle, nodes within 1e-9."""
    _say("firk swirling flow", code)
    spec = cases.swirl(0.01, dt=0.01)
    o = _opts(code, newton_abstol=1e-12)
    (g,) = _lib.solve_batch(ctx, o, [spec])
    r = O.solve(_oopts(code, newton_abstol=1e-12), spec)
    _compare(code, [g], [r], [spec], abstol=1e-6)
    assert g.retcode == 0 and g.newton_iters == r.newton_iters
