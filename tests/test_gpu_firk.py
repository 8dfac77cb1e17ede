"""FIRK collocation (Radau IIA, Lobatto IIIA/B/C; expanded form) on the GPU through the
C ABI, against the oracle (oracle/refcpu.cpp's FIRK restatement, pinned in
tests/test_firk_oracle.py).

Tolerances: residual 1e-12 (same arithmetic up to contraction and summation order),
Newton correction and defect 1e-9, dense output 1e-11, full solves within 1e-9 at
every node and stage (Newton tolerance 1e-12 on both sides) with the same mesh
sequence.
"""
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib
import pyoracle as O
import cases

pytestmark = pytest.mark.gpu

ALL = [101, 102, 103, 105, 107] + [k * 100 + s for k in (2, 3, 4) for s in (2, 3, 4, 5)]


def _name(code):
    return {1: "RadauIIa", 2: "LobattoIIIa", 3: "LobattoIIIb", 4: "LobattoIIIc"}[code // 100] + str(code % 100)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


def _opts(code, **kw):
    o = _abi.default_opts(code)
    o.nlsolve_stages = _abi.STAGE_NEWTON | _abi.STAGE_BACKTRACKING
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _oopts(code, **kw):
    kw.setdefault("nlsolve_stages", 3)
    return O.opts(code, **kw)


def _specs(rng):
    """Random states on four problem kinds: M = 2 two-point (Bratu), M = 3 two-point
    linear + tanh, M = 2 STANDARD rows at the ends, M = 1 scalar with a row at t0."""
    A3 = rng.standard_normal((3, 3)) * 0.5
    out = [
        cases.random_state(cases.bratu(2.0, N=7), rng),
        cases.random_state(cases.linear(A3, c=0.3, tspan=(0.0, 1.0), bc=([(0, 1.0), (2, -1.0)], [(1, 0.5)]), dt=0.1),
                           rng),
        cases.random_mesh(cases.random_state(cases.linear([[0, 1], [-1, 0]], bc=[(0.0, 0, 5.0), (5.0, 0, 0.0)],
                                                          dt=0.5, u0=[5.0, -3.5], standard=True), rng), rng),
        cases.random_state(scalar_tu0(0.125), rng),
    ]
    return out


def scalar_tu0(dt):
    import boundaryvaluediffeq_amd as bvp
    return bvp.extract(bvp.BVProblem(bvp.ScalarTUFunction(), [(0.0, 0, 1.0)], 0.5, (0.0, 1.0), []), dt=dt)


def _stages(code, spec, rng, scale=0.3):
    return rng.standard_normal((spec.N, code % 100, spec.M)) * scale


def _by_m(specs):
    groups = {}
    for s in specs:
        groups.setdefault(s.M, []).append(s)
    return groups.values()


@pytest.mark.parametrize("code", ALL, ids=_name)
def test_firk_residual_matches_oracle(ctx, code):
    _say("firk residual", _name(code))
    rng = np.random.default_rng(code)
    for group in _by_m(_specs(rng)):
        K = [_stages(code, s, rng) for s in group]
        got = _lib.firk_residual(ctx, code, group, K)
        for s, k, g in zip(group, K, got):
            ref = O.firk_residual(code, s, k)
            assert np.max(np.abs(g - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
        # no stages given: extend_y's
        got = _lib.firk_residual(ctx, code, group)
        for s, g in zip(group, got):
            ref = O.firk_residual(code, s)
            assert np.max(np.abs(g - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("code", ALL, ids=_name)
def test_firk_newton_step_matches_oracle(ctx, code):
    _say("firk newton step", _name(code))
    rng = np.random.default_rng(100 + code)
    for group in _by_m(_specs(rng)):
        K = [_stages(code, s, rng, 0.1) for s in group]
        got = _lib.firk_newton_step(ctx, code, group, K)
        for s, k, g in zip(group, K, got):
            ref = O.firk_newton_step(code, s, k)
            assert np.max(np.abs(g - ref)) <= 1e-9 * max(1.0, np.max(np.abs(ref))), (s.M, np.max(np.abs(g - ref)))


@pytest.mark.parametrize("code", ALL, ids=_name)
def test_firk_defect_matches_oracle(ctx, code):
    _say("firk defect", _name(code))
    rng = np.random.default_rng(200 + code)
    for group in _by_m(_specs(rng)):
        K = [_stages(code, s, rng, 0.1) for s in group]
        got = _lib.firk_defect(ctx, code, group, K)
        for s, k, (err, dn) in zip(group, K, got):
            er, dr = O.firk_defect(code, s, k)
            assert np.max(np.abs(err - er)) <= 1e-9 * max(1.0, np.max(np.abs(er)))
            assert abs(dn - dr) <= 1e-9 * max(1.0, dr)


def _compare(code, got, ref, specs, tol=1e-9, abstol=None):
    """Same return code, mesh size and outer iterations; nodes and stages within tol of
    the oracle's where the final meshes agree to 1e-12.  Else the GPU state must be the
    collocation solution on the GPU's mesh: the oracle's Newton correction of the
    expanded system there within tol, and with abstol the oracle's defect there within
    it.  (A redistribution after several passes amplifies the 1e-12 differences the
    Newton tolerance leaves: a 2e-16 relative change of abstol moves the oracle's own
    final mesh of Bratu lambda = 1.5 under LobattoIIIb3 by 1e-4.)"""
    for s, g, r in zip(specs, got, ref):
        assert g.retcode == r.retcode, (g.retcode, r.retcode)
        assert g.N == r.N and g.outer_iters == r.outer_iters, (g.N, r.N, g.outer_iters, r.outer_iters)
        if g.retcode != 0:
            continue
        dt = float(np.max(np.abs(g.t - r.t)))
        if dt <= 1e-12:
            eu = float(np.max(np.abs(g.u - r.u) / np.maximum(1.0, np.abs(r.u))))
            ek = float(np.max(np.abs(g.stages - r.stages) / np.maximum(1.0, np.abs(r.stages))))
            err = max(eu, ek)
        else:
            st = s.with_state(g.t, g.u.reshape(-1))
            d = O.firk_newton_step(code, st, g.stages)
            err = float(np.max(np.abs(d)) / max(1.0, np.max(np.abs(g.u))))
            if abstol is not None:
                assert O.firk_defect(code, st, g.stages)[1] <= abstol
        assert err <= tol, err


@pytest.mark.parametrize("code", ALL, ids=_name)
def test_firk_fixed_mesh_solve_matches_oracle(ctx, code):
    """adaptive = false: Newton on the given meshes (the affineness and convergence
    tests of the reference run this way)."""
    _say("firk fixed-mesh solve", _name(code))
    rng = np.random.default_rng(300 + code)
    specs = [cases.bratu(lam, N=int(n)) for lam, n in zip(rng.uniform(0.5, 3.0, 6), rng.integers(4, 40, 6))]
    specs.append(cases.linear([[0, 1], [0, 0]], bc=([(0, 5.0)], [(0, 0.0)]), dt=0.2, u0=[5.0, -3.5]))
    o = _opts(code, adaptive=0, newton_abstol=1e-12)
    got = _lib.solve_batch(ctx, o, specs)
    ref = [O.solve(_oopts(code, adaptive=0, newton_abstol=1e-12), s) for s in specs]
    _compare(code, got, ref, specs)
    assert all(g.newton_iters == r.newton_iters for g, r in zip(got, ref))
    # affineness (FIRK test/expanded/firk_affineness_tests.jl): exact linear solution
    u1 = got[-1].u[:, 0]
    assert np.max(np.abs(np.diff(u1) + 0.2)) + abs(u1[0] - 5) < 1e-10


ADAPTIVE = [102, 103, 105, 107, 202, 203, 204, 205, 303, 304, 305, 403, 404, 405]


@pytest.mark.parametrize("code", ADAPTIVE, ids=_name)
def test_firk_adaptive_solve_matches_oracle(ctx, code):
    """Adaptive solves (defect control, mesh_selector!, transfer) of a Bratu batch and a
    singular-perturbation instance that needs several mesh passes, in one launch."""
    _say("firk adaptive solve", _name(code))
    specs = [cases.bratu(lam, N=8) for lam in (0.5, 1.5, 3.0)] + [cases.singpert(0.02, N=16)]
    tol = 1e-6 if code % 100 == 2 else 1e-7  # the two-stage methods need > 4000 intervals at 1e-7
    o = _opts(code, abstol=tol, newton_abstol=1e-12, max_num_subintervals=4000)
    got = _lib.solve_batch(ctx, o, specs)
    ref = [O.solve(_oopts(code, abstol=tol, newton_abstol=1e-12, max_num_subintervals=4000), s) for s in specs]
    _compare(code, got, ref, specs, abstol=tol)
    assert got[-1].retcode == 0 and (got[-1].outer_iters > 1 or code == 107)
    for g in got:
        assert g.defect_norm <= tol


@pytest.mark.parametrize("code", [102, 103, 105, 107], ids=_name)
def test_firk_interpolation_reference_case(ctx, code):
    """FIRK test/expanded/firk_basic_tests.jl:60-116: u'' = u, u(0) = 1, u(1) = 0 at
    dt = 0.001, adaptive; sol(0.001) = [0.998687464, -1.312035941] and the analytic
    value and slope at t = 0.04 (atol 1e-6), solved and interpolated on the GPU; the
    dense output within 1e-11 of the oracle's at the GPU solution."""
    _say("firk interpolation", _name(code))
    import math
    spec = cases.linear([[0, 1], [1, 0]], tspan=(0.0, 1.0), bc=[(0.0, 0, 1.0), (1.0, 0, 0.0)], dt=0.001,
                        u0=[1.0, 0.0], standard=True)
    o = _opts(code)
    (g,) = _lib.solve_batch(ctx, o, [spec])
    assert g.retcode == 0
    st = spec.with_state(g.t, g.u.reshape(-1))
    tq = np.array([0.001, 0.04, 0.5, 0.99999])
    ((u, du),) = _lib.firk_interpolate(ctx, code, [st], [tq], [g.stages])
    assert np.max(np.abs(u[0] - [0.998687464, -1.312035941])) < 1e-6
    a, d = 1.0, 1 - math.exp(-2.0)
    t = 0.04
    ua = [(math.exp(-a * t) - math.exp((t - 2) * a)) / d, (-a * math.exp(-t * a) - a * math.exp((t - 2) * a)) / d]
    assert np.max(np.abs(u[1] - ua)) < 1e-6 and np.max(np.abs(du[1] - [ua[1], ua[0]])) < 1e-6
    ur, dur = O.firk_interpolate(code, st, tq, g.stages)
    assert np.max(np.abs(u - ur)) <= 1e-11 and np.max(np.abs(du - dur)) <= 1e-10


def test_firk_plan_results_and_limits(ctx):
    """The plan API serves FIRK batches (results with stages, repeated runs give the
    same bits); what the FIRK path does not serve is refused with BVP_ENOTSUP."""
    _say("firk plan and limits")
    specs = [cases.bratu(1.0 + 0.1 * i, N=10 + i) for i in range(8)]
    o = _opts(103, abstol=1e-8, newton_abstol=1e-12)
    b = _lib.Batch(ctx, o, specs)
    b.run()
    b.sync()
    r1 = b.results()
    b.run()
    b.sync()
    r2 = b.results()
    b.close()
    for x, y in zip(r1, r2):
        assert np.array_equal(x.u, y.u) and np.array_equal(x.stages, y.stages)
    ref = [O.solve(_oopts(103, abstol=1e-8, newton_abstol=1e-12), s) for s in specs]
    _compare(103, r1, ref, specs)
    with pytest.raises(_lib.BVPError):  # RadauIIa1 does not adapt
        _lib.solve_batch(ctx, _opts(101), [cases.bratu(1.0, N=8)])
    with pytest.raises(_lib.BVPError):  # TrustRegion stage: expanded form only
        o = _opts(103 + _abi.FIRK_NESTED)
        o.nlsolve_stages = _abi.STAGES_DEFAULT
        _lib.solve_batch(ctx, o, [cases.bratu(1.0, N=8)])
    with pytest.raises(_lib.BVPError):  # FIRK codes go through bvp_firk_*
        _lib.residual(ctx, 103, [cases.bratu(1.0, N=8)])


def test_user_api_solve_and_dense_output(ctx):
    """solve(prob, RadauIIa5(); dt) and sol(t), sol(t, Val{1}) through the Python mirror
    of the reference interface, for a FIRK and a MIRK method (FIRK
    test/expanded/firk_basic_tests.jl:60-116; MIRKt/Core/mirk_basic_tests.jl:173-313)."""
    _say("firk user api")
    import math
    import boundaryvaluediffeq_amd as bvp
    f = bvp.LinearTanhFunction(2)
    prob = bvp.BVProblem(f, [(0.0, 0, 1.0), (1.0, 0, 0.0)], [1.0, 0.0], (0.0, 1.0), [0.0, 0, 1, 1, 0])
    a, d = 1.0, 1 - math.exp(-2.0)
    ua = lambda t: [(math.exp(-a * t) - math.exp((t - 2) * a)) / d, (-a * math.exp(-t * a) - a * math.exp((t - 2) * a)) / d]
    for alg in (bvp.RadauIIa5(), bvp.LobattoIIIa4(), bvp.MIRK4()):
        sol = bvp.solve(prob, alg, dt=0.01)
        assert bvp.successful_retcode(sol), alg
        assert np.max(np.abs(sol(0.001) - [0.998687464, -1.312035941])) < 1e-6
        assert np.max(np.abs(sol(0.04) - ua(0.04))) < 1e-6
        assert np.max(np.abs(sol(0.04, deriv=1) - ua(0.04)[::-1])) < 1e-6
        assert abs(sol(0.5, idxs=0) - ua(0.5)[0]) < 1e-6
        if isinstance(alg, bvp.RadauIIa5):
            assert sol.stages.shape == (len(sol.t) - 1, 5, 2)


@pytest.mark.parametrize("code", [102, 103, 105, 204, 304, 404], ids=_name)
def test_firk_trust_region_matches_oracle(ctx, code):
    """The TrustRegion stage of the polyalgorithm on the expanded form: alone (nlsolve_stages
    = 4) on a fixed mesh and through the adaptive outer loop, and last after the
    NewtonRaphson stages with an iteration cap that makes every stage fail; Bratu instances
    and a linear problem, same return codes, iteration counts and stages as the oracle,
    nodes within 1e-9."""
    _say("firk trust region", _name(code))
    specs = [cases.bratu(lam, N=n) for lam, n in ((0.7, 12), (2.0, 20), (3.2, 16))]
    specs.append(cases.linear([[0, 1], [0, 0]], bc=([(0, 5.0)], [(0, 0.0)]), dt=0.2, u0=[5.0, -3.5]))
    for kw in (dict(adaptive=0, nlsolve_stages=4), dict(nlsolve_stages=4, abstol=1e-7),
               dict(adaptive=0, nlsolve_stages=7, max_newton_iters=2)):
        o = _opts(code, newton_abstol=1e-12, **kw)
        got = _lib.solve_batch(ctx, o, specs)
        ref = [O.solve(_oopts(code, newton_abstol=1e-12, **kw), s) for s in specs]
        _compare(code, got, ref, specs)
        for g, r in zip(got, ref):
            assert g.newton_iters == r.newton_iters and g.fallback_used == r.fallback_used
    # the pendulum's row at pi/4 (inside the span) in J'F and |Jg|^2: the trust region
    # alone on a fixed mesh takes some 80 dogleg steps there, whose accept/shrink decisions
    # round-off can flip, so the check is the solution and the return code
    spec = cases.pendulum(0.05)
    o = _opts(code, newton_abstol=1e-12, adaptive=0, nlsolve_stages=4)
    (g,) = _lib.solve_batch(ctx, o, [spec])
    r = O.solve(_oopts(code, newton_abstol=1e-12, adaptive=0, nlsolve_stages=4), spec)
    assert g.retcode == r.retcode == 0 and (g.fallback_used & 16)
    assert np.max(np.abs(g.u - r.u)) <= 1e-9 * max(1.0, np.max(np.abs(r.u)))


@pytest.mark.parametrize("code", [202, 203, 204, 205, 303, 304, 305, 403, 404, 405, 102, 103, 105, 107], ids=_name)
def test_reference_singular_bvp_case(ctx, code):
    """FIRK test/expanded/singular_bvp_tests.jl:36-97 as the reference runs it (the problem
    passes no singular_term, so the ODE is y1' = y2, y2' = -y1): y(0) = 1, y(1) = sin(1),
    u0 = [1, 0], dt = 0.01, default options; Success, and the oracle's return code, mesh size
    and nodes (within 1e-9)."""
    import math
    _say("firk singular-bvp case", _name(code))
    spec = cases.linear([[0, 1], [-1, 0]], tspan=(0.0, 1.0), bc=([(0, 1.0)], [(0, math.sin(1.0))]), dt=0.01,
                        u0=[1.0, 0.0])
    (g,) = _lib.solve_batch(ctx, _opts(code, newton_abstol=1e-12), [spec])
    r = O.solve(_oopts(code, newton_abstol=1e-12), spec)
    _compare(code, [g], [r], [spec])
    assert g.retcode == 0
