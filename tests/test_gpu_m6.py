"""State dimension M = 6 on the wave-per-instance kernels, through the C ABI, against
the oracle: the reference's Swirling Flow III problem (MIRKt/Core/mirk_basic_tests.jl:
315-345) and a 6 x 6 linear + tanh system.

Tolerances as for M <= 4: residual and Jacobian blocks 1e-12, Newton correction and
defect 1e-9, full solves within 1e-9 at every mesh point with the same mesh sequence.
"""
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib
import pyoracle as O
import cases

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


def _lt6(dt=0.1, seed=2):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((6, 6)) * 0.5
    return cases.linear(A, c=0.3, tspan=(0, 1), bc=([(0, 1.0), (1, -1.0), (2, 0.5)], [(3, 0.2), (4, 0.0), (5, 1.0)]),
                        dt=dt)


def _specs():
    rng = np.random.default_rng(4)
    out = [cases.random_state(cases.swirl(0.05, 0.05), rng, 0.3), cases.random_state(_lt6(0.07), rng, 0.4)]
    out.append(cases.random_mesh(cases.random_state(cases.swirl(0.01, 0.02), rng, 0.2), rng))
    return out


@pytest.mark.parametrize("method", [4, 6])
def test_m6_residual_blocks_newton_defect(ctx, method):
    _say("M=6 components", method)
    specs = _specs()
    for s, g in zip(specs, _lib.residual(ctx, method, specs)):
        ref = O.residual(method, s)
        assert np.max(np.abs(g - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    for s, (S, R, node) in zip(specs, _lib.jacobian_blocks(ctx, method, specs)):
        Sr, Rr, nr = O.jacobian_blocks(method, s)
        sc = max(1.0, np.max(np.abs(Sr)), np.max(np.abs(Rr)))
        assert np.max(np.abs(S - Sr)) <= 1e-12 * sc and np.max(np.abs(R - Rr)) <= 1e-12 * sc
        assert np.array_equal(node, nr)
    for s, d in zip(specs, _lib.newton_step(ctx, method, specs)):
        ref = O.newton_step(method, s)
        assert np.max(np.abs(d - ref)) <= 1e-9 * max(1.0, np.max(np.abs(ref)))
    for s, (err, dn) in zip(specs, _lib.defect(ctx, method, specs)):
        er, dr = O.defect(method, s)
        assert np.max(np.abs(err - er)) <= 1e-9 * max(1.0, np.max(np.abs(er)))
        assert abs(dn - dr) <= 1e-9 * max(1.0, dr)


def _compare(got, ref, specs, method=4, tol=1e-9):
    """Same return code, mesh size and outer iterations; the solution within tol of
    the oracle's where the final meshes agree to 1e-12, else (a redistribution
    rounded its nodes differently, by 1e-11 here) the oracle's own Newton correction
    at the GPU state on the GPU's mesh within tol."""
    for s, g, r in zip(specs, got, ref):
        assert g.retcode == r.retcode == 0, (g.retcode, r.retcode)
        assert g.N == r.N and g.outer_iters == r.outer_iters, (g.N, r.N)
        dt = float(np.max(np.abs(g.t - r.t)))
        assert dt <= 1e-9, dt
        if dt <= 1e-12:
            err = float(np.max(np.abs(g.u - r.u) / np.maximum(1.0, np.abs(r.u))))
        else:
            d = O.newton_step(method, s.with_state(g.t, g.u.reshape(-1))).reshape(-1, s.M)
            err = float(np.max(np.abs(d) / np.maximum(1.0, np.abs(g.u))))
        assert err <= tol, err


@pytest.mark.parametrize("method", [4, 6])
def test_swirling_flow_reference_case(ctx, method):
    """solve(prob, MIRK4(); dt = 0.01, abstol = 1e-4) of the reference's test (and
    MIRK6), Newton tolerance tightened on both sides."""
    _say("swirling flow", method)
    spec = cases.swirl(0.01, 0.01)
    o = _abi.default_opts(method)
    o.abstol = 1e-4
    o.newton_abstol = 1e-10
    got = _lib.solve_batch(ctx, o, [spec])
    ref = [O.solve(O.opts(method, abstol=1e-4, newton_abstol=1e-10), spec)]
    _compare(got, ref, [spec], method)
    assert got[0].defect_norm <= 1e-4


def test_swirling_flow_adaptive_batch(ctx):
    """Adaptive refinement at M = 6 in one batch (meshes grow from 10 to 147 and 80
    intervals in the oracle), with the 6 x 6 linear + tanh system beside them."""
    _say("swirling flow adaptive")
    specs = [cases.swirl(3e-3, 0.1), cases.swirl(1e-2, 0.1), _lt6()]
    o = _abi.default_opts(4)
    o.abstol = 1e-6
    o.newton_abstol = 1e-11
    o.max_num_subintervals = 5000
    got = _lib.solve_batch(ctx, o, specs)
    ref = [O.solve(O.opts(4, abstol=1e-6, newton_abstol=1e-11, max_num_subintervals=5000), s) for s in specs]
    assert ref[0].N > 100 and ref[1].N > 50
    _compare(got, ref, specs)


def test_m6_mixed_request_runs_fp64_records(ctx):
    """FP32 records are not built at M = 6: the request runs FP64 records (the plan
    reports it) and gives the FP64 bits."""
    _say("M=6 precision")
    specs = [cases.swirl(1e-2, 0.05)]
    o = _abi.default_opts(4)
    o.abstol = 1e-6
    a = _lib.solve_batch(ctx, o, specs)
    o.precision = _abi.PREC_FP32_REFINE64
    b = _lib.Batch(ctx, o, specs)
    assert b.precision() == _abi.PREC_FP64
    b.run()
    b.sync()
    g = b.results()
    b.close()
    assert np.array_equal(a[0].u, g[0].u) and a[0].N == g[0].N


def test_m6_dense_output_and_limits(ctx):
    _say("M=6 dense output")
    sol = O.solve(O.opts(4, abstol=1e-6, newton_abstol=1e-11), cases.swirl(1e-2, 0.05))
    spec = cases.swirl(1e-2, 0.05).with_state(sol.t, sol.u.reshape(-1))
    tq = np.linspace(0.0, 1.0, 37)[1:-1] + 1e-3
    (u, du), = _lib.interpolate(ctx, 4, [spec], [tq])
    ur, dur = O.interpolate(4, spec, tq)
    assert np.max(np.abs(u - ur)) <= 1e-11 * max(1.0, np.max(np.abs(ur)))
    assert np.max(np.abs(du - dur)) <= 1e-10 * max(1.0, np.max(np.abs(dur)))
    o = _abi.default_opts(4)
    o.controller = _abi.CONTROL_GLOBAL
    with pytest.raises(_lib.BVPError):
        _lib.solve_batch(ctx, o, [cases.swirl(1e-2, 0.1)])
