"""Parity of the HIP library (through its C ABI) with the oracle.

Tolerances (FP64 on both sides; the GPU contracts a*b+c into FMA and sums in a
different order, so agreement is to rounding, amplified by conditioning):
  residual, Jacobian blocks ........ 1e-12 relative to the block scale
  Newton correction ................ 1e-9 relative (conditioning of J)
  defect vectors / norm ............ 1e-9 relative
  full solves ...................... every mesh point within 1e-9 relative
                                     (scale max(1, |u|)), same mesh size sequence;
                                     node positions within 1e-10 (they follow the
                                     defect, whose low digits are rounding noise)
"""
import math
import sys

import numpy as np
import pytest

import boundaryvaluediffeq_amd as bvp
from boundaryvaluediffeq_amd import _abi, _lib
import pyoracle as O
import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


FAMS = {
    "bratu": lambda: cases.bratu(2.5, N=37),
    "pendulum": lambda: cases.pendulum(dt=0.05),
    "singpert": lambda: cases.singpert(0.05, N=70),
    "turning": lambda: cases.turning(0.05, dt=0.02),
    "linear3": lambda: cases.linear([[0.3, 1.0, -0.2], [-1.0, 0.1, 0.4], [0.5, -0.3, -0.6]], c=0.7,
                                    tspan=(0, 1), bc=([(0, 1.0), (2, -1.0)], [(1, 0.5)]), dt=1 / 130),
    "scalar": lambda: cases.linear([[0.5]], tspan=(0, 1), bc=([(0, 1.0)], []), dt=1 / 9),
}


def _specs(name, rng, k=3):
    out = []
    for i in range(k):
        s = cases.random_state(FAMS[name](), rng, 0.3)
        if i == 1:
            s = cases.random_mesh(s, rng)
        out.append(s)
    return out


@pytest.mark.parametrize("method", [4, 6])
@pytest.mark.parametrize("name", sorted(FAMS))
def test_residual_parity(ctx, method, name):
    _say("residual", name, method)
    rng = np.random.default_rng(1)
    specs = _specs(name, rng)
    got = _lib.residual(ctx, method, specs)
    for s, g in zip(specs, got):
        ref = O.residual(method, s)
        scale = max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(g - ref)) <= 1e-12 * scale


@pytest.mark.parametrize("method", [4, 6])
@pytest.mark.parametrize("name", sorted(FAMS))
def test_jacobian_blocks_parity(ctx, method, name):
    _say("jacobian", name, method)
    rng = np.random.default_rng(2)
    specs = _specs(name, rng)
    got = _lib.jacobian_blocks(ctx, method, specs)
    for s, (S, R, node) in zip(specs, got):
        Sr, Rr, nr = O.jacobian_blocks(method, s)
        scale = max(1.0, np.max(np.abs(Sr)), np.max(np.abs(Rr)))
        assert np.max(np.abs(S - Sr)) <= 1e-12 * scale
        assert np.max(np.abs(R - Rr)) <= 1e-12 * scale
        assert np.array_equal(node, nr)


@pytest.mark.parametrize("method", [4, 6])
@pytest.mark.parametrize("name", sorted(FAMS))
def test_newton_step_parity(ctx, method, name):
    _say("newton_step", name, method)
    rng = np.random.default_rng(3)
    specs = _specs(name, rng)
    got = _lib.newton_step(ctx, method, specs)
    for s, d in zip(specs, got):
        ref = O.newton_step(method, s)
        assert np.max(np.abs(d - ref)) <= 1e-9 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("N", [1, 2, 5, 63, 64, 65, 127, 200, 1000])
def test_newton_step_chunking_edges(ctx, N):
    """Chunk partitions with N below, at and above the wave width; a TWO_POINT case and
    a STANDARD case whose BC point pi/4 is a node (N even) or inside an interval (N odd)."""
    _say("newton_step N", N)
    rng = np.random.default_rng(N)
    two = cases.random_state(cases.bratu(1.5, N=N), rng, 0.2)
    pend = cases.pendulum(dt=0.05)
    std = pend.with_state(bvp.uniform_mesh(0.0, math.pi / 2, N), rng.standard_normal(2 * (N + 1)) * 0.1)
    for s in (two, std):
        got = _lib.newton_step(ctx, 4, [s])[0]
        ref = O.newton_step(4, s)
        assert np.max(np.abs(got - ref)) <= 1e-9 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("method", [4, 6])
@pytest.mark.parametrize("name", sorted(FAMS))
def test_defect_parity(ctx, method, name):
    _say("defect", name, method)
    rng = np.random.default_rng(4)
    specs = _specs(name, rng)
    got = _lib.defect(ctx, method, specs)
    for s, (err, dn) in zip(specs, got):
        er, dr = O.defect(method, s)
        assert np.max(np.abs(err - er)) <= 1e-9 * max(1.0, np.max(np.abs(er)))
        assert abs(dn - dr) <= 1e-9 * max(1.0, dr)


@pytest.mark.parametrize("mode", [_abi.TRANSFER_REFERENCE, _abi.TRANSFER_CLEAN])
@pytest.mark.parametrize("method", [4, 6])
def test_mesh_select_parity(ctx, mode, method):
    """Defects well above rounding (coarse meshes), so that the selected meshes are
    determined by the algorithm and not by the noise in the last digits."""
    _say("mesh_select", mode, method)
    specs = []
    bratu_cases = ((1.0, 6), (3.3, 12), (3.3, 25), (2.0, 9)) if method == 4 else ((3.3, 5), (3.0, 7), (2.5, 4))
    sp_cases = ((1e-2, 20), (3e-2, 8)) if method == 4 else ((3e-2, 8),)
    for lam, N in bratu_cases:
        s = cases.bratu(lam, N=N)
        sol = O.solve(O.opts(method, adaptive=0, abstol=1e-12), s)
        specs.append(s.with_state(sol.t, sol.u.reshape(-1)))
    for eps, N in sp_cases:
        s = cases.singpert(eps, N=N)
        sol = O.solve(O.opts(method, adaptive=0, abstol=1e-12), s)
        specs.append(s.with_state(sol.t, sol.u.reshape(-1)))
    o = _abi.default_opts(method)
    o.abstol = 1e-8
    o.transfer_mode = mode
    oo = O.opts(method, abstol=1e-8, transfer_mode=mode)
    got = _lib.mesh_select(ctx, o, specs)
    for s, (info, n, m, y) in zip(specs, got):
        ri, rn, rm, ry = O.mesh_select(oo, s)
        assert info == ri and n == rn
        if info == 0:
            assert n != s.N
            # The defect is a difference of O(1) quantities, so its relative rounding
            # noise is ~1e-15/defect; the equidistributed nodes move by up to ~100x that
            # (the smallest interval defect sets the scale).
            e, _ = O.defect(method, s)
            tol_t = max(1e-10, 1e-13 / np.min(np.max(np.abs(e), axis=1)))
            assert np.max(np.abs(m - rm)) <= tol_t
            assert np.max(np.abs(y - ry)) <= (1e-9 + 10 * tol_t) * max(1.0, np.max(np.abs(ry)))


def _compare_solutions(g, r, tol=1e-9):
    assert g.retcode == r.retcode
    assert g.N == r.N
    assert np.max(np.abs(g.t - r.t)) <= 1e-10 * max(1.0, np.max(np.abs(r.t)))
    scale = np.maximum(1.0, np.abs(r.u))
    err = np.max(np.abs(g.u - r.u) / scale)
    assert err <= tol, err
    return err


@pytest.mark.parametrize("method", [4, 6])
def test_solve_bratu_batch_parity(ctx, method):
    """Config-2 family at small size: uneven meshes, abstol 1e-8, adaptive."""
    _say("solve bratu", method)
    rng = np.random.default_rng(20)
    lam = rng.uniform(0.1, 3.4, 48)
    Ns = rng.integers(64, 300, 48)
    specs = [cases.bratu(float(l), N=int(n)) for l, n in zip(lam, Ns)]
    o = _abi.default_opts(method)
    o.abstol = 1e-8
    got = _lib.solve_batch(ctx, o, specs)
    oo = O.opts(method, abstol=1e-8)
    for s, g in zip(specs, got):
        r = O.solve(oo, s)
        _compare_solutions(g, r)
        assert g.newton_iters == r.newton_iters and g.outer_iters == r.outer_iters
        assert g.retcode == 0 and g.defect_norm <= 1e-8


def test_solve_pendulum_readme(ctx):
    """Config 1: README pendulum, STANDARD BCs through the continuous extension."""
    _say("solve pendulum")
    s = cases.pendulum(0.05)
    for method in (4, 6):
        o = _abi.default_opts(method)
        g = _lib.solve_batch(ctx, o, [s])[0]
        r = O.solve(O.opts(method), s)
        _compare_solutions(g, r)
        assert g.newton_iters == r.newton_iters


def test_solve_public_api_pendulum(ctx):
    prob = bvp.BVProblem(bvp.PendulumFunction(), [(math.pi / 4, 0, -math.pi / 2), (math.pi / 2, 0, math.pi / 2)],
                         [math.pi / 2, math.pi / 2], (0.0, math.pi / 2), [9.81])
    sol = bvp.solve(prob, bvp.MIRK4(), dt=0.05)
    assert bvp.successful_retcode(sol)
    assert abs(sol.u[-1][0] - math.pi / 2) < 1e-6


def _parity_on_same_mesh(spec, g, method, tol=1e-9):
    """The GPU's final solution is within `tol` (relative, scale max(1,|u|)) of the
    oracle's discrete solution on the same mesh: the oracle's Newton correction at the
    GPU state (J^-1 F of the oracle's own residual) is that distance to first order."""
    s = spec.with_state(g.t, g.u.reshape(-1))
    d = O.newton_step(method, s).reshape(-1, spec.M)
    err = float(np.max(np.abs(d) / np.maximum(1.0, np.abs(g.u))))
    assert err <= tol, err
    return err


def _tight(method, **kw):
    # 1e-10: tight, yet above the rounding floor of the stiff cases (at 1e-12 the
    # eps=3e-4 problem stalls and both sides fall into the halving branch)
    o = _abi.default_opts(method)
    o.newton_abstol = 1e-10
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def test_solve_singpert_adaptive(ctx):
    """Config-5 family: adaptive from N=32, growing per instance.  Several
    redistributions amplify rounding differences in the node positions, so parity
    is checked on the same mesh (B6 of the survey: tight Newton on both sides)."""
    _say("solve singpert")
    eps = [1e-1, 3e-2, 1e-2, 3e-3, 1e-3, 3e-4]
    specs = [cases.singpert(e, N=32) for e in eps]
    got = _lib.solve_batch(ctx, _tight(4, max_num_subintervals=16384), specs)
    oo = O.opts(4, max_num_subintervals=16384, newton_abstol=1e-10)
    for s, g in zip(specs, got):
        r = O.solve(oo, s)
        assert g.retcode == r.retcode == 0 and g.defect_norm <= 1e-6
        assert abs(g.N - r.N) <= max(4, r.N // 5)  # adaptive sequences are rounding-sensitive
        _parity_on_same_mesh(s, g, 4)


def test_solve_turning_point_reference_case(ctx):
    _say("solve turning")
    from scipy.special import erf
    s = cases.turning(0.001, dt=0.01)
    g = _lib.solve_batch(ctx, _tight(4), [s])[0]
    r = O.solve(O.opts(4, newton_abstol=1e-10), s)
    assert g.retcode == r.retcode == 0 and g.defect_norm <= 1e-6
    assert abs(g.N - r.N) <= max(4, r.N // 5)  # adaptive sequences are rounding-sensitive
    _parity_on_same_mesh(s, g, 4)
    exact = np.cos(np.pi * g.t) + erf(g.t / math.sqrt(0.002)) / erf(1 / math.sqrt(0.002))
    assert np.max(np.abs(g.u[:, 0] - exact)) < 1e-4


def test_solve_clean_transfer_and_nonadaptive(ctx):
    _say("solve modes")
    specs = [cases.bratu(3.0, N=70), cases.singpert(1e-2, N=32)]
    for adaptive in (0, 1):
        for mode in (0, 1):
            o = _tight(4, adaptive=adaptive, transfer_mode=mode, max_num_subintervals=16384)
            got = _lib.solve_batch(ctx, o, specs[:1]) + _lib.solve_batch(ctx, o, specs[1:])
            oo = O.opts(4, adaptive=adaptive, transfer_mode=mode, max_num_subintervals=16384, newton_abstol=1e-10)
            for s, g in zip(specs, got):
                r = O.solve(oo, s)
                assert g.retcode == r.retcode
                assert abs(g.N - r.N) <= max(4, r.N // 5)  # adaptive sequences are rounding-sensitive
                if g.N == r.N and np.max(np.abs(g.t - r.t)) <= 1e-12:
                    _compare_solutions(g, r)
                _parity_on_same_mesh(s, g, 4)


def test_failure_paths(ctx):
    """maxiters=0 returns the guess; the subinterval cap gives Failure; an infeasible
    Bratu (lambda > lambda_c) fails like the oracle."""
    _say("failure paths")
    rng = np.random.default_rng(5)
    s = cases.random_state(cases.bratu(1.0, N=10), rng)
    o = _abi.default_opts(4)
    o.adaptive = 0
    o.max_newton_iters = 0
    g = _lib.solve_batch(ctx, o, [s])[0]
    assert np.array_equal(g.u.reshape(-1), s.guess) and g.retcode == _abi.RET_MAXITERS
    o = _abi.default_opts(4)
    o.max_num_subintervals = 40
    s2 = cases.singpert(1e-3, N=32)
    g = _lib.solve_batch(ctx, o, [s2])[0]
    r = O.solve(O.opts(4, max_num_subintervals=40), s2)
    assert g.retcode == r.retcode != 0 and g.N == r.N
    o = _abi.default_opts(4)
    o.max_newton_iters = 50
    s3 = cases.bratu(4.0, N=16)
    g = _lib.solve_batch(ctx, o, [s3])[0]
    r = O.solve(O.opts(4, max_newton_iters=50), s3)
    assert g.retcode == r.retcode != 0 and g.N == r.N and g.outer_iters == r.outer_iters


def test_linesearch_restarts_match_oracle(ctx):
    """Newton capped at one step: the plain attempt fails on every mesh, so the
    line-search attempt restarts -- from the initial guess on the first mesh and
    from zeros after each halving -- and the final state, mesh and counts match
    the oracle's, which restarts from a copy of the starting state."""
    _say("line-search restarts")
    for lam, N in ((1.0, 8), (2.5, 12)):
        s = cases.bratu(lam, N=N)
        o = _abi.default_opts(4)
        o.abstol = 1e-8
        o.max_newton_iters = 1
        o.max_num_subintervals = 64
        g = _lib.solve_batch(ctx, o, [s])[0]
        r = O.solve(O.opts(4, abstol=1e-8, max_newton_iters=1, max_num_subintervals=64), s)
        assert g.retcode == r.retcode and g.N == r.N and g.outer_iters == r.outer_iters
        assert g.newton_iters == r.newton_iters
        assert np.max(np.abs(g.u - r.u) / np.maximum(1.0, np.abs(r.u))) <= 1e-9


def test_solve_linear_tanh_small_M(ctx):
    _say("solve linear M=3,4")
    rng = np.random.default_rng(8)
    for M in (3, 4):
        A = rng.standard_normal((M, M))
        half = M // 2
        specs = [cases.linear(A, c=0.1, tspan=(0, 1), bc=([(k, 0.3 * k) for k in range(half)],
                                                             [(k, -0.2) for k in range(half, M)]), dt=1 / n)
                 for n in (16, 90)]
        for method in (4, 6):
            o = _abi.default_opts(method)
            got = _lib.solve_batch(ctx, o, specs)
            for s, g in zip(specs, got):
                _compare_solutions(g, O.solve(O.opts(method), s))
