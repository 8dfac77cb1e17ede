"""GPU parity of the nonlinear polyalgorithm (NewtonRaphson -> NewtonRaphson +
BackTracking -> TrustRegion, Core/default_internal_solve.jl:31-45) and of the
per-result error estimate, through the C ABI, against the oracle.

The line search and the trust region take decisions on sums of squares whose
low digits differ between the two sides (summation order), so a trajectory may
part from the oracle's in its last digits; the checks are the return code, the
stage that succeeded, the branch reached and the solution at 1e-9 on the same
mesh.  Converging undamped Newton is deterministic to rounding, so its iteration
counts must agree exactly.
"""
import json
import pathlib
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib
import pyoracle as O
import cases

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "closed_forms.json").read_text())


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _opts(stages, method=4, **kw):
    o = _abi.default_opts(method)
    o.adaptive = 0
    o.abstol = 1e-10
    o.nlsolve_stages = stages
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _close(g, r, tol=1e-9):
    err = float(np.max(np.abs(g.u - r.u) / np.maximum(1.0, np.abs(r.u))))
    assert err <= tol, err
    return err


# (lam, c): start u = c; NR alone diverges (Unstable) in every case, the line search
# rescues the first two, the trust region the last three
CASES = [(1.0, 4.0), (3.4, 3.0), (1.0, 6.0), (3.4, 4.0), (3.4, 6.0)]


@pytest.mark.parametrize("lam,c", CASES)
def test_polyalgorithm_matches_oracle(ctx, lam, c):
    _say("polyalgorithm", lam, c)
    spec = cases.bratu(lam, N=64, u0=(c, 0.0))
    for stages in (1, 2, 4, 7):
        g = _lib.solve_batch(ctx, _opts(stages), [spec])[0]
        r = O.solve(O.opts(4, adaptive=0, abstol=1e-10, nlsolve_stages=stages), spec)
        assert g.retcode == r.retcode, (stages, g.retcode, r.retcode)
        assert (g.fallback_used & 17) == (r.fallback_used & 17), stages
        if stages == 1 and r.retcode == 0:
            # converging Newton is deterministic to rounding; a diverging one may reach a
            # non-finite residual one step apart
            assert g.newton_iters == r.newton_iters
        if r.retcode == 0:
            _close(g, r)
            up = [e for e in GOLDEN["bratu_upper"] if e["lam"] == lam][0]
            assert abs(g.u[32, 0] - up["u_half"]) < 1e-4 * up["u_half"]


@pytest.mark.parametrize("entry", GOLDEN["bratu"], ids=lambda e: f"lam{e['lam']}")
def test_each_stage_alone_lower_branch(ctx, entry):
    _say("stage alone", entry["lam"])
    spec = cases.bratu(entry["lam"], N=64)  # dt = 1/64 is exact: node 32 is t = 1/2
    for stages in (1, 2, 4):
        g = _lib.solve_batch(ctx, _opts(stages), [spec])[0]
        r = O.solve(O.opts(4, adaptive=0, abstol=1e-10, nlsolve_stages=stages), spec)
        assert g.retcode == r.retcode == 0
        _close(g, r)
        assert g.t[32] == 0.5
        assert abs(g.u[32, 0] - entry["u_half"]) < 2e-6 * max(1.0, entry["u_half"])


def test_polyalgorithm_batch_mixed_outcomes(ctx):
    """One batch whose instances end in different stages (and one with no solution)
    solves each like the oracle (the persistent kernel's stage paths side by side)."""
    _say("polyalgorithm batch")
    specs = [cases.bratu(lam, N=n, u0=(c, 0.0)) for lam, c in CASES for n in (40, 64)]
    specs += [cases.bratu(1.5, N=50), cases.bratu(3.7, N=24)]
    o = _opts(7, max_newton_iters=100)
    got = _lib.solve_batch(ctx, o, specs)
    for s, g in zip(specs, got):
        r = O.solve(O.opts(4, adaptive=0, abstol=1e-10, max_newton_iters=100), s)
        assert g.retcode == r.retcode
        if r.retcode == 0:
            _close(g, r)


def test_adaptive_solve_with_polyalgorithm(ctx):
    """Adaptive solves from starts where NR fails on the first mesh."""
    _say("adaptive polyalgorithm")
    specs = [cases.bratu(1.0, N=40, u0=(4.0, 0.0)), cases.bratu(3.4, N=48, u0=(3.0, 0.0))]
    o = _abi.default_opts(4)
    o.abstol = 1e-8
    o.newton_abstol = 1e-11
    got = _lib.solve_batch(ctx, o, specs)
    for s, g in zip(specs, got):
        r = O.solve(O.opts(4, abstol=1e-8, newton_abstol=1e-11), s)
        assert g.retcode == r.retcode == 0 and g.N == r.N
        _close(g, r)


@pytest.mark.parametrize("prec", [_abi.PREC_FP64, _abi.PREC_FP32_REFINE64])
def test_err_vs_fp64_is_the_oracle_correction(ctx, prec):
    """err_vs_fp64 is the max-norm of the FP64 Newton correction at the returned state;
    the oracle's own correction there gives the same number (to the correction's
    rounding), and asking for it leaves the solution bits unchanged."""
    _say("err_vs_fp64", prec)
    rng = np.random.default_rng(3)
    specs = [cases.bratu(float(l), N=int(n)) for l, n in zip(rng.uniform(0.5, 3.3, 12), rng.integers(64, 400, 12))]
    o = _abi.default_opts(4)
    o.abstol = 1e-8
    o.precision = prec
    plain = _lib.solve_batch(ctx, o, specs)
    o.report_err = 1
    rep = _lib.solve_batch(ctx, o, specs)
    for s, a, b in zip(specs, plain, rep):
        assert np.isnan(a.err_vs_fp64)
        assert np.array_equal(a.u, b.u) and a.N == b.N and a.newton_iters == b.newton_iters
        d = O.newton_step(4, s.with_state(b.t, b.u.reshape(-1))).reshape(-1, 2)
        ref = float(np.max(np.abs(d) / np.maximum(1.0, np.abs(b.u))))
        assert b.err_vs_fp64 >= 0
        assert abs(b.err_vs_fp64 - ref) <= 1e-3 * ref + 1e-15, (b.err_vs_fp64, ref)
