"""GPU parity of the global-error controllers (GlobalErrorControl with the HO and RE
estimates, SequentialErrorControl, HybridErrorControl; Core/calc_errors.jl:54-163,
MIRK/adaptivity.jl:77-243, 463-567) through the C ABI, against the oracle.

The estimate alone (bvp_mirk_global_error) runs a nested Newton solve on the halved
mesh (RE) or with MIRK6 (HO) from the same start on both sides, so both end on the
same discrete solution to the Newton tolerance: the per-interval errors agree to a
small multiple of that tolerance.  Full adaptive solves must take the same mesh
sequence (same N, outer iterations and return code) and end within 1e-9 of the
oracle at every node.
"""
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib
from boundaryvaluediffeq_amd import MIRK4, GlobalErrorControl, SequentialErrorControl, HybridErrorControl, REErrorControl
from boundaryvaluediffeq_amd.solve import make_opts
import pyoracle as O
import cases
from test_oracle_dense import prob_tp, prob_mp

pytestmark = pytest.mark.gpu

NTOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


def _gopts(method, controller, gm, **kw):
    o = _abi.default_opts(method)
    o.newton_abstol = NTOL
    o.controller = controller
    o.global_method = gm
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _states():
    """Instances at converged plain solves and at perturbed states (the nested solve
    then starts further from its root)."""
    rng = np.random.default_rng(5)
    out = []
    for spec, method in [(cases.bratu(1.3, N=24), 4), (cases.bratu(2.9, N=41), 6), (prob_tp(1.0 / 20), 4),
                         (prob_mp(0.05), 6), (cases.singpert(1e-2, N=48), 4), (cases.pendulum(0.05), 4)]:
        sol = O.solve(O.opts(method, adaptive=0, newton_abstol=NTOL), spec)
        assert sol.retcode == 0
        u = sol.u.reshape(-1)
        out.append((spec.with_state(sol.t, u), method))
        out.append((spec.with_state(sol.t, u * (1.0 + 1e-3 * rng.standard_normal(u.size))), method))
    return out


@pytest.mark.parametrize("gm", [_abi.GLOBAL_RE, _abi.GLOBAL_HO], ids=["RE", "HO"])
def test_global_error_estimate_matches_oracle(ctx, gm):
    _say("global error estimate", gm)
    for spec, method in _states():
        if gm == _abi.GLOBAL_HO and method != 4:
            continue
        o = _gopts(method, _abi.CONTROL_GLOBAL, gm)
        (ge, gn), = _lib.global_error(ctx, o, [spec])
        re, rn = O.global_error(O.opts(method, newton_abstol=NTOL, controller=1, global_method=gm), spec)
        assert ge.shape == re.shape
        assert abs(gn - rn) <= 1e-6 * rn + 1e-11, (gn, rn, spec.family, method)
        assert np.max(np.abs(ge - re)) <= 1e-6 * np.max(np.abs(re)) + 1e-11
        # every interval carries its larger end vector: the same side as the oracle's
        # wherever the two ends are not within rounding of each other
        assert np.array_equal(np.sign(ge), np.sign(re)) or np.max(np.abs(ge - re)) < 1e-11


def test_global_error_batch_mixed_sizes(ctx):
    """A batch of instances of different sizes at once (the persistent queue and the
    halved meshes up to twice the largest one) gives each instance's own estimate."""
    _say("global error batch")
    specs = []
    for lam, N in [(0.7, 12), (1.9, 300), (3.1, 64), (2.2, 1000), (1.1, 5)]:
        sol = O.solve(O.opts(4, adaptive=0, newton_abstol=NTOL), cases.bratu(lam, N=N))
        specs.append(cases.bratu(lam, N=N).with_state(sol.t, sol.u.reshape(-1)))
    o = _gopts(4, _abi.CONTROL_GLOBAL, _abi.GLOBAL_RE, max_num_subintervals=1000)
    got = _lib.global_error(ctx, o, specs)
    for s, (ge, gn) in zip(specs, got):
        re, rn = O.global_error(O.opts(4, newton_abstol=NTOL, controller=1, global_method=1), s)
        assert abs(gn - rn) <= 1e-6 * rn + 1e-12, (s.N, gn, rn)


CONTROLLERS = [(_abi.CONTROL_GLOBAL, _abi.GLOBAL_RE), (_abi.CONTROL_GLOBAL, _abi.GLOBAL_HO),
               (_abi.CONTROL_SEQUENTIAL, _abi.GLOBAL_RE), (_abi.CONTROL_SEQUENTIAL, _abi.GLOBAL_HO),
               (_abi.CONTROL_HYBRID, _abi.GLOBAL_RE), (_abi.CONTROL_HYBRID, _abi.GLOBAL_HO)]


def _compare(specs, got, ref, tol=1e-9):
    for s, g, r in zip(specs, got, ref):
        assert g.retcode == r.retcode, (g.retcode, r.retcode)
        if r.retcode != 0:
            continue
        assert g.N == r.N and g.outer_iters == r.outer_iters, (g.N, r.N, g.outer_iters, r.outer_iters)
        assert np.max(np.abs(g.t - r.t)) <= 1e-12
        err = float(np.max(np.abs(g.u - r.u) / np.maximum(1.0, np.abs(r.u))))
        assert err <= tol, err
        if np.isnan(r.defect_norm):
            assert np.isnan(g.defect_norm)
        else:
            assert abs(g.defect_norm - r.defect_norm) <= 1e-6 * r.defect_norm + 1e-14


@pytest.mark.parametrize("controller,gm", CONTROLLERS, ids=["G-RE", "G-HO", "S-RE", "S-HO", "H-RE", "H-HO"])
def test_adaptive_solves_match_oracle(ctx, controller, gm):
    """Adaptive MIRK4 solves under each controller: the reference's two-point test
    problem and Bratu instances, same mesh sequence and solution as the oracle."""
    _say("controller", controller, gm)
    specs = [prob_tp(1.0 / 8), cases.bratu(1.0, N=10), cases.bratu(3.0, N=16), cases.singpert(5e-2, N=32)]
    kw = dict(abstol=1e-8, max_num_subintervals=4000)
    o = _gopts(4, controller, gm, **kw)
    got = _lib.solve_batch(ctx, o, specs)
    ref = [O.solve(O.opts(4, newton_abstol=NTOL, controller=controller, global_method=gm, **kw), s) for s in specs]
    assert all(r.retcode == 0 for r in ref)
    _compare(specs, got, ref)


def test_re_control_mirk6_and_hybrid_weights(ctx):
    """RE control with MIRK6, and hybrid control with unequal weights."""
    _say("MIRK6 RE, hybrid weights")
    specs = [prob_tp(1.0 / 6), cases.bratu(2.0, N=8)]
    kw = dict(abstol=1e-9, max_num_subintervals=4000)
    got = _lib.solve_batch(ctx, _gopts(6, _abi.CONTROL_GLOBAL, _abi.GLOBAL_RE, **kw), specs)
    ref = [O.solve(O.opts(6, newton_abstol=NTOL, controller=1, global_method=1, **kw), s) for s in specs]
    _compare(specs, got, ref)
    o = _gopts(4, _abi.CONTROL_HYBRID, _abi.GLOBAL_RE, hybrid_de=0.25, hybrid_ge=3.0, **kw)
    got = _lib.solve_batch(ctx, o, specs)
    ref = [O.solve(O.opts(4, newton_abstol=NTOL, controller=3, global_method=1, hybrid_de=0.25, hybrid_ge=3.0, **kw),
                   s) for s in specs]
    _compare(specs, got, ref)


def test_global_control_plan_and_public_api(ctx):
    """The plan API (graph replay, autotuned launch) runs the controller kernel, and
    the host mirror's controller objects select it."""
    _say("plan + public API")
    specs = [cases.bratu(l, N=n) for l, n in [(0.5, 12), (1.5, 20), (2.5, 30), (3.3, 40)]]
    kw = dict(abstol=1e-7, max_num_subintervals=3000)
    o = _gopts(4, _abi.CONTROL_SEQUENTIAL, _abi.GLOBAL_RE, **kw)
    b = _lib.Batch(ctx, o, specs)
    b.set_graph(True)
    b.autotune(repeats=1)
    b.run()
    b.sync()
    got = b.results()
    ref = [O.solve(O.opts(4, newton_abstol=NTOL, controller=2, global_method=1, **kw), s) for s in specs]
    _compare(specs, got, ref)
    b.close()
    # solve(prob, MIRK4(); controller = ...) fills the controller fields
    for ctl, code, gmc in [(GlobalErrorControl(method=REErrorControl()), 1, 1),
                           (SequentialErrorControl(), 2, 0), (HybridErrorControl(DE=2.0), 3, 0)]:
        mo = make_opts(MIRK4(), abstol=1e-7, controller=ctl)
        assert (mo.controller, mo.global_method) == (code, gmc)


def test_unsupported_combinations_raise(ctx):
    o = _gopts(6, _abi.CONTROL_GLOBAL, _abi.GLOBAL_HO)
    with pytest.raises(_lib.BVPError):
        _lib.solve_batch(ctx, o, [cases.bratu(1.0, N=8)])
