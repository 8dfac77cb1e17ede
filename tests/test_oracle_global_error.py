"""Pins the oracle's global-error controllers (GlobalErrorControl with HO and RE
estimates, SequentialErrorControl, HybridErrorControl; Core/calc_errors.jl:54-110,
MIRK/adaptivity.jl:77-243, 463-567) by what the mathematics fixes:

* Richardson extrapolation: (u_{h/2} - u_h) 2^p / (2^p - 1) is the leading term of
  the global error u - u_h, so on a problem with a known solution the estimate
  approaches the true error as h -> 0 (a missing 2^p / (2^p - 1) factor is 6%
  off for MIRK4, a wrong exponent far more);
* the higher-order estimate u_{MIRK6} - u_{MIRK4} approaches the MIRK4 error;
* each adaptive controller ends with its own error measure below abstol, and the
  global controllers leave a true global error of the order of abstol.
"""
import math

import numpy as np
import pytest

import pyoracle as O
from test_oracle_dense import prob_tp, tp_exact


def _true_error(sol):
    ex = np.array([tp_exact(t)[0] for t in sol.t])
    return float(np.max(np.abs(sol.u - ex) / (1.0 + np.abs(sol.u))))


def _solve(N, method=4, **kw):
    return O.solve(O.opts(method, abstol=1e-13, newton_abstol=1e-13, **kw), prob_tp(1.0 / N))


@pytest.mark.parametrize("gm", [1, 0], ids=["RE", "HO"])
def test_global_estimate_approaches_the_true_error(gm):
    """The oracle's estimate at a converged non-adaptive solution against the true
    error there; the ratio tends to 1 under refinement."""
    ratios = []
    for N in (16, 32):
        spec = prob_tp(1.0 / N)
        low = O.solve(O.opts(4, adaptive=0, newton_abstol=1e-13), spec)
        true = _true_error(low)
        o = O.opts(4, newton_abstol=1e-13, controller=1, global_method=gm)
        errs, est = O.global_error(o, spec.with_state(low.t, low.u.reshape(-1)))
        assert errs.shape == (N, 2)
        assert est == pytest.approx(np.max(np.abs(errs)) * (16.0 / 15.0 if gm == 1 else 1.0), rel=1e-14)
        ratios.append(est / true)
    assert abs(ratios[1] - 1.0) < 0.01, ratios
    assert abs(ratios[1] - 1.0) < abs(ratios[0] - 1.0) + 1e-12, ratios


def test_per_interval_errors_pick_the_larger_end_node():
    """GE_subinterval!: each interval carries the end-node error vector with the larger
    max-abs (checked against the node errors recomputed from the two solutions)."""
    spec = prob_tp(1.0 / 12)
    low = O.solve(O.opts(4, adaptive=0, newton_abstol=1e-13), spec)
    high = O.solve(O.opts(6, adaptive=0, newton_abstol=1e-13), spec)
    node = (high.u - low.u) / (1.0 + np.abs(low.u))
    errs, _ = O.global_error(O.opts(4, newton_abstol=1e-13, controller=1, global_method=0),
                             spec.with_state(low.t, low.u.reshape(-1)))
    for i in range(12):
        a, b = np.max(np.abs(node[i])), np.max(np.abs(node[i + 1]))
        assert np.allclose(errs[i], node[i] if a >= b else node[i + 1], rtol=1e-6, atol=1e-18)


@pytest.mark.parametrize("controller,gm", [(1, 1), (1, 0), (2, 1), (2, 0), (3, 1), (3, 0)])
def test_controllers_reach_abstol(controller, gm):
    abstol = 1e-8
    o = O.opts(4, abstol=abstol, newton_abstol=1e-13, controller=controller, global_method=gm,
               max_num_subintervals=4000)
    sol = O.solve(o, prob_tp(1.0 / 8))
    assert sol.retcode == 0
    assert sol.N > 8  # the error measure forced refinement
    true = _true_error(sol)
    assert true < 10 * abstol, true
    if controller in (2, 3):  # the defect is part of both measures
        assert sol.defect_norm <= abstol


def test_global_control_differs_from_defect_control():
    """The two families of controllers select different meshes for the same problem."""
    base = dict(abstol=1e-8, newton_abstol=1e-13, max_num_subintervals=4000)
    d = O.solve(O.opts(4, **base), prob_tp(1.0 / 8))
    g = O.solve(O.opts(4, controller=1, global_method=1, **base), prob_tp(1.0 / 8))
    assert d.retcode == g.retcode == 0
    assert d.N != g.N


def test_ho_with_mirk6_is_unsupported():
    o = O.opts(6, controller=1, global_method=0)
    with pytest.raises(RuntimeError):
        O.solve(o, prob_tp(1.0 / 8))
