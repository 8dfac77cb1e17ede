"""GPU dense output (bvp_mirk_interpolate: sol(t), sol(t, Val{1}) of
MIRKInterpolation, MIRK/interpolation.jl:17-206) against the oracle, and the
reference's interpolation tests (MIRKt/Core/mirk_basic_tests.jl:173-313) run
end to end on the GPU: solve, then evaluate the extension.

Tolerance against the oracle, relative to max(1, |u|): 1e-11 for u, 1e-10 for u'
(FP64 both sides with different FMA contraction and summation order; the MIRK6
weight polynomials have coefficients of order 10-100, and the points include
extrapolation beyond the end nodes, tau > 1, where they grow further).
"""
import math
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib
import pyoracle as O
import cases
from test_oracle_dense import prob_tp, prob_mp, tp_exact

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
    "linear3": lambda: cases.linear([[0.3, 1.0, -0.2], [-1.0, 0.1, 0.4], [0.5, -0.3, -0.6]], c=0.7,
                                    tspan=(0, 1), bc=([(0, 1.0), (2, -1.0)], [(1, 0.5)]), dt=1 / 13),
    "scalar": lambda: cases.linear([[0.5]], tspan=(0, 1), bc=([(0, 1.0)], []), dt=1 / 9),
}


@pytest.mark.parametrize("method", [4, 6])
@pytest.mark.parametrize("name", sorted(FAMS))
def test_interpolate_parity(ctx, method, name):
    _say("dense output", name, method)
    rng = np.random.default_rng(11)
    specs = [cases.random_state(cases.random_mesh(FAMS[name](), rng), rng, 0.3) for _ in range(3)]
    tvals = []
    for s in specs:
        m = s.full_mesh()
        t = np.concatenate([rng.uniform(m[0], m[-1], 40), m[::3], [m[0], m[-1], m[0] - 0.01, m[-1] + 0.02]])
        tvals.append(t)
    tvals[2] = tvals[2][:0]  # an instance with no points
    got = _lib.interpolate(ctx, method, specs, tvals)
    for s, t, (u, du) in zip(specs, tvals, got):
        ru, rdu = O.interpolate(method, s, t)
        assert u.shape == ru.shape
        if len(t):
            assert np.max(np.abs(u - ru) / np.maximum(1.0, np.abs(ru))) <= 1e-11
            assert np.max(np.abs(du - rdu) / np.maximum(1.0, np.abs(rdu))) <= 1e-10


@pytest.mark.parametrize("method", [4, 6])
def test_reference_interpolation_cases_on_gpu(ctx, method):
    """Solve the reference's two interpolation problems on the GPU and evaluate the
    extension there: values and derivatives within 1e-6 of the analytic solutions."""
    _say("reference interpolation cases", method)
    o = _abi.default_opts(method)
    g = _lib.solve_batch(ctx, o, [prob_tp()])[0]
    assert g.retcode == 0
    s = prob_tp().with_state(g.t, g.u.reshape(-1))
    pts = [0.001, 0.04, 0.0375, 0.5123]
    (u, du), = _lib.interpolate(ctx, method, [s], [pts])
    for k, t in enumerate(pts):
        ex, dex = tp_exact(t)
        assert np.max(np.abs(u[k] - ex)) < 1e-6 and np.max(np.abs(du[k] - dex)) < 1e-6
    g = _lib.solve_batch(ctx, o, [prob_mp()])[0]
    assert g.retcode == 0
    s = prob_mp().with_state(g.t, g.u.reshape(-1))
    (u, _), = _lib.interpolate(ctx, method, [s], [[math.pi / 6, math.pi / 3]], derivative=False)
    assert abs(u[0, 0] - 0.5) < 1e-6 and abs(u[1, 1] - 0.5) < 1e-6
    assert abs(u[0, 1] - math.cos(math.pi / 6)) < 1e-6 and abs(u[1, 0] - math.sin(math.pi / 3)) < 1e-6


def test_interpolate_rejects_block_sizes(ctx):
    rng = np.random.default_rng(1)
    s = cases.linear_tanh_c3(rng, M=8, N=16)
    with pytest.raises(_lib.BVPError):
        _lib.interpolate(ctx, 4, [s], [[0.5]])
