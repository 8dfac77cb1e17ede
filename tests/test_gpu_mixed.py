"""Mixed precision on the wave kernels (M <= 4): FP32 storage of the chunk
elimination factors, FP64 residuals and convergence test.

The FP32 factors make every Newton correction inexact by the rounding of X and
Z (relative 2^-24, amplified by the conditioning of the interval blocks), while
the residual that decides convergence is the same FP64 residual as the oracle's.
Tolerances:
  one Newton correction ........... 1e-5 relative to max|delta| (FP32 factors)
  full solves ..................... same return code and mesh sequence as the
                                    oracle where the mesh is fixed by the first
                                    defect test, Newton iterations within one, the
                                    solution within 1e-6 relative (both sides
                                    stop at max|F| <= abstol = 1e-8 with
                                    |J^-1| ~ 1e2), defect <= abstol
and the mode is shown to be engaged: the mixed corrections are not bitwise the
FP64 ones (every case here has more than 64 intervals, so that lanes hold chunks
of two or more intervals and write factor records; with one interval per lane
all factors stay in the FP64 merge tree).
"""
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib, synthetic
import pyoracle as O
import cases

pytestmark = pytest.mark.gpu

MIXED = _abi.PREC_FP32_REFINE64


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


FAMS = {
    "bratu": lambda: cases.bratu(2.5, N=137),
    "pendulum": lambda: cases.pendulum(dt=0.01),
    "singpert": lambda: cases.singpert(0.05, N=170),
    "linear3": lambda: cases.linear([[0.3, 1.0, -0.2], [-1.0, 0.1, 0.4], [0.5, -0.3, -0.6]], c=0.7,
                                    tspan=(0, 1), bc=([(0, 1.0), (2, -1.0)], [(1, 0.5)]), dt=1 / 230),
}


@pytest.mark.parametrize("method", [4, 6])
@pytest.mark.parametrize("name", sorted(FAMS))
def test_newton_step_mixed(ctx, method, name):
    _say("newton_step mixed", name, method)
    rng = np.random.default_rng(7)
    specs = [cases.random_state(FAMS[name](), rng, 0.3) for _ in range(3)]
    specs[1] = cases.random_mesh(specs[1], rng)
    got = _lib.newton_step(ctx, method, specs, precision=MIXED)
    full = _lib.newton_step(ctx, method, specs)
    differs = False
    for s, d, d64 in zip(specs, got, full):
        ref = O.newton_step(method, s)
        scale = max(1.0, np.max(np.abs(ref)))
        assert np.max(np.abs(d - ref)) <= 1e-5 * scale
        differs |= not np.array_equal(d, d64)
    assert differs, "FP32 factor storage not engaged"


@pytest.mark.parametrize("method", [4, 6])
def test_solve_bratu_batch_mixed(ctx, method):
    _say("solve bratu mixed", method)
    rng = np.random.default_rng(21)
    lam = rng.uniform(0.1, 3.4, 48)
    Ns = rng.integers(64, 300, 48)
    specs = [cases.bratu(float(l), N=int(n)) for l, n in zip(lam, Ns)]
    o = _abi.default_opts(method)
    o.abstol = 1e-8
    o.precision = MIXED
    got = _lib.solve_batch(ctx, o, specs)
    oo = O.opts(method, abstol=1e-8)
    for s, g in zip(specs, got):
        r = O.solve(oo, s)
        assert g.retcode == r.retcode == 0
        assert g.N == r.N and g.outer_iters == r.outer_iters
        assert abs(g.newton_iters - r.newton_iters) <= 1
        assert g.defect_norm <= 1e-8 and g.resid_norm <= 1e-8
        scale = np.maximum(1.0, np.abs(r.u))
        assert np.max(np.abs(g.u - r.u) / scale) <= 1e-6


def test_solve_singpert_adaptive_mixed(ctx):
    """Mesh refinement driven by the defect, with FP32 factors.  The defect
    estimates carry the 1e-7-level differences of the inexact corrections, and
    the number of points mesh selection places rounds them: the mesh sequence
    may part from the oracle's, so the check is on what defect control
    guarantees (success, defect <= abstol) and on the final size staying within
    a factor of two of the oracle's."""
    _say("solve singpert mixed")
    specs = [cases.singpert(eps, N=32) for eps in (0.05, 0.01, 0.003)]
    o = _abi.default_opts(4)
    o.abstol = 1e-6
    o.precision = MIXED
    got = _lib.solve_batch(ctx, o, specs)
    oo = O.opts(4, abstol=1e-6)
    for s, g in zip(specs, got):
        r = O.solve(oo, s)
        assert g.retcode == r.retcode == 0 and g.defect_norm <= 1e-6
        assert r.N // 2 <= g.N <= 2 * r.N, (g.N, r.N)


def test_plan_mixed_matches_fp64_plan(ctx):
    """The plan API on the specialised Bratu kernel with FP32 factors and the
    learned-guess slot empty: same successes and meshes as the FP64 plan."""
    _say("plan mixed")
    specs, _, _ = synthetic.bratu_batch(1024)
    o = _abi.default_opts(_abi.MIRK4)
    o.abstol = 1e-8
    b64 = _lib.Batch(ctx, o, specs)
    b64.run()
    b64.sync()
    r64 = b64.results()
    o.precision = MIXED
    bm = _lib.Batch(ctx, o, specs)
    bm.run()
    bm.sync()
    rm = bm.results()
    for a, c in zip(r64, rm):
        assert a.retcode == c.retcode and a.N == c.N
        if a.retcode == 0:
            assert c.defect_norm <= 1e-8 and abs(a.newton_iters - c.newton_iters) <= 1
            assert np.max(np.abs(a.u - c.u) / np.maximum(1.0, np.abs(a.u))) <= 1e-6
