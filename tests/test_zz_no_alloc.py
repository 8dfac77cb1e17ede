"""bvp_mirk_solve_batch_noalloc against bvp_mirk_solve_batch, through the C ABI:
the same bits on two calls in a row (the second with a smaller batch, which
reuses the static scratch the first one left), and a clear error when the batch
does not fit the scratch."""
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib
import cases

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


def _same(a, b):
    for x, y in zip(a, b):
        assert x.retcode == y.retcode and x.N == y.N and x.newton_iters == y.newton_iters
        assert x.outer_iters == y.outer_iters and x.fallback_used == y.fallback_used
        assert np.array_equal(x.t, y.t) and np.array_equal(x.u, y.u)
        assert x.defect_norm == y.defect_norm and x.resid_norm == y.resid_norm


def test_noalloc_matches_solve_batch_twice(ctx):
    _say("no-alloc: two calls, second smaller")
    rng = np.random.default_rng(9)
    big = [cases.bratu(float(l), N=int(n)) for l, n in zip(rng.uniform(0.2, 3.3, 48), rng.integers(20, 300, 48))]
    small = [cases.singpert(e, N=32) for e in (1e-1, 1e-2, 3e-3)] + [cases.bratu(1.0, N=11)]
    o = _abi.default_opts(_abi.MIRK4)
    o.abstol = 1e-8
    o.max_num_subintervals = 2048
    ref1 = _lib.solve_batch(ctx, o, big)
    got1 = _lib.solve_batch_noalloc(ctx, o, big)
    _same(ref1, got1)
    o2 = _abi.default_opts(_abi.MIRK6)
    o2.abstol = 1e-6
    o2.max_num_subintervals = 2048
    o2.precision = _abi.PREC_FP32_REFINE64
    ref2 = _lib.solve_batch(ctx, o2, small)
    got2 = _lib.solve_batch_noalloc(ctx, o2, small)
    _same(ref2, got2)
    assert all(r.retcode == 0 for r in got1 + got2)


def test_noalloc_rejects_batches_beyond_the_scratch(ctx):
    _say("no-alloc: too large")
    specs = [cases.bratu(1.0, N=64) for _ in range(400)]
    o = _abi.default_opts(_abi.MIRK4)
    o.max_num_subintervals = 3000
    with pytest.raises(_lib.BVPError, match="no-allocation scratch"):
        _lib.solve_batch_noalloc(ctx, o, specs)
    # the context stays usable
    g = _lib.solve_batch_noalloc(ctx, o, specs[:2])
    assert all(r.retcode == 0 for r in g)
