"""GPU checks of the plan API around the persistent solve: hipGraph replay, the
autotuner's launch variants, and the timing split.

Every configuration runs the same arithmetic, so the solutions must agree to the
bit across graph/no-graph runs and across the solve-kernel variants the
autotuner chooses from; the oracle parity of the solutions themselves is covered
by test_gpu_parity.py.
"""
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib, synthetic

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


def _same(ra, rb):
    for a, b in zip(ra, rb):
        assert a.retcode == b.retcode and a.N == b.N and a.newton_iters == b.newton_iters
        assert np.array_equal(a.t, b.t) and np.array_equal(a.u, b.u)


def _opts():
    o = _abi.default_opts(_abi.MIRK4)
    o.abstol = 1e-8
    return o


def test_graph_replay_matches_direct_launches(ctx):
    _say("graph replay")
    specs, _, _ = synthetic.bratu_batch(600)
    b = _lib.Batch(ctx, _opts(), specs)
    b.set_graph(False)
    b.run()
    b.sync()
    direct = b.results()
    b.set_graph(True)
    for _ in range(3):  # replays with eager work and synchronisation in between
        b.run()
        b.sync()
        _same(direct, b.results())
        g, s = b.timing()
        assert g == 0.0 and s > 0.0


def test_autotuner_variants_agree_and_choice_is_fastest(ctx):
    _say("autotune")
    specs, _, _ = synthetic.bratu_batch(2048)
    b = _lib.Batch(ctx, _opts(), specs)
    b.run()
    b.sync()
    ref = b.results()
    ms, (variant, blocks) = b.autotune(repeats=2)
    minc = b.chunking()
    _say("  candidates ms", ms, "chosen", variant, blocks, "min intervals per lane", minc)
    # M = 2: 1, 2 or 3 waves per SIMD, then at least 2, 4, 8 intervals per lane for the fastest
    assert len(ms) == 6
    assert min(ms) == ms[variant] if minc == 1 else min(ms) == ms[3 + [2, 4, 8].index(minc)]
    assert ms[variant] == min(ms[:3])
    assert blocks >= 1
    # every variant computes the same bits, at any number of persistent waves, for a
    # given interval-per-lane mapping; mappings agree to rounding
    for m in sorted({1, minc}):
        b.set_chunking(m)
        b.set_config(variant, blocks)
        b.run()
        b.sync()
        first = b.results()
        if m == 1:
            _same(ref, first)
        else:
            for x, y in zip(ref, first):
                assert x.retcode == y.retcode and x.N == y.N
                assert np.max(np.abs(x.u - y.u) / np.maximum(1.0, np.abs(x.u))) <= 1e-10
        for v in range(3):
            for nb in (blocks, 7):
                b.set_config(v, nb)
                b.run()
                b.sync()
                _same(first, b.results())


def test_block_plan_autotune(ctx):
    """The block kernels (M = 8) take 1, 2, 4 ... waves per CU."""
    _say("autotune block")
    specs, _ = synthetic.lintanh_batch(64, M=8, Nrange=(40, 80))
    o = _abi.default_opts(_abi.MIRK6)
    o.precision = _abi.PREC_FP32_REFINE64
    o.refine_iters = 2
    b = _lib.Batch(ctx, o, specs)
    b.run()
    b.sync()
    ref = b.results()
    ms, (variant, blocks) = b.autotune(repeats=1)
    _say("  candidates ms", ms, "chosen", variant, blocks)
    assert len(ms) >= 1 and variant == 0
    b.run()
    b.sync()
    _same(ref, b.results())


def test_autotune_chooses_record_precision(ctx):
    """A mixed-precision wave plan is also timed with FP64 records; the choice is
    the fastest candidate, and the plan then computes the same bits as a plan
    built for that precision with the same launch configuration."""
    _say("autotune precision")
    specs, _ = synthetic.singpert_batch(1024)
    o = _abi.default_opts(_abi.MIRK4)
    o.abstol = 1e-6
    o.max_num_subintervals = 16384
    o.precision = _abi.PREC_FP32_REFINE64
    b = _lib.Batch(ctx, o, specs)
    ms, (variant, blocks) = b.autotune(repeats=1)
    _say("  candidates ms", ms, "chosen", variant, blocks, "precision", b.precision(), "chunking", b.chunking())
    # variants 0..2 with FP32 records, then with FP64 records, then the mappings
    assert len(ms) == 9
    best = int(np.argmin(ms[:6]))
    assert variant == best % 3
    assert b.precision() == (_abi.PREC_FP32_REFINE64 if best < 3 else _abi.PREC_FP64)
    b.run()
    b.sync()
    got = b.results()
    o2 = _abi.default_opts(_abi.MIRK4)
    o2.abstol = 1e-6
    o2.max_num_subintervals = 16384
    o2.precision = b.precision()
    ref_ctx = _lib.Context(0)  # a fresh context: no tuned entry for this kind
    b2 = _lib.Batch(ref_ctx, o2, specs)
    b2.set_config(variant, blocks)
    b2.set_chunking(b.chunking())
    b2.run()
    b2.sync()
    _same(got, b2.results())
