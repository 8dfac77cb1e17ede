"""GPU checks of the warm-start network and of the warm-started solve.

* forward pass and FP32 gradient of the HIP training kernels against the CPU
  oracle that rounds at the same places (BF16 parameters/activations/dL/dz,
  FP32 sums): outputs within 5e-3 relative, gradient within 2e-2 per block and
  1e-2 overall (relative Frobenius norm; the residual difference is FP32
  summation order flipping single BF16 roundings), loss within 1e-4;
* training lowers the loss, and a captured (hipGraph) step replays to the same
  bits as direct launches;
* a warm start whose first Newton step does not contract falls back to the
  constant guess and then reproduces the plain solve bit for bit;
* a network trained on the method's own Bratu (u(0)=a, u(1)=b) solves saves
  Newton iterations on unseen instances, and the warm-started solutions satisfy
  the oracle's Newton correction test on their mesh (<= 1e-9 relative).
"""
import sys

import numpy as np
import pytest

from boundaryvaluediffeq_amd import _abi, _lib, learned, synthetic
import mlp_oracle as MO
import pyoracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return _lib.context(0)


def _net(ctx, seed=3):
    rng = np.random.default_rng(seed)
    w = MO.glorot(seed).astype(np.float32)
    norm = (rng.normal(0, 0.3, 4), rng.uniform(0.5, 2, 4), rng.normal(0, 0.3, 2), rng.uniform(0.5, 2, 2))
    net = _lib.MLP(ctx, seed=seed)
    net.set_params(w)
    net.set_normalization(*norm)
    return net, w, norm


def test_predict_matches_oracle(ctx):
    _say("mlp predict")
    net, w, norm = _net(ctx)
    x = np.random.default_rng(1).uniform(-1.5, 1.5, (1000, 4)).astype(np.float32)  # 1000: ragged last tile
    got = net.predict(x)
    ref, _ = MO.forward(w, x, norm, emulate=True)
    scale = np.asarray(norm[3], np.float32)
    assert np.all(np.abs(got - ref) <= 5e-3 * (np.abs(ref) + scale)), np.max(np.abs(got - ref))


def test_params_round_trip(ctx):
    net, w, _ = _net(ctx, seed=4)
    assert np.array_equal(net.params(), w)


@pytest.mark.parametrize("n", [64, 333])
def test_gradient_matches_oracle(ctx, n):
    _say("mlp gradient", n)
    net, w, norm = _net(ctx, seed=5)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (n, 4)).astype(np.float32)
    y = rng.normal(0, 1, (n, 2)).astype(np.float32)
    loss, g = net.gradient(x, y)
    lr, gr = MO.loss_and_grad(w, x, y, norm, emulate=True)
    assert abs(loss - lr) <= 1e-4 * lr
    assert np.linalg.norm(g - gr) <= 1e-2 * np.linalg.norm(gr)
    o = 0
    for name, shp in MO.SHAPES:
        k = int(np.prod(shp))
        a, b = g[o:o + k], gr[o:o + k]
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7, name
        o += k


def _toy(n=16384, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 4)).astype(np.float32)
    y = np.stack([np.sin(2 * x[:, 0] + x[:, 3]) * x[:, 1], np.cos(x[:, 2] * x[:, 3])], 1).astype(np.float32)
    return x, y


def test_training_lowers_loss_and_graph_replay_matches(ctx):
    _say("mlp training")
    x, y = _toy()
    nets = []
    for g in (True, False):
        net = _lib.MLP(ctx, seed=9)
        net.set_normalization(*learned.fit_normalization(x, y))
        loss, ms = net.train(x, y, epochs=30, batch=2048, lr=2e-3, lr_min=1e-4, warmup=20, seed=3, use_graph=g)
        _say("  graph", g, "steps", len(loss), "first", loss[0], "last", loss[-1], "ms", ms)
        nets.append((net, loss))
    (n1, l1), (n2, l2) = nets
    assert len(l1) == 30 * 8
    assert l1[-1] < 0.05 * l1[0]
    assert np.array_equal(l1, l2)
    assert np.array_equal(n1.params(), n2.params())


def test_chunked_training_equals_one_call(ctx):
    """A schedule split over several calls (first_epoch / total_epochs) gives the
    same bits as one call."""
    x, y = _toy(8192, seed=4)
    norm = learned.fit_normalization(x, y)
    a = _lib.MLP(ctx, seed=2)
    a.set_normalization(*norm)
    la, _ = a.train(x, y, epochs=6, batch=1024, lr=2e-3, warmup=5, seed=5)
    b = _lib.MLP(ctx, seed=2)
    b.set_normalization(*norm)
    parts = [b.train(x, y, epochs=k, batch=1024, lr=2e-3, warmup=5, seed=5, first_epoch=e0, total_epochs=6)[0]
             for e0, k in ((0, 2), (2, 3), (5, 1))]
    assert np.array_equal(la, np.concatenate(parts))
    assert np.array_equal(a.params(), b.params())


def _plain_opts(abstol=1e-8):
    o = _abi.default_opts(_abi.MIRK4)
    o.abstol = abstol
    o.newton_abstol = 1e-10
    return o


def test_bad_warm_start_falls_back_to_plain_bitwise(ctx):
    _say("warm fallback")
    specs, _, _ = synthetic.bratu_ab_batch(24, seed=123, Nrange=(16, 80))
    plain = _lib.solve_batch(ctx, _plain_opts(), specs)
    bad = [s.with_state(s.full_mesh(), np.full(2 * (s.N + 1), 40.0)) for s in specs]  # e^40: no contraction
    o = _plain_opts()
    o.warm_start = _abi.WARM_MLP
    warm = _lib.solve_batch(ctx, o, bad)
    for p, w in zip(plain, warm):
        assert (w.fallback_used >> 1) & 1 == 1
        assert w.retcode == p.retcode and w.N == p.N
        assert np.array_equal(w.t, p.t) and np.array_equal(w.u, p.u)


@pytest.fixture(scope="module")
def bratu_net(ctx):
    """A network trained on 4,096 plain solves of the Bratu (a, b) family."""
    _say("warm start training")
    opts = _plain_opts()
    opts.max_newton_iters = 50
    train_specs, _, _ = synthetic.bratu_ab_batch(4096, seed=synthetic.SEED_BRATU_AB_TRAIN)
    x, y, sols = learned.harvest(ctx, train_specs, opts, per_solve=64)
    _say("  samples", len(x), "solved", sum(r.retcode == 0 for r in sols))
    net, loss, ms = learned.train(ctx, x, y, epochs=40, batch=4096, lr=2e-3, lr_min=2e-5, warmup=50)
    _say("  loss", loss[0], "->", loss[-1], "ms", ms)
    assert loss[-1] < 0.02 * loss[0]
    return net


def _bratu_f(lam, y):
    return np.stack([y[:, 1], -lam * np.exp(y[:, 0])], 1)


@pytest.mark.parametrize("anchors", [0, 2, 9])
def test_learned_guess_is_network_then_hermite(ctx, bratu_net, anchors):
    """The device guess equals the network at every node (anchors = 0), or the
    network at the anchors followed by cubic Hermite interpolation with the slopes
    f(t, y) of the problem (checked against numpy, 1e-12 relative)."""
    specs, _, _ = synthetic.bratu_ab_batch(40, seed=77, Nrange=(3, 90))
    b = _lib.Batch(ctx, _plain_opts(), specs)
    b.set_warm_start(bratu_net, _abi.FEAT_PARAM, anchors)
    got = b.learned_guess()
    for s, g in zip(specs, got):
        lam, a, bb = float(s.params[0]), float(s.bc_val[0]), float(s.bc_val[1])
        t = s.full_mesh()
        if anchors == 0:
            x = np.stack([np.full_like(t, lam), np.full_like(t, a), np.full_like(t, bb), t], 1)
            ref = bratu_net.predict(x).astype(np.float64)
            assert np.array_equal(g, ref)
            continue
        K = anchors
        ta = np.array([s.t0 + (s.t1 - s.t0) * (j / (K - 1)) for j in range(K)])
        ta[-1] = s.t1
        xa = np.stack([np.full(K, lam), np.full(K, a), np.full(K, bb), ta], 1)
        ya = bratu_net.predict(xa).astype(np.float64)
        ma = _bratu_f(lam, ya)
        sc = (t - s.t0) / (s.t1 - s.t0) * (K - 1)
        j = np.clip(np.floor(sc).astype(int), 0, K - 2)
        th = (sc - j)[:, None]
        h = (s.t1 - s.t0) / (K - 1)
        ref = ((2 * th ** 3 - 3 * th ** 2 + 1) * ya[j] + (th ** 3 - 2 * th ** 2 + th) * h * ma[j]
               + (-2 * th ** 3 + 3 * th ** 2) * ya[j + 1] + (th ** 3 - th ** 2) * h * ma[j + 1])
        assert np.all(np.abs(g - ref) <= 1e-12 * (1 + np.abs(ref)))


@pytest.mark.parametrize("anchors", [0, 16])
def test_learned_warm_start_saves_newton_iterations(ctx, bratu_net, anchors):
    # Newton tolerance 1e-12: the (a, b) family has Jacobians with norm(J^-1) ~ 1e2,
    # so that the solutions sit within 1e-9 of the discrete solution
    opts = _plain_opts()
    opts.newton_abstol = 1e-12
    eval_specs, _, _ = synthetic.bratu_ab_batch(512, seed=synthetic.SEED_BRATU_AB_EVAL)
    b = _lib.Batch(ctx, opts, eval_specs)
    b.run()
    b.sync()
    plain = b.results()
    b.set_warm_start(bratu_net, _abi.FEAT_PARAM, anchors)
    b.run()
    b.sync()
    warm = b.results()
    sp, sw = learned.summary(plain), learned.summary(warm)
    _say("  anchors", anchors, "plain", sp, "\n  warm ", sw)
    # instances without a solution (part of the (lambda, a, b) box) run to maxiters
    # from either guess; the saving shows on the solved ones
    assert sw["mean_newton_iters_solved"] < 0.85 * sp["mean_newton_iters_solved"]
    assert sw["fallback_rate"] < 0.2
    assert sw["solved"] >= sp["solved"] - 2
    # every warm-started success is a solution of the collocation equations on its mesh
    checked = 0
    for s, w in zip(eval_specs, warm):
        if w.retcode != 0 or checked >= 40:
            continue
        d = O.newton_step(4, s.with_state(w.t, w.u.reshape(-1)))
        assert np.max(np.abs(d) / np.maximum(1.0, np.abs(w.u.reshape(-1)))) <= 1e-9
        assert w.defect_norm <= opts.abstol
        checked += 1
    assert checked >= 20


def test_learned_mixed_graph_plan(ctx, bratu_net):
    """The bench's configuration: learned guess (K = 4 anchors), FP32 elimination
    records, graph replay, autotuned launch.  Against the plain FP64 plan on the
    same unseen instances: as many solved, fewer Newton iterations, replays give
    the same bits, and every success passes the oracle's Newton-correction test on
    its mesh (max|F| <= abstol = 1e-8 with norm(J^-1) ~ 1e2: corrections <= 1e-6)."""
    _say("learned + mixed + graph plan")
    eval_specs, _, _ = synthetic.bratu_ab_batch(512, seed=synthetic.SEED_BRATU_AB_EVAL)
    opts = _plain_opts()
    b = _lib.Batch(ctx, opts, eval_specs)
    b.run()
    b.sync()
    plain = b.results()
    om = _plain_opts()
    om.precision = _abi.PREC_FP32_REFINE64
    bm = _lib.Batch(ctx, om, eval_specs)
    bm.set_graph(True)
    bm.set_warm_start(bratu_net, _abi.FEAT_PARAM, 4)
    ms, (variant, blocks) = bm.autotune(repeats=1)
    _say("  candidates", ms, "chosen", variant, "precision", bm.precision())
    bm.run()
    bm.sync()
    first = bm.results()
    bm.run()
    bm.sync()
    again = bm.results()
    for a, c in zip(first, again):
        assert a.retcode == c.retcode and a.N == c.N and np.array_equal(a.u, c.u, equal_nan=True)
    sp, sm = learned.summary(plain), learned.summary(first)
    _say("  plain", sp, "\n  learned+mixed", sm)
    assert sm["solved"] >= sp["solved"] - 2
    assert sm["mean_newton_iters_solved"] < 0.85 * sp["mean_newton_iters_solved"]
    checked = 0
    for s, w in zip(eval_specs, first):
        if w.retcode != 0 or checked >= 40:
            continue
        d = O.newton_step(4, s.with_state(w.t, w.u.reshape(-1)))
        assert np.max(np.abs(d) / np.maximum(1.0, np.abs(w.u.reshape(-1)))) <= 1e-6
        assert w.defect_norm <= opts.abstol
        checked += 1
    assert checked >= 20
