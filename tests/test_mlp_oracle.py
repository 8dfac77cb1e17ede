"""CPU checks of the warm-start network's oracle (oracle/mlp_oracle.py).

The float64 mode is pinned by central finite differences (a dropped tanh
derivative, a transposed matrix or a wrong scale fails them); BF16 rounding is
pinned against torch's bfloat16 conversion and hand-computed ties; the
emulating mode is the float64 function up to that rounding."""
import numpy as np
import pytest
import torch

import mlp_oracle as MO


def test_bf16_rounding_ties_to_even():
    one = np.float32(1.0)
    cases = {
        1.0: 1.0,
        1.0 + 2.0 ** -8: 1.0,                      # halfway, even neighbour is 1.0
        1.0 + 3 * 2.0 ** -8: 1.0 + 2.0 ** -6,      # halfway, even neighbour above
        1.0 + 2.0 ** -8 + 2.0 ** -20: 1.0 + 2.0 ** -7,
        -2.5: -2.5,
    }
    for a, b in cases.items():
        assert MO.bf16(np.array([a], np.float32))[0] == np.float32(b), a
    assert one == MO.bf16(np.array([one]))[0]


def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    a = (rng.standard_normal(20000) * 10.0 ** rng.uniform(-6, 6, 20000)).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).float().numpy()
    assert np.array_equal(MO.bf16(a), ref)


def _problem(n=7, seed=3):
    rng = np.random.default_rng(seed)
    w = MO.glorot(seed)
    x = rng.uniform(-1, 1, (n, 4))
    y = rng.normal(0, 1, (n, 2))
    norm = (rng.normal(0, 0.3, 4), rng.uniform(0.5, 2, 4), rng.normal(0, 0.3, 2), rng.uniform(0.5, 2, 2))
    return w, x, y, norm


def test_gradient_matches_finite_differences_per_block():
    w, x, y, norm = _problem()
    _, g = MO.loss_and_grad(w, x, y, norm, emulate=False)
    rng = np.random.default_rng(11)
    o = 0
    for name, shp in MO.SHAPES:
        k = int(np.prod(shp))
        for j in rng.choice(k, size=min(k, 4), replace=False):
            e = np.zeros_like(w)
            e[o + j] = 1.0
            h = 1e-5
            fd = (MO.loss_only(w + h * e, x, y, norm) - MO.loss_only(w - h * e, x, y, norm)) / (2 * h)
            assert abs(fd - g[o + j]) <= 1e-8 + 1e-6 * abs(fd), (name, j, fd, g[o + j])
        o += k


def test_gradient_matches_finite_differences_random_directions():
    w, x, y, norm = _problem(n=13, seed=5)
    _, g = MO.loss_and_grad(w, x, y, norm, emulate=False)
    rng = np.random.default_rng(2)
    for _ in range(4):
        d = rng.standard_normal(w.shape)
        d /= np.linalg.norm(d)
        h = 1e-5
        fd = (MO.loss_only(w + h * d, x, y, norm) - MO.loss_only(w - h * d, x, y, norm)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-9 + 1e-6 * abs(fd)


def test_emulated_mode_is_the_same_function_up_to_bf16():
    w, x, y, norm = _problem(n=64, seed=9)
    l64, g64 = MO.loss_and_grad(w, x, y, norm, emulate=False)
    l32, g32 = MO.loss_and_grad(w, x, y, norm, emulate=True)
    assert abs(l32 - l64) <= 0.05 * l64
    assert np.linalg.norm(g32 - g64) <= 0.05 * np.linalg.norm(g64)
    y64, _ = MO.forward(w, x, norm, emulate=False)
    y32, _ = MO.forward(w, x, norm, emulate=True)
    assert np.max(np.abs(y32 - y64)) <= 0.05 * (1 + np.max(np.abs(y64)))


def test_pack_unpack_layout():
    w = np.arange(MO.NPARAM, dtype=np.float64)
    d = MO.unpack(w)
    assert d["W1"][1, 0] == 256                 # input-major: W1[i][o] at i*256 + o
    assert d["b1"][0] == 1024
    assert d["W2"][0, 1] == 1024 + 256 + 1
    assert d["b4"][1] == MO.NPARAM - 1
    assert np.array_equal(MO.pack(d), w)
    assert MO.NPARAM == 133378
