"""CPU oracle of the warm-start network (test infrastructure only: imported by
tests/, never by the library).

The learned warm start has no counterpart in the reference; it produces the
initial guess that the reference takes from the caller
(lib/BoundaryValueDiffEqCore/src/utils.jl:750-777).  This module restates the
network of the HIP kernels in numpy so that their forward pass and gradients
can be checked:

    x_n = (x - in_shift) * in_scale
    h1 = tanh(x_n W1 + b1),  h2 = tanh(h1 W2 + b2),  h3 = tanh(h2 W3 + b3)
    out = h3 W4 + b4,  target t = (y - out_shift) / out_scale
    L = sum_points 0.5 |out - t|^2 / n

Parameter vector layout (matrices input-major W[i][o]):
W1[4][256] b1[256] W2[256][256] b2[256] W3[256][256] b3[256] W4[256][2] b4[2].

``emulate=True`` rounds where the kernels round (parameters, activations and
hidden-layer gradients dL/dz to BF16, round to nearest even; sums in FP32);
``emulate=False`` computes in float64 without rounding, and that mode is pinned
against central finite differences in tests/test_mlp_oracle.py.
"""
import numpy as np

NIN, H, NOUT = 4, 256, 2
SHAPES = [("W1", (NIN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (H, NOUT)), ("b4", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


def unpack(w):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = w[o:o + k].reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in SHAPES])


def glorot(seed=0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            r = np.sqrt(6.0 / (shp[0] + shp[1]))
            d[name] = rng.uniform(-r, r, shp)
        else:
            d[name] = rng.normal(0.0, 0.1, shp)
    return pack(d)


def _norm(norm):
    if norm is None:
        return np.zeros(NIN), np.ones(NIN), np.zeros(NOUT), np.ones(NOUT)
    return tuple(np.asarray(v, dtype=np.float64) for v in norm)


def forward(w, x, norm=None, emulate=True):
    """Network output in raw units, and the intermediate values."""
    ish, isc, osh, osc = _norm(norm)
    if emulate:
        f = np.float32
        P = {k: bf16(v.astype(f)) for k, v in unpack(np.asarray(w, dtype=f)).items()}
        xn = ((np.asarray(x, f) - ish.astype(f)) * isc.astype(f)).astype(f)
        act = lambda z: bf16(np.tanh(z.astype(f)))
    else:
        f = np.float64
        P = unpack(np.asarray(w, dtype=f))
        xn = (np.asarray(x, f) - ish) * isc
        act = np.tanh
    h1 = act(xn @ P["W1"] + P["b1"])
    h2 = act(h1 @ P["W2"] + P["b2"])
    h3 = act(h2 @ P["W3"] + P["b3"])
    out = h3 @ P["W4"] + P["b4"]
    y = out * osc.astype(f) + osh.astype(f)
    return y, dict(P=P, xn=xn, h1=h1, h2=h2, h3=h3, out=out)


def loss_and_grad(w, x, y, norm=None, emulate=True):
    """L and dL/dw (parameter-vector layout) over the points (x, y)."""
    ish, isc, osh, osc = _norm(norm)
    _, c = forward(w, x, norm, emulate)
    f = np.float32 if emulate else np.float64
    rz = bf16 if emulate else (lambda a: a)
    P, xn, h1, h2, h3, out = c["P"], c["xn"], c["h1"], c["h2"], c["h3"], c["out"]
    n = len(x)
    t = ((np.asarray(y, f) - osh.astype(f)) / osc.astype(f)).astype(f)
    d = out - t
    loss = float(np.sum(0.5 * d.astype(np.float64) ** 2) / n)
    dout = (d / f(n)).astype(f)
    g = {}
    g["W4"] = h3.T @ dout
    g["b4"] = dout.sum(0)
    dz3 = rz((dout @ P["W4"].T) * (1 - h3 * h3))
    g["W3"] = h2.T @ dz3
    g["b3"] = dz3.sum(0)
    dz2 = rz((dz3 @ P["W3"].T) * (1 - h2 * h2))
    g["W2"] = h1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = rz((dz2 @ P["W2"].T) * (1 - h1 * h1))
    g["W1"] = xn.T @ dz1
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def loss_only(w, x, y, norm=None):
    """float64 loss without rounding (for finite differences)."""
    ish, isc, osh, osc = _norm(norm)
    _, c = forward(w, x, norm, emulate=False)
    t = (np.asarray(y, np.float64) - osh) / osc
    return float(np.sum(0.5 * (c["out"] - t) ** 2) / len(x))
