"""Problem builders shared by the oracle and GPU tests."""
import math

import numpy as np

import boundaryvaluediffeq_amd as bvp
from boundaryvaluediffeq_amd import _abi


def bratu(lam, N=None, dt=None, a=0.0, b=0.0, u0=(0.0, 0.0), mesh=None, guess=None):
    prob = bvp.TwoPointBVProblem(bvp.BratuFunction(), ([(0, a)], [(0, b)]), list(u0), (0.0, 1.0), [lam])
    if N is not None:
        dt = 1.0 / N
    spec = bvp.extract(prob, dt=dt)
    if mesh is not None or guess is not None:
        m = mesh if mesh is not None else spec.full_mesh()
        g = guess if guess is not None else spec.full_guess()
        spec = spec.with_state(m, g)
    return spec


def pendulum(dt=0.05):
    """README simple pendulum (README.md:23-38; benchmark/simple_pendulum.jl:1-34)."""
    prob = bvp.BVProblem(bvp.PendulumFunction(), [(math.pi / 4, 0, -math.pi / 2), (math.pi / 2, 0, math.pi / 2)],
                         [math.pi / 2, math.pi / 2], (0.0, math.pi / 2), [9.81])
    return bvp.extract(prob, dt=dt)


def singpert(eps, N=32, u0=(0.0, 0.0)):
    prob = bvp.TwoPointBVProblem(bvp.SingPertFunction(), ([(0, 1.0)], [(0, 2.0)]), list(u0), (-1.0, 1.0), [eps])
    return bvp.extract(prob, dt=2.0 / N)


def turning(eps, dt=0.01):
    """test/misc/adaptivity_tests.jl:4-17 (Standard BCs at the end points)."""
    prob = bvp.BVProblem(bvp.TurningFunction(), [(-1.0, 0, -2.0), (1.0, 0, 0.0)], [1.0, 0.0], (-1.0, 1.0), [eps])
    return bvp.extract(prob, dt=dt)


def swirl(eps=0.01, dt=0.01, u0=None):
    """Swirling Flow III (MIRKt/Core/mirk_basic_tests.jl:315-345): M = 6 on [0, 1],
    u(0)[1] = -1, u(0)[3] = u(0)[4] = 0, u(1)[1] = 1, u(1)[3] = u(1)[4] = 0, zero guess."""
    bc = [(0.0, 0, -1.0), (0.0, 2, 0.0), (0.0, 3, 0.0), (1.0, 0, 1.0), (1.0, 2, 0.0), (1.0, 3, 0.0)]
    prob = bvp.BVProblem(bvp.SwirlFunction(), bc, list(u0 if u0 is not None else np.zeros(6)), (0.0, 1.0), [eps])
    return bvp.extract(prob, dt=dt)


def linear(A, c=0.0, tspan=(0.0, 5.0), bc=None, dt=0.2, u0=None, standard=False):
    A = np.asarray(A, dtype=np.float64)
    M = A.shape[0]
    p = np.concatenate([[c], A.reshape(-1)])
    u0 = np.zeros(M) if u0 is None else u0
    if standard:
        prob = bvp.BVProblem(bvp.LinearTanhFunction(M), bc, list(u0), tspan, p)
    else:
        prob = bvp.TwoPointBVProblem(bvp.LinearTanhFunction(M), bc, list(u0), tspan, p)
    return bvp.extract(prob, dt=dt)


def linear_tanh_c3(rng, M=32, N=64, spread=(0.5, 50.0)):
    """One instance of the n=32 family u' = A u + 0.1 tanh(u) with A = Q diag(mu) Q^T,
    mu log-spread in +-[0.5, 50], separated BCs (first M/2 at t=0, last M/2 at t=1)."""
    G = rng.standard_normal((M, M))
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diag(R))
    half = M // 2
    mags = np.logspace(math.log10(spread[0]), math.log10(spread[1]), half)
    mu = np.concatenate([-mags, mags])
    A = (Q * mu) @ Q.T
    alpha = rng.standard_normal(half)
    beta = rng.standard_normal(M - half)
    bca = [(k, alpha[k]) for k in range(half)]
    bcb = [(half + k, beta[k]) for k in range(M - half)]
    return linear(A, c=0.1, tspan=(0.0, 1.0), bc=(bca, bcb), dt=1.0 / N)


def random_state(spec, rng, scale=0.3):
    """Same instance with a random guess (for residual / Jacobian checks)."""
    m = spec.full_mesh()
    y = rng.standard_normal(spec.M * (spec.N + 1)) * scale
    return spec.with_state(m, y)


def random_mesh(spec, rng):
    """Same instance on a random strictly increasing mesh with the same end points."""
    N = spec.N
    h = rng.uniform(0.5, 1.5, N)
    t = np.concatenate([[0.0], np.cumsum(h)])
    t = spec.t0 + (spec.t1 - spec.t0) * t / t[-1]
    t[0], t[-1] = spec.t0, spec.t1
    return spec.with_state(t, spec.full_guess())
