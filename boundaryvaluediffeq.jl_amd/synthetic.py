"""Seeded synthetic workloads (the batched problem sets the benchmark and the
tests solve).

Generator: splitmix64, one stream per instance, state0 = seed ^ (i * 0x9E3779B97F4A7C15);
U[a,b) = a + (b-a) * (u53 / 2^53);  U{lo..hi} = lo + floor(u * (hi - lo + 1)).
"""
import math

import numpy as np

from . import _abi
from .problems import InstanceSpec, uniform_mesh

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
SEED_BRATU16K = 0xB0A70002
SEED_LINEAR_TANH = 0xB0A70003
SEED_SINGPERT = 0xB0A70005
SEED_BRATU_AB_TRAIN = 0xB0A70004
SEED_BRATU_AB_EVAL = 0xB0A70044
SEED_SINGPERT_TRAIN = 0xB0A70055

# Newton tolerance max|F| of the Bratu workloads when solutions are compared at 1e-9:
# the reference stops Newton at max|F| <= abstol (MIRK/mirk.jl:51), which leaves two
# solves of one instance up to norm(J^-1) * abstol apart (norm(J^-1) reaches a few
# hundred near the fold, lambda -> 3.51); both sides stop here instead
HEADLINE_NEWTON_ABSTOL = 1e-12


def _u01(seed, n, ndraw):
    """n streams x ndraw uniforms in [0, 1) (splitmix64, vectorised with wrapping uint64)."""
    with np.errstate(over="ignore"):
        i = np.arange(n, dtype=np.uint64)
        state = np.uint64(seed) ^ (i * GOLDEN)
        out = np.empty((n, ndraw))
        for k in range(ndraw):
            state = state + GOLDEN
            z = state.copy()
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            out[:, k] = (z >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
    return out


def _uint(u, lo, hi):
    return (lo + np.floor(u * (hi - lo + 1))).astype(np.int64)


def bratu_batch(n=16384, seed=SEED_BRATU16K, lam=(0.1, 3.4), Nrange=(64, 1024)):
    """u'' + lambda e^u = 0 on [0,1], u(0)=u(1)=0 as a first-order system;
    lambda ~ U(0.1, 3.4) and N ~ U{64..1024} per instance, uniform meshes,
    constant zero initial guess."""
    u = _u01(seed, n, 2)
    lams = lam[0] + (lam[1] - lam[0]) * u[:, 0]
    Ns = _uint(u[:, 1], Nrange[0], Nrange[1])
    zero = np.zeros(2)
    specs = []
    for k in range(n):
        specs.append(InstanceSpec(_abi.FAMILY_BRATU, 2, 0.0, 1.0, [lams[k]], _abi.BC_TWO_POINT,
                                  [0.0, 1.0], [0, 0], [0.0, 0.0], int(Ns[k]), None, None, zero))
    return specs, lams, Ns


def bratu_ab_batch(n, seed=SEED_BRATU_AB_EVAL, lam=(0.1, 3.4), ab=(-1.0, 1.0), Nrange=(64, 1024)):
    """Bratu with u(0) = a, u(1) = b: lambda ~ U(0.1, 3.4), a, b ~ U(-1, 1),
    N ~ U{64..1024}, uniform meshes, constant zero guess (u0 = 0).  Part of this
    box has no solution (the solve then ends in Failure or MaxIters)."""
    u = _u01(seed, n, 4)
    lams = lam[0] + (lam[1] - lam[0]) * u[:, 0]
    a = ab[0] + (ab[1] - ab[0]) * u[:, 1]
    b = ab[0] + (ab[1] - ab[0]) * u[:, 2]
    Ns = _uint(u[:, 3], Nrange[0], Nrange[1])
    zero = np.zeros(2)
    specs = [InstanceSpec(_abi.FAMILY_BRATU, 2, 0.0, 1.0, [lams[k]], _abi.BC_TWO_POINT, [0.0, 1.0], [0, 0],
                          [a[k], b[k]], int(Ns[k]), None, None, zero) for k in range(n)]
    return specs, np.stack([lams, a, b], axis=1), Ns


def singpert_batch(n=8192, seed=SEED_SINGPERT, N0=32):
    """eps u'' + t u' - u = 0 on [-1,1], u(-1)=1, u(1)=2; eps ~ LogUniform(1e-4, 1e-1)."""
    u = _u01(seed, n, 1)
    eps = 10.0 ** (-4.0 + 3.0 * u[:, 0])
    zero = np.zeros(2)
    specs = [InstanceSpec(_abi.FAMILY_SINGPERT, 2, -1.0, 1.0, [eps[k]], _abi.BC_TWO_POINT, [-1.0, 1.0], [0, 0],
                          [1.0, 2.0], N0, None, None, zero) for k in range(n)]
    return specs, eps


def pendulum_instance(dt=0.05):
    """README simple pendulum (README.md:23-38): theta'' = -9.81 sin(theta) on
    [0, pi/2], u(pi/4)[1] = -pi/2, u(pi/2)[1] = pi/2, guess [pi/2, pi/2]."""
    from fractions import Fraction
    t1 = math.pi / 2
    N = math.ceil(Fraction(t1) / Fraction(dt))
    return InstanceSpec(_abi.FAMILY_PENDULUM, 2, 0.0, t1, [9.81], _abi.BC_STANDARD, [math.pi / 4, t1], [0, 0],
                        [-math.pi / 2, math.pi / 2], N, None, None, [math.pi / 2, math.pi / 2])


def bratu_exact(lam, t):
    """Lower-branch Bratu solution u(t) = -2 ln[cosh((t-1/2) th/2)/cosh(th/4)], th = sqrt(2 lam) cosh(th/4)."""
    th = 0.0
    for _ in range(400):
        th = math.sqrt(2 * lam) * math.cosh(th / 4)
    return -2 * np.log(np.cosh((np.asarray(t) - 0.5) * th / 2) / math.cosh(th / 4))


def lintanh_batch(n=2048, seed=SEED_LINEAR_TANH, M=32, Nrange=(512, 4096), spread=(0.5, 50.0), c=0.1):
    """u' = A u + c tanh(u) on [0,1] with M components: A = Q diag(mu) Q^T, Q from the
    QR factorisation of a seeded Gaussian matrix (signs fixed), mu = -/+ magnitudes
    log-spaced in [0.5, 50] (half each); separated BCs, the first M/2 components fixed
    at t = 0 and the last M/2 at t = 1 to N(0,1) values; N ~ U{512..4096}; zero guess.
    Gaussian draws use numpy's generator seeded with (seed, i); N uses splitmix64."""
    u = _u01(seed, n, 1)
    Ns = _uint(u[:, 0], Nrange[0], Nrange[1])
    half = M // 2
    mags = np.logspace(math.log10(spread[0]), math.log10(spread[1]), half)
    mu = np.concatenate([-mags, mags])
    zero = np.zeros(M)
    bc_t = np.concatenate([np.zeros(half), np.ones(M - half)])
    bc_c = np.arange(M, dtype=np.int32)
    specs = []
    for i in range(n):
        rng = np.random.default_rng([seed, i])
        G = rng.standard_normal((M, M))
        Q, R = np.linalg.qr(G)
        Q = Q * np.sign(np.diag(R))
        A = (Q * mu) @ Q.T
        bcv = rng.standard_normal(M)
        p = np.concatenate([[c], A.reshape(-1)])
        specs.append(InstanceSpec(_abi.FAMILY_LINEAR_TANH, M, 0.0, 1.0, p, _abi.BC_TWO_POINT, bc_t, bc_c, bcv,
                                  int(Ns[i]), None, None, zero))
    return specs, Ns
