"""Problem types of the host-side mirror of the reference interface.

The reference builds problems as ``BVProblem(f!, bc!, u0, tspan, p)`` and
``TwoPointBVProblem(f!, (bca!, bcb!), u0, tspan, p)`` (SciMLBase), where
``f!`` and ``bc!`` are Julia closures.  A closure cannot run on the GPU, so
here ``f`` is a :class:`BVPFunction` descriptor naming a right-hand side that
is compiled into the device code, and a boundary condition is data: rows
``(t, component, value)`` meaning ``sol(t)[component] - value``.

Problem details are extracted as the reference does in
``__extract_problem_details`` / ``__extract_mesh`` / ``__initial_guess_on_mesh``
(Core/utils.jl:330-388, 694-704, 750-777).
"""
from fractions import Fraction
import inspect
import math
import warnings

import numpy as np

from . import _abi


class BVPFunction:
    """Right-hand side ``f(u, p, t)`` resident in the device code.

    ``family`` is one of the ``_abi.FAMILY_*`` codes; ``M`` the state size.
    Calling it evaluates ``f`` on the host with numpy (for users and for
    building inputs), not on the GPU.
    """

    def __init__(self, family, M, name):
        self.family = int(family)
        self.M = int(M)
        self.name = name

    def __repr__(self):
        return f"BVPFunction({self.name}, M={self.M})"

    def __call__(self, u, p, t):
        u = np.asarray(u, dtype=np.float64)
        p = np.asarray(p, dtype=np.float64)
        fam = self.family
        if fam == _abi.FAMILY_LINEAR_TANH:
            M = self.M
            A = p[1:1 + M * M].reshape(M, M)
            return A @ u + p[0] * np.tanh(u)
        if fam == _abi.FAMILY_PENDULUM:
            return np.array([u[1], -p[0] * np.sin(u[0])])
        if fam == _abi.FAMILY_BRATU:
            return np.array([u[1], -p[0] * np.exp(u[0])])
        if fam == _abi.FAMILY_SINGPERT:
            return np.array([u[1], (u[0] - t * u[1]) / p[0]])
        if fam == _abi.FAMILY_SCALAR_TU:
            return np.array([t * u[0]])
        if fam == _abi.FAMILY_TURNING:
            pi = math.pi
            return np.array([u[1], -t / p[0] * u[1] - pi * pi * math.cos(pi * t)
                             - pi * t / p[0] * math.sin(pi * t)])
        if fam == _abi.FAMILY_SWIRL:
            return np.array([u[1], (u[0] * u[3] - u[2] * u[1]) / p[0], u[3], u[4], u[5],
                             (-u[2] * u[5] - u[0] * u[1]) / p[0]])
        raise ValueError(f"unknown family {fam}")


def LinearTanhFunction(M):
    """u' = A u + c tanh(u); p = [c, A (row-major)]."""
    return BVPFunction(_abi.FAMILY_LINEAR_TANH, M, "linear_tanh")


def PendulumFunction():
    """theta'' = -g sin(theta) as a first-order system; p = [g]."""
    return BVPFunction(_abi.FAMILY_PENDULUM, 2, "pendulum")


def BratuFunction():
    """u'' + lambda e^u = 0 as a first-order system; p = [lambda]."""
    return BVPFunction(_abi.FAMILY_BRATU, 2, "bratu")


def SingPertFunction():
    """eps u'' + t u' - u = 0 as a first-order system; p = [eps]."""
    return BVPFunction(_abi.FAMILY_SINGPERT, 2, "singpert")


def ScalarTUFunction():
    """u' = t u (scalar)."""
    return BVPFunction(_abi.FAMILY_SCALAR_TU, 1, "scalar_tu")


def TurningFunction():
    """u'' = -t u'/eps - pi^2 cos(pi t) - pi t sin(pi t)/eps; p = [eps]."""
    return BVPFunction(_abi.FAMILY_TURNING, 2, "turning")


def SwirlFunction():
    """Swirling Flow III (MIRKt/Core/mirk_basic_tests.jl:318-327), M = 6; p = [eps]."""
    return BVPFunction(_abi.FAMILY_SWIRL, 6, "swirl")


class DiffEqArray:
    """Initial guess on a user mesh (RecursiveArrayTools.DiffEqArray)."""

    def __init__(self, u, t):
        self.u = [np.asarray(x, dtype=np.float64) for x in u]
        self.t = np.asarray(t, dtype=np.float64)


class BVProblem:
    """Standard BVP: ``bc`` is a sequence of rows ``(t, comp, value)``, each
    the residual ``sol(t)[comp] - value`` of the continuous MIRK solution."""

    bc_kind = _abi.BC_STANDARD

    def __init__(self, f, bc, u0, tspan, p=()):
        self.f = f
        self.bc_rows = [(float(t), int(c), float(v)) for (t, c, v) in bc]
        self.u0 = u0
        self.tspan = (float(tspan[0]), float(tspan[1]))
        self.p = np.atleast_1d(np.asarray(p, dtype=np.float64))


class TwoPointBVProblem(BVProblem):
    """Two-point BVP: ``bc = (bca, bcb)``, each a sequence of rows
    ``(comp, value)`` evaluated at ``tspan[0]`` and ``tspan[1]``."""

    bc_kind = _abi.BC_TWO_POINT

    def __init__(self, f, bc, u0, tspan, p=()):
        bca, bcb = bc
        t0, t1 = float(tspan[0]), float(tspan[1])
        rows = [(t0, int(c), float(v)) for (c, v) in bca] + [(t1, int(c), float(v)) for (c, v) in bcb]
        super().__init__(f, rows, u0, tspan, p)


def cld(x, y):
    """Julia ``cld`` on floats: ceil of the exact quotient."""
    return math.ceil(Fraction(x) / Fraction(y))


def uniform_mesh(t0, t1, N):
    """``collect(range(t0; stop=t1, length=N+1))`` (Core/utils.jl:694-700)."""
    k = np.arange(N + 1, dtype=np.longdouble)
    m = ((np.longdouble(t0) * (N - k) + np.longdouble(t1) * k) / np.longdouble(N)).astype(np.float64)
    m[0] = t0
    m[N] = t1
    return m


class InstanceSpec:
    """One problem instance as plain arrays, ready for the C ABI."""

    def __init__(self, family, M, t0, t1, params, bc_kind, bc_t, bc_comp, bc_val, N,
                 mesh=None, guess=None, u0=None):
        self.family = int(family)
        self.M = int(M)
        self.t0 = float(t0)
        self.t1 = float(t1)
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.bc_kind = int(bc_kind)
        self.bc_t = np.ascontiguousarray(bc_t, dtype=np.float64)
        self.bc_comp = np.ascontiguousarray(bc_comp, dtype=np.int32)
        self.bc_val = np.ascontiguousarray(bc_val, dtype=np.float64)
        self.N = int(N)
        self.mesh = None if mesh is None else np.ascontiguousarray(mesh, dtype=np.float64)
        self.guess = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64).reshape(-1)
        self.u0 = None if u0 is None else np.ascontiguousarray(u0, dtype=np.float64).reshape(-1)
        if self.mesh is not None and self.mesh.shape[0] != self.N + 1:
            raise ValueError("mesh must have N+1 nodes")
        if self.guess is not None and self.guess.shape[0] != self.M * (self.N + 1):
            raise ValueError("guess must have M*(N+1) entries")
        if self.bc_t.shape[0] != self.M:
            raise ValueError("the number of BC rows must equal M")

    def with_state(self, mesh, y):
        """Copy of this instance with a new mesh and guess (N from the mesh)."""
        mesh = np.ascontiguousarray(mesh, dtype=np.float64)
        return InstanceSpec(self.family, self.M, self.t0, self.t1, self.params, self.bc_kind,
                            self.bc_t, self.bc_comp, self.bc_val, mesh.shape[0] - 1, mesh,
                            np.ascontiguousarray(y, dtype=np.float64).reshape(-1), self.u0)

    def full_mesh(self):
        return self.mesh if self.mesh is not None else uniform_mesh(self.t0, self.t1, self.N)

    def full_guess(self):
        if self.guess is not None:
            return self.guess
        return np.tile(self.u0, self.N + 1)

    def ctypes(self):
        """An ``_abi.Instance`` pointing into this object's arrays (keep ``self`` alive)."""
        s = _abi.Instance()
        s.family = self.family
        s.M = self.M
        s.t0 = self.t0
        s.t1 = self.t1
        s.params = _abi.dptr(self.params)
        s.n_params = self.params.shape[0]
        s.bc_kind = self.bc_kind
        s.n_bc = self.bc_t.shape[0]
        s.bc_t = _abi.dptr(self.bc_t)
        s.bc_comp = _abi.iptr(self.bc_comp)
        s.bc_val = _abi.dptr(self.bc_val)
        s.N = self.N
        s.mesh = _abi.dptr(self.mesh)
        s.guess = _abi.dptr(self.guess)
        s.u0 = _abi.dptr(self.u0)
        return s


def instance_array(specs):
    """ctypes array of ``_abi.Instance`` for a list of specs."""
    arr = (_abi.Instance * len(specs))()
    for i, s in enumerate(specs):
        arr[i] = s.ctypes()
    return arr


def _initial_guess(f, p, t):
    """An initial-guess function at t with the (p, t) calling convention, or the
    deprecated single-argument t form (Core/utils.jl:395-410)."""
    try:
        sig = inspect.signature(f)
        npos = sum(1 for q in sig.parameters.values()
                   if q.kind in (q.POSITIONAL_ONLY, q.POSITIONAL_OR_KEYWORD) and q.default is q.empty)
        var = any(q.kind == q.VAR_POSITIONAL for q in sig.parameters.values())
    except (TypeError, ValueError):
        npos, var = 2, False
    if npos == 2 or var:
        return f(p, t)
    if npos == 1:
        warnings.warn("initial guess function must take 2 inputs `(p, t)` instead of just `t`",
                      DeprecationWarning, stacklevel=3)
        return f(t)
    raise TypeError("`initial_guess` must be a function of the form `f(p, t)`")


def extract(prob, dt=0.0):
    """Problem details, mesh and guess of ``prob`` (Core/utils.jl:330-388, 694-704, 750-777).

    u0 may be a constant vector (or number), a vector of per-node vectors, a
    DiffEqArray or any solution object carrying ``t`` and ``u`` (ODESolution: its
    nodes become the mesh), or a function ``f(p, t)`` (or the deprecated ``f(t)``)
    evaluated at the nodes of the mesh from dt (Core/utils.jl:750-777)."""
    t0, t1 = prob.tspan
    M = prob.f.M
    u0 = prob.u0
    mesh = None
    guess = None
    const = None
    if isinstance(u0, DiffEqArray) or (hasattr(u0, "t") and hasattr(u0, "u")):
        N = len(u0.u) - 1
        mesh = np.asarray(u0.t, dtype=np.float64).copy()
        guess = np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in u0.u])
    elif callable(u0):
        if not dt > 0:
            raise ValueError("dt must be positive")
        N = cld(t1 - t0, dt)
        nodes = uniform_mesh(t0, t1, N)
        guess = np.concatenate([np.asarray(_initial_guess(u0, prob.p, float(t)), dtype=np.float64).reshape(M)
                                for t in nodes])
    else:
        a = np.asarray(u0, dtype=np.float64)
        if a.ndim == 2 or (a.ndim == 1 and isinstance(u0, (list, tuple)) and len(u0) > 0
                           and np.ndim(u0[0]) == 1):
            a = np.asarray(u0, dtype=np.float64).reshape(-1, M)
            N = a.shape[0] - 1
            guess = a.reshape(-1).copy()
        else:
            if not dt > 0:
                raise ValueError("dt must be positive")
            N = cld(t1 - t0, dt)
            const = np.atleast_1d(a).astype(np.float64)
    bc_t = [r[0] for r in prob.bc_rows]
    bc_c = [r[1] for r in prob.bc_rows]
    bc_v = [r[2] for r in prob.bc_rows]
    return InstanceSpec(prob.f.family, M, t0, t1, prob.p, prob.bc_kind, bc_t, bc_c, bc_v, N,
                        mesh=mesh, guess=guess, u0=const)
