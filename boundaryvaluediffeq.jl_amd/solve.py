"""``solve`` of the host-side mirror: the reference's entry point
``solve(prob, MIRK4()/MIRK6(); dt, abstol, adaptive, controller, nlsolve_kwargs)``
(Core/BoundaryValueDiffEqCore.jl:48-54 -> MIRK/mirk.jl:49-265 __init,
MIRK/mirk.jl:286-332 solve!), served by the HIP library; and the batched
``EnsembleProblem`` form (MIRKt/Core/ensemble_tests.jl:7-41) as one launch.
"""
import numpy as np

from . import _abi, _lib
from .problems import BVProblem, extract


class DefectControl:
    """Core/calc_errors.jl:54-60."""
    code = _abi.CONTROL_DEFECT

    def __init__(self, defect_threshold=0.1):
        self.defect_threshold = float(defect_threshold)


class HOErrorControl:
    """Core/calc_errors.jl:131: MIRK(p+2) on the same mesh (MIRK4 only here: there is no MIRK8)."""
    code = _abi.GLOBAL_HO


class REErrorControl:
    """Core/calc_errors.jl:145: Richardson extrapolation with the same method on the halved mesh."""
    code = _abi.GLOBAL_RE


class GlobalErrorControl:
    """Core/calc_errors.jl:67-73."""
    code = _abi.CONTROL_GLOBAL

    def __init__(self, method=None):
        self.method = method if method is not None else HOErrorControl()


class SequentialErrorControl:
    """Core/calc_errors.jl:80-87: the defect first, then the global error once the defect is satisfied."""
    code = _abi.CONTROL_SEQUENTIAL

    def __init__(self, defect=None, global_error=None):
        self.defect = defect if defect is not None else DefectControl()
        self.global_error = global_error if global_error is not None else GlobalErrorControl()


class HybridErrorControl:
    """Core/calc_errors.jl:94-107: DE * defect + GE * global error."""
    code = _abi.CONTROL_HYBRID

    def __init__(self, DE=1.0, GE=1.0, defect=None, global_error=None):
        self.DE = float(DE)
        self.GE = float(GE)
        self.defect = defect if defect is not None else DefectControl()
        self.global_error = global_error if global_error is not None else GlobalErrorControl()


def _controller_fields(controller):
    """(controller, global_method, defect_threshold, DE, GE) of bvp_opts for a controller object."""
    c = controller if controller is not None else DefectControl()
    if isinstance(c, DefectControl):
        return c.code, _abi.GLOBAL_HO, c.defect_threshold, 1.0, 1.0
    if isinstance(c, GlobalErrorControl):
        return c.code, c.method.code, DefectControl().defect_threshold, 1.0, 1.0
    if isinstance(c, SequentialErrorControl):
        return c.code, c.global_error.method.code, c.defect.defect_threshold, 1.0, 1.0
    if isinstance(c, HybridErrorControl):
        return c.code, c.global_error.method.code, c.defect.defect_threshold, c.DE, c.GE
    raise TypeError("controller must be DefectControl, GlobalErrorControl, SequentialErrorControl or "
                    "HybridErrorControl")


class BackTracking:
    """LineSearch.jl's BackTracking (Armijo c1 = 1e-4, quadratic then cubic
    interpolation, steps kept within [0.1, 0.5] of the previous one)."""


class NewtonRaphson:
    """NonlinearSolve's NewtonRaphson, undamped or with a BackTracking line search."""

    def __init__(self, linesearch=None):
        if linesearch is not None and not isinstance(linesearch, BackTracking):
            raise ValueError("the line search must be BackTracking()")
        self.linesearch = linesearch

    @property
    def stages(self):
        return _abi.STAGE_BACKTRACKING if self.linesearch is not None else _abi.STAGE_NEWTON


class TrustRegion:
    """NonlinearSolve's TrustRegion (dogleg step, "Simple" radius update, defaults)."""
    stages = _abi.STAGE_TRUST_REGION


class NonlinearSolvePolyAlgorithm:
    """Stages tried in order, each from the same start; the default nonlinear solver
    is NewtonRaphson -> NewtonRaphson(BackTracking) -> TrustRegion
    (Core/default_internal_solve.jl:31-45).  The stages run in that fixed order."""

    def __init__(self, algs):
        self.algs = tuple(algs)

    @property
    def stages(self):
        m = 0
        for a in self.algs:
            if a.stages <= m:
                raise ValueError("stages must be NewtonRaphson, NewtonRaphson(BackTracking), TrustRegion in order")
            m |= a.stages
        return m


def nlsolve_stages(nlsolve):
    """bvp_opts.nlsolve_stages of an `nlsolve` choice (None: the default polyalgorithm)."""
    if nlsolve is None:
        return _abi.STAGES_DEFAULT
    return nlsolve.stages


class _MIRK:
    order = 0

    def __init__(self, nlsolve=None, jac_alg=None, defect_threshold=0.1, max_num_subintervals=3000,
                 precision=_abi.PREC_FP64, refine_iters=0, transfer_mode=_abi.TRANSFER_REFERENCE,
                 warm_start=None):
        # fields of MIRK/algorithms.jl:55-61; `defect_threshold` is stored but, as in the
        # reference, the controller's threshold is the one used.
        self.nlsolve = nlsolve
        self.jac_alg = jac_alg
        self.defect_threshold = defect_threshold
        self.max_num_subintervals = int(max_num_subintervals)
        self.precision = precision
        self.refine_iters = refine_iters
        self.transfer_mode = transfer_mode
        self.warm_start = warm_start

    def __repr__(self):
        return f"MIRK{self.order}()"


class MIRK4(_MIRK):
    order = 4
    code = _abi.MIRK4


class MIRK6(_MIRK):
    order = 6
    code = _abi.MIRK6


class _FIRK:
    """A FIRK collocation method (lib/BoundaryValueDiffEqFIRK/src/algorithms.jl:4-411).
    nested_nlsolve = false (the reference's default) solves the expanded system of node
    values and stages; nested_nlsolve = true solves for the node values only, each
    interval's stages from an inner NewtonRaphson (lib/BoundaryValueDiffEqFIRK/src/
    collocation.jl:118-205), whose tolerance nested_nlsolve_kwargs["abstol"] sets
    (default eps^(4/5)).  On the GPU nlsolve = None runs the NewtonRaphson stages of the
    polyalgorithm (plain, then with BackTracking); a nlsolve that names the TrustRegion
    stage adds it (expanded form)."""
    code = 0
    stages = 0
    order = 0

    def __init__(self, nlsolve=None, jac_alg=None, nested_nlsolve=False, nested_nlsolve_kwargs=None,
                 defect_threshold=0.1, max_num_subintervals=3000):
        self.nlsolve = nlsolve
        self.jac_alg = jac_alg
        self.nested_nlsolve = bool(nested_nlsolve)
        self.nested_nlsolve_kwargs = dict(nested_nlsolve_kwargs or {})
        unknown = set(self.nested_nlsolve_kwargs) - {"abstol"}
        if unknown:
            raise ValueError(f"nested_nlsolve_kwargs: only abstol is supported, got {sorted(unknown)}")
        if self.nested_nlsolve:
            self.code = type(self).code + _abi.FIRK_NESTED
        self.defect_threshold = float(defect_threshold)
        self.max_num_subintervals = int(max_num_subintervals)
        self.precision = _abi.PREC_FP64
        self.refine_iters = 0
        self.transfer_mode = _abi.TRANSFER_CLEAN
        self.warm_start = None

    def __repr__(self):
        return type(self).__name__ + ("(nested_nlsolve=True)" if self.nested_nlsolve else "()")


def _firk_class(name, code, order):
    return type(name, (_FIRK,), {"code": code, "stages": code % 100, "order": order,
                                 "__doc__": f"{name}: {code % 100} stages, order {order}."})


RadauIIa1, RadauIIa2, RadauIIa3, RadauIIa5, RadauIIa7 = (
    _firk_class(f"RadauIIa{s}", _abi.RADAU_IIA(s), 2 * s - 1) for s in (1, 2, 3, 5, 7))
LobattoIIIa2, LobattoIIIa3, LobattoIIIa4, LobattoIIIa5 = (
    _firk_class(f"LobattoIIIa{s}", _abi.LOBATTO_IIIA(s), 2 * s - 2) for s in (2, 3, 4, 5))
LobattoIIIb2, LobattoIIIb3, LobattoIIIb4, LobattoIIIb5 = (
    _firk_class(f"LobattoIIIb{s}", _abi.LOBATTO_IIIB(s), 2 * s - 2) for s in (2, 3, 4, 5))
LobattoIIIc2, LobattoIIIc3, LobattoIIIc4, LobattoIIIc5 = (
    _firk_class(f"LobattoIIIc{s}", _abi.LOBATTO_IIIC(s), 2 * s - 2) for s in (2, 3, 4, 5))


class BVPSolution:
    def __init__(self, raw, prob, alg, spec=None, device=0):
        self.prob = prob
        self.alg = alg
        self.t = raw.t
        self.u = [raw.u[i].copy() for i in range(raw.u.shape[0])]
        self.stages = raw.stages  # FIRK: the stage values (N, s, M) of the solution
        self.retcode = _abi.RETCODE_NAMES.get(raw.retcode, str(raw.retcode))
        self.resid = raw.resid_norm
        self.err_vs_fp64 = raw.err_vs_fp64
        self.stats = {"newton_iters": raw.newton_iters, "outer_iters": raw.outer_iters,
                      "defect_norm": raw.defect_norm, "N": raw.N, "fallback_used": raw.fallback_used}
        self._spec = spec
        self._device = device

    def __array__(self, dtype=None):
        return np.asarray(self.u, dtype=dtype)

    def __call__(self, t, deriv=0, idxs=None):
        """Dense output sol(t) (deriv = 0) and sol(t, Val{1}) (deriv = 1) on the GPU: the
        method's continuous extension (MIRK/interpolation.jl:17-206,
        FIRK/interpolation.jl:229-316); idxs selects components."""
        if self._spec is None:
            raise ValueError("this solution carries no problem description for dense output")
        scalar = np.ndim(t) == 0
        tq = np.atleast_1d(np.asarray(t, dtype=np.float64))
        spec = self._spec.with_state(self.t, np.asarray(self.u, dtype=np.float64).reshape(-1))
        ctx = _lib.context(self._device)
        if isinstance(self.alg, _FIRK):
            ((u, du),) = _lib.firk_interpolate(ctx, self.alg.code, [spec], [tq], [self.stages])
        else:
            ((u, du),) = _lib.interpolate(ctx, self.alg.code, [spec], [tq])
        v = du if deriv == 1 else u
        if idxs is not None:
            v = v[:, idxs]
        return v[0] if scalar else v


def successful_retcode(sol):
    return sol.retcode == "Success"


class EnsembleProblem:
    """SciMLBase.EnsembleProblem: trajectory i solves prob_func(prob, i, repeat)."""

    def __init__(self, prob, prob_func=None, output_func=None):
        self.prob = prob
        self.prob_func = prob_func or (lambda p, i, repeat: p)
        self.output_func = output_func


class EnsembleSolution(list):
    @property
    def converged(self):
        return all(successful_retcode(s) for s in self)


def make_opts(alg, abstol=1e-6, adaptive=True, controller=None, nlsolve_kwargs=None, report_err=False):
    o = _abi.default_opts(alg.code)
    o.abstol = float(abstol)
    nk = dict(nlsolve_kwargs or {})
    o.newton_abstol = float(nk.get("abstol", abstol))   # MIRK/mirk.jl:51
    o.max_newton_iters = int(nk.get("maxiters", 1000))
    o.adaptive = 1 if adaptive else 0
    (o.controller, o.global_method, o.defect_threshold, o.hybrid_de, o.hybrid_ge) = _controller_fields(controller)
    o.max_num_subintervals = alg.max_num_subintervals
    o.precision = alg.precision
    o.refine_iters = alg.refine_iters
    o.transfer_mode = alg.transfer_mode
    o.nlsolve_stages = nlsolve_stages(alg.nlsolve)
    o.report_err = 1 if report_err else 0
    if isinstance(alg, _FIRK):
        # the FIRK path keeps its own defect threshold (FIRK/firk.jl:596) and runs the
        # NewtonRaphson stages
        if controller is not None and not isinstance(controller, DefectControl):
            raise ValueError("FIRK methods use DefectControl")
        o.defect_threshold = alg.defect_threshold
        o.nest_abstol = float(alg.nested_nlsolve_kwargs.get("abstol", 0.0))
        if alg.nlsolve is None:
            o.nlsolve_stages = _abi.STAGE_NEWTON | _abi.STAGE_BACKTRACKING
    return o


def solve(prob, alg, *, dt=0.0, abstol=1e-6, adaptive=True, controller=None, nlsolve_kwargs=None,
          trajectories=None, device=0, report_err=False):
    """Solve a BVProblem / TwoPointBVProblem (or an EnsembleProblem) with MIRK4/MIRK6 or a
    FIRK method (RadauIIa*, LobattoIII*) on the GPU.
    report_err: also return, per solution, err_vs_fp64 (the FP64 Newton correction at the result)."""
    opts = make_opts(alg, abstol, adaptive, controller, nlsolve_kwargs, report_err)
    ctx = _lib.context(device)
    if isinstance(prob, EnsembleProblem):
        if trajectories is None:
            raise ValueError("trajectories is required for an EnsembleProblem")
        probs = [prob.prob_func(prob.prob, i + 1, 1) for i in range(trajectories)]
        specs = [extract(p, dt) for p in probs]
        raws = _lib.solve_batch(ctx, opts, specs)
        return EnsembleSolution(BVPSolution(r, p, alg, s, device) for r, p, s in zip(raws, probs, specs))
    if not isinstance(prob, BVProblem):
        raise TypeError("prob must be a BVProblem, TwoPointBVProblem or EnsembleProblem")
    spec = extract(prob, dt)
    raw = _lib.solve_batch(ctx, opts, [spec])[0]
    return BVPSolution(raw, prob, alg, spec, device)
