"""MI355X-native MIRK and FIRK collocation solvers behind the reference's interface.

Host-side mirror of ``solve(prob, MIRK4()/MIRK6()/RadauIIa5()/...; dt, abstol)`` from
SciML/BoundaryValueDiffEq.jl; the compute runs in the HIP library
``_build/libbvp_mirk.so`` through the C ABI of ``include/bvp_mirk.h``.
"""
from . import _abi
from .problems import (BVPFunction, BVProblem, TwoPointBVProblem, DiffEqArray, InstanceSpec,
                       LinearTanhFunction, PendulumFunction, BratuFunction, SingPertFunction,
                       ScalarTUFunction, TurningFunction, SwirlFunction, extract, uniform_mesh, cld, instance_array)
from .solve import (MIRK4, MIRK6, DefectControl, GlobalErrorControl, SequentialErrorControl, HybridErrorControl,
                    HOErrorControl, REErrorControl, EnsembleProblem, EnsembleSolution, BVPSolution, solve,
                    successful_retcode, make_opts, NewtonRaphson, BackTracking, TrustRegion,
                    NonlinearSolvePolyAlgorithm, nlsolve_stages, RadauIIa1, RadauIIa2, RadauIIa3, RadauIIa5,
                    RadauIIa7, LobattoIIIa2, LobattoIIIa3, LobattoIIIa4, LobattoIIIa5, LobattoIIIb2, LobattoIIIb3,
                    LobattoIIIb4, LobattoIIIb5, LobattoIIIc2, LobattoIIIc3, LobattoIIIc4, LobattoIIIc5)

__all__ = [
    "BVPFunction", "BVProblem", "TwoPointBVProblem", "DiffEqArray", "InstanceSpec",
    "LinearTanhFunction", "PendulumFunction", "BratuFunction", "SingPertFunction",
    "ScalarTUFunction", "TurningFunction", "SwirlFunction", "extract", "uniform_mesh", "cld", "instance_array",
    "MIRK4", "MIRK6", "DefectControl", "GlobalErrorControl", "SequentialErrorControl", "HybridErrorControl",
    "HOErrorControl", "REErrorControl", "EnsembleProblem", "EnsembleSolution", "BVPSolution", "solve",
    "successful_retcode", "make_opts", "NewtonRaphson", "BackTracking", "TrustRegion",
    "NonlinearSolvePolyAlgorithm", "nlsolve_stages",
    "RadauIIa1", "RadauIIa2", "RadauIIa3", "RadauIIa5", "RadauIIa7", "LobattoIIIa2", "LobattoIIIa3", "LobattoIIIa4",
    "LobattoIIIa5", "LobattoIIIb2", "LobattoIIIb3", "LobattoIIIb4", "LobattoIIIb5", "LobattoIIIc2", "LobattoIIIc3",
    "LobattoIIIc4", "LobattoIIIc5",
]
