"""ctypes mirror of ``include/bvp_mirk.h`` (structs and constants).

The constants are checked against the header by ``tests/test_abi.py``.
"""
import ctypes as C

MIRK4 = 4
MIRK6 = 6


FIRK_NESTED = 1000


def RADAU_IIA(s):
    return 100 + s


def LOBATTO_IIIA(s):
    return 200 + s


def LOBATTO_IIIB(s):
    return 300 + s


def LOBATTO_IIIC(s):
    return 400 + s


def is_firk(method):
    return method >= 100


def firk_stages(method):
    return method % 100


def is_firk_nested(method):
    return method >= FIRK_NESTED

FAMILY_LINEAR_TANH = 0
FAMILY_PENDULUM = 1
FAMILY_BRATU = 2
FAMILY_SINGPERT = 3
FAMILY_SCALAR_TU = 4
FAMILY_TURNING = 5
FAMILY_SWIRL = 6

BC_TWO_POINT = 0
BC_STANDARD = 1

RET_SUCCESS = 0
RET_FAILURE = 1
RET_MAXITERS = 2
RET_UNSTABLE = 3
RETCODE_NAMES = {0: "Success", 1: "Failure", 2: "MaxIters", 3: "Unstable"}

PREC_FP64 = 0
PREC_FP32_REFINE64 = 1

STAGE_NEWTON = 1
STAGE_BACKTRACKING = 2
STAGE_TRUST_REGION = 4
STAGES_DEFAULT = 7

CONTROL_DEFECT = 0
CONTROL_GLOBAL = 1
CONTROL_SEQUENTIAL = 2
CONTROL_HYBRID = 3
GLOBAL_HO = 0
GLOBAL_RE = 1

TRANSFER_REFERENCE = 0
TRANSFER_CLEAN = 1

WARM_NONE = 0
WARM_MLP = 1

MLP_NPARAM = 133378
FEAT_PARAM = 0
FEAT_LOG10_PARAM = 1

OK = 0
EINVAL = -1
EHIP = -2
ENOMEM = -3
ENOTSUP = -4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Instance(C.Structure):
    _fields_ = [
        ("family", C.c_int32),
        ("M", C.c_int32),
        ("t0", C.c_double),
        ("t1", C.c_double),
        ("params", _dp),
        ("n_params", C.c_int32),
        ("bc_kind", C.c_int32),
        ("n_bc", C.c_int32),
        ("bc_t", _dp),
        ("bc_comp", _ip),
        ("bc_val", _dp),
        ("N", C.c_int32),
        ("mesh", _dp),
        ("guess", _dp),
        ("u0", _dp),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("abstol", C.c_double),
        ("newton_abstol", C.c_double),
        ("adaptive", C.c_int32),
        ("defect_threshold", C.c_double),
        ("max_num_subintervals", C.c_int32),
        ("max_newton_iters", C.c_int32),
        ("nlsolve_stages", C.c_int32),
        ("transfer_mode", C.c_int32),
        ("precision", C.c_int32),
        ("refine_iters", C.c_int32),
        ("warm_start", C.c_int32),
        ("report_err", C.c_int32),
        ("controller", C.c_int32),
        ("global_method", C.c_int32),
        ("hybrid_de", C.c_double),
        ("hybrid_ge", C.c_double),
        ("nest_abstol", C.c_double),
    ]


class Result(C.Structure):
    _fields_ = [
        ("retcode", C.c_int32),
        ("N_final", C.c_int32),
        ("newton_iters", C.c_int32),
        ("outer_iters", C.c_int32),
        ("fallback_used", C.c_int32),
        ("capacity", C.c_int32),
        ("defect_norm", C.c_double),
        ("resid_norm", C.c_double),
        ("err_vs_fp64", C.c_double),
        ("mesh", _dp),
        ("u", _dp),
        ("stages", _dp),
    ]


class TrainOpts(C.Structure):
    _fields_ = [
        ("epochs", C.c_int32),
        ("batch", C.c_int32),
        ("lr", C.c_float),
        ("lr_min", C.c_float),
        ("weight_decay", C.c_float),
        ("warmup", C.c_int32),
        ("seed", C.c_uint64),
        ("use_graph", C.c_int32),
        ("first_epoch", C.c_int32),
        ("total_epochs", C.c_int32),
    ]


def default_opts(method=MIRK4):
    """Defaults of the reference (MIRK/mirk.jl:49-53, MIRK/algorithms.jl:55-61,
    Core/calc_errors.jl:54-60)."""
    o = Opts()
    o.method = method
    o.abstol = 1e-6
    o.newton_abstol = 0.0
    o.adaptive = 1
    o.defect_threshold = 0.1
    o.max_num_subintervals = 3000
    o.max_newton_iters = 1000
    o.nlsolve_stages = STAGES_DEFAULT
    o.transfer_mode = TRANSFER_REFERENCE
    o.precision = PREC_FP64
    o.refine_iters = 0
    o.warm_start = WARM_NONE
    o.report_err = 0
    o.controller = CONTROL_DEFECT
    o.global_method = GLOBAL_HO
    o.hybrid_de = 1.0
    o.hybrid_ge = 1.0
    o.nest_abstol = 0.0
    return o


def dptr(a):
    """Pointer to a contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return _dp()
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return _ip()
    return a.ctypes.data_as(_ip)
