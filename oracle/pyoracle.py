"""ctypes wrapper of the oracle library (``oracle/_build/librefcpu.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by the smoke check of
``__graft_entry__`` and by the CPU-baseline leg of ``bench.py``, always as the
checker or the timed CPU baseline, never by the product.

Instances are passed as any object with the attributes of the C
``bvp_instance`` (family, M, t0, t1, params, bc_kind, bc_t, bc_comp, bc_val,
N, mesh, guess, u0) holding numpy arrays.
"""
import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "librefcpu.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class _Instance(C.Structure):
    _fields_ = [("family", C.c_int32), ("M", C.c_int32), ("t0", C.c_double), ("t1", C.c_double),
                ("params", _dp), ("n_params", C.c_int32), ("bc_kind", C.c_int32), ("n_bc", C.c_int32),
                ("bc_t", _dp), ("bc_comp", _ip), ("bc_val", _dp), ("N", C.c_int32), ("mesh", _dp),
                ("guess", _dp), ("u0", _dp)]


class Opts(C.Structure):
    _fields_ = [("method", C.c_int32), ("abstol", C.c_double), ("newton_abstol", C.c_double),
                ("adaptive", C.c_int32), ("defect_threshold", C.c_double),
                ("max_num_subintervals", C.c_int32), ("max_newton_iters", C.c_int32),
                ("nlsolve_stages", C.c_int32), ("transfer_mode", C.c_int32),
                ("precision", C.c_int32), ("refine_iters", C.c_int32), ("warm_start", C.c_int32),
                ("report_err", C.c_int32), ("controller", C.c_int32), ("global_method", C.c_int32),
                ("hybrid_de", C.c_double), ("hybrid_ge", C.c_double), ("nest_abstol", C.c_double)]


class _Result(C.Structure):
    _fields_ = [("retcode", C.c_int32), ("N_final", C.c_int32), ("newton_iters", C.c_int32),
                ("outer_iters", C.c_int32), ("fallback_used", C.c_int32), ("capacity", C.c_int32),
                ("defect_norm", C.c_double), ("resid_norm", C.c_double), ("err_vs_fp64", C.c_double),
                ("mesh", _dp), ("u", _dp), ("stages", _dp)]


def _d(a):
    return _dp() if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.ref_interp_weights.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.ref_uniform_mesh.argtypes = [C.c_double, C.c_double, C.c_int, _dp]
        L.ref_interval.argtypes = [_dp, C.c_int, C.c_double]
        L.ref_tableau.argtypes = [C.c_int, C.POINTER(C.c_int), _dp]
        L.ref_f.argtypes = [C.POINTER(_Instance), C.c_double, _dp, _dp]
        L.ref_residual.argtypes = [C.c_int, C.POINTER(_Instance), _dp]
        L.ref_jacobian_blocks.argtypes = [C.c_int, C.POINTER(_Instance), _dp, _dp, _ip]
        L.ref_newton_step.argtypes = [C.c_int, C.POINTER(_Instance), _dp]
        L.ref_dense_system.argtypes = [C.c_int, C.POINTER(_Instance), _dp, _dp]
        L.ref_defect.argtypes = [C.c_int, C.POINTER(_Instance), _dp, _dp]
        L.ref_mesh_select.argtypes = [C.POINTER(Opts), C.POINTER(_Instance), C.c_int, _ip, _ip, _dp, _dp]
        L.ref_mesh_select_errors.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, _dp, C.c_int, _dp,
                                             C.c_int, _dp]
        L.ref_solve.argtypes = [C.POINTER(Opts), C.POINTER(_Instance), C.POINTER(_Result)]
        L.ref_solve_batch.argtypes = [C.POINTER(Opts), C.c_int, C.POINTER(_Instance), C.POINTER(_Result),
                                      C.c_int]
        L.ref_opts_default.argtypes = [C.POINTER(Opts)]
        L.ref_backtracking_poly.argtypes = [_dp, C.c_int, _dp, _dp, C.c_int, _ip]
        L.ref_dogleg.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp]
        L.ref_interpolate.argtypes = [C.c_int, C.POINTER(_Instance), C.c_int, _dp, _dp, _dp]
        L.ref_global_error.argtypes = [C.POINTER(Opts), C.POINTER(_Instance), _dp, _dp]
        L.ref_firk_tableau.argtypes = [C.c_int, C.POINTER(C.c_int), _dp]
        L.ref_firk_residual.argtypes = [C.c_int, C.POINTER(_Instance), _dp, _dp]
        L.ref_firk_dense_system.argtypes = [C.c_int, C.POINTER(_Instance), _dp, _dp, _dp]
        L.ref_firk_newton_step.argtypes = [C.c_int, C.POINTER(_Instance), _dp, _dp]
        L.ref_firk_defect.argtypes = [C.c_int, C.POINTER(_Instance), _dp, _dp, _dp]
        L.ref_firk_interpolate.argtypes = [C.c_int, C.POINTER(_Instance), _dp, C.c_int, _dp, _dp, _dp]
        L.ref_firk_evalsol.argtypes = [C.c_int, C.POINTER(_Instance), C.c_double, C.c_double, C.c_double, C.c_int,
                                       _dp, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


class _Pinned:
    """Holds the numpy arrays behind a ctypes instance."""

    def __init__(self, spec):
        self.keep = []

        def arr(x, dt=np.float64):
            if x is None:
                return None
            a = np.ascontiguousarray(x, dtype=dt)
            self.keep.append(a)
            return a

        s = _Instance()
        s.family, s.M, s.t0, s.t1 = spec.family, spec.M, spec.t0, spec.t1
        p = arr(spec.params)
        s.params, s.n_params = _d(p), p.shape[0]
        s.bc_kind = spec.bc_kind
        bt, bcmp, bv = arr(spec.bc_t), arr(spec.bc_comp, np.int32), arr(spec.bc_val)
        s.n_bc = bt.shape[0]
        s.bc_t, s.bc_val = _d(bt), _d(bv)
        s.bc_comp = bcmp.ctypes.data_as(_ip)
        s.N = spec.N
        s.mesh = _d(arr(spec.mesh))
        s.guess = _d(arr(spec.guess))
        s.u0 = _d(arr(spec.u0))
        self.inst = s


def opts(method=4, **kw):
    o = Opts()
    lib().ref_opts_default(C.byref(o))
    o.method = method
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def interp_weights(method, tau):
    n = 4 if method == 4 else 9
    w, wp = np.zeros(n), np.zeros(n)
    lib().ref_interp_weights(method, tau, _d(w), _d(wp))
    return w, wp


def tableau(method):
    ints = (C.c_int * 3)()
    dbl = np.zeros(128)
    lib().ref_tableau(method, ints, _d(dbl))
    s, ss, order = ints[0], ints[1], ints[2]
    o = 0
    out = {"s": s, "sstar": ss, "order": order}
    for name, size in (("c", 5), ("v", 5), ("b", 5)):
        out[name] = dbl[o:o + size][:s].copy(); o += size
    out["x"] = dbl[o:o + 25].reshape(5, 5)[:s, :s].copy(); o += 25
    ns = ss - s
    out["cs"] = dbl[o:o + 4][:ns].copy(); o += 4
    out["vs"] = dbl[o:o + 4][:ns].copy(); o += 4
    out["xs"] = dbl[o:o + 36].reshape(4, 9)[:ns, :ss].copy(); o += 36
    out["tau"] = dbl[o]
    return out


def uniform_mesh(t0, t1, N):
    m = np.zeros(N + 1)
    lib().ref_uniform_mesh(t0, t1, N, _d(m))
    return m


def interval(mesh, t):
    mesh = np.ascontiguousarray(mesh, dtype=np.float64)
    return lib().ref_interval(_d(mesh), mesh.shape[0] - 1, t)


def f(spec, t, y):
    P = _Pinned(spec)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(spec.M)
    lib().ref_f(C.byref(P.inst), t, _d(y), _d(out))
    return out


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} returned {rc}")


def residual(method, spec):
    P = _Pinned(spec)
    out = np.zeros(spec.M * (spec.N + 1))
    _check(lib().ref_residual(method, C.byref(P.inst), _d(out)), "residual")
    return out


def jacobian_blocks(method, spec):
    P = _Pinned(spec)
    M, N = spec.M, spec.N
    S, R = np.zeros((N, M, M)), np.zeros((N, M, M))
    node = np.zeros(M, dtype=np.int32)
    _check(lib().ref_jacobian_blocks(method, C.byref(P.inst), _d(S), _d(R), node.ctypes.data_as(_ip)),
           "jacobian_blocks")
    return S, R, node


def newton_step(method, spec):
    P = _Pinned(spec)
    out = np.zeros(spec.M * (spec.N + 1))
    _check(lib().ref_newton_step(method, C.byref(P.inst), _d(out)), "newton_step")
    return out


def dense_system(method, spec):
    P = _Pinned(spec)
    n = spec.M * (spec.N + 1)
    J, F = np.zeros((n, n)), np.zeros(n)
    _check(lib().ref_dense_system(method, C.byref(P.inst), _d(J), _d(F)), "dense_system")
    return J, F


def defect(method, spec):
    P = _Pinned(spec)
    err = np.zeros(spec.N * spec.M)
    dn = np.zeros(1)
    _check(lib().ref_defect(method, C.byref(P.inst), _d(err), _d(dn)), "defect")
    return err.reshape(spec.N, spec.M), float(dn[0])


def mesh_select(o, spec, capacity=None):
    P = _Pinned(spec)
    cap = capacity or (o.max_num_subintervals + 1)
    nN = np.zeros(1, dtype=np.int32)
    info = np.zeros(1, dtype=np.int32)
    nm = np.zeros(cap)
    ny = np.zeros(cap * spec.M)
    _check(lib().ref_mesh_select(C.byref(o), C.byref(P.inst), cap, nN.ctypes.data_as(_ip),
                                 info.ctypes.data_as(_ip), _d(nm), _d(ny)), "mesh_select")
    n = int(nN[0])
    return int(info[0]), n, nm[:n + 1].copy(), ny[:spec.M * (n + 1)].copy()


def mesh_select_errors(order, abstol, max_sub, mesh, errors):
    mesh = np.ascontiguousarray(mesh, dtype=np.float64)
    errors = np.ascontiguousarray(errors, dtype=np.float64)
    N = mesh.shape[0] - 1
    M = errors.size // N
    cap = 4 * N + 2 if 4 * N + 2 > 2 * N + 1 else 2 * N + 1
    out = np.zeros(max(cap, 2 * N + 1))
    r = lib().ref_mesh_select_errors(order, abstol, max_sub, N, _d(mesh), M, _d(errors), out.shape[0], _d(out))
    if r < 0:
        return None
    return out[:r + 1].copy()


def interpolate(method, spec, t):
    """Dense output of (mesh, guess) at the points t: (u (k, M), du (k, M))."""
    P = _Pinned(spec)
    t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
    u, du = np.zeros(max(1, len(t)) * spec.M), np.zeros(max(1, len(t)) * spec.M)
    _check(lib().ref_interpolate(method, C.byref(P.inst), len(t), _d(t), _d(u), _d(du)), "interpolate")
    return u[:len(t) * spec.M].reshape(-1, spec.M), du[:len(t) * spec.M].reshape(-1, spec.M)


def global_error(o, spec):
    """The global error estimate of o.global_method at (mesh, guess): (per-interval
    errors (N, M), the norm compared with abstol)."""
    P = _Pinned(spec)
    e, nrm = np.zeros(spec.N * spec.M), np.zeros(1)
    _check(lib().ref_global_error(C.byref(o), C.byref(P.inst), _d(e), _d(nrm)), "global_error")
    return e.reshape(spec.N, spec.M), float(nrm[0])


def backtracking_poly(coef, max_trace=64):
    """The oracle's BackTracking line search on phi(a) = sum_k coef[k] a^k:
    (return code, accepted step, trial steps in order)."""
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    a = np.zeros(1)
    tr = np.zeros(max_trace)
    nt = np.zeros(1, dtype=np.int32)
    rc = lib().ref_backtracking_poly(_d(coef), coef.shape[0], _d(a), _d(tr), max_trace, nt.ctypes.data_as(_ip))
    return rc, float(a[0]), tr[:min(int(nt[0]), max_trace)].copy()


def dogleg(J, F, radius):
    """The oracle's dogleg step of a dense system: (step, predicted reduction)."""
    J = np.ascontiguousarray(J, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.float64)
    n = F.shape[0]
    d, pred = np.zeros(n), np.zeros(1)
    _check(lib().ref_dogleg(n, _d(J), _d(F), radius, _d(d), _d(pred)), "dogleg")
    return d, float(pred[0])


# ---- FIRK (expanded form).  Method codes: 100 + s (Radau IIA), 200 + s / 300 + s /
# 400 + s (Lobatto IIIA / IIIB / IIIC).  `stages` is an (N, s, M) array or None for
# the reference's initial stages K_ij = y_{i+1}.

def firk_tableau(code):
    ints = (C.c_int * 3)()
    dbl = np.zeros(7 * 7 * 2 + 15)
    _check(lib().ref_firk_tableau(code, ints, _d(dbl)), "firk_tableau")
    s = ints[0]
    A = dbl[:49].reshape(7, 7)[:s, :s].copy()
    b = dbl[49:56][:s].copy()
    c = dbl[56:63][:s].copy()
    Q = dbl[63:112].reshape(7, 7)[:s, :s].copy()
    return {"s": s, "order": ints[1], "adaptive": bool(ints[2]), "A": A, "b": b, "c": c, "Q": Q,
            "tau": float(dbl[112])}


def _stg(stages):
    return None if stages is None else np.ascontiguousarray(stages, dtype=np.float64).reshape(-1)


def firk_nx(code, spec):
    """Length of the state vector of a FIRK method: the expanded form's M (N (s+1) + 1),
    the nested form's (code + 1000) M (N + 1)."""
    if code >= 1000:
        return spec.M * (spec.N + 1)
    s = code % 100
    return spec.M * (spec.N * (s + 1) + 1)


def firk_residual(code, spec, stages=None):
    P = _Pinned(spec)
    st = _stg(stages)
    out = np.zeros(firk_nx(code, spec))
    _check(lib().ref_firk_residual(code, C.byref(P.inst), _d(st), _d(out)), "firk_residual")
    return out


def firk_dense_system(code, spec, stages=None):
    P = _Pinned(spec)
    st = _stg(stages)
    n = firk_nx(code, spec)
    J, F = np.zeros((n, n)), np.zeros(n)
    _check(lib().ref_firk_dense_system(code, C.byref(P.inst), _d(st), _d(J), _d(F)), "firk_dense_system")
    return J, F


def firk_newton_step(code, spec, stages=None):
    P = _Pinned(spec)
    st = _stg(stages)
    out = np.zeros(firk_nx(code, spec))
    _check(lib().ref_firk_newton_step(code, C.byref(P.inst), _d(st), _d(out)), "firk_newton_step")
    return out


def firk_defect(code, spec, stages=None):
    P = _Pinned(spec)
    st = _stg(stages)
    err, dn = np.zeros(spec.N * spec.M), np.zeros(1)
    _check(lib().ref_firk_defect(code, C.byref(P.inst), _d(st), _d(err), _d(dn)), "firk_defect")
    return err.reshape(spec.N, spec.M), float(dn[0])


def firk_interpolate(code, spec, t, stages=None):
    P = _Pinned(spec)
    st = _stg(stages)
    t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
    u, du = np.zeros(max(1, len(t)) * spec.M), np.zeros(max(1, len(t)) * spec.M)
    _check(lib().ref_firk_interpolate(code, C.byref(P.inst), _d(st), len(t), _d(t), _d(u), _d(du)),
           "firk_interpolate")
    return u[:len(t) * spec.M].reshape(-1, spec.M), du[:len(t) * spec.M].reshape(-1, spec.M)


def firk_evalsol(code, spec, tj, tj1, t, comp, yi, yj, K):
    """EvalSol of a FIRK cache (FIRK/interpolation.jl:338-441) on the interval [tj, tj1] at t,
    component comp, from nodes yi, yj and stages K (s, M).  Returns (value, g): g holds the
    derivatives wrt y_j (M), y_{j+1} (M) and K_j (s M); the nested form's stages carry none."""
    s, M = code % 100, spec.M
    P = _Pinned(spec)
    yi = np.ascontiguousarray(yi, dtype=np.float64).reshape(M)
    yj = np.ascontiguousarray(yj, dtype=np.float64).reshape(M)
    K = np.ascontiguousarray(K, dtype=np.float64).reshape(s * M)
    v = np.zeros(1)
    g = np.zeros((s + 2) * M)
    _check(lib().ref_firk_evalsol(code, C.byref(P.inst), tj, tj1, t, comp, _d(yi), _d(yj), _d(K), _d(v), _d(g)),
           "firk_evalsol")
    return float(v[0]), g


def firk_unpack(code, spec, x):
    """Split an expanded vector [y_0, K_0., y_1, ...] into nodes (N+1, M) and stages (N, s, M)."""
    s, M, N = code % 100, spec.M, spec.N
    x = np.asarray(x).reshape(-1)
    y = np.stack([x[i * (s + 1) * M:i * (s + 1) * M + M] for i in range(N + 1)])
    K = np.stack([x[i * (s + 1) * M + M:(i + 1) * (s + 1) * M].reshape(s, M) for i in range(N)])
    return y, K


class Solution:
    def __init__(self, res, M, mesh, u):
        self.retcode = res.retcode
        self.N = res.N_final
        self.newton_iters = res.newton_iters
        self.outer_iters = res.outer_iters
        self.fallback_used = res.fallback_used
        self.defect_norm = res.defect_norm
        self.resid_norm = res.resid_norm
        self.err_vs_fp64 = res.err_vs_fp64
        self.t = mesh[:self.N + 1].copy()
        self.u = u[:M * (self.N + 1)].reshape(self.N + 1, M).copy()
        self.stages = None


def solve(o, spec, capacity=None):
    P = _Pinned(spec)
    cap = capacity or (o.max_num_subintervals + 1)
    mesh, u = np.zeros(cap), np.zeros(cap * spec.M)
    s = o.method % 100 if o.method >= 100 else 0  # stages (FIRK, either form)
    st = np.zeros(max(1, cap * s * spec.M))
    r = _Result()
    r.capacity = cap
    r.mesh, r.u = _d(mesh), _d(u)
    if s:
        r.stages = _d(st)
    _check(lib().ref_solve(C.byref(o), C.byref(P.inst), C.byref(r)), "solve")
    sol = Solution(r, spec.M, mesh, u)
    if s:
        sol.stages = st[:sol.N * s * spec.M].reshape(sol.N, s, spec.M).copy()
    return sol


def solve_batch(o, specs, nthreads=0, capacity=None, keep_solutions=True):
    n = len(specs)
    pins = [_Pinned(s) for s in specs]
    arr = (_Instance * n)()
    for i, p in enumerate(pins):
        arr[i] = p.inst
    cap = capacity or (o.max_num_subintervals + 1)
    res = (_Result * n)()
    bufs = []
    for i, s in enumerate(specs):
        mesh, u = np.zeros(cap), np.zeros(cap * s.M)
        bufs.append((mesh, u))
        res[i].capacity = cap
        res[i].mesh, res[i].u = _d(mesh), _d(u)
    _check(lib().ref_solve_batch(C.byref(o), n, arr, res, nthreads), "solve_batch")
    if not keep_solutions:
        return [int(res[i].retcode) for i in range(n)]
    return [Solution(res[i], specs[i].M, bufs[i][0], bufs[i][1]) for i in range(n)]


def max_threads():
    return lib().ref_max_threads()
