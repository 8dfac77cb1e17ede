"""ctypes binding of the HIP library ``_build/libbvp_mirk.so`` (C ABI of
``include/bvp_mirk.h``).  There is no CPU fallback: a missing library or a
missing GPU raises."""
import ctypes as C
import pathlib

import numpy as np

from . import _abi
from .problems import instance_array

LIB_PATH = pathlib.Path(__file__).resolve().parent / "_build" / "libbvp_mirk.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_longlong)
_fp = C.POINTER(C.c_float)


class BVPError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise BVPError(f"HIP library not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    Inst = C.POINTER(_abi.Instance)
    Opts = C.POINTER(_abi.Opts)
    Res = C.POINTER(_abi.Result)
    sig = {
        "bvp_opts_default": (None, [Opts]),
        "bvp_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "bvp_ctx_destroy": (None, [vp]),
        "bvp_last_error": (C.c_char_p, [vp]),
        "bvp_mirk_solve_batch": (C.c_int, [vp, Opts, C.c_int32, Inst, Res]),
        "bvp_mirk_solve_batch_noalloc": (C.c_int, [vp, Opts, C.c_int32, Inst, Res]),
        "bvp_mirk_residual": (C.c_int, [vp, C.c_int32, C.c_int32, Inst, _dp]),
        "bvp_mirk_jacobian_blocks": (C.c_int, [vp, C.c_int32, C.c_int32, Inst, _dp, _dp, _ip]),
        "bvp_mirk_newton_step": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, Inst, _dp]),
        "bvp_mirk_defect": (C.c_int, [vp, C.c_int32, C.c_int32, Inst, _dp, _dp]),
        "bvp_mirk_mesh_select": (C.c_int, [vp, Opts, C.c_int32, Inst, C.c_int32, _ip, _ip, _dp, _dp]),
        "bvp_mirk_interpolate": (C.c_int, [vp, C.c_int32, C.c_int32, Inst, _ip, _dp, _dp, _dp]),
        "bvp_mirk_global_error": (C.c_int, [vp, Opts, C.c_int32, Inst, _dp, _dp]),
        "bvp_batch_create": (C.c_int, [vp, Opts, C.c_int32, Inst, C.POINTER(vp)]),
        "bvp_batch_run": (C.c_int, [vp, vp]),
        "bvp_batch_sync": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
        "bvp_batch_results": (C.c_int, [vp, vp, Res]),
        "bvp_batch_counters": (C.c_int, [vp, vp, _lp]),
        "bvp_batch_info": (C.c_int, [vp, _ip]),
        "bvp_batch_destroy": (None, [vp]),
        "bvp_copy_probe": (C.c_int, [vp, C.c_longlong, C.c_int, C.POINTER(C.c_float)]),
        "bvp_batch_set_graph": (C.c_int, [vp, C.c_int32]),
        "bvp_batch_timing": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "bvp_batch_learned_guess": (C.c_int, [vp, vp, _dp]),
        "bvp_batch_autotune": (C.c_int, [vp, vp, C.c_int32, _fp, _ip]),
        "bvp_batch_config": (C.c_int, [vp, _ip, _ip]),
        "bvp_batch_precision": (C.c_int, [vp, _ip]),
        "bvp_batch_chunking": (C.c_int, [vp, _ip]),
        "bvp_batch_set_chunking": (C.c_int, [vp, C.c_int32]),
        "bvp_batch_set_config": (C.c_int, [vp, vp, C.c_int32, C.c_int32]),
        "bvp_mlp_create": (C.c_int, [vp, C.c_uint64, C.POINTER(vp)]),
        "bvp_mlp_destroy": (None, [vp]),
        "bvp_mlp_set_normalization": (C.c_int, [vp, _fp, _fp, _fp, _fp]),
        "bvp_mlp_get_params": (C.c_int, [vp, _fp]),
        "bvp_mlp_set_params": (C.c_int, [vp, _fp]),
        "bvp_mlp_gradient": (C.c_int, [vp, C.c_int32, _fp, _fp, _fp, _fp]),
        "bvp_mlp_predict": (C.c_int, [vp, C.c_int32, _fp, _fp]),
        "bvp_mlp_train": (C.c_int, [vp, C.c_int64, _fp, _fp, C.POINTER(_abi.TrainOpts), _fp, _ip,
                                    C.POINTER(C.c_float)]),
        "bvp_batch_set_warm_start": (C.c_int, [vp, vp, C.c_int32, C.c_int32]),
        "bvp_firk_residual": (C.c_int, [vp, C.c_int32, C.c_int32, Inst, _dp, _dp]),
        "bvp_firk_newton_step": (C.c_int, [vp, C.c_int32, C.c_int32, Inst, _dp, _dp]),
        "bvp_firk_defect": (C.c_int, [vp, C.c_int32, C.c_int32, Inst, _dp, _dp, _dp]),
        "bvp_firk_interpolate": (C.c_int, [vp, C.c_int32, C.c_int32, Inst, _dp, _ip, _dp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


EXPORTED = ["bvp_opts_default", "bvp_ctx_create", "bvp_ctx_destroy", "bvp_last_error",
            "bvp_mirk_solve_batch", "bvp_mirk_residual", "bvp_mirk_jacobian_blocks",
            "bvp_mirk_newton_step", "bvp_mirk_defect", "bvp_mirk_mesh_select", "bvp_batch_create",
            "bvp_batch_run", "bvp_batch_sync", "bvp_batch_results", "bvp_batch_counters",
            "bvp_batch_info", "bvp_batch_destroy", "bvp_copy_probe", "bvp_batch_set_graph",
            "bvp_mlp_create", "bvp_mlp_destroy", "bvp_mlp_set_normalization", "bvp_mlp_get_params",
            "bvp_mlp_set_params", "bvp_mlp_gradient", "bvp_mlp_predict", "bvp_mlp_train",
            "bvp_batch_set_warm_start", "bvp_batch_timing", "bvp_batch_learned_guess", "bvp_batch_autotune",
            "bvp_batch_config", "bvp_batch_set_config", "bvp_batch_precision", "bvp_mirk_interpolate",
            "bvp_mirk_global_error",
            "bvp_mirk_solve_batch_noalloc", "bvp_batch_chunking", "bvp_batch_set_chunking",
            "bvp_firk_residual", "bvp_firk_newton_step", "bvp_firk_defect", "bvp_firk_interpolate"]


class Context:
    """A device context (stream, events, device memory of the calls)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().bvp_ctx_create(device, C.byref(self._h))
        if rc != 0:
            raise BVPError(f"bvp_ctx_create(device={device}) failed with {rc} (no usable GPU?)")
        self.device = device

    def close(self):
        if self._h:
            lib().bvp_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def check(self, rc, what):
        if rc != 0:
            msg = lib().bvp_last_error(self._h)
            raise BVPError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


_ctx = {}


def context(device=0):
    if device not in _ctx:
        _ctx[device] = Context(device)
    return _ctx[device]


def _nodes(specs):
    return [s.N + 1 for s in specs]


def residual(ctx, method, specs):
    arr = instance_array(specs)
    M = specs[0].M
    tot = sum(_nodes(specs)) * M
    out = np.zeros(tot)
    ctx.check(lib().bvp_mirk_residual(ctx.handle, method, len(specs), arr, out.ctypes.data_as(_dp)), "residual")
    res, o = [], 0
    for s in specs:
        k = M * (s.N + 1)
        res.append(out[o:o + k].copy())
        o += k
    return res


def jacobian_blocks(ctx, method, specs):
    arr = instance_array(specs)
    M = specs[0].M
    nint = sum(s.N for s in specs)
    S = np.zeros(nint * M * M)
    R = np.zeros(nint * M * M)
    node = np.zeros(len(specs) * M, dtype=np.int32)
    ctx.check(lib().bvp_mirk_jacobian_blocks(ctx.handle, method, len(specs), arr, S.ctypes.data_as(_dp),
                                             R.ctypes.data_as(_dp), node.ctypes.data_as(_ip)), "jacobian_blocks")
    out, o = [], 0
    for i, s in enumerate(specs):
        k = s.N * M * M
        out.append((S[o:o + k].reshape(s.N, M, M).copy(), R[o:o + k].reshape(s.N, M, M).copy(),
                    node[i * M:(i + 1) * M].copy()))
        o += k
    return out


def newton_step(ctx, method, specs, precision=_abi.PREC_FP64, refine_iters=0):
    arr = instance_array(specs)
    M = specs[0].M
    out = np.zeros(sum(_nodes(specs)) * M)
    ctx.check(lib().bvp_mirk_newton_step(ctx.handle, method, precision, refine_iters, len(specs), arr,
                                         out.ctypes.data_as(_dp)), "newton_step")
    res, o = [], 0
    for s in specs:
        k = M * (s.N + 1)
        res.append(out[o:o + k].copy())
        o += k
    return res


def defect(ctx, method, specs):
    arr = instance_array(specs)
    M = specs[0].M
    nint = sum(s.N for s in specs)
    err = np.zeros(nint * M)
    dn = np.zeros(len(specs))
    ctx.check(lib().bvp_mirk_defect(ctx.handle, method, len(specs), arr, err.ctypes.data_as(_dp),
                                    dn.ctypes.data_as(_dp)), "defect")
    out, o = [], 0
    for i, s in enumerate(specs):
        out.append((err[o:o + s.N * M].reshape(s.N, M).copy(), float(dn[i])))
        o += s.N * M
    return out


def global_error(ctx, opts, specs):
    """The global-error estimate of opts.global_method at each instance's (mesh,
    guess): [(errors (N, M), norm)]."""
    arr = instance_array(specs)
    M = specs[0].M
    err = np.zeros(sum(s.N for s in specs) * M)
    nrm = np.zeros(len(specs))
    ctx.check(lib().bvp_mirk_global_error(ctx.handle, C.byref(opts), len(specs), arr, err.ctypes.data_as(_dp),
                                          nrm.ctypes.data_as(_dp)), "global_error")
    out, o = [], 0
    for i, s in enumerate(specs):
        out.append((err[o:o + s.N * M].reshape(s.N, M).copy(), float(nrm[i])))
        o += s.N * M
    return out


def interpolate(ctx, method, specs, tvals, derivative=True):
    """Dense output: the continuous extension of each instance's (mesh, guess) at the
    points tvals[i]; returns [(u (k, M), du (k, M) or None)] per instance."""
    arr = instance_array(specs)
    M = specs[0].M
    ts = [np.ascontiguousarray(t, dtype=np.float64).reshape(-1) for t in tvals]
    nt = np.array([len(t) for t in ts], dtype=np.int32)
    tt = np.concatenate(ts) if len(ts) else np.zeros(0)
    u = np.zeros(max(1, len(tt)) * M)
    du = np.zeros(max(1, len(tt)) * M) if derivative else None
    ctx.check(lib().bvp_mirk_interpolate(ctx.handle, method, len(specs), arr, nt.ctypes.data_as(_ip),
                                         tt.ctypes.data_as(_dp), u.ctypes.data_as(_dp),
                                         du.ctypes.data_as(_dp) if derivative else _dp()), "interpolate")
    out, o = [], 0
    for k in nt:
        out.append((u[o * M:(o + k) * M].reshape(k, M).copy(),
                    du[o * M:(o + k) * M].reshape(k, M).copy() if derivative else None))
        o += k
    return out


def mesh_select(ctx, opts, specs, capacity=None):
    arr = instance_array(specs)
    M = specs[0].M
    n = len(specs)
    cap = capacity or (opts.max_num_subintervals + 1)
    nN = np.zeros(n, dtype=np.int32)
    info = np.zeros(n, dtype=np.int32)
    nm = np.zeros(n * cap)
    ny = np.zeros(n * cap * M)
    ctx.check(lib().bvp_mirk_mesh_select(ctx.handle, C.byref(opts), n, arr, cap, nN.ctypes.data_as(_ip),
                                         info.ctypes.data_as(_ip), nm.ctypes.data_as(_dp),
                                         ny.ctypes.data_as(_dp)), "mesh_select")
    out = []
    for i in range(n):
        k = int(nN[i])
        out.append((int(info[i]), k, nm[i * cap:i * cap + k + 1].copy(),
                    ny[i * cap * M:i * cap * M + M * (k + 1)].copy()))
    return out


# ---- FIRK components.  stages: None (the reference's initial stages K_ij = y_{i+1}) or
# a list of (N, s, M) arrays, one per instance.
def _firk_stages_arg(method, specs, stages):
    if stages is None:
        return None
    s = _abi.firk_stages(method)
    parts = [np.ascontiguousarray(k, dtype=np.float64).reshape(sp.N * s * sp.M) for k, sp in zip(stages, specs)]
    return np.concatenate(parts) if parts else np.zeros(1)


def _nx(method, spec):
    if _abi.is_firk_nested(method):
        return spec.M * (spec.N + 1)
    return spec.M * (spec.N * (_abi.firk_stages(method) + 1) + 1)


def firk_residual(ctx, method, specs, stages=None):
    arr = instance_array(specs)
    st = _firk_stages_arg(method, specs, stages)
    out = np.zeros(sum(_nx(method, s) for s in specs))
    ctx.check(lib().bvp_firk_residual(ctx.handle, method, len(specs), arr,
                                      st.ctypes.data_as(_dp) if st is not None else _dp(),
                                      out.ctypes.data_as(_dp)), "firk_residual")
    res, o = [], 0
    for s in specs:
        k = _nx(method, s)
        res.append(out[o:o + k].copy())
        o += k
    return res


def firk_newton_step(ctx, method, specs, stages=None):
    arr = instance_array(specs)
    st = _firk_stages_arg(method, specs, stages)
    out = np.zeros(sum(_nx(method, s) for s in specs))
    ctx.check(lib().bvp_firk_newton_step(ctx.handle, method, len(specs), arr,
                                         st.ctypes.data_as(_dp) if st is not None else _dp(),
                                         out.ctypes.data_as(_dp)), "firk_newton_step")
    res, o = [], 0
    for s in specs:
        k = _nx(method, s)
        res.append(out[o:o + k].copy())
        o += k
    return res


def firk_defect(ctx, method, specs, stages=None):
    arr = instance_array(specs)
    st = _firk_stages_arg(method, specs, stages)
    M = specs[0].M
    err = np.zeros(sum(s.N for s in specs) * M)
    dn = np.zeros(len(specs))
    ctx.check(lib().bvp_firk_defect(ctx.handle, method, len(specs), arr,
                                    st.ctypes.data_as(_dp) if st is not None else _dp(),
                                    err.ctypes.data_as(_dp), dn.ctypes.data_as(_dp)), "firk_defect")
    out, o = [], 0
    for i, s in enumerate(specs):
        out.append((err[o:o + s.N * M].reshape(s.N, M).copy(), float(dn[i])))
        o += s.N * M
    return out


def firk_interpolate(ctx, method, specs, tvals, stages=None, derivative=True):
    arr = instance_array(specs)
    st = _firk_stages_arg(method, specs, stages)
    M = specs[0].M
    ts = [np.ascontiguousarray(t, dtype=np.float64).reshape(-1) for t in tvals]
    nt = np.array([len(t) for t in ts], dtype=np.int32)
    tt = np.concatenate(ts) if len(ts) else np.zeros(0)
    u = np.zeros(max(1, len(tt)) * M)
    du = np.zeros(max(1, len(tt)) * M) if derivative else None
    ctx.check(lib().bvp_firk_interpolate(ctx.handle, method, len(specs), arr,
                                         st.ctypes.data_as(_dp) if st is not None else _dp(),
                                         nt.ctypes.data_as(_ip), tt.ctypes.data_as(_dp), u.ctypes.data_as(_dp),
                                         du.ctypes.data_as(_dp) if derivative else _dp()), "firk_interpolate")
    out, o = [], 0
    for k in nt:
        out.append((u[o * M:(o + k) * M].reshape(k, M).copy(),
                    du[o * M:(o + k) * M].reshape(k, M).copy() if derivative else None))
        o += k
    return out


def copy_probe(ctx, nbytes, repeats=3):
    ms = C.c_float()
    ctx.check(lib().bvp_copy_probe(ctx.handle, int(nbytes), int(repeats), C.byref(ms)), "copy_probe")
    return float(ms.value)


class RawSolution:
    """Per-instance output of a batched solve."""

    def __init__(self, r, M, mesh, u):
        self.retcode = int(r.retcode)
        self.N = int(r.N_final)
        self.newton_iters = int(r.newton_iters)
        self.outer_iters = int(r.outer_iters)
        self.fallback_used = int(r.fallback_used)
        self.defect_norm = float(r.defect_norm)
        self.resid_norm = float(r.resid_norm)
        self.err_vs_fp64 = float(r.err_vs_fp64)
        self.t = mesh[:self.N + 1].copy()
        self.u = u[:M * (self.N + 1)].reshape(self.N + 1, M).copy()
        self.stages = None


def _result_array(specs, cap, stages=0):
    n = len(specs)
    res = (_abi.Result * n)()
    bufs = []
    for i, s in enumerate(specs):
        m, u = np.zeros(cap), np.zeros(cap * s.M)
        k = np.zeros(max(1, cap * stages * s.M)) if stages else None
        bufs.append((m, u, k))
        res[i].capacity = cap
        res[i].mesh = m.ctypes.data_as(_dp)
        res[i].u = u.ctypes.data_as(_dp)
        if stages:
            res[i].stages = k.ctypes.data_as(_dp)
    return res, bufs


def _raw(res, spec, buf, stages):
    r = RawSolution(res, spec.M, buf[0], buf[1])
    if stages:
        r.stages = buf[2][:r.N * stages * spec.M].reshape(r.N, stages, spec.M).copy()
    return r


def solve_batch(ctx, opts, specs):
    """bvp_mirk_solve_batch; for a FIRK method each solution also carries its final
    stage values (N, s, M)."""
    arr = instance_array(specs)
    cap = max(opts.max_num_subintervals, max(s.N for s in specs)) + 1
    st = _abi.firk_stages(opts.method) if _abi.is_firk(opts.method) else 0
    res, bufs = _result_array(specs, cap, st)
    ctx.check(lib().bvp_mirk_solve_batch(ctx.handle, C.byref(opts), len(specs), arr, res), "solve_batch")
    return [_raw(res[i], specs[i], bufs[i], st) for i in range(len(specs))]


def solve_batch_noalloc(ctx, opts, specs):
    """bvp_mirk_solve_batch_noalloc: the same solve with no device allocation (the
    library's static scratch array holds the call's device memory)."""
    arr = instance_array(specs)
    cap = max(opts.max_num_subintervals, max(s.N for s in specs)) + 1
    st = _abi.firk_stages(opts.method) if _abi.is_firk(opts.method) else 0
    res, bufs = _result_array(specs, cap, st)
    ctx.check(lib().bvp_mirk_solve_batch_noalloc(ctx.handle, C.byref(opts), len(specs), arr, res),
              "solve_batch_noalloc")
    return [_raw(res[i], specs[i], bufs[i], st) for i in range(len(specs))]


class Batch:
    """Uploaded batch that can be solved repeatedly (bvp_batch_* plan API)."""

    def __init__(self, ctx, opts, specs):
        self.ctx = ctx
        self.specs = specs
        self.opts = opts
        self._arr = instance_array(specs)
        self._h = C.c_void_p()
        ctx.check(lib().bvp_batch_create(ctx.handle, C.byref(opts), len(specs), self._arr, C.byref(self._h)),
                  "batch_create")

    def run(self):
        self.ctx.check(lib().bvp_batch_run(self.ctx.handle, self._h), "batch_run")

    def sync(self):
        ms = C.c_float()
        self.ctx.check(lib().bvp_batch_sync(self.ctx.handle, self._h, C.byref(ms)), "batch_sync")
        return float(ms.value)

    def results(self):
        cap = max(self.opts.max_num_subintervals, max(s.N for s in self.specs)) + 1
        st = _abi.firk_stages(self.opts.method) if _abi.is_firk(self.opts.method) else 0
        res, bufs = _result_array(self.specs, cap, st)
        self.ctx.check(lib().bvp_batch_results(self.ctx.handle, self._h, res), "batch_results")
        return [_raw(res[i], self.specs[i], bufs[i], st) for i in range(len(self.specs))]

    def learned_guess(self):
        """The attached network's guess at every node, one (N+1, M) array per instance."""
        M = self.specs[0].M
        tot = sum(s.N + 1 for s in self.specs)
        y = np.zeros(tot * M)
        self.ctx.check(lib().bvp_batch_learned_guess(self.ctx.handle, self._h, y.ctypes.data_as(_dp)),
                       "batch_learned_guess")
        out, o = [], 0
        for s in self.specs:
            out.append(y[o * M:(o + s.N + 1) * M].reshape(s.N + 1, M).copy())
            o += s.N + 1
        return out

    def autotune(self, repeats=3):
        """Pick the fastest launch configuration by measurement; returns the median
        device ms of every candidate and the chosen (variant, persistent waves)."""
        ms = np.zeros(16, np.float32)
        n = C.c_int32()
        self.ctx.check(lib().bvp_batch_autotune(self.ctx.handle, self._h, repeats, ms.ctypes.data_as(_fp),
                                                C.byref(n)), "batch_autotune")
        return [float(x) for x in ms[:n.value]], self.config()

    def set_config(self, variant, blocks):
        self.ctx.check(lib().bvp_batch_set_config(self.ctx.handle, self._h, variant, blocks), "batch_set_config")

    def config(self):
        v, b = C.c_int32(), C.c_int32()
        lib().bvp_batch_config(self._h, C.byref(v), C.byref(b))
        return int(v.value), int(b.value)

    def chunking(self):
        """Minimum intervals per lane of the wave kernels' chunk mapping."""
        m = C.c_int32()
        lib().bvp_batch_chunking(self._h, C.byref(m))
        return int(m.value)

    def set_chunking(self, min_chunk):
        rc = lib().bvp_batch_set_chunking(self._h, int(min_chunk))
        if rc != 0:
            raise BVPError(f"batch_set_chunking failed ({rc})")

    def precision(self):
        """The record / factor precision the plan runs with (BVP_PREC_*)."""
        p = C.c_int32()
        lib().bvp_batch_precision(self._h, C.byref(p))
        return int(p.value)

    def timing(self):
        """(guess ms, solve ms) of the last synced run."""
        g, s = C.c_float(), C.c_float()
        lib().bvp_batch_timing(self._h, C.byref(g), C.byref(s))
        return float(g.value), float(s.value)

    def set_graph(self, enable):
        """Capture the run's launches in a hipGraph and replay it (default on)."""
        rc = lib().bvp_batch_set_graph(self._h, 1 if enable else 0)
        if rc != 0:
            raise BVPError(f"batch_set_graph failed ({rc})")

    def set_warm_start(self, net, feature_kind=_abi.FEAT_PARAM, anchors=16):
        """Attach an MLP (or detach with None): each run writes the learned guess
        of every node on the device before the solve (the network at `anchors`
        points per instance and Hermite interpolation, or at every node for 0)."""
        self._net = net
        h = net.handle if net is not None else C.c_void_p()
        rc = lib().bvp_batch_set_warm_start(self._h, h, feature_kind, anchors)
        if rc != 0:
            raise BVPError(f"batch_set_warm_start failed ({rc}): {lib().bvp_last_error(self.ctx.handle).decode()}")

    def counters(self):
        c = (C.c_longlong * 4)()
        self.ctx.check(lib().bvp_batch_counters(self.ctx.handle, self._h, c), "batch_counters")
        return [int(x) for x in c]

    def info(self):
        c = (C.c_int32 * 4)()
        lib().bvp_batch_info(self._h, c)
        return [int(x) for x in c]

    def close(self):
        if self._h:
            lib().bvp_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


class MLP:
    """The warm-start network (bvp_mlp_* entry points): 4 -> 256 -> 256 -> 256 -> 2,
    tanh, BF16 weights, FP32 accumulation, trained and evaluated on the GPU."""

    NPARAM = _abi.MLP_NPARAM

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        self._h = C.c_void_p()
        ctx.check(lib().bvp_mlp_create(ctx.handle, C.c_uint64(seed), C.byref(self._h)), "mlp_create")
        self.norm = (np.zeros(4, np.float32), np.ones(4, np.float32), np.zeros(2, np.float32),
                     np.ones(2, np.float32))

    @property
    def handle(self):
        return self._h

    def set_normalization(self, in_shift, in_scale, out_shift, out_scale):
        v = [_f32(in_shift, 4), _f32(in_scale, 4), _f32(out_shift, 2), _f32(out_scale, 2)]
        self.ctx.check(lib().bvp_mlp_set_normalization(self._h, *[a.ctypes.data_as(_fp) for a in v]),
                       "mlp_set_normalization")
        self.norm = tuple(v)

    def params(self):
        w = np.zeros(self.NPARAM, np.float32)
        self.ctx.check(lib().bvp_mlp_get_params(self._h, w.ctypes.data_as(_fp)), "mlp_get_params")
        return w

    def set_params(self, w):
        w = _f32(w, self.NPARAM)
        self.ctx.check(lib().bvp_mlp_set_params(self._h, w.ctypes.data_as(_fp)), "mlp_set_params")

    def gradient(self, x, y):
        x = _f32(x).reshape(-1, 4)
        y = _f32(y).reshape(-1, 2)
        g = np.zeros(self.NPARAM, np.float32)
        loss = C.c_float()
        self.ctx.check(lib().bvp_mlp_gradient(self._h, len(x), x.ctypes.data_as(_fp), y.ctypes.data_as(_fp),
                                              g.ctypes.data_as(_fp), C.byref(loss)), "mlp_gradient")
        return float(loss.value), g

    def predict(self, x):
        x = _f32(x).reshape(-1, 4)
        y = np.zeros((len(x), 2), np.float32)
        self.ctx.check(lib().bvp_mlp_predict(self._h, len(x), x.ctypes.data_as(_fp), y.ctypes.data_as(_fp)),
                       "mlp_predict")
        return y

    def train(self, x, y, epochs=10, batch=8192, lr=2e-3, lr_min=2e-5, weight_decay=0.0, warmup=100, seed=7,
              use_graph=True, first_epoch=0, total_epochs=0):
        """Adam on (x, y) for epochs [first_epoch, first_epoch + epochs) of a schedule of
        total_epochs (0: epochs); returns (loss per step of this call, device ms)."""
        x = _f32(x).reshape(-1, 4)
        y = _f32(y).reshape(-1, 2)
        o = _abi.TrainOpts(epochs, batch, lr, lr_min, weight_decay, warmup, seed, 1 if use_graph else 0,
                           first_epoch, total_epochs)
        nsteps = (len(x) // batch) * epochs
        loss = np.zeros(max(nsteps, 1), np.float32)
        steps = C.c_int32()
        ms = C.c_float()
        self.ctx.check(lib().bvp_mlp_train(self._h, len(x), x.ctypes.data_as(_fp), y.ctypes.data_as(_fp),
                                           C.byref(o), loss.ctypes.data_as(_fp), C.byref(steps), C.byref(ms)),
                       "mlp_train")
        return loss[:steps.value], float(ms.value)

    def close(self):
        if self._h:
            lib().bvp_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
