"""CPU checks of the C ABI: the library loads and exports every symbol the
header declares, and the Python mirror's constants and struct sizes match."""
import ctypes as C
import pathlib
import re

import pytest

import boundaryvaluediffeq_amd as bvp
from boundaryvaluediffeq_amd import _abi, _lib

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include").glob("*.h"))


def _declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?[\s\*]+(bvp_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def _defines():
    out = {}
    for h in HEADERS:
        for m in re.finditer(r"#define\s+(BVP_[A-Z0-9_]+)\s+(-?\d+)", h.read_text()):
            out[m.group(1)] = int(m.group(2))
    return out


def test_library_exports_every_declared_symbol():
    L = C.CDLL(str(_lib.LIB_PATH))
    declared = _declared_functions()
    assert len(declared) >= 15
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert declared <= set(_lib.EXPORTED) | {"bvp_opts_default"}


def test_constants_match_header():
    d = _defines()
    pairs = {
        "BVP_MIRK4": _abi.MIRK4, "BVP_MIRK6": _abi.MIRK6,
        "BVP_FAMILY_LINEAR_TANH": _abi.FAMILY_LINEAR_TANH, "BVP_FAMILY_PENDULUM": _abi.FAMILY_PENDULUM,
        "BVP_FAMILY_BRATU": _abi.FAMILY_BRATU, "BVP_FAMILY_SINGPERT": _abi.FAMILY_SINGPERT,
        "BVP_FAMILY_SCALAR_TU": _abi.FAMILY_SCALAR_TU, "BVP_FAMILY_TURNING": _abi.FAMILY_TURNING,
        "BVP_FAMILY_SWIRL": _abi.FAMILY_SWIRL,
        "BVP_BC_TWO_POINT": _abi.BC_TWO_POINT, "BVP_BC_STANDARD": _abi.BC_STANDARD,
        "BVP_RET_SUCCESS": _abi.RET_SUCCESS, "BVP_RET_FAILURE": _abi.RET_FAILURE,
        "BVP_RET_MAXITERS": _abi.RET_MAXITERS, "BVP_RET_UNSTABLE": _abi.RET_UNSTABLE,
        "BVP_PREC_FP64": _abi.PREC_FP64, "BVP_PREC_FP32_REFINE64": _abi.PREC_FP32_REFINE64,
        "BVP_TRANSFER_REFERENCE": _abi.TRANSFER_REFERENCE, "BVP_TRANSFER_CLEAN": _abi.TRANSFER_CLEAN,
        "BVP_EINVAL": _abi.EINVAL, "BVP_EHIP": _abi.EHIP, "BVP_ENOTSUP": _abi.ENOTSUP,
        "BVP_WARM_NONE": _abi.WARM_NONE, "BVP_WARM_MLP": _abi.WARM_MLP,
        "BVP_MLP_NPARAM": _abi.MLP_NPARAM, "BVP_FEAT_PARAM": _abi.FEAT_PARAM,
        "BVP_FEAT_LOG10_PARAM": _abi.FEAT_LOG10_PARAM,
        "BVP_STAGE_NEWTON": _abi.STAGE_NEWTON, "BVP_STAGE_BACKTRACKING": _abi.STAGE_BACKTRACKING,
        "BVP_STAGE_TRUST_REGION": _abi.STAGE_TRUST_REGION, "BVP_STAGES_DEFAULT": _abi.STAGES_DEFAULT,
        "BVP_CONTROL_DEFECT": _abi.CONTROL_DEFECT, "BVP_CONTROL_GLOBAL": _abi.CONTROL_GLOBAL,
        "BVP_CONTROL_SEQUENTIAL": _abi.CONTROL_SEQUENTIAL, "BVP_CONTROL_HYBRID": _abi.CONTROL_HYBRID,
        "BVP_GLOBAL_HO": _abi.GLOBAL_HO, "BVP_GLOBAL_RE": _abi.GLOBAL_RE,
    }
    for k, v in pairs.items():
        assert d[k] == v, k


def test_firk_method_codes_match_header():
    """BVP_RADAU_IIA(s) and BVP_LOBATTO_IIIA/B/C(s) are function-like macros of the header;
    the Python mirror and the algorithm classes give the same codes."""
    text = (ROOT / "include" / "bvp_mirk.h").read_text()
    pairs = {"BVP_RADAU_IIA": _abi.RADAU_IIA, "BVP_LOBATTO_IIIA": _abi.LOBATTO_IIIA,
             "BVP_LOBATTO_IIIB": _abi.LOBATTO_IIIB, "BVP_LOBATTO_IIIC": _abi.LOBATTO_IIIC}
    for name, fn in pairs.items():
        m = re.search(r"#define " + name + r"\(s\)\s+\((\d+) \+ \(s\)\)", text)
        assert m, name
        for st in (1, 2, 3, 4, 5, 7):
            assert fn(st) == int(m.group(1)) + st
    assert bvp.RadauIIa5().code == _abi.RADAU_IIA(5) and bvp.LobattoIIIc4().code == _abi.LOBATTO_IIIC(4)
    assert bvp.RadauIIa5.order == 9 and bvp.LobattoIIIb3.order == 4 and bvp.RadauIIa7().stages == 7
    o = bvp.make_opts(bvp.LobattoIIIa4(defect_threshold=0.2), abstol=1e-7)
    assert o.method == 204 and o.nlsolve_stages == 3 and o.defect_threshold == 0.2 and o.abstol == 1e-7
    nested = bvp.RadauIIa3(nested_nlsolve=True, nested_nlsolve_kwargs={"abstol": 1e-10})
    assert nested.code == _abi.FIRK_NESTED + _abi.RADAU_IIA(3) and _abi.is_firk_nested(nested.code)
    assert _abi.firk_stages(nested.code) == 3 and bvp.RadauIIa3().code == _abi.RADAU_IIA(3)
    o = bvp.make_opts(nested, abstol=1e-7)
    assert o.method == 1103 and o.nest_abstol == 1e-10
    assert bvp.make_opts(bvp.RadauIIa3(nested_nlsolve=True), abstol=1e-7).nest_abstol == 0.0
    with pytest.raises(ValueError):
        bvp.RadauIIa3(nested_nlsolve=True, nested_nlsolve_kwargs={"maxiters": 5})


def test_struct_layouts(tmp_path):
    """ctypes mirrors (product and oracle) have the C compiler's sizes and offsets."""
    import subprocess
    import pyoracle
    structs = {"bvp_instance": (_abi.Instance, pyoracle._Instance), "bvp_opts": (_abi.Opts, pyoracle.Opts),
               "bvp_result": (_abi.Result, pyoracle._Result), "bvp_train_opts": (_abi.TrainOpts, None)}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADERS[0]}"', "int main(void) {"]
    for cname, (py, _) in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _t in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            a, b, c = line.split()
            got[(a, b)] = int(c)
    for cname, (py, orc) in structs.items():
        assert C.sizeof(py) == got[(cname, "size")], cname
        for f, _t in py._fields_:
            assert getattr(py, f).offset == got[(cname, f)], (cname, f)
            if orc is not None:
                assert getattr(orc, f).offset == got[(cname, f)], (cname, f)
                assert C.sizeof(orc) == got[(cname, "size")], cname


def test_opts_default_matches_reference_defaults():
    o = _abi.Opts()
    _lib.lib().bvp_opts_default(C.byref(o))
    ref = _abi.default_opts()
    for name, _ in _abi.Opts._fields_:
        assert getattr(o, name) == getattr(ref, name), name
    assert o.abstol == 1e-6 and o.defect_threshold == 0.1 and o.max_num_subintervals == 3000


def test_controller_objects_fill_opts():
    """solve(...; controller) maps the reference's controller structs (Core/calc_errors.jl:54-107,
    their defaults included) onto bvp_opts."""
    from boundaryvaluediffeq_amd.solve import make_opts
    cases = [(bvp.DefectControl(defect_threshold=0.3), (0, 0, 0.3, 1.0, 1.0)),
             (bvp.GlobalErrorControl(), (1, 0, 0.1, 1.0, 1.0)),
             (bvp.GlobalErrorControl(method=bvp.REErrorControl()), (1, 1, 0.1, 1.0, 1.0)),
             (bvp.SequentialErrorControl(defect=bvp.DefectControl(defect_threshold=0.2)), (2, 0, 0.2, 1.0, 1.0)),
             (bvp.HybridErrorControl(DE=0.5, GE=2.0, global_error=bvp.GlobalErrorControl(method=bvp.REErrorControl())),
              (3, 1, 0.1, 0.5, 2.0))]
    for ctl, want in cases:
        o = make_opts(bvp.MIRK4(), controller=ctl)
        assert (o.controller, o.global_method, o.defect_threshold, o.hybrid_de, o.hybrid_ge) == want, ctl
    with pytest.raises(TypeError):
        make_opts(bvp.MIRK4(), controller=object())


def test_dt_to_mesh_like_reference():
    """N = cld(t1 - t0, dt); nodes of range(t0, t1, N+1) (Core/utils.jl:354-362, 694-700)."""
    import math
    assert bvp.cld(math.pi / 2, 0.05) == 32
    assert bvp.cld(1.0, 0.1) == 10
    m = bvp.uniform_mesh(0.0, math.pi / 2, 32)
    assert m[16] == math.pi / 4
    import pyoracle
    assert (pyoracle.uniform_mesh(-1.0, 1.0, 37) == bvp.uniform_mesh(-1.0, 1.0, 37)).all()


def test_context_fails_loudly_without_gpu():
    """No CPU fallback: on a machine without a GPU, creating a context raises."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.BVPError):
        _lib.Context(0)


def test_mlp_parameter_count():
    """4 -> 256 -> 256 -> 256 -> 2 with biases."""
    assert _abi.MLP_NPARAM == 4 * 256 + 256 + 2 * (256 * 256 + 256) + 256 * 2 + 2


_JULIA_TYPES = {"Int32": "int32_t", "Int64": "int64_t", "UInt64": "uint64_t", "Float64": "double",
                "Float32": "float", "Cint": "int32_t"}


def _c_struct_fields(name):
    """Field names and C types of a typedef'd struct of the header, in order."""
    text = re.sub(r"/\*.*?\*/", "", HEADERS[0].read_text(), flags=re.S)
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r"\s*;", text)
    assert m, name
    out = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ptr = "*" in decl
        decl = decl.replace("*", " ").replace("const ", "")
        parts = decl.replace(",", " , ").split()
        ctype = parts[0]
        names = [p for p in parts[1:] if p != ","]
        for nm in names:
            out.append((nm, "ptr" if ptr else ctype))
    return out


def _julia_struct_fields(jname):
    text = (ROOT / "INTEGRATION.md").read_text()
    m = re.search(r"struct " + jname + r"\b[^\n]*\n(.*?)\nend", text, re.S)
    assert m, jname
    out = []
    for nm, jt in re.findall(r"(\w+)::([\w{}]+)", m.group(1)):
        out.append((nm, "ptr" if jt.startswith("Ptr") else _JULIA_TYPES[jt]))
    return out


@pytest.mark.parametrize("jname,cname", [("BvpInstance", "bvp_instance"), ("BvpOpts", "bvp_opts"),
                                         ("BvpResult", "bvp_result"), ("BvpTrainOpts", "bvp_train_opts")])
def test_integration_julia_structs_match_header(jname, cname):
    """The Julia mirrors in INTEGRATION.md have the header's fields, in order, with matching types
    (a Julia caller passes these structs by reference to the library)."""
    assert _julia_struct_fields(jname) == _c_struct_fields(cname)


def test_initial_guess_forms_like_reference():
    """u0 as a function f(p, t) (and the deprecated f(t)) is evaluated at the nodes of the
    mesh from dt; a solution object (t, u) gives the mesh and the guess
    (Core/utils.jl:395-410, 683-704, 750-777); the oracle solves all forms alike."""
    import math
    import warnings
    import numpy as np
    import pyoracle as O
    f = bvp.BratuFunction()
    bc = ([(0, 0.0)], [(0, 0.0)])
    guess_pt = lambda p, t: [0.1 * p[0] * t * (1 - t), 0.1 * p[0] * (1 - 2 * t)]
    prob = bvp.TwoPointBVProblem(f, bc, guess_pt, (0.0, 1.0), [2.0])
    s = bvp.extract(prob, dt=0.1)
    assert s.N == 10 and s.u0 is None
    nodes = bvp.uniform_mesh(0.0, 1.0, 10)
    assert np.allclose(s.guess.reshape(-1, 2), np.array([guess_pt([2.0], t) for t in nodes]), rtol=0, atol=0)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        s1 = bvp.extract(bvp.TwoPointBVProblem(f, bc, lambda t: guess_pt([2.0], t), (0.0, 1.0), [2.0]), dt=0.1)
    assert any(issubclass(x.category, DeprecationWarning) for x in w)
    assert np.array_equal(s1.guess, s.guess)
    with pytest.raises(TypeError):
        bvp.extract(bvp.TwoPointBVProblem(f, bc, lambda a, b, c: 0, (0.0, 1.0), [2.0]), dt=0.1)
    # a solution object as the guess: its nodes become the mesh
    r = O.solve(O.opts(4, abstol=1e-8), s)
    class Sol:
        pass
    sol = Sol()
    sol.t = r.t
    sol.u = [r.u[k] for k in range(r.u.shape[0])]
    s2 = bvp.extract(bvp.TwoPointBVProblem(f, bc, sol, (0.0, 1.0), [2.0]), dt=0.5)
    assert s2.N == r.N and np.array_equal(s2.mesh, r.t)
    r2 = O.solve(O.opts(4, abstol=1e-8, max_newton_iters=0, adaptive=0), s2)
    assert r2.retcode == 0 and np.array_equal(r2.u, r.u)
    # a number as the constant guess of a scalar problem
    s3 = bvp.extract(bvp.BVProblem(bvp.ScalarTUFunction(), [(0.5, 0, 1.0)], 0.0, (0.0, 1.0), []), dt=0.25)
    assert s3.N == 4 and s3.u0.shape == (1,)
    assert math.isfinite(float(s3.u0[0]))
