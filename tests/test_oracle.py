"""Pins the oracle (CPU restatement of the reference MIRK path) against
closed forms, finite differences, a dense LU and the reference's own test
expectations.  Every check is chosen so that a plausible slip in the oracle
(a wrong tableau entry, sign, index or transposed block) fails one of them.
"""
import json
import math
import pathlib

import numpy as np
import pytest

import pyoracle as O
import cases

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "closed_forms.json").read_text())


# --------------------------------------------------------------- tableaus
@pytest.mark.parametrize("method", [4, 6])
def test_tableau_row_sums(method):
    """c_r = v_r + sum_j x_rj for the discrete and the interpolation stages; sum b = 1."""
    T = O.tableau(method)
    assert abs(T["b"].sum() - 1.0) < 1e-15
    for r in range(T["s"]):
        assert abs(T["c"][r] - (T["v"][r] + T["x"][r].sum())) < 1e-15
    for r in range(T["sstar"] - T["s"]):
        assert abs(T["cs"][r] - (T["vs"][r] + T["xs"][r].sum())) < 1e-15


def _pade(method, z):
    if method == 4:
        return (1 + z / 2 + z * z / 12) / (1 - z / 2 + z * z / 12)
    return (1 + z / 2 + z * z / 10 + z ** 3 / 120) / (1 - z / 2 + z * z / 10 - z ** 3 / 120)


@pytest.mark.parametrize("method", [4, 6])
@pytest.mark.parametrize("z", [-3.0, -0.7, 0.25, 1.3])
def test_stability_function_is_pade(method, z):
    """On y' = lam y one MIRK step maps y_i to P(h lam) y_i with P the (2,2) (MIRK4)
    or (3,3) (MIRK6) Pade approximant of e^z."""
    lam, h = z / 0.5, 0.5
    spec0 = cases.linear([[lam]], tspan=(0.0, h), bc=([(0, 1.0)], []), dt=h)

    def phi(y1):
        s = spec0.with_state(np.array([0.0, h]), np.array([1.0, y1]))
        return O.residual(method, s)[1]  # [bca; Phi_1]

    p0, p1 = phi(0.0), phi(1.0)
    r = -p0 / (p1 - p0)
    assert abs(r - _pade(method, z)) < 1e-13 * max(1, abs(r))


@pytest.mark.parametrize("method", [4, 6])
def test_interp_weights_endpoints(method):
    """w(0)=0, w(1)=(b,0..), w'(0)=e1, w'(1)=e2."""
    T = O.tableau(method)
    w0, wp0 = O.interp_weights(method, 0.0)
    w1, wp1 = O.interp_weights(method, 1.0)
    n = T["sstar"]
    b = np.zeros(n)
    b[:T["s"]] = T["b"]
    e1, e2 = np.eye(n)[0], np.eye(n)[1]
    assert np.max(np.abs(w0)) < 1e-14
    assert np.max(np.abs(w1 - b)) < 1e-12
    assert np.max(np.abs(wp0 - e1)) < 1e-12
    assert np.max(np.abs(wp1 - e2)) < 1e-11


@pytest.mark.parametrize("method", [4, 6])
def test_interp_weights_derivative(method):
    """w' is the derivative of w (central differences)."""
    for tau in (0.1, 0.226, 0.5, 0.7156, 0.9):
        e = 1e-6
        wa, _ = O.interp_weights(method, tau - e)
        wb, _ = O.interp_weights(method, tau + e)
        _, wp = O.interp_weights(method, tau)
        assert np.max(np.abs((wb - wa) / (2 * e) - wp)) < 1e-6 * max(1, np.max(np.abs(wp)))


# --------------------------------------------------------------- residual
def test_phi_mirk4_is_hermite_simpson():
    """MIRK4: K3 = f((y_i+y_{i+1})/2 + h/8 (K1-K2), t+h/2), Phi = y_{i+1}-y_i-h/6 (K1+K2+4K3)."""
    rng = np.random.default_rng(1)
    spec = cases.random_mesh(cases.random_state(cases.bratu(2.0, N=7), rng), rng)
    F = O.residual(4, spec)
    t, y = spec.full_mesh(), spec.full_guess().reshape(-1, 2)
    lam = 2.0

    def f(u):
        return np.array([u[1], -lam * math.exp(u[0])])

    for i in range(spec.N):
        h = t[i + 1] - t[i]
        K1, K2 = f(y[i]), f(y[i + 1])
        K3 = f((y[i] + y[i + 1]) / 2 + h / 8 * (K1 - K2))
        phi = y[i + 1] - y[i] - h / 6 * (K1 + K2 + 4 * K3)
        assert np.max(np.abs(F[1 + 2 * i:3 + 2 * i] - phi)) < 1e-14
    # two-point layout: [bca; Phi; bcb]
    assert F[0] == y[0, 0] - 0.0 and F[-1] == y[-1, 0] - 0.0


def test_affineness_reference_case():
    """MIRKt/Core/mirk_basic_tests.jl:103-118: u'=(u2,0), u(0)=5, u(5)=0 at dt=0.2 is
    reproduced exactly by every MIRK method (reference bound 1e-2)."""
    for method in (4, 6):
        for standard in (False, True):
            if standard:
                bc = [(0.0, 0, 5.0), (5.0, 0, 0.0)]
            else:
                bc = ([(0, 5.0)], [(0, 0.0)])
            spec = cases.linear([[0, 1], [0, 0]], bc=bc, dt=0.2, u0=[5.0, -3.5], standard=standard)
            sol = O.solve(O.opts(method), spec)
            assert sol.retcode == 0
            u1 = sol.u[:, 0]
            assert np.max(np.abs(np.diff(u1) + 0.2)) + abs(u1[0] - 5) < 1e-12


# --------------------------------------------------------------- Jacobian
def _fd_blocks(method, spec, eps=1e-6):
    """Central differences of Phi_i with respect to y_i and y_{i+1}."""
    M, N = spec.M, spec.N
    y = spec.full_guess().copy()
    m = spec.full_mesh()
    La = sum(1 for k in range(M) if spec.bc_kind == 0 and spec.bc_t[k] == spec.t0) if spec.bc_kind == 0 else M

    def phis(yv):
        F = O.residual(method, spec.with_state(m, yv))
        return F[La:La + N * M].reshape(N, M)

    S = np.zeros((N, M, M))
    R = np.zeros((N, M, M))
    for k in range(N + 1):
        for c in range(M):
            yp, ym = y.copy(), y.copy()
            yp[k * M + c] += eps
            ym[k * M + c] -= eps
            d = (phis(yp) - phis(ym)) / (2 * eps)
            if k < N:
                S[k, :, c] = d[k]
            if k > 0:
                R[k - 1, :, c] = d[k - 1]
    return S, R


FAMILY_CASES = {
    "bratu": lambda: cases.bratu(2.5, N=5),
    "pendulum": lambda: cases.pendulum(dt=0.4),
    "singpert": lambda: cases.singpert(0.05, N=6),
    "turning": lambda: cases.turning(0.05, dt=0.4),
    "swirl": lambda: cases.swirl(0.05, dt=0.25),
    "linear_tanh3": lambda: cases.linear([[0.3, 1.0, -0.2], [-1.0, 0.1, 0.4], [0.5, -0.3, -0.6]], c=0.7,
                                         tspan=(0, 1), bc=([(0, 1.0), (2, -1.0)], [(1, 0.5)]), dt=0.25),
}


@pytest.mark.parametrize("method", [4, 6])
@pytest.mark.parametrize("name", sorted(FAMILY_CASES))
def test_jacobian_blocks_vs_finite_differences(method, name):
    rng = np.random.default_rng(7)
    spec = cases.random_mesh(cases.random_state(FAMILY_CASES[name](), rng, 0.4), rng)
    S, R, node = O.jacobian_blocks(method, spec)
    Sf, Rf = _fd_blocks(method, spec)
    scale = max(1.0, np.max(np.abs(S)), np.max(np.abs(R)))
    assert np.max(np.abs(S - Sf)) < 2e-7 * scale
    assert np.max(np.abs(R - Rf)) < 2e-7 * scale


@pytest.mark.parametrize("method", [4, 6])
def test_dense_system_two_point_is_exact_jacobian(method):
    """TwoPoint: the assembled Newton matrix equals the finite-difference Jacobian of
    the full residual, BC rows and layout included."""
    rng = np.random.default_rng(3)
    spec = cases.random_state(FAMILY_CASES["linear_tanh3"](), rng, 0.5)
    J, F = O.dense_system(method, spec)
    assert np.array_equal(F, O.residual(method, spec))
    y, m = spec.full_guess(), spec.full_mesh()
    n = y.size
    Jf = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1e-6
        Jf[:, j] = (O.residual(method, spec.with_state(m, y + e)) - O.residual(method, spec.with_state(m, y - e))) / 2e-6
    assert np.max(np.abs(J - Jf)) < 1e-7 * max(1, np.max(np.abs(J)))


def test_standard_bc_rows_are_frozen_stage():
    """STANDARD BCs: the Jacobian row of sol(t)[c] is the unit entry at the left node of
    the interval holding t (stages frozen, MIRK/mirk.jl:565-572; EvalSol
    MIRK/interpolation.jl:214-231); end points use the shortcut nodes."""
    spec = cases.pendulum(dt=0.05)
    assert spec.N == 32
    J, F = O.dense_system(4, spec)
    n = 2 * 33
    row0 = np.zeros(n)
    row0[2 * 15 + 0] = 1.0    # pi/4 is node 16 -> interval 15 (0-based), tau = 1
    row1 = np.zeros(n)
    row1[2 * 32 + 0] = 1.0    # t_end shortcut
    assert np.array_equal(J[0], row0) and np.array_equal(J[1], row1)


# --------------------------------------------------------------- linear solve
@pytest.mark.parametrize("method", [4, 6])
@pytest.mark.parametrize("name", sorted(FAMILY_CASES))
def test_banded_newton_step_vs_dense_lu(method, name):
    rng = np.random.default_rng(11)
    spec = cases.random_state(FAMILY_CASES[name](), rng, 0.4)
    J, F = O.dense_system(method, spec)
    d_dense = np.linalg.solve(J, -F)
    d = O.newton_step(method, spec)
    assert np.max(np.abs(d - d_dense)) < 1e-10 * max(1, np.max(np.abs(d_dense)))


def test_banded_newton_step_c3_shape():
    """n=32 separated BCs (16 at each end): banded solve vs dense LU."""
    rng = np.random.default_rng(5)
    spec = cases.random_state(cases.linear_tanh_c3(rng, N=6), rng, 0.3)
    J, F = O.dense_system(6, spec)
    d = O.newton_step(6, spec)
    d_dense = np.linalg.solve(J, -F)
    assert np.max(np.abs(d - d_dense)) < 1e-9 * max(1, np.max(np.abs(d_dense)))


# --------------------------------------------------------------- end to end
@pytest.mark.parametrize("entry", GOLDEN["bratu"], ids=lambda e: f"lam{e['lam']}")
@pytest.mark.parametrize("method,order", [(4, 4), (6, 6)])
def test_bratu_closed_form_and_order(entry, method, order):
    lam = entry["lam"]
    errs = []
    for N in ((8, 16) if method == 4 else (4, 8)):
        spec = cases.bratu(lam, N=N)
        sol = O.solve(O.opts(method, adaptive=0, abstol=1e-13), spec)
        assert sol.retcode == 0
        i = N // 2
        assert sol.t[i] == 0.5
        errs.append(abs(sol.u[i, 0] - entry["u_half"]))
    if errs[1] < 1e-12:          # already at round-off: nothing left to measure
        assert errs[0] < 1e-11
        return
    rate = math.log2(errs[0] / errs[1])
    assert abs(rate - order) < 0.4, (errs, rate)


@pytest.mark.parametrize("method,order", [(4, 4), (6, 6)])
def test_convergence_linear_reference_case(method, order):
    """MIRKt/Core/mirk_basic_tests.jl:122-139 with problem f2 (u''=-u, u(0)=5, u(5)=0):
    observed order within 0.4 of the method order (dts 1/8, 1/4, 1/2)."""
    exact = lambda t: 5 * (np.cos(t) - np.sin(t) / np.tan(5.0))
    errs = []
    for dt in (0.5, 0.25, 0.125):
        spec = cases.linear([[0, 1], [-1, 0]], bc=([(0, 5.0)], [(0, 0.0)]), dt=dt, u0=[5.0, -3.5])
        sol = O.solve(O.opts(method, adaptive=0, abstol=1e-12), spec)
        errs.append(np.max(np.abs(sol.u[:, 0] - exact(sol.t))))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert abs(rates[-1] - order) < 0.4, (errs, rates)


@pytest.mark.parametrize("method", [4, 6])
def test_defect_converges_at_interpolant_order(method):
    """Defect of the continuous extension on a converged solve falls by ~2^p when h halves."""
    dn = []
    for N in (8, 16):
        spec = cases.bratu(3.0, N=N)
        sol = O.solve(O.opts(method, adaptive=0, abstol=1e-13), spec)
        _, d = O.defect(method, spec.with_state(sol.t, sol.u.reshape(-1)))
        dn.append(d)
    rate = math.log2(dn[0] / dn[1])
    assert abs(rate - method) < 0.6, (dn, rate)


def test_maxiters_zero_returns_guess():
    """MIRKt/Core/mirk_basic_tests.jl:678-720: maxiters=0, adaptive=false returns the guess."""
    rng = np.random.default_rng(2)
    spec = cases.random_state(cases.bratu(1.0, N=10), rng)
    sol = O.solve(O.opts(4, adaptive=0, max_newton_iters=0), spec)
    assert np.array_equal(sol.u.reshape(-1), spec.guess)
    assert sol.retcode == 2  # MaxIters


@pytest.mark.parametrize("entry", GOLDEN["singpert"], ids=lambda e: f"eps{e['eps']}")
def test_singpert_adaptive_closed_form(entry):
    """eps u'' + t u' - u = 0: adaptive MIRK4 from N=32 reaches defect <= abstol and the
    closed-form u(0) (the turning-point layer value)."""
    eps = entry["eps"]
    spec = cases.singpert(eps, N=32)
    sol = O.solve(O.opts(4, abstol=1e-6, max_num_subintervals=16384), spec)
    assert sol.retcode == 0 and sol.defect_norm <= 1e-6
    i0 = np.argmin(np.abs(sol.t))
    if sol.t[i0] == 0.0:
        assert abs(sol.u[i0, 0] - entry["u0"]) < 1e-4


def test_turning_point_reference_case():
    """test/misc/adaptivity_tests.jl: eps=0.001, dt=0.01, MIRK4; mesh is redistributed
    and the solve succeeds; compare with the erf closed form."""
    from scipy.special import erf
    eps = 0.001
    spec = cases.turning(eps, dt=0.01)
    sol = O.solve(O.opts(4), spec)
    assert sol.retcode == 0 and sol.defect_norm <= 1e-6
    exact = np.cos(np.pi * sol.t) + erf(sol.t / math.sqrt(2 * eps)) / erf(1 / math.sqrt(2 * eps))
    assert np.max(np.abs(sol.u[:, 0] - exact)) < 1e-4


def test_pendulum_readme_case():
    """README pendulum: converges; the BC through the continuous extension holds."""
    spec = cases.pendulum(0.05)
    sol = O.solve(O.opts(4), spec)
    assert sol.retcode == 0
    assert sol.defect_norm <= 1e-6
    assert abs(sol.u[-1, 0] - math.pi / 2) < 1e-6


# --------------------------------------------------------------- mesh selection
def test_mesh_selector_uniform_defect_halves():
    """r1 <= r3 (uniform defect) -> half_mesh!: 2N intervals at exact midpoints."""
    mesh = np.array([0.0, 0.1, 0.3, 0.6, 1.0])
    errors = np.full((4, 2), 1e-4)
    nm = O.mesh_select_errors(4, 1e-6, 3000, mesh, errors)
    assert nm is not None and nm.shape[0] == 9
    assert np.array_equal(nm[::2], mesh)
    assert np.array_equal(nm[1::2], (mesh[1:] + mesh[:-1]) / 2)


def test_mesh_selector_redistributes_and_equidistributes():
    """Non-uniform defect: N* = clamp(round(1.3 r2 + 1), (N+1)/2, 4N) intervals (ties to
    even) placed so every new interval holds the same integral of s_hat/h."""
    rng = np.random.default_rng(4)
    N = 40
    mesh = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, N))])
    mesh /= mesh[-1]
    errors = (10.0 ** rng.uniform(-9, -4, N))[:, None] * np.array([[1.0, 0.5]])
    abstol = 1e-6
    shat = (np.max(np.abs(errors), axis=1) / abstol) ** (1 / 5)
    r2 = shat.sum()
    npred = int(np.rint(1.3 * r2 + 1))
    if abs(npred - N) < 0.1 * N:
        npred = int(np.rint(N + 0.1 * N))
    Nstar = min(max(npred, (N + 1) // 2), 4 * N)
    nm = O.mesh_select_errors(4, abstol, 3000, mesh, errors)
    assert nm.shape[0] == Nstar + 1
    assert nm[0] == mesh[0] and nm[-1] == mesh[-1]
    assert np.all(np.diff(nm) > 0)
    dens = shat / np.diff(mesh)
    # integral of the piecewise constant density over each new interval
    cum_old = np.concatenate([[0.0], np.cumsum(dens * np.diff(mesh))])
    cum_new = np.interp(nm, mesh, cum_old)
    pieces = np.diff(cum_new)
    zeta = cum_old[-1] / Nstar
    assert np.max(np.abs(pieces - zeta)) < 1e-9 * zeta


def test_mesh_selector_cap_fails():
    mesh = np.linspace(0, 1, 11)
    errors = np.full((10, 1), 1.0)
    assert O.mesh_select_errors(4, 1e-6, 15, mesh, errors) is None


def test_transfer_modes():
    """Clean transfer evaluates the old continuous extension; on a halved mesh the even
    nodes reproduce the old nodes up to Phi.  The reference's in-place transfer reads
    the base value from the array it writes, so it agrees with the clean one exactly
    where the old interval index ii is >= the new node index."""
    spec = cases.bratu(2.0, N=8)
    sol = O.solve(O.opts(4, adaptive=0, abstol=1e-13), spec)
    st = spec.with_state(sol.t, sol.u.reshape(-1))
    o_clean = O.opts(4, abstol=1e-14, transfer_mode=1)
    o_ref = O.opts(4, abstol=1e-14, transfer_mode=0)
    info, n1, m1, y1 = O.mesh_select(o_clean, st)
    info2, n2, m2, y2 = O.mesh_select(o_ref, st)
    assert info == info2 == 0 and n1 == n2 and np.array_equal(m1, m2)
    Y1, Y2 = y1.reshape(-1, 2), y2.reshape(-1, 2)
    old = sol.u
    for j, t in enumerate(m1):
        ii = O.interval(sol.t, t)
        if t in sol.t:
            k = int(np.where(sol.t == t)[0][0])
            assert np.max(np.abs(Y1[j] - old[k])) < 1e-12
        if ii >= j:
            assert np.array_equal(Y1[j], Y2[j])


def test_ensemble_matches_sequential():
    rng = np.random.default_rng(9)
    specs = [cases.bratu(float(l), N=int(n)) for l, n in zip(rng.uniform(0.1, 3.4, 12), rng.integers(8, 40, 12))]
    o = O.opts(4, abstol=1e-8)
    batch = O.solve_batch(o, specs, nthreads=4)
    for s, b in zip(specs, batch):
        one = O.solve(o, s)
        assert one.retcode == b.retcode and one.N == b.N
        assert np.array_equal(one.u, b.u)
