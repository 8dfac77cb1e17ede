"""Writes closed_forms.json: known-answer values of the test problems' exact
solutions, from the mathematics alone (no solver involved).

Bratu u'' + lam e^u = 0, u(0)=u(1)=0, both branches (lam < lam_c):
    u(t) = -2 ln[cosh((t-1/2) th/2) / cosh(th/4)],  th = sqrt(2 lam) cosh(th/4),
the lower branch with the smaller root th (fixed-point iteration from 0), the upper
branch with the larger one (bisection between the tangency point, where
th / cosh(th/4) is largest, and a bracket end).
Singular perturbation eps u'' + t u' - u = 0, u(-1)=1, u(1)=2:
    u(t) = t/2 + (3/2) g(t)/g(1),  g(t) = e^{-t^2/(2 eps)} + t sqrt(pi/(2 eps)) erf(t/sqrt(2 eps)).
Turning point (reference test/misc/adaptivity_tests.jl): u(-1)=-2, u(1)=0:
    u(t) = cos(pi t) + erf(t/sqrt(2 eps))/erf(1/sqrt(2 eps)).
Linear harmonic u''=-u, u(0)=5, u(5)=0: u = 5(cos t - cot 5 sin t).
"""
import json
import math
import pathlib

from scipy.special import erf


def bratu_theta(lam):
    th = 0.0
    for _ in range(500):
        th = math.sqrt(2 * lam) * math.cosh(th / 4)
    return th


def bratu_theta_upper(lam):
    # roots of q(th) = sqrt(2 lam) cosh(th/4) - th; q is convex, negative between
    # the two roots; its minimum sits where sinh(th/4) = 4 / sqrt(2 lam)
    c = math.sqrt(2 * lam)
    lo = 4 * math.asinh(4 / c)
    hi = lo
    while c * math.cosh(hi / 4) - hi < 0:
        hi *= 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if c * math.cosh(mid / 4) - mid < 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def bratu_u(lam, t, upper=False):
    th = bratu_theta_upper(lam) if upper else bratu_theta(lam)
    return -2 * math.log(math.cosh((t - 0.5) * th / 2) / math.cosh(th / 4))


def singpert_g(eps, t):
    return math.exp(-t * t / (2 * eps)) + t * math.sqrt(math.pi / (2 * eps)) * erf(t / math.sqrt(2 * eps))


def singpert_u(eps, t):
    return t / 2 + 1.5 * singpert_g(eps, t) / singpert_g(eps, 1.0)


def main():
    out = {
        "bratu": [{"lam": lam, "theta": bratu_theta(lam), "u_half": bratu_u(lam, 0.5),
                   "u_quarter": bratu_u(lam, 0.25)} for lam in (0.1, 1.0, 3.4)],
        "bratu_upper": [{"lam": lam, "theta": bratu_theta_upper(lam), "u_half": bratu_u(lam, 0.5, True),
                         "u_quarter": bratu_u(lam, 0.25, True)} for lam in (1.0, 3.4)],
        "singpert": [{"eps": eps, "g1": singpert_g(eps, 1.0), "u0": singpert_u(eps, 0.0),
                      "u_half": singpert_u(eps, 0.5)} for eps in (1e-4, 1e-2)],
        "lambda_critical": 3.513830719125162,
    }
    path = pathlib.Path(__file__).with_name("closed_forms.json")
    path.write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
