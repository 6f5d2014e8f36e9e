"""Exact solution of the 1-D Riemann problem for an ideal gas (Toro, Riemann Solvers and
Numerical Methods for Fluid Dynamics, 3rd ed., ch. 4: pressure function f_K, Newton iteration for
p*, sampling of the self-similar solution).  Test helper for the HLLC / MUSCL-Hancock
convergence check."""
import numpy as np


def _fk(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        q = np.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    # rarefaction
    f = 2.0 * ck / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)
    df = 1.0 / (rho * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dfl = _fk(p, rl, pl, cl, g)
        fr, dfr = _fk(p, rr, pr, cr, g)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-12)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, cl, g)
    fr, _ = _fk(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, left, right, g):
    """Density, velocity and pressure at positions x and time t; left/right = (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    gm = (g - 1.0) / (g + 1.0)
    out = np.zeros((3, x.size))
    for k, xv in enumerate(x):
        s = (xv - x0) / t
        if s <= us:  # left of the contact
            if ps > pl:  # left shock
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    out[:, k] = (rl, ul, pl)
                else:
                    out[:, k] = (rl * (ps / pl + gm) / (gm * ps / pl + 1), us, ps)
            else:  # left rarefaction
                csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
                if s <= ul - cl:
                    out[:, k] = (rl, ul, pl)
                elif s >= us - csl:
                    out[:, k] = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[:, k] = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:  # right of the contact
            if ps > pr:  # right shock
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    out[:, k] = (rr, ur, pr)
                else:
                    out[:, k] = (rr * (ps / pr + gm) / (gm * ps / pr + 1), us, ps)
            else:  # right rarefaction
                csr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                if s >= ur + cr:
                    out[:, k] = (rr, ur, pr)
                elif s <= us + csr:
                    out[:, k] = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[:, k] = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out
