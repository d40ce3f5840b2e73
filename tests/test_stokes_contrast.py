"""Convergence of the Stokes2D pseudo-transient solver against viscosity contrast, on the oracle.

miniapps/Stokes2D.jl:60-71 fixes the solver's parameters (Vdmp = 4, Vsc = 1, Ptsc = 1/4) for one
inclusion of viscosity 0.1 in a matrix of 1 (log10 contrast -1). Its pseudo-time steps
(compute_timesteps!, :11-15) use the local cell-averaged viscosity. With the same parameters the
iteration diverges once the contrast passes about 10^+-1.25; taking the 3x3 maximum viscosity
for the two velocity steps ("maxV", the pressure step keeps the local value) converges up to
10^+-2.5 with no other change. These tests keep both facts; run as a script, the module prints
the table of iteration counts:

    python -m tests.test_stokes_contrast 127      # from the repository root; nx = ny = 127

Everything here runs the oracle's restatement of the miniapp's kernels (:18-55, :109-124) on
the CPU."""
import sys
import time

import numpy as np
import pytest

from oracle import ps_oracle as O


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def maxloc3x3(M):
    """Maximum over each cell's 3x3 neighbourhood, edges replicated."""
    P = np.pad(M, 1, mode="edge")
    n, m = M.shape
    return _f(np.max([P[1 + i:1 + i + n, 1 + j:1 + j + m] for i in (-1, 0, 1) for j in (-1, 0, 1)], axis=0))


def pt_solve(nx, mu_incl, ts="av", eps=1e-8, itmax=80000, nout=200, Vdmp=4.0, Vsc=1.0, Ptsc=0.25):
    """The miniapp's iteration from rest on its inclusion setup (:60-107) with the inclusion
    viscosity `mu_incl`: ("conv" | "div" | "nc", iterations, last residual)."""
    ny = nx
    lx = ly = 10.0
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    z = lambda *s: _f(np.zeros(s))  # noqa: E731
    a = dict(Pt=z(nx, ny), dtPt=z(nx, ny), divV=z(nx, ny), Vx=z(nx + 1, ny), Vy=z(nx, ny + 1), txx=z(nx, ny), tyy=z(nx, ny),
             txy=z(nx - 1, ny - 1), Rx=z(nx - 1, ny - 2), Ry=z(nx - 2, ny - 1), dVxdt=z(nx - 1, ny - 2), dVydt=z(nx - 2, ny - 1),
             dtVx=z(nx - 1, ny - 2), dtVy=z(nx - 2, ny - 1))
    x = (np.arange(nx) * dx)[:, None]
    y = (np.arange(ny) * dy)[None, :]
    rad = (x - 0.5 * lx) ** 2 + (y - 0.5 * ly) ** 2
    Mus = _f(np.where(rad < 1.0, mu_incl, 1.0))
    Rog = _f(np.where(rad < 1.0, 1.0, 0.0))
    O.stokes2d_timesteps(a["dtVx"], a["dtVy"], a["dtPt"], Mus, Vsc, Ptsc, min(dx, dy) ** 2, max(nx, ny))
    if ts == "maxV":
        t = z(nx, ny)
        O.stokes2d_timesteps(a["dtVx"], a["dtVy"], t, maxloc3x3(Mus), Vsc, Ptsc, min(dx, dy) ** 2, max(nx, ny))
    err = float("nan")
    for it in range(1, itmax + 1):
        O.stokes2d_p(a["divV"], a["Pt"], a["Vx"], a["Vy"], a["dtPt"], dx, dy)
        O.stokes2d_tau(a["divV"], a["txx"], a["tyy"], a["txy"], a["Vx"], a["Vy"], Mus, dx, dy)
        O.stokes2d_dv(a["Rx"], a["Ry"], a["dVxdt"], a["dVydt"], a["Pt"], Rog, a["txx"], a["tyy"], a["txy"], 1 - Vdmp / nx, 1 - Vdmp / ny,
                      dx, dy)
        O.stokes2d_v(a["Vx"], a["Vy"], a["dVxdt"], a["dVydt"], a["dtVx"], a["dtVy"])
        O.bc_y(a["Vx"])
        O.bc_x(a["Vy"])
        if it % nout == 0:
            err = max(O.mean_abs(a["Rx"]), O.mean_abs(a["Ry"]), O.mean_abs(a["divV"]))
            if not err < 1e10:
                return "div", it, err
            if err <= eps:
                return "conv", it, err
    return "nc", itmax, err


@pytest.mark.parametrize("log10c", [-2.0, 2.0])
def test_local_steps_diverge_beyond_moderate_contrast(log10c):
    st, it, err = pt_solve(63, 10.0 ** log10c, ts="av", itmax=40000)
    assert st == "div", (st, it, err)


@pytest.mark.parametrize("log10c", [-2.0, -1.0, 1.0, 2.0])
def test_max_viscosity_velocity_steps_converge(log10c):
    st, it, err = pt_solve(63, 10.0 ** log10c, ts="maxV", itmax=40000)
    assert st == "conv" and err <= 1e-8, (st, it, err)


def test_miniapp_contrast_converges_with_either_step():
    # the miniapp's own setup (mu_incl = 0.1): both step rules converge
    for ts in ("av", "maxV"):
        st, it, err = pt_solve(63, 0.1, ts=ts, itmax=40000)
        assert st == "conv", (ts, st, it, err)


if __name__ == "__main__":
    nx = int(sys.argv[1]) if len(sys.argv) > 1 else 127
    O.set_threads(4)
    print(f"nx = ny = {nx}, eps = 1e-8, nout = 200, Vdmp = 4, Vsc = 1, Ptsc = 1/4, at most 120000 iterations")
    print("| log10 contrast | local steps (miniapp) | 3x3-max velocity steps |")
    print("|---|---|---|")
    for lc in (-3.0, -2.5, -2.0, -1.5, -1.25, -1.0, 1.0, 1.25, 1.5, 2.0, 2.5, 3.0):
        row = []
        for ts in ("av", "maxV"):
            t0 = time.time()
            st, it, err = pt_solve(nx, 10.0 ** lc, ts=ts, itmax=120000)
            row.append(f"{st} {it} (err {err:.1e}, {time.time() - t0:.0f} s)")
        print(f"| {lc:+.2f} | {row[0]} | {row[1]} |", flush=True)
