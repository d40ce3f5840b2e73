"""Closed-form pins of the oracle's acoustic and Stokes2D restatements.

Every check compares the oracle with the derivative of a continuous polynomial field evaluated at
the staggered position the statement's result lives at (cell centre, x/y/z face, grid node),
not with the statement's formula. Central differences of a quadratic at a midpoint and averages
of a linear field at a midpoint are exact, so each expected value is known in closed form:
  * a flipped sign gives the negated increment;
  * a difference shifted by one cell gives 2a(i +- 1 - 0.5) instead of 2a(i - 0.5);
  * a wrong mask updates a boundary face or misses an inner one;
  * a dropped term, a swapped spacing or a wrong damping changes the value.
Staggered positions (cells of size dx, cell i centred at (i - 1/2) dx, 1-based):
  P, Pt, Mus, txx, tyy, divV      at cell centres ((i - 1/2) dx, (j - 1/2) dy)
  Vx (nx+1 in x)                  at x-faces ((I - 1) dx, (j - 1/2) dy)
  Vy (ny+1 in y)                  at y-faces ((i - 1/2) dx, (J - 1) dy)
  txy ((nx-1) x (ny-1))           at inner nodes (i dx, j dy)
  Rx, dVxdt, dtVx ((nx-1)(ny-2))  at the inner x-faces they update: Vx[i+1, j+1]
  Ry, dVydt, dtVy ((nx-2)(ny-1))  at the inner y-faces Vy[i+1, j+1]
(miniapps/acoustic3D.jl:11-21, miniapps/acoustic2D.jl:11-20, miniapps/Stokes2D.jl:11-55).
"""
import numpy as np
import pytest

from oracle import ps_oracle as O


def F(a):
    return np.asfortranarray(a, dtype=np.float64)


def close(got, want, rtol=1e-13):
    scale = max(1.0, float(np.abs(want).max()))
    return np.abs(np.asarray(got) - want).max() <= rtol * scale


# ------------------------------------------------------------------ acoustic 3-D
def _centres(n, d):
    return (np.arange(1, n + 1) - 0.5) * d


def _faces(n, d):
    return np.arange(0, n + 1) * d  # n+1 faces, face I (1-based) at (I - 1) d


@pytest.mark.parametrize("seed", [0, 1])
def test_acoustic3d_velocity_is_minus_dt_over_rho_grad_p_on_inner_faces(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = 9, 8, 7
    dx, dy, dz = 0.5, 0.25, 0.75
    dt, rho = 0.1, 1.3
    a, b, c = rng.uniform(0.5, 2.0, 3)
    x, y, z = _centres(nx, dx), _centres(ny, dy), _centres(nz, dz)
    # P(x, y, z) = a x^2 + b y^2 + c z^2 sampled at the cell centres
    P = F(a * x[:, None, None] ** 2 + b * y[None, :, None] ** 2 + c * z[None, None, :] ** 2)
    V0 = [F(rng.random(s)) for s in ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))]
    Vx, Vy, Vz = [v.copy(order="F") for v in V0]
    O.acoustic3d_v(Vx, Vy, Vz, P, dt, rho, dx, dy, dz)
    # dP/dx at the x-face (I - 1) dx is 2 a (I - 1) dx; updated faces: I = 2..nx, j = 2..ny-1, k = 2..nz-1
    fx, fy, fz = _faces(nx, dx), _faces(ny, dy), _faces(nz, dz)
    for V, Vold, f, coef, axis, n_ in ((Vx, V0[0], fx, a, 0, nx), (Vy, V0[1], fy, b, 1, ny), (Vz, V0[2], fz, c, 2, nz)):
        want = Vold.copy()
        grad = 2.0 * coef * f  # dP/d(axis) at every face position of that axis
        sl = [slice(1, -1)] * 3
        sl[axis] = slice(1, n_)  # faces 2..n (1-based)
        sh = [1, 1, 1]
        sh[axis] = -1
        want[tuple(sl)] = Vold[tuple(sl)] - dt / rho * grad[1:n_].reshape(sh)
        assert close(V, want), axis
    # the boundary faces of every component are never touched
    assert np.array_equal(Vx[[0, -1]], V0[0][[0, -1]])
    assert np.array_equal(Vx[:, [0, -1]], V0[0][:, [0, -1]])
    assert np.array_equal(Vz[:, :, [0, -1]], V0[2][:, :, [0, -1]])


@pytest.mark.parametrize("seed", [2, 3])
def test_acoustic3d_pressure_is_minus_dt_k_div_v_at_centres(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = 8, 9, 6
    dx, dy, dz = 0.25, 0.5, 0.125
    dt, k = 0.07, 1.7
    a, b, c = rng.uniform(-2.0, 2.0, 3)
    fx, fy, fz = _faces(nx, dx), _faces(ny, dy), _faces(nz, dz)
    x, y, z = _centres(nx, dx), _centres(ny, dy), _centres(nz, dz)
    # V = (a x^2, b y^2, c z^2) sampled on the faces of its own component, with a transverse
    # variation that a divergence must ignore
    Vx = F(a * fx[:, None, None] ** 2 + 0.3 * y[None, :, None] + 0.0 * z[None, None, :])
    Vy = F(b * fy[None, :, None] ** 2 + 0.2 * z[None, None, :] + 0.0 * x[:, None, None])
    Vz = F(c * fz[None, None, :] ** 2 + 0.1 * x[:, None, None] + 0.0 * y[None, :, None])
    P0 = F(rng.random((nx, ny, nz)))
    P = P0.copy(order="F")
    O.acoustic3d_p(P, Vx, Vy, Vz, dt, k, dx, dy, dz)
    div = 2 * a * x[:, None, None] + 2 * b * y[None, :, None] + 2 * c * z[None, None, :]
    assert close(P, P0 - dt * k * div)


def test_acoustic2d_closed_forms():
    rng = np.random.default_rng(7)
    nx, ny, dx, dy, dt, rho, k = 10, 7, 0.5, 0.25, 0.05, 0.8, 1.4
    a, b = rng.uniform(0.5, 2.0, 2)
    x, y, fx, fy = _centres(nx, dx), _centres(ny, dy), _faces(nx, dx), _faces(ny, dy)
    P = F(a * x[:, None] ** 2 + b * y[None, :] ** 2)
    Vx0, Vy0 = F(rng.random((nx + 1, ny))), F(rng.random((nx, ny + 1)))
    Vx, Vy = Vx0.copy(order="F"), Vy0.copy(order="F")
    O.acoustic2d_v(Vx, Vy, P, dt, rho, dx, dy)
    wx, wy = Vx0.copy(), Vy0.copy()
    wx[1:nx, 1:-1] -= dt / rho * (2 * a * fx[1:nx])[:, None]
    wy[1:-1, 1:ny] -= dt / rho * (2 * b * fy[1:ny])[None, :]
    assert close(Vx, wx) and close(Vy, wy)
    Vx = F(a * fx[:, None] ** 2 + 0.5 * y[None, :])
    Vy = F(b * fy[None, :] ** 2 + 0.25 * x[:, None])
    P0 = F(rng.random((nx, ny)))
    P = P0.copy(order="F")
    O.acoustic2d_p(P, Vx, Vy, dt, k, dx, dy)
    assert close(P, P0 - dt * k * (2 * a * x[:, None] + 2 * b * y[None, :]))


# ------------------------------------------------------------------ Stokes 2-D
NX, NY, DX, DY = 11, 9, 0.5, 0.25
XC, YC = _centres(NX, DX), _centres(NY, DY)
XF, YF = _faces(NX, DX), _faces(NY, DY)
XN, YN = XF[1:NX], YF[1:NY]  # inner nodes i dx, j dy (i = 1..nx-1, j = 1..ny-1)


def _z(*s):
    return F(np.zeros(s))


def _mu(x, y, m0=2.0, m1=0.3, m2=-0.2):
    return m0 + m1 * x + m2 * y


def test_stokes2d_timesteps_are_the_viscosity_at_the_face():
    # Mus linear: @av_xi(Mus) at Rx[i, j] is mu at the x-face (i dx, (j + 1/2) dy) exactly
    Mus = F(_mu(XC[:, None], YC[None, :]))
    dtVx, dtVy, dtPt = _z(NX - 1, NY - 2), _z(NX - 2, NY - 1), _z(NX, NY)
    Vsc, Ptsc, min_dxy2, max_nxy = 0.9, 0.3, min(DX, DY) ** 2, max(NX, NY)
    O.stokes2d_timesteps(dtVx, dtVy, dtPt, Mus, Vsc, Ptsc, min_dxy2, max_nxy)
    mu_x = _mu(XN[:, None], YC[None, 1:NY - 1])  # x-faces 2..nx, rows 2..ny-1
    mu_y = _mu(XC[1:NX - 1, None], YN[None, :])
    assert close(dtVx, Vsc * min_dxy2 / mu_x / 4.1, 1e-14)
    assert close(dtVy, Vsc * min_dxy2 / mu_y / 4.1, 1e-14)
    assert close(dtPt, Ptsc * 4.1 * Mus / max_nxy, 1e-14)


def _velocity(a, b, c, e):
    """Vx = a x^2 + c y on the x-faces, Vy = b y^2 + e x on the y-faces."""
    Vx = F(a * XF[:, None] ** 2 + c * YC[None, :])
    Vy = F(b * YF[None, :] ** 2 + e * XC[:, None])
    return Vx, Vy


@pytest.mark.parametrize("abce", [(1.0, -0.5, 0.3, 0.7), (0.25, 2.0, -1.5, 0.0)])
def test_stokes2d_divergence_pressure_and_stresses(abce):
    a, b, c, e = abce
    Vx, Vy = _velocity(a, b, c, e)
    Mus = F(_mu(XC[:, None], YC[None, :]))
    divV, Pt0, dtPt = _z(NX, NY), F(np.random.default_rng(1).random((NX, NY))), F(0.1 + 0.01 * np.arange(NX * NY).reshape(NX, NY))
    Pt = Pt0.copy(order="F")
    O.stokes2d_p(divV, Pt, Vx, Vy, dtPt, DX, DY)
    div = 2 * a * XC[:, None] + 2 * b * YC[None, :]  # dVx/dx + dVy/dy at the centres
    assert close(divV, div)
    assert close(Pt, Pt0 - dtPt * div)
    txx, tyy, txy = _z(NX, NY), _z(NX, NY), _z(NX - 1, NY - 1)
    O.stokes2d_tau(divV, txx, tyy, txy, Vx, Vy, Mus, DX, DY)
    assert close(txx, 2 * Mus * (2 * a * XC[:, None] - div / 3))
    assert close(tyy, 2 * Mus * (2 * b * YC[None, :] - div / 3))
    # at the inner nodes: 2 mu (1/2)(dVx/dy + dVy/dx) with mu linear, so @av(Mus) is mu at the node
    assert close(txy, 2 * _mu(XN[:, None], YN[None, :]) * 0.5 * (c + e))


def test_stokes2d_uniform_strain_rates():
    # linear Vx = a x, Vy = b y: divV = a + b, txx = 2 mu (a - (a + b)/3), txy = 0 exactly
    a, b = 3.0, -1.0
    Vx = F(a * XF[:, None] + 0.0 * YC[None, :])
    Vy = F(b * YF[None, :] + 0.0 * XC[:, None])
    Mus = F(_mu(XC[:, None], YC[None, :]))
    divV, Pt, dtPt = _z(NX, NY), _z(NX, NY), _z(NX, NY)
    O.stokes2d_p(divV, Pt, Vx, Vy, dtPt, DX, DY)
    assert np.all(divV == a + b)
    txx, tyy, txy = _z(NX, NY), _z(NX, NY), _z(NX - 1, NY - 1)
    O.stokes2d_tau(divV, txx, tyy, txy, Vx, Vy, Mus, DX, DY)
    assert close(txx, 2 * Mus * (a - (a + b) / 3), 1e-15) and close(tyy, 2 * Mus * (b - (a + b) / 3), 1e-15)
    assert np.all(txy == 0)


def test_stokes2d_residuals_are_momentum_balance_at_the_faces():
    rng = np.random.default_rng(5)
    p, q, s, t, u, w, g0, g1 = rng.uniform(-1.5, 1.5, 8)
    # stresses and pressure as polynomials sampled at their own positions
    txx = F(p * XC[:, None] ** 2 + 0.4 * YC[None, :])    # d/dx = 2 p x
    tyy = F(q * YC[None, :] ** 2 + 0.6 * XC[:, None])    # d/dy = 2 q y
    txy = F(s * YN[None, :] ** 2 + t * XN[:, None] ** 2)  # d/dy = 2 s y, d/dx = 2 t x
    Pt = F(u * XC[:, None] ** 2 + w * YC[None, :] ** 2)   # d/dx = 2 u x, d/dy = 2 w y
    Rog = F(g0 + g1 * YC[None, :] + 0.0 * XC[:, None])    # linear: @av_yi is rho g at the y-face
    Rx, Ry = _z(NX - 1, NY - 2), _z(NX - 2, NY - 1)
    D0x, D0y = F(rng.random((NX - 1, NY - 2))), F(rng.random((NX - 2, NY - 1)))
    dVxdt, dVydt = D0x.copy(order="F"), D0y.copy(order="F")
    dampX, dampY = 0.9, 0.8
    O.stokes2d_dv(Rx, Ry, dVxdt, dVydt, Pt, Rog, txx, tyy, txy, dampX, dampY, DX, DY)
    xfx, yfx = XN[:, None], YC[None, 1:NY - 1]  # x-faces of Rx
    xfy, yfy = XC[1:NX - 1, None], YN[None, :]  # y-faces of Ry
    rx = 2 * p * xfx + 2 * s * yfx - 2 * u * xfx
    ry = 2 * q * yfy + 2 * t * xfy - 2 * w * yfy + (g0 + g1 * yfy)
    assert close(Rx, rx + 0 * yfx) and close(Ry, ry + 0 * xfy)
    assert close(dVxdt, dampX * D0x + Rx, 1e-15) and close(dVydt, dampY * D0y + Ry, 1e-15)


def test_stokes2d_hydrostatic_pressure_balances_buoyancy():
    # Pt = rho g y with uniform rho g and no stress: Ry = 0 exactly (dyadic values), Rx = 0
    rg = 0.5
    Pt = F(rg * YC[None, :] + 0.0 * XC[:, None])
    Rog = F(np.full((NX, NY), rg))
    Rx, Ry, dVxdt, dVydt = _z(NX - 1, NY - 2), _z(NX - 2, NY - 1), _z(NX - 1, NY - 2), _z(NX - 2, NY - 1)
    z = _z(NX, NY)
    O.stokes2d_dv(Rx, Ry, dVxdt, dVydt, Pt, Rog, z, z.copy(order="F"), _z(NX - 1, NY - 1), 0.5, 0.5, DX, DY)
    assert np.all(Rx == 0) and np.all(Ry == 0)


def test_stokes2d_velocity_update_lands_on_the_inner_faces():
    rng = np.random.default_rng(9)
    Vx0, Vy0 = F(rng.random((NX + 1, NY))), F(rng.random((NX, NY + 1)))
    Vx, Vy = Vx0.copy(order="F"), Vy0.copy(order="F")
    # accelerations and steps as functions of the face position they belong to
    xfx, yfx = XN[:, None], YC[None, 1:NY - 1]
    xfy, yfy = XC[1:NX - 1, None], YN[None, :]
    dVxdt, dtVx = F(1.0 + xfx + 10 * yfx), F(0.5 + 0.0 * xfx + 0.1 * yfx)
    dVydt, dtVy = F(2.0 - xfy + 7 * yfy), F(0.25 + 0.2 * xfy + 0.0 * yfy)
    O.stokes2d_v(Vx, Vy, dVxdt, dVydt, dtVx, dtVy)
    wx, wy = Vx0.copy(), Vy0.copy()
    wx[1:NX, 1:NY - 1] += (0.5 + 0.1 * yfx) * (1.0 + xfx + 10 * yfx)
    wy[1:NX - 1, 1:NY] += (0.25 + 0.2 * xfy) * (2.0 - xfy + 7 * yfy)
    assert close(Vx, wx, 1e-15) and close(Vy, wy, 1e-15)


def test_stokes2d_pt_iteration_reaches_the_hydrostatic_state():
    # uniform viscosity and body force, free-slip walls: the steady state is V = 0 and
    # dPt/dy = rho g; the whole pseudo-transient chain of miniapps/Stokes2D.jl:109-125 must get there
    nx = ny = 31
    lx = ly = 10.0
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    rg = 1.0
    Mus, Rog = F(np.ones((nx, ny))), F(np.full((nx, ny), rg))
    a = dict(Pt=_z(nx, ny), dtPt=_z(nx, ny), divV=_z(nx, ny), Vx=_z(nx + 1, ny), Vy=_z(nx, ny + 1), txx=_z(nx, ny),
             tyy=_z(nx, ny), txy=_z(nx - 1, ny - 1), Rx=_z(nx - 1, ny - 2), Ry=_z(nx - 2, ny - 1), dVxdt=_z(nx - 1, ny - 2),
             dVydt=_z(nx - 2, ny - 1), dtVx=_z(nx - 1, ny - 2), dtVy=_z(nx - 2, ny - 1))
    O.stokes2d_timesteps(a["dtVx"], a["dtVy"], a["dtPt"], Mus, 1.0, 0.25, min(dx, dy) ** 2, max(nx, ny))
    err = 1.0
    for it in range(1, 20001):
        O.stokes2d_p(a["divV"], a["Pt"], a["Vx"], a["Vy"], a["dtPt"], dx, dy)
        O.stokes2d_tau(a["divV"], a["txx"], a["tyy"], a["txy"], a["Vx"], a["Vy"], Mus, dx, dy)
        O.stokes2d_dv(a["Rx"], a["Ry"], a["dVxdt"], a["dVydt"], a["Pt"], Rog, a["txx"], a["tyy"], a["txy"],
                      1 - 4.0 / nx, 1 - 4.0 / ny, dx, dy)
        O.stokes2d_v(a["Vx"], a["Vy"], a["dVxdt"], a["dVydt"], a["dtVx"], a["dtVy"])
        O.bc_y(a["Vx"])
        O.bc_x(a["Vy"])
        if it % 100 == 0:
            err = max(O.mean_abs(a["Rx"]), O.mean_abs(a["Ry"]), O.mean_abs(a["divV"]))
            if err < 1e-10:
                break
    assert err < 1e-10
    assert np.abs(a["Vx"]).max() < 1e-8 and np.abs(a["Vy"]).max() < 1e-8
    dPdy = np.diff(a["Pt"], axis=1) / dy
    assert np.abs(dPdy[1:-1] - rg).max() < 1e-7


# ------------------------------------------------------------------ CellArray copy
def test_copy3d_cells_closed_form():
    # examples/memcopyCellArray3D.jl:10-13 adds cells value by value: with T holding the cell's
    # position and Ci the value's index inside the cell, T2 = T + Ci holds both, for every value
    # of every cell (a transposed cell, a swapped or dropped operand fails)
    nx, ny, nz, c1, c2 = 3, 4, 5, 2, 3
    ix, iy, iz, a, b = np.meshgrid(*(np.arange(k, dtype=np.float64) for k in (nx, ny, nz, c1, c2)), indexing="ij")
    T = F(ix + 10 * iy + 100 * iz + 0 * a)
    Ci = F(1000 * (a + c1 * b + 1) + 0 * ix)
    got = O.copy3d_cells(T, Ci)
    assert got.shape == (nx, ny, nz, c1, c2)
    assert np.array_equal(got, ix + 10 * iy + 100 * iz + 1000 * (a + c1 * b + 1))
