"""Pins the CPU oracle before it is trusted as the parity checker.

Each check is chosen so that a plausible mistake in the oracle (a dropped term,
a wrong sign or offset, a transposed operand) fails it: closed forms, invariants,
brute-force single-point evaluation and the reference's own known-answer tests.
"""
import numpy as np
import pytest

from oracle import ps_oracle as O


def F(a):
    return np.asfortranarray(a, dtype=np.float64)


# ------------------------------------------------------------ FD macro KATs
@pytest.mark.parametrize("name", sorted(O.FD3_ARG_OFFSET))
def test_fd3_shapes_match_reference_field_types(name):
    # R is an IField of nxyz=(9,7,8) in test_FiniteDifferences3D.jl:58; each macro's argument is the
    # field type whose size minus the macro's footprint gives R's size.
    R = (7, 5, 6)
    off = O.FD3_ARG_OFFSET[name]
    A = np.random.default_rng(1).random(tuple(R[i] + off[i] for i in range(3))) + 0.5
    assert O.fd3(name, F(A)).shape == R


def test_fd3_closed_forms():
    nx, ny, nz = 9, 7, 8
    ix, iy, iz = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1), indexing="ij")
    lin = F(3.0 * ix - 2.0 * iy + 5.0 * iz)
    assert np.all(O.fd3("d_xa", lin) == 3.0)
    assert np.all(O.fd3("d_ya", lin) == -2.0)
    assert np.all(O.fd3("d_za", lin) == 5.0)
    assert np.all(O.fd3("d_xi", lin) == 3.0) and O.fd3("d_xi", lin).shape == (8, 5, 6)
    assert np.all(O.fd3("d_zi", lin) == 5.0) and O.fd3("d_zi", lin).shape == (7, 5, 7)
    quad = F(1.0 * ix ** 2 + 3.0 * iy ** 2 - 2.0 * iz ** 2)
    assert np.all(O.fd3("d2_xi", quad) == 2.0)
    assert np.all(O.fd3("d2_yi", quad) == 6.0)
    assert np.all(O.fd3("d2_zi", quad) == -4.0)
    # averages of a linear field are its value at the staggered point
    assert np.all(O.fd3("av_xa", lin) == (3.0 * (np.arange(1, nx) + 0.5))[:, None, None] - 2.0 * iy[:-1, :, :] + 5.0 * iz[:-1, :, :])
    c = F(np.full((nx, ny, nz), 4.0))
    for n in ("av", "av_xya", "av_yzi", "harm", "harm_xza", "harm_zi"):
        assert np.all(O.fd3(n, c) == 4.0), n
    # harmonic mean of (1, 3) along x is 1.5
    a = F(np.where(ix % 2 == 1, 1.0, 3.0))
    assert np.all(O.fd3("harm_xa", a) == 1.5)
    # maxloc/minloc of a single peak / pit
    p = np.zeros((nx, ny, nz))
    p[4, 3, 4] = 7.0
    mx = O.fd3("maxloc", F(p))
    assert mx[3, 2, 3] == 7.0 and mx[2, 2, 3] == 7.0 and mx[3, 2, 2] == 7.0 and mx[2, 1, 2] == 0.0
    assert np.all(O.fd3("minloc", F(-p)) == -mx)
    # selection: inner
    assert np.array_equal(O.fd3("inn", lin), lin[1:-1, 1:-1, 1:-1])
    assert np.array_equal(O.fd3("inn_yz", lin), lin[:, 1:-1, 1:-1])


def test_fd2_fd1_closed_forms():
    nx, ny = 9, 7
    ix, iy = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), indexing="ij")
    lin = F(3.0 * ix - 2.0 * iy)
    assert np.all(O.fd2("d_xa", lin) == 3.0) and np.all(O.fd2("d_yi", lin) == -2.0)
    quad = F(1.0 * ix ** 2 - 3.0 * iy ** 2)
    assert np.all(O.fd2("d2_xa", quad) == 2.0) and np.all(O.fd2("d2_yi", quad) == -6.0)
    assert np.all(O.fd2("av", F(np.full((nx, ny), 2.5))) == 2.5)
    assert np.all(O.fd2("harm_yi", F(np.full((nx, ny), 2.5))) == 2.5)
    x = F(np.arange(1, 10, dtype=np.float64) ** 2)
    assert np.all(O.fd1("d2", x) == 2.0)
    assert np.all(O.fd1("d", F(np.arange(5.0))) == 1.0)
    assert np.all(O.fd1("av", F(np.arange(5.0))) == np.arange(4) + 0.5)


# ------------------------------------------------------------ diffusion
def test_otf_diffusion_matches_reference_kat():
    # test/test_parallel.jl:411-434: integer-cubed T, unit constants, Ci = 1, exact ==.
    nx, ny, nz = 32, 8, 8
    T = O.cubed_indices((nx, ny, nz))
    Ci = F(np.ones((nx, ny, nz)))
    T2 = F(np.zeros((nx, ny, nz)))
    O.diffusion3d_otf_test(T2, T, Ci, 1.0, 1.0, 1.0, 1.0, 1.0)
    ref = np.zeros_like(T)
    c = T[1:-1, 1:-1, 1:-1]
    ref[1:-1, 1:-1, 1:-1] = c + 1.0 * (1.0 * Ci[1:-1, 1:-1, 1:-1] * (
        ((T[2:, 1:-1, 1:-1] - c) - (c - T[:-2, 1:-1, 1:-1])) * 1.0
        + ((T[1:-1, 2:, 1:-1] - c) - (c - T[1:-1, :-2, 1:-1])) * 1.0
        + ((T[1:-1, 1:-1, 2:] - c) - (c - T[1:-1, 1:-1, :-2])) * 1.0))
    assert np.array_equal(T2, ref)


def test_diffusion_linear_field_is_stationary():
    n = 12
    x = np.arange(1, n + 1, dtype=np.float64)
    T = F(np.broadcast_to((2.0 * x - 7.0)[:, None, None], (n, n, n)))
    Cp = F(1.0 + np.random.default_rng(3).random((n, n, n)))
    for f in (O.diffusion3d_flux, O.diffusion3d_d2):
        T2 = F(np.full((n, n, n), -1.0))
        f(T2, T, Cp, 1.0, 0.01, 0.5, 0.5, 0.5)
        assert np.array_equal(T2[1:-1, 1:-1, 1:-1], T[1:-1, 1:-1, 1:-1])
        assert np.all(T2[0] == -1.0) and np.all(T2[:, -1] == -1.0)  # boundaries untouched


def test_diffusion_flux_single_point_and_d2_agreement():
    rng = np.random.default_rng(7)
    n = 10
    T = F(rng.random((n, n, n)))
    Cp = F(1.0 + rng.random((n, n, n)))
    lam, dt, dx, dy, dz = 1.3, 0.01, 0.5, 0.6, 0.7
    T2 = F(np.zeros_like(T))
    O.diffusion3d_flux(T2, T, Cp, lam, dt, dx, dy, dz)
    i, j, k = 4, 6, 3  # 0-based interior point
    qxp = -lam * (T[i + 1, j, k] - T[i, j, k]) / dx
    qxm = -lam * (T[i, j, k] - T[i - 1, j, k]) / dx
    qyp = -lam * (T[i, j + 1, k] - T[i, j, k]) / dy
    qym = -lam * (T[i, j, k] - T[i, j - 1, k]) / dy
    qzp = -lam * (T[i, j, k + 1] - T[i, j, k]) / dz
    qzm = -lam * (T[i, j, k] - T[i, j, k - 1]) / dz
    ref = T[i, j, k] + dt / Cp[i, j, k] * (-(qxp - qxm) / dx - (qyp - qym) / dy - (qzp - qzm) / dz)
    assert abs(T2[i, j, k] - ref) <= 1e-15 * abs(ref)
    # same physics in the d2 form with Ci = 1/Cp
    T3 = F(np.zeros_like(T))
    O.diffusion3d_d2(T3, T, F(1.0 / Cp), lam, dt, dx, dy, dz)
    assert np.allclose(T3[1:-1, 1:-1, 1:-1], T2[1:-1, 1:-1, 1:-1], rtol=1e-13, atol=0)


def test_diffusion_conserves_heat_with_no_flux():
    # sum(Cp*T) changes only through boundary fluxes: a field constant on the boundary
    # layers has zero boundary flux in the first step -> interior heat is conserved.
    rng = np.random.default_rng(11)
    n = 12
    T = np.full((n, n, n), 2.0)
    T[2:-2, 2:-2, 2:-2] += rng.random((n - 4, n - 4, n - 4))
    T = F(T)
    Cp = F(1.0 + rng.random((n, n, n)))
    T2 = F(T.copy())
    O.diffusion3d_flux(T2, T, Cp, 1.0, 1e-3, 1.0, 1.0, 1.0)
    s0 = np.sum((Cp * T)[1:-1, 1:-1, 1:-1])
    s1 = np.sum((Cp * T2)[1:-1, 1:-1, 1:-1])
    assert abs(s1 - s0) < 1e-11 * abs(s0)


# ------------------------------------------------------------ acoustic
def _acoustic3d_state(n, rng):
    P = F(rng.random((n, n, n)))
    Vx = F(np.zeros((n + 1, n, n)))
    Vy = F(np.zeros((n, n + 1, n)))
    Vz = F(np.zeros((n, n, n + 1)))
    return P, Vx, Vy, Vz


def test_acoustic3d_pressure_is_conserved():
    rng = np.random.default_rng(5)
    n = 10
    P, Vx, Vy, Vz = _acoustic3d_state(n, rng)
    s0 = P.sum()
    for _ in range(5):
        O.acoustic3d_v(Vx, Vy, Vz, P, 0.1, 1.0, 1.0, 1.0, 1.0)
        O.acoustic3d_p(P, Vx, Vy, Vz, 0.1, 1.0, 1.0, 1.0, 1.0)
    # velocities on the outer faces are never updated (masks) -> telescoping divergence
    assert np.all(Vx[0] == 0) and np.all(Vx[-1] == 0) and np.all(Vy[:, 0] == 0) and np.all(Vz[:, :, -1] == 0)
    assert abs(P.sum() - s0) < 1e-12 * abs(s0)


def test_acoustic3d_single_point():
    rng = np.random.default_rng(6)
    n = 8
    P, Vx, Vy, Vz = _acoustic3d_state(n, rng)
    Vx[:] = rng.random(Vx.shape)
    Vy[:] = rng.random(Vy.shape)
    Vz[:] = rng.random(Vz.shape)
    dt, rho, k, dx, dy, dz = 0.1, 1.2, 0.9, 0.5, 0.6, 0.7
    Vx0, Vy0, Vz0, P0 = Vx.copy(), Vy.copy(), Vz.copy(), P.copy()
    O.acoustic3d_v(Vx, Vy, Vz, P, dt, rho, dx, dy, dz)
    # Vx[ix+1,iy+1,iz+1] updated for ix in 1..n-1: 0-based i in 1..n-1
    i, j, k3 = 3, 2, 5
    assert Vx[i, j, k3] == Vx0[i, j, k3] - dt / rho * (P[i, j, k3] - P[i - 1, j, k3]) / dx
    assert Vy[i, j, k3] == Vy0[i, j, k3] - dt / rho * (P[i, j, k3] - P[i, j - 1, k3]) / dy
    assert Vz[i, j, k3] == Vz0[i, j, k3] - dt / rho * (P[i, j, k3] - P[i, j, k3 - 1]) / dz
    assert Vx[n, j, k3] == Vx0[n, j, k3]  # last face untouched
    O.acoustic3d_p(P, Vx, Vy, Vz, dt, k, dx, dy, dz)
    ref = P0[i, j, k3] - dt * k * ((Vx[i + 1, j, k3] - Vx[i, j, k3]) / dx + (Vy[i, j + 1, k3] - Vy[i, j, k3]) / dy
                                    + (Vz[i, j, k3 + 1] - Vz[i, j, k3]) / dz)
    assert P[i, j, k3] == ref


def test_acoustic_precisions_track_f64():
    rng = np.random.default_rng(9)
    n = 12
    x = (np.arange(n) - n / 2) / 2
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    P0 = np.exp(-(X ** 2 + Y ** 2 + Z ** 2))
    res = {}
    for prec, dt_ in (("f64", np.float64), ("f32", np.float32), ("bf16", np.uint16)):
        if prec == "bf16":
            from parallelstencil.core import f32_to_bf16, bf16_to_f32  # host conversion helpers only
            P = np.asfortranarray(f32_to_bf16(P0.astype(np.float32)))
            Vs = [np.asfortranarray(np.zeros(s, dtype=np.uint16)) for s in ((n + 1, n, n), (n, n + 1, n), (n, n, n + 1))]
        else:
            P = np.asfortranarray(P0.astype(dt_))
            Vs = [np.asfortranarray(np.zeros(s, dtype=dt_)) for s in ((n + 1, n, n), (n, n + 1, n), (n, n, n + 1))]
        for _ in range(10):
            O.acoustic3d_v(*Vs, P, 0.08, 1.0, 0.5, 0.5, 0.5, prec=prec)
            O.acoustic3d_p(P, *Vs, 0.08, 1.0, 0.5, 0.5, 0.5, prec=prec)
        res[prec] = bf16_to_f32(P).astype(np.float64) if prec == "bf16" else P.astype(np.float64)
    e32 = np.abs(res["f32"] - res["f64"]).max()
    e16 = np.abs(res["bf16"] - res["f64"]).max()
    assert 0 < e32 < 1e-5 and e32 < e16 < 5e-2


def test_acoustic2d_pressure_is_conserved():
    rng = np.random.default_rng(12)
    n = 16
    P = F(rng.random((n, n)))
    Vx, Vy = F(np.zeros((n + 1, n))), F(np.zeros((n, n + 1)))
    s0 = P.sum()
    for _ in range(4):
        O.acoustic2d_v(Vx, Vy, P, 0.1, 1.0, 1.0, 1.0)
        O.acoustic2d_p(P, Vx, Vy, 0.1, 1.0, 1.0, 1.0)
    assert abs(P.sum() - s0) < 1e-12 * abs(s0)


# ------------------------------------------------------------ Stokes
def _stokes_arrays(nx, ny):
    z = lambda *s: F(np.zeros(s))  # noqa: E731
    return dict(Pt=z(nx, ny), dtPt=z(nx, ny), divV=z(nx, ny), Vx=z(nx + 1, ny), Vy=z(nx, ny + 1), txx=z(nx, ny),
                tyy=z(nx, ny), txy=z(nx - 1, ny - 1), Rx=z(nx - 1, ny - 2), Ry=z(nx - 2, ny - 1), dVxdt=z(nx - 1, ny - 2),
                dVydt=z(nx - 2, ny - 1), dtVx=z(nx - 1, ny - 2), dtVy=z(nx - 2, ny - 1))


def test_stokes_uniform_velocity_has_no_strain():
    nx, ny = 12, 10
    a = _stokes_arrays(nx, ny)
    a["Vx"][:] = 0.7
    a["Vy"][:] = -0.3
    Mus = F(1.0 + np.random.default_rng(2).random((nx, ny)))
    O.stokes2d_p(a["divV"], a["Pt"], a["Vx"], a["Vy"], a["dtPt"], 0.1, 0.1)
    O.stokes2d_tau(a["divV"], a["txx"], a["tyy"], a["txy"], a["Vx"], a["Vy"], Mus, 0.1, 0.1)
    assert np.all(a["divV"] == 0) and np.all(a["txx"] == 0) and np.all(a["txy"] == 0)


def test_stokes_single_points():
    rng = np.random.default_rng(4)
    nx, ny = 9, 8
    a = _stokes_arrays(nx, ny)
    for k in a:
        a[k][:] = rng.random(a[k].shape)
    Mus = F(1.0 + rng.random((nx, ny)))
    Rog = F(rng.random((nx, ny)))
    dx, dy = 0.3, 0.4
    b = {k: v.copy() for k, v in a.items()}
    O.stokes2d_timesteps(a["dtVx"], a["dtVy"], a["dtPt"], Mus, 1.0, 0.25, min(dx, dy) ** 2, max(nx, ny))
    i, j = 2, 3
    assert a["dtVx"][i, j] == 1.0 * min(dx, dy) ** 2 / ((Mus[i, j + 1] + Mus[i + 1, j + 1]) * 0.5) / 4.1
    assert a["dtPt"][i, j] == 0.25 * 4.1 * Mus[i, j] / max(nx, ny)
    O.stokes2d_p(a["divV"], a["Pt"], a["Vx"], a["Vy"], a["dtPt"], dx, dy)
    dv = (a["Vx"][i + 1, j] - a["Vx"][i, j]) / dx + (a["Vy"][i, j + 1] - a["Vy"][i, j]) / dy
    assert a["divV"][i, j] == dv and a["Pt"][i, j] == b["Pt"][i, j] - a["dtPt"][i, j] * dv
    O.stokes2d_tau(a["divV"], a["txx"], a["tyy"], a["txy"], a["Vx"], a["Vy"], Mus, dx, dy)
    assert a["txx"][i, j] == 2.0 * Mus[i, j] * ((a["Vx"][i + 1, j] - a["Vx"][i, j]) / dx - 1.0 / 3.0 * a["divV"][i, j])
    av = (Mus[i, j] + Mus[i + 1, j] + Mus[i, j + 1] + Mus[i + 1, j + 1]) * 0.25
    assert a["txy"][i, j] == 2.0 * av * (0.5 * ((a["Vx"][i + 1, j + 1] - a["Vx"][i + 1, j]) / dy
                                                 + (a["Vy"][i + 1, j + 1] - a["Vy"][i, j + 1]) / dx))
    O.stokes2d_dv(a["Rx"], a["Ry"], a["dVxdt"], a["dVydt"], a["Pt"], Rog, a["txx"], a["tyy"], a["txy"], 0.9, 0.8, dx, dy)
    rx = (a["txx"][i + 1, j + 1] - a["txx"][i, j + 1]) / dx + (a["txy"][i, j + 1] - a["txy"][i, j]) / dy \
        - (a["Pt"][i + 1, j + 1] - a["Pt"][i, j + 1]) / dx
    assert a["Rx"][i, j] == rx and a["dVxdt"][i, j] == 0.9 * b["dVxdt"][i, j] + rx
    ry = (a["tyy"][i + 1, j + 1] - a["tyy"][i + 1, j]) / dy + (a["txy"][i + 1, j] - a["txy"][i, j]) / dx \
        - (a["Pt"][i + 1, j + 1] - a["Pt"][i + 1, j]) / dy + (Rog[i + 1, j] + Rog[i + 1, j + 1]) * 0.5
    assert a["Ry"][i, j] == ry
    O.stokes2d_v(a["Vx"], a["Vy"], a["dVxdt"], a["dVydt"], a["dtVx"], a["dtVy"])
    assert a["Vx"][i + 1, j + 1] == b["Vx"][i + 1, j + 1] + a["dtVx"][i, j] * a["dVxdt"][i, j]
    assert a["Vx"][0, j] == b["Vx"][0, j]


def test_bc_kernels():
    A = O.linear_indices((5, 4))[:, :, 0].copy(order="F")
    B = A.copy(order="F")
    O.bc_x(B)
    assert np.array_equal(B[0], A[1]) and np.array_equal(B[-1], A[-2]) and np.array_equal(B[1:-1], A[1:-1])
    C = A.copy(order="F")
    O.bc_y(C)
    assert np.array_equal(C[:, 0], A[:, 1]) and np.array_equal(C[:, -1], A[:, -2])


def test_mean_abs():
    a = F(np.array([[1.0, -2.0], [3.0, -4.0]]))
    assert O.mean_abs(a) == 2.5


# ------------------------------------------------- Stokes3D: reduction to Stokes2D on a plane
def _stokes3d_arrays(nx, ny, nz):
    z = lambda *s: F(np.zeros(s))  # noqa: E731
    sx, sy, sz = (nx - 1, ny - 2, nz - 2), (nx - 2, ny - 1, nz - 2), (nx - 2, ny - 2, nz - 1)
    return dict(Pt=z(nx, ny, nz), dtPt=z(nx, ny, nz), divV=z(nx, ny, nz), Mus=z(nx, ny, nz), Rog=z(nx, ny, nz),
                Vx=z(nx + 1, ny, nz), Vy=z(nx, ny + 1, nz), Vz=z(nx, ny, nz + 1),
                txx=z(nx, ny - 2, nz - 2), tyy=z(nx - 2, ny, nz - 2), tzz=z(nx - 2, ny - 2, nz),
                txy=z(nx - 1, ny - 1, nz - 2), txz=z(nx - 1, ny - 2, nz - 1), tyz=z(nx - 2, ny - 1, nz - 1),
                Rx=z(*sx), Ry=z(*sy), Rz=z(*sz), dVxdt=z(*sx), dVydt=z(*sy), dVzdt=z(*sz), dtVx=z(*sx), dtVy=z(*sy), dtVz=z(*sz))


def _crop(A, shape):
    """The centre of A with the given shape (staggered and inner ranges differ by 0 or 2)."""
    return A[tuple(slice((a - b) // 2, a - (a - b) // 2) for a, b in zip(A.shape, shape))]


@pytest.mark.parametrize("plane", [(0, 2), (1, 2), (0, 1)])
def test_stokes3d_reduces_to_stokes2d_on_a_plane(plane):
    """Fields invariant along the third axis c, with no velocity along it: every kernel of the 3-D
    restatement must give, on each plane, what the 2-D C oracle gives on the plane's axes (a, b):
    a dropped or misplaced term, a wrong stagger or a swapped spacing fails. Iterated with the
    free-slip boundaries; the off-plane stresses, residual and velocity must stay zero."""
    a, b = plane
    c = 3 - a - b
    X = "xyz"
    A, B, C = X[a], X[b], X[c]
    n2 = (9, 7)
    m = 5
    n3 = [0, 0, 0]
    n3[a], n3[b], n3[c] = n2[0], n2[1], m
    rng = np.random.default_rng(17 + a * 3 + b)
    s2 = _stokes_arrays(*n2)
    for k in ("Pt", "Vx", "Vy", "dVxdt", "dVydt"):
        s2[k][:] = rng.standard_normal(s2[k].shape)
    Mus2 = F(1.0 + rng.random(n2))
    Rog2 = F(rng.random(n2)) if b == 2 else F(np.zeros(n2))  # gravity acts along z in 3-D, along y in 2-D
    s3 = _stokes3d_arrays(*n3)

    def embed(A2, name3):
        shp = s3[name3].shape
        s3[name3][...] = np.broadcast_to(np.expand_dims(_crop(A2, (shp[a], shp[b])), c), shp)

    embed(s2["Pt"], "Pt"), embed(Mus2, "Mus"), embed(Rog2, "Rog")
    embed(s2["Vx"], "V" + A), embed(s2["Vy"], "V" + B)
    embed(s2["dVxdt"], f"dV{A}dt"), embed(s2["dVydt"], f"dV{B}dt")
    sp2 = (0.3, 0.4)
    sp3 = [0.7, 0.7, 0.7]
    sp3[a], sp3[b] = sp2
    damp2 = (0.9, 0.8)
    damp3 = [0.5, 0.5, 0.5]
    damp3[a], damp3[b] = damp2
    Vsc, Ptsc, mind2, maxn = 1.0, 0.25, min(sp2) ** 2, max(n3)
    O.stokes2d_timesteps(s2["dtVx"], s2["dtVy"], s2["dtPt"], Mus2, Vsc, Ptsc, mind2, maxn)
    O.stokes3d_timesteps(s3["dtVx"], s3["dtVy"], s3["dtVz"], s3["dtPt"], s3["Mus"], Vsc, Ptsc, mind2, maxn)
    # the 3-D pseudo-time steps use 6.1 where the 2-D ones use 4.1
    sl = [slice(None)] * 3
    sl[c] = 0
    sl = tuple(sl)
    np.testing.assert_allclose(s3["dtV" + A][sl] * 6.1, _crop(s2["dtVx"], s3["dtV" + A][sl].shape) * 4.1, rtol=1e-14)
    np.testing.assert_allclose(s3["dtV" + B][sl] * 6.1, _crop(s2["dtVy"], s3["dtV" + B][sl].shape) * 4.1, rtol=1e-14)
    np.testing.assert_allclose(s3["dtPt"][sl] / 6.1, s2["dtPt"] / 4.1, rtol=1e-14)
    embed(s2["dtVx"], "dtV" + A), embed(s2["dtVy"], "dtV" + B), embed(s2["dtPt"], "dtPt")
    prm = dict(dx=sp3[0], dy=sp3[1], dz=sp3[2], dampX=damp3[0], dampY=damp3[1], dampZ=damp3[2])
    pair = lambda u, v: "t" + "".join(sorted(u + v))  # noqa: E731
    match = {"Pt": "Pt", "divV": "divV", "V" + A: "Vx", "V" + B: "Vy", "t" + A + A: "txx", "t" + B + B: "tyy",
             pair(A, B): "txy", "R" + A: "Rx", "R" + B: "Ry", f"dV{A}dt": "dVxdt", f"dV{B}dt": "dVydt"}
    zero = [pair(A, C), pair(B, C), "R" + C, f"dV{C}dt", "V" + C]
    bc2 = {(0, 1): O.bc_x, (1, 0): O.bc_y}  # (boundary axis, velocity component) of the 2-D solver
    for it in range(3):
        O.stokes2d_p(s2["divV"], s2["Pt"], s2["Vx"], s2["Vy"], s2["dtPt"], *sp2)
        O.stokes2d_tau(s2["divV"], s2["txx"], s2["tyy"], s2["txy"], s2["Vx"], s2["Vy"], Mus2, *sp2)
        O.stokes2d_dv(s2["Rx"], s2["Ry"], s2["dVxdt"], s2["dVydt"], s2["Pt"], Rog2, s2["txx"], s2["tyy"], s2["txy"], *damp2, *sp2)
        O.stokes2d_v(s2["Vx"], s2["Vy"], s2["dVxdt"], s2["dVydt"], s2["dtVx"], s2["dtVy"])
        bc2[(1, 0)](s2["Vx"])
        bc2[(0, 1)](s2["Vy"])
        O.stokes3d_iterate(s3, prm, 1)
        for n3name, n2name in match.items():
            V3 = s3[n3name]
            assert np.all(V3 == V3[sl][tuple(np.newaxis if d == c else slice(None) for d in range(3))]), (it, n3name)
            np.testing.assert_allclose(V3[sl], _crop(s2[n2name], V3[sl].shape), rtol=1e-13, atol=1e-300,
                                       err_msg=f"iteration {it}: {n3name} vs 2-D {n2name}")
        for name in zero:
            assert not np.any(s3[name]), (it, name)


def test_stokes3d_rigid_motion_has_no_strain():
    """A uniform velocity field: no divergence, no stress, and with a uniform pressure and no
    buoyancy no residual."""
    s3 = _stokes3d_arrays(8, 7, 6)
    s3["Vx"][:], s3["Vy"][:], s3["Vz"][:], s3["Pt"][:] = 0.7, -0.3, 0.2, 1.5
    s3["Mus"][:] = 1.0 + np.random.default_rng(3).random(s3["Mus"].shape)
    O.stokes3d_iterate(s3, dict(dx=0.1, dy=0.2, dz=0.3, dampX=0.9, dampY=0.9, dampZ=0.9), 2)
    for k in ("divV", "txx", "tyy", "tzz", "txy", "txz", "tyz", "Rx", "Ry", "Rz", "dVxdt"):
        assert not np.any(s3[k]), k
    assert np.all(s3["Pt"] == 1.5)


# ------------------------------------------------- ThermalConvection2D
def test_thermal2d_stokes_part_is_stokes2d():
    """With rho = dtau_iter = 1 the PT kernels of ThermalConvection2D are Stokes2D's (compute_P!
    with dτPt = dτ_iter/β, compute_τ! on Eta, compute_dV! on RogT, compute_V! with dτV = 1):
    equal to the C restatement bit for bit; rho = 2 halves the residuals exactly."""
    rng = np.random.default_rng(23)
    nx, ny = 11, 9
    s2 = _stokes_arrays(nx, ny)
    for k in ("Pt", "Vx", "Vy", "dVxdt", "dVydt"):
        s2[k][:] = rng.standard_normal(s2[k].shape)
    t = {k: v.copy(order="F") for k, v in s2.items()}
    T = F(rng.standard_normal((nx, ny)))
    z = lambda *sh: F(np.zeros(sh))  # noqa: E731
    RogT, Eta, sxy = z(nx, ny), z(nx, ny), z(nx - 1, ny - 1)
    dx, dy, beta, rho0ga = 0.3, 0.4, 3.7, 2.5
    O.thermal2d_0(RogT, Eta, t["divV"], T, t["Vx"], t["Vy"], rho0ga, 1.3, 0.2, 1.0, dx, dy)
    O.thermal2d_1(t["Pt"], t["txx"], t["tyy"], sxy, Eta, t["divV"], t["Vx"], t["Vy"], 1.0, beta, dx, dy)
    s2["dtPt"][:] = 1.0 / beta
    O.stokes2d_p(s2["divV"], s2["Pt"], s2["Vx"], s2["Vy"], s2["dtPt"], dx, dy)
    O.stokes2d_tau(s2["divV"], s2["txx"], s2["tyy"], s2["txy"], s2["Vx"], s2["Vy"], Eta, dx, dy)
    for a, b in (("divV", "divV"), ("Pt", "Pt"), ("txx", "txx"), ("tyy", "tyy")):
        assert np.array_equal(t[a], s2[b]), a
    assert np.array_equal(sxy, s2["txy"])
    r2 = {k: t[k].copy(order="F") for k in ("Rx", "Ry", "dVxdt", "dVydt")}
    O.thermal2d_2(t["Rx"], t["Ry"], t["dVxdt"], t["dVydt"], t["Pt"], RogT, t["txx"], t["tyy"], sxy, 1.0, 0.9, 0.8, 1.0, dx, dy)
    O.stokes2d_dv(s2["Rx"], s2["Ry"], s2["dVxdt"], s2["dVydt"], s2["Pt"], RogT, s2["txx"], s2["tyy"], s2["txy"], 0.9, 0.8, dx, dy)
    for k in ("Rx", "Ry", "dVxdt", "dVydt"):
        assert np.array_equal(t[k], s2[k]), k
    O.thermal2d_2(r2["Rx"], r2["Ry"], r2["dVxdt"], r2["dVydt"], t["Pt"], RogT, t["txx"], t["tyy"], sxy, 2.0, 0.9, 0.8, 1.0, dx, dy)
    assert np.array_equal(r2["Rx"] * 2.0, t["Rx"]) and np.array_equal(r2["Ry"] * 2.0, t["Ry"])
    O.thermal2d_v(t["Vx"], t["Vy"], t["dVxdt"], t["dVydt"], 1.0)
    s2["dtVx"][:], s2["dtVy"][:] = 1.0, 1.0
    O.stokes2d_v(s2["Vx"], s2["Vy"], s2["dVxdt"], s2["dVydt"], s2["dtVx"], s2["dtVy"])
    assert np.array_equal(t["Vx"], s2["Vx"]) and np.array_equal(t["Vy"], s2["Vy"])
    Tm = F(np.full((nx, ny), -0.5))  # T = -dT/2: Eta = eta0 exactly, RogT = rho0ga * T
    O.thermal2d_0(RogT, Eta, t["divV"], Tm, t["Vx"], t["Vy"], rho0ga, 1.3, 0.2, 1.0, dx, dy)
    assert np.all(Eta == 1.3) and np.all(RogT == rho0ga * -0.5)


@pytest.mark.parametrize("u", [0.5, -0.5])
def test_thermal2d_upwind_advection_closed_forms(u):
    """On an integer grid (dx = dy = 1, DcT = 1) T = x^2 has backward and forward differences
    2x-1 and 2x+1 and a constant flux divergence -2: dT_dt = 2 - u (2x-1) for u > 0 (upwind from
    the left) and 2 - u (2x+1) for u < 0."""
    nx, ny = 9, 7
    x = np.arange(nx, dtype=np.float64)
    T = F(np.repeat((x * x)[:, None], ny, 1))
    z = lambda *sh: F(np.zeros(sh))  # noqa: E731
    qTx, qTy, dT_dt = z(nx - 1, ny - 2), z(nx - 2, ny - 1), z(nx - 2, ny - 2)
    Vx, Vy = F(np.full((nx + 1, ny), u)), z(nx, ny + 1)
    O.thermal2d_qt(qTx, qTy, T, 1.0, 1.0, 1.0)
    O.thermal2d_advect(dT_dt, qTx, qTy, T, Vx, Vy, 1.0, 1.0)
    r = x[1:-1][:, None]  # the T row of each dT_dt row
    want = 2.0 - u * (2 * r - 1) if u > 0 else 2.0 - u * (2 * r + 1)
    assert np.array_equal(dT_dt, np.broadcast_to(want, dT_dt.shape))
    # the same in y, and the 2-D flux divergence of x^2 + y^2 with no velocity: -(-4) = 4
    y = np.arange(ny, dtype=np.float64)
    T2 = F((x * x)[:, None] + (y * y)[None, :])
    O.thermal2d_qt(qTx, qTy, T2, 1.0, 1.0, 1.0)
    O.thermal2d_advect(dT_dt, qTx, qTy, T2, z(nx + 1, ny), z(nx, ny + 1), 1.0, 1.0)
    assert np.all(dT_dt == 4.0)


def test_thermal2d_strong_zero():
    """Julia's false * Inf is 0: an infinite velocity upwind-excluded at a node adds nothing there."""
    nx, ny = 6, 5
    T = F(np.repeat((np.arange(nx, dtype=np.float64) ** 2)[:, None], ny, 1))
    z = lambda *sh: F(np.zeros(sh))  # noqa: E731
    qTx, qTy, dT_dt = z(nx - 1, ny - 2), z(nx - 2, ny - 1), z(nx - 2, ny - 2)
    Vx = z(nx + 1, ny)
    Vx[3, 2] = np.inf  # the right face of dT_dt row 1 (Vx < 0 false there), the left face of row 2
    O.thermal2d_advect(dT_dt, qTx, qTy, T, Vx, z(nx, ny + 1), 1.0, 1.0)
    assert np.isfinite(dT_dt[1, 1]) and dT_dt[2, 1] == -np.inf
    assert O.julia_bool_mul(np.array([False]), np.array([-np.inf]))[0] == 0.0


# ------------------------------------------------- scalar porosity waves
def _poro_fields(nx, ny):
    z = lambda *sh: F(np.zeros(sh))  # noqa: E731
    return dict(Eta=z(nx - 2, ny - 2), EtaR=z(nx - 2, ny - 2), dPedt=z(nx - 2, ny - 2), dPhidt=z(nx - 2, ny - 2), Pe=z(nx, ny),
                K_mu=z(nx, ny), Phi=z(nx, ny), qx=z(nx - 1, ny - 2), qy=z(nx - 2, ny - 1))


def test_poro_porosity_changes_by_the_boundary_fluxes_only():
    """No compaction term (EtaR = Inf): dPhidt = -β dPedt = div q, so the interior sum of Phi
    changes by dt times the net boundary flux (the telescoping sums of d_xa(qx), d_ya(qy))."""
    rng = np.random.default_rng(29)
    nx, ny = 12, 10
    f = _poro_fields(nx, ny)
    f["Phi"][:] = 1.0 + rng.random((nx, ny))
    f["Pe"][1:-1, 1:-1] = rng.standard_normal((nx - 2, ny - 2))
    f["K_mu"][:] = 0.5 + rng.random((nx, ny))
    f["EtaR"][:] = np.inf
    dx, dy, dt, beta = 0.5, 0.25, 1e-3, 0.2
    O.poro_flux(f["qx"], f["qy"], f["K_mu"], f["Pe"], 1.0, dx, dy)
    boundary = (f["qx"][-1, :].sum() - f["qx"][0, :].sum()) / dx + (f["qy"][:, -1].sum() - f["qy"][:, 0].sum()) / dy
    s0 = f["Phi"][1:-1, 1:-1].sum()
    O.poro_pe_phi(f["dPedt"], f["Pe"], f["Phi"], f["dPhidt"], f["qx"], f["qy"], f["EtaR"], beta, dt, dx, dy)
    assert abs((f["Phi"][1:-1, 1:-1].sum() - s0) - dt * boundary) <= 1e-12 * (abs(s0) + 1)
    assert abs(f["dPedt"].sum() * -beta - boundary) <= 1e-10 * (abs(boundary) + 1)


def test_poro_fluxes_and_clamp():
    nx, ny = 8, 7
    f = _poro_fields(nx, ny)
    f["K_mu"][:] = 2.0
    x, y = np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64)
    f["Pe"][:] = 3.0 * x[:, None] - 1.0 * y[None, :]
    O.poro_flux(f["qx"], f["qy"], f["K_mu"], f["Pe"], 0.5, 1.0, 1.0)
    assert np.all(f["qx"] == -2.0 * 3.0) and np.all(f["qy"] == -2.0 * (-1.0 + 0.5))
    # a large outflow drives the porosity below zero: clamped at 0
    f["Phi"][:] = 0.1
    f["EtaR"][:] = 1.0
    f["qx"][:] = 0.0
    f["qy"][:] = 0.0
    f["Pe"][:] = -100.0  # dPedt = 500, Pe -> 400, dPhidt = -100 - 400
    O.poro_pe_phi(f["dPedt"], f["Pe"], f["Phi"], f["dPhidt"], f["qx"], f["qy"], f["EtaR"], 0.2, 1.0, 1.0, 1.0)
    assert np.all(f["Phi"][1:-1, 1:-1] == 0.0) and np.all(f["Phi"][0] == 0.1)


def test_poro_etar_strict_guard():
    nx, ny = 7, 6
    EtaR = F(np.full((nx - 2, ny - 2), 5.0))
    Pe = F(-np.ones((nx, ny)))
    O.poro_etar(EtaR, Pe, 0.25)
    assert np.all(EtaR[:-1, :-1] == 0.25) and np.all(EtaR[-1, :] == 5.0) and np.all(EtaR[:, -1] == 5.0)
    Pe[2, 2] = 1.0  # EtaR[1, 1] reads Pe[2, 2]
    EtaR[:] = 5.0
    O.poro_etar(EtaR, Pe, 0.25)
    assert EtaR[1, 1] == 5.0 and EtaR[0, 0] == 0.25


# ------------------------------------------------- HydroMech2D
def test_hydromech_stokes_part_is_stokes2d():
    rng = np.random.default_rng(31)
    nx, ny = 11, 9
    s2 = _stokes_arrays(nx, ny)
    for k in ("Pt", "Vx", "Vy", "dVxdt", "dVydt"):
        s2[k][:] = rng.standard_normal(s2[k].shape)
    p = dict(dx=0.3, dy=0.4, mus=0.5, beta_n=0.0, dtauPt=0.0, dampX=0.9, dampY=0.8)
    z = lambda *sh: F(np.zeros(sh))  # noqa: E731
    h = {k: v.copy(order="F") for k, v in s2.items()}
    RPt, RPf, dtPf, Rog = F(rng.standard_normal((nx, ny))), z(nx, ny), z(nx, ny), F(rng.random((nx, ny)))
    O.stokes2d_p(s2["divV"], s2["Pt"], s2["Vx"], s2["Vy"], s2["dtPt"], p["dx"], p["dy"])  # dtPt = 0: Pt unchanged
    h["divV"][...] = s2["divV"]
    O.hm_ptau(h["Pt"], z(nx, ny), h["txx"], h["tyy"], h["txy"], RPt, RPf, dtPf, h["Vx"], h["Vy"], h["divV"], p)
    O.stokes2d_tau(s2["divV"], s2["txx"], s2["tyy"], s2["txy"], s2["Vx"], s2["Vy"], F(np.full((nx, ny), 0.5)), p["dx"], p["dy"])
    for k in ("Pt", "txx", "tyy", "txy"):
        assert np.array_equal(h[k], s2[k]), k
    O.hm_res(h["Rx"], h["Ry"], h["dVxdt"], h["dVydt"], h["txx"], h["tyy"], h["txy"], h["Pt"], Rog, p)
    O.stokes2d_dv(s2["Rx"], s2["Ry"], s2["dVxdt"], s2["dVydt"], s2["Pt"], F(-Rog), s2["txx"], s2["tyy"], s2["txy"], 0.9, 0.8, p["dx"], p["dy"])
    for k in ("Rx", "Ry", "dVxdt", "dVydt"):
        assert np.array_equal(h[k], s2[k]), k
    # relaxation limits of compute_params_∇!
    q = dict(p, theta_e=0.0, theta_k=1.0, eta2mus=10.0, R=500.0, lamPe=0.01, k_muf0=1.0, phi0=0.01, nperm=3.0, rhofg=1.0, rhosg=2.0,
             rhogBG=1.99)
    Phi = F(0.01 + 0.01 * rng.random((nx, ny)))
    EtaC, K, divqD = F(rng.random((nx, ny))), z(nx, ny), z(nx, ny)
    E0 = EtaC.copy()
    Pt = F(rng.random((nx, ny)))
    O.hm_params(EtaC, K, Rog, h["divV"], divqD, Phi, Pt.copy(), Pt, h["Vx"], h["Vy"], z(nx + 1, ny), z(nx, ny + 1), q)
    assert np.array_equal(EtaC, E0) and np.allclose(K, (Phi / 0.01) ** 3, rtol=1e-15) and np.array_equal(h["divV"], s2["divV"])
    q["theta_e"] = 1.0
    O.hm_params(EtaC, K, Rog, h["divV"], divqD, Phi, Pt.copy(), Pt, h["Vx"], h["Vy"], z(nx + 1, ny), z(nx, ny + 1), q)
    np.testing.assert_allclose(EtaC, 0.5 / Phi * 10.0 * (1.0 + 0.5 * (1.0 / 500.0 - 1.0)), rtol=1e-15)


# ------------------------------------------------- acoustic 2-D adjoint (no reference counterpart)
def _ac2_case(nx=22, ny=19, nt=50):
    rng = np.random.default_rng(5)
    K = 1.0 + 0.1 * rng.random((nx, ny))
    prm = dict(dt=0.05, rho=1.0, dx=0.3, dy=0.35)
    t = (np.arange(nt) + 1) * prm["dt"] - 0.8
    src = prm["dt"] * (1 - 2 * (3 * t) ** 2) * np.exp(-(3 * t) ** 2)
    geo = dict(isx=9, isy=7, rx0=3, rdx=2, ry=15, nrec=8)
    return K, prm, src, geo, nt


def _zero_boundary_v(Vx, Vy):
    Vx[0], Vx[-1], Vx[:, 0], Vx[:, -1] = 0, 0, 0, 0
    Vy[0], Vy[-1], Vy[:, 0], Vy[:, -1] = 0, 0, 0, 0


def test_acoustic2d_adjoint_step_dot_product():
    """<A x, y> = <x, A^T y> for one step (velocity boundary nodes zero: they are constant)."""
    nx, ny = 22, 19
    rng = np.random.default_rng(1)
    K = 1.0 + 0.1 * rng.random((nx, ny))
    x = [rng.standard_normal((nx, ny)), rng.standard_normal((nx + 1, ny)), rng.standard_normal((nx, ny + 1))]
    y = [rng.standard_normal((nx, ny)), rng.standard_normal((nx + 1, ny)), rng.standard_normal((nx, ny + 1))]
    _zero_boundary_v(x[1], x[2])
    _zero_boundary_v(y[1], y[2])
    Ax, ATy = [a.copy() for a in x], [a.copy() for a in y]
    O.acoustic2d_step(*Ax, K, 0.05, 1.0, 0.3, 0.35)
    O.acoustic2d_step_adjoint(*ATy, K, 0.05, 1.0, 0.3, 0.35)
    lhs = sum(float(np.sum(a * b)) for a, b in zip(Ax, y))
    rhs = sum(float(np.sum(a * b)) for a, b in zip(x, ATy))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # a plausible mistake (sign of the adjoint velocity update) breaks the identity
    bad = [a.copy() for a in y]
    O.acoustic2d_step_adjoint(bad[0], -bad[1], -bad[2], K, 0.05, 1.0, 0.3, 0.35)
    assert abs(sum(float(np.sum(a * b)) for a, b in zip(x, bad)) - lhs) > 1e-6 * abs(lhs)


def test_acoustic2d_gradient_finite_differences():
    K, prm, src, geo, nt = _ac2_case()
    args = (nt, prm["dt"], prm["rho"], prm["dx"], prm["dy"], src, geo["isx"], geo["isy"], geo["rx0"], geo["rdx"], geo["ry"],
            geo["nrec"])
    Dobs = O.acoustic2d_traces(np.ones_like(K), *args)
    J, g = O.acoustic2d_misfit_gradient(K, Dobs, *args)
    assert J > 0
    rng = np.random.default_rng(2)
    for _ in range(4):
        d = rng.standard_normal(K.shape)
        h = 1e-4
        r = lambda Kp: 0.5 * float(np.sum((O.acoustic2d_traces(Kp, *args) - Dobs) ** 2))
        fd = (r(K + h * d) - r(K - h * d)) / (2 * h)
        assert abs(fd - float(np.sum(g * d))) <= 1e-6 * abs(fd)


# ------------------------------------------------------------ kernel language
def test_diffusion2d_shmem_linear_field_is_stationary():
    nx, ny = 13, 9
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    T = F(3.0 * ix - 2.0 * iy + 7.0)
    Ci = F(np.random.default_rng(2).random((nx, ny)) + 0.5)
    T2 = F(np.full((nx, ny), -1.0))
    O.diffusion2d_shmem_step(T2, T, Ci, 1.3, 0.01, 4.0, 2.0)
    assert np.array_equal(T2[1:-1, 1:-1], T[1:-1, 1:-1])
    assert np.all(T2[0, :] == -1.0) and np.all(T2[:, -1] == -1.0)  # boundary untouched


def test_diffusion2d_shmem_quadratic_field():
    # T = x^2 + 3 y^2 on the integer grid: second differences are exactly 2 and 6, so with lam=1,
    # _dx=_dy=1, Ci=1 the update is T + dt*(2 + 6) (a sign or neighbour mistake changes it)
    nx, ny = 11, 10
    ix, iy = np.meshgrid(np.arange(nx, dtype=float), np.arange(ny, dtype=float), indexing="ij")
    T = F(ix * ix + 3.0 * iy * iy)
    T2 = F(np.zeros((nx, ny)))
    O.diffusion2d_shmem_step(T2, T, F(np.ones((nx, ny))), 1.0, 0.125, 1.0, 1.0)
    assert np.array_equal(T2[1:-1, 1:-1], T[1:-1, 1:-1] + 0.125 * 8.0)


def test_diffusion2d_shmem_brute_force():
    rng = np.random.default_rng(4)
    T, Ci = F(rng.random((8, 7))), F(rng.random((8, 7)) + 0.5)
    T2 = F(np.zeros((8, 7)))
    O.diffusion2d_shmem_step(T2, T, Ci, 0.7, 0.003, 7.0, 6.0)
    for i in range(1, 7):
        for j in range(1, 6):
            assert T2[i, j] == O.diffusion2d_shmem_cell(T, Ci, 0.7, 0.003, 7.0, 6.0, i, j)


def test_wavefront_model_closed_forms():
    rng = np.random.default_rng(9)
    v = rng.integers(-100, 100, 64).astype(np.float64)
    s = v.copy()
    for m in (32, 16, 8, 4, 2, 1):  # xor butterfly: every lane ends with the wave sum
        s = s + O.wave_shfl_xor(s, m, 64)
    assert np.all(s == v.sum())
    for w in (8, 16, 64):  # a partial butterfly sums each partition
        s = v.copy()
        m = w // 2
        while m:
            s = s + O.wave_shfl_xor(s, m, w)
            m //= 2
        assert np.array_equal(s, np.repeat(v.reshape(-1, w).sum(axis=1), w))
    lanes = np.arange(64)
    assert np.array_equal(O.wave_shfl(v, lanes % 16 + 1, 16), v)  # own lane: identity
    assert np.array_equal(O.wave_shfl(v, np.ones(64, dtype=int), 16), np.repeat(v[::16], 16))  # broadcast lane 1
    up = O.wave_shfl_up(v, 3, 32)
    assert np.array_equal(up[3:32], v[:29]) and np.array_equal(up[:3], v[:3]) and np.array_equal(up[32:35], v[32:35])
    dn = O.wave_shfl_down(v, 5, 32)
    assert np.array_equal(dn[:27], v[5:32]) and np.array_equal(dn[27:32], v[27:32])
    assert np.array_equal(O.wave_shfl_xor(v, 24, 16), v)  # partner in another partition: own value
    assert O.wave_ballot(lanes % 2 == 0) == np.array([0x5555555555555555], dtype=np.uint64).view(np.int64)[0]
    assert O.wave_ballot(np.ones(64, bool)) == -1 and O.wave_ballot(np.zeros(64, bool)) == 0
    assert bin(int(np.array([O.wave_ballot(lanes < 10)]).view(np.uint64)[0])).count("1") == 10


# ------------------------------------------------------------ CellArrays
def test_cellarray_layout_closed_forms():
    N, nc = 10, 6
    for I in range(N):
        for j in range(nc):
            assert O.cellarray_data_offset(I, j, 1, nc) == I * nc + j  # cells stored whole
            assert O.cellarray_data_offset(I, j, N, nc) == j * N + I  # each value contiguous
    for B in (1, 3, 4, 10, 16):
        nb = -(-N // B)
        offs = sorted(O.cellarray_data_offset(I, j, B, nc) for I in range(N) for j in range(nc))
        assert len(set(offs)) == N * nc and offs[-1] < B * nc * nb  # one slot per value, inside the data


def test_cells_to_data_against_numpy_views():
    rng = np.random.default_rng(8)
    dims, celldims = (5, 3, 2), (2, 3)
    cells = np.asfortranarray(rng.integers(0, 1000, dims + celldims).astype(np.float64))
    N, nc = 30, 6
    flat = cells.reshape((N, nc), order="F")
    for B in (1, 4, 30):
        nb = -(-N // B)
        data = O.cells_to_data(cells, celldims, B)
        pad = np.zeros((B * nb, nc))
        pad[:N] = flat
        # (B, nb, nc) view of the cells, values moved to the middle axis
        assert np.array_equal(data, pad.reshape((B, nb, nc), order="F").transpose(0, 2, 1))


def test_linear_jacobian_closed_form():
    n = 7

    def d2(x):  # interior second difference, boundary rows zero
        y = np.zeros_like(x)
        y[1:-1] = x[2:] - 2 * x[1:-1] + x[:-2]
        return y

    J = O.linear_jacobian(d2, n)
    want = np.zeros((n, n))
    for i in range(1, n - 1):
        want[i, i - 1:i + 2] = (1, -2, 1)
    assert np.array_equal(J, want)
    M = np.arange(12.0).reshape(3, 4)
    assert np.array_equal(O.linear_jacobian(lambda x: M @ x, 4), M)
    assert np.array_equal(O.linear_jacobian(lambda x: M.T @ x, 3), M.T)


def test_complex_step_closed_forms():
    x = np.linspace(-1.3, 1.7, 11)
    one = np.ones_like(x)
    assert np.allclose(O.complex_step_jvp(np.exp, [x], [one]), np.exp(x), rtol=1e-15, atol=0)
    assert np.allclose(O.complex_step_jvp(lambda z: z ** 3, [x], [one]), 3 * x ** 2, rtol=1e-15, atol=1e-300)
    assert np.allclose(O.complex_step_jvp(lambda z: np.log(1 + z * z), [x], [one]), 2 * x / (1 + x * x), rtol=1e-15, atol=1e-300)
    xp = x + 2.0
    assert np.allclose(O.complex_step_jvp(np.sqrt, [xp], [one]), 0.5 / np.sqrt(xp), rtol=1e-15, atol=0)
    v = np.cos(x)
    got = O.complex_step_jvp(lambda a, b: a * b, [x, xp], [v, None])
    assert np.allclose(got, v * xp, rtol=1e-15, atol=1e-300)
    # vjp of a linear map is J^T y
    M = np.random.default_rng(5).standard_normal((5, 4))
    y = np.arange(1.0, 6.0)
    (g,) = O.complex_step_vjp(lambda a: M @ a, [np.zeros(4)], [0], y)
    assert np.allclose(g, O.linear_jacobian(lambda a: M @ a, 4).T @ y, rtol=1e-14, atol=1e-14)


def _periodic_samples(shape, nxyz, overlaps, halowidths, periods):
    """f at the periodic image of every cell: cell i of a dimension of size s (overlap
    ol = overlap + s - nx) sits at global position i - 1 - (s - nx)/2 - 1 (staggered arrays are
    shifted by half a cell), taken modulo the global size nx - overlap."""
    pos = []
    for d in range(3):
        s = shape[d] if d < len(shape) else 1
        i = np.arange(1, s + 1, dtype=float)
        if periods[d] and s > 1:
            ng = nxyz[d] - overlaps[d]
            p = i - 1 - 0.5 * (s - nxyz[d]) - 1
            pos.append(np.mod(p, ng) / ng)
        else:
            pos.append(i / 1000.0)
    X, Y, Z = np.meshgrid(*pos, indexing="ij")
    f = np.sin(2 * np.pi * X) + 0.5 * np.cos(2 * np.pi * Y) + 0.25 * np.sin(2 * np.pi * Z + 0.3) + X * 0 + 3.0
    return f.reshape(shape, order="F")


@pytest.mark.parametrize("case", ["ol2", "ol4", "staggered", "nonperiodic_y"])
def test_update_halo_periodic_restores_periodic_images(case):
    nxyz = (14, 12, 10)
    overlaps, halowidths, periods, shape = (2, 2, 2), (1, 1, 1), (1, 1, 1), nxyz
    if case == "ol4":
        overlaps, halowidths = (4, 4, 4), (2, 2, 2)
    if case == "staggered":
        shape = (15, 12, 11)
    if case == "nonperiodic_y":
        periods = (1, 0, 1)
    want = _periodic_samples(shape, nxyz, overlaps, halowidths, periods)
    A = want.copy()
    for d in range(3):  # the halo cells of the periodic dimensions are unknown before the update
        if not periods[d]:
            continue
        s, hw = shape[d], halowidths[d]
        sl = [slice(None)] * 3
        for i in list(range(hw)) + list(range(s - hw, s)):
            sl[d] = i
            A[tuple(sl)] = np.nan
    got = O.update_halo_periodic(A, nxyz, overlaps, halowidths, periods)
    assert np.array_equal(got, want)
