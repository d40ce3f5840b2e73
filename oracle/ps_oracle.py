"""CPU oracle for the ParallelStencil test-bed kernels -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker (and, for bench.py, as the timed CPU baseline). The
product path (parallelstencil.jl_amd) never imports it.

What it contains
  * ctypes wrappers of ps_oracle.c: the Threads-backend loop nest of the
    reference (src/ParallelKernel/parallel.jl:715-768) for the test-bed kernels,
    compiled with -ffp-contract=off (Julia semantics);
  * numpy restatements of the reference's finite-difference known-answer
    formulas (test/test_FiniteDifferences{1,2,3}D.jl) and of the
    closed-form-input generator of test/test_parallel.jl;
  * seeded input generators (SplitMix64) for the benchmark configurations.

How each function is pinned (a check that a plausible mistake fails), so
that parity is not anchored on a formula merely retyped:
  fd1/fd2/fd3       reference KAT formulas themselves (test_FiniteDifferences*D.jl:
                    the broadcast expressions are the golden values), plus
                    closed forms in tests/test_oracle.py: d(linear)=slope,
                    d2(quadratic)=2a, av/harm of constants, maxloc of a known peak.
  diffusion3d_otf_test   test/test_parallel.jl:411-434 KAT (integer-cubed input,
                    unit constants: exact), and against diffusion3d_flux with
                    _d = 1/d where both are exact, and the d2 form.
  diffusion3d_flux  linear-in-x field is stationary (d2 = 0 exactly); agreement with
                    the d2 form to rounding for lam=1 constant Cp (mathematically
                    identical); conservation of sum(T*Cp) over the interior
                    except boundary fluxes; single-cell brute-force evaluation.
  diffusion3d_d2    mathematically equal to the flux form (tolerance), closed form on
                    a quadratic field (d2 = 2 exactly on integer grids).
  acoustic2d/3d     closed forms (tests/test_oracle_closed_forms.py): a quadratic P
                    gives every inner face exactly V - dt/rho * dP/dx of the continuous
                    field at the face position, and leaves the boundary faces alone; a
                    quadratic V gives P - dt*K*div V at the centres (a flipped sign, a
                    difference shifted by one cell or a wrong mask fails, shown by planted
                    mistakes); sum(P) is conserved with zero boundary velocities;
                    FP32/BF16 variants against FP64.
  stokes2d_*        closed forms on polynomial fields at the staggered positions
                    (tests/test_oracle_closed_forms.py): pseudo-time steps of a linear
                    viscosity are the face viscosity's; divergence, pressure update and
                    stresses of quadratic/shear velocities; residuals of polynomial
                    stresses and pressure are the momentum balance at the faces with the
                    buoyancy average; hydrostatic pressure balances a uniform buoyancy
                    exactly; the velocity update lands on the inner faces; the whole PT
                    chain with free-slip walls converges to the hydrostatic state
                    (V = 0, dPt/dy = rho g). Fifteen planted mistakes (signs, offsets,
                    masks, spacings, damping, the 1/3) each fail at least one of them.
  bc_x/bc_y         KAT on an index-valued array.
  thermal2d_0/1/2/v the Stokes part of ThermalConvection2D against stokes2d_* (exact:
                    with rho = dtau_iter = 1 compute_1!/compute_2!/update_V! are compute_τ!,
                    compute_P! with dτPt = dτ_iter/β, compute_dV!, compute_V!; rho = 2 halves
                    the residuals exactly); Eta = eta0 exactly where T = -dT/2.
  thermal2d_qt/advect  closed forms on integer grids: T = x^2 gives the backward and forward
                    differences 2x-1 and 2x+1, so the upwind choice for either sign of the
                    velocity is visible; the flux divergence of T = x^2 + y^2 is -4 DcT
                    exactly; Julia's strong zero (false * Inf == 0).
  poro_*            conservation: with no compaction term (EtaR = Inf) the interior sum of
                    Phi changes by dt times the boundary fluxes alone (telescoping sums of
                    d_xa/d_ya), which a wrong sign or index breaks; uniform K and a linear
                    Pe give the fluxes -K slope and -K (slope + Δρg); the clamp at 0; the
                    strict guard of compute_EtaR! on a KAT (last row and column untouched).
  hm_*              the Stokes part against stokes2d_* (exact): compute_res! is compute_dV!
                    with the buoyancy negated; compute_P_τ! with β_n = 0 and μs = 0.5 is
                    compute_τ! on a uniform viscosity; the divergences are compute_P!'s;
                    relaxation limits θ = 0 (unchanged) and θ = 1 with Pf = Pt (tanh 0 = 0).
  stokes3d_*        reduction to the 2-D C restatement: fields invariant along one axis,
                    with no velocity along it, must give on each of the three planes
                    (x-z, y-z, x-y) exactly what stokes2d_* gives on the plane's axes,
                    through three iterations with the free-slip boundaries, while the
                    off-plane stresses, residual and velocity stay zero (every term of
                    every statement is active on one plane; a swapped spacing, dropped
                    term, wrong damping or missing boundary launch fails, checked by
                    planted mistakes); a uniform velocity gives no strain.
  diffusion2d_shmem_step  a linear field is stationary (flux differences are exactly 0);
                    single-cell brute force in Python floats (diffusion2d_shmem_cell);
                    a wrong sign or swapped neighbour changes a quadratic field's
                    exactly known discrete Laplacian (2 on integer grids).
  cellarray_data_offset / cells_to_data  closed forms: B = 1 stores cells whole
                    (I*nc + j), B >= N stores each value contiguously (j*B + I); a
                    blocked case against numpy's own reshape/transpose of the
                    (B, nblocks, nc) view; every offset hit exactly once.
  copy3d_cells      closed form: cells holding their position plus values holding their
                    index inside the cell add to a known integer per value (a transposed
                    cell or a dropped operand fails).
  wave_*            closed forms in tests/test_oracle.py: the xor butterfly over
                    masks 32..1 leaves every lane with the wave sum; shfl_up then
                    shfl_down by the same delta restores the lanes away from the
                    partition edges; shfl with lane = own lane is the identity;
                    the ballot of "lane is even" is 0x5555...; popcounts.
  linear_jacobian   the 1-D second difference gives the tridiagonal (1, -2, 1)
                    matrix; a map and its transpose give transposed matrices.
  update_halo_periodic  a periodic function sampled on the grid, with the halo cells set to
                    NaN, is restored exactly at every cell (its periodic images), for overlap 2
                    / halo width 1, overlap 4 / halo width 2, staggered arrays (overlap 3)
                    and all three dimensions; non-periodic dimensions are left alone.
  complex_step_*    closed forms: d exp = exp, d x^3 = 3x^2, d log(1+x^2) = 2x/(1+x^2),
                    d sqrt, a product rule; the vjp of a linear map equals J^T y with J
                    from linear_jacobian; a dropped imaginary part or wrong h scaling
                    fails them.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


def _lib(fast=False):
    key = "fast" if fast else "exact"
    if key not in _libs:
        path = os.path.join(_HERE, "libpsoracle_fast.so" if fast else "libpsoracle.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        _libs[key] = L
    return _libs[key]


def _p(a):
    assert a.flags.f_contiguous or a.ndim == 1 or a.flags.c_contiguous and a.ndim <= 1
    return a.ctypes.data_as(ctypes.c_void_p)


def _f(a, dtype=np.float64):
    if not (isinstance(a, np.ndarray) and a.dtype == dtype and a.flags.f_contiguous):
        raise TypeError("oracle arrays must be Fortran-ordered numpy arrays of the right dtype")
    return a.ctypes.data_as(ctypes.c_void_p)


D = ctypes.c_double
F = ctypes.c_float
I = ctypes.c_int


def set_threads(n):
    """Threads used by the OpenMP loops (the Threads backend's nthreads)."""
    os.environ["OMP_NUM_THREADS"] = str(n)
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(int(n))
    except OSError:
        pass


# ---------------------------------------------------------------- kernels
def triad(T2, T, Ci, fast=False):
    L = _lib(fast)
    L.pso_triad(ctypes.c_int64(T.size), _f(T2), _f(T), _f(Ci))


def diffusion3d_flux(T2, T, Cp, lam, dt, dx, dy, dz, fast=False):
    nx, ny, nz = T.shape
    _lib(fast).pso_diffusion3d_flux(I(nx), I(ny), I(nz), _f(T2), _f(T), _f(Cp), D(lam), D(dt), D(dx), D(dy), D(dz))


def diffusion3d_otf_test(T2, T, Ci, lam, dt, _dx, _dy, _dz):
    nx, ny, nz = T.shape
    _lib().pso_diffusion3d_otf_test(I(nx), I(ny), I(nz), _f(T2), _f(T), _f(Ci), D(lam), D(dt), D(_dx), D(_dy), D(_dz))


def diffusion3d_d2(T2, T, Ci, lam, dt, dx, dy, dz):
    nx, ny, nz = T.shape
    _lib().pso_diffusion3d_d2(I(nx), I(ny), I(nz), _f(T2), _f(T), _f(Ci), D(lam), D(dt), D(dx), D(dy), D(dz))


def acoustic3d_v(Vx, Vy, Vz, P, dt, rho, dx, dy, dz, prec="f64"):
    nx, ny, nz = P.shape
    if prec == "f64":
        _lib().pso_acoustic3d_v_f64(I(nx), I(ny), I(nz), _f(Vx), _f(Vy), _f(Vz), _f(P), D(dt), D(rho), D(dx), D(dy), D(dz))
    elif prec == "f32":
        t = np.float32
        _lib().pso_acoustic3d_v_f32(I(nx), I(ny), I(nz), _f(Vx, t), _f(Vy, t), _f(Vz, t), _f(P, t), F(dt), F(rho), F(dx), F(dy), F(dz))
    else:
        t = np.uint16
        _lib().pso_acoustic3d_v_bf16(I(nx), I(ny), I(nz), _f(Vx, t), _f(Vy, t), _f(Vz, t), _f(P, t), F(dt), F(rho), F(dx), F(dy), F(dz))


def acoustic3d_p(P, Vx, Vy, Vz, dt, k, dx, dy, dz, prec="f64"):
    nx, ny, nz = P.shape
    if prec == "f64":
        _lib().pso_acoustic3d_p_f64(I(nx), I(ny), I(nz), _f(P), _f(Vx), _f(Vy), _f(Vz), D(dt), D(k), D(dx), D(dy), D(dz))
    elif prec == "f32":
        t = np.float32
        _lib().pso_acoustic3d_p_f32(I(nx), I(ny), I(nz), _f(P, t), _f(Vx, t), _f(Vy, t), _f(Vz, t), F(dt), F(k), F(dx), F(dy), F(dz))
    else:
        t = np.uint16
        _lib().pso_acoustic3d_p_bf16(I(nx), I(ny), I(nz), _f(P, t), _f(Vx, t), _f(Vy, t), _f(Vz, t), F(dt), F(k), F(dx), F(dy), F(dz))


def acoustic2d_v(Vx, Vy, P, dt, rho, dx, dy):
    nx, ny = P.shape
    _lib().pso_acoustic2d_v_f64(I(nx), I(ny), _f(Vx), _f(Vy), _f(P), D(dt), D(rho), D(dx), D(dy))


def acoustic2d_p(P, Vx, Vy, dt, k, dx, dy):
    nx, ny = P.shape
    _lib().pso_acoustic2d_p_f64(I(nx), I(ny), _f(P), _f(Vx), _f(Vy), D(dt), D(k), D(dx), D(dy))


def stokes2d_timesteps(dtVx, dtVy, dtPt, Mus, Vsc, Ptsc, min_dxy2, max_nxy):
    nx, ny = Mus.shape
    _lib().pso_stokes2d_timesteps(I(nx), I(ny), _f(dtVx), _f(dtVy), _f(dtPt), _f(Mus), D(Vsc), D(Ptsc), D(min_dxy2), ctypes.c_int64(max_nxy))


def stokes2d_p(divV, Pt, Vx, Vy, dtPt, dx, dy):
    nx, ny = Pt.shape
    _lib().pso_stokes2d_p(I(nx), I(ny), _f(divV), _f(Pt), _f(Vx), _f(Vy), _f(dtPt), D(dx), D(dy))


def stokes2d_tau(divV, txx, tyy, txy, Vx, Vy, Mus, dx, dy):
    nx, ny = Mus.shape
    _lib().pso_stokes2d_tau(I(nx), I(ny), _f(divV), _f(txx), _f(tyy), _f(txy), _f(Vx), _f(Vy), _f(Mus), D(dx), D(dy))


def stokes2d_dv(Rx, Ry, dVxdt, dVydt, Pt, Rog, txx, tyy, txy, dampX, dampY, dx, dy):
    nx, ny = Pt.shape
    _lib().pso_stokes2d_dv(I(nx), I(ny), _f(Rx), _f(Ry), _f(dVxdt), _f(dVydt), _f(Pt), _f(Rog), _f(txx), _f(tyy), _f(txy),
                           D(dampX), D(dampY), D(dx), D(dy))


def stokes2d_v(Vx, Vy, dVxdt, dVydt, dtVx, dtVy):
    nx, ny = Vy.shape[0], Vx.shape[1]
    _lib().pso_stokes2d_v(I(nx), I(ny), _f(Vx), _f(Vy), _f(dVxdt), _f(dVydt), _f(dtVx), _f(dtVy))


def bc_x(A):
    _lib().pso_bc_x(I(A.shape[0]), I(A.shape[1]), _f(A))


def bc_y(A):
    _lib().pso_bc_y(I(A.shape[0]), I(A.shape[1]), _f(A))


def mean_abs(A):
    L = _lib()
    L.pso_mean_abs.restype = D
    a = np.asfortranarray(A, dtype=np.float64)
    return L.pso_mean_abs(ctypes.c_int64(a.size), _f(a))


# ------------------------------------------------- Stokes3D (numpy, over the pinned fd3 helpers)
def stokes3d_timesteps(dtVx, dtVy, dtVz, dtPt, Mus, Vsc, Ptsc, min_dxyz2, max_nxyz):
    """compute_timesteps! of miniapps/Stokes3D.jl:11-17."""
    dtVx[...] = Vsc * min_dxyz2 / fd3("av_xi", Mus) / 6.1
    dtVy[...] = Vsc * min_dxyz2 / fd3("av_yi", Mus) / 6.1
    dtVz[...] = Vsc * min_dxyz2 / fd3("av_zi", Mus) / 6.1
    dtPt[...] = Ptsc * 6.1 * Mus / float(max_nxyz)


def stokes3d_p(divV, Pt, Vx, Vy, Vz, dtPt, dx, dy, dz):
    """compute_P! of miniapps/Stokes3D.jl:19-23."""
    divV[...] = fd3("d_xa", Vx) / dx + fd3("d_ya", Vy) / dy + fd3("d_za", Vz) / dz
    Pt[...] = Pt - dtPt * divV


def stokes3d_tau(divV, txx, tyy, tzz, txy, txz, tyz, Vx, Vy, Vz, Mus, dx, dy, dz):
    """compute_τ! of miniapps/Stokes3D.jl:25-33."""
    third = 1.0 / 3.0
    txx[...] = 2.0 * fd3("inn_yz", Mus) * (fd3("d_xi", Vx) / dx - third * fd3("inn_yz", divV))
    tyy[...] = 2.0 * fd3("inn_xz", Mus) * (fd3("d_yi", Vy) / dy - third * fd3("inn_xz", divV))
    tzz[...] = 2.0 * fd3("inn_xy", Mus) * (fd3("d_zi", Vz) / dz - third * fd3("inn_xy", divV))
    txy[...] = 2.0 * fd3("av_xyi", Mus) * (0.5 * (fd3("d_yi", Vx) / dy + fd3("d_xi", Vy) / dx))
    txz[...] = 2.0 * fd3("av_xzi", Mus) * (0.5 * (fd3("d_zi", Vx) / dz + fd3("d_xi", Vz) / dx))
    tyz[...] = 2.0 * fd3("av_yzi", Mus) * (0.5 * (fd3("d_zi", Vy) / dz + fd3("d_yi", Vz) / dy))


def stokes3d_dv(Rx, Ry, Rz, dVxdt, dVydt, dVzdt, Pt, Rog, txx, tyy, tzz, txy, txz, tyz, dampX, dampY, dampZ, dx, dy, dz):
    """compute_dV! of miniapps/Stokes3D.jl:35-43."""
    Rx[...] = fd3("d_xa", txx) / dx + fd3("d_ya", txy) / dy + fd3("d_za", txz) / dz - fd3("d_xi", Pt) / dx
    Ry[...] = fd3("d_ya", tyy) / dy + fd3("d_xa", txy) / dx + fd3("d_za", tyz) / dz - fd3("d_yi", Pt) / dy
    Rz[...] = fd3("d_za", tzz) / dz + fd3("d_xa", txz) / dx + fd3("d_ya", tyz) / dy - fd3("d_zi", Pt) / dz + fd3("av_zi", Rog)
    dVxdt[...] = dampX * dVxdt + Rx
    dVydt[...] = dampY * dVydt + Ry
    dVzdt[...] = dampZ * dVzdt + Rz


def stokes3d_v(Vx, Vy, Vz, dVxdt, dVydt, dVzdt, dtVx, dtVy, dtVz):
    """compute_V! of miniapps/Stokes3D.jl:45-50."""
    inn = (slice(1, -1),) * 3
    Vx[inn] = Vx[inn] + dtVx * dVxdt
    Vy[inn] = Vy[inn] + dtVy * dVydt
    Vz[inn] = Vz[inn] + dtVz * dVzdt


def stokes3d_bc(A, axis):
    """bc_x! / bc_y! / bc_z! of miniapps/Stokes3D.jl:52-68 (axis 0, 1, 2): free slip, the first
    and last plane copy their inner neighbours."""
    lo, lo1, hi, hi1 = ([slice(None)] * 3 for _ in range(4))
    lo[axis], lo1[axis], hi[axis], hi1[axis] = 0, 1, -1, -2
    A[tuple(lo)] = A[tuple(lo1)]
    A[tuple(hi)] = A[tuple(hi1)]


def stokes3d_iterate(F, prm, n):
    """n PT iterations (miniapps/Stokes3D.jl:133-142) on a dict of Fortran-ordered float64 arrays
    named as the fields of the GPU solver (Pt, divV, Vx, ..., dtVz); prm holds dx, dy, dz and
    dampX, dampY, dampZ."""
    dx, dy, dz = prm["dx"], prm["dy"], prm["dz"]
    for _ in range(n):
        stokes3d_p(F["divV"], F["Pt"], F["Vx"], F["Vy"], F["Vz"], F["dtPt"], dx, dy, dz)
        stokes3d_tau(F["divV"], F["txx"], F["tyy"], F["tzz"], F["txy"], F["txz"], F["tyz"], F["Vx"], F["Vy"], F["Vz"], F["Mus"],
                     dx, dy, dz)
        stokes3d_dv(F["Rx"], F["Ry"], F["Rz"], F["dVxdt"], F["dVydt"], F["dVzdt"], F["Pt"], F["Rog"], F["txx"], F["tyy"],
                    F["tzz"], F["txy"], F["txz"], F["tyz"], prm["dampX"], prm["dampY"], prm["dampZ"], dx, dy, dz)
        stokes3d_v(F["Vx"], F["Vy"], F["Vz"], F["dVxdt"], F["dVydt"], F["dVzdt"], F["dtVx"], F["dtVy"], F["dtVz"])
        for axis, name in ((0, "Vy"), (0, "Vz"), (1, "Vx"), (1, "Vz"), (2, "Vx"), (2, "Vy")):
            stokes3d_bc(F[name], axis)


# ------------------------------------------------- ThermalConvection2D (numpy, over the pinned fd2 helpers)
def julia_bool_mul(b, x):
    """Julia's Bool * x for floats: x where b, else copysign(0, x) (a strong zero)."""
    return np.where(b, x, np.copysign(0.0, x))


def thermal2d_0(RogT, Eta, divV, T, Vx, Vy, rho0ga, eta0, deta_dT, dT, dx, dy):
    """compute_0! of miniapps/ThermalConvection2D.jl:21-26."""
    RogT[...] = rho0ga * T
    Eta[...] = eta0 * (1.0 - deta_dT * (T + dT / 2.0))
    divV[...] = fd2("d_xa", Vx) / dx + fd2("d_ya", Vy) / dy


def thermal2d_1(Pt, txx, tyy, sxy, Eta, divV, Vx, Vy, dtau_iter, beta, dx, dy):
    """compute_1! of miniapps/ThermalConvection2D.jl:28-34."""
    Pt[...] = Pt - dtau_iter / beta * divV
    txx[...] = 2.0 * Eta * (fd2("d_xa", Vx) / dx - 1.0 / 3.0 * divV)
    tyy[...] = 2.0 * Eta * (fd2("d_ya", Vy) / dy - 1.0 / 3.0 * divV)
    sxy[...] = 2.0 * fd2("av", Eta) * (0.5 * (fd2("d_yi", Vx) / dy + fd2("d_xi", Vy) / dx))


def thermal2d_2(Rx, Ry, dVxdt, dVydt, Pt, RogT, txx, tyy, sxy, rho, dampX, dampY, dtau_iter, dx, dy):
    """compute_2! of miniapps/ThermalConvection2D.jl:36-42."""
    Rx[...] = 1.0 / rho * (fd2("d_xi", txx) / dx + fd2("d_ya", sxy) / dy - fd2("d_xi", Pt) / dx)
    Ry[...] = 1.0 / rho * (fd2("d_yi", tyy) / dy + fd2("d_xa", sxy) / dx - fd2("d_yi", Pt) / dy + fd2("av_yi", RogT))
    dVxdt[...] = dampX * dVxdt + Rx * dtau_iter
    dVydt[...] = dampY * dVydt + Ry * dtau_iter


def thermal2d_v(Vx, Vy, dVxdt, dVydt, dtau_iter):
    """update_V! of miniapps/ThermalConvection2D.jl:44-48."""
    Vx[1:-1, 1:-1] = Vx[1:-1, 1:-1] + dVxdt * dtau_iter
    Vy[1:-1, 1:-1] = Vy[1:-1, 1:-1] + dVydt * dtau_iter


def thermal2d_qt(qTx, qTy, T, DcT, dx, dy):
    """compute_qT! of miniapps/ThermalConvection2D.jl:50-54."""
    qTx[...] = -DcT * fd2("d_xi", T) / dx
    qTy[...] = -DcT * fd2("d_yi", T) / dy


def thermal2d_advect(dT_dt, qTx, qTy, T, Vx, Vy, dx, dy):
    """advect_T! of miniapps/ThermalConvection2D.jl:56-63: flux divergence and first-order
    upwind advection; dT_dt[i, j] (1-based) belongs to T[i+1, j+1]."""
    n1, n2 = dT_dt.shape
    c = (slice(1, n1 + 1), slice(1, n2 + 1))  # T[ix+1, iy+1]
    xm, xp = (slice(0, n1), slice(1, n2 + 1)), (slice(2, n1 + 2), slice(1, n2 + 1))
    ym, yp = (slice(1, n1 + 1), slice(0, n2)), (slice(1, n1 + 1), slice(2, n2 + 2))
    Vxl, Vxr = Vx[1:n1 + 1, 1:n2 + 1], Vx[2:n1 + 2, 1:n2 + 1]
    Vyl, Vyr = Vy[1:n1 + 1, 1:n2 + 1], Vy[1:n1 + 1, 2:n2 + 2]
    div = (qTx[1:n1 + 1, :n2] - qTx[:n1, :n2]) / dx + (qTy[:n1, 1:n2 + 1] - qTy[:n1, :n2]) / dy
    dT_dt[...] = ((((-div) - julia_bool_mul(Vxl > 0, Vxl) * (T[c] - T[xm]) / dx)
                   - julia_bool_mul(Vxr < 0, Vxr) * (T[xp] - T[c]) / dx)
                  - julia_bool_mul(Vyl > 0, Vyl) * (T[c] - T[ym]) / dy) - julia_bool_mul(Vyr < 0, Vyr) * (T[yp] - T[c]) / dy


def thermal2d_t(T, T_old, dT_dt, dt):
    """update_T! of miniapps/ThermalConvection2D.jl:65-68."""
    T[1:-1, 1:-1] = T_old[1:-1, 1:-1] + dT_dt * dt


def thermal2d_noflux(T):
    """no_fluxY_T! of miniapps/ThermalConvection2D.jl:70-74 (the last row first, then the first)."""
    T[-1, :] = T[-2, :]
    T[0, :] = T[1, :]


def thermal2d_pt_iterate(F, prm, n):
    """n PT iterations of miniapps/ThermalConvection2D.jl:160-169 on a dict of float64 arrays
    named as the GPU solver's fields; prm holds the solver's scalars."""
    p = prm
    for _ in range(n):
        F["ErrV"][...] = F["Vy"]
        F["ErrP"][...] = F["Pt"]
        thermal2d_0(F["RogT"], F["Eta"], F["divV"], F["T"], F["Vx"], F["Vy"], p["rho0ga"], p["eta0"], p["deta_dT"], p["dT"], p["dx"], p["dy"])
        thermal2d_1(F["Pt"], F["txx"], F["tyy"], F["sxy"], F["Eta"], F["divV"], F["Vx"], F["Vy"], p["dtau_iter"], p["beta"], p["dx"], p["dy"])
        thermal2d_2(F["Rx"], F["Ry"], F["dVxdt"], F["dVydt"], F["Pt"], F["RogT"], F["txx"], F["tyy"], F["sxy"], p["rho"], p["dampX"],
                    p["dampY"], p["dtau_iter"], p["dx"], p["dy"])
        thermal2d_v(F["Vx"], F["Vy"], F["dVxdt"], F["dVydt"], p["dtau_iter"])
        F["Vx"][:, 0], F["Vx"][:, -1] = F["Vx"][:, 1], F["Vx"][:, -2]
        F["Vy"][0, :], F["Vy"][-1, :] = F["Vy"][1, :], F["Vy"][-2, :]
        F["ErrV"][...] = F["ErrV"] - F["Vy"]
        F["ErrP"][...] = F["ErrP"] - F["Pt"]


def thermal2d_time_step(F, prm):
    """One time step (miniapps/ThermalConvection2D.jl:155-183): PT iterations in blocks of nerr
    until errV and errP are at or below eps (or itermax), then the thermal step. Returns
    (iterations, errV, errP, dt)."""
    p = prm
    F["T_old"][...] = F["T"]
    errV = errP = 2 * p["eps"]
    niter = 0
    while (errV > p["eps"] or errP > p["eps"]) and niter < p["itermax"]:
        thermal2d_pt_iterate(F, p, p["nerr"])
        niter += p["nerr"]
        errV = np.abs(F["ErrV"]).max() / (1e-12 + np.abs(F["Vy"]).max())
        errP = np.abs(F["ErrP"]).max() / (1e-12 + np.abs(F["Pt"]).max())
    thermal2d_qt(F["qTx"], F["qTy"], F["T"], p["DcT"], p["dx"], p["dy"])
    thermal2d_advect(F["dT_dt"], F["qTx"], F["qTy"], F["T"], F["Vx"], F["Vy"], p["dx"], p["dy"])
    with np.errstate(divide="ignore"):
        dt_adv = min(p["dx"] / np.abs(F["Vx"]).max(), p["dy"] / np.abs(F["Vy"]).max()) / 2.1
    dt = min(p["dt_diff"], dt_adv)
    thermal2d_t(F["T"], F["T_old"], F["dT_dt"], dt)
    thermal2d_noflux(F["T"])
    return niter, errV, errP, dt


# ------------------------------------------------- scalar porosity waves (numpy, over the pinned fd2 helpers)
def poro_nonlin(K_mu, EtaR, Phi, Eta, k_mu0, n):
    """compute_nonlin! of miniapps/scalar_porowaves2D.jl:11-15."""
    K_mu[...] = k_mu0 * np.power(Phi, n)
    EtaR[...] = Eta


def poro_etar(EtaR, Pe, etaR):
    """compute_EtaR! of miniapps/scalar_porowaves2D.jl:17-20: the guard is ix < size(EtaR, 1)
    (strict), so the last row and column of EtaR are never set."""
    n1, n2 = EtaR.shape
    sub = EtaR[:n1 - 1, :n2 - 1]
    sub[Pe[1:n1, 1:n2] < 0.0] = etaR


def poro_flux(qx, qy, K_mu, Pe, drhog, dx, dy):
    """compute_flux! of miniapps/scalar_porowaves2D.jl:22-26."""
    qx[...] = -fd2("av_xi", K_mu) * fd2("d_xi", Pe) / dx
    qy[...] = -fd2("av_yi", K_mu) * (fd2("d_yi", Pe) / dy + drhog)


def poro_pe_phi(dPedt, Pe, Phi, dPhidt, qx, qy, EtaR, beta, dt, dx, dy):
    """compute_Pe_Phi! of miniapps/scalar_porowaves2D.jl:28-34 (Pe updated before its second use)."""
    inn = (slice(1, -1), slice(1, -1))
    dPedt[...] = 1.0 / beta * ((-fd2("d_xa", qx) / dx - fd2("d_ya", qy) / dy) - 1.0 / EtaR * Pe[inn])
    Pe[inn] = Pe[inn] + dt * dPedt
    dPhidt[...] = -beta * dPedt - 1.0 / EtaR * Pe[inn]
    Phi[inn] = np.maximum(0.0, Phi[inn] + dt * dPhidt)


def poro_steps(F, prm, it0, dt, n):
    """n time steps (miniapps/scalar_porowaves2D.jl:84-89) from step number it0 with time step dt;
    returns (last step number, dt)."""
    p = prm
    it = it0
    for _ in range(n):
        it += 1
        poro_nonlin(F["K_mu"], F["EtaR"], F["Phi"], F["Eta"], p["k_mu0"], p["n"])
        poro_etar(F["EtaR"], F["Pe"], p["etaR"])
        if it % 100 == 1:
            dt = p["dx"] ** 2 / (F["K_mu"].max() / p["beta"]) / 5.0
        poro_flux(F["qx"], F["qy"], F["K_mu"], F["Pe"], p["drhog"], p["dx"], p["dy"])
        poro_pe_phi(F["dPedt"], F["Pe"], F["Phi"], F["dPhidt"], F["qx"], F["qy"], F["EtaR"], p["beta"], dt, p["dx"], p["dy"])
    return it, dt


# ------------------------------------------------- HydroMech2D (numpy, over the pinned fd2 helpers)
def hm_params(EtaC, K_muf, Rog, divV, divqD, Phi, Pf, Pt, Vx, Vy, qDx, qDy, p):
    """compute_params_∇! of miniapps/HydroMech2D.jl:17-24."""
    th_e, th_k = p["theta_e"], p["theta_k"]
    EtaC[...] = (1.0 - th_e) * EtaC + th_e * (p["mus"] / Phi * p["eta2mus"] * (1.0 + 0.5 * (1.0 / p["R"] - 1.0) * (1.0 + np.tanh((Pf - Pt) / p["lamPe"]))))
    K_muf[...] = (1.0 - th_k) * K_muf + th_k * (p["k_muf0"] * np.power(Phi / p["phi0"], p["nperm"]))
    Rog[...] = p["rhofg"] * Phi + p["rhosg"] * (1.0 - Phi) - p["rhogBG"]
    divV[...] = fd2("d_xa", Vx) / p["dx"] + fd2("d_ya", Vy) / p["dy"]
    divqD[...] = fd2("d_xa", qDx) / p["dx"] + fd2("d_ya", qDy) / p["dy"]


def hm_rp(dtPf, RPt, RPf, K_muf, divV, divqD, Pt, Pf, EtaC, Phi, p):
    """compute_RP! of miniapps/HydroMech2D.jl:26-31."""
    dtPf[1:-1, 1:-1] = p["min_dxy2"] / fd2("maxloc", K_muf) / 4.1 / p["Pfsc"]
    c = (Pt - Pf) / (EtaC * (1.0 - Phi))
    RPt[...] = -divV - c
    RPf[...] = RPf * p["Pfdmp"] - divqD + c


def hm_ptau(Pt, Pf, txx, tyy, sxy, RPt, RPf, dtPf, Vx, Vy, divV, p):
    """compute_P_τ! of miniapps/HydroMech2D.jl:33-40."""
    mus, bn = p["mus"], p["beta_n"]
    Pt[...] = Pt + p["dtauPt"] * RPt
    Pf[...] = Pf + dtPf * RPf
    txx[...] = 2.0 * mus * (fd2("d_xa", Vx) / p["dx"] - 1.0 / 3.0 * divV - bn * RPt)
    tyy[...] = 2.0 * mus * (fd2("d_ya", Vy) / p["dy"] - 1.0 / 3.0 * divV - bn * RPt)
    sxy[...] = 2.0 * mus * (0.5 * (fd2("d_yi", Vx) / p["dy"] + fd2("d_xi", Vy) / p["dx"]))


def hm_res(Rx, Ry, dVxdt, dVydt, txx, tyy, sxy, Pt, Rog, p):
    """compute_res! of miniapps/HydroMech2D.jl:42-48."""
    dx, dy = p["dx"], p["dy"]
    Rx[...] = fd2("d_xi", txx) / dx + fd2("d_ya", sxy) / dy - fd2("d_xi", Pt) / dx
    Ry[...] = fd2("d_yi", tyy) / dy + fd2("d_xa", sxy) / dx - fd2("d_yi", Pt) / dy - fd2("av_yi", Rog)
    dVxdt[...] = p["dampX"] * dVxdt + Rx
    dVydt[...] = p["dampY"] * dVydt + Ry


def hm_update(Vx, Vy, qDx, qDy, Phi, dVxdt, dVydt, K_muf, Pf, Phi_o, divV, divV_o, dt, p):
    """compute_update! of miniapps/HydroMech2D.jl:50-57."""
    inn = (slice(1, -1), slice(1, -1))
    Vx[inn] = Vx[inn] + p["dtauV"] * dVxdt
    Vy[inn] = Vy[inn] + p["dtauV"] * dVydt
    qDx[inn] = -fd2("av_xi", K_muf) * (fd2("d_xi", Pf) / p["dx"])
    qDy[inn] = -fd2("av_yi", K_muf) * (fd2("d_yi", Pf) / p["dy"] + (p["rhofg"] - p["rhogBG"]))
    CN = p["CN"]
    Phi[...] = Phi_o + (1.0 - Phi) * (CN * divV_o + (1.0 - CN) * divV) * dt


def hm_iterate(F, p, dt, n):
    """n PT iterations of miniapps/HydroMech2D.jl:157-169 (the boundary launches with one index)."""
    for _ in range(n):
        hm_params(F["EtaC"], F["K_muf"], F["Rog"], F["divV"], F["divqD"], F["Phi"], F["Pf"], F["Pt"], F["Vx"], F["Vy"], F["qDx"], F["qDy"], p)
        hm_rp(F["dtPf"], F["RPt"], F["RPf"], F["K_muf"], F["divV"], F["divqD"], F["Pt"], F["Pf"], F["EtaC"], F["Phi"], p)
        F["dtPf"][0, :], F["dtPf"][-1, :] = F["dtPf"][1, :], F["dtPf"][-2, :]
        F["dtPf"][:, 0], F["dtPf"][:, -1] = F["dtPf"][:, 1], F["dtPf"][:, -2]
        hm_ptau(F["Pt"], F["Pf"], F["txx"], F["tyy"], F["sxy"], F["RPt"], F["RPf"], F["dtPf"], F["Vx"], F["Vy"], F["divV"], p)
        hm_res(F["Rx"], F["Ry"], F["dVxdt"], F["dVydt"], F["txx"], F["tyy"], F["sxy"], F["Pt"], F["Rog"], p)
        hm_update(F["Vx"], F["Vy"], F["qDx"], F["qDy"], F["Phi"], F["dVxdt"], F["dVydt"], F["K_muf"], F["Pf"], F["Phi_o"], F["divV"],
                  F["divV_o"], dt, p)
        for A in (F["Vx"], F["qDx"]):
            A[:, 0], A[:, -1] = A[:, 1], A[:, -2]
        for A in (F["Vy"], F["qDy"]):
            A[0, :], A[-1, :] = A[1, :], A[-2, :]


def hm_time_step(F, p, dt):
    """One time step (miniapps/HydroMech2D.jl:151-197); returns (iterations, error, next dt)."""
    F["Phi_o"][...] = F["Phi"]
    F["divV_o"][...] = F["divV"]
    err, niter = 2 * p["eps"], 0
    while err > p["eps"] and niter < p["itermax"]:
        hm_iterate(F, p, dt, p["nout"])
        niter += p["nout"]
        err = max(np.linalg.norm(F["Ry"]) / F["Ry"].size, np.linalg.norm(F["RPf"]) / F["RPf"].size)
    return niter, err, p["dt_red"] / (1e-10 + np.abs(F["divV"]).max())


# ------------------------------------------------- finite-difference KATs
def _s(a, b):
    """Julia range a:end-b (1-based, inclusive) as a python slice on one axis."""
    return slice(a - 1, None if b == 0 else -b)


def fd3(name, A):
    """Value of @name(A) over its natural 3-D result range, as the broadcast
    formulas of test/test_FiniteDifferences3D.jl:65-163 compute it."""
    a, i, o = _s(1, 0), _s(2, 1), None  # 1:end, 2:end-1
    lo, hi = _s(1, 1), _s(2, 0)  # 1:end-1, 2:end
    l2, h2 = _s(1, 2), _s(3, 0)  # 1:end-2, 3:end
    A = np.asarray(A)
    if name == "d_xa": return A[hi, a, a] - A[lo, a, a]
    if name == "d_ya": return A[a, hi, a] - A[a, lo, a]
    if name == "d_za": return A[a, a, hi] - A[a, a, lo]
    if name == "d_xi": return A[hi, i, i] - A[lo, i, i]
    if name == "d_yi": return A[i, hi, i] - A[i, lo, i]
    if name == "d_zi": return A[i, i, hi] - A[i, i, lo]
    if name == "d2_xi": return (A[h2, i, i] - A[i, i, i]) - (A[i, i, i] - A[l2, i, i])
    if name == "d2_yi": return (A[i, h2, i] - A[i, i, i]) - (A[i, i, i] - A[i, l2, i])
    if name == "d2_zi": return (A[i, i, h2] - A[i, i, i]) - (A[i, i, i] - A[i, i, l2])
    if name == "all": return A.copy()
    if name == "inn": return A[i, i, i]
    if name == "inn_x": return A[i, a, a]
    if name == "inn_y": return A[a, i, a]
    if name == "inn_z": return A[a, a, i]
    if name == "inn_xy": return A[i, i, a]
    if name == "inn_xz": return A[i, a, i]
    if name == "inn_yz": return A[a, i, i]
    if name == "av":
        return (((((((A[lo, lo, lo] + A[hi, lo, lo]) + A[lo, hi, lo]) + A[hi, hi, lo]) + A[lo, lo, hi]) + A[hi, lo, hi])
                 + A[lo, hi, hi]) + A[hi, hi, hi]) * 0.125
    if name == "av_xa": return (A[hi, a, a] + A[lo, a, a]) * 0.5
    if name == "av_ya": return (A[a, hi, a] + A[a, lo, a]) * 0.5
    if name == "av_za": return (A[a, a, hi] + A[a, a, lo]) * 0.5
    if name == "av_xi": return (A[hi, i, i] + A[lo, i, i]) * 0.5
    if name == "av_yi": return (A[i, hi, i] + A[i, lo, i]) * 0.5
    if name == "av_zi": return (A[i, i, hi] + A[i, i, lo]) * 0.5
    if name == "av_xya": return (((A[lo, lo, a] + A[hi, lo, a]) + A[lo, hi, a]) + A[hi, hi, a]) * 0.25
    if name == "av_xza": return (((A[lo, a, lo] + A[hi, a, lo]) + A[lo, a, hi]) + A[hi, a, hi]) * 0.25
    if name == "av_yza": return (((A[a, lo, lo] + A[a, hi, lo]) + A[a, lo, hi]) + A[a, hi, hi]) * 0.25
    if name == "av_xyi": return (((A[lo, lo, i] + A[hi, lo, i]) + A[lo, hi, i]) + A[hi, hi, i]) * 0.25
    if name == "av_xzi": return (((A[lo, i, lo] + A[hi, i, lo]) + A[lo, i, hi]) + A[hi, i, hi]) * 0.25
    if name == "av_yzi": return (((A[i, lo, lo] + A[i, hi, lo]) + A[i, lo, hi]) + A[i, hi, hi]) * 0.25
    inv = lambda x: 1.0 / x  # noqa: E731
    if name == "harm":
        s = inv(A[lo, lo, lo]) + inv(A[hi, lo, lo]) + inv(A[lo, hi, lo]) + inv(A[hi, hi, lo]) + inv(A[lo, lo, hi]) \
            + inv(A[hi, lo, hi]) + inv(A[lo, hi, hi]) + inv(A[hi, hi, hi])
        return 8.0 / s
    if name == "harm_xa": return 2.0 / (inv(A[hi, a, a]) + inv(A[lo, a, a]))
    if name == "harm_ya": return 2.0 / (inv(A[a, hi, a]) + inv(A[a, lo, a]))
    if name == "harm_za": return 2.0 / (inv(A[a, a, hi]) + inv(A[a, a, lo]))
    if name == "harm_xi": return 2.0 / (inv(A[hi, i, i]) + inv(A[lo, i, i]))
    if name == "harm_yi": return 2.0 / (inv(A[i, hi, i]) + inv(A[i, lo, i]))
    if name == "harm_zi": return 2.0 / (inv(A[i, i, hi]) + inv(A[i, i, lo]))
    if name == "harm_xya": return 4.0 / (inv(A[lo, lo, a]) + inv(A[hi, lo, a]) + inv(A[lo, hi, a]) + inv(A[hi, hi, a]))
    if name == "harm_xza": return 4.0 / (inv(A[lo, a, lo]) + inv(A[hi, a, lo]) + inv(A[lo, a, hi]) + inv(A[hi, a, hi]))
    if name == "harm_yza": return 4.0 / (inv(A[a, lo, lo]) + inv(A[a, hi, lo]) + inv(A[a, lo, hi]) + inv(A[a, hi, hi]))
    if name == "harm_xyi": return 4.0 / (inv(A[lo, lo, i]) + inv(A[hi, lo, i]) + inv(A[lo, hi, i]) + inv(A[hi, hi, i]))
    if name == "harm_xzi": return 4.0 / (inv(A[lo, i, lo]) + inv(A[hi, i, lo]) + inv(A[lo, i, hi]) + inv(A[hi, i, hi]))
    if name == "harm_yzi": return 4.0 / (inv(A[i, lo, lo]) + inv(A[i, hi, lo]) + inv(A[i, lo, hi]) + inv(A[i, hi, hi]))
    if name == "maxloc":
        m = np.maximum
        return m(m(m(m(m(m(A[l2, i, i], A[h2, i, i]), A[i, i, i]), A[i, l2, i]), A[i, h2, i]), A[i, i, l2]), A[i, i, h2])
    if name == "minloc":
        m = np.minimum
        return m(m(m(m(m(m(A[l2, i, i], A[h2, i, i]), A[i, i, i]), A[i, l2, i]), A[i, h2, i]), A[i, i, l2]), A[i, i, h2])
    raise KeyError(name)


def fd2(name, A):
    """2-D counterparts (test/test_FiniteDifferences2D.jl:53-138)."""
    a, i = _s(1, 0), _s(2, 1)
    lo, hi, l2, h2 = _s(1, 1), _s(2, 0), _s(1, 2), _s(3, 0)
    A = np.asarray(A)
    inv = lambda x: 1.0 / x  # noqa: E731
    tab = {
        "d_xa": lambda: A[hi, a] - A[lo, a],
        "d_ya": lambda: A[a, hi] - A[a, lo],
        "d_xi": lambda: A[hi, i] - A[lo, i],
        "d_yi": lambda: A[i, hi] - A[i, lo],
        "d2_xa": lambda: (A[h2, a] - A[i, a]) - (A[i, a] - A[l2, a]),
        "d2_ya": lambda: (A[a, h2] - A[a, i]) - (A[a, i] - A[a, l2]),
        "d2_xi": lambda: (A[h2, i] - A[i, i]) - (A[i, i] - A[l2, i]),
        "d2_yi": lambda: (A[i, h2] - A[i, i]) - (A[i, i] - A[i, l2]),
        "all": lambda: A.copy(),
        "inn": lambda: A[i, i],
        "inn_x": lambda: A[i, a],
        "inn_y": lambda: A[a, i],
        "av": lambda: (((A[lo, lo] + A[hi, lo]) + A[lo, hi]) + A[hi, hi]) * 0.25,
        "av_xa": lambda: (A[lo, a] + A[hi, a]) * 0.5,
        "av_ya": lambda: (A[a, lo] + A[a, hi]) * 0.5,
        "av_xi": lambda: (A[lo, i] + A[hi, i]) * 0.5,
        "av_yi": lambda: (A[i, lo] + A[i, hi]) * 0.5,
        "harm": lambda: 4.0 / (inv(A[lo, lo]) + inv(A[hi, lo]) + inv(A[lo, hi]) + inv(A[hi, hi])),
        "harm_xa": lambda: 2.0 / (inv(A[lo, a]) + inv(A[hi, a])),
        "harm_ya": lambda: 2.0 / (inv(A[a, lo]) + inv(A[a, hi])),
        "harm_xi": lambda: 2.0 / (inv(A[lo, i]) + inv(A[hi, i])),
        "harm_yi": lambda: 2.0 / (inv(A[i, lo]) + inv(A[i, hi])),
        "maxloc": lambda: np.maximum(np.maximum(np.maximum(np.maximum(A[l2, i], A[h2, i]), A[i, i]), A[i, l2]), A[i, h2]),
        "minloc": lambda: np.minimum(np.minimum(np.minimum(np.minimum(A[l2, i], A[h2, i]), A[i, i]), A[i, l2]), A[i, h2]),
    }
    return tab[name]()


def fd1(name, A):
    """1-D counterparts (test/test_FiniteDifferences1D.jl:47-97)."""
    A = np.asarray(A)
    tab = {
        "d": lambda: A[1:] - A[:-1],
        "d2": lambda: (A[2:] - A[1:-1]) - (A[1:-1] - A[:-2]),
        "all": lambda: A.copy(),
        "inn": lambda: A[1:-1],
        "av": lambda: (A[:-1] + A[1:]) * 0.5,
        "harm": lambda: 2.0 / (1.0 / A[:-1] + 1.0 / A[1:]),
        "maxloc": lambda: np.maximum(np.maximum(A[:-2], A[2:]), A[1:-1]),
        "minloc": lambda: np.minimum(np.minimum(A[:-2], A[2:]), A[1:-1]),
    }
    return tab[name]()


# Natural argument size of each macro relative to the result size R
# (which field type the reference's KAT feeds it: test_FiniteDifferences3D.jl:38-57).
FD3_ARG_OFFSET = {
    "d_xa": (1, 0, 0), "d_ya": (0, 1, 0), "d_za": (0, 0, 1), "d_xi": (1, 2, 2), "d_yi": (2, 1, 2), "d_zi": (2, 2, 1),
    "d2_xi": (2, 2, 2), "d2_yi": (2, 2, 2), "d2_zi": (2, 2, 2), "all": (0, 0, 0), "inn": (2, 2, 2), "inn_x": (2, 0, 0),
    "inn_y": (0, 2, 0), "inn_z": (0, 0, 2), "inn_xy": (2, 2, 0), "inn_xz": (2, 0, 2), "inn_yz": (0, 2, 2),
    "av": (1, 1, 1), "av_xa": (1, 0, 0), "av_ya": (0, 1, 0), "av_za": (0, 0, 1), "av_xi": (1, 2, 2), "av_yi": (2, 1, 2),
    "av_zi": (2, 2, 1), "av_xya": (1, 1, 0), "av_xza": (1, 0, 1), "av_yza": (0, 1, 1), "av_xyi": (1, 1, 2),
    "av_xzi": (1, 2, 1), "av_yzi": (2, 1, 1), "maxloc": (2, 2, 2), "minloc": (2, 2, 2),
}
for _k in list(FD3_ARG_OFFSET):
    if _k.startswith("av"):
        FD3_ARG_OFFSET["harm" + _k[2:]] = FD3_ARG_OFFSET[_k]

FD2_ARG_OFFSET = {
    "d_xa": (1, 0), "d_ya": (0, 1), "d_xi": (1, 2), "d_yi": (2, 1), "d2_xa": (2, 0), "d2_ya": (0, 2), "d2_xi": (2, 2),
    "d2_yi": (2, 2), "all": (0, 0), "inn": (2, 2), "inn_x": (2, 0), "inn_y": (0, 2), "av": (1, 1), "av_xa": (1, 0),
    "av_ya": (0, 1), "av_xi": (1, 2), "av_yi": (2, 1), "maxloc": (2, 2), "minloc": (2, 2),
}
for _k in list(FD2_ARG_OFFSET):
    if _k.startswith("av"):
        FD2_ARG_OFFSET["harm" + _k[2:]] = FD2_ARG_OFFSET[_k]

FD1_ARG_OFFSET = {"d": 1, "d2": 2, "all": 0, "inn": 2, "av": 1, "harm": 1, "maxloc": 2, "minloc": 2}


# ---------------------------------------------------------------- inputs
def cubed_indices(shape):
    """(ix + (iy-1)*nx + (iz-1)*nx*ny)^3 as Float64, the closed-form input of the
    reference's KATs (test/test_parallel.jl:416, 443-446)."""
    shape = tuple(shape) + (1,) * (3 - len(shape))
    nx, ny, nz = shape
    ix, iy, iz = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1), indexing="ij")
    v = (ix + (iy - 1) * nx + (iz - 1) * nx * ny).astype(np.float64)
    return np.asfortranarray(v ** 3)


def linear_indices(shape):
    shape = tuple(shape) + (1,) * (3 - len(shape))
    nx, ny, nz = shape
    ix, iy, iz = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1), indexing="ij")
    return np.asfortranarray((ix + (iy - 1) * nx + (iz - 1) * nx * ny).astype(np.float64))


def splitmix64_uniform(n, seed):
    """n uniform [0,1) doubles from SplitMix64 (53-bit mantissa), deterministic across hosts."""
    out = np.empty(n, dtype=np.float64)
    x = np.uint64(seed)
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = x + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out[:] = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return out


def diffusion_inputs(n, seed=1, lx=10.0):
    """3-D heat diffusion configuration: cells of size lx/n, T0 = 100 exp(-r^2) about the centre,
    Cp = 1 + U[0,1), dt = min(dx^2,dy^2,dz^2)*cp_min/lam/8.1 (cp_min = 1, lam = 1)."""
    dx = lx / n
    x = (np.arange(1, n + 1) - 0.5) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    T = np.asfortranarray(100.0 * np.exp(-((X - lx / 2) ** 2 + (Y - lx / 2) ** 2 + (Z - lx / 2) ** 2)))
    Cp = np.asfortranarray((1.0 + splitmix64_uniform(n ** 3, seed)).reshape((n, n, n), order="F"))
    dt = min(dx * dx, dx * dx, dx * dx) * 1.0 / 1.0 / 8.1
    return T, Cp, dict(lam=1.0, dt=dt, dx=dx, dy=dx, dz=dx)


# ------------------------------------------------- acoustic 2-D inversion gradient (adjoint)
# Forward: miniapps/acoustic2D.jl:11-20 with a bulk-modulus field K (P -= dt*K*div V), a point
# source added after the pressure update and receivers sampling P^{n+1}. The adjoint has no
# reference counterpart (the reference differentiates with Enzyme, not available here); it is
# pinned in tests/test_oracle.py by the dot-product identity <A x, y> = <x, A^T y> of one step and
# by central finite differences of the misfit along random directions.
def _ac2_fwd_step(P, Vx, Vy, K, dt, rho, dx, dy):
    nx, ny = P.shape
    Vx[1:nx, 1:ny - 1] = Vx[1:nx, 1:ny - 1] - ((dt / rho) * (P[1:nx, 1:ny - 1] - P[0:nx - 1, 1:ny - 1])) / dx
    Vy[1:nx - 1, 1:ny] = Vy[1:nx - 1, 1:ny] - ((dt / rho) * (P[1:nx - 1, 1:ny] - P[1:nx - 1, 0:ny - 1])) / dy
    P[:, :] = P - (dt * K) * (((Vx[1:, :] - Vx[:-1, :]) / dx) + ((Vy[:, 1:] - Vy[:, :-1]) / dy))


def _ac2_adj_step(lP, lVx, lVy, K, dt, rho, dx, dy):
    nx, ny = lP.shape
    W = K * lP
    lVx[1:nx, 1:ny - 1] = lVx[1:nx, 1:ny - 1] + (dt * (W[1:nx, 1:ny - 1] - W[0:nx - 1, 1:ny - 1])) / dx
    lVy[1:nx - 1, 1:ny] = lVy[1:nx - 1, 1:ny] + (dt * (W[1:nx - 1, 1:ny] - W[1:nx - 1, 0:ny - 1])) / dy
    lP[:, :] = lP + (dt / rho) * (((lVx[1:, :] - lVx[:-1, :]) / dx) + ((lVy[:, 1:] - lVy[:, :-1]) / dy))


def acoustic2d_step(P, Vx, Vy, K, dt, rho, dx, dy):
    """One source-free forward step (in place): the linear map A of the adjoint tests."""
    _ac2_fwd_step(P, Vx, Vy, K, dt, rho, dx, dy)


def acoustic2d_step_adjoint(lP, lVx, lVy, K, dt, rho, dx, dy):
    """A^T of acoustic2d_step (in place)."""
    _ac2_adj_step(lP, lVx, lVy, K, dt, rho, dx, dy)


def acoustic2d_traces(K, nt, dt, rho, dx, dy, src, isx, isy, rx0, rdx, ry, nrec, keep=False):
    """Forward run from rest; traces D (nrec x nt) of P^{n+1} at the receivers (1-based grid
    positions rx0 + (r-1) rdx, ry), optionally with every V^{n+1} kept."""
    nx, ny = K.shape
    P, Vx, Vy = np.zeros((nx, ny)), np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1))
    rx = rx0 - 1 + np.arange(nrec) * rdx
    D = np.zeros((nrec, nt))
    Vs = []
    for n in range(nt):
        _ac2_fwd_step(P, Vx, Vy, K, dt, rho, dx, dy)
        P[isx - 1, isy - 1] = P[isx - 1, isy - 1] + src[n]
        D[:, n] = P[rx, ry - 1]
        if keep:
            Vs.append((Vx.copy(), Vy.copy()))
    return (D, Vs) if keep else D


def acoustic2d_misfit_gradient(K, Dobs, nt, dt, rho, dx, dy, src, isx, isy, rx0, rdx, ry, nrec):
    """(J, dJ/dK) for J = 1/2 sum (D - Dobs)^2: every state kept, adjoint swept backwards."""
    nx, ny = K.shape
    D, Vs = acoustic2d_traces(K, nt, dt, rho, dx, dy, src, isx, isy, rx0, rdx, ry, nrec, keep=True)
    r = D - Dobs
    J = 0.5 * float(np.sum(r * r))
    rx = rx0 - 1 + np.arange(nrec) * rdx
    lP, lVx, lVy = np.zeros((nx, ny)), np.zeros((nx + 1, ny)), np.zeros((nx, ny + 1))
    gK = np.zeros((nx, ny))
    for n in reversed(range(nt)):
        lP[rx, ry - 1] = lP[rx, ry - 1] + (D[:, n] - Dobs[:, n])
        Vx1, Vy1 = Vs[n]
        gK[:, :] = gK - (dt * lP) * (((Vx1[1:, :] - Vx1[:-1, :]) / dx) + ((Vy1[:, 1:] - Vy1[:, :-1]) / dy))
        _ac2_adj_step(lP, lVx, lVy, K, dt, rho, dx, dy)
    return J, gK


# ------------------------------------------------- kernel language (@parallel_indices)
# examples/diffusion2D_shmem_novis.jl:11-27: the shared-memory tile T_l holds exactly the values of
# T at the block's cells and halo, so on the CPU (one thread per block, kernel_language.jl:622-628)
# and on the GPU the update is the 5-point flux form below, in the example's operator order, on the
# cells 1 < ix < nx, 1 < iy < ny; T2 keeps its other values.
def diffusion2d_shmem_step(T2, T, Ci, lam, dt, _dx, _dy):
    c = T[1:-1, 1:-1]
    qxp = (-lam * (T[2:, 1:-1] - c)) * _dx
    qxm = (-lam * (c - T[:-2, 1:-1])) * _dx
    qyp = (-lam * (T[1:-1, 2:] - c)) * _dy
    qym = (-lam * (c - T[1:-1, :-2])) * _dy
    T2[1:-1, 1:-1] = c + dt * (Ci[1:-1, 1:-1] * ((-(qxp - qxm)) * _dx - (qyp - qym) * _dy))


def diffusion2d_shmem_cell(T, Ci, lam, dt, _dx, _dy, ix, iy):
    """One cell of the same update in plain Python floats (0-based ix, iy): the brute-force check."""
    t = lambda i, j: float(T[i, j])
    a = ((-lam * (t(ix + 1, iy) - t(ix, iy))) * _dx) - ((-lam * (t(ix, iy) - t(ix - 1, iy))) * _dx)
    b = ((-lam * (t(ix, iy + 1) - t(ix, iy))) * _dy) - ((-lam * (t(ix, iy) - t(ix, iy - 1))) * _dy)
    return t(ix, iy) + dt * (float(Ci[ix, iy]) * (-a * _dx - b * _dy))


# Wavefront model of the warp-level primitives (kernel_language.jl:123-193 docstrings), for one
# wavefront of `w` lanes holding values v[0..w-1] (lane l = laneid() - 1). Lanes without a valid
# partner keep their own value.
def wave_shfl(v, lane, width):
    """@shfl_sync(mask, v, lane, width): lane-1 of the lane's own width-partition (modulo width)."""
    n = len(v)
    return np.array([v[(l // width) * width + ((lane[l] - 1) % width)] for l in range(n)], dtype=np.asarray(v).dtype)


def wave_shfl_up(v, delta, width):
    n = len(v)
    return np.array([v[l - delta] if (l % width) >= delta else v[l] for l in range(n)], dtype=np.asarray(v).dtype)


def wave_shfl_down(v, delta, width):
    n = len(v)
    return np.array([v[l + delta] if (l % width) + delta < width else v[l] for l in range(n)], dtype=np.asarray(v).dtype)


def wave_shfl_xor(v, lane_mask, width):
    n = len(v)
    out = []
    for l in range(n):
        p = l ^ lane_mask
        out.append(v[p] if (p // width) == (l // width) else v[l])
    return np.array(out, dtype=np.asarray(v).dtype)


def wave_ballot(pred):
    """Bit l set iff lane l's predicate holds (64-bit mask as int64 bits)."""
    m = 0
    for l, p in enumerate(pred):
        if p:
            m |= 1 << l
    return np.array([m], dtype=np.uint64).view(np.int64)[0]


# ------------------------------------------------- CellArrays
# CellArrays.jl (a dependency of the reference, Project.toml; not vendored) stores a CellArray of
# N cells with nc values each in a data array of size (B, nc, ceil(N/B)), B the block length
# (src/ParallelKernel/allocators.jl:449: CellArray{T,N,B}(rand(eltype(T), blocklen, prod(size(T)),
# ceil(Int, prod(dims)/blocklen)), dims)): value j of cell I sits at data[I % B, j, I // B]
# (0-based). Restated here element by element.
def cellarray_data_offset(I, j, B, nc):
    """0-based offset in the column-major data array of value j of cell I."""
    return (I % B) + j * B + (I // B) * B * nc


def cells_to_data(cells, celldims, B):
    """Host cells (shape dims + celldims, column-major) -> the data array (B, nc, nblocks), by loops."""
    cells = np.asarray(cells)
    nd = cells.ndim - len(celldims)
    dims = cells.shape[:nd]
    N = int(np.prod(dims))
    nc = int(np.prod(celldims))
    nb = -(-N // B) if N else 0
    flat = cells.reshape((N, nc), order="F")
    data = np.zeros(B * nc * nb, dtype=cells.dtype)
    for I in range(N):
        for j in range(nc):
            data[cellarray_data_offset(I, j, B, nc)] = flat[I, j]
    return data.reshape((B, nc, nb), order="F")


# examples/memcopyCellArray3D.jl:10-27: T2 = T + Ci cell by cell (every value of the cells).
def copy3d_cells(T, Ci):
    return np.asarray(T) + np.asarray(Ci)


# ------------------------------------------------- derivatives (@parallel ∇=...)
# The reference checks its derivative kernels against Enzyme on a CPU loop restatement of the
# kernel (test/ParallelKernel/test_parallel.jl:236-283: A[ix] += a*B[ix]*100.65, Ā = 1 and
# B̄ = 1 + a*100.65). Enzyme is not available here; derivatives are obtained from primal
# evaluations instead, independently of any derivative formula:
#   * a linear kernel's Jacobian, column by column, from the oracle's own primal step applied
#     to unit vectors (exact up to the step's rounding);
#   * a nonlinear kernel's directional derivative by the complex step (Squire & Trapp 1998):
#     f'(x) v = Im f(x + i h v) / h, free of cancellation, so exact to rounding for tiny h, for
#     restatements written with analytic operations and real-part comparisons.
def linear_jacobian(step, n_in):
    """Dense Jacobian of a linear map given as step(flat_input) -> flat_output."""
    cols = []
    for j in range(n_in):
        e = np.zeros(n_in)
        e[j] = 1.0
        cols.append(np.asarray(step(e), dtype=np.float64).ravel(order="F"))
    return np.stack(cols, axis=1)


COMPLEX_STEP = 1e-200


def complex_step_jvp(f, xs, vs, h=COMPLEX_STEP):
    """Directional derivative of f(*xs) along vs (one tangent per input, None for none)."""
    zs = [np.asarray(x, dtype=np.complex128) + (0 if v is None else 1j * h * np.asarray(v)) for x, v in zip(xs, vs)]
    return np.imag(f(*zs)) / h


def complex_step_vjp(f, xs, wrt, y, h=COMPLEX_STEP):
    """Jacobian-transpose product: for each input index in wrt, the array sum_i y_i df_i/dx_j,
    one column of the Jacobian (one complex evaluation) per element x_j."""
    y = np.asarray(y, dtype=np.float64)
    out = []
    for k in wrt:
        x = np.asarray(xs[k])
        g = np.zeros(x.shape, order="F")
        for j in range(x.size):
            v = np.zeros(x.shape, order="F")
            v.flat[j] = 1.0
            vs = [None] * len(xs)
            vs[k] = v
            g.flat[j] = np.sum(y * complex_step_jvp(f, xs, vs, h))
        out.append(g)
    return out


# ------------------------------------------------- one-process halo update (ImplicitGlobalGrid)
# update_halo! of ImplicitGlobalGrid (a dependency of the reference's multi-XPU examples, e.g.
# examples/diffusion3D_multigpucpu_hidecomm_novis.jl:59-62; not vendored) on one process: in a
# periodic dimension the process is its own neighbour, so the halo cells 1:hw take the cells
# s-ol+1 : s-ol+hw and s-hw+1 : s take ol-hw+1 : ol, dimension after dimension; ol is the
# array's overlap (overlap + size - nx, larger for staggered arrays).
def update_halo_periodic(A, nxyz, overlaps=(2, 2, 2), halowidths=(1, 1, 1), periods=(1, 1, 1)):
    A = np.array(A, order="F", copy=True)
    shape = A.shape + (1,) * (3 - A.ndim)
    B = A.reshape(shape, order="F")
    for d in range(3):
        if not periods[d] or nxyz[d] <= 1 or shape[d] <= 1:
            continue
        s, ol, hw = shape[d], overlaps[d] + shape[d] - nxyz[d], halowidths[d]
        for i in range(hw):
            src_l, dst_l = s - ol + i, i          # 0-based: cell s-ol+1+i -> cell 1+i
            src_r, dst_r = ol - hw + i, s - hw + i  # cell ol-hw+1+i -> cell s-hw+1+i
            sl = [slice(None)] * 3
            sl[d] = dst_l
            sr = [slice(None)] * 3
            sr[d] = src_l
            B[tuple(sl)] = B[tuple(sr)]
            sl[d] = dst_r
            sr[d] = src_r
            B[tuple(sl)] = B[tuple(sr)]
    return B.reshape(A.shape, order="F")
