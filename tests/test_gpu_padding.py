"""padding=true on the GPU: the reference's padded-Field KATs (test/test_parallel.jl:1196-1247) and
the miniapps' staggered kernels on padded Fields from @allocate, bit for bit against the oracle on
dense arrays. On padded Fields the kernels address the parents, so the staggered reads shift by the
parents' size differences (size(B.parent, d) > size(A.parent, d)): the acoustic velocity update
reads P one cell to the left of the face's parent index, and the Stokes residuals mix Field,
XField, YField, XYField, BXField and BYField parents."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import FieldAllocators as FA
from parallelstencil import workloads as W
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, padding=True)
MOD2 = __name__ + ".2d"
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, padding=True, module=MOD2)


def test_write_indices_all_and_inn():
    # test/test_parallel.jl:1202-1217
    A = FA.Field((4, 5, 6))
    assert A.shape == (4, 5, 6) and A.parent.shape == (4, 5, 6)
    k = ps.parallel("""
    @parallel function write_indices!(A)
        @all(A) = $ix + ($iy-1)*size(A,1) + ($iz-1)*size(A,1)*size(A,2);
        return
    end""")
    k(A)
    assert np.array_equal(A.to_host(), O.linear_indices((4, 5, 6)))
    B = FA.Field((4, 5, 6))
    k = ps.parallel("""
    @parallel function write_indices!(A)
        @inn(A) = $ixi + ($iyi-1)*size(A,1) + ($izi-1)*size(A,1)*size(A,2);
        return
    end""")
    k(B)
    got, ref = B.to_host(), O.linear_indices((4, 5, 6))
    assert np.array_equal(got[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1])
    assert np.all(got[0] == 0) and np.all(got[:, :, -1] == 0)


OTF = """
@parallel function diffusion3D_step!(T2, T, Ci, lam::Data.Number, dt::Float64, _dx, _dy, _dz)
    @all(qx)   = -lam*@d_xi(T)*_dx
    @all(qy)   = -lam*@d_yi(T)*_dy
    @all(qz)   = -lam*@d_zi(T)*_dz
    @all(dTdt) = @inn(Ci)*(-@d_xa(qx)*_dx - @d_ya(qy)*_dy - @d_za(qz)*_dz)
    @inn(T2)   = @inn(T) + dt*@all(dTdt)
    return
end
"""


@pytest.mark.parametrize("memopt", [False, True])
def test_onthefly_diffusion_on_fields(memopt):
    # test/test_parallel.jl:1218-1243 (integer-cubed T, unit constants, Ci = 1, exact)
    nxyz = (32, 8, 8)
    T, T2, Ci = FA.Field(nxyz), FA.Field(nxyz), FA.Field(nxyz, ps.ones)
    hT = O.cubed_indices(nxyz)
    T.copy_from(hT)
    k = ps.parallel(OTF, memopt=memopt)
    k(T2, T, Ci, 1.0, 1.0, 1.0, 1.0, 1.0)
    ref = np.asfortranarray(np.zeros(nxyz))
    O.diffusion3d_otf_test(ref, hT, np.asfortranarray(np.ones(nxyz)), 1.0, 1.0, 1.0, 1.0, 1.0)
    assert np.array_equal(T2.to_host(), ref)


def test_acoustic3d_on_padded_fields():
    # miniapps/acoustic3D.jl:11-21 on P = @Field, V = @BVectorField: the parents of Vx..Vz are one
    # cell larger than P's in their own dimension
    n, nt = 20, 6
    dt, rho, k_, dx = 0.05, 1.0, 1.0, 0.5
    f = FA.allocate(gridsize=(n, n, n), fields={"Field": "P", "BVectorField": "V"})
    P, V = f["P"], f["V"]
    assert V["x"].parent.shape == (n + 1, n, n) and V["z"].indices[2].stop == n + 1
    rng = np.random.default_rng(4)
    hP = np.asfortranarray(rng.random((n, n, n)))
    P.copy_from(hP)
    kv, kp = ps.parallel(W.ACOUSTIC3D_V), ps.parallel(W.ACOUSTIC3D_P)
    hV = [np.asfortranarray(np.zeros(s)) for s in ((n + 1, n, n), (n, n + 1, n), (n, n, n + 1))]
    for _ in range(nt):
        kv(V["x"], V["y"], V["z"], P, dt, rho, dx, dx, dx)
        kp(P, V["x"], V["y"], V["z"], dt, k_, dx, dx, dx)
        O.acoustic3d_v(*hV, hP, dt, rho, dx, dx, dx)
        O.acoustic3d_p(hP, *hV, dt, k_, dx, dx, dx)
    assert np.array_equal(P.to_host(), hP)
    for c, h in zip("xyz", hV):
        assert np.array_equal(V[c].to_host(), h), c


def _stokes_oracle_iter(a, Mus, Rog, p):
    O.stokes2d_p(a["divV"], a["Pt"], a["Vx"], a["Vy"], a["dtPt"], p["dx"], p["dy"])
    O.stokes2d_tau(a["divV"], a["txx"], a["tyy"], a["txy"], a["Vx"], a["Vy"], Mus, p["dx"], p["dy"])
    O.stokes2d_dv(a["Rx"], a["Ry"], a["dVxdt"], a["dVydt"], a["Pt"], Rog, a["txx"], a["tyy"], a["txy"], p["dampX"], p["dampY"],
                  p["dx"], p["dy"])
    O.stokes2d_v(a["Vx"], a["Vy"], a["dVxdt"], a["dVydt"], a["dtVx"], a["dtVy"])
    O.bc_y(a["Vx"])
    O.bc_x(a["Vy"])


def test_stokes2d_on_padded_fields():
    # miniapps/Stokes2D.jl:11-55,109-116 with the miniapp's field types from @allocate: every
    # statement reads arrays whose parents differ in size from the written array's
    nx, ny, nit = 33, 27, 37
    dx, dy = 10.0 / (nx - 1), 10.0 / (ny - 1)
    p = dict(dx=dx, dy=dy, dampX=1 - 4.0 / nx, dampY=1 - 4.0 / ny)
    f = _allocate_stokes(nx, ny)
    assert f["txy"].parent.shape == (nx + 1, ny + 1) and f["Rx"].indices[0].start == 2
    rng = np.random.default_rng(8)
    hMus = np.asfortranarray(1.0 + rng.random((nx, ny)))
    hRog = np.asfortranarray(rng.random((nx, ny)))
    f["Mus"].copy_from(hMus)
    f["Rog"].copy_from(hRog)
    K = _stokes_kernels()
    K["ts"](f["dtVx"], f["dtVy"], f["dtPt"], f["Mus"], 1.0, 0.25, min(dx, dy) ** 2, max(nx, ny))
    a = {k: np.asfortranarray(np.zeros(f[k].shape)) for k in ("Pt", "dtPt", "divV", "Vx", "Vy", "txx", "tyy", "txy", "Rx", "Ry",
                                                               "dVxdt", "dVydt", "dtVx", "dtVy")}
    O.stokes2d_timesteps(a["dtVx"], a["dtVy"], a["dtPt"], hMus, 1.0, 0.25, min(dx, dy) ** 2, max(nx, ny))
    for it in range(1, nit + 1):
        K["p"](f["divV"], f["Pt"], f["Vx"], f["Vy"], f["dtPt"], dx, dy)
        K["tau"](f["divV"], f["txx"], f["tyy"], f["txy"], f["Vx"], f["Vy"], f["Mus"], dx, dy)
        K["dv"](f["Rx"], f["Ry"], f["dVxdt"], f["dVydt"], f["Pt"], f["Rog"], f["txx"], f["tyy"], f["txy"], p["dampX"], p["dampY"], dx, dy)
        K["v"](f["Vx"], f["Vy"], f["dVxdt"], f["dVydt"], f["dtVx"], f["dtVy"])
        K["bcy"](f["Vx"], ranges=((1, nx + 1),))
        K["bcx"](f["Vy"], ranges=((1, ny + 1),))
        _stokes_oracle_iter(a, hMus, hRog, p)
        if it in (1, nit):
            for k, h in a.items():
                assert np.array_equal(f[k].to_host(), h), (it, k)


def _allocate_stokes(nx, ny):
    # the miniapp's arrays as FieldAllocators types (Field, BXField, BYField, XYField, XField, YField)
    f = FA.allocate(gridsize=(nx, ny), fields={"Field": ("Pt", "dtPt", "divV", "txx", "tyy", "Mus", "Rog"), "BXField": "Vx",
                                               "BYField": "Vy", "XYField": "txy", "XField": ("Rx", "dVxdt", "dtVx"),
                                               "YField": ("Ry", "dVydt", "dtVy")})
    return f


def _stokes_kernels():
    d = lambda text, kind="parallel": ps.core._define(kind, text, {}, MOD2)  # noqa: E731
    return dict(ts=d(W.STOKES2D_TIMESTEPS), p=d(W.STOKES2D_P), tau=d(W.STOKES2D_TAU), dv=d(W.STOKES2D_DV), v=d(W.STOKES2D_V),
                bcx=d(W.STOKES2D_BC_X, "parallel_indices"), bcy=d(W.STOKES2D_BC_Y, "parallel_indices"))
