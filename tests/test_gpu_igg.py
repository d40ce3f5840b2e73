"""GPU tests of update_halo! on one process (periodic self-exchange) against the oracle, and of the
loop of examples/diffusion3D_multigpucpu_hidecomm_novis.jl (@hide_communication around a
@parallel step and update_halo!) against oracle steps with periodic halos: exact."""
import numpy as np
import pytest

import parallelstencil as ps
from oracle import ps_oracle as O
from parallelstencil import igg
from parallelstencil import workloads as W

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=MOD)


@pytest.fixture(autouse=True)
def _fresh_grid():
    igg._grid = None
    yield
    igg._grid = None


def _dev(h):
    return ps.zeros(*h.shape).copy_from(np.asfortranarray(h))


@pytest.mark.parametrize("case", ["ol2", "ol4", "staggered", "2d", "nonperiodic"])
def test_update_halo_against_oracle(case):
    print(f"igg: update_halo {case}", flush=True)
    nxyz, kw, shape = (33, 20, 12), {}, None
    periods = (1, 1, 1)
    if case == "ol4":
        kw = dict(overlaps=(4, 4, 4), halowidths=(2, 2, 2))
    if case == "2d":
        nxyz = (40, 24, 1)
    if case == "nonperiodic":
        periods = (0, 1, 0)
    igg.init_global_grid(*nxyz, periodx=periods[0], periody=periods[1], periodz=periods[2], **kw)
    shape = (34, 20, 13) if case == "staggered" else (nxyz[:2] if case == "2d" else nxyz)
    rng = np.random.default_rng(7)
    h = np.asfortranarray(rng.standard_normal(shape))
    A = _dev(h)
    igg.update_halo_(A)
    ov = kw.get("overlaps", (2, 2, 2))
    hw = kw.get("halowidths", (1, 1, 1))
    assert np.array_equal(A.to_host(), O.update_halo_periodic(h, nxyz, ov, hw, periods))


def test_hidecomm_example_loop_with_periodic_halos():
    """The example's loop on a periodic grid: every step updates T2 and its halo inside
    @hide_communication; the oracle runs the same step and the restated halo update. The halo
    copy reads the cells 2 and n-1 of every dimension: an @inn kernel writes cell i+1 from thread
    i, so the boundary width must reach thread n-2 (width >= 3) for those cells to be computed
    before the copy starts, while the inner region runs concurrently."""
    n, nt = 48, 6
    igg.init_global_grid(n, n, n, periodx=1, periody=1, periodz=1)
    T0, Cp, p = W.diffusion_inputs(n, seed=3)
    args = (p["lam"], p["dt"], p["dx"], p["dy"], p["dz"])
    k = W.define_diffusion3d(False, MOD)
    T, T2, C = _dev(T0), _dev(T0), _dev(Cp)
    a, b = np.asfortranarray(T0.copy()), np.asfortranarray(T0.copy())
    for it in range(nt):
        ps.hide_communication((16, 4, 4), [(k, (T2, T, C) + args)], lambda: igg.update_halo_(T2))
        T, T2 = T2, T
        O.diffusion3d_flux(b, a, Cp, *args)
        b = O.update_halo_periodic(b, (n, n, n))
        a, b = b, a
    ps.synchronize()
    assert np.array_equal(T.to_host(), a)
    # periodicity holds after the last update
    h = T.to_host()
    assert np.array_equal(h[0], h[-2]) and np.array_equal(h[-1], h[1])


def test_hidecomm_example_kernel_nonperiodic():
    """examples/diffusion3D_multigpucpu_hidecomm_novis.jl as written (non-periodic grid:
    update_halo! has no neighbour): @hide_communication gives the plain loop's fields bit for bit."""
    n, nt = 40, 5
    me, dims, *_ = igg.init_global_grid(n, n, n)
    dx = dy = dz = 1.0 / (igg.nx_g() - 1)
    k = ps.core._define("parallel", """
    @parallel function diffusion3D_step!(T2, T, Ci, lam, dt, _dx, _dy, _dz)
        @inn(T2) = @inn(T) + dt*(lam*@inn(Ci)*(@d2_xi(T)*_dx^2 + @d2_yi(T)*_dy^2 + @d2_zi(T)*_dz^2));
        return
    end""", {}, MOD)
    rng = np.random.default_rng(8)
    T0 = np.asfortranarray(1.7 + rng.random((n, n, n)))
    dt = min(dx ** 2, dy ** 2, dz ** 2) / 1.0 / 0.5 / 8.1
    sc = (1.0, dt, 1 / dx, 1 / dy, 1 / dz)
    out = []
    for hide in (True, False):
        T, T2, Ci = _dev(T0), _dev(T0), ps.fill(0.5, n, n, n)
        for it in range(nt):
            if hide:
                ps.hide_communication((16, 2, 2), [(k, (T2, T, Ci) + sc)], lambda: igg.update_halo_(T2))
            else:
                k(T2, T, Ci, *sc)
                igg.update_halo_(T2)
            T, T2 = T2, T
        out.append(T.to_host())
    assert np.array_equal(out[0], out[1])
    g = np.empty((n, n, n))
    igg.gather_(T, g)
    assert np.array_equal(g, out[0])
    igg.finalize_global_grid()
