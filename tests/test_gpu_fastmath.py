"""fastmath kernels (FMA contraction, division through reciprocals) against the exact oracle:
within 1e-12 relative per field (relative to the field's largest magnitude) after the stated
step counts, the parity bar for FP64 with contraction allowed."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import workloads as W
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=MOD)


def _rel(got, ref):
    return float(np.abs(got - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize("memopt", [False, True])
def test_diffusion_fastmath_within_1e12(memopt):
    n, steps = 64, 100
    T0, Cp, p = W.diffusion_inputs(n, seed=1)
    args = (p["lam"], p["dt"], p["dx"], p["dy"], p["dz"])
    a, b = np.asfortranarray(T0.copy()), np.asfortranarray(T0.copy())
    for _ in range(steps):
        O.diffusion3d_flux(b, a, Cp, *args)
        a, b = b, a
    k = W.define_diffusion3d(memopt, MOD, fastmath=True)
    T, T2, C = ps.zeros(n, n, n), ps.zeros(n, n, n), ps.zeros(n, n, n)
    T.copy_from(T0)
    T2.copy_from(T0)
    C.copy_from(Cp)
    for _ in range(steps):
        k(T2, T, C, *args)
        T, T2 = T2, T
    got = T.to_host()
    rel = _rel(got, a)
    print(f"diffusion fastmath memopt={memopt}: max rel deviation {rel:.2e} after {steps} steps")
    assert rel <= 1e-12
    # the boundary cells are never written: identical
    assert np.array_equal(got[0], a[0]) and np.array_equal(got[:, :, -1], a[:, :, -1])


def test_acoustic3d_fastmath_within_1e12():
    n, steps = 33, 50
    P0, prm = W.acoustic3d_inputs(n)
    kV, kP = W.define_acoustic3d(MOD, False, fastmath=True)
    P, Vx, Vy, Vz = ps.zeros(n, n, n), ps.zeros(n + 1, n, n), ps.zeros(n, n + 1, n), ps.zeros(n, n, n + 1)
    P.copy_from(P0)
    ref = [np.asfortranarray(P0.copy()), np.zeros((n + 1, n, n), order="F"), np.zeros((n, n + 1, n), order="F"),
           np.zeros((n, n, n + 1), order="F")]
    for _ in range(steps):
        kV(Vx, Vy, Vz, P, prm["dt"], prm["rho"], prm["dx"], prm["dy"], prm["dz"])
        kP(P, Vx, Vy, Vz, prm["dt"], prm["k"], prm["dx"], prm["dy"], prm["dz"])
        O.acoustic3d_v(ref[1], ref[2], ref[3], ref[0], prm["dt"], prm["rho"], prm["dx"], prm["dy"], prm["dz"])
        O.acoustic3d_p(ref[0], ref[1], ref[2], ref[3], prm["dt"], prm["k"], prm["dx"], prm["dy"], prm["dz"])
    for A, r in zip((P, Vx, Vy, Vz), ref):
        assert _rel(A.to_host(), r) <= 1e-12
