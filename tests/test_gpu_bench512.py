"""The exact kernels that bench.py times, at the benched size: triad and 3-D heat diffusion
(plain and memopt) on 512^3 FP64 fields, bit for bit against the oracle after a few steps.

The memopt launch at this size is the one the benchmark runs: 256x4 blocks, LDS plane tiles,
nontemporal stores, loopsize 32 (examples/memcopy3D.jl:10-13,
examples/diffusion3D_novis_noperf.jl:10-13 in the flux form of test/test_parallel.jl:419-426)."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import workloads as W
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=MOD)
N = 512


def test_triad_512():
    rng = np.random.default_rng(11)
    B = np.asfortranarray(rng.random((N, N, N)))
    C = np.asfortranarray(rng.random((N, N, N)))
    k = W.define_triad3d(MOD)
    A, Bd, Cd = ps.zeros(N, N, N), ps.zeros(N, N, N), ps.zeros(N, N, N)
    Bd.copy_from(B)
    Cd.copy_from(C)
    k(A, Bd, Cd)
    got = A.to_host()
    assert np.array_equal(got, B + C)
    del B, C, got


@pytest.mark.parametrize("memopt", [False, True])
def test_diffusion_512_bitexact(memopt):
    nt = 3
    T0, Cp, p = O.diffusion_inputs(N, seed=1)
    args = (p["lam"], p["dt"], p["dx"], p["dy"], p["dz"])
    k = W.define_diffusion3d(memopt, MOD)
    geo = k.geometry(*(ps.ArgSpec((N, N, N)) for _ in range(3)), *args)
    if memopt:
        assert geo.memopt and geo.nthreads == (256, 4, 1) and geo.nblocks == (2, 128, 16), geo
    T, T2, Cd = ps.zeros(N, N, N), ps.zeros(N, N, N), ps.zeros(N, N, N)
    T.copy_from(T0)
    T2.copy_from(T0)
    Cd.copy_from(Cp)
    for _ in range(nt):
        k(T2, T, Cd, *args, async_=True)
        T, T2 = T2, T
    ps.synchronize()
    a, b = np.asfortranarray(T0.copy()), np.asfortranarray(T0.copy())
    for _ in range(nt):
        O.diffusion3d_flux(b, a, Cp, *args)
        a, b = b, a
    got = T.to_host()
    bad = np.argwhere(got != a)
    assert bad.shape[0] == 0, f"{bad.shape[0]} cells differ, first {bad[:4].tolist()}, max |d| = {np.abs(got - a).max()}"


def test_diffusion_512_fastmath_200_steps_within_1e12():
    # the benched fastmath kernel over the 200 steps of the 512^3 configuration: within 1e-12
    # relative (to the field's largest magnitude) of the exact oracle, FMA contraction and
    # reciprocal division allowed; boundary cells untouched
    nt = 200
    T0, Cp, p = O.diffusion_inputs(N, seed=1)
    args = (p["lam"], p["dt"], p["dx"], p["dy"], p["dz"])
    k = W.define_diffusion3d(True, MOD, fastmath=True)
    T, T2, Cd = ps.zeros(N, N, N), ps.zeros(N, N, N), ps.zeros(N, N, N)
    T.copy_from(T0)
    T2.copy_from(T0)
    Cd.copy_from(Cp)
    for _ in range(nt):
        k(T2, T, Cd, *args, async_=True)
        T, T2 = T2, T
    ps.synchronize()
    got = T.to_host()
    del T, T2, Cd
    a, b = np.asfortranarray(T0.copy()), np.asfortranarray(T0.copy())
    for _ in range(nt):
        O.diffusion3d_flux(b, a, Cp, *args)
        a, b = b, a
    rel = float(np.abs(got - a).max() / np.abs(a).max())
    print(f"diffusion 512^3 fastmath memopt, {nt} steps: max rel deviation {rel:.2e}")
    assert rel <= 1e-12
    assert np.array_equal(got[0], a[0]) and np.array_equal(got[:, :, -1], a[:, :, -1])
