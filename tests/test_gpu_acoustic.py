"""Acoustic wave kernels (miniapps/acoustic3D.jl:11-21, miniapps/acoustic2D.jl:11-20) on the GPU
against the oracle: bit-exact in FP64, in FP32 and in BFloat16 storage with FP32 arithmetic,
for the plain and the memopt (z-streaming, LDS-tiled) paths; plus the size of the FP32 and
BF16 errors against the FP64 run."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import workloads as W
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

MOD64 = __name__
MOD32 = __name__ + ".f32"
MOD2D = __name__ + ".2d"
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=MOD64)
ps.init_parallel_stencil(ps.AMDGPU, ps.Float32, 3, module=MOD32)
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD2D)


def _rand(shape, seed, scale=1.0):
    n = int(np.prod(shape))
    return np.asfortranarray((scale * (W.splitmix64_uniform(n, seed) - 0.5)).reshape(shape, order="F"))


def _state3(n, seed):
    nx, ny, nz = n
    return (_rand((nx, ny, nz), seed), _rand((nx + 1, ny, nz), seed + 1, 0.1), _rand((nx, ny + 1, nz), seed + 2, 0.1),
            _rand((nx, ny, nz + 1), seed + 3, 0.1))


def _to(a, prec):
    if prec == "f64":
        return np.asfortranarray(a, dtype=np.float64)
    if prec == "f32":
        return np.asfortranarray(a, dtype=np.float32)
    return np.asfortranarray(ps.f32_to_bf16(np.asarray(a, dtype=np.float32)))


ELT = {"f64": ps.Float64, "f32": ps.Float32, "bf16": ps.BFloat16}


@pytest.mark.parametrize("memopt", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
def test_acoustic3d_matches_oracle(prec, memopt):
    n, steps = (21, 18, 17), 8
    dt, rho, k, dx, dy, dz = 0.03, 1.0, 1.0, 0.21, 0.19, 0.23
    host = [_to(a, prec) for a in _state3(n, 7)]
    mod = MOD64 if prec == "f64" else MOD32
    kV, kP = W.define_acoustic3d(mod, memopt)
    dev = [ps.zeros(*a.shape, eltype=ELT[prec]) for a in host]
    for d, h in zip(dev, host):
        d.copy_from(h)
    P, Vx, Vy, Vz = dev
    ref = [h.copy(order="F") for h in host]
    for _ in range(steps):
        kV(Vx, Vy, Vz, P, dt, rho, dx, dy, dz)
        kP(P, Vx, Vy, Vz, dt, k, dx, dy, dz)
        O.acoustic3d_v(ref[1], ref[2], ref[3], ref[0], dt, rho, dx, dy, dz, prec=prec)
        O.acoustic3d_p(ref[0], ref[1], ref[2], ref[3], dt, k, dx, dy, dz, prec=prec)
    for name, d, r in zip(("P", "Vx", "Vy", "Vz"), dev, ref):
        got = d.to_host()
        assert got.dtype == r.dtype
        assert np.array_equal(got, r), f"{prec} memopt={memopt}: {name} differs from the oracle"


@pytest.mark.parametrize("memopt", [False, True])
def test_acoustic2d_matches_oracle(memopt):
    nx, ny, steps = 40, 33, 10
    dt, rho, k, dx, dy = 0.05, 1.0, 1.0, 0.3, 0.27
    P0, Vx0, Vy0 = _rand((nx, ny), 3), _rand((nx + 1, ny), 4, 0.1), _rand((nx, ny + 1), 5, 0.1)
    kV, kP = W.define_acoustic2d(MOD2D, memopt)
    P, Vx, Vy = ps.zeros(nx, ny, eltype=ps.Float64), ps.zeros(nx + 1, ny, eltype=ps.Float64), ps.zeros(nx, ny + 1, eltype=ps.Float64)
    P.copy_from(P0)
    Vx.copy_from(Vx0)
    Vy.copy_from(Vy0)
    rP, rVx, rVy = P0.copy(order="F"), Vx0.copy(order="F"), Vy0.copy(order="F")
    for _ in range(steps):
        kV(Vx, Vy, P, dt, rho, dx, dy)
        kP(P, Vx, Vy, dt, k, dx, dy)
        O.acoustic2d_v(rVx, rVy, rP, dt, rho, dx, dy)
        O.acoustic2d_p(rP, rVx, rVy, dt, k, dx, dy)
    assert np.array_equal(P.to_host(), rP)
    assert np.array_equal(Vx.to_host(), rVx)
    assert np.array_equal(Vy.to_host(), rVy)


def test_acoustic3d_reduced_precision_errors():
    """The configuration's precision study at a small size: FP32 and BF16-storage runs against
    the FP64 run, per field (relative max and L2 errors)."""
    n, steps = 33, 60
    P0, prm = W.acoustic3d_inputs(n)
    r64 = W.acoustic3d_run(MOD64, P0, prm, steps, ps.Float64)
    r32 = W.acoustic3d_run(MOD32, P0, prm, steps, ps.Float32)
    r16 = W.acoustic3d_run(MOD32, P0, prm, steps, ps.BFloat16)
    for name, a64, a32, a16 in zip(("P", "Vx", "Vy", "Vz"), r64, r32, r16):
        m32, l32 = W.field_errors(a32, a64)
        m16, l16 = W.field_errors(a16, a64)
        print(f"{name}: fp32 max {m32:.2e} l2 {l32:.2e}; bf16 max {m16:.2e} l2 {l16:.2e}")
        assert l32 < 1e-5 and m32 < 1e-5
        assert l16 < 5e-2 and m16 < 5e-2
        assert l16 > l32  # BF16 storage is the coarser one
