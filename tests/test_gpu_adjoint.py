"""Hand-written adjoint of the acoustic 2-D time step (transposed-stencil kernels, sqrt(N)
checkpointing) on the GPU: traces and gradient against the FP64 oracle, the dot-product identity
of one step, central finite differences along random directions, independence of the
checkpoint spacing, and the FP32 gradient against the FP64 one."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import acoustic_adjoint as AA
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

MOD64 = __name__
MOD32 = __name__ + ".f32"
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD64)
ps.init_parallel_stencil(ps.AMDGPU, ps.Float32, 2, module=MOD32)


def _oracle_args(s):
    return (s.nt, s.dt, s.rho, s.dx, s.dy, s.src, s.isx, s.isy, s.rx0, s.rdx, s.ry, s.nrec)


def _inv(mod, nx=70, ny=58, nt=120, seg=None, nrec=16):
    s = AA.Setup(nx, ny, nt, seed=3, nrec=nrec, sigma=nx / 5)
    inv = AA.Acoustic2DInversion(mod, s, seg=seg)
    Dobs = O.acoustic2d_traces(s.K0, *_oracle_args(s))
    inv.set_observed(Dobs)
    return s, inv, Dobs


def test_traces_and_gradient_match_oracle():
    s, inv, Dobs = _inv(MOD64)
    D = inv.forward(s.K)
    assert np.array_equal(D, O.acoustic2d_traces(s.K, *_oracle_args(s)))
    J, g = inv.gradient(s.K)
    Jr, gr = O.acoustic2d_misfit_gradient(s.K, Dobs, *_oracle_args(s))
    assert J > 0 and abs(J - Jr) <= 1e-12 * Jr
    assert np.abs(g - gr).max() <= 1e-12 * np.abs(gr).max()


def test_adjoint_step_dot_product():
    s, inv, _ = _inv(MOD64, nx=40, ny=33, nt=4)
    nx, ny = s.nx, s.ny
    rng = np.random.default_rng(4)
    x = [rng.standard_normal((nx, ny)), rng.standard_normal((nx + 1, ny)), rng.standard_normal((nx, ny + 1))]
    y = [rng.standard_normal((nx, ny)), rng.standard_normal((nx + 1, ny)), rng.standard_normal((nx, ny + 1))]
    for a in x[1:] + y[1:]:  # velocity boundary nodes are constant: zero in both vectors
        a[0], a[-1], a[:, 0], a[:, -1] = 0, 0, 0, 0
    Ax = inv.step_linear(*x, s.K)
    ATy = inv.step_adjoint(*y, s.K)
    lhs = sum(float(np.sum(a * b)) for a, b in zip(Ax, y))
    rhs = sum(float(np.sum(a * b)) for a, b in zip(x, ATy))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_gradient_finite_differences():
    s, inv, _ = _inv(MOD64)
    J, g = inv.gradient(s.K)
    rng = np.random.default_rng(8)
    for _ in range(8):
        d = rng.standard_normal(s.K.shape)
        h = 1e-4
        fd = (inv.misfit(s.K + h * d) - inv.misfit(s.K - h * d)) / (2 * h)
        assert abs(fd - float(np.sum(g * d))) <= 1e-6 * abs(fd)


def test_checkpoint_spacing_does_not_change_the_gradient():
    s, inv_a, _ = _inv(MOD64, seg=7)
    _, inv_b, _ = _inv(MOD64, seg=s.nt)
    Ja, ga = inv_a.gradient(s.K)
    Jb, gb = inv_b.gradient(s.K)
    assert Ja == Jb and np.array_equal(ga, gb)


def test_fp32_gradient_against_fp64():
    s, inv64, _ = _inv(MOD64)
    _, inv32, _ = _inv(MOD32)
    _, g64 = inv64.gradient(s.K)
    _, g32 = inv32.gradient(s.K)
    rel = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    print(f"FP32 gradient relative L2 error vs FP64: {rel:.3e}")
    assert rel < 1e-3


def test_hand_written_adjoint_matches_derivative_kernels():
    """The gradient of the hand-written adjoint (transposed stencils, checkpointing) against the
    one obtained by differentiating the forward kernels themselves with @parallel ∇ (the
    reference's route to this gradient, src/ParallelKernel/parallel.jl:265-301), step by step
    backwards through stored states: record!, inject!, compute_PK! and compute_V! in reverse."""
    s, inv, Dobs = _inv(MOD64, nx=48, ny=40, nt=60)
    J, g = inv.gradient(s.K)
    dt, rho, dx, dy = s.dt, s.rho, s.dx, s.dy
    inv.K.copy_from(s.K)
    inv._reset_state()
    states = []
    for n in range(s.nt):  # the state before each step
        states.append(tuple(ps.zeros(*A.shape).copy_(A) for A in (inv.P, inv.Vx, inv.Vy)))
        inv._step(n)
    ps.synchronize()
    D = inv.D.to_host()
    Db = ps.zeros(s.nrec, s.nt).copy_from(np.asfortranarray(D - Dobs))
    lP, lVx, lVy, gK = ps.zeros(s.nx, s.ny), ps.zeros(s.nx + 1, s.ny), ps.zeros(s.nx, s.ny + 1), ps.zeros(s.nx, s.ny)
    P, Vx, Vy = ps.zeros(s.nx, s.ny), ps.zeros(s.nx + 1, s.ny), ps.zeros(s.nx, s.ny + 1)
    for n in reversed(range(s.nt)):
        for A, B in zip((P, Vx, Vy), states[n]):
            A.copy_(B)
        inv.k_v(Vx, Vy, P, dt, rho, dx, dy)  # V^{n+1} (compute_PK! reads it), P^n
        inv.k_rec(inv.D, P, s.rx0, s.rdx, s.ry, n + 1, ranges=((1, s.nrec),), grad=[(inv.D, Db), (P, lP)])
        inv.k_inj(P, s.src[n], s.isx, s.isy, ranges=((1, 1),), grad=[(P, lP)])
        inv.k_p(P, Vx, Vy, inv.K, dt, dx, dy, grad=[(P, lP), (Vx, lVx), (Vy, lVy), (inv.K, gK)])
        inv.k_v(Vx, Vy, P, dt, rho, dx, dy, grad=[(Vx, lVx), (Vy, lVy), (P, lP)])
    ga = gK.to_host()
    assert np.abs(ga).max() > 0
    assert np.abs(ga - g).max() <= 1e-12 * np.abs(g).max()
    assert np.array_equal(Db.to_host(), np.zeros((s.nrec, s.nt)))  # every residual consumed
