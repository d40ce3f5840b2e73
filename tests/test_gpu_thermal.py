"""ThermalConvection2D (miniapps/ThermalConvection2D.jl) on the GPU against the oracle's numpy
restatement: every field after PT iterations on a ragged grid, and whole time steps (PT to
tolerance through graph replays, upwind advection, the advective time step) with the oracle's
iteration counts, errors and time steps. FP64 bar: 1e-12 relative to each field's largest
magnitude for iterations; 1e-10 after whole time steps (the error checks compare maxima of
differences of nearly equal fields)."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import workloads as W
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD)


def _ref(S):
    F = {k: np.zeros(v.shape, order="F") for k, v in S.fields().items()}
    F["T"][...] = S.T0
    return F


def _compare(S, F, tol, what):
    for f, A in S.fields().items():
        got, want = A.to_host(), F[f]
        scale = max(np.abs(want).max(), 1e-300)
        err = np.abs(got - want).max() / scale
        assert err <= tol, f"{f}: relative error {err:.3e} ({what})"


def test_thermal2d_pt_iterations_match_oracle():
    nx, ny = 50, 17
    print(f"thermal2d PT iterations {nx}x{ny}", flush=True)
    S = W.ThermalConvection2D(MOD, nx, ny, graph=False)
    F = _ref(S)
    for n in (1, 40, 59):
        S.iterate(n)
        O.thermal2d_pt_iterate(F, S.prm, n)
        ps.synchronize()
        _compare(S, F, 1e-12, f"after {n} more iterations")


def test_thermal2d_time_steps_match_oracle():
    nx, ny = 47, 15
    print(f"thermal2d time steps {nx}x{ny}", flush=True)
    S = W.ThermalConvection2D(MOD, nx, ny, graph=True)
    F = _ref(S)
    for step in range(3):
        niter, errV, errP, dt = S.time_step()
        r_niter, r_errV, r_errP, r_dt = O.thermal2d_time_step(F, S.prm)
        ps.synchronize()
        assert niter == r_niter, (step, niter, r_niter)
        assert abs(errV - r_errV) <= 1e-6 * r_errV and abs(errP - r_errP) <= 1e-6 * r_errP, (step, errV, r_errV, errP, r_errP)
        assert abs(dt - r_dt) <= 1e-12 * r_dt, (step, dt, r_dt)
        _compare(S, F, 1e-10, f"time step {step}")
