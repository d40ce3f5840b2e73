"""Scalar porosity waves (miniapps/scalar_porowaves2D.jl) on the GPU against the oracle's numpy
restatement: every field after 250 time steps (three time-step updates from the GPU maximum of
K_mu) on a ragged grid, FP64 within 1e-12 relative to each field's largest magnitude."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import workloads as W
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD)


def test_porowaves_steps_match_oracle():
    nx, ny = 40, 52
    print(f"porowaves {nx}x{ny}", flush=True)
    S = W.Porowaves2D(MOD, nx, ny)
    F = {k: np.zeros(v.shape, order="F") for k, v in S.fields().items()}
    F["Phi"][...] = S.Phi0
    F["Eta"][...] = S.eta
    it, dt = 0, 1.0
    for n in (1, 120, 129):
        S.step(n)
        it, dt = O.poro_steps(F, S.prm, it, dt, n)
        ps.synchronize()
        assert S.it == it and abs(S.dt - dt) <= 1e-14 * dt, (S.it, it, S.dt, dt)
        for f, A in S.fields().items():
            got, want = A.to_host(), F[f]
            err = np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)
            assert err <= 1e-12, f"{f}: relative error {err:.3e} after {it} steps"

