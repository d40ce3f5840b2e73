"""HydroMech2D (miniapps/HydroMech2D.jl) on the GPU against the oracle's numpy restatement at
res = 64 (63 x 127): every field after PT iterations, and two whole time steps (PT to the norm
tolerance through graph replays, dt from the GPU maximum of |∇V|). FP64 tolerances relative to
each field's largest magnitude: 1e-10 for iterations (2.2e-11 measured after 300), 1e-9 after
whole time steps (1e-7 for the residual-type fields, which converge towards zero); the effective-pressure switch tanh((Pf - Pt) / 0.01) amplifies last-bit
differences of tanh and pow (GPU against numpy) a hundredfold, and they accumulate with the
iteration count. The L2-norm reduction is checked against numpy."""
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
    for k, v in S.init.items():
        F[k][...] = v
    return F


RESIDUALS = ("Rx", "Ry", "dVxdt", "dVydt", "RPt", "RPf", "divV", "divV_o", "divqD")


def _compare(S, F, tol, what, tol_res=None):
    """tol_res: tolerance of the residual-type fields, which shrink towards zero as the solve
    converges (differences of nearly equal terms)."""
    errs, bad = {}, []
    for f, A in S.fields().items():
        got, want = A.to_host(), F[f]
        errs[f] = np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)
        t = tol_res if (tol_res is not None and f in RESIDUALS) else tol
        if not errs[f] <= t:
            bad.append(f)
    assert not bad, f"{bad} ({what}): " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items())


def test_norm_reduction():
    A = ps.zeros(37, 29)
    h = np.asfortranarray(np.random.default_rng(5).standard_normal((37, 29)))
    A.copy_from(h)
    assert abs(A.reduce("norm") - np.linalg.norm(h)) <= 1e-13 * np.linalg.norm(h)


def test_hydromech_pt_iterations_match_oracle():
    print("hydromech PT iterations res 64", flush=True)
    S = W.HydroMech2D(MOD, 64, graph=False)
    F = _ref(S)
    S.k_old(S.Phi_o, S.divV_o, S.Phi, S.divV)  # the time step's update_old! (Phi_o = Phi, ∇V_o = ∇V)
    F["Phi_o"][...], F["divV_o"][...] = F["Phi"], F["divV"]
    for n in (1, 99, 200):
        S.iterate(n)
        O.hm_iterate(F, S.prm, S.dt0, n)
        ps.synchronize()
        _compare(S, F, 1e-10, f"after {n} more iterations")


def test_hydromech_time_steps_match_oracle():
    print("hydromech time steps res 64", flush=True)
    S = W.HydroMech2D(MOD, 64, graph=True)
    F = _ref(S)
    dt = S.dt0
    for step in range(2):
        niter, err, dt_used = S.time_step()
        r_niter, r_err, r_dt = O.hm_time_step(F, S.prm, dt)
        ps.synchronize()
        assert niter == r_niter and dt_used == dt, (step, niter, r_niter)
        assert abs(err - r_err) <= 1e-6 * r_err and abs(S.dt - r_dt) <= 1e-9 * r_dt, (step, err, r_err, S.dt, r_dt)
        _compare(S, F, 1e-9, f"time step {step}", tol_res=1e-7)
        dt = r_dt
