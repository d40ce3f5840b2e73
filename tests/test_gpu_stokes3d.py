"""Stokes3D pseudo-transient solver (miniapps/Stokes3D.jl:11-147) on the GPU against the oracle's
numpy restatement: every field after a number of PT iterations on a ragged grid with the
miniapp's spherical inclusion, the residual check, graph replays, and a solve to tolerance.
FP64 bar: 1e-12 relative to each field's largest magnitude."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import workloads as W
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=MOD)

FIELDS = ("Pt", "dtPt", "divV", "Vx", "Vy", "Vz", "txx", "tyy", "tzz", "txy", "txz", "tyz", "Rx", "Ry", "Rz",
          "dVxdt", "dVydt", "dVzdt", "dtVx", "dtVy", "dtVz")


def _ref(S, Mus, Rog):
    """The oracle's state for solver S: zero fields, the material and the pseudo-time steps."""
    F = {k: np.zeros(v.shape, order="F") for k, v in S.fields().items()}
    F["Mus"][...], F["Rog"][...] = Mus, Rog
    O.stokes3d_timesteps(F["dtVx"], F["dtVy"], F["dtVz"], F["dtPt"], F["Mus"], S.Vsc, S.Ptsc, S.min_dxyz2, S.max_nxyz)
    prm = dict(dx=S.dx, dy=S.dy, dz=S.dz, dampX=S.dampX, dampY=S.dampY, dampZ=S.dampZ)
    return F, prm


def _compare(S, F, what):
    for f in FIELDS:
        got, want = getattr(S, f).to_host(), F[f]
        scale = max(np.abs(want).max(), 1e-300)
        err = np.abs(got - want).max() / scale
        assert err <= 1e-12, f"{f}: relative error {err:.3e} ({what})"


def test_stokes3d_iterations_match_oracle():
    nx, ny, nz = 24, 22, 20
    Mus, Rog = W.stokes3d_material(nx, ny, nz, musi=0.3)
    print(f"stokes3d iterations {nx}x{ny}x{nz}", flush=True)
    S = W.Stokes3D(MOD, nx, ny, nz, graph=False)
    S.set_material(Mus, Rog)
    S.reset()
    F, prm = _ref(S, Mus, Rog)
    for n in (1, 29, 30):
        S.iterate(n)
        O.stokes3d_iterate(F, prm, n)
        ps.synchronize()
        _compare(S, F, f"after {n} more iterations")
    ref = max(float(np.mean(np.abs(F[k]))) for k in ("Rx", "Ry", "Rz", "divV"))
    got = S.error()
    assert abs(got - ref) <= 1e-12 * abs(ref)


def test_stokes3d_graph_replays_match_oracle():
    """Blocks of iterations replayed from one HIP graph, with eager work between replays."""
    nx, ny, nz = 19, 17, 21
    Mus, Rog = W.stokes3d_material(nx, ny, nz, musi=0.5)
    print(f"stokes3d graph {nx}x{ny}x{nz}", flush=True)
    S = W.Stokes3D(MOD, nx, ny, nz, nout=15, graph=True)
    S.set_material(Mus, Rog)
    S.reset()
    F, prm = _ref(S, Mus, Rog)
    for rep in range(3):
        S.iterate_block()
        O.stokes3d_iterate(F, prm, S.nout)
        S.error()
        ps.synchronize()
        _compare(S, F, f"replay {rep}")
        if rep == 0:
            S.iterate(2)
            O.stokes3d_iterate(F, prm, 2)


def test_stokes3d_solves_to_tolerance():
    """31^3 with an inclusion of viscosity 0.3: the GPU solver stops at the same check as the
    oracle (1400 iterations, error below 1e-6) and its final state is the oracle's."""
    n = 31
    Mus, Rog = W.stokes3d_material(n, n, n, musi=0.3)
    print(f"stokes3d solve {n}^3", flush=True)
    S = W.Stokes3D(MOD, n, n, n, eps=1e-6, nout=100, itermax=5000)
    S.set_material(Mus, Rog)
    S.reset()
    it, err = S.solve()
    F, prm = _ref(S, Mus, Rog)
    it_ref, err_ref = 0, float("inf")
    while it_ref < S.itermax:
        O.stokes3d_iterate(F, prm, S.nout)
        it_ref += S.nout
        err_ref = max(float(np.mean(np.abs(F[k]))) for k in ("Rx", "Ry", "Rz", "divV"))
        if err_ref <= S.eps:
            break
    assert err <= S.eps and it == it_ref == 1400, (it, err, it_ref, err_ref)
    assert abs(err - err_ref) <= 1e-9 * err_ref
    for f in ("Pt", "Vx", "Vy", "Vz"):
        got, want = getattr(S, f).to_host(), F[f]
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max(), f
