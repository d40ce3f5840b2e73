"""Stokes2D pseudo-transient solver (miniapps/Stokes2D.jl:11-125) on the GPU against the oracle:
every field bit-exact after a number of PT iterations on a ragged grid with inclusions, the FP64
residual reduction on the GPU against the oracle's, and a solve to the 1e-8 tolerance."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import workloads as W
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD)


def _maxloc3x3(M):
    """3x3 local maximum, clipped at the edges (two 1-D passes, x then y)."""
    Mx = np.maximum(np.maximum(np.concatenate([M[:1], M[:-1]], 0), M), np.concatenate([M[1:], M[-1:]], 0))
    return np.asfortranarray(np.maximum(np.maximum(np.concatenate([Mx[:, :1], Mx[:, :-1]], 1), Mx),
                                        np.concatenate([Mx[:, 1:], Mx[:, -1:]], 1)))


class _Ref:
    """The same iteration on the oracle."""

    def __init__(self, S, Mus, Rog):
        nx, ny = S.nx, S.ny
        z = lambda *s: np.zeros(s, order="F")
        self.S = S
        self.Pt, self.dtPt, self.divV = z(nx, ny), z(nx, ny), z(nx, ny)
        self.Vx, self.Vy = z(nx + 1, ny), z(nx, ny + 1)
        self.txx, self.tyy, self.txy = z(nx, ny), z(nx, ny), z(nx - 1, ny - 1)
        self.Rx, self.Ry = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.dVxdt, self.dVydt = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.dtVx, self.dtVy = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.Mus, self.Rog = np.asfortranarray(Mus), np.asfortranarray(Rog)
        Mts = _maxloc3x3(self.Mus) if S.eta_ts == "maxloc" else self.Mus
        O.stokes2d_timesteps(self.dtVx, self.dtVy, self.dtPt, Mts, S.Vsc, S.Ptsc, S.min_dxy2, S.max_nxy)

    def iterate(self, n):
        S = self.S
        for _ in range(n):
            O.stokes2d_p(self.divV, self.Pt, self.Vx, self.Vy, self.dtPt, S.dx, S.dy)
            O.stokes2d_tau(self.divV, self.txx, self.tyy, self.txy, self.Vx, self.Vy, self.Mus, S.dx, S.dy)
            O.stokes2d_dv(self.Rx, self.Ry, self.dVxdt, self.dVydt, self.Pt, self.Rog, self.txx, self.tyy, self.txy,
                          S.dampX, S.dampY, S.dx, S.dy)
            O.stokes2d_v(self.Vx, self.Vy, self.dVxdt, self.dVydt, self.dtVx, self.dtVy)
            O.bc_y(self.Vx)
            O.bc_x(self.Vy)


FIELDS = ("Pt", "dtPt", "divV", "Vx", "Vy", "txx", "tyy", "txy", "Rx", "Ry", "dVxdt", "dVydt", "dtVx", "dtVy")


@pytest.mark.parametrize("shape,seed,eta_ts", [((33, 29), 11, "av"), ((64, 48), 5, "maxloc")])
def test_stokes2d_iterations_match_oracle(shape, seed, eta_ts):
    nx, ny = shape
    Mus, Rog = W.stokes2d_inclusions(nx, ny, seed)
    S = W.Stokes2D(MOD, nx, ny, eta_ts=eta_ts)
    S.set_material(Mus, Rog)
    S.reset()
    R = _Ref(S, Mus, Rog)
    for n in (1, 37):
        S.iterate(n)
        R.iterate(n)
        ps.synchronize()
        for f in FIELDS:
            assert np.array_equal(getattr(S, f).to_host(), getattr(R, f)), f"{f} differs after iterations (shape {shape})"
    # residual check: FP64 GPU reduction vs the oracle's sequential mean (summation order differs)
    ref = max(O.mean_abs(R.Rx), O.mean_abs(R.Ry), O.mean_abs(R.divV))
    got = S.error()
    assert abs(got - ref) <= 1e-12 * abs(ref)


def test_stokes2d_graph_replays_match_oracle():
    """Each block of iterations replayed from one HIP graph, with eager launches, reductions and
    synchronisations in between, equals the oracle after every replay."""
    nx, ny = 47, 39
    Mus, Rog = W.stokes2d_inclusions(nx, ny, 9)
    S = W.Stokes2D(MOD, nx, ny, nout=25, graph=True)
    S.set_material(Mus, Rog)
    S.reset()
    R = _Ref(S, Mus, Rog)
    for rep in range(4):
        S.iterate_block()
        R.iterate(S.nout)
        S.error()  # eager work between replays
        ps.synchronize()
        for f in FIELDS:
            assert np.array_equal(getattr(S, f).to_host(), getattr(R, f)), f"{f} differs after replay {rep}"
        if rep == 1:  # eager iterations between replays
            S.iterate(3)
            R.iterate(3)


def test_stokes2d_solves_to_tolerance():
    nx = ny = 63
    Mus, Rog = W.stokes2d_inclusions(nx, ny, 3, log10_contrast=1.0)
    S = W.Stokes2D(MOD, nx, ny, eps=1e-8, nout=200, itermax=200000)
    S.set_material(Mus, Rog)
    S.reset()
    it, err = S.solve()
    print(f"Stokes2D {nx}x{ny}: {it} PT iterations, err {err:.3e}")
    assert err <= 1e-8
    # the converged state is a fixed point: more iterations keep the residual below tolerance
    S.iterate(200)
    assert S.error() <= 1e-8
