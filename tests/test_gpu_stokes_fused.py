"""The two-launch Stokes2D iteration (compute_P!+compute_τ!, compute_dV!+compute_V!+free-slip copies)
against the oracle's six statements of miniapps/Stokes2D.jl:111-116, bit for bit, eager and through
graph replays with the in-graph residual check; the 3x3-maximum velocity steps against their
restatement."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil.stokes_fused import Stokes2DFused
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD)
NAMES = ("Pt", "dtPt", "divV", "Vx", "Vy", "txx", "tyy", "txy", "Rx", "Ry", "dVxdt", "dVydt", "dtVx", "dtVy")


def _oracle_state(S, Mus, Rog):
    a = {k: np.asfortranarray(np.zeros(getattr(S, k).shape)) for k in NAMES}
    O.stokes2d_timesteps(a["dtVx"], a["dtVy"], a["dtPt"], Mus, S.Vsc, S.Ptsc, S.min_dxy2, S.max_nxy)
    return a


def _oracle_iter(a, S, Mus, Rog):
    O.stokes2d_p(a["divV"], a["Pt"], a["Vx"], a["Vy"], a["dtPt"], S.dx, S.dy)
    O.stokes2d_tau(a["divV"], a["txx"], a["tyy"], a["txy"], a["Vx"], a["Vy"], Mus, S.dx, S.dy)
    O.stokes2d_dv(a["Rx"], a["Ry"], a["dVxdt"], a["dVydt"], a["Pt"], Rog, a["txx"], a["tyy"], a["txy"], S.dampX, S.dampY, S.dx, S.dy)
    O.stokes2d_v(a["Vx"], a["Vy"], a["dVxdt"], a["dVydt"], a["dtVx"], a["dtVy"])
    O.bc_y(a["Vx"])
    O.bc_x(a["Vy"])


def _material(nx, ny, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(1.0 + 3.0 * rng.random((nx, ny))), np.asfortranarray(rng.random((nx, ny)))


@pytest.mark.parametrize("shape", [(47, 39), (33, 64)])
def test_fused_iterations_bitexact(shape):
    nx, ny = shape
    S = Stokes2DFused(MOD, nx, ny, graph=False)
    Mus, Rog = _material(nx, ny, 3)
    S.set_material(Mus, Rog)
    S.reset()
    a = _oracle_state(S, Mus, Rog)
    # a warm start with arbitrary wall values: the free-slip copies must reproduce bc_x!/bc_y!
    rng = np.random.default_rng(5)
    Vx0, Vy0, Pt0 = rng.random(S.Vx.shape), rng.random(S.Vy.shape), rng.random(S.Pt.shape)
    S.reset(Vx0, Vy0, Pt0)
    for k, v in (("Vx", S.Vx), ("Vy", S.Vy), ("Pt", S.Pt)):
        a[k][:] = v.to_host()
    done = 0
    for n in (1, 36):
        S.iterate(n)
        for _ in range(n):
            _oracle_iter(a, S, Mus, Rog)
        done += n
        for k in NAMES:
            assert np.array_equal(getattr(S, k).to_host(), a[k]), (done, k)


def test_fused_graph_blocks_and_residual():
    nx, ny = 40, 36
    S = Stokes2DFused(MOD, nx, ny, nout=20, graph=True)
    Mus, Rog = _material(nx, ny, 7)
    S.set_material(Mus, Rog)
    S.reset()
    a = _oracle_state(S, Mus, Rog)
    for blk in range(3):  # eager work and reads between the replays
        S.iterate_block()
        for _ in range(S.nout):
            _oracle_iter(a, S, Mus, Rog)
        err = S.error()
        want = max(O.mean_abs(a["Rx"]), O.mean_abs(a["Ry"]), O.mean_abs(a["divV"]))
        assert abs(err - want) <= 1e-12 * max(want, 1e-300), (blk, err, want)
        for k in ("Vx", "Vy", "Pt", "Rx"):
            assert np.array_equal(getattr(S, k).to_host(), a[k]), (blk, k)
    S.close()


def test_maxloc_velocity_steps():
    nx, ny = 30, 26
    S = Stokes2DFused(MOD, nx, ny, graph=False, eta_ts="maxV")
    Mus, Rog = _material(nx, ny, 9)
    S.set_material(Mus, Rog)
    P = np.pad(Mus, 1, mode="edge")
    M = np.asfortranarray(np.max([P[1 + i:1 + i + nx, 1 + j:1 + j + ny] for i in (-1, 0, 1) for j in (-1, 0, 1)], axis=0))
    tVx, tVy, tPt = (np.asfortranarray(np.zeros(s)) for s in ((nx - 1, ny - 2), (nx - 2, ny - 1), (nx, ny)))
    O.stokes2d_timesteps(tVx, tVy, tPt, M, S.Vsc, S.Ptsc, S.min_dxy2, S.max_nxy)
    assert np.array_equal(S.dtVx.to_host(), tVx) and np.array_equal(S.dtVy.to_host(), tVy)
    O.stokes2d_timesteps(tVx, tVy, tPt, Mus, S.Vsc, S.Ptsc, S.min_dxy2, S.max_nxy)
    assert np.array_equal(S.dtPt.to_host(), tPt)
