"""Stokes2DBatch (S samples as the third dimension, two launches per PT iteration for all of
them) against Stokes2DFused solving each sample alone: every field after n iterations and every
residual bit for bit (both pseudo-time step rules), and the iteration counts and errors of solve
from rest and from warm starts (each sample's fields frozen where its solve ends), with graph
replays and a diverging sample in the batch, and the slot queue that refills finished planes."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import stokes_fused as F
from parallelstencil import workloads as W

pytestmark = pytest.mark.gpu
MOD = __name__ + ".f64"
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD)

FIELDS = ("Pt", "dtPt", "divV", "Vx", "Vy", "txx", "tyy", "txy", "Rx", "Ry", "dVxdt", "dVydt", "dtVx", "dtVy")


def _materials(nx, ny, seeds, contrasts):
    ms = [W.stokes2d_inclusions(nx, ny, s, log10_contrast=c) for s, c in zip(seeds, contrasts)]
    return np.asfortranarray(np.stack([m for m, _ in ms], axis=2)), np.asfortranarray(np.stack([r for _, r in ms], axis=2))


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.uint64)


@pytest.mark.parametrize("eta_ts", ["av", "maxV"])
def test_batch_fields_match_single_sample_solver(eta_ts):
    nx, ny, seeds = 47, 39, [3, 11, 12]
    Mus, Rog = _materials(nx, ny, seeds, [1.0, 1.0, 2.0])
    B = F.Stokes2DBatch(MOD, nx, ny, len(seeds), eta_ts=eta_ts, graph=False)
    B.set_material(Mus, Rog)
    B.reset()
    B.iterate(57)
    B._check()
    ps.synchronize()
    eb = B.errors()
    one = F.Stokes2DFused(MOD, nx, ny, eta_ts=eta_ts, graph=False)
    for s in range(len(seeds)):
        one.set_material(np.asfortranarray(Mus[:, :, s]), np.asfortranarray(Rog[:, :, s]))
        one.reset()
        one.iterate(57)
        one._check()
        ps.synchronize()
        for name in FIELDS:
            a, b = getattr(B, name).to_host()[:, :, s], getattr(one, name).to_host()
            assert np.array_equal(_bits(a), _bits(b)), (eta_ts, name, s)
        assert _bits([eb[s]])[0] == _bits([one.error()])[0], (eta_ts, s)
    B.close()
    one.close()


def test_batch_solve_matches_single_sample_solves():
    # three samples that converge and one of contrast 10^3 (the miniapp's iteration diverges
    # there); graph replays of nout iterations; then warm starts from scaled solutions
    nx = ny = 31
    seeds, contrasts = [1, 2, 5, 9], [1.0, 1.0, 1.0, 3.0]
    Mus, Rog = _materials(nx, ny, seeds, contrasts)
    kw = dict(eps=1e-6, nout=100, itermax=20000)
    B = F.Stokes2DBatch(MOD, nx, ny, len(seeds), **kw)
    one = F.Stokes2DFused(MOD, nx, ny, **kw)
    B.set_material(Mus, Rog)
    B.reset()
    itb, erb = B.solve()
    sols = []
    for s in range(len(seeds)):
        one.set_material(np.asfortranarray(Mus[:, :, s]), np.asfortranarray(Rog[:, :, s]))
        one.reset()
        it, err = one.solve()
        assert itb[s] == it and _bits([erb[s]])[0] == _bits([err])[0], (s, itb[s], it, erb[s], err)
        sols.append((one.Vx.to_host(), one.Vy.to_host(), one.Pt.to_host()))
        for name in ("Vx", "Vy", "Pt"):  # each sample frozen where its solve ended
            assert np.array_equal(_bits(getattr(B, name).to_host()[:, :, s]), _bits(getattr(one, name).to_host())), (name, s)
    assert (itb[:3] < kw["itermax"]).all() and (erb[:3] <= kw["eps"]).all()
    warm = [np.asfortranarray(np.stack([0.9 * np.asarray(sol[f]) for sol in sols], axis=2)) for f in range(3)]
    B.reset(*warm)
    itw, erw = B.solve()
    for s in range(len(seeds)):
        one.set_material(np.asfortranarray(Mus[:, :, s]), np.asfortranarray(Rog[:, :, s]))
        one.reset(*[w[:, :, s] for w in warm])
        it, err = one.solve()
        assert itw[s] == it and _bits([erw[s]])[0] == _bits([err])[0], ("warm", s, itw[s], it)
    assert (itw[:3] < itb[:3]).all()
    B.close()
    one.close()


def test_solve_queue_matches_single_sample_solves():
    # 8 samples through 3 slots: contrasts up to 10^3 (where the miniapp's iteration may diverge),
    # a warm start, a sample with its own iteration limit; each result (iterations, error,
    # fields) bit for bit that of Stokes2DFused alone
    nx, ny = 33, 29
    kw = dict(eps=1e-6, nout=100, itermax=12000)
    seeds, contrasts = [4, 6, 7, 8, 13, 14, 15], [1.0, 3.0, 1.0, 1.0, 2.0, 1.0, 1.0]
    mats = [W.stokes2d_inclusions(nx, ny, s, log10_contrast=c) for s, c in zip(seeds, contrasts)]
    one = F.Stokes2DFused(MOD, nx, ny, **kw)
    one.set_material(*mats[0])
    one.reset()
    one.solve()
    warm = tuple(0.8 * np.asarray(a) for a in (one.Vx.to_host(), one.Vy.to_host(), one.Pt.to_host()))
    items = [(m[0], m[1]) for m in mats]
    items[3] = (mats[3][0], mats[3][1], None, 300)  # its own limit of 300 iterations
    items[5] = (mats[5][0], mats[5][1], None)
    items.append((mats[0][0], mats[0][1], warm))  # a warm start of the first sample
    B = F.Stokes2DBatch(MOD, nx, ny, 3, **kw)
    got = {i: r for i, *r in B.solve_queue(iter(items))}
    assert sorted(got) == list(range(len(items)))
    for i, item in enumerate(items):
        one.set_material(item[0], item[1])
        init = item[2] if len(item) > 2 else None
        one.reset(*init) if init is not None else one.reset()
        itm = one.itermax
        if len(item) > 3:
            one.itermax = item[3]
        it, err = one.solve()
        one.itermax = itm
        git, gerr, Vx, Vy, Pt = got[i]
        assert git == it and _bits([gerr])[0] == _bits([err])[0], (i, git, it, gerr, err)
        for a, b in ((Vx, one.Vx), (Vy, one.Vy), (Pt, one.Pt)):
            assert np.array_equal(_bits(a), _bits(b.to_host())), i
    assert got[3][0] == 300 and got[7][0] < got[0][0]
    B.close()
    one.close()


def test_batch_fields_at_511():
    # the warm start's resolution: two samples (contrast 10^[-2,2], maxV steps) after 40 fused
    # iterations, every field bit for bit the single-sample solver's
    n = 511
    seeds = [100003, 100004]
    Mus, Rog = _materials(n, n, seeds, [2.0, 2.0])
    B = F.Stokes2DBatch(MOD, n, n, 2, eta_ts="maxV", graph=False)
    B.set_material(Mus, Rog)
    B.reset()
    B.iterate(40)
    B._check()
    ps.synchronize()
    eb = B.errors()
    one = F.Stokes2DFused(MOD, n, n, eta_ts="maxV", graph=False)
    for s in range(2):
        one.set_material(np.asfortranarray(Mus[:, :, s]), np.asfortranarray(Rog[:, :, s]))
        one.reset()
        one.iterate(40)
        one._check()
        ps.synchronize()
        for name in ("Pt", "Vx", "Vy", "txy", "Rx", "Ry"):
            assert np.array_equal(_bits(getattr(B, name).to_host()[:, :, s]), _bits(getattr(one, name).to_host())), (name, s)
        assert _bits([eb[s]])[0] == _bits([one.error()])[0]
    B.close()
    one.close()
