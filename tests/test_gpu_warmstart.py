"""The learned Stokes warm start end to end on the GPU: solver-generated samples, U-Net training
with the library's kernels, then held-out samples solved from the prediction. The warm-started
iteration must reach the same residual tolerance, and the prediction must be closer to the
converged solution than the plain iteration's start from rest (what the network gains)."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import unet as U
from parallelstencil import workloads as W

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD)


def test_warm_start_gains_and_converges():
    n = 63
    solver = W.Stokes2D(MOD, n, n, eps=1e-8, nout=200)
    data, stats = U.generate_samples(solver, range(1, 81), log10_contrast=0.5)
    data = [d for d, s in zip(data, stats) if s[2] <= solver.eps]
    assert len(data) >= 60
    ws = U.WarmStart(n, n, base=16, seed=0)
    X, T, Mk = U.dataset(ws, data)
    ws.set_normalisation(T, Mk)
    hist = ws.train(X, T, Mk, epochs=25, batch=8, lr=3e-3, cosine=True)
    assert hist[-1] < 0.5 * hist[1]
    held = range(5001, 5007)
    ev = U.evaluate(ws, solver, held, log10_contrast=0.5)
    gains = []
    for e, s in zip(ev, held):
        if not e["err_cold"] <= solver.eps:
            continue
        assert e["err_warm"] <= solver.eps  # the plain iteration still decides the answer
        Mus, Rog = W.stokes2d_inclusions(n, n, s, log10_contrast=0.5)
        solver.set_material(Mus, Rog)
        solver.reset()
        solver.solve()
        ref = [solver.Vx.to_host(), solver.Vy.to_host()]
        pred = ws.predict(Mus, Rog)[:2]
        err = np.sqrt(sum(np.sum((p - r) ** 2) for p, r in zip(pred, ref)) / sum(np.sum(r ** 2) for r in ref))
        gains.append(err)
        print(f"seed {s}: velocity error of the prediction {err:.3f} (start from rest: 1.0); "
              f"PT iterations cold {e['cold_iters']} warm {e['warm_iters']}")
    assert len(gains) >= 4
    # typical sample: the prediction beats the start from rest; on average it saves iterations
    assert np.median(gains) < 1.0
    conv = [e for e in ev if e["err_cold"] <= solver.eps]
    assert np.mean([e["warm_iters"] for e in conv]) <= np.mean([e["cold_iters"] for e in conv])
