"""The slot queue of Stokes2DBatch (solve_queue) on the host: the GPU work is replaced by a
stand-in in which each sample's residual follows a fixed sequence per block of iterations while
its plane is active. Every sample must come out with the iterations and residual that
Stokes2DFused.solve's loop gives for the same sequence: stop at the first check at or below eps,
at a residual that is not below 1e30 (diverged, NaN included), or once its iteration limit is
reached; planes whose sample is done or that hold none must not advance."""
import numpy as np

from parallelstencil.stokes_fused import Stokes2DBatch


class _FakePlane:
    def __init__(self, owner, s):
        self.owner, self.s = owner, s

    def to_host(self):
        return np.full((2, 2, 1), float(self.owner.sample[self.s]))


class FakeBatch(Stokes2DBatch):
    """Stokes2DBatch without a GPU: plane s holds sample `sample[s]`, whose residual after k
    active blocks is traj[sample][k-1] (the last value repeats)."""

    def __init__(self, S, traj, nout=100, itermax=1000, eps=1e-8):
        self.S, self.nout, self.itermax, self.eps = S, nout, itermax, eps
        self.traj = traj
        self.sample = [-1] * S
        self.blocks = [0] * S
        self.active = [False] * S
        self.loads = []

    def load_sample(self, s, Mus, Rog, init=None):
        self.sample[s] = int(Mus)
        self.blocks[s] = 0
        self.loads.append((s, int(Mus), init))

    def set_active(self, s, on):
        self.active[s] = bool(on)

    def iterate_block(self):
        for s in range(self.S):
            if self.active[s]:
                self.blocks[s] += 1

    def errors(self):
        out = []
        for s in range(self.S):
            t = self.traj.get(self.sample[s])
            k = self.blocks[s]
            out.append(np.inf if t is None or k == 0 else t[min(k, len(t)) - 1])
        return np.array(out)

    def _plane(self, name, s):
        return _FakePlane(self, s)


def _single(traj, nout, itermax, eps):
    """Stokes2DFused.solve's loop on one residual sequence."""
    it, err, k = 0, float("inf"), 0
    while it < itermax:
        k += 1
        it += nout
        err = traj[min(k, len(traj)) - 1]
        if err <= eps or not err < 1e30:
            break
    return it, err


def test_solve_queue_bookkeeping():
    rng = np.random.default_rng(3)
    traj = {}
    for i in range(23):
        kind = i % 5
        if kind == 3:  # diverges after a few blocks
            t = [10.0 ** (2 * k) for k in range(1, 30)]
        elif kind == 4:  # a NaN residual
            t = [1e-2, 1e-3, float("nan")]
        else:  # geometric decay at its own rate, some too slow for the limit
            r = rng.uniform(0.2, 2.0)
            t = [10.0 ** (-r * k) for k in range(1, 60)]
        traj[i] = t
    limits = {5: 300, 11: 200}  # samples with their own iteration limit
    items = [(i, None, None, limits.get(i)) for i in range(len(traj))]
    for S in (1, 3, 8, 40):
        fb = FakeBatch(S, traj, nout=100, itermax=1500)
        got = {}
        order = []
        for i, its, err, Vx, Vy, Pt in fb.solve_queue(iter(items)):
            assert i not in got
            got[i] = (its, err)
            order.append(i)
            assert float(Vx[0, 0]) == i  # the fields come from the plane that held sample i
        assert sorted(got) == list(range(len(traj)))
        for i, t in traj.items():
            it, err = _single(t, 100, limits.get(i, 1500), 1e-8)
            assert got[i][0] == it, (S, i, got[i], it)
            assert (got[i][1] == err) or (np.isnan(got[i][1]) and np.isnan(err)), (S, i)
        # samples start in input order, each in the slot that the previous one freed
        assert [m for _, m, _ in fb.loads] == list(range(len(traj)))
        assert not any(fb.active)  # every slot is off once the queue is empty


def test_solve_queue_passes_warm_starts():
    fb = FakeBatch(2, {0: [1e-9], 1: [1e-9], 2: [1e-9]}, nout=50)
    warm = ("Vx", "Vy", "Pt")
    res = list(fb.solve_queue(iter([(0, None), (1, None, warm), (2, None, None)])))
    assert [r[:2] for r in res] == [(0, 50), (1, 50), (2, 50)]
    assert [init for _, _, init in fb.loads] == [None, warm, None]


def test_solve_freezes_each_sample_where_it_ends():
    traj = {0: [1e-1, 1e-5, 1e-9], 1: [1e3, 1e31], 2: [10.0 ** (-0.1 * k) for k in range(1, 200)], 3: [1e-9]}
    fb = FakeBatch(4, traj, nout=100, itermax=900)
    fb.sample = [0, 1, 2, 3]
    fb.active = [True] * 4
    iters, errs = fb.solve()
    for s in range(4):
        it, err = _single(traj[s], 100, 900, 1e-8)
        assert iters[s] == it and errs[s] == err, (s, iters[s], it)
    # a finished sample stops advancing: its block count is where its solve ended
    assert fb.blocks[:2] == [3, 2] and fb.blocks[3] == 1 and fb.blocks[2] == 9


def test_generate_samples_batch_orders_results_by_seed(monkeypatch):
    # generate_samples_batch over the stand-in: results in seed order whatever order the solves
    # finish in, iterations and residuals of the single-sample loop, time shared by iterations
    from parallelstencil import unet as U
    from parallelstencil import workloads as W
    traj = {s: [10.0 ** (-(s % 7 + 1) * k / 3.0) for k in range(1, 40)] for s in range(10, 30)}
    monkeypatch.setattr(W, "stokes2d_inclusions", lambda nx, ny, seed, log10_contrast=1.0: (np.float64(seed), np.float64(-seed)))
    monkeypatch.setattr(U.core, "synchronize", lambda *a, **k: None)  # no device here
    fb = FakeBatch(4, traj, nout=100, itermax=3000)
    fb.nx = fb.ny = 8
    data, stats = U.generate_samples_batch(fb, range(10, 30))
    assert [int(d[0]) for d in data] == list(range(10, 30)) and [int(-d[1]) for d in data] == list(range(10, 30))
    for s, (it, sec, err) in zip(range(10, 30), stats):
        assert (it, err) == _single(traj[s], 100, 3000, 1e-8)
    its = np.array([st[0] for st in stats], float)
    secs = np.array([st[1] for st in stats])
    assert np.allclose(secs / secs.sum(), its / its.sum())
