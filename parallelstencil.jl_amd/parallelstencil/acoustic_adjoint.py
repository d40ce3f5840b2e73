"""Gradient of an acoustic 2-D misfit with respect to the bulk modulus, by a hand-written adjoint.

Forward model (miniapps/acoustic2D.jl:11-20 with a bulk-modulus field K instead of the scalar k):
step n maps (P^n, V^n) to (P^{n+1}, V^{n+1}):

    V^{n+1} = V^n - dt/rho * G P^n                 (compute_V!, inner velocity nodes)
    P^{n+1} = P^n - dt * K o (D V^{n+1}) + s_n e_src  (compute_PK!, then the source)
    d_n     = P^{n+1} at the receivers

with G the staggered gradient (@d_xi, @d_yi on inner nodes) and D the divergence (@d_xa, @d_ya).
The misfit is J(K) = 1/2 sum_n sum_r (d_n,r - dobs_n,r)^2.

Adjoint: for step n, backwards, with lP the adjoint of P^{n+1} and lV the adjoint of V^{n+1}:

    lP += d_n - dobs_n at the receivers                 (inject_residual!)
    gK -= dt * lP o (D V^{n+1})                          (grad_K!)
    lV(inner) += dt * G (K o lP)      [= -dt D^T (K o lP)] (adjoint_V!, K o lP on the fly)
    lP += dt/rho * D lV               [= -dt/rho G^T lV]   (adjoint_P!)

The transposes follow from D^T = -G on the inner velocity nodes and G^T = -D on pressure cells
(velocity nodes outside the inner range are constant, so their adjoint stays zero). Every
kernel is an FD-macro kernel of the same codegen as the forward ones.

Time loop: sqrt(N) checkpointing. The forward pass stores (P, Vx, Vy) every `seg` steps; the
backward pass restores each checkpoint, recomputes its segment storing V^{n+1} of each step,
and runs the adjoint steps of the segment in reverse order. Cost: two forward passes and one
adjoint pass; memory: N/seg checkpoints plus seg velocity snapshots.

The reference differentiates @parallel kernels with Enzyme (src/ParallelKernel/parallel.jl:265-301),
which is not available here; `rrule` mirrors the shape of a ChainRulesCore rrule.
"""
import math

import numpy as np

from . import core

FWD_V = """
@parallel function compute_V!(Vx::Data.Array, Vy::Data.Array, P::Data.Array, dt::Data.Number, ρ::Data.Number, dx::Data.Number, dy::Data.Number)
    @inn(Vx) = @inn(Vx) - dt/ρ*@d_xi(P)/dx
    @inn(Vy) = @inn(Vy) - dt/ρ*@d_yi(P)/dy
    return
end
"""
FWD_P = """
@parallel function compute_PK!(P::Data.Array, Vx::Data.Array, Vy::Data.Array, K::Data.Array, dt::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(P) = @all(P) - dt*@all(K)*(@d_xa(Vx)/dx + @d_ya(Vy)/dy)
    return
end
"""
INJECT = """
@parallel_indices (i) function inject!(P::Data.Array, s::Data.Number, isx::Int, isy::Int)
    P[isx + i - 1, isy] = P[isx + i - 1, isy] + s
    return
end
"""
RECORD = """
@parallel_indices (ir) function record!(D::Data.Array, P::Data.Array, rx0::Int, rdx::Int, ry::Int, it::Int)
    D[ir, it] = P[rx0 + (ir - 1)*rdx, ry]
    return
end
"""
INJECT_RES = """
@parallel_indices (ir) function inject_residual!(λP::Data.Array, D::Data.Array, Dobs::Data.Array, rx0::Int, rdx::Int, ry::Int, it::Int)
    λP[rx0 + (ir - 1)*rdx, ry] = λP[rx0 + (ir - 1)*rdx, ry] + (D[ir, it] - Dobs[ir, it])
    return
end
"""
GRAD_K = """
@parallel function grad_K!(gK::Data.Array, λP::Data.Array, Vx::Data.Array, Vy::Data.Array, dt::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(gK) = @all(gK) - dt*@all(λP)*(@d_xa(Vx)/dx + @d_ya(Vy)/dy)
    return
end
"""
ADJ_V = """
@parallel function adjoint_V!(λVx::Data.Array, λVy::Data.Array, λP::Data.Array, K::Data.Array, dt::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(W)   = @all(K)*@all(λP)
    @inn(λVx) = @inn(λVx) + dt*@d_xi(W)/dx
    @inn(λVy) = @inn(λVy) + dt*@d_yi(W)/dy
    return
end
"""
ADJ_P = """
@parallel function adjoint_P!(λP::Data.Array, λVx::Data.Array, λVy::Data.Array, dt::Data.Number, ρ::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(λP) = @all(λP) + dt/ρ*(@d_xa(λVx)/dx + @d_ya(λVy)/dy)
    return
end
"""


def ricker(nt, dt, f0):
    """Ricker wavelet samples s_n = (1 - 2 (pi f0 tau)^2) exp(-(pi f0 tau)^2), tau = (n+1) dt - 1.5/f0."""
    t = (np.arange(nt) + 1) * dt - 1.5 / f0
    a = (math.pi * f0 * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


class Setup:
    """Geometry of the inversion experiment on an nx x ny pressure grid (lx = ly = 40 as the
    miniapp): K = 1 + amp * exp(-r^2 / (2 sigma^2)) (sigma in cells), a Ricker source with `ppw`
    cells per wavelength at a random interior point, `nrec` receivers on a horizontal line,
    dt = min(dx,dy)/sqrt(max K/rho)/4.1. A wave crosses about nt/4.1 cells during the run, so the
    receiver line lies `reach` = 0.6 nt/4.1 cells from the source (at most ny/3), spans 2 reach
    cells, and the blob centre is drawn between the two (random from `seed`)."""

    def __init__(self, nx, ny, nt, seed=1, nrec=64, sigma=100.0, amp=0.1, ppw=10.0, lx=40.0, ly=40.0, rho=1.0):
        u = _uniform(8, seed)
        self.nx, self.ny, self.nt, self.rho = nx, ny, nt, rho
        self.dx, self.dy = lx / (nx - 1), ly / (ny - 1)
        reach = max(4, min(int(0.6 * nt / 4.1), ny // 3))
        self.isx = int(nx // 2 + (u[2] - 0.5) * (nx // 4))
        self.isy = int(ny // 3 + (u[3] - 0.5) * (ny // 8))
        self.nrec = nrec
        self.rdx = max(1, min(2 * reach // nrec, (nx - 4) // nrec))
        self.rx0 = min(max(2, self.isx - (nrec - 1) * self.rdx // 2), nx - 1 - (nrec - 1) * self.rdx)
        self.ry = min(self.isy + reach, ny - 2)
        assert 2 <= self.rx0 and self.rx0 + (nrec - 1) * self.rdx <= nx - 1
        cx, cy = self.isx + (u[0] - 0.5) * reach, self.isy + (0.3 + 0.4 * u[1]) * reach
        ix = np.arange(1, nx + 1)[:, None]
        iy = np.arange(1, ny + 1)[None, :]
        self.K = np.asfortranarray(1.0 + amp * np.exp(-((ix - cx) ** 2 + (iy - cy) ** 2) / (2.0 * sigma * sigma)))
        self.K0 = np.asfortranarray(np.ones((nx, ny)))
        self.dt = min(self.dx, self.dy) / math.sqrt(float(self.K.max()) / rho) / 4.1
        f0 = 1.0 / (ppw * self.dx)  # wave speed 1 (K = rho = 1 background)
        self.src = self.dt * ricker(nt, self.dt, f0)


def _uniform(n, seed):
    from .workloads import splitmix64_uniform
    return splitmix64_uniform(n, seed)


class Acoustic2DInversion:
    """Forward model, misfit and adjoint gradient on the GPU (numbertype of `module`)."""

    def __init__(self, module, setup, seg=None):
        s = self.s = setup
        self.module = module
        d = lambda text, kind="parallel": core._define(kind, text, {}, module)
        self.k_v, self.k_p = d(FWD_V), d(FWD_P)
        self.k_inj, self.k_rec, self.k_res = d(INJECT, "parallel_indices"), d(RECORD, "parallel_indices"), d(INJECT_RES, "parallel_indices")
        self.k_gk, self.k_av, self.k_ap = d(GRAD_K), d(ADJ_V), d(ADJ_P)
        nx, ny, nt = s.nx, s.ny, s.nt
        elt = core._numbertype_of(module)
        z = lambda *sh: core.zeros(*sh, eltype=elt)
        self.P, self.Vx, self.Vy = z(nx, ny), z(nx + 1, ny), z(nx, ny + 1)
        self.lP, self.lVx, self.lVy = z(nx, ny), z(nx + 1, ny), z(nx, ny + 1)
        self.K, self.gK = z(nx, ny), z(nx, ny)
        self.D, self.Dobs = z(s.nrec, nt), z(s.nrec, nt)
        self.seg = seg or max(1, int(math.ceil(math.sqrt(nt))))
        nck = (nt + self.seg - 1) // self.seg
        self.ck = [(z(nx, ny), z(nx + 1, ny), z(nx, ny + 1)) for _ in range(nck)]
        self.snap = [(z(nx + 1, ny), z(nx, ny + 1)) for _ in range(self.seg)]
        # numbertype scalars
        self._t = np.float32 if core._numbertype_of(module) == core.Float32 else np.float64

    # -------------------------------------------------------------- forward
    def _step(self, n, record=True):
        s, t = self.s, self._t
        self.k_v(self.Vx, self.Vy, self.P, t(s.dt), t(s.rho), t(s.dx), t(s.dy), async_=True)
        self.k_p(self.P, self.Vx, self.Vy, self.K, t(s.dt), t(s.dx), t(s.dy), async_=True)
        self.k_inj(self.P, t(s.src[n]), s.isx, s.isy, ranges=((1, 1),), async_=True)
        if record:
            self.k_rec(self.D, self.P, s.rx0, s.rdx, s.ry, n + 1, ranges=((1, s.nrec),), async_=True)

    def _reset_state(self):
        for A in (self.P, self.Vx, self.Vy):
            A.fill_(0.0)

    def forward(self, K, checkpoint=False):
        """Runs the forward model for K (host array); returns the recorded traces (nrec x nt)."""
        self.K.copy_from(np.asarray(K, dtype=self._t, order="F"))
        self._reset_state()
        for n in range(self.s.nt):
            if checkpoint and n % self.seg == 0:
                for A, B in zip(self.ck[n // self.seg], (self.P, self.Vx, self.Vy)):
                    A.copy_(B, async_=True)
            self._step(n)
        core.synchronize()
        return self.D.to_host().astype(np.float64)

    def set_observed(self, Dobs):
        self.Dobs.copy_from(np.asarray(Dobs, dtype=self._t, order="F"))
        self._dobs = np.asarray(Dobs, dtype=np.float64)

    def misfit(self, K):
        D = self.forward(K)
        r = D - self._dobs
        return 0.5 * float(np.sum(r * r))

    # -------------------------------------------------------------- adjoint
    def gradient(self, K):
        """(J, dJ/dK) by the adjoint with sqrt(N) checkpointing."""
        s, t = self.s, self._t
        D = self.forward(K, checkpoint=True)
        r = D - self._dobs
        J = 0.5 * float(np.sum(r * r))
        for A in (self.lP, self.lVx, self.lVy, self.gK):
            A.fill_(0.0)
        nt, seg = s.nt, self.seg
        for c in reversed(range(len(self.ck))):
            n0, n1 = c * seg, min(nt, (c + 1) * seg)
            for A, B in zip((self.P, self.Vx, self.Vy), self.ck[c]):
                A.copy_(B, async_=True)
            for n in range(n0, n1):  # recompute the segment, keep V^{n+1}
                self._step(n, record=False)
                self.snap[n - n0][0].copy_(self.Vx, async_=True)
                self.snap[n - n0][1].copy_(self.Vy, async_=True)
            for n in reversed(range(n0, n1)):
                Vx1, Vy1 = self.snap[n - n0]
                self.k_res(self.lP, self.D, self.Dobs, s.rx0, s.rdx, s.ry, n + 1, ranges=((1, s.nrec),), async_=True)
                self.k_gk(self.gK, self.lP, Vx1, Vy1, t(s.dt), t(s.dx), t(s.dy), async_=True)
                self.k_av(self.lVx, self.lVy, self.lP, self.K, t(s.dt), t(s.dx), t(s.dy), async_=True)
                self.k_ap(self.lP, self.lVx, self.lVy, t(s.dt), t(s.rho), t(s.dx), t(s.dy), async_=True)
        core.synchronize()
        return J, self.gK.to_host().astype(np.float64)

    def rrule(self, K):
        """(J, pullback) with pullback(dJ) = (dJ * dJ/dK,), the shape of a ChainRulesCore rrule."""
        J, g = self.gradient(K)
        return J, (lambda dJ: (dJ * g,))

    # -------------------------------------------------------------- one linear step (tests)
    def step_linear(self, P, Vx, Vy, K):
        """A x: one forward step without source, for host arrays (P, Vx, Vy); returns host arrays."""
        s, t = self.s, self._t
        self.K.copy_from(np.asarray(K, dtype=self._t, order="F"))
        for A, h in zip((self.P, self.Vx, self.Vy), (P, Vx, Vy)):
            A.copy_from(np.asarray(h, dtype=self._t, order="F"))
        self.k_v(self.Vx, self.Vy, self.P, t(s.dt), t(s.rho), t(s.dx), t(s.dy))
        self.k_p(self.P, self.Vx, self.Vy, self.K, t(s.dt), t(s.dx), t(s.dy))
        return tuple(A.to_host().astype(np.float64) for A in (self.P, self.Vx, self.Vy))

    def step_adjoint(self, lP, lVx, lVy, K):
        """A^T y: the adjoint of step_linear, for host arrays; returns host arrays."""
        s, t = self.s, self._t
        self.K.copy_from(np.asarray(K, dtype=self._t, order="F"))
        for A, h in zip((self.lP, self.lVx, self.lVy), (lP, lVx, lVy)):
            A.copy_from(np.asarray(h, dtype=self._t, order="F"))
        self.k_av(self.lVx, self.lVy, self.lP, self.K, t(s.dt), t(s.dx), t(s.dy))
        self.k_ap(self.lP, self.lVx, self.lVy, t(s.dt), t(s.rho), t(s.dx), t(s.dy))
        return tuple(A.to_host().astype(np.float64) for A in (self.lP, self.lVx, self.lVy))
