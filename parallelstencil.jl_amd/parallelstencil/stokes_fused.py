"""The pseudo-transient Stokes2D solver of miniapps/Stokes2D.jl with its six launches per iteration
fused into two, for the launch-bound sizes of the learned warm start (511^2, 1023^2).

One PT iteration of the miniapp (:111-116) is compute_P!, compute_τ!, compute_dV!, compute_V!,
bc_y!(Vx), bc_x!(Vy). Here:
  * STOKES2D_PTAU = compute_P! + compute_τ!: both write per cell from the velocities, and compute_τ!
    reads ∇V only at its own cell, so one thread computes the divergence once and writes ∇V, Pt,
    τxx, τyy (and τxy at the node it owns);
  * STOKES2D_DVV = compute_dV! + compute_V! + the free-slip copies: compute_V! updates the velocity
    node whose residual the same thread has just formed, and the boundary copies of bc_y!/bc_x!
    move values no other thread of the launch reads.
Each statement keeps the miniapp's operations in the miniapp's order (the FiniteDifferences2D
macros expanded by hand: @d_xa(A) = A[ix+1,iy]-A[ix,iy], @av(A) = (((A[ix,iy]+A[ix+1,iy])+A[ix,iy+1])
+A[ix+1,iy+1])*0.25, ...), so the fields are bit-identical to the six-launch iteration.

Pseudo-time steps: `eta_ts="av"` is compute_timesteps! on the viscosity (:11-16). `eta_ts="maxV"`
feeds the velocity steps dτVx, dτVy the 3x3 local maximum of the viscosity (as the HydroMech2D
miniapp does with @maxloc) and keeps dτPt on the local viscosity; the iteration then converges for
inclusion contrasts where the miniapp's steps diverge (10^±2 and 10^+3 measured).

The residual check (max of mean |Rx|, mean |Ry|, mean |∇V|, :117-123) runs on the GPU through
ps_field_reduce_noalloc inside the graph of each block of `nout` iterations, so a block is one graph
replay followed by one 24-byte copy to the host.
"""
import ctypes
import re

import numpy as np

from . import core
from ._lib import check, lib
from . import workloads as W

STOKES2D_PTAU = """
@parallel_indices (ix, iy) function compute_Pτ!(∇V::Data.Array, Pt::Data.Array, τxx::Data.Array, τyy::Data.Array, τxy::Data.Array, Vx::Data.Array, Vy::Data.Array, dτPt::Data.Array, Mus::Data.Array, dx::Data.Number, dy::Data.Number)
    if (ix <= size(Pt, 1) && iy <= size(Pt, 2))
        dVx = (Vx[ix+1, iy] - Vx[ix, iy])/dx
        dVy = (Vy[ix, iy+1] - Vy[ix, iy])/dy
        dV  = dVx + dVy
        ∇V[ix, iy]  = dV
        Pt[ix, iy]  = Pt[ix, iy] - dτPt[ix, iy]*dV
        τxx[ix, iy] = 2.0*Mus[ix, iy]*(dVx - 1.0/3.0*dV)
        τyy[ix, iy] = 2.0*Mus[ix, iy]*(dVy - 1.0/3.0*dV)
    end
    if (ix <= size(τxy, 1) && iy <= size(τxy, 2))
        τxy[ix, iy] = 2.0*((Mus[ix, iy] + Mus[ix+1, iy] + Mus[ix, iy+1] + Mus[ix+1, iy+1])*0.25)*(0.5*((Vx[ix+1, iy+1] - Vx[ix+1, iy])/dy + (Vy[ix+1, iy+1] - Vy[ix, iy+1])/dx))
    end
    return
end
"""

STOKES2D_DVV = """
@parallel_indices (ix, iy) function compute_dVV!(Rx::Data.Array, Ry::Data.Array, dVxdτ::Data.Array, dVydτ::Data.Array, Vx::Data.Array, Vy::Data.Array, Pt::Data.Array, Rog::Data.Array, τxx::Data.Array, τyy::Data.Array, τxy::Data.Array, dτVx::Data.Array, dτVy::Data.Array, dampX::Data.Number, dampY::Data.Number, dx::Data.Number, dy::Data.Number)
    if (ix <= size(Rx, 1) && iy <= size(Rx, 2))
        rx = (τxx[ix+1, iy+1] - τxx[ix, iy+1])/dx + (τxy[ix, iy+1] - τxy[ix, iy])/dy - (Pt[ix+1, iy+1] - Pt[ix, iy+1])/dx
        Rx[ix, iy] = rx
        ax = dampX*dVxdτ[ix, iy] + rx
        dVxdτ[ix, iy] = ax
        vx = Vx[ix+1, iy+1] + dτVx[ix, iy]*ax
        Vx[ix+1, iy+1] = vx
        if (iy == 1) Vx[ix+1, 1] = vx end
        if (iy == size(Rx, 2)) Vx[ix+1, iy+2] = vx end
    end
    if (ix <= size(Ry, 1) && iy <= size(Ry, 2))
        ry = (τyy[ix+1, iy+1] - τyy[ix+1, iy])/dy + (τxy[ix+1, iy] - τxy[ix, iy])/dx - (Pt[ix+1, iy+1] - Pt[ix+1, iy])/dy + (Rog[ix+1, iy] + Rog[ix+1, iy+1])*0.5
        Ry[ix, iy] = ry
        ay = dampY*dVydτ[ix, iy] + ry
        dVydτ[ix, iy] = ay
        vy = Vy[ix+1, iy+1] + dτVy[ix, iy]*ay
        Vy[ix+1, iy+1] = vy
        if (ix == 1) Vy[1, iy+1] = vy end
        if (ix == size(Ry, 1)) Vy[ix+2, iy+1] = vy end
    end
    if (ix == 1 && iy == 1)
        Vx[1, 1] = Vx[1, 2]
        Vx[1, size(Vx, 2)] = Vx[1, size(Vx, 2)-1]
        Vx[size(Vx, 1), 1] = Vx[size(Vx, 1), 2]
        Vx[size(Vx, 1), size(Vx, 2)] = Vx[size(Vx, 1), size(Vx, 2)-1]
        Vy[1, 1] = Vy[2, 1]
        Vy[size(Vy, 1), 1] = Vy[size(Vy, 1)-1, 1]
        Vy[1, size(Vy, 2)] = Vy[2, size(Vy, 2)]
        Vy[size(Vy, 1), size(Vy, 2)] = Vy[size(Vy, 1)-1, size(Vy, 2)]
    end
    return
end
"""


class Stokes2DFused:
    """miniapps/Stokes2D.jl:62-125 with two launches per PT iteration (module docstring). Same
    interface as workloads.Stokes2D: set_material, reset, iterate, iterate_block, error, solve."""

    def __init__(self, module, nx, ny, lx=10.0, ly=10.0, Vdmp=4.0, Vsc=1.0, Ptsc=0.25, eps=1e-8, nout=200, itermax=500000,
                 eta_ts="av", graph=True, history=False):
        if eta_ts not in ("av", "maxV"):
            raise core.ArgumentError("eta_ts must be 'av' or 'maxV'")
        self.nx, self.ny = nx, ny
        self.dx, self.dy = lx / (nx - 1), ly / (ny - 1)
        self.min_dxy2 = min(self.dx, self.dy) ** 2
        self.max_nxy = max(nx, ny)
        self.dampX, self.dampY = 1.0 - Vdmp / nx, 1.0 - Vdmp / ny
        self.Vsc, self.Ptsc, self.eps, self.nout, self.itermax = Vsc, Ptsc, eps, nout, itermax
        self.eta_ts, self.use_graph = eta_ts, graph
        self.history = [] if history else None
        d = lambda text, kind="parallel": core._define(kind, text, {}, module)  # noqa: E731
        self.k_ts = d(W.STOKES2D_TIMESTEPS)
        self.k_mx, self.k_my = d(W.STOKES2D_MAXLOC_X, "parallel_indices"), d(W.STOKES2D_MAXLOC_Y, "parallel_indices")
        self.k_ptau, self.k_dvv = d(STOKES2D_PTAU, "parallel_indices"), d(STOKES2D_DVV, "parallel_indices")
        z = lambda *s: core.zeros(*s, eltype=core.Float64)  # noqa: E731
        self.Pt, self.dtPt, self.divV = z(nx, ny), z(nx, ny), z(nx, ny)
        self.Vx, self.Vy = z(nx + 1, ny), z(nx, ny + 1)
        self.txx, self.tyy, self.txy = z(nx, ny), z(nx, ny), z(nx - 1, ny - 1)
        self.Rx, self.Ry = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.dVxdt, self.dVydt = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.dtVx, self.dtVy = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.Mus, self.Rog, self.Mts, self.Mtmp = z(nx, ny), z(nx, ny), z(nx, ny), z(nx, ny)
        self.err_d = z(4)  # mean |Rx|, mean |Ry|, mean |∇V| of the last check, in device memory
        self._red_fields = (core.ps_field * 3)(self.Rx.f, self.Ry.f, self.divV.f)
        rng = ((1, nx), (1, ny))  # covers every cell, node and face the two kernels own
        self._steps = [
            self.k_ptau.launcher(self.divV, self.Pt, self.txx, self.tyy, self.txy, self.Vx, self.Vy, self.dtPt, self.Mus, self.dx,
                                 self.dy, ranges=rng),
            self.k_dvv.launcher(self.Rx, self.Ry, self.dVxdt, self.dVydt, self.Vx, self.Vy, self.Pt, self.Rog, self.txx, self.tyy,
                                self.txy, self.dtVx, self.dtVy, self.dampX, self.dampY, self.dx, self.dy, ranges=rng),
        ]
        self._graph = None
        self._host_err = np.zeros(4)
        self.iterate(1)  # compiles and loads both kernels (zero fields stay zero)
        self._check()
        core.synchronize()

    def set_material(self, Mus, Rog):
        self.Mus.copy_from(Mus)
        self.Rog.copy_from(Rog)
        self.k_ts(self.dtVx, self.dtVy, self.dtPt, self.Mus, self.Vsc, self.Ptsc, self.min_dxy2, self.max_nxy)
        if self.eta_ts == "maxV":
            self.k_mx(self.Mtmp, self.Mus)
            self.k_my(self.Mts, self.Mtmp)
            # velocity steps from the 3x3 maximum; the pressure step written to scratch and dropped
            self.k_ts(self.dtVx, self.dtVy, self.Mtmp, self.Mts, self.Vsc, self.Ptsc, self.min_dxy2, self.max_nxy)

    def reset(self, Vx=None, Vy=None, Pt=None):
        """Zero state, or a warm start (host arrays of the shapes of Vx, Vy, Pt); the wall-normal
        velocity nodes get the zero of the free-slip condition."""
        for A in (self.Pt, self.Vx, self.Vy, self.divV, self.txx, self.tyy, self.txy, self.Rx, self.Ry, self.dVxdt, self.dVydt):
            A.fill_(0.0)
        if Vx is not None:
            Vx = np.array(Vx, dtype=np.float64, order="F")
            Vy = np.array(Vy, dtype=np.float64, order="F")
            Vx[0, :] = Vx[-1, :] = 0.0
            Vy[:, 0] = Vy[:, -1] = 0.0
            self.Vx.copy_from(Vx)
            self.Vy.copy_from(Vy)
            self.Pt.copy_from(Pt)

    def iterate(self, n):
        """n PT iterations (miniapps/Stokes2D.jl:111-116), asynchronous."""
        a, b = self._steps
        for _ in range(n):
            a()
            b()

    def _check(self):
        check(lib.ps_field_reduce_noalloc(self._red_fields, 3, 3, ctypes.c_void_p(self.err_d.ptr), None))

    def iterate_block(self):
        """nout iterations and the residual reductions: one graph replay (captured on first use)."""
        if not self.use_graph:
            self.iterate(self.nout)
            self._check()
            return
        if self._graph is None:
            core.synchronize()
            g = core.Graph()
            with g:
                self.iterate(self.nout)
                self._check()
            self._graph = g
        self._graph.launch()

    def error(self):
        """max(mean |Rx|, mean |Ry|, mean |∇V|) of the last block (FP64 on the GPU)."""
        e = self.err_d.to_host()
        return float(max(e[0], e[1], e[2]))

    def solve(self):
        """Iterates in blocks of nout until the residual is below eps (or diverges); returns
        (iterations, final error). With history=True the error after every block is recorded."""
        it, err = 0, float("inf")
        if self.history is not None:
            self.history = []
        while it < self.itermax:
            self.iterate_block()
            it += self.nout
            err = self.error()
            if self.history is not None:
                self.history.append(err)
            if err <= self.eps or not err < 1e30:
                break
        return it, err

    def close(self):
        if self._graph is not None:
            self._graph.destroy()
            self._graph = None


def batched_text(text, guard=False):
    """The (ix, iy) kernel `text` over a batch of samples: the kernel gains the index iz (the
    sample), every array access A[i, j] becomes A[i, j, iz] and the name gets the suffix _batch.
    Sample iz then reads and writes plane iz of each (nx, ny, S) array only, with the statements,
    operations and their order of the 2-D kernel. With `guard` the kernel takes a first argument
    Act (1 x 1 x S) and runs its statements only for the samples whose Act is nonzero. The
    texts it applies to hold no nested subscripts and end with `return` and `end`."""
    t = re.sub(r"@parallel_indices \(ix, ?iy\)", "@parallel_indices (ix, iy, iz)", text)
    t = re.sub(r"function (\S+?)!\(", r"function \1_batch!(", t, count=1)
    t = re.sub(r"(\w+)\[([^\[\]]+)\]", r"\1[\2, iz]", t)
    if guard:
        _, head, body = t.split("\n", 2)
        head = head.replace("_batch!(", "_batch!(Act::Data.Array, ", 1)
        k = body.rindex("    return")
        t = "\n" + head + "\n    if (Act[1, 1, iz] != 0.0)\n" + body[:k] + "    end\n" + body[k:]
    return t


# compute_timesteps! of miniapps/Stokes2D.jl:11-16 over a batch, the FiniteDifferences2D macros
# expanded (@av_xi(A) = (A[ix,iy+1] + A[ix+1,iy+1])*0.5, @av_yi(A) = (A[ix+1,iy] + A[ix+1,iy+1])*0.5,
# src/FiniteDifferences.jl:175-176) and @all as the array's own bounds
STOKES2D_TIMESTEPS_BATCH = """
@parallel_indices (ix, iy, iz) function compute_timesteps_batch!(dτVx::Data.Array, dτVy::Data.Array, dτPt::Data.Array, Mus::Data.Array, Vsc::Data.Number, Ptsc::Data.Number, min_dxy2::Data.Number, max_nxy::Int)
    if (ix <= size(dτVx, 1) && iy <= size(dτVx, 2))
        dτVx[ix, iy, iz] = Vsc*min_dxy2/((Mus[ix, iy+1, iz] + Mus[ix+1, iy+1, iz])*0.5)/4.1
    end
    if (ix <= size(dτVy, 1) && iy <= size(dτVy, 2))
        dτVy[ix, iy, iz] = Vsc*min_dxy2/((Mus[ix+1, iy, iz] + Mus[ix+1, iy+1, iz])*0.5)/4.1
    end
    if (ix <= size(dτPt, 1) && iy <= size(dτPt, 2))
        dτPt[ix, iy, iz] = Ptsc*4.1*Mus[ix, iy, iz]/max_nxy
    end
    return
end
"""


class Stokes2DBatch:
    """S independent Stokes2D problems on one nx x ny grid (miniapps/Stokes2D.jl:62-125), solved
    together: every array gains a third dimension over the samples, (nx, ny, S) with each sample's
    plane contiguous, and a PT iteration is the two launches of Stokes2DFused over all S planes.
    At 511^2 one sample's launch is latency-bound (about 6 us for 261k cells); S samples share
    that latency. Per sample the fields after n iterations and the residuals of every check are
    bit for bit those of Stokes2DFused: the kernels are its own texts with a sample index, and
    each sample's residual is reduced over its plane in the 2-D field's order.

    solve() iterates until every sample has converged or diverged (or itermax) and returns each
    sample's iteration count and error at its first check below eps, as Stokes2DFused.solve would;
    samples that are done keep iterating with the rest of the batch."""

    def __init__(self, module, nx, ny, nsamples, lx=10.0, ly=10.0, Vdmp=4.0, Vsc=1.0, Ptsc=0.25, eps=1e-8, nout=200,
                 itermax=500000, eta_ts="av", graph=True):
        if eta_ts not in ("av", "maxV"):
            raise core.ArgumentError("eta_ts must be 'av' or 'maxV'")
        if nsamples < 1:
            raise core.ArgumentError("nsamples must be at least 1")
        S = int(nsamples)
        self.nx, self.ny, self.S = nx, ny, S
        self.dx, self.dy = lx / (nx - 1), ly / (ny - 1)
        self.min_dxy2 = min(self.dx, self.dy) ** 2
        self.max_nxy = max(nx, ny)
        self.dampX, self.dampY = 1.0 - Vdmp / nx, 1.0 - Vdmp / ny
        self.Vsc, self.Ptsc, self.eps, self.nout, self.itermax = Vsc, Ptsc, eps, nout, itermax
        self.eta_ts, self.use_graph = eta_ts, graph
        d = lambda text: core._define("parallel_indices", text, {}, module)  # noqa: E731
        self.k_ts = d(STOKES2D_TIMESTEPS_BATCH)
        self.k_mx, self.k_my = d(batched_text(W.STOKES2D_MAXLOC_X)), d(batched_text(W.STOKES2D_MAXLOC_Y))
        self.k_ptau, self.k_dvv = d(batched_text(STOKES2D_PTAU, guard=True)), d(batched_text(STOKES2D_DVV, guard=True))
        z = lambda a, b: core.zeros(a, b, S, eltype=core.Float64)  # noqa: E731
        self.Act = z(1, 1)  # 1: the sample of this plane iterates; 0: its fields stay as they are
        self.Pt, self.dtPt, self.divV = z(nx, ny), z(nx, ny), z(nx, ny)
        self.Vx, self.Vy = z(nx + 1, ny), z(nx, ny + 1)
        self.txx, self.tyy, self.txy = z(nx, ny), z(nx, ny), z(nx - 1, ny - 1)
        self.Rx, self.Ry = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.dVxdt, self.dVydt = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.dtVx, self.dtVy = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.Mus, self.Rog, self.Mts, self.Mtmp = z(nx, ny), z(nx, ny), z(nx, ny), z(nx, ny)
        # the residual check: mean |Rx|, mean |Ry|, mean |∇V| of every sample's plane, reduced in
        # calls of at most 4 plane views (one stream, so the calls never overlap)
        planes = [A.view((1, 1, s), (A.shape[0], A.shape[1], s)) for s in range(1, S + 1) for A in (self.Rx, self.Ry, self.divV)]
        self._planes = planes  # the views hold their parents
        self.err_d = core.zeros(4 * ((3 * S + 3) // 4), eltype=core.Float64)
        self._red = [((core.ps_field * len(planes[c:c + 4]))(*[v.f for v in planes[c:c + 4]]), len(planes[c:c + 4]), self.err_d.ptr + 8 * c)
                     for c in range(0, 3 * S, 4)]
        rng = ((1, nx), (1, ny), (1, S))
        self._rng = rng
        self._steps = [
            self.k_ptau.launcher(self.Act, self.divV, self.Pt, self.txx, self.tyy, self.txy, self.Vx, self.Vy, self.dtPt, self.Mus,
                                 self.dx, self.dy, ranges=rng),
            self.k_dvv.launcher(self.Act, self.Rx, self.Ry, self.dVxdt, self.dVydt, self.Vx, self.Vy, self.Pt, self.Rog, self.txx,
                                self.tyy, self.txy, self.dtVx, self.dtVy, self.dampX, self.dampY, self.dx, self.dy, ranges=rng),
        ]
        self._graph = None
        self.iterate(1)  # compiles and loads both kernels (no plane is active yet)
        self._check()
        core.synchronize()

    def set_active(self, s, on):
        """Lets the sample of plane s (0-based) iterate, or freezes its fields."""
        self._plane("Act", s).fill_(1.0 if on else 0.0)

    def set_material(self, Mus, Rog):
        """Viscosity and buoyancy of every sample: host arrays (nx, ny, S)."""
        self.Mus.copy_from(Mus)
        self.Rog.copy_from(Rog)
        rng = self._rng
        self.k_ts(self.dtVx, self.dtVy, self.dtPt, self.Mus, self.Vsc, self.Ptsc, self.min_dxy2, self.max_nxy, ranges=rng)
        if self.eta_ts == "maxV":
            self.k_mx(self.Mtmp, self.Mus, ranges=rng)
            self.k_my(self.Mts, self.Mtmp, ranges=rng)
            self.k_ts(self.dtVx, self.dtVy, self.Mtmp, self.Mts, self.Vsc, self.Ptsc, self.min_dxy2, self.max_nxy, ranges=rng)

    def reset(self, Vx=None, Vy=None, Pt=None):
        """Zero state, or warm starts (host arrays (nx+1, ny, S), (nx, ny+1, S), (nx, ny, S)); the
        wall-normal velocity nodes get the zero of the free-slip condition. Every sample is active."""
        for A in (self.Pt, self.Vx, self.Vy, self.divV, self.txx, self.tyy, self.txy, self.Rx, self.Ry, self.dVxdt, self.dVydt):
            A.fill_(0.0)
        self.Act.fill_(1.0)
        if Vx is not None:
            Vx = np.array(Vx, dtype=np.float64, order="F")
            Vy = np.array(Vy, dtype=np.float64, order="F")
            Vx[0, :, :] = Vx[-1, :, :] = 0.0
            Vy[:, 0, :] = Vy[:, -1, :] = 0.0
            self.Vx.copy_from(Vx)
            self.Vy.copy_from(Vy)
            self.Pt.copy_from(Pt)

    def iterate(self, n):
        """n PT iterations of every sample, asynchronous."""
        a, b = self._steps
        for _ in range(n):
            a()
            b()

    def _check(self):
        for fields, n, out in self._red:
            check(lib.ps_field_reduce_noalloc(fields, n, 3, ctypes.c_void_p(out), None))

    def iterate_block(self):
        """nout iterations and the residual reductions: one graph replay (captured on first use)."""
        if not self.use_graph:
            self.iterate(self.nout)
            self._check()
            return
        if self._graph is None:
            core.synchronize()
            g = core.Graph()
            with g:
                self.iterate(self.nout)
                self._check()
            self._graph = g
        self._graph.launch()

    def errors(self):
        """Per sample max(mean |Rx|, mean |Ry|, mean |∇V|) of the last check (shape (S,))."""
        e = self.err_d.to_host()[:3 * self.S].reshape(self.S, 3)
        return np.array([max(r[0], r[1], r[2]) for r in e])  # Stokes2DFused.error's max, NaNs included

    def solve(self):
        """Blocks of nout iterations until every sample is below eps or diverged (or itermax);
        returns (iterations, errors), arrays of shape (S,). A sample is frozen at the check that
        ends it, so its fields are then those Stokes2DFused.solve leaves."""
        it = 0
        iters = np.zeros(self.S, dtype=np.int64)
        errs = np.full(self.S, np.inf)
        done = np.zeros(self.S, dtype=bool)
        while it < self.itermax:
            self.iterate_block()
            it += self.nout
            e = self.errors()
            for s in np.nonzero(~done)[0]:
                if e[s] <= self.eps or not e[s] < 1e30:
                    done[s], iters[s], errs[s] = True, it, e[s]
                    self.set_active(int(s), False)
            if done.all():
                break
        iters[~done] = it
        errs[~done] = self.errors()[~done]
        return iters, errs

    _STATE = ("Pt", "Vx", "Vy", "divV", "txx", "tyy", "txy", "Rx", "Ry", "dVxdt", "dVydt")

    def _plane(self, name, s):
        """The view of plane s (0-based) of array `name`, made once."""
        if not hasattr(self, "_pv"):
            self._pv = {}
        v = self._pv.get((name, s))
        if v is None:
            A = getattr(self, name)
            v = self._pv[(name, s)] = A.view((1, 1, s + 1), (A.shape[0], A.shape[1], s + 1))
        return v

    def load_sample(self, s, Mus, Rog, init=None):
        """Puts one sample into plane s (0-based): its viscosity and body force (host (nx, ny)),
        its pseudo-time steps, and state from rest or the warm start init = (Vx, Vy, Pt); the
        other planes are left as they are. Ordered on the library stream before later launches."""
        self._plane("Mus", s).copy_from(Mus)
        self._plane("Rog", s).copy_from(Rog)
        rng = ((1, self.nx), (1, self.ny), (s + 1, s + 1))
        self.k_ts(self.dtVx, self.dtVy, self.dtPt, self.Mus, self.Vsc, self.Ptsc, self.min_dxy2, self.max_nxy, ranges=rng)
        if self.eta_ts == "maxV":
            self.k_mx(self.Mtmp, self.Mus, ranges=rng)
            self.k_my(self.Mts, self.Mtmp, ranges=rng)
            self.k_ts(self.dtVx, self.dtVy, self.Mtmp, self.Mts, self.Vsc, self.Ptsc, self.min_dxy2, self.max_nxy, ranges=rng)
        for name in self._STATE:
            self._plane(name, s).fill_(0.0)
        if init is not None:
            Vx = np.array(init[0], dtype=np.float64, order="F")
            Vy = np.array(init[1], dtype=np.float64, order="F")
            Vx[0, :] = Vx[-1, :] = 0.0
            Vy[:, 0] = Vy[:, -1] = 0.0
            self._plane("Vx", s).copy_from(Vx)
            self._plane("Vy", s).copy_from(Vy)
            self._plane("Pt", s).copy_from(init[2])

    def solve_queue(self, samples):
        """Solves a stream of samples S at a time, each plane a slot: whenever a slot's sample is
        below eps, diverged or at its iteration limit (at a residual check, as Stokes2DFused.solve
        stops), its result is taken from the plane and the next sample starts there. `samples`
        yields (Mus, Rog), (Mus, Rog, init) or (Mus, Rog, init, itermax) with init None or a warm
        start (Vx, Vy, Pt), and itermax None for the solver's. Yields (index, iterations, error,
        Vx, Vy, Pt) per sample in the order they finish: bit for bit what Stokes2DFused.solve
        returns and leaves in its fields for that sample alone. A slot left without a sample is
        switched off (its plane no longer iterates), so the last samples run at nearly the cost of
        a batch of one."""
        src = enumerate(samples)
        slot = [None] * self.S  # [index, iterations so far, itermax]

        def fill(s):
            nxt = next(src, None)
            if nxt is None:
                slot[s] = None
                self.set_active(s, False)
                return
            i, item = nxt
            self.load_sample(s, item[0], item[1], item[2] if len(item) > 2 else None)
            self.set_active(s, True)
            imax = item[3] if len(item) > 3 and item[3] else self.itermax
            slot[s] = [i, 0, imax]

        for s in range(self.S):
            fill(s)
        while any(x is not None for x in slot):
            self.iterate_block()
            e = self.errors()
            for s in range(self.S):
                if slot[s] is None:
                    continue
                slot[s][1] += self.nout
                i, its, imax = slot[s]
                if e[s] <= self.eps or not e[s] < 1e30 or its >= imax:
                    yield (i, its, float(e[s]), self._plane("Vx", s).to_host()[:, :, 0], self._plane("Vy", s).to_host()[:, :, 0],
                           self._plane("Pt", s).to_host()[:, :, 0])
                    fill(s)

    def close(self):
        if self._graph is not None:
            self._graph.destroy()
            self._graph = None
