"""Kernel texts of the test-bed applications, written as the reference writes them
(examples/ and miniapps/ of ParallelStencil.jl) and defined through the public API.
"""
from . import core

# 3-D heat diffusion, flux form with on-the-fly fluxes (the diffusion configuration:
# qTx = -lam*@d_xi(T)/dx ...; T2[inn] = T[inn] + dt/Cp*(-@d_xa(qTx)/dx - ...)).
DIFFUSION3D = """
@parallel function diffusion3D_step!(T2, T, Cp, lam, dt, dx, dy, dz)
    @all(qTx) = -lam*@d_xi(T)/dx
    @all(qTy) = -lam*@d_yi(T)/dy
    @all(qTz) = -lam*@d_zi(T)/dz
    @inn(T2)  = @inn(T) + dt/@inn(Cp)*(-@d_xa(qTx)/dx - @d_ya(qTy)/dy - @d_za(qTz)/dz)
    return
end
"""

# examples/diffusion2D_shmem_novis.jl:11-27: the kernel language (@sharedMem tile with a one-cell
# halo, @sync_threads); launched with explicit nblocks/nthreads and shmem=prod(nthreads .+ 2)*8
DIFFUSION2D_SHMEM = """
@parallel_indices (ix,iy) function diffusion2D_step!(T2, T, Ci, lam, dt, _dx, _dy)
    tx  = @threadIdx().x + 1
    ty  = @threadIdx().y + 1
    T_l = @sharedMem(eltype(T), (@blockDim().x+2, @blockDim().y+2))
    T_l[tx,ty] = T[ix,iy]
    if (1<ix<size(T2,1) && 1<iy<size(T2,2))
        if (@threadIdx().x == 1)             T_l[tx-1,ty] = T[ix-1,iy] end
        if (@threadIdx().x == @blockDim().x) T_l[tx+1,ty] = T[ix+1,iy] end
        if (@threadIdx().y == 1)             T_l[tx,ty-1] = T[ix,iy-1] end
        if (@threadIdx().y == @blockDim().y) T_l[tx,ty+1] = T[ix,iy+1] end
        @sync_threads()
        T2[ix,iy] = T_l[tx,ty] + dt*(Ci[ix,iy]*(
                    - ((-lam*(T_l[tx+1,ty] - T_l[tx,ty])*_dx) - (-lam*(T_l[tx,ty] - T_l[tx-1,ty])*_dx))*_dx
                    - ((-lam*(T_l[tx,ty+1] - T_l[tx,ty])*_dy) - (-lam*(T_l[tx,ty] - T_l[tx,ty-1])*_dy))*_dy
                    ));
    end
    return
end
"""

# The same update with FiniteDifferences2D macros and on-the-fly fluxes: the same operations in
# the same order, so both kernels give bit-identical fields.
DIFFUSION2D_FD = """
@parallel function diffusion2D_fd!(T2, T, Ci, lam, dt, _dx, _dy)
    @all(qx) = -lam*@d_xi(T)*_dx
    @all(qy) = -lam*@d_yi(T)*_dy
    @inn(T2) = @inn(T) + dt*(@inn(Ci)*(-@d_xa(qx)*_dx - @d_ya(qy)*_dy))
    return
end
"""

# examples/memcopy3D.jl:10-13
TRIAD3D = """
@parallel function copy3D!(T2, T, Ci)
    @all(T2) = @all(T) + @all(Ci)
    return
end
"""

# miniapps/acoustic3D.jl:11-21
ACOUSTIC3D_V = """
@parallel function compute_V!(Vx::Data.Array, Vy::Data.Array, Vz::Data.Array, P::Data.Array, dt::Data.Number, ρ::Data.Number, dx::Data.Number, dy::Data.Number, dz::Data.Number)
    @inn(Vx) = @inn(Vx) - dt/ρ*@d_xi(P)/dx
    @inn(Vy) = @inn(Vy) - dt/ρ*@d_yi(P)/dy
    @inn(Vz) = @inn(Vz) - dt/ρ*@d_zi(P)/dz
    return
end
"""
ACOUSTIC3D_P = """
@parallel function compute_P!(P::Data.Array, Vx::Data.Array, Vy::Data.Array, Vz::Data.Array, dt::Data.Number, k::Data.Number, dx::Data.Number, dy::Data.Number, dz::Data.Number)
    @all(P) = @all(P) - dt*k*(@d_xa(Vx)/dx + @d_ya(Vy)/dy + @d_za(Vz)/dz)
    return
end
"""

# miniapps/acoustic2D.jl:11-20
ACOUSTIC2D_V = """
@parallel function compute_V!(Vx::Data.Array, Vy::Data.Array, P::Data.Array, dt::Data.Number, ρ::Data.Number, dx::Data.Number, dy::Data.Number)
    @inn(Vx) = @inn(Vx) - dt/ρ*@d_xi(P)/dx
    @inn(Vy) = @inn(Vy) - dt/ρ*@d_yi(P)/dy
    return
end
"""
ACOUSTIC2D_P = """
@parallel function compute_P!(P::Data.Array, Vx::Data.Array, Vy::Data.Array, dt::Data.Number, k::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(P) = @all(P) - dt*k*(@d_xa(Vx)/dx + @d_ya(Vy)/dy)
    return
end
"""


def splitmix64_uniform(n, seed):
    """n uniform [0,1) doubles from SplitMix64 (53-bit mantissa), identical on every host."""
    import numpy as np
    x = np.uint64(seed)
    with np.errstate(over="ignore"):
        z = x + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def diffusion_inputs(n, seed=1, lx=10.0):
    """Initial state of the 3-D heat diffusion configuration on n^3 cells of size lx/n:
    T0 = 100 exp(-|x - c|^2) on cell centres, Cp = 1 + U[0,1) (SplitMix64, seed),
    dt = min(dx^2, dy^2, dz^2) * cp_min / lam / 8.1 with cp_min = lam = 1."""
    import numpy as np
    dx = lx / n
    x = (np.arange(1, n + 1) - 0.5) * dx
    # exp(-(a+b+c)) evaluated as one exp per cell, as the configuration states it
    X2 = (x - lx / 2) ** 2
    T = np.empty((n, n, n), order="F")
    for k in range(n):
        T[:, :, k] = 100.0 * np.exp(-(X2[:, None] + X2[None, :] + X2[k]))
    Cp = np.asfortranarray((1.0 + splitmix64_uniform(n ** 3, seed)).reshape((n, n, n), order="F"))
    dt = min(dx * dx, dx * dx, dx * dx) * 1.0 / 1.0 / 8.1
    return T, Cp, dict(lam=1.0, dt=dt, dx=dx, dy=dx, dz=dx)


def define_diffusion3d(memopt, module, **kw):
    return core._define("parallel", DIFFUSION3D, dict(kw, memopt=memopt), module)


def define_diffusion2d_shmem(module, inbounds=False):
    return core._define("parallel_indices", DIFFUSION2D_SHMEM, {"inbounds": inbounds}, module)


def define_diffusion2d_fd(module, memopt=False):
    return core._define("parallel", DIFFUSION2D_FD, {"memopt": memopt}, module)


def define_triad3d(module):
    return core._define("parallel", TRIAD3D, {}, module)


# ------------------------------------------------------------------ Stokes 2-D (pseudo-transient)
# miniapps/Stokes2D.jl:11-42 (2-D module). The boundary kernels of :44-54 declare (ix,iy) and use
# one index, which src/parallel.jl:706-713 rejects; they are written with the one index they use
# and launched over that dimension of the array.
STOKES2D_TIMESTEPS = """
@parallel function compute_timesteps!(dτVx::Data.Array, dτVy::Data.Array, dτPt::Data.Array, Mus::Data.Array, Vsc::Data.Number, Ptsc::Data.Number, min_dxy2::Data.Number, max_nxy::Int)
    @all(dτVx) = Vsc*min_dxy2/@av_xi(Mus)/4.1
    @all(dτVy) = Vsc*min_dxy2/@av_yi(Mus)/4.1
    @all(dτPt) = Ptsc*4.1*@all(Mus)/max_nxy
    return
end
"""
STOKES2D_P = """
@parallel function compute_P!(∇V::Data.Array, Pt::Data.Array, Vx::Data.Array, Vy::Data.Array, dτPt::Data.Array, dx::Data.Number, dy::Data.Number)
    @all(∇V)  = @d_xa(Vx)/dx + @d_ya(Vy)/dy
    @all(Pt)  = @all(Pt) - @all(dτPt)*@all(∇V)
    return
end
"""
STOKES2D_TAU = """
@parallel function compute_τ!(∇V::Data.Array, τxx::Data.Array, τyy::Data.Array, τxy::Data.Array, Vx::Data.Array, Vy::Data.Array, Mus::Data.Array, dx::Data.Number, dy::Data.Number)
    @all(τxx) = 2.0*@all(Mus)*(@d_xa(Vx)/dx - 1.0/3.0*@all(∇V))
    @all(τyy) = 2.0*@all(Mus)*(@d_ya(Vy)/dy - 1.0/3.0*@all(∇V))
    @all(τxy) = 2.0*@av(Mus)*(0.5*(@d_yi(Vx)/dy + @d_xi(Vy)/dx))
    return
end
"""
STOKES2D_DV = """
@parallel function compute_dV!(Rx::Data.Array, Ry::Data.Array, dVxdτ::Data.Array, dVydτ::Data.Array, Pt::Data.Array, Rog::Data.Array, τxx::Data.Array, τyy::Data.Array, τxy::Data.Array, dampX::Data.Number, dampY::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(Rx)    = @d_xi(τxx)/dx + @d_ya(τxy)/dy - @d_xi(Pt)/dx
    @all(Ry)    = @d_yi(τyy)/dy + @d_xa(τxy)/dx - @d_yi(Pt)/dy + @av_yi(Rog)
    @all(dVxdτ) = dampX*@all(dVxdτ) + @all(Rx)
    @all(dVydτ) = dampY*@all(dVydτ) + @all(Ry)
    return
end
"""
STOKES2D_V = """
@parallel function compute_V!(Vx::Data.Array, Vy::Data.Array, dVxdτ::Data.Array, dVydτ::Data.Array, dτVx::Data.Array, dτVy::Data.Array)
    @inn(Vx) = @inn(Vx) + @all(dτVx)*@all(dVxdτ)
    @inn(Vy) = @inn(Vy) + @all(dτVy)*@all(dVydτ)
    return
end
"""
STOKES2D_BC_X = """
@parallel_indices (iy) function bc_x!(A::Data.Array)
    A[1  , iy] = A[2    , iy]
    A[end, iy] = A[end-1, iy]
    return
end
"""
STOKES2D_BC_Y = """
@parallel_indices (ix) function bc_y!(A::Data.Array)
    A[ix, 1  ] = A[ix, 2    ]
    A[ix, end] = A[ix, end-1]
    return
end
"""


def stokes2d_inclusions(nx, ny, seed, lx=10.0, ly=10.0, mu0=1.0, log10_contrast=3.0):
    """Viscosity and body force of one Stokes sample on the miniapp's node grid
    (x = (ix-1)dx, dx = lx/(nx-1)): 1-8 circular inclusions with centres uniform in the box,
    radii uniform [0.3, 1.5], log10 viscosity contrast uniform [-c, c] (c = log10_contrast),
    rho*g = 1 inside and 0 outside (later inclusions overwrite earlier ones). SplitMix64 stream
    `seed`. The miniapp's pseudo-transient iteration converges for contrasts up to about one
    order of magnitude (c = 1) and diverges from about 10^1.5 on."""
    import numpy as np
    u = splitmix64_uniform(1 + 4 * 8, seed)
    ninc = 1 + int(u[0] * 8)
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    x = (np.arange(nx) * dx)[:, None]
    y = (np.arange(ny) * dy)[None, :]
    Mus = np.full((nx, ny), mu0, order="F")
    Rog = np.zeros((nx, ny), order="F")
    for i in range(ninc):
        cx, cy, r = u[1 + 4 * i] * lx, u[2 + 4 * i] * ly, 0.3 + 1.2 * u[3 + 4 * i]
        c = log10_contrast * (2.0 * u[4 + 4 * i] - 1.0)
        m = (x - cx) ** 2 + (y - cy) ** 2 < r * r
        Mus[m] = mu0 * 10.0 ** c
        Rog[m] = 1.0
    return Mus, Rog


STOKES2D_MAXLOC_X = """
@parallel_indices (ix,iy) function maxloc_x!(Mx::Data.Array, Mus::Data.Array)
    Mx[ix,iy] = max(max(Mus[max(ix-1,1),iy], Mus[ix,iy]), Mus[min(ix+1,size(Mus,1)),iy])
    return
end
"""
STOKES2D_MAXLOC_Y = """
@parallel_indices (ix,iy) function maxloc_y!(My::Data.Array, Mx::Data.Array)
    My[ix,iy] = max(max(Mx[ix,max(iy-1,1)], Mx[ix,iy]), Mx[ix,min(iy+1,size(Mx,2))])
    return
end
"""


class Stokes2D:
    """The pseudo-transient Stokes solver of miniapps/Stokes2D.jl:62-125 on the GPU.

    Launches are prepared once (Kernel.launcher) and issued asynchronously; with `graph=True`
    each block of `nout` iterations between two residual checks is one HIP graph replay. The
    residual check (mean |Rx|, mean |Ry|, mean |∇V|, :117-123) is a deterministic FP64 reduction
    on the GPU.

    Pseudo-time steps: `eta_ts="av"` (default) is the miniapp's compute_timesteps! on the
    viscosity (two-point averages @av_xi / @av_yi, :11-16). `eta_ts="maxloc"` feeds the same
    kernel the 3x3 local maximum of the viscosity instead. Neither makes the iteration converge
    for inclusion contrasts beyond about one order of magnitude."""

    def __init__(self, module, nx, ny, lx=10.0, ly=10.0, Vdmp=4.0, Vsc=1.0, Ptsc=0.25, eps=1e-8, nout=200, itermax=500000,
                 eta_ts="av", graph=True):
        self.nx, self.ny = nx, ny
        self.dx, self.dy = lx / (nx - 1), ly / (ny - 1)
        self.min_dxy2 = min(self.dx, self.dy) ** 2
        self.max_nxy = max(nx, ny)
        self.dampX, self.dampY = 1.0 - Vdmp / nx, 1.0 - Vdmp / ny
        self.Vsc, self.Ptsc, self.eps, self.nout, self.itermax = Vsc, Ptsc, eps, nout, itermax
        self.eta_ts, self.use_graph = eta_ts, graph
        d = lambda text, kind="parallel": core._define(kind, text, {}, module)
        self.k_ts, self.k_p, self.k_tau = d(STOKES2D_TIMESTEPS), d(STOKES2D_P), d(STOKES2D_TAU)
        self.k_dv, self.k_v = d(STOKES2D_DV), d(STOKES2D_V)
        self.k_bcx, self.k_bcy = d(STOKES2D_BC_X, "parallel_indices"), d(STOKES2D_BC_Y, "parallel_indices")
        self.k_mx, self.k_my = d(STOKES2D_MAXLOC_X, "parallel_indices"), d(STOKES2D_MAXLOC_Y, "parallel_indices")
        z = lambda *s: core.zeros(*s, eltype=core.Float64)
        self.Pt, self.dtPt, self.divV = z(nx, ny), z(nx, ny), z(nx, ny)
        self.Vx, self.Vy = z(nx + 1, ny), z(nx, ny + 1)
        self.txx, self.tyy, self.txy = z(nx, ny), z(nx, ny), z(nx - 1, ny - 1)
        self.Rx, self.Ry = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.dVxdt, self.dVydt = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.dtVx, self.dtVy = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.Mus, self.Rog, self.Mts, self.Mtmp = z(nx, ny), z(nx, ny), z(nx, ny), z(nx, ny)
        self._steps = [
            self.k_p.launcher(self.divV, self.Pt, self.Vx, self.Vy, self.dtPt, self.dx, self.dy),
            self.k_tau.launcher(self.divV, self.txx, self.tyy, self.txy, self.Vx, self.Vy, self.Mus, self.dx, self.dy),
            self.k_dv.launcher(self.Rx, self.Ry, self.dVxdt, self.dVydt, self.Pt, self.Rog, self.txx, self.tyy, self.txy,
                               self.dampX, self.dampY, self.dx, self.dy),
            self.k_v.launcher(self.Vx, self.Vy, self.dVxdt, self.dVydt, self.dtVx, self.dtVy),
            self.k_bcy.launcher(self.Vx, ranges=((1, nx + 1),)),
            self.k_bcx.launcher(self.Vy, ranges=((1, ny + 1),)),
        ]
        self._graph = None
        # compile and load every kernel now (all fields are zero: the iteration maps zeros to
        # zeros), so that graphs can be captured later
        self.iterate(1)
        core.synchronize()

    def set_material(self, Mus, Rog):
        self.Mus.copy_from(Mus)
        self.Rog.copy_from(Rog)
        M = self.Mus
        if self.eta_ts == "maxloc":
            self.k_mx(self.Mtmp, self.Mus)
            self.k_my(self.Mts, self.Mtmp)
            M = self.Mts
        self.k_ts(self.dtVx, self.dtVy, self.dtPt, M, self.Vsc, self.Ptsc, self.min_dxy2, self.max_nxy)

    def reset(self, Vx=None, Vy=None, Pt=None):
        """Zero state, or a warm start (host arrays of the shapes of Vx, Vy, Pt). The iteration never
        updates the wall-normal velocity nodes (Vx at the x walls, Vy at the y walls), so a warm
        start gets them set to the zero of the free-slip condition."""
        import numpy as np
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
        steps = self._steps
        for _ in range(n):
            for s in steps:
                s()

    def iterate_block(self):
        """nout iterations: one graph replay (captured on first use) or nout eager iterations."""
        if not self.use_graph:
            self.iterate(self.nout)
            return
        if self._graph is None:
            core.synchronize()
            g = core.Graph()
            with g:
                self.iterate(self.nout)
            self._graph = g
        self._graph.launch()

    def error(self):
        """max(mean |Rx|, mean |Ry|, mean |∇V|), FP64 on the GPU."""
        return max(self.Rx.reduce("mean_abs"), self.Ry.reduce("mean_abs"), self.divV.reduce("mean_abs"))

    def solve(self):
        """Iterates until the error check (every nout iterations) is below eps; returns the
        number of iterations and the final error."""
        it, err = 0, float("inf")
        while it < self.itermax:
            self.iterate_block()
            it += self.nout
            err = self.error()
            if err <= self.eps or not err < 1e30:  # converged, or diverged (inf / NaN)
                break
        return it, err


# ------------------------------------------------------------------ Stokes3D (miniapps/Stokes3D.jl)
STOKES3D_TIMESTEPS = """
@parallel function compute_timesteps!(dτVx::Data.Array, dτVy::Data.Array, dτVz::Data.Array, dτPt::Data.Array, Mus::Data.Array, Vsc::Data.Number, Ptsc::Data.Number, min_dxyz2::Data.Number, max_nxyz::Int)
    @all(dτVx) = Vsc*min_dxyz2/@av_xi(Mus)/6.1
    @all(dτVy) = Vsc*min_dxyz2/@av_yi(Mus)/6.1
    @all(dτVz) = Vsc*min_dxyz2/@av_zi(Mus)/6.1
    @all(dτPt) = Ptsc*6.1*@all(Mus)/max_nxyz
    return
end
"""
STOKES3D_P = """
@parallel function compute_P!(∇V::Data.Array, Pt::Data.Array, Vx::Data.Array, Vy::Data.Array, Vz::Data.Array, dτPt::Data.Array, dx::Data.Number, dy::Data.Number, dz::Data.Number)
    @all(∇V) = @d_xa(Vx)/dx + @d_ya(Vy)/dy + @d_za(Vz)/dz
    @all(Pt) = @all(Pt) - @all(dτPt)*@all(∇V)
    return
end
"""
STOKES3D_TAU = """
@parallel function compute_τ!(∇V::Data.Array, τxx::Data.Array, τyy::Data.Array, τzz::Data.Array, τxy::Data.Array, τxz::Data.Array, τyz::Data.Array, Vx::Data.Array, Vy::Data.Array, Vz::Data.Array, Mus::Data.Array, dx::Data.Number, dy::Data.Number, dz::Data.Number)
    @all(τxx) = 2.0*@inn_yz(Mus)*(@d_xi(Vx)/dx - 1.0/3.0*@inn_yz(∇V))
    @all(τyy) = 2.0*@inn_xz(Mus)*(@d_yi(Vy)/dy - 1.0/3.0*@inn_xz(∇V))
    @all(τzz) = 2.0*@inn_xy(Mus)*(@d_zi(Vz)/dz - 1.0/3.0*@inn_xy(∇V))
    @all(τxy) = 2.0*@av_xyi(Mus)*(0.5*(@d_yi(Vx)/dy + @d_xi(Vy)/dx))
    @all(τxz) = 2.0*@av_xzi(Mus)*(0.5*(@d_zi(Vx)/dz + @d_xi(Vz)/dx))
    @all(τyz) = 2.0*@av_yzi(Mus)*(0.5*(@d_zi(Vy)/dz + @d_yi(Vz)/dy))
    return
end
"""
STOKES3D_DV = """
@parallel function compute_dV!(Rx::Data.Array, Ry::Data.Array, Rz::Data.Array, dVxdτ::Data.Array, dVydτ::Data.Array, dVzdτ::Data.Array, Pt::Data.Array, Rog::Data.Array, τxx::Data.Array, τyy::Data.Array, τzz::Data.Array, τxy::Data.Array, τxz::Data.Array, τyz::Data.Array, dampX::Data.Number, dampY::Data.Number, dampZ::Data.Number, dx::Data.Number, dy::Data.Number, dz::Data.Number)
    @all(Rx) = @d_xa(τxx)/dx + @d_ya(τxy)/dy + @d_za(τxz)/dz - @d_xi(Pt)/dx
    @all(Ry) = @d_ya(τyy)/dy + @d_xa(τxy)/dx + @d_za(τyz)/dz - @d_yi(Pt)/dy
    @all(Rz) = @d_za(τzz)/dz + @d_xa(τxz)/dx + @d_ya(τyz)/dy - @d_zi(Pt)/dz + @av_zi(Rog)
    @all(dVxdτ) = dampX*@all(dVxdτ) + @all(Rx)
    @all(dVydτ) = dampY*@all(dVydτ) + @all(Ry)
    @all(dVzdτ) = dampZ*@all(dVzdτ) + @all(Rz)
    return
end
"""
STOKES3D_V = """
@parallel function compute_V!(Vx::Data.Array, Vy::Data.Array, Vz::Data.Array, dVxdτ::Data.Array, dVydτ::Data.Array, dVzdτ::Data.Array, dτVx::Data.Array, dτVy::Data.Array, dτVz::Data.Array)
    @inn(Vx) = @inn(Vx) + @all(dτVx)*@all(dVxdτ)
    @inn(Vy) = @inn(Vy) + @all(dτVy)*@all(dVydτ)
    @inn(Vz) = @inn(Vz) + @all(dτVz)*@all(dVzdτ)
    return
end
"""
STOKES3D_BC = {
    "x": """
@parallel_indices (iy,iz) function bc_x!(A::Data.Array)
    A[1  , iy, iz] = A[2    , iy, iz]
    A[end, iy, iz] = A[end-1, iy, iz]
    return
end
""",
    "y": """
@parallel_indices (ix,iz) function bc_y!(A::Data.Array)
    A[ix, 1  , iz] = A[ix, 2    , iz]
    A[ix, end, iz] = A[ix, end-1, iz]
    return
end
""",
    "z": """
@parallel_indices (ix,iy) function bc_z!(A::Data.Array)
    A[ix, iy, 1  ] = A[ix, iy, 2    ]
    A[ix, iy, end] = A[ix, iy, end-1]
    return
end
""",
}


def stokes3d_material(nx, ny, nz, lx=10.0, ly=10.0, lz=10.0, mus0=1.0, musi=0.1, rogi=1.0):
    """Viscosity and density*gravity of miniapps/Stokes3D.jl:110-115: a sphere of radius 1 at the
    centre of the box (squared radius from ((ix-1)dx - lx/2)^2 + ..., summed left to right)."""
    import numpy as np
    dx, dy, dz = lx / (nx - 1), ly / (ny - 1), lz / (nz - 1)
    x = (np.arange(nx) * dx - 0.5 * lx) ** 2
    y = (np.arange(ny) * dy - 0.5 * ly) ** 2
    z = (np.arange(nz) * dz - 0.5 * lz) ** 2
    Radc = np.asfortranarray((x[:, None, None] + y[None, :, None]) + z[None, None, :])
    Mus = np.asfortranarray(np.full((nx, ny, nz), mus0))
    Rog = np.zeros((nx, ny, nz), order="F")
    Mus[Radc <= 1.0] = musi
    Rog[Radc <= 1.0] = rogi
    return Mus, Rog


class Stokes3D:
    """The pseudo-transient Stokes solver of miniapps/Stokes3D.jl:71-143 on the GPU: compute_P!,
    compute_τ!, compute_dV!, compute_V! and the six free-slip boundary launches per iteration
    (:133-142), prepared once and issued asynchronously; the residual check every nout
    iterations (:144-147) is max(mean |Rx|, mean |Ry|, mean |Rz|, mean |∇V|), a deterministic FP64
    reduction on the GPU. `graph=True` replays each block of nout iterations as one HIP graph."""

    def __init__(self, module, nx, ny, nz, lx=10.0, ly=10.0, lz=10.0, Vdmp=4.0, Vsc=1.0, Ptsc=0.25, eps=1e-6, nout=500,
                 itermax=10000, graph=True):
        self.n = (nx, ny, nz)
        self.dx, self.dy, self.dz = lx / (nx - 1), ly / (ny - 1), lz / (nz - 1)
        self.min_dxyz2 = min(self.dx, self.dy, self.dz) ** 2
        self.max_nxyz = max(nx, ny, nz)
        self.dampX, self.dampY, self.dampZ = 1.0 - Vdmp / nx, 1.0 - Vdmp / ny, 1.0 - Vdmp / nz
        self.Vsc, self.Ptsc, self.eps, self.nout, self.itermax, self.use_graph = Vsc, Ptsc, eps, nout, itermax, graph
        d = lambda text, kind="parallel": core._define(kind, text, {}, module)  # noqa: E731
        self.k_ts, self.k_p, self.k_tau = d(STOKES3D_TIMESTEPS), d(STOKES3D_P), d(STOKES3D_TAU)
        self.k_dv, self.k_v = d(STOKES3D_DV), d(STOKES3D_V)
        self.k_bc = {a: d(t, "parallel_indices") for a, t in STOKES3D_BC.items()}
        z = lambda *s: core.zeros(*s, eltype=core.Float64)  # noqa: E731
        self.Pt, self.dtPt, self.divV = z(nx, ny, nz), z(nx, ny, nz), z(nx, ny, nz)
        self.Vx, self.Vy, self.Vz = z(nx + 1, ny, nz), z(nx, ny + 1, nz), z(nx, ny, nz + 1)
        self.txx, self.tyy, self.tzz = z(nx, ny - 2, nz - 2), z(nx - 2, ny, nz - 2), z(nx - 2, ny - 2, nz)
        self.txy, self.txz, self.tyz = z(nx - 1, ny - 1, nz - 2), z(nx - 1, ny - 2, nz - 1), z(nx - 2, ny - 1, nz - 1)
        sx, sy, sz = (nx - 1, ny - 2, nz - 2), (nx - 2, ny - 1, nz - 2), (nx - 2, ny - 2, nz - 1)
        self.Rx, self.Ry, self.Rz = z(*sx), z(*sy), z(*sz)
        self.dVxdt, self.dVydt, self.dVzdt = z(*sx), z(*sy), z(*sz)
        self.dtVx, self.dtVy, self.dtVz = z(*sx), z(*sy), z(*sz)
        self.Mus, self.Rog = z(nx, ny, nz), z(nx, ny, nz)
        s = self
        bc = lambda axis, A, r: self.k_bc[axis].launcher(A, ranges=r)  # noqa: E731
        self._steps = [
            s.k_p.launcher(s.divV, s.Pt, s.Vx, s.Vy, s.Vz, s.dtPt, s.dx, s.dy, s.dz),
            s.k_tau.launcher(s.divV, s.txx, s.tyy, s.tzz, s.txy, s.txz, s.tyz, s.Vx, s.Vy, s.Vz, s.Mus, s.dx, s.dy, s.dz),
            s.k_dv.launcher(s.Rx, s.Ry, s.Rz, s.dVxdt, s.dVydt, s.dVzdt, s.Pt, s.Rog, s.txx, s.tyy, s.tzz, s.txy, s.txz, s.tyz,
                            s.dampX, s.dampY, s.dampZ, s.dx, s.dy, s.dz),
            s.k_v.launcher(s.Vx, s.Vy, s.Vz, s.dVxdt, s.dVydt, s.dVzdt, s.dtVx, s.dtVy, s.dtVz),
            bc("x", s.Vy, ((1, ny + 1), (1, nz))),
            bc("x", s.Vz, ((1, ny), (1, nz + 1))),
            bc("y", s.Vx, ((1, nx + 1), (1, nz))),
            bc("y", s.Vz, ((1, nx), (1, nz + 1))),
            bc("z", s.Vx, ((1, nx + 1), (1, ny))),
            bc("z", s.Vy, ((1, nx), (1, ny + 1))),
        ]
        self._graph = None
        self.iterate(1)  # compile and load every kernel (zeros map to zeros)
        core.synchronize()

    def fields(self):
        """Name -> device array, for comparisons and checkpoints."""
        return {k: v for k, v in vars(self).items() if isinstance(v, core.DeviceArray)}

    def set_material(self, Mus, Rog):
        self.Mus.copy_from(Mus)
        self.Rog.copy_from(Rog)
        self.k_ts(self.dtVx, self.dtVy, self.dtVz, self.dtPt, self.Mus, self.Vsc, self.Ptsc, self.min_dxyz2, self.max_nxyz)

    def reset(self):
        for A in (self.Pt, self.divV, self.Vx, self.Vy, self.Vz, self.txx, self.tyy, self.tzz, self.txy, self.txz, self.tyz,
                  self.Rx, self.Ry, self.Rz, self.dVxdt, self.dVydt, self.dVzdt):
            A.fill_(0.0)

    def iterate(self, n):
        """n PT iterations (miniapps/Stokes3D.jl:133-142), asynchronous."""
        steps = self._steps
        for _ in range(n):
            for st in steps:
                st()

    def iterate_block(self):
        if not self.use_graph:
            self.iterate(self.nout)
            return
        if self._graph is None:
            core.synchronize()
            g = core.Graph()
            with g:
                self.iterate(self.nout)
            self._graph = g
        self._graph.launch()

    def error(self):
        """max(mean |Rx|, mean |Ry|, mean |Rz|, mean |∇V|) (miniapps/Stokes3D.jl:144-147)."""
        return max(A.reduce("mean_abs") for A in (self.Rx, self.Ry, self.Rz, self.divV))

    def solve(self):
        """Iterates until the check every nout iterations is at or below eps (or itermax);
        returns (iterations, error)."""
        it, err = 0, float("inf")
        while it < self.itermax:
            self.iterate_block()
            it += self.nout
            err = self.error()
            if err <= self.eps or not err < 1e30:
                break
        return it, err


# ------------------------------------------------------------------ ThermalConvection2D (miniapps/ThermalConvection2D.jl)
TC2D_ASSIGN = """
@parallel function assign!(A::Data.Array, B::Data.Array)
    @all(A) = @all(B)
    return
end
"""
TC2D_ERROR = """
@parallel function compute_error!(Err_A::Data.Array, A::Data.Array)
    @all(Err_A) = @all(Err_A) - @all(A)
    return
end
"""
TC2D_0 = """
@parallel function compute_0!(RogT::Data.Array, Eta::Data.Array, ∇V::Data.Array, T::Data.Array, Vx::Data.Array, Vy::Data.Array, ρ0gα::Data.Number, η0::Data.Number, dη_dT::Data.Number, ΔT::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(RogT) = ρ0gα*@all(T)
    @all(Eta)  = η0*(1.0 - dη_dT*(@all(T) + ΔT/2.0))
    @all(∇V)   = @d_xa(Vx)/dx + @d_ya(Vy)/dy
    return
end
"""
TC2D_1 = """
@parallel function compute_1!(Pt::Data.Array, τxx::Data.Array, τyy::Data.Array, σxy::Data.Array, Eta::Data.Array, ∇V::Data.Array, Vx::Data.Array, Vy::Data.Array, dτ_iter::Data.Number, β::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(Pt)  = @all(Pt) - dτ_iter/β*@all(∇V)
    @all(τxx) = 2.0*@all(Eta)*(@d_xa(Vx)/dx - 1.0/3.0*@all(∇V))
    @all(τyy) = 2.0*@all(Eta)*(@d_ya(Vy)/dy - 1.0/3.0*@all(∇V))
    @all(σxy) = 2.0*@av(Eta)*(0.5*(@d_yi(Vx)/dy + @d_xi(Vy)/dx))
    return
end
"""
TC2D_2 = """
@parallel function compute_2!(Rx::Data.Array, Ry::Data.Array, dVxdτ::Data.Array, dVydτ::Data.Array, Pt::Data.Array, RogT::Data.Array, τxx::Data.Array, τyy::Data.Array, σxy::Data.Array, ρ::Data.Number, dampX::Data.Number, dampY::Data.Number, dτ_iter::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(Rx)    = 1.0/ρ*(@d_xi(τxx)/dx + @d_ya(σxy)/dy - @d_xi(Pt)/dx)
    @all(Ry)    = 1.0/ρ*(@d_yi(τyy)/dy + @d_xa(σxy)/dx - @d_yi(Pt)/dy + @av_yi(RogT))
    @all(dVxdτ) = dampX*@all(dVxdτ) + @all(Rx)*dτ_iter
    @all(dVydτ) = dampY*@all(dVydτ) + @all(Ry)*dτ_iter
    return
end
"""
TC2D_V = """
@parallel function update_V!(Vx::Data.Array, Vy::Data.Array, dVxdτ::Data.Array, dVydτ::Data.Array, dτ_iter::Data.Number)
    @inn(Vx) = @inn(Vx) + @all(dVxdτ)*dτ_iter
    @inn(Vy) = @inn(Vy) + @all(dVydτ)*dτ_iter
    return
end
"""
TC2D_QT = """
@parallel function compute_qT!(qTx::Data.Array, qTy::Data.Array, T::Data.Array, DcT::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(qTx) = -DcT*@d_xi(T)/dx
    @all(qTy) = -DcT*@d_yi(T)/dy
    return
end
"""
TC2D_ADVECT = """
@parallel_indices (ix,iy) function advect_T!(dT_dt::Data.Array, qTx::Data.Array, qTy::Data.Array, T::Data.Array, Vx::Data.Array, Vy::Data.Array, dx::Data.Number, dy::Data.Number)
    if (ix<=size(dT_dt, 1) && iy<=size(dT_dt, 2)) dT_dt[ix,iy] = -((qTx[ix+1,iy]-qTx[ix,iy])/dx + (qTy[ix,iy+1]-qTy[ix,iy])/dy) -
                                                                  (Vx[ix+1,iy+1]>0)*Vx[ix+1,iy+1]*(T[ix+1,iy+1]-T[ix  ,iy+1])/dx -
                                                                  (Vx[ix+2,iy+1]<0)*Vx[ix+2,iy+1]*(T[ix+2,iy+1]-T[ix+1,iy+1])/dx -
                                                                  (Vy[ix+1,iy+1]>0)*Vy[ix+1,iy+1]*(T[ix+1,iy+1]-T[ix+1,iy  ])/dy -
                                                                  (Vy[ix+1,iy+2]<0)*Vy[ix+1,iy+2]*(T[ix+1,iy+2]-T[ix+1,iy+1])/dy   end
    return
end
"""
TC2D_T = """
@parallel function update_T!(T::Data.Array, T_old::Data.Array, dT_dt::Data.Array, dt::Data.Number)
    @inn(T) = @inn(T_old) + @all(dT_dt)*dt
    return
end
"""
TC2D_NOFLUX = """
@parallel_indices (ix,iy) function no_fluxY_T!(T::Data.Array)
    if (ix==size(T, 1) && iy<=size(T ,2)) T[ix,iy] = T[ix-1,iy] end
    if (ix==1          && iy<=size(T ,2)) T[ix,iy] = T[ix+1,iy] end
    return
end
"""


def _div_inf(a, b):
    """a / b with IEEE semantics for b == 0 (Julia's, where dx / 0.0 is Inf)."""
    return a / b if b != 0.0 else (float("inf") if a > 0 else float("-inf") if a < 0 else float("nan"))


def thermal_convection2d_setup(nx=None, ny=None, ar=3, ly=1.0, eta0=1.0, DcT=1.0, dT=1.0, Ra=1e7, Pra=1e3, itermax=5 * 10 ** 4,
                               nerr=100, eps=1e-4, dmp=2.0):
    """Scalars and initial temperature of miniapps/ThermalConvection2D.jl:90-124 (grid
    (96 ar - 1) x 95 by default): returns (prm, T0) with prm holding nx, ny, dx, dy, rho0ga,
    eta0, deta_dT, dT, DcT, rho, dt_diff, dtau_iter, beta, dampX, dampY, itermax, nerr, eps."""
    import numpy as np
    nx = 96 * ar - 1 if nx is None else nx
    ny = 96 - 1 if ny is None else ny
    lx = ar * ly
    w = 1e-2 * ly
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    mind = min(dx, dy)
    rho = 1.0 / Pra * eta0 / DcT
    dtau_iter = 1.0 / 6.1 * mind / float(np.sqrt(eta0 / rho))
    prm = dict(nx=nx, ny=ny, dx=dx, dy=dy, rho0ga=Ra * eta0 * DcT / dT / ly ** 3, eta0=eta0, deta_dT=1e-10 / dT, dT=dT, DcT=DcT,
               rho=rho, dt_diff=1.0 / 4.1 * mind ** 2 / DcT, dtau_iter=dtau_iter, beta=6.1 * dtau_iter ** 2 / mind ** 2 / rho,
               dampX=1.0 - dmp / nx, dampY=1.0 - dmp / ny, itermax=itermax, nerr=nerr, eps=eps)
    x = (np.arange(nx) * dx - 0.5 * lx) / w
    y = (np.arange(ny) * dy - 0.5 * ly) / w
    T0 = np.asfortranarray(dT * np.exp(-(x * x)[:, None] - (y * y)[None, :]))
    T0[:, 0] = dT / 2.0
    T0[:, -1] = -dT / 2.0
    return prm, T0


class ThermalConvection2D:
    """Thermal convection of miniapps/ThermalConvection2D.jl:88-184 on the GPU: each time step
    runs the pseudo-transient Stokes iteration (10 launches per iteration, :160-169) until the
    relative changes errV and errP checked every nerr iterations (:170-174) are at or below eps,
    then the thermal step (:177-183) with dt = min(dt_diff, dt_adv). The error check and the
    advective time step use FP64 max-|.| reductions on the GPU. `graph=True` replays each block of
    nerr iterations as one HIP graph."""

    def __init__(self, module, nx=None, ny=None, graph=True, **physics):
        prm, T0 = thermal_convection2d_setup(nx, ny, **physics)
        self.prm, self.T0 = prm, T0
        for k, v in prm.items():
            setattr(self, k, v)
        nx, ny = self.nx, self.ny
        self.use_graph = graph
        d = lambda text, kind="parallel": core._define(kind, text, {}, module)  # noqa: E731
        self.k_assign, self.k_err, self.k_0, self.k_1, self.k_2, self.k_v = (d(t) for t in (TC2D_ASSIGN, TC2D_ERROR, TC2D_0, TC2D_1, TC2D_2, TC2D_V))
        self.k_qt, self.k_t = d(TC2D_QT), d(TC2D_T)
        self.k_adv, self.k_noflux = d(TC2D_ADVECT, "parallel_indices"), d(TC2D_NOFLUX, "parallel_indices")
        self.k_bcx, self.k_bcy = d(STOKES2D_BC_X, "parallel_indices"), d(STOKES2D_BC_Y, "parallel_indices")
        z = lambda *s: core.zeros(*s, eltype=core.Float64)  # noqa: E731
        for name in ("T", "T_old", "Pt", "divV", "RogT", "Eta", "txx", "tyy", "ErrP"):
            setattr(self, name, z(nx, ny))
        self.Vx, self.Vy, self.ErrV, self.sxy = z(nx + 1, ny), z(nx, ny + 1), z(nx, ny + 1), z(nx - 1, ny - 1)
        self.Rx, self.Ry, self.dVxdt, self.dVydt = z(nx - 1, ny - 2), z(nx - 2, ny - 1), z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.qTx, self.qTy, self.dT_dt = z(nx - 1, ny - 2), z(nx - 2, ny - 1), z(nx - 2, ny - 2)
        s = self
        self._pt = [
            s.k_assign.launcher(s.ErrV, s.Vy), s.k_assign.launcher(s.ErrP, s.Pt),
            s.k_0.launcher(s.RogT, s.Eta, s.divV, s.T, s.Vx, s.Vy, s.rho0ga, s.eta0, s.deta_dT, s.dT, s.dx, s.dy),
            s.k_1.launcher(s.Pt, s.txx, s.tyy, s.sxy, s.Eta, s.divV, s.Vx, s.Vy, s.dtau_iter, s.beta, s.dx, s.dy),
            s.k_2.launcher(s.Rx, s.Ry, s.dVxdt, s.dVydt, s.Pt, s.RogT, s.txx, s.tyy, s.sxy, s.rho, s.dampX, s.dampY, s.dtau_iter, s.dx, s.dy),
            s.k_v.launcher(s.Vx, s.Vy, s.dVxdt, s.dVydt, s.dtau_iter),
            s.k_bcy.launcher(s.Vx, ranges=((1, nx + 1),)),
            s.k_bcx.launcher(s.Vy, ranges=((1, ny + 1),)),
            s.k_err.launcher(s.ErrV, s.Vy), s.k_err.launcher(s.ErrP, s.Pt),
        ]
        self._graph = None
        self.reset()
        self.iterate(1)  # compile and load every kernel
        self.k_qt(s.qTx, s.qTy, s.T, s.DcT, s.dx, s.dy)
        self.k_adv(s.dT_dt, s.qTx, s.qTy, s.T, s.Vx, s.Vy, s.dx, s.dy)
        self.k_t(s.T, s.T_old, s.dT_dt, 0.0)
        self.k_noflux(s.T)
        self.reset()
        core.synchronize()

    def fields(self):
        return {k: v for k, v in vars(self).items() if isinstance(v, core.DeviceArray)}

    def reset(self):
        """The initial state: the Gaussian temperature anomaly with the hot bottom and cold top
        rows (:121-124), every other field zero."""
        for A in self.fields().values():
            A.fill_(0.0)
        self.T.copy_from(self.T0)

    def iterate(self, n):
        """n PT iterations, asynchronous."""
        steps = self._pt
        for _ in range(n):
            for st in steps:
                st()

    def _block(self):
        if not self.use_graph:
            self.iterate(self.nerr)
            return
        if self._graph is None:
            core.synchronize()
            g = core.Graph()
            with g:
                self.iterate(self.nerr)
            self._graph = g
        self._graph.launch()

    def errors(self):
        """(errV, errP) of :171-172."""
        errV = self.ErrV.reduce("max_abs") / (1e-12 + self.Vy.reduce("max_abs"))
        errP = self.ErrP.reduce("max_abs") / (1e-12 + self.Pt.reduce("max_abs"))
        return errV, errP

    def time_step(self):
        """One time step (:155-183); returns (PT iterations, errV, errP, dt)."""
        s = self
        s.k_assign(s.T_old, s.T, async_=True)
        errV = errP = 2 * s.eps
        niter = 0
        while (errV > s.eps or errP > s.eps) and niter < s.itermax:
            self._block()
            niter += s.nerr
            errV, errP = self.errors()
        s.k_qt(s.qTx, s.qTy, s.T, s.DcT, s.dx, s.dy, async_=True)
        s.k_adv(s.dT_dt, s.qTx, s.qTy, s.T, s.Vx, s.Vy, s.dx, s.dy, async_=True)
        dt_adv = min(_div_inf(s.dx, s.Vx.reduce("max_abs")), _div_inf(s.dy, s.Vy.reduce("max_abs"))) / 2.1
        dt = min(s.dt_diff, dt_adv)
        s.k_t(s.T, s.T_old, s.dT_dt, dt, async_=True)
        s.k_noflux(s.T, async_=True)
        return niter, errV, errP, dt


# ------------------------------------------------------------------ scalar porosity waves (miniapps/scalar_porowaves2D.jl)
PORO_NONLIN = """
@parallel function compute_nonlin!(K_mu::Data.Array, EtaR::Data.Array, Phi::Data.Array, Eta::Data.Array, k_μ0::Data.Number, n::Data.Number)
    @all(K_mu) = k_μ0 * @all(Phi)^n
    @all(EtaR) = @all(Eta)
    return
end
"""
PORO_ETAR = """
@parallel_indices (ix,iy) function compute_EtaR!(EtaR::Data.Array, Pe::Data.Array, ηR::Data.Number)
    if (ix<size(EtaR,1) && iy<size(EtaR,2)) if (Pe[ix+1,iy+1]<0.0) EtaR[ix,iy] = ηR; end;  end
    return
end
"""
PORO_FLUX = """
@parallel function compute_flux!(qx::Data.Array, qy::Data.Array, K_mu::Data.Array, Pe::Data.Array, Δρg::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(qx) = -@av_xi(K_mu)* @d_xi(Pe)/dx
    @all(qy) = -@av_yi(K_mu)*(@d_yi(Pe)/dy + Δρg)
    return
end
"""
PORO_PE_PHI = """
@parallel function compute_Pe_Phi!(dPedt::Data.Array, Pe::Data.Array, Phi::Data.Array, dPhidt::Data.Array, qx::Data.Array, qy::Data.Array, EtaR::Data.Array, β::Data.Number, dt::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(dPedt)  = 1.0/β*(-@d_xa(qx)/dx -@d_ya(qy)/dy -1.0/@all(EtaR)*@inn(Pe))
    @inn(Pe)     = @inn(Pe) + dt*@all(dPedt)
    @all(dPhidt) = -β*@all(dPedt) - 1.0/@all(EtaR)*@inn(Pe)
    @inn(Phi)    = max(0.0, @inn(Phi) + dt*@all(dPhidt))
    return
end
"""


def porowaves2d_setup(nx=256, ny=512, chan=0):
    """Scalars and initial porosity of miniapps/scalar_porowaves2D.jl:36-78 (blob setup by
    default): returns (prm, Phi0). The initial porosity is the background plus the elliptic blob,
    smoothed once by 1/4.1 of its discrete Laplacian in the interior."""
    import numpy as np
    sc = 1.0 + chan * 30
    lx, ly = 250.0 / sc, 500.0 / sc
    eta = 1.0
    lam = 20.0 / sc
    prm = dict(nx=nx, ny=ny, k_mu0=1.0, drhog=1.0, n=3.0, eta=eta, etaR=eta / (1.0 + chan * 1000), beta=0.2,
               dx=lx / (nx - 1), dy=ly / (ny - 1))
    x = (np.arange(nx) * prm["dx"] - 0.5 * lx) / lam / 4.0
    y = (np.arange(ny) * prm["dy"] - 0.2 * ly) / lam
    Radc = (x * x)[:, None] + (y * y)[None, :]
    Phi = np.ones((nx, ny))
    Phi[Radc < 1.0] = Phi[Radc < 1.0] + 2.0
    lap = np.diff(np.diff(Phi[:, 1:-1], axis=0), axis=0) + np.diff(np.diff(Phi[1:-1, :], axis=1), axis=1)
    Phi[1:-1, 1:-1] = Phi[1:-1, 1:-1] + 1.0 / 4.1 * lap
    return prm, np.asfortranarray(Phi)


class Porowaves2D:
    """Scalar porosity waves of miniapps/scalar_porowaves2D.jl:79-97 on the GPU: four launches per
    time step; every 100 steps (when mod(it, 100) == 1) dt = dx^2 / (maximum(K_mu) / β) / 5 from
    an FP64 reduction on the GPU."""

    def __init__(self, module, nx=256, ny=512, chan=0):
        prm, Phi0 = porowaves2d_setup(nx, ny, chan)
        self.prm, self.Phi0 = prm, Phi0
        for k, v in prm.items():
            setattr(self, k, v)
        d = lambda text, kind="parallel": core._define(kind, text, {}, module)  # noqa: E731
        self.k_nonlin, self.k_flux, self.k_pe = d(PORO_NONLIN), d(PORO_FLUX), d(PORO_PE_PHI)
        self.k_etar = d(PORO_ETAR, "parallel_indices")
        z = lambda *s: core.zeros(*s, eltype=core.Float64)  # noqa: E731
        self.Eta, self.EtaR, self.dPedt, self.dPhidt = z(nx - 2, ny - 2), z(nx - 2, ny - 2), z(nx - 2, ny - 2), z(nx - 2, ny - 2)
        self.Pe, self.K_mu, self.Phi = z(nx, ny), z(nx, ny), z(nx, ny)
        self.qx, self.qy = z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.reset()

    def fields(self):
        return {k: v for k, v in vars(self).items() if isinstance(v, core.DeviceArray)}

    def reset(self):
        for A in self.fields().values():
            A.fill_(0.0)
        self.Eta.fill_(self.eta)
        self.Phi.copy_from(self.Phi0)
        self.it, self.dt, self.t = 0, 1.0, 0.0

    def step(self, n=1):
        """n time steps, asynchronous except for the time-step reduction every 100 steps."""
        s = self
        for _ in range(n):
            s.it += 1
            s.k_nonlin(s.K_mu, s.EtaR, s.Phi, s.Eta, s.k_mu0, s.n, async_=True)
            s.k_etar(s.EtaR, s.Pe, s.etaR, async_=True)
            if s.it % 100 == 1:
                s.dt = s.dx ** 2 / (s.K_mu.reduce("max_abs") / s.beta) / 5.0
            s.k_flux(s.qx, s.qy, s.K_mu, s.Pe, s.drhog, s.dx, s.dy, async_=True)
            s.k_pe(s.dPedt, s.Pe, s.Phi, s.dPhidt, s.qx, s.qy, s.EtaR, s.beta, s.dt, s.dx, s.dy, async_=True)
            s.t += s.dt


# ------------------------------------------------------------------ HydroMech2D (miniapps/HydroMech2D.jl)
HM_OLD = """
@parallel function update_old!(Phi_o::Data.Array, ∇V_o::Data.Array, Phi::Data.Array, ∇V::Data.Array)
    @all(Phi_o) = @all(Phi)
    @all(∇V_o)  = @all(∇V)
    return
end
"""
HM_PARAMS = """
@parallel function compute_params_∇!(EtaC::Data.Array, K_muf::Data.Array, Rog::Data.Array, ∇V::Data.Array, ∇qD::Data.Array, Phi::Data.Array, Pf::Data.Array, Pt::Data.Array, Vx::Data.Array, Vy::Data.Array, qDx::Data.Array, qDy::Data.Array, μs::Data.Number, η2μs::Data.Number, R::Data.Number, λPe::Data.Number, k_μf0::Data.Number, ϕ0::Data.Number, nperm::Data.Number, θ_e::Data.Number, θ_k::Data.Number, ρfg::Data.Number, ρsg::Data.Number, ρgBG::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(EtaC)  = (1.0-θ_e)*@all(EtaC)  + θ_e*( μs/@all(Phi)*η2μs*(1.0+0.5*(1.0/R-1.0)*(1.0+tanh((@all(Pf)-@all(Pt))/λPe))) )
    @all(K_muf) = (1.0-θ_k)*@all(K_muf) + θ_k*( k_μf0 * (@all(Phi)/ϕ0)^nperm )
    @all(Rog)   = ρfg*@all(Phi) + ρsg*(1.0-@all(Phi)) - ρgBG
    @all(∇V)    = @d_xa(Vx)/dx  + @d_ya(Vy)/dy
    @all(∇qD)   = @d_xa(qDx)/dx + @d_ya(qDy)/dy
    return
end
"""
HM_RP = """
@parallel function compute_RP!(dτPf::Data.Array, RPt::Data.Array, RPf::Data.Array, K_muf::Data.Array, ∇V::Data.Array, ∇qD::Data.Array, Pt::Data.Array, Pf::Data.Array, EtaC::Data.Array, Phi::Data.Array, Pfsc::Data.Number, Pfdmp::Data.Number, min_dxy2::Data.Number, dx::Data.Number, dy::Data.Number)
    @inn(dτPf) = min_dxy2/@maxloc(K_muf)/4.1/Pfsc
    @all(RPt)  =                 - @all(∇V)  - (@all(Pt) - @all(Pf))/(@all(EtaC)*(1.0-@all(Phi)))
    @all(RPf)  = @all(RPf)*Pfdmp - @all(∇qD) + (@all(Pt) - @all(Pf))/(@all(EtaC)*(1.0-@all(Phi)))
    return
end
"""
HM_PTAU = """
@parallel function compute_P_τ!(Pt::Data.Array, Pf::Data.Array, τxx::Data.Array, τyy::Data.Array, σxy::Data.Array, RPt::Data.Array, RPf::Data.Array, dτPf::Data.Array, Vx::Data.Array, Vy::Data.Array, ∇V::Data.Array, dτPt::Data.Number, μs::Data.Number, β_n::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(Pt)  = @all(Pt) +      dτPt *@all(RPt)
    @all(Pf)  = @all(Pf) + @all(dτPf)*@all(RPf)
    @all(τxx) = 2.0*μs*( @d_xa(Vx)/dx - 1.0/3.0*@all(∇V) - β_n*@all(RPt) )
    @all(τyy) = 2.0*μs*( @d_ya(Vy)/dy - 1.0/3.0*@all(∇V) - β_n*@all(RPt) )
    @all(σxy) = 2.0*μs*(0.5*( @d_yi(Vx)/dy + @d_xi(Vy)/dx ))
    return
end
"""
HM_RES = """
@parallel function compute_res!(Rx::Data.Array, Ry::Data.Array, dVxdτ::Data.Array, dVydτ::Data.Array, τxx::Data.Array, τyy::Data.Array, σxy::Data.Array, Pt::Data.Array, Rog::Data.Array, dampX::Data.Number, dampY::Data.Number, dx::Data.Number, dy::Data.Number)
    @all(Rx)    = @d_xi(τxx)/dx + @d_ya(σxy)/dy - @d_xi(Pt)/dx
    @all(Ry)    = @d_yi(τyy)/dy + @d_xa(σxy)/dx - @d_yi(Pt)/dy - @av_yi(Rog)
    @all(dVxdτ) = dampX*@all(dVxdτ) + @all(Rx)
    @all(dVydτ) = dampY*@all(dVydτ) + @all(Ry)
    return
end
"""
HM_UPDATE = """
@parallel function compute_update!(Vx::Data.Array, Vy::Data.Array, qDx::Data.Array, qDy::Data.Array, Phi::Data.Array, dVxdτ::Data.Array, dVydτ::Data.Array, K_muf::Data.Array, Pf::Data.Array, Phi_o::Data.Array, ∇V::Data.Array, ∇V_o::Data.Array, dτV::Data.Number, ρfg::Data.Number, ρgBG::Data.Number, CN::Data.Number, dt::Data.Number, dx::Data.Number, dy::Data.Number)
    @inn(Vx)  =  @inn(Vx) + dτV*@all(dVxdτ)
    @inn(Vy)  =  @inn(Vy) + dτV*@all(dVydτ)
    @inn(qDx) = -@av_xi(K_muf)*(@d_xi(Pf)/dx)
    @inn(qDy) = -@av_yi(K_muf)*(@d_yi(Pf)/dy + (ρfg - ρgBG))
    @all(Phi) =  @all(Phi_o) + (1.0-@all(Phi))*(CN*@all(∇V_o) + (1.0-CN)*@all(∇V))*dt
    return
end
"""


def hydromech2d_setup(res=128):
    """Scalars and initial fields of miniapps/HydroMech2D.jl:63-150 ((res-1) x (2 res - 1)):
    returns (prm, init) with init holding Phi, EtaC, K_muf and qDy as float64 host arrays."""
    import numpy as np
    rhofg, k_muf0, etaC0, eta2mus, R, nperm, phi0, ra, lam0 = 1.0, 1.0, 1.0, 10.0, 500.0, 3.0, 0.01, 2, 1.0
    rhosg = 2.0 * rhofg
    lx = 20.0
    ly = ra * lx
    nx, ny = res - 1, ra * res - 1
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    mus = etaC0 * phi0 / eta2mus
    lam = lam0 * float(np.sqrt(k_muf0 * etaC0))
    beta_n, Vdmp, Vsc, Ptsc = 1.0, 5.0, 2.0, 2.0
    min_dxy2 = min(dx, dy) ** 2
    prm = dict(nx=nx, ny=ny, dx=dx, dy=dy, rhofg=rhofg, rhosg=rhosg, k_muf0=k_muf0, eta2mus=eta2mus, R=R, nperm=nperm, phi0=phi0,
               lamPe=0.01, mus=mus, rhogBG=rhofg * phi0 + rhosg * (1.0 - phi0), CN=0.5, eps=1e-5, itermax=5000, nout=200,
               beta_n=beta_n, Pfdmp=0.8, Pfsc=4.0, theta_e=9e-1, theta_k=1e-1, dt_red=1e-3, min_dxy2=min_dxy2,
               dtauV=min_dxy2 / mus / (1.0 + beta_n) / 4.1 / Vsc, dtauPt=4.1 * mus * (1.0 + beta_n) / max(nx, ny) / Ptsc,
               dampX=1.0 - Vdmp / nx, dampY=1.0 - Vdmp / ny, dt0=1e-5, t_tot=0.02)
    x = ((np.arange(nx) * dx - 0.5 * lx) / lam / 4.0) ** 2
    y = ((np.arange(ny) * dy - 0.25 * ly) / lam) ** 2
    Radc = x[:, None] + y[None, :]
    Phi = np.full((nx, ny), phi0)
    Phi[Radc < 1.0] = Phi[Radc < 1.0] + 2 * phi0
    phibc = Phi[:, -1].copy()
    qDy = np.zeros((nx, ny + 1))
    qDy[:, 0] = qDy[:, -1] = (rhosg - rhofg) * (1.0 - phibc) * k_muf0 * (phibc / phi0) ** nperm
    init = dict(Phi=Phi, EtaC=mus / Phi * eta2mus, K_muf=k_muf0 * (Phi / phi0), qDy=qDy)
    return prm, {k: np.asfortranarray(v) for k, v in init.items()}


class HydroMech2D:
    """Two-phase hydro-mechanical flow of miniapps/HydroMech2D.jl:150-197 on the GPU: per time
    step update_old! and PT iterations (five compute launches and six boundary launches each,
    :157-169) until norm(Ry)/length(Ry) and norm(RPf)/length(RPf), checked every nout
    iterations, are at or below eps; then dt = dt_red / (1e-10 + maximum(abs.(∇V))). The norms and
    the maximum are FP64 reductions on the GPU; `graph=True` replays blocks of nout iterations."""

    def __init__(self, module, res=128, graph=True):
        prm, init = hydromech2d_setup(res)
        self.prm, self.init = prm, init
        for k, v in prm.items():
            setattr(self, k, v)
        nx, ny = self.nx, self.ny
        self.use_graph = graph
        d = lambda text, kind="parallel": core._define(kind, text, {}, module)  # noqa: E731
        self.k_old, self.k_params, self.k_rp, self.k_ptau, self.k_res, self.k_upd = (
            d(t) for t in (HM_OLD, HM_PARAMS, HM_RP, HM_PTAU, HM_RES, HM_UPDATE))
        self.k_bcx, self.k_bcy = d(STOKES2D_BC_X, "parallel_indices"), d(STOKES2D_BC_Y, "parallel_indices")
        z = lambda *s: core.zeros(*s, eltype=core.Float64)  # noqa: E731
        for name in ("Phi_o", "Pt", "Pf", "Rog", "divV", "divV_o", "divqD", "dtPf", "RPt", "RPf", "txx", "tyy", "Phi", "EtaC", "K_muf"):
            setattr(self, name, z(nx, ny))
        self.sxy = z(nx - 1, ny - 1)
        self.dVxdt, self.dVydt, self.Rx, self.Ry = z(nx - 1, ny - 2), z(nx - 2, ny - 1), z(nx - 1, ny - 2), z(nx - 2, ny - 1)
        self.Vx, self.Vy, self.qDx, self.qDy = z(nx + 1, ny), z(nx, ny + 1), z(nx + 1, ny), z(nx, ny + 1)
        s = self
        self._pt = [
            s.k_params.launcher(s.EtaC, s.K_muf, s.Rog, s.divV, s.divqD, s.Phi, s.Pf, s.Pt, s.Vx, s.Vy, s.qDx, s.qDy, s.mus, s.eta2mus, s.R,
                                s.lamPe, s.k_muf0, s.phi0, s.nperm, s.theta_e, s.theta_k, s.rhofg, s.rhosg, s.rhogBG, s.dx, s.dy),
            s.k_rp.launcher(s.dtPf, s.RPt, s.RPf, s.K_muf, s.divV, s.divqD, s.Pt, s.Pf, s.EtaC, s.Phi, s.Pfsc, s.Pfdmp, s.min_dxy2, s.dx, s.dy),
            s.k_bcx.launcher(s.dtPf, ranges=((1, ny),)),
            s.k_bcy.launcher(s.dtPf, ranges=((1, nx),)),
            s.k_ptau.launcher(s.Pt, s.Pf, s.txx, s.tyy, s.sxy, s.RPt, s.RPf, s.dtPf, s.Vx, s.Vy, s.divV, s.dtauPt, s.mus, s.beta_n, s.dx, s.dy),
            s.k_res.launcher(s.Rx, s.Ry, s.dVxdt, s.dVydt, s.txx, s.tyy, s.sxy, s.Pt, s.Rog, s.dampX, s.dampY, s.dx, s.dy),
            s.k_upd.launcher(s.Vx, s.Vy, s.qDx, s.qDy, s.Phi, s.dVxdt, s.dVydt, s.K_muf, s.Pf, s.Phi_o, s.divV, s.divV_o, s.dtauV, s.rhofg,
                             s.rhogBG, s.CN, s.dt0, s.dx, s.dy),
            s.k_bcy.launcher(s.Vx, ranges=((1, nx + 1),)),
            s.k_bcx.launcher(s.Vy, ranges=((1, ny + 1),)),
            s.k_bcy.launcher(s.qDx, ranges=((1, nx + 1),)),
            s.k_bcx.launcher(s.qDy, ranges=((1, ny + 1),)),
        ]
        self._graph = None
        self.reset()
        self.iterate(1)  # compile and load every kernel
        self.k_old(s.Phi_o, s.divV_o, s.Phi, s.divV)
        self.reset()
        core.synchronize()

    def fields(self):
        return {k: v for k, v in vars(self).items() if isinstance(v, core.DeviceArray)}

    def reset(self):
        for A in self.fields().values():
            A.fill_(0.0)
        for k, v in self.init.items():
            getattr(self, k).copy_from(v)
        self.dt, self.t, self.it = self.dt0, 0.0, 1
        self._set_dt(self.dt0)

    def _set_dt(self, dt):
        """The physical time step is an argument of compute_update!: replaced in the prepared
        launch (the graph is captured again when it changes)."""
        self._pt[6].set(16, dt)
        self._graph = None

    def iterate(self, n):
        steps = self._pt
        for _ in range(n):
            for st in steps:
                st()

    def _block(self):
        if not self.use_graph:
            self.iterate(self.nout)
            return
        if self._graph is None:
            core.synchronize()
            g = core.Graph()
            with g:
                self.iterate(self.nout)
            self._graph = g
        self._graph.launch()

    def error(self):
        """max(norm(Ry)/length(Ry), norm(RPf)/length(RPf)) (:171)."""
        n = lambda A: A.shape[0] * A.shape[1]  # noqa: E731
        return max(self.Ry.reduce("norm") / n(self.Ry), self.RPf.reduce("norm") / n(self.RPf))

    def time_step(self):
        """One time step (:151-197); returns (PT iterations, error, dt used)."""
        s = self
        s.k_old(s.Phi_o, s.divV_o, s.Phi, s.divV, async_=True)
        err, niter = 2 * s.eps, 0
        while err > s.eps and niter < s.itermax:
            s._block()
            niter += s.nout
            err = s.error()
        dt_used = s.dt
        s.dt = s.dt_red / (1e-10 + s.divV.reduce("max_abs"))
        s.t += s.dt
        s.it += 1
        s._set_dt(s.dt)
        return niter, err, dt_used


# ------------------------------------------------------------------ acoustic waves
def acoustic3d_inputs(n, lx=40.0, k=1.0, rho=1.0):
    """Initial state of miniapps/acoustic3D.jl:24-40 on n^3 pressure nodes (dx = lx/(n-1)):
    P0 = exp(-((ix-1)dx - lx/2)^2 - ...), V0 = 0, dt = min(dx,dy,dz)/sqrt(k/rho)/6.1."""
    import numpy as np
    dx = lx / (n - 1)
    x = np.arange(n) * dx - 0.5 * lx
    X2 = x * x
    P = np.empty((n, n, n), order="F")
    for kz in range(n):
        # exp(-a^2 - b^2 - c^2), left to right as the miniapp writes it
        P[:, :, kz] = np.exp(((-X2[:, None]) - X2[None, :]) - X2[kz])
    dt = min(dx, dx, dx) / np.sqrt(k / rho) / 6.1
    return P, dict(dt=dt, rho=rho, k=k, dx=dx, dy=dx, dz=dx)


def acoustic2d_inputs(n, lx=40.0, k=1.0, rho=1.0):
    """Initial state of miniapps/acoustic2D.jl (same construction as the 3-D case)."""
    import numpy as np
    dx = lx / (n - 1)
    x = np.arange(n) * dx - 0.5 * lx
    P = np.asfortranarray(np.exp((-(x * x)[:, None]) - (x * x)[None, :]))
    dt = min(dx, dx) / np.sqrt(k / rho) / 6.1
    return P, dict(dt=dt, rho=rho, k=k, dx=dx, dy=dx)


def define_acoustic3d(module, memopt=False, **kw):
    """(compute_V!, compute_P!) of miniapps/acoustic3D.jl:11-21."""
    kw = dict(kw, memopt=memopt)
    return core._define("parallel", ACOUSTIC3D_V, kw, module), core._define("parallel", ACOUSTIC3D_P, kw, module)


def define_acoustic2d(module, memopt=False, **kw):
    """(compute_V!, compute_P!) of miniapps/acoustic2D.jl:11-20."""
    kw = dict(kw, memopt=memopt)
    return core._define("parallel", ACOUSTIC2D_V, kw, module), core._define("parallel", ACOUSTIC2D_P, kw, module)


def acoustic3d_run(module, P0, prm, steps, eltype, memopt=False, **kw):
    """Runs `steps` acoustic time steps (V then P, miniapps/acoustic3D.jl:48-52) on the GPU from
    P0 (float64 host array) in storage type `eltype` and returns host copies (P, Vx, Vy, Vz) as
    float64 (BFloat16 widened); `kw` are kernel keywords (fastmath, cells_x, ...)."""
    n1, n2, n3 = P0.shape
    kV, kP = define_acoustic3d(module, memopt, **kw)
    P = core.zeros(n1, n2, n3, eltype=eltype)
    Vx, Vy, Vz = (core.zeros(n1 + 1, n2, n3, eltype=eltype), core.zeros(n1, n2 + 1, n3, eltype=eltype),
                  core.zeros(n1, n2, n3 + 1, eltype=eltype))
    P.copy_from(_to_storage(P0, eltype))
    for _ in range(steps):
        kV(Vx, Vy, Vz, P, prm["dt"], prm["rho"], prm["dx"], prm["dy"], prm["dz"], async_=True)
        kP(P, Vx, Vy, Vz, prm["dt"], prm["k"], prm["dx"], prm["dy"], prm["dz"], async_=True)
    core.synchronize()
    return tuple(_from_storage(A.to_host(), eltype) for A in (P, Vx, Vy, Vz))


def _to_storage(a, eltype):
    import numpy as np
    if eltype == core.BFloat16:
        return np.asfortranarray(core.f32_to_bf16(np.asarray(a, dtype=np.float32)))
    if eltype == core.Float32:
        return np.asfortranarray(a, dtype=np.float32)
    return np.asfortranarray(a, dtype=np.float64)


def _from_storage(a, eltype):
    import numpy as np
    if eltype == core.BFloat16:
        return np.asfortranarray(core.bf16_to_f32(a).astype(np.float64))
    return np.asfortranarray(a, dtype=np.float64)


def field_errors(got, ref):
    """(max |got - ref| / max |ref|, ||got - ref||_2 / ||ref||_2) of one field."""
    import numpy as np
    d = got - ref
    mref = float(np.abs(ref).max()) or 1.0
    nref = float(np.linalg.norm(ref.ravel())) or 1.0
    return float(np.abs(d).max()) / mref, float(np.linalg.norm(d.ravel())) / nref
