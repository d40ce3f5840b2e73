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


def define_triad3d(module):
    return core._define("parallel", TRIAD3D, {}, module)
