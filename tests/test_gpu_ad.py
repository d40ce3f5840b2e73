"""GPU tests of derivative kernels (@parallel ∇=(A->Ā, ...) f!(...), ad_mode, ad_annotations and
the generic AD.autodiff_deferred! form), test/ParallelKernel/test_parallel.jl:224-283.

Expected derivatives never come from a derivative formula: the reference's closed form for
A[ix] += a*B[ix]*100.65, a dense Jacobian built from the oracle's primal diffusion step applied
to unit vectors, and complex-step derivatives of numpy restatements of the test kernels."""
import numpy as np
import pytest

import parallelstencil as ps
from oracle import ps_oracle as O
from parallelstencil import workloads as W

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=MOD)

F_REF = """
@parallel_indices (ix) function f!(A, B, a)
    A[ix] += a * B[ix] * 100.65
    return
end"""


def _dev(h):
    d = ps.zeros(*h.shape)
    return d.copy_from(np.asfortranarray(h))


def _close(got, want, rtol=1e-12):
    scale = max(1.0, float(np.max(np.abs(want)))) if np.size(want) else 1.0
    np.testing.assert_allclose(got, want, rtol=rtol, atol=rtol * scale)


# ---------------------------------------------------------------- reference closed form
def test_reference_ad_numerical():
    """test/ParallelKernel/test_parallel.jl:236-283: Ā stays 1, B̄ = 1 + a*100.65."""
    print("ad: reference closed form N=16", flush=True)
    N, a = 16, 6.5
    ps.seed(3)
    k = ps.core._define("parallel_indices", F_REF, {}, MOD)
    A, B = ps.rand(N), ps.rand(N)
    hA, hB = A.to_host(), B.to_host()
    Ab, Bb = ps.ones(N), ps.ones(N)
    k(A, B, a, **{"∇": [(A, Ab), (B, Bb)]})
    assert np.allclose(Ab.to_host(), np.ones(N), rtol=1e-15)
    assert np.allclose(Bb.to_host(), 1 + a * 100.65, rtol=1e-15)
    assert np.array_equal(A.to_host(), hA)  # DuplicatedNoNeed: primal update dropped
    # the generic form, with the scalar promoted to Const
    Ab, Bb = ps.ones(N), ps.ones(N)
    ps.AD.autodiff_deferred_(ps.Reverse, k, ps.AD.DuplicatedNoNeed(A, Ab), ps.AD.DuplicatedNoNeed(B, Bb), ps.AD.Const(a))
    assert np.allclose(Ab.to_host(), 1.0, rtol=1e-15) and np.allclose(Bb.to_host(), 1 + a * 100.65, rtol=1e-15)
    Ab, Bb = ps.ones(N), ps.ones(N)
    ps.AD.autodiff_deferred_(ps.Reverse, k, ps.AD.DuplicatedNoNeed(A, Ab), ps.AD.DuplicatedNoNeed(B, Bb), a)
    assert np.allclose(Bb.to_host(), 1 + a * 100.65, rtol=1e-15)
    # ∇=B->B̄ only: A is Const, so its update is kept and nothing reaches B̄
    Bb = ps.ones(N)
    k(A, B, a, grad=(B, Bb))
    assert np.array_equal(Bb.to_host(), np.ones(N))
    assert np.array_equal(A.to_host(), hA + a * hB * 100.65)
    # Duplicated keeps the primal update as well
    A.copy_from(hA)
    Ab, Bb = ps.ones(N), ps.zeros(N)
    k(A, B, a, grad={"A": Ab, "B": Bb}, ad_annotations={"Duplicated": ("A", "B")})
    assert np.array_equal(A.to_host(), hA + a * hB * 100.65)
    assert np.allclose(Bb.to_host(), a * 100.65, rtol=1e-15)
    with pytest.raises(ps.ArgumentError, match="explicit insertion of the return type activity"):
        ps.AD.autodiff_deferred_(ps.Reverse, k, ps.AD.Const, ps.AD.DuplicatedNoNeed(A, Ab), ps.AD.DuplicatedNoNeed(B, Bb), a)


def test_reference_ad_forward_mode():
    """ad_mode=Enzyme.Forward: Ȧ = Ȧ + a*100.65*Ḃ."""
    N, a = 16, 6.5
    k = ps.core._define("parallel_indices", F_REF, {}, MOD)
    rng = np.random.default_rng(4)
    hA, hB, hAd, hBd = (rng.standard_normal(N) for _ in range(4))
    A, B, Ad, Bd = _dev(hA), _dev(hB), _dev(hAd), _dev(hBd)
    k(A, B, a, **{"∇": [(A, Ad), (B, Bd)]}, ad_mode="Enzyme.Forward")
    _close(Ad.to_host(), hAd + a * 100.65 * hBd, rtol=1e-15)
    assert np.array_equal(Bd.to_host(), hBd) and np.array_equal(A.to_host(), hA)


# ---------------------------------------------------------------- headline kernel (linear)
def _diffusion_case(shape, seed):
    rng = np.random.default_rng(seed)
    T = np.asfortranarray(rng.standard_normal(shape))
    Cp = np.asfortranarray(1.0 + rng.random(shape))
    return T, Cp, dict(lam=0.7, dt=0.05, dx=0.9, dy=1.1, dz=1.3)


@pytest.mark.parametrize("form", ["gather", "scatter"])
@pytest.mark.parametrize("memopt", [False, True])
def test_diffusion3d_reverse_against_oracle_jacobian(memopt, form):
    """The diffusion step is linear in T: its Jacobian is the oracle's step applied to unit
    vectors. Reverse mode must give T̄ += Jᵀ T̄2 (inner points) and zero T̄2 where T2 is written,
    in the gather form (two kernels, no atomics) and in the atomic scatter form."""
    print(f"ad: diffusion3D reverse memopt={memopt} form={form}", flush=True)
    shape = (9, 8, 7)
    T, Cp, c = _diffusion_case(shape, 11)
    n = T.size

    def step(x):
        out = np.zeros(shape, order="F")
        O.diffusion3d_flux(out, np.asfortranarray(x.reshape(shape, order="F")), Cp, c["lam"], c["dt"], c["dx"], c["dy"], c["dz"])
        return out

    J = O.linear_jacobian(step, n)
    inner = np.zeros(shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    rng = np.random.default_rng(12)
    y = np.asfortranarray(rng.standard_normal(shape))
    z = np.asfortranarray(rng.standard_normal(shape))
    k = W.define_diffusion3d(memopt, MOD)
    d = k._derivative("Reverse", {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})
    assert d._gather is not None
    d.use_gather = form == "gather"
    T2d, Td, Cpd = _dev(T), _dev(T), _dev(Cp)
    T2b, Tb = _dev(y), _dev(z)
    k(T2d, Td, Cpd, c["lam"], c["dt"], c["dx"], c["dy"], c["dz"], grad=[(T2d, T2b), (Td, Tb)])
    d.use_gather = True
    want_Tb = z.ravel(order="F") + J.T @ np.where(inner, y, 0.0).ravel(order="F")
    _close(Tb.to_host().ravel(order="F"), want_Tb)
    assert np.array_equal(T2b.to_host(), np.where(inner, 0.0, y))
    assert np.array_equal(T2d.to_host(), T)  # T2 DuplicatedNoNeed: not updated


def test_diffusion3d_forward_against_oracle_jacobian():
    shape = (9, 8, 7)
    T, Cp, c = _diffusion_case(shape, 13)

    def step(x):
        out = np.zeros(shape, order="F")
        O.diffusion3d_flux(out, np.asfortranarray(x.reshape(shape, order="F")), Cp, c["lam"], c["dt"], c["dx"], c["dy"], c["dz"])
        return out

    J = O.linear_jacobian(step, T.size)
    inner = np.zeros(shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    rng = np.random.default_rng(14)
    x = np.asfortranarray(rng.standard_normal(shape))
    w = np.asfortranarray(rng.standard_normal(shape))
    k = W.define_diffusion3d(True, MOD)
    T2d, Td, Cpd, T2t, Tt = _dev(np.zeros(shape)), _dev(T), _dev(Cp), _dev(w), _dev(x)
    k(T2d, Td, Cpd, c["lam"], c["dt"], c["dx"], c["dy"], c["dz"], grad=[(T2d, T2t), (Td, Tt)], ad_mode=ps.Forward,
      ad_annotations={"Duplicated": T2d})
    Jx = (J @ x.ravel(order="F")).reshape(shape, order="F")
    _close(T2t.to_host(), np.where(inner, Jx, w))
    # Duplicated T2: the primal update is kept and equals the primal kernel's, bit for bit
    P = _dev(np.zeros(shape))
    k(P, Td, Cpd, c["lam"], c["dt"], c["dx"], c["dy"], c["dz"])
    assert np.array_equal(T2d.to_host(), P.to_host())


def test_diffusion3d_adjoint_dot_product_full_size():
    """At a production size the adjoint identity <J x, y> = <x, Jᵀ y> ties forward and reverse
    mode together (size-independent property; 1e-12 relative)."""
    n = 128
    print(f"ad: diffusion3D dot-product test {n}^3", flush=True)
    T, Cp, c = W.diffusion_inputs(n)
    rng = np.random.default_rng(15)
    x = np.asfortranarray(rng.standard_normal(T.shape))
    y = np.asfortranarray(rng.standard_normal(T.shape))
    k = W.define_diffusion3d(True, MOD)
    T2d, Td, Cpd = _dev(T), _dev(T), _dev(Cp)
    args = (T2d, Td, Cpd, c["lam"], c["dt"], c["dx"], c["dy"], c["dz"])
    Jx, Xt = _dev(np.zeros(T.shape)), _dev(x)
    k(*args, grad=[(T2d, Jx), (Td, Xt)], ad_mode=ps.Forward)
    Yb, Tb = _dev(y), _dev(np.zeros(T.shape))
    k(*args, grad=[(T2d, Yb), (Td, Tb)])
    inner = np.zeros(T.shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    lhs = float(np.sum(np.where(inner, Jx.to_host(), 0.0) * y))
    rhs = float(np.sum(x * Tb.to_host()))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))


# ---------------------------------------------------------------- nonlinear kernel
NL = """
@parallel_indices (ix, iy) function nl!(R, U, V, K, s)
    if ix > 1 && ix < size(U,1) && iy > 1 && iy < size(U,2)
        u  = U[ix,iy]
        gx = (U[ix+1,iy] - U[ix-1,iy]) * 0.5
        k  = exp(-u*u) * K[ix,iy] + sqrt(V[ix,iy]^2 + 1.0)
        w  = u > 0.3 ? log(1.0 + u*u) : tanh(u)
        R[ix,iy] = k*gx/(1.0 + abs(u)) + s*max(w, V[ix,iy-1]) + u^3
    else
        R[ix,iy] = 2.0*U[ix,iy] - min(V[ix,iy], 0.5)
    end
    return
end"""


def nl_ref(U, V, K, s):
    """numpy restatement of nl! with analytic operations (comparisons on real parts)."""
    R = np.zeros(U.shape, dtype=np.result_type(U, V, K))
    u = U[1:-1, 1:-1]
    gx = (U[2:, 1:-1] - U[:-2, 1:-1]) * 0.5
    kk = np.exp(-u * u) * K[1:-1, 1:-1] + np.sqrt(V[1:-1, 1:-1] ** 2 + 1.0)
    w = np.where(u.real > 0.3, np.log(1.0 + u * u), np.tanh(u))
    au = np.where(u.real >= 0, u, -u)
    vm = V[1:-1, :-2]
    mx = np.where(w.real >= vm.real, w, vm)
    R[...] = 2.0 * U - np.where(V.real <= 0.5, V, 0.5)
    R[1:-1, 1:-1] = kk * gx / (1.0 + au) + s * mx + u ** 3
    return R


def _nl_inputs(shape, seed):
    rng = np.random.default_rng(seed)
    return (np.asfortranarray(rng.uniform(-1, 1, shape)), np.asfortranarray(rng.uniform(-1, 1.5, shape)),
            np.asfortranarray(rng.uniform(0.5, 2, shape)), 0.8)


def _init(suffix, numbertype, ndims):
    if not ps.is_initialized(MOD + suffix):
        ps.init_parallel_stencil(ps.AMDGPU, numbertype, ndims, module=MOD + suffix)


def _nl_kernel():
    _init(".nl", ps.Float64, 2)
    return ps.core._define("parallel_indices", NL, {}, MOD + ".nl")


def test_nonlinear_forward_against_complex_step():
    shape = (12, 10)
    U, V, K, s = _nl_inputs(shape, 21)
    rng = np.random.default_rng(22)
    dU, dV, dK = (np.asfortranarray(rng.standard_normal(shape)) for _ in range(3))
    k = _nl_kernel()
    R = _dev(np.zeros(shape))
    Rt = _dev(np.zeros(shape))
    Ud, Vd, Kd = _dev(U), _dev(V), _dev(K)
    Ut, Vt, Kt = _dev(dU), _dev(dV), _dev(dK)
    k(R, Ud, Vd, Kd, s, grad=[(R, Rt), (Ud, Ut), (Vd, Vt), (Kd, Kt)], ad_mode=ps.Forward)
    want = O.complex_step_jvp(lambda a, b, c: nl_ref(a, b, c, s), [U, V, K], [dU, dV, dK])
    _close(Rt.to_host(), want)
    assert np.array_equal(R.to_host(), np.zeros(shape))


@pytest.mark.parametrize("form", ["gather", "scatter"])
def test_nonlinear_reverse_against_complex_step(form):
    """@parallel_indices kernel with locals and branches: the gather form (locals inlined,
    branch conditions shifted to the source point) and the atomic scatter form."""
    print(f"ad: nonlinear reverse form={form}", flush=True)
    shape = (12, 10)
    U, V, K, s = _nl_inputs(shape, 23)
    rng = np.random.default_rng(24)
    y = np.asfortranarray(rng.standard_normal(shape))
    zU, zV, zK = (np.asfortranarray(rng.standard_normal(shape)) for _ in range(3))
    k = _nl_kernel()
    R, Ud, Vd, Kd = _dev(np.zeros(shape)), _dev(U), _dev(V), _dev(K)
    Rb, Ub, Vb, Kb = _dev(y), _dev(zU), _dev(zV), _dev(zK)
    d = k._derivative("Reverse", {"R": "Duplicated", "U": "DuplicatedNoNeed", "V": "DuplicatedNoNeed", "K": "DuplicatedNoNeed"})
    assert d._gather is not None
    d.use_gather = form == "gather"
    k(R, Ud, Vd, Kd, s, grad={"R": Rb, "U": Ub, "V": Vb, "K": Kb}, ad_annotations={"Duplicated": "R", "Active": "s"})
    d.use_gather = True
    gU, gV, gK = O.complex_step_vjp(lambda a, b, c: nl_ref(a, b, c, s), [U, V, K], [0, 1, 2], y)
    _close(Ub.to_host(), zU + gU)
    _close(Vb.to_host(), zV + gV)
    _close(Kb.to_host(), zK + gK)
    assert np.array_equal(Rb.to_host(), np.zeros(shape))
    # Duplicated R: the kept primal update equals the primal kernel's, bit for bit
    P = _dev(np.zeros(shape))
    k(P, Ud, Vd, Kd, s)
    assert np.array_equal(R.to_host(), P.to_host())
    _close(P.to_host(), nl_ref(U, V, K, s))


# ---------------------------------------------------------------- loops
ACC = """
@parallel_indices (ix) function acc!(A, B, C)
    s = 0.0
    for k = 1:size(B,2)
        t = B[ix, k] * C[ix, k]
        s += t * t
    end
    A[ix] = sqrt(s) + (ix > 2 ? B[ix,1] : C[ix,2])
    return
end"""

OVERWRITE = """
@parallel_indices (ix) function ow!(A, B)
    for k = 1:3
        A[ix] = B[ix, k] * 2.0
    end
    return
end"""


def acc_ref(B, C):
    t = B * C
    ix = np.arange(1, B.shape[0] + 1)
    return np.sqrt(np.sum(t * t, axis=1)) + np.where(ix > 2, B[:, 0], C[:, 1])


def test_loop_accumulation_both_modes():
    n = 37
    rng = np.random.default_rng(31)
    B, C = (np.asfortranarray(rng.uniform(0.5, 1.5, (n, 4))) for _ in range(2))
    _init(".acc", ps.Float64, 1)
    k = ps.core._define("parallel_indices", ACC, {}, MOD + ".acc")
    A, Bd, Cd = _dev(np.zeros(n)), _dev(B), _dev(C)
    y = rng.standard_normal(n)
    Ab, Bb, Cb = _dev(y), _dev(np.zeros((n, 4))), _dev(np.zeros((n, 4)))
    k(A, Bd, Cd, grad=[(A, Ab), (Bd, Bb), (Cd, Cb)], ranges=range(1, n + 1))
    gB, gC = O.complex_step_vjp(acc_ref, [B, C], [0, 1], y)
    _close(Bb.to_host(), gB)
    _close(Cb.to_host(), gC)
    dB, dC = (np.asfortranarray(rng.standard_normal((n, 4))) for _ in range(2))
    At, Bt, Ct = _dev(np.zeros(n)), _dev(dB), _dev(dC)
    k(A, Bd, Cd, grad=[(A, At), (Bd, Bt), (Cd, Ct)], ad_mode=ps.Forward, ranges=range(1, n + 1))
    _close(At.to_host(), O.complex_step_jvp(acc_ref, [B, C], [dB, dC]))


def test_loop_overwrite_takes_the_last_iteration():
    """A[ix] is written in every iteration; only the last write survives, so only B[:, 3] gets
    an adjoint (the reverse sweep runs the iterations backwards)."""
    n = 20
    rng = np.random.default_rng(32)
    B = np.asfortranarray(rng.standard_normal((n, 3)))
    y = rng.standard_normal(n)
    _init(".ow", ps.Float64, 1)
    k = ps.core._define("parallel_indices", OVERWRITE, {}, MOD + ".ow")
    A, Bd, Ab, Bb = _dev(np.zeros(n)), _dev(B), _dev(y), _dev(np.zeros((n, 3)))
    k(A, Bd, grad=[(A, Ab), (Bd, Bb)], ranges=range(1, n + 1))
    want = np.zeros((n, 3))
    want[:, 2] = 2.0 * y
    assert np.array_equal(Bb.to_host(), want)
    assert np.array_equal(Ab.to_host(), np.zeros(n))


@pytest.mark.parametrize("form", ["gather", "scatter"])
def test_float32_stencil_adjoint(form):
    """Float32 arrays (atomics on 32-bit shadows in the scatter form); against the float64
    Jacobian (1e-5)."""
    _init(".f32", ps.Float32, 2)
    k = ps.core._define("parallel", """
    @parallel function lap!(B, A, c)
        @inn(B) = c*(@d2_xi(A) + @d2_yi(A))
        return
    end""", {}, MOD + ".f32")
    shape = (17, 13)
    rng = np.random.default_rng(33)
    y = np.asfortranarray(rng.standard_normal(shape).astype(np.float32))
    c = 0.25

    def lap(x):
        a = x.reshape(shape, order="F")
        o = np.zeros(shape)
        o[1:-1, 1:-1] = c * ((a[2:, 1:-1] - 2 * a[1:-1, 1:-1] + a[:-2, 1:-1]) + (a[1:-1, 2:] - 2 * a[1:-1, 1:-1] + a[1:-1, :-2]))
        return o

    J = O.linear_jacobian(lap, int(np.prod(shape)))
    B, A = ps.zeros(*shape, eltype=ps.Float32), ps.zeros(*shape, eltype=ps.Float32)
    Bb = ps.zeros(*shape, eltype=ps.Float32).copy_from(y)
    Ab = ps.zeros(*shape, eltype=ps.Float32)
    d = k._derivative("Reverse", {"B": "DuplicatedNoNeed", "A": "DuplicatedNoNeed"})
    d.use_gather = form == "gather"
    k(B, A, c, grad=[(B, Bb), (A, Ab)])
    d.use_gather = True
    inner = np.zeros(shape, dtype=bool)
    inner[1:-1, 1:-1] = True
    want = (J.T @ np.where(inner, y, 0).astype(np.float64).ravel(order="F")).reshape(shape, order="F")
    np.testing.assert_allclose(Ab.to_host(), want, rtol=1e-5, atol=1e-5)


def test_staggered_grids_gather_scatter_and_oracle():
    """Arrays of different sizes (the staggered velocities of miniapps/acoustic2D.jl). With the
    default ranges the gather form runs, with explicit ranges the scatter form; both against the
    dense Jacobian of the oracle's compute_V! (zero initial velocities). The kernel of the
    miniapp itself reads Vx and Vy before updating them (in-place updates): its seeds pass
    through unchanged (dVx_new/dVx_old = 1), in the gather form and in the scatter form."""
    _init(".ac", ps.Float64, 2)
    grad_p = ps.core._define("parallel", """
    @parallel function grad_P!(Vx, Vy, P, dt, rho, dx, dy)
        @inn(Vx) = -dt/rho*@d_xi(P)/dx
        @inn(Vy) = -dt/rho*@d_yi(P)/dy
        return
    end""", {}, MOD + ".ac")
    compute_v = ps.core._define("parallel", W.ACOUSTIC2D_V, {}, MOD + ".ac")
    nx, ny, sc = 23, 17, (0.1, 2.0, 0.5, 0.7)
    rng = np.random.default_rng(41)
    hP = np.asfortranarray(rng.standard_normal((nx, ny)))
    yx = np.asfortranarray(rng.standard_normal((nx + 1, ny)))
    yy = np.asfortranarray(rng.standard_normal((nx, ny + 1)))

    def step(p):
        Vx = np.zeros((nx + 1, ny), order="F")
        Vy = np.zeros((nx, ny + 1), order="F")
        O.acoustic2d_v(Vx, Vy, np.asfortranarray(p.reshape((nx, ny), order="F")), *sc)
        return np.concatenate([Vx.ravel(order="F"), Vy.ravel(order="F")])

    J = O.linear_jacobian(step, nx * ny)
    inx = np.zeros((nx + 1, ny), dtype=bool)
    inx[1:-1, 1:-1] = True
    iny = np.zeros((nx, ny + 1), dtype=bool)
    iny[1:-1, 1:-1] = True
    want = J.T @ np.concatenate([np.where(inx, yx, 0).ravel(order="F"), np.where(iny, yy, 0).ravel(order="F")])
    annots = {"P": "DuplicatedNoNeed", "Vx": "DuplicatedNoNeed", "Vy": "DuplicatedNoNeed"}
    assert grad_p._derivative("Reverse", annots)._gather is not None
    assert compute_v._derivative("Reverse", annots)._gather is not None  # in-place update: the gather form applies
    for k, explicit in ((grad_p, False), (grad_p, True), (compute_v, False), (compute_v, True)):
        print(f"ad: staggered {k.name} explicit_ranges={explicit}", flush=True)
        Vx, Vy, P = _dev(np.zeros((nx + 1, ny))), _dev(np.zeros((nx, ny + 1))), _dev(hP)
        Vxb, Vyb, Pb = _dev(yx), _dev(yy), _dev(np.zeros((nx, ny)))
        kw = {"ranges": ((1, nx + 1), (1, ny + 1))} if explicit else {}
        k(Vx, Vy, P, *sc, grad=[(P, Pb), (Vx, Vxb), (Vy, Vyb)], **kw)
        _close(Pb.to_host().ravel(order="F"), want)
        if k is compute_v:
            assert np.array_equal(Vxb.to_host(), yx) and np.array_equal(Vyb.to_host(), yy)
        else:
            assert np.array_equal(Vxb.to_host(), np.where(inx, 0.0, yx)) and np.array_equal(Vyb.to_host(), np.where(iny, 0.0, yy))


def test_named_tuple_arguments_async_and_cellarray_error():
    """∇ through NamedTuple arguments (the components are differentiated; test_parallel.jl:230
    differentiates V.x, V.y) launched with @parallel_async, and the error for CellArrays."""
    import collections
    _init(".nt", ps.Float64, 2)
    k = ps.core._define("parallel_indices", """
    @parallel_indices (ix, iy) function sq!(R, V, a)
        R[ix, iy] = a * V.x[ix, iy] * V.y[ix, iy]
        return
    end""", {}, MOD + ".nt")
    VT = collections.namedtuple("VT", "x y")
    shape = (19, 11)
    rng = np.random.default_rng(51)
    hx, hy, y = (np.asfortranarray(rng.standard_normal(shape)) for _ in range(3))
    V = VT(_dev(hx), _dev(hy))
    Vb = VT(_dev(np.zeros(shape)), _dev(np.zeros(shape)))
    R, Rb = _dev(np.zeros(shape)), _dev(y)
    ps.parallel_async(k, R, V, 1.5, grad={"R": Rb, "V": Vb})
    ps.synchronize()
    _close(Vb.x.to_host(), 1.5 * y * hy)
    _close(Vb.y.to_host(), 1.5 * y * hx)
    assert np.array_equal(Rb.to_host(), np.zeros(shape))
    C = ps.zeros(*shape, celldims=(2,))
    Cb = ps.zeros(*shape, celldims=(2,))
    kc = ps.core._define("parallel_indices", """
    @parallel_indices (ix, iy) function c!(R, C)
        R[ix, iy] = C[ix, iy][1]
        return
    end""", {}, MOD + ".nt")
    with pytest.raises(ps.ArgumentError, match="CellArray"):
        kc(R, C, grad=[(C, Cb)])


def test_branch_local_gather_matches_complex_step():
    """A local set in both branches (merged as an ifelse in the gather form)."""
    _init(".br", ps.Float64, 2)
    k = ps.core._define("parallel_indices", """
    @parallel_indices (ix, iy) function g!(R, U, V)
        if ix > 1 && ix < size(U,1)
            a = U[ix+1,iy] - U[ix-1,iy]*U[ix,iy]
        else
            a = 2.0*U[ix,iy]
        end
        R[ix,iy] = a*V[ix,iy]
        return
    end""", {}, MOD + ".br")

    def ref(U, V):
        a = 2.0 * U
        a[1:-1, :] = U[2:, :] - U[:-2, :] * U[1:-1, :]
        return a * V

    shape = (13, 9)
    rng = np.random.default_rng(61)
    U, V, y = (np.asfortranarray(rng.standard_normal(shape)) for _ in range(3))
    for form in ("gather", "scatter"):
        d = k._derivative("Reverse", {"R": "DuplicatedNoNeed", "U": "DuplicatedNoNeed", "V": "DuplicatedNoNeed"})
        d.use_gather = form == "gather"
        R, Ud, Vd = _dev(np.zeros(shape)), _dev(U), _dev(V)
        Rb, Ub, Vb = _dev(y), _dev(np.zeros(shape)), _dev(np.zeros(shape))
        k(R, Ud, Vd, grad=[(R, Rb), (Ud, Ub), (Vd, Vb)])
        d.use_gather = True
        gU, gV = O.complex_step_vjp(ref, [U, V], [0, 1], y)
        _close(Ub.to_host(), gU)
        _close(Vb.to_host(), gV)


def test_untaken_branch_derivatives_stay_out():
    """The derivative of the untaken branch of a selection is not evaluated: sqrt of a negative
    value there would make 0 * NaN = NaN, and a neighbour read at the array's edge an
    out-of-bounds access (the fixture's synchronisation after each test would report it)."""
    _init(".nan", ps.Float64, 1)
    k = ps.core._define("parallel_indices", """
    @parallel_indices (ix) function h!(R, X)
        R[ix] = X[ix] > 0.0 ? sqrt(X[ix]) : (ix > 1 ? X[ix-1]*X[ix] : 0.0)
        return
    end""", {}, MOD + ".nan")
    hX = np.array([-1.0, 4.0, -2.0, 9.0, -0.5, 0.25])
    y = np.arange(1.0, 7.0)
    want = np.zeros(6)
    for i in range(6):  # by hand: d sqrt(x) = 1/(2 sqrt x); d (x[i-1] x[i]) by the product rule
        if hX[i] > 0:
            want[i] += y[i] * 0.5 / np.sqrt(hX[i])
        elif i > 0:
            want[i - 1] += y[i] * hX[i]
            want[i] += y[i] * hX[i - 1]
    for form in ("gather", "scatter"):
        d = k._derivative("Reverse", {"R": "DuplicatedNoNeed", "X": "DuplicatedNoNeed"})
        d.use_gather = form == "gather"
        R, X, Rb, Xb = _dev(np.zeros(6)), _dev(hX), _dev(y), _dev(np.zeros(6))
        k(R, X, grad=[(R, Rb), (X, Xb)])
        d.use_gather = True
        ps.synchronize()
        assert np.all(np.isfinite(Xb.to_host()))
        _close(Xb.to_host(), want)
    R, X, Rt, Xt = _dev(np.zeros(6)), _dev(hX), _dev(np.zeros(6)), _dev(np.ones(6))
    k(R, X, grad=[(R, Rt), (X, Xt)], ad_mode=ps.Forward)
    t = Rt.to_host()
    assert np.all(np.isfinite(t))
    _close(t, np.where(hX > 0, 0.5 / np.sqrt(np.abs(hX)), np.r_[0.0, hX[:-1] + hX[1:]]))


def test_memopt_parallel_indices_reverse_both_forms():
    """A memopt @parallel_indices 3-D kernel (stencil offsets on both sides, a variable
    coefficient): its gather kernel is a memopt kernel with launch-box guards on the streamed
    dimension. Against the dense Jacobian of a numpy restatement (linear in A)."""
    _init(".mo", ps.Float64, 3)
    k = ps.core._define("parallel_indices", """
    @parallel_indices (ix,iy,iz) memopt=true function st!(A2, A, B)
        if ix > 1 && ix < size(A,1) && iy > 1 && iy < size(A,2) && iz > 1 && iz < size(A,3)
            A2[ix,iy,iz] = A[ix,iy,iz+1] - 2.0*A[ix,iy,iz] + A[ix,iy,iz-1] + B[ix,iy,iz]*A[ix+1,iy,iz] - A[ix,iy-1,iz]
        end
        return
    end""", {}, MOD + ".mo")
    shape = (16, 8, 16)
    rng = np.random.default_rng(71)
    hB = np.asfortranarray(rng.standard_normal(shape))
    y = np.asfortranarray(rng.standard_normal(shape))

    def st(a):
        A = a.reshape(shape, order="F")
        out = np.zeros(shape)
        s = (slice(1, -1),) * 3
        out[s] = (A[1:-1, 1:-1, 2:] - 2.0 * A[1:-1, 1:-1, 1:-1] + A[1:-1, 1:-1, :-2] + hB[s] * A[2:, 1:-1, 1:-1]
                  - A[1:-1, :-2, 1:-1])
        return out

    J = O.linear_jacobian(st, int(np.prod(shape)))
    inner = np.zeros(shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    want = (J.T @ np.where(inner, y, 0.0).ravel(order="F")).reshape(shape, order="F")
    d = k._derivative("Reverse", {"A2": "DuplicatedNoNeed", "A": "DuplicatedNoNeed"})
    assert d._gather is not None and d._gather[0].info["memopt"]
    for form in ("gather", "scatter"):
        d.use_gather = form == "gather"
        A2, A, B = _dev(np.zeros(shape)), _dev(np.zeros(shape)), _dev(hB)
        A2b, Ab = _dev(y), _dev(np.zeros(shape))
        k(A2, A, B, grad=[(A2, A2b), (A, Ab)])
        d.use_gather = True
        _close(Ab.to_host(), want)
        assert np.array_equal(A2b.to_host(), np.where(inner, 0.0, y))
