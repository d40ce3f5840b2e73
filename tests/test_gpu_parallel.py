"""GPU execution KATs of @parallel / @parallel_indices (test/test_parallel.jl:402-990)
and of the diffusion workload against the oracle."""
import numpy as np
import pytest

import parallelstencil as ps
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=__name__)


def _dev(a):
    A = ps.zeros(*a.shape, eltype=ps.Float64 if a.dtype == np.float64 else ps.Float32)
    return A.copy_from(a)


def test_write_indices():
    # test/test_parallel.jl:402-410
    A = ps.zeros(4, 5, 6)
    k = ps.parallel("""
    @parallel function write_indices!(A)
        @all(A) = $ix + ($iy-1)*size(A,1) + ($iz-1)*size(A,1)*size(A,2);
        return
    end""")
    k(A)
    assert np.array_equal(A.to_host(), O.linear_indices((4, 5, 6)))


OTF = """
@parallel function diffusion3D_step!(T2, T, Ci, lam::Data.Number, dt::Float64, _dx, _dy, _dz)
    @all(qx)   = -lam*@d_xi(T)*_dx                                          # Fourier's law of heat conduction
    @all(qy)   = -lam*@d_yi(T)*_dy                                          # ...
    @all(qz)   = -lam*@d_zi(T)*_dz                                          # ...
    @all(dTdt) = @inn(Ci)*(-@d_xa(qx)*_dx - @d_ya(qy)*_dy - @d_za(qz)*_dz)  # Conservation of energy
    @inn(T2)   = @inn(T) + dt*@all(dTdt)                                    # Update of temperature
    return
end
"""


@pytest.mark.parametrize("memopt", [False, True])
def test_onthefly_diffusion_kat(memopt):
    # test/test_parallel.jl:411-434 (integer-cubed input, unit constants, exact)
    nx, ny, nz = 32, 8, 8
    k = ps.parallel(OTF, memopt=memopt)
    T = O.cubed_indices((nx, ny, nz))
    Ci = np.asfortranarray(np.ones((nx, ny, nz)))
    T2d = ps.zeros(nx, ny, nz)
    k(T2d, _dev(T), _dev(Ci), 1.0, 1.0, 1.0, 1.0, 1.0)
    ref = np.asfortranarray(np.zeros_like(T))
    O.diffusion3d_otf_test(ref, T, Ci, 1.0, 1.0, 1.0, 1.0, 1.0)
    assert np.array_equal(T2d.to_host(), ref)


DIFF = """
@parallel function diffusion3D_step!(T2, T, Cp, lam, dt, dx, dy, dz)
    @all(qTx) = -lam*@d_xi(T)/dx
    @all(qTy) = -lam*@d_yi(T)/dy
    @all(qTz) = -lam*@d_zi(T)/dz
    @inn(T2)  = @inn(T) + dt/@inn(Cp)*(-@d_xa(qTx)/dx - @d_ya(qTy)/dy - @d_za(qTz)/dz)
    return
end
"""


@pytest.mark.parametrize("memopt", [False, True])
def test_diffusion_64_100_steps_bitexact(memopt):
    # 3-D heat diffusion, 64^3 cells, 100 explicit steps, Dirichlet boundaries: bit-exact vs the oracle
    n, nt = 64, 100
    T0, Cp, p = O.diffusion_inputs(n, seed=1)
    k = ps.parallel(DIFF, memopt=memopt)
    T, T2, Cpd = _dev(T0), _dev(T0), _dev(Cp)
    for _ in range(nt):
        k(T2, T, Cpd, p["lam"], p["dt"], p["dx"], p["dy"], p["dz"])
        T, T2 = T2, T
    a, b = np.asfortranarray(T0.copy()), np.asfortranarray(T0.copy())
    for _ in range(nt):
        O.diffusion3d_flux(b, a, Cp, p["lam"], p["dt"], p["dx"], p["dy"], p["dz"])
        a, b = b, a
    got = T.to_host()
    assert np.array_equal(got, a), f"max abs diff {np.abs(got - a).max()}"


@pytest.mark.parametrize("shape", [(32, 8, 8), (31, 7, 9), (33, 9, 9), (65, 5, 40)])
def test_diffusion_ragged_shapes(shape):
    # ragged sizes: partial blocks in x, y and z, memopt chunks not dividing nz
    T0 = np.asfortranarray(np.random.default_rng(3).random(shape))
    Cp = np.asfortranarray(1.0 + np.random.default_rng(4).random(shape))
    ref = np.asfortranarray(T0.copy())
    O.diffusion3d_flux(ref, T0, Cp, 1.0, 1e-3, 0.1, 0.1, 0.1)
    for memopt in (False, True):
        k = ps.parallel(DIFF, memopt=memopt, loopsize=7)
        T2 = _dev(T0)
        k(T2, _dev(T0), _dev(Cp), 1.0, 1e-3, 0.1, 0.1, 0.1)
        assert np.array_equal(T2.to_host(), ref), (shape, memopt)


def test_triad():
    n = (64, 33, 17)
    rng = np.random.default_rng(2)
    B, C = np.asfortranarray(rng.random(n)), np.asfortranarray(rng.random(n))
    k = ps.parallel("@parallel copy3D!(T2, T, Ci) = (@all(T2) = @all(T) + @all(Ci); return)")
    A = ps.zeros(*n)
    k(A, _dev(B), _dev(C))
    ref = np.asfortranarray(np.zeros(n))
    O.triad(ref, B, C)
    assert np.array_equal(A.to_host(), ref)


# --------------------------------------------------------------- memopt KATs (test_parallel.jl:438-990)
def _cubed_dev(shape):
    return _dev(O.cubed_indices(shape))


MEMOPT_CASES = [
    ("copy_memopt!", "@parallel_indices (ix,iy,iz) memopt=true loopsize=3 optvars=A optranges=(A=(0:0,0:0,0:0),) function copy_memopt!(A2, A)\n A2[ix,iy,iz] = A[ix,iy,iz]\n return\nend",
     lambda A: A.copy()),
    ("d2z", "@parallel_indices (ix,iy,iz) memopt=true loopsize=3 function d2_memopt!(A2, A)\n if (iz>1 && iz<size(A2,3))\n A2[ix,iy,iz] = A[ix,iy,iz+1] - 2*A[ix,iy,iz] + A[ix,iy,iz-1]\n end\n return\nend",
     lambda A: _pad_z(A[:, :, 2:] - 2 * A[:, :, 1:-1] + A[:, :, :-2])),
    ("d2y", "@parallel_indices (ix,iy,iz) memopt=true function d2_memopt!(A2, A)\n if (iy>1 && iy<size(A2,2))\n A2[ix,iy,iz] = A[ix,iy+1,iz] - 2*A[ix,iy,iz] + A[ix,iy-1,iz]\n end\n return\nend",
     lambda A: _pad_y(A[:, 2:, :] - 2 * A[:, 1:-1, :] + A[:, :-2, :])),
    ("d2zi", "@parallel memopt=true loopsize=3 function d2_memopt!(A2, A)\n @inn(A2) = @d2_zi(A)\n return\nend",
     lambda A: _pad_all(O.fd3("d2_zi", A))),
    ("d2xi", "@parallel memopt=true loopsize=3 function d2_memopt!(A2, A)\n @inn(A2) = @d2_xi(A)\n return\nend",
     lambda A: _pad_all(O.fd3("d2_xi", A))),
    ("higher_order", "@parallel_indices (ix,iy,iz) memopt=true loopsize=3 function higher_order_memopt!(A2, A)\n if (ix-4>1 && ix-1<size(A2,1) && iy+2>1 && iy+2<=size(A2,2) && iz-2>=1 && iz+3<=size(A2,3))\n A2[ix-1,iy+2,iz] = A[ix-1,iy+2,iz+3] - 2*A[ix-3,iy+2,iz] + A[ix-4,iy+2,iz-2]\n end\n return\nend",
     None),
]


def _pad_z(x):
    out = np.zeros((x.shape[0], x.shape[1], x.shape[2] + 2))
    out[:, :, 1:-1] = x
    return out


def _pad_y(x):
    out = np.zeros((x.shape[0], x.shape[1] + 2, x.shape[2]))
    out[:, 1:-1, :] = x
    return out


def _pad_all(x):
    out = np.zeros(tuple(s + 2 for s in x.shape))
    out[1:-1, 1:-1, 1:-1] = x
    return out


@pytest.mark.parametrize("case", MEMOPT_CASES, ids=[c[0] for c in MEMOPT_CASES])
def test_memopt_kats(case):
    name, text, ref_fn = case
    nxyz = (32, 8, 8)
    A = O.cubed_indices(nxyz)
    if text.startswith("@parallel_indices"):
        k = ps.parallel_indices(text)
    else:
        k = ps.parallel(text)
    A2 = ps.zeros(*nxyz)
    k(A2, _dev(A))
    if ref_fn is None:  # higher order: test_parallel.jl:565-579
        ref = np.zeros(nxyz)
        ref[4:-1, 2:, 2:-3] = A[4:-1, 2:, 5:] - 2 * A[2:-3, 2:, 2:-3] + A[1:-4, 2:, 0:-5]
    else:
        ref = ref_fn(A)
    assert np.array_equal(A2.to_host(), ref)


def test_parallel_indices_bc_kernels_2d():
    mod = __name__ + ".bc2d"
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=mod)
    # miniapps/Stokes2D.jl:44-54 with the unused index dropped: ParallelStencil v0.16 rejects
    # @parallel_indices kernels that leave a parallel index unused (src/parallel.jl:304,706)
    bcx = ps.core._define("parallel_indices", "@parallel_indices (iy) function bc_x!(A::Data.Array)\n A[1  , iy] = A[2    , iy]\n A[end, iy] = A[end-1, iy]\n return\nend", {}, mod)
    bcy = ps.core._define("parallel_indices", "@parallel_indices (ix) function bc_y!(A::Data.Array)\n A[ix, 1  ] = A[ix, 2    ]\n A[ix, end] = A[ix, end-1]\n return\nend", {}, mod)
    with pytest.raises(ps.ArgumentError, match="all parallel indices must be used"):
        ps.core._define("parallel_indices", "@parallel_indices (ix,iy) function bc_x!(A)\n A[1, iy] = A[2, iy]\n return\nend", {}, mod)
    a = np.asfortranarray(O.linear_indices((9, 6))[:, :, 0])
    for k, f, r in ((bcx, O.bc_x, ((1, 6),)), (bcy, O.bc_y, ((1, 9),))):
        A = ps.zeros(9, 6)
        A.copy_from(a)
        k(A, ranges=r)
        ref = a.copy(order="F")
        f(ref)
        assert np.array_equal(A.to_host(), ref)


def test_out_of_bounds_access_raises_bounds_error():
    # an @parallel_indices kernel that reads past the end: the host cannot exclude it, so the
    # checked variant runs, suppresses the access and reports a BoundsError
    k = ps.parallel_indices("@parallel_indices (ix,iy,iz) function shift!(A2, A)\n A2[ix,iy,iz] = A[ix+1,iy,iz]\n return\nend")
    A = ps.ones(8, 4, 4)
    A2 = ps.zeros(8, 4, 4)
    with pytest.raises(ps.BoundsError):
        k(A2, A)
    k(A2, A, ranges=((1, 7), (1, 4), (1, 4)))  # in-bounds launch proves safe and runs unchecked
    r = A2.to_host()
    assert np.all(r[:7] == 1.0)


def test_hide_communication_kat():
    # test/ParallelKernel/test_hide_communication.jl:117-124
    k = ps.parallel_indices("@parallel_indices (ix,iy,iz) function add_indices!(A)\n A[ix,iy,iz] = A[ix,iy,iz] + ix + (iy-1)*size(A,1) + (iz-1)*size(A,1)*size(A,2)\n return\nend")
    halo = ps.parallel_indices("""@parallel_indices (iy,iz) function halo_x!(A)
        A[1, iy, iz] = A[size(A,1)-1, iy, iz]
        A[size(A,1), iy, iz] = A[2, iy, iz]
        return
    end""")
    halo_y = ps.parallel_indices("""@parallel_indices (ix,iz) function halo_y!(A)
        A[ix, 1, iz] = A[ix, size(A,2)-1, iz]
        A[ix, size(A,2), iz] = A[ix, 2, iz]
        return
    end""")
    halo_z = ps.parallel_indices("""@parallel_indices (ix,iy) function halo_z!(A)
        A[ix, iy, 1] = A[ix, iy, size(A,3)-1]
        A[ix, iy, size(A,3)] = A[ix, iy, 2]
        return
    end""")
    A = ps.zeros(6, 7, 8)

    def communication():
        halo(A, ranges=((1, 7), (1, 8)))
        halo_y(A, ranges=((1, 6), (1, 8)))
        halo_z(A, ranges=((1, 6), (1, 7)))

    ps.hide_communication((2, 2, 3), [(k, (A,))], communication)
    ref = O.linear_indices((6, 7, 8))
    ref[0] = ref[-2]
    ref[-1] = ref[1]
    ref[:, 0] = ref[:, -2]
    ref[:, -1] = ref[:, 1]
    ref[:, :, 0] = ref[:, :, -2]
    ref[:, :, -1] = ref[:, :, 1]
    assert np.array_equal(A.to_host(), ref)


# --------------------------------------------------------------- nested functions (ParallelKernel/test_parallel.jl:368-433)
NESTED = {
    "long_modify": """@parallel_indices (ix,iy,iz) function write_indices!(A)
    function compute_indices!(A)
        A[ix,iy,iz] = ix + (iy-1)*size(A,1) + (iz-1)*size(A,1)*size(A,2);
        return
    end
    compute_indices!(A)
    return
end""",
    "short_modify": """@parallel_indices (ix,iy,iz) function write_indices!(A)
    compute_indices!(A) = (A[ix,iy,iz] = ix + (iy-1)*size(A,1) + (iz-1)*size(A,1)*size(A,2); return)
    compute_indices!(A)
    return
end""",
    "long_value": """@parallel_indices (ix,iy,iz) function write_indices!(A)
    function compute_indices(A)
        return ix + (iy-1)*size(A,1) + (iz-1)*size(A,1)*size(A,2)
    end
    A[ix,iy,iz] = compute_indices(A)
    return
end""",
    "short_value": """@parallel_indices (ix,iy,iz) function write_indices!(A)
    compute_indices(A) = return ix + (iy-1)*size(A,1) + (iz-1)*size(A,1)*size(A,2)
    A[ix,iy,iz] = compute_indices(A)
    return
end""",
}


@pytest.mark.parametrize("form", sorted(NESTED))
def test_nested_functions(form):
    A = ps.zeros(4, 5, 6)
    ps.parallel_indices(NESTED[form])(A)
    assert np.array_equal(A.to_host(), O.linear_indices((4, 5, 6)))


def test_write_indices_1d_2d():
    # ParallelKernel/test_parallel.jl:368-386
    A1, A2 = ps.zeros(4), ps.zeros(4, 5)
    ps.parallel_indices("@parallel_indices (ix) function write_indices!(A)\n    A[ix] = ix;\n    return\nend")(A1)
    ps.parallel_indices("@parallel_indices (ix,iy) function write_indices!(A)\n    A[ix,iy] = ix + (iy-1)*size(A,1);\n    return\nend")(A2)
    assert np.array_equal(A1.to_host(), np.arange(1, 5, dtype=float))
    assert np.array_equal(A2.to_host(), O.linear_indices((4, 5)).reshape(4, 5, order="F"))


# --------------------------------------------------------------- tuples and NamedTuples of arrays, @∀
def test_named_array_tuple_example():
    # examples/multimemcopyNamedArrayTuple3D.jl:10-15 (staggered components)
    import collections
    n = 17
    rng = np.random.default_rng(12)
    k = ps.parallel("""@parallel function copy3D!(V2, V, dV)
    @all(V2.x) = @all(V.x) + @all(dV.x)
    @all(V2.y) = @all(V.y) + @all(dV.y)
    @all(V2.z) = @all(V.z) + @all(dV.z)
    return
end""")
    shapes = {"x": (n + 1, n, n), "y": (n, n + 1, n), "z": (n, n, n + 1)}
    h = {c: np.asfortranarray(rng.integers(-99, 99, s).astype(float)) for c, s in shapes.items()}
    d = {c: np.asfortranarray(rng.integers(-99, 99, s).astype(float)) for c, s in shapes.items()}
    V = {c: _dev(h[c]) for c in shapes}
    NT = collections.namedtuple("NT", "x y z")
    dV = NT(*(_dev(d[c]) for c in "xyz"))
    V2 = {c: ps.zeros(*shapes[c]) for c in shapes}
    k(V2, V, dV)
    for c in shapes:
        assert np.array_equal(V2[c].to_host(), h[c] + d[c])
    assert ps.get_ranges(V2, V) == (ps.UnitRange(1, n + 1),) * 3  # maxsize over the components


def test_forall_over_names_ranges_and_tuples():
    # kernel_language.jl:197-237 (@∀ forms) with NamedTuple and plain-tuple arguments
    import collections
    n = (9, 8, 7)
    rng = np.random.default_rng(13)
    k = ps.parallel_indices("""@parallel_indices (ix,iy,iz) function f!(C, A, B, T)
    @∀ i ∈ (x,z) C.i[ix,iy,iz] = A.i[ix,iy,iz] + B.i[ix,iy,iz]
    @∀ i ∈ 1:2 T[i][ix,iy,iz] = T[i][ix,iy,iz] * 2.0
    @∀ (ij,i,j) ∈ ((x,x,y), (z,y,z)) C.ij[ix,iy,iz] = C.ij[ix,iy,iz] + A.i[ix,iy,iz] * B.j[ix,iy,iz]
    return
end""")
    NT = collections.namedtuple("NT", "x y z")
    ha = NT(*(rng.integers(-9, 9, n).astype(float) for _ in range(3)))
    hb = NT(*(rng.integers(-9, 9, n).astype(float) for _ in range(3)))
    ht = [rng.integers(-9, 9, n).astype(float) for _ in range(2)]
    A, B = NT(*map(_dev, ha)), NT(*map(_dev, hb))
    C = NT(*(ps.zeros(*n) for _ in range(3)))
    T = (_dev(ht[0]), _dev(ht[1]))
    k(C, A, B, T)
    assert np.array_equal(C.x.to_host(), ha.x + hb.x + ha.x * hb.y)
    assert np.array_equal(C.y.to_host(), np.zeros(n))
    assert np.array_equal(C.z.to_host(), ha.z + hb.z + ha.y * hb.z)
    assert np.array_equal(T[0].to_host(), 2 * ht[0]) and np.array_equal(T[1].to_host(), 2 * ht[1])
