"""GPU tests of @parallel_indices (I...) (test/test_parallel.jl:1283-1325) and of kernels
defined for several ndims (N substitution, test/test_parallel.jl:268-300): results are index
values and integer powers, exact."""
import numpy as np
import pytest

import parallelstencil as ps

pytestmark = pytest.mark.gpu

MOD = __name__


def _mod(suffix, nd):
    m = MOD + suffix
    if not ps.is_initialized(m):
        ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, nd, module=m)
    return m


def test_splat_indices_1d():
    m = _mod(".1", 1)
    k = ps.core._define("parallel_indices", """
    @parallel_indices (I...) function write_indices!(A, one)
        A[I...] = sum((I .- (1,)) .* (one));
        return
    end""", {}, m)
    A = ps.zeros(4 * 5 * 6)
    k(A, 1.0)
    assert np.array_equal(A.to_host(), np.arange(4 * 5 * 6, dtype=float))


def test_splat_indices_2d():
    m = _mod(".2", 2)
    k = ps.core._define("parallel_indices", """
    @parallel_indices (I...) function write_indices!(A, one)
        A[I...] = sum((I .- (1,)) .* (one, size(A,1)));
        return
    end""", {}, m)
    A = ps.zeros(4, 5 * 6)
    k(A, 1.0)
    ix, iy = np.meshgrid(np.arange(1, 5), np.arange(1, 31), indexing="ij")
    assert np.array_equal(A.to_host(), (ix - 1) + (iy - 1) * 4.0)


def test_splat_indices_3d():
    m = _mod(".3", 3)
    k = ps.core._define("parallel_indices", """
    @parallel_indices (I...) function write_indices!(A, one)
        A[I...] = sum((I .- (1,)) .* (one, size(A,1), size(A,1)*size(A,2)));
        return
    end""", {}, m)
    A = ps.zeros(4, 5, 6)
    k(A, 1.0)
    want = np.arange(4 * 5 * 6, dtype=float).reshape((4, 5, 6), order="F")
    assert np.array_equal(A.to_host(), want)


def test_ndims_tuple_methods_dispatch():
    m = _mod(".nt", 3)
    km = ps.core._define("parallel", "@parallel ndims=(1,2,3) N=ndims f(A::Data.Array{T,N}, B::Data.Array{T,N}, c::Integer) "
                                     "where T <: PSNumber = (@all(A) = @all(B)^c; return)", {}, m)
    rng = np.random.default_rng(3)
    for shape in ((37,), (11, 7), (5, 6, 7)):
        hB = np.asfortranarray(rng.integers(-5, 6, shape).astype(float))
        A, B = ps.zeros(*shape), ps.zeros(*shape).copy_from(hB)
        ps.parallel(km, A, B, 3)
        assert np.array_equal(A.to_host(), hB ** 3)
