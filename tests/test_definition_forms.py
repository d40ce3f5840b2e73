"""CPU tests of kernel-definition forms: N substitution and ndims tuple expansion
(test/test_parallel.jl:268-300 and 1346-1410), @parallel_indices (I...) splat indices with tuple
broadcasts (test/test_parallel.jl:1283-1325, lowered here), @init_parallel_stencil without
numbertype and ndims, and runtime hardware selection (test/ParallelKernel/test_select_hardware.jl)."""
import re

import pytest

import parallelstencil as ps

MOD = __name__
SIG = "f(A::Data.Array{T,N}, B::Data.Array{T,N}, c::Integer) where T <: PSNumber = (@all(A) = @all(B)^c; return)"
SIG_I = "f(A::Data.Array{T,N}, B::Data.Array{T,N}, c::Integer) where T <: PSNumber = (A[I...] = B[I...]^c; return)"


def _mod(suffix, *a, **kw):
    m = MOD + suffix
    ps.reset_parallel_stencil(module=m)
    ps.init_parallel_stencil(ps.AMDGPU, *a, module=m, **kw)
    return m


def _sig(k):
    return re.sub(r"\s+", "", k.text.split("=", 1)[0] if "function" not in k.text else k.text)


@pytest.mark.parametrize("N,want", [(3, 3), (2, 2), (1, 1), ("ndims", 2), ("ndims+2", 4)])
def test_N_substitution(N, want):
    m = _mod(".N")  # numbertype and ndims omitted
    k = ps.core._define("parallel", SIG, {"ndims": 2, "N": N}, m)
    assert f"f(A::Data.Array{{T,{want}}},B::Data.Array{{T,{want}}}," in _sig(k)
    ki = ps.core._define("parallel_indices", SIG_I, {"ndims": 2, "N": N}, m, indices=["I..."])
    assert f"f(A::Data.Array{{T,{want}}},B::Data.Array{{T,{want}}}," in _sig(ki)
    assert ki.info["nb_parallel_indices"] == 2


def test_N_substitution_in_the_text_prefix():
    m = _mod(".Np", ps.Float64, 3)
    k = ps.core._define("parallel", "@parallel N=3 " + SIG.replace("{T,N}", "{N}"), {}, m)
    assert "f(A::Data.Array{3},B::Data.Array{3}," in _sig(k)
    k = ps.core._define("parallel", "@parallel ndims=2 N=ndims+2 " + SIG, {}, m)
    assert "{T,4}" in _sig(k) and k.info["nb_parallel_indices"] == 2
    k = ps.core._define("parallel_indices", "@parallel_indices I... ndims=2 N=ndims " + SIG_I, {}, m)
    assert "{T,2}" in _sig(k) and k.info["nb_parallel_indices"] == 2


def test_ndims_tuple_expansion():
    m = _mod(".nt")
    km = ps.core._define("parallel", SIG, {"ndims": (1, 2, 3), "N": "ndims"}, m)
    assert isinstance(km, ps.KernelMethods) and set(km.methods) == {1, 2, 3}
    for d, (n, k) in km.methods.items():
        assert n == d and f"{{T,{d}}}" in _sig(k) and k.info["nb_parallel_indices"] == d
    km = ps.core._define("parallel", "@parallel ndims=(1,3) N=ndims.+1 " + SIG, {}, m)
    assert set(km.methods) == {1, 3}
    assert "{T,2}" in _sig(km.methods[1][1]) and "{T,4}" in _sig(km.methods[3][1])
    ki = ps.core._define("parallel_indices", SIG_I, {"ndims": (1, 3), "N": "ndims.+1"}, m, indices=["I..."])
    assert {d: k.info["nb_parallel_indices"] for d, (n, k) in ki.methods.items()} == {1: 1, 3: 3}
    # dispatch on the number of dimensions of the array arguments
    A2, A4 = ps.ArgSpec((4, 5)), ps.ArgSpec((4, 5, 6))
    assert km.method(A2, A2, 2) is km.methods[1][1]
    with pytest.raises(ps.ArgumentError, match="no method"):
        km.method(A4, A4, 2)


def test_N_and_ndims_errors():
    m = _mod(".ne")
    with pytest.raises(ps.KeywordArgumentError, match="'N' is mandatory"):
        ps.core._define("parallel", SIG, {"ndims": (1, 2)}, m)
    with pytest.raises(ps.KeywordArgumentError, match="same length as 'ndims'"):
        ps.core._define("parallel", SIG, {"ndims": (1, 2), "N": "ndims+1"}, m)
    with pytest.raises(ps.KeywordArgumentError, match="'ndims' is invalid or missing"):
        ps.core._define("parallel", SIG, {"N": 3}, m)
    with pytest.raises(ps.IncoherentArgumentError, match="where clause"):
        ps.core._define("parallel", SIG.replace("where T <: PSNumber", "where {T, N}"), {"ndims": 2, "N": 2}, m)
    with pytest.raises(ps.ArgumentError, match="'eltype' is mandatory"):
        ps.core._numbertype_of(m)


def test_splat_indices_lowering():
    m = _mod(".I", ps.Float64, 2)
    k = ps.core._define("parallel_indices", """
    @parallel_indices (I...) function write_indices!(A, one)
        A[I...] = sum((I .- (1,)) .* (one, size(A,1)));
        return
    end""", {}, m)
    assert k.info["nb_parallel_indices"] == 2
    src = k.source(ps.ArgSpec((4, 30)), 1.0)
    # the folded expression: (ix-1)*one + (iy-1)*size(A,1)
    assert "a0_nx" in src and "ix" in src and "iy" in src
    m3 = _mod(".I3", ps.Float64, 3)
    k3 = ps.core._define("parallel_indices", """
    @parallel_indices (I...) function w!(A, one)
        A[I...] = sum((I .- (1,)) .* (one, size(A,1), size(A,1)*size(A,2))) + length(I) - 3 + maximum(I .* 0)
        return
    end""", {}, m3)
    assert k3.info["nb_parallel_indices"] == 3
    with pytest.raises(ps.ArgumentError, match="different lengths"):
        ps.core._define("parallel_indices", """
        @parallel_indices (I...) function w!(A)
            A[I...] = sum(I .+ (1, 2, 3))
            return
        end""", {}, m)


def test_select_hardware():
    m = _mod(".hw", ps.Float64, 3)
    assert ps.get_hardware(module=m) == "gpu_amd" and ps.current_hardware(module=m) == "gpu_amd"
    ps.select_hardware(":gpu_amd", module=m)
    assert ps.current_hardware(module=m) == "gpu_amd"
    for bad in ("cpu", "gpu_cuda", "gpu_metal", "gpu_oneapi"):
        with pytest.raises(ps.ArgumentError, match="unsupported hardware selection"):
            ps.select_hardware(bad, module=m)
        assert ps.current_hardware(module=m) == "gpu_amd"
    with pytest.raises(ps.ArgumentError):
        ps.handle("cpu", ps.AMDGPU)
    ps.reset_parallel_stencil(module=m)
    with pytest.raises(ps.ArgumentError, match="has not been initialized"):
        ps.select_hardware("gpu_amd", module=m)
