"""Host-side behaviour without a GPU: the C-ABI library exports every declared symbol, kernel
definition (parsing, FD-macro lowering, memopt analysis and metadata as in
test/test_parallel.jl:1418-1528), definition-time errors, launch geometry helpers, the
@hide_communication range split and the FieldAllocators sizes (test_allocators.jl:899-960)."""
import ctypes
import itertools
import os
import re

import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import FieldAllocators as FA
from parallelstencil._lib import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=__name__)
MOD2 = __name__ + ".2d"
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD2)


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ps_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", ["ps_amd.h", "ps_nn.h"])
def test_library_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(LIB_PATH)
    names = _declared(header)
    assert len(names) >= (30 if header == "ps_amd.h" else 20)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared in include/{header} but not exported: {missing}"


def test_version_and_no_device_errors():
    lib = ctypes.CDLL(LIB_PATH)
    assert lib.ps_version() > 0
    # launching or allocating without an initialised device fails with a typed error, never silently
    with pytest.raises((ps.NoDeviceError, ps.NotInitializedError)):
        ps.zeros(4, 4, 4)


def _k(text, **kw):
    return ps.core._define("parallel_indices" if text.lstrip().startswith("@parallel_indices") else "parallel", text, kw, __name__)


def test_metadata_standard():
    k = _k("@parallel_indices (ix, iy, iz) function metadata_probe!(A, B, D)\n A[ix, iy, iz] = 2.0 * B[ix, iy, iz]\n return\nend")
    assert k.info["memopt"] is False and k.info["nb_parallel_indices"] == 3
    assert k.info["is_parallel_kernel"] is False


def test_metadata_memopt():
    k = _k("@parallel_indices (ix, iy, iz) memopt=true loopsize=3 optvars=B optranges=(B=(0:0,0:0,0:0),) "
           "function metadata_memopt_probe!(A, B, D)\n A[ix, iy, iz] = 2.0 * B[ix, iy, iz]\n return\nend")
    m = k.info
    assert (m["loopdim"], m["loopsize"], m["loopsizes"], m["memopt"], m["nb_parallel_indices"]) == (3, 3, [1, 1, 3], True, 3)
    assert m["optvars"] == ["B"] and m["optranges"]["B"] == [[0, 0], [0, 0], [0, 0]]
    assert m["offsets"]["B"] == [[0, 0, 0, 1]]
    assert (m["shmem_dim1"], m["shmem_dim2"], m["shmem_optvars"], m["use_any_shmem"]) == (1, 2, [], False)
    assert m["shmem_spans"] == {"B": [0, 0]} and m["stencilranges"]["B"] == [[0, 0], [0, 0], [0, 0]]
    assert m["use_shmemhalos"]["B"] is True


def test_metadata_memopt_z_and_full_stencils():
    z = _k("@parallel_indices (ix, iy, iz) memopt=true loopsize=3 optvars=B optranges=(B=(0:0,0:0,-1:1),) "
           "function metadata_memopt_zstencil_probe!(A, B, D)\n A[ix, iy, iz] = B[ix, iy, iz-1] + B[ix, iy, iz] + B[ix, iy, iz+1] + D[ix, iy, iz, 1]\n return\nend").info
    assert z["shmem_optvars"] == [] and z["shmem_spans"] == {"B": [0, 0]} and z["use_any_shmem"] is False
    assert z["stencilranges"]["B"] == [[0, 0], [0, 0], [-1, 1]]
    f = _k("@parallel_indices (ix, iy, iz) memopt=true loopsize=3 optvars=B function metadata_memopt_fullstencil_probe!(A, B, D)\n"
           " A[ix, iy, iz] = B[ix-1, iy-1, iz-1] + B[ix, iy, iz] + B[ix+1, iy+1, iz+1] + D[ix, iy, iz, 1]\n return\nend").info
    # the GPU branch of the reference test
    assert f["shmem_optvars"] == ["B"] and f["shmem_spans"] == {"B": [2, 2]} and f["use_any_shmem"] is True
    assert f["stencilranges"]["B"] == [[-1, 1], [-1, 1], [-1, 1]]


def test_metadata_memopt_2d():
    k = ps.core._define("parallel_indices", "@parallel_indices (ix, iy) ndims=2 memopt=true loopdim=2 loopsize=5 optvars=B "
                        "optranges=(B=(0:0,-1:1,0:0),) function metadata_memopt_2d_probe!(A, B)\n if (1 < iy < size(A, 2))\n"
                        " A[ix, iy] = B[ix, iy - 1] + B[ix, iy] + B[ix, iy + 1]\n end\n return\nend", {}, MOD2)
    m = k.info
    assert (m["nb_parallel_indices"], m["loopdim"], m["loopsize"], m["loopsizes"]) == (2, 2, 5, [1, 5, 1])
    assert m["optranges"]["B"] == [[0, 0], [-1, 1], [0, 0]] and (m["shmem_dim1"], m["shmem_dim2"]) == (1, 3)
    assert m["shmem_optvars"] == [] and m["use_any_shmem"] is False and m["use_shmemhalos"]["B"] is False


def test_onthefly_and_fd_lowering():
    """The flux-form diffusion kernel: qTx, qTy, qTz become on-the-fly arrays, and the lowered
    body reads T at the seven-point stencil and Cp at the centre."""
    from parallelstencil import workloads as W
    k = W.define_diffusion3d(False, __name__)
    assert k.info["onthefly"] == ["qTx", "qTy", "qTz"]
    assert k.info["written"] == ["T2"] and set(k.info["accessed"]) >= {"T", "Cp", "T2"}
    A = [ps.ArgSpec((9, 7, 8)) for _ in range(3)]
    src = k.source(*A, 1.0, 0.1, 0.3, 0.4, 0.5)
    assert "psk_diffusion3D_step" in src and "__global__" in src
    g = k.geometry(*A, 1.0, 0.1, 0.3, 0.4, 0.5)
    assert g.memopt is False and g.checked is False  # the bounds proof clears every access
    assert g.ranges == ((1, 9), (1, 7), (1, 8))


@pytest.mark.parametrize("text,exc,match", [
    ("@parallel function f!(A, B)\n a = B + 1\n return\nend", ps.ArgumentError, None),
    ("@parallel_indices (ix, iy, iz) function write_xy_plane!(A)\n A[ix, iy, 1] = 2.0 * A[ix, iy, 1]\n return\nend",
     ps.ArgumentError, "all parallel indices must be used"),
    ("@parallel_indices (ix, iy, iz) memopt=true function w!(A, B)\n A[ix, iy, 1] = 2.0 * A[ix, iy, 1] + B[ix, iy, 1]\n return\nend",
     ps.ArgumentError, "all parallel indices must be used"),
    ("@parallel_indices (ix, iy, iz) memopt=true optvars=A function w!(A, B)\n A[ix, iy, iz] = B[ix, iy, iz]\n return\nend",
     ps.IncoherentArgumentError, "only read"),
    ("@parallel_indices (ix) memopt=true function w!(A, B)\n A[ix] = B[ix] + B[ix+1]\n return\nend",
     ps.IncoherentArgumentError, "2-D and 3-D"),
    ("@parallel function f!(A, B)\n @all(A) = @nosuchmacro(B)\n return\nend", ps.PSException, None),
    ("@parallel function f!(A, B)\n @all(A) = (@all(B) + 1\n return\nend", ps.PSException, None),
])
def test_definition_errors(text, exc, match):
    with pytest.raises(exc, match=match):
        _k(text)


def test_compute_nthreads_and_nblocks():
    assert ps.compute_nthreads((512, 512, 512)) == (64, 4, 1)
    assert ps.compute_nblocks((512, 512, 512), (64, 4, 1)) == (8, 128, 512)
    nt = ps.compute_nthreads((20, 3, 1))
    assert nt[0] == 20 and nt[1] <= 3 and nt[0] * nt[1] <= 256


def _cells(r):
    return set(itertools.product(*[range(a, b + 1) for a, b in r]))


@pytest.mark.parametrize("bw,ms", [((2, 2, 3), (6, 7, 8)), ((1, 1, 2), (6, 7, 8)), ((2, 2), (6, 7)), ((2,), (9,))])
def test_hide_communication_ranges_partition_the_box(bw, ms):
    full = tuple((1, m) for m in ms)
    outer = ps.get_ranges_outer(bw, full)
    inner = ps.get_ranges_inner(bw, full)
    box = _cells(tuple(full) + ((1, 1),) * (3 - len(ms)))
    parts = [_cells(r) for r in outer] + [_cells(r) for r in inner]
    assert sum(len(p) for p in parts) == len(box) and set().union(*parts) == box  # disjoint cover
    bw3 = tuple(bw) + (0,) * (3 - len(bw))
    ms3 = tuple(ms) + (1,) * (3 - len(ms))
    exp_inner = tuple((b + 1, m - b) if m > 1 else (1, 1) for b, m in zip(bw3, ms3))
    assert inner == (exp_inner,)


def test_hide_communication_errors():
    with pytest.raises(ps.IncoherentArgumentError):
        ps.get_ranges_outer((3, 2, 3), ((1, 6), (1, 7), (1, 8)))


@pytest.mark.parametrize("grid", [(9, 7, 8), (9, 7), (9,)])
def test_field_allocator_sizes(grid):
    n = len(grid)
    exp = {None: (0, 0, 0), "X": (-1, -2, -2), "Y": (-2, -1, -2), "Z": (-2, -2, -1), "BX": (1, 0, 0), "BY": (0, 1, 0),
           "BZ": (0, 0, 1), "XX": (0, -2, -2), "YY": (-2, 0, -2), "ZZ": (-2, -2, 0), "XY": (-1, -1, -2), "XZ": (-1, -2, -1),
           "YZ": (-2, -1, -1)}
    for t, off in exp.items():
        if n < 3 and t in ("BZ",):
            continue
        assert FA.field_size(grid, t) == tuple(grid[i] + off[i] for i in range(n)), t


def test_bf16_conversion_rounds_to_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -2.5, 3.0e38, np.inf, np.nan], np.float32)
    b = ps.f32_to_bf16(x)
    assert b[0] == 0x3F80 and b[1] == 0x3F80 and b[2] == 0x3F82  # ties to even
    assert b[5] == 0x7F80 and b[6] == 0x7FC0
    assert np.array_equal(ps.bf16_to_f32(b[:4]), np.array([1.0, 1.0, 1.015625, -2.5], np.float32))


# ---------------------------------------------------------------- kernel language and @overlap
KL_SHMEM = """
@parallel_indices (ix,iy) function sm!(B, A)
    T_l = @sharedMem(eltype(A), (@blockDim().x+2, @blockDim().y))
    T_l[@threadIdx().x+1, @threadIdx().y] = A[ix,iy]
    @sync_threads()
    B[ix,iy] = T_l[@threadIdx().x+1, @threadIdx().y] + T_l[end, 1]
    return
end
"""


def test_kernel_language_lowering():
    k = _kl(KL_SHMEM)
    assert k.info["kernel_language"] and k.info["shared_arrays"] == 1
    A = ps.ArgSpec((64, 8))
    src = k.source(A, A, checked=True)
    assert "extern __shared__" in src and "__syncthreads();" in src and "ps_shmem" in src
    assert "((ps_int)threadIdx.x + 1)" in src and "((ps_int)blockDim.y)" in src
    # `end` inside a @sharedMem index is that dimension's extent
    assert "sm_T_l[((sm_T_l_n1) - 1)" in src
    # without shmem the launch is refused before anything runs (no device needed)
    with pytest.raises(ps.ArgumentError, match="shmem"):
        k.geometry(A, A, nblocks=(2, 2), nthreads=(32, 4))
    g = k.geometry(A, A, nblocks=(2, 2), nthreads=(32, 4), shmem=34 * 4 * 8)
    assert g.checked  # shared-memory accesses are not provable on the host: bounds-checked variant


def _kl(text):
    return ps.parallel_indices(text)


@pytest.mark.parametrize("body,match", [
    ("B[ix] = @threadIdx(1).x", "no arguments"),
    ("B[ix] = @threadIdx().w", "x, y and z"),
    ("B[ix] = @threadIdx()", "use @threadIdx"),
    ("B[ix] = @shfl_sync(1, 2.0)", "wrong number"),
    ("B[ix] = @shfl_sync(@active_mask(), 2.0, 1, 48)", "power of two"),
    ("B[ix] = @vote_any_sync(@active_mask(), 2.0)", "Bool"),
    ("B[ix] = @sharedMem(Float64, 4)[1]", "local variable"),
    ("S = @sharedMem(Float64, (1,2,3,4))\n    B[ix] = 1.0", "1 to 3"),
    ("S = @sharedMem(Float64, 4)\n    B[ix] = S[1,1]", "1 or 1 indices"),
    ("B[ix] = S[1]\n    S = @sharedMem(Float64, 4)", "before its @sharedMem"),
    ("S = @sharedMem(Complex, 4)\n    B[ix] = 1.0", "element type"),
    ("B[ix] = @sync_threads()", "statement"),
    ("@foo(B)\n    B[ix] = 1.0", "not supported"),
])
def test_kernel_language_errors(body, match):
    text = f"@parallel_indices (ix) function kl!(B)\n    {body}\n    return\nend"
    with pytest.raises(ps.PSException, match=match):
        k = ps.parallel_indices(text)
        k.source(ps.ArgSpec((64,)))


def test_kernel_language_not_with_memopt():
    with pytest.raises(ps.ArgumentError, match="memopt"):
        ps.parallel_indices("@parallel_indices (ix,iy,iz) function m!(B, A)\n"
                            "    B[ix,iy,iz] = A[ix,iy,iz-1] + A[ix,iy,iz+1] + @threadIdx().x\n    return\nend", memopt=True)


def test_overlap_errors_on_the_host():
    k = ps.parallel_indices("@parallel_indices (ix) function f!(A)\n    A[ix] = 1.0\n    return\nend")
    with pytest.raises(ps.ArgumentError, match="stream"):
        with ps.overlap():
            k(None, stream=1)
    with pytest.raises(ps.ArgumentError, match="only contain @parallel"):
        with ps.overlap():
            pass
    with pytest.raises(ps.ArgumentError, match="only contain @parallel"):
        with ps.overlap():
            ps.synchronize()
    with pytest.raises(ps.ArgumentError, match="only contain @parallel"):
        with ps.overlap():
            k(None, async_=True)
    with pytest.raises(ps.ArgumentError, match="stream"):
        ps.overlap((k, (None,), {"stream": 2}))
    with pytest.raises(ps.ArgumentError, match="only contain @parallel"):
        ps.overlap((k,))


# ---------------------------------------------------------------- CellArrays
def _cellspec(shape, celldims, blocklength=0, celltype=None):
    return ps.ArgSpec(shape, celldims=celldims, blocklength=blocklength, celltype=celltype)


def test_celltype_checks():
    with pytest.raises(ps.ArgumentError, match="fieldnames"):
        ps.CellType("C")
    with pytest.raises(ps.ArgumentError, match="prod"):
        ps.CellType("C", fieldnames=("a", "b", "c"), dims=(2, 2))
    with pytest.raises(ps.ArgumentError, match="parametric"):
        ps.CellType("C", fieldnames=("a", "b"), eltype=ps.Float32, parametric=True)
    T = ps.CellType("T", fieldnames=("xx", "yy", "xy"))
    assert T.dims == (3,)
    P = ps.CellType("P", fieldnames=("xx", "yy"), parametric=True)
    assert P[ps.Float32].eltype == ps.Float32


def test_cellarray_lowering():
    k = ps.parallel("@parallel function c!(T2, T, Ci)\n    @all(T2) = @all(T) + 2.0*@all(Ci)\n    return\nend")
    A = _cellspec((8, 8, 8), (4, 4))
    src = k.source(A, A, A)
    assert src.count("a0[") == 16  # the statement once per value of the 4x4 cells
    assert "a1[((a1_o)) + (15) * a1_B]" in src  # value 16 of a cell field stored contiguously
    B = _cellspec((8, 8, 8), (4, 4), blocklength=1)
    src = k.source(B, B, B)
    assert "a0[((a0_o)) * 16 + (15)]" in src  # cells stored whole
    assert not k.geometry(A, A, A).checked  # whole-cell accesses are proven like field accesses


@pytest.mark.parametrize("body,match", [
    ("T2[ix,iy,iz] = T[ix,iy,iz] + 1.0", "cannot be combined"),
    ("T2[ix,iy,iz] = T[ix,iy,iz] * T[ix,iy,iz]", "products of two cells"),
    ("T2[ix,iy,iz] = 1.0 / T[ix,iy,iz]", "division by a cell"),
    ("T2[ix,iy,iz] = 3.0", "number is assigned"),
    ("field(T2,1,1)[ix,iy,iz] = T[ix,iy,iz]", "cell is assigned"),
    ("T2[ix,iy,iz] = sqrt(T[ix,iy,iz])", "where a number is expected"),
    ("field(T2,3,1)[ix,iy,iz] = 1.0", "BoundsError"),
    ("field(T2,1,1)[ix,iy,iz] = T[ix,iy,iz].zz", "no cell field"),
])
def test_cellarray_errors(body, match):
    k = ps.parallel_indices(f"@parallel_indices (ix,iy,iz) function c!(T2, T)\n    {body}\n    return\nend")
    Tn = ps.CellType("Tn", fieldnames=("xx", "yx", "xy", "yy"), dims=(2, 2))
    with pytest.raises(ps.PSException, match=match):
        k.source(_cellspec((4, 4, 4), (2, 2)), _cellspec((4, 4, 4), None, celltype=Tn))


def test_cellarray_value_forms_are_checked_unless_literal():
    k = ps.parallel_indices("""
    @parallel_indices (ix,iy,iz) function c!(T2, T)
        for c = 1:cellsize(T2,1)
            field(T2,c,1)[ix,iy,iz] = T[ix,iy,iz][c,2]
        end
        return
    end""")
    A = _cellspec((4, 4, 4), (2, 2))
    assert not k.geometry(A, A).checked  # c runs over 1:cellsize(T2,1) = 1:2: inside the 2x2 cells
    k2 = ps.parallel_indices("@parallel_indices (ix,iy,iz) function c!(T2, T)\n    field(T2,2,1)[ix,iy,iz] = T[ix,iy,iz][1,2]\n    return\nend")
    assert not k2.geometry(A, A).checked
    k3 = ps.parallel_indices("@parallel_indices (ix,iy,iz) function c!(T2, T)\n    for c = 1:3\n"
                             "        field(T2,c,1)[ix,iy,iz] = T[ix,iy,iz][c,1]\n    end\n    return\nend")
    assert k3.geometry(A, A).checked  # c = 3 lies outside the cells
    k4 = ps.parallel_indices("@parallel_indices (ix,iy,iz) function c!(T2, T, n)\n    for c = 1:n\n"
                             "        field(T2,c,1)[ix,iy,iz] = T[ix,iy,iz][c,1]\n    end\n    return\nend")
    assert not k4.geometry(A, A, 2).checked and k4.geometry(A, A, 3).checked  # integer arguments enter the proof
    with pytest.raises(ps.ArgumentError, match="loop variable"):
        ps.parallel_indices("@parallel_indices (ix) function c!(A)\n    for c = 1:3\n        c = 2\n"
                            "        A[ix] = c\n    end\n    return\nend").source(ps.ArgSpec((8,)))


# ---------------------------------------------------------------- nested functions
@pytest.mark.parametrize("body,match", [
    ("f(A) = (A[ix] = 1.0; return)\n    g(A) = f(A)\n    A[ix] = g(A)", "single `return expr`"),
    ("f(A) = f(A)\n    A[ix] = f(A)", "recursion"),
    ("function f(x)\n        x = x + 1\n        return x\n    end\n    A[ix] = f(1.0)", "parameter"),
    ("function f(A)\n        if ix > 2\n            return\n        end\n        A[ix] = 1.0\n    end\n    f(A)", "inside a branch"),
    ("f(a, b) = a + b\n    A[ix] = f(1.0)", "MethodError"),
    ("if ix > 1\n        f(A) = 2.0\n    end\n    A[ix] = 1.0", "top level"),
])
def test_nested_function_errors(body, match):
    with pytest.raises(ps.PSException, match=match):
        k = ps.parallel_indices(f"@parallel_indices (ix) function k!(A)\n    {body}\n    return\nend")
        k.source(ps.ArgSpec((8,)))


def test_nested_function_locals_get_their_own_scope():
    with pytest.raises(ps.ArgumentError, match="single `return expr`"):  # refused rather than misread
        ps.parallel_indices("""@parallel_indices (ix) function k!(A)
    function twice(a)
        t = 2.0*a
        return t
    end
    t = 1.0
    A[ix] = twice(A[ix]) + t
    return
end""")
    k2 = ps.parallel_indices("""@parallel_indices (ix) function k!(A, B)
    function step!(A, B)
        t = 2.0*B[ix]
        A[ix] = t
    end
    t = 1.0
    step!(A, B)
    A[ix] = A[ix] + t
    return
end""")
    src = k2.source(ps.ArgSpec((8,)), ps.ArgSpec((8,)))
    assert "v_t__step_21" in src and "v_t = " in src  # the function's t is not the kernel's t


def test_for_with_several_variables():
    k = ps.parallel_indices("@parallel_indices (ix) function k!(A)\n    for a = 1:2, b = 1:3\n        A[ix] = A[ix] + a*b\n    end\n    return\nend")
    src = k.source(ps.ArgSpec((8,)))
    assert "for (ps_int f_a" in src and "for (ps_int f_b" in src


@pytest.mark.parametrize("text,match", [
    ("@parallel_indices (ix) function f!(A)\n    A[ix] = 1.0\nend", "last statement must be a `return nothing`"),
    ("@parallel_indices (ix) function f!(A)\n    A[ix] = 1.0\n    return A\nend", "last statement must be a `return nothing`"),
    ("@parallel_indices (ix) function f!(A)\n    if ix > 1\n        return\n    end\n    A[ix] = 1.0\n    return\nend", "only one return statement"),
])
def test_return_statement_rules(text, match):
    # remove_return, src/ParallelKernel/shared.jl:200-209
    with pytest.raises(ps.ArgumentError, match=match):
        ps.parallel_indices(text)
    ps.parallel_indices("@parallel_indices (ix) function f!(A)\n    A[ix] = 1.0\n    nothing\nend")
    ps.parallel_indices("@parallel_indices (ix) function f!(A)\n    A[ix] = 1.0\n    return nothing\nend")


# ---------------------------------------------------------------- tuple arguments and @∀
def test_tuple_arguments_and_forall_lowering():
    k = ps.parallel_indices("""@parallel_indices (ix) function f!(C, A)
    @∀ i ∈ (x,y) C.i[ix] = 2.0*A.i[ix]
    return
end""")
    assert k.info["def_params"] == ["C", "A"]
    s = lambda: {"x": ps.ArgSpec((8,)), "y": ps.ArgSpec((8,))}
    src = k.source(s(), s())
    assert re.search(r"a0\[[^\n]*\] = \(0x1p\+1 \* a2\[", src) and re.search(r"a1\[[^\n]*\] = \(0x1p\+1 \* a3\[", src)
    # the structure of the call decides the components: a third component is a parameter too
    # (C.x, C.y, C.z are a0..a2, A.x, A.y, A.z a3..a5)
    s3 = lambda: {"x": ps.ArgSpec((8,)), "y": ps.ArgSpec((8,)), "z": ps.ArgSpec((8,))}
    assert re.search(r"a1\[[^\n]*\] = \(0x1p\+1 \* a4\[", k.source(s3(), s3()))
    with pytest.raises(ps.ArgumentError, match="has no component"):
        k.source({"x": ps.ArgSpec((8,))}, s())
    with pytest.raises(ps.ArgumentError, match="tuple arguments"):
        k.source(ps.ArgSpec((8,)), ps.ArgSpec((8,)))


@pytest.mark.parametrize("text,match", [
    ("@∀ i ∈ 1:n A[ix] = 1.0", "literal bounds"),
    ("@∀ (i,j) ∈ ((x,y), (z,)) A[ix] = 1.0", "tuple of 2"),
    ("@∀ i ⊂ (x,y) A[ix] = 1.0", "x ∈ X"),
])
def test_forall_errors(text, match):
    with pytest.raises(ps.PSException, match=match):
        ps.parallel_indices(f"@parallel_indices (ix) function f!(A, n)\n    {text}\n    return\nend")


def test_launch_refuses_argspecs_and_freed_arrays():
    # an ArgSpec is a shape without memory, and a freed array has none left: both are refused on
    # the host, before anything reaches a kernel (no GPU is touched)
    k = _k("@parallel function launch_probe!(A, B)\n @all(A) = @all(B)\n return\nend")
    spec = ps.ArgSpec((8, 8, 8))
    with pytest.raises(ps.ArgumentError, match="ArgSpec"):
        k(spec, ps.ArgSpec((8, 8, 8)))
    with pytest.raises(ps.ArgumentError, match="ArgSpec"):
        k.launcher(spec, ps.ArgSpec((8, 8, 8)))
    f = ps.core.ps_field()
    for d in range(3):
        f.dims[d] = 8
    f.ndims, f.dtype, f.pitch, f.plane = 3, ps.core.PS_F64, 8, 64
    freed = ps.core.DeviceArray(f)  # ptr and base NULL: what free() leaves behind
    with pytest.raises(ps.ArgumentError, match="freed"):
        k(freed, freed)
    view = ps.core.DeviceArray(ps.core.ps_field(), owner=False, parent=freed)
    view.f.ptr, view.f.ndims, view.f.dtype = 1 << 44, 3, ps.core.PS_F64
    for d in range(3):
        view.f.dims[d] = 4
    with pytest.raises(ps.ArgumentError, match="freed"):
        k(view, view)


def test_uniform_subexpressions_become_host_arguments():
    """dt/ρ of the acoustic velocity update depends on scalar arguments only: the generated kernel
    receives it as one hidden argument, and ρ is no longer a device divisor. Under fastmath the
    device keeps its own form."""
    from parallelstencil import workloads as W
    kV, _ = W.define_acoustic3d(__name__, False)
    n = 12
    F = [ps.ArgSpec((n + 1, n, n)), ps.ArgSpec((n, n + 1, n)), ps.ArgSpec((n, n, n + 1)), ps.ArgSpec((n, n, n))]
    src = kV.source(*F, 0.1, 2.0, 0.3, 0.3, 0.3)
    assert "double ps_u0" in src and "ps_u1" not in src and "ps_u0 * " in src and "a5_r" not in src
    kf = ps.parallel(W.ACOUSTIC3D_V, fastmath=True)
    assert "ps_u0" not in kf.source(*F, 0.1, 2.0, 0.3, 0.3, 0.3)


def test_x_group_lowering():
    """Plain kernels made of masked array assignments run as x groups: a fast path for groups whose
    masks all hold, with the cells' accesses one row base apart by constant offsets, and a cell by
    cell path for the rest. Kernels with local variables keep one cell per thread."""
    from parallelstencil import workloads as W
    kV, _ = W.define_acoustic3d(__name__, False)
    n = 20
    F = [ps.ArgSpec((n + 1, n, n)), ps.ArgSpec((n, n + 1, n)), ps.ArgSpec((n, n, n + 1)), ps.ArgSpec((n, n, n))]
    src = kV.source(*F, 0.1, 2.0, 0.3, 0.3, 0.3, off32=True, cells_x=4)
    assert "if (ps_grp) {" in src and "const ps_int ps_cx = 3;" in src and "(int)ps_cx + 1)" in src
    # one cell per thread: x offsets in the uniform part (no merged loads that overlap between lanes)
    assert "ps_at(a3, (ps_int)1 + (1)*a3_sx + (1)*a3_sxy, a3_vo, 0)" in kV.source(*F, 0.1, 2.0, 0.3, 0.3, 0.3, off32=True)
    assert "ps_grp" not in kV.source(*F, 0.1, 2.0, 0.3, 0.3, 0.3, off32=True)
    k = ps.parallel_indices("@parallel_indices (ix,iy,iz) function loc!(A, B)\n t = B[ix,iy,iz]*2\n A[ix,iy,iz] = t\n return\nend")
    assert "ps_grp" not in k.source(ps.ArgSpec((n, n, n)), ps.ArgSpec((n, n, n)), off32=True, cells_x=4)
    k = ps.parallel_indices("@parallel_indices (ix,iy,iz) function cp!(A, B)\n A[ix,iy,iz] = B[ix,iy,iz]*2\n return\nend")
    assert "if (ps_grp) {" in k.source(ps.ArgSpec((n, n, n)), ps.ArgSpec((n, n, n)), off32=True, cells_x=2)


def test_x_group_row_split():
    """x groups split each row: the interior, where every statement's x mask holds (the maximum of
    the masks' lower bounds and the minimum of their upper bounds), in groups from its first cell;
    the cells outside it one at a time in the threads past the interior. A mask bound that is not
    an integer keeps groups from the start of the launch range."""
    from parallelstencil import workloads as W
    kV, _ = W.define_acoustic3d(__name__, False)
    n = 20
    F = [ps.ArgSpec((n + 1, n, n)), ps.ArgSpec((n, n + 1, n)), ps.ArgSpec((n, n, n + 1)), ps.ArgSpec((n, n, n))]
    src = kV.source(*F, 0.1, 2.0, 0.3, 0.3, 0.3, off32=True, cells_x=4)
    for a in ("a0", "a1", "a2"):  # @inn of each velocity: 1 <= ix <= n_x - 2 in the kernel's indices
        assert f"ps_xh = min(ps_xh, (ps_int)(({a}_nx - ((ps_int)2))));" in src
    assert src.count("ps_xl = max(ps_xl, (ps_int)(((ps_int)1)));") == 3
    # groups start where Vx[ix+1, ...] (x offset 1) falls on a 128-byte line: 16 FP64 cells
    assert "const ps_int ps_e0 = ps_xl - 1 + (1);" in src and "ps_g0 = ps_e0 - ((ps_e0 % 16) + 16) % 16 + 1 - (1);" in src
    assert "const ps_int ps_ix0 = ps_g0 + tx * 4;" in src and "for (ps_int ps_e = tx - ps_ngp; ps_e < ps_ne; ps_e += ps_es)" in src
    strict = ps.parallel_indices("@parallel_indices (ix,iy,iz) function st!(A, B)\n if ix > 2 && 7 > ix\n A[ix,iy,iz] = B[ix,iy,iz]\n end\n return\nend")
    s2 = strict.source(ps.ArgSpec((n, n, n)), ps.ArgSpec((n, n, n)), off32=True, cells_x=2)
    assert "ps_xl = max(ps_xl, (ps_int)(((ps_int)2)) + 1);" in s2 and "ps_xh = min(ps_xh, (ps_int)(((ps_int)7)) - 1);" in s2
    half = ps.parallel_indices("@parallel_indices (ix,iy,iz) function hf!(A, B)\n if ix <= size(A, 1) / 2\n A[ix,iy,iz] = B[ix,iy,iz]\n end\n return\nend")
    s3 = half.source(ps.ArgSpec((n, n, n)), ps.ArgSpec((n, n, n)), off32=True, cells_x=2)
    assert "ps_ng" not in s3 and "const ps_int ps_ix0 = ps_r0x + tx * 2;" in s3


def test_z_reach_of_kernels():
    """The z reach (largest spread of z offsets into one array) that sets the default z cells of
    FP64 groups: 2 for the 7-point diffusion stencil, 1 for acoustic3D's staggered updates, 0
    without z neighbours; kernels with local variables are not groupable."""
    from parallelstencil import workloads as W
    kd = W.define_diffusion3d(False, __name__)
    kV, kP = W.define_acoustic3d(__name__, False)
    assert kd.info["groupable"] and kd.info["z_reach"] == 2 and kd.info["x_reach"] == 2
    assert kV.info["z_reach"] == 1 and kP.info["z_reach"] == 1 and kV.info["x_reach"] == 1
    k0 = ps.parallel_indices("@parallel_indices (ix,iy,iz) function cp!(A, B)\n A[ix,iy,iz] = B[ix+1,iy,iz]*2\n return\nend")
    assert k0.info["groupable"] and k0.info["z_reach"] == 0 and k0.info["x_reach"] == 0
    kl = ps.parallel_indices("@parallel_indices (ix,iy,iz) function loc!(A, B)\n t = B[ix,iy,iz]*2\n A[ix,iy,iz] = t\n return\nend")
    assert not kl.info["groupable"]


def _row_split_cells(r0, rl, xl, xh, k, A, CX, bx, nthreads):
    """The cells that the threads of one row take in the row split of x groups, as the emitted
    prologue computes them: (interior groups, edge cells)."""
    xl, xh = max(r0, xl), min(r0 + rl - 1, xh)  # the mask bounds within the launch range
    e0 = xl - 1 + k
    g0 = e0 - ((e0 % A) + A) % A + 1 - k
    s = (xl - g0 + CX - 1) // CX if xl > g0 else 0
    gs = g0 + s * CX
    ng = (xh - gs + 1) // CX if xh >= gs else 0
    ngp = (s + ng + bx - 1) // bx * bx
    groups, edges = [], []
    for tx in range(nthreads):
        if s <= tx < s + ng:
            groups.append(g0 + tx * CX)
        elif tx >= ngp:
            nl = gs - r0 if ng > 0 else rl
            ne = rl - ng * CX
            e = tx - ngp
            while e < ne:
                edges.append(r0 + e if e < nl else gs + ng * CX + (e - nl))
                e += nthreads - ngp
    return groups, edges, g0


def test_x_group_row_split_covers_every_cell_once():
    """Every cell of the launch range goes to exactly one thread of the row split, interior groups
    lie inside the interval where all masks hold, and the groups' leading accesses start on
    128-byte lines; the host's x grid leaves room for the leading threads and the edge column."""
    import itertools
    rng = np.random.default_rng(5)
    for es, CX, bx in itertools.product((2, 4), (2, 4, 8), (8, 32, 64)):
        A = 128 // es
        for _ in range(40):
            r0 = int(rng.integers(1, 6))
            rl = int(rng.integers(2 * CX, 700))
            xl = int(rng.integers(r0 - 3, r0 + 20))
            xh = int(rng.integers(r0 + rl - 30, r0 + rl + 3))
            k = int(rng.integers(-1, 3))
            ng_host = (rl + CX - 1) // CX
            lead = max(1, 128 // (es * CX))
            nthreads = ((ng_host + lead + bx - 1) // bx + 1) * bx
            groups, edges, g0 = _row_split_cells(r0, rl, xl, xh, k, A, CX, bx, nthreads)
            cells = [g + c for g in groups for c in range(CX)] + edges
            assert sorted(cells) == list(range(r0, r0 + rl)), (es, CX, bx, r0, rl, xl, xh, k)
            for g in groups:
                assert max(xl, r0) <= g and g + CX - 1 <= min(xh, r0 + rl - 1)
            assert ((g0 - 1 + k) * es) % 128 == 0


def test_quotient_checks_per_statement():
    """Statements whose value is NaN whenever a quotient is NaN check their quotients once (the
    dividends' exponent range and the value); a comparison keeps the check of each quotient."""
    k = ps.parallel("@parallel function q1!(R, A, b)\n @all(R) = @all(A)/b + 2.0*@all(A)/b\n return\nend")
    body = k.source(ps.ArgSpec((8, 8, 8)), ps.ArgSpec((8, 8, 8)), 0.3).split("__global__")[1]
    assert "ps_divx(" in body and "ps_bad" not in body and "(ps_lo < -599) | (ps_hi > 600)" in body
    k = ps.parallel("@parallel function q2!(R, A, b)\n @all(R) = (@all(A)/b > 1.0) ? 1.0 : 0.0\n return\nend")
    body = k.source(ps.ArgSpec((8, 8, 8)), ps.ArgSpec((8, 8, 8)), 0.3).split("__global__")[1]
    assert "ps_divp(" in body and "ps_divx(" not in body


def test_batched_stokes_kernel_texts():
    # every array access of the 2-D fused kernels gains the sample index, nothing else changes
    import re
    from parallelstencil import stokes_fused as F
    from parallelstencil import workloads as W
    for text in (F.STOKES2D_PTAU, F.STOKES2D_DVV, W.STOKES2D_MAXLOC_X, W.STOKES2D_MAXLOC_Y):
        b = F.batched_text(text)
        assert re.search(r"@parallel_indices \(ix, iy, iz\) function \S+_batch!\(", b)
        n = len(re.findall(r"\w+\[[^\[\]]+\]", text))
        assert n > 0 and len(re.findall(r"\w+\[[^\[\]]+, iz\]", b)) == n
        assert re.sub(r", iz\]", "]", b.replace("(ix, iy, iz)", "(ix, iy)").replace("_batch!(", "!(", 1)).replace("(ix,iy)", "(ix, iy)") \
            == text.replace("(ix,iy)", "(ix, iy)")


def test_batched_stokes_kernel_guard():
    # the guarded form: Act first, every statement inside one `if (Act[1, 1, iz] != 0.0)`
    from parallelstencil import stokes_fused as F
    for text in (F.STOKES2D_PTAU, F.STOKES2D_DVV):
        plain, g = F.batched_text(text), F.batched_text(text, guard=True)
        assert "_batch!(Act::Data.Array, " in g
        body = g.split("\n", 2)[2]
        assert body.startswith("    if (Act[1, 1, iz] != 0.0)\n") and body.endswith("    end\n    return\nend\n")
        assert body[len("    if (Act[1, 1, iz] != 0.0)\n"):-len("    end\n    return\nend\n")] == plain.split("\n", 2)[2][:-len("    return\nend\n")]
