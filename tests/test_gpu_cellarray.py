"""GPU tests of CellArrays (@zeros(...; celldims / celltype / blocklength), @CellType, the kernel
forms of examples/memcopyCellArray3D.jl) against the oracle's restatement of the CellArrays.jl
data layout and of the cell-by-cell update."""
import numpy as np
import pytest

import parallelstencil as ps
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=__name__)

COPY_FD = """
@parallel function copy3D!(T2, T, Ci)
    @all(T2) = @all(T) + @all(Ci);
    return
end
"""
COPY_EXPLICIT = """
@parallel_indices (ix,iy,iz) function copy3D_explicit!(T2, T, Ci)
    T2[ix,iy,iz] = T[ix,iy,iz] + Ci[ix,iy,iz]
    return
end
"""
COPY_LARGE = """
@parallel_indices (ix,iy,iz) function copy3D_large_cells!(T2, T, Ci)
    for cy = 1:cellsize(T2,2), cx = 1:cellsize(T2,1)
        field(T2,cx,cy)[ix,iy,iz] = field(T,cx,cy)[ix,iy,iz] + field(Ci,cx,cy)[ix,iy,iz]
    end
    return
end
"""


def _cells(dims, celldims, rng, **kw):
    h = np.asfortranarray(rng.integers(-1000, 1000, dims + celldims).astype(np.float64))
    A = ps.zeros(*dims, celldims=celldims, **kw) if "celltype" not in kw else ps.zeros(*dims, **kw)
    return A.copy_from(h), h


@pytest.mark.parametrize("blocklength", [0, 1, 5])
def test_cellarray_data_layout(blocklength):
    rng = np.random.default_rng(1)
    dims, celldims = (7, 6, 5), (4, 4)
    A, h = _cells(dims, celldims, rng, blocklength=blocklength)
    B = 7 * 6 * 5 if blocklength == 0 else blocklength
    assert A.blocklength == B and A.celldims == celldims and A.shape == dims
    assert np.array_equal(A.data.to_host(), O.cells_to_data(h, celldims, B))
    assert np.array_equal(A.to_host(), h)


@pytest.mark.parametrize("text", [COPY_FD, COPY_EXPLICIT, COPY_LARGE], ids=["fd", "explicit", "large_cells"])
@pytest.mark.parametrize("blocklength", [0, 1, 3])
def test_memcopy_cellarray_example(text, blocklength):
    print(f"memcopy cells blocklength={blocklength}", flush=True)
    rng = np.random.default_rng(2)
    dims, celldims = (33, 17, 9), (4, 4)
    Cell4x4 = ps.CellType("Cell4x4", fieldnames=tuple(f"v{i}" for i in range(16)), dims=(4, 4))
    T, hT = _cells(dims, celldims, rng, celltype=Cell4x4, blocklength=blocklength)
    Ci, hC = _cells(dims, celldims, rng, blocklength=blocklength)
    T2 = ps.zeros(*dims, celltype=Cell4x4, blocklength=blocklength)
    k = ps.parallel(text) if text is COPY_FD else ps.parallel_indices(text)
    k(T2, T, Ci)
    assert np.array_equal(T2.to_host(), O.copy3d_cells(hT, hC))


def test_cell_fields_by_name_and_index_with_plain_fields():
    # a symmetric 2-D tensor cell updated from a scalar field: named fields, A[I][i] and
    # number * cell expressions over a cell array with cells stored whole
    rng = np.random.default_rng(3)
    dims = (40, 30)
    Sym = ps.CellType("SymmetricTensor2D", fieldnames=("xx", "zz", "xz"))
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=__name__ + ".2d")
    k = ps.core._define("parallel_indices", """
    @parallel_indices (ix,iy) function upd!(S2, S, P, E)
        S2[ix,iy] = 2.0*S[ix,iy] - P[ix,iy]*E[ix,iy]
        field(S2, 3)[ix,iy] = S[ix,iy].xx + S[ix,iy][2] + E[ix,iy].xz
        return
    end""", {}, __name__ + ".2d")
    S, hS = _cells(dims, (3,), rng, celltype=Sym, blocklength=1)
    E, hE = _cells(dims, (3,), rng, celltype=Sym)
    P = ps.zeros(*dims)
    hP = np.asfortranarray(rng.integers(-9, 9, dims).astype(np.float64))
    P.copy_from(hP)
    S2 = ps.zeros(*dims, celltype=Sym, blocklength=1)
    k(S2, S, P, E)
    ref = 2.0 * hS - hP[..., None] * hE
    ref[..., 2] = hS[..., 0] + hS[..., 1] + hE[..., 2]
    assert np.array_equal(S2.to_host(), ref)


def test_cell_value_index_out_of_range_raises():
    k = ps.parallel_indices("""
    @parallel_indices (ix,iy,iz) function c!(T2, T)
        for c = 1:3
            field(T2,c,1)[ix,iy,iz] = T[ix,iy,iz][c,1]
        end
        return
    end""")
    T2, T = ps.zeros(4, 4, 4, celldims=(2, 2)), ps.ones(4, 4, 4, celldims=(2, 2))
    with pytest.raises(ps.BoundsError):
        k(T2, T)  # c = 3 lies outside the 2x2 cells: the checked variant skips it and raises
    h = T2.to_host()
    assert np.all(h[..., :, 0] == 1.0) and np.all(h[..., :, 1] == 0.0)


def test_cellarray_allocators_copy_and_fill():
    A = ps.ones(5, 4, 3, celldims=(2, 3))
    assert np.all(A.to_host() == 1.0)
    B = ps.fill(2.5, 5, 4, 3, celldims=(2, 3), blocklength=1)
    assert B.to_host().shape == (5, 4, 3, 2, 3) and np.all(B.to_host() == 2.5)
    C = ps.zeros(5, 4, 3, celldims=(2, 3))
    C.data.copy_(A.data)
    assert np.all(C.to_host() == 1.0)
    ps.fill_(C, -1.0)
    assert np.all(C.to_host() == -1.0)
    ps.seed(4)
    R = ps.rand(5, 4, 3, celldims=(2, 3))
    r = R.to_host()
    assert r.shape == (5, 4, 3, 2, 3) and 0.0 <= r.min() and r.max() < 1.0
    assert ps.trues(3, 3, celldims=2).to_host().all()
    with pytest.raises(ps.ArgumentError, match="memopt"):
        k = ps.parallel("@parallel function m!(T2, T)\n    @inn(T2) = @d2_xi(T) + @d2_yi(T) + @d2_zi(T)\n    return\nend", memopt=True)
        T = ps.zeros(8, 8, 8, celldims=(2,))
        k(T, T)
