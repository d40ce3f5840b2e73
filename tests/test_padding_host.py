"""padding=true without a GPU: the padded Field tables (parent sizes and view ranges of
test/ParallelKernel/test_allocators.jl:942-1046), the lowering of masks and index roles
(test/test_parallel.jl:1068-1079,1261-1276) and @allocate's argument checks."""
import pytest

import parallelstencil as ps
from parallelstencil import FieldAllocators as FA

MODP = __name__ + ".pad"
MODN = __name__ + ".nopad"
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, padding=True, module=MODP)
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=MODN)
nx, ny, nz = 7, 5, 8

# test/ParallelKernel/test_allocators.jl:946-969 (size(@XField(...).parent) == ...)
PARENTS_3D = {None: (nx, ny, nz), "X": (nx + 1, ny, nz), "Y": (nx, ny + 1, nz), "Z": (nx, ny, nz + 1),
              "BX": (nx + 1, ny, nz), "BY": (nx, ny + 1, nz), "BZ": (nx, ny, nz + 1), "XX": (nx, ny, nz),
              "YY": (nx, ny, nz), "ZZ": (nx, ny, nz), "XY": (nx + 1, ny + 1, nz), "XZ": (nx + 1, ny, nz + 1),
              "YZ": (nx, ny + 1, nz + 1)}
PARENTS_2D = {None: (nx, ny), "X": (nx + 1, ny), "Y": (nx, ny + 1), "Z": (nx, ny), "BX": (nx + 1, ny), "BY": (nx, ny + 1),
              "BZ": (nx, ny), "XX": (nx, ny), "YY": (nx, ny), "ZZ": (nx, ny), "XY": (nx + 1, ny + 1), "XZ": (nx + 1, ny),
              "YZ": (nx, ny + 1)}
# :1001-1030 (@XField(...).indices == ...)
VIEWS_3D = {None: ((1, nx), (1, ny), (1, nz)), "X": ((2, nx), (2, ny - 1), (2, nz - 1)), "Y": ((2, nx - 1), (2, ny), (2, nz - 1)),
            "Z": ((2, nx - 1), (2, ny - 1), (2, nz)), "BX": ((1, nx + 1), (1, ny), (1, nz)), "BY": ((1, nx), (1, ny + 1), (1, nz)),
            "BZ": ((1, nx), (1, ny), (1, nz + 1)), "XX": ((1, nx), (2, ny - 1), (2, nz - 1)), "YY": ((2, nx - 1), (1, ny), (2, nz - 1)),
            "ZZ": ((2, nx - 1), (2, ny - 1), (1, nz)), "XY": ((2, nx), (2, ny), (2, nz - 1)), "XZ": ((2, nx), (2, ny - 1), (2, nz)),
            "YZ": ((2, nx - 1), (2, ny), (2, nz))}
VIEWS_2D = {None: ((1, nx), (1, ny)), "X": ((2, nx), (2, ny - 1)), "Y": ((2, nx - 1), (2, ny)), "Z": ((2, nx - 1), (2, ny - 1)),
            "BX": ((1, nx + 1), (1, ny)), "BY": ((1, nx), (1, ny + 1)), "BZ": ((1, nx), (1, ny)), "XX": ((1, nx), (2, ny - 1)),
            "YY": ((2, nx - 1), (1, ny)), "ZZ": ((2, nx - 1), (2, ny - 1)), "XY": ((2, nx), (2, ny)), "XZ": ((2, nx), (2, ny - 1)),
            "YZ": ((2, nx - 1), (2, ny))}


@pytest.mark.parametrize("t", list(PARENTS_3D))
def test_padded_parent_sizes_and_views(t):
    assert FA.parent_size((nx, ny, nz), t) == PARENTS_3D[t]
    assert FA.parent_size((nx, ny), t) == PARENTS_2D[t]
    assert FA.view_ranges((nx, ny, nz), t) == VIEWS_3D[t]
    assert FA.view_ranges((nx, ny), t) == VIEWS_2D[t]
    # the view has the size of the field without padding (FieldAllocators.jl:584-603)
    assert tuple(b - a + 1 for a, b in VIEWS_3D[t]) == FA.field_size((nx, ny, nz), t)


def _body(text, mod):
    return ps.core._define("parallel", text, {}, mod).info["body"]


SUM_ALL = "@parallel function sum!(A, B)\n @all(A) = @all(A) + @all(B)\n return\nend"
SUM_INN = "@parallel function sum!(A, B)\n @inn(A) = @inn(A) + @inn(B)\n return\nend"


def test_masks_and_indices_with_padding():
    # test/test_parallel.jl:1075-1079: masks from the views' boxes, parent accesses without shifts
    b = _body(SUM_ALL, MODP)
    assert "(__first(A, 1) <= __ix) && (__ix <= __last(A, 1))" in b and "__first(A, 3) <= __iz" in b
    assert "A[__ix, __iy, __iz] = (A[__ix, __iy, __iz] + B[__ix, __iy, __iz])" in b
    b = _body(SUM_INN, MODP)
    assert "(__first(A, 1) < __ix) && (__ix < __last(A, 1))" in b
    assert "A[__ix, __iy, __iz] = (A[__ix, __iy, __iz] + B[__ix, __iy, __iz])" in b
    # padding=false keeps the shifted inner indices (:1068-1073)
    b = _body(SUM_INN, MODN)
    assert "A[(__ix + 1), (__iy + 1), (__iz + 1)]" in b and "__first" not in b


def test_staggered_reads_compare_parent_sizes():
    # substitute_indices_dir (src/ParallelKernel/parallel.jl:514-548): a staggered read of another
    # array is shifted by size(B.parent, d) > size(A.parent, d); one of the written array is not
    k = ps.core._define("parallel", "@parallel function v!(Vx, P, dx)\n @inn(Vx) = @inn(Vx) - @d_xi(P)/dx\n return\nend", {}, MODP)
    b = k.info["body"]
    assert "P[((__ix - __pgt(Vx, P, 1)) + 1), __iy, __iz] - P[(__ix - __pgt(Vx, P, 1)), __iy, __iz]" in b
    k = ps.core._define("parallel", "@parallel function d!(A)\n @inn(A) = @d2_xi(A)\n return\nend", {}, MODP)
    assert "__pgt" not in k.info["body"]


def test_parallel_indices_view_and_parent_accesses():
    # substitute_view_accesses: an access whose every index involves a parallel index goes to the
    # parent; any other addresses the view (shifted by its first index)
    k = ps.core._define("parallel_indices", "@parallel_indices (ix, iy, iz) function f!(A, B)\n"
                        " A[ix, iy, iz] = B[ix + 1, iy, iz] + B[1, iy, iz]\n return\nend", {}, MODP)
    b = k.info["body"]
    assert "A[__ix, __iy, __iz]" in b and "B[(__ix + 1), __iy, __iz]" in b
    assert "B[((1 + __first(B, 1)) - 1), ((__iy + __first(B, 2)) - 1), ((__iz + __first(B, 3)) - 1)]" in b


def test_allocate_argument_checks():
    with pytest.raises(ps.ArgumentError):
        FA.allocate(gridsize=(4, 4, 4))
    with pytest.raises(ps.ArgumentError):
        FA.allocate(gridsize=(4, 4, 4), fields={"QField": ("A",)})
    with pytest.raises(ps.ArgumentError):
        FA.allocate(gridsize=(4, 4, 4), fields={"Field": ("A",)}, allocator="@empty")
