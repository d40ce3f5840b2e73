"""Plain kernels whose threads take several adjacent x cells (x groups): bit for bit the same
fields as one cell per thread, for FP64, FP32 and BFloat16 arrays, x extents that the group size
does not divide, staggered arrays, masks of @all and @inn, neighbours in x, y and z and uniform
divisions with their IEEE fallback."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import workloads as W

pytestmark = pytest.mark.gpu

M64, M32 = __name__ + ".f64", __name__ + ".f32"
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=M64)
ps.init_parallel_stencil(ps.AMDGPU, ps.Float32, 3, module=M32)

# race-free: no statement reads, at another cell, an array that the kernel writes
STENCIL = """
@parallel function mix!(A, B, C, D, s::Data.Number, h::Data.Number)
    @inn(A) = @inn(A) + s/h*(@d2_xi(B) + @d2_yi(B) + @d2_zi(B)) - @inn(C)/h
    @all(D) = @all(C)*s + @all(B)/h
    return
end
"""


def _fields(shape, elt, rng):
    t = np.float64 if elt is ps.Float64 else np.float32
    out = []
    for k in range(4):
        a = np.asfortranarray(rng.standard_normal(shape).astype(t))
        if k == 0:
            a.reshape(-1)[::97] = 0.0  # zero quotients take the fast path, as they must
        if k == 2:  # quotients below the exact range: the group's IEEE recomputation runs
            a.reshape(-1)[5::211] = 1e-200 if t is np.float64 else 1e-20
        d = ps.zeros(*shape, eltype=elt)
        d.copy_from(ps.f32_to_bf16(a) if elt is ps.BFloat16 else a)
        out.append((a, d))
    return out


@pytest.mark.parametrize("elt,mod", [(ps.Float64, M64), (ps.Float32, M32), (ps.BFloat16, M32)])
@pytest.mark.parametrize("shape", [(130, 7, 6), (67, 5, 9), (256, 3, 4)])
@pytest.mark.parametrize("cx", [2, 4, 8])
def test_x_groups_match_one_cell_per_thread(elt, mod, shape, cx):
    rng = np.random.default_rng(sum(shape) + cx)
    ref = ps.core._define("parallel", STENCIL, {"cells_x": 1}, mod)
    grp = ps.core._define("parallel", STENCIL, {"cells_x": cx}, mod)
    f1 = _fields(shape, elt, rng)
    f2 = [(a, ps.zeros(*shape, eltype=elt)) for a, _ in f1]
    for (_, d1), (_, d2) in zip(f1, f2):
        d2.copy_from(d1.to_host())
    for it in range(3):
        ref(*[d for _, d in f1], 0.37, 1e-3 if it == 1 else 0.9)
        grp(*[d for _, d in f2], 0.37, 1e-3 if it == 1 else 0.9)
    for (_, d1), (_, d2) in zip(f1, f2):
        a, b = np.ascontiguousarray(d1.to_host()), np.ascontiguousarray(d2.to_host())
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (elt, shape, cx)


@pytest.mark.parametrize("elt,mod", [(ps.Float64, M64), (ps.Float32, M32), (ps.BFloat16, M32)])
@pytest.mark.parametrize("shape", [(130, 7, 6), (300, 2, 3), (9, 40, 5)])
@pytest.mark.parametrize("width", [1, 8, 64, 256])
def test_x_group_block_shapes_match_one_cell_per_thread(elt, mod, shape, width):
    # the launched block of x groups: `width` threads along x, the rest along y (clamped to the
    # y extent, which then widens x again)
    rng = np.random.default_rng(sum(shape) + width)
    ref = ps.core._define("parallel", STENCIL, {"cells_x": 1}, mod)
    grp = ps.core._define("parallel", STENCIL, {"cells_x": 2, "nthreads_x_groups": width}, mod)
    f1 = _fields(shape, elt, rng)
    f2 = [(a, ps.zeros(*shape, eltype=elt)) for a, _ in f1]
    for (_, d1), (_, d2) in zip(f1, f2):
        d2.copy_from(d1.to_host())
    for it in range(2):
        ref(*[d for _, d in f1], 0.37, 1e-3 if it == 1 else 0.9)
        grp(*[d for _, d in f2], 0.37, 1e-3 if it == 1 else 0.9)
    for (_, d1), (_, d2) in zip(f1, f2):
        a, b = np.ascontiguousarray(d1.to_host()), np.ascontiguousarray(d2.to_host())
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (elt, shape, width)


# row split: the interior (every x mask holds) is [5, nx - 60], 64 edge cells per row; with a
# second statement under ix <= 3 no cell of the row passes both masks (no interior); a float
# bound leaves the kernel without the split
EDGES = """
@parallel_indices (ix, iy, iz) function edges!(A, B, C, s::Data.Number)
    if ix >= 5 && ix <= size(A, 1) - 60
        A[ix, iy, iz] = B[ix, iy, iz] * s + C[ix + 1, iy, iz]
    end
    return
end
"""
NO_INTERIOR = """
@parallel_indices (ix, iy, iz) function nointerior!(A, B, C, D, s::Data.Number)
    if ix >= 5 && ix <= size(A, 1) - 60
        A[ix, iy, iz] = B[ix, iy, iz] * s + C[ix + 1, iy, iz]
    end
    if ix <= 3
        D[ix, iy, iz] = B[ix, iy, iz] - s
    end
    return
end
"""
FLOAT_BOUND = """
@parallel_indices (ix, iy, iz) function halfrow!(A, B, C, s::Data.Number)
    if ix <= size(A, 1) / 2
        A[ix, iy, iz] = B[ix, iy, iz] * s + C[ix + 1, iy, iz]
    end
    return
end
"""


@pytest.mark.parametrize("elt,mod", [(ps.Float64, M64), (ps.Float32, M32), (ps.BFloat16, M32)])
@pytest.mark.parametrize("src,nf", [(EDGES, 3), (NO_INTERIOR, 4), (FLOAT_BOUND, 3)])
@pytest.mark.parametrize("cx,width", [(2, 8), (4, 64), (8, 4)])
def test_x_group_row_split_edges(elt, mod, src, nf, cx, width):
    # more edge cells than edge threads (the edge threads loop), rows without interior, and the
    # unsplit layout for a float bound: the same fields as one cell per thread
    shape = (300, 40, 3)
    rng = np.random.default_rng(cx * 7 + width + nf)
    ref = ps.core._define("parallel_indices", src, {"cells_x": 1}, mod)
    grp = ps.core._define("parallel_indices", src, {"cells_x": cx, "nthreads_x_groups": width}, mod)
    f1 = _fields(shape, elt, rng)[:nf]
    f2 = [(a, ps.zeros(*shape, eltype=elt)) for a, _ in f1]
    for (_, d1), (_, d2) in zip(f1, f2):
        d2.copy_from(d1.to_host())
    box = ((2, 299), (1, 40), (1, 3))  # the launch box starts at x index 2
    ref(*[d for _, d in f1], 0.37, ranges=box)
    grp(*[d for _, d in f2], 0.37, ranges=box)
    for (_, d1), (_, d2) in zip(f1, f2):
        a, b = np.ascontiguousarray(d1.to_host()), np.ascontiguousarray(d2.to_host())
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (elt, cx, width)


@pytest.mark.parametrize("elt,mod", [(ps.Float64, M64), (ps.Float32, M32), (ps.BFloat16, M32)])
@pytest.mark.parametrize("shape", [(130, 7, 6), (67, 5, 9), (33, 8, 11)])
@pytest.mark.parametrize("cx,cz", [(1, 2), (4, 2), (2, 4), (8, 2), (1, 8)])
def test_z_groups_match_one_cell_per_thread(elt, mod, shape, cx, cz):
    # groups of cx x cz cells: z extents that cz does not divide, z neighbours (@d2_zi) shared
    # between a thread's cells
    rng = np.random.default_rng(sum(shape) + 10 * cx + cz)
    ref = ps.core._define("parallel", STENCIL, {"cells_x": 1, "cells_z": 1}, mod)
    grp = ps.core._define("parallel", STENCIL, {"cells_x": cx, "cells_z": cz}, mod)
    f1 = _fields(shape, elt, rng)
    f2 = [(a, ps.zeros(*shape, eltype=elt)) for a, _ in f1]
    for (_, d1), (_, d2) in zip(f1, f2):
        d2.copy_from(d1.to_host())
    for it in range(3):
        ref(*[d for _, d in f1], 0.37, 1e-3 if it == 1 else 0.9)
        grp(*[d for _, d in f2], 0.37, 1e-3 if it == 1 else 0.9)
    for (_, d1), (_, d2) in zip(f1, f2):
        a, b = np.ascontiguousarray(d1.to_host()), np.ascontiguousarray(d2.to_host())
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (elt, shape, cx, cz)


@pytest.mark.parametrize("elt,mod", [(ps.Float32, M32), (ps.BFloat16, M32)])
@pytest.mark.parametrize("src,nf", [(EDGES, 3), (NO_INTERIOR, 4)])
@pytest.mark.parametrize("cz", [2, 4])
def test_z_group_row_split_edges(elt, mod, src, nf, cz):
    # row ends of groups with several z cells: each edge thread takes its cell's z cells
    shape = (300, 40, 7)
    rng = np.random.default_rng(cz + nf)
    ref = ps.core._define("parallel_indices", src, {"cells_x": 1, "cells_z": 1}, mod)
    grp = ps.core._define("parallel_indices", src, {"cells_x": 4, "cells_z": cz, "nthreads_x_groups": 8}, mod)
    f1 = _fields(shape, elt, rng)[:nf]
    f2 = [(a, ps.zeros(*shape, eltype=elt)) for a, _ in f1]
    for (_, d1), (_, d2) in zip(f1, f2):
        d2.copy_from(d1.to_host())
    box = ((2, 299), (1, 40), (2, 7))
    ref(*[d for _, d in f1], 0.37, ranges=box)
    grp(*[d for _, d in f2], 0.37, ranges=box)
    for (_, d1), (_, d2) in zip(f1, f2):
        a, b = np.ascontiguousarray(d1.to_host()), np.ascontiguousarray(d2.to_host())
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (elt, cz)


def test_x_groups_acoustic3d_bf16_against_oracle_run():
    # the acoustic kernels pick x groups of 4 cells by themselves for FP32 and BFloat16;
    # the reported geometry stays the reference's
    n, steps = 45, 20
    P0, prm = W.acoustic3d_inputs(n)
    r16 = W.acoustic3d_run(M32, P0, prm, steps, ps.BFloat16)
    r1 = W.acoustic3d_run(M32, P0, prm, steps, ps.BFloat16, cells_x=1)
    for a, b in zip(r16, r1):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    F = [ps.ArgSpec((n + 1, n, n), ps.BFloat16), ps.ArgSpec((n, n + 1, n), ps.BFloat16), ps.ArgSpec((n, n, n + 1), ps.BFloat16),
         ps.ArgSpec((n, n, n), ps.BFloat16)]
    g = W.define_acoustic3d(M32, False)[0].geometry(*F, 0.1, 1.0, 1.0, 1.0, 1.0)
    g1 = W.define_acoustic3d(M32, False, cells_x=1)[0].geometry(*F, 0.1, 1.0, 1.0, 1.0, 1.0)
    assert g == g1, (g, g1)
