"""ps_field_reduce_noalloc (no counterpart in the reference; for callers that cannot allocate device
memory): called through the C-ABI, checked against ps_field_reduce on every field and every
statistic, on two calls in a row whose second has fewer and smaller fields, so stale partials of
the first call would show; plus the limit on the number of fields."""
import ctypes

import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil._lib import lib

pytestmark = pytest.mark.gpu

ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=__name__)
OPS = ("sum", "sum_abs", "max_abs", "mean_abs", "norm")


def _fields(shapes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for shape, et in shapes:
        a = rng.standard_normal(shape)
        A = ps.zeros(*shape, eltype=et)
        A.copy_from(a.astype(np.float32) if et == ps.Float32 else a)
        out.append(A)
    return out


def _noalloc(fields, op, out):
    arr = (ps.core.ps_field * len(fields))(*[f.f for f in fields])
    rc = lib.ps_field_reduce_noalloc(arr, len(fields), op, ctypes.c_void_p(out.ptr), None)
    assert rc == 0, lib.ps_last_error()
    ps.synchronize()
    return out.to_host()[:len(fields)]


def test_noalloc_reduce_matches_ps_field_reduce():
    out = ps.zeros(4)
    first = _fields([((513, 300, 7), ps.Float64), ((64, 64, 64), ps.Float32), ((1000, 3), ps.Float64), ((9,), ps.Float64)], 1)
    second = _fields([((31, 17, 5), ps.Float64), ((40, 2), ps.Float32)], 2)
    for fields in (first, second):  # the second call reduces fewer, smaller fields
        for code, op in enumerate(OPS):
            got = _noalloc(fields, code, out)
            want = np.array([f.reduce(op) for f in fields])
            if op == "norm":
                assert np.allclose(got, want, rtol=2e-16, atol=0), (op, got, want)
            else:
                assert np.array_equal(got, want), (op, got, want)


def test_noalloc_reduce_limits():
    out = ps.zeros(8)
    fields = _fields([((8, 8), ps.Float64)] * 5, 3)
    arr = (ps.core.ps_field * 5)(*[f.f for f in fields])
    assert lib.ps_field_reduce_noalloc(arr, 5, 0, ctypes.c_void_p(out.ptr), None) == -1
    assert b"nfields must be 1..4" in lib.ps_last_error()
    assert lib.ps_field_reduce_noalloc(arr, 0, 0, ctypes.c_void_p(out.ptr), None) == -1
    assert lib.ps_field_reduce_noalloc(arr, 2, 7, ctypes.c_void_p(out.ptr), None) == -1
    assert lib.ps_field_reduce_noalloc(arr, 2, 0, None, None) == -1
