"""Runtime behaviour on the GPU: reductions keep NaN as Julia's maximum does, allocation is refused
while the library stream is being captured into a graph, freed fields never reach a kernel, and
graphs can be destroyed explicitly."""
import ctypes

import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil._lib import lib

pytestmark = pytest.mark.gpu

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD)

COPY = """
@parallel function copy2D!(A, B)
    @all(A) = @all(B)
    return
end
"""


def test_reductions_propagate_nan():
    # miniapps/Stokes2D.jl:119 takes maximum(abs.(R)): a diverged field gives NaN, not a finite value
    a = np.random.default_rng(0).random((300, 70)) - 0.5
    A = ps.zeros(300, 70)
    A.copy_from(a)
    assert A.reduce("max_abs") == np.abs(a).max()
    a[217, 35] = np.nan
    A.copy_from(a)
    for op in ("max_abs", "sum", "sum_abs", "mean_abs", "norm"):
        assert np.isnan(A.reduce(op)), op
    B = ps.zeros(5, 3)
    b = np.zeros((5, 3))
    b[4, 2] = np.nan
    B.copy_from(b)
    assert np.isnan(B.reduce("max_abs"))


def test_allocation_during_capture_is_refused():
    k = ps.parallel(COPY)
    A, B = ps.zeros(16, 8), ps.ones(16, 8)
    k(A, B)  # compiled before the capture
    g = ps.Graph()
    with pytest.raises(ps.ArgumentError, match="captured"):
        with g:
            k(A, B, async_=True)
            ps.zeros(16, 8)
    # the library stream is usable afterwards and the kernel still runs
    A.fill_(0.0)
    k(A, B)
    assert np.all(A.to_host() == 1.0)


def test_freed_fields_never_reach_a_kernel():
    k = ps.parallel(COPY)
    A, B = ps.zeros(16, 8), ps.ones(16, 8)
    k(A, B)
    run = k.launcher(A, B)
    run()
    ps.synchronize()
    B.free()
    with pytest.raises(ps.ArgumentError, match="freed"):
        k(A, B)
    with pytest.raises(ps.ArgumentError, match="freed"):
        run()
    # the C-ABI refuses a field without memory on its own
    args = (ps.core.ps_arg * 2)()
    for i, f in enumerate((A.f, B.f)):
        args[i].kind, args[i].dtype, args[i].field = ps.core.PS_ARG_FIELD, ps.core.PS_F64, f
    L = ps.core.ps_launch()
    L.memopt = -1
    rc = lib.ps_kernel_launch(k._h, args, 2, ctypes.byref(L))
    assert rc == -1 and b"no device memory" in lib.ps_last_error()


def test_graph_destroy():
    k = ps.parallel(COPY)
    A, B = ps.zeros(16, 8), ps.fill(3.0, 16, 8)
    k(A, B)
    A.fill_(0.0)
    g = ps.Graph()
    with g:
        k(A, B, async_=True)
    g.launch()
    ps.synchronize()
    assert np.all(A.to_host() == 3.0)
    g.destroy()
    with pytest.raises(ps.ArgumentError, match="destroyed"):
        g.launch()
    g.destroy()  # a second destroy does nothing
