"""CPU tests of derivative kernels: the generated code (atomics only where threads share a shadow
element, dropped primal writes for DuplicatedNoNeed), hiprtc compilation, and the keyword and
kernel errors of @parallel ∇=... (src/ParallelKernel/parallel.jl:265-301)."""
import re

import pytest

import parallelstencil as ps
from parallelstencil import workloads as W

MOD = __name__
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=MOD)
S3 = ps.ArgSpec((9, 8, 7))

F_REF = """
@parallel_indices (ix) function f!(A, B, a)
    A[ix] += a * B[ix] * 100.65
    return
end"""


def _k(text, ndims=1):
    if not ps.is_initialized(MOD + f".{ndims}"):
        ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, ndims, module=MOD + f".{ndims}")
    kind = "parallel" if text.lstrip().startswith("@parallel function") else "parallel_indices"
    return ps.core._define(kind, text, {}, MOD + f".{ndims}")


def _body(src):
    return src.split('extern "C"')[1]


def test_derivative_parameters_and_names():
    k = _k(F_REF)
    d = k._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})
    assert d.params == ["A", "B", "a", "A#shadow", "B#shadow"]
    assert d.name == "f!_adjoint" and not d.info["memopt"]
    assert set(d.info["written"]) == {"A#shadow", "B#shadow"}
    f = k._derivative("Forward", {"B": "Duplicated"})
    assert f.params == ["A", "B", "a", "B#shadow"] and f.name == "f!_tangent"
    assert set(f.info["written"]) == {"A"}  # A is Const: its update is kept; B's tangent reaches no shadow


def test_atomics_only_for_shared_shadow_elements():
    k = W.define_diffusion3d(True, MOD)
    d = k._derivative("Reverse", {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})
    src = _body(d.source(S3, S3, S3, 1.0, 0.1, 1.0, 1.0, 1.0, S3, S3))
    # T is read at 7 points: one atomic per distinct element, none for the seed of T2
    assert len(re.findall(r"unsafeAtomicAdd\(&a9\[", src)) == 7
    assert "unsafeAtomicAdd(&a8[" not in src
    # the seed of T2 is read and cleared where T2 is written
    assert re.search(r"= a8\[", src) and re.search(r"a8\[[^\]]*\] = 0x0p\+0;", src)
    # the primal arrays stay read-only (T2 DuplicatedNoNeed: no primal update)
    assert "const double* __restrict__ a0" in d.source(S3, S3, S3, 1.0, 0.1, 1.0, 1.0, 1.0, S3, S3)
    # reading only the thread's own element: plain read-modify-write
    r = _k(F_REF)._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})
    S1 = ps.ArgSpec((16,))
    assert "unsafeAtomicAdd" not in r.source(S1, S1, 6.5, S1, S1)
    # forward mode never needs atomics
    f = k._derivative("Forward", {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})
    assert "unsafeAtomicAdd" not in f.source(S3, S3, S3, 1.0, 0.1, 1.0, 1.0, 1.0, S3, S3)


def test_duplicated_keeps_the_primal_update():
    k = _k(F_REF)
    S1 = ps.ArgSpec((16,))
    nn = _body(k._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"}).source(S1, S1, 6.5, S1, S1))
    dd = _body(k._derivative("Reverse", {"A": "Duplicated", "B": "DuplicatedNoNeed"}).source(S1, S1, 6.5, S1, S1))
    assert not re.search(r"\ba0\[[^\]]*\] =", nn)
    assert re.search(r"\ba0\[[^\]]*\] =", dd)
    # the primal update comes last: after the shadow of B received its adjoint
    assert dd.rfind("a0[") > dd.rfind("a4[")


def test_derivative_kernels_compile_for_gfx950():
    k = W.define_diffusion3d(True, MOD)
    for mode in ("Reverse", "Forward"):
        d = k._derivative(mode, {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})
        d.precompile(S3, S3, S3, 1.0, 0.1, 1.0, 1.0, 1.0, S3, S3)
    S1, S2 = ps.ArgSpec((10,)), ps.ArgSpec((10, 3))
    acc = _k("""
    @parallel_indices (ix) function acc!(A, B, C)
        s = 0.0
        for k = 1:size(B,2)
            t = B[ix, k] * C[ix, k]
            s += t * t
        end
        A[ix] = sqrt(s) + (ix > 2 ? B[ix,1] : C[ix,2])
        return
    end""")
    acc._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed", "C": "DuplicatedNoNeed"}).precompile(S1, S2, S2, S1, S2, S2)


def test_reverse_mode_local_rules():
    twice = _k("""
    @parallel_indices (ix) function r!(A, B)
        x = B[ix]
        x = x * x
        A[ix] = x
        return
    end""")
    with pytest.raises(ps.ArgumentError, match="assigned more than once"):
        twice._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})
    twice._derivative("Forward", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})  # forward mode has no such rule
    early = _k("""
    @parallel_indices (ix) function r!(A, B)
        s = B[ix]
        y = 2.0 * s
        s += B[ix]
        A[ix] = y + s
        return
    end""")
    with pytest.raises(ps.ArgumentError, match="read before an update"):
        early._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})
    across = _k("""
    @parallel_indices (ix) function r!(A, B)
        s = 0.0
        for k = 1:3
            A[ix] = s
            s += B[ix]
        end
        return
    end""")
    with pytest.raises(ps.ArgumentError, match="read before an update"):  # iteration k+1 reads what k updated
        across._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})
    branches = _k("""
    @parallel_indices (ix) function r!(A, B)
        if ix > 2
            x = B[ix]
        else
            x = 2.0 * B[ix]
        end
        A[ix] = x
        return
    end""")
    branches._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})  # one definition per path


def test_read_after_write_is_rejected():
    k = _k("""
    @parallel_indices (ix) function r!(A, B)
        A[ix] = B[ix]
        B[ix] = A[ix] * 2.0
        return
    end""")
    for mode in ("Reverse", "Forward"):
        with pytest.raises(ps.ArgumentError, match="reads array 'A' after writing it"):
            k._derivative(mode, {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})


def test_kernel_language_and_unknown_functions_are_rejected():
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD + ".kl")
    shm = W.define_diffusion2d_shmem(MOD + ".kl")
    with pytest.raises(ps.ArgumentError, match="kernel language"):
        shm._derivative("Reverse", {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})
    k = _k("""
    @parallel_indices (ix) function r!(A, B)
        A[ix] = erfcx(B[ix])
        return
    end""")
    with pytest.raises(ps.ArgumentError, match="erfcx"):
        k._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})


def test_keyword_errors():
    k = _k(F_REF)
    A, B, Ab, Bb = (ps.ArgSpec((16,)) for _ in range(4))
    with pytest.raises(ps.IncoherentArgumentError, match="AD keywords are only valid"):
        k(A, B, 6.5, ad_mode=ps.Forward)
    with pytest.raises(ps.IncoherentArgumentError):
        k(A, B, 6.5, ad_annotations={"Duplicated": "A"})
    with pytest.raises(ps.KeywordArgumentError, match="is not a positional argument"):
        k(A, B, 6.5, grad={"C": Ab})
    with pytest.raises(ps.KeywordArgumentError, match="is not a positional argument"):
        k(A, B, 6.5, grad=(ps.ArgSpec((16,)), Ab))
    with pytest.raises(ps.KeywordArgumentError, match="annotation Mixed is not \\(yet\\) supported"):
        k(A, B, 6.5, grad=(A, Ab), ad_annotations={"Mixed": "A"})
    with pytest.raises(ps.KeywordArgumentError, match="more than one annotation"):
        k(A, B, 6.5, grad=[(A, Ab), (B, Bb)], ad_annotations={"Duplicated": ("A", "B"), "DuplicatedNoNeed": "B"})
    with pytest.raises(ps.KeywordArgumentError, match="has no duplicate"):
        k(A, B, 6.5, grad=(A, Ab), ad_annotations={"Duplicated": "B"})
    with pytest.raises(ps.KeywordArgumentError, match="ad_mode"):
        k(A, B, 6.5, grad=(A, Ab), ad_mode="Sideways")
    with pytest.raises(ps.KeywordArgumentError, match="invalid keyword"):
        k(A, B, 6.5, gradient=(A, Ab))


def test_gather_form_of_parallel_kernels():
    k = W.define_diffusion3d(True, MOD)
    d = k._derivative("Reverse", {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})
    assert d._gather is not None
    g, c = d._gather
    assert g.name.endswith("_adjoint_gather") and c.name.endswith("_adjoint_clear")
    args = (S3, S3, S3, 1.0, 0.1, 1.0, 1.0, 1.0, S3, S3)
    gs, cs = _body(g.source(*args)), _body(c.source(*args))
    assert "unsafeAtomicAdd" not in gs + cs
    # the gather kernel reads the seeds and never writes them; the clear kernel zeroes them
    assert not re.search(r"\ba8\[[^\]]*\] =", gs) and re.search(r"\ba9\[[^\]]*\] =", gs)
    assert re.search(r"\ba8\[[^\]]*\] = 0x0p\+0;", cs) and not re.search(r"\ba9\[", cs)
    # 7 neighbour contributions (one per read offset of T), each under the shifted mask
    assert gs.count("if (") >= 8
    g.precompile(*args)
    c.precompile(*args)
    # no gather form: @parallel_indices kernels, forward mode, arrays read and written
    assert _k(F_REF)._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})._gather is not None  # in-place
    assert k._derivative("Forward", {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})._gather is None
    # an in-place update (the element it writes is the only one of A it reads) keeps the gather
    # form: the clear kernel applies A's own derivative
    inplace = _k("""
    @parallel function upd!(A, B)
        @inn(A) = 2.0*@inn(A) + @d2(B)
        return
    end""")
    d = inplace._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})
    assert d._gather is not None
    S1 = ps.ArgSpec((20,))
    clear = _body(d._gather[1].source(S1, S1, S1, S1))
    assert re.search(r"a2\[[^\]]*\] = \(v_ad_23g\d+ \* 0x1p\+1\)", clear) or re.search(r"a2\[[^\]]*\] = \(0x1p\+1 \* v_ad_23g\d+\)", clear)
    neighbour = _k("""
    @parallel_indices (ix) function nb!(A, B)
        A[ix] = A[ix] + B[ix]*A[ix]
        return
    end""")
    assert neighbour._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})._gather is not None
    other = _k("""
    @parallel_indices (ix) function ot!(A, B, C)
        A[ix] = B[ix]
        C[ix] = A[ix] + 1.0
        return
    end""")
    with pytest.raises(ps.ArgumentError, match="after writing it"):
        other._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed", "C": "DuplicatedNoNeed"})


def test_memopt_primal_gives_streaming_derivatives():
    """A memopt primal gets a memopt gather kernel and a memopt tangent (both read their inputs at
    constant offsets of their own point); the atomic scatter form stays a plain kernel."""
    for memopt in (True, False):
        k = W.define_diffusion3d(memopt, MOD)
        r = k._derivative("Reverse", {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})
        f = k._derivative("Forward", {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})
        assert r._gather[0].info["memopt"] is memopt and f.info["memopt"] is memopt
        assert r.info["memopt"] is False and r._gather[1].info["memopt"] is False
    g = W.define_diffusion3d(True, MOD)._derivative("Reverse", {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"})._gather[0]
    assert set(g.info["optvars"]) == {"T2#shadow", "Cp"}
    assert g.info["stencilranges"]["T2#shadow"] == [[-1, 1], [-1, 1], [-1, 1]]


def test_gather_form_of_parallel_indices_kernels():
    """Loop-free @parallel_indices kernels get the gather form too: locals are inlined, a local
    set in both branches becomes an ifelse, and every contribution is guarded by the source
    point's existence in the launch box and by its path condition."""
    k = _k("""
    @parallel_indices (ix, iy) function g!(R, U, V)
        if ix > 1 && ix < size(U,1)
            a = U[ix+1,iy] - U[ix-1,iy]
        else
            a = 2.0*U[ix,iy]
        end
        R[ix,iy] = a*V[ix,iy]
        return
    end""", ndims=2)
    d = k._derivative("Reverse", {"R": "DuplicatedNoNeed", "U": "DuplicatedNoNeed", "V": "DuplicatedNoNeed"})
    assert d._gather is not None
    S = ps.ArgSpec((9, 7))
    g = _body(d._gather[0].source(S, S, S, S, S, S))
    assert "unsafeAtomicAdd" not in g and "ps_r0x" in g and "ps_rlx" in g
    loop = _k("""
    @parallel_indices (ix) function l!(A, B)
        for k = 1:2
            A[ix] = B[ix]
        end
        return
    end""")
    assert loop._derivative("Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"})._gather is None
    twice = _k("""
    @parallel_indices (ix) function t!(A, B)
        A[ix] = B[ix]
        if ix > 1
            A[ix] = 2.0*B[ix]
        end
        return
    end""")
    with pytest.raises(ps.ArgumentError, match="written twice on one path"):
        ps.core._DerivativeKernel(twice, "Reverse", {"A": "DuplicatedNoNeed", "B": "DuplicatedNoNeed"}, "gather")
