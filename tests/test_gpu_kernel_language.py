"""GPU tests of the kernel language of @parallel_indices kernels
(src/ParallelKernel/kernel_language.jl) and of @overlap (src/ParallelKernel/overlap.jl).

The shared-memory diffusion example (examples/diffusion2D_shmem_novis.jl:11-27) is checked
bit-exactly against the oracle's restatement. The reference's CPU backend degenerates the
wavefront primitives to identities (kernel_language.jl:630-680), so their GPU results are checked
against the oracle's 64-lane wavefront model and closed forms (wave sums, ballot bits)."""
import ctypes

import numpy as np
import pytest

import parallelstencil as ps
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=__name__)

SHMEM2D = """
@parallel_indices (ix,iy) function diffusion2D_step!(T2, T, Ci, lam, dt, _dx, _dy)
    tx  = @threadIdx().x + 1
    ty  = @threadIdx().y + 1
    T_l = @sharedMem(eltype(T), (@blockDim().x+2, @blockDim().y+2))
    T_l[tx,ty] = T[ix,iy]
    if (1<ix<size(T2,1) && 1<iy<size(T2,2))
        if (@threadIdx().x == 1)             T_l[tx-1,ty] = T[ix-1,iy] end
        if (@threadIdx().x == @blockDim().x) T_l[tx+1,ty] = T[ix+1,iy] end
        if (@threadIdx().y == 1)             T_l[tx,ty-1] = T[ix,iy-1] end
        if (@threadIdx().y == @blockDim().y) T_l[tx,ty+1] = T[ix,iy+1] end
        @sync_threads()
        T2[ix,iy] = T_l[tx,ty] + dt*(Ci[ix,iy]*(
                    - ((-lam*(T_l[tx+1,ty] - T_l[tx,ty])*_dx) - (-lam*(T_l[tx,ty] - T_l[tx-1,ty])*_dx))*_dx
                    - ((-lam*(T_l[tx,ty+1] - T_l[tx,ty])*_dy) - (-lam*(T_l[tx,ty] - T_l[tx,ty-1])*_dy))*_dy
                    ));
    end
    return
end
"""


def _dev(a, eltype=None):
    et = eltype or {np.dtype(np.float64): ps.Float64, np.dtype(np.int64): ps.Int64, np.dtype(np.bool_): ps.Bool}[a.dtype]
    return ps.zeros(*a.shape, eltype=et).copy_from(a)


@pytest.mark.parametrize("n,threads,steps", [((64, 64), (32, 8), 3), ((512, 256), (32, 8), 10), ((256, 256), (64, 4), 5)])
def test_diffusion2d_shmem_example(n, threads, steps):
    print(f"diffusion2D shmem n={n} threads={threads}", flush=True)
    rng = np.random.default_rng(7)
    nx, ny = n
    T0 = np.asfortranarray(rng.random((nx, ny)) + 1.0)
    Ci = np.asfortranarray(0.5 + 0.1 * rng.random((nx, ny)))
    lam, dx, dy = 1.0, 1.0 / (nx - 1), 1.0 / (ny - 1)
    _dx, _dy = 1.0 / dx, 1.0 / dy
    dt = min(dx * dx, dy * dy) / lam / Ci.max() / 4.1
    k = ps.parallel_indices(SHMEM2D)
    T, T2, C = _dev(T0), _dev(T0), _dev(Ci)
    blocks = (nx // threads[0], ny // threads[1])
    shmem = (threads[0] + 2) * (threads[1] + 2) * 8
    a, b = T0.copy(order="F"), T0.copy(order="F")
    for _ in range(steps):
        k(T2, T, C, lam, dt, _dx, _dy, nblocks=blocks, nthreads=threads, shmem=shmem)
        T, T2 = T2, T
        O.diffusion2d_shmem_step(b, a, Ci, lam, dt, _dx, _dy)
        a, b = b, a
    assert np.array_equal(T.to_host(), a)
    # inbounds=true (@inbounds over the body): the same launches without bounds checks, after the
    # checked runs above have shown that every access of this geometry is inside its array
    ki = ps.parallel_indices(SHMEM2D, inbounds=True)
    T, T2 = _dev(T0), _dev(T0)
    for _ in range(steps):
        ki(T2, T, C, lam, dt, _dx, _dy, nblocks=blocks, nthreads=threads, shmem=shmem)
        T, T2 = T2, T
    assert np.array_equal(T.to_host(), a)


WARP = """
@parallel_indices (ix) function warp!(Bsh, Bup, Bdn, Bx, Bx16, Bsum, Bl, Bw, Bany, Ball, Bbal, Bt, Bb, Bd, Bg, A, L, P)
    m = @active_mask()
    l = @laneid()
    Bl[ix] = l
    Bw[ix] = @warpsize()
    Bt[ix] = @threadIdx().x
    Bb[ix] = @blockIdx().x
    Bd[ix] = @blockDim().x
    Bg[ix] = @gridDim().x
    v = A[ix]
    Bsh[ix] = @shfl_sync(m, v, L[ix], 16)
    Bup[ix] = @shfl_up_sync(m, v, 3)
    Bdn[ix] = @shfl_down_sync(m, v, 5, 32)
    Bx[ix] = @shfl_xor_sync(m, v, 5, 8)
    Bx16[ix] = @shfl_xor_sync(m, v, 24, 16)
    s = v
    s = s + @shfl_xor_sync(m, s, 32)
    s = s + @shfl_xor_sync(m, s, 16)
    s = s + @shfl_xor_sync(m, s, 8)
    s = s + @shfl_xor_sync(m, s, 4)
    s = s + @shfl_xor_sync(m, s, 2)
    s = s + @shfl_xor_sync(m, s, 1)
    Bsum[ix] = s
    p = P[ix]
    Bany[ix] = @vote_any_sync(m, p)
    Ball[ix] = @vote_all_sync(m, p)
    Bbal[ix] = @vote_ballot_sync(m, p)
    return
end
"""


def test_wavefront_primitives():
    n, nt = 512, 128
    print(f"wavefront primitives n={n} nthreads={nt}", flush=True)
    rng = np.random.default_rng(3)
    A = rng.integers(-1000, 1000, n).astype(np.float64)
    L = rng.integers(1, 17, n).astype(np.int64)
    P = rng.random(n) < 0.5
    P[128:192] = True   # a wavefront where the predicate holds everywhere
    P[192:256] = False  # and one where it holds nowhere
    k = ps.parallel_indices(WARP)
    outs_f = [ps.zeros(n) for _ in range(6)]
    outs_i = [ps.zeros(n, eltype=ps.Int64) for _ in range(2)]
    Bany, Ball = ps.falses(n), ps.falses(n)
    Bbal = ps.zeros(n, eltype=ps.Int64)
    ids = [ps.zeros(n, eltype=ps.Int64) for _ in range(4)]
    k(*outs_f, *outs_i, Bany, Ball, Bbal, *ids, _dev(A), _dev(L), _dev(P), nblocks=(n // nt,), nthreads=(nt,))
    Bsh, Bup, Bdn, Bx, Bx16, Bsum = [o.to_host() for o in outs_f]
    Bl, Bw = [o.to_host() for o in outs_i]
    Bt, Bb, Bd, Bg = [o.to_host() for o in ids]
    i = np.arange(n)
    assert np.array_equal(Bt, i % nt + 1) and np.array_equal(Bb, i // nt + 1)
    assert np.all(Bd == nt) and np.all(Bg == n // nt)
    assert np.all(Bw == 64) and np.array_equal(Bl, i % 64 + 1)
    for w in range(n // 64):
        s = slice(64 * w, 64 * (w + 1))
        v = A[s]
        assert np.array_equal(Bsh[s], O.wave_shfl(v, L[s], 16)), w
        assert np.array_equal(Bup[s], O.wave_shfl_up(v, 3, 64)), w
        assert np.array_equal(Bdn[s], O.wave_shfl_down(v, 5, 32)), w
        assert np.array_equal(Bx[s], O.wave_shfl_xor(v, 5, 8)), w
        assert np.array_equal(Bx16[s], O.wave_shfl_xor(v, 24, 16)), w
        assert np.all(Bsum[s] == v.sum()), w
        p = P[s]
        assert np.all(Bany.to_host()[s] == p.any()) and np.all(Ball.to_host()[s] == p.all()), w
        assert np.all(Bbal.to_host()[s] == O.wave_ballot(p)), w


BLOCKSUM = """
@parallel_indices (ix) function blocksum!(S, C, A)
    t = @threadIdx().x
    n = @blockDim().x
    buf = @sharedMem(eltype(A), n)
    cnt = @sharedMem(Int64, (2,), n*8)
    buf[t] = A[ix]
    if t == 1
        cnt[1] = n
        cnt[2] = @blockIdx().x
    end
    @sync_threads()
    for k = 1:8
        h = n ÷ (2^k)
        if t <= h
            buf[t] = buf[t] + buf[t + h]
        end
        @sync_threads()
    end
    if t == 1
        S[@blockIdx().x] = buf[1]
    end
    C[ix] = cnt[1] * 1000 + cnt[2]
    return
end
"""


def test_shared_arrays_with_offsets_and_barriers():
    nt, nb = 256, 24
    n = nt * nb
    print(f"block sums n={n}", flush=True)
    rng = np.random.default_rng(5)
    A = rng.integers(-50, 50, n).astype(np.float64)
    k = ps.parallel_indices(BLOCKSUM)
    S, C = ps.zeros(nb), ps.zeros(n, eltype=ps.Int64)
    k(S, C, _dev(A), nblocks=(nb,), nthreads=(nt,), shmem=nt * 8 + 16)
    assert np.array_equal(S.to_host(), A.reshape(nb, nt).sum(axis=1))
    # the Int64 array lives nt*8 bytes in: it must not alias buf
    assert np.array_equal(C.to_host(), nt * 1000 + np.arange(n) // nt + 1)


def test_shared_array_bounds_are_checked():
    # indices past the @sharedMem extents, and an array larger than the launch's dynamic LDS, are
    # caught by the bounds-checked variant (the access does not happen) and raised on the host
    k = ps.parallel_indices("""
    @parallel_indices (ix) function over!(B, A)
        buf = @sharedMem(eltype(A), @blockDim().x)
        buf[@threadIdx().x + 1] = A[ix]
        B[ix] = 1.0
        return
    end""")
    A, B = ps.ones(128), ps.zeros(128)
    with pytest.raises(ps.BoundsError):
        k(B, A, nblocks=(2,), nthreads=(64,), shmem=64 * 8 + 64)
    k2 = ps.parallel_indices("""
    @parallel_indices (ix) function big!(B, A)
        buf = @sharedMem(eltype(A), 2*@blockDim().x)
        buf[@threadIdx().x] = A[ix]
        B[ix] = buf[@threadIdx().x]
        return
    end""")
    with pytest.raises(ps.BoundsError):
        k2(B, A, nblocks=(2,), nthreads=(64,), shmem=64 * 8)
    k2(B, A, nblocks=(2,), nthreads=(64,), shmem=2 * 64 * 8)
    assert np.all(B.to_host() == 1.0)


def test_missing_shmem_is_an_error():
    k = ps.parallel_indices(SHMEM2D)
    T = ps.zeros(64, 64)
    with pytest.raises(ps.ArgumentError, match="shmem"):
        k(T, T, T, 1.0, 0.1, 1.0, 1.0, nblocks=(2, 8), nthreads=(32, 8))


def test_pk_println_and_show(capfd):
    k = ps.parallel_indices("""
    @parallel_indices (ix) function hello!(A)
        if ix == 2
            @pk_println("ix = ", ix, ", a = ", A[ix], ", ", A[ix] > 0.0)
            @pk_show(ix, A[ix])
        end
        return
    end""")
    A = _dev(np.array([0.5, 0.25, 3.0]))
    k(A)
    ps.synchronize()
    ctypes.CDLL(None).fflush(None)  # device printf goes through the C stdio buffer of the process
    out = capfd.readouterr().out
    assert "ix = 2, a = 0.25, true" in out
    assert "ix = 2" in out and "A[ix] = 0.25" in out


def test_overlap_runs_each_call_on_its_own_stream():
    k = ps.parallel_indices("""
    @parallel_indices (ix,iy) function axpy!(Y, X, a)
        Y[ix,iy] = Y[ix,iy] + a*X[ix,iy]
        return
    end""")
    rng = np.random.default_rng(11)
    xs = [rng.random((300, 200)) for _ in range(3)]
    ys = [rng.random((300, 200)) for _ in range(3)]
    X = [_dev(x) for x in xs]
    Y = [_dev(y) for y in ys]
    with ps.overlap():
        for i in range(3):
            k(Y[i], X[i], float(i + 1))
    for i in range(3):
        assert np.array_equal(Y[i].to_host(), ys[i] + (i + 1) * xs[i])
    ps.overlap((k, (Y[0], X[0], 2.0)), (k, (Y[1], X[1], 3.0), {"ranges": ((1, 100), (1, 200))}))
    ref0 = ys[0] + 1 * xs[0] + 2.0 * xs[0]
    ref1 = ys[1] + 2 * xs[1]
    ref1[:100] = ref1[:100] + 3.0 * xs[1][:100]
    assert np.array_equal(Y[0].to_host(), ref0)
    assert np.array_equal(Y[1].to_host(), ref1)
