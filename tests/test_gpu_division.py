"""Division by a uniform divisor (scalar argument or literal) must be the IEEE quotient bit for bit:
the generated code divides through a precomputed reciprocal with one fma correction and falls back
to plain division for divisors and magnitudes outside the exact range."""
import numpy as np
import pytest

import parallelstencil as ps

pytestmark = pytest.mark.gpu

ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=__name__)
MOD32 = __name__ + ".f32"
ps.init_parallel_stencil(ps.AMDGPU, ps.Float32, 3, module=MOD32)


def _same(got, ref, utype):
    """Bit-identical except that any NaN matches any NaN (NaN payloads are not specified by IEEE)."""
    gn, rn = np.isnan(got), np.isnan(ref)
    return np.array_equal(gn, rn) and np.array_equal(got[~gn].view(utype), ref[~rn].view(utype))


def _wide(shape, rng, dtype=np.float64):
    emax = 300 if dtype == np.float64 else 30
    m = rng.random(shape) + 0.5
    e = rng.integers(-emax, emax, size=shape)
    s = np.where(rng.random(shape) < 0.5, -1.0, 1.0)
    a = (s * np.ldexp(m, e)).astype(dtype)
    flat = a.reshape(-1)
    flat[:4] = [0.0, -0.0, np.inf, -np.inf]
    flat[4] = np.nan
    # subnormal dividends and dividends next to the overflow threshold
    if dtype == np.float64:
        flat[5:10] = [5e-324, -1e-310, 2.2250738585072014e-308, 1.7e308, -1.0e308]
    else:
        flat[5:10] = [1e-45, -1e-40, 1.1754944e-38, 3.3e38, -1.0e38]
    return np.asfortranarray(a)


DIVISORS64 = [10.0 / 512, 40.0 / 383, 3.0, 0.1, 4.1, 1e-300, 1e300, 5e-324, -7.25,
              float.fromhex("0x1.fffffffffffffp+0"), float.fromhex("0x1.fffffffffffffp-3"), np.inf, 0.0]


@pytest.mark.parametrize("b", DIVISORS64)
def test_scalar_divisor_f64(b):
    k = ps.parallel("@parallel divs!(R, A, b) = (@all(R) = @all(A)/b; return)")
    a = _wide((37, 11, 5), np.random.default_rng(1))
    A = ps.zeros(*a.shape)
    A.copy_from(a)
    R = ps.zeros(*a.shape)
    k(R, A, float(b))
    with np.errstate(all="ignore"):
        ref = a / b
    got = R.to_host()
    assert _same(got, ref, np.uint64)


def test_literal_divisors_f64():
    k = ps.parallel("@parallel divl!(R, A) = (@all(R) = @all(A)/4.1 + @all(A)/0.1 - @all(A)/3.0; return)")
    a = _wide((64, 9, 3), np.random.default_rng(2))
    A = ps.zeros(*a.shape)
    A.copy_from(a)
    R = ps.zeros(*a.shape)
    k(R, A)
    with np.errstate(all="ignore"):
        ref = (a / 4.1 + a / 0.1) - a / 3.0
    assert _same(R.to_host(), ref, np.uint64)


@pytest.mark.parametrize("b", [0.3, 3.0, 1e-30, float.fromhex("0x1.fffffep+0"), 1e-40])
def test_scalar_divisor_f32(b):
    k = ps.core._define("parallel", "@parallel divs!(R, A, b::Data.Number) = (@all(R) = @all(A)/b + @all(A)/0.7; return)", {}, MOD32)
    a = _wide((33, 7, 4), np.random.default_rng(3), np.float32)
    A = ps.zeros(*a.shape, eltype=ps.Float32)
    A.copy_from(a)
    R = ps.zeros(*a.shape, eltype=ps.Float32)
    k(R, A, b)
    with np.errstate(all="ignore"):
        ref = (a / np.float32(b)) + (a / np.float32(0.7))
    assert _same(R.to_host(), ref.astype(np.float32), np.uint32)


HOIST = """@parallel function hoist!(R, A, a::Data.Number, b::Data.Number, n::Int)
    @all(R) = @all(A)*(a/b) + (n*a - b^2)/(3*pi) + (-a)^-2 + a^3/n - @all(A)/(2*b) + @all(A)*(a*b^-1)
    return
end"""


def _hoist_ref(x, a, b, n, t):
    """The statement of HOIST evaluated left to right in type t with numpy's IEEE scalars."""
    a, b, n3, pi = t(a), t(b), t(n), t(np.pi)
    ab = a / b
    c = (n3 * a - b * b) / (t(3) * pi)
    d = (t(1) / -a) * (t(1) / -a)
    e = (a * a * a) / n3
    f = x / (t(2) * b)
    g = a * (t(1) / b)
    with np.errstate(all="ignore"):
        return ((((x * ab + c) + d) + e) - f) + x * g


@pytest.mark.parametrize("mod,t,ut", [(__name__, np.float64, np.uint64), (MOD32, np.float32, np.uint32)])
@pytest.mark.parametrize("a,b,n", [(0.3, 7.1, 5), (-2.5, 1e-3, -3), (1e-20, 3e15, 1)])
def test_hoisted_scalar_subexpressions(mod, t, ut, a, b, n):
    # a/b, (n*a - b^2)/(3*pi), (-a)^-2, a^3/n, 2*b and a*b^-1 are evaluated once by the host; every
    # cell must still see the value its own IEEE evaluation would give
    k = ps.core._define("parallel", HOIST, {}, mod)
    x = _wide((29, 6, 5), np.random.default_rng(5), t)
    A = ps.zeros(*x.shape, eltype=ps.Float64 if t is np.float64 else ps.Float32)
    A.copy_from(x)
    R = ps.zeros(*x.shape, eltype=ps.Float64 if t is np.float64 else ps.Float32)
    k(R, A, a, b, n)
    ref = _hoist_ref(x, a, b, n, t).astype(t)
    assert _same(R.to_host(), ref, ut)


def test_bf16_stores_round_to_nearest_even():
    # FP32 -> BFloat16 stores: nearest even at every magnitude (ties, subnormals, the overflow
    # threshold) and every NaN to 0x7fc0, as BFloat16(::Float32) of BFloat16s.jl
    special = [0x3F808000, 0x3F818000, 0x3F808001, 0x3F807FFF, 0x00000001, 0x00008000, 0x00018000, 0x007FFFFF, 0x00007FFF,
               0x7F7FFFFF, 0x7F7F7FFF, 0x7F7F8000, 0x7F800000, 0xFF800000, 0x7FC00000, 0x7F800001, 0xFFC00001, 0x7FFFFFFF,
               0xBF808000, 0xBF818000, 0x80008000, 0x80018000, 0x00000000, 0x80000000]
    rng = np.random.default_rng(17)
    u = rng.integers(0, 2 ** 32, size=(97, 31, 13), dtype=np.uint64).astype(np.uint32)
    u.reshape(-1)[:len(special)] = special
    a = np.asfortranarray(u.view(np.float32))
    k = ps.core._define("parallel", "@parallel function cvt!(B, A)\n @all(B) = @all(A)\n return\nend", {}, MOD32)
    A = ps.zeros(*a.shape, eltype=ps.Float32)
    A.copy_from(a)
    B = ps.zeros(*a.shape, eltype=ps.BFloat16)
    k(B, A)
    got = np.asarray(B.to_host()).view(np.uint16)
    assert np.array_equal(got, ps.f32_to_bf16(a))


@pytest.mark.parametrize("mod,t,ut", [(__name__, np.float64, np.uint64), (MOD32, np.float32, np.uint32)])
def test_quotients_under_comparisons_keep_per_quotient_checks(mod, t, ut):
    # a comparison can hide a NaN quotient, so this statement tests each quotient itself;
    # infinite dividends must still give IEEE results
    k = ps.core._define("parallel", "@parallel divc!(R, A, b::Data.Number) = (@all(R) = (@all(A)/b > 1.0) ? @all(A)/b : -@all(A)/b; return)",
                        {}, mod)
    a = _wide((31, 9, 4), np.random.default_rng(9), t)
    A = ps.zeros(*a.shape, eltype=ps.Float64 if t is np.float64 else ps.Float32)
    A.copy_from(a)
    R = ps.zeros(*a.shape, eltype=ps.Float64 if t is np.float64 else ps.Float32)
    b = t(0.3)
    k(R, A, float(b))
    with np.errstate(all="ignore"):
        q = a / b
        ref = np.where(q > 1.0, q, -q).astype(t)
    assert _same(R.to_host(), ref, ut)
