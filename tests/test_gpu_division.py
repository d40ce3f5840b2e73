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
