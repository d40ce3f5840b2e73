"""GPU parity of the finite-difference macros and masks against the reference's
known-answer formulas (test/test_FiniteDifferences{1,2,3}D.jl), exact equality."""
import numpy as np
import pytest

import parallelstencil as ps
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=__name__)

NXYZ = (9, 7, 8)  # the reference's grid size for the 3-D KATs
R3 = tuple(n - 2 for n in NXYZ)  # IField


def _dev(a):
    A = ps.zeros(*a.shape)
    A.copy_from(a)
    return A


def _rand(shape, seed):
    return np.asfortranarray(np.random.default_rng(seed).random(shape) + 0.25)


_K3 = {}


def _k3(name):
    if name not in _K3:
        _K3[name] = ps.parallel(f"@parallel {name}!(R, A) = (@all(R) = @{name}(A); return)", ndims=3)
    return _K3[name]


@pytest.mark.parametrize("name", sorted(O.FD3_ARG_OFFSET))
def test_fd3_macro(name):
    off = O.FD3_ARG_OFFSET[name]
    a = _rand(tuple(R3[i] + off[i] for i in range(3)), seed=hash(name) % 1000)
    R = ps.zeros(*R3)
    _k3(name)(R, _dev(a))
    assert np.array_equal(R.to_host(), O.fd3(name, a)), name


@pytest.mark.parametrize("name", ["d2_xi", "d_yi", "av", "inn"])
def test_fd3_macro_cubed_indices(name):
    # the closed-form integer input of the reference's execution KATs
    off = O.FD3_ARG_OFFSET[name]
    a = O.cubed_indices(tuple(R3[i] + off[i] for i in range(3)))
    R = ps.zeros(*R3)
    _k3(name)(R, _dev(a))
    assert np.array_equal(R.to_host(), O.fd3(name, a))


def test_fd3_masks_leave_boundaries_untouched():
    # test_FiniteDifferences3D.jl:165-190
    A = _rand(R3, 1)
    Axxyyzz = _rand(NXYZ, 2)
    Ax = _rand((R3[0] + 1, R3[1], R3[2]), 3)
    Axxyzz = _rand((NXYZ[0], NXYZ[1] - 1, NXYZ[2]), 4)
    k_all = ps.parallel("@parallel inn_all!(Rxxyyzz, A) = (@inn(Rxxyyzz) = @all(A); return)")
    k_inn = ps.parallel("@parallel inn_inn!(Rxxyyzz, Axxyyzz) = (@inn(Rxxyyzz) = @inn(Axxyyzz); return)")
    k_dxa = ps.parallel("@parallel inn_d_xa!(Rxxyyzz, Ax) = (@inn(Rxxyyzz) = @d_xa(Ax); return)")
    k_dyi = ps.parallel("@parallel inn_d_yi!(Rxxyyzz, Axxyzz) = (@inn(Rxxyyzz) = @d_yi(Axxyzz); return)")
    k_d2yi = ps.parallel("@parallel inn_d2_yi!(Rxxyyzz, Axxyyzz) = (@inn(Rxxyyzz) = @d2_yi(Axxyyzz); return)")
    for k, a, expect in ((k_all, A, A), (k_inn, Axxyyzz, O.fd3("inn", Axxyyzz)), (k_dxa, Ax, O.fd3("d_xa", Ax)),
                         (k_dyi, Axxyzz, O.fd3("d_yi", Axxyzz)), (k_d2yi, Axxyyzz, O.fd3("d2_yi", Axxyyzz))):
        Rf = ps.zeros(*NXYZ)
        k(Rf, _dev(a))
        r = Rf.to_host()
        assert np.array_equal(r[1:-1, 1:-1, 1:-1], expect)
        r[1:-1, 1:-1, 1:-1] = 0
        assert np.all(r == 0)


def test_fd2_macros():
    ps.reset_parallel_stencil(module=__name__ + ".fd2")
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=__name__ + ".fd2")
    nxy = (9, 7)
    R2 = (7, 5)
    for i, name in enumerate(sorted(O.FD2_ARG_OFFSET)):
        off = O.FD2_ARG_OFFSET[name]
        a = _rand((R2[0] + off[0], R2[1] + off[1]), 100 + i)
        k = ps.core._define("parallel", f"{name}!(R, A) = (@all(R) = @{name}(A); return)", {}, __name__ + ".fd2")
        R = ps.zeros(*R2)
        A = ps.zeros(*a.shape)
        A.copy_from(a)
        k(R, A)
        assert np.array_equal(R.to_host(), O.fd2(name, a)), name
    Rf = ps.zeros(*nxy)
    k = ps.core._define("parallel", "inn_inn!(R, A) = (@inn(R) = @inn(A); return)", {}, __name__ + ".fd2")
    a = _rand(nxy, 5)
    A = ps.zeros(*nxy)
    A.copy_from(a)
    k(Rf, A)
    r = Rf.to_host()
    assert np.array_equal(r[1:-1, 1:-1], a[1:-1, 1:-1]) and np.all(r[0] == 0) and np.all(r[:, -1] == 0)


def test_fd1_macros():
    mod = __name__ + ".fd1"
    ps.reset_parallel_stencil(module=mod)
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 1, module=mod)
    R1 = 7
    for i, name in enumerate(sorted(O.FD1_ARG_OFFSET)):
        a = _rand((R1 + O.FD1_ARG_OFFSET[name],), 200 + i)
        k = ps.core._define("parallel", f"{name}!(R, A) = (@all(R) = @{name}(A); return)", {}, mod)
        R = ps.zeros(R1)
        A = ps.zeros(a.shape[0])
        A.copy_from(a)
        k(R, A)
        assert np.array_equal(R.to_host(), O.fd1(name, a)), name
