"""The memopt known-answer tests of the reference (test/test_parallel.jl:438-1060 for 3-D,
:1084-1195 for 2-D), on the closed-form inputs they use (cubed linear indices, scaled by 2 or 3
for the second and third array, staggered fields sized as the reference's @XXYYZField /
@XYYZZField). The expected fields are the reference's broadcast formulas restated in numpy in the
same operation order; the inputs are integers, so most are exact, and the rest round identically
because the kernels are compiled without contraction. Where the reference's 2-D tests use zero
fields, these use the cubed indices instead, which pins the stencils."""
import numpy as np
import pytest

import parallelstencil as ps
from oracle import ps_oracle as O

pytestmark = pytest.mark.gpu

ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=__name__)
MOD2 = __name__ + ".2d"
ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=MOD2)

N3 = (32, 8, 8)


def cub(shape, f=1.0):
    a = O.cubed_indices(shape)
    return np.asfortranarray(f * a.reshape(shape, order="F"))


def dev(a):
    return ps.zeros(*a.shape).copy_from(a)


def pi(text):
    return ps.parallel_indices(text)


def par(text):
    return ps.parallel(text)


def inn(x):
    return x[1:-1, 1:-1, 1:-1]


def d2x(T):
    c = inn(T)
    return (T[2:, 1:-1, 1:-1] - c) - (c - T[:-2, 1:-1, 1:-1])


def d2y(T):
    c = inn(T)
    return (T[1:-1, 2:, 1:-1] - c) - (c - T[1:-1, :-2, 1:-1])


def d2z(T):
    c = inn(T)
    return (T[1:-1, 1:-1, 2:] - c) - (c - T[1:-1, 1:-1, :-2])


def run(k, outs, ins, *scalars, **kw):
    O_ = [ps.zeros(*o) for o in outs]
    k(*O_, *[dev(a) for a in ins], *scalars, **kw)
    return [o.to_host() for o in O_]


# ----------------------------------------------------------------- one array (:443-579)
def test_copy_parallel_optranges():
    # :454-464
    A = cub(N3)
    k = par("@parallel memopt=true loopsize=3 optvars=A optranges=(A=(0:0,0:0,0:0),) function copy_memopt!(A2, A)\n"
            "    @all(A2) = @all(A)\n    return\nend")
    (A2,) = run(k, [N3], [A])
    assert np.array_equal(A2, A)


def test_d2_yi_parallel():
    # :508-520
    A = cub(N3)
    k = par("@parallel memopt=true loopsize=3 function d2_memopt!(A2, A)\n    @inn(A2) = @d2_yi(A)\n    return\nend")
    (A2,) = run(k, [N3], [A])
    ref = np.zeros(N3)
    ref[1:-1, 1:-1, 1:-1] = d2y(A)
    assert np.array_equal(A2, ref)


def test_diffusion_parallel_indices_stencil_m1_1():
    # :521-545, lam = dt = _dx = _dy = _dz = 1
    T, Ci = cub(N3), np.ones(N3)
    k = pi("""
    @parallel_indices (ix,iy,iz) memopt=true loopsize=3 function diffusion3D_step!(T2, T, Ci, lam, dt, _dx, _dy, _dz)
        if (1<ix<size(T2,1) && 1<iy<size(T2,2) && 1<iz<size(T2,3))
            T2[ix,iy,iz] = T[ix,iy,iz] + dt*(Ci[ix,iy,iz]*(
                            - ((-lam*(T[ix+1,iy,iz] - T[ix,iy,iz])*_dx) - (-lam*(T[ix,iy,iz] - T[ix-1,iy,iz])*_dx))*_dx
                            - ((-lam*(T[ix,iy+1,iz] - T[ix,iy,iz])*_dy) - (-lam*(T[ix,iy,iz] - T[ix,iy-1,iz])*_dy))*_dy
                            - ((-lam*(T[ix,iy,iz+1] - T[ix,iy,iz])*_dz) - (-lam*(T[ix,iy,iz] - T[ix,iy,iz-1])*_dz))*_dz)
                            );
        end
        return
    end""")
    (T2,) = run(k, [N3], [T, Ci], 1.0, 1.0, 1.0, 1.0, 1.0)
    c = inn(T)
    q = lambda p, m: ((-1.0 * (p - c)) * 1.0) - ((-1.0 * (c - m)) * 1.0)
    ref = np.zeros(N3)
    ref[1:-1, 1:-1, 1:-1] = c + 1.0 * (inn(Ci) * (-q(T[2:, 1:-1, 1:-1], T[:-2, 1:-1, 1:-1]) * 1.0
                                                  - q(T[1:-1, 2:, 1:-1], T[1:-1, :-2, 1:-1]) * 1.0
                                                  - q(T[1:-1, 1:-1, 2:], T[1:-1, 1:-1, :-2]) * 1.0))
    assert np.array_equal(T2, ref)


DIFF02 = ("@parallel memopt=true loopsize=3 function diffusion3D_step!(T2, T, Ci, lam, dt, _dx, _dy, _dz)\n"
          "    @inn(T2) = @inn(T) + dt*(lam*@inn(Ci)*(@d2_xi(T)*_dx^2 + @d2_yi(T)*_dy^2 + @d2_zi(T)*_dz^2))\n    return\nend")


@pytest.mark.parametrize("n", [N3, (33, 7, 8)])
def test_diffusion_parallel_stencil_0_2(n):
    # :546-564 (integer scalars, as there) and :1018-1036 (33 x 7 x 8)
    T, Ci = cub(n), np.ones(n)
    (T2,) = run(par(DIFF02), [n], [T, Ci], 1, 1, 1, 1, 1)
    ref = np.zeros(n)
    ref[1:-1, 1:-1, 1:-1] = inn(T) + 1.0 * (1.0 * inn(Ci) * (d2x(T) * 1.0 + d2y(T) * 1.0 + d2z(T) * 1.0))
    assert np.array_equal(T2, ref)


# ----------------------------------------------------------------- two and three arrays (:603-761)
def test_two_arrays_copy():
    # :603-615
    A, B = cub(N3), cub(N3, 2.0)
    k = par("@parallel memopt=true loopsize=3 optvars=(A, B) optranges=(A=(0:0,0:0,0:0), B=(0:0,0:0,0:0)) function copy_memopt!(A2, A, B)\n"
            "    @all(A2) = @all(A) + @all(B)\n    return\nend")
    (A2,) = run(k, [N3], [A, B])
    assert np.array_equal(A2, A + B)


def test_two_arrays_z_stencil():
    # :616-632, B an XXYYZ field (nz - 1 planes)
    A, B = cub(N3), cub((32, 8, 7), 2.0)
    k = pi("@parallel_indices (ix,iy,iz) memopt=true loopsize=3 function d2_memopt!(A2, A, B)\n"
           "    if (iz>1 && iz<size(A2,3))\n"
           "        A2[ix,iy,iz] = A[ix,iy,iz+1] - 2*A[ix,iy,iz] + A[ix,iy,iz-1] + B[ix,iy,iz] - B[ix,iy,iz-1]\n"
           "    end\n    return\nend")
    (A2,) = run(k, [N3], [A, B])
    ref = np.zeros(N3)
    ref[:, :, 1:-1] = A[:, :, 2:] - 2 * A[:, :, 1:-1] + A[:, :, :-2] + B[:, :, 1:] - B[:, :, :-1]
    assert np.array_equal(A2, ref)


def test_two_arrays_y_stencil():
    # :633-649
    A, B = cub(N3), cub(N3, 2.0)
    k = pi("@parallel_indices (ix,iy,iz) memopt=true loopsize=3 function d2_memopt!(A2, A, B)\n"
           "    if (iy>1 && iy<size(A2,2))\n"
           "        A2[ix,iy,iz] = A[ix,iy+1,iz] - 2*A[ix,iy,iz] + A[ix,iy-1,iz] + B[ix,iy+1,iz] - 2*B[ix,iy,iz] + B[ix,iy-1,iz]\n"
           "    end\n    return\nend")
    (A2,) = run(k, [N3], [A, B])
    ref = np.zeros(N3)
    ref[:, 1:-1, :] = (((A[:, 2:, :] - 2 * A[:, 1:-1, :]) + A[:, :-2, :] + B[:, 2:, :]) - 2 * B[:, 1:-1, :]) + B[:, :-2, :]
    assert np.array_equal(A2, ref)


def test_two_arrays_x_stencil():
    # :650-666, B an XYYZZ field (nx - 1 columns), default loopsize
    A, B = cub(N3), cub((31, 8, 8), 2.0)
    k = pi("@parallel_indices (ix,iy,iz) memopt=true function d2_memopt!(A2, A, B)\n"
           "    if (ix>1 && ix<size(A2,1))\n"
           "        A2[ix,iy,iz] = A[ix+1,iy,iz] - 2*A[ix,iy,iz] + A[ix-1,iy,iz] + B[ix,iy,iz] - B[ix-1,iy,iz]\n"
           "    end\n    return\nend")
    (A2,) = run(k, [N3], [A, B])
    ref = np.zeros(N3)
    ref[1:-1, :, :] = A[2:, :, :] - 2 * A[1:-1, :, :] + A[:-2, :, :] + B[1:, :, :] - B[:-1, :, :]
    assert np.array_equal(A2, ref)


def _diff_mod(T, dCi):
    return inn(T) + 1.0 * (1.0 * dCi * (d2x(T) * 1.0 + d2y(T) * 1.0 + d2z(T) * 1.0))


def test_two_arrays_xyz_and_z_stencil():
    # :667-686, Ci an XXYYZ field
    T, Ci = cub(N3), cub((32, 8, 7), 2.0)
    k = par("@parallel memopt=true loopsize=3 function diffusion3D_step_modified!(T2, T, Ci, lam, dt, _dx, _dy, _dz)\n"
            "    @inn(T2) = @inn(T) + dt*(lam*@d_zi(Ci)*(@d2_xi(T)*_dx^2 + @d2_yi(T)*_dy^2 + @d2_zi(T)*_dz^2))\n    return\nend")
    (T2,) = run(k, [N3], [T, Ci], 1.0, 1.0, 1.0, 1.0, 1.0)
    ref = np.zeros(N3)
    ref[1:-1, 1:-1, 1:-1] = _diff_mod(T, Ci[1:-1, 1:-1, 1:] - Ci[1:-1, 1:-1, :-1])
    assert np.array_equal(T2, ref)


@pytest.mark.parametrize("n", [N3, (33, 7, 8)])
def test_three_arrays_xyz_y_x(n):
    # :707-728 and :1037-1060, Ci an XYYZZ field
    T, Ci, B = cub(n), cub((n[0] - 1,) + n[1:], 2.0), cub(n, 3.0)
    k = par("@parallel memopt=true loopsize=3 function diffusion3D_step_modified!(T2, T, Ci, B, lam, dt, _dx, _dy, _dz)\n"
            "    @inn(T2) = @inn(T) + dt*(lam*@d_xi(Ci)*(@d2_xi(T)*_dx^2 + @d2_yi(T)*_dy^2 + @d2_zi(T)*_dz^2)) + @d2_yi(B)\n"
            "    return\nend")
    (T2,) = run(k, [n], [T, Ci, B], 1.0, 1.0, 1.0, 1.0, 1.0)
    ref = np.zeros(n)
    ref[1:-1, 1:-1, 1:-1] = _diff_mod(T, Ci[1:, 1:-1, 1:-1] - Ci[:-1, 1:-1, 1:-1]) + d2y(B)
    assert np.array_equal(T2, ref)


def test_two_arrays_xyz_and_x_stencil():
    # :687-706
    T, Ci = cub(N3), cub((31, 8, 8), 2.0)
    k = par("@parallel memopt=true loopsize=3 function diffusion3D_step_modified!(T2, T, Ci, lam, dt, _dx, _dy, _dz)\n"
            "    @inn(T2) = @inn(T) + dt*(lam*@d_xi(Ci)*(@d2_xi(T)*_dx^2 + @d2_yi(T)*_dy^2 + @d2_zi(T)*_dz^2))\n    return\nend")
    (T2,) = run(k, [N3], [T, Ci], 1.0, 1.0, 1.0, 1.0, 1.0)
    ref = np.zeros(N3)
    ref[1:-1, 1:-1, 1:-1] = _diff_mod(T, Ci[1:, 1:-1, 1:-1] - Ci[:-1, 1:-1, 1:-1])
    assert np.array_equal(T2, ref)


# ------------------------------------------------------------ higher order, three arrays (:729-990)
HO_A = ("    if (ix-4>1 && ix-1<size(A2,1) && iy+2>1 && iy+2<=size(A2,2) && iz-2>=1 && iz+3<=size(A2,3))\n"
        "        A2[ix-1,iy+2,iz] = A[ix-1,iy+2,iz+3] - 2*A[ix-3,iy+2,iz] + A[ix-4,iy+2,iz-2]\n    end\n")
HO_B_SAME = HO_A.replace("A2", "B2").replace("A[", "B[")
HO_C_SAME = HO_A.replace("A2", "C2").replace("A[", "C[")
HO_B_SHIFT = ("    if (ix-4>1 && ix-1<size(B2,1) && iy+2>1 && iy+2<=size(B2,2) && iz+1>=1 && iz+2<=size(B2,3))\n"
              "        B2[ix-1,iy+2,iz+1] = B[ix-1,iy+2,iz+2] - 2*B[ix-3,iy+2,iz+1] + B[ix-4,iy+2,iz+1]\n    end\n")
HO_C_SHIFT = ("    if (ix-4>1 && ix-1<size(C2,1) && iy+2>1 && iy+2<=size(C2,2) && iz-1>=1 && iz<=size(C2,3))\n"
              "        C2[ix-1,iy+2,iz-1] = C[ix-1,iy+2,iz] - 2*C[ix-3,iy+2,iz-1] + C[ix-4,iy+2,iz-1]\n    end\n")
HO_B_WIDE = ("    if (ix-4>1 && ix+1<size(B2,1) && iy+2>1 && iy+2<=size(B2,2) && iz+1>=1 && iz+2<=size(B2,3))\n"
             "        B2[ix+1,iy+2,iz+1] = B[ix+1,iy+2,iz+2] - 2*B[ix-3,iy+2,iz+1] + B[ix-4,iy+2,iz+1]\n    end\n")
HO_C_NARROW = ("    if (ix-1>1 && ix-1<size(C2,1) && iy+2>1 && iy+2<=size(C2,2) && iz-1>=1 && iz<=size(C2,3))\n"
               "        C2[ix-1,iy+2,iz-1] = C[ix-1,iy+2,iz] - 2*C[ix-1,iy+2,iz-1] + C[ix-1,iy+2,iz-1]\n    end\n")


def _ho_ref_a(A):
    r = np.zeros(N3)
    r[4:-1, 2:, 2:-3] = A[4:-1, 2:, 5:] - 2 * A[2:-3, 2:, 2:-3] + A[1:-4, 2:, 0:-5]
    return r


def _ho_ref_b_shift(B):
    r = np.zeros(N3)
    r[4:-1, 2:, 1:-1] = B[4:-1, 2:, 2:] - 2 * B[2:-3, 2:, 1:-1] + B[1:-4, 2:, 1:-1]
    return r


def _ho_ref_c_shift(C):
    r = np.zeros(N3)
    r[4:-1, 2:, 0:-1] = C[4:-1, 2:, 1:] - 2 * C[2:-3, 2:, 0:-1] + C[1:-4, 2:, 0:-1]
    return r


def _ho_ref_b_wide(B):
    r = np.zeros(N3)
    r[6:-1, 2:, 1:-1] = B[6:-1, 2:, 2:] - 2 * B[2:-5, 2:, 1:-1] + B[1:-6, 2:, 1:-1]
    return r


def _ho_ref_c_narrow(C):
    r = np.zeros(N3)
    r[1:-1, 2:, 0:-1] = C[1:-1, 2:, 1:] - 2 * C[1:-1, 2:, 0:-1] + C[1:-1, 2:, 0:-1]
    return r


HO_CASES = [
    # :729-761 the same stencil on three arrays
    ("same", "", HO_B_SAME, HO_C_SAME, (_ho_ref_a, _ho_ref_a, _ho_ref_a)),
    # :762-794 z-shifted stencils
    ("shifted", "", HO_B_SHIFT, HO_C_SHIFT, (_ho_ref_a, _ho_ref_b_shift, _ho_ref_c_shift)),
    # :795-827 different x extents
    ("wide_narrow", "", HO_B_WIDE, HO_C_NARROW, (_ho_ref_a, _ho_ref_b_wide, _ho_ref_c_narrow)),
    # :828-941 optvars=(A, C) with optranges
    ("optvars_AC", "optvars=(A, C) loopdim=3 optranges=(A=(-4:-1, 2:2, -2:3), B=(-4:1, 2:2, 1:2), C=(-1:-1, 2:2, -1:0)) ",
     HO_B_WIDE, HO_C_NARROW, (_ho_ref_a, _ho_ref_b_wide, _ho_ref_c_narrow)),
    # :942-990 optvars=(A, B) with narrower optranges than the stencils
    ("optvars_AB", "optvars=(A, B) loopdim=3 optranges=(A=(-1:-1, 2:2, -2:3), B=(-4:-3, 2:2, 1:1)) ",
     HO_B_WIDE, HO_C_NARROW, (_ho_ref_a, _ho_ref_b_wide, _ho_ref_c_narrow)),
]


@pytest.mark.parametrize("case", HO_CASES, ids=[c[0] for c in HO_CASES])
def test_higher_order_three_arrays(case):
    name, opts, b_stmt, c_stmt, refs = case
    print(f"higher order three arrays: {name}", flush=True)
    A, B, C = cub(N3), cub(N3, 2.0), cub(N3, 3.0)
    k = pi(f"@parallel_indices (ix,iy,iz) memopt=true {opts}loopsize=3 function higher_order_memopt!(A2, B2, C2, A, B, C)\n"
           + HO_A + b_stmt + c_stmt + "    return\nend")
    got = run(k, [N3, N3, N3], [A, B, C])
    for g, ref_fn, X in zip(got, refs, (A, B, C)):
        assert np.array_equal(g, ref_fn(X))


# ------------------------------------------------------------ sizes that are not multiples of the block (:991-1036)
def test_ragged_copy_with_larger_ranges():
    # :993-1006: ranges (1:64, 1:64, 1:8) over a 33 x 7 x 8 field, guarded in the kernel
    n = (33, 7, 8)
    A = cub(n)
    k = pi("@parallel_indices (ix,iy,iz) memopt=true loopsize=3 optvars=A optranges=(A=(0:0,0:0,0:0),) function copy_memopt!(A2, A)\n"
           "    if ix>0 && ix<=size(A2,1) && iy>0 && iy<=size(A2,2)\n        A2[ix,iy,iz] = A[ix,iy,iz]\n    end\n    return\nend")
    (A2,) = run(k, [n], [A], ranges=((1, 64), (1, 64), (1, 8)))
    assert np.array_equal(A2, A)


def test_ragged_copy_parallel():
    # :1007-1017
    n = (33, 7, 8)
    A = cub(n)
    k = par("@parallel memopt=true loopsize=3 optvars=A optranges=(A=(0:0,0:0,0:0),) function copy_memopt!(A2, A)\n"
            "    @all(A2) = @all(A)\n    return\nend")
    (A2,) = run(k, [n], [A])
    assert np.array_equal(A2, A)


# ----------------------------------------------------------------- 2-D (:1084-1195)
N2 = (32, 8)


def pi2(text):
    return ps.core._define("parallel_indices", text, {}, MOD2)


def par2(text, **kw):
    return ps.core._define("parallel", text, kw, MOD2)


def _d2y2(A):
    return A[:, 2:] - 2 * A[:, 1:-1] + A[:, :-2]


D2Y_BODY = "    if (1 < iy < size(A2, 2))\n        A2[ix, iy] = A[ix, iy + 1] - 2 * A[ix, iy] + A[ix, iy - 1]\n    end\n    return\nend"


@pytest.mark.parametrize("opts", ["", "optvars=A optranges=(A=(0:0, -1:1, 0:0),) ", "loopsize=4 ",
                                  "optvars=A optranges=(A=(0:0, -1:1, 0:0),) optimize_halo_read=false "],
                         ids=["plain", "optranges", "loopsize4", "no_halo_opt"])
def test_2d_y_stencil(opts):
    # :1090-1104, :1148-1195
    A = cub(N2)
    k = pi2(f"@parallel_indices (ix, iy) ndims=2 memopt=true loopdim=2 {opts}function d2_memopt_indices!(A2, A)\n" + D2Y_BODY)
    (A2,) = run(k, [N2], [A])
    ref = np.zeros(N2)
    ref[:, 1:-1] = _d2y2(A)
    assert np.array_equal(A2, ref)


def test_2d_parallel_d2_yi():
    # :1105-1130
    B = cub(N2, 2.0)
    k = par2("@parallel ndims=2 memopt=true optvars=B loopdim=2 function d2_memopt_kernel!(A2, B)\n    @inn(A2) = @d2_yi(B)\n    return\nend")
    (A2,) = run(k, [N2], [B])
    ref = np.zeros(N2)
    ref[1:-1, 1:-1] = B[1:-1, 2:] - 2 * B[1:-1, 1:-1] + B[1:-1, :-2]
    assert np.array_equal(A2, ref)


def test_2d_mixed_optvars():
    # :1131-1147: A streamed (optvar), B read through the cache
    A, B = cub(N2), cub(N2, 2.0)
    k = pi2("@parallel_indices (ix, iy) ndims=2 memopt=true loopdim=2 optvars=A function mixed_memopt_indices!(A2, A, B)\n"
            "    if (1 < ix < size(A2, 1) && 1 < iy < size(A2, 2))\n"
            "        A2[ix, iy] = (A[ix, iy + 1] - 2 * A[ix, iy] + A[ix, iy - 1]) + (B[ix + 1, iy] - 2 * B[ix, iy] + B[ix - 1, iy])\n"
            "    end\n    return\nend")
    (A2,) = run(k, [N2], [A, B])
    ref = np.zeros(N2)
    ref[1:-1, 1:-1] = (A[1:-1, 2:] - 2 * A[1:-1, 1:-1] + A[1:-1, :-2]) + (B[2:, 1:-1] - 2 * B[1:-1, 1:-1] + B[:-2, 1:-1])
    assert np.array_equal(A2, ref)
