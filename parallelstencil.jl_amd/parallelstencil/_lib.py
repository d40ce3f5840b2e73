"""ctypes binding of libpsamd.so (the C-ABI in include/ps_amd.h).

The shared library is the product: it lowers kernels to HIP for gfx950,
compiles them with hiprtc and launches them. This module only declares the
C signatures and turns negative return codes into Python exceptions that
mirror the reference's exception types (src/ParallelKernel/Exceptions.jl:2-67).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PS_AMD_LIB", os.path.join(os.path.dirname(_HERE), "libpsamd.so"))

PS_F64, PS_F32, PS_BF16, PS_I32, PS_I64, PS_BOOL = 1, 2, 3, 4, 5, 6
PS_ARG_FIELD, PS_ARG_SCALAR = 1, 2


class PSException(Exception):
    """Base class of the errors raised by the backend."""


class ArgumentError(PSException, ValueError):
    pass


class KeywordArgumentError(PSException, TypeError):
    pass


class IncoherentArgumentError(PSException, ValueError):
    pass


class NotInitializedError(PSException, RuntimeError):
    pass


class ModuleInternalError(PSException, RuntimeError):
    pass


class ParseError(ArgumentError):
    pass


class HIPError(PSException, RuntimeError):
    pass


class CompileError(PSException, RuntimeError):
    pass


class OutOfMemoryError(PSException, MemoryError):
    pass


class BoundsError(PSException, IndexError):
    pass


class NoDeviceError(PSException, RuntimeError):
    pass


class MethodPluginError(PSException, RuntimeError):
    pass


_CODE_TO_EXC = {
    -1: ArgumentError,
    -2: KeywordArgumentError,
    -3: IncoherentArgumentError,
    -4: NotInitializedError,
    -5: ModuleInternalError,
    -6: ParseError,
    -7: HIPError,
    -8: CompileError,
    -9: OutOfMemoryError,
    -10: BoundsError,
    -11: NoDeviceError,
    -12: MethodPluginError,
}


class ps_field(ctypes.Structure):
    _fields_ = [
        ("ptr", ctypes.c_void_p),
        ("dims", ctypes.c_int64 * 3),
        ("pitch", ctypes.c_int64),
        ("plane", ctypes.c_int64),
        ("dtype", ctypes.c_int32),
        ("ndims", ctypes.c_int32),
        ("base", ctypes.c_void_p),
        ("alloc_bytes", ctypes.c_int64),
        ("ncomp", ctypes.c_int32),
        ("celldims", ctypes.c_int32 * 4),
        ("blocklen", ctypes.c_int64),
        ("cellnames", ctypes.c_char_p),
        ("voff", ctypes.c_int64 * 3),
        ("pdims", ctypes.c_int64 * 3),
    ]


class ps_arg(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("dtype", ctypes.c_int32),
        ("field", ps_field),
        ("f", ctypes.c_double),
        ("i", ctypes.c_int64),
    ]


class ps_launch(ctypes.Structure):
    _fields_ = [
        ("ranges", (ctypes.c_int64 * 2) * 3),
        ("has_ranges", ctypes.c_int32),
        ("nthreads", ctypes.c_int32 * 3),
        ("nblocks", ctypes.c_int32 * 3),
        ("stream", ctypes.c_void_p),
        ("async_", ctypes.c_int32),
        ("memopt", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("shmem", ctypes.c_int64),
    ]


class ps_devinfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
        ("cu_count", ctypes.c_int32),
        ("wave_size", ctypes.c_int32),
        ("lds_bytes_per_block", ctypes.c_int64),
        ("total_mem_bytes", ctypes.c_int64),
        ("l2_bytes", ctypes.c_int64),
        ("hbm_peak_gbs", ctypes.c_double),
        ("clock_khz", ctypes.c_int32),
        ("mem_clock_khz", ctypes.c_int32),
        ("mem_bus_width", ctypes.c_int32),
    ]


# name -> (restype, argtypes); every function returns int status unless noted
_P = ctypes.POINTER
_SIGS = {
    "ps_last_error": (ctypes.c_char_p, []),
    "ps_version": (ctypes.c_int, []),
    "ps_init": (ctypes.c_int, [ctypes.c_int]),
    "ps_device_info": (ctypes.c_int, [_P(ps_devinfo)]),
    "ps_device_sync": (ctypes.c_int, []),
    "ps_field_alloc": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_int64), ctypes.c_int, _P(ps_field)]),
    "ps_cellarray_alloc": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_int64), ctypes.c_int, _P(ctypes.c_int32), ctypes.c_int64, ctypes.c_int,
                                          _P(ps_field)]),
    "ps_field_free": (ctypes.c_int, [_P(ps_field)]),
    "ps_field_view": (ctypes.c_int, [_P(ps_field), _P(ctypes.c_int64), _P(ctypes.c_int64), _P(ps_field)]),
    "ps_field_h2d": (ctypes.c_int, [_P(ps_field), ctypes.c_void_p, ctypes.c_void_p]),
    "ps_field_d2h": (ctypes.c_int, [_P(ps_field), ctypes.c_void_p, ctypes.c_void_p]),
    "ps_field_fill": (ctypes.c_int, [_P(ps_field), ctypes.c_double, ctypes.c_void_p]),
    "ps_field_copy": (ctypes.c_int, [_P(ps_field), _P(ps_field), ctypes.c_void_p]),
    "ps_field_reduce": (ctypes.c_int, [_P(ps_field), ctypes.c_int, _P(ctypes.c_double), ctypes.c_void_p]),
    "ps_field_reduce_noalloc": (ctypes.c_int, [_P(ps_field), ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "ps_stream_get": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _P(ctypes.c_void_p)]),
    "ps_stream_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_kernel_define": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _P(ctypes.c_void_p)]),
    "ps_kernel_derivative": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _P(ctypes.c_void_p)]),
    "ps_kernel_free": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_kernel_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    "ps_kernel_source": (ctypes.c_int, [ctypes.c_void_p, _P(ps_arg), ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]),
    "ps_kernel_precompile": (ctypes.c_int, [ctypes.c_void_p, _P(ps_arg), ctypes.c_int, ctypes.c_int]),
    "ps_kernel_precompile_launch": (ctypes.c_int, [ctypes.c_void_p, _P(ps_arg), ctypes.c_int, _P(ps_launch)]),
    "ps_kernel_launch": (ctypes.c_int, [ctypes.c_void_p, _P(ps_arg), ctypes.c_int, _P(ps_launch)]),
    "ps_kernel_geometry": (ctypes.c_int, [ctypes.c_void_p, _P(ps_arg), ctypes.c_int, _P(ps_launch), _P(ctypes.c_int32),
                                          _P(ctypes.c_int32), _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    "ps_compute_nthreads": (ctypes.c_int, [_P(ctypes.c_int64), ctypes.c_int, ctypes.c_int, ctypes.c_int, _P(ctypes.c_int32)]),
    "ps_compute_nblocks": (ctypes.c_int, [_P(ctypes.c_int64), _P(ctypes.c_int32), _P(ctypes.c_int32)]),
    "ps_ranges_outer": (ctypes.c_int, [_P(ctypes.c_int64), _P(ctypes.c_int64), _P(ctypes.c_int64), _P(ctypes.c_int)]),
    "ps_ranges_inner": (ctypes.c_int, [_P(ctypes.c_int64), _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    "ps_event_create": (ctypes.c_int, [_P(ctypes.c_void_p)]),
    "ps_event_record": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ps_event_elapsed_ms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _P(ctypes.c_float)]),
    "ps_event_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_set_cache_dir": (ctypes.c_int, [ctypes.c_char_p]),
    "ps_graph_begin": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_graph_end": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_void_p)]),
    "ps_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ps_graph_destroy": (ctypes.c_int, [ctypes.c_void_p]),
}

# include/ps_nn.h (kernels of the Stokes warm-start network)
_V, _I64, _F = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float
_I = ctypes.c_int
_SIGS.update({
    "ps_nn_gemm": (_I, [_I, _I, _I, _V, _I64, _I64, _V, _I64, _I64, _V, _I64, _V, _I, _I, _V]),
    "ps_nn_conv_flip": (_I, [_V, _I, _I, _V, _V]),
    "ps_nn_gemm_im2col": (_I, [_I, _I, _I, _I, _V, _I, _I, _I, _V, _I64, _I64, _V, _I64, _V, _I, _I, _V]),
    "ps_nn_im2col": (_I, [_V, _I, _I, _I, _I, _V, _V]),
    "ps_nn_col2im": (_I, [_V, _I, _I, _I, _I, _V, _V]),
    "ps_nn_pool2": (_I, [_V, _I, _I, _I, _I, _V, _V]),
    "ps_nn_pool2_bwd": (_I, [_V, _I, _I, _I, _I, _V, _V]),
    "ps_nn_up2": (_I, [_V, _I, _I, _I, _I, _V, _V]),
    "ps_nn_up2_bwd": (_I, [_V, _I, _I, _I, _I, _V, _V]),
    "ps_nn_concat": (_I, [_V, _I, _V, _I, _I64, _V, _V]),
    "ps_nn_split": (_I, [_V, _I, _I, _I64, _V, _V, _V]),
    "ps_nn_relu_bwd": (_I, [_V, _V, _I64, _V]),
    "ps_nn_bias_grad": (_I, [_V, _I64, _I, _V, _V]),
    "ps_nn_mse": (_I, [_V, _V, _V, _I64, _F, _V, _V, _V]),
    "ps_nn_adam": (_I, [_V, _V, _V, _V, _I64, _F, _F, _F, _F, _F, _F, _V, _V]),
    "ps_nn_f32_to_bf16": (_I, [_V, _I64, _V, _V]),
    "ps_nn_bf16_to_f32": (_I, [_V, _I64, _V, _V]),
    "ps_nn_fill_f32": (_I, [_V, _I64, _F, _V]),
    "ps_nn_add_bf16": (_I, [_V, _V, _I64, _V]),
    "ps_nn_alloc": (_I, [_I64, _P(_V)]),
    "ps_nn_free": (_I, [_V]),
    "ps_nn_h2d": (_I, [_V, _V, _I64, _V]),
    "ps_nn_d2h": (_I, [_V, _V, _I64, _V]),
    "ps_nn_d2d": (_I, [_V, _V, _I64, _V]),
})


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libpsamd.so not found at {LIB_PATH}: build it with `make -C parallelstencil.jl_amd` "
            "(or __graft_entry__.build()); the backend has no fallback path")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    """Raise the exception matching a negative status code of the C-ABI."""
    if rc < 0:
        msg = lib.ps_last_error().decode("utf-8", "replace")
        raise _CODE_TO_EXC.get(rc, PSException)(msg)
    return rc
