"""Host-side mirror of the ParallelStencil API over libpsamd.so.

Names, argument meaning and error behaviour follow the reference:
  init_parallel_stencil     src/init_parallel_stencil.jl:20-151
  parallel / parallel_indices / parallel_async
                            src/parallel.jl:75-215, src/ParallelKernel/parallel.jl:149-164
  synchronize               src/ParallelKernel/parallel.jl:175-184
  zeros / ones / rand / fill / fill! / falses / trues
                            src/ParallelKernel/allocators.jl:21-172
  hide_communication        src/ParallelKernel/hide_communication.jl:147-258
  get_stream / get_priority_stream
                            src/ParallelKernel/hide_communication.jl:158-174
Kernels are written in the FiniteDifferences macro language of the reference
and passed as text; all parsing, lowering, compilation and launching happens
in the shared library. There is no CPU execution path: without the library
or without a GPU the calls raise.
"""
import atexit
import contextlib
import ctypes
import inspect
import json
import re
import sys
import threading
import warnings
import weakref

import numpy as np

from . import _lib
from . import ad as _ad
from ._lib import (PS_ARG_FIELD, PS_ARG_SCALAR, PS_BF16, PS_BOOL, PS_F32, PS_F64, PS_I32, PS_I64, ArgumentError,
                   IncoherentArgumentError, KeywordArgumentError, NotInitializedError, check, lib, ps_arg, ps_devinfo,
                   ps_field, ps_launch)

AMDGPU = "AMDGPU"
Float64 = "Float64"
Float32 = "Float32"
BFloat16 = "BFloat16"
Int32 = "Int32"
Int64 = "Int64"
Bool = "Bool"

_DTYPES = {
    "Float64": PS_F64, "Float32": PS_F32, "BFloat16": PS_BF16, "Int32": PS_I32, "Int64": PS_I64, "Int": PS_I64,
    "Bool": PS_BOOL,
}
_NP = {PS_F64: np.float64, PS_F32: np.float32, PS_BF16: np.uint16, PS_I32: np.int32, PS_I64: np.int64, PS_BOOL: np.bool_}
_NP_TO_CODE = {np.dtype(np.float64): PS_F64, np.dtype(np.float32): PS_F32, np.dtype(np.int32): PS_I32,
               np.dtype(np.int64): PS_I64, np.dtype(np.bool_): PS_BOOL}
_CODE_NAME = {v: k for k, v in _DTYPES.items() if k != "Int"}


def dtype_code(t):
    if isinstance(t, int) and t in _NP:
        return t
    if isinstance(t, str):
        if t in _DTYPES:
            return _DTYPES[t]
        raise ArgumentError(f"unsupported element type {t}")
    try:
        return _NP_TO_CODE[np.dtype(t)]
    except (TypeError, KeyError):
        raise ArgumentError(f"unsupported element type {t}") from None


# ---------------------------------------------------------------------------
# initialisation (per caller module, like the reference's metadata module)
# ---------------------------------------------------------------------------
class _Config:
    __slots__ = ("package", "numbertype", "ndims", "inbounds", "padding", "memopt", "nonconst_metadata")

    def __init__(self, **kw):
        for k in self.__slots__:
            setattr(self, k, kw[k])

    def astuple(self):
        return tuple(getattr(self, k) for k in self.__slots__)


_configs = {}
_device_ready = False
_lock = threading.Lock()


def _caller(depth):
    f = inspect.currentframe()
    for _ in range(depth + 1):
        f = f.f_back
    return f.f_globals.get("__name__", "__main__")


def init_parallel_stencil(package=AMDGPU, numbertype=None, ndims=None, *, inbounds=False, padding=False, memopt=False,
                          nonconst_metadata=False, module=None):
    """@init_parallel_stencil(package, numbertype, ndims; inbounds, padding, memopt, nonconst_metadata).
    numbertype and ndims may be omitted (src/ParallelKernel/init_parallel_kernel.jl:114): kernels
    then give `ndims` themselves and allocations their `eltype`."""
    mod = module or _caller(1)
    if package not in ("AMDGPU", "HIP"):
        raise ArgumentError(f"unsupported package {package}: this backend runs on AMD MI355X GPUs (package AMDGPU)")
    if numbertype is not None and numbertype not in ("Float64", "Float32"):
        raise ArgumentError(f"unsupported numbertype {numbertype} (Float64 or Float32)")
    if ndims is not None and ndims not in (1, 2, 3):
        raise ArgumentError(f"ndims must be evaluatable at parse time (e.g. literal or constant) and has to be one of the following Integers: 1, 2, 3 (obtained: {ndims}).")
    for name, val in (("memopt", memopt), ("nonconst_metadata", nonconst_metadata), ("inbounds", inbounds), ("padding", padding)):
        if not isinstance(val, bool):
            raise ArgumentError(f"{name} must be evaluatable at parse time (e.g. literal or constant) and has to be of type Bool.")
    cfg = _Config(package="AMDGPU", numbertype=numbertype, ndims=ndims, inbounds=inbounds, padding=padding, memopt=memopt,
                  nonconst_metadata=nonconst_metadata)
    if mod in _configs:
        if _configs[mod].astuple() == cfg.astuple():
            warnings.warn(f"ParallelStencil has already been initialized for the module {mod}, with the same arguments.")
            return Data(mod)
        raise IncoherentCallError(
            f"ParallelStencil has already been initialized for the module {mod}, with different arguments. "
            "Call reset_parallel_stencil() first.")
    _configs[mod] = cfg
    return Data(mod)


class IncoherentCallError(IncoherentArgumentError):
    pass


def is_initialized(module=None):
    return (module or _caller(1)) in _configs


def reset_parallel_stencil(module=None):
    mod = module or _caller(1)
    _configs.pop(mod, None)
    _hardware.pop(mod, None)


# Runtime hardware selection (src/ParallelKernel/select_hardware.jl:28-105): the AMDGPU package
# has one hardware symbol, gpu_amd; the selection exists for code written against backends with
# several (KernelAbstractions).
SUPPORTED_HARDWARE = {"AMDGPU": ("gpu_amd",)}
_hardware = {}


def _hw_name(h):
    return str(h).lstrip(":")


def select_hardware(hardware, module=None):
    """@select_hardware(hardware): only :gpu_amd is valid for this backend."""
    mod = module or _caller(1)
    if mod not in _configs:
        raise ArgumentError(f"ParallelKernel has not been initialized in module {mod} (missing @init_parallel_kernel call).")
    supported = SUPPORTED_HARDWARE[_configs[mod].package]
    if _hw_name(hardware) not in supported:
        raise ArgumentError(f"unsupported hardware selection for package {_configs[mod].package} (obtained: {hardware}). Supported hardware "
                            f"symbols are: {', '.join(supported)}.")
    _hardware[mod] = _hw_name(hardware)


def current_hardware(module=None):
    """@current_hardware(): the selected hardware symbol (gpu_amd)."""
    mod = module or _caller(1)
    _config(mod)
    return _hardware.get(mod, SUPPORTED_HARDWARE[_configs[mod].package][0])


def get_hardware(module=None):
    """@get_hardware(): the hardware of the module (the same as current_hardware here)."""
    return current_hardware(module or _caller(1))


def handle(hardware, package=AMDGPU):
    """Backend handle of a hardware symbol: defined for KernelAbstractions only (an ArgumentError
    for every other package, as in the reference)."""
    raise ArgumentError(f"handle: hardware handles exist for the KernelAbstractions package only (obtained: {hardware}, package {package}).")


def _config(mod):
    if mod not in _configs:
        raise NotInitializedError(f"no ParallelStencil macro or function can be called before @init_parallel_stencil in each module (missing call in {mod}).")
    return _configs[mod]


def ensure_device(device=0):
    """Initialises the GPU once (ps_init); raises if none is present."""
    global _device_ready
    with _lock:
        if not _device_ready:
            check(lib.ps_init(device))
            _device_ready = True


def device_info():
    ensure_device()
    d = ps_devinfo()
    check(lib.ps_device_info(ctypes.byref(d)))
    return {f: (getattr(d, f).decode() if isinstance(getattr(d, f), bytes) else getattr(d, f)) for f, _ in d._fields_}


class Data:
    """The Data module of the reference (src/ParallelKernel/Data.jl): Number is the numbertype."""

    def __init__(self, mod):
        self.module = mod
        self.Number = _configs[mod].numbertype
        self.Array = DeviceArray
        self.Field = DeviceArray


# ---------------------------------------------------------------------------
# device arrays
# ---------------------------------------------------------------------------
# number of explicit DeviceArray.free() calls so far: cached argument blocks built before the
# latest free are checked again before they are launched
_FREED = [0]


def _check_launch_args(name, args):
    """A launch needs device memory behind every array argument: ArgSpecs (shapes only) and freed
    arrays are refused before they reach a kernel."""
    for i, v in enumerate(args):
        if isinstance(v, ArgSpec):
            raise ArgumentError(f"argument {i + 1} of kernel {name} is an ArgSpec (a shape without memory); "
                                "ArgSpecs only serve compilation and geometry queries")
        if isinstance(v, DeviceArray) and not v._alive():
            raise ArgumentError(f"argument {i + 1} of kernel {name}: the array has been freed")


class DeviceArray:
    """A pitch-aligned device field (x fastest, 1-based in kernels)."""

    __slots__ = ("f", "_owner", "_parent", "__weakref__")

    def __init__(self, f, owner=True, parent=None):
        self.f = f
        self._owner = owner
        self._parent = parent

    @property
    def ndim(self):
        return self.f.ndims

    @property
    def shape(self):
        return tuple(int(self.f.dims[i]) for i in range(self.f.ndims))

    @property
    def dtype_code(self):
        return self.f.dtype

    @property
    def eltype(self):
        return _CODE_NAME[self.f.dtype]

    @property
    def pitch(self):
        return int(self.f.pitch)

    @property
    def ptr(self):
        return int(self.f.ptr or 0)

    def __len__(self):
        return self.shape[0]

    def __repr__(self):
        return f"DeviceArray({self.eltype}, {self.shape}, pitch={self.pitch})"

    def to_host(self, stream=None):
        """Array(A): a Fortran-ordered numpy copy (BFloat16 comes back as raw uint16 bits)."""
        out = np.empty(self.shape, dtype=_NP[self.f.dtype], order="F")
        if out.size:
            check(lib.ps_field_d2h(ctypes.byref(self.f), out.ctypes.data_as(ctypes.c_void_p), _stream_handle(stream)))
        return out

    def __array__(self, dtype=None, copy=None):
        a = self.to_host()
        return a if dtype is None else a.astype(dtype)

    def copy_from(self, host, stream=None):
        """copy!(A, host)."""
        h = np.asarray(host)
        if self.f.dtype == PS_BF16 and h.dtype != np.uint16:
            h = f32_to_bf16(h)
        h = np.asfortranarray(h, dtype=_NP[self.f.dtype])
        if h.shape != self.shape:
            if h.size == int(np.prod(self.shape)) and h.ndim != self.ndim:
                h = np.asfortranarray(h.reshape(self.shape, order="F"))
            else:
                raise ArgumentError(f"copy!: shape mismatch (device {self.shape}, host {h.shape})")
        if h.size:
            check(lib.ps_field_h2d(ctypes.byref(self.f), h.ctypes.data_as(ctypes.c_void_p), _stream_handle(stream)))
        return self

    def fill_(self, value, stream=None):
        """fill!(A, x)."""
        check(lib.ps_field_fill(ctypes.byref(self.f), float(value), _stream_handle(stream)))
        if stream is None:
            check(lib.ps_stream_sync(None))
        return self

    def copy_(self, other, stream=None, async_=False):
        """self .= other on the device; blocks unless a stream is given or async_ is set (then it
        is ordered with the kernels of that stream, the library stream by default)."""
        check(lib.ps_field_copy(ctypes.byref(self.f), ctypes.byref(other.f), _stream_handle(stream)))
        if stream is None and not async_:
            check(lib.ps_stream_sync(None))
        return self

    def view(self, lo, hi):
        """A view of the 1-based inclusive index box lo..hi."""
        lo3 = (ctypes.c_int64 * 3)(*(list(lo) + [1] * (3 - len(lo))))
        hi3 = (ctypes.c_int64 * 3)(*(list(hi) + [1] * (3 - len(hi))))
        v = ps_field()
        check(lib.ps_field_view(ctypes.byref(self.f), lo3, hi3, ctypes.byref(v)))
        v.ndims = self.f.ndims
        return DeviceArray(v, owner=False, parent=self)

    @property
    def parent(self):
        """A.parent of a SubArray: the array this one is a view of (itself if it is none)."""
        a = self
        while a._parent is not None:
            a = a._parent
        return a

    @property
    def indices(self):
        """A.indices of a SubArray: the index ranges of this array in its parent (1-based)."""
        off = [int(self.f.voff[d]) if self.f.pdims[0] else 0 for d in range(self.f.ndims)]
        return tuple(UnitRange(off[d] + 1, off[d] + self.shape[d]) for d in range(self.f.ndims))

    def reduce(self, op, stream=None):
        out = ctypes.c_double()
        code = {"sum": 0, "sum_abs": 1, "max_abs": 2, "mean_abs": 3, "norm": 4}[op]
        check(lib.ps_field_reduce(ctypes.byref(self.f), code, ctypes.byref(out), _stream_handle(stream)))
        return out.value

    def free(self):
        if self._owner and self.f.base:
            check(lib.ps_field_free(ctypes.byref(self.f)))
            _FREED[0] += 1

    def _alive(self):
        """False once the memory of this array (or of the array it views) has been freed."""
        a = self
        while a is not None:
            if not a.f.ptr and a.f.dims[0] * a.f.dims[1] * a.f.dims[2] > 0:
                return False
            if a._owner and not a.f.base and a.f.dims[0] * a.f.dims[1] * a.f.dims[2] > 0:
                return False
            a = a._parent
        return True

    def __del__(self):
        # at interpreter teardown the process is about to release its device memory anyway, and
        # the runtime may already be shutting down: no frees from finalisers then
        try:
            if self._owner and self.f.base and lib is not None and not sys.is_finalizing():
                lib.ps_field_free(ctypes.byref(self.f))
        except Exception:
            pass


def Array(A):
    """Array(A): host copy of a device array."""
    return A.to_host()


def f32_to_bf16(x):
    """Round float32 values to BFloat16 bits (nearest even; NaN -> 0x7fc0)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = 0x7FC0
    return np.asarray(r).reshape(np.shape(x))


def bf16_to_f32(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


class ArgSpec:
    """Shape and element type of a field, without memory (for ahead-of-time compilation); with
    `celldims` (or a `celltype`) that of a CellArray."""

    def __init__(self, shape, eltype=Float64, celldims=None, blocklength=0, celltype=None):
        self.shape = tuple(shape)
        self.f = ps_field()
        for i in range(3):
            self.f.dims[i] = self.shape[i] if i < len(self.shape) else 1
        self.f.ndims = len(self.shape)
        self.f.dtype = dtype_code(eltype)
        self.f.pitch = self.shape[0] if self.shape else 1
        self.f.plane = self.f.pitch * (self.shape[1] if len(self.shape) > 1 else 1)
        if celltype is not None:
            celldims = celltype.dims
            self._names = ",".join(celltype.fieldnames).encode()
            self.f.cellnames = self._names
        if celldims is not None:
            cd = tuple(celldims) if isinstance(celldims, (tuple, list)) else (celldims,)
            self.f.ncomp = int(np.prod(cd))
            for i in range(4):
                self.f.celldims[i] = cd[i] if i < len(cd) else 1
            n = int(np.prod(self.shape))
            self.f.blocklen = max(n, 1) if not blocklength else blocklength
        # distinct fake addresses (never dereferenced) so that aliasing analysis sees separate fields
        ArgSpec._next += 1
        self.f.ptr = ArgSpec._next << 40

    _next = 0


def _numbertype_of(mod):
    nt = _configs[mod].numbertype if mod in _configs else Float64
    if nt is None:
        raise ArgumentError("the keyword argument 'eltype' is mandatory in @zeros, @ones, @rand and @fill when no default is set.")
    return nt


# ---------------------------------------------------------------------------
# CellArrays (src/ParallelKernel/allocators.jl:1-17,176-224,389-454; CellArrays.jl)
# ---------------------------------------------------------------------------
class CellType:
    """@CellType(name; fieldnames, dims=length(fieldnames), eltype, parametric): a cell of named
    values for the `celltype` keyword of the allocators. Values can always also be addressed by
    index. A parametric cell type takes its element type as `T[eltype]`."""

    def __init__(self, name, fieldnames=None, dims=None, eltype=None, parametric=False):
        if fieldnames is None:
            raise ArgumentError("@CellType: the keyword argument 'fieldnames' is mandatory.")
        if isinstance(fieldnames, str):
            fieldnames = (fieldnames,)
        self.name = str(name)
        self.fieldnames = tuple(str(f) for f in fieldnames)
        self.dims = (len(self.fieldnames),) if dims is None else tuple(int(d) for d in (dims if isinstance(dims, (tuple, list)) else (dims,)))
        if int(np.prod(self.dims)) != len(self.fieldnames):
            raise ArgumentError(f"@CellType: prod(dims) == length(fieldnames) must hold (dims {self.dims}, {len(self.fieldnames)} fieldnames).")
        if not 1 <= len(self.dims) <= 4:
            raise ArgumentError("@CellType: cells of 1 to 4 dimensions are supported.")
        if parametric and eltype is not None:
            raise ArgumentError("@CellType: the keyword argument 'eltype' is invalid when `parametric=true` is set.")
        self.parametric = bool(parametric)
        self.eltype = eltype

    def __getitem__(self, eltype):
        if not self.parametric:
            raise ArgumentError(f"@CellType {self.name} is not parametric")
        return CellType(self.name, self.fieldnames, self.dims, eltype)

    def __repr__(self):
        return f"CellType({self.name}, fieldnames={self.fieldnames}, dims={self.dims})"


class CellArray(DeviceArray):
    """An array of cells (a CellArray of CellArrays.jl): `shape` is the grid of cells, `celldims`
    the dimensions of each cell. The values of one cell field are stored in blocks of
    `blocklength` cells; the default 0 stores each cell field contiguously over all cells."""

    __slots__ = ("celltype", "_names")

    def __init__(self, f, celltype=None):
        super().__init__(f)
        self.celltype = celltype
        self._names = None
        if celltype is not None:
            self._names = ",".join(celltype.fieldnames).encode()
            self.f.cellnames = self._names

    @property
    def celldims(self):
        nd = 4
        while nd > 1 and self.f.celldims[nd - 1] == 1:
            nd -= 1
        return tuple(int(self.f.celldims[i]) for i in range(nd))

    @property
    def blocklength(self):
        return int(self.f.blocklen)

    @property
    def data(self):
        """The data array, of size (blocklength, prod(celldims), ceil(length / blocklength))."""
        n = int(np.prod(self.shape))
        B, nc = int(self.f.blocklen), int(self.f.ncomp)
        v = ps_field()
        v.ptr, v.dtype, v.ndims = self.f.ptr, self.f.dtype, 3
        for i, d in enumerate((B, nc, -(-n // B) if n else 0)):
            v.dims[i] = d
        v.pitch, v.plane = B, B * nc
        return DeviceArray(v, owner=False, parent=self)

    def _raw_shape(self):
        n = int(np.prod(self.shape))
        B = int(self.f.blocklen)
        return (B, int(self.f.ncomp), -(-n // B) if n else 0)

    def to_host(self, stream=None):
        """The cells as a Fortran-ordered numpy array of shape shape + celldims."""
        raw = self.data.to_host(stream)
        B, nc, nb = self._raw_shape()
        n = int(np.prod(self.shape))
        cells = raw.transpose(0, 2, 1).reshape((B * nb, nc), order="F")[:n]
        return np.asfortranarray(cells.reshape(self.shape + self.celldims, order="F"))

    def copy_from(self, host, stream=None):
        """copy!(A, host) with host of shape shape + celldims."""
        h = np.asarray(host)
        if h.shape != self.shape + self.celldims:
            raise ArgumentError(f"copy!: shape mismatch (CellArray {self.shape} with cells {self.celldims}, host {h.shape})")
        B, nc, nb = self._raw_shape()
        n = int(np.prod(self.shape))
        cells = np.zeros((B * nb, nc), dtype=h.dtype)
        cells[:n] = h.reshape((n, nc), order="F")
        raw = np.asfortranarray(cells.reshape((B, nb, nc), order="F").transpose(0, 2, 1))
        self.data.copy_from(raw, stream)
        return self

    def fill_(self, value, stream=None):
        self.data.fill_(value, stream)
        return self

    def copy_(self, other, stream=None, async_=False):
        check(lib.ps_field_copy(ctypes.byref(self.f), ctypes.byref(other.f), _stream_handle(stream)))
        if not async_ and stream is None:
            check(lib.ps_stream_sync(None))
        return self

    def view(self, lo, hi):
        raise ArgumentError("views of CellArrays are not supported")

    def reduce(self, op, stream=None):
        return self.data.reduce(op, stream)


def _cell_spec(mod, eltype, celldims, celltype):
    """(eltype, celldims, celltype) after the checks of determine_celltype (allocators.jl:415-430)."""
    if celltype is not None:
        if celldims is not None or eltype is not None:
            raise ArgumentError("the keyword argument 'celltype' is incompatible with the other keyword arguments.")
        if not isinstance(celltype, CellType):
            raise ArgumentError("celltype must be created with CellType (@CellType)")
        if celltype.parametric:
            raise ArgumentError(f"the parametric cell type {celltype.name} needs an element type: {celltype.name}[eltype]")
        return celltype.eltype or _numbertype_of(mod), celltype.dims, celltype
    if celldims is None:
        return eltype if eltype is not None else _numbertype_of(mod), None, None
    cd = tuple(int(d) for d in (celldims if isinstance(celldims, (tuple, list)) else (celldims,)))
    return eltype if eltype is not None else _numbertype_of(mod), cd, None


def _alloc(dims, eltype, celldims=None, celltype=None, blocklength=None):
    ensure_device()
    if len(dims) == 1 and isinstance(dims[0], (tuple, list)):
        dims = tuple(dims[0])
    dims = tuple(int(d) for d in dims)
    if not 1 <= len(dims) <= 3:
        raise ArgumentError("fields have 1, 2 or 3 dimensions")
    f = ps_field()
    d3 = (ctypes.c_int64 * 3)(*(list(dims) + [1] * (3 - len(dims))))
    if celldims is None:  # blocklength only matters for cells (allocators.jl:15)
        check(lib.ps_field_alloc(len(dims), d3, dtype_code(eltype), ctypes.byref(f)))
        return DeviceArray(f)
    if not 1 <= len(celldims) <= 4:
        raise ArgumentError("celldims must have 1 to 4 dimensions")
    c4 = (ctypes.c_int32 * 4)(*(list(celldims) + [1] * (4 - len(celldims))))
    check(lib.ps_cellarray_alloc(len(dims), d3, len(celldims), c4, int(blocklength or 0), dtype_code(eltype), ctypes.byref(f)))
    return CellArray(f, celltype)


def _allocate(dims, kw, depth=2):
    eltype, celldims, celltype = _cell_spec(_caller(depth), kw.pop("eltype", None), kw.pop("celldims", None), kw.pop("celltype", None))
    blocklength = kw.pop("blocklength", None)
    if kw:
        raise KeywordArgumentError(f"invalid keyword argument(s): {', '.join(sorted(kw))}")
    return _alloc(dims, eltype, celldims, celltype, blocklength)


def zeros(*dims, **kw):
    """@zeros(dims...; eltype=numbertype, celldims, celltype, blocklength)."""
    return _allocate(dims, kw, 2)


def ones(*dims, **kw):
    return _allocate(dims, kw, 2).fill_(1.0)


def fill(x, *dims, **kw):
    return _allocate(dims, kw, 2).fill_(x)


def fill_(A, x):
    return A.fill_(x)


def falses(*dims, **kw):
    kw.setdefault("eltype", Bool)
    return _allocate(dims, kw, 2)


def trues(*dims, **kw):
    kw.setdefault("eltype", Bool)
    return _allocate(dims, kw, 2).fill_(1)


_rng = np.random.default_rng(0)


def seed(s):
    """Seeds the host generator behind rand()."""
    global _rng
    _rng = np.random.default_rng(s)


def rand(*dims, **kw):
    """@rand: uniform [0,1) values drawn on the host (as the reference's AMDGPU allocator does) and copied."""
    A = _allocate(dims, kw, 2)
    code = A.dtype_code
    shape = A.shape + (A.celldims if isinstance(A, CellArray) else ())
    if code in (PS_F64, PS_F32, PS_BF16):
        h = _rng.random(shape)
        h = f32_to_bf16(h.astype(np.float32)) if code == PS_BF16 else h.astype(_NP[code])
    elif code == PS_BOOL:
        h = _rng.random(shape) < 0.5
    else:
        h = _rng.integers(0, 2**31 - 1, size=shape).astype(_NP[code])
    return A.copy_from(np.asfortranarray(h))


# ---------------------------------------------------------------------------
# streams and events
# ---------------------------------------------------------------------------
class Stream:
    __slots__ = ("handle", "id", "priority")

    def __init__(self, handle, id_, priority):
        self.handle = handle
        self.id = id_
        self.priority = priority

    def synchronize(self):
        check(lib.ps_stream_sync(self.handle))

    def __repr__(self):
        return f"Stream(id={self.id}, {'priority' if self.priority else 'normal'} pool)"


_streams = {}


def get_stream(id_):
    """get_rocstream(id) equivalent: stream `id` of the normal pool."""
    return _get_stream(id_, 0)


def get_priority_stream(id_):
    """get_priority_rocstream(id) equivalent. Streams of both pools run at the default
    hardware priority; the pools only separate boundary and inner work."""
    return _get_stream(id_, 1)


def _get_stream(id_, prio):
    ensure_device()
    key = (int(id_), prio)
    if key not in _streams:
        h = ctypes.c_void_p()
        check(lib.ps_stream_get(int(id_), prio, ctypes.byref(h)))
        _streams[key] = Stream(h.value, int(id_), prio)
    return _streams[key]


_tls = threading.local()


def _default_stream():
    return getattr(_tls, "stream", None)


def _stream_handle(s):
    if s is None:
        s = _default_stream()
    if s is None:
        return None
    return s.handle if isinstance(s, Stream) else s


@contextlib.contextmanager
def stream_context(s):
    """Makes `s` the default stream of @parallel calls in this thread."""
    prev = getattr(_tls, "stream", None)
    _tls.stream = s
    try:
        yield s
    finally:
        _tls.stream = prev


def synchronize(stream=None):
    """@synchronize() / @synchronize(stream)."""
    if getattr(_tls, "overlap", None) is not None:
        raise ArgumentError(ERRMSG_OVERLAP)
    ensure_device()
    if stream is None:
        check(lib.ps_device_sync())
    else:
        check(lib.ps_stream_sync(_stream_handle(stream)))


class Event:
    def __init__(self):
        ensure_device()
        h = ctypes.c_void_p()
        check(lib.ps_event_create(ctypes.byref(h)))
        self.h = h

    def record(self, stream=None):
        check(lib.ps_event_record(self.h, _stream_handle(stream)))
        return self

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        check(lib.ps_event_elapsed_ms(self.h, later.h, ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if not sys.is_finalizing():
                lib.ps_event_destroy(self.h)
        except Exception:
            pass


_LIVE_GRAPHS = weakref.WeakSet()
_GRAPH_ATEXIT = [False]


def _destroy_graphs():
    """Destroys every graph still alive. Registered with atexit once the device is initialised,
    so it runs while the process is whole: before interpreter teardown (where objects are freed in
    no fixed order) and before the C-level exit handlers of the HIP runtime and of tools loaded
    into the process."""
    for g in list(_LIVE_GRAPHS):
        g.destroy()


class Graph:
    """HIP graph of the asynchronous launches issued on a stream inside `with graph:`; `launch()`
    replays them with one launch. Kernels must have run once (compiled) before the capture.
    `destroy()` releases the executable graph; graphs still alive at exit are destroyed by an
    atexit hook, never from a finaliser during interpreter teardown."""

    def __init__(self, stream=None):
        ensure_device()
        self.stream = stream
        self.h = None
        if not _GRAPH_ATEXIT[0]:
            atexit.register(_destroy_graphs)
            _GRAPH_ATEXIT[0] = True
        _LIVE_GRAPHS.add(self)

    def __enter__(self):
        check(lib.ps_graph_begin(_stream_handle(self.stream)))
        return self

    def __exit__(self, exc_type, exc, tb):
        h = ctypes.c_void_p()
        rc = lib.ps_graph_end(_stream_handle(self.stream), ctypes.byref(h))
        if exc_type is None:
            check(rc)
            self.h = h
        elif rc == 0:
            lib.ps_graph_destroy(h)
        return False

    def launch(self, stream=None):
        if self.h is None:
            raise ArgumentError("the graph has not been captured (or has been destroyed)")
        check(lib.ps_graph_launch(self.h, _stream_handle(self.stream if stream is None else stream)))

    def destroy(self):
        h, self.h = self.h, None
        if h is not None:
            check(lib.ps_graph_destroy(h))

    def __del__(self):
        try:
            if self.h is not None and not sys.is_finalizing():
                lib.ps_graph_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------
class BoundLaunch:
    """A kernel launch with a fixed argument block (see Kernel.launcher). With `weak`, the
    arrays are held through weak references only (`alive()` tells whether they all still exist)."""

    def __init__(self, kernel, args, launch, weak=False):
        self.kernel = kernel
        self.args = list(args)
        self._launch = launch
        self._build()
        self._refs = None
        if weak:
            self._refs = [(i, weakref.ref(a)) for i, a in enumerate(self.args) if isinstance(a, DeviceArray)]
            self.args = [None if isinstance(a, DeviceArray) else a for a in self.args]

    def _current_args(self):
        if self._refs is None:
            return self.args
        a = list(self.args)
        for i, r in self._refs:
            a[i] = r()
            if a[i] is None:
                raise ArgumentError(f"kernel {self.kernel.name}: an array of this prepared launch no longer exists")
        return a

    def alive(self):
        return self._refs is None or all(r() is not None for _, r in self._refs)

    def _build(self):
        k = self.kernel
        self._arr = k._args(self.args, launch=True)
        self._gen = _FREED[0]
        self._L = k._launch_struct(**self._launch)
        self._Lp = ctypes.byref(self._L)
        self._n = len(self.args)

    def set(self, i, value):
        """Replaces argument i (a field or a scalar)."""
        if self._refs is not None:
            raise ArgumentError("arguments of a prepared launch with weak references cannot be replaced")
        self.args[i] = value
        self._build()

    def __call__(self):
        if self._gen != _FREED[0]:
            _check_launch_args(self.kernel.name, self._current_args())
            self._gen = _FREED[0]
        rc = lib.ps_kernel_launch(self.kernel._h, self._arr, self._n, self._Lp)
        if rc < 0:
            check(rc)
def _range3(r):
    """Normalises ranges: (start, stop) pairs (Julia start:stop) per dimension, 1-D/2-D promoted."""
    if isinstance(r, (UnitRange, range)):
        r = (r,)
    if len(r) == 2 and all(isinstance(x, (int, np.integer)) for x in r):
        r = ((r[0], r[1]),)
    out = []
    for x in r:
        if isinstance(x, UnitRange):
            out.append((x.start, x.stop))
        elif isinstance(x, range):
            if x.step != 1:
                raise ArgumentError("ranges must have unit step")
            out.append((x.start, x.stop - 1))
        else:
            a, b = x
            out.append((int(a), int(b)))
    if not 1 <= len(out) <= 3:
        raise ArgumentError(f"ranges must be a Tuple of UnitRange of size 1, 2 or 3 (obtained: {r}).")
    while len(out) < 3:
        out.append((1, 1))
    return out


class UnitRange:
    """Julia's start:stop (inclusive, 1-based)."""
    __slots__ = ("start", "stop")

    def __init__(self, start, stop):
        self.start, self.stop = int(start), int(stop)

    def __len__(self):
        return max(0, self.stop - self.start + 1)

    def __eq__(self, o):
        return isinstance(o, UnitRange) and (self.start, self.stop) == (o.start, o.stop)

    def __repr__(self):
        return f"{self.start}:{self.stop}"


def _triple(x, name):
    if x is None:
        return None
    x = tuple(int(v) for v in (x if isinstance(x, (tuple, list)) else (x,)))
    if not 1 <= len(x) <= 3:
        raise ArgumentError(f"{name} must be an Array or Tuple of size 3 (obtained: {x}).")
    return x + (1,) * (3 - len(x))


class Geometry(tuple):
    """(memopt, checked, nthreads, nblocks, ranges) of a launch; `checked` is True when the host could
    not prove every access in bounds and the bounds-checked code runs."""
    __slots__ = ()

    def __new__(cls, memopt, checked, nthreads, nblocks, ranges):
        return tuple.__new__(cls, (memopt, checked, nthreads, nblocks, ranges))

    memopt = property(lambda s: s[0])
    checked = property(lambda s: s[1])
    nthreads = property(lambda s: s[2])
    nblocks = property(lambda s: s[3])
    ranges = property(lambda s: s[4])


def _tuple_names(v):
    """Component names of a tuple-like kernel argument (None for arrays and scalars)."""
    if isinstance(v, dict):
        return [str(n) for n in v]
    if isinstance(v, tuple) and hasattr(v, "_fields"):
        return list(v._fields)
    if isinstance(v, (tuple, list)):
        return [str(i + 1) for i in range(len(v))]
    return None


def _tuple_values(v):
    return list(v.values()) if isinstance(v, dict) else list(v)


class Kernel:
    """A kernel defined with @parallel / @parallel_indices; calling it is `@parallel kernel(args...)`."""

    def __init__(self, text, opts):
        self._h = ctypes.c_void_p()
        self.opts = opts
        self.text = text
        check(lib.ps_kernel_define(text.encode(), json.dumps(opts).encode(), ctypes.byref(self._h)))
        buf = ctypes.create_string_buffer(1 << 20)
        check(lib.ps_kernel_info(self._h, buf, len(buf)))
        self.info = json.loads(buf.value.decode())
        self.name = self.info["name"]
        self.params = [p["name"] for p in self.info["params"]]
        self._annots = [p["annot"] for p in self.info["params"]]
        self.numbertype = self.info["numbertype"]
        self._expanded = {}
        self._derivs = {}

    def _resolve(self, args):
        """(kernel, arguments) with tuple and NamedTuple arguments (Data.ArrayTuple,
        Data.NamedArrayTuple) split into their components; each distinct structure of such
        arguments gets a kernel of its own, whose parameters are the components."""
        dp = self.info.get("def_params", self.params)
        if len(dp) == len(self.params) and not any(isinstance(v, (tuple, list, dict)) for v in args):
            return self, args  # no tuple arguments
        spec, flat = {}, []
        for i, v in enumerate(args):
            names = _tuple_names(v)
            if names is None or i >= len(dp):
                flat.append(v)
                continue
            spec[dp[i]] = names
            flat.extend(_tuple_values(v))
        if not spec:
            if len(dp) != len(self.params):  # the body uses tuple arguments the call does not pass
                raise ArgumentError(f"kernel {self.name} takes tuple arguments ({', '.join(p for p in dp if p + '.' in ' '.join(self.params))})")
            return self, args
        key = tuple((p, tuple(n)) for p, n in spec.items())
        k = self._expanded.get(key)
        if k is None:
            k = self._expanded[key] = Kernel(self.text, dict(self.opts, tuple_params=spec))
        return k, tuple(flat)

    def __del__(self):
        try:
            lib.ps_kernel_free(self._h)
        except Exception:
            pass

    def __repr__(self):
        return f"Kernel({self.name}, {self.info['kind']}, params={self.params})"

    @property
    def metadata(self):
        return self.info

    def _scalar_code(self, i, v):
        ann = self._annots[i]
        if isinstance(v, (bool, np.bool_)):
            return PS_BOOL
        if ann.endswith("Number") or ann.endswith("Data.Number"):
            return dtype_code(self.numbertype)
        if ann in ("Float64", "Float32", "Int32", "Int64", "Int", "Bool"):
            return dtype_code(ann)
        if ann in ("Integer",):
            return PS_I64
        if isinstance(v, np.generic):
            return dtype_code(v.dtype)
        if isinstance(v, (int, np.integer)):
            return PS_I64
        if isinstance(v, (float, np.floating)):
            return PS_F64
        raise ArgumentError(f"argument {self.params[i]}: unsupported type {type(v).__name__}")

    def _args(self, args, launch=False):
        """The argument block of a launch. Blocks are kept for the last few argument lists
        (arrays by identity, through weak references, and scalars by type and value): a loop
        that launches with the same arrays skips the rebuild. For a launch (`launch`), the arrays
        are checked to have memory when the block is built and again after any free()."""
        key = tuple(id(v) if isinstance(v, (DeviceArray, ArgSpec)) else (type(v), v) for v in args)
        cache = self.__dict__.setdefault("_arg_blocks", {})
        hit = cache.get(key)
        if hit is not None:
            arr, refs, gen = hit
            if all(r() is v for r, v in zip(refs, (v for v in args if isinstance(v, (DeviceArray, ArgSpec))))):
                if launch and gen != _FREED[0]:
                    _check_launch_args(self.name, args)
                    cache[key] = (arr, refs, _FREED[0])
                return arr
        if launch:
            _check_launch_args(self.name, args)
        arr = self._build_args(args)
        if len(cache) >= 16:
            cache.pop(next(iter(cache)))
        cache[key] = (arr, [weakref.ref(v) for v in args if isinstance(v, (DeviceArray, ArgSpec))], _FREED[0] if launch else -1)
        return arr

    def _build_args(self, args):
        if len(args) != len(self.params):
            raise ArgumentError(f"MethodError: kernel {self.name} takes {len(self.params)} arguments ({len(args)} given)")
        arr = (ps_arg * max(1, len(args)))()
        for i, v in enumerate(args):
            a = arr[i]
            if isinstance(v, (DeviceArray, ArgSpec)):
                a.kind = PS_ARG_FIELD
                a.field = v.f
                a.dtype = v.f.dtype
            else:
                a.kind = PS_ARG_SCALAR
                code = self._scalar_code(i, v)
                a.dtype = code
                if code in (PS_F64, PS_F32):
                    a.f = float(v)
                else:
                    a.i = int(v)
        return arr

    def _launch_struct(self, ranges, nblocks, nthreads, stream, async_, memopt, shmem=None):
        L = ps_launch()
        if ranges is not None:
            for d, (a, b) in enumerate(_range3(ranges)):
                L.ranges[d][0] = a
                L.ranges[d][1] = b
            L.has_ranges = 1
        nb, nt = _triple(nblocks, "nblocks"), _triple(nthreads, "nthreads")
        if (nb is None) != (nt is None):
            raise ArgumentError("nblocks and nthreads must be given together")
        if nb:
            for d in range(3):
                L.nblocks[d], L.nthreads[d] = nb[d], nt[d]
        L.stream = _stream_handle(stream)
        L.async_ = 1 if async_ else 0
        L.memopt = -1 if memopt is None else int(bool(memopt))
        if shmem is not None:
            if int(shmem) < 0:
                raise ArgumentError("shmem must be a non-negative number of bytes")
            L.shmem = int(shmem)
        return L

    def _derivative(self, mode, annots):
        """The derivative kernel (reverse or forward mode) for these annotations {param: kind}."""
        key = (mode, tuple(sorted(annots.items())))
        d = self._derivs.get(key)
        if d is None:
            d = self._derivs[key] = _DerivativeKernel(self, mode, annots)
        return d

    def __call__(self, *args, ranges=None, nblocks=None, nthreads=None, stream=None, async_=False, memopt=None, shmem=None, **adkw):
        """Launches the kernel (@parallel); `shmem` is the number of bytes of dynamic LDS that
        @sharedMem arrays of the kernel use. With `∇` (or `grad`), `ad_mode` and
        `ad_annotations` the derivative of the kernel is launched instead (`parallelstencil.ad`)."""
        if adkw:
            bad = set(adkw) - {"∇", "grad", "ad_mode", "ad_annotations"}
            if bad:
                raise KeywordArgumentError(f"invalid keyword argument(s) in @parallel <kernelcall>: {', '.join(sorted(bad))}")
            grad = adkw.get("∇", adkw.get("grad"))
            if "∇" in adkw and "grad" in adkw:
                raise KeywordArgumentError("∇ and grad are the same keyword argument")
            if grad is None:
                _ad.incoherent()
            if memopt is not None:
                raise KeywordArgumentError("memopt cannot be chosen for a derivative launch (derivative kernels are plain kernels)")
            return _ad.call_with_gradient(self, args, grad, adkw.get("ad_mode"), adkw.get("ad_annotations"), ranges=ranges,
                                          nblocks=nblocks, nthreads=nthreads, stream=stream, async_=async_, shmem=shmem)
        ov = getattr(_tls, "overlap", None)
        if ov is not None:  # inside an @overlap block: stream i for the i-th call
            if stream is not None:
                raise ArgumentError(ERRMSG_INVALID_STREAM_OVERLAP)
            if async_:
                raise ArgumentError(ERRMSG_OVERLAP)
            stream, async_ = get_stream(len(ov) + 1), True
            ov.append(stream)
        k, args = self._resolve(args)
        arr = k._args(args, launch=True)
        ensure_device()
        L = k._launch_struct(ranges, nblocks, nthreads, stream, async_, memopt, shmem)
        check(lib.ps_kernel_launch(k._h, arr, len(args), ctypes.byref(L)))

    def _launcher_weak(self, args, kw):
        k, args = self._resolve(tuple(args))
        _check_launch_args(k.name, args)
        ensure_device()
        full = dict(ranges=None, nblocks=None, nthreads=None, stream=None, async_=True, memopt=None, shmem=None)
        full.update(kw)
        return BoundLaunch(k, args, full, weak=True)

    def launcher(self, *args, ranges=None, nblocks=None, nthreads=None, stream=None, async_=True, memopt=None, shmem=None):
        """A prepared launch of this kernel with these arguments (asynchronous by default):
        calling it launches again without rebuilding the argument block, for launch-bound loops
        and graph capture. It keeps its arguments alive; `set(i, value)` replaces one."""
        k, args = self._resolve(args)
        _check_launch_args(k.name, args)
        ensure_device()
        return BoundLaunch(k, args, dict(ranges=ranges, nblocks=nblocks, nthreads=nthreads, stream=stream,
                                            async_=async_, memopt=memopt, shmem=shmem))

    def geometry(self, *args, ranges=None, nblocks=None, nthreads=None, memopt=None, shmem=None):
        """(memopt, nthreads, nblocks, ranges) the launch would use."""
        k, args = self._resolve(args)
        arr = k._args(args)
        L = k._launch_struct(ranges, nblocks, nthreads, None, True, memopt, shmem)
        nt = (ctypes.c_int32 * 3)()
        nb = (ctypes.c_int32 * 3)()
        rg = ((ctypes.c_int64 * 2) * 3)()
        lds = ctypes.c_int64()
        m = check(lib.ps_kernel_geometry(k._h, arr, len(args), ctypes.byref(L), nt, nb,
                                         ctypes.cast(rg, ctypes.POINTER(ctypes.c_int64)), ctypes.byref(lds)))
        return Geometry(bool(m & 1), bool(m & 2), tuple(nt), tuple(nb), tuple((rg[d][0], rg[d][1]) for d in range(3)))

    def source(self, *args, memopt=None, checked=False, idx64=False, off32=False, cells_x=1, cells_z=1):
        """The generated HIP source for these argument kinds (plain kernels: with `off32`, 32-bit
        per-thread byte offsets and 24-bit offset multiplies, and groups of `cells_x` x `cells_z`
        cells per thread)."""
        k, args = self._resolve(args)
        arr = k._args(args)
        mo = k.info["memopt"] if memopt is None else memopt
        flags = (1 if mo else 0) | (2 if checked else 0) | (4 if idx64 else 0)
        if off32 and not mo:
            flags |= 64 | 128 | ((cells_x & 0xF) << 8 if cells_x > 1 else 0) | ((cells_z & 0xF) << 12 if cells_z > 1 else 0)
        buf = ctypes.create_string_buffer(1 << 22)
        check(lib.ps_kernel_source(k._h, arr, len(args), flags, buf, len(buf)))
        return buf.value.decode()

    def precompile(self, *args, memopt=None, checked=False, idx64=False, maxthreads=256):
        """Compiles the code object for these argument kinds into the disk cache (no GPU needed)."""
        k, args = self._resolve(args)
        arr = k._args(args)
        mo = k.info["memopt"] if memopt is None else memopt
        flags = (1 if mo else 0) | (2 if checked else 0) | (4 if idx64 else 0) | (16 if maxthreads > 256 else 0)
        check(lib.ps_kernel_precompile(k._h, arr, len(args), flags))

    def precompile_launch(self, *args, ranges=None, nblocks=None, nthreads=None, memopt=None, shmem=None):
        """Compiles exactly the variant a launch with these arguments would use (no GPU needed);
        fields may be ArgSpecs."""
        k, args = self._resolve(args)
        arr = k._args(args)
        L = k._launch_struct(ranges, nblocks, nthreads, None, True, memopt, shmem)
        check(lib.ps_kernel_precompile_launch(k._h, arr, len(args), ctypes.byref(L)))


class _DerivativeKernel(Kernel):
    """A derivative kernel made by ps_kernel_derivative from a defined kernel. Reverse mode of a
    loop-free kernel also gets the two-kernel gather form (each point sums the adjoints that
    reach it, then the seeds are cleared): it is used for launches with the default geometry,
    the atomic (scatter) form otherwise or where the gather form does not apply."""

    def __init__(self, primal, mode, annots, form="scatter"):
        self._h = ctypes.c_void_p()
        self.opts = dict(primal.opts, memopt=False)
        self.text = primal.text
        self.primal = primal
        self.mode = mode
        self.form = form
        self._gather = None
        self.use_gather = True
        check(lib.ps_kernel_derivative(primal._h, _ad.spec_json(mode, annots, form).encode(), ctypes.byref(self._h)))
        buf = ctypes.create_string_buffer(1 << 20)
        check(lib.ps_kernel_info(self._h, buf, len(buf)))
        self.info = json.loads(buf.value.decode())
        self.info["def_params"] = [p["name"] for p in self.info["params"]]
        self.name = self.info["name"]
        self.params = [p["name"] for p in self.info["params"]]
        self._annots = [p["annot"] for p in self.info["params"]]
        self.numbertype = self.info["numbertype"]
        self._expanded = {}
        self._derivs = {}
        if form == "scatter" and mode == "Reverse":
            try:
                self._gather = (_DerivativeKernel(primal, mode, annots, "gather"), _DerivativeKernel(primal, mode, annots, "clear"))
            except ArgumentError:
                self._gather = None

    def _resolve(self, args):
        return self, args

    def __call__(self, *args, ranges=None, nblocks=None, nthreads=None, stream=None, async_=False, memopt=None, shmem=None):
        if self._gather is not None and self.use_gather and ranges is None and nblocks is None and nthreads is None:
            g, c = self._gather
            Kernel.__call__(g, *args, stream=stream, async_=True, memopt=memopt, shmem=shmem)
            return Kernel.__call__(c, *args, stream=stream, async_=async_, memopt=memopt, shmem=shmem)
        return Kernel.__call__(self, *args, ranges=ranges, nblocks=nblocks, nthreads=nthreads, stream=stream, async_=async_,
                               memopt=memopt, shmem=shmem)


_KW_DEF = ("ndims", "N", "inbounds", "padding", "memopt", "optvars", "loopdim", "loopsize", "optranges", "useshmemhalos",
           "optimize_halo_read", "fp_contract", "fastmath", "prefetch", "lds", "xcd_remap", "nt_store", "nthreads_max_memopt",
           "nthreads_x_max_memopt", "cells_x", "cells_z", "nthreads_x_groups", "xcd_spread", "addr32")


_PREFIX_RE = re.compile(r"^(\s*@parallel(?:_indices)?\b\s*(?:\([^()]*\)|\w+\.\.\.)?)((?:\s*\w+\s*=\s*(?:\([^()]*\)|[^\s(]+))*)", re.S)
_KW_PAIR_RE = re.compile(r"(\w+)\s*=\s*(\([^()]*\)|[^\s(]+)")


def _prefix_ndims(text, kw):
    """Moves `ndims=...` and `N=...` of the text prefix (`@parallel ndims=(1,3) N=ndims.+1 f(...)`)
    into the keyword arguments: both select kernel methods, which the definition expands here."""
    m = _PREFIX_RE.match(text)
    if not m or not m.group(2).strip():
        return text, kw
    kept, kw = [], dict(kw)
    for k, v in _KW_PAIR_RE.findall(m.group(2)):
        if k in ("ndims", "N"):
            if k in kw:
                raise KeywordArgumentError(f"keyword argument '{k}' is given twice")
            kw[k] = _literal(v) if k == "ndims" else v
        else:
            kept.append(f"{k}={v}")
    return m.group(1) + (" " + " ".join(kept) if kept else "") + text[m.end():], kw


def _literal(v):
    v = v.strip()
    if re.fullmatch(r"\d+", v):
        return int(v)
    if re.fullmatch(r"\(\s*\d+\s*(,\s*\d+\s*)*,?\s*\)", v):
        return tuple(int(x) for x in re.findall(r"\d+", v))
    raise KeywordArgumentError(f"keyword argument 'ndims' must be an integer or a tuple of integers (obtained: {v}).")


def _eval_N(N, ndims, macro):
    """The value(s) of N (an integer, a tuple, or an expression of ndims such as ndims+2 or
    ndims.+1, broadcast over a tuple of ndims), src/ParallelKernel/parallel.jl:334-359."""
    if isinstance(N, (int, np.integer)) and not isinstance(N, bool):
        vals = int(N)
    elif isinstance(N, (tuple, list)):
        vals = tuple(int(x) for x in N)
    elif isinstance(N, str):
        expr = N.strip()
        if "ndims" not in expr:
            ndims = 0 if not isinstance(ndims, tuple) else ndims
        if not re.fullmatch(r"[0-9ndims+\-*/(). ,]*", expr):
            raise KeywordArgumentError(f"{macro}: keyword argument 'N' must be an integer expression of ndims (obtained: {N}).")
        broadcast = isinstance(ndims, tuple)
        if broadcast and expr != "ndims" and re.search(r"(?<!\.)[+\-*/]", expr.replace("ndims", "")):
            raise KeywordArgumentError(f"{macro}: keyword argument 'N' must be a tuple of integers of the same length as 'ndims' when "
                                       f"'ndims' is a tuple (use broadcasting operators, e.g. ndims.+1; obtained: N={N}, ndims={ndims}).")
        plain = expr.replace(".+", "+").replace(".-", "-").replace(".*", "*").replace("./", "//")
        ev = lambda d: eval(plain, {"__builtins__": {}}, {"ndims": d})  # noqa: S307 (digits, ndims and operators only)
        vals = tuple(ev(d) for d in ndims) if broadcast and "ndims" in expr else ev(ndims)
    else:
        raise KeywordArgumentError(f"{macro}: keyword argument 'N' must be an integer (or a tuple if 'ndims' is a tuple; obtained: {N}).")
    return vals


def _subst_N(text, n, macro):
    """N substituted in the type parameters of the kernel signature (`Data.Array{T,N}`)."""
    if re.search(r"\bwhere\b[^=\n]*\bN\b", text):
        raise IncoherentArgumentError(f"{macro}: 'N' must not appear in the where clause of the kernel signature when the keyword argument "
                                      "'N' is used.")
    return re.sub(r"\{[^{}]*\}", lambda m: re.sub(r"\bN\b", str(n), m.group(0)), text)


class KernelMethods:
    """A kernel defined for several ndims (`@parallel ndims=(1,3) N=ndims.+1 f(...)`): one
    method per ndims, chosen at the call by the number of dimensions of the array arguments
    (the value of N of each method)."""

    def __init__(self, methods):
        self.methods = dict(methods)  # ndims -> (N, Kernel)
        self.name = next(iter(self.methods.values()))[1].name

    def __repr__(self):
        return f"KernelMethods({self.name}, ndims={tuple(self.methods)})"

    def method(self, *args):
        nds = {len(a.shape) for a in args if isinstance(a, (DeviceArray, ArgSpec))}
        for nd, (n, k) in self.methods.items():
            if nds and max(nds) == n:
                return k
        raise ArgumentError(f"MethodError: no method of kernel {self.name} for arrays of {max(nds) if nds else 0} dimensions "
                            f"(methods for N = {', '.join(str(n) for n, _ in self.methods.values())})")

    def __call__(self, *args, **kw):
        return self.method(*args)(*args, **kw)


def _define(kind, text, kw, mod, indices=None):
    bad = set(kw) - set(_KW_DEF)
    if bad:
        raise KeywordArgumentError(f"invalid keyword argument(s) in @{kind} <kernel>: {', '.join(sorted(bad))}")
    cfg = _config(mod)
    macro = f"@{kind}"
    text, kw = _prefix_ndims(text, kw)
    nd = kw.get("ndims", cfg.ndims)
    if isinstance(nd, (tuple, list)):
        nd = tuple(nd)
        if not all(isinstance(d, (int, np.integer)) for d in nd):
            raise KeywordArgumentError(f"{macro}: keyword argument 'ndims' must be an integer or a tuple of integers (obtained: {nd}).")
        if "N" not in kw:
            raise KeywordArgumentError(f"{macro}: keyword argument 'N' is mandatory when 'ndims' is a tuple ('N' must also be present as type "
                                       "parameter in the function signature enabling to dispatch on). ")
        Ns = _eval_N(kw["N"], nd, macro)
        if not isinstance(Ns, tuple) or len(Ns) != len(nd):
            raise KeywordArgumentError(f"{macro}: keyword argument 'N' must be a tuple of integers of the same length as 'ndims' when 'ndims' "
                                       f"is a tuple (obtained: N={Ns}, ndims={nd}).")
        return KernelMethods({d: (n, _define(kind, text, dict(kw, ndims=d, N=n), mod, indices)) for d, n in zip(nd, Ns)})
    if "N" in kw:
        if nd is None or not 1 <= nd <= 3:
            raise KeywordArgumentError(f"{macro}: keyword argument 'ndims' is invalid or missing (valid values are 1, 2 or 3; 'ndims' an be set "
                                       "globally in @init_parallel_stencil and overwritten per kernel if needed).")
        n = _eval_N(kw["N"], nd, macro)
        if not isinstance(n, int):
            raise KeywordArgumentError(f"{macro}: keyword argument 'N' must be an integer (or a tuple if 'ndims' is a tuple; obtained: {n}).")
        text = _subst_N(text, n, macro)
        kw = {k: v for k, v in kw.items() if k != "N"}
    opts = {"kind": kind, "ndims": nd if nd is not None else 0, "numbertype": cfg.numbertype or "Float64",
            "inbounds": kw.get("inbounds", cfg.inbounds), "padding": kw.get("padding", cfg.padding),
            "memopt": kw.get("memopt", cfg.memopt)}
    if indices is not None:
        opts["indices"] = list(indices)
    for k in ("optvars", "loopdim", "loopsize", "fp_contract", "fastmath", "prefetch", "lds", "xcd_remap", "nt_store", "nthreads_max_memopt",
              "nthreads_x_max_memopt", "cells_x", "cells_z", "nthreads_x_groups", "xcd_spread", "addr32"):
        if k in kw:
            v = kw[k]
            opts[k] = list(v) if k == "optvars" and isinstance(v, (tuple, list)) else v
    if "optranges" in kw and kw["optranges"] is not None:
        opts["optranges"] = {k: [list(_range3((r,))[0]) if not isinstance(r, (tuple, list)) else list(r) for r in v]
                             for k, v in kw["optranges"].items()}
    return Kernel(text, opts)


def _split_call(args):
    """(ranges, nblocks, nthreads, kernel, kernel_args) from the positional forms of @parallel."""
    for i, a in enumerate(args):
        if isinstance(a, (Kernel, KernelMethods)):
            pre, k, rest = args[:i], a, args[i + 1:]
            break
    else:
        raise ArgumentError("@parallel: the last argument must be a kernel call")
    ranges = nblocks = nthreads = None
    if len(pre) == 1:
        (ranges,) = pre
    elif len(pre) == 2:
        nblocks, nthreads = pre
    elif len(pre) == 3:
        ranges, nblocks, nthreads = pre
    elif len(pre) > 3:
        raise ArgumentError("too many positional arguments before the kernel call")
    return ranges, nblocks, nthreads, k, rest


def parallel(*args, **kw):
    """@parallel: `parallel(text, **kw)` defines a kernel; `parallel([ranges,] [nblocks, nthreads,] kernel, args...)` runs it."""
    if args and isinstance(args[0], str):
        if len(args) != 1:
            raise ArgumentError("@parallel <kernel>: the kernel text must be the only positional argument")
        return _define("parallel", args[0], kw, _caller(1))
    ranges, nblocks, nthreads, k, rest = _split_call(args)
    ranges = kw.pop("ranges", ranges)
    return k(*rest, ranges=ranges, nblocks=nblocks, nthreads=nthreads, **kw)


def parallel_async(*args, **kw):
    """@parallel_async: launch without synchronising the stream."""
    ranges, nblocks, nthreads, k, rest = _split_call(args)
    ranges = kw.pop("ranges", ranges)
    return k(*rest, ranges=ranges, nblocks=nblocks, nthreads=nthreads, async_=True, **kw)


def parallel_indices(indices_or_text, text=None, **kw):
    """@parallel_indices (ix,iy,...) kernel."""
    mod = _caller(1)
    if text is None:
        return _define("parallel_indices", indices_or_text, kw, mod)
    idx = indices_or_text
    if isinstance(idx, str):
        idx = [s.strip() for s in idx.strip("() ").split(",") if s.strip()]
    return _define("parallel_indices", text, kw, mod, indices=idx)


# ---------------------------------------------------------------------------
# @overlap (src/ParallelKernel/overlap.jl)
# ---------------------------------------------------------------------------
ERRMSG_OVERLAP = "@overlap: the code block must only contain @parallel calls. Type `?@overlap` for more information."
ERRMSG_INVALID_STREAM_OVERLAP = ("@overlap: invalid keyword argument in @parallel <kernelcall>: stream. Specifying the stream is not "
                                 "allowed in a @parallel call inside the code block argument of @overlap.")


class _OverlapBlock:
    def __enter__(self):
        if getattr(_tls, "overlap", None) is not None:
            raise ArgumentError(ERRMSG_OVERLAP)
        self.streams = _tls.overlap = []
        return self

    def __exit__(self, exc_type, exc, tb):
        _tls.overlap = None
        if exc_type is not None:
            for st in self.streams:  # leave no launch of the block running behind the error
                st.synchronize()
            return False
        if not self.streams:
            raise ArgumentError(ERRMSG_OVERLAP)
        for st in self.streams:
            st.synchronize()
        return False


def overlap(*calls):
    """@overlap (src/ParallelKernel/overlap.jl:47-58): runs the i-th @parallel call of the block
    asynchronously on stream i of the normal pool (get_stream(i)) and synchronises those streams
    at the end of the block. Two forms:

        with overlap():
            kernel1(A)
            kernel2(B)

        overlap((kernel1, (A,)), (kernel2, (B,), {"nblocks": ..., "nthreads": ...}))

    Inside the block, a call may not name a stream, be asynchronous or synchronise."""
    if not calls:
        return _OverlapBlock()
    norm = []
    for c in calls:
        if not (isinstance(c, tuple) and 2 <= len(c) <= 3 and isinstance(c[0], Kernel)):
            raise ArgumentError(ERRMSG_OVERLAP)
        kw = dict(c[2]) if len(c) == 3 else {}
        if "stream" in kw:
            raise ArgumentError(ERRMSG_INVALID_STREAM_OVERLAP)
        if kw.get("async_"):
            raise ArgumentError(ERRMSG_OVERLAP)
        norm.append((c[0], tuple(c[1]), kw))
    with _OverlapBlock():
        for k, args, kw in norm:
            k(*args, **kw)


# ---------------------------------------------------------------------------
# @hide_communication
# ---------------------------------------------------------------------------
def _promote_bw(boundary_width):
    """promote_boundary_width: an Int or 1-3-tuple -> (bx, by, bz) padded with 0."""
    bwt = tuple(boundary_width) if isinstance(boundary_width, (tuple, list)) else (boundary_width,)
    if not 1 <= len(bwt) <= 3 or not all(isinstance(b, (int, np.integer)) for b in bwt):
        raise ArgumentError(f"@hide_communication: boundary_width must be a Tuple of Integer of size 1, 2 or 3 (obtained: {boundary_width}).")
    return tuple(int(bwt[i]) if i < len(bwt) else 0 for i in range(3))


def _c_ranges(r):
    return ((ctypes.c_int64 * 2) * 3)(*[(ctypes.c_int64 * 2)(a, b) for a, b in r])


def get_ranges_outer(boundary_width, ranges):
    """get_ranges_outer: the boundary slabs computed first (src/ParallelKernel/hide_communication.jl:350-381)."""
    rr = _c_ranges(_range3(ranges))
    out = (((ctypes.c_int64 * 2) * 3) * 6)()
    n = ctypes.c_int()
    check(lib.ps_ranges_outer(ctypes.cast(rr, ctypes.POINTER(ctypes.c_int64)), (ctypes.c_int64 * 3)(*_promote_bw(boundary_width)),
                              ctypes.cast(out, ctypes.POINTER(ctypes.c_int64)), ctypes.byref(n)))
    return tuple(tuple((out[i][d][0], out[i][d][1]) for d in range(3)) for i in range(n.value))


def get_ranges_inner(boundary_width, ranges):
    """get_ranges_inner (src/ParallelKernel/hide_communication.jl:383-394)."""
    rr = _c_ranges(_range3(ranges))
    out = ((ctypes.c_int64 * 2) * 3)()
    check(lib.ps_ranges_inner(ctypes.cast(rr, ctypes.POINTER(ctypes.c_int64)), (ctypes.c_int64 * 3)(*_promote_bw(boundary_width)),
                              ctypes.cast(out, ctypes.POINTER(ctypes.c_int64))))
    return (tuple((out[d][0], out[d][1]) for d in range(3)),)


def _maxsize_ranges(args):
    """1:maxsize per dimension over the array arguments, tuples and NamedTuples of arrays included
    (maxsize, src/ParallelKernel/parallel.jl:773-803)."""
    ms = [1, 1, 1]
    for a in args:
        parts = _tuple_values(a) if _tuple_names(a) is not None else [a]
        for x in parts:
            if isinstance(x, DeviceArray):
                for d in range(3):
                    ms[d] = max(ms[d], int(x.f.dims[d]))
    return tuple((1, m) for m in ms)


_HC_PLANS = {}


def _hc_key(calls, boundary_width, ranges_inner):
    def arg_key(x):
        return ("v", x) if isinstance(x, (bool, int, float, np.number)) else ("o", id(x))
    return (tuple((id(k), tuple(arg_key(x) for x in a), tuple(sorted(kw.items()))) for k, a, kw in calls), repr(boundary_width), repr(ranges_inner))


def hide_communication(boundary_width, computations, communication, ranges_inner=None, computation_calls=None):
    """@hide_communication boundary_width begin <computation calls>; <boundary/communication code> end.

    computations: a list of (kernel, args) or (kernel, args, kwargs) tuples (the first
    `computation_calls` statements of the reference's block). communication: a callable
    running the boundary-condition / communication code; its @parallel calls go to
    priority stream 1. With `ranges_inner` given, `boundary_width` is the tuple of outer ranges
    (the explicit form `@hide_communication ranges_outer ranges_inner block`).

    The boundary slabs all run on priority stream 1 and the inner region on stream 1 (the
    reference gives each slab a stream of its own): a process has few hardware queues (4 by
    default), and streams beyond them share queues, which serialises the boundary work behind
    the inner region. The launches of a call are prepared once and replayed when the same
    kernels are called with the same arguments.
    """
    ensure_device()
    if callable(computations) or isinstance(computations, tuple) and computations and isinstance(computations[0], Kernel):
        computations = [computations]
    calls = []
    for c in computations:
        k, a = c[0], c[1]
        kw = dict(c[2]) if len(c) > 2 else {}
        if "stream" in kw:
            raise ArgumentError("@hide_communication: the keyword argument 'stream' is not allowed in computation calls")
        if "ranges" in kw:
            raise ArgumentError("@hide_communication: computation calls must not be given explicit ranges")
        calls.append((k, tuple(a), kw))
    if computation_calls is not None and computation_calls < 1:
        raise KeywordArgumentError("Invalid keyword argument in @hide_communication: computation_calls must be >= 1.")
    key = _hc_key(calls, boundary_width, ranges_inner)
    plan = _HC_PLANS.get(key)
    if plan is not None and not all(L.alive() for L in plan[0] + plan[1]):
        plan = None  # an array of the cached plan was released (its id may now name another array)
    if plan is None:
        if ranges_inner is None:
            rngs = [_maxsize_ranges(a) for _, a, _ in calls]
            if any(r != rngs[0] for r in rngs):
                raise ArgumentError(f"@hide_communication: the ranges of the computation calls are not all equal (obtained: {rngs}).")
            outer = get_ranges_outer(boundary_width, rngs[0])
            inner = get_ranges_inner(boundary_width, rngs[0])
        else:
            outer, inner = boundary_width, ranges_inner
        bs, ins = get_priority_stream(1), get_stream(1)
        # prepared launches hold their arrays through weak references, so a cached plan never keeps
        # fields of an earlier call signature alive in HBM
        plan = ([k._launcher_weak(a, dict(ranges=r, stream=bs, async_=True, **kw)) for r in outer for k, a, kw in calls],
                [k._launcher_weak(a, dict(ranges=r, stream=ins, async_=True, **kw)) for r in inner for k, a, kw in calls], bs, ins)
        if len(_HC_PLANS) >= 64:
            _HC_PLANS.pop(next(iter(_HC_PLANS)))
        _HC_PLANS[key] = plan
    outer_l, inner_l, bs, ins = plan
    for L in outer_l:
        L()
    for L in inner_l:
        L()
    bs.synchronize()
    with stream_context(bs):
        communication()
    bs.synchronize()
    ins.synchronize()


# ---------------------------------------------------------------------------
# launch-geometry helpers (host, mirror src/ParallelKernel/parallel.jl:773-819)
# ---------------------------------------------------------------------------
def compute_nthreads(maxsize, nthreads_x_max=64, nthreads_max=256, flatdim=0):
    ms = _triple(maxsize, "maxsize")
    out = (ctypes.c_int32 * 3)()
    check(lib.ps_compute_nthreads((ctypes.c_int64 * 3)(*ms), nthreads_x_max, nthreads_max, flatdim, out))
    return tuple(out)


def compute_nblocks(maxsize, nthreads):
    ms = _triple(maxsize, "maxsize")
    nt = _triple(nthreads, "nthreads")
    out = (ctypes.c_int32 * 3)()
    check(lib.ps_compute_nblocks((ctypes.c_int64 * 3)(*ms), (ctypes.c_int32 * 3)(*nt), out))
    return tuple(out)


def get_ranges(*args):
    """get_ranges: 1:maxsize over the array arguments."""
    return tuple(UnitRange(a, b) for a, b in _maxsize_ranges(args))
