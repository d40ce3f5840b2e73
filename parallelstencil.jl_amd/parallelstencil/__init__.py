"""ParallelStencil programming model on AMD MI355X (gfx950).

Kernels are written in the FiniteDifferences macro language of ParallelStencil.jl
and passed as text; libpsamd.so lowers them to HIP, compiles them with hiprtc and
launches them. See `core` for the API mirror.
"""
from ._lib import (ArgumentError, BoundsError, CompileError, HIPError, IncoherentArgumentError, KeywordArgumentError,
                   MethodPluginError, ModuleInternalError, NoDeviceError, NotInitializedError, OutOfMemoryError,
                   ParseError, PSException, LIB_PATH)
from .core import (AMDGPU, Float64, Float32, BFloat16, Int32, Int64, Bool, Array, ArgSpec, BoundLaunch, CellArray, CellType, Data,
                   DeviceArray, Event, Graph,
                   IncoherentCallError, Kernel, Stream, UnitRange, bf16_to_f32, compute_nblocks, compute_nthreads,
                   device_info, ensure_device, f32_to_bf16, falses, fill, fill_, get_priority_stream, get_ranges,
                   get_ranges_inner, get_ranges_outer, get_stream, hide_communication, init_parallel_stencil,
                   is_initialized, ones, overlap, parallel, parallel_async, parallel_indices, rand, reset_parallel_stencil, seed,
                   stream_context, synchronize, trues, zeros, select_hardware, current_hardware, get_hardware, handle,
                   KernelMethods)
from . import ad as AD  # noqa: E402
from . import igg  # noqa: E402
from .ad import Forward, Reverse  # noqa: E402
from . import FieldAllocators, FiniteDifferences1D, FiniteDifferences2D, FiniteDifferences3D  # noqa: E402
