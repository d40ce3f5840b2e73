/*
 * ps_amd.h -- C-ABI of the MI355X (gfx950) stencil backend for the
 * ParallelStencil programming model.
 *
 * One shared library, libpsamd.so, replaces the AMDGPU extension of the
 * reference (ext/ParallelStencil_AMDGPUExt.jl, src/ParallelKernel/AMDGPUExt/)
 * together with the GPU code generation path of @parallel / @parallel_indices.
 * A kernel is handed over as its source text in the FiniteDifferences macro
 * language; the library parses it, lowers it to HIP C++ for gfx950, compiles
 * it with hiprtc (content-addressed disk cache) and launches it.
 *
 * Conventions (kept from the reference, src/ParallelKernel/parallel.jl):
 *   - indices are 1-based and column-major (x fastest);
 *   - ranges are inclusive [start, stop] pairs, 1-D/2-D promoted to 3-D
 *     with 1:1 in the missing dimensions (parallel.jl:773-776);
 *   - a synchronous launch blocks on its stream, an async one does not
 *     (parallel.jl:64-86, 376, 884).
 * Every function returns 0 on success or a negative PS_E* code; the message
 * of the last failure on the calling thread is returned by ps_last_error().
 * No function takes or returns PyTorch or Julia types.
 */
#ifndef PS_AMD_H
#define PS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes: one per exception type of src/ParallelKernel/Exceptions.jl:2-67 ---- */
#define PS_OK            0
#define PS_EARGUMENT    -1  /* ArgumentError                                  */
#define PS_EKEYWORD     -2  /* KeywordArgumentError                           */
#define PS_EINCOHERENT  -3  /* IncoherentArgumentError / IncoherentCallError  */
#define PS_ENOTINIT     -4  /* NotInitializedError                            */
#define PS_EINTERNAL    -5  /* ModuleInternalError                            */
#define PS_ESYNTAX      -6  /* kernel text does not parse (Julia ParseError)  */
#define PS_EHIP         -7  /* HIP runtime failure                            */
#define PS_ECOMPILE     -8  /* hiprtc compilation failure                     */
#define PS_EOOM         -9  /* device allocation failed (OutOfMemoryError)    */
#define PS_EBOUNDS     -10  /* out-of-bounds array access (BoundsError)       */
#define PS_ENODEVICE   -11  /* no usable GPU                                  */
#define PS_EPLUGIN     -12  /* MethodPluginError (e.g. unknown macro)         */

/* ---- element / scalar types ---- */
#define PS_F64   1
#define PS_F32   2
#define PS_BF16  3  /* storage only: loads widen to FP32, stores round to nearest even */
#define PS_I32   4
#define PS_I64   5
#define PS_BOOL  6  /* stored as one byte */

#define PS_ARG_FIELD  1
#define PS_ARG_SCALAR 2

/* A device field: a pitch-aligned 1-D/2-D/3-D array in HBM.
 * ptr addresses element (1,1,1); element (i,j,k) lives at
 * ptr + (i-1) + (j-1)*pitch + (k-1)*plane   (in elements). */
typedef struct ps_field {
  void*   ptr;
  int64_t dims[3];     /* logical sizes, 1 for absent dimensions            */
  int64_t pitch;       /* elements between consecutive x-rows               */
  int64_t plane;       /* elements between consecutive z-planes             */
  int32_t dtype;       /* PS_F64, PS_F32, PS_BF16, PS_I32, PS_I64, PS_BOOL  */
  int32_t ndims;
  void*   base;        /* allocation base (NULL for views)                  */
  int64_t alloc_bytes;
  /* CellArray (ncomp > 0): dims, pitch and plane describe the dense grid of cells;
   * value j (0-based, column-major over celldims) of the cell with linear index I
   * (0-based) lives at ptr + (I / blocklen) * blocklen * ncomp + j * blocklen + I % blocklen,
   * the layout of CellArrays.jl (data of size (blocklen, ncomp, ceil(N / blocklen))). */
  int32_t ncomp;        /* values per cell; 0 for a plain field               */
  int32_t celldims[4];  /* cell dimensions, 1 for absent ones                  */
  int64_t blocklen;     /* values of one cell field stored contiguously        */
  const char* cellnames;/* comma-separated field names of a @CellType, or NULL */
  /* Views (ps_field_view): the view's first element sits at index voff+1 of its parent
   * (per dimension, 0-based offset) and the parent holds pdims elements. pdims all 0: the field
   * is not a view (its own parent). Kernels defined with padding=true address the parents of
   * their field arguments, as the padded Fields of src/ParallelKernel/FieldAllocators.jl:622-634
   * (SubArray views of a larger parent) are addressed through A.parent
   * (src/ParallelKernel/parallel.jl:572-581). */
  int64_t voff[3];
  int64_t pdims[3];
} ps_field;

typedef struct ps_arg {
  int32_t  kind;       /* PS_ARG_FIELD or PS_ARG_SCALAR */
  int32_t  dtype;      /* element type of the field / type of the scalar */
  ps_field field;
  double   f;          /* value of a float scalar */
  int64_t  i;          /* value of an integer / bool scalar */
} ps_arg;

typedef struct ps_launch {
  int64_t ranges[3][2];  /* inclusive 1-based [start, stop] per dimension  */
  int32_t has_ranges;    /* 0: ranges = 1:maxsize over all field arguments */
  int32_t nthreads[3];   /* all 0: automatic (64,4,1)-style heuristic      */
  int32_t nblocks[3];    /* all 0: ceil(range length / nthreads)           */
  void*   stream;        /* hipStream_t, NULL = the library's stream 0      */
  int32_t async;         /* 0: block on the stream after the launch         */
  int32_t memopt;        /* -1: kernel default, 0 / 1: override              */
  int32_t reserved;
  int64_t shmem;         /* bytes of dynamic LDS for @sharedMem arrays: the
                            shmem keyword of @parallel
                            (src/ParallelKernel/parallel.jl:151,883-895)    */
} ps_launch;

typedef struct ps_devinfo {
  char    name[64];
  char    arch[32];        /* gcnArchName, e.g. gfx950:sramecc+:xnack-      */
  int32_t cu_count;
  int32_t wave_size;
  int64_t lds_bytes_per_block;
  int64_t total_mem_bytes;
  int64_t l2_bytes;
  double  hbm_peak_gbs;    /* 4 x reported memory clock x bus width (HBM3E)  */
  int32_t clock_khz;
  int32_t mem_clock_khz;
  int32_t mem_bus_width;
} ps_devinfo;

typedef struct ps_kernel_s* ps_kernel;

/* ---- errors and version ---- */
/* Replaces the typed exceptions of src/ParallelKernel/Exceptions.jl:2-67. */
const char* ps_last_error(void);
int ps_version(void);

/* ---- device ---- */
/* Replaces backend initialisation in src/ParallelKernel/init_parallel_kernel.jl:34-105
 * (package import + Data module creation) for the one supported device. */
int ps_init(int device);
/* Replaces get_amdgpu_compute_capability, src/ParallelKernel/AMDGPUExt/shared.jl:36-40. */
int ps_device_info(ps_devinfo* out);
/* Replaces @synchronize() without arguments, src/ParallelKernel/parallel.jl:175-184. */
int ps_device_sync(void);

/* ---- fields (allocators) ---- */
/* Replaces zeros_amdgpu / ones_amdgpu / rand_amdgpu / fill_amdgpu,
 * src/ParallelKernel/AMDGPUExt/allocators.jl:3-27, behind @zeros/@ones/@rand/@fill
 * (src/ParallelKernel/allocators.jl:279-387) and the Field allocators
 * (src/ParallelKernel/FieldAllocators.jl:569-637). The x-pitch is rounded up
 * to a multiple of 256 bytes so that every x-row starts on a 256-byte boundary. */
int ps_field_alloc(int ndims, const int64_t dims[3], int dtype, ps_field* out);
/* A CellArray: @zeros(dims...; celldims, blocklength) and @zeros(dims...; celltype)
 * (src/ParallelKernel/allocators.jl:1-17,436-454; the CellArray type of CellArrays.jl).
 * blocklength 0 stores each cell field contiguously over all cells (the reference's GPU
 * default); 1 stores cells whole. h2d/d2h/fill/copy of a CellArray move its data array. */
int ps_cellarray_alloc(int ndims, const int64_t dims[3], int cellndims, const int32_t celldims[4], int64_t blocklength, int dtype,
                       ps_field* out);
int ps_field_free(ps_field* f);
/* A view of the index box lo..hi (1-based, inclusive) of a parent field; replaces
 * the subarray views of padded Fields (FieldAllocators.jl:622-634). */
int ps_field_view(const ps_field* parent, const int64_t lo[3], const int64_t hi[3], ps_field* out);
/* Host <-> device copies of a dense column-major host array of the field's logical
 * size (Array(A) / copy!(A, host) in the reference); for a CellArray, of its data array.
 * stream NULL = stream 0; both block. */
int ps_field_h2d(const ps_field* f, const void* host, void* stream);
int ps_field_d2h(const ps_field* f, void* host, void* stream);
/* fill!(A, x): replaces fill!_amdgpu (AMDGPUExt/allocators.jl). Asynchronous on stream. */
int ps_field_fill(const ps_field* f, double value, void* stream);
/* copy!(dst, src) between fields of equal logical size. Asynchronous on stream. */
int ps_field_copy(const ps_field* dst, const ps_field* src, void* stream);

/* ---- streams ---- */
/* Replaces get_priority_rocstream(id) (priority=1) and get_rocstream(id) (priority=0),
 * src/ParallelKernel/AMDGPUExt/shared.jl:17-31. Both pools hold non-blocking streams at
 * the default hardware priority: the two pools are kept apart for the launch order of
 * @hide_communication, not by queue priority. */
int ps_stream_get(int id, int priority, void** out);
/* Replaces synchronize_amdgpu(stream), src/ParallelKernel/parallel.jl:376. */
int ps_stream_sync(void* stream);

/* ---- kernels ---- */
/* Defines a kernel from its source text. Replaces the kernel-definition paths
 * parallel()/parallel_kernel() (src/parallel.jl:192-215, 386-446) and
 * parallel_indices() (src/parallel.jl:292-370): the body is validated
 * (src/shared.jl:51-57), on-the-fly arrays are extracted (src/parallel.jl:789-819),
 * masks applied (src/parallel.jl:654-680), FD macros expanded
 * (src/FiniteDifferences.jl), padding/index rewrites done
 * (src/ParallelKernel/parallel.jl:487-557). opts_json is a JSON object with the
 * @init_parallel_stencil / per-kernel keywords: kind ("parallel" |
 * "parallel_indices"), indices, ndims, numbertype, inbounds, padding, memopt,
 * loopsize, optvars, optranges, nthreads_max_memopt. */
int ps_kernel_define(const char* text, const char* opts_json, ps_kernel* out);
int ps_kernel_free(ps_kernel k);
/* Derivative kernel of `@parallel ∇=(A->Ā, ...) ad_mode=... ad_annotations=... f!(A, ...)`
 * (replaces parallel_call_ad, src/ParallelKernel/parallel.jl:265-301, and the Enzyme call
 * AD.autodiff_deferred! it expands to, src/ParallelKernel/EnzymeExt/autodiff_gpu.jl).
 * spec_json: {"mode": "Reverse"|"Forward" (Enzyme. prefix allowed; default Reverse),
 * "annotations": {param: "Const"|"Duplicated"|"DuplicatedNoNeed"}} with unnamed parameters
 * Const (defaults of src/ParallelKernel/shared.jl:73-76 are applied by the caller). The new
 * kernel takes the primal's arguments followed by one shadow array per Duplicated* parameter,
 * in parameter order, and is launched with the primal's ranges. Reverse mode adds into the
 * shadows of read arrays and consumes (zeroes) the shadows of written ones; forward mode writes
 * the tangents of written arrays. Writes to DuplicatedNoNeed arrays are dropped. */
int ps_kernel_derivative(ps_kernel primal, const char* spec_json, ps_kernel* out);
/* JSON description of a defined kernel: name, params, written/on-the-fly arrays,
 * nb_parallel_indices and the memopt metadata of src/memopt.jl:1179-1236
 * (stencilranges, loopsize(s), loopdim, optvars). */
int ps_kernel_info(ps_kernel k, char* out, size_t cap);
/* The generated HIP source for the given argument kinds (values are ignored): the
 * counterpart of inspecting the expansion with @prettystring
 * (src/ParallelKernel/shared.jl:553-570). flags: bit0 memopt, bit1 checked, bit2 64-bit
 * indices, bit3 no __restrict__, bit4 1024-thread bound, bit5 LDS tiles (block shape in bits
 * 8..19 / 20..31), bit6 32-bit per-thread byte offsets and bit7 24-bit cell-offset multiplies
 * (plain kernels), bits 8..11 and 12..15 with bit6 and without bit5: x and z cells per thread
 * (groups of cells). */
int ps_kernel_source(ps_kernel k, const ps_arg* args, int nargs, int flags, char* out, size_t cap);
/* Compiles (or loads from the cache) the code object for the given argument kinds
 * without touching a GPU. Used to fill the kernel cache ahead of time. */
int ps_kernel_precompile(ps_kernel k, const ps_arg* args, int nargs, int flags);
/* Same, for exactly the variant that ps_kernel_launch would pick for these arguments (field
 * pointers may be NULL) and launch parameters: geometry, bounds proof, LDS tiling. */
int ps_kernel_precompile_launch(ps_kernel k, const ps_arg* args, int nargs, const ps_launch* launch);
/* Launches the kernel. Replaces parallel_call_gpu_or_xpu / create_gpu_or_xpu_call /
 * AMDGPU.@roc (src/ParallelKernel/parallel.jl:303-334, 883-927) and, for memopt
 * kernels, parallel_call_memopt (src/parallel.jl:451-476). Out-of-bounds accesses
 * that cannot be excluded on the host are checked in the kernel and reported as
 * PS_EBOUNDS at the next synchronisation. */
int ps_kernel_launch(ps_kernel k, const ps_arg* args, int nargs, const ps_launch* launch);
/* Launch geometry the library would use: nthreads and nblocks, 0/1 memopt. */
int ps_kernel_geometry(ps_kernel k, const ps_arg* args, int nargs, const ps_launch* launch,
                       int32_t nthreads[3], int32_t nblocks[3], int64_t ranges[3][2], int64_t* lds_bytes);

/* ---- launch-geometry helpers (host only) ---- */
/* compute_nthreads, src/ParallelKernel/parallel.jl:805-811. */
int ps_compute_nthreads(const int64_t maxsize[3], int nthreads_x_max, int nthreads_max, int flatdim,
                        int32_t out[3]);
/* compute_nblocks, src/ParallelKernel/parallel.jl:813-817. */
int ps_compute_nblocks(const int64_t maxsize[3], const int32_t nthreads[3], int32_t out[3]);
/* get_ranges_outer / get_ranges_inner, src/ParallelKernel/hide_communication.jl:350-394.
 * out receives up to 6 boxes; *nout their number. */
int ps_ranges_outer(const int64_t ranges[3][2], const int64_t boundary_width[3], int64_t out[6][3][2], int* nout);
int ps_ranges_inner(const int64_t ranges[3][2], const int64_t boundary_width[3], int64_t out[3][2]);

/* ---- reductions (residual norms of the pseudo-transient solvers) ---- */
/* FP64 reductions over a field on the GPU; replaces the host-side
 * mean(abs.(R)) / maximum(abs.(R)) of miniapps/Stokes2D.jl:117-123 and the norm(R) of
 * miniapps/HydroMech2D.jl:171.
 * op: 0 sum, 1 sum of |x|, 2 max |x|, 3 mean |x|, 4 L2 norm sqrt(sum x^2). Blocks on the stream. */
int ps_field_reduce(const ps_field* f, int op, double* out, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight solver loops, harnesses that forbid hipMalloc): ps_field_reduce of nfields
 * fields (1..4) in one call, asynchronous on `stream`, out[i] (DEVICE memory, nfields doubles)
 * receiving the statistic of fields[i]. The partial sums live in a static device array of the
 * library (256 KiB per GPU, held until the process ends); the call allocates nothing, copies
 * nothing to the host and does not wait, so it can be captured into a graph. More than 4 fields
 * fail with PS_EARGUMENT. Calls that overlap on one GPU share that array and race. */
int ps_field_reduce_noalloc(const ps_field* fields, int nfields, int op, double* out, void* stream);

/* ---- events (measurement; no reference counterpart: the reference times with
 * wall-clock T_eff prints, README.md:457-464) ---- */
int ps_event_create(void** ev);
int ps_event_record(void* ev, void* stream);
int ps_event_elapsed_ms(void* ev0, void* ev1, float* ms);
int ps_event_destroy(void* ev);

/* ---- graphs (no reference counterpart: the reference launches each @parallel call on its
 * own, src/ParallelKernel/parallel.jl:883-927; these batch launch-bound loops such as the
 * pseudo-transient iteration of miniapps/Stokes2D.jl:109-125) ----
 * ps_graph_begin starts capturing the launches issued on `stream` (launches must be
 * asynchronous and their kernels already compiled: a blocking launch fails during capture);
 * ps_graph_end instantiates the captured work; ps_graph_launch replays it on a stream. */
int ps_graph_begin(void* stream);
int ps_graph_end(void* stream, void** exec);
int ps_graph_launch(void* exec, void* stream);
int ps_graph_destroy(void* exec);

/* ---- kernel cache ---- */
/* Directory of the content-addressed code-object cache (default: $PS_AMD_CACHE or
 * the kcache/ folder beside the library). */
int ps_set_cache_dir(const char* dir);

#ifdef __cplusplus
}
#endif
#endif /* PS_AMD_H */
