// Stencil IR and HIP lowering (host-only C++).
#pragma once
#include <array>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "ps_ast.h"
#include "ps_json.h"

namespace ps {

enum class Ty { Int, F64, F32, Bool };

struct ArgKind {
  bool is_field = false;
  int dtype = 0;   // PS_F64 ...
  int ndims = 0;   // fields only
  // CellArray fields: values per cell (0: plain field), cell dimensions, storage layout
  // (1 one block of all cells, 2 cells stored whole, 3 blocks of blocklen cells) and the
  // comma-separated field names of a @CellType
  int ncomp = 0;
  int celldims[4] = {1, 1, 1, 1};
  int cell_layout = 0;
  std::string cellnames;
};

struct MemoptMeta {
  bool enabled = false;
  int loopdim = 0;      // 1-based dimension streamed by each thread
  int loopsize = 32;
  std::vector<std::string> optvars;
  // per optvar: stencil range [min,max] per dimension (3 entries)
  std::map<std::string, std::vector<std::pair<int, int>>> stencilranges;
  std::map<std::string, std::vector<std::pair<int, int>>> optranges;  // explicit optranges
  // per optvar: offsets keyed by (ox,oy) -> oz -> count (loopdim 3) or (ox) -> oy (loopdim 2)
  std::map<std::string, std::map<std::pair<int, int>, std::map<int, int>>> offsets;
  // same for the read-only arrays that are not optvars (register prefetch only)
  std::map<std::string, std::map<std::pair<int, int>, std::map<int, int>>> ro_offsets;
  std::vector<std::string> shmem_optvars;
  std::map<std::string, bool> use_shmemhalos;
};

struct Options {
  std::string kind = "parallel";          // parallel | parallel_indices
  std::vector<std::string> indices;       // parallel_indices names
  int ndims = 3;                          // @parallel kernels
  std::string numbertype = "Float64";
  bool inbounds = false;
  bool padding = false;
  bool memopt = false;
  int loopsize = 32;
  int loopdim = 0;
  std::vector<std::string> optvars;
  std::map<std::string, std::vector<std::pair<int, int>>> optranges;
  // memopt block shape limits; 0: automatic (128 x 4, or 256 x 4 for 3-D streams whose x range
  // is a multiple of 256 of at least 512 cells: wider rows of HBM per wave and a smaller LDS halo
  // share than 64 x 4, measured fastest at 512^3)
  int nthreads_max_memopt = 0;
  int nthreads_x_max_memopt = 0;
  int nthreads_x_max = 64;
  int nthreads_max = 256;
  bool fp_contract = false;               // Julia semantics: no FMA contraction
  // FMA contraction, division by uniform divisors as multiplication by the host reciprocal and
  // by other values through a refined hardware reciprocal (within a few ulp per operation)
  bool fastmath = false;
  int prefetch = 1;                       // memopt: loads are issued this many steps ahead
  // plain kernels: adjacent x cells per thread (x groups); 0: automatic (4 for arrays of 2 and 4
  // bytes, 1 for 8 bytes), 1: off
  int cells_x = 0;
  // plain kernels run as x groups: threads along x of the launched block (the block keeps its
  // thread count); 0: automatic (128 cells along x per block row in 3-D, 256 in 2-D)
  int nthreads_x_groups = 0;
  // plain kernels run as groups: adjacent z cells per thread (a z plane that several of a thread's
  // cells read is loaded once); 0: automatic (2 for arrays of 8 bytes, 1 otherwise), 1: off
  int cells_z = 0;
  // plain kernels: 32-bit per-thread byte offsets on scalar bases where they fit (Variant::off32)
  bool addr32 = true;
  bool lds = true;                        // memopt: stage x-y neighbours of optvars through LDS
  bool xcd_remap = false;                 // XCD-aware block ordering for memopt kernels
  // 3-D plain kernels run as x groups: an odd number of block columns along x, so that the
  // round-robin dispatch of blocks to the 8 XCDs rotates the x columns over them
  bool xcd_spread = true;
  // nontemporal stores for array writes: 1 on, 0 off, -1 automatic (on for the 256-wide
  // LDS-tiled memopt blocks, where they measured faster; off elsewhere)
  int nt_store = -1;
  // arguments that are tuples / NamedTuples of arrays (Data.ArrayTuple, Data.NamedArrayTuple):
  // name -> component names ("1", "2", ... for plain tuples); each component becomes a
  // parameter of its own, named name.component
  std::vector<std::pair<std::string, std::vector<std::string>>> tuple_params;
};

struct Kernel {
  KernelDef def;
  Options opt;
  int nidx = 3;                           // number of parallel indices
  NodeP body;                             // processed statements
  std::set<std::string> written;          // params assigned to
  std::set<std::string> accessed;         // params indexed
  std::vector<std::string> otf_vars;
  std::map<std::string, int> param_index;
  MemoptMeta mm;
  // kernel language (@parallel_indices kernels): local arrays created by @sharedMem (name ->
  // number of dimensions), and whether the kernel needs 64-bit integers (wavefront masks)
  std::map<std::string, int> shared_arrays;
  bool kernel_language = false;
  bool needs_idx64 = false;
  std::vector<std::string> def_params;  // parameters as written (tuple arguments not yet split)
  // padding=true: the (written array B, read array A, dimension 1..3) triples of the index shifts
  // `size(B.parent, d) > size(A.parent, d)` left in the body as __pgt(B, A, d) calls; a launch
  // replaces them by their values for its arguments (specialize_padding)
  std::vector<std::array<int, 3>> pgt_terms;
};

// padding=true kernels whose body compares parent sizes (pgt_terms): the kernel with every
// __pgt(B, A, d) replaced by bits[i] of its term i, indices canonicalised and the memopt analysis
// redone on the result.
Kernel specialize_padding(const Kernel& K, const std::string& bits);

struct Variant {
  bool memopt = false;
  bool checked = false;   // bounds-checked accesses + error flag
  bool idx64 = false;     // 64-bit index arithmetic
  bool noalias = true;    // __restrict__ on array pointers
  int maxthreads = 256;   // __launch_bounds__
  bool lds = false;       // memopt: stage x-y neighbours of optvars through LDS tiles
  bool nt = false;        // nontemporal stores for array writes
  // plain kernels: affine accesses address an SGPR base (array + uniform offset) plus a 32-bit
  // per-thread byte offset; the host sets it when every thread's cell offset is non-negative and
  // every byte offset of the launch box fits in 31 bits
  bool off32 = false;
  // plain kernels: cell offsets from 24-bit multiplies (every stride and index below 2^24)
  bool u24 = false;
  // plain kernels: per argument, a letter shared by fields with equal strides (one cell offset
  // serves them all), '-' for scalars; empty: every field computes its own
  std::string sclass;
  // plain kernels: adjacent x cells per thread (1: one cell per thread)
  int cx = 1;
  // plain kernels: adjacent z cells per thread (groups of cx x cz cells; 1: one z cell)
  int cz = 1;
  int bx = 0, by = 0;     // block shape baked into LDS variants
  // per argument: 'p' / 'n' a positive / negative float scalar accepted as reciprocal divisor,
  // 'x' a float scalar that must be divided by exactly, '-' anything else
  std::string divmode;
  // padding=true: results of the kernel's parent-size comparisons ('0'/'1' per term)
  std::string padbits;
  std::string key() const;
};

struct Emitted {
  std::string source;
  std::string entry;
  // kernel parameter layout: one entry per parameter in order
  struct P {
    int kind;      // 0 field ptr, 1 int size, 2 scalar, 3 range start, 4 range len, 5 err ptr,
                   // 6 RN(1/|b|) of a uniform divisor or 0 (exact division), 7 RN(1/b) (fastmath),
                   // 8 bytes of dynamic LDS of the launch (@sharedMem), 9 blocklen of a CellArray,
                   // 10 view box of a field in its parent (padding=true): which 0..2 first, 3..5 last index
                   // 11 value of the host-evaluated scalar subexpression uprogs[arg]
    int arg;       // arg index (fields/scalars)
    int which;     // for sizes: 0 nx 1 ny 2 nz 3 sx 4 sxy; for ranges: dim
    int bytes;     // 4 or 8
    int dtype;     // scalar dtype
  };
  std::vector<P> params;
  int lds_bytes = 0;
  // Subexpressions of scalar arguments and literals only (dt/ρ, dt*k): the host evaluates them
  // once per launch, in the kernel's arithmetic, as postfix programs over a typed stack.
  struct UOp {
    enum Code { Arg, Lit, Neg, Add, Sub, Mul, Div, Pow0, Pow2, Pow3, PowM1, PowM2 } op;
    int arg = -1;         // Arg: argument index
    int ty = 0;           // type of the result: 0 int (ps_int), 1 float, 2 double
    double lit = 0.0;     // Lit of a float type
    long long ilit = 0;   // Lit of the integer type
  };
  struct UProg {
    std::vector<UOp> ops;
    int ty = 2;           // 1 float, 2 double
    bool idx64 = false;   // ps_int is 64-bit (integer arguments and literals are truncated to it)
  };
  std::vector<UProg> uprogs;
};

// Divisors for which the reciprocal-based division of the generated code is exact.
bool rcp_safe64(double b);
bool rcp_safe32(float b);

Kernel define_kernel(const std::string& text, const JVal& opts);
// Recomputes the parameter index and the accessed / written sets after the body or the
// parameters of a kernel changed.
void refresh_kernel(Kernel& k);
// memopt metadata (z-streaming plan) of a kernel whose body is final.
void analyse_memopt(Kernel& k);
// Rewrites the index expressions of every array access to the canonical form var +/- c.
NodeP canonical_indices(const NodeP& n);
// Derivative kernel of `@parallel ∇=...` (reverse or forward mode): spec is
// {"mode": "Reverse"|"Forward", "annotations": {param: "Const"|"Duplicated"|"DuplicatedNoNeed"},
// "form": "scatter"|"gather"|"clear"}; the derivative takes the primal's parameters followed by
// one shadow per duplicated array. Reverse mode of @parallel kernels also has a two-kernel form:
// "gather" (each point sums the adjoints that reach it) then "clear" (seeds and kept primal
// updates); it raises an ArgumentError where it does not apply.
Kernel derivative_kernel(const Kernel& k, const JVal& spec);
Emitted lower_kernel(const Kernel& k, const std::vector<ArgKind>& kinds, const Variant& v);
// Whether a plain kernel's body can run as groups of cells (Variant::cx, Variant::cz): a sequence
// of array assignments, each optionally under an if without else whose condition bounds the x
// index (and, for z groups, the z index) by an interval; no local variables, loops, kernel
// language or atomic updates.
bool groupable_body(const Kernel& k, bool z = false);
// Largest spread of offsets along dimension q (0 x, 2 z) with which the kernel body accesses one
// array (0: no neighbours along q).
int idx_span(const Kernel& k, int q);
std::string kernel_info_json(const Kernel& k);

// Bounds proof: true if every array access of the kernel stays inside its
// array for the given launch box and argument shapes.
struct LaunchShape {
  std::vector<ArgKind> kinds;
  std::vector<std::array<long long, 3>> dims;  // per arg (fields); the parents' for padding=true
  std::vector<std::array<long long, 3>> vlo, vhi;  // padding=true: each field's box in its parent
  std::vector<long long> ivals;                 // integer scalar values
  long long lo[3], hi[3];                       // index box (1-based inclusive)
};
bool prove_in_bounds(const Kernel& k, const LaunchShape& s, bool memopt, std::string* why);

}  // namespace ps
