// Abstract syntax tree and parser for the kernel language accepted by
// @parallel / @parallel_indices: the Julia subset that stencil kernels are
// written in (array assignments through finite-difference macros, explicit
// indexing, scalar arithmetic, comparisons, if/elseif/else, return).
//
// The reference receives these kernels as Julia `Expr` trees
// (src/parallel.jl:386-446, src/ParallelKernel/parallel.jl:189-217); this
// parser produces the equivalent tree from the kernel's source text so that
// the lowering can run without a Julia toolchain.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ps {

struct PsError : std::runtime_error {
  int code;
  PsError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Node;
using NodeP = std::shared_ptr<Node>;

enum class NK {
  Num,      // numeric literal: is_int / is_f32 flags
  Ident,    // identifier (UTF-8 allowed)
  Interp,   // $name  (ParallelStencil.INDICES interpolation)
  EndKw,    // `end` inside an index expression
  Index,    // a[0] indexed by a[1..]
  Call,     // s(a...)
  Macro,    // @s(a...)
  Unary,    // s in {-, +, !}  a[0]
  Binary,   // s in {+,-,*,/,^,÷,%}  a[0] s a[1]
  Cmp,      // chained comparison: a[0] ops[0] a[1] ops[1] a[2] ...
  And,      // a[0] && a[1]
  Or,       // a[0] || a[1]
  Tern,     // a[0] ? a[1] : a[2]
  Tuple,    // (a...)
  Range,    // a[0]:a[1]
  Dot,      // a[0].s  (e.g. Data.Array)
  Typed,    // a[0]::a[1]
  Kw,       // s = a[0]  (keyword inside tuples / calls)
  Str,      // "..."
  // statements
  Assign,   // a[0] = a[1]   (s holds the update operator for +=, -=, ...)
  If,       // a[0]=cond, a[1]=then Block, a[2]=else (Block or If) or null
  Block,    // a... statements
  Return,   // a[0] optional
  For,      // s = loop var, a[0] = Range, a[1] = Block
  FuncDef,  // nested function: s = name, a[0] = Tuple of parameter Idents, a[1] = Block
  Forall,   // @∀ x ∈ X statement: a[0] = variable (Ident or Tuple of Idents), a[1] = X, a[2] = statement
  Splat,    // a[0]...  (the parallel indices I... of @parallel_indices (I...))
};

struct Node {
  NK k;
  std::string s;
  std::vector<std::string> ops;  // comparison chain operators
  std::vector<NodeP> a;
  double num = 0.0;
  long long ival = 0;
  bool is_int = false;
  bool is_f32 = false;
  int line = 0;
  Node(NK k_) : k(k_) {}
};

NodeP mk(NK k, const std::string& s = "", std::vector<NodeP> a = {});
NodeP mknum_int(long long v);
NodeP mknum_float(double v, bool f32 = false);
NodeP clone(const NodeP& n);
std::string to_julia(const NodeP& n);  // re-emit as Julia-like text (for diagnostics / tests)

struct Param {
  std::string name;
  std::string annot;  // type annotation text ("" if none), e.g. "Data.Array"
};

struct KernelDef {
  std::string kind;                  // "parallel" or "parallel_indices" (from the text prefix, may be empty)
  std::vector<std::string> indices;  // parallel indices named in @parallel_indices (ix,iy,...)
  std::vector<std::pair<std::string, NodeP>> prefix_kwargs;  // kw=val given in the text prefix
  std::string name;
  std::vector<Param> params;
  NodeP body;  // Block
};

// Parses "[@parallel[_indices] (ix,iy) kw=val ...] function name(args) body end"
// or the short form "name(args) = (stmt; stmt; return)".
KernelDef parse_kernel(const std::string& text);
// Parses a single expression (used for macro templates).
NodeP parse_expression(const std::string& text);

}  // namespace ps
