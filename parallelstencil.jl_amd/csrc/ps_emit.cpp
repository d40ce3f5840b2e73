// HIP C++ emission for gfx950 and the host-side bounds proof.
//
// The emitted code reproduces the arithmetic of the reference's kernels
// bit for bit where IEEE allows it:
//   - operators are emitted in the parser's (Julia's) association order and
//     the code object is compiled without FMA contraction (Julia does not
//     contract a*b+c unless asked to with muladd/@fastmath);
//   - float literals are retyped to the kernel's number type
//     (literaltypes, src/ParallelKernel/parallel.jl:400-425) and printed as
//     exact hexadecimal literals;
//   - literal integer powers follow Base.literal_pow (x^2 = x*x, x^3 = x*x*x,
//     x^-1 = inv(x), x^-2 = inv(x)*inv(x));
//   - max/min follow Julia's NaN and signed-zero rules;
//   - Int/Int division yields Float64, as in Julia.
// Thread geometry: one thread per (ix,iy,iz) of the launch ranges with an
// early return outside them (add_threadids, src/ParallelKernel/parallel.jl:603-713),
// or, for memopt kernels, one thread per column that streams `loopsize`
// planes along the last parallel dimension with a register window per cached
// stencil column (the register queues of src/memopt.jl:190-464).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>

#include "ps_amd.h"
#include "ps_lower.h"

namespace ps {

static const char* kIdx[3] = {"__ix", "__iy", "__iz"};
static const char* kIdxC[3] = {"ix", "iy", "iz"};

static Ty value_ty(int dtype) {
  switch (dtype) {
    case PS_F64: return Ty::F64;
    case PS_F32: return Ty::F32;
    case PS_BF16: return Ty::F32;
    case PS_I32:
    case PS_I64: return Ty::Int;
    case PS_BOOL: return Ty::Bool;
  }
  throw PsError(PS_EARGUMENT, "unsupported element type");
}
static const char* elem_cty(int dtype) {
  switch (dtype) {
    case PS_F64: return "double";
    case PS_F32: return "float";
    case PS_BF16: return "unsigned short";
    case PS_I32: return "int";
    case PS_I64: return "long long";
    case PS_BOOL: return "unsigned char";
  }
  throw PsError(PS_EARGUMENT, "unsupported element type");
}
static const char* cty(Ty t) {
  switch (t) {
    case Ty::Int: return "ps_int";
    case Ty::F64: return "double";
    case Ty::F32: return "float";
    case Ty::Bool: return "bool";
  }
  return "void";
}
static bool is_float(Ty t) { return t == Ty::F64 || t == Ty::F32; }
static Ty promote(Ty a, Ty b) {
  if (a == Ty::F64 || b == Ty::F64) return Ty::F64;
  if (a == Ty::F32 || b == Ty::F32) return Ty::F32;
  return Ty::Int;
}

static std::string hexlit(double v, Ty t) {
  char b[64];
  if (t == Ty::F32) {
    float f = (float)v;
    if (std::isinf(f)) return f > 0 ? "__builtin_huge_valf()" : "(-__builtin_huge_valf())";
    std::snprintf(b, sizeof b, "%af", (double)f);
  } else {
    if (std::isinf(v)) return v > 0 ? "__builtin_huge_val()" : "(-__builtin_huge_val())";
    std::snprintf(b, sizeof b, "%a", v);
  }
  return b;
}

struct EV {
  std::string c;
  Ty t;
};

static std::string sanitize(const std::string& s) {
  std::string o;
  for (unsigned char c : s) {
    if (std::isalnum(c) || c == '_') o += (char)c;
    else {
      char b[8];
      std::snprintf(b, sizeof b, "_%02x", c);
      o += b;
    }
  }
  return o;
}

// device helpers prepended to every generated kernel
static const char* kPrelude = R"PSD(
#define PS_DI __device__ __forceinline__
PS_DI float ps_bf16_ld(unsigned short v) { return __uint_as_float(((unsigned)v) << 16); }
// FP32 to BFloat16 bits: nearest even, every NaN to 0x7fc0 (the BFloat16(::Float32) of
// BFloat16s.jl); gfx950 rounds in one instruction
PS_DI unsigned short ps_bf16_st(float f) {
  const unsigned short u = __builtin_bit_cast(unsigned short, (__bf16)f);
  return f != f ? (unsigned short)0x7fc0u : u;
}
// the rounding alone, for values that are not NaN or are the canonical NaN 0x7fc00000 (which
// rounds to 0x7fc0); ps_nan_canon maps every NaN to it
PS_DI unsigned short ps_bf16_cvt(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
PS_DI float ps_nan_canon(float f) { return f != f ? __uint_as_float(0x7fc00000u) : f; }
PS_DI double ps_max(double x, double y) { bool py = (y > x) || (__builtin_signbit(y) < __builtin_signbit(x)); return py ? (x != x ? x : y) : (y != y ? y : x); }
PS_DI float  ps_max(float x, float y)   { bool py = (y > x) || (__builtin_signbit(y) < __builtin_signbit(x)); return py ? (x != x ? x : y) : (y != y ? y : x); }
PS_DI double ps_min(double x, double y) { bool py = (y < x) || (__builtin_signbit(y) > __builtin_signbit(x)); return py ? (x != x ? x : y) : (y != y ? y : x); }
PS_DI float  ps_min(float x, float y)   { bool py = (y < x) || (__builtin_signbit(y) > __builtin_signbit(x)); return py ? (x != x ? x : y) : (y != y ? y : x); }
PS_DI ps_int ps_max(ps_int x, ps_int y) { return x > y ? x : y; }
PS_DI ps_int ps_min(ps_int x, ps_int y) { return x < y ? x : y; }
PS_DI double ps_p2(double x) { return x * x; }
PS_DI float  ps_p2(float x)  { return x * x; }
PS_DI double ps_p3(double x) { return x * x * x; }
PS_DI float  ps_p3(float x)  { return x * x * x; }
PS_DI double ps_pm2(double x) { double i = 1.0 / x; return i * i; }
PS_DI float  ps_pm2(float x)  { float i = 1.0f / x; return i * i; }
PS_DI ps_int ps_ipow(ps_int b, ps_int e) { ps_int r = 1; while (e > 0) { if (e & 1) r *= b; b *= b; e >>= 1; } return r; }
// Correctly rounded a/b for a positive divisor b that is uniform over the launch (a scalar
// argument or a literal; a negative divisor is handled as (-a)/(-b)), from r = RN(1/b):
// q0 = RN(a*r), e = q0*b - a (exact by fma), q = RN(q0 - e*r)
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990). The host accepts b only in [2^-400, 2^400] with a
// significand that is not all ones; results outside [2^-500, 2^500] (other than exact zeros,
// whose sign this form gets right), infinities and NaN set `bad`, and the statement is then
// recomputed with IEEE division.
PS_DI double ps_divp(double a, double b, double r, bool& bad) {
  double q0 = a * r;
  double e = __builtin_fma(q0, b, -a);
  double q = __builtin_fma(-e, r, q0);
  double aq = __builtin_fabs(q);
  bad |= !(aq <= 0x1p500) | ((aq < 0x1p-500) & (aq != 0.0));
  return q;
}
PS_DI float ps_divp(float a, float b, float r, bool& bad) {
  float q0 = a * r;
  float e = __builtin_fmaf(q0, b, -a);
  float q = __builtin_fmaf(-e, r, q0);
  float aq = __builtin_fabsf(q);
  bad |= !(aq <= 0x1p50f) | ((aq < 0x1p-50f) & (aq != 0.0f));
  return q;
}
// The same quotient for statements whose value is NaN whenever one of their quotients is NaN
// (NaN-transparent statements): instead of testing every quotient, the frexp exponent of the
// dividend joins the statement's running minimum and maximum. A dividend with |a| in
// [2^-600, 2^600) (double; [2^-80, 2^80) float) or zero keeps every intermediate of the
// division normal for the accepted divisors, so the quotient is exact; an infinite or NaN
// dividend gives a NaN quotient, which reaches the statement's value. The statement falls back
// to IEEE division when an exponent left [-599, 600] ([-79, 80]) or its value is NaN.
PS_DI double ps_divx(double a, double b, double r, int& lo, int& hi) {
  const double q0 = a * r;
  const double e = __builtin_fma(q0, b, -a);
  const int x = __builtin_amdgcn_frexp_exp(a);
  lo = x < lo ? x : lo;
  hi = x > hi ? x : hi;
  return __builtin_fma(-e, r, q0);
}
PS_DI float ps_divx(float a, float b, float r, int& lo, int& hi) {
  const float q0 = a * r;
  const float e = __builtin_fmaf(q0, b, -a);
  const int x = __builtin_amdgcn_frexp_expf(a);
  lo = x < lo ? x : lo;
  hi = x > hi ? x : hi;
  return __builtin_fmaf(-e, r, q0);
}
// fastmath reciprocal: hardware estimate refined by Newton steps r += r (1 - b r) (two for
// double, one for float); the raw estimate is kept where the refinement is not finite
// (b = 0, b = inf, NaN)
PS_DI double ps_frcp(double b) {
  const double r0 = __builtin_amdgcn_rcp(b);
  double r = __builtin_fma(r0, __builtin_fma(-b, r0, 1.0), r0);
  r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
  return r == r ? r : r0;
}
PS_DI float ps_frcp(float b) {
  const float r0 = __builtin_amdgcn_rcpf(b);
  const float r = __builtin_fmaf(r0, __builtin_fmaf(-b, r0, 1.0f), r0);
  return r == r ? r : r0;
}
// Element through a uniform element offset u and a per-thread 32-bit byte offset vo: the address
// form of global loads and stores with an SGPR base and a VGPR offset
// plus a constant element offset k (x neighbours share one base and differ in immediate offsets)
template <class T>
PS_DI T& ps_at(T* p, ps_int u, unsigned vo, int k) { return ((T*)((char*)(p + u) + vo))[k]; }
template <class T>
PS_DI const T& ps_at(const T* p, ps_int u, unsigned vo, int k) { return ((const T*)((const char*)(p + u) + vo))[k]; }
// Guarded load without a branch: element 0 of the array stands in for an index outside it
// (the runtime only launches these kernels on non-empty arrays), and the value is then zeroed.
template <class T>
PS_DI T ps_ldz(const T* p, bool c, ps_int i) {
  T v = p[c ? i : (ps_int)0];
  return c ? v : (T)0;
}
// Same without zeroing, for kernels whose accesses are proven in bounds: every use of such a
// value happens where the access is inside the array, so the stand-in value is never used.
template <class T>
PS_DI T ps_ldc(const T* p, bool c, ps_int i) {
  return p[c ? i : (ps_int)0];
}
// Wavefront shuffles within partitions of w lanes (w a power of two <= 64); a lane whose partner
// lies outside its partition keeps its own value. Every form reads through one full-wave
// permute, so any index wraps inside the wavefront.
template <class T>
PS_DI T ps_shfl(T v, int lane, int w) {
  const int s = __lane_id();
  return __shfl(v, (s & ~(w - 1)) + ((lane - 1) & (w - 1)), 64);
}
template <class T>
PS_DI T ps_shfl_up(T v, unsigned d, int w) {
  const int s = __lane_id();
  return __shfl(v, (d >= (unsigned)w || (unsigned)(s & (w - 1)) < d) ? s : s - (int)d, 64);
}
template <class T>
PS_DI T ps_shfl_down(T v, unsigned d, int w) {
  const int s = __lane_id();
  return __shfl(v, (d >= (unsigned)w || (unsigned)(s & (w - 1)) + d >= (unsigned)w) ? s : s + (int)d, 64);
}
template <class T>
PS_DI T ps_shfl_xor(T v, int m, int w) {
  const int s = __lane_id(), i = s ^ m;
  return __shfl(v, ((i ^ s) & ~(w - 1)) ? s : i, 64);
}
PS_DI bool ps_chk(ps_int i, ps_int j, ps_int k, ps_int nx, ps_int ny, ps_int nz, int* err) {
  bool ok = i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz;
  if (!ok) *err = 1;
  return ok;
}
)PSD";

class Emitter {
 public:
  Emitter(const Kernel& K, const std::vector<ArgKind>& kinds, const Variant& V) : K(K), kinds(kinds), V(V) {
    NT = K.opt.numbertype == "Float32" ? Ty::F32 : Ty::F64;
    if (kinds.size() != K.def.params.size())
      throw PsError(PS_EARGUMENT, "kernel " + K.def.name + " expects " + std::to_string(K.def.params.size()) + " arguments, got " +
                                      std::to_string(kinds.size()));
    for (auto& a : K.accessed) {
      int i = K.param_index.at(a);
      if (!kinds[i].is_field) throw PsError(PS_EARGUMENT, "argument '" + a + "' is indexed in the kernel but is not an array");
    }
    for (auto& w : K.written) {
      int i = K.param_index.at(w);
      if (!kinds[i].is_field) throw PsError(PS_EARGUMENT, "argument '" + w + "' is assigned to but is not an array");
    }
    for (auto& k : kinds) any_cell = any_cell || k.ncomp > 0;
    if (any_cell && V.memopt) throw PsError(PS_EARGUMENT, "memopt=true is not supported with CellArray arguments");
  }

  Emitted run();

 private:
  const Kernel& K;
  const std::vector<ArgKind>& kinds;
  Variant V;
  Ty NT;
  std::ostringstream o;
  int ind = 1;
  std::map<std::string, Ty> locals;
  std::map<std::string, std::string> local_c;
  std::vector<std::string> local_order;
  bool pass1 = false;
  bool in_loop = false;  // inside the memopt stream loop
  // memopt register cache: key (arg, ox, oy, oz) -> register name
  std::map<std::tuple<int, int, int, int>, std::string> cache;
  // scalar arguments used as uniform divisors: their reciprocal travels as a hidden argument
  std::map<int, Ty> rcp_params;
  bool fast_div = false;   // emitting the reciprocal form of a statement
  bool used_fast = false;  // the current statement used it
  // the current statement is NaN-transparent: its quotients use ps_divx, whose checks (per type:
  // [0] double, [1] float) are made once on the statement's value
  bool div_x = false;
  bool used_x[2] = {false, false};
  bool nan_transparent(const NodeP& n) const;
  std::string fast_decls() const;
  std::string fast_bad(const std::string& v) const;
  // scalar subexpressions the host evaluates once per launch: canonical text -> program index
  std::map<std::string, int> uni_index;
  std::vector<Emitted::UProg> uprogs;
  bool uniform_prog(const NodeP& n, std::vector<Emitted::UOp>& ops, Ty& t, int& nops) const;
  int tmp_counter = 0;
  int E_lds_bytes = 0;  // static LDS of the emitted kernel
  EV divide(const NodeP& n);
  EV rhs_value(const NodeP& rhs, bool* opened);
  // kernel language: element dtype of each @sharedMem array; the extent `end` stands for while
  // an index of such an array is emitted
  std::map<std::string, int> sm_dtype;
  std::string end_subst;
  int type_dtype(const NodeP& t) const;
  EV kl_expr(const NodeP& n);
  void kl_stmt(const NodeP& n);
  void sm_create(const std::string& v, const NodeP& mac);
  std::string sm_address(const NodeP& n, std::string* check);
  // CellArrays: the component (0-based) of the cell assignment being emitted, -1 outside one,
  // and the number of values per cell of its left-hand side
  int comp_ctx = -1;
  int comp_nc = 0;
  std::set<std::string> loop_vars;  // variables of the enclosing for loops
  bool any_cell = false;
  bool is_cell(int ai) const { return ai >= 0 && kinds[ai].ncomp > 0; }
  int cell_shape(const NodeP& n);
  bool cell_ref(const NodeP& n, int& ai, NodeP& cellidx, std::string& comp, std::string& compchk);
  std::string cell_offset(int ai, const std::string& I, const std::string& j) const;
  EV cell_load(int ai, const NodeP& cellidx, const std::string& comp, const std::string& compchk);
  void cell_store(int ai, const NodeP& cellidx, const std::string& comp, const std::string& compchk, const EV& r);

  std::string pad() const { return std::string(ind * 2, ' '); }
  std::string an(int i) const { return "a" + std::to_string(i); }
  int field_index(const std::string& name) const {
    auto it = K.param_index.find(name);
    if (it == K.param_index.end() || !kinds[it->second].is_field) return -1;
    return it->second;
  }

  EV cast(const EV& v, Ty t) const {
    if (v.t == t) return v;
    return EV{std::string("((") + cty(t) + ")(" + v.c + "))", t};
  }
  EV lit(double x, Ty t) const { return EV{hexlit(x, t), t}; }

  EV expr(const NodeP& n);
  EV load(const NodeP& idxnode);
  std::string offset(int ai, const std::vector<NodeP>& idx, std::string idxc[3], bool* affine_ok, std::string* ref32 = nullptr);
  std::vector<NodeP> norm_indices(const NodeP& idxnode, int ai);
  EV call(const NodeP& n);
  EV binary(const std::string& op, EV a, EV b, const NodeP& rhsnode);
  EV cmp(const NodeP& n);
  void stmt(const NodeP& s);
  void atomic_add(const NodeP& s);
  void block(const NodeP& b);
  void assign(const NodeP& s);
  // x groups (V.cx > 1): array assignments for all cells of the group, in batches, loads first
  void group_statements(const std::vector<NodeP>& sts);
  std::string group_scope(int c) const;
  bool x_bounds(const NodeP& c, std::vector<std::string>& lo, std::vector<std::string>& hi);
  std::string store_value(int dt, const EV& r) const;
  void emit_memopt_body(int L);
};

std::vector<NodeP> Emitter::norm_indices(const NodeP& n, int ai) {
  std::vector<NodeP> idx(n->a.begin() + 1, n->a.end());
  int fnd = kinds[ai].ndims;
  if ((int)idx.size() > 3) {
    // trailing literal-1 indices beyond the third are allowed
    for (size_t i = 3; i < idx.size(); ++i)
      if (!(idx[i]->k == NK::Num && idx[i]->is_int && idx[i]->ival == 1))
        throw PsError(PS_EARGUMENT, "arrays with more than 3 dimensions are not supported");
    idx.resize(3);
  }
  if (idx.size() == 1 && fnd > 1) {
    // linear indexing of a multi-dimensional array is not supported by the stencil backend
    throw PsError(PS_EARGUMENT, "linear indexing of multi-dimensional arrays is not supported");
  }
  while (idx.size() < 3) idx.push_back(mknum_int(1));
  return idx;
}

// Element offset of A[i,j,k] and the C expressions of the three indices.
// With ref32, an affine access of a plain off32 variant also gets its element reference in the
// SGPR-base form (ps_at).
std::string Emitter::offset(int ai, const std::vector<NodeP>& idx, std::string idxc[3], bool* affine_ok, std::string* ref32) {
  static const char* szn[3] = {"_nx", "_ny", "_nz"};
  std::string A = an(ai);
  long long c[3] = {0, 0, 0};
  bool aff = true;
  for (int d = 0; d < 3; ++d) {
    const NodeP& e = idx[d];
    // affine analysis
    bool ok = false;
    int var = -1;
    long long cc = 0;
    std::function<bool(const NodeP&, int&, long long&)> af = [&](const NodeP& x, int& v, long long& k) -> bool {
      if (x->k == NK::Num && x->is_int) {
        k += x->ival;
        return true;
      }
      if (x->k == NK::Ident) {
        for (int q = 0; q < 3; ++q)
          if (x->s == kIdx[q]) {
            if (v != -1) return false;
            v = q;
            return true;
          }
        return false;
      }
      if (x->k == NK::EndKw) {
        if (v != -1) return false;
        v = 3;
        return true;
      }
      if (x->k == NK::Binary && (x->s == "+" || x->s == "-")) {
        if (!af(x->a[0], v, k)) return false;
        if (x->s == "+") return af(x->a[1], v, k);
        int v2 = -1;
        long long k2 = 0;
        if (!af(x->a[1], v2, k2) || v2 != -1) return false;
        k -= k2;
        return true;
      }
      return false;
    };
    ok = af(e, var, cc);
    if (ok && var == 3) {  // `end` -> size of this dimension
      idxc[d] = "(" + A + szn[d] + " + (" + std::to_string(cc) + "))";
      aff = false;
      continue;
    }
    if (ok && var >= 0) {
      idxc[d] = std::string("(") + kIdxC[var] + (cc >= 0 ? " + " : " - ") + std::to_string(cc >= 0 ? cc : -cc) + ")";
      if (var != d) aff = false;
      c[d] = cc;
      continue;
    }
    if (ok && var == -1) {
      idxc[d] = "((ps_int)" + std::to_string(cc) + ")";
      aff = false;
      continue;
    }
    NodeP ee = e;
    // replace `end` inside a general index expression
    std::function<NodeP(const NodeP&)> rep = [&](const NodeP& x) -> NodeP {
      if (x->k == NK::EndKw) return mk(NK::Call, "size", {mk(NK::Ident, K.def.params[ai].name), mknum_int(d + 1)});
      auto y = std::make_shared<Node>(*x);
      for (auto& ch : y->a) ch = rep(ch);
      return y;
    };
    ee = rep(e);
    EV v = expr(ee);
    if (v.t != Ty::Int) throw PsError(PS_EARGUMENT, "array index must be an integer (got a " + std::string(cty(v.t)) + " expression)");
    idxc[d] = "(" + v.c + ")";
    aff = false;
  }
  if (affine_ok) *affine_ok = aff;
  if (aff) {
    std::string off = A + "_o";
    std::string k;
    if (c[0]) k += " + (" + std::to_string(c[0]) + ")";
    if (c[1]) k += " + (" + std::to_string(c[1]) + ")*" + A + "_sx";
    if (c[2]) k += " + (" + std::to_string(c[2]) + ")*" + A + "_sxy";
    // In x groups the y and z parts form the uniform offset, and the x offset plus the cell's
    // offset within its group (ps_cx, a constant in each cell's scope) stay constant element
    // offsets: the accesses of one row differ by immediate offsets only and merge into wide
    // loads of the group's cells. With one cell per thread the x offset joins the uniform part:
    // merged x neighbours would be 16-byte loads that overlap between lanes (measured 20% slower
    // on the 7-point diffusion step).
    if (ref32 && V.off32 && !V.memopt && kinds[ai].ncomp == 0) {
      std::string kyz;
      if (c[1]) kyz += " + (" + std::to_string(c[1]) + ")*" + A + "_sx";
      if (c[2]) kyz += " + (" + std::to_string(c[2]) + ")*" + A + "_sxy";
      if (V.cz > 1) {  // the cell's z offset within its group joins the uniform part
        kyz.clear();
        if (c[1]) kyz += " + (" + std::to_string(c[1]) + ")*" + A + "_sx";
        kyz += " + (ps_cz + (" + std::to_string(c[2]) + "))*" + A + "_sxy";
      }
      if (V.cx > 1 || V.cz > 1) *ref32 = "ps_at(" + A + ", (ps_int)0" + kyz + ", " + A + "_vo, (int)ps_cx + " + std::to_string(c[0]) + ")";
      else *ref32 = "ps_at(" + A + ", (ps_int)" + std::to_string(c[0]) + kyz + ", " + A + "_vo, 0)";
    }
    return "(" + off + k + ")";
  }
  return "((" + idxc[0] + " - 1) + (" + idxc[1] + " - 1)*" + A + "_sx + (" + idxc[2] + " - 1)*" + A + "_sxy)";
}

EV Emitter::load(const NodeP& n) {
  if (any_cell) {
    int cai;
    NodeP cidx;
    std::string comp, cchk;
    if (cell_ref(n, cai, cidx, comp, cchk)) return cell_load(cai, cidx, comp, cchk);
    if (n->a[0]->k == NK::Ident && is_cell(field_index(n->a[0]->s))) {
      const int ai = field_index(n->a[0]->s);
      if (comp_ctx < 0)
        throw PsError(PS_EARGUMENT, "a cell of CellArray '" + n->a[0]->s + "' is used where a number is expected; select one value with field(" +
                                        n->a[0]->s + ", i)[...] or " + n->a[0]->s + "[...][i]");
      if (kinds[ai].ncomp != comp_nc)
        throw PsError(PS_EARGUMENT, "CellArrays with cells of " + std::to_string(kinds[ai].ncomp) + " and " + std::to_string(comp_nc) +
                                        " values meet in one cell assignment");
      return cell_load(ai, n, std::to_string(comp_ctx), "");
    }
  }
  if (n->a[0]->k != NK::Ident) throw PsError(PS_EARGUMENT, "unsupported indexing expression " + to_julia(n));
  const std::string& name = n->a[0]->s;
  if (K.shared_arrays.count(name)) {  // @sharedMem array in LDS
    std::string chk;
    std::string addr = sm_address(n, &chk);
    int dt = sm_dtype.at(name);
    std::string raw = chk.empty() ? addr : "(" + chk + " ? " + addr + " : (" + elem_cty(dt) + ")0)";
    if (dt == PS_BF16) return EV{"ps_bf16_ld(" + raw + ")", Ty::F32};
    if (dt == PS_BOOL) return EV{"((bool)" + raw + ")", Ty::Bool};
    if (dt == PS_I64 || dt == PS_I32) return EV{"((ps_int)" + raw + ")", Ty::Int};
    return EV{raw, value_ty(dt)};
  }
  int ai = field_index(name);
  if (ai < 0) throw PsError(PS_EARGUMENT, "'" + name + "' is indexed but is not an array argument of the kernel");
  auto idx = norm_indices(n, ai);
  // memopt register cache
  if (in_loop && !cache.empty()) {
    long long off[3];
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
      int v = -1;
      long long k = 0;
      const NodeP& e = idx[d];
      std::function<bool(const NodeP&)> af = [&](const NodeP& x) -> bool {
        if (x->k == NK::Num && x->is_int) {
          k += x->ival;
          return true;
        }
        if (x->k == NK::Ident) {
          for (int q = 0; q < 3; ++q)
            if (x->s == kIdx[q] && v == -1) {
              v = q;
              return true;
            }
          return false;
        }
        if (x->k == NK::Binary && x->s == "+") return af(x->a[0]) && af(x->a[1]);
        if (x->k == NK::Binary && x->s == "-" && x->a[1]->k == NK::Num && x->a[1]->is_int) {
          bool r = af(x->a[0]);
          k -= x->a[1]->ival;
          return r;
        }
        return false;
      };
      if (d < K.nidx) {
        if (!af(e) || v != d) ok = false;
      } else {
        if (!(e->k == NK::Num && e->is_int && e->ival == 1)) ok = false;
        k = 0;
      }
      off[d] = k;
    }
    if (ok) {
      auto it = cache.find(std::make_tuple(ai, (int)off[0], (int)off[1], (int)off[2]));
      if (it != cache.end()) return EV{it->second, value_ty(kinds[ai].dtype)};
    }
  }
  std::string ic[3];
  bool aff = false;
  std::string ref32;
  std::string off = offset(ai, idx, ic, &aff, &ref32);
  std::string A = an(ai);
  std::string raw = ref32.empty() ? A + "[" + off + "]" : ref32;
  int dt = kinds[ai].dtype;
  std::string val;
  if (V.checked) {
    std::string chk = "ps_chk(" + ic[0] + ", " + ic[1] + ", " + ic[2] + ", " + A + "_nx, " + A + "_ny, " + A + "_nz, ps_err)";
    raw = "(" + chk + " ? " + raw + " : (" + elem_cty(dt) + ")0)";
  }
  if (dt == PS_BF16) val = "ps_bf16_ld(" + raw + ")";
  else if (dt == PS_BOOL) val = "((bool)" + raw + ")";
  else if (dt == PS_I64 || dt == PS_I32) val = "((ps_int)" + raw + ")";
  else val = raw;
  return EV{val, value_ty(dt)};
}

EV Emitter::binary(const std::string& op, EV a, EV b, const NodeP& rhsnode) {
  if (op == "^") {
    // literal integer exponent -> Base.literal_pow
    long long k = 0;
    bool literal = false;
    if (rhsnode && rhsnode->k == NK::Num && rhsnode->is_int) {
      k = rhsnode->ival;
      literal = true;
    } else if (rhsnode && rhsnode->k == NK::Unary && rhsnode->s == "-" && rhsnode->a[0]->k == NK::Num && rhsnode->a[0]->is_int) {
      k = -rhsnode->a[0]->ival;
      literal = true;
    }
    if (a.t == Ty::Bool) a = cast(a, Ty::Int);
    if (literal) {
      if (a.t == Ty::Int) {
        if (k < 0) {
          EV f = cast(a, Ty::F64);
          return binary("^", f, b, rhsnode);
        }
        return EV{"ps_ipow(" + a.c + ", " + std::to_string(k) + ")", Ty::Int};
      }
      const char* one = a.t == Ty::F32 ? "1.0f" : "1.0";
      switch (k) {
        case 0: return EV{one, a.t};
        case 1: return a;
        case 2: return EV{"ps_p2(" + a.c + ")", a.t};
        case 3: return EV{"ps_p3(" + a.c + ")", a.t};
        case -1: return EV{"(" + std::string(one) + " / " + a.c + ")", a.t};
        case -2: return EV{"ps_pm2(" + a.c + ")", a.t};
        default:
          return EV{std::string(a.t == Ty::F32 ? "powf(" : "pow(") + a.c + ", " + (a.t == Ty::F32 ? "(float)" : "(double)") + std::to_string(k) + ")", a.t};
      }
    }
    Ty t = promote(a.t == Ty::Int ? Ty::F64 : a.t, b.t == Ty::Int ? (a.t == Ty::Int ? Ty::F64 : a.t) : b.t);
    if (a.t == Ty::Int && b.t == Ty::Int) return EV{"ps_ipow(" + a.c + ", " + b.c + ")", Ty::Int};
    EV x = cast(a, t), y = cast(b, t);
    return EV{std::string(t == Ty::F32 ? "powf(" : "pow(") + x.c + ", " + y.c + ")", t};
  }
  if (op == "*" && (a.t == Ty::Bool) != (b.t == Ty::Bool) && is_float(a.t == Ty::Bool ? b.t : a.t)) {
    // Julia's Bool * x = ifelse(b, x, copysign(zero(x), x)): a strong zero (false * Inf == 0)
    const EV& c = a.t == Ty::Bool ? a : b;
    const EV& x = a.t == Ty::Bool ? b : a;
    bool f32 = x.t == Ty::F32;
    return EV{"(" + c.c + " ? " + x.c + " : " + (f32 ? "copysignf(0.0f, " : "copysign(0.0, ") + x.c + "))", x.t};
  }
  if (a.t == Ty::Bool) a = cast(a, Ty::Int);
  if (b.t == Ty::Bool) b = cast(b, Ty::Int);
  if (op == "/") {
    Ty t = promote(a.t, b.t);
    if (t == Ty::Int) t = Ty::F64;
    EV x = cast(a, t), y = cast(b, t);
    return EV{"(" + x.c + " / " + y.c + ")", t};
  }
  if (op == "÷") {
    if (a.t != Ty::Int || b.t != Ty::Int) throw PsError(PS_EARGUMENT, "÷ is only supported for integers");
    return EV{"(" + a.c + " / " + b.c + ")", Ty::Int};
  }
  if (op == "%") {
    Ty t = promote(a.t, b.t);
    EV x = cast(a, t), y = cast(b, t);
    if (t == Ty::Int) return EV{"(" + x.c + " % " + y.c + ")", t};
    return EV{std::string(t == Ty::F32 ? "fmodf(" : "fmod(") + x.c + ", " + y.c + ")", t};
  }
  if (op == "+" || op == "-" || op == "*") {
    Ty t = promote(a.t, b.t);
    EV x = cast(a, t), y = cast(b, t);
    return EV{"(" + x.c + " " + op + " " + y.c + ")", t};
  }
  if (op == "&" || op == "|") {
    return EV{"(" + a.c + " " + op + " " + b.c + ")", Ty::Int};
  }
  throw PsError(PS_EARGUMENT, "unsupported operator " + op);
}

bool rcp_safe64(double b) {
  double ab = std::fabs(b);
  if (!(ab >= 0x1p-400 && ab <= 0x1p400)) return false;
  unsigned long long bits;
  std::memcpy(&bits, &b, 8);
  return (bits & ((1ULL << 52) - 1)) != ((1ULL << 52) - 1);
}
bool rcp_safe32(float b) {
  float ab = std::fabs(b);
  if (!(ab >= 0x1p-40f && ab <= 0x1p40f)) return false;
  unsigned bits;
  std::memcpy(&bits, &b, 4);
  return (bits & 0x7fffffu) != 0x7fffffu;
}

// a / b. Uniform float divisors (scalar arguments the host accepted, literals) are divided
// through their reciprocal (ps_divp) when the enclosing statement provides a fallback.
EV Emitter::divide(const NodeP& n) {
  EV a = expr(n->a[0]);
  const NodeP& d = n->a[1];
  EV b = expr(d);
  if (a.t == Ty::Bool) a = cast(a, Ty::Int);
  if (b.t == Ty::Bool) b = cast(b, Ty::Int);
  Ty t = promote(a.t, b.t);
  if (t == Ty::Int) t = Ty::F64;
  if (K.opt.fastmath && (t == Ty::F64 || t == Ty::F32)) {
    // a * RN(1/b): the host reciprocal of a scalar argument, the rounded reciprocal of a
    // literal, otherwise the hardware reciprocal refined by Newton steps
    if (d->k == NK::Ident && !locals.count(d->s)) {
      auto it = K.param_index.find(d->s);
      int k = it == K.param_index.end() ? -1 : it->second;
      if (k >= 0 && !kinds[k].is_field && value_ty(kinds[k].dtype) == t) {
        rcp_params[k] = t;
        return EV{"(" + cast(a, t).c + " * " + an(k) + "_r)", t};
      }
    }
    if (d->k == NK::Num && !d->is_int) {
      double v = (d->is_f32 || NT == Ty::F32) ? (double)(float)d->num : d->num;
      double r = t == Ty::F32 ? (double)(1.0f / (float)v) : 1.0 / v;
      return EV{"(" + cast(a, t).c + " * " + hexlit(r, t) + ")", t};
    }
    return EV{"(" + cast(a, t).c + " * ps_frcp(" + cast(b, t).c + "))", t};
  }
  if (fast_div && d->k == NK::Ident && !locals.count(d->s)) {
    auto it = K.param_index.find(d->s);
    int k = it == K.param_index.end() ? -1 : it->second;
    if (k >= 0 && !kinds[k].is_field && value_ty(kinds[k].dtype) == t) {
      char mode = (size_t)k < V.divmode.size() ? V.divmode[k] : 'x';
      if (mode == 'p' || mode == 'n') {
        rcp_params[k] = t;
        used_fast = true;
        std::string A = an(k);
        std::string tail = ", ps_bad)";
        if (div_x) {
          used_x[t == Ty::F64 ? 0 : 1] = true;
          tail = t == Ty::F64 ? ", ps_lo, ps_hi)" : ", ps_lo32, ps_hi32)";
        }
        const std::string fn = div_x ? "ps_divx(" : "ps_divp(";
        if (mode == 'p') return EV{fn + cast(a, t).c + ", " + A + ", " + A + "_r" + tail, t};
        return EV{fn + "-(" + cast(a, t).c + "), -" + A + ", " + A + "_r" + tail, t};
      }
    }
  }
  if (fast_div && d->k == NK::Num && !d->is_int) {
    double v = (d->is_f32 || NT == Ty::F32) ? (double)(float)d->num : d->num;
    bool ok = t == Ty::F32 ? rcp_safe32((float)v) : rcp_safe64(v);
    if (ok && v > 0) {
      used_fast = true;
      double r = t == Ty::F32 ? (double)(1.0f / (float)v) : 1.0 / v;
      if (div_x) {
        used_x[t == Ty::F64 ? 0 : 1] = true;
        return EV{"ps_divx(" + cast(a, t).c + ", " + hexlit(v, t) + ", " + hexlit(r, t) + (t == Ty::F64 ? ", ps_lo, ps_hi)" : ", ps_lo32, ps_hi32)"), t};
      }
      return EV{"ps_divp(" + cast(a, t).c + ", " + hexlit(v, t) + ", " + hexlit(r, t) + ", ps_bad)", t};
    }
  }
  return binary("/", a, b, nullptr);
}

EV Emitter::cmp(const NodeP& n) {
  std::vector<EV> v;
  for (auto& x : n->a) v.push_back(expr(x));
  std::string c;
  for (size_t i = 0; i + 1 < v.size(); ++i) {
    EV a = v[i], b = v[i + 1];
    if (a.t == Ty::Bool && b.t != Ty::Bool) a = cast(a, Ty::Int);
    if (b.t == Ty::Bool && a.t != Ty::Bool) b = cast(b, Ty::Int);
    if (a.t != Ty::Bool) {
      Ty t = promote(a.t, b.t);
      a = cast(a, t);
      b = cast(b, t);
    }
    std::string part = "(" + a.c + " " + n->ops[i] + " " + b.c + ")";
    c = c.empty() ? part : "(" + c + " && " + part + ")";
  }
  return EV{c, Ty::Bool};
}

EV Emitter::call(const NodeP& n) {
  const std::string& f = n->s;
  auto argv = [&](size_t i) { return expr(n->a.at(i)); };
  if (f == "size") {
    if (n->a.empty() || n->a[0]->k != NK::Ident) throw PsError(PS_EARGUMENT, "size() expects an array argument");
    int ai = field_index(n->a[0]->s);
    if (ai < 0) throw PsError(PS_EARGUMENT, "size() of '" + n->a[0]->s + "': not an array argument");
    if (n->a.size() != 2 || !(n->a[1]->k == NK::Num && n->a[1]->is_int)) throw PsError(PS_EARGUMENT, "size(A, d) requires a literal dimension d");
    long long d = n->a[1]->ival;
    if (d < 1) throw PsError(PS_EARGUMENT, "size(A, d): d must be >= 1");
    if (d > 3) return EV{"((ps_int)1)", Ty::Int};
    static const char* s[3] = {"_nx", "_ny", "_nz"};
    // padding=true: the kernel holds the parent; size() is the view's (its box in the parent)
    if (K.opt.padding) return EV{"(" + an(ai) + "_hi" + "xyz"[d - 1] + " - " + an(ai) + "_lo" + "xyz"[d - 1] + " + 1)", Ty::Int};
    return EV{an(ai) + s[d - 1], Ty::Int};
  }
  if (f == "__psize" || f == "__first" || f == "__last") {  // padding=true: parent extent, view box
    int ai = n->a.size() == 2 && n->a[0]->k == NK::Ident ? field_index(n->a[0]->s) : -1;
    if (ai < 0 || !(n->a[1]->k == NK::Num && n->a[1]->is_int) || n->a[1]->ival < 1 || n->a[1]->ival > 3)
      throw PsError(PS_EINTERNAL, f + ": expects an array argument and a dimension 1..3");
    const int d = (int)n->a[1]->ival - 1;
    static const char* s[3] = {"_nx", "_ny", "_nz"};
    if (f == "__psize") return EV{an(ai) + s[d], Ty::Int};
    return EV{an(ai) + (f == "__first" ? "_lo" : "_hi") + "xyz"[d], Ty::Int};
  }
  if (f == "__pgt") throw PsError(PS_EINTERNAL, "padding: a parent-size comparison was not specialised before code generation");
  if (f == "length") {
    int ai = field_index(n->a.at(0)->s);
    if (ai < 0) throw PsError(PS_EARGUMENT, "length(): not an array argument");
    std::string A = an(ai);
    if (K.opt.padding)
      return EV{"((" + A + "_hix - " + A + "_lox + 1)*(" + A + "_hiy - " + A + "_loy + 1)*(" + A + "_hiz - " + A + "_loz + 1))", Ty::Int};
    return EV{"(" + A + "_nx*" + A + "_ny*" + A + "_nz)", Ty::Int};
  }
  if (f == "cellsize") {
    if (n->a.size() != 2 || n->a[0]->k != NK::Ident || !(n->a[1]->k == NK::Num && n->a[1]->is_int))
      throw PsError(PS_EARGUMENT, "cellsize(A, d) requires a CellArray argument and a literal dimension d");
    int ai = field_index(n->a[0]->s);
    if (!is_cell(ai)) throw PsError(PS_EARGUMENT, "cellsize(" + n->a[0]->s + ", d): not a CellArray argument");
    long long d = n->a[1]->ival;
    if (d < 1) throw PsError(PS_EARGUMENT, "cellsize(A, d): d must be >= 1");
    return EV{"((ps_int)" + std::to_string(d <= 3 ? kinds[ai].celldims[d - 1] : 1) + ")", Ty::Int};
  }
  if ((f == "firstindex" || f == "lastindex") && K.opt.padding && n->a.size() == 2) {
    // padding=true: A.indices[d][1] / A.indices[d][end] (src/ParallelKernel/kernel_language.jl:605-617)
    return call(mk(NK::Call, f == "firstindex" ? "__first" : "__last", n->a));
  }
  if (f == "firstindex") return EV{"((ps_int)1)", Ty::Int};
  if (f == "lastindex") {
    NodeP s = mk(NK::Call, "size", n->a);
    return call(s);
  }
  if (f == "max" || f == "min") {
    if (n->a.size() < 2) throw PsError(PS_EARGUMENT, f + " expects at least two arguments");
    EV acc = argv(0);
    for (size_t i = 1; i < n->a.size(); ++i) {
      EV b = argv(i);
      if (acc.t == Ty::Bool) acc = cast(acc, Ty::Int);
      if (b.t == Ty::Bool) b = cast(b, Ty::Int);
      Ty t = promote(acc.t, b.t);
      acc = EV{"ps_" + f + "(" + cast(acc, t).c + ", " + cast(b, t).c + ")", t};
    }
    return acc;
  }
  if (f == "inv") {
    EV a = argv(0);
    if (!is_float(a.t)) a = cast(a, Ty::F64);
    return EV{std::string("(") + (a.t == Ty::F32 ? "1.0f" : "1.0") + " / " + a.c + ")", a.t};
  }
  if (f == "abs") {
    EV a = argv(0);
    if (a.t == Ty::Int) return EV{"((" + a.c + ") < 0 ? -(" + a.c + ") : (" + a.c + "))", Ty::Int};
    return EV{std::string(a.t == Ty::F32 ? "fabsf(" : "fabs(") + a.c + ")", a.t};
  }
  if (f == "Float64" || f == "Float32" || f == "Int" || f == "Int64" || f == "Int32") {
    EV a = argv(0);
    Ty t = f == "Float64" ? Ty::F64 : f == "Float32" ? Ty::F32 : Ty::Int;
    return cast(a, t);
  }
  if (f == "one" || f == "zero") {
    EV a = argv(0);
    double v = f == "one" ? 1.0 : 0.0;
    if (a.t == Ty::Int) return EV{f == "one" ? "((ps_int)1)" : "((ps_int)0)", Ty::Int};
    return lit(v, a.t);
  }
  if (f == "ifelse") {
    EV c = argv(0), a = argv(1), b = argv(2);
    Ty t = promote(a.t, b.t);
    return EV{"(" + c.c + " ? " + cast(a, t).c + " : " + cast(b, t).c + ")", t};
  }
  if (f == "muladd" || f == "fma") {
    EV a = argv(0), b = argv(1), c = argv(2);
    Ty t = promote(promote(a.t, b.t), c.t);
    if (t == Ty::Int) return EV{"(" + a.c + "*" + b.c + " + " + c.c + ")", Ty::Int};
    return EV{std::string(t == Ty::F32 ? "fmaf(" : "fma(") + cast(a, t).c + ", " + cast(b, t).c + ", " + cast(c, t).c + ")", t};
  }
  if (f == "sign") {
    EV a = argv(0);
    if (a.t == Ty::Int) return EV{"((ps_int)((" + a.c + ") > 0) - (ps_int)((" + a.c + ") < 0))", Ty::Int};
    std::string z = a.t == Ty::F32 ? "0.0f" : "0.0", one = a.t == Ty::F32 ? "1.0f" : "1.0";
    return EV{"((" + a.c + ") > " + z + " ? " + one + " : ((" + a.c + ") < " + z + " ? -" + one + " : (" + a.c + ")))", a.t};
  }
  if (f == "copysign" || f == "hypot" || (f == "atan" && n->a.size() == 2)) {
    EV a = argv(0), b = argv(1);
    Ty t = promote(a.t == Ty::Int ? Ty::F64 : a.t, b.t == Ty::Int ? Ty::F64 : b.t);
    std::string fn = f == "atan" ? "atan2" : f;
    return EV{fn + (t == Ty::F32 ? "f(" : "(") + cast(a, t).c + ", " + cast(b, t).c + ")", t};
  }
  static const char* unary_math[] = {"exp", "exp2", "exp10", "expm1", "log", "log2", "log10", "log1p", "sqrt", "cbrt", "sin", "cos", "tan",
                                     "asin", "acos", "atan", "sinh", "cosh", "tanh", "floor", "ceil", "trunc", "erf", "erfc"};
  for (auto um : unary_math) {
    if (f == um) {
      EV a = argv(0);
      if (!is_float(a.t)) a = cast(a, Ty::F64);
      return EV{std::string(um) + (a.t == Ty::F32 ? "f(" : "(") + a.c + ")", a.t};
    }
  }
  if (f == "round") {
    EV a = argv(0);
    if (!is_float(a.t)) return a;
    return EV{std::string(a.t == Ty::F32 ? "rintf(" : "rint(") + a.c + ")", a.t};
  }
  throw PsError(PS_EARGUMENT, "unsupported function '" + f + "' in kernel " + K.def.name);
}

static int uprog_ty(Ty t) { return t == Ty::Int ? 0 : t == Ty::F32 ? 1 : 2; }

// Postfix program of an expression built only from scalar arguments, literals and + - * /
// (and literal powers), typed by the rules expr() applies; false for anything else, and for
// integer arithmetic, which stays on the device. Without contraction and with every quotient
// correctly rounded, the host's IEEE evaluation of the program is the value the device would
// compute for the expression in every cell.
bool Emitter::uniform_prog(const NodeP& n, std::vector<Emitted::UOp>& ops, Ty& t, int& nops) const {
  using U = Emitted::UOp;
  switch (n->k) {
    case NK::Num: {
      U u;
      u.op = U::Lit;
      if (n->is_int) {
        if (!V.idx64 && (n->ival > INT32_MAX || n->ival < INT32_MIN)) return false;
        t = Ty::Int;
        u.ilit = n->ival;
      } else {
        t = n->is_f32 ? Ty::F32 : NT;
        u.lit = n->num;
      }
      u.ty = uprog_ty(t);
      ops.push_back(u);
      return true;
    }
    case NK::Ident: {
      for (int d = 0; d < 3; ++d)
        if (n->s == kIdx[d]) return false;
      if (locals.count(n->s)) return false;
      U u;
      if (n->s == "pi" || n->s == "π") {
        t = NT;
        u.op = U::Lit;
        u.lit = M_PI;
      } else {
        auto it = K.param_index.find(n->s);
        if (it == K.param_index.end()) return false;
        const ArgKind& k = kinds[it->second];
        if (k.is_field || !(k.dtype == PS_F64 || k.dtype == PS_F32 || k.dtype == PS_I32 || k.dtype == PS_I64)) return false;
        t = value_ty(k.dtype);
        u.op = U::Arg;
        u.arg = it->second;
      }
      u.ty = uprog_ty(t);
      ops.push_back(u);
      return true;
    }
    case NK::Unary: {
      if (n->s != "-" && n->s != "+") return false;
      if (!uniform_prog(n->a[0], ops, t, nops)) return false;
      if (n->s == "-") {
        if (t == Ty::Int) return false;
        U u;
        u.op = U::Neg;
        u.ty = uprog_ty(t);
        ops.push_back(u);
      }
      return true;
    }
    case NK::Binary: {
      const std::string& op = n->s;
      U u;
      if (op == "^") {  // Base.literal_pow, for the exponents binary() inlines
        const NodeP& e = n->a[1];
        long long k = 99;
        if (e->k == NK::Num && e->is_int) k = e->ival;
        else if (e->k == NK::Unary && e->s == "-" && e->a[0]->k == NK::Num && e->a[0]->is_int) k = -e->a[0]->ival;
        if (k < -2 || k > 3) return false;
        if (!uniform_prog(n->a[0], ops, t, nops) || t == Ty::Int) return false;
        if (k == 1) return true;
        u.op = k == 0 ? U::Pow0 : k == 2 ? U::Pow2 : k == 3 ? U::Pow3 : k == -1 ? U::PowM1 : U::PowM2;
      } else {
        if (op != "+" && op != "-" && op != "*" && op != "/") return false;
        Ty ta, tb;
        if (!uniform_prog(n->a[0], ops, ta, nops) || !uniform_prog(n->a[1], ops, tb, nops)) return false;
        t = promote(ta, tb);
        if (op == "/" && t == Ty::Int) t = Ty::F64;
        if (t == Ty::Int) return false;
        u.op = op == "+" ? U::Add : op == "-" ? U::Sub : op == "*" ? U::Mul : U::Div;
      }
      u.ty = uprog_ty(t);
      ops.push_back(u);
      ++nops;
      return true;
    }
    default: return false;
  }
}

EV Emitter::expr(const NodeP& n) {
  // a float subexpression of scalar arguments and literals becomes one hidden argument that the
  // host evaluates (not under fastmath or contraction, whose device forms differ from IEEE)
  if ((n->k == NK::Binary || n->k == NK::Unary) && !K.opt.fastmath && !K.opt.fp_contract) {
    std::vector<Emitted::UOp> ops;
    Ty t = Ty::Int;
    int nops = 0;
    if (uniform_prog(n, ops, t, nops) && nops > 0 && is_float(t)) {
      const std::string key = to_julia(n);
      auto it = uni_index.find(key);
      int u;
      if (it == uni_index.end()) {
        u = (int)uprogs.size();
        uni_index[key] = u;
        Emitted::UProg p;
        p.ops = std::move(ops);
        p.ty = uprog_ty(t);
        p.idx64 = V.idx64;
        uprogs.push_back(std::move(p));
      } else {
        u = it->second;
      }
      return EV{"ps_u" + std::to_string(u), t};
    }
  }
  switch (n->k) {
    case NK::Num:
      if (n->is_int) return EV{"((ps_int)" + std::to_string(n->ival) + ")", Ty::Int};
      return lit(n->num, n->is_f32 ? Ty::F32 : NT);
    case NK::Ident: {
      for (int d = 0; d < 3; ++d)
        if (n->s == kIdx[d]) return EV{kIdxC[d], Ty::Int};
      // the launch box (first index and length per dimension), used by generated kernels that
      // ask whether another point of the box exists
      if (n->s.size() == 5 && n->s.compare(0, 3, "__r") == 0 && (n->s[3] == '0' || n->s[3] == 'l') && n->s[4] >= 'x' && n->s[4] <= 'z')
        return EV{std::string("ps_r") + n->s[3] + n->s[4], Ty::Int};
      if (n->s == "true" || n->s == "false") return EV{n->s, Ty::Bool};
      if (n->s == "pi" || n->s == "π") return lit(M_PI, NT);
      auto lt = locals.find(n->s);
      if (lt != locals.end()) return EV{local_c.at(n->s), lt->second};
      auto it = K.param_index.find(n->s);
      if (it != K.param_index.end()) {
        const ArgKind& k = kinds[it->second];
        if (k.is_field) throw PsError(PS_EARGUMENT, "array '" + n->s + "' is used as a scalar (index it or use a finite-difference macro)");
        Ty t = value_ty(k.dtype);
        return EV{an(it->second), t};
      }
      throw PsError(PS_EARGUMENT, "UndefVarError: `" + n->s + "` not defined in kernel " + K.def.name);
    }
    case NK::Interp: throw PsError(PS_EARGUMENT, "unsupported interpolation $" + n->s);
    case NK::EndKw:
      if (!end_subst.empty()) return EV{end_subst, Ty::Int};
      throw PsError(PS_EARGUMENT, "`end` is only valid inside array indexing");
    case NK::Index: return load(n);
    case NK::Call: return call(n);
    case NK::Macro:
      if (K.kernel_language) return kl_expr(n);
      throw PsError(PS_EPLUGIN, "macro @" + n->s + " is not available here");
    case NK::Dot:
      if (K.kernel_language && n->a[0]->k == NK::Macro) return kl_expr(n);
      if (any_cell) {
        int cai;
        NodeP cidx;
        std::string comp, cchk;
        if (cell_ref(n, cai, cidx, comp, cchk)) return cell_load(cai, cidx, comp, cchk);
      }
      throw PsError(PS_EARGUMENT, "unsupported expression: " + to_julia(n));
    case NK::Unary: {
      EV a = expr(n->a[0]);
      if (n->s == "+") return a;
      if (n->s == "!") {
        if (a.t != Ty::Bool) throw PsError(PS_EARGUMENT, "! expects a Bool");
        return EV{"(!" + a.c + ")", Ty::Bool};
      }
      if (a.t == Ty::Bool) a = cast(a, Ty::Int);
      return EV{"(-" + a.c + ")", a.t};
    }
    case NK::Binary: {
      if (n->s == "/") return divide(n);
      EV a = expr(n->a[0]);
      if (n->s == "^") {
        bool lit_exp = (n->a[1]->k == NK::Num && n->a[1]->is_int) ||
                       (n->a[1]->k == NK::Unary && n->a[1]->s == "-" && n->a[1]->a[0]->k == NK::Num && n->a[1]->a[0]->is_int);
        EV b = lit_exp ? EV{"0", Ty::Int} : expr(n->a[1]);
        return binary("^", a, b, n->a[1]);
      }
      EV b = expr(n->a[1]);
      return binary(n->s, a, b, nullptr);
    }
    case NK::Cmp: return cmp(n);
    case NK::And:
    case NK::Or: {
      EV a = expr(n->a[0]), b = expr(n->a[1]);
      if (a.t != Ty::Bool || b.t != Ty::Bool) throw PsError(PS_EARGUMENT, "&& and || expect Bool operands");
      return EV{"(" + a.c + (n->k == NK::And ? " && " : " || ") + b.c + ")", Ty::Bool};
    }
    case NK::Tern: {
      EV c = expr(n->a[0]), a = expr(n->a[1]), b = expr(n->a[2]);
      if (c.t != Ty::Bool) throw PsError(PS_EARGUMENT, "the condition of ?: must be a Bool");
      Ty t = (a.t == Ty::Bool && b.t == Ty::Bool) ? Ty::Bool : promote(a.t == Ty::Bool ? Ty::Int : a.t, b.t == Ty::Bool ? Ty::Int : b.t);
      return EV{"(" + c.c + " ? " + cast(a, t).c + " : " + cast(b, t).c + ")", t};
    }
    case NK::Typed: return expr(n->a[0]);
    default: throw PsError(PS_EARGUMENT, "unsupported expression: " + to_julia(n));
  }
}

void Emitter::assign(const NodeP& s) {
  const NodeP& lhs = s->a[0];
  NodeP rhs = s->a[1];
  if (!s->s.empty()) rhs = mk(NK::Binary, s->s, {clone(lhs), rhs});  // A op= x  ->  A = A op x
  if (any_cell) {
    int cai;
    NodeP cidx;
    std::string comp, cchk;
    if (lhs->k == NK::Index && cell_ref(lhs, cai, cidx, comp, cchk)) {  // one value of a cell
      if (cell_shape(rhs)) throw PsError(PS_EARGUMENT, "a cell is assigned to one value of CellArray '" + K.def.params[cai].name + "'");
      bool blk = false;
      EV r = rhs_value(rhs, &blk);
      cell_store(cai, cidx, comp, cchk, r);
      if (blk) {
        --ind;
        o << pad() << "}\n";
      }
      return;
    }
    if (lhs->k == NK::Index && lhs->a[0]->k == NK::Ident && is_cell(field_index(lhs->a[0]->s))) {  // a whole cell
      const int ai = field_index(lhs->a[0]->s);
      if (!cell_shape(rhs))
        throw PsError(PS_EARGUMENT, "a number is assigned to a cell of CellArray '" + lhs->a[0]->s + "'; assign a cell, or one value with field(" +
                                        lhs->a[0]->s + ", i)[...]");
      comp_nc = kinds[ai].ncomp;
      for (int j = 0; j < comp_nc; ++j) {  // the statement once per value of the cell
        comp_ctx = j;
        bool blk = false;
        EV r = rhs_value(rhs, &blk);
        cell_store(ai, lhs, std::to_string(j), "", r);
        if (blk) {
          --ind;
          o << pad() << "}\n";
        }
      }
      comp_ctx = -1;
      return;
    }
    if (cell_shape(rhs)) throw PsError(PS_EARGUMENT, "a cell is used where a number is expected in " + to_julia(s));
  }
  if (lhs->k == NK::Ident && rhs->k == NK::Macro && rhs->s == "sharedMem") {
    if (!s->s.empty()) throw PsError(PS_EARGUMENT, "@sharedMem: use A = @sharedMem(T, dims)");
    sm_create(lhs->s, rhs);
    return;
  }
  if (lhs->k == NK::Index && lhs->a[0]->k == NK::Ident && K.shared_arrays.count(lhs->a[0]->s)) {
    bool blk = false;
    EV r = rhs_value(rhs, &blk);
    std::string chk;
    std::string addr = sm_address(lhs, &chk);
    int dt = sm_dtype.at(lhs->a[0]->s);
    std::string v;
    if (dt == PS_BF16) v = "ps_bf16_st(" + cast(r, Ty::F32).c + ")";
    else if (dt == PS_F64) v = cast(r, Ty::F64).c;
    else if (dt == PS_F32) v = cast(r, Ty::F32).c;
    else if (dt == PS_BOOL) v = "((unsigned char)(" + r.c + "))";
    else v = std::string("((") + elem_cty(dt) + ")(" + r.c + "))";
    o << pad() << (chk.empty() ? "" : "if (" + chk + ") ") << addr << " = " << v << ";\n";
    if (blk) {
      --ind;
      o << pad() << "}\n";
    }
    return;
  }
  if (lhs->k == NK::Ident) {
    const std::string& v = lhs->s;
    if (K.shared_arrays.count(v)) throw PsError(PS_EARGUMENT, "'" + v + "' is a @sharedMem array; assign to its elements");
    // the bounds proof relies on a loop variable taking exactly the values of its range
    if (loop_vars.count(v)) throw PsError(PS_EARGUMENT, "assignment to the for-loop variable '" + v + "' is not supported");
    if (K.param_index.count(v)) throw PsError(PS_EARGUMENT, "cannot assign to kernel argument '" + v + "' (arguments are passed by value)");
    for (int d = 0; d < 3; ++d)
      if (v == kIdx[d]) throw PsError(PS_EARGUMENT, "cannot assign to a parallel index");
    bool blk = false;
    EV r = rhs_value(rhs, &blk);
    if (!locals.count(v)) {
      if (!pass1) throw PsError(PS_EINTERNAL, "local variable not hoisted: " + v);
      locals[v] = r.t;
      local_c[v] = "v_" + sanitize(v);
      local_order.push_back(v);
    }
    o << pad() << local_c[v] << " = " << cast(r, locals[v]).c << ";\n";
    if (blk) {
      --ind;
      o << pad() << "}\n";
    }
    return;
  }
  if (lhs->k != NK::Index || lhs->a[0]->k != NK::Ident) throw PsError(PS_EARGUMENT, "unsupported assignment target " + to_julia(lhs));
  int ai = field_index(lhs->a[0]->s);
  if (ai < 0) throw PsError(PS_EARGUMENT, "'" + lhs->a[0]->s + "' is assigned to but is not an array argument");
  bool blk = false;
  EV r = rhs_value(rhs, &blk);
  auto idx = norm_indices(lhs, ai);
  std::string ic[3], ref32;
  std::string off = offset(ai, idx, ic, nullptr, &ref32);
  int dt = kinds[ai].dtype;
  const std::string v = store_value(dt, r);
  std::string A = an(ai);
  const std::string ref = ref32.empty() ? A + "[" + off + "]" : ref32;
  if (V.checked) {
    o << pad() << "if (ps_chk(" << ic[0] << ", " << ic[1] << ", " << ic[2] << ", " << A << "_nx, " << A << "_ny, " << A
      << "_nz, ps_err)) " << ref << " = " << v << ";\n";
  } else {
    if (V.nt) o << pad() << "__builtin_nontemporal_store(" << v << ", &" << ref << ");\n";
    else o << pad() << ref << " = " << v << ";\n";
  }
  if (blk) {
    --ind;
    o << pad() << "}\n";
  }
}

// A[i...] += v as one atomic read-modify-write of the element (the shadow accumulations of
// derivative kernels, where several threads add into one element).
void Emitter::atomic_add(const NodeP& s) {
  const NodeP& lhs = s->a.at(0);
  if (lhs->k != NK::Index || lhs->a[0]->k != NK::Ident) throw PsError(PS_EINTERNAL, "@ps_atomic_add expects an array element");
  int ai = field_index(lhs->a[0]->s);
  if (ai < 0 || is_cell(ai)) throw PsError(PS_EINTERNAL, "@ps_atomic_add: '" + lhs->a[0]->s + "' is not a plain array argument");
  int dt = kinds[ai].dtype;
  if (dt != PS_F64 && dt != PS_F32)
    throw PsError(PS_EARGUMENT, "∇: shadow array '" + K.def.params[ai].name + "' must hold Float64 or Float32 values");
  bool blk = false;
  EV r = rhs_value(s->a.at(1), &blk);
  auto idx = norm_indices(lhs, ai);
  std::string ic[3];
  std::string off = offset(ai, idx, ic, nullptr);
  std::string A = an(ai);
  std::string v = cast(r, dt == PS_F64 ? Ty::F64 : Ty::F32).c;
  std::string op = "unsafeAtomicAdd(&" + A + "[" + off + "], " + v + ");\n";
  if (V.checked)
    o << pad() << "if (ps_chk(" << ic[0] << ", " << ic[1] << ", " << ic[2] << ", " << A << "_nx, " << A << "_ny, " << A << "_nz, ps_err)) " << op;
  else
    o << pad() << op;
  if (blk) {
    --ind;
    o << pad() << "}\n";
  }
}

// The stored form of a value for an array of element type dt.
std::string Emitter::store_value(int dt, const EV& r) const {
  if (dt == PS_BF16) return "ps_bf16_st(" + cast(r, Ty::F32).c + ")";
  if (dt == PS_F64) return cast(r, Ty::F64).c;
  if (dt == PS_F32) return cast(r, Ty::F32).c;
  if (dt == PS_BOOL) return "((unsigned char)(" + r.c + "))";
  return std::string("((") + elem_cty(dt) + ")(" + r.c + "))";
}

// Cell q of a group (x offset q % cx, z offset q / cx): its x and z indices and its offsets
// within the group, shadowing the group's.
std::string Emitter::group_scope(int q) const {
  const int c = q % V.cx, d = q / V.cx;
  std::string r = "const ps_int ix = ps_ix0 + " + std::to_string(c) + "; const ps_int ps_cx = " + std::to_string(c) + "; (void)ix; (void)ps_cx; ";
  if (V.cz > 1) r += "const ps_int iz = ps_iz0 + " + std::to_string(d) + "; const ps_int ps_cz = " + std::to_string(d) + "; (void)iz; (void)ps_cz; ";
  return r;
}

// The constant c of an x index of the form ix + c or ix - c (k unchanged otherwise).
static void x_index_offset(const NodeP& e, long long& k) {
  if (e->k == NK::Ident && e->s == kIdx[0]) {
    k = 0;
    return;
  }
  if (e->k == NK::Binary && (e->s == "+" || e->s == "-") && e->a.size() == 2) {
    const NodeP &l = e->a[0], &r = e->a[1];
    if (l->k == NK::Ident && l->s == kIdx[0] && r->k == NK::Num && r->is_int) k = e->s == "+" ? r->ival : -r->ival;
    else if (e->s == "+" && r->k == NK::Ident && r->s == kIdx[0] && l->k == NK::Num && l->is_int) k = l->ival;
  }
}

// Bounds on the x index that a mask (a conjunction of comparisons, idx_interval) puts: for each
// comparison of ix with an integer expression, the expression (adjusted for strictness) as a lower
// or upper bound. False if a bound is not an integer.
bool Emitter::x_bounds(const NodeP& c, std::vector<std::string>& lo, std::vector<std::string>& hi) {
  bool has_ix = false;
  std::function<bool(const NodeP&)> uses = [&](const NodeP& n) {
    if (!n) return false;
    if (n->k == NK::Ident && n->s == kIdx[0]) return true;
    for (auto& a : n->a)
      if (uses(a)) return true;
    return false;
  };
  has_ix = uses(c);
  if (!has_ix) return true;
  if (c->k == NK::And) return x_bounds(c->a[0], lo, hi) && x_bounds(c->a[1], lo, hi);
  if (c->k != NK::Cmp) return false;
  for (size_t i = 0; i + 1 < c->a.size(); ++i) {
    const NodeP &l = c->a[i], &r = c->a[i + 1];
    const bool lx = l->k == NK::Ident && l->s == kIdx[0], rx = r->k == NK::Ident && r->s == kIdx[0];
    if (!lx && !rx) continue;
    const std::string& op = c->ops[i];
    if (op != "<" && op != "<=" && op != ">" && op != ">=") return false;
    EV b = expr(lx ? r : l);
    if (b.t != Ty::Int) return false;
    const std::string v = "(ps_int)(" + b.c + ")";
    // ix op b, or b op ix
    const bool less = op == "<" || op == "<=", strict = op == "<" || op == ">";
    if (lx == less) hi.push_back(strict ? v + " - 1" : v);
    else lo.push_back(strict ? v + " + 1" : v);
  }
  return true;
}

// Array names that an expression reads (indexed argument names).
static void read_arrays(const NodeP& n, std::set<std::string>& out) {
  if (!n) return;
  if (n->k == NK::Index && n->a[0]->k == NK::Ident) out.insert(n->a[0]->s);
  for (auto& c : n->a) read_arrays(c, out);
}

// Array assignments for every cell of an x group whose masks all hold, a batch at a time: a batch
// is a run of statements none of which reads or rewrites an array that an earlier statement of
// the batch writes. For a batch: the right-hand sides of all its statements and cells in one
// basic block (the group's loads are adjacent and merge into wide accesses, and the statements'
// loads overlap), the IEEE recomputation of a statement's cells when a quotient of any of them
// left the exact range (IEEE division gives the cells that were in range the same values), then
// the stores. As in a wavefront whose lanes run one cell each, every load of a statement
// precedes its stores, and a statement that reads what an earlier one writes sees it.
void Emitter::group_statements(const std::vector<NodeP>& sts) {
  struct G {
    std::string g, ref, decl, badv, bad;
    EV f, e;
    bool uf;
    bool nan_in_bad;  // the flag of the IEEE recomputation is set by any NaN value of the group
    int dt;
  };
  const int CX = V.cx * V.cz;  // cells of the group
  size_t i = 0;
  while (i < sts.size()) {
    std::set<std::string> written;
    size_t j = i;
    for (; j < sts.size(); ++j) {
      const NodeP& s = sts[j];
      std::set<std::string> rd;
      read_arrays(s->a[1], rd);
      if (!s->s.empty()) rd.insert(s->a[0]->a[0]->s);
      for (size_t q = 1; q < s->a[0]->a.size(); ++q) read_arrays(s->a[0]->a[q], rd);
      bool dep = written.count(s->a[0]->a[0]->s) > 0;
      for (auto& r : rd) dep = dep || written.count(r) > 0;
      if (j > i && dep) break;
      written.insert(s->a[0]->a[0]->s);
    }
    std::vector<G> batch;
    for (size_t k = i; k < j; ++k) {
      const NodeP& s = sts[k];
      const NodeP& lhs = s->a[0];
      NodeP rhs = s->a[1];
      if (!s->s.empty()) rhs = mk(NK::Binary, s->s, {clone(lhs), rhs});
      const int ai = field_index(lhs->a[0]->s);
      if (ai < 0) throw PsError(PS_EARGUMENT, "'" + lhs->a[0]->s + "' is assigned to but is not an array argument");
      G x;
      fast_div = true;
      used_fast = false;
      div_x = nan_transparent(rhs);
      used_x[0] = used_x[1] = false;
      x.f = expr(rhs);
      fast_div = false;
      x.uf = used_fast;
      x.e = x.uf ? expr(rhs) : x.f;
      x.decl = fast_decls();
      x.badv = "ps_v";  // the cell's value, named in its scope
      x.bad = fast_bad(x.badv);
      x.nan_in_bad = x.uf && div_x;
      div_x = false;
      auto idx = norm_indices(lhs, ai);
      std::string ic[3], ref32;
      const std::string off = offset(ai, idx, ic, nullptr, &ref32);
      x.ref = ref32.empty() ? an(ai) + "[" + off + "]" : ref32;
      x.g = std::to_string(++tmp_counter);
      x.dt = kinds[ai].dtype;
      batch.push_back(x);
    }
    for (auto& x : batch) {
      for (int c = 0; c < CX; ++c) o << pad() << cty(x.f.t) << " ps_g" << x.g << "_" << c << ";\n";
      if (x.uf) o << pad() << "bool ps_g" << x.g << "_bad = false;\n";
    }
    for (auto& x : batch)
      for (int c = 0; c < CX; ++c) {
        const std::string vc = "ps_g" + x.g + "_" + std::to_string(c);
        if (x.uf)
          o << pad() << "{ " << group_scope(c) << x.decl << cty(x.f.t) << " ps_v = " << x.f.c << "; " << vc << " = ps_v; ps_g" << x.g
            << "_bad |= " << x.bad << "; }\n";
        else
          o << pad() << "{ " << group_scope(c) << vc << " = " << x.f.c << "; }\n";
      }
    // BFloat16 stores round in one instruction; a group's NaNs become the canonical NaN first, in
    // a branch that only groups holding a NaN take (for statements whose recomputation flag is
    // set by NaN values, inside the recomputation)
    auto canon = [&](const G& x, const std::string& in) {
      for (int c = 0; c < CX; ++c) {
        const std::string vc = "ps_g" + x.g + "_" + std::to_string(c);
        o << in << vc << " = ps_nan_canon(" << vc << ");\n";
      }
    };
    for (auto& x : batch) {
      const bool bf = x.dt == PS_BF16 && x.f.t == Ty::F32;
      if (x.uf) {
        o << pad() << "if (ps_g" << x.g << "_bad) {\n";
        o << pad() << "  __asm__ volatile(\"\");\n";
        for (int c = 0; c < CX; ++c)
          o << pad() << "  { " << group_scope(c) << "ps_g" << x.g << "_" << c << " = " << cast(x.e, x.f.t).c << "; }\n";
        if (bf && x.nan_in_bad) canon(x, pad() + "  ");
        o << pad() << "}\n";
      }
      if (bf && !x.nan_in_bad) {
        std::string any;
        for (int c = 0; c < CX; ++c) {
          const std::string vc = "ps_g" + x.g + "_" + std::to_string(c);
          any += std::string(c ? " | " : "") + "(" + vc + " != " + vc + ")";
        }
        o << pad() << "if (" << any << ") {\n";
        canon(x, pad() + "  ");
        o << pad() << "}\n";
      }
    }
    for (auto& x : batch)
      for (int c = 0; c < CX; ++c) {
        const bool bf = x.dt == PS_BF16 && x.f.t == Ty::F32;
        const std::string vn = "ps_g" + x.g + "_" + std::to_string(c);
        const std::string v = bf ? "ps_bf16_cvt(" + vn + ")" : store_value(x.dt, EV{vn, x.f.t});
        o << pad() << "{ " << group_scope(c);
        if (V.nt) o << "__builtin_nontemporal_store(" << v << ", &" << x.ref << ");";
        else o << x.ref << " = " << v << ";";
        o << " }\n";
      }
    i = j;
  }
}

// Whether a NaN anywhere in the expression reaches its value: arithmetic, literal powers other
// than 0, NaN-propagating max/min and elementary functions of non-Bool operands; not comparisons,
// conditionals, Bool operands (false * NaN is 0), conversions or other calls.
bool Emitter::nan_transparent(const NodeP& n) const {
  static const std::set<std::string> fns = {"max", "min", "abs", "sqrt", "exp", "log", "log10", "log2", "exp2", "expm1", "log1p",
                                            "sin", "cos", "tan", "sinh", "cosh", "tanh", "asin", "acos", "atan"};
  switch (n->k) {
    case NK::Num: return true;
    case NK::Ident: {
      if (n->s == "true" || n->s == "false") return false;
      auto lt = locals.find(n->s);
      if (lt != locals.end()) return lt->second != Ty::Bool;
      auto it = K.param_index.find(n->s);
      if (it != K.param_index.end()) return kinds[it->second].is_field || kinds[it->second].dtype != PS_BOOL;
      return true;
    }
    case NK::Index: {  // a load of a non-Bool array
      if (n->a[0]->k != NK::Ident) return false;
      const int ai = field_index(n->a[0]->s);
      return ai >= 0 && kinds[ai].dtype != PS_BOOL;
    }
    case NK::Unary: return (n->s == "-" || n->s == "+") && nan_transparent(n->a[0]);
    case NK::Binary:
      if (n->s == "^") {
        const NodeP& e = n->a[1];
        const bool lit = (e->k == NK::Num && e->is_int && e->ival != 0) ||
                         (e->k == NK::Unary && e->s == "-" && e->a[0]->k == NK::Num && e->a[0]->is_int && e->a[0]->ival != 0);
        return lit && nan_transparent(n->a[0]);
      }
      if (n->s != "+" && n->s != "-" && n->s != "*" && n->s != "/") return false;
      return nan_transparent(n->a[0]) && nan_transparent(n->a[1]);
    case NK::Call:
      if (!fns.count(n->s)) return false;
      for (auto& c : n->a)
        if (!nan_transparent(c)) return false;
      return true;
    default: return false;
  }
}

// Declarations and the fallback test of the fast form of a statement (ps_bad, or the exponent
// ranges and the NaN test of ps_divx).
std::string Emitter::fast_decls() const {
  std::string d;
  if (div_x) {
    if (used_x[0]) d += "int ps_lo = 0, ps_hi = 0; ";
    if (used_x[1]) d += "int ps_lo32 = 0, ps_hi32 = 0; ";
  } else {
    d = "bool ps_bad = false; ";
  }
  return d;
}
std::string Emitter::fast_bad(const std::string& v) const {
  if (!div_x) return "ps_bad";
  std::string c = "(" + v + " != " + v + ")";
  if (used_x[0]) c += " | (ps_lo < -599) | (ps_hi > 600)";
  if (used_x[1]) c += " | (ps_lo32 < -79) | (ps_hi32 > 80)";
  return "(" + c + ")";
}

// Value of a right-hand side. When it divides by uniform divisors, the fast (reciprocal) form
// is evaluated first and the IEEE form only if a quotient may have left the exact range; the
// emitted code then opens a block that the caller closes after using the value.
EV Emitter::rhs_value(const NodeP& rhs, bool* opened) {
  fast_div = true;
  used_fast = false;
  div_x = nan_transparent(rhs);
  used_x[0] = used_x[1] = false;
  EV f = expr(rhs);
  fast_div = false;
  *opened = false;
  if (!used_fast) {
    div_x = false;
    return f;
  }
  EV e = expr(rhs);
  std::string var = "ps_v" + std::to_string(++tmp_counter);
  o << pad() << "{\n";
  ++ind;
  const std::string decl = fast_decls();
  o << pad() << decl << "\n";
  o << pad() << cty(f.t) << " " << var << " = " << f.c << ";\n";
  // the empty volatile asm keeps the rarely taken IEEE form behind its branch; without it the
  // compiler may evaluate that form in every cell and select
  o << pad() << "if (" << fast_bad(var) << ") { __asm__ volatile(\"\"); " << var << " = " << cast(e, f.t).c << "; }\n";
  div_x = false;
  *opened = true;
  return EV{var, f.t};
}

void Emitter::stmt(const NodeP& s) {
  switch (s->k) {
    case NK::Assign: assign(s); return;
    case NK::If: {
      EV c = expr(s->a[0]);
      if (c.t != Ty::Bool) throw PsError(PS_EARGUMENT, "if-condition must be a Bool");
      o << pad() << "if (" << c.c << ") {\n";
      ++ind;
      block(s->a[1]);
      --ind;
      if (s->a.size() > 2 && s->a[2]) {
        o << pad() << "} else {\n";
        ++ind;
        if (s->a[2]->k == NK::If) stmt(s->a[2]);
        else block(s->a[2]);
        --ind;
      }
      o << pad() << "}\n";
      return;
    }
    case NK::Block: block(s); return;
    case NK::For: {
      EV a = expr(s->a[0]->a[0]), b = expr(s->a[0]->a[1]);
      if (a.t != Ty::Int || b.t != Ty::Int) throw PsError(PS_EARGUMENT, "for-loop bounds must be integers");
      std::string v = s->s;
      std::string cv = "f_" + sanitize(v);
      bool had = locals.count(v);
      Ty prev = had ? locals[v] : Ty::Int;
      std::string prevc = had ? local_c[v] : "";
      locals[v] = Ty::Int;
      local_c[v] = cv;
      o << pad() << "for (ps_int " << cv << " = " << a.c << "; " << cv << " <= " << b.c << "; ++" << cv << ") {\n";
      ++ind;
      const bool outer = loop_vars.count(v) != 0;
      loop_vars.insert(v);
      block(s->a[1]);
      if (!outer) loop_vars.erase(v);
      --ind;
      o << pad() << "}\n";
      if (had) {
        locals[v] = prev;
        local_c[v] = prevc;
      } else {
        locals.erase(v);
        local_c.erase(v);
      }
      return;
    }
    case NK::Return:
      if (in_loop) throw PsError(PS_EARGUMENT, "early return is not supported in memopt kernels");
      o << pad() << "return;\n";
      return;
    default: {
      if (s->k == NK::Macro && s->s == "ps_atomic_add") {
        atomic_add(s);
        return;
      }
      if (K.kernel_language && s->k == NK::Macro && (s->s == "sync_threads" || s->s == "pk_println" || s->s == "pk_show")) {
        kl_stmt(s);
        return;
      }
      EV e = expr(s);
      o << pad() << "(void)(" << e.c << ");\n";
    }
  }
}

void Emitter::block(const NodeP& b) {
  for (auto& s : b->a) stmt(s);
}

// ---------------------------------------------------------------------------
// CellArrays (src/ParallelKernel/allocators.jl:1-17,436-454 and the CellArray type of
// CellArrays.jl): a cell assignment is emitted once per value of the cell; field(A, i...)[I],
// A[I][i...] and A[I].name select one value.
// ---------------------------------------------------------------------------

// Offset of value j of the cell with linear index I, in the layout of the argument.
std::string Emitter::cell_offset(int ai, const std::string& I, const std::string& j) const {
  const std::string A = an(ai), nc = std::to_string(kinds[ai].ncomp);
  switch (kinds[ai].cell_layout) {
    case 1: return "(" + I + ") + (" + j + ") * " + A + "_B";
    case 2: return "(" + I + ") * " + nc + " + (" + j + ")";
    default:
      return "((" + I + ") / " + A + "_B) * (" + A + "_B * " + nc + ") + (" + j + ") * " + A + "_B + (" + I + ") % " + A + "_B";
  }
}

// Recognises the forms that select one value of a cell: field(A, i...)[I], A[I][i...] and
// A[I].name. Returns the argument, the cell access A[I], the 0-based value index and, for
// value indices not known at compile time, the check of that index (checked variants).
bool Emitter::cell_ref(const NodeP& n, int& ai, NodeP& cellidx, std::string& comp, std::string& compchk) {
  std::vector<NodeP> cs;
  if (n->k == NK::Index && n->a[0]->k == NK::Call && n->a[0]->s == "field") {
    const NodeP& f = n->a[0];
    if (f->a.size() < 2 || f->a[0]->k != NK::Ident) throw PsError(PS_EARGUMENT, "field(A, i...) expects a CellArray argument and value indices");
    ai = field_index(f->a[0]->s);
    if (!is_cell(ai)) throw PsError(PS_EARGUMENT, "field(" + f->a[0]->s + ", ...): not a CellArray argument");
    std::vector<NodeP> a{f->a[0]};
    a.insert(a.end(), n->a.begin() + 1, n->a.end());
    cellidx = mk(NK::Index, "", a);
    cs.assign(f->a.begin() + 1, f->a.end());
  } else if (n->k == NK::Index && n->a[0]->k == NK::Index && n->a[0]->a[0]->k == NK::Ident && is_cell(field_index(n->a[0]->a[0]->s))) {
    ai = field_index(n->a[0]->a[0]->s);
    cellidx = n->a[0];
    cs.assign(n->a.begin() + 1, n->a.end());
  } else if (n->k == NK::Dot && n->a[0]->k == NK::Index && n->a[0]->a[0]->k == NK::Ident && is_cell(field_index(n->a[0]->a[0]->s))) {
    ai = field_index(n->a[0]->a[0]->s);
    cellidx = n->a[0];
    const std::string& names = kinds[ai].cellnames;
    int j = 0;
    size_t p = 0;
    for (;;) {
      size_t q = names.find(',', p);
      if (names.substr(p, q == std::string::npos ? std::string::npos : q - p) == n->s) break;
      if (q == std::string::npos) {
        throw PsError(PS_EARGUMENT, "CellArray '" + n->a[0]->a[0]->s + "' has no cell field named " + n->s +
                                        (names.empty() ? " (its cells have no field names)" : " (fields: " + names + ")"));
      }
      p = q + 1;
      ++j;
    }
    if (j >= kinds[ai].ncomp) throw PsError(PS_EARGUMENT, "cell field " + n->s + " lies outside the cell");
    comp = std::to_string(j);
    compchk.clear();
    return true;
  } else {
    return false;
  }
  const int* cd = kinds[ai].celldims;
  if (cs.empty() || cs.size() > 4) throw PsError(PS_EARGUMENT, "a cell value is selected with 1 to 4 indices");
  // extents per given index; the last index runs over the remaining cell dimensions (Julia's
  // trailing linear indexing), so one index runs over the whole cell
  long long ext[4] = {1, 1, 1, 1};
  for (size_t d = 0; d < cs.size(); ++d) ext[d] = cd[d];
  for (size_t d = cs.size(); d < 4; ++d) ext[cs.size() - 1] *= cd[d];
  std::string ic[4] = {"((ps_int)1)", "((ps_int)1)", "((ps_int)1)", "((ps_int)1)"};
  bool literal = true;
  long long lin = 0, stride = 1;
  comp.clear();
  for (size_t d = 0; d < cs.size(); ++d) {
    EV e = expr(cs[d]);
    if (e.t != Ty::Int) throw PsError(PS_EARGUMENT, "cell value indices must be integers");
    ic[d] = e.c;
    if (cs[d]->k == NK::Num && cs[d]->is_int) {
      const long long v = cs[d]->ival;
      if (v < 1 || v > ext[d])
        throw PsError(PS_EBOUNDS, "BoundsError: value index " + std::to_string(v) + " of a cell of CellArray '" + K.def.params[ai].name +
                                      "' is outside 1:" + std::to_string(ext[d]));
      lin += (v - 1) * stride;
    } else {
      literal = false;
    }
    comp += std::string(comp.empty() ? "" : " + ") + "((" + e.c + ") - 1) * " + std::to_string(stride);
    stride *= ext[d];
  }
  if (literal) {
    comp = std::to_string(lin);
    compchk.clear();
  } else {
    comp = "(" + comp + ")";
    compchk = V.checked ? "(ps_chk(" + ic[0] + ", " + ic[1] + ", " + ic[2] + ", " + std::to_string(ext[0]) + ", " + std::to_string(ext[1]) + ", " +
                              std::to_string(ext[2]) + ", ps_err) && ps_chk(" + ic[3] + ", 1, 1, " + std::to_string(ext[3]) + ", 1, 1, ps_err))"
                        : "";
  }
  return true;
}

EV Emitter::cell_load(int ai, const NodeP& cellidx, const std::string& comp, const std::string& compchk) {
  auto idx = norm_indices(cellidx, ai);
  std::string ic[3];
  bool aff = false;
  const std::string I = offset(ai, idx, ic, &aff);
  const std::string A = an(ai);
  std::string raw = A + "[" + cell_offset(ai, I, comp) + "]";
  const int dt = kinds[ai].dtype;
  if (V.checked) {
    std::string chk = "ps_chk(" + ic[0] + ", " + ic[1] + ", " + ic[2] + ", " + A + "_nx, " + A + "_ny, " + A + "_nz, ps_err)";
    if (!compchk.empty()) chk = "(" + chk + " && " + compchk + ")";
    raw = "(" + chk + " ? " + raw + " : (" + elem_cty(dt) + ")0)";
  }
  if (dt == PS_BF16) return EV{"ps_bf16_ld(" + raw + ")", Ty::F32};
  if (dt == PS_BOOL) return EV{"((bool)" + raw + ")", Ty::Bool};
  if (dt == PS_I64 || dt == PS_I32) return EV{"((ps_int)" + raw + ")", Ty::Int};
  return EV{raw, value_ty(dt)};
}

void Emitter::cell_store(int ai, const NodeP& cellidx, const std::string& comp, const std::string& compchk, const EV& r) {
  auto idx = norm_indices(cellidx, ai);
  std::string ic[3];
  const std::string I = offset(ai, idx, ic, nullptr);
  const std::string A = an(ai);
  const int dt = kinds[ai].dtype;
  std::string v;
  if (dt == PS_BF16) v = "ps_bf16_st(" + cast(r, Ty::F32).c + ")";
  else if (dt == PS_F64) v = cast(r, Ty::F64).c;
  else if (dt == PS_F32) v = cast(r, Ty::F32).c;
  else if (dt == PS_BOOL) v = "((unsigned char)(" + r.c + "))";
  else v = std::string("((") + elem_cty(dt) + ")(" + r.c + "))";
  std::string chk;
  if (V.checked) {
    chk = "ps_chk(" + ic[0] + ", " + ic[1] + ", " + ic[2] + ", " + A + "_nx, " + A + "_ny, " + A + "_nz, ps_err)";
    if (!compchk.empty()) chk += " && " + compchk;
  }
  o << pad() << (chk.empty() ? "" : "if (" + chk + ") ") << A << "[" << cell_offset(ai, I, comp) << "] = " << v << ";\n";
}

// 1 if the expression is a whole cell, 0 if it is a number. Cells combine as the static arrays
// of CellArrays.jl do: cell +/- cell, number * cell, cell * number, cell / number and -cell;
// anything else (cell + number, cell * cell, functions of cells, cells in conditions) is an
// error, as the corresponding Julia method does not exist or is not elementwise.
int Emitter::cell_shape(const NodeP& n) {
  switch (n->k) {
    case NK::Index: {
      int ai;
      NodeP ci;
      std::string c, cc;
      if ((n->a[0]->k == NK::Call && n->a[0]->s == "field") ||
          (n->a[0]->k == NK::Index && n->a[0]->a[0]->k == NK::Ident && is_cell(field_index(n->a[0]->a[0]->s))))
        return 0;
      (void)ai;
      (void)ci;
      (void)c;
      (void)cc;
      return (n->a[0]->k == NK::Ident && is_cell(field_index(n->a[0]->s))) ? 1 : 0;
    }
    case NK::Binary: {
      const int a = cell_shape(n->a[0]), b = cell_shape(n->a[1]);
      const std::string& op = n->s;
      if (op == "+" || op == "-") {
        if (a != b) throw PsError(PS_EARGUMENT, "MethodError: a cell and a number cannot be combined with " + op + " in " + to_julia(n));
        return a;
      }
      if (op == "*") {
        if (a && b) throw PsError(PS_EARGUMENT, "products of two cells are not supported (" + to_julia(n) + ")");
        return a | b;
      }
      if (op == "/") {
        if (b) throw PsError(PS_EARGUMENT, "division by a cell is not supported (" + to_julia(n) + ")");
        return a;
      }
      if (a || b) throw PsError(PS_EARGUMENT, "operator " + op + " is not defined for cells (" + to_julia(n) + ")");
      return 0;
    }
    case NK::Unary: {
      const int a = cell_shape(n->a[0]);
      if (a && n->s == "!") throw PsError(PS_EARGUMENT, "! is not defined for cells");
      return a;
    }
    case NK::Dot:
    case NK::Num:
    case NK::Ident:
    case NK::Str:
    case NK::EndKw:
      return 0;
    default:
      for (auto& c : n->a)
        if (c && cell_shape(c)) throw PsError(PS_EARGUMENT, "a cell is used where a number is expected in " + to_julia(n));
      return 0;
  }
}

// ---------------------------------------------------------------------------
// Kernel language (src/ParallelKernel/kernel_language.jl:9-193, AMDGPU mappings :300-580)
// ---------------------------------------------------------------------------
static int dtype_bytes(int dt) {
  switch (dt) {
    case PS_F64:
    case PS_I64: return 8;
    case PS_F32:
    case PS_I32: return 4;
    case PS_BF16: return 2;
    default: return 1;
  }
}

// Element type of @sharedMem(T, dims): eltype(A) of an array argument or a type name.
int Emitter::type_dtype(const NodeP& t) const {
  if (t->k == NK::Call && t->s == "eltype" && t->a.size() == 1 && t->a[0]->k == NK::Ident) {
    int ai = field_index(t->a[0]->s);
    if (ai < 0) throw PsError(PS_EARGUMENT, "eltype(" + t->a[0]->s + "): not an array argument of the kernel");
    return kinds[ai].dtype;
  }
  std::string s;
  if (t->k == NK::Ident) s = t->s;
  else if (t->k == NK::Dot && t->a[0]->k == NK::Ident && t->a[0]->s == "Data") s = "Data." + t->s;
  if (s == "Float64") return PS_F64;
  if (s == "Float32") return PS_F32;
  if (s == "BFloat16") return PS_BF16;
  if (s == "Int64" || s == "Int") return PS_I64;
  if (s == "Int32") return PS_I32;
  if (s == "Bool") return PS_BOOL;
  if (s == "Data.Number") return NT == Ty::F32 ? PS_F32 : PS_F64;
  if (s == "Data.Index") return PS_I64;
  throw PsError(PS_EARGUMENT, "@sharedMem: unsupported element type " + to_julia(t));
}

// A = @sharedMem(T, dims[, offset]): a column-major view of the launch's dynamic LDS starting
// `offset` bytes in. The offset is honoured (CUDA.jl's @cuDynamicSharedMem semantics), so several
// arrays can share the dynamic LDS without aliasing.
void Emitter::sm_create(const std::string& v, const NodeP& mac) {
  int dt = type_dtype(mac->a[0]);
  auto it = sm_dtype.find(v);
  if (it != sm_dtype.end() && it->second != dt) throw PsError(PS_EARGUMENT, "@sharedMem: '" + v + "' is created with different element types");
  sm_dtype[v] = dt;
  const NodeP& dn = mac->a[1];
  std::vector<NodeP> dims = dn->k == NK::Tuple ? dn->a : std::vector<NodeP>{dn};
  std::string d[3] = {"(ps_int)1", "(ps_int)1", "(ps_int)1"};
  for (size_t i = 0; i < dims.size() && i < 3; ++i) {
    EV e = expr(dims[i]);
    if (e.t != Ty::Int) throw PsError(PS_EARGUMENT, "@sharedMem: dimensions must be integers");
    d[i] = e.c;
  }
  std::string off = "(ps_int)0";
  if (mac->a.size() == 3) {
    EV e = expr(mac->a[2]);
    if (e.t != Ty::Int) throw PsError(PS_EARGUMENT, "@sharedMem: the offset must be an integer (bytes)");
    off = e.c;
  }
  const std::string P = "sm_" + sanitize(v);
  o << pad() << P << "_n1 = " << d[0] << "; " << P << "_n2 = " << d[1] << "; " << P << "_n3 = " << d[2] << ";\n";
  o << pad() << P << " = (" << elem_cty(dt) << "*)(ps_dyn_lds + (" << off << "));\n";
  if (V.checked) {
    const int es = dtype_bytes(dt);
    // an array that does not fit the launch's dynamic LDS gets no elements: every access fails
    // the bounds check instead of reaching past the allocation
    o << pad() << "if ((" << off << ") < 0 || (" << off << ") % " << es << " != 0 || " << P << "_n1 < 0 || " << P << "_n2 < 0 || " << P
      << "_n3 < 0 || (" << off << ") + " << P << "_n1 * " << P << "_n2 * " << P << "_n3 * " << es << " > ps_shmem) { *ps_err = 1; " << P
      << "_n1 = 0; }\n";
  }
}

// Address of an element of a @sharedMem array (1-based, column-major; one index = linear).
std::string Emitter::sm_address(const NodeP& n, std::string* check) {
  const std::string& v = n->a[0]->s;
  if (!sm_dtype.count(v)) throw PsError(PS_EARGUMENT, "'" + v + "' is used before its @sharedMem");
  const int nd = K.shared_arrays.at(v);
  const std::string P = "sm_" + sanitize(v);
  const std::string N[3] = {P + "_n1", P + "_n2", P + "_n3"};
  const size_t ni = n->a.size() - 1;
  if (ni != 1 && (int)ni != nd)
    throw PsError(PS_EARGUMENT, "'" + v + "' has " + std::to_string(nd) + " dimensions; index it with 1 or " + std::to_string(nd) + " indices");
  const std::string saved = end_subst;
  std::string ic[3] = {"(ps_int)1", "(ps_int)1", "(ps_int)1"};
  auto index = [&](const NodeP& e, const std::string& extent) {
    end_subst = extent;
    EV r = expr(e);
    end_subst = saved;
    if (r.t != Ty::Int) throw PsError(PS_EARGUMENT, "indices of '" + v + "' must be integers");
    return r.c;
  };
  check->clear();
  if (ni == 1 && nd > 1) {
    const std::string len = "(" + N[0] + " * " + N[1] + " * " + N[2] + ")";
    ic[0] = index(n->a[1], len);
    if (V.checked) *check = "ps_chk(" + ic[0] + ", 1, 1, " + len + ", 1, 1, ps_err)";
    return P + "[(" + ic[0] + ") - 1]";
  }
  for (int d = 0; d < nd; ++d) ic[d] = index(n->a[1 + d], N[d]);
  std::string off = "((" + ic[0] + ") - 1)";
  if (nd > 1) off += " + ((" + ic[1] + ") - 1) * " + N[0];
  if (nd > 2) off += " + ((" + ic[2] + ") - 1) * " + N[0] + " * " + N[1];
  if (V.checked) *check = "ps_chk(" + ic[0] + ", " + ic[1] + ", " + ic[2] + ", " + N[0] + ", " + N[1] + ", " + N[2] + ", ps_err)";
  return P + "[" + off + "]";
}

// Value-producing kernel-language macros. Thread and block indices are 1-based, as AMDGPU.jl's
// workitemIdx / workgroupIdx; wavefront operations work on gfx950's 64 lanes.
EV Emitter::kl_expr(const NodeP& n) {
  if (n->k == NK::Dot) {
    const std::string& m = n->a[0]->s;
    const std::string& c = n->s;
    if (!n->a[0]->a.empty()) throw PsError(PS_EARGUMENT, "@" + m + ": no arguments allowed.");
    if (c != "x" && c != "y" && c != "z") throw PsError(PS_EARGUMENT, "@" + m + "()." + c + ": only the fields x, y and z exist");
    if (m == "threadIdx") return EV{"((ps_int)threadIdx." + c + " + 1)", Ty::Int};
    if (m == "blockIdx") return EV{"((ps_int)blockIdx." + c + " + 1)", Ty::Int};
    if (m == "blockDim") return EV{"((ps_int)blockDim." + c + ")", Ty::Int};
    if (m == "gridDim") return EV{"((ps_int)gridDim." + c + ")", Ty::Int};
    throw PsError(PS_EARGUMENT, "unsupported expression: " + to_julia(n));
  }
  const std::string& m = n->s;
  if (m == "threadIdx" || m == "blockIdx" || m == "blockDim" || m == "gridDim")
    throw PsError(PS_EARGUMENT, "@" + m + "() returns a triple; use @" + m + "().x, .y or .z");
  if (m == "warpsize") return EV{"((ps_int)64)", Ty::Int};
  if (m == "laneid") return EV{"((ps_int)__lane_id() + 1)", Ty::Int};
  if (m == "active_mask") return EV{"((ps_int)__ballot(1))", Ty::Int};
  auto mask = [&]() {
    EV k = expr(n->a[0]);
    if (k.t != Ty::Int) throw PsError(PS_EARGUMENT, "@" + m + ": the mask must be an integer");
    return "((unsigned long long)(" + k.c + "))";
  };
  auto intarg = [&](size_t i, const char* what) {
    EV e = expr(n->a[i]);
    if (e.t != Ty::Int) throw PsError(PS_EARGUMENT, "@" + m + ": " + what + " must be an integer");
    if (n->a[i]->k == NK::Num && std::string(what) == "width") {
      long long w = n->a[i]->ival;
      if (w < 1 || w > 64 || (w & (w - 1))) throw PsError(PS_EARGUMENT, "@" + m + ": width must be a power of two between 1 and 64");
    }
    return "((int)(" + e.c + "))";
  };
  if (m == "shfl_sync" || m == "shfl_up_sync" || m == "shfl_down_sync" || m == "shfl_xor_sync") {
    (void)mask();  // every lane of the wavefront takes part; the mask names no subset here
    EV v = expr(n->a[1]);
    const std::string w = n->a.size() == 4 ? intarg(3, "width") : "64";
    std::string fn, sel;
    if (m == "shfl_sync") {
      fn = "ps_shfl";
      sel = intarg(2, "lane");  // 1-based source lane within the partition
    } else if (m == "shfl_up_sync") {
      fn = "ps_shfl_up";
      sel = "(unsigned int)" + intarg(2, "delta");
    } else if (m == "shfl_down_sync") {
      fn = "ps_shfl_down";
      sel = "(unsigned int)" + intarg(2, "delta");
    } else {
      fn = "ps_shfl_xor";
      sel = intarg(2, "lane_mask");  // partner lane = (laneid - 1) xor lane_mask
    }
    const std::string tail = ", " + sel + ", " + w + ")";
    switch (v.t) {
      case Ty::Bool: return EV{"((bool)" + fn + "((int)(" + v.c + ")" + tail + ")", Ty::Bool};
      case Ty::Int: return EV{"((ps_int)" + fn + "((ps_int)(" + v.c + ")" + tail + ")", Ty::Int};
      case Ty::F32: return EV{fn + "((float)(" + v.c + ")" + tail, Ty::F32};
      default: return EV{fn + "((double)(" + v.c + ")" + tail, Ty::F64};
    }
  }
  if (m == "vote_any_sync" || m == "vote_all_sync" || m == "vote_ballot_sync") {
    const std::string k = mask();
    EV p = expr(n->a[1]);
    if (p.t != Ty::Bool) throw PsError(PS_EARGUMENT, "@" + m + ": the predicate must be a Bool");
    const std::string b = "__ballot((int)(" + p.c + "))";
    if (m == "vote_any_sync") return EV{"((" + b + " & " + k + ") != 0ull)", Ty::Bool};
    if (m == "vote_all_sync") return EV{"((" + b + " & " + k + ") == (" + k + " & __ballot(1)))", Ty::Bool};
    return EV{"((ps_int)(" + b + " & " + k + "))", Ty::Int};
  }
  throw PsError(PS_EARGUMENT, "@" + m + " is a statement, not a value");
}

// Statement macros: @sync_threads() is a workgroup barrier; @pk_println(args...) prints its
// arguments concatenated (Base.println); @pk_show(args...) prints `expr = value` per argument.
// Floats print with 16 (Float64) or 8 (Float32) significant digits.
void Emitter::kl_stmt(const NodeP& s) {
  if (s->s == "sync_threads") {
    if (!s->a.empty()) throw PsError(PS_EARGUMENT, "@sync_threads: no arguments allowed.");
    o << pad() << "__syncthreads();\n";
    return;
  }
  auto esc = [](const std::string& t) {
    std::string r;
    for (char c : t) {
      if (c == '%') r += "%%";
      else if (c == '"') r += "\\\"";
      else r += c;
    }
    return r;
  };
  auto add = [&](const NodeP& a, std::string& f, std::vector<std::string>& vals) {
    if (a->k == NK::Str) {
      f += esc(a->s);
      return;
    }
    EV e = expr(a);
    switch (e.t) {
      case Ty::Int:
        f += "%lld";
        vals.push_back("(long long)(" + e.c + ")");
        break;
      case Ty::F64:
        f += "%.16g";
        vals.push_back("(double)(" + e.c + ")");
        break;
      case Ty::F32:
        f += "%.8g";
        vals.push_back("(double)(" + e.c + ")");
        break;
      case Ty::Bool:
        f += "%s";
        vals.push_back("((" + e.c + ") ? \"true\" : \"false\")");
        break;
    }
  };
  auto emit = [&](const std::string& f, const std::vector<std::string>& vals) {
    o << pad() << "printf(\"" << f << "\\n\"";
    for (auto& v : vals) o << ", " << v;
    o << ");\n";
  };
  if (s->s == "pk_println") {
    std::string f;
    std::vector<std::string> vals;
    for (auto& a : s->a) add(a, f, vals);
    emit(f, vals);
    return;
  }
  for (auto& a : s->a) {
    std::string text = to_julia(a);
    for (int d = 0; d < 3 && d < (int)K.opt.indices.size(); ++d) {  // the kernel's own index names
      const std::string from = kIdx[d];
      for (size_t p = text.find(from); p != std::string::npos; p = text.find(from, p + K.opt.indices[d].size()))
        text.replace(p, from.size(), K.opt.indices[d]);
    }
    std::string f = esc(text) + " = ";
    std::vector<std::string> vals;
    add(a, f, vals);
    emit(f, vals);
  }
}

// memopt: thread per column, stream along dimension L (0-based) in chunks of loopsize.
// The loop is software-pipelined: every value a step takes from a register column or from a
// halo of an LDS plane tile is loaded during the previous step, so each load has a whole step
// to arrive and the barrier that publishes a tile only waits for LDS writes.
void Emitter::emit_memopt_body(int L) {
  const MemoptMeta& mm = K.mm;
  const int PD = std::max(1, K.opt.prefetch);  // prefetch distance in steps
  // points read per read-only array: (ox, oy) -> planes oz (oy = 0 when streaming dimension 2)
  struct Pts {
    int ai;
    std::string name;
    bool optvar;
    std::map<std::pair<int, int>, std::set<int>> cols;
  };
  std::vector<Pts> pts;
  auto collect = [&](const std::string& A, const std::map<std::pair<int, int>, std::map<int, int>>& offs, bool optvar) {
    Pts P;
    P.ai = K.param_index.at(A);
    P.name = A;
    P.optvar = optvar;
    bool explicit_r = optvar && mm.optranges.count(A) > 0;
    for (auto& c : offs)
      for (auto& z : c.second) {
        if (explicit_r) {
          const auto& r = mm.optranges.at(A);
          int xy[2] = {c.first.first, c.first.second};
          bool in = true;
          for (int d = 0; d < L; ++d)
            if (xy[d] < r[d].first || xy[d] > r[d].second) in = false;
          if (z.first < r[L].first || z.first > r[L].second) in = false;
          if (!in) continue;
        }
        P.cols[c.first].insert(z.first);
      }
    pts.push_back(P);
  };
  for (auto& A : mm.optvars) collect(A, mm.offsets.at(A), true);
  // read-only arrays outside optvars are only prefetched into registers; a bounds-checked
  // variant keeps their plain (checked) loads
  if (!V.checked)
    for (auto& kv : mm.ro_offsets) collect(kv.first, kv.second, false);

  // LDS tiles: planes of optvars read at several (ox, oy) columns are staged in LDS once per
  // step; each thread contributes its register value ("own" column), the halo runs are loaded
  // by the first lanes of the block.
  struct Run {
    bool horiz;
    int x0, y0, len;
  };
  struct Tile {
    int ai, oz, xlo, xhi, ylo, yhi, W, H;
    std::pair<int, int> own;
    int own_col = -1;
    std::vector<Run> runs;
    int nhalo = 0, nslots = 0;
    std::string name;
    std::set<std::pair<int, int>> cols_at;
  };
  std::vector<Tile> tiles;
  const int BX = V.bx, BY = V.by;
  if (V.lds && L == 2) {
    for (auto& P : pts) {
      if (!P.optvar || std::find(mm.shmem_optvars.begin(), mm.shmem_optvars.end(), P.name) == mm.shmem_optvars.end()) continue;
      auto orig = P.cols;
      std::set<int> planes;
      for (auto& c : orig)
        for (int z : c.second) planes.insert(z);
      for (int oz : planes) {
        Tile t;
        t.ai = P.ai;
        t.oz = oz;
        for (auto& c : orig)
          if (c.second.count(oz)) t.cols_at.insert(c.first);
        if (t.cols_at.size() < 2) continue;
        t.xlo = t.ylo = 1 << 30;
        t.xhi = t.yhi = -(1 << 30);
        size_t best = 0;
        for (auto& p : t.cols_at) {
          t.xlo = std::min(t.xlo, p.first);
          t.xhi = std::max(t.xhi, p.first);
          t.ylo = std::min(t.ylo, p.second);
          t.yhi = std::max(t.yhi, p.second);
          if (orig.at(p).size() > best) {
            best = orig.at(p).size();
            t.own = p;
          }
        }
        for (auto& p : t.cols_at)
          if (p != t.own) P.cols[p].erase(oz);  // served by the tile
        t.W = BX + t.xhi - t.xlo;
        t.H = BY + t.yhi - t.ylo;
        std::vector<char> need((size_t)t.W * t.H, 0);
        for (auto& p : t.cols_at)
          for (int ty = 0; ty < BY; ++ty)
            for (int tx = 0; tx < BX; ++tx) need[(size_t)(ty + p.second - t.ylo) * t.W + (tx + p.first - t.xlo)] = 1;
        for (int ty = 0; ty < BY; ++ty)
          for (int tx = 0; tx < BX; ++tx) need[(size_t)(ty + t.own.second - t.ylo) * t.W + (tx + t.own.first - t.xlo)] = 0;
        // horizontal runs of length >= 2, then vertical runs of what is left
        for (int ly = 0; ly < t.H; ++ly)
          for (int lx = 0; lx < t.W;) {
            int e = lx;
            while (e < t.W && need[(size_t)ly * t.W + e] == 1) ++e;
            if (e - lx >= 2) {
              t.runs.push_back({true, lx, ly, e - lx});
              for (int q = lx; q < e; ++q) need[(size_t)ly * t.W + q] = 2;
            }
            lx = e > lx ? e : lx + 1;
          }
        for (int lx = 0; lx < t.W; ++lx)
          for (int ly = 0; ly < t.H;) {
            int e = ly;
            while (e < t.H && need[(size_t)e * t.W + lx] == 1) ++e;
            if (e > ly) {
              t.runs.push_back({false, lx, ly, e - ly});
              for (int q = ly; q < e; ++q) need[(size_t)q * t.W + lx] = 2;
            }
            ly = e > ly ? e : ly + 1;
          }
        for (auto& r : t.runs) t.nhalo += r.len;
        t.nslots = (t.nhalo + BX * BY - 1) / (BX * BY);
        t.name = "ps_t" + std::to_string(P.ai) + "_" + std::to_string(tiles.size());
        tiles.push_back(t);
      }
    }
  }
  // register columns: a window of planes [zlo, zhi] per (array, ox, oy) plus the next plane
  struct Col {
    int ai;
    int o[3];  // o[L] unused
    int zlo, zhi;
    std::string name;
  };
  std::vector<Col> cols;
  for (auto& P : pts)
    for (auto& c : P.cols) {
      if (c.second.empty()) continue;
      Col col;
      col.ai = P.ai;
      col.o[0] = c.first.first;
      col.o[1] = L == 2 ? c.first.second : 0;
      col.o[2] = 0;
      col.zlo = *c.second.begin();
      col.zhi = *c.second.rbegin();
      col.name = "q" + std::to_string(P.ai) + "_" + std::to_string(cols.size());
      cols.push_back(col);
    }
  for (auto& t : tiles)
    for (size_t ci = 0; ci < cols.size(); ++ci)
      if (cols[ci].ai == t.ai && cols[ci].o[0] == t.own.first && cols[ci].o[1] == t.own.second && cols[ci].zlo <= t.oz &&
          t.oz <= cols[ci].zhi)
        t.own_col = (int)ci;
  for (auto& t : tiles)
    if (t.own_col < 0) throw PsError(PS_EINTERNAL, "memopt: tile without register column");

  std::string lc = kIdxC[L];
  static const char* szn[3] = {"_nx", "_ny", "_nz"};
  static const char* str[3] = {"1", "_sx", "_sxy"};
  // per column: base offset (without the streamed dimension) and validity
  for (auto& c : cols) {
    std::string A = an(c.ai);
    std::ostringstream base, ok;
    base << "(ps_int)0";
    ok << "true";
    for (int d = 0; d < 3; ++d) {
      if (d == L) continue;
      std::string idx = d < K.nidx ? std::string("(") + kIdxC[d] + " + (" + std::to_string(c.o[d]) + "))" : std::string("((ps_int)1)");
      base << " + (" << idx << " - 1)" << (d == 0 ? "" : "*" + A + str[d]);
      ok << " && " << idx << " >= 1 && " << idx << " <= " << A << szn[d];
    }
    o << pad() << "const ps_int " << c.name << "_b = " << base.str() << ";\n";
    o << pad() << "const bool " << c.name << "_ok = " << ok.str() << ";\n";
    int w = c.zhi - c.zlo + 1;
    Ty vt = value_ty(kinds[c.ai].dtype);
    for (int k = 0; k < w; ++k) o << pad() << cty(vt) << " " << c.name << "_" << k << ";\n";
    for (int d = 0; d < PD; ++d) o << pad() << cty(vt) << " " << c.name << "_n" << d << ";\n";
  }
  // guarded load of column c at plane z (an expression); `extra` is an additional condition
  auto ldq = [&](const Col& c, const std::string& z, const std::string& extra) {
    std::string A = an(c.ai);
    std::string idx = c.name + "_b + (" + z + " - 1)*" + A + str[L];
    std::string cond = extra + c.name + "_ok && (" + z + ") >= 1 && (" + z + ") <= " + A + szn[L];
    int dt = kinds[c.ai].dtype;
    std::string v = std::string(V.checked ? "ps_ldz(" : "ps_ldc(") + A + ", " + cond + ", " + idx + ")";
    if (dt == PS_BF16) v = "ps_bf16_ld(" + v + ")";
    return v;
  };
  for (auto& t : tiles) {
    Ty vt = value_ty(kinds[t.ai].dtype);
    E_lds_bytes += 2 * t.W * t.H * (vt == Ty::F64 ? 8 : 4);
    o << pad() << "__shared__ " << cty(vt) << " " << t.name << "[2][" << t.W * t.H << "];\n";
  }
  if (!tiles.empty()) {
    o << pad() << "const int ps_lx = threadIdx.x, ps_ly = threadIdx.y, ps_tid = ps_lx + ps_ly * " << BX << ";\n";
    o << pad() << "const ps_int ps_ix0 = ix - ps_lx, ps_iy0 = iy - ps_ly;\n";
  }
  // halo slots: slot j of a thread covers halo cell ps_tid + j*BX*BY of its tile; the position
  // in the tile and the x-y part of the global offset are fixed for the whole stream
  for (auto& t : tiles) {
    std::string A = an(t.ai);
    Ty vt = value_ty(kinds[t.ai].dtype);
    for (int j = 0; j < t.nslots; ++j) {
      std::string S = t.name + "_s" + std::to_string(j);
      o << pad() << "int " << S << "_x = 0, " << S << "_y = 0;\n";
      o << pad() << "const int " << S << "_c = ps_tid + " << j * BX * BY << ";\n";
      o << pad() << "const bool " << S << "_v = " << S << "_c < " << t.nhalo << ";\n";
      int base = 0;
      for (size_t r = 0; r < t.runs.size(); ++r) {
        const Run& run = t.runs[r];
        o << pad() << (r ? "else " : "") << "if (" << S << "_c < " << base + run.len << ") { ";
        if (run.horiz) o << S << "_x = " << run.x0 << " + (" << S << "_c - " << base << "); " << S << "_y = " << run.y0 << "; }\n";
        else o << S << "_x = " << run.x0 << "; " << S << "_y = " << run.y0 << " + (" << S << "_c - " << base << "); }\n";
        base += run.len;
      }
      o << pad() << "const ps_int " << S << "_gx = ps_ix0 + " << S << "_x + (" << t.xlo << "), " << S << "_gy = ps_iy0 + " << S << "_y + ("
        << t.ylo << ");\n";
      o << pad() << "const bool " << S << "_ok = " << S << "_v && " << S << "_gx >= 1 && " << S << "_gx <= " << A << "_nx && " << S
        << "_gy >= 1 && " << S << "_gy <= " << A << "_ny;\n";
      o << pad() << "const ps_int " << S << "_b = (" << S << "_gx - 1) + (" << S << "_gy - 1)*" << A << "_sx;\n";
      o << pad() << "const int " << S << "_l = " << S << "_y * " << t.W << " + " << S << "_x;\n";
      for (int d = 0; d < PD; ++d) o << pad() << cty(vt) << " " << S << "_h" << d << ";\n";
    }
  }
  auto ldh = [&](const Tile& t, int j, const std::string& z, const std::string& extra) {
    std::string A = an(t.ai);
    std::string S = t.name + "_s" + std::to_string(j);
    std::string cond = extra + S + "_ok && (" + z + ") >= 1 && (" + z + ") <= " + A + "_nz";
    std::string idx = S + "_b + (" + z + " - 1)*" + A + "_sxy";
    int dt = kinds[t.ai].dtype;
    std::string v = std::string(V.checked ? "ps_ldz(" : "ps_ldc(") + A + ", " + cond + ", " + idx + ")";
    if (dt == PS_BF16) v = "ps_bf16_ld(" + v + ")";
    return v;
  };
  // prologue: the window below the first step, and the values of the first PD steps
  auto within = [&](int d) { return d ? "(ps_zb + " + std::to_string(d) + " < ps_ze) && " : std::string(); };
  for (auto& c : cols) {
    int w = c.zhi - c.zlo + 1;
    for (int k = 0; k + 1 < w; ++k)
      o << pad() << c.name << "_" << k << " = " << ldq(c, "ps_zb + (" + std::to_string(c.zlo + k) + ")", "") << ";\n";
    for (int d = 0; d < PD; ++d)
      o << pad() << c.name << "_n" << d << " = " << ldq(c, "ps_zb + (" + std::to_string(c.zhi + d) + ")", within(d)) << ";\n";
  }
  for (auto& t : tiles)
    for (int j = 0; j < t.nslots; ++j)
      for (int d = 0; d < PD; ++d)
        o << pad() << t.name << "_s" << j << "_h" << d << " = " << ldh(t, j, "ps_zb + (" + std::to_string(t.oz + d) + ")", within(d)) << ";\n";
  // Unchecked variants split the stream in two loops: in the first, every load issued for a later
  // step (plane lc + offset + PD) lies inside its array, so those loads need neither the end-of-
  // chunk nor the z-range guard, and their x-y clamp is hoisted out of the loop; the second loop
  // (the last PD steps of the chunk, and planes near the array's end) keeps the guarded loads.
  const bool split = !V.checked && (!cols.empty() || !tiles.empty());
  if (split) {
    // the first loop ends where the furthest load of any array would leave it, and is empty when
    // the chunk starts so low that a load would fall below the array
    int minoff = 1 << 30;
    std::string zs = "ps_ze - " + std::to_string(PD);
    for (auto& c : cols) {
      zs = "ps_min(" + zs + ", " + an(c.ai) + szn[L] + " - (" + std::to_string(c.zhi + PD) + ") + 1)";
      minoff = std::min(minoff, c.zhi + PD);
    }
    for (auto& t : tiles) {
      zs = "ps_min(" + zs + ", " + an(t.ai) + "_nz - (" + std::to_string(t.oz + PD) + ") + 1)";
      minoff = std::min(minoff, t.oz + PD);
    }
    o << pad() << "const ps_int ps_zs = (ps_zb >= (ps_int)(" << (1 - minoff) << ")) ? ps_max(ps_zb, " << zs << ") : ps_zb;\n";
    for (auto& c : cols) o << pad() << "const ps_int " << c.name << "_bs = " << c.name << "_ok ? " << c.name << "_b : (ps_int)0;\n";
    for (auto& t : tiles)
      for (int j = 0; j < t.nslots; ++j) {
        std::string S = t.name + "_s" + std::to_string(j);
        o << pad() << "const ps_int " << S << "_bs = " << S << "_ok ? " << S << "_b : (ps_int)0;\n";
      }
  }
  auto ldq_fast = [&](const Col& c, const std::string& z) {
    std::string A = an(c.ai);
    std::string v = A + "[" + c.name + "_bs + (" + z + " - 1)*" + A + str[L] + "]";
    if (kinds[c.ai].dtype == PS_BF16) v = "ps_bf16_ld(" + v + ")";
    return v;
  };
  auto ldh_fast = [&](const Tile& t, int j, const std::string& z) {
    std::string A = an(t.ai);
    std::string S = t.name + "_s" + std::to_string(j);
    std::string v = A + "[" + S + "_bs + (" + z + " - 1)*" + A + "_sxy]";
    if (kinds[t.ai].dtype == PS_BF16) v = "ps_bf16_ld(" + v + ")";
    return v;
  };
  auto emit_loop = [&](const std::string& to, bool fast) {
    o << pad() << "for (; " << lc << " < " << to << "; ++" << lc << ") {\n";
    ++ind;
    // the loads issued in this step serve step lc + PD
    if (!fast) o << pad() << "const bool ps_more = " << lc << " + " << PD << " < ps_ze;\n";
    for (auto& c : cols) {
      int w = c.zhi - c.zlo + 1;
      o << pad() << c.name << "_" << (w - 1) << " = " << c.name << "_n0;\n";
      for (int d = 0; d + 1 < PD; ++d) o << pad() << c.name << "_n" << d << " = " << c.name << "_n" << d + 1 << ";\n";
      const std::string z = lc + " + (" + std::to_string(c.zhi + PD) + ")";
      o << pad() << c.name << "_n" << PD - 1 << " = " << (fast ? ldq_fast(c, z) : ldq(c, z, "ps_more && ")) << ";\n";
    }
    if (!tiles.empty()) {
      o << pad() << "const int ps_buf = (int)((" << lc << " - ps_zb) & 1);\n";
      for (auto& t : tiles) {
        const Col& c = cols[t.own_col];
        o << pad() << t.name << "[ps_buf][(ps_ly + " << (c.o[1] - t.ylo) << ") * " << t.W << " + ps_lx + " << (c.o[0] - t.xlo)
          << "] = " << c.name << "_" << (t.oz - c.zlo) << ";\n";
        for (int j = 0; j < t.nslots; ++j) {
          std::string S = t.name + "_s" + std::to_string(j);
          o << pad() << "if (" << S << "_v) " << t.name << "[ps_buf][" << S << "_l] = " << S << "_h0;\n";
          for (int d = 0; d + 1 < PD; ++d) o << pad() << S << "_h" << d << " = " << S << "_h" << d + 1 << ";\n";
          const std::string z = lc + " + (" + std::to_string(t.oz + PD) + ")";
          // lanes without a halo slot skip the load (whole waves of them branch around it)
          o << pad() << (fast ? "if (" + S + "_v) " : std::string()) << S << "_h" << PD - 1 << " = "
            << (fast ? ldh_fast(t, j, z) : ldh(t, j, z, "ps_more && ")) << ";\n";
        }
      }
      o << pad() << "__syncthreads();\n";
    }
    // center offsets of every field at this step
    for (size_t i = 0; i < kinds.size(); ++i) {
      if (!kinds[i].is_field) continue;
      std::string A = an((int)i);
      o << pad() << "const ps_int " << A << "_o = (ix - 1) + (iy - 1)*" << A << "_sx + (iz - 1)*" << A << "_sxy;\n";
    }
    cache.clear();
    for (auto& c : cols) {
      int w = c.zhi - c.zlo + 1;
      for (int k = 0; k < w; ++k) {
        int off[3] = {c.o[0], c.o[1], c.o[2]};
        off[L] = c.zlo + k;
        cache[std::make_tuple(c.ai, off[0], off[1], off[2])] = c.name + "_" + std::to_string(k);
      }
    }
    for (auto& t : tiles)
      for (auto& p : t.cols_at) {
        auto key = std::make_tuple(t.ai, p.first, p.second, t.oz);
        if (cache.count(key)) continue;  // a register of the thread's own column
        cache[key] = t.name + "[ps_buf][(ps_ly + " + std::to_string(p.second - t.ylo) + ") * " + std::to_string(t.W) + " + ps_lx + " +
                     std::to_string(p.first - t.xlo) + "]";
      }
    in_loop = true;
    if (V.lds) {
      o << pad() << "if (ps_act) {\n";
      ++ind;
    }
    block(K.body);
    if (V.lds) {
      --ind;
      o << pad() << "}\n";
    }
    in_loop = false;
    for (auto& c : cols) {
      int w = c.zhi - c.zlo + 1;
      for (int k = 0; k + 1 < w; ++k) o << pad() << c.name << "_" << k << " = " << c.name << "_" << (k + 1) << ";\n";
    }
    --ind;
    o << pad() << "}\n";
  };
  o << pad() << "ps_int " << lc << " = ps_zb;\n";
  if (split) emit_loop("ps_zs", true);
  emit_loop("ps_ze", false);
}

Emitted Emitter::run() {
  Emitted E;
  // variants carry their execution path in the symbol so that profiles tell them apart; kernels
  // whose array arguments are all FP32 or all BFloat16 also carry their storage type
  std::string entry = "psk_" + sanitize(K.def.name) + (V.memopt ? (V.lds ? "_memopt_lds" : "_memopt") : "") + (V.checked ? "_checked" : "") +
                      (K.opt.fastmath ? "_fastmath" : "");
  {
    int fdt = 0;
    for (auto& k : kinds)
      if (k.is_field) fdt = (fdt == 0 || fdt == k.dtype) ? k.dtype : -1;
    if (fdt == PS_F32) entry += "_f32";
    if (fdt == PS_BF16) entry += "_bf16";
  }
  E.entry = entry;
  const char* I = V.idx64 ? "long long" : "int";
  int ib = V.idx64 ? 8 : 4;
  // pass 1: discover local variables, their types and the uniform divisors
  {
    pass1 = true;
    std::ostringstream saved;
    saved.swap(o);
    in_loop = false;
    block(K.body);
    o.swap(saved);
    pass1 = false;
    o.str("");
  }
  std::ostringstream sig;
  for (size_t i = 0; i < kinds.size(); ++i) {
    const ArgKind& k = kinds[i];
    std::string A = an((int)i);
    if (k.is_field) {
      bool ro = !K.written.count(K.def.params[i].name);
      sig << "    " << (ro ? "const " : "") << elem_cty(k.dtype) << "* " << (V.noalias ? "__restrict__ " : "") << A << ",  // "
          << K.def.params[i].name << "\n";
      sig << "    ps_int " << A << "_nx, ps_int " << A << "_ny, ps_int " << A << "_nz, ps_int " << A << "_sx, ps_int " << A << "_sxy,\n";
      E.params.push_back({0, (int)i, 0, 8, 0});
      for (int w = 0; w < 5; ++w) E.params.push_back({1, (int)i, w, ib, 0});
      if (K.opt.padding) {  // the argument's box in its parent: firstindex / lastindex per dimension
        sig << "    ps_int " << A << "_lox, ps_int " << A << "_loy, ps_int " << A << "_loz, ps_int " << A << "_hix, ps_int " << A
            << "_hiy, ps_int " << A << "_hiz,\n";
        for (int w = 0; w < 6; ++w) E.params.push_back({10, (int)i, w, ib, 0});
      }
      if (k.ncomp > 0) {  // CellArray: values of one cell field stored contiguously
        sig << "    ps_int " << A << "_B,\n";
        E.params.push_back({9, (int)i, 0, ib, 0});
      }
    } else {
      Ty t = value_ty(k.dtype);
      std::string ct = t == Ty::Bool ? "int" : cty(t);
      sig << "    " << ct << " " << A << ",  // " << K.def.params[i].name << "\n";
      int bytes = t == Ty::F64 ? 8 : t == Ty::F32 ? 4 : t == Ty::Bool ? 4 : ib;
      E.params.push_back({2, (int)i, 0, bytes, k.dtype});
    }
  }
  sig << "    ps_int ps_r0x, ps_int ps_r0y, ps_int ps_r0z, ps_int ps_rlx, ps_int ps_rly, ps_int ps_rlz,\n";
  for (int d = 0; d < 3; ++d) E.params.push_back({3, -1, d, ib, 0});
  for (int d = 0; d < 3; ++d) E.params.push_back({4, -1, d, ib, 0});
  sig << "    int* ps_err";
  E.params.push_back({5, -1, 0, 8, 0});
  if (!K.shared_arrays.empty()) {  // bytes of dynamic LDS the launch provides (@sharedMem)
    sig << ",\n    ps_int ps_shmem";
    E.params.push_back({8, -1, 0, ib, 0});
  }
  for (auto& rp : rcp_params) {  // reciprocals of uniform divisors (exact mode: 0 = use IEEE division)
    sig << ",\n    " << cty(rp.second) << " " << an(rp.first) << "_r";
    E.params.push_back({K.opt.fastmath ? 7 : 6, rp.first, 0, rp.second == Ty::F64 ? 8 : 4, rp.second == Ty::F64 ? PS_F64 : PS_F32});
  }
  for (size_t u = 0; u < uprogs.size(); ++u) {  // host-evaluated scalar subexpressions
    const bool f64 = uprogs[u].ty == 2;
    sig << ",\n    " << (f64 ? "double" : "float") << " ps_u" << u;
    E.params.push_back({11, (int)u, 0, f64 ? 8 : 4, f64 ? PS_F64 : PS_F32});
  }
  E.uprogs = uprogs;

  std::ostringstream src;
  src << "// generated by libpsamd for kernel " << K.def.name << " (variant " << V.key() << ")\n";
  src << "typedef " << I << " ps_int;\n" << kPrelude << "\n";
  if (!K.shared_arrays.empty()) src << "extern __shared__ __attribute__((aligned(16))) unsigned char ps_dyn_lds[];\n";
  src << "extern \"C\" __global__ void __launch_bounds__(" << V.maxthreads << ") " << entry << "(\n" << sig.str() << ") {\n";
  if (!V.memopt) {
    const int CX = V.cx, CZ = V.cz;
    const bool GR = CX > 1 || CZ > 1;  // groups of CX x CZ cells per thread
    if (GR && !groupable_body(K, CZ > 1)) throw PsError(PS_EINTERNAL, "groups of cells requested for a kernel body that cannot run as groups");
    // groups: thread (tx, tz) owns the CX x CZ cells from (ps_ix0, ps_iz0); the cell offsets below
    // are the group's first cell's (each cell adds its x offset ps_cx as a constant element offset
    // and its z offset ps_cz to the uniform part of an access)
    const char* X = GR ? "ps_ix0" : "ix";
    o << "  const ps_int ps_bx = blockIdx.x, ps_by = blockIdx.y, ps_bz = blockIdx.z;\n";
    o << "  const ps_int tx = ps_bx * (ps_int)blockDim.x + (ps_int)threadIdx.x;\n";
    o << "  const ps_int ty = ps_by * (ps_int)blockDim.y + (ps_int)threadIdx.y;\n";
    o << "  const ps_int tz = ps_bz * (ps_int)blockDim.z + (ps_int)threadIdx.z;\n";
    // x groups with the row split: the x interval in which every statement's mask holds (from
    // the masks' bounds on ix) holds the row's interior groups, which the first threads of the
    // row take in order; the cells outside it (the row ends) go one at a time to the threads of
    // the last block column, so that interior wavefronts never run the per-cell path
    std::vector<std::string> xlo, xhi;
    bool split = GR;
    if (GR)
      for (auto& st : K.body->a)
        if (st->k == NK::If && !x_bounds(st->a[0], xlo, xhi)) split = false;
    if (split) {
      // line phase: groups start where the first statement's left-hand side access (x offset
      // k) falls on a 128-byte line of the narrowest arrays (A cells), so that a block row of
      // groups covers whole lines of the rows (256-byte aligned); threads before the first
      // interior group take none, and the interior cells before it are edge cells
      long long k = 0;
      for (auto& st : K.body->a) {
        const NodeP* a = nullptr;
        if (st->k == NK::Assign) a = &st;
        else if (st->k == NK::If && !st->a[1]->a.empty()) a = &st->a[1]->a[0];
        if (!a) continue;
        const NodeP& lhs = (*a)->a[0];
        if (lhs->a.size() > 1) x_index_offset(lhs->a[1], k);
        break;
      }
      int es = 8;
      for (auto& kd : kinds)
        if (kd.is_field) es = std::min(es, dtype_bytes(kd.dtype));
      const int A = 128 / es;
      o << "  if (ty >= ps_rly || tz * " << CZ << " >= ps_rlz) return;\n";
      o << "  const ps_int iy = ps_r0y + ty;\n  const ps_int ps_iz0 = ps_r0z + tz * " << CZ << ";\n  const ps_int iz = ps_iz0;\n  (void)iy; (void)iz;\n";
      o << "  ps_int ps_xl = ps_r0x, ps_xh = ps_r0x + ps_rlx - 1;\n";
      for (auto& b : xlo) o << "  ps_xl = max(ps_xl, " << b << ");\n";
      for (auto& b : xhi) o << "  ps_xh = min(ps_xh, " << b << ");\n";
      o << "  const ps_int ps_e0 = ps_xl - 1 + (" << k << ");  // element of the first interior cell's access\n";
      o << "  const ps_int ps_g0 = ps_e0 - ((ps_e0 % " << A << ") + " << A << ") % " << A << " + 1 - (" << k << ");  // thread 0's first cell\n";
      o << "  const ps_int ps_s = ps_xl > ps_g0 ? (ps_xl - ps_g0 + " << (CX - 1) << ") / " << CX << " : 0;  // threads without a group\n";
      o << "  const ps_int ps_gs = ps_g0 + ps_s * " << CX << ";  // first cell of the interior groups\n";
      o << "  const ps_int ps_ng = ps_xh >= ps_gs ? (ps_xh - ps_gs + 1) / " << CX << " : 0;  // interior groups\n";
      o << "  const ps_int ps_ngp = (ps_s + ps_ng + (ps_int)blockDim.x - 1) / (ps_int)blockDim.x * (ps_int)blockDim.x;\n";
    } else if (GR) {
      o << "  if (tx * " << CX << " >= ps_rlx || ty >= ps_rly || tz * " << CZ << " >= ps_rlz) return;\n";
      o << "  const ps_int ps_ix0 = ps_r0x + tx * " << CX << ";\n  const ps_int iy = ps_r0y + ty;\n";
      o << "  const ps_int ps_iz0 = ps_r0z + tz * " << CZ << ";\n  const ps_int iz = ps_iz0;\n";
      o << "  (void)iy; (void)iz;\n";
    } else {
      o << "  if (tx >= ps_rlx || ty >= ps_rly || tz >= ps_rlz) return;\n";
      o << "  const ps_int ix = ps_r0x + tx;\n  const ps_int iy = ps_r0y + ty;\n  const ps_int iz = ps_r0z + tz;\n";
      o << "  (void)ix; (void)iy; (void)iz;\n";
    }
    // per-field cell offsets of the thread's (first) cell X
    auto offsets = [&](const std::string& in) {
      for (size_t i = 0; i < kinds.size(); ++i) {
        if (!kinds[i].is_field) continue;
        std::string A = an((int)i);
        // fields with the strides of an earlier one share its cell offset
        int same = -1;
        if (V.sclass.size() == kinds.size() && V.sclass[i] != '-')
          for (size_t j = 0; j < i && same < 0; ++j)
            if (V.sclass[j] == V.sclass[i]) same = (int)j;
        if (same >= 0) {
          o << in << "const ps_int " << A << "_o = " << an(same) << "_o; (void)" << A << "_o;\n";
        } else if (V.u24) {
          o << in << "const ps_int " << A << "_o = (" << X << " - 1) + (ps_int)__umul24((unsigned)(iy - 1), (unsigned)" << A
            << "_sx) + (ps_int)__umul24((unsigned)(iz - 1), (unsigned)" << A << "_sxy); (void)" << A << "_o;\n";
        } else {
          o << in << "const ps_int " << A << "_o = (" << X << " - 1) + (iy - 1)*" << A << "_sx + (iz - 1)*" << A << "_sxy; (void)" << A << "_o;\n";
        }
        if (V.off32 && kinds[i].ncomp == 0) {
          if (same >= 0 && kinds[same].ncomp == 0 && dtype_bytes(kinds[same].dtype) == dtype_bytes(kinds[i].dtype))
            o << in << "const unsigned " << A << "_vo = " << an(same) << "_vo; (void)" << A << "_vo;\n";
          else
            o << in << "const unsigned " << A << "_vo = (unsigned)" << A << "_o * " << dtype_bytes(kinds[i].dtype) << "u; (void)" << A << "_vo;\n";
        }
      }
    };
    if (!split) offsets("  ");
    for (auto& v : local_order) o << "  " << cty(locals[v]) << " " << local_c[v] << " = 0;\n";
    for (auto& sm : sm_dtype) {
      const std::string P = "sm_" + sanitize(sm.first);
      o << "  " << elem_cty(sm.second) << "* " << P << " = nullptr; ps_int " << P << "_n1 = 0, " << P << "_n2 = 1, " << P << "_n3 = 1;\n";
      o << "  (void)" << P << "; (void)" << P << "_n1; (void)" << P << "_n2; (void)" << P << "_n3;\n";
    }
    ind = 1;
    // the group from (ps_ix0, ps_iz0): all statements for all cells in batches when the masks hold
    // at the group's corner cells (every mask is an interval in x and z, groupable_body), else
    // statement by statement, each for the group's cells in the launch range, in the order in
    // which the lanes of a wavefront run one cell each (a trailing return would end the group
    // after its first cell and is left out)
    const int NC = CX * CZ;
    std::vector<int> corners{0};
    for (int q : {CX - 1, (CZ - 1) * CX, NC - 1})
      if (std::find(corners.begin(), corners.end(), q) == corners.end()) corners.push_back(q);
    auto group = [&](int b, const std::string& first) {
      const std::string in(2 * b, ' ');
      o << in << "bool ps_grp = " << first << ";\n";
      for (int c : corners) {
        std::string conds;
        for (auto& st : K.body->a)
          if (st->k == NK::If) conds += " && (" + expr(st->a[0]).c + ")";
        if (!conds.empty()) o << in << "{ " << group_scope(c) << "ps_grp = ps_grp" << conds << "; }\n";
      }
      o << in << "if (ps_grp) {\n";
      ind = b + 1;
      std::vector<NodeP> sts;
      for (auto& st : K.body->a) {
        if (st->k == NK::If)
          for (auto& t : st->a[1]->a) sts.push_back(t);
        else if (st->k == NK::Assign)
          sts.push_back(st);
      }
      group_statements(sts);
      o << in << "} else {\n";
      for (auto& st : K.body->a) {
        if (st->k == NK::Return) continue;
        for (int c = 0; c < NC; ++c) {
          o << in << "  { " << group_scope(c) << "if (ix < ps_r0x + ps_rlx && iz < ps_r0z + ps_rlz) {\n";
          ind = b + 2;
          stmt(st);
          o << in << "  } }\n";
        }
      }
      o << in << "}\n";
    };
    if (split) {
      o << "  if (tx >= ps_s && tx < ps_s + ps_ng) {\n";
      o << "    const ps_int ps_ix0 = ps_g0 + tx * " << CX << ";\n";
      offsets("    ");
      group(2, CZ > 1 ? "tz * " + std::to_string(CZ) + " + " + std::to_string(CZ - 1) + " < ps_rlz" : std::string("true"));
      // row ends: edge cell e of the row is the e-th cell left of the interior groups or right of
      // them; the edge threads of the row take them in turn
      o << "  } else if (tx >= ps_ngp) {\n";
      o << "    const ps_int ps_nl = ps_ng > 0 ? ps_gs - ps_r0x : ps_rlx;\n";
      o << "    const ps_int ps_ne = ps_rlx - ps_ng * " << CX << ";\n";
      o << "    const ps_int ps_es = (ps_int)gridDim.x * (ps_int)blockDim.x - ps_ngp;\n";
      o << "    for (ps_int ps_e = tx - ps_ngp; ps_e < ps_ne; ps_e += ps_es) {\n";
      o << "      const ps_int ps_ix0 = ps_e < ps_nl ? ps_r0x + ps_e : ps_gs + ps_ng * " << CX << " + (ps_e - ps_nl);\n";
      offsets("      ");
      for (int d = 0; d < CZ; ++d)  // the edge cell's z cells, each with all statements
        for (auto& st : K.body->a) {
          if (st->k == NK::Return) continue;
          o << "      { " << group_scope(d * CX) << (CZ > 1 ? "if (iz < ps_r0z + ps_rlz) " : "") << "{\n";
          ind = 4;
          stmt(st);
          o << "      } }\n";
        }
      o << "    }\n  }\n";
    } else if (GR) {
      group(1, "tx * " + std::to_string(CX) + " + " + std::to_string(CX - 1) + " < ps_rlx && tz * " + std::to_string(CZ) + " + " +
                   std::to_string(CZ - 1) + " < ps_rlz");
    } else {
      block(K.body);
    }
  } else {
    int L = K.nidx - 1;  // streamed dimension
    int ls = K.mm.loopsize;
    // block indices, optionally remapped so that each XCD works on a contiguous part of the grid
    if (K.opt.xcd_remap) {
      o << "  const unsigned ps_nb = gridDim.x * gridDim.y * gridDim.z;\n";
      o << "  const unsigned ps_lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);\n";
      o << "  const unsigned ps_q = ps_nb / 8u, ps_r = ps_nb % 8u, ps_x = ps_lin % 8u, ps_l = ps_lin / 8u;\n";
      o << "  const unsigned ps_id = (ps_x < ps_r ? ps_x * (ps_q + 1u) : ps_r * (ps_q + 1u) + (ps_x - ps_r) * ps_q) + ps_l;\n";
      o << "  const ps_int ps_bx = (ps_int)(ps_id % gridDim.x);\n";
      o << "  const ps_int ps_by = (ps_int)((ps_id / gridDim.x) % gridDim.y);\n";
      o << "  const ps_int ps_bz = (ps_int)(ps_id / (gridDim.x * gridDim.y));\n";
    } else {
      o << "  const ps_int ps_bx = blockIdx.x, ps_by = blockIdx.y, ps_bz = blockIdx.z;\n";
    }
    const char* bl[3] = {"ps_bx", "ps_by", "ps_bz"};
    const char* td[3] = {"threadIdx.x", "threadIdx.y", "threadIdx.z"};
    const char* bd[3] = {"blockDim.x", "blockDim.y", "blockDim.z"};
    const char* rl[3] = {"ps_rlx", "ps_rly", "ps_rlz"};
    const char* r0[3] = {"ps_r0x", "ps_r0y", "ps_r0z"};
    std::string guard;
    for (int d = 0; d < 3; ++d) {
      if (d == L) continue;
      o << "  const ps_int t" << "xyz"[d] << " = " << bl[d] << " * (ps_int)" << bd[d] << " + (ps_int)" << td[d] << ";\n";
      guard += std::string(guard.empty() ? "" : " || ") + "t" + "xyz"[d] + " >= " + rl[d];
    }
    // with LDS tiles every thread of the block must reach the barriers: no early return
    if (V.lds) o << "  const bool ps_act = !(" << guard << ");\n";
    else o << "  if (" << guard << ") return;\n";
    for (int d = 0; d < 3; ++d) {
      if (d == L) continue;
      o << "  const ps_int " << kIdxC[d] << " = " << r0[d] << " + t" << "xyz"[d] << ";\n";
    }
    o << "  const ps_int ps_zb = " << r0[L] << " + " << bl[L] << " * (ps_int)" << ls << ";\n";
    o << "  const ps_int ps_ze = (ps_zb + (ps_int)" << ls << " < " << r0[L] << " + " << rl[L] << ") ? ps_zb + (ps_int)" << ls << " : "
      << r0[L] << " + " << rl[L] << ";\n";
    for (auto& v : local_order) o << "  " << cty(locals[v]) << " " << local_c[v] << " = 0;\n";
    ind = 1;
    emit_memopt_body(L);
  }
  src << o.str() << "}\n";
  E.source = src.str();
  E.lds_bytes = E_lds_bytes;
  return E;
}

static bool uses_idx(const NodeP& n, int q) {
  if (!n) return false;
  if (n->k == NK::Ident && n->s == kIdx[q]) return true;
  for (auto& c : n->a)
    if (uses_idx(c, q)) return true;
  return false;
}

// A conjunction of comparisons in which index q (0 x, 2 z) stands alone on one side of <, <=, >
// or >= wherever it appears.
static bool idx_interval(const NodeP& c, int q) {
  if (!uses_idx(c, q)) return true;
  if (c->k == NK::And) return idx_interval(c->a[0], q) && idx_interval(c->a[1], q);
  if (c->k != NK::Cmp) return false;
  for (size_t i = 0; i + 1 < c->a.size(); ++i) {
    const NodeP &l = c->a[i], &r = c->a[i + 1];
    if (!uses_idx(l, q) && !uses_idx(r, q)) continue;
    const std::string& op = c->ops[i];
    if (op != "<" && op != "<=" && op != ">" && op != ">=") return false;
    const bool lx = l->k == NK::Ident && l->s == kIdx[q], rx = r->k == NK::Ident && r->s == kIdx[q];
    if (!((lx && !uses_idx(r, q)) || (rx && !uses_idx(l, q)))) return false;
  }
  return true;
}

static bool array_assign(const NodeP& s) {
  return s->k == NK::Assign && s->a[0]->k == NK::Index && s->a[0]->a[0]->k == NK::Ident;
}

bool groupable_body(const Kernel& K, bool z) {
  if (K.kernel_language || !K.shared_arrays.empty() || !K.body || K.nidx < 1) return false;
  if (z && K.nidx < 3) return false;
  const auto& st = K.body->a;
  for (size_t i = 0; i < st.size(); ++i) {
    const NodeP& s = st[i];
    if (array_assign(s)) continue;
    if (s->k == NK::Return && s->a.empty() && i + 1 == st.size()) continue;
    if (s->k == NK::If && (s->a.size() < 3 || !s->a[2]) && idx_interval(s->a[0], 0) && (!z || idx_interval(s->a[0], 2))) {
      for (auto& t : s->a[1]->a)
        if (!array_assign(t)) return false;
      continue;
    }
    return false;
  }
  return true;
}

// Largest spread of offsets along dimension q (0 x, 2 z; max - min of c in an index i + c)
// with which the kernel body accesses one array: the number of neighbours beyond its own cell
// that a cell of that array reads along q.
int idx_span(const Kernel& K, int q) {
  std::map<std::string, std::pair<long long, long long>> r;
  std::function<void(const NodeP&)> walk = [&](const NodeP& n) {
    if (!n) return;
    if (n->k == NK::Index && (int)n->a.size() >= q + 2 && n->a[0]->k == NK::Ident) {
      const NodeP& e = n->a[q + 1];
      long long c = 0;
      bool ok = false;
      if (e->k == NK::Ident && e->s == kIdx[q]) ok = true;
      else if (e->k == NK::Binary && (e->s == "+" || e->s == "-") && e->a.size() == 2 && e->a[0]->k == NK::Ident && e->a[0]->s == kIdx[q] &&
               e->a[1]->k == NK::Num && e->a[1]->is_int) {
        ok = true;
        c = e->s == "+" ? e->a[1]->ival : -e->a[1]->ival;
      }
      if (ok) {
        auto it = r.find(n->a[0]->s);
        if (it == r.end()) r[n->a[0]->s] = {c, c};
        else it->second = {std::min(it->second.first, c), std::max(it->second.second, c)};
      }
    }
    for (auto& a : n->a) walk(a);
  };
  walk(K.body);
  long long span = 0;
  for (auto& kv : r) span = std::max(span, kv.second.second - kv.second.first);
  return (int)span;
}

Emitted lower_kernel(const Kernel& k, const std::vector<ArgKind>& kinds, const Variant& v) {
  Emitter e(k, kinds, v);
  return e.run();
}

// ---------------------------------------------------------------------------
// Bounds proof (interval analysis over the launch box)
// ---------------------------------------------------------------------------
namespace {
struct Prover {
  const Kernel& K;
  const LaunchShape& S;
  std::string why;
  bool ok = true;
  // for-loop variables with constant bounds -> [first, last] (in scope while their body is scanned)
  std::map<std::string, std::pair<long long, long long>> loopvars;

  // The range of a cell value index: a literal or a loop variable with constant bounds.
  bool value_range(const NodeP& e, long long& a, long long& b) {
    long long v;
    if (eval(e, v)) {
      a = b = v;
      return true;
    }
    if (e->k == NK::Ident) {
      auto it = loopvars.find(e->s);
      if (it != loopvars.end()) {
        a = it->second.first;
        b = it->second.second;
        return true;
      }
    }
    return false;
  }
  // Value indices of a cell of argument ai (extents as the emitter counts them: the last index
  // runs over the remaining cell dimensions) lie inside the cell.
  bool values_inside(int ai, const std::vector<NodeP>& cs, const NodeP& n) {
    if (ai < 0 || ai >= (int)S.kinds.size() || S.kinds[ai].ncomp <= 0 || cs.empty() || cs.size() > 4) return false;
    const int* cd = S.kinds[ai].celldims;
    long long ext[4] = {1, 1, 1, 1};
    for (size_t d = 0; d < cs.size(); ++d) ext[d] = cd[d];
    for (size_t d = cs.size(); d < 4; ++d) ext[cs.size() - 1] *= cd[d];
    for (size_t d = 0; d < cs.size(); ++d) {
      long long a, b;
      if (!value_range(cs[d], a, b) || a < 1 || b > ext[d]) {
        why = "cell value index of " + to_julia(n) + " is not provably inside the cell";
        return false;
      }
    }
    return true;
  }

  bool eval(const NodeP& e, long long& v) {
    switch (e->k) {
      case NK::Num:
        if (!e->is_int) return false;
        v = e->ival;
        return true;
      case NK::Ident: {
        if (e->s.size() == 5 && e->s.compare(0, 3, "__r") == 0 && e->s[4] >= 'x' && e->s[4] <= 'z') {  // the launch box
          const int d = e->s[4] - 'x';
          if (e->s[3] == '0') v = S.lo[d];
          else if (e->s[3] == 'l') v = S.hi[d] - S.lo[d] + 1;
          else return false;
          return true;
        }
        auto it = K.param_index.find(e->s);
        if (it == K.param_index.end()) return false;
        int i = it->second;
        if (S.kinds[i].is_field) return false;
        if (S.kinds[i].dtype != PS_I32 && S.kinds[i].dtype != PS_I64) return false;
        v = S.ivals[i];
        return true;
      }
      case NK::Call: {
        if ((e->s == "size" || e->s == "__psize" || e->s == "__first" || e->s == "__last") && e->a.size() == 2 &&
            e->a[0]->k == NK::Ident && e->a[1]->k == NK::Num && e->a[1]->is_int) {
          auto it = K.param_index.find(e->a[0]->s);
          if (it == K.param_index.end() || !S.kinds[it->second].is_field) return false;
          long long d = e->a[1]->ival;
          const int i = it->second;
          const bool box = K.opt.padding && i < (int)S.vlo.size();
          if (d < 1 || d > 3) {
            if (e->s == "__first" || e->s == "__last") return false;
            v = 1;
          } else if (e->s == "__psize" || (e->s == "size" && !K.opt.padding)) {
            v = S.dims[i][d - 1];
          } else if (!box) {
            return false;
          } else if (e->s == "size") {
            v = S.vhi[i][d - 1] - S.vlo[i][d - 1] + 1;
          } else {
            v = e->s == "__first" ? S.vlo[i][d - 1] : S.vhi[i][d - 1];
          }
          return true;
        }
        if (e->s == "cellsize" && e->a.size() == 2 && e->a[0]->k == NK::Ident && e->a[1]->k == NK::Num && e->a[1]->is_int) {
          auto it = K.param_index.find(e->a[0]->s);
          if (it == K.param_index.end() || S.kinds[it->second].ncomp <= 0) return false;
          long long d = e->a[1]->ival;
          if (d < 1) return false;
          v = d <= 4 ? S.kinds[it->second].celldims[d - 1] : 1;
          return true;
        }
        return false;
      }
      case NK::Binary: {
        long long a, b;
        if (!eval(e->a[0], a) || !eval(e->a[1], b)) return false;
        if (e->s == "+") v = a + b;
        else if (e->s == "-") v = a - b;
        else if (e->s == "*") v = a * b;
        else if (e->s == "÷" && b != 0) v = a / b;
        else return false;
        return true;
      }
      case NK::Unary: {
        long long a;
        if (!eval(e->a[0], a)) return false;
        if (e->s == "-") v = -a;
        else if (e->s == "+") v = a;
        else return false;
        return true;
      }
      default: return false;
    }
  }
  // var + c with var an index (0..2); returns false otherwise
  bool idxaff(const NodeP& e, int& var, long long& c) {
    var = -1;
    c = 0;
    std::function<bool(const NodeP&, int)> f = [&](const NodeP& x, int sign) -> bool {
      if (x->k == NK::Ident) {
        for (int d = 0; d < 3; ++d)
          if (x->s == kIdx[d]) {
            if (var != -1 || sign < 0) return false;
            var = d;
            return true;
          }
      }
      if (x->k == NK::Binary && x->s == "+") return f(x->a[0], sign) && f(x->a[1], sign);
      if (x->k == NK::Binary && x->s == "-") return f(x->a[0], sign) && f(x->a[1], -sign);
      long long v;
      if (eval(x, v)) {
        c += sign * v;
        return true;
      }
      return false;
    };
    return f(e, 1) && var >= 0;
  }
  void narrow(const NodeP& cond, long long lo[3], long long hi[3]) {
    if (!cond) return;
    if (cond->k == NK::And) {
      narrow(cond->a[0], lo, hi);
      narrow(cond->a[1], lo, hi);
      return;
    }
    if (cond->k != NK::Cmp) return;
    for (size_t i = 0; i + 1 < cond->a.size(); ++i) {
      const NodeP& l = cond->a[i];
      const NodeP& r = cond->a[i + 1];
      std::string op = cond->ops[i];
      int var;
      long long c, b;
      if (idxaff(l, var, c) && eval(r, b)) {
        // var + c op b
        if (op == "<") hi[var] = std::min(hi[var], b - c - 1);
        else if (op == "<=") hi[var] = std::min(hi[var], b - c);
        else if (op == ">") lo[var] = std::max(lo[var], b - c + 1);
        else if (op == ">=") lo[var] = std::max(lo[var], b - c);
        else if (op == "==") {
          lo[var] = std::max(lo[var], b - c);
          hi[var] = std::min(hi[var], b - c);
        }
      } else if (eval(l, b) && idxaff(r, var, c)) {
        // b op var + c
        if (op == "<") lo[var] = std::max(lo[var], b - c + 1);
        else if (op == "<=") lo[var] = std::max(lo[var], b - c);
        else if (op == ">") hi[var] = std::min(hi[var], b - c - 1);
        else if (op == ">=") hi[var] = std::min(hi[var], b - c);
        else if (op == "==") {
          lo[var] = std::max(lo[var], b - c);
          hi[var] = std::min(hi[var], b - c);
        }
      }
    }
  }
  void check_access(const NodeP& n, long long lo[3], long long hi[3]) {
    for (int d = 0; d < 3; ++d)
      if (lo[d] > hi[d]) return;  // unreachable
    // one value of a CellArray cell: field(A, i...)[I] or A[I][i...]; value indices must be
    // literals (their range is checked when the kernel is lowered), the cell access as usual
    if (n->a[0]->k == NK::Call && n->a[0]->s == "field") {
      const NodeP& f = n->a[0];
      if (f->a.empty() || f->a[0]->k != NK::Ident) return;
      auto it = K.param_index.find(f->a[0]->s);
      if (!values_inside(it == K.param_index.end() ? -1 : it->second, std::vector<NodeP>(f->a.begin() + 1, f->a.end()), n)) {
        ok = false;
        return;
      }
      std::vector<NodeP> a{f->a[0]};
      a.insert(a.end(), n->a.begin() + 1, n->a.end());
      check_access(mk(NK::Index, "", a), lo, hi);
      return;
    }
    if (n->a[0]->k == NK::Index) {
      const NodeP& c = n->a[0];
      auto it = c->a[0]->k == NK::Ident ? K.param_index.find(c->a[0]->s) : K.param_index.end();
      if (!values_inside(it == K.param_index.end() ? -1 : it->second, std::vector<NodeP>(n->a.begin() + 1, n->a.end()), n)) {
        ok = false;
        return;
      }
      check_access(c, lo, hi);
      return;
    }
    auto it = K.param_index.find(n->a[0]->s);
    if (it == K.param_index.end()) {
      if (K.shared_arrays.count(n->a[0]->s)) {  // extents known only on the device
        ok = false;
        why = "access " + to_julia(n) + " to a @sharedMem array";
      }
      return;
    }
    int ai = it->second;
    if (!S.kinds[ai].is_field) return;
    size_t ni = n->a.size() - 1;
    for (size_t d = 0; d < std::max<size_t>(ni, 3); ++d) {
      long long dim = d < 3 ? S.dims[ai][d] : 1;
      long long a, b;
      if (d >= ni) {
        a = b = 1;
      } else {
        NodeP e = n->a[1 + d];
        // `end` -> size of this dimension
        std::function<NodeP(const NodeP&)> rep = [&](const NodeP& x) -> NodeP {
          if (x->k == NK::EndKw) return mknum_int(dim);
          auto y = std::make_shared<Node>(*x);
          for (auto& ch : y->a) ch = rep(ch);
          return y;
        };
        e = rep(e);
        int var;
        long long c, v;
        if (eval(e, v)) {
          a = b = v;
        } else if (idxaff(e, var, c)) {
          a = lo[var] + c;
          b = hi[var] + c;
        } else {
          ok = false;
          why = "non-affine index in " + to_julia(n);
          return;
        }
      }
      if (a < 1 || b > dim) {
        ok = false;
        why = "access " + to_julia(n) + " may reach index " + std::to_string(a < 1 ? a : b) + " of dimension " + std::to_string(d + 1) +
              " (size " + std::to_string(dim) + ")";
        return;
      }
    }
    for (size_t d = 1; d < n->a.size(); ++d) scan_expr(n->a[d], lo, hi);
  }
  void scan_expr(const NodeP& e, long long lo[3], long long hi[3]) {
    if (!e || !ok) return;
    if (e->k == NK::Index) {
      check_access(e, lo, hi);
      return;
    }
    for (auto& c : e->a) scan_expr(c, lo, hi);
  }
  void scan_block(const NodeP& b, long long lo[3], long long hi[3]) {
    for (auto& s : b->a) {
      if (!ok) return;
      scan_stmt(s, lo, hi);
    }
  }
  void scan_stmt(const NodeP& s, long long lo[3], long long hi[3]) {
    switch (s->k) {
      case NK::Assign:
        scan_expr(s->a[0], lo, hi);
        scan_expr(s->a[1], lo, hi);
        return;
      case NK::If: {
        scan_expr(s->a[0], lo, hi);
        long long l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
        narrow(s->a[0], l2, h2);
        scan_block(s->a[1], l2, h2);
        if (s->a.size() > 2 && s->a[2]) {
          if (s->a[2]->k == NK::If) scan_stmt(s->a[2], lo, hi);
          else scan_block(s->a[2], lo, hi);
        }
        return;
      }
      case NK::Block: scan_block(s, lo, hi); return;
      case NK::For: {
        // accesses indexed by the loop variable are not affine in a parallel index; a loop with
        // constant bounds gives its variable a known range (cell value indices)
        long long a, b;
        const bool known = eval(s->a[0]->a[0], a) && eval(s->a[0]->a[1], b);
        auto saved = loopvars;
        if (known) loopvars[s->s] = {a, b};
        else loopvars.erase(s->s);
        scan_block(s->a[1], lo, hi);
        loopvars = saved;
        return;
      }
      default: scan_expr(s, lo, hi);
    }
  }
};
}  // namespace

bool prove_in_bounds(const Kernel& k, const LaunchShape& s, bool memopt, std::string* why) {
  (void)memopt;
  Prover P{k, s, "", true};
  long long lo[3] = {s.lo[0], s.lo[1], s.lo[2]}, hi[3] = {s.hi[0], s.hi[1], s.hi[2]};
  P.scan_block(k.body, lo, hi);
  if (why) *why = P.why;
  return P.ok;
}

}  // namespace ps
