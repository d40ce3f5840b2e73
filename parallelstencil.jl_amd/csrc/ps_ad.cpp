// Derivative kernels: the transformation behind `@parallel ∇=(A->Ā, B->B̄) f!(A, B, a)`
// (src/ParallelKernel/parallel.jl:265-301). The reference hands the kernel to Enzyme
// (AD.autodiff_deferred!, src/ParallelKernel/EnzymeExt/autodiff_gpu.jl); here the lowered
// kernel body (finite-difference macros expanded, masks applied, nested functions inlined) is
// differentiated at the source level and the derivative is compiled as a kernel of its own,
// whose parameters are the primal's followed by one shadow array per duplicated argument.
//
// Annotations (defaults src/ParallelKernel/shared.jl:73-76): arrays named in ∇ are
// DuplicatedNoNeed unless annotated Duplicated; every other argument is Const. The primal
// update of an array is kept for Const and Duplicated arrays and dropped for DuplicatedNoNeed
// ones (Enzyme is free to skip it; this backend always does).
//
// Reverse mode. Per thread, the local values of the body are recomputed, then its statements
// are swept in reverse order. A write W[i] = e of a duplicated array takes its adjoint from the
// shadow (g = W̄[i]; W̄[i] = 0) and adds g * ∂e/∂R[j] to the shadow of every duplicated element
// R[j] that e reads; a local x = e passes x̄ * ∂e/∂y on to the locals y it reads. Additions go
// through atomics unless every access of the array is at the thread's own point shifted by one
// fixed offset, so that no other thread reaches the same shadow element. Kept primal updates
// are applied at the end of the thread's sweep: a race-free kernel reads no element that
// another thread writes, so only this thread, which has finished reading, sees the update.
//
// Forward mode. Every value carries a tangent, computed before the value itself; a write
// W[i] = e of a duplicated array also writes Ẇ[i] = ė. No atomics are involved.
//
// Limits (each raised as an ArgumentError): a kernel may not read an array after writing it
// (the derivative would need the overwritten value); in reverse mode a local may be assigned
// at most once along each path through the body (each iteration of a loop counts apart); the
// kernel language (@sharedMem, shuffles, ...) and memopt metadata are not differentiated (the
// derivative of a memopt kernel is a plain kernel with the same geometry).
#include <algorithm>
#include <functional>
#include <map>
#include <set>

#include "ps_amd.h"
#include "ps_lower.h"

namespace ps {
namespace {

const char* kPIdx[3] = {"__ix", "__iy", "__iz"};

template <class F>
void walk(const NodeP& n, F&& f) {
  if (!n) return;
  f(n);
  for (auto& c : n->a) walk(c, f);
}

NodeP ident(const std::string& s) { return mk(NK::Ident, s); }
NodeP fnum(double v) { return mknum_float(v); }
NodeP call(const std::string& f, std::vector<NodeP> a) { return mk(NK::Call, f, std::move(a)); }
NodeP assign(NodeP l, NodeP r) { return mk(NK::Assign, "", {std::move(l), std::move(r)}); }
NodeP cmp(NodeP a, const std::string& op, NodeP b) {
  auto c = mk(NK::Cmp, "", {std::move(a), std::move(b)});
  c->ops = {op};
  return c;
}

bool is_num(const NodeP& n, double v) {
  if (!n || n->k != NK::Num) return false;
  return n->is_int ? (double)n->ival == v : n->num == v;
}
NodeP neg(const NodeP& a) {
  if (a->k == NK::Num) return a->is_int ? mknum_int(-a->ival) : fnum(-a->num);
  if (a->k == NK::Unary && a->s == "-") return a->a[0];
  return mk(NK::Unary, "-", {a});
}
NodeP mul(const NodeP& a, const NodeP& b) {
  if (is_num(a, 0) || is_num(b, 0)) return fnum(0.0);
  if (is_num(a, 1)) return b;
  if (is_num(b, 1)) return a;
  if (is_num(a, -1)) return neg(b);
  if (is_num(b, -1)) return neg(a);
  return mk(NK::Binary, "*", {a, b});
}
NodeP add(const NodeP& a, const NodeP& b) {
  if (is_num(a, 0)) return b;
  if (is_num(b, 0)) return a;
  if (b->k == NK::Unary && b->s == "-") return mk(NK::Binary, "-", {a, b->a[0]});
  return mk(NK::Binary, "+", {a, b});
}
NodeP sub(const NodeP& a, const NodeP& b) {
  if (is_num(b, 0)) return a;
  if (is_num(a, 0)) return neg(b);
  return mk(NK::Binary, "-", {a, b});
}
NodeP dv(const NodeP& a, const NodeP& b) {
  if (is_num(a, 0)) return fnum(0.0);
  if (is_num(b, 1)) return a;
  return mk(NK::Binary, "/", {a, b});
}
NodeP sel(const NodeP& c, const NodeP& a, const NodeP& b) {
  if (is_num(a, 0) && is_num(b, 0)) return fnum(0.0);
  return call("ifelse", {c, a, b});
}

// Effective right-hand side of an assignment: A op= x is A = A op x.
NodeP eff_rhs(const NodeP& s) { return s->s.empty() ? s->a[1] : mk(NK::Binary, s->s, {clone(s->a[0]), s->a[1]}); }

// Functions whose value is piecewise constant or integer: their derivative is zero.
bool flat_function(const std::string& f) {
  static const std::set<std::string> z = {"Int", "Int64", "Int32", "floor", "ceil", "trunc", "round", "sign", "size", "length",
                                          "one", "zero", "cellsize", "firstindex", "lastindex", "isnan", "isinf", "isfinite"};
  return z.count(f) != 0;
}

class Ad {
 public:
  Ad(const Kernel& K, bool reverse, const std::map<std::string, std::string>& annot, const std::string& form = "scatter")
      : K(K), reverse(reverse), form(form) {
    for (auto& p : K.def.params) {
      auto it = annot.find(p.name);
      if (it == annot.end() || it->second == "Const") continue;
      shadow[p.name] = p.name + "#shadow";
      if (it->second == "DuplicatedNoNeed") noneed.insert(p.name);
    }
  }

  Kernel run() {
    if (K.kernel_language || !K.shared_arrays.empty())
      throw PsError(PS_EARGUMENT, "∇: kernel " + K.def.name + " uses the kernel language (@sharedMem, @sync_threads, shuffles, ...), which is not differentiated");
    walk(K.body, [&](const NodeP& x) {
      if (x->k == NK::Index && x->a[0]->k != NK::Ident)
        throw PsError(PS_EARGUMENT, "∇: kernel " + K.def.name + " indexes " + to_julia(x->a[0]) + "; only plain arrays are differentiated (no CellArray values)");
      if (x->k == NK::Assign && x->a[0]->k == NK::Ident && K.param_index.count(x->a[0]->s))
        throw PsError(PS_EARGUMENT, "∇: cannot assign to kernel argument '" + x->a[0]->s + "'");
      if (x->k == NK::Assign && x->a[0]->k == NK::Ident) locals.insert(x->a[0]->s);
    });
    for (auto& s : shadow)
      if (!K.accessed.count(s.first))
        throw PsError(PS_EARGUMENT, "∇: '" + s.first + "' is not an array indexed by kernel " + K.def.name);
    {
      std::set<std::string> w;
      check_read_after_write(K.body, w);
    }
    if (reverse) {
      LocalUse u;
      check_locals(K.body, u);
    }
    find_active();
    find_private_shadows();

    Kernel D = K;
    D.def.name = K.def.name + (reverse ? "_adjoint" : "_tangent") + (form == "scatter" ? "" : "_" + form);
    for (auto& p : K.def.params)
      if (shadow.count(p.name)) D.def.params.push_back(Param{shadow[p.name], p.annot});
    D.opt.memopt = false;  // set again below for the forms that stream
    D.opt.optvars.clear();
    D.opt.optranges.clear();
    D.mm = MemoptMeta{};
    D.opt.tuple_params.clear();
    auto body = mk(NK::Block);
    if (form != "scatter") {
      check_gather();
      body = form == "gather" ? gather_body() : clear_body();
    } else if (reverse) {
      append(body, rev_block(K.body, {}));
      bool tail = false;
      for (auto& w : K.written)
        if (!noneed.count(w)) tail = true;
      if (tail) append(body, primal_tail(K.body));
    } else {
      for (auto& x : active) body->a.push_back(assign(ident(tan(x)), fnum(0.0)));
      append(body, fwd_block(K.body));
    }
    D.body = body;
    refresh_kernel(D);
    // the gather form and the tangent read the seeds, tangents and primal arrays at constant
    // offsets of their own point: a memopt primal gives memopt (z-streaming) derivative kernels
    if ((form == "gather" || !reverse) && K.opt.memopt && K.nidx >= 2) {
      D.opt.memopt = true;
      analyse_memopt(D);
    }
    return D;
  }

 private:
  const Kernel& K;
  bool reverse;
  std::string form;                           // scatter | gather | clear
  bool no_temps = false;                      // contributions as pure expressions (gather form)
  std::map<std::string, std::string> shadow;  // duplicated argument -> shadow parameter
  std::set<std::string> noneed;               // DuplicatedNoNeed arguments
  std::set<std::string> locals, active;
  std::set<std::string> private_shadow;       // shadows that only the owning thread reaches
  int ntmp = 0;

  static std::string adj(const std::string& x) { return x + "#adj"; }
  static std::string tan(const std::string& x) { return x + "#tan"; }
  std::string fresh(const char* w) { return std::string("ad#") + w + std::to_string(++ntmp); }
  static void append(const NodeP& b, const NodeP& more) {
    for (auto& s : more->a) b->a.push_back(s);
  }
  bool is_param_array(const NodeP& x) const { return x->k == NK::Index && x->a[0]->k == NK::Ident && K.param_index.count(x->a[0]->s); }

  // ---- checks --------------------------------------------------------------------------------
  void reads_of(const NodeP& e, std::set<std::string>& r) const {
    walk(e, [&](const NodeP& x) {
      if (is_param_array(x)) r.insert(x->a[0]->s);
    });
  }
  void check_reads(const NodeP& e, const std::set<std::string>& written) const {
    std::set<std::string> r;
    reads_of(e, r);
    for (auto& a : r)
      if (written.count(a))
        throw PsError(PS_EARGUMENT, "∇: kernel " + K.def.name + " reads array '" + a +
                                        "' after writing it; derivative kernels support kernels that read each written array only before writing it");
  }
  void check_read_after_write(const NodeP& blk, std::set<std::string>& written) const {
    for (auto& s : blk->a) check_raw_stmt(s, written);
  }
  void check_raw_stmt(const NodeP& s, std::set<std::string>& written) const {
    switch (s->k) {
      case NK::Assign: {
        check_reads(s->a[1], written);
        if (s->a[0]->k == NK::Index) {
          for (size_t i = 1; i < s->a[0]->a.size(); ++i) check_reads(s->a[0]->a[i], written);
          if (!s->s.empty()) check_reads(s->a[0], written);
          if (is_param_array(s->a[0])) written.insert(s->a[0]->a[0]->s);
        }
        return;
      }
      case NK::If: {
        check_reads(s->a[0], written);
        auto w1 = written, w2 = written;
        check_read_after_write(s->a[1], w1);
        if (s->a.size() > 2 && s->a[2]) {
          if (s->a[2]->k == NK::If) check_raw_stmt(s->a[2], w2);
          else check_read_after_write(s->a[2], w2);
        }
        written.insert(w1.begin(), w1.end());
        written.insert(w2.begin(), w2.end());
        return;
      }
      case NK::For:
        check_reads(s->a[0], written);
        check_read_after_write(s->a[1], written);  // twice: a later iteration reads what an earlier one wrote
        check_read_after_write(s->a[1], written);
        return;
      case NK::Block: check_read_after_write(s, written); return;
      default: check_reads(s, written); return;
    }
  }
  // Reverse mode: each local is defined once along every path (a loop iteration counts apart)
  // and afterwards only incremented (x = x ± e, x ±= e with e not reading x); it is read only
  // after its last increment. The recomputed final value then serves every read, and an
  // increment passes the adjoint of x through unchanged.
  struct LocalUse {
    std::set<std::string> defined, read;
  };
  void mark_reads(const NodeP& e, LocalUse& u) const {
    walk(e, [&](const NodeP& x) {
      if (x->k == NK::Ident && locals.count(x->s)) u.read.insert(x->s);
    });
  }
  void check_locals(const NodeP& blk, LocalUse& u) const {
    for (auto& s : blk->a) check_local_stmt(s, u);
  }
  void check_local_stmt(const NodeP& s, LocalUse& u) const {
    switch (s->k) {
      case NK::Assign: {
        if (s->a[0]->k != NK::Ident) {
          mark_reads(s->a[1], u);
          mark_reads(s->a[0], u);
          return;
        }
        const std::string& x = s->a[0]->s;
        NodeP e;
        bool minus;
        if (is_increment(s, e, minus)) {
          mark_reads(e, u);
          if (u.read.count(x))
            throw PsError(PS_EARGUMENT, "∇ (reverse mode): local variable '" + x + "' of kernel " + K.def.name +
                                            " is read before an update of it; read it only after its last update");
          return;
        }
        mark_reads(s->a[1], u);
        if (!s->s.empty() || u.defined.count(x))
          throw PsError(PS_EARGUMENT, "∇ (reverse mode): local variable '" + x + "' of kernel " + K.def.name +
                                          " is assigned more than once; give each value its own variable (updates x = x ± e are allowed)");
        u.defined.insert(x);
        u.read.erase(x);
        return;
      }
      case NK::If: {
        mark_reads(s->a[0], u);
        LocalUse u1 = u, u2 = u;
        check_locals(s->a[1], u1);
        if (s->a.size() > 2 && s->a[2]) {
          if (s->a[2]->k == NK::If) check_local_stmt(s->a[2], u2);
          else check_locals(s->a[2], u2);
        }
        u.defined = u1.defined;
        u.defined.insert(u2.defined.begin(), u2.defined.end());
        u.read = u1.read;
        u.read.insert(u2.read.begin(), u2.read.end());
        return;
      }
      case NK::For: {
        mark_reads(s->a[0], u);
        // twice: a read in one iteration followed by an update in the next is caught; locals
        // defined in the body belong to one iteration
        LocalUse it = u;
        for (int pass = 0; pass < 2; ++pass) {
          it.defined = u.defined;
          check_locals(s->a[1], it);
        }
        u.read = it.read;
        return;
      }
      case NK::Block: check_locals(s, u); return;
      default: mark_reads(s, u); return;
    }
  }
  // x = x ± e, x = e + x, x += e, x -= e with e not reading x: sets e and the sign.
  static bool refers(const NodeP& n, const std::string& x) {
    bool hit = false;
    walk(n, [&](const NodeP& y) {
      if (y->k == NK::Ident && y->s == x) hit = true;
    });
    return hit;
  }
  static bool is_increment(const NodeP& s, NodeP& e, bool& minus) {
    if (s->k != NK::Assign || s->a[0]->k != NK::Ident) return false;
    const std::string& x = s->a[0]->s;
    const NodeP& r = s->a[1];
    if (s->s == "+" || s->s == "-") {
      if (refers(r, x)) return false;
      e = r, minus = s->s == "-";
      return true;
    }
    if (!s->s.empty() || r->k != NK::Binary || (r->s != "+" && r->s != "-")) return false;
    if (r->a[0]->k == NK::Ident && r->a[0]->s == x && !refers(r->a[1], x)) {
      e = r->a[1], minus = r->s == "-";
      return true;
    }
    if (r->s == "+" && r->a[1]->k == NK::Ident && r->a[1]->s == x && !refers(r->a[0], x)) {
      e = r->a[0], minus = false;
      return true;
    }
    return false;
  }
  // Locals defined (not merely updated) anywhere inside a block.
  void defs_in(const NodeP& blk, std::set<std::string>& d) const {
    walk(blk, [&](const NodeP& s) {
      NodeP e;
      bool m;
      if (s->k == NK::Assign && s->a[0]->k == NK::Ident && !is_increment(s, e, m)) d.insert(s->a[0]->s);
    });
  }
  static bool has_increment(const NodeP& blk) {
    bool hit = false;
    walk(blk, [&](const NodeP& s) {
      NodeP e;
      bool m;
      if (is_increment(s, e, m)) hit = true;
    });
    return hit;
  }

  // ---- activity ------------------------------------------------------------------------------
  // True if the value of e depends differentiably on a duplicated array or an active local.
  bool dep(const NodeP& e) const {
    switch (e->k) {
      case NK::Num: case NK::Str: case NK::EndKw: case NK::Cmp: case NK::And: case NK::Or: case NK::Macro: return false;
      case NK::Ident: return active.count(e->s) != 0;
      case NK::Index: return e->a[0]->k == NK::Ident && shadow.count(e->a[0]->s) != 0;
      case NK::Unary: return e->s != "!" && dep(e->a[0]);
      case NK::Binary: return e->s != "÷" && e->s != "%" && (dep(e->a[0]) || dep(e->a[1]));
      case NK::Tern: return dep(e->a[1]) || dep(e->a[2]);
      case NK::Call:
        if (flat_function(e->s)) return false;
        if (e->s == "ifelse") return dep(e->a.at(1)) || dep(e->a.at(2));
        return std::any_of(e->a.begin(), e->a.end(), [&](const NodeP& c) { return dep(c); });
      default:
        for (auto& c : e->a)
          if (dep(c)) throw PsError(PS_EARGUMENT, "∇: cannot differentiate " + to_julia(e) + " in kernel " + K.def.name);
        return false;
    }
  }
  void find_active() {
    for (bool changed = true; changed;) {
      changed = false;
      walk(K.body, [&](const NodeP& s) {
        if (s->k == NK::Assign && s->a[0]->k == NK::Ident && !active.count(s->a[0]->s) && dep(eff_rhs(s))) {
          active.insert(s->a[0]->s);
          changed = true;
        }
      });
    }
  }
  // Shadows whose every access (reads, and writes in reverse mode) is at the thread's own point
  // shifted by one fixed offset: distinct threads then reach distinct shadow elements.
  void find_private_shadows() {
    std::map<std::string, std::set<std::string>> forms;
    std::set<std::string> bad;
    walk(K.body, [&](const NodeP& x) {
      if (!is_param_array(x) || !shadow.count(x->a[0]->s)) return;
      const std::string& A = x->a[0]->s;
      if ((int)x->a.size() - 1 != K.nidx) {
        bad.insert(A);
        return;
      }
      std::string f;
      for (int d = 0; d < K.nidx; ++d) {
        const NodeP& e = x->a[d + 1];
        const NodeP& v = (e->k == NK::Binary && (e->s == "+" || e->s == "-") && e->a[1]->k == NK::Num && e->a[1]->is_int) ? e->a[0] : e;
        if (!(v->k == NK::Ident && v->s == kPIdx[d])) bad.insert(A);
        f += to_julia(e) + ",";
      }
      forms[A].insert(f);
    });
    for (auto& kv : forms)
      if (!bad.count(kv.first) && kv.second.size() == 1) private_shadow.insert(kv.first);
  }

  // ---- local derivatives ---------------------------------------------------------------------
  // d f(x)/dx for the functions of one argument; e is the call itself (reused for exp, tanh ...)
  NodeP coef1(const std::string& f, const NodeP& x, const NodeP& e) const {
    auto X = [&] { return clone(x); };
    auto E = [&] { return clone(e); };
    if (f == "exp" || f == "expm1") return f == "exp" ? E() : call("exp", {X()});
    if (f == "exp2") return mul(E(), fnum(0.69314718055994530942));
    if (f == "exp10") return mul(E(), fnum(2.30258509299404568402));
    if (f == "log") return dv(fnum(1.0), X());
    if (f == "log2") return dv(fnum(1.0), mul(X(), fnum(0.69314718055994530942)));
    if (f == "log10") return dv(fnum(1.0), mul(X(), fnum(2.30258509299404568402)));
    if (f == "log1p") return dv(fnum(1.0), add(fnum(1.0), X()));
    if (f == "sqrt") return dv(fnum(0.5), E());
    if (f == "cbrt") return dv(fnum(1.0), mul(fnum(3.0), mul(E(), E())));
    if (f == "sin") return call("cos", {X()});
    if (f == "cos") return neg(call("sin", {X()}));
    if (f == "tan") return add(fnum(1.0), mul(E(), E()));
    if (f == "asin") return dv(fnum(1.0), call("sqrt", {sub(fnum(1.0), mul(X(), X()))}));
    if (f == "acos") return neg(dv(fnum(1.0), call("sqrt", {sub(fnum(1.0), mul(X(), X()))})));
    if (f == "atan") return dv(fnum(1.0), add(fnum(1.0), mul(X(), X())));
    if (f == "sinh") return call("cosh", {X()});
    if (f == "cosh") return call("sinh", {X()});
    if (f == "tanh") return sub(fnum(1.0), mul(E(), E()));
    if (f == "erf") return mul(fnum(1.12837916709551257390), call("exp", {neg(mul(X(), X()))}));
    if (f == "erfc") return mul(fnum(-1.12837916709551257390), call("exp", {neg(mul(X(), X()))}));
    if (f == "abs") return call("sign", {X()});
    if (f == "inv") return neg(mul(E(), E()));
    if (f == "Float64" || f == "Float32" || f == "+") return fnum(1.0);
    throw PsError(PS_EARGUMENT, "∇: the derivative of " + f + " is not available (kernel " + K.def.name + ")");
  }
  // Partial derivatives of an expression with respect to its direct operands: value of e =
  // f(c1, c2, ...), de = sum coef_i * dc_i. Selections (ifelse, ?:, max, min) are not handled
  // here: their untaken operand must contribute nothing, not 0 times its derivative.
  std::vector<std::pair<NodeP, NodeP>> partials(const NodeP& e) const {
    std::vector<std::pair<NodeP, NodeP>> p;
    switch (e->k) {
      case NK::Unary:
        if (e->s == "-") p.push_back({e->a[0], fnum(-1.0)});
        else p.push_back({e->a[0], fnum(1.0)});
        return p;
      case NK::Binary: {
        const NodeP &a = e->a[0], &b = e->a[1];
        if (e->s == "+") return {{a, fnum(1.0)}, {b, fnum(1.0)}};
        if (e->s == "-") return {{a, fnum(1.0)}, {b, fnum(-1.0)}};
        if (e->s == "*") return {{a, clone(b)}, {b, clone(a)}};
        if (e->s == "/") return {{a, dv(fnum(1.0), clone(b))}, {b, neg(dv(clone(e), clone(b)))}};
        if (e->s == "^") {
          NodeP ca;
          if (b->k == NK::Num && b->is_int) {
            long long n = b->ival;
            ca = n == 0 ? fnum(0.0) : n == 1 ? fnum(1.0) : n == 2 ? mul(fnum(2.0), clone(a))
                                                          : mul(fnum((double)n), mk(NK::Binary, "^", {clone(a), mknum_int(n - 1)}));
          } else {
            ca = mul(clone(b), mk(NK::Binary, "^", {clone(a), sub(clone(b), fnum(1.0))}));
          }
          p.push_back({a, ca});
          if (dep(b)) p.push_back({b, mul(clone(e), call("log", {clone(a)}))});
          return p;
        }
        throw PsError(PS_EARGUMENT, "∇: cannot differentiate operator " + e->s + " in kernel " + K.def.name);
      }
      case NK::Call: {
        const std::string& f = e->s;
        if (f == "muladd" || f == "fma")
          return {{e->a.at(0), clone(e->a.at(1))}, {e->a.at(1), clone(e->a.at(0))}, {e->a.at(2), fnum(1.0)}};
        if (f == "hypot") return {{e->a.at(0), dv(clone(e->a[0]), clone(e))}, {e->a.at(1), dv(clone(e->a[1]), clone(e))}};
        if (f == "atan" && e->a.size() == 2) {  // atan(y, x)
          NodeP r2 = add(mul(clone(e->a[0]), clone(e->a[0])), mul(clone(e->a[1]), clone(e->a[1])));
          return {{e->a[0], dv(clone(e->a[1]), r2)}, {e->a[1], neg(dv(clone(e->a[0]), clone(r2)))}};
        }
        if (e->a.size() != 1) throw PsError(PS_EARGUMENT, "∇: the derivative of " + f + " is not available (kernel " + K.def.name + ")");
        return {{e->a[0], coef1(f, e->a[0], e)}};
      }
      default: throw PsError(PS_EARGUMENT, "∇: cannot differentiate " + to_julia(e) + " in kernel " + K.def.name);
    }
  }
  // Selections: (condition, taken-if-true operand, taken-if-false operand).
  bool selection(const NodeP& e, NodeP& c, NodeP& a, NodeP& b) const {
    if (e->k == NK::Tern) {
      c = e->a[0], a = e->a[1], b = e->a[2];
      return true;
    }
    if (e->k != NK::Call) return false;
    if (e->s == "ifelse") {
      c = e->a.at(0), a = e->a.at(1), b = e->a.at(2);
      return true;
    }
    if (e->s == "max" || e->s == "min") {
      if (e->a.size() < 2) return false;
      NodeP x = e->a[0];  // max(x, y, z) = max(max(x, y), z)
      for (size_t i = 1; i + 1 < e->a.size(); ++i) x = call(e->s, {x, e->a[i]});
      const NodeP& y = e->a.back();
      c = cmp(clone(x), e->s == "max" ? ">=" : "<=", clone(y));
      a = x, b = y;
      return true;
    }
    return false;
  }

  // ---- reverse mode --------------------------------------------------------------------------
  // Conditions of the selections (ifelse, ?:, max, min) whose branch is being differentiated:
  // a contribution made inside a branch is evaluated only where the branch is taken, as in the
  // primal (its coefficients may read elements or take values that exist only there).
  std::vector<NodeP> guards;
  NodeP guard_cond() const {
    NodeP c;
    for (auto& g : guards) c = c ? mk(NK::And, "&&", {c, clone(g)}) : clone(g);
    return c;
  }
  NodeP guarded(const NodeP& v) const {
    if (guards.empty() || is_num(v, 0)) return v;
    return call("ifelse", {guard_cond(), v, fnum(0.0)});
  }
  NodeP share(const NodeP& v, const NodeP& out) {
    if (v->k == NK::Ident || v->k == NK::Num || no_temps) return v;
    std::string t = fresh("a");
    out->a.push_back(assign(ident(t), guarded(v)));
    return ident(t);
  }
  // Contributions of one statement to shadow elements, summed per element before they are
  // added (one atomic per element read, however often the statement reads it).
  struct Pending {
    std::string key, array;  // shadow element and branch conditions as text, the duplicated array
    NodeP target, sum, cond;  // cond: the branch conditions of the read (null outside selections)
    NodeP value() const { return cond ? call("ifelse", {clone(cond), sum, fnum(0.0)}) : sum; }
  };
  std::vector<Pending> pending;
  void accumulate(const NodeP& r, const NodeP& v, const NodeP&) {
    std::vector<NodeP> ia{ident(shadow.at(r->a[0]->s))};
    for (size_t i = 1; i < r->a.size(); ++i) ia.push_back(clone(r->a[i]));
    NodeP target = mk(NK::Index, "", ia);
    NodeP c = guard_cond();
    const std::string key = to_julia(target) + "|" + (c ? to_julia(c) : "");
    for (auto& p : pending)
      if (p.key == key) {
        p.sum = add(p.sum, v);
        return;
      }
    pending.push_back({key, r->a[0]->s, target, v, c});
  }
  // Scatter form: one addition per element and branch condition; a read made inside a branch
  // adds only where the branch is taken (its element may not exist elsewhere).
  void flush(const NodeP& out) {
    for (auto& p : pending) {
      NodeP st = private_shadow.count(p.array) ? assign(p.target, add(clone(p.target), p.sum))
                                               : mk(NK::Macro, "ps_atomic_add", {p.target, p.sum});
      if (p.cond) out->a.push_back(mk(NK::If, "", {p.cond, mk(NK::Block, "", {st}), nullptr}));
      else out->a.push_back(st);
    }
    pending.clear();
  }
  // Adds the adjoint g of the value of e to everything e depends on.
  void back(const NodeP& e, NodeP g, const NodeP& out) {
    if (!dep(e)) return;
    if (e->k == NK::Ident) {
      out->a.push_back(assign(ident(adj(e->s)), add(ident(adj(e->s)), guarded(g))));
      return;
    }
    if (e->k == NK::Index) {
      accumulate(e, g, out);  // under the branch conditions (the element itself may exist only there)
      return;
    }
    NodeP c, a, b;
    if (selection(e, c, a, b)) {
      g = share(g, out);
      guards.push_back(clone(c));
      back(a, g, out);
      guards.back() = mk(NK::Unary, "!", {clone(c)});
      back(b, g, out);
      guards.pop_back();
      return;
    }
    auto ps = partials(e);
    int n = 0;
    for (auto& pc : ps) n += dep(pc.first);
    if (n > 1) g = share(g, out);
    for (auto& pc : ps)
      if (dep(pc.first)) back(pc.first, mul(g, pc.second), out);
  }
  // Local assignments of a block in program order (values only, no array writes), without the
  // updates of the frozen locals (defined outside the block being reversed); a loop is
  // included where it updates locals.
  NodeP recompute(const NodeP& blk, const std::set<std::string>& frozen) {
    auto out = mk(NK::Block);
    for (auto& s : blk->a) {
      NodeP e;
      bool m;
      if (s->k == NK::Assign && s->a[0]->k == NK::Ident) {
        if (!(is_increment(s, e, m) && frozen.count(s->a[0]->s))) out->a.push_back(clone(s));
      } else if (s->k == NK::If) {
        out->a.push_back(recompute_if(s, frozen));
      } else if (s->k == NK::Block) {
        append(out, recompute(s, frozen));
      } else if (s->k == NK::For && has_increment(s->a[1])) {
        out->a.push_back(mk(NK::For, s->s, {clone(s->a[0]), recompute(s->a[1], frozen)}));
      }
    }
    return out;
  }
  NodeP recompute_if(const NodeP& s, const std::set<std::string>& frozen) {
    NodeP e;
    if (s->a.size() > 2 && s->a[2]) e = s->a[2]->k == NK::If ? recompute_if(s->a[2], frozen) : recompute(s->a[2], frozen);
    return mk(NK::If, "", {clone(s->a[0]), recompute(s->a[1], frozen), e});
  }
  void adjoint_decls(const NodeP& blk, const NodeP& out, std::set<std::string>& done) {
    for (auto& s : blk->a) {
      NodeP e;
      bool m;
      if (s->k == NK::Assign && s->a[0]->k == NK::Ident && active.count(s->a[0]->s) && !done.count(s->a[0]->s) && !is_increment(s, e, m)) {
        done.insert(s->a[0]->s);
        out->a.push_back(assign(ident(adj(s->a[0]->s)), fnum(0.0)));
      } else if (s->k == NK::If) {
        for (const NodeP* q = &s; *q && (*q)->k == NK::If;) {
          adjoint_decls((*q)->a[1], out, done);
          if ((*q)->a.size() > 2 && (*q)->a[2]) {
            if ((*q)->a[2]->k == NK::If) {
              q = &(*q)->a[2];
              continue;
            }
            adjoint_decls((*q)->a[2], out, done);
          }
          break;
        }
      } else if (s->k == NK::Block) {
        adjoint_decls(s, out, done);
      }
    }
  }
  NodeP rev_block(const NodeP& blk, const std::set<std::string>& frozen) {
    auto out = recompute(blk, frozen);
    std::set<std::string> done;
    adjoint_decls(blk, out, done);
    append(out, sweep(blk));
    return out;
  }
  NodeP sweep(const NodeP& blk) {
    auto out = mk(NK::Block);
    for (auto it = blk->a.rbegin(); it != blk->a.rend(); ++it) sweep_stmt(*it, out);
    return out;
  }
  void sweep_stmt(const NodeP& s, const NodeP& out) {
    switch (s->k) {
      case NK::Assign: {
        const NodeP& l = s->a[0];
        if (l->k == NK::Ident) {
          NodeP e;
          bool minus;
          if (active.count(l->s)) {
            if (is_increment(s, e, minus)) back(e, minus ? neg(ident(adj(l->s))) : ident(adj(l->s)), out);
            else back(eff_rhs(s), ident(adj(l->s)), out);
          }
          flush(out);
          return;
        }
        if (!is_param_array(l) || !shadow.count(l->a[0]->s)) return;
        std::vector<NodeP> ia{ident(shadow.at(l->a[0]->s))};
        for (size_t i = 1; i < l->a.size(); ++i) ia.push_back(clone(l->a[i]));
        NodeP W = mk(NK::Index, "", ia);
        NodeP rhs = eff_rhs(s);
        if (!dep(rhs)) {
          out->a.push_back(assign(W, fnum(0.0)));
          return;
        }
        std::string g = fresh("g");
        out->a.push_back(assign(ident(g), clone(W)));
        out->a.push_back(assign(W, fnum(0.0)));
        back(rhs, ident(g), out);
        flush(out);
        return;
      }
      case NK::If: out->a.push_back(sweep_if(s)); return;
      case NK::Block: append(out, sweep(s)); return;
      case NK::For: {
        // the iterations in reverse order: a later iteration's write of an element overrides an
        // earlier one's, so its adjoint is taken first
        const NodeP& r = s->a[0];
        std::string k = fresh("r");
        auto body = mk(NK::Block);
        body->a.push_back(assign(ident(s->s), sub(add(clone(r->a[0]), clone(r->a[1])), ident(k))));
        std::set<std::string> own, frozen;
        defs_in(s->a[1], own);
        for (auto& x : locals)
          if (!own.count(x)) frozen.insert(x);
        append(body, rev_block(s->a[1], frozen));
        auto f = mk(NK::For, k, {mk(NK::Range, "", {clone(r->a[0]), clone(r->a[1])}), body});
        out->a.push_back(f);
        return;
      }
      default: return;
    }
  }
  NodeP sweep_if(const NodeP& s) {
    NodeP e;
    if (s->a.size() > 2 && s->a[2]) e = s->a[2]->k == NK::If ? sweep_if(s->a[2]) : sweep(s->a[2]);
    return mk(NK::If, "", {clone(s->a[0]), sweep(s->a[1]), e});
  }
  // The primal body without the writes of DuplicatedNoNeed arrays.
  NodeP primal_tail(const NodeP& blk) {
    auto out = mk(NK::Block);
    for (auto& s : blk->a) {
      if (s->k == NK::Assign && is_param_array(s->a[0]) && noneed.count(s->a[0]->a[0]->s)) continue;
      if (s->k == NK::If) out->a.push_back(tail_if(s));
      else if (s->k == NK::For) out->a.push_back(mk(NK::For, s->s, {clone(s->a[0]), primal_tail(s->a[1])}));
      else if (s->k == NK::Block) out->a.push_back(primal_tail(s));
      else out->a.push_back(clone(s));
    }
    return out;
  }
  NodeP tail_if(const NodeP& s) {
    NodeP e;
    if (s->a.size() > 2 && s->a[2]) e = s->a[2]->k == NK::If ? tail_if(s->a[2]) : primal_tail(s->a[2]);
    return mk(NK::If, "", {clone(s->a[0]), primal_tail(s->a[1]), e});
  }

  // ---- reverse mode, gather form (@parallel kernels) ------------------------------------------
  // Every statement of a @parallel kernel is `if mask(i): W[i + w] = e(i)`. The adjoint that a
  // read R[i + o] of e receives is g(i) * ∂e/∂R[i + o](i) with g(i) = W̄[i + w]; the point
  // j = i + o of R̄ thus collects, over the statements and read offsets o, the contribution
  // evaluated at i = j - o where mask(j - o) holds. Summed by the thread of point j, without
  // atomics. The seeds W̄ are read by several threads, so they are cleared by a second kernel
  // (the clear form), which also applies the kept primal updates.
  static bool offset_of(const NodeP& e, int d, long long& c) {
    if (e->k == NK::Ident && e->s == kPIdx[d]) {
      c = 0;
      return true;
    }
    if (e->k == NK::Binary && (e->s == "+" || e->s == "-") && e->a[0]->k == NK::Ident && e->a[0]->s == kPIdx[d] && e->a[1]->k == NK::Num &&
        e->a[1]->is_int) {
      c = e->s == "+" ? e->a[1]->ival : -e->a[1]->ival;
      return true;
    }
    return false;
  }
  // The body as a list of guarded writes: every array write with the condition of its path
  // and its right-hand side, locals inlined (a local set in branches becomes an ifelse of the
  // branch values). Loop-free bodies only; a duplicated array may be written once per path.
  struct FlatWrite {
    NodeP cond, target, rhs;
  };
  std::vector<FlatWrite> flat;
  std::string flat_why;
  NodeP inline_locals(const NodeP& e, const std::map<std::string, NodeP>& env) const {
    std::function<NodeP(const NodeP&)> f = [&](const NodeP& n) -> NodeP {
      if (!n) return n;
      if (n->k == NK::Ident) {
        auto it = env.find(n->s);
        if (it != env.end()) {
          if (!it->second) throw std::string("local '" + n->s + "' is not defined on every path");
          return clone(it->second);
        }
        return clone(n);
      }
      auto c = std::make_shared<Node>(*n);
      for (auto& x : c->a) x = f(x);
      return c;
    };
    return canonical_indices(f(e));
  }
  static NodeP conj(const NodeP& a, const NodeP& b) { return !a ? b : mk(NK::And, "&&", {a, b}); }
  void flatten(const NodeP& blk, const NodeP& cond, std::map<std::string, NodeP>& env, std::set<std::string>& written) {
    for (auto& st : blk->a) {
      switch (st->k) {
        case NK::Assign:
          if (st->a[0]->k == NK::Ident) {
            env[st->a[0]->s] = inline_locals(eff_rhs(st), env);
          } else {
            NodeP t = inline_locals(st->a[0], env);
            const std::string& W = t->a[0]->s;
            if (shadow.count(W) && written.count(W)) throw std::string("duplicated array '" + W + "' is written twice on one path");
            written.insert(W);
            flat.push_back({cond ? clone(cond) : nullptr, t, inline_locals(eff_rhs(st), env)});
          }
          break;
        case NK::If: {
          NodeP c = inline_locals(st->a[0], env);
          auto e1 = env, e2 = env;
          auto w1 = written, w2 = written;
          flatten(st->a[1], conj(cond, c), e1, w1);
          if (st->a.size() > 2 && st->a[2]) {
            auto eb = st->a[2]->k == NK::If ? mk(NK::Block, "", {st->a[2]}) : st->a[2];
            flatten(eb, conj(cond, mk(NK::Unary, "!", {clone(c)})), e2, w2);
          }
          std::set<std::string> names;
          for (auto& kv : e1) names.insert(kv.first);
          for (auto& kv : e2) names.insert(kv.first);
          for (auto& x : names) {
            NodeP v1 = e1.count(x) ? e1[x] : nullptr, v2 = e2.count(x) ? e2[x] : nullptr;
            if (env.count(x) && v1 == env[x] && v2 == env[x]) continue;
            env[x] = (v1 && v2) ? call("ifelse", {clone(c), v1, v2}) : nullptr;  // undefined on one path
          }
          written.insert(w1.begin(), w1.end());
          written.insert(w2.begin(), w2.end());
          break;
        }
        case NK::Block: flatten(st, cond, env, written); break;
        case NK::For: throw std::string("the kernel has loops");
        default: break;
      }
    }
  }
  void check_gather() {
    auto no = [&](const std::string& why) {
      throw PsError(PS_EARGUMENT, "∇: the gather form does not apply to kernel " + K.def.name + ": " + why);
    };
    if (!reverse) no("forward mode");
    flat.clear();
    try {
      std::map<std::string, NodeP> env;
      std::set<std::string> written;
      flatten(K.body, nullptr, env, written);
    } catch (const std::string& why) {
      no(why);
    }
    for (auto& w : flat) {
      walk(w.rhs, [&](const NodeP& x) {
        if (!is_param_array(x) || !shadow.count(x->a[0]->s)) return;
        if (K.written.count(x->a[0]->s)) {
          // an in-place update reads the element it writes: its derivative stays with the
          // element (applied by the clear kernel); any other read of a written array is refused
          if (!(x->a[0]->s == w.target->a[0]->s && to_julia(x) == to_julia(w.target)))
            no("duplicated array '" + x->a[0]->s + "' is read and written at different elements");
          return;
        }
        if ((int)x->a.size() - 1 != K.nidx) no("'" + x->a[0]->s + "' has other dimensions than the kernel");
        long long c;
        for (int d = 0; d < K.nidx; ++d)
          if (!offset_of(x->a[d + 1], d, c)) no("'" + x->a[0]->s + "' is read at a non-constant offset");
      });
    }
  }
  // Copy of e at the point i = j - o: parallel index d becomes (index - o_d).
  NodeP shift(const NodeP& e, const long long o[3]) const {
    std::function<NodeP(const NodeP&)> f = [&](const NodeP& n) -> NodeP {
      if (n->k == NK::Ident)
        for (int d = 0; d < 3; ++d)
          if (n->s == kPIdx[d] && o[d] != 0) return mk(NK::Binary, o[d] > 0 ? "-" : "+", {ident(n->s), mknum_int(o[d] > 0 ? o[d] : -o[d])});
      auto c = std::make_shared<Node>(*n);
      for (auto& x : c->a) x = x ? f(x) : x;
      return c;
    };
    return canonical_indices(f(e));
  }
  NodeP gather_body() {
    no_temps = true;
    auto body = mk(NK::Block);
    for (auto& prm : K.def.params) {
      const std::string& R = prm.name;
      if (!shadow.count(R) || K.written.count(R) || !K.accessed.count(R)) continue;
      std::string acc = fresh("acc");
      body->a.push_back(assign(ident(acc), fnum(0.0)));
      for (auto& w : flat) {
        const std::string& W = w.target->a[0]->s;
        if (!shadow.count(W)) continue;
        std::vector<NodeP> ia{ident(shadow.at(W))};
        for (size_t i = 1; i < w.target->a.size(); ++i) ia.push_back(clone(w.target->a[i]));
        auto sink = mk(NK::Block);
        back(w.rhs, mk(NK::Index, "", ia), sink);
        for (auto& p : pending) {
          if (p.array != R) continue;
          long long o[3] = {0, 0, 0};
          for (int d = 0; d < K.nidx; ++d) offset_of(p.target->a[d + 1], d, o[d]);
          // the source point i = j - o exists (it lies in the launch box) and takes this path
          NodeP cond;
          for (int d = 0; d < K.nidx; ++d) {
            const std::string r0 = std::string("__r0") + "xyz"[d], rl = std::string("__rl") + "xyz"[d];
            NodeP i = shift(ident(kPIdx[d]), o);
            cond = conj(cond, cmp(ident(r0), "<=", clone(i)));
            cond = conj(cond, cmp(i, "<", mk(NK::Binary, "+", {ident(r0), ident(rl)})));
          }
          if (w.cond) cond = conj(cond, shift(w.cond, o));
          auto then = mk(NK::Block, "", {assign(ident(acc), add(ident(acc), shift(p.value(), o)))});
          body->a.push_back(mk(NK::If, "", {cond, then, nullptr}));
        }
        pending.clear();
      }
      NodeP guard;
      std::vector<NodeP> ja{ident(shadow.at(R))};
      for (int d = 0; d < K.nidx; ++d) {
        guard = conj(guard, cmp(ident(kPIdx[d]), "<=", call("size", {ident(R), mknum_int(d + 1)})));
        ja.push_back(ident(kPIdx[d]));
      }
      NodeP target = mk(NK::Index, "", ja);
      body->a.push_back(mk(NK::If, "", {guard, mk(NK::Block, "", {assign(target, add(clone(target), ident(acc)))}), nullptr}));
    }
    no_temps = false;
    return body;
  }
  NodeP clear_body() {
    no_temps = true;
    auto body = mk(NK::Block);
    for (auto& w : flat) {
      const std::string& W = w.target->a[0]->s;
      auto blk = mk(NK::Block);
      if (shadow.count(W)) {
        // the seed is consumed: it becomes the adjoint of the element's old value, g * de/dW[i]
        // for an in-place update (e reads W[i]) and zero otherwise
        std::vector<NodeP> ia{ident(shadow.at(W))};
        for (size_t i = 1; i < w.target->a.size(); ++i) ia.push_back(clone(w.target->a[i]));
        NodeP seed = mk(NK::Index, "", ia);
        std::string g = fresh("g");
        auto sink = mk(NK::Block);
        back(w.rhs, ident(g), sink);
        NodeP own = fnum(0.0);
        for (auto& p : pending)
          if (p.array == W) own = add(own, p.value());
        pending.clear();
        if (!(own->k == NK::Ident && own->s == g)) {  // a plain in-place update leaves the seed as it is
          if (!is_num(own, 0)) blk->a.push_back(assign(ident(g), clone(seed)));
          blk->a.push_back(assign(seed, own));
        }
      }
      if (!noneed.count(W)) blk->a.push_back(assign(clone(w.target), clone(w.rhs)));
      if (blk->a.empty()) continue;
      if (w.cond) body->a.push_back(mk(NK::If, "", {clone(w.cond), blk, nullptr}));
      else append(body, blk);
    }
    no_temps = false;
    return body;
  }

  // ---- forward mode --------------------------------------------------------------------------
  NodeP tangent(const NodeP& e) {
    if (!dep(e)) return fnum(0.0);
    if (e->k == NK::Ident) return ident(tan(e->s));
    if (e->k == NK::Index) {
      std::vector<NodeP> ia{ident(shadow.at(e->a[0]->s))};
      for (size_t i = 1; i < e->a.size(); ++i) ia.push_back(clone(e->a[i]));
      return mk(NK::Index, "", ia);
    }
    NodeP c, a, b;
    if (selection(e, c, a, b)) return sel(clone(c), tangent(a), tangent(b));
    NodeP t = fnum(0.0);
    for (auto& pc : partials(e))
      if (dep(pc.first)) t = add(t, mul(pc.second, tangent(pc.first)));
    return t;
  }
  NodeP fwd_block(const NodeP& blk) {
    auto out = mk(NK::Block);
    for (auto& s : blk->a) fwd_stmt(s, out);
    return out;
  }
  void fwd_stmt(const NodeP& s, const NodeP& out) {
    switch (s->k) {
      case NK::Assign: {
        const NodeP& l = s->a[0];
        NodeP rhs = eff_rhs(s);
        if (l->k == NK::Ident) {
          if (!active.count(l->s)) {
            out->a.push_back(clone(s));
            return;
          }
          std::string t = fresh("t");
          out->a.push_back(assign(ident(t), tangent(rhs)));  // before the value: rhs may read l
          out->a.push_back(clone(s));
          out->a.push_back(assign(ident(tan(l->s)), ident(t)));
          return;
        }
        const bool dup = is_param_array(l) && shadow.count(l->a[0]->s);
        const bool keep = !(is_param_array(l) && noneed.count(l->a[0]->s));
        std::string t;
        if (dup) {
          t = fresh("t");
          out->a.push_back(assign(ident(t), tangent(rhs)));
        }
        if (keep) out->a.push_back(clone(s));
        if (dup) {
          std::vector<NodeP> ia{ident(shadow.at(l->a[0]->s))};
          for (size_t i = 1; i < l->a.size(); ++i) ia.push_back(clone(l->a[i]));
          out->a.push_back(assign(mk(NK::Index, "", ia), ident(t)));
        }
        return;
      }
      case NK::If: out->a.push_back(fwd_if(s)); return;
      case NK::For: out->a.push_back(mk(NK::For, s->s, {clone(s->a[0]), fwd_block(s->a[1])})); return;
      case NK::Block: append(out, fwd_block(s)); return;
      default: out->a.push_back(clone(s)); return;
    }
  }
  NodeP fwd_if(const NodeP& s) {
    NodeP e;
    if (s->a.size() > 2 && s->a[2]) e = s->a[2]->k == NK::If ? fwd_if(s->a[2]) : fwd_block(s->a[2]);
    return mk(NK::If, "", {clone(s->a[0]), fwd_block(s->a[1]), e});
  }
};

}  // namespace

Kernel derivative_kernel(const Kernel& K, const JVal& spec) {
  std::string mode = "Reverse";
  if (auto* m = spec.get("mode")) mode = m->s;
  if (mode.rfind("Enzyme.", 0) == 0) mode = mode.substr(7);
  if (mode != "Reverse" && mode != "Forward")
    throw PsError(PS_EKEYWORD, "∇: ad_mode must be Enzyme.Reverse or Enzyme.Forward (obtained: " + mode + ")");
  std::map<std::string, std::string> annot;
  if (auto* a = spec.get("annotations")) {
    for (auto& kv : a->obj) {
      std::string v = kv.second.s;
      if (v.rfind("Enzyme.", 0) == 0) v = v.substr(7);
      if (v != "Const" && v != "Duplicated" && v != "DuplicatedNoNeed")
        throw PsError(PS_EKEYWORD, "annotation " + v + " is not (yet) supported with high-level syntax; use the generic syntax calling directly `autodiff_deferred!`.");
      if (!K.param_index.count(kv.first))
        throw PsError(PS_EKEYWORD, "variable " + kv.first + " is not a positional argument of the kernel call " + K.def.name + ".");
      annot[kv.first] = v;
    }
  }
  std::string form = "scatter";
  if (auto* f = spec.get("form")) form = f->s;
  if (form != "scatter" && form != "gather" && form != "clear")
    throw PsError(PS_EARGUMENT, "∇: unknown derivative form '" + form + "' (scatter, gather or clear)");
  return Ad(K, mode == "Reverse", annot, form).run();
}

}  // namespace ps
