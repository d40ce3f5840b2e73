// Kernel definition: the passes that the reference runs at macro-expansion
// time, restated over the parsed kernel text.
//
//   validate_body            src/shared.jl:51-57
//   extract_onthefly_arrays! src/parallel.jl:789-819
//   apply_masks / @within    src/parallel.jl:654-680, src/FiniteDifferences.jl:397-407
//   FD macro expansion       src/FiniteDifferences.jl:54-61, 160-196, 328-396
//   handle_padding (false)   src/ParallelKernel/parallel.jl:487-557  (ixi -> ix+1, ixd -> ix)
//   insert_onthefly!         src/parallel.jl:821-859
//   memopt offsets/metadata  src/memopt.jl:57-87, 606-700, 1179-1236
#include "ps_lower.h"

#include <algorithm>
#include <functional>
#include <cstdio>
#include <sstream>

#include "ps_amd.h"

namespace ps {

static const char* kIdx[3] = {"__ix", "__iy", "__iz"};
// padding=true: the staggered index roles (ixd) stay symbolic until the statement's written array
// is known (src/ParallelKernel/parallel.jl:508-549)
static const char* kIdxD[3] = {"__ixd", "__iyd", "__izd"};

// ---------------------------------------------------------------------------
// Finite-difference macro tables. Each entry restates the index algebra of
// one macro of src/FiniteDifferences.jl with the reference's index roles:
// ixa (all), ixi (inner, = ix+1 without padding), ixd (staggered, = ix).
// ---------------------------------------------------------------------------
struct MacroDef {
  const char* name;
  const char* tmpl;
};

static const MacroDef kFD1[] = {
    {"d", "A[ixd+1] - A[ixd]"},
    {"d2", "(A[ixi+1] - A[ixi]) - (A[ixi] - A[ixi-1])"},
    {"all", "A[ixa]"},
    {"inn", "A[ixi]"},
    {"av", "(A[ixd] + A[ixd+1])*0.5"},
    {"harm", "1.0/(1.0/A[ixd] + 1.0/A[ixd+1])*2.0"},
    {"maxloc", "max(max(A[ixi-1], A[ixi+1]), A[ixi])"},
    {"minloc", "min(min(A[ixi-1], A[ixi+1]), A[ixi])"},
};

static const MacroDef kFD2[] = {
    {"d_xa", "A[ixd+1,iya] - A[ixd,iya]"},
    {"d_ya", "A[ixa,iyd+1] - A[ixa,iyd]"},
    {"d_xi", "A[ixd+1,iyi] - A[ixd,iyi]"},
    {"d_yi", "A[ixi,iyd+1] - A[ixi,iyd]"},
    {"d2_xa", "(A[ixi+1,iya] - A[ixi,iya]) - (A[ixi,iya] - A[ixi-1,iya])"},
    {"d2_ya", "(A[ixa,iyi+1] - A[ixa,iyi]) - (A[ixa,iyi] - A[ixa,iyi-1])"},
    {"d2_xi", "(A[ixi+1,iyi] - A[ixi,iyi]) - (A[ixi,iyi] - A[ixi-1,iyi])"},
    {"d2_yi", "(A[ixi,iyi+1] - A[ixi,iyi]) - (A[ixi,iyi] - A[ixi,iyi-1])"},
    {"all", "A[ixa,iya]"},
    {"inn", "A[ixi,iyi]"},
    {"inn_x", "A[ixi,iya]"},
    {"inn_y", "A[ixa,iyi]"},
    {"av", "(A[ixd,iyd] + A[ixd+1,iyd] + A[ixd,iyd+1] + A[ixd+1,iyd+1])*0.25"},
    {"av_xa", "(A[ixd,iya] + A[ixd+1,iya])*0.5"},
    {"av_ya", "(A[ixa,iyd] + A[ixa,iyd+1])*0.5"},
    {"av_xi", "(A[ixd,iyi] + A[ixd+1,iyi])*0.5"},
    {"av_yi", "(A[ixi,iyd] + A[ixi,iyd+1])*0.5"},
    {"harm", "1.0/(1.0/A[ixd,iyd] + 1.0/A[ixd+1,iyd] + 1.0/A[ixd,iyd+1] + 1.0/A[ixd+1,iyd+1])*4.0"},
    {"harm_xa", "1.0/(1.0/A[ixd,iya] + 1.0/A[ixd+1,iya])*2.0"},
    {"harm_ya", "1.0/(1.0/A[ixa,iyd] + 1.0/A[ixa,iyd+1])*2.0"},
    {"harm_xi", "1.0/(1.0/A[ixd,iyi] + 1.0/A[ixd+1,iyi])*2.0"},
    {"harm_yi", "1.0/(1.0/A[ixi,iyd] + 1.0/A[ixi,iyd+1])*2.0"},
    {"maxloc", "max(max(max(A[ixi-1,iyi], A[ixi+1,iyi]), A[ixi,iyi]), max(A[ixi,iyi-1], A[ixi,iyi+1]))"},
    {"minloc", "min(min(min(A[ixi-1,iyi], A[ixi+1,iyi]), A[ixi,iyi]), min(A[ixi,iyi-1], A[ixi,iyi+1]))"},
};

static const MacroDef kFD3[] = {
    {"d_xa", "A[ixd+1,iya,iza] - A[ixd,iya,iza]"},
    {"d_ya", "A[ixa,iyd+1,iza] - A[ixa,iyd,iza]"},
    {"d_za", "A[ixa,iya,izd+1] - A[ixa,iya,izd]"},
    {"d_xi", "A[ixd+1,iyi,izi] - A[ixd,iyi,izi]"},
    {"d_yi", "A[ixi,iyd+1,izi] - A[ixi,iyd,izi]"},
    {"d_zi", "A[ixi,iyi,izd+1] - A[ixi,iyi,izd]"},
    {"d2_xi", "(A[ixi+1,iyi,izi] - A[ixi,iyi,izi]) - (A[ixi,iyi,izi] - A[ixi-1,iyi,izi])"},
    {"d2_yi", "(A[ixi,iyi+1,izi] - A[ixi,iyi,izi]) - (A[ixi,iyi,izi] - A[ixi,iyi-1,izi])"},
    {"d2_zi", "(A[ixi,iyi,izi+1] - A[ixi,iyi,izi]) - (A[ixi,iyi,izi] - A[ixi,iyi,izi-1])"},
    {"all", "A[ixa,iya,iza]"},
    {"inn", "A[ixi,iyi,izi]"},
    {"inn_x", "A[ixi,iya,iza]"},
    {"inn_y", "A[ixa,iyi,iza]"},
    {"inn_z", "A[ixa,iya,izi]"},
    {"inn_xy", "A[ixi,iyi,iza]"},
    {"inn_xz", "A[ixi,iya,izi]"},
    {"inn_yz", "A[ixa,iyi,izi]"},
    {"av", "(A[ixd,iyd,izd] + A[ixd+1,iyd,izd] + A[ixd,iyd+1,izd] + A[ixd+1,iyd+1,izd] + "
           "A[ixd,iyd,izd+1] + A[ixd+1,iyd,izd+1] + A[ixd,iyd+1,izd+1] + A[ixd+1,iyd+1,izd+1])*0.125"},
    {"av_xa", "(A[ixd,iya,iza] + A[ixd+1,iya,iza])*0.5"},
    {"av_ya", "(A[ixa,iyd,iza] + A[ixa,iyd+1,iza])*0.5"},
    {"av_za", "(A[ixa,iya,izd] + A[ixa,iya,izd+1])*0.5"},
    {"av_xi", "(A[ixd,iyi,izi] + A[ixd+1,iyi,izi])*0.5"},
    {"av_yi", "(A[ixi,iyd,izi] + A[ixi,iyd+1,izi])*0.5"},
    {"av_zi", "(A[ixi,iyi,izd] + A[ixi,iyi,izd+1])*0.5"},
    {"av_xya", "(A[ixd,iyd,iza] + A[ixd+1,iyd,iza] + A[ixd,iyd+1,iza] + A[ixd+1,iyd+1,iza])*0.25"},
    {"av_xza", "(A[ixd,iya,izd] + A[ixd+1,iya,izd] + A[ixd,iya,izd+1] + A[ixd+1,iya,izd+1])*0.25"},
    {"av_yza", "(A[ixa,iyd,izd] + A[ixa,iyd+1,izd] + A[ixa,iyd,izd+1] + A[ixa,iyd+1,izd+1])*0.25"},
    {"av_xyi", "(A[ixd,iyd,izi] + A[ixd+1,iyd,izi] + A[ixd,iyd+1,izi] + A[ixd+1,iyd+1,izi])*0.25"},
    {"av_xzi", "(A[ixd,iyi,izd] + A[ixd+1,iyi,izd] + A[ixd,iyi,izd+1] + A[ixd+1,iyi,izd+1])*0.25"},
    {"av_yzi", "(A[ixi,iyd,izd] + A[ixi,iyd+1,izd] + A[ixi,iyd,izd+1] + A[ixi,iyd+1,izd+1])*0.25"},
    {"harm", "1.0/(1.0/A[ixd,iyd,izd] + 1.0/A[ixd+1,iyd,izd] + 1.0/A[ixd,iyd+1,izd] + 1.0/A[ixd+1,iyd+1,izd] + "
             "1.0/A[ixd,iyd,izd+1] + 1.0/A[ixd+1,iyd,izd+1] + 1.0/A[ixd,iyd+1,izd+1] + 1.0/A[ixd+1,iyd+1,izd+1])*8.0"},
    {"harm_xa", "1.0/(1.0/A[ixd,iya,iza] + 1.0/A[ixd+1,iya,iza])*2.0"},
    {"harm_ya", "1.0/(1.0/A[ixa,iyd,iza] + 1.0/A[ixa,iyd+1,iza])*2.0"},
    {"harm_za", "1.0/(1.0/A[ixa,iya,izd] + 1.0/A[ixa,iya,izd+1])*2.0"},
    {"harm_xi", "1.0/(1.0/A[ixd,iyi,izi] + 1.0/A[ixd+1,iyi,izi])*2.0"},
    {"harm_yi", "1.0/(1.0/A[ixi,iyd,izi] + 1.0/A[ixi,iyd+1,izi])*2.0"},
    {"harm_zi", "1.0/(1.0/A[ixi,iyi,izd] + 1.0/A[ixi,iyi,izd+1])*2.0"},
    {"harm_xya", "1.0/(1.0/A[ixd,iyd,iza] + 1.0/A[ixd+1,iyd,iza] + 1.0/A[ixd,iyd+1,iza] + 1.0/A[ixd+1,iyd+1,iza])*4.0"},
    {"harm_xza", "1.0/(1.0/A[ixd,iya,izd] + 1.0/A[ixd+1,iya,izd] + 1.0/A[ixd,iya,izd+1] + 1.0/A[ixd+1,iya,izd+1])*4.0"},
    {"harm_yza", "1.0/(1.0/A[ixa,iyd,izd] + 1.0/A[ixa,iyd+1,izd] + 1.0/A[ixa,iyd,izd+1] + 1.0/A[ixa,iyd+1,izd+1])*4.0"},
    {"harm_xyi", "1.0/(1.0/A[ixd,iyd,izi] + 1.0/A[ixd+1,iyd,izi] + 1.0/A[ixd,iyd+1,izi] + 1.0/A[ixd+1,iyd+1,izi])*4.0"},
    {"harm_xzi", "1.0/(1.0/A[ixd,iyi,izd] + 1.0/A[ixd+1,iyi,izd] + 1.0/A[ixd,iyi,izd+1] + 1.0/A[ixd+1,iyi,izd+1])*4.0"},
    {"harm_yzi", "1.0/(1.0/A[ixi,iyd,izd] + 1.0/A[ixi,iyd+1,izd] + 1.0/A[ixi,iyd,izd+1] + 1.0/A[ixi,iyd+1,izd+1])*4.0"},
    {"maxloc", "max(max(max(max(A[ixi-1,iyi,izi], A[ixi+1,iyi,izi]), A[ixi,iyi,izi]), max(A[ixi,iyi-1,izi], A[ixi,iyi+1,izi])), "
               "max(A[ixi,iyi,izi-1], A[ixi,iyi,izi+1]))"},
    {"minloc", "min(min(min(min(A[ixi-1,iyi,izi], A[ixi+1,iyi,izi]), A[ixi,iyi,izi]), min(A[ixi,iyi-1,izi], A[ixi,iyi+1,izi])), "
               "min(A[ixi,iyi,izi-1], A[ixi,iyi,izi+1]))"},
};

// @within(macroname, A): per dimension 'a' (firstindex <= ia <= lastindex) or
// 'i' (firstindex < ii < lastindex), src/FiniteDifferences.jl:63-70,184-196,397-407.
struct WithinDef {
  const char* name;
  const char* roles;
};
static const WithinDef kWithin1[] = {{"all", "a"}, {"inn", "i"}};
static const WithinDef kWithin2[] = {{"all", "aa"}, {"inn", "ii"}, {"inn_x", "ia"}, {"inn_y", "ai"}};
static const WithinDef kWithin3[] = {{"all", "aaa"}, {"inn", "iii"}, {"inn_x", "iaa"}, {"inn_y", "aia"},
                                     {"inn_z", "aai"}, {"inn_xy", "iia"}, {"inn_xz", "iai"}, {"inn_yz", "aii"}};

// ---------------------------------------------------------------------------
// tree helpers
// ---------------------------------------------------------------------------
template <class F>
static void walk(const NodeP& n, F&& f) {
  if (!n) return;
  f(n);
  for (auto& c : n->a) walk(c, f);
}

template <class F>
static NodeP rewrite(const NodeP& n, F&& f) {
  if (!n) return n;
  NodeP r = f(n);
  if (r != n) return r;
  for (auto& c : n->a) c = rewrite(c, f);
  return n;
}

static bool refers_to(const NodeP& n, const std::string& name) {
  bool hit = false;
  walk(n, [&](const NodeP& x) {
    if ((x->k == NK::Ident && x->s == name) || (x->k == NK::Macro && !x->a.empty() && x->a[0]->k == NK::Ident && x->a[0]->s == name))
      hit = true;
  });
  return hit;
}

// Returns a substituted copy; `n` itself (e.g. a cached macro template) is left untouched.
static NodeP subst_idents(const NodeP& n, const std::map<std::string, NodeP>& m) {
  return rewrite(clone(n), [&](const NodeP& x) -> NodeP {
    if (x->k == NK::Ident) {
      auto it = m.find(x->s);
      if (it != m.end()) return clone(it->second);
    }
    return x;
  });
}

// Affine form of an index expression: var (0..2 index, 3 = `end`, -1 none) + c.
struct Aff {
  bool ok = false;
  int var = -1;
  long long c = 0;
};

static Aff affine(const NodeP& e) {
  Aff r;
  if (!e) return r;
  switch (e->k) {
    case NK::Num:
      if (e->is_int) {
        r.ok = true;
        r.c = e->ival;
      }
      return r;
    case NK::Ident:
      for (int d = 0; d < 3; ++d)
        if (e->s == kIdx[d]) {
          r.ok = true;
          r.var = d;
        }
      return r;
    case NK::EndKw:
      r.ok = true;
      r.var = 3;
      return r;
    case NK::Binary: {
      if (e->s != "+" && e->s != "-") return r;
      Aff a = affine(e->a[0]), b = affine(e->a[1]);
      if (!a.ok || !b.ok) return r;
      if (e->s == "+") {
        if (a.var >= 0 && b.var >= 0) return r;
        r.ok = true;
        r.var = a.var >= 0 ? a.var : b.var;
        r.c = a.c + b.c;
        return r;
      }
      if (b.var >= 0) return r;
      r.ok = true;
      r.var = a.var;
      r.c = a.c - b.c;
      return r;
    }
    case NK::Unary:
      if (e->s == "+") return affine(e->a[0]);
      if (e->s == "-") {
        Aff a = affine(e->a[0]);
        if (a.ok && a.var < 0) {
          r.ok = true;
          r.c = -a.c;
        }
      }
      return r;
    default:
      return r;
  }
}

static NodeP aff_node(const Aff& a) {
  if (a.var < 0) return mknum_int(a.c);
  NodeP v = a.var == 3 ? mk(NK::EndKw, "end") : mk(NK::Ident, kIdx[a.var]);
  if (a.c == 0) return v;
  if (a.c > 0) return mk(NK::Binary, "+", {v, mknum_int(a.c)});
  return mk(NK::Binary, "-", {v, mknum_int(-a.c)});
}

// Canonicalise the index expressions of every array access to var +/- c.
static NodeP canon_indices(const NodeP& n) {
  return rewrite(n, [&](const NodeP& x) -> NodeP {
    if (x->k == NK::Index) {
      for (size_t i = 1; i < x->a.size(); ++i) {
        x->a[i] = canon_indices(x->a[i]);
        Aff a = affine(x->a[i]);
        if (a.ok) x->a[i] = aff_node(a);
      }
      x->a[0] = canon_indices(x->a[0]);
      return x;  // children handled
    }
    return x;
  });
}

NodeP canonical_indices(const NodeP& n) { return canon_indices(n); }

static std::vector<std::pair<int, int>> parse_optrange_tuple(const NodeP& t) {
  // (a:b, c:d, e:f)
  std::vector<std::pair<int, int>> r;
  if (!t || t->k != NK::Tuple) throw PsError(PS_EKEYWORD, "optranges entries must be tuples of ranges");
  for (auto& x : t->a) {
    if (x->k != NK::Range) throw PsError(PS_EKEYWORD, "optranges entries must be tuples of ranges");
    Aff lo = affine(x->a[0]), hi = affine(x->a[1]);
    if (!lo.ok || !hi.ok || lo.var >= 0 || hi.var >= 0) throw PsError(PS_EKEYWORD, "optranges bounds must be integer literals");
    r.emplace_back((int)lo.c, (int)hi.c);
  }
  return r;
}

// ---------------------------------------------------------------------------
// options
// ---------------------------------------------------------------------------
static void read_options(Options& o, const JVal& j) {
  if (j.t != JVal::Obj) throw PsError(PS_EARGUMENT, "kernel options must be a JSON object");
  for (auto& kv : j.obj) {
    const std::string& k = kv.first;
    const JVal& v = kv.second;
    if (v.t == JVal::Null) continue;
    if (k == "kind") o.kind = v.s;
    else if (k == "indices") {
      o.indices.clear();
      for (auto& x : v.arr) o.indices.push_back(x.s);
    } else if (k == "ndims") o.ndims = (int)v.n;
    else if (k == "numbertype") o.numbertype = v.s;
    else if (k == "inbounds") o.inbounds = v.b;
    else if (k == "padding") o.padding = v.b;
    else if (k == "memopt") o.memopt = v.b;
    else if (k == "loopsize") o.loopsize = (int)v.n;
    else if (k == "loopdim") o.loopdim = (int)v.n;
    else if (k == "optvars") {
      o.optvars.clear();
      if (v.t == JVal::Str) o.optvars.push_back(v.s);
      else
        for (auto& x : v.arr) o.optvars.push_back(x.s);
    } else if (k == "optranges") {
      for (auto& e : v.obj) {
        std::vector<std::pair<int, int>> r;
        for (auto& d : e.second.arr) r.emplace_back((int)d.arr.at(0).n, (int)d.arr.at(1).n);
        o.optranges[e.first] = r;
      }
    } else if (k == "nthreads_max_memopt") o.nthreads_max_memopt = (int)v.n;
    else if (k == "nthreads_x_max_memopt") o.nthreads_x_max_memopt = (int)v.n;
    else if (k == "fp_contract") o.fp_contract = v.b;
    else if (k == "fastmath") o.fastmath = v.b;
    else if (k == "prefetch") {
      o.prefetch = (int)v.n;
      if (o.prefetch < 1 || o.prefetch > 4) throw PsError(PS_EKEYWORD, "prefetch must be 1, 2, 3 or 4");
    } else if (k == "cells_x") {
      o.cells_x = (int)v.n;
      if (!(o.cells_x == 0 || o.cells_x == 1 || o.cells_x == 2 || o.cells_x == 4 || o.cells_x == 8))
        throw PsError(PS_EKEYWORD, "cells_x must be 0 (automatic), 1, 2, 4 or 8");
    } else if (k == "cells_z") {
      o.cells_z = (int)v.n;
      if (!(o.cells_z == 0 || o.cells_z == 1 || o.cells_z == 2 || o.cells_z == 4 || o.cells_z == 8))
        throw PsError(PS_EKEYWORD, "cells_z must be 0 (automatic), 1, 2, 4 or 8");
    } else if (k == "nthreads_x_groups") {
      o.nthreads_x_groups = (int)v.n;
      if (!(o.nthreads_x_groups == 0 || (o.nthreads_x_groups >= 1 && o.nthreads_x_groups <= 1024 &&
                                         (o.nthreads_x_groups & (o.nthreads_x_groups - 1)) == 0)))
        throw PsError(PS_EKEYWORD, "nthreads_x_groups must be 0 (automatic) or a power of 2 up to 1024");
    } else if (k == "addr32") o.addr32 = v.b;
    else if (k == "lds") o.lds = v.b;
    else if (k == "xcd_remap") o.xcd_remap = v.b;
    else if (k == "xcd_spread") o.xcd_spread = v.b;
    else if (k == "nt_store") o.nt_store = v.b ? 1 : 0;
    else if (k == "tuple_params") {
      o.tuple_params.clear();
      for (auto& e : v.obj) {
        std::vector<std::string> names;
        for (auto& x : e.second.arr) names.push_back(x.s);
        if (names.empty()) throw PsError(PS_EARGUMENT, "tuple argument '" + e.first + "' has no components");
        o.tuple_params.emplace_back(e.first, names);
      }
    }
    else if (k == "useshmemhalos" || k == "optimize_halo_read" || k == "nonconst_metadata" || k == "metadata_module" || k == "N") {
      // accepted for interface compatibility; they only steer the reference's shared-memory path
    } else {
      throw PsError(PS_EKEYWORD, "unknown keyword argument '" + k + "' in kernel definition");
    }
  }
}

static void apply_prefix_kwargs(Options& o, const KernelDef& kd) {
  for (auto& kw : kd.prefix_kwargs) {
    const std::string& k = kw.first;
    const NodeP& v = kw.second;
    auto as_bool = [&]() {
      if (v->k == NK::Ident && (v->s == "true" || v->s == "false")) return v->s == "true";
      throw PsError(PS_EKEYWORD, "keyword '" + k + "' expects true or false");
    };
    auto as_int = [&]() {
      Aff a = affine(v);
      if (!a.ok || a.var >= 0) throw PsError(PS_EKEYWORD, "keyword '" + k + "' expects an integer literal");
      return (int)a.c;
    };
    if (k == "memopt") o.memopt = as_bool();
    else if (k == "fastmath") o.fastmath = as_bool();
    else if (k == "inbounds") o.inbounds = as_bool();
    else if (k == "padding") o.padding = as_bool();
    else if (k == "ndims") o.ndims = as_int();
    else if (k == "loopsize") o.loopsize = as_int();
    else if (k == "loopdim") o.loopdim = as_int();
    else if (k == "optvars") {
      o.optvars.clear();
      if (v->k == NK::Ident) o.optvars.push_back(v->s);
      else if (v->k == NK::Tuple)
        for (auto& x : v->a) o.optvars.push_back(x->s);
    } else if (k == "optranges") {
      if (v->k != NK::Tuple) throw PsError(PS_EKEYWORD, "optranges must be a named tuple");
      for (auto& e : v->a) {
        if (e->k != NK::Kw) throw PsError(PS_EKEYWORD, "optranges must be a named tuple");
        o.optranges[e->s] = parse_optrange_tuple(e->a[0]);
      }
    } else if (k == "useshmemhalos" || k == "optimize_halo_read") {
    } else {
      throw PsError(PS_EKEYWORD, "unknown keyword argument '" + k + "' in kernel definition");
    }
  }
}

// ---------------------------------------------------------------------------
// macro expansion
// ---------------------------------------------------------------------------
static const MacroDef* find_macro(int nd, const std::string& name) {
  const MacroDef* tab = nd == 1 ? kFD1 : nd == 2 ? kFD2 : kFD3;
  size_t n = nd == 1 ? sizeof(kFD1) / sizeof(MacroDef) : nd == 2 ? sizeof(kFD2) / sizeof(MacroDef) : sizeof(kFD3) / sizeof(MacroDef);
  for (size_t i = 0; i < n; ++i)
    if (name == tab[i].name) return &tab[i];
  return nullptr;
}

static std::map<std::string, NodeP>& template_cache() {
  static std::map<std::string, NodeP> c;
  return c;
}

static NodeP expand_macro(int nd, const std::string& name, const std::string& arr, bool padding = false) {
  const MacroDef* m = find_macro(nd, name);
  if (!m) {
    throw PsError(PS_EPLUGIN, "macro @" + name + " is not defined for " + std::to_string(nd) +
                                  "-D kernels (FiniteDifferences" + std::to_string(nd) + "D)");
  }
  std::string key = std::to_string(nd) + ":" + name;
  auto& cache = template_cache();
  if (!cache.count(key)) cache[key] = parse_expression(m->tmpl);
  std::map<std::string, NodeP> sub;
  const char* dn[3] = {"x", "y", "z"};
  for (int d = 0; d < 3; ++d) {
    NodeP v = mk(NK::Ident, kIdx[d]);
    sub[std::string("i") + dn[d] + "a"] = v;
    if (padding) {  // substitute_indices_inn / _dir with padding: ixi -> ix, ixd resolved per statement
      sub[std::string("i") + dn[d] + "d"] = mk(NK::Ident, kIdxD[d]);
      sub[std::string("i") + dn[d] + "i"] = mk(NK::Ident, kIdx[d]);
    } else {
      sub[std::string("i") + dn[d] + "d"] = v;  // padding=false: ixd -> ix
      sub[std::string("i") + dn[d] + "i"] = mk(NK::Binary, "+", {mk(NK::Ident, kIdx[d]), mknum_int(1)});  // ixi -> ix+1
    }
  }
  sub["A"] = mk(NK::Ident, arr);
  return subst_idents(cache[key], sub);
}

static NodeP expand_all_macros(int nd, const NodeP& n, bool padding = false) {
  return rewrite(n, [&](const NodeP& x) -> NodeP {
    if (x->k == NK::Macro) {
      if (x->a.size() != 1 || x->a[0]->k != NK::Ident)
        throw PsError(PS_EARGUMENT, "finite-difference macro @" + x->s + " expects a single array argument");
      return expand_macro(nd, x->s, x->a[0]->s, padding);
    }
    return x;
  });
}

static NodeP within_mask(int nd, const std::string& mname, const std::string& arr, bool padding = false) {
  const WithinDef* tab = nd == 1 ? kWithin1 : nd == 2 ? kWithin2 : kWithin3;
  size_t n = nd == 1 ? 2 : nd == 2 ? 4 : 8;
  const WithinDef* w = nullptr;
  for (size_t i = 0; i < n; ++i)
    if (mname == tab[i].name) w = &tab[i];
  if (!w)
    throw PsError(PS_EARGUMENT, "unkown macroname: @" + mname +
                                    ". Only @all and the @inn family can be assigned to in @parallel kernels.");
  NodeP cond;
  for (int d = 0; d < nd; ++d) {
    NodeP sz = mk(NK::Call, "size", {mk(NK::Ident, arr), mknum_int(d + 1)});
    NodeP v = mk(NK::Ident, kIdx[d]);
    NodeP lo, hi;
    if (padding) {
      // firstindex(A,d) = A.indices[d][1], lastindex(A,d) = A.indices[d][end]
      // (src/ParallelKernel/kernel_language.jl:605-617): the view's box in its parent
      NodeP f = mk(NK::Call, "__first", {mk(NK::Ident, arr), mknum_int(d + 1)});
      NodeP l = mk(NK::Call, "__last", {mk(NK::Ident, arr), mknum_int(d + 1)});
      const char* op = w->roles[d] == 'a' ? "<=" : "<";
      lo = mk(NK::Cmp, "", {f, clone(v)});
      lo->ops = {op};
      hi = mk(NK::Cmp, "", {clone(v), l});
      hi->ops = {op};
    } else if (w->roles[d] == 'a') {  // 1 <= ix <= size(A,d)
      lo = mk(NK::Cmp, "", {mknum_int(1), clone(v)});
      lo->ops = {"<="};
      hi = mk(NK::Cmp, "", {clone(v), sz});
      hi->ops = {"<="};
    } else {  // 1 < ix+1 < size(A,d)   <=>   1 <= ix <= size(A,d)-2
      lo = mk(NK::Cmp, "", {mknum_int(1), clone(v)});
      lo->ops = {"<="};
      hi = mk(NK::Cmp, "", {clone(v), mk(NK::Binary, "-", {sz, mknum_int(2)})});
      hi->ops = {"<="};
    }
    NodeP c = mk(NK::And, "&&", {lo, hi});
    cond = cond ? mk(NK::And, "&&", {cond, c}) : c;
  }
  return cond;
}

// Substitutes on-the-fly arrays: A[i,j,k] -> expr_A with the parallel indices
// replaced by (i,j,k) (src/parallel.jl:821-859, src/shared.jl:59-76).
static NodeP insert_onthefly(const NodeP& n, const std::map<std::string, NodeP>& otf, int nd, int depth = 0) {
  if (depth > 32) throw PsError(PS_EARGUMENT, "on-the-fly arrays are nested too deeply (cyclic definition?)");
  return rewrite(n, [&](const NodeP& x) -> NodeP {
    if (x->k == NK::Index && x->a[0]->k == NK::Ident) {
      auto it = otf.find(x->a[0]->s);
      if (it == otf.end()) return x;
      if ((int)x->a.size() - 1 != nd)
        throw PsError(PS_EARGUMENT, "unsupported kernel statements in @parallel kernel definition: wrong number of indices in " +
                                        x->a[0]->s + " (expected " + std::to_string(nd) + " indices).");
      std::map<std::string, NodeP> sub;
      for (int d = 0; d < nd; ++d) sub[kIdx[d]] = insert_onthefly(x->a[1 + d], otf, nd, depth + 1);
      NodeP e = subst_idents(it->second, sub);
      return insert_onthefly(e, otf, nd, depth + 1);
    }
    return x;
  });
}

// padding=true: the local index roles of an on-the-fly array read at index expression e in
// dimension d (determine_local_index_dir, src/parallel.jl:829-835): its ix is e, and its ixd is e
// with every __dirf(d, n) raised to __dirf(d, n+1), every ixd turned into __dirf(d, 2) and every
// ix into ixd. __dirf(d, n) stands for ix - n÷2 until the statement's written array is known.
static NodeP local_index_dir(const NodeP& e, int d) {
  NodeP r = rewrite(clone(e), [&](const NodeP& x) -> NodeP {
    if (x->k == NK::Call && x->s == "__dirf" && x->a.size() == 2 && x->a[0]->ival == d)
      return mk(NK::Call, "__dirf", {mknum_int(d), mknum_int(x->a[1]->ival + 1)});
    return x;
  });
  r = rewrite(r, [&](const NodeP& x) -> NodeP {
    if (x->k == NK::Ident && x->s == kIdxD[d]) return mk(NK::Call, "__dirf", {mknum_int(d), mknum_int(2)});
    return x;
  });
  return rewrite(r, [&](const NodeP& x) -> NodeP {
    if (x->k == NK::Ident && x->s == kIdx[d]) return mk(NK::Ident, kIdxD[d]);
    return x;
  });
}

static NodeP insert_onthefly_padded(const NodeP& n, const std::map<std::string, NodeP>& otf, int nd, int depth = 0) {
  if (depth > 32) throw PsError(PS_EARGUMENT, "on-the-fly arrays are nested too deeply (cyclic definition?)");
  return rewrite(n, [&](const NodeP& x) -> NodeP {
    if (x->k == NK::Index && x->a[0]->k == NK::Ident) {
      auto it = otf.find(x->a[0]->s);
      if (it == otf.end()) return x;
      if ((int)x->a.size() - 1 != nd)
        throw PsError(PS_EARGUMENT, "unsupported kernel statements in @parallel kernel definition: wrong number of indices in " +
                                        x->a[0]->s + " (expected " + std::to_string(nd) + " indices).");
      std::map<std::string, NodeP> sub;
      for (int d = 0; d < nd; ++d) {
        NodeP e = insert_onthefly_padded(x->a[1 + d], otf, nd, depth + 1);
        sub[kIdxD[d]] = local_index_dir(e, d);
        sub[kIdx[d]] = e;
      }
      NodeP e = subst_idents(it->second, sub);
      return insert_onthefly_padded(e, otf, nd, depth + 1);
    }
    return x;
  });
}

// padding=true: resolves the staggered index roles of one statement whose written array is B
// (substitute_indices_dir, src/ParallelKernel/parallel.jl:514-548): in a read of array A, ixd is
// ix and __dirf(d, n) is ix - n÷2, each minus (size(B.parent,d) > size(A.parent,d)) when A is not
// B. The comparison stays a __pgt(B, A, d) call, recorded in K.pgt_terms, until a launch knows the
// parents' sizes.
static NodeP resolve_dir_indices(Kernel& K, const NodeP& rhs, const std::string& B) {
  auto pgt = [&](const std::string& A, int d) -> NodeP {
    std::array<int, 3> t{K.param_index.at(B), K.param_index.at(A), d + 1};
    if (std::find(K.pgt_terms.begin(), K.pgt_terms.end(), t) == K.pgt_terms.end()) K.pgt_terms.push_back(t);
    return mk(NK::Call, "__pgt", {mk(NK::Ident, B), mk(NK::Ident, A), mknum_int(d + 1)});
  };
  NodeP r = rewrite(rhs, [&](const NodeP& x) -> NodeP {
    if (x->k != NK::Index || x->a[0]->k != NK::Ident || !K.param_index.count(x->a[0]->s)) return x;
    const std::string A = x->a[0]->s;
    for (size_t i = 1; i < x->a.size(); ++i) {
      x->a[i] = rewrite(x->a[i], [&](const NodeP& e) -> NodeP {
        for (int d = 0; d < 3; ++d) {
          NodeP base;
          if (e->k == NK::Ident && e->s == kIdxD[d]) base = mk(NK::Ident, kIdx[d]);
          else if (e->k == NK::Call && e->s == "__dirf" && e->a[0]->ival == d)
            base = mk(NK::Binary, "-", {mk(NK::Ident, kIdx[d]), mknum_int(e->a[1]->ival / 2)});
          if (!base) continue;
          return A == B ? base : mk(NK::Binary, "-", {base, pgt(A, d)});
        }
        return e;
      });
    }
    return x;
  });
  walk(r, [&](const NodeP& x) {
    if ((x->k == NK::Ident && (x->s == kIdxD[0] || x->s == kIdxD[1] || x->s == kIdxD[2])) || (x->k == NK::Call && x->s == "__dirf"))
      throw PsError(PS_EINTERNAL, "padding: a staggered index remains outside an array access");
  });
  return r;
}

// padding=true in @parallel_indices kernels (substitute_view_accesses,
// src/ParallelKernel/parallel.jl:572-581, is_access in src/ParallelKernel/shared.jl:328-330): a read
// or write of an array whose every index involves a parallel index addresses the array's parent
// (`end` there is the parent's extent); any other access addresses the view itself, which here
// means its indices are shifted by the view's first parent index.
static NodeP padded_view_accesses(const Kernel& K, const NodeP& body) {
  auto involves_index = [&](const NodeP& e) {
    bool hit = false;
    walk(e, [&](const NodeP& y) {
      for (int d = 0; d < K.nidx; ++d)
        if (y->k == NK::Ident && y->s == kIdx[d]) hit = true;
    });
    return hit;
  };
  return rewrite(body, [&](const NodeP& x) -> NodeP {
    if (x->k != NK::Index || x->a[0]->k != NK::Ident || !K.param_index.count(x->a[0]->s)) return x;
    const std::string A = x->a[0]->s;
    const size_t ni = x->a.size() - 1;
    bool promoted = (int)ni == K.nidx;
    for (size_t i = 1; i < x->a.size() && promoted; ++i) promoted = involves_index(x->a[i]);
    for (size_t i = 1; i < x->a.size(); ++i) {
      const int d = (int)i;
      NodeP e = padded_view_accesses(K, x->a[i]);
      e = rewrite(e, [&](const NodeP& y) -> NodeP {
        if (y->k == NK::EndKw) return mk(NK::Call, promoted ? "__psize" : "size", {mk(NK::Ident, A), mknum_int(d)});
        return y;
      });
      if (!promoted && ni > 1)
        e = mk(NK::Binary, "-", {mk(NK::Binary, "+", {e, mk(NK::Call, "__first", {mk(NK::Ident, A), mknum_int(d)})}), mknum_int(1)});
      x->a[i] = e;
    }
    return x;
  });
}

static void memopt_analysis(Kernel& K);

Kernel specialize_padding(const Kernel& K0, const std::string& bits) {
  Kernel K = K0;
  if (bits.size() != K.pgt_terms.size()) throw PsError(PS_EINTERNAL, "specialize_padding: wrong number of bits");
  K.body = rewrite(clone(K0.body), [&](const NodeP& x) -> NodeP {
    if (x->k != NK::Call || x->s != "__pgt") return x;
    std::array<int, 3> t{K.param_index.at(x->a[0]->s), K.param_index.at(x->a[1]->s), (int)x->a[2]->ival};
    size_t i = std::find(K.pgt_terms.begin(), K.pgt_terms.end(), t) - K.pgt_terms.begin();
    if (i >= bits.size()) throw PsError(PS_EINTERNAL, "specialize_padding: unknown parent-size comparison");
    return mknum_int(bits[i] == '1' ? 1 : 0);
  });
  K.body = canon_indices(K.body);
  K.pgt_terms.clear();
  if (K.opt.memopt) {
    K.mm = MemoptMeta();
    memopt_analysis(K);
  }
  return K;
}

// ---------------------------------------------------------------------------
// define
// ---------------------------------------------------------------------------
// The argument an indexing expression reads: A[I], field(A, i...)[I] (CellArrays) or A[I][i...].
static std::string indexed_param(const NodeP& x, const std::map<std::string, int>& params) {
  const NodeP& b = x->a[0];
  if (b->k == NK::Ident) return params.count(b->s) ? b->s : "";
  if (b->k == NK::Call && b->s == "field" && !b->a.empty() && b->a[0]->k == NK::Ident) return params.count(b->a[0]->s) ? b->a[0]->s : "";
  if (b->k == NK::Index) return indexed_param(b, params);
  return "";
}

static void collect_access(const NodeP& n, std::set<std::string>& accessed, std::set<std::string>& written,
                           const std::map<std::string, int>& params) {
  walk(n, [&](const NodeP& x) {
    if (x->k == NK::Index) {
      std::string p = indexed_param(x, params);
      if (!p.empty()) accessed.insert(p);
    }
    if ((x->k == NK::Assign || (x->k == NK::Macro && x->s == "ps_atomic_add")) && !x->a.empty() && x->a[0]->k == NK::Index) {
      std::string p = indexed_param(x->a[0], params);
      if (!p.empty()) written.insert(p);
    }
  });
}

void refresh_kernel(Kernel& K) {
  K.param_index.clear();
  for (size_t i = 0; i < K.def.params.size(); ++i) K.param_index[K.def.params[i].name] = (int)i;
  K.accessed.clear();
  K.written.clear();
  collect_access(K.body, K.accessed, K.written, K.param_index);
}

static void memopt_analysis(Kernel& K) {
  MemoptMeta& mm = K.mm;
  mm.enabled = true;
  int nd = K.nidx;
  if (nd != 2 && nd != 3)
    throw PsError(PS_EINCOHERENT,
                  "incoherent arguments memopt in @parallel[_indices] <kernel>: optimization can only be applied in 2-D and 3-D @parallel kernels and @parallel_indices kernels.");
  int loopdim = K.opt.loopdim ? K.opt.loopdim : nd;
  if (loopdim != nd)
    throw PsError(PS_EINCOHERENT,
                  "incoherent arguments memopt in @parallel[_indices] <kernel>: two-index kernels require `loopdim=2` and three-index kernels require `loopdim=3`.");
  mm.loopdim = loopdim;
  mm.loopsize = K.opt.loopsize;
  if (mm.loopsize < 1) throw PsError(PS_EKEYWORD, "loopsize must be >= 1");
  // read-only arrays in access order
  std::vector<std::string> readonly;
  walk(K.body, [&](const NodeP& x) {
    if (x->k == NK::Index && x->a[0]->k == NK::Ident && K.param_index.count(x->a[0]->s) && !K.written.count(x->a[0]->s) &&
        std::find(readonly.begin(), readonly.end(), x->a[0]->s) == readonly.end())
      readonly.push_back(x->a[0]->s);
  });
  std::vector<std::string> optvars = K.opt.optvars;
  if (optvars.empty()) optvars = readonly;
  for (auto& A : optvars)
    if (std::find(readonly.begin(), readonly.end(), A) == readonly.end())
      throw PsError(PS_EINCOHERENT, "incoherent argument memopt in @parallel[_indices] <kernel>: optimization can only be applied to arrays that are only read within the kernel (not applicable to: " + A + ").");
  // offsets of stencil accesses (of every read-only array)
  for (auto& A : readonly) mm.offsets[A];
  walk(K.body, [&](const NodeP& x) {
    if (!(x->k == NK::Index && x->a[0]->k == NK::Ident && mm.offsets.count(x->a[0]->s))) return;
    if ((int)x->a.size() - 1 < nd) return;
    int off[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d) {
      Aff a = affine(x->a[1 + d]);
      if (!a.ok || a.var != d) return;  // not a stencil access (e.g. A[1,iy])
      off[d] = (int)a.c;
    }
    for (size_t d = nd; d + 1 < x->a.size(); ++d) {  // extra indices must be literal 1
      Aff a = affine(x->a[1 + d]);
      if (!a.ok || a.var >= 0 || a.c != 1) return;
    }
    std::pair<int, int> col = nd == 3 ? std::make_pair(off[0], off[1]) : std::make_pair(off[0], 0);
    int oz = off[nd - 1];
    mm.offsets[x->a[0]->s][col][oz] += 1;
  });
  // remove optvars that are accessed at a single point only (no reuse along the loop
  // dimension) unless their optrange is given explicitly (src/memopt.jl:83, 1012-1025)
  for (auto& A : optvars) {
    auto& offs = mm.offsets[A];
    bool multi = false;
    int npoints = 0;
    for (auto& c : offs) {
      npoints += (int)c.second.size();
      if (c.second.size() > 1) multi = true;
    }
    if (npoints == 0) continue;
    if (npoints > 1 || K.opt.optranges.count(A)) {
      (void)multi;
      mm.optvars.push_back(A);
    }
  }
  if (mm.optvars.empty())
    throw PsError(PS_EINCOHERENT,
                  "incoherent argument memopt in @parallel[_indices] <kernel>: optimization can only be applied if there is at least one array that is read-only within the kernel (and accessed with a multi-point stencil). Set memopt=false for this kernel.");
  for (auto it = mm.offsets.begin(); it != mm.offsets.end();) {
    if (std::find(mm.optvars.begin(), mm.optvars.end(), it->first) == mm.optvars.end()) {
      if (!it->second.empty()) mm.ro_offsets[it->first] = it->second;
      it = mm.offsets.erase(it);
    } else {
      ++it;
    }
  }
  for (auto& A : mm.optvars) {
    std::vector<std::pair<int, int>> sr(3, {0, 0});
    bool first = true;
    for (auto& c : mm.offsets[A]) {
      for (auto& z : c.second) {
        int o[3] = {c.first.first, nd == 3 ? c.first.second : z.first, nd == 3 ? z.first : 0};
        for (int d = 0; d < 3; ++d) {
          if (first) sr[d] = {o[d], o[d]};
          else {
            sr[d].first = std::min(sr[d].first, o[d]);
            sr[d].second = std::max(sr[d].second, o[d]);
          }
        }
        first = false;
      }
    }
    if (K.opt.optranges.count(A)) {
      auto r = K.opt.optranges.at(A);
      while (r.size() < 3) r.push_back({0, 0});
      mm.optranges[A] = r;
      mm.stencilranges[A] = r;
    } else {
      mm.stencilranges[A] = sr;
    }
    bool xyhalo = nd == 3 && (mm.stencilranges[A][0].second > mm.stencilranges[A][0].first ||
                              mm.stencilranges[A][1].second > mm.stencilranges[A][1].first);
    if (xyhalo) mm.shmem_optvars.push_back(A);
    mm.use_shmemhalos[A] = nd == 3;
  }
}

// @∀ x ∈ X statement (src/ParallelKernel/kernel_language.jl:197-237, 543-566): the statement once per
// element of X, with x replaced by the element: as a name (also as a field name, `C.x`), a
// number (X = a:b with literal bounds) or, for a tuple of variables, element-wise.
void analyse_memopt(Kernel& K) {
  K.mm = MemoptMeta{};
  memopt_analysis(K);
}

static NodeP expand_forall(const NodeP& n);
static void expand_forall_into(const NodeP& st, NodeP& out) {
  if (st->k != NK::Forall) {
    out->a.push_back(expand_forall(st));
    return;
  }
  const NodeP& var = st->a[0];
  const NodeP& coll = st->a[1];
  std::vector<NodeP> elems;
  if (coll->k == NK::Range) {
    Aff a = affine(coll->a[0]), b = affine(coll->a[1]);
    if (!a.ok || a.var >= 0 || !b.ok || b.var >= 0) throw PsError(PS_EARGUMENT, "@∀: a range must have literal bounds");
    for (long long v = a.c; v <= b.c; ++v) elems.push_back(mknum_int(v));
  } else if (coll->k == NK::Tuple) {
    elems = coll->a;
  } else if (coll->k == NK::Ident) {
    elems = {coll};
  } else {
    throw PsError(PS_EARGUMENT, "@∀: the set must be a tuple of names or a range a:b");
  }
  std::vector<std::string> vars;
  if (var->k == NK::Ident) vars.push_back(var->s);
  else
    for (auto& v : var->a) {
      if (v->k != NK::Ident) throw PsError(PS_EARGUMENT, "@∀: variables must be names");
      vars.push_back(v->s);
    }
  for (auto& e : elems) {
    std::vector<NodeP> vals;
    if (vars.size() == 1) vals.push_back(e);
    else {
      if (e->k != NK::Tuple || e->a.size() != vars.size())
        throw PsError(PS_EARGUMENT, "@∀: each element must be a tuple of " + std::to_string(vars.size()) + " values");
      vals = e->a;
    }
    std::map<std::string, NodeP> sub;
    for (size_t i = 0; i < vars.size(); ++i) sub[vars[i]] = vals[i];
    NodeP b = subst_idents(st->a[2], sub);
    walk(b, [&](const NodeP& x) {  // field names: C.i -> C.x
      if (x->k == NK::Dot) {
        auto it = sub.find(x->s);
        if (it != sub.end()) {
          if (it->second->k != NK::Ident) throw PsError(PS_EARGUMENT, "@∀: field ." + x->s + " needs a name, not a number");
          x->s = it->second->s;
        }
      }
    });
    expand_forall_into(b, out);
  }
}
static NodeP expand_forall(const NodeP& n) {
  if (!n) return n;
  if (n->k == NK::Block) {
    auto out = mk(NK::Block);
    for (auto& st : n->a) expand_forall_into(st, out);
    return out;
  }
  auto c = std::make_shared<Node>(*n);
  for (auto& x : c->a) x = expand_forall(x);
  return c;
}

// Tuple and NamedTuple arguments: C.x and C[k] (k a literal) name the component parameter.
static NodeP expand_tuple_args(const NodeP& body, const Options& o) {
  std::map<std::string, const std::vector<std::string>*> tp;
  for (auto& t : o.tuple_params) tp[t.first] = &t.second;
  if (tp.empty()) return body;
  NodeP b = rewrite(clone(body), [&](const NodeP& x) -> NodeP {
    if (x->k == NK::Dot && x->a[0]->k == NK::Ident && tp.count(x->a[0]->s)) {
      const auto& names = *tp[x->a[0]->s];
      if (std::find(names.begin(), names.end(), x->s) == names.end())
        throw PsError(PS_EARGUMENT, "tuple argument '" + x->a[0]->s + "' has no component ." + x->s);
      return mk(NK::Ident, x->a[0]->s + "." + x->s);
    }
    if (x->k == NK::Index && x->a[0]->k == NK::Ident && tp.count(x->a[0]->s) && x->a.size() == 2 && x->a[1]->k == NK::Num && x->a[1]->is_int) {
      const auto& names = *tp[x->a[0]->s];
      long long k = x->a[1]->ival;
      if (k < 1 || k > (long long)names.size())
        throw PsError(PS_EBOUNDS, "BoundsError: tuple argument '" + x->a[0]->s + "' has " + std::to_string(names.size()) + " components (index " +
                                      std::to_string(k) + ")");
      return mk(NK::Ident, x->a[0]->s + "." + names[(size_t)k - 1]);
    }
    if (x->k == NK::Call && x->s == "length" && x->a.size() == 1 && x->a[0]->k == NK::Ident && tp.count(x->a[0]->s))
      return mknum_int((long long)tp[x->a[0]->s]->size());
    return x;
  });
  walk(b, [&](const NodeP& x) {
    if (x->k == NK::Ident && tp.count(x->s))
      throw PsError(PS_EARGUMENT, "tuple argument '" + x->s + "' must be used through its components (" + x->s + ".name or " + x->s + "[k])");
  });
  return b;
}

// Nested functions of @parallel_indices kernels (test/ParallelKernel/test_parallel.jl:387-433): a
// function defined at the top of the body is inlined where it is called. Called as a statement,
// its body runs in place (a final `return` ends it); called as a value, its body must be a single
// `return expr`. Parameters are replaced by the call's arguments, and variables the body assigns
// get names of their own, as in Julia's function scope.
namespace {
struct NestedFn {
  std::vector<std::string> params;
  NodeP body;
};
}  // namespace

static NodeP inline_nested_functions(const NodeP& body) {
  std::map<std::string, NestedFn> fns;
  auto rest = mk(NK::Block);
  for (auto& st : body->a) {
    if (st->k == NK::FuncDef) {
      NestedFn f;
      for (auto& q : st->a[0]->a) f.params.push_back(q->s);
      f.body = st->a[1];
      fns[st->s] = f;
    } else {
      rest->a.push_back(st);
    }
  }
  walk(rest, [&](const NodeP& x) {
    if (x->k == NK::FuncDef) throw PsError(PS_EARGUMENT, "nested functions must be defined at the top level of the kernel body (" + x->s + ")");
  });
  if (fns.empty()) return body;
  for (auto& f : fns)
    walk(f.second.body, [&](const NodeP& x) {
      if (x->k == NK::Assign && x->a[0]->k == NK::Ident &&
          std::find(f.second.params.begin(), f.second.params.end(), x->a[0]->s) != f.second.params.end())
        throw PsError(PS_EARGUMENT, "nested function " + f.first + ": assignment to its parameter " + x->a[0]->s + " is not supported");
      if (x->k == NK::FuncDef) throw PsError(PS_EARGUMENT, "nested function " + f.first + " defines a function itself");
    });
  int counter = 0;
  auto instantiate = [&](const std::string& name, const std::vector<NodeP>& args) {
    const NestedFn& f = fns.at(name);
    if (args.size() != f.params.size())
      throw PsError(PS_EARGUMENT, "MethodError: nested function " + name + " takes " + std::to_string(f.params.size()) + " arguments (" +
                                      std::to_string(args.size()) + " given)");
    std::map<std::string, NodeP> sub;
    for (size_t i = 0; i < args.size(); ++i) sub[f.params[i]] = args[i];
    ++counter;
    std::map<std::string, std::string> ren;
    walk(f.body, [&](const NodeP& x) {
      if (x->k == NK::Assign && x->a[0]->k == NK::Ident) ren[x->a[0]->s] = x->a[0]->s + "__" + name + "_" + std::to_string(counter);
      if (x->k == NK::For) ren[x->s] = x->s + "__" + name + "_" + std::to_string(counter);
    });
    for (auto& r : ren) sub[r.first] = mk(NK::Ident, r.second);
    NodeP b = subst_idents(f.body, sub);
    walk(b, [&](const NodeP& x) {
      if (x->k == NK::For && ren.count(x->s)) x->s = ren[x->s];
    });
    return b;
  };
  std::function<NodeP(const NodeP&, int)> ex = [&](const NodeP& n, int depth) -> NodeP {
    return rewrite(n, [&](const NodeP& x) -> NodeP {
      if (x->k == NK::Call && fns.count(x->s)) {
        if (depth > 16) throw PsError(PS_EARGUMENT, "nested function " + x->s + " calls itself (recursion is not supported)");
        std::vector<NodeP> args;
        for (auto& a : x->a) args.push_back(ex(a, depth));
        NodeP b = instantiate(x->s, args);
        if (b->a.size() != 1 || b->a[0]->k != NK::Return || b->a[0]->a.empty())
          throw PsError(PS_EARGUMENT, "nested function " + x->s + " is used as a value, so its body must be a single `return expr`");
        return ex(b->a[0]->a[0], depth + 1);
      }
      return x;
    });
  };
  std::function<void(const NodeP&, NodeP&, int)> st;
  auto blk = [&](const NodeP& b, int depth) {
    auto out = mk(NK::Block);
    for (auto& t : b->a) st(t, out, depth);
    return out;
  };
  st = [&](const NodeP& t, NodeP& out, int depth) {
    if (t->k == NK::Call && fns.count(t->s)) {
      if (depth > 16) throw PsError(PS_EARGUMENT, "nested function " + t->s + " calls itself (recursion is not supported)");
      std::vector<NodeP> args;
      for (auto& a : t->a) args.push_back(ex(a, depth));
      NodeP b = instantiate(t->s, args);
      for (size_t i = 0; i < b->a.size(); ++i) {
        if (b->a[i]->k == NK::Return) {
          if (i + 1 != b->a.size())
            throw PsError(PS_EARGUMENT, "nested function " + t->s + ": a return before its last statement is not supported");
          break;  // the value of a function called as a statement is discarded
        }
        walk(b->a[i], [&](const NodeP& x) {
          if (x->k == NK::Return) throw PsError(PS_EARGUMENT, "nested function " + t->s + ": a return inside a branch or loop is not supported");
        });
        st(b->a[i], out, depth + 1);
      }
      return;
    }
    if (t->k == NK::If) {
      NodeP els;
      if (t->a.size() > 2 && t->a[2]) {
        if (t->a[2]->k == NK::If) {
          auto tmp = mk(NK::Block);
          st(t->a[2], tmp, depth);
          els = tmp->a.size() == 1 ? tmp->a[0] : tmp;
        } else {
          els = blk(t->a[2], depth);
        }
      }
      out->a.push_back(mk(NK::If, "", {ex(t->a[0], depth), blk(t->a[1], depth), els}));
      return;
    }
    if (t->k == NK::For) {
      auto r = mk(NK::Range, t->a[0]->s, {ex(t->a[0]->a[0], depth), ex(t->a[0]->a[1], depth)});
      auto f = mk(NK::For, t->s, {r, blk(t->a[1], depth)});
      f->line = t->line;
      out->a.push_back(f);
      return;
    }
    if (t->k == NK::Block) {
      out->a.push_back(blk(t, depth));
      return;
    }
    out->a.push_back(ex(t, depth));
  };
  return blk(rest, 0);
}

// Kernel language of @parallel_indices kernels (src/ParallelKernel/kernel_language.jl:9-193, argument
// checks :264-295). Validates the argument count of each macro call; the lowering happens in the
// emitter.
// Tuples of index expressions (the splatted parallel indices I of @parallel_indices (I...)):
// broadcast arithmetic (.+ .- .* ./ .^, and + - * / with a scalar), unary minus, sum, prod,
// length, maximum, minimum and literal indexing are evaluated at definition, so that only
// scalar expressions reach the code generator (src/ParallelKernel/parallel.jl:189-217).
static NodeP fold_tuples(const NodeP& n) {
  if (!n) return n;
  for (auto& c : n->a) c = fold_tuples(c);
  auto is_tuple = [](const NodeP& x) { return x && x->k == NK::Tuple && std::none_of(x->a.begin(), x->a.end(), [](const NodeP& y) { return y->k == NK::Kw; }); };
  if (n->k == NK::Binary) {
    std::string op = n->s;
    const bool dotted = op.size() == 2 && op[0] == '.';
    if (dotted) op = op.substr(1);
    const NodeP &l = n->a[0], &r = n->a[1];
    if (!is_tuple(l) && !is_tuple(r)) {
      if (dotted) return mk(NK::Binary, op, {l, r});
      return n;
    }
    if (!dotted && op != "+" && op != "-" && !(is_tuple(l) != is_tuple(r)))
      throw PsError(PS_EARGUMENT, "tuple arithmetic needs broadcasting operators (.+ .- .* ./): " + to_julia(n));
    size_t nl = is_tuple(l) ? l->a.size() : 1, nr = is_tuple(r) ? r->a.size() : 1;
    if (nl != nr && nl != 1 && nr != 1) throw PsError(PS_EARGUMENT, "tuples of different lengths in " + to_julia(n));
    size_t m = std::max(nl, nr);
    auto t = mk(NK::Tuple);
    for (size_t i = 0; i < m; ++i) {
      NodeP a = is_tuple(l) ? clone(l->a[nl == 1 ? 0 : i]) : clone(l);
      NodeP b = is_tuple(r) ? clone(r->a[nr == 1 ? 0 : i]) : clone(r);
      t->a.push_back(mk(NK::Binary, op, {a, b}));
    }
    return t;
  }
  if (n->k == NK::Unary && is_tuple(n->a[0])) {
    auto t = mk(NK::Tuple);
    for (auto& x : n->a[0]->a) t->a.push_back(mk(NK::Unary, n->s, {x}));
    return t;
  }
  if (n->k == NK::Call && n->a.size() == 1 && is_tuple(n->a[0])) {
    const auto& xs = n->a[0]->a;
    if (n->s == "length") return mknum_int((long long)xs.size());
    if ((n->s == "sum" || n->s == "prod") && xs.empty()) return mknum_int(n->s == "sum" ? 0 : 1);
    if (n->s == "sum" || n->s == "prod") {
      NodeP acc = xs[0];
      for (size_t i = 1; i < xs.size(); ++i) acc = mk(NK::Binary, n->s == "sum" ? "+" : "*", {acc, xs[i]});
      return acc;
    }
    if ((n->s == "maximum" || n->s == "minimum") && !xs.empty())
      return xs.size() == 1 ? xs[0] : mk(NK::Call, n->s == "maximum" ? "max" : "min", xs);
  }
  if (n->k == NK::Index && n->a.size() == 2 && is_tuple(n->a[0]) && n->a[1]->k == NK::Num && n->a[1]->is_int) {
    long long i = n->a[1]->ival;
    if (i < 1 || i > (long long)n->a[0]->a.size()) throw PsError(PS_EARGUMENT, "tuple index out of range in " + to_julia(n));
    return n->a[0]->a[i - 1];
  }
  return n;
}

// @parallel_indices (I...): I stands for the kernel's ndims parallel indices; A[I...] and
// f(I...) spread them, a bare I is their tuple.
static NodeP expand_splat_indices(const NodeP& n, const std::string& I, int nd) {
  if (!n) return n;
  auto idx = [&] {
    std::vector<NodeP> v;
    for (int d = 0; d < nd; ++d) v.push_back(mk(NK::Ident, kIdx[d]));
    return v;
  };
  if (n->k == NK::Ident && n->s == I) return mk(NK::Tuple, "", idx());
  std::vector<NodeP> out;
  for (auto& c : n->a) {
    if (c && c->k == NK::Splat && c->a[0]->k == NK::Ident && c->a[0]->s == I && (n->k == NK::Index || n->k == NK::Call || n->k == NK::Tuple)) {
      for (auto& x : idx()) out.push_back(x);
    } else {
      out.push_back(expand_splat_indices(c, I, nd));
    }
  }
  n->a = out;
  return n;
}

static void check_kernel_language(Kernel& K, const NodeP& x) {
  const std::string& m = x->s;
  const size_t n = x->a.size();
  auto nargs = [&](size_t lo, size_t hi) {
    if (n < lo || n > hi) throw PsError(PS_EARGUMENT, "@" + m + ": " + (hi == 0 ? "no arguments allowed." : "wrong number of arguments."));
  };
  if (m == "gridDim" || m == "blockIdx" || m == "blockDim" || m == "threadIdx" || m == "sync_threads" || m == "warpsize" || m == "laneid") {
    nargs(0, 0);
  } else if (m == "active_mask") {
    nargs(0, 0);
    K.needs_idx64 = true;
  } else if (m == "sharedMem") {
    nargs(2, 3);
  } else if (m == "shfl_sync" || m == "shfl_up_sync" || m == "shfl_down_sync" || m == "shfl_xor_sync") {
    nargs(3, 4);
  } else if (m == "vote_any_sync" || m == "vote_all_sync") {
    nargs(2, 2);
  } else if (m == "vote_ballot_sync") {
    nargs(2, 2);
    K.needs_idx64 = true;
  } else if (m == "pk_println" || m == "pk_show") {
  } else {
    throw PsError(PS_EARGUMENT, "@parallel_indices kernels index arrays explicitly; macro @" + m + " is not supported here");
  }
  K.kernel_language = true;
}

Kernel define_kernel(const std::string& text, const JVal& opts) {
  Kernel K;
  K.def = parse_kernel(text);
  read_options(K.opt, opts);
  if (!K.def.kind.empty()) K.opt.kind = K.def.kind;
  if (!K.def.indices.empty()) K.opt.indices = K.def.indices;
  apply_prefix_kwargs(K.opt, K.def);
  if (K.opt.numbertype != "Float64" && K.opt.numbertype != "Float32")
    throw PsError(PS_EARGUMENT, "numbertype must be Float64 or Float32 (obtained: " + K.opt.numbertype + ")");
  K.def.body = expand_forall(K.def.body);
  for (auto& prm : K.def.params) K.def_params.push_back(prm.name);
  if (K.opt.tuple_params.empty()) {
    // without the call's structure, parameters used as P.name or P[k][...] are taken to be tuples
    // of the components the body names (the launch defines the kernel again with the structure
    // of its arguments)
    std::set<std::string> pn;
    for (auto& prm : K.def.params) pn.insert(prm.name);
    std::map<std::string, std::vector<std::string>> inferred;
    std::vector<std::string> order;
    auto add = [&](const std::string& P, const std::string& c) {
      if (!inferred.count(P)) order.push_back(P);
      auto& v = inferred[P];
      if (std::find(v.begin(), v.end(), c) == v.end()) v.push_back(c);
    };
    walk(K.def.body, [&](const NodeP& x) {
      if (x->k == NK::Dot && x->a[0]->k == NK::Ident && pn.count(x->a[0]->s)) add(x->a[0]->s, x->s);
      if (x->k == NK::Index && x->a[0]->k == NK::Index && x->a[0]->a[0]->k == NK::Ident && pn.count(x->a[0]->a[0]->s) &&
          x->a[0]->a.size() == 2 && x->a[0]->a[1]->k == NK::Num && x->a[0]->a[1]->is_int)
        for (long long i = 1; i <= x->a[0]->a[1]->ival; ++i) add(x->a[0]->a[0]->s, std::to_string(i));
    });
    for (auto& P : order) K.opt.tuple_params.emplace_back(P, inferred[P]);
  }
  if (!K.opt.tuple_params.empty()) {  // one parameter per component, in order
    std::vector<Param> ps;
    for (auto& prm : K.def.params) {
      auto it = std::find_if(K.opt.tuple_params.begin(), K.opt.tuple_params.end(), [&](const auto& t) { return t.first == prm.name; });
      if (it == K.opt.tuple_params.end()) {
        ps.push_back(prm);
        continue;
      }
      for (auto& c : it->second) ps.push_back(Param{prm.name + "." + c, prm.annot});
    }
    for (auto& t : K.opt.tuple_params)
      if (std::none_of(K.def.params.begin(), K.def.params.end(), [&](const Param& q) { return q.name == t.first; }))
        throw PsError(PS_EARGUMENT, "tuple argument '" + t.first + "' is not a parameter of kernel " + K.def.name);
    K.def.params = ps;
    K.def.body = expand_tuple_args(K.def.body, K.opt);
  }
  for (size_t i = 0; i < K.def.params.size(); ++i) K.param_index[K.def.params[i].name] = (int)i;

  // remove_return (src/ParallelKernel/shared.jl:200-209): the last statement must return
  // nothing and is dropped (the generated kernel returns at its end); no other return statement
  // may appear, except inside nested function definitions.
  auto strip_returns = [](NodeP body) {
    auto nothing_ret = [](const NodeP& s) {
      if (s->k == NK::Return) return s->a.empty() || (s->a[0]->k == NK::Ident && s->a[0]->s == "nothing");
      return s->k == NK::Ident && s->s == "nothing";
    };
    if (body->a.empty() || !nothing_ret(body->a.back()))
      throw PsError(PS_EARGUMENT, "invalid kernel in @parallel kernel definition: the last statement must be a `return nothing` statement ('return' or 'return nothing' or 'nothing') as required for any GPU kernels.");
    auto b = mk(NK::Block);
    b->a.assign(body->a.begin(), body->a.end() - 1);
    std::function<void(const NodeP&)> scan = [&](const NodeP& x) {
      if (!x || x->k == NK::FuncDef) return;
      if (x->k == NK::Return)
        throw PsError(PS_EARGUMENT, "invalid kernel in @parallel kernel definition: only one return statement is allowed in the kernel (exception: nested function definitions) and it must return nothing and be the last statement (required to ensure equal behaviour with different packages for parallellization).");
      for (auto& c : x->a) scan(c);
    };
    scan(b);
    return b;
  };

  if (K.opt.kind == "parallel") {
    int nd = K.opt.ndims;
    if (nd < 1 || nd > 3)
      throw PsError(PS_EKEYWORD, "@parallel: keyword argument 'ndims' is invalid or missing (valid values are 1, 2 or 3; 'ndims' an be set globally in @init_parallel_stencil and overwritten per kernel if needed).");
    K.nidx = nd;
    NodeP body = strip_returns(K.def.body);
    // validate_body: exclusively macro array assignments
    for (auto& s : body->a) {
      if (!(s->k == NK::Assign && s->s.empty() && s->a[0]->k == NK::Macro))
        throw PsError(PS_EARGUMENT,
                      "unsupported kernel statements in @parallel kernel definition: @parallel is only applicable to kernels that contain exclusively array assignments using macros from FiniteDifferences{1|2|3}D or from another compatible computation submodule. @parallel_indices supports any kind of statements in the kernels.");
      if (s->a[0]->a.size() != 1 || s->a[0]->a[0]->k != NK::Ident)
        throw PsError(PS_EARGUMENT, "unsupported kernel statements in @parallel kernel definition: the left-hand side must be @macro(A) with A an array.");
    }
    // interpolated indices: $ix, $iy, $iz name the parallel indices (ParallelStencil.INDICES) and
    // $ixi, $iyi, $izi the inner ones (INDICES_INN: ix+1 without padding, ix with padding,
    // src/ParallelKernel/parallel.jl:498-504)
    body = rewrite(body, [&](const NodeP& x) -> NodeP {
      if (x->k == NK::Interp) {
        for (int d = 0; d < 3; ++d) {
          const std::string nm = std::string("i") + "xyz"[d];
          if (x->s == nm) return mk(NK::Ident, kIdx[d]);
          if (x->s == nm + "i")
            return K.opt.padding ? mk(NK::Ident, kIdx[d]) : mk(NK::Binary, "+", {mk(NK::Ident, kIdx[d]), mknum_int(1)});
        }
      }
      return x;
    });
    // write vars and on-the-fly extraction
    std::vector<std::string> write_vars;
    for (auto& s : body->a) {
      const std::string& A = s->a[0]->a[0]->s;
      if (K.param_index.count(A) && std::find(write_vars.begin(), write_vars.end(), A) == write_vars.end()) write_vars.push_back(A);
    }
    std::map<std::string, NodeP> otf;
    auto remaining = mk(NK::Block);
    for (auto& s : body->a) {
      const std::string& m = s->a[0]->s;
      const std::string& A = s->a[0]->a[0]->s;
      if (K.param_index.count(A)) {
        remaining->a.push_back(s);
        continue;
      }
      if (m != "all")
        throw PsError(PS_EARGUMENT, "unsupported kernel statements in @parallel kernel definition: partial assignments are not possible for arrays that are not stored in global memory (arrays that are not among the arguments of the kernel); use '@all' instead.");
      if (refers_to(s->a[1], A))
        throw PsError(PS_EARGUMENT, "unsupported kernel statements in @parallel kernel definition: auto-dependency is not possible for arrays that are not stored in global memory (arrays that are not among the arguments of the kernel).");
      for (auto& w : write_vars)
        if (refers_to(s->a[1], w))
          throw PsError(PS_EARGUMENT, "unsupported kernel statements in @parallel kernel definition: the assignment of " + A +
                                          " should be done on the fly as it is not among the arguments of the kernel; however, this is not possible because it depends on at least one variable that is not read-only within the scope of the kernel (any of: " + w + ").");
      otf[A] = expand_all_macros(nd, s->a[1], K.opt.padding);
      K.otf_vars.push_back(A);
    }
    // masks, macro expansion, on-the-fly insertion
    auto out = mk(NK::Block);
    for (auto& s : remaining->a) {
      const std::string& m = s->a[0]->s;
      const std::string& A = s->a[0]->a[0]->s;
      NodeP mask = within_mask(nd, m, A, K.opt.padding);
      NodeP lhs = expand_macro(nd, m, A, K.opt.padding);
      if (lhs->k != NK::Index) throw PsError(PS_EARGUMENT, "@" + m + " cannot be assigned to");
      NodeP rhs = expand_all_macros(nd, s->a[1], K.opt.padding);
      if (K.opt.padding) {
        rhs = resolve_dir_indices(K, insert_onthefly_padded(rhs, otf, nd), A);
      } else {
        rhs = insert_onthefly(rhs, otf, nd);
      }
      auto as = mk(NK::Assign, "", {canon_indices(lhs), canon_indices(rhs)});
      auto blk = mk(NK::Block, "", {as});
      out->a.push_back(mk(NK::If, "", {mask, blk, nullptr}));
    }
    K.body = out;
  } else if (K.opt.kind == "parallel_indices") {
    if (K.opt.indices.empty()) throw PsError(PS_EARGUMENT, "@parallel_indices: argument 'indices' must be a tuple of indices");
    std::string splat;
    if (K.opt.indices.size() == 1 && K.opt.indices[0].size() > 3 && K.opt.indices[0].compare(K.opt.indices[0].size() - 3, 3, "...") == 0) {
      splat = K.opt.indices[0].substr(0, K.opt.indices[0].size() - 3);
      if (K.opt.ndims < 1 || K.opt.ndims > 3)
        throw PsError(PS_EKEYWORD, "@parallel_indices (I...): keyword argument 'ndims' is invalid or missing (valid values are 1, 2 or 3).");
      K.opt.indices.clear();
      for (int d = 0; d < K.opt.ndims; ++d) K.opt.indices.push_back(kIdx[d]);
    }
    if (K.opt.indices.size() > 3) throw PsError(PS_EARGUMENT, "@parallel_indices: at most 3 parallel indices are supported");
    K.nidx = (int)K.opt.indices.size();
    std::map<std::string, NodeP> sub;
    for (int d = 0; d < K.nidx; ++d) sub[K.opt.indices[d]] = mk(NK::Ident, kIdx[d]);
    NodeP body = inline_nested_functions(strip_returns(K.def.body));
    if (!splat.empty()) body = expand_splat_indices(body, splat, K.nidx);
    body = fold_tuples(body);
    body = subst_idents(body, sub);
    body = rewrite(body, [&](const NodeP& x) -> NodeP {
      if (x->k == NK::Interp) {
        for (int d = 0; d < 3; ++d)
          if (x->s == std::string(1, "xyz"[d]).insert(0, "i")) return mk(NK::Ident, kIdx[d]);
      }
      if (x->k == NK::Macro) check_kernel_language(K, x);
      return x;
    });
    // local arrays created by @sharedMem
    walk(body, [&](const NodeP& x) {
      if (x->k == NK::Assign && x->a[0]->k == NK::Ident && x->a[1]->k == NK::Macro && x->a[1]->s == "sharedMem") {
        const std::string& v = x->a[0]->s;
        if (K.param_index.count(v)) throw PsError(PS_EARGUMENT, "@sharedMem: cannot assign to kernel argument '" + v + "'");
        const NodeP& dims = x->a[1]->a[1];
        int nd = dims->k == NK::Tuple ? (int)dims->a.size() : 1;
        if (nd < 1 || nd > 3) throw PsError(PS_EARGUMENT, "@sharedMem: arrays of 1 to 3 dimensions are supported");
        auto it = K.shared_arrays.find(v);
        if (it != K.shared_arrays.end() && it->second != nd)
          throw PsError(PS_EARGUMENT, "@sharedMem: '" + v + "' is created with different numbers of dimensions");
        K.shared_arrays[v] = nd;
      } else if (x->k == NK::Macro && x->s == "sharedMem") {
        // a @sharedMem anywhere but as the whole right-hand side of a local assignment
        bool ok = false;
        walk(body, [&](const NodeP& y) {
          if (y->k == NK::Assign && y->a[1].get() == x.get() && y->a[0]->k == NK::Ident) ok = true;
        });
        if (!ok) throw PsError(PS_EARGUMENT, "@sharedMem must be assigned to a local variable: A = @sharedMem(T, dims)");
      }
    });
    if (K.kernel_language && K.opt.memopt)
      throw PsError(PS_EARGUMENT, "memopt kernels stage their own tiles through LDS; the kernel language (@sharedMem, @sync_threads, @threadIdx, ...) is only available with memopt=false");
    // all parallel indices must be used (src/parallel.jl:696-704)
    for (int d = 0; d < K.nidx; ++d) {
      bool used = false;
      walk(body, [&](const NodeP& x) {
        if (x->k == NK::Ident && x->s == kIdx[d]) used = true;
      });
      if (!used) throw PsError(PS_EARGUMENT, "@parallel_indices: all parallel indices must be used in the kernel body (unused indices: " + K.opt.indices[d] + ").");
    }
    if (K.opt.padding) body = padded_view_accesses(K, body);
    K.body = canon_indices(body);
  } else {
    throw PsError(PS_EARGUMENT, "unknown kernel kind '" + K.opt.kind + "'");
  }
  collect_access(K.body, K.accessed, K.written, K.param_index);
  // a kernel whose indices still compare parent sizes is analysed for memopt per launch, once the
  // comparisons have values (specialize_padding); its definition-time metadata take them as 0
  if (K.opt.memopt) {
    if (K.pgt_terms.empty()) memopt_analysis(K);
    else {
      Kernel Z = specialize_padding(K, std::string(K.pgt_terms.size(), '0'));
      K.mm = Z.mm;
    }
  }
  return K;
}

// ---------------------------------------------------------------------------
// info
// ---------------------------------------------------------------------------
std::string kernel_info_json(const Kernel& K) {
  std::ostringstream o;
  o << "{\"name\":" << json_str(K.def.name) << ",\"kind\":" << json_str(K.opt.kind)
    << ",\"is_parallel_kernel\":" << (K.opt.kind == "parallel" ? "true" : "false") << ",\"nb_parallel_indices\":" << K.nidx
    << ",\"kernel_language\":" << (K.kernel_language ? "true" : "false") << ",\"shared_arrays\":" << K.shared_arrays.size()
    << ",\"def_params\":[" << [&] {
         std::string r;
         for (size_t i = 0; i < K.def_params.size(); ++i) r += (i ? "," : "") + json_str(K.def_params[i]);
         return r;
       }() << "]"
    << ",\"numbertype\":" << json_str(K.opt.numbertype) << ",\"memopt\":" << (K.opt.memopt ? "true" : "false")
    << ",\"groupable\":" << (groupable_body(K) ? "true" : "false") << ",\"x_reach\":" << (K.body ? idx_span(K, 0) : 0) << ",\"z_reach\":" << (K.body ? idx_span(K, 2) : 0) << ",\"params\":[";
  for (size_t i = 0; i < K.def.params.size(); ++i)
    o << (i ? "," : "") << "{\"name\":" << json_str(K.def.params[i].name) << ",\"annot\":" << json_str(K.def.params[i].annot) << "}";
  o << "],\"written\":[";
  bool f = true;
  for (auto& w : K.written) {
    o << (f ? "" : ",") << json_str(w);
    f = false;
  }
  o << "],\"accessed\":[";
  f = true;
  for (auto& w : K.accessed) {
    o << (f ? "" : ",") << json_str(w);
    f = false;
  }
  o << "],\"onthefly\":[";
  for (size_t i = 0; i < K.otf_vars.size(); ++i) o << (i ? "," : "") << json_str(K.otf_vars[i]);
  o << "]";
  if (K.mm.enabled) {
    const auto& mm = K.mm;
    o << ",\"loopdim\":" << mm.loopdim << ",\"loopsize\":" << mm.loopsize << ",\"loopsizes\":["
      << (mm.loopdim == 3 ? "1,1," + std::to_string(mm.loopsize) : mm.loopdim == 2 ? "1," + std::to_string(mm.loopsize) + ",1" : std::to_string(mm.loopsize) + ",1,1")
      << "],\"optvars\":[";
    for (size_t i = 0; i < mm.optvars.size(); ++i) o << (i ? "," : "") << json_str(mm.optvars[i]);
    o << "],\"stencilranges\":{";
    for (size_t i = 0; i < mm.optvars.size(); ++i) {
      auto& r = mm.stencilranges.at(mm.optvars[i]);
      o << (i ? "," : "") << json_str(mm.optvars[i]) << ":[[" << r[0].first << "," << r[0].second << "],[" << r[1].first << ","
        << r[1].second << "],[" << r[2].first << "," << r[2].second << "]]";
    }
    o << "},\"optranges\":{";
    f = true;
    for (auto& kv : mm.optranges) {
      auto& r = kv.second;
      o << (f ? "" : ",") << json_str(kv.first) << ":[[" << r[0].first << "," << r[0].second << "],[" << r[1].first << ","
        << r[1].second << "],[" << r[2].first << "," << r[2].second << "]]";
      f = false;
    }
    o << "},\"offsets\":{";
    f = true;
    for (auto& kv : mm.offsets) {
      o << (f ? "" : ",") << json_str(kv.first) << ":[";
      bool g = true;
      for (auto& c : kv.second)
        for (auto& z : c.second) {
          o << (g ? "" : ",") << "[" << c.first.first << "," << c.first.second << "," << z.first << "," << z.second << "]";
          g = false;
        }
      o << "]";
      f = false;
    }
    o << "},\"shmem_optvars\":[";
    for (size_t i = 0; i < mm.shmem_optvars.size(); ++i) o << (i ? "," : "") << json_str(mm.shmem_optvars[i]);
    o << "],\"use_any_shmem\":" << (mm.shmem_optvars.empty() ? "false" : "true") << ",\"use_shmemhalos\":{";
    f = true;
    for (auto& kv : mm.use_shmemhalos) {
      o << (f ? "" : ",") << json_str(kv.first) << ":" << (kv.second ? "true" : "false");
      f = false;
    }
    o << "},\"shmem_dim1\":1,\"shmem_dim2\":" << (mm.loopdim == 3 ? 2 : 3) << ",\"shmem_spans\":{";
    for (size_t i = 0; i < mm.optvars.size(); ++i) {
      auto& r = mm.stencilranges.at(mm.optvars[i]);
      int d2 = mm.loopdim == 3 ? 1 : 2;
      o << (i ? "," : "") << json_str(mm.optvars[i]) << ":[" << (r[0].second - r[0].first) << "," << (r[d2].second - r[d2].first) << "]";
    }
    o << "}";
  }
  o << ",\"body\":" << json_str(to_julia(K.body)) << "}";
  return o.str();
}

std::string Variant::key() const {
  char b[96];
  std::snprintf(b, sizeof b, "m%dc%di%da%dt%dl%d_%dx%dn%do%du%dg%dz%d", memopt, checked, idx64, noalias, maxthreads, lds, bx, by, nt, off32, u24, cx,
                cz);
  return std::string(b) + "d" + divmode + (padbits.empty() ? "" : "p" + padbits) + (sclass.empty() ? "" : "s" + sclass);
}

}  // namespace ps
