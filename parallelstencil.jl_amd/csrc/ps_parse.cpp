// Lexer and recursive-descent parser for the stencil kernel language.
// Operator precedence and associativity follow Julia's parser, because the
// evaluation order of the generated code must match what the reference's
// kernels compute (n-ary `+`/`*` fold left, unary minus binds tighter than
// `*` but looser than `^`, `^` is right-associative, comparisons chain).
#include "ps_ast.h"
#include "ps_amd.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ps {

NodeP mk(NK k, const std::string& s, std::vector<NodeP> a) {
  auto n = std::make_shared<Node>(k);
  n->s = s;
  n->a = std::move(a);
  return n;
}
NodeP mknum_int(long long v) {
  auto n = mk(NK::Num);
  n->is_int = true;
  n->ival = v;
  n->num = (double)v;
  return n;
}
NodeP mknum_float(double v, bool f32) {
  auto n = mk(NK::Num);
  n->num = v;
  n->is_f32 = f32;
  return n;
}
NodeP clone(const NodeP& n) {
  if (!n) return n;
  auto c = std::make_shared<Node>(*n);
  for (auto& x : c->a) x = clone(x);
  return c;
}

namespace {

enum class TK { Eof, NL, Ident, Num, Op, Str };

struct Tok {
  TK k;
  std::string s;
  int line;
};

bool is_ident_start(unsigned char c) { return std::isalpha(c) || c == '_' || c >= 0x80; }
bool is_ident_char(unsigned char c) { return std::isalnum(c) || c == '_' || c == '!' || c >= 0x80; }

// UTF-8 operator sequences that must not be absorbed into identifiers.
const char* kUniOps[][2] = {
    {"\xC3\xB7", "÷"},          // ÷
    {"\xE2\x89\xA4", "<="},     // ≤
    {"\xE2\x89\xA5", ">="},     // ≥
    {"\xE2\x89\xA0", "!="},     // ≠
    {"\xE2\x88\x88", "∈"},      // ∈
};

int match_uniop(const std::string& t, size_t i, std::string* canon) {
  for (auto& u : kUniOps) {
    size_t L = std::strlen(u[0]);
    if (t.compare(i, L, u[0]) == 0) {
      if (canon) *canon = u[1];
      return (int)L;
    }
  }
  return 0;
}

std::vector<Tok> lex(const std::string& t) {
  std::vector<Tok> out;
  int line = 1;
  int depth = 0;  // () and [] nesting: newlines inside are not statement separators
  size_t i = 0, n = t.size();
  auto push = [&](TK k, std::string s) { out.push_back(Tok{k, std::move(s), line}); };
  while (i < n) {
    unsigned char c = t[i];
    if (c == '\n') {
      if (depth == 0) push(TK::NL, "\n");
      ++line;
      ++i;
      continue;
    }
    if (c == ' ' || c == '\t' || c == '\r') {
      ++i;
      continue;
    }
    if (c == '#') {
      if (i + 1 < n && t[i + 1] == '=') {  // #= block comment =#
        size_t e = t.find("=#", i + 2);
        if (e == std::string::npos) throw PsError(PS_ESYNTAX, "unterminated block comment");
        for (size_t j = i; j < e; ++j)
          if (t[j] == '\n') ++line;
        i = e + 2;
      } else {
        while (i < n && t[i] != '\n') ++i;
      }
      continue;
    }
    std::string canon;
    int ul = match_uniop(t, i, &canon);
    if (ul) {
      push(TK::Op, canon);
      i += ul;
      continue;
    }
    if (std::isdigit(c) || (c == '.' && i + 1 < n && std::isdigit((unsigned char)t[i + 1]))) {
      size_t j = i;
      bool isf = false;
      while (j < n && std::isdigit((unsigned char)t[j])) ++j;
      if (j < n && t[j] == '.' && !(j + 1 < n && (t[j + 1] == '.' || t[j + 1] == '*' || t[j + 1] == '/' || t[j + 1] == '^' || t[j + 1] == '+' || t[j + 1] == '-'))) {
        isf = true;
        ++j;
        while (j < n && std::isdigit((unsigned char)t[j])) ++j;
      }
      if (j < n && (t[j] == 'e' || t[j] == 'E' || t[j] == 'f') &&
          (j + 1 < n && (std::isdigit((unsigned char)t[j + 1]) || ((t[j + 1] == '-' || t[j + 1] == '+') && j + 2 < n && std::isdigit((unsigned char)t[j + 2]))))) {
        isf = true;
        ++j;
        if (t[j] == '-' || t[j] == '+') ++j;
        while (j < n && std::isdigit((unsigned char)t[j])) ++j;
      }
      (void)isf;
      push(TK::Num, t.substr(i, j - i));
      i = j;
      continue;
    }
    if (is_ident_start(c)) {
      size_t j = i;
      while (j < n) {
        if (match_uniop(t, j, nullptr)) break;
        unsigned char d = t[j];
        if (!is_ident_char(d)) break;
        if (d == '!' && j + 1 < n && t[j + 1] == '=') break;  // a!=b
        ++j;
      }
      push(TK::Ident, t.substr(i, j - i));
      i = j;
      continue;
    }
    if (c == '"') {
      size_t j = i + 1;
      while (j < n && t[j] != '"') ++j;
      push(TK::Str, t.substr(i + 1, j - i - 1));
      i = j + 1;
      continue;
    }
    if (t.compare(i, 3, "...") == 0) {
      push(TK::Op, "...");
      i += 3;
      continue;
    }
    static const char* ops2[] = {"&&", "||", "==", "!=", "<=", ">=", "::", "+=", "-=", "*=", "/=", "^=", "->", "<:",
                                 ".+", ".-", ".*", "./", ".^"};
    bool done = false;
    for (auto o : ops2) {
      if (t.compare(i, 2, o) == 0) {
        push(TK::Op, o);
        i += 2;
        done = true;
        break;
      }
    }
    if (done) continue;
    if (std::strchr("+-*/^()[],;=<>!?:.$@%{}&|", c)) {
      if (c == '(' || c == '[' || c == '{') ++depth;
      if (c == ')' || c == ']' || c == '}') --depth;
      push(TK::Op, std::string(1, (char)c));
      ++i;
      continue;
    }
    char buf[96];
    std::snprintf(buf, sizeof buf, "unexpected character '%c' (0x%02x) at line %d", c, c, line);
    throw PsError(PS_ESYNTAX, buf);
  }
  out.push_back(Tok{TK::Eof, "", line});
  return out;
}

struct Parser {
  std::vector<Tok> t;
  size_t p = 0;
  int bracket = 0;    // >0 while parsing inside [...]: `end` is an index
  int no_range = 0;   // >0 while parsing the true branch of a ternary

  const Tok& cur() const { return t[p]; }
  const Tok& peek(int k = 1) const { return t[std::min(p + k, t.size() - 1)]; }
  bool is_op(const char* s) const { return cur().k == TK::Op && cur().s == s; }
  bool is_kw(const char* s) const { return cur().k == TK::Ident && cur().s == s; }
  [[noreturn]] void fail(const std::string& m) const {
    throw PsError(PS_ESYNTAX, "syntax error at line " + std::to_string(cur().line) + " near '" + cur().s + "': " + m);
  }
  void expect_op(const char* s) {
    if (!is_op(s)) fail(std::string("expected '") + s + "'");
    ++p;
  }
  void skip_nl() {
    while (cur().k == TK::NL || is_op(";")) ++p;
  }
  void skip_nl_only() {
    while (cur().k == TK::NL) ++p;
  }

  // ---------------- expressions ----------------
  NodeP expr() { return ternary(); }

  NodeP ternary() {
    NodeP c = oror();
    if (is_op("?")) {
      ++p;
      skip_nl_only();
      ++no_range;
      NodeP a = ternary();
      --no_range;
      skip_nl_only();
      expect_op(":");
      skip_nl_only();
      NodeP b = ternary();
      return mk(NK::Tern, "", {c, a, b});
    }
    return c;
  }
  NodeP oror() {
    NodeP l = andand();
    while (is_op("||")) {
      ++p;
      skip_nl_only();
      l = mk(NK::Or, "||", {l, andand()});
    }
    return l;
  }
  NodeP andand() {
    NodeP l = cmp();
    while (is_op("&&")) {
      ++p;
      skip_nl_only();
      l = mk(NK::And, "&&", {l, cmp()});
    }
    return l;
  }
  static bool is_cmp_op(const Tok& k) {
    return k.k == TK::Op && (k.s == "<" || k.s == "<=" || k.s == ">" || k.s == ">=" || k.s == "==" || k.s == "!=");
  }
  NodeP cmp() {
    NodeP l = range();
    if (!is_cmp_op(cur())) return l;
    auto c = mk(NK::Cmp);
    c->a.push_back(l);
    while (is_cmp_op(cur())) {
      c->ops.push_back(cur().s);
      ++p;
      skip_nl_only();
      c->a.push_back(range());
    }
    return c;
  }
  NodeP range() {
    NodeP l = add();
    if (no_range == 0 && is_op(":") && bracket_or_paren_ok()) {
      ++p;
      NodeP r = add();
      return mk(NK::Range, ":", {l, r});
    }
    return l;
  }
  bool bracket_or_paren_ok() const { return true; }
  NodeP add() {
    NodeP l = mul();
    while ((is_op("+") || is_op("-") || is_op("|") || is_op(".+") || is_op(".-")) && !(cur().k == TK::Op && peek().k == TK::NL && false)) {
      std::string o = cur().s;
      ++p;
      skip_nl_only();
      l = mk(NK::Binary, o, {l, mul()});
    }
    return l;
  }
  NodeP mul() {
    NodeP l = unary();
    while (is_op("*") || is_op("/") || is_op("÷") || is_op("%") || is_op("&") || is_op(".*") || is_op("./")) {
      std::string o = cur().s;
      ++p;
      skip_nl_only();
      l = mk(NK::Binary, o, {l, unary()});
    }
    return l;
  }
  NodeP unary() {
    if (is_op("-") || is_op("+") || is_op("!")) {
      std::string o = cur().s;
      ++p;
      NodeP x = unary();
      return mk(NK::Unary, o, {x});
    }
    return power();
  }
  NodeP power() {
    NodeP b = postfix();
    if (is_op("^") || is_op(".^")) {
      std::string op = cur().s;
      ++p;
      NodeP e;
      if (is_op("-") || is_op("+")) {
        std::string o = cur().s;
        ++p;
        e = mk(NK::Unary, o, {power()});
      } else {
        e = power();
      }
      return mk(NK::Binary, op, {b, e});
    }
    return b;
  }
  std::vector<NodeP> arglist(const char* close) {
    std::vector<NodeP> args;
    skip_nl_only();
    while (!is_op(close)) {
      if (cur().k == TK::Ident && peek().k == TK::Op && peek().s == "=") {
        std::string kw = cur().s;
        p += 2;
        args.push_back(mk(NK::Kw, kw, {expr()}));
      } else if (is_op(";")) {  // keyword-argument separator in calls
        ++p;
        continue;
      } else {
        args.push_back(expr());
      }
      skip_nl_only();
      if (is_op(",")) {
        ++p;
        skip_nl_only();
      } else if (!is_op(close) && !is_op(";")) {
        fail(std::string("expected ',' or '") + close + "'");
      }
    }
    ++p;
    return args;
  }
  NodeP postfix() {
    NodeP e = atom();
    for (;;) {
      if (is_op("(")) {
        ++p;
        ++bracket_guard;
        int saved = bracket;
        bracket = 0;
        auto args = arglist(")");
        bracket = saved;
        --bracket_guard;
        e = mk(NK::Call, e->k == NK::Ident ? e->s : to_julia(e), args);
      } else if (is_op("[")) {
        ++p;
        ++bracket;
        auto idx = arglist("]");
        --bracket;
        std::vector<NodeP> a{e};
        a.insert(a.end(), idx.begin(), idx.end());
        e = mk(NK::Index, "", a);
      } else if (is_op(".") && peek().k == TK::Ident) {
        ++p;
        std::string f = cur().s;
        ++p;
        e = mk(NK::Dot, f, {e});
      } else if (is_op("::")) {
        ++p;
        NodeP ty = postfix_type();
        e = mk(NK::Typed, "", {e, ty});
      } else if (is_op("...")) {
        ++p;
        e = mk(NK::Splat, "...", {e});
      } else {
        break;
      }
    }
    return e;
  }
  int bracket_guard = 0;
  NodeP postfix_type() {
    // a type expression: Ident(.Ident)*({...})?  or $interp
    NodeP e = atom();
    for (;;) {
      if (is_op(".") && peek().k == TK::Ident) {
        ++p;
        e = mk(NK::Dot, cur().s, {e});
        ++p;
      } else if (is_op("{")) {
        int d = 0;
        do {
          if (is_op("{")) ++d;
          if (is_op("}")) --d;
          ++p;
        } while (d > 0 && cur().k != TK::Eof);
      } else {
        break;
      }
    }
    return e;
  }
  NodeP atom() {
    const Tok& k = cur();
    if (k.k == TK::Num) {
      ++p;
      return parse_number(k.s);
    }
    if (k.k == TK::Str) {
      ++p;
      return mk(NK::Str, k.s);
    }
    if (k.k == TK::Ident) {
      if (k.s == "end" && bracket > 0) {
        ++p;
        return mk(NK::EndKw, "end");
      }
      if (k.s == "true" || k.s == "false") {
        ++p;
        auto n = mk(NK::Ident, k.s);
        return n;
      }
      ++p;
      auto n = mk(NK::Ident, k.s);
      n->line = k.line;
      return n;
    }
    if (k.k == TK::Op && k.s == "$") {
      ++p;
      if (cur().k != TK::Ident) fail("expected identifier after '$'");
      auto n = mk(NK::Interp, cur().s);
      ++p;
      return n;
    }
    if (k.k == TK::Op && k.s == "@") {
      ++p;
      if (cur().k != TK::Ident) fail("expected macro name after '@'");
      std::string name = cur().s;
      ++p;
      if (!is_op("(")) fail("macro calls must use parentheses: @" + name + "(...)");
      ++p;
      int saved = bracket;
      bracket = 0;
      auto args = arglist(")");
      bracket = saved;
      return mk(NK::Macro, name, args);
    }
    if (k.k == TK::Op && k.s == "(") {
      ++p;
      int saved = bracket;
      bracket = 0;
      skip_nl_only();
      if (is_op(")")) {
        ++p;
        bracket = saved;
        return mk(NK::Tuple);
      }
      std::vector<NodeP> items;
      bool tuple = false;
      for (;;) {
        if (cur().k == TK::Ident && peek().k == TK::Op && peek().s == "=") {
          std::string kw = cur().s;
          p += 2;
          items.push_back(mk(NK::Kw, kw, {expr()}));
          tuple = true;
        } else {
          items.push_back(expr());
        }
        skip_nl_only();
        if (is_op(",")) {
          tuple = true;
          ++p;
          skip_nl_only();
          if (is_op(")")) break;
          continue;
        }
        break;
      }
      expect_op(")");
      bracket = saved;
      if (!tuple && items.size() == 1) return items[0];
      return mk(NK::Tuple, "", items);
    }
    fail("unexpected token");
  }
  NodeP parse_number(const std::string& s) {
    bool f32 = s.find('f') != std::string::npos;
    bool isfloat = f32 || s.find('.') != std::string::npos || s.find('e') != std::string::npos || s.find('E') != std::string::npos;
    if (!isfloat) {
      errno = 0;
      long long v = std::strtoll(s.c_str(), nullptr, 10);
      if (errno) fail("integer literal out of range");
      return mknum_int(v);
    }
    std::string t = s;
    if (f32) t[t.find('f')] = 'e';
    double v = std::strtod(t.c_str(), nullptr);
    if (f32) v = (double)(float)v;
    return mknum_float(v, f32);
  }

  // ---------------- statements ----------------
  bool at_block_end() const {
    return cur().k == TK::Eof || is_kw("end") || is_kw("else") || is_kw("elseif") || is_op(")");
  }
  NodeP block() {
    auto b = mk(NK::Block);
    skip_nl();
    while (!at_block_end()) {
      b->a.push_back(statement());
      if (!(cur().k == TK::NL || is_op(";") || at_block_end())) fail("expected end of statement");
      skip_nl();
    }
    return b;
  }
  NodeP statement() {
    int line = cur().line;
    if (is_kw("return")) {
      ++p;
      auto r = mk(NK::Return);
      if (!(cur().k == TK::NL || is_op(";") || at_block_end())) r->a.push_back(expr());
      return r;
    }
    if (is_kw("if")) {
      ++p;
      return if_rest();
    }
    if (is_kw("begin")) {
      ++p;
      auto b = block();
      if (!is_kw("end")) fail("expected 'end'");
      ++p;
      return b;
    }
    if (is_kw("local")) ++p;
    if (is_kw("for")) {
      ++p;
      // for a = r1, b = r2 ... end  is  for a = r1; for b = r2 ... end; end
      std::vector<std::pair<std::string, NodeP>> loops;
      for (;;) {
        if (cur().k != TK::Ident) fail("expected loop variable");
        std::string v = cur().s;
        ++p;
        if (is_op("=") || is_kw("in") || is_op("∈")) ++p;
        else fail("expected '=' or 'in' in for loop");
        NodeP r = expr();
        if (r->k != NK::Range) fail("for loops must iterate over a range a:b");
        loops.emplace_back(v, r);
        if (!is_op(",")) break;
        ++p;
      }
      auto b = block();
      if (!is_kw("end")) fail("expected 'end' closing 'for'");
      ++p;
      NodeP f = b;
      for (size_t i = loops.size(); i-- > 0;) {
        f = mk(NK::For, loops[i].first, {loops[i].second, i + 1 == loops.size() ? f : mk(NK::Block, "", {f})});
        f->line = line;
      }
      return f;
    }
    if (is_op("@") && peek().k == TK::Ident && peek().s == "∀") {  // @∀ x ∈ X statement
      p += 2;
      NodeP var = atom();
      if (!(var->k == NK::Ident || var->k == NK::Tuple)) fail("@∀: expected a variable name or a tuple of names");
      if (is_op("∈") || is_kw("in") || is_op("=")) ++p;
      else fail("@∀: the first argument must be of the form `x ∈ X`, `x in X` or `x = X`");
      NodeP coll = expr();
      NodeP body = statement();
      auto f = mk(NK::Forall, "", {var, coll, body});
      f->line = line;
      return f;
    }
    if (is_kw("function")) {  // nested function (long form)
      ++p;
      NodeP sig = expr();
      if (sig->k != NK::Call) fail("expected a function signature name(args...)");
      auto b = block();
      if (!is_kw("end")) fail("expected 'end' closing 'function'");
      ++p;
      return funcdef(sig, b);
    }
    if (is_op("@") && peek().k == TK::Ident && (peek().s == "inbounds" || peek().s == "inline" || peek().s == "fastmath") && !(peek(2).k == TK::Op && peek(2).s == "(")) {
      p += 2;  // statement-level annotations carry no meaning for generated code
      return statement();
    }
    NodeP lhs = expr();
    if (lhs->k == NK::Call && is_op("=")) {  // nested function (short form): name(args...) = body
      ++p;
      skip_nl_only();
      return funcdef(lhs, short_body());
    }
    if (is_op("=") || is_op("+=") || is_op("-=") || is_op("*=") || is_op("/=")) {
      std::string o = cur().s;
      ++p;
      skip_nl_only();
      NodeP rhs = expr();
      auto a = mk(NK::Assign, o == "=" ? "" : o.substr(0, 1), {lhs, rhs});
      a->line = line;
      return a;
    }
    return lhs;  // expression statement
  }
  // Body of a short-form function: `return e`, `(stmt; stmt; ...)` or an expression (its value).
  NodeP short_body() {
    if (is_kw("return")) return mk(NK::Block, "", {statement()});
    if (is_op("(")) {
      ++p;
      int saved = bracket;
      bracket = 0;
      auto b = block();
      bracket = saved;
      if (!is_op(")")) fail("expected ')'");
      ++p;
      if (b->a.size() == 1) {
        NK k = b->a[0]->k;
        if (k != NK::Assign && k != NK::Return && k != NK::If && k != NK::For && k != NK::Block && k != NK::FuncDef)
          return mk(NK::Block, "", {mk(NK::Return, "", {b->a[0]})});
      }
      return b;
    }
    return mk(NK::Block, "", {mk(NK::Return, "", {expr()})});
  }
  NodeP funcdef(const NodeP& sig, const NodeP& body) {
    auto params = mk(NK::Tuple);
    for (auto& a : sig->a) {
      NodeP x = a->k == NK::Typed ? a->a[0] : a;
      if (x->k != NK::Ident) fail("nested function parameters must be names");
      params->a.push_back(x);
    }
    auto f = mk(NK::FuncDef, sig->s, {params, body});
    f->line = sig->line;
    return f;
  }
  NodeP if_rest() {
    NodeP c = expr();
    NodeP th = block();
    NodeP el;
    if (is_kw("elseif")) {
      ++p;
      el = if_rest();
      return mk(NK::If, "", {c, th, el});
    }
    if (is_kw("else")) {
      ++p;
      el = block();
    }
    if (!is_kw("end")) fail("expected 'end' closing 'if'");
    ++p;
    return mk(NK::If, "", {c, th, el});
  }

  KernelDef kernel() {
    KernelDef kd;
    skip_nl();
    if (is_op("@") && peek().k == TK::Ident &&
        (peek().s == "parallel" || peek().s == "parallel_indices" || peek().s == "parallel_async")) {
      ++p;
      kd.kind = cur().s == "parallel_indices" ? "parallel_indices" : "parallel";
      ++p;
      if (kd.kind == "parallel_indices") {
        NodeP idx = postfix();
        if (idx->k == NK::Ident) kd.indices.push_back(idx->s);
        else if (idx->k == NK::Splat && idx->a[0]->k == NK::Ident) kd.indices.push_back(idx->a[0]->s + "...");
        else if (idx->k == NK::Tuple) {
          for (auto& x : idx->a) {
            if (x->k != NK::Ident) fail("parallel indices must be identifiers");
            kd.indices.push_back(x->s);
          }
        } else fail("@parallel_indices expects a tuple of indices");
      }
      while (cur().k == TK::Ident && peek().k == TK::Op && peek().s == "=") {
        std::string kw = cur().s;
        p += 2;
        kd.prefix_kwargs.emplace_back(kw, expr());
      }
    }
    skip_nl_only();
    bool longform = false;
    if (is_kw("function")) {
      longform = true;
      ++p;
    }
    if (cur().k != TK::Ident) fail("expected kernel name");
    kd.name = cur().s;
    ++p;
    expect_op("(");
    auto params = arglist(")");
    for (auto& pr : params) {
      Param q;
      if (pr->k == NK::Ident) q.name = pr->s;
      else if (pr->k == NK::Typed && pr->a[0]->k == NK::Ident) {
        q.name = pr->a[0]->s;
        q.annot = to_julia(pr->a[1]);
      } else fail("unsupported kernel parameter");
      kd.params.push_back(q);
    }
    if (is_kw("where")) {  // where-clauses constrain Julia dispatch only
      while (!(cur().k == TK::NL || cur().k == TK::Eof || is_op("="))) ++p;
    }
    if (longform) {
      kd.body = block();
      if (!is_kw("end")) fail("expected 'end' closing 'function'");
      ++p;
    } else {
      expect_op("=");
      skip_nl_only();
      if (is_op("(")) {
        ++p;
        kd.body = block();
        expect_op(")");
      } else {
        auto b = mk(NK::Block);
        b->a.push_back(statement());
        kd.body = b;
      }
    }
    skip_nl();
    if (cur().k != TK::Eof) fail("trailing text after kernel definition");
    return kd;
  }
};

void emit(const NodeP& n, std::string& o) {
  if (!n) return;
  switch (n->k) {
    case NK::Num: {
      char b[64];
      if (n->is_int) std::snprintf(b, sizeof b, "%lld", n->ival);
      else std::snprintf(b, sizeof b, "%.17g%s", n->num, n->is_f32 ? "f0" : "");
      o += b;
      if (!n->is_int && !n->is_f32 && std::string(b).find_first_of(".e") == std::string::npos) o += ".0";
      break;
    }
    case NK::Ident: o += n->s; break;
    case NK::Interp: o += "$" + n->s; break;
    case NK::EndKw: o += "end"; break;
    case NK::Str: o += "\"" + n->s + "\""; break;
    case NK::Index:
      emit(n->a[0], o);
      o += "[";
      for (size_t i = 1; i < n->a.size(); ++i) {
        if (i > 1) o += ", ";
        emit(n->a[i], o);
      }
      o += "]";
      break;
    case NK::Call:
    case NK::Macro:
      o += (n->k == NK::Macro ? "@" : "") + n->s + "(";
      for (size_t i = 0; i < n->a.size(); ++i) {
        if (i) o += ", ";
        emit(n->a[i], o);
      }
      o += ")";
      break;
    case NK::Unary: o += n->s + "("; emit(n->a[0], o); o += ")"; break;
    case NK::Binary: o += "("; emit(n->a[0], o); o += " " + n->s + " "; emit(n->a[1], o); o += ")"; break;
    case NK::Cmp:
      o += "(";
      for (size_t i = 0; i < n->a.size(); ++i) {
        if (i) o += " " + n->ops[i - 1] + " ";
        emit(n->a[i], o);
      }
      o += ")";
      break;
    case NK::And: o += "("; emit(n->a[0], o); o += " && "; emit(n->a[1], o); o += ")"; break;
    case NK::Or: o += "("; emit(n->a[0], o); o += " || "; emit(n->a[1], o); o += ")"; break;
    case NK::Tern: o += "("; emit(n->a[0], o); o += " ? "; emit(n->a[1], o); o += " : "; emit(n->a[2], o); o += ")"; break;
    case NK::Tuple:
      o += "(";
      for (size_t i = 0; i < n->a.size(); ++i) {
        if (i) o += ", ";
        emit(n->a[i], o);
      }
      if (n->a.size() == 1) o += ",";
      o += ")";
      break;
    case NK::Range: emit(n->a[0], o); o += ":"; emit(n->a[1], o); break;
    case NK::Dot: emit(n->a[0], o); o += "." + n->s; break;
    case NK::Typed: emit(n->a[0], o); o += "::"; emit(n->a[1], o); break;
    case NK::Kw: o += n->s + "="; emit(n->a[0], o); break;
    case NK::Assign: emit(n->a[0], o); o += " " + n->s + "= "; emit(n->a[1], o); break;
    case NK::Return: o += "return"; break;
    case NK::Block:
      for (auto& s : n->a) { emit(s, o); o += "\n"; }
      break;
    case NK::If:
      o += "if "; emit(n->a[0], o); o += "\n"; emit(n->a[1], o);
      if (n->a.size() > 2 && n->a[2]) { o += "else\n"; emit(n->a[2], o); }
      o += "end";
      break;
    case NK::For:
      o += "for " + n->s + " = "; emit(n->a[0], o); o += "\n"; emit(n->a[1], o); o += "end";
      break;
    case NK::FuncDef:
      o += "function " + n->s; emit(n->a[0], o); o += "\n"; emit(n->a[1], o); o += "end";
      break;
    case NK::Forall:
      o += "@∀ "; emit(n->a[0], o); o += " ∈ "; emit(n->a[1], o); o += " "; emit(n->a[2], o);
      break;
    case NK::Splat:
      emit(n->a[0], o); o += "...";
      break;
  }
}

}  // namespace

std::string to_julia(const NodeP& n) {
  std::string o;
  emit(n, o);
  return o;
}

KernelDef parse_kernel(const std::string& text) {
  Parser P;
  P.t = lex(text);
  return P.kernel();
}

NodeP parse_expression(const std::string& text) {
  Parser P;
  P.t = lex(text);
  NodeP e = P.expr();
  P.skip_nl();
  if (P.cur().k != TK::Eof) P.fail("trailing text after expression");
  return e;
}

}  // namespace ps
