// Interpreter of the Lua subset of FIESTA input decks: lexer, recursive-
// descent parser to a syntax tree, tree-walking evaluator, library.
#include "fiesta_lua.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_set>

namespace fiesta_lua {

const char *Value::type_name() const {
  switch (k) {
    case Nil: return "nil";
    case Bool: return "boolean";
    case Num: return "number";
    case Str: return "string";
    case Tab: return "table";
    default: return "function";
  }
}

// ------------------------------------------------------------------ tables

namespace {

// a whole number that fits the array part / an exact key
bool whole(double d, long long *out) {
  if (!(d == std::floor(d)) || std::fabs(d) > 9.0e15) return false;
  *out = (long long)d;
  return true;
}

bool raw_equal(const Value &a, const Value &b) {
  if (a.k != b.k) return false;
  switch (a.k) {
    case Value::Nil: return true;
    case Value::Bool: return a.b == b.b;
    case Value::Num: return a.n == b.n;
    case Value::Str: return *a.s == *b.s;
    case Value::Tab: return a.t == b.t;
    default: return a.f == b.f;
  }
}

}  // namespace

Value Table::get(const Value &key) const {
  switch (key.k) {
    case Value::Str: {
      auto it = str.find(*key.s);
      return it == str.end() ? Value() : it->second;
    }
    case Value::Num: {
      long long i;
      if (whole(key.n, &i) && i >= 1 && (size_t)i <= arr.size()) return arr[(size_t)i - 1];
      auto it = num.find(key.n);
      return it == num.end() ? Value() : it->second;
    }
    case Value::Nil: return Value();
    default:
      for (const auto &kv : other)
        if (raw_equal(kv.first, key)) return kv.second;
      return Value();
  }
}

void Table::set(const Value &key, const Value &v) {
  switch (key.k) {
    case Value::Str:
      if (v.k == Value::Nil) str.erase(*key.s);
      else str[*key.s] = v;
      return;
    case Value::Num: {
      if (std::isnan(key.n)) throw Error("table index is NaN");
      long long i;
      if (whole(key.n, &i) && i >= 1 && (size_t)i <= arr.size() + 1) {
        if ((size_t)i == arr.size() + 1) {
          if (v.k == Value::Nil) { num.erase((double)i); return; }
          arr.push_back(v);
          num.erase((double)i);
          // migrate following keys from the hash part
          for (auto it = num.find((double)(arr.size() + 1)); it != num.end(); it = num.find((double)(arr.size() + 1))) {
            arr.push_back(it->second);
            num.erase(it);
          }
        } else {
          arr[(size_t)i - 1] = v;
          if (v.k == Value::Nil && (size_t)i == arr.size()) {
            while (!arr.empty() && arr.back().k == Value::Nil) arr.pop_back();
          }
        }
        return;
      }
      if (v.k == Value::Nil) num.erase(key.n);
      else num[key.n] = v;
      return;
    }
    case Value::Nil: throw Error("table index is nil");
    default:
      for (auto &kv : other)
        if (raw_equal(kv.first, key)) { kv.second = v; return; }
      if (v.k != Value::Nil) other.emplace_back(key, v);
  }
}

Value Table::field(const std::string &k) const {
  auto it = str.find(k);
  return it == str.end() ? Value() : it->second;
}

void Table::set_field(const std::string &k, const Value &v) { set(Value::string(k), v); }

bool Table::next(const Value &key, Value *k, Value *v) const {
  // order: array part, number keys, string keys (sorted), other keys
  size_t stage = 0, pos = 0;
  std::vector<const std::string *> skeys;
  for (const auto &kv : str) skeys.push_back(&kv.first);
  std::sort(skeys.begin(), skeys.end(), [](const std::string *a, const std::string *b) { return *a < *b; });
  if (key.k != Value::Nil) {
    long long i;
    bool found = false;
    if (key.k == Value::Num && whole(key.n, &i) && i >= 1 && (size_t)i <= arr.size()) {
      stage = 0; pos = (size_t)i; found = true;
    } else if (key.k == Value::Num) {
      auto it = num.find(key.n);
      if (it != num.end()) { stage = 1; pos = (size_t)std::distance(num.begin(), it) + 1; found = true; }
    } else if (key.k == Value::Str) {
      for (size_t q = 0; q < skeys.size(); ++q)
        if (*skeys[q] == *key.s) { stage = 2; pos = q + 1; found = true; break; }
    } else {
      for (size_t q = 0; q < other.size(); ++q)
        if (raw_equal(other[q].first, key)) { stage = 3; pos = q + 1; found = true; break; }
    }
    if (!found) throw Error("invalid key to 'next'");
  }
  for (; stage < 4; ++stage, pos = 0) {
    if (stage == 0) {
      for (; pos < arr.size(); ++pos)
        if (arr[pos].k != Value::Nil) { *k = Value::number((double)(pos + 1)); *v = arr[pos]; return true; }
    } else if (stage == 1) {
      auto it = num.begin();
      std::advance(it, (long)std::min(pos, num.size()));
      if (it != num.end()) { *k = Value::number(it->first); *v = it->second; return true; }
    } else if (stage == 2) {
      if (pos < skeys.size()) { *k = Value::string(*skeys[pos]); *v = str.at(*skeys[pos]); return true; }
    } else {
      if (pos < other.size()) { *k = other[pos].first; *v = other[pos].second; return true; }
    }
  }
  return false;
}

// ------------------------------------------------------------------ lexer

namespace {

enum Tok {
  T_EOF, T_NAME, T_NUMBER, T_STRING,
  // keywords
  T_AND, T_BREAK, T_DO, T_ELSE, T_ELSEIF, T_END, T_FALSE, T_FOR, T_FUNCTION, T_GOTO, T_IF, T_IN, T_LOCAL, T_NIL,
  T_NOT, T_OR, T_REPEAT, T_RETURN, T_THEN, T_TRUE, T_UNTIL, T_WHILE,
  // symbols
  T_PLUS, T_MINUS, T_STAR, T_SLASH, T_DSLASH, T_PERCENT, T_CARET, T_HASH, T_CONCAT, T_DOTS, T_EQ, T_NE, T_LT, T_LE,
  T_GT, T_GE, T_ASSIGN, T_LPAREN, T_RPAREN, T_LBRACE, T_RBRACE, T_LBRACKET, T_RBRACKET, T_SEMI, T_COLON,
  T_DCOLON, T_COMMA, T_DOT, T_BITOP
};

struct Token {
  Tok t = T_EOF;
  std::string s;  // name or string contents
  double n = 0.0;
  int line = 1;
};

const std::unordered_map<std::string, Tok> &keywords() {
  static const std::unordered_map<std::string, Tok> k = {
      {"and", T_AND}, {"break", T_BREAK}, {"do", T_DO}, {"else", T_ELSE}, {"elseif", T_ELSEIF}, {"end", T_END},
      {"false", T_FALSE}, {"for", T_FOR}, {"function", T_FUNCTION}, {"goto", T_GOTO}, {"if", T_IF}, {"in", T_IN},
      {"local", T_LOCAL}, {"nil", T_NIL}, {"not", T_NOT}, {"or", T_OR}, {"repeat", T_REPEAT},
      {"return", T_RETURN}, {"then", T_THEN}, {"true", T_TRUE}, {"until", T_UNTIL}, {"while", T_WHILE}};
  return k;
}

class Lexer {
 public:
  Lexer(const std::string &src, const std::string &chunk) : s_(src), chunk_(chunk) {}
  std::vector<Token> run() {
    std::vector<Token> out;
    for (;;) {
      Token t = next();
      out.push_back(t);
      if (t.t == T_EOF) break;
    }
    return out;
  }

 private:
  const std::string &s_;
  std::string chunk_;
  size_t p_ = 0;
  int line_ = 1;

  [[noreturn]] void fail(const std::string &m) {
    throw Error(chunk_ + ":" + std::to_string(line_) + ": " + m);
  }
  char at(size_t q) const { return q < s_.size() ? s_[q] : '\0'; }

  // [[ ]] / [==[ ]==] long bracket starting at p_ (which holds '['); returns
  // false if this is not a long bracket
  bool long_bracket(std::string *out) {
    size_t q = p_ + 1;
    int level = 0;
    while (at(q) == '=') { ++level; ++q; }
    if (at(q) != '[') return false;
    q++;
    if (at(q) == '\r') ++q;
    if (at(q) == '\n') { ++line_; ++q; }
    const std::string close = "]" + std::string((size_t)level, '=') + "]";
    size_t e = s_.find(close, q);
    if (e == std::string::npos) fail("unfinished long string or comment");
    for (size_t r = q; r < e; ++r)
      if (s_[r] == '\n') ++line_;
    if (out) *out = s_.substr(q, e - q);
    p_ = e + close.size();
    return true;
  }

  Token next() {
    for (;;) {
      char c = at(p_);
      if (c == '\n') { ++line_; ++p_; continue; }
      if (c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v') { ++p_; continue; }
      if (c == '-' && at(p_ + 1) == '-') {
        p_ += 2;
        if (at(p_) == '[' && long_bracket(nullptr)) continue;
        while (at(p_) != '\n' && at(p_) != '\0') ++p_;
        continue;
      }
      break;
    }
    Token t;
    t.line = line_;
    char c = at(p_);
    if (c == '\0') { t.t = T_EOF; return t; }
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t q = p_;
      while (std::isalnum((unsigned char)at(q)) || at(q) == '_') ++q;
      t.s = s_.substr(p_, q - p_);
      p_ = q;
      auto it = keywords().find(t.s);
      t.t = it == keywords().end() ? T_NAME : it->second;
      return t;
    }
    if (std::isdigit((unsigned char)c) || (c == '.' && std::isdigit((unsigned char)at(p_ + 1)))) {
      const char *b = s_.c_str() + p_;
      char *e = nullptr;
      if (c == '0' && (at(p_ + 1) == 'x' || at(p_ + 1) == 'X')) {
        t.n = (double)std::strtoull(b, &e, 16);
      } else {
        t.n = std::strtod(b, &e);
      }
      if (e == b) fail("malformed number");
      p_ += (size_t)(e - b);
      if (std::isalpha((unsigned char)at(p_)) || at(p_) == '_') fail("malformed number");
      t.t = T_NUMBER;
      return t;
    }
    if (c == '"' || c == '\'') {
      const char quote = c;
      ++p_;
      std::string v;
      for (;;) {
        char d = at(p_);
        if (d == '\0' || d == '\n') fail("unfinished string");
        ++p_;
        if (d == quote) break;
        if (d == '\\') {
          char e = at(p_++);
          switch (e) {
            case 'n': v += '\n'; break;
            case 't': v += '\t'; break;
            case 'r': v += '\r'; break;
            case 'a': v += '\a'; break;
            case 'b': v += '\b'; break;
            case 'f': v += '\f'; break;
            case 'v': v += '\v'; break;
            case '\\': v += '\\'; break;
            case '"': v += '"'; break;
            case '\'': v += '\''; break;
            case '\n': v += '\n'; ++line_; break;
            default:
              if (std::isdigit((unsigned char)e)) {
                int code = e - '0', nd = 1;
                while (nd < 3 && std::isdigit((unsigned char)at(p_))) { code = code * 10 + (at(p_++) - '0'); ++nd; }
                if (code > 255) fail("decimal escape too large");
                v += (char)code;
              } else {
                fail(std::string("invalid escape sequence '\\") + e + "'");
              }
          }
        } else {
          v += d;
        }
      }
      t.t = T_STRING;
      t.s = v;
      return t;
    }
    if (c == '[' && (at(p_ + 1) == '[' || at(p_ + 1) == '=')) {
      std::string v;
      if (long_bracket(&v)) { t.t = T_STRING; t.s = v; return t; }
    }
    auto two = [&](char a, char b) { return at(p_) == a && at(p_ + 1) == b; };
    struct Sym { const char *s; Tok t; };
    static const Sym syms[] = {{"...", T_DOTS}, {"..", T_CONCAT}, {"==", T_EQ}, {"~=", T_NE}, {"<=", T_LE},
                               {">=", T_GE}, {"//", T_DSLASH}, {"::", T_DCOLON}, {"<<", T_BITOP}, {">>", T_BITOP},
                               {"+", T_PLUS}, {"-", T_MINUS}, {"*", T_STAR}, {"/", T_SLASH}, {"%", T_PERCENT},
                               {"^", T_CARET}, {"#", T_HASH}, {"<", T_LT}, {">", T_GT}, {"=", T_ASSIGN},
                               {"(", T_LPAREN}, {")", T_RPAREN}, {"{", T_LBRACE}, {"}", T_RBRACE},
                               {"[", T_LBRACKET}, {"]", T_RBRACKET}, {";", T_SEMI}, {":", T_COLON}, {",", T_COMMA},
                               {".", T_DOT}, {"&", T_BITOP}, {"|", T_BITOP}, {"~", T_BITOP}};
    (void)two;
    for (const Sym &sy : syms) {
      const size_t n = std::strlen(sy.s);
      if (s_.compare(p_, n, sy.s) == 0) {
        p_ += n;
        t.t = sy.t;
        t.s = sy.s;
        return t;
      }
    }
    fail(std::string("unexpected symbol '") + c + "'");
  }
};

// ------------------------------------------------------------------ syntax tree

struct Expr;
struct Stmt;
using ExprP = std::unique_ptr<Expr>;
using StmtP = std::unique_ptr<Stmt>;
using Block = std::vector<StmtP>;

enum class BinOp { Add, Sub, Mul, Div, IDiv, Mod, Pow, Concat, Eq, Ne, Lt, Le, Gt, Ge, And, Or };
enum class UnOp { Neg, Not, Len };

struct Expr {
  enum K { Nil, True, False, Number, String, Func, Name, Index, Call, Method, Bin, Un, Table, Paren, Dots } k;
  int line = 0;
  double num = 0.0;
  Value strv;                       // String constant
  const std::string *name = nullptr; // Name, Method (method name), Index with constant string key
  std::vector<ExprP> kids;          // Index: obj, key; Call: fn, args...; Method: obj, args...; Bin/Un/Paren
  BinOp bop = BinOp::Add;
  UnOp uop = UnOp::Neg;
  // Table: positional items kids[i] with keys[i] == nullptr, keyed items with keys[i]
  std::vector<ExprP> keys;
  const FuncBody *body = nullptr;   // Func
  explicit Expr(K kind) : k(kind) {}
};

}  // namespace

struct FuncBody {
  std::vector<const std::string *> params;
  bool is_method = false;
  Block body;
  std::string name;
  int line = 0;
};

namespace {

struct Stmt {
  enum K { Local, Assign, CallS, If, While, Repeat, NumFor, GenFor, FuncDef, LocalFunc, Return, Break, Do } k;
  int line = 0;
  std::vector<const std::string *> names;   // Local, NumFor (1), GenFor, LocalFunc (1)
  std::vector<ExprP> targets;               // Assign
  std::vector<ExprP> exprs;                 // Local, Assign, Return, GenFor, NumFor (start, limit, step), If conds
  ExprP e;                                  // CallS, While / Repeat condition, FuncDef target
  std::vector<Block> blocks;                // If (then blocks + optional else), loops (one)
  const FuncBody *body = nullptr;           // FuncDef, LocalFunc
  explicit Stmt(K kind) : k(kind) {}
};

}  // namespace

struct Chunk {
  std::vector<std::unique_ptr<FuncBody>> bodies;  // every function of the script
  std::unordered_set<std::string> names;          // interned identifiers
  const FuncBody *main = nullptr;
};

namespace {

class Parser {
 public:
  Parser(std::vector<Token> toks, Chunk &c, const std::string &chunkname)
      : t_(std::move(toks)), c_(c), chunk_(chunkname) {}

  const FuncBody *parse_main() {
    auto fb = std::make_unique<FuncBody>();
    fb->name = "main chunk";
    fb->body = block();
    expect(T_EOF, "'<eof>'");
    const FuncBody *r = fb.get();
    c_.bodies.push_back(std::move(fb));
    return r;
  }

 private:
  std::vector<Token> t_;
  size_t p_ = 0;
  Chunk &c_;
  std::string chunk_;

  const Token &cur() const { return t_[p_]; }
  const Token &peek(size_t a = 1) const { return t_[std::min(p_ + a, t_.size() - 1)]; }
  bool is(Tok t) const { return cur().t == t; }
  bool accept(Tok t) {
    if (!is(t)) return false;
    ++p_;
    return true;
  }
  [[noreturn]] void fail(const std::string &m) {
    throw Error(chunk_ + ":" + std::to_string(cur().line) + ": " + m + " near '" +
                (cur().t == T_EOF ? std::string("<eof>") : (cur().s.empty() ? std::to_string(cur().n) : cur().s)) + "'");
  }
  void expect(Tok t, const char *what) {
    if (!accept(t)) fail(std::string(what) + " expected");
  }
  const std::string *intern(const std::string &s) { return &*c_.names.insert(s).first; }
  const std::string *name() {
    if (!is(T_NAME)) fail("<name> expected");
    return intern(t_[p_++].s);
  }

  static bool block_end(Tok t) { return t == T_EOF || t == T_END || t == T_ELSE || t == T_ELSEIF || t == T_UNTIL; }

  Block block() {
    Block b;
    while (!block_end(cur().t)) {
      if (is(T_RETURN)) {
        b.push_back(ret());
        break;
      }
      if (accept(T_SEMI)) continue;
      b.push_back(statement());
    }
    return b;
  }

  StmtP ret() {
    auto s = std::make_unique<Stmt>(Stmt::Return);
    s->line = cur().line;
    expect(T_RETURN, "'return'");
    if (!block_end(cur().t) && !is(T_SEMI)) s->exprs = exprlist();
    accept(T_SEMI);
    return s;
  }

  StmtP statement() {
    const int line = cur().line;
    StmtP s;
    switch (cur().t) {
      case T_IF: {
        ++p_;
        s = std::make_unique<Stmt>(Stmt::If);
        s->exprs.push_back(expr());
        expect(T_THEN, "'then'");
        s->blocks.push_back(block());
        while (is(T_ELSEIF)) {
          ++p_;
          s->exprs.push_back(expr());
          expect(T_THEN, "'then'");
          s->blocks.push_back(block());
        }
        if (accept(T_ELSE)) s->blocks.push_back(block());
        expect(T_END, "'end'");
        break;
      }
      case T_WHILE: {
        ++p_;
        s = std::make_unique<Stmt>(Stmt::While);
        s->e = expr();
        expect(T_DO, "'do'");
        s->blocks.push_back(block());
        expect(T_END, "'end'");
        break;
      }
      case T_DO: {
        ++p_;
        s = std::make_unique<Stmt>(Stmt::Do);
        s->blocks.push_back(block());
        expect(T_END, "'end'");
        break;
      }
      case T_REPEAT: {
        ++p_;
        s = std::make_unique<Stmt>(Stmt::Repeat);
        s->blocks.push_back(block());
        expect(T_UNTIL, "'until'");
        s->e = expr();
        break;
      }
      case T_FOR: {
        ++p_;
        const std::string *n1 = name();
        if (accept(T_ASSIGN)) {
          s = std::make_unique<Stmt>(Stmt::NumFor);
          s->names.push_back(n1);
          s->exprs.push_back(expr());
          expect(T_COMMA, "','");
          s->exprs.push_back(expr());
          if (accept(T_COMMA)) s->exprs.push_back(expr());
        } else {
          s = std::make_unique<Stmt>(Stmt::GenFor);
          s->names.push_back(n1);
          while (accept(T_COMMA)) s->names.push_back(name());
          expect(T_IN, "'in'");
          s->exprs = exprlist();
        }
        expect(T_DO, "'do'");
        s->blocks.push_back(block());
        expect(T_END, "'end'");
        break;
      }
      case T_FUNCTION: {
        ++p_;
        s = std::make_unique<Stmt>(Stmt::FuncDef);
        // funcname: Name {'.' Name} [':' Name]
        auto e = std::make_unique<Expr>(Expr::Name);
        e->line = line;
        e->name = name();
        std::string full = *e->name;
        bool method = false;
        while (is(T_DOT) || is(T_COLON)) {
          method = is(T_COLON);
          ++p_;
          auto ix = std::make_unique<Expr>(Expr::Index);
          ix->line = line;
          ix->name = name();
          full += (method ? ":" : ".") + *ix->name;
          auto key = std::make_unique<Expr>(Expr::String);
          key->strv = Value::string(*ix->name);
          ix->kids.push_back(std::move(e));
          ix->kids.push_back(std::move(key));
          e = std::move(ix);
          if (method) break;
        }
        s->e = std::move(e);
        s->body = funcbody(method, full, line);
        break;
      }
      case T_LOCAL: {
        ++p_;
        if (accept(T_FUNCTION)) {
          s = std::make_unique<Stmt>(Stmt::LocalFunc);
          s->names.push_back(name());
          s->body = funcbody(false, *s->names[0], line);
        } else {
          s = std::make_unique<Stmt>(Stmt::Local);
          s->names.push_back(name());
          while (accept(T_COMMA)) s->names.push_back(name());
          if (accept(T_ASSIGN)) s->exprs = exprlist();
        }
        break;
      }
      case T_RETURN: fail("'return' must be the last statement of a block");
      case T_BREAK: ++p_; s = std::make_unique<Stmt>(Stmt::Break); break;
      case T_GOTO: case T_DCOLON: fail("goto and labels are not supported");
      default: {
        ExprP e = suffixedexp();
        if (is(T_ASSIGN) || is(T_COMMA)) {
          s = std::make_unique<Stmt>(Stmt::Assign);
          s->targets.push_back(std::move(e));
          while (accept(T_COMMA)) s->targets.push_back(suffixedexp());
          expect(T_ASSIGN, "'='");
          s->exprs = exprlist();
          for (auto &tg : s->targets)
            if (tg->k != Expr::Name && tg->k != Expr::Index) fail("syntax error (cannot assign)");
        } else {
          if (e->k != Expr::Call && e->k != Expr::Method) fail("syntax error (statement expected)");
          s = std::make_unique<Stmt>(Stmt::CallS);
          s->e = std::move(e);
        }
      }
    }
    s->line = line;
    return s;
  }

  const FuncBody *funcbody(bool method, const std::string &nm, int line) {
    auto fb = std::make_unique<FuncBody>();
    fb->name = nm;
    fb->line = line;
    fb->is_method = method;
    if (method) fb->params.push_back(intern("self"));
    expect(T_LPAREN, "'('");
    if (!is(T_RPAREN)) {
      do {
        if (is(T_DOTS)) fail("varargs (...) are not supported");
        fb->params.push_back(name());
      } while (accept(T_COMMA));
    }
    expect(T_RPAREN, "')'");
    fb->body = block();
    expect(T_END, "'end'");
    const FuncBody *r = fb.get();
    c_.bodies.push_back(std::move(fb));
    return r;
  }

  std::vector<ExprP> exprlist() {
    std::vector<ExprP> v;
    v.push_back(expr());
    while (accept(T_COMMA)) v.push_back(expr());
    return v;
  }

  ExprP primaryexp() {
    const int line = cur().line;
    if (is(T_NAME)) {
      auto e = std::make_unique<Expr>(Expr::Name);
      e->line = line;
      e->name = name();
      return e;
    }
    if (accept(T_LPAREN)) {
      auto e = std::make_unique<Expr>(Expr::Paren);
      e->line = line;
      e->kids.push_back(expr());
      expect(T_RPAREN, "')'");
      return e;
    }
    fail("unexpected symbol");
  }

  ExprP suffixedexp() {
    ExprP e = primaryexp();
    for (;;) {
      const int line = cur().line;
      if (accept(T_DOT)) {
        auto ix = std::make_unique<Expr>(Expr::Index);
        ix->line = line;
        auto key = std::make_unique<Expr>(Expr::String);
        key->strv = Value::string(*name());
        ix->kids.push_back(std::move(e));
        ix->kids.push_back(std::move(key));
        e = std::move(ix);
      } else if (accept(T_LBRACKET)) {
        auto ix = std::make_unique<Expr>(Expr::Index);
        ix->line = line;
        ix->kids.push_back(std::move(e));
        ix->kids.push_back(expr());
        expect(T_RBRACKET, "']'");
        e = std::move(ix);
      } else if (accept(T_COLON)) {
        auto m = std::make_unique<Expr>(Expr::Method);
        m->line = line;
        m->name = name();
        m->kids.push_back(std::move(e));
        callargs(m->kids);
        e = std::move(m);
      } else if (is(T_LPAREN) || is(T_STRING) || is(T_LBRACE)) {
        auto c = std::make_unique<Expr>(Expr::Call);
        c->line = line;
        c->kids.push_back(std::move(e));
        callargs(c->kids);
        e = std::move(c);
      } else {
        return e;
      }
    }
  }

  void callargs(std::vector<ExprP> &out) {
    if (is(T_STRING)) {
      auto s = std::make_unique<Expr>(Expr::String);
      s->strv = Value::string(t_[p_++].s);
      out.push_back(std::move(s));
    } else if (is(T_LBRACE)) {
      out.push_back(table());
    } else {
      expect(T_LPAREN, "function arguments");
      if (!is(T_RPAREN)) {
        out.push_back(expr());
        while (accept(T_COMMA)) out.push_back(expr());
      }
      expect(T_RPAREN, "')'");
    }
  }

  ExprP table() {
    auto t = std::make_unique<Expr>(Expr::Table);
    t->line = cur().line;
    expect(T_LBRACE, "'{'");
    while (!is(T_RBRACE)) {
      if (is(T_LBRACKET)) {
        ++p_;
        t->keys.push_back(expr());
        expect(T_RBRACKET, "']'");
        expect(T_ASSIGN, "'='");
        t->kids.push_back(expr());
      } else if (is(T_NAME) && peek().t == T_ASSIGN) {
        auto k = std::make_unique<Expr>(Expr::String);
        k->strv = Value::string(t_[p_].s);
        p_ += 2;
        t->keys.push_back(std::move(k));
        t->kids.push_back(expr());
      } else {
        t->keys.push_back(nullptr);
        t->kids.push_back(expr());
      }
      if (!accept(T_COMMA) && !accept(T_SEMI)) break;
    }
    expect(T_RBRACE, "'}'");
    return t;
  }

  ExprP simpleexp() {
    const int line = cur().line;
    ExprP e;
    switch (cur().t) {
      case T_NUMBER: e = std::make_unique<Expr>(Expr::Number); e->num = cur().n; ++p_; break;
      case T_STRING: e = std::make_unique<Expr>(Expr::String); e->strv = Value::string(cur().s); ++p_; break;
      case T_NIL: e = std::make_unique<Expr>(Expr::Nil); ++p_; break;
      case T_TRUE: e = std::make_unique<Expr>(Expr::True); ++p_; break;
      case T_FALSE: e = std::make_unique<Expr>(Expr::False); ++p_; break;
      case T_DOTS: fail("varargs (...) are not supported");
      case T_LBRACE: return table();
      case T_FUNCTION:
        ++p_;
        e = std::make_unique<Expr>(Expr::Func);
        e->body = funcbody(false, "anonymous function", line);
        break;
      default: return suffixedexp();
    }
    e->line = line;
    return e;
  }

  // binary operator priorities (left, right): right < left makes the
  // operator right-associative
  static bool binop(Tok t, BinOp *op, int *lp, int *rp) {
    switch (t) {
      case T_OR: *op = BinOp::Or; *lp = 1; *rp = 1; return true;
      case T_AND: *op = BinOp::And; *lp = 2; *rp = 2; return true;
      case T_LT: *op = BinOp::Lt; *lp = 3; *rp = 3; return true;
      case T_GT: *op = BinOp::Gt; *lp = 3; *rp = 3; return true;
      case T_LE: *op = BinOp::Le; *lp = 3; *rp = 3; return true;
      case T_GE: *op = BinOp::Ge; *lp = 3; *rp = 3; return true;
      case T_NE: *op = BinOp::Ne; *lp = 3; *rp = 3; return true;
      case T_EQ: *op = BinOp::Eq; *lp = 3; *rp = 3; return true;
      case T_CONCAT: *op = BinOp::Concat; *lp = 9; *rp = 8; return true;
      case T_PLUS: *op = BinOp::Add; *lp = 10; *rp = 10; return true;
      case T_MINUS: *op = BinOp::Sub; *lp = 10; *rp = 10; return true;
      case T_STAR: *op = BinOp::Mul; *lp = 11; *rp = 11; return true;
      case T_SLASH: *op = BinOp::Div; *lp = 11; *rp = 11; return true;
      case T_DSLASH: *op = BinOp::IDiv; *lp = 11; *rp = 11; return true;
      case T_PERCENT: *op = BinOp::Mod; *lp = 11; *rp = 11; return true;
      case T_CARET: *op = BinOp::Pow; *lp = 14; *rp = 13; return true;
      default: return false;
    }
  }
  static constexpr int UNARY_PRIORITY = 12;

  ExprP subexpr(int limit) {
    ExprP e;
    const int line = cur().line;
    if (is(T_NOT) || is(T_MINUS) || is(T_HASH)) {
      const Tok u = cur().t;
      ++p_;
      e = std::make_unique<Expr>(Expr::Un);
      e->line = line;
      e->uop = u == T_NOT ? UnOp::Not : (u == T_MINUS ? UnOp::Neg : UnOp::Len);
      e->kids.push_back(subexpr(UNARY_PRIORITY));
    } else if (is(T_BITOP)) {
      fail("bitwise operators are not supported");
    } else {
      e = simpleexp();
    }
    BinOp op;
    int lp, rp;
    while (binop(cur().t, &op, &lp, &rp) && lp > limit) {
      const int l2 = cur().line;
      ++p_;
      auto b = std::make_unique<Expr>(Expr::Bin);
      b->line = l2;
      b->bop = op;
      b->kids.push_back(std::move(e));
      b->kids.push_back(subexpr(rp));
      e = std::move(b);
    }
    if (is(T_BITOP)) fail("bitwise operators are not supported");
    return e;
  }

  ExprP expr() { return subexpr(0); }
};

}  // namespace

// ------------------------------------------------------------------ evaluation

struct Env {
  std::vector<std::pair<const std::string *, std::shared_ptr<Value>>> vars;
  std::shared_ptr<Env> up;
  explicit Env(std::shared_ptr<Env> parent) : up(std::move(parent)) {}
  std::shared_ptr<Value> *find(const std::string *n) {
    for (Env *e = this; e; e = e->up.get())
      for (auto it = e->vars.rbegin(); it != e->vars.rend(); ++it)
        if (it->first == n) return &it->second;
    return nullptr;
  }
  void declare(const std::string *n, Value v) { vars.emplace_back(n, std::make_shared<Value>(std::move(v))); }
};

namespace {

using EnvP = std::shared_ptr<Env>;

std::string num_to_string(double d) {
  char buf[64];
  long long i;
  if (whole(d, &i)) {
    std::snprintf(buf, sizeof buf, "%lld", i);
  } else if (std::isinf(d)) {
    return d > 0 ? "inf" : "-inf";
  } else if (std::isnan(d)) {
    return d > 0 ? "nan" : "-nan";
  } else {
    std::snprintf(buf, sizeof buf, "%.14g", d);
  }
  return buf;
}

bool str_to_number(const std::string &s, double *out) {
  const char *b = s.c_str();
  while (std::isspace((unsigned char)*b)) ++b;
  if (!*b) return false;
  char *e = nullptr;
  double v;
  if (b[0] == '0' && (b[1] == 'x' || b[1] == 'X')) v = (double)std::strtoull(b, &e, 16);
  else v = std::strtod(b, &e);
  if (e == b) return false;
  while (std::isspace((unsigned char)*e)) ++e;
  if (*e) return false;
  *out = v;
  return true;
}

enum class Flow { Normal, Break, Return };

class Exec {
 public:
  explicit Exec(std::shared_ptr<Table> g) : g_(std::move(g)) {}

  Values call(const Value &fv, Values &args, int line) {
    if (fv.k != Value::Fn) fail(line, std::string("attempt to call a ") + fv.type_name() + " value");
    const Function &fn = *fv.f;
    if (fn.builtin) return fn.builtin(args);
    if (++depth_ > 200) { --depth_; fail(line, "stack overflow (more than 200 nested calls)"); }
    auto env = std::make_shared<Env>(fn.env);
    const FuncBody &b = *fn.body;
    for (size_t i = 0; i < b.params.size(); ++i) env->declare(b.params[i], i < args.size() ? args[i] : Value());
    Values ret;
    Flow fl = block(b.body, env, &ret, false);
    --depth_;
    if (fl == Flow::Break) fail(b.line, "break outside a loop");
    return ret;
  }

  Flow block(const Block &b, const EnvP &parent, Values *ret, bool new_scope = true) {
    EnvP env = new_scope ? std::make_shared<Env>(parent) : parent;
    for (const auto &s : b) {
      Flow f = stmt(*s, env, ret);
      if (f != Flow::Normal) return f;
    }
    return Flow::Normal;
  }

 private:
  std::shared_ptr<Table> g_;
  int depth_ = 0;

  [[noreturn]] static void fail(int line, const std::string &m) {
    throw Error("line " + std::to_string(line) + ": " + m);
  }

  double tonum(const Value &v, int line, const char *what) {
    if (v.k == Value::Num) return v.n;
    double d;
    if (v.k == Value::Str && str_to_number(*v.s, &d)) return d;
    fail(line, std::string("attempt to perform arithmetic on a ") + v.type_name() + " value" +
                   (what ? std::string(" (") + what + ")" : ""));
  }

  Value lookup(const std::string *n, const EnvP &env) {
    if (auto *p = env ? env->find(n) : nullptr) return **p;
    return g_->field(*n);
  }

  void assign(const Expr &t, const Value &v, const EnvP &env) {
    if (t.k == Expr::Name) {
      if (auto *p = env ? env->find(t.name) : nullptr) {
        **p = v;
        return;
      }
      g_->set_field(*t.name, v);
      return;
    }
    const Value obj = eval(*t.kids[0], env);
    if (obj.k != Value::Tab) fail(t.line, std::string("attempt to index a ") + obj.type_name() + " value");
    obj.t->set(eval(*t.kids[1], env), v);
  }

  // values of an expression list, the last one expanded
  Values evals(const std::vector<ExprP> &es, const EnvP &env, size_t from = 0) {
    Values out;
    for (size_t i = from; i < es.size(); ++i) {
      const Expr &e = *es[i];
      if (i + 1 == es.size() && (e.k == Expr::Call || e.k == Expr::Method)) {
        Values r = multi(e, env);
        out.insert(out.end(), r.begin(), r.end());
      } else {
        out.push_back(eval(e, env));
      }
    }
    return out;
  }

  Values multi(const Expr &e, const EnvP &env) {
    if (e.k == Expr::Call) {
      Value fn = eval(*e.kids[0], env);
      Values args = evals(e.kids, env, 1);
      return call(fn, args, e.line);
    }
    if (e.k == Expr::Method) {
      Value obj = eval(*e.kids[0], env);
      if (obj.k != Value::Tab) fail(e.line, std::string("attempt to index a ") + obj.type_name() + " value");
      Value fn = obj.t->field(*e.name);
      Values args;
      args.push_back(obj);
      Values rest = evals(e.kids, env, 1);
      args.insert(args.end(), rest.begin(), rest.end());
      return call(fn, args, e.line);
    }
    return {eval(e, env)};
  }

  static bool less(const Value &a, const Value &b, int line, bool eq) {
    if (a.k == Value::Num && b.k == Value::Num) return eq ? a.n <= b.n : a.n < b.n;
    if (a.k == Value::Str && b.k == Value::Str) return eq ? *a.s <= *b.s : *a.s < *b.s;
    fail(line, std::string("attempt to compare ") + a.type_name() + " with " + b.type_name());
  }

  Value eval(const Expr &e, const EnvP &env) {
    switch (e.k) {
      case Expr::Nil: return Value();
      case Expr::True: return Value::boolean(true);
      case Expr::False: return Value::boolean(false);
      case Expr::Number: return Value::number(e.num);
      case Expr::String: return e.strv;
      case Expr::Name: return lookup(e.name, env);
      case Expr::Paren: return eval(*e.kids[0], env);
      case Expr::Index: {
        const Value obj = eval(*e.kids[0], env);
        if (obj.k == Value::Str) {  // string methods through the string table ("abc"):upper() aside
          const Value st = g_->field("string");
          if (st.k == Value::Tab) return st.t->get(eval(*e.kids[1], env));
        }
        if (obj.k != Value::Tab) fail(e.line, std::string("attempt to index a ") + obj.type_name() + " value");
        return obj.t->get(eval(*e.kids[1], env));
      }
      case Expr::Call:
      case Expr::Method: {
        Values r = multi(e, env);
        return r.empty() ? Value() : r[0];
      }
      case Expr::Func: {
        auto f = std::make_shared<Function>();
        f->body = e.body;
        f->env = env;
        return Value::function(f);
      }
      case Expr::Table: {
        auto t = std::make_shared<Table>();
        double pos = 1;
        for (size_t i = 0; i < e.kids.size(); ++i) {
          if (e.keys[i]) {
            t->set(eval(*e.keys[i], env), eval(*e.kids[i], env));
          } else if (i + 1 == e.kids.size() && (e.kids[i]->k == Expr::Call || e.kids[i]->k == Expr::Method)) {
            for (const Value &v : multi(*e.kids[i], env)) t->set(Value::number(pos++), v);
          } else {
            t->set(Value::number(pos++), eval(*e.kids[i], env));
          }
        }
        return Value::table(t);
      }
      case Expr::Un: {
        const Value a = eval(*e.kids[0], env);
        switch (e.uop) {
          case UnOp::Not: return Value::boolean(!a.truthy());
          case UnOp::Neg: return Value::number(-tonum(a, e.line, nullptr));
          default:
            if (a.k == Value::Str) return Value::number((double)a.s->size());
            if (a.k == Value::Tab) return Value::number((double)a.t->length());
            fail(e.line, std::string("attempt to get length of a ") + a.type_name() + " value");
        }
      }
      case Expr::Bin: {
        if (e.bop == BinOp::And) {
          Value a = eval(*e.kids[0], env);
          return a.truthy() ? eval(*e.kids[1], env) : a;
        }
        if (e.bop == BinOp::Or) {
          Value a = eval(*e.kids[0], env);
          return a.truthy() ? a : eval(*e.kids[1], env);
        }
        const Value a = eval(*e.kids[0], env);
        const Value b = eval(*e.kids[1], env);
        switch (e.bop) {
          case BinOp::Add: return Value::number(tonum(a, e.line, nullptr) + tonum(b, e.line, nullptr));
          case BinOp::Sub: return Value::number(tonum(a, e.line, nullptr) - tonum(b, e.line, nullptr));
          case BinOp::Mul: return Value::number(tonum(a, e.line, nullptr) * tonum(b, e.line, nullptr));
          case BinOp::Div: return Value::number(tonum(a, e.line, nullptr) / tonum(b, e.line, nullptr));
          case BinOp::IDiv:
            return Value::number(std::floor(tonum(a, e.line, nullptr) / tonum(b, e.line, nullptr)));
          case BinOp::Mod: {
            const double x = tonum(a, e.line, nullptr), y = tonum(b, e.line, nullptr);
            double m = std::fmod(x, y);
            if (m != 0 && ((m < 0) != (y < 0))) m += y;  // Lua: the sign of the divisor
            return Value::number(m);
          }
          case BinOp::Pow: return Value::number(std::pow(tonum(a, e.line, nullptr), tonum(b, e.line, nullptr)));
          case BinOp::Concat: {
            auto part = [&](const Value &v) -> std::string {
              if (v.k == Value::Str) return *v.s;
              if (v.k == Value::Num) return num_to_string(v.n);
              fail(e.line, std::string("attempt to concatenate a ") + v.type_name() + " value");
            };
            return Value::string(part(a) + part(b));
          }
          case BinOp::Eq: return Value::boolean(raw_equal(a, b));
          case BinOp::Ne: return Value::boolean(!raw_equal(a, b));
          case BinOp::Lt: return Value::boolean(less(a, b, e.line, false));
          case BinOp::Le: return Value::boolean(less(a, b, e.line, true));
          case BinOp::Gt: return Value::boolean(less(b, a, e.line, false));
          case BinOp::Ge: return Value::boolean(less(b, a, e.line, true));
          default: break;
        }
        fail(e.line, "bad operator");
      }
      case Expr::Dots: fail(e.line, "varargs are not supported");
    }
    fail(e.line, "bad expression");
  }

  Flow stmt(const Stmt &s, EnvP &env, Values *ret) {
    switch (s.k) {
      case Stmt::Local: {
        Values v = evals(s.exprs, env);
        for (size_t i = 0; i < s.names.size(); ++i) env->declare(s.names[i], i < v.size() ? v[i] : Value());
        return Flow::Normal;
      }
      case Stmt::Assign: {
        Values v = evals(s.exprs, env);
        for (size_t i = 0; i < s.targets.size(); ++i) assign(*s.targets[i], i < v.size() ? v[i] : Value(), env);
        return Flow::Normal;
      }
      case Stmt::CallS: (void)multi(*s.e, env); return Flow::Normal;
      case Stmt::If: {
        for (size_t i = 0; i < s.exprs.size(); ++i)
          if (eval(*s.exprs[i], env).truthy()) return block(s.blocks[i], env, ret);
        if (s.blocks.size() > s.exprs.size()) return block(s.blocks.back(), env, ret);
        return Flow::Normal;
      }
      case Stmt::While:
        while (eval(*s.e, env).truthy()) {
          Flow f = block(s.blocks[0], env, ret);
          if (f == Flow::Break) break;
          if (f == Flow::Return) return f;
        }
        return Flow::Normal;
      case Stmt::Repeat:
        for (;;) {
          // the condition sees the body's locals
          auto inner = std::make_shared<Env>(env);
          Flow f = Flow::Normal;
          for (const auto &st : s.blocks[0]) {
            f = stmt(*st, inner, ret);
            if (f != Flow::Normal) break;
          }
          if (f == Flow::Break) break;
          if (f == Flow::Return) return f;
          if (eval(*s.e, inner).truthy()) break;
        }
        return Flow::Normal;
      case Stmt::NumFor: {
        const double a = tonum(eval(*s.exprs[0], env), s.line, "'for' initial value");
        const double lim = tonum(eval(*s.exprs[1], env), s.line, "'for' limit");
        const double st = s.exprs.size() > 2 ? tonum(eval(*s.exprs[2], env), s.line, "'for' step") : 1.0;
        if (st == 0) fail(s.line, "'for' step is zero");
        for (double i = a; st > 0 ? i <= lim : i >= lim; i += st) {
          auto inner = std::make_shared<Env>(env);
          inner->declare(s.names[0], Value::number(i));
          Flow f = block(s.blocks[0], inner, ret, false);
          if (f == Flow::Break) break;
          if (f == Flow::Return) return f;
        }
        return Flow::Normal;
      }
      case Stmt::GenFor: {
        Values init = evals(s.exprs, env);
        init.resize(3);
        const Value fn = init[0], state = init[1];
        Value ctl = init[2];
        for (;;) {
          Values args{state, ctl};
          Values r = call(fn, args, s.line);
          if (r.empty() || r[0].k == Value::Nil) break;
          ctl = r[0];
          auto inner = std::make_shared<Env>(env);
          for (size_t i = 0; i < s.names.size(); ++i) inner->declare(s.names[i], i < r.size() ? r[i] : Value());
          Flow f = block(s.blocks[0], inner, ret, false);
          if (f == Flow::Break) break;
          if (f == Flow::Return) return f;
        }
        return Flow::Normal;
      }
      case Stmt::FuncDef: {
        auto f = std::make_shared<Function>();
        f->body = s.body;
        f->env = env;
        assign(*s.e, Value::function(f), env);
        return Flow::Normal;
      }
      case Stmt::LocalFunc: {
        auto f = std::make_shared<Function>();
        f->body = s.body;
        env->declare(s.names[0], Value());
        f->env = env;  // the function sees its own name (recursion)
        **env->find(s.names[0]) = Value::function(f);
        return Flow::Normal;
      }
      case Stmt::Return: *ret = evals(s.exprs, env); return Flow::Return;
      case Stmt::Break: return Flow::Break;
      case Stmt::Do: return block(s.blocks[0], env, ret);
    }
    return Flow::Normal;
  }
};

// ------------------------------------------------------------------ library

Value builtin(std::function<Values(Values &)> f) {
  auto fn = std::make_shared<Function>();
  fn->builtin = std::move(f);
  return Value::function(fn);
}

double arg_num(const Values &a, size_t i, const char *fname) {
  if (i < a.size() && a[i].k == Value::Num) return a[i].n;
  double d;
  if (i < a.size() && a[i].k == Value::Str && str_to_number(*a[i].s, &d)) return d;
  throw Error(std::string("bad argument #") + std::to_string(i + 1) + " to '" + fname + "' (number expected, got " +
              (i < a.size() ? a[i].type_name() : "no value") + ")");
}

std::string arg_str(const Values &a, size_t i, const char *fname) {
  if (i < a.size() && a[i].k == Value::Str) return *a[i].s;
  if (i < a.size() && a[i].k == Value::Num) return num_to_string(a[i].n);
  throw Error(std::string("bad argument #") + std::to_string(i + 1) + " to '" + fname + "' (string expected, got " +
              (i < a.size() ? a[i].type_name() : "no value") + ")");
}

std::string format(const Values &a) {
  const std::string f = arg_str(a, 0, "format");
  std::string out;
  size_t ai = 1;
  for (size_t i = 0; i < f.size(); ++i) {
    if (f[i] != '%') { out += f[i]; continue; }
    if (i + 1 < f.size() && f[i + 1] == '%') { out += '%'; ++i; continue; }
    size_t j = i + 1;
    while (j < f.size() && std::strchr("-+ #0123456789.", f[j])) ++j;
    if (j >= f.size()) throw Error("invalid conversion in format string");
    const char conv = f[j];
    std::string spec = f.substr(i, j - i);
    char buf[512];
    switch (conv) {
      case 'd': case 'i': case 'x': case 'X': case 'o': case 'c':
        std::snprintf(buf, sizeof buf, (spec + "ll" + conv).c_str(), (long long)arg_num(a, ai++, "format"));
        break;
      case 'f': case 'F': case 'e': case 'E': case 'g': case 'G': case 'a': case 'A':
        std::snprintf(buf, sizeof buf, (spec + conv).c_str(), arg_num(a, ai++, "format"));
        break;
      case 's': case 'q':
        std::snprintf(buf, sizeof buf, (spec + 's').c_str(), arg_str(a, ai++, "format").c_str());
        break;
      default: throw Error(std::string("invalid conversion '%") + conv + "' to 'format'");
    }
    out += buf;
    i = j;
  }
  return out;
}

void open_libs(const std::shared_ptr<Table> &g, Exec *ex) {
  auto math = std::make_shared<Table>();
  auto fn1 = [&](const char *n, double (*f)(double)) {
    math->set_field(n, builtin([f, n](Values &a) { return Values{Value::number(f(arg_num(a, 0, n)))}; }));
  };
  fn1("abs", [](double x) { return std::fabs(x); });
  fn1("sqrt", [](double x) { return std::sqrt(x); });
  fn1("exp", [](double x) { return std::exp(x); });
  fn1("sin", [](double x) { return std::sin(x); });
  fn1("cos", [](double x) { return std::cos(x); });
  fn1("tan", [](double x) { return std::tan(x); });
  fn1("asin", [](double x) { return std::asin(x); });
  fn1("acos", [](double x) { return std::acos(x); });
  fn1("floor", [](double x) { return std::floor(x); });
  fn1("ceil", [](double x) { return std::ceil(x); });
  fn1("sinh", [](double x) { return std::sinh(x); });
  fn1("cosh", [](double x) { return std::cosh(x); });
  fn1("tanh", [](double x) { return std::tanh(x); });
  fn1("log10", [](double x) { return std::log10(x); });
  fn1("rad", [](double x) { return x * (M_PI / 180.0); });
  fn1("deg", [](double x) { return x * (180.0 / M_PI); });
  math->set_field("log", builtin([](Values &a) {
    const double x = arg_num(a, 0, "log");
    if (a.size() > 1 && a[1].k != Value::Nil) return Values{Value::number(std::log(x) / std::log(arg_num(a, 1, "log")))};
    return Values{Value::number(std::log(x))};
  }));
  math->set_field("atan", builtin([](Values &a) {
    const double y = arg_num(a, 0, "atan");
    const double x = a.size() > 1 && a[1].k != Value::Nil ? arg_num(a, 1, "atan") : 1.0;
    return Values{Value::number(std::atan2(y, x))};
  }));
  math->set_field("pow", builtin([](Values &a) {
    return Values{Value::number(std::pow(arg_num(a, 0, "pow"), arg_num(a, 1, "pow")))};
  }));
  math->set_field("fmod", builtin([](Values &a) {
    return Values{Value::number(std::fmod(arg_num(a, 0, "fmod"), arg_num(a, 1, "fmod")))};
  }));
  math->set_field("modf", builtin([](Values &a) {
    double ip;
    const double fr = std::modf(arg_num(a, 0, "modf"), &ip);
    return Values{Value::number(ip), Value::number(fr)};
  }));
  math->set_field("max", builtin([](Values &a) {
    double m = arg_num(a, 0, "max");
    for (size_t i = 1; i < a.size(); ++i) m = std::max(m, arg_num(a, i, "max"));
    return Values{Value::number(m)};
  }));
  math->set_field("min", builtin([](Values &a) {
    double m = arg_num(a, 0, "min");
    for (size_t i = 1; i < a.size(); ++i) m = std::min(m, arg_num(a, i, "min"));
    return Values{Value::number(m)};
  }));
  math->set_field("tointeger", builtin([](Values &a) {
    long long i;
    if (!a.empty() && a[0].k == Value::Num && whole(a[0].n, &i)) return Values{a[0]};
    return Values{Value()};
  }));
  math->set_field("pi", Value::number(M_PI));
  math->set_field("huge", Value::number(HUGE_VAL));
  math->set_field("maxinteger", Value::number(9007199254740991.0));
  math->set_field("mininteger", Value::number(-9007199254740991.0));
  g->set_field("math", Value::table(math));

  auto str = std::make_shared<Table>();
  str->set_field("format", builtin([](Values &a) { return Values{Value::string(format(a))}; }));
  str->set_field("rep", builtin([](Values &a) {
    std::string s = arg_str(a, 0, "rep"), o;
    for (long long i = 0; i < (long long)arg_num(a, 1, "rep"); ++i) o += s;
    return Values{Value::string(o)};
  }));
  str->set_field("len", builtin([](Values &a) { return Values{Value::number((double)arg_str(a, 0, "len").size())}; }));
  str->set_field("upper", builtin([](Values &a) {
    std::string s = arg_str(a, 0, "upper");
    for (auto &c : s) c = (char)std::toupper((unsigned char)c);
    return Values{Value::string(s)};
  }));
  str->set_field("lower", builtin([](Values &a) {
    std::string s = arg_str(a, 0, "lower");
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return Values{Value::string(s)};
  }));
  str->set_field("sub", builtin([](Values &a) {
    const std::string s = arg_str(a, 0, "sub");
    const long long n = (long long)s.size();
    long long i = a.size() > 1 ? (long long)arg_num(a, 1, "sub") : 1;
    long long j = a.size() > 2 && a[2].k != Value::Nil ? (long long)arg_num(a, 2, "sub") : -1;
    if (i < 0) i = std::max(n + i + 1, 1LL);
    if (i == 0) i = 1;
    if (j < 0) j = n + j + 1;
    if (j > n) j = n;
    return Values{Value::string(i > j ? std::string() : s.substr((size_t)(i - 1), (size_t)(j - i + 1)))};
  }));
  g->set_field("string", Value::table(str));

  g->set_field("print", builtin([](Values &a) {
    for (size_t i = 0; i < a.size(); ++i) std::printf("%s%s", i ? "\t" : "", Interpreter::to_string(a[i]).c_str());
    std::printf("\n");
    std::fflush(stdout);
    return Values{};
  }));
  g->set_field("type", builtin([](Values &a) {
    if (a.empty()) throw Error("bad argument #1 to 'type' (value expected)");
    return Values{Value::string(a[0].type_name())};
  }));
  g->set_field("tostring", builtin([](Values &a) {
    return Values{Value::string(Interpreter::to_string(a.empty() ? Value() : a[0]))};
  }));
  g->set_field("tonumber", builtin([](Values &a) {
    if (!a.empty() && a[0].k == Value::Num) return Values{a[0]};
    double d;
    if (!a.empty() && a[0].k == Value::Str && str_to_number(*a[0].s, &d)) return Values{Value::number(d)};
    return Values{Value()};
  }));
  g->set_field("error", builtin([](Values &a) -> Values {
    throw Error(a.empty() ? std::string("error") : Interpreter::to_string(a[0]));
  }));
  g->set_field("assert", builtin([](Values &a) -> Values {
    if (a.empty() || !a[0].truthy())
      throw Error(a.size() > 1 ? Interpreter::to_string(a[1]) : std::string("assertion failed!"));
    return a;
  }));
  Value next = builtin([](Values &a) {
    if (a.empty() || a[0].k != Value::Tab) throw Error("bad argument #1 to 'next' (table expected)");
    Value k, v;
    if (!a[0].t->next(a.size() > 1 ? a[1] : Value(), &k, &v)) return Values{Value()};
    return Values{k, v};
  });
  g->set_field("next", next);
  g->set_field("pairs", builtin([next](Values &a) {
    if (a.empty() || a[0].k != Value::Tab) throw Error("bad argument #1 to 'pairs' (table expected)");
    return Values{next, a[0], Value()};
  }));
  Value inext = builtin([](Values &a) {
    const double i = arg_num(a, 1, "ipairs") + 1;
    Value v = a[0].t->get(Value::number(i));
    if (v.k == Value::Nil) return Values{Value()};
    return Values{Value::number(i), v};
  });
  g->set_field("ipairs", builtin([inext](Values &a) {
    if (a.empty() || a[0].k != Value::Tab) throw Error("bad argument #1 to 'ipairs' (table expected)");
    return Values{inext, a[0], Value::number(0)};
  }));
  g->set_field("select", builtin([](Values &a) {
    if (!a.empty() && a[0].k == Value::Str && *a[0].s == "#") return Values{Value::number((double)a.size() - 1)};
    const long long n = (long long)arg_num(a, 0, "select");
    if (n < 1) throw Error("bad argument #1 to 'select' (index out of range)");
    return Values(a.begin() + std::min((size_t)n, a.size()), a.end());
  }));
  auto unpack = builtin([](Values &a) {
    if (a.empty() || a[0].k != Value::Tab) throw Error("bad argument #1 to 'unpack' (table expected)");
    Values out(a[0].t->arr.begin(), a[0].t->arr.end());
    return out;
  });
  g->set_field("unpack", unpack);
  auto tab = std::make_shared<Table>();
  tab->set_field("unpack", unpack);
  g->set_field("table", Value::table(tab));
  (void)ex;
}

}  // namespace

// ------------------------------------------------------------------ interpreter

std::string Interpreter::to_string(const Value &v) {
  char buf[64];
  switch (v.k) {
    case Value::Nil: return "nil";
    case Value::Bool: return v.b ? "true" : "false";
    case Value::Num: return num_to_string(v.n);
    case Value::Str: return *v.s;
    case Value::Tab: std::snprintf(buf, sizeof buf, "table: %p", (void *)v.t.get()); return buf;
    default: std::snprintf(buf, sizeof buf, "function: %p", (void *)v.f.get()); return buf;
  }
}

Interpreter::Interpreter() : globals_(std::make_shared<Table>()) {
  Exec ex(globals_);
  open_libs(globals_, &ex);
  globals_->set_field("_G", Value::table(globals_));
}

Interpreter::~Interpreter() = default;

void Interpreter::run(const std::string &source, const std::string &chunkname) {
  auto c = std::make_unique<Chunk>();
  Parser p(Lexer(source, chunkname).run(), *c, chunkname);
  c->main = p.parse_main();
  const FuncBody *main = c->main;
  chunks_.push_back(std::move(c));
  Exec ex(globals_);
  Values ret;
  try {
    if (ex.block(main->body, nullptr, &ret, true) == Flow::Break) throw Error("break outside a loop");
  } catch (const Error &e) {
    throw Error(chunkname + ": " + e.what());
  }
}

Values Interpreter::call(const Value &fn, Values args) {
  Exec ex(globals_);
  return ex.call(fn, args, 0);
}

}  // namespace fiesta_lua
