// A small interpreter for the Lua subset FIESTA input decks use (host code).
//
// Values: nil, boolean, number (double), string, table, function.  Tables
// keep keys 1..n in an array part and other keys in hash parts; numbers
// that are whole are the same key whatever their spelling (t[1], t[1.0]).
// Statements: local, assignment (several targets and values), calls,
// function definitions (global, field and method names; local functions),
// if / elseif / else, while, repeat / until, numeric and generic for, do
// blocks, return and break.  Expressions: the arithmetic (+ - * / // % ^
// unary -), comparison, logical (and, or, not with Lua's value semantics),
// concatenation (..) and length (#) operators with Lua's precedence and
// associativity, table constructors, anonymous functions, calls (f(...),
// f{...}, f"..."), method calls (o:m(...)) and indexing.  Library: math
// (the Lua 5.3 set with its 5.2-compatibility functions), print, type,
// tostring, tonumber, pairs, ipairs, next, select, error, assert, unpack,
// string.format / rep / sub / len / upper / lower.  Not supported (an error
// names them): goto and labels, varargs, metatables, coroutines, io, os,
// require, bitwise operators.
#pragma once
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace fiesta_lua {

struct Table;
struct Function;
struct FuncBody;
struct Env;

struct Value {
  enum Kind : unsigned char { Nil, Bool, Num, Str, Tab, Fn };
  Kind k = Nil;
  bool b = false;
  double n = 0.0;
  std::shared_ptr<const std::string> s;
  std::shared_ptr<Table> t;
  std::shared_ptr<Function> f;

  static Value boolean(bool v) { Value x; x.k = Bool; x.b = v; return x; }
  static Value number(double v) { Value x; x.k = Num; x.n = v; return x; }
  static Value string(std::string v) { Value x; x.k = Str; x.s = std::make_shared<const std::string>(std::move(v)); return x; }
  static Value table(std::shared_ptr<Table> v) { Value x; x.k = Tab; x.t = std::move(v); return x; }
  static Value function(std::shared_ptr<Function> v) { Value x; x.k = Fn; x.f = std::move(v); return x; }
  bool truthy() const { return !(k == Nil || (k == Bool && !b)); }
  const char *type_name() const;
};

using Values = std::vector<Value>;

struct Table {
  std::vector<Value> arr;                          // keys 1..arr.size()
  std::unordered_map<std::string, Value> str;      // string keys
  std::map<double, Value> num;                     // other number keys
  std::vector<std::pair<Value, Value>> other;      // boolean, table and function keys
  Value get(const Value &key) const;
  void set(const Value &key, const Value &v);
  Value field(const std::string &k) const;
  void set_field(const std::string &k, const Value &v);
  size_t length() const { return arr.size(); }
  // the entry after `key` (nil key: the first); false at the end
  bool next(const Value &key, Value *k, Value *v) const;
};

struct Function {
  std::function<Values(Values &)> builtin;   // set for library functions
  const FuncBody *body = nullptr;            // set for Lua functions
  std::shared_ptr<Env> env;                  // their enclosing scope
};

class Error : public std::runtime_error {
 public:
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

struct Chunk;  // a parsed script (opaque)

class Interpreter {
 public:
  Interpreter();
  ~Interpreter();
  // parse and run a script; chunkname appears in error messages
  void run(const std::string &source, const std::string &chunkname);
  std::shared_ptr<Table> globals() const { return globals_; }
  Values call(const Value &fn, Values args);
  static std::string to_string(const Value &v);

 private:
  std::shared_ptr<Table> globals_;
  std::vector<std::unique_ptr<Chunk>> chunks_;  // parsed scripts stay alive for their closures
};

}  // namespace fiesta_lua
