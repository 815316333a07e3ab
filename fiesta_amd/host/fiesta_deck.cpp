// Lua input decks (include/fiesta_deck.h): the script runs in the build's own
// interpreter (fiesta_lua.hpp); the keys map onto fst_config as
// executeConfiguration does (src/input.cpp:130-443), the callbacks fill the
// grid and the initial state as loadGrid / loadInitialConditions do
// (src/input.cpp:445-557).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "fiesta_deck.h"
#include "fiesta_lua.hpp"

using fiesta_lua::Table;
using fiesta_lua::Value;
using fiesta_lua::Values;

namespace {

struct DeckError {
  int code;
  std::string msg;
};

// The root table and its sub-tables as luaReader creates them before the
// script runs (src/luaReader.cpp:36-73).
const char *const kSubtables[] = {"ceq", "noise", "restart", "bc", "time", "viscosity", "buoyancy",
                                  "grid", "mpi", "hdf5", "eos", "progress", "status"};

std::unique_ptr<fiesta_lua::Interpreter> run_script(const std::string &path, const std::string &source) {
  auto L = std::make_unique<fiesta_lua::Interpreter>();
  auto root = std::make_shared<Table>();
  for (const char *s : kSubtables) root->set_field(s, Value::table(std::make_shared<Table>()));
  L->globals()->set_field("fiesta", Value::table(root));
  try {
    L->run(source, path);
  } catch (const fiesta_lua::Error &e) {
    throw DeckError{FDK_ERR_SCRIPT, std::string("Cannot run config file: ") + e.what()};
  }
  return L;
}

std::string key_name(std::initializer_list<const char *> keys) {
  std::string s = "fiesta";
  for (const char *k : keys) s += std::string(".") + k;
  return s;
}

}  // namespace

struct fdk_deck {
  std::string path, source;
  std::unique_ptr<fiesta_lua::Interpreter> L;
  std::string err;
  // filled by fdk_read_config
  bool configured = false;
  fst_config cf{};
  fdk_run run{};
  std::vector<std::string> species;
  std::vector<fdk_ioview> views;

  Value root() const { return L->globals()->field("fiesta"); }

  // lua_getfield chain from the root; nil when any level is missing
  Value find(std::initializer_list<const char *> keys) const {
    Value v = root();
    for (const char *k : keys) {
      if (v.k != Value::Tab) return Value();
      v = v.t->field(k);
    }
    return v;
  }
  [[noreturn]] void missing(std::initializer_list<const char *> keys) const {
    throw DeckError{FDK_ERR_MISSING,
                    "Could not find required parameter '" + key_name(keys) + "' in '" + path + "'"};
  }
  [[noreturn]] void bad(std::initializer_list<const char *> keys, const std::string &what) const {
    throw DeckError{FDK_ERR_ARG, "'" + key_name(keys) + "' in '" + path + "': " + what};
  }

  // luaReader::getValue conversions
  double as_number(const Value &v, std::initializer_list<const char *> keys) const {
    if (v.k == Value::Num) return v.n;
    if (v.k == Value::Str) {
      char *e = nullptr;
      const double d = std::strtod(v.s->c_str(), &e);
      if (e != v.s->c_str() && *e == '\0') return d;
    }
    bad(keys, std::string("a number is required, got ") + v.type_name());
  }
  long long as_int(const Value &v, std::initializer_list<const char *> keys) const {
    const double d = as_number(v, keys);
    if (d != std::floor(d) || std::fabs(d) > 9.0e15) bad(keys, "an integer is required");
    return (long long)d;
  }
  bool as_bool(const Value &v) const { return v.truthy(); }  // lua_toboolean
  std::string as_string(const Value &v, std::initializer_list<const char *> keys) const {
    if (v.k == Value::Str) return *v.s;
    if (v.k == Value::Num) return fiesta_lua::Interpreter::to_string(v);
    bad(keys, std::string("a string is required, got ") + v.type_name());
  }

  double num(std::initializer_list<const char *> k) const {
    Value v = find(k);
    if (v.k == Value::Nil) missing(k);
    return as_number(v, k);
  }
  double num(std::initializer_list<const char *> k, double d) const {
    Value v = find(k);
    return v.k == Value::Nil ? d : as_number(v, k);
  }
  long long integer(std::initializer_list<const char *> k) const {
    Value v = find(k);
    if (v.k == Value::Nil) missing(k);
    return as_int(v, k);
  }
  long long integer(std::initializer_list<const char *> k, long long d) const {
    Value v = find(k);
    return v.k == Value::Nil ? d : as_int(v, k);
  }
  bool boolean(std::initializer_list<const char *> k, bool d) const {
    Value v = find(k);
    return v.k == Value::Nil ? d : as_bool(v);
  }
  std::string str(std::initializer_list<const char *> k) const {
    Value v = find(k);
    if (v.k == Value::Nil) missing(k);
    return as_string(v, k);
  }
  std::string str(std::initializer_list<const char *> k, const std::string &d) const {
    Value v = find(k);
    return v.k == Value::Nil ? d : as_string(v, k);
  }
  // luaReader::get(keys, vector, n): the first n entries of an array
  std::vector<double> array(std::initializer_list<const char *> k, int n) const {
    Value v = find(k);
    if (v.k == Value::Nil) missing(k);
    if (v.k != Value::Tab) bad(k, "a table is required");
    std::vector<double> out;
    for (int i = 1; i <= n; ++i) {
      Value e = v.t->get(Value::number(i));
      if (e.k == Value::Nil) bad(k, "needs " + std::to_string(n) + " entries");
      out.push_back(as_number(e, k));
    }
    return out;
  }
};

namespace {

void copy_str(char *dst, size_t n, const std::string &s) {
  std::snprintf(dst, n, "%s", s.c_str());
}

int fail(fdk_deck *d, const DeckError &e) {
  if (d) d->err = e.msg;
  return e.code;
}

// parseBC (src/bc.cpp:252-270)
int parse_bc(const fdk_deck &d, std::initializer_list<const char *> k) {
  const std::string s = d.str(k);
  if (s == "outflow") return FST_BC_OUTFLOW;
  if (s == "reflective") return FST_BC_REFLECTIVE;
  if (s == "noslip") return FST_BC_NOSLIP;
  if (s == "hydrostatic") return FST_BC_HYDROSTATIC;
  d.bad(k, "invalid boundary condition type '" + s + "'");
}

// executeConfiguration, src/input.cpp:130-443
void read_config(fdk_deck &d) {
  fst_config c;
  fst_config_init(&c);
  fdk_run r;
  std::memset(&r, 0, sizeof r);
  copy_str(r.title, sizeof r.title, d.str({"title"}));
  const long long nt = d.integer({"time", "nt"}, 0);
  c.dt = d.num({"time", "dt"});
  r.tstart = d.integer({"time", "start_index"}, 0);
  r.time0 = d.num({"time", "start_time"}, 0.0);
  if (nt == 0) {
    r.tinterval = 0;
    r.tend = d.integer({"time", "tend"}, 0);
    if (r.tend == 0)
      throw DeckError{FDK_ERR_MISSING, "Must specify either 'fiesta.time.nt' or 'fiesta.time.tend' in '" + d.path + "'"};
    r.nt = r.tend - r.tstart;
  } else {
    r.tinterval = 1;
    r.nt = nt;
    r.tend = r.tstart + nt;
  }
  copy_str(r.metadata, sizeof r.metadata, d.str({"metadata"}, "none"));
  c.R = d.num({"eos", "R"}, 8.314462);
  c.ndim = (int)d.integer({"grid", "ndim"}, 2);
  if (c.ndim != 2 && c.ndim != 3) d.bad({"grid", "ndim"}, "must be 2 or 3");
  c.visc = d.boolean({"viscosity", "enabled"}, false);
  c.ng = (int)d.integer({"ng"}, 3);
  c.xPer = d.boolean({"bc", "xperiodic"}, false);
  c.yPer = d.boolean({"bc", "yperiodic"}, false);
  r.out_freq = d.integer({"progress", "frequency"}, 0);
  r.write_freq = d.integer({"write_frequency"}, 0);
  r.stat_freq = d.integer({"status", "frequency"}, 0);
  copy_str(r.terrain_name, sizeof r.terrain_name, d.str({"terrain", "name"}, "terrain.h5"));
  r.restart = d.boolean({"restart", "enabled"}, false);
  std::string rname = d.str({"restart", "name"}, "restart-0000000.h5");
  const std::string rpath = d.str({"restart", "path"}, ".");
  r.restart_reset = d.boolean({"restart", "reset"}, false);
  r.auto_restart = d.boolean({"restart", "auto"}, false);
  std::string aname = d.str({"restart", "auto_name"}, "restart-auto");
  r.restart_time_remaining = d.integer({"restart", "time_remaining"}, 0);
  if (r.auto_restart) {
    r.restart = 1;
    rname = rpath + "/" + aname + ".h5";
    std::ifstream probe(rname);
    if (!probe) r.restart = 0;  // no auto restart file yet: a fresh run
    r.restart_freq = d.integer({"restart", "frequency"}, r.tend);
  } else {
    aname = "restart";
    r.restart_freq = d.integer({"restart", "frequency"}, 0);
  }
  copy_str(r.restart_name, sizeof r.restart_name, rname);
  copy_str(r.restart_path, sizeof r.restart_path, rpath);
  copy_str(r.auto_restart_name, sizeof r.auto_restart_name, aname);
  copy_str(r.mpi_type, sizeof r.mpi_type, d.str({"mpi", "type"}, "host"));
  r.hdf5_chunk = d.boolean({"hdf5", "chunk"}, false);
  r.hdf5_compress = d.boolean({"hdf5", "compress"}, false);
  if (!r.hdf5_chunk && r.hdf5_compress) r.hdf5_compress = 0;  // compression requires chunking
  const std::string scheme = d.str({"advection_scheme"}, "weno5");
  const std::string grid = d.str({"grid", "type"}, "cartesian");
  const std::vector<double> dx = d.array({"grid", "dx"}, c.ndim);
  const std::vector<double> ni = d.array({"grid", "ni"}, c.ndim);
  for (double v : ni)
    if (v < 1 || v != std::floor(v) || v > 2.0e9) d.bad({"grid", "ni"}, "cell counts must be positive integers");
  c.nci = (int)ni[0];
  c.ncj = (int)ni[1];
  c.nck = c.ndim == 3 ? (int)ni[2] : 1;
  const std::vector<double> procs = d.array({"mpi", "procs"}, c.ndim);
  for (int a = 0; a < 3; ++a) r.procs[a] = a < c.ndim ? (int)procs[a] : 1;
  bool several = false;
  for (int a = 0; a < c.ndim; ++a) {
    if (r.procs[a] < 1 || procs[a] != std::floor(procs[a])) d.bad({"mpi", "procs"}, "positive integers are required");
    several = several || r.procs[a] > 1;
  }
  if (several)  // the decomposition only splits the work; one process owns the whole grid here
    std::fprintf(stderr, "fiesta deck: mpi.procs = {%d,%d,%d}: this build runs the whole grid in one process on one GPU\n",
                 r.procs[0], r.procs[1], r.procs[2]);
  if (!c.xPer) {
    c.bcL = parse_bc(d, {"bc", "xmin"});
    c.bcR = parse_bc(d, {"bc", "xmax"});
  }
  if (!c.yPer) {
    c.bcB = parse_bc(d, {"bc", "ymin"});
    c.bcT = parse_bc(d, {"bc", "ymax"});
  }
  c.buoyancy = d.boolean({"buoyancy", "enabled"}, false);
  if (c.buoyancy) {
    c.gAccel = d.num({"buoyancy", "acceleration"}, 9.81);
    c.rhoRef = d.num({"buoyancy", "rho_reference"}, 0.0);
  }
  c.ceq = d.boolean({"ceq", "enabled"}, false);
  if (c.ceq) {
    c.kap = d.num({"ceq", "kappa"}, 1.23456);
    c.eps = d.num({"ceq", "epsilon"});
    c.alpha = d.num({"ceq", "alpha"});
    r.beta = d.num({"ceq", "beta"});
    r.betae = d.num({"ceq", "betae"});
    r.st = d.integer({"ceq", "st"}, 0);
  }
  c.noise = d.boolean({"noise", "enabled"}, false);
  if (c.noise) {
    c.n_dh = d.num({"noise", "dh"});
    c.n_eta = d.num({"noise", "eta"});
    c.n_coff = d.num({"noise", "coff"});
    c.n_nt = (int)d.integer({"noise", "nt"}, 1);
    c.n_mode = (int)d.integer({"noise", "mode"}, 1);
  }
  if (c.ndim == 3) {
    c.zPer = d.boolean({"bc", "zperiodic"}, false);
    if (!c.zPer) {
      c.bcH = parse_bc(d, {"bc", "zmin"});
      c.bcF = parse_bc(d, {"bc", "zmax"});
    }
  }
  if (grid == "cartesian") {
    c.grid = 0;
    c.dx = dx[0];
    c.dy = dx[1];
    if (c.ndim == 3) c.dz = dx[2];
  } else if (grid == "terrain") {
    if (c.ndim != 3) d.bad({"grid", "type"}, "ndim must be equal to 3 for terrain");
    c.grid = 2;
    c.dx = c.dy = c.dz = 1.0;
    r.tdx = d.num({"tdx"});
    r.tdy = d.num({"tdy"});
    r.h = d.num({"h"});
  } else if (grid == "generalized") {
    c.grid = 1;
    c.dx = c.dy = 1.0;
    if (c.ndim == 3) c.dz = 1.0;
  } else {
    // the reference keeps its default (cartesian numbering, unset spacing) here
    d.bad({"grid", "type"}, "'" + grid + "' is not cartesian, generalized or terrain");
  }
  // getSpeciesData, src/luaReader.cpp:163-205
  Value sp = d.find({"species"});
  if (sp.k == Value::Nil) throw DeckError{FDK_ERR_MISSING, "Could not find fiesta.species in '" + d.path + "'"};
  if (sp.k != Value::Tab) d.bad({"species"}, "a table of species is required");
  const size_t ns = sp.t->length();
  if (ns < 1 || ns > FST_MAX_SPECIES) d.bad({"species"}, "1 to 8 species are supported");
  c.ns = (int)ns;
  d.species.clear();
  for (size_t s = 0; s < ns; ++s) {
    Value e = sp.t->get(Value::number((double)(s + 1)));
    if (e.k != Value::Tab) d.bad({"species"}, "each species is a table {name, gamma, M, mu}");
    auto field = [&](const char *f) { return e.t->field(f); };
    d.species.push_back(field("name").k == Value::Nil ? std::string() : d.as_string(field("name"), {"species", "name"}));
    c.gamma[s] = d.as_number(field("gamma"), {"species", "gamma"});
    c.M[s] = d.as_number(field("M"), {"species", "M"});
    c.mu[s] = d.as_number(field("mu"), {"species", "mu"});
  }
  c.scheme = FST_WENO5;  // src/input.cpp:305-311: unknown names stay weno5
  if (scheme == "centered4") c.scheme = FST_CENTERED4;
  if (scheme == "quick") c.scheme = FST_QUICK;
  if (c.ndim == 2) {  // src/input.cpp:351-360
    c.dz = c.dx;
    c.nck = 1;
    c.bcH = c.bcF = FST_BC_OUTFLOW;
    c.zPer = 0;
  }
  // options of this build (no counterpart in the reference decks)
  const std::string prec = d.str({"precision"}, "fp64");
  if (prec == "fp64") c.precision = FST_FP64;
  else if (prec == "fp32") c.precision = FST_FP32_BULK;
  else if (prec == "fp64_fast") c.precision = FST_FP64_FAST;
  else d.bad({"precision"}, "fp64, fp32 or fp64_fast");
  c.learned_c = d.boolean({"ceq", "learned"}, false);
  const std::string integ = d.str({"time", "integrator"}, "rk2");
  if (integ == "rk2") c.rk_order = 2;
  else if (integ == "ssp-rk3") c.rk_order = 3;
  else d.bad({"time", "integrator"}, "rk2 or ssp-rk3");
  // getIOBlock, src/luaReader.cpp:208-312
  Value iv = d.find({"ioviews"});
  if (iv.k == Value::Nil) throw DeckError{FDK_ERR_MISSING, "Could not find fiesta.ioviews in '" + d.path + "'"};
  if (iv.k != Value::Tab) d.bad({"ioviews"}, "a table of views is required");
  d.views.clear();
  for (size_t i = 1; i <= iv.t->length(); ++i) {
    Value e = iv.t->get(Value::number((double)i));
    if (e.k != Value::Tab) d.bad({"ioviews"}, "each view is a table");
    fdk_ioview w;
    std::memset(&w, 0, sizeof w);
    copy_str(w.name, sizeof w.name, d.as_string(e.t->field("name"), {"ioviews", "name"}));
    Value pth = e.t->field("path");
    copy_str(w.path, sizeof w.path, pth.k == Value::Nil ? std::string("./") : d.as_string(pth, {"ioviews", "path"}));
    Value fq = e.t->field("frequency");
    w.frequency = fq.k == Value::Nil ? 0 : d.as_int(fq, {"ioviews", "frequency"});
    Value av = e.t->field("average");
    w.average = av.k == Value::Nil ? 1 : (int)d.as_int(av, {"ioviews", "average"});
    bool whole = true;
    const int64_t cells[3] = {c.nci, c.ncj, c.nck};
    auto list = [&](const char *f, int64_t *out, int64_t def, int64_t shift) {
      Value a = e.t->field(f);
      if (a.k == Value::Tab) {
        whole = false;
        for (int j = 0; j < 3; ++j) {
          Value x = a.t->get(Value::number(j + 1));
          out[j] = x.k == Value::Nil ? (j < c.ndim ? def : 0) : d.as_int(x, {"ioviews", f}) - shift;
        }
      } else {
        for (int j = 0; j < 3; ++j) out[j] = def;
      }
    };
    list("start", w.start, 0, 1);
    list("stride", w.stride, 1, 0);
    Value lim = e.t->field("limit");
    if (lim.k == Value::Tab) {
      whole = false;
      for (int j = 0; j < 3; ++j) {
        Value x = lim.t->get(Value::number(j + 1));
        w.limit[j] = x.k == Value::Nil ? cells[j] - 1 : d.as_int(x, {"ioviews", "limit"}) - 1;
      }
    } else {
      for (int j = 0; j < 3; ++j) w.limit[j] = cells[j] - 1;
    }
    // the reference's defaultable flag: start absent, limit and stride absent
    w.whole = whole ? 1 : 0;
    if (e.t->field("start").k == Value::Tab) w.whole = 0;
    d.views.push_back(w);
  }
  r.n_ioviews = (int)d.views.size();
  char why[512];
  if (fst_config_check(&c, why, sizeof why) != FST_OK)
    throw DeckError{FDK_ERR_ARG, std::string("the deck's configuration is not supported by this build: ") + why};
  d.cf = c;
  d.run = r;
  d.configured = true;
}

// fiesta.<fname>(args) as a number (luaReader::call, src/luaReader.cpp:316-345)
double call_number(fiesta_lua::Interpreter &L, const std::string &path, const Value &fn, const std::string &fname,
                   Values args) {
  Values r;
  try {
    r = L.call(fn, std::move(args));
  } catch (const fiesta_lua::Error &e) {
    throw DeckError{FDK_ERR_SCRIPT, "error running function '" + fname + "': " + e.what()};
  }
  if (r.empty() || r[0].k != Value::Num) {
    if (!r.empty() && r[0].k == Value::Str) {
      char *e = nullptr;
      const double v = std::strtod(r[0].s->c_str(), &e);
      if (e != r[0].s->c_str() && *e == '\0') return v;
    }
    throw DeckError{FDK_ERR_SCRIPT, "function '" + fname + "' should return a number (in '" + path + "')"};
  }
  return r[0].n;
}

Value root_function(fiesta_lua::Interpreter &L, const std::string &path, const std::string &fname) {
  Value root = L.globals()->field("fiesta");
  Value fn = root.k == Value::Tab ? root.t->field(fname) : Value();
  if (fn.k == Value::Nil) throw DeckError{FDK_ERR_MISSING, "Could not find fiesta." + fname + " in '" + path + "'"};
  if (fn.k != Value::Fn) throw DeckError{FDK_ERR_SCRIPT, "fiesta." + fname + " is not a function"};
  return fn;
}

int thread_count(int requested, long long work) {
  int n = requested;
  if (n <= 0) {
    const char *env = std::getenv("OMP_NUM_THREADS");
    n = env && std::atoi(env) > 0 ? std::atoi(env) : (int)std::thread::hardware_concurrency();
  }
  n = std::max(1, std::min(n, 64));
  return (int)std::max(1LL, std::min((long long)n, work));
}

// Evaluate body(L, fn, lo, hi) over [0, n) on `threads` host threads, each
// with its own run of the script (the reference runs the script afresh for
// every pass too: src/input.cpp:453, 533).
template <typename F>
void parallel_calls(fdk_deck &d, const char *fname, long long n, int threads, F body) {
  const int nt = thread_count(threads, n);
  std::vector<DeckError> errs((size_t)nt, DeckError{0, ""});
  auto work = [&](int w) {
    try {
      std::unique_ptr<fiesta_lua::Interpreter> own;
      fiesta_lua::Interpreter *L = d.L.get();
      if (w > 0) {
        own = run_script(d.path, d.source);
        L = own.get();
      }
      const Value fn = root_function(*L, d.path, fname);
      const long long lo = n * w / nt, hi = n * (w + 1) / nt;
      body(*L, fn, lo, hi);
    } catch (const DeckError &e) {
      errs[(size_t)w] = e;
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < nt; ++w) pool.emplace_back(work, w);
  work(0);
  for (auto &t : pool) t.join();
  for (const auto &e : errs)
    if (e.code) throw e;
}

}  // namespace

extern "C" {

int fdk_open(const char *path, fdk_deck **out) {
  if (!path || !out) return FDK_ERR_ARG;
  *out = nullptr;
  auto d = std::make_unique<fdk_deck>();
  d->path = path;
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    std::fprintf(stderr, "fdk_open: cannot read '%s'\n", path);
    return FDK_ERR_IO;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  d->source = ss.str();
  try {
    d->L = run_script(d->path, d->source);
  } catch (const DeckError &e) {
    std::fprintf(stderr, "fdk_open: %s\n", e.msg.c_str());
    return e.code;
  }
  *out = d.release();
  return FDK_OK;
}

int fdk_close(fdk_deck *d) {
  if (!d) return FDK_ERR_ARG;
  delete d;
  return FDK_OK;
}

const char *fdk_last_error(const fdk_deck *d) { return d ? d->err.c_str() : "null deck"; }

int fdk_read_config(fdk_deck *d, fst_config *cf, fdk_run *run) {
  if (!d || !cf || !run) return FDK_ERR_ARG;
  try {
    read_config(*d);
  } catch (const DeckError &e) {
    return fail(d, e);
  }
  *cf = d->cf;
  *run = d->run;
  return FDK_OK;
}

int fdk_species_name(fdk_deck *d, int s, char *buf, size_t len) {
  if (!d || !buf || len == 0) return FDK_ERR_ARG;
  if (!d->configured) return fail(d, DeckError{FDK_ERR_ARG, "fdk_species_name: call fdk_read_config first"});
  if (s < 0 || s >= (int)d->species.size()) return fail(d, DeckError{FDK_ERR_ARG, "fdk_species_name: no such species"});
  copy_str(buf, len, d->species[(size_t)s]);
  return FDK_OK;
}

int fdk_get_ioview(fdk_deck *d, int i, fdk_ioview *out) {
  if (!d || !out) return FDK_ERR_ARG;
  if (!d->configured) return fail(d, DeckError{FDK_ERR_ARG, "fdk_get_ioview: call fdk_read_config first"});
  if (i < 0 || i >= (int)d->views.size()) return fail(d, DeckError{FDK_ERR_ARG, "fdk_get_ioview: no such view"});
  *out = d->views[(size_t)i];
  return FDK_OK;
}

int fdk_grid(fdk_deck *d, const fst_config *cf, double *grid, size_t count, int threads) {
  if (!d || !cf || !grid) return FDK_ERR_ARG;
  const long long ni = cf->nci + 1, nj = cf->ncj + 1, nk = cf->ndim == 3 ? cf->nck + 1 : 1;
  if (count != (size_t)(ni * nj * nk * 3)) return fail(d, DeckError{FDK_ERR_ARG, "fdk_grid: wrong node count"});
  std::fill(grid, grid + count, 0.0);
  const long long plane = ni * nj * nk;
  try {
    // loadGrid, src/input.cpp:527-557: v < ndim, every node, (i, j, k, v) as numbers
    parallel_calls(*d, "initialize_grid", nk, threads, [&](fiesta_lua::Interpreter &L, const Value &fn, long long lo,
                                                           long long hi) {
      for (int v = 0; v < cf->ndim; ++v)
        for (long long k = lo; k < hi; ++k)
          for (long long j = 0; j < nj; ++j)
            for (long long i = 0; i < ni; ++i)
              grid[i + ni * (j + nj * k) + plane * v] =
                  call_number(L, d->path, fn, "initialize_grid",
                              {Value::number((double)i), Value::number((double)j), Value::number((double)k),
                               Value::number((double)v)});
    });
  } catch (const DeckError &e) {
    return fail(d, e);
  }
  return FDK_OK;
}

int fdk_initial_conditions(fdk_deck *d, const fst_config *cf, const double *grid, double *var, size_t count,
                           int threads) {
  if (!d || !cf || !var) return FDK_ERR_ARG;
  fst_dims dm;
  if (fst_config_dims(cf, &dm) != FST_OK) return fail(d, DeckError{FDK_ERR_ARG, "fdk_initial_conditions: bad config"});
  if (count != (size_t)dm.ncells_ghost * (size_t)dm.nvt)
    return fail(d, DeckError{FDK_ERR_ARG, "fdk_initial_conditions: wrong value count"});
  if (cf->grid != 0 && !grid)
    return fail(d, DeckError{FDK_ERR_ARG, "fdk_initial_conditions: a generalized grid needs its node coordinates"});
  std::fill(var, var + count, 0.0);
  const int ng = cf->ng, ndim = cf->ndim;
  const long long ngi = dm.ngi, ngj = dm.ngj, N = dm.ncells_ghost;
  const long long nci = cf->nci, ncj = cf->ncj, nck = ndim == 3 ? cf->nck : 1;
  const long long ni = nci + 1, nj = ncj + 1, nk = ndim == 3 ? nck + 1 : 1;
  auto node = [&](long long i, long long j, long long k, int v) {
    return grid[i + ni * (j + nj * k) + ni * nj * nk * v];
  };
  try {
    // loadInitialConditions, src/input.cpp:445-516: variables v < nv of the
    // interior cells; the C-equation variables stay zero
    parallel_calls(*d, "initial_conditions", nck, threads, [&](fiesta_lua::Interpreter &L, const Value &fn,
                                                               long long lo, long long hi) {
      for (int v = 0; v < dm.nv; ++v)
        for (long long kk = lo; kk < hi; ++kk)
          for (long long jj = 0; jj < ncj; ++jj)
            for (long long ii = 0; ii < nci; ++ii) {
              double x, y, z;
              if (cf->grid == 0) {
                x = cf->dx * (double)ii + 0.5 * cf->dx;
                y = cf->dy * (double)jj + 0.5 * cf->dy;
                z = ndim == 3 ? cf->dz * (double)kk + 0.5 * cf->dz : 0.0;
              } else if (ndim == 3) {
                x = y = z = 0;
                for (int ix = 0; ix < 2; ++ix)
                  for (int iy = 0; iy < 2; ++iy)
                    for (int iz = 0; iz < 2; ++iz) {
                      x += node(ii + ix, jj + iy, kk + iz, 0);
                      y += node(ii + ix, jj + iy, kk + iz, 1);
                      z += node(ii + ix, jj + iy, kk + iz, 2);
                    }
                x = x / 8.0;
                y = y / 8.0;
                z = z / 8.0;
              } else {
                x = y = 0;
                for (int ix = 0; ix < 2; ++ix)
                  for (int iy = 0; iy < 2; ++iy) {
                    x += node(ii + ix, jj + iy, 0, 0);
                    y += node(ii + ix, jj + iy, 0, 1);
                  }
                x = x / 4.0;
                y = y / 4.0;
                z = 0.0;
              }
              const long long i = ii + ng, j = jj + ng, k = ndim == 3 ? kk + ng : 0;
              var[i + ngi * (j + ngj * k) + N * v] =
                  call_number(L, d->path, fn, "initial_conditions",
                              {Value::number(x), Value::number(y), Value::number(z), Value::number((double)v)});
            }
    });
  } catch (const DeckError &e) {
    return fail(d, e);
  }
  return FDK_OK;
}

int fdk_call(fdk_deck *d, const char *fname, int nargs, const double *args, double *out) {
  if (!d || !fname || !out || nargs < 0 || (nargs > 0 && !args)) return FDK_ERR_ARG;
  try {
    const Value fn = root_function(*d->L, d->path, fname);
    Values a;
    for (int i = 0; i < nargs; ++i) a.push_back(Value::number(args[i]));
    *out = call_number(*d->L, d->path, fn, fname, std::move(a));
  } catch (const DeckError &e) {
    return fail(d, e);
  }
  return FDK_OK;
}

int fdk_get(fdk_deck *d, const char *key, int *kind, double *num, char *str, size_t len) {
  if (!d || !key || !kind) return FDK_ERR_ARG;
  Value v = d->root();
  std::string k(key), part;
  std::stringstream ss(k);
  while (v.k != Value::Nil && std::getline(ss, part, '.')) {
    if (v.k != Value::Tab) {
      v = Value();
      break;
    }
    char *e = nullptr;
    const double idx = std::strtod(part.c_str(), &e);
    v = (!part.empty() && *e == '\0') ? v.t->get(Value::number(idx)) : v.t->field(part);
  }
  switch (v.k) {
    case Value::Nil: *kind = 0; break;
    case Value::Bool: *kind = 1; if (num) *num = v.b ? 1.0 : 0.0; break;
    case Value::Num: *kind = 2; if (num) *num = v.n; break;
    case Value::Str: *kind = 3; if (str && len) copy_str(str, len, *v.s); break;
    case Value::Tab: *kind = 4; break;
    default: *kind = 5; break;
  }
  return FDK_OK;
}

}  // extern "C"
