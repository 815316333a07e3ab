// Host-side rk_func interface over the C-ABI; see fiesta_host.hpp.
#include "fiesta_host.hpp"

#include <cmath>
#include <string>

namespace fiesta_hip {

rk_func::rk_func(const fst_config &cf) : cf_(cf) {
  check(fst_config_dims(&cf_, &d_), "fst_config_dims");
  const int rc = fst_create(&cf_, &ctx_);
  if (rc != FST_OK) throw Error("fst_create failed (invalid configuration or no device memory)", rc);
  var.resize(d_.ngi, d_.ngj, d_.ngk, d_.nvt);
  dvar.resize(d_.ngi, d_.ngj, d_.ngk, d_.nvt);
  tmp1.resize(d_.ngi, d_.ngj, d_.ngk, d_.nvt);
  varx.resize(d_.ngi, d_.ngj, d_.ngk, d_.nvx);
  if (cf_.grid == 1) grid.resize(cf_.nci + 1, cf_.ncj + 1, cf_.ndim == 3 ? cf_.nck + 1 : 1, 3);
}

rk_func::~rk_func() {
  if (ctx_) fst_destroy(ctx_);
}

void rk_func::check(int rc, const char *what) const {
  if (rc != FST_OK) {
    std::string msg = std::string(what) + " failed";
    if (ctx_) msg += std::string(": ") + fst_last_error(ctx_);
    throw Error(msg, rc);
  }
}

HostArray *rk_func::mirror(int array) {
  switch (array) {
    case FST_VAR: return &var;
    case FST_VARX: return &varx;
    case FST_DVAR: return &dvar;
    case FST_TMP1: return &tmp1;
    case FST_GRID: return grid.size() ? &grid : nullptr;
    default: return nullptr;
  }
}

void rk_func::pull(int array) {
  HostArray *h = mirror(array);
  if (!h) throw Error("pull: no host mirror for this array", FST_ERR_ARG);
  check(fst_download(ctx_, array, h->data(), h->size()), "fst_download");
}

void rk_func::push(int array) {
  HostArray *h = mirror(array);
  if (!h) throw Error("push: no host mirror for this array", FST_ERR_ARG);
  check(fst_upload(ctx_, array, h->data(), h->size()), "fst_upload");
}

void rk_func::sync() { check(fst_sync(ctx_), "fst_sync"); }

namespace {
// variable names as src/cart3d.cpp:98-121 and src/cart2d.cpp:41-77 set them
void names(rk_func &f, const fst_config &cf, const fst_dims &d, const std::vector<std::string> &sp) {
  auto sname = [&](int s) { return s < (int)sp.size() ? sp[s] : "Species " + std::to_string(s); };
  if (cf.ndim == 3) {
    f.varNames = {"X-Momentum", "Y-Momentum", "Z-Momentum", "Energy"};
    for (int s = 0; s < cf.ns; ++s) f.varNames.push_back(sname(s) + " Density");
    if (d.nvt > d.nv) for (const char *n : {"C", "C_hat", "C_1", "C_2", "C_3"}) f.varNames.push_back(n);
    f.varxNames = {"X-Velocity", "Y-Velocity", "Z-Velocity", "Pressure", "Temperature", "Total Density"};
  } else {
    f.varNames = {"X-Momentum", "Y-Momentum", "Energy"};
    for (int s = 0; s < cf.ns; ++s) f.varNames.push_back(sname(s) + " Density");
    if (d.nvt > d.nv) for (const char *n : {"C", "C_hat", "Tau_1", "Tau_2", "Debug"}) f.varNames.push_back(n);
    f.varxNames = {"X-Velocity", "Y-Velocity", "Pressure", "Temperature", "Total Density"};
  }
}
}  // namespace

cart3d_func::cart3d_func(const fst_config &cf, const std::vector<std::string> &sp) : rk_func(cf) {
  if (cf.ndim != 3) throw Error("cart3d_func needs ndim = 3", FST_ERR_ARG);
  names(*this, cf_, d_, sp);
}
void cart3d_func::compute() { check(fst_compute(ctx_), "fst_compute"); }
void cart3d_func::preStep() { check(fst_pre_step(ctx_), "fst_pre_step"); }
void cart3d_func::postStep() { check(fst_post_step(ctx_), "fst_post_step"); }
void cart3d_func::preSim() { check(fst_pre_sim(ctx_), "fst_pre_sim"); }
void cart3d_func::postSim() {}

cart2d_func::cart2d_func(const fst_config &cf, const std::vector<std::string> &sp) : rk_func(cf) {
  if (cf.ndim != 2) throw Error("cart2d_func needs ndim = 2", FST_ERR_ARG);
  names(*this, cf_, d_, sp);
}
void cart2d_func::compute() { check(fst_compute(ctx_), "fst_compute"); }
void cart2d_func::preStep() { check(fst_pre_step(ctx_), "fst_pre_step"); }
void cart2d_func::postStep() { check(fst_post_step(ctx_), "fst_post_step"); }
void cart2d_func::preSim() { check(fst_pre_sim(ctx_), "fst_pre_sim"); }
void cart2d_func::postSim() {}

// variable names as src/gen3d.cpp:67-86 and src/gen2d.cpp:37-78 set them
void gen_names(rk_func &f, const fst_config &cf, const std::vector<std::string> &sp) {
  auto sname = [&](int s) { return s < (int)sp.size() ? sp[s] : "Species " + std::to_string(s); };
  if (cf.ndim == 3) {
    f.varNames = {"X-Momentum", "Y-Momentum", "Z-Momentum", "Energy"};
    f.varxNames = {"X-Velocity", "Y-Velocity", "Z-Velocity", "Pressure", "Temperature", "Density"};
  } else {
    f.varNames = {"X-Momentum", "Y-Momentum", "Energy"};
    f.varxNames = {"X-Velocity", "Y-Velocity", "Pressure", "Temperature", "Density"};
  }
  for (int s = 0; s < cf.ns; ++s) f.varNames.push_back("Density " + sname(s));
}

gen3d_func::gen3d_func(const fst_config &cf, const std::vector<std::string> &sp) : rk_func(cf) {
  if (cf.ndim != 3 || cf.grid != 1) throw Error("gen3d_func needs ndim = 3 and grid = 1", FST_ERR_ARG);
  gen_names(*this, cf_, sp);
}
void gen3d_func::compute() { check(fst_compute(ctx_), "fst_compute"); }
void gen3d_func::preStep() { check(fst_pre_step(ctx_), "fst_pre_step"); }
void gen3d_func::postStep() { check(fst_post_step(ctx_), "fst_post_step"); }
void gen3d_func::preSim() { check(fst_pre_sim(ctx_), "fst_pre_sim"); }
void gen3d_func::postSim() {}

gen2d_func::gen2d_func(const fst_config &cf, const std::vector<std::string> &sp) : rk_func(cf) {
  if (cf.ndim != 2 || cf.grid != 1) throw Error("gen2d_func needs ndim = 2 and grid = 1", FST_ERR_ARG);
  gen_names(*this, cf_, sp);
}
void gen2d_func::compute() { check(fst_compute(ctx_), "fst_compute"); }
void gen2d_func::preStep() { check(fst_pre_step(ctx_), "fst_pre_step"); }
void gen2d_func::postStep() { check(fst_post_step(ctx_), "fst_post_step"); }
void gen2d_func::preSim() { check(fst_pre_sim(ctx_), "fst_pre_sim"); }
void gen2d_func::postSim() {}

rk_func *make_module(const fst_config &cf, const std::vector<std::string> &sp) {
  if (cf.grid == 1) {
    if (cf.ndim == 3) return new gen3d_func(cf, sp);
    return new gen2d_func(cf, sp);
  }
  if (cf.ndim == 3) return new cart3d_func(cf, sp);
  return new cart2d_func(cf, sp);
}

void rkAdvance(rk_func &f) {
  fst_ctx *c = f.context();
  const int nstages = f.config().rk_order == 3 ? 3 : 2;  // RK2 midpoint, or the SSP-RK3 option
  for (int s = 1; s <= nstages; ++s) {
    f.compute();
    if (int rc = fst_rk_stage(c, s))
      throw Error("fst_rk_stage(" + std::to_string(s) + "): " + fst_last_error(c), rc);
  }
}

void applyBCs(rk_func &f) {
  fst_ctx *c = f.context();
  if (int rc = fst_apply_bcs(c)) throw Error(std::string("fst_apply_bcs: ") + fst_last_error(c), rc);
}

void step(rk_func &f, int nsteps) {
  fst_ctx *c = f.context();
  if (int rc = fst_step(c, nsteps)) throw Error(std::string("fst_step: ") + fst_last_error(c), rc);
  double ms = 0.0;
  if (fst_last_step_ms(c, &ms) == FST_OK) {
    Timer &t = f.timers["step"];
    t.description = "Fused RK steps (device time)";
    t.ms += ms;
  }
}

bool statusIsBad(double v) {  // isBad, src/status.cpp:100-105
  return !((std::isnormal(v) || v == 0) && (v > -1.0e+200 && v < 1.0e+200));
}

bool statusIsConcern(double v) { return v < -1.0e+16 || v > 1.0e+16; }  // isConcern, src/status.cpp:107-112

std::vector<VarStatus> statusCheck(rk_func &f, std::FILE *out, bool color, long long t, long long tend, double time) {
  std::vector<VarStatus> st;
  const struct {
    int array;
    const std::vector<std::string> *names;
  } sets[2] = {{FST_VAR, &f.varNames}, {FST_VARX, &f.varxNames}};
  for (const auto &set : sets) {
    const int n = (int)set.names->size();
    std::vector<double> mn(n), mx(n);
    if (int rc = fst_status(f.context(), set.array, mn.data(), mx.data(), n))
      throw Error(std::string("fst_status: ") + fst_last_error(f.context()), rc);
    for (int v = 0; v < n; ++v) st.push_back({(*set.names)[v], mn[v], mx[v]});
  }
  if (!out) return st;
  auto col = [&](double v) {
    if (!color) return "";
    return statusIsBad(v) ? "\033[0;31m" : statusIsConcern(v) ? "\033[0;33m" : "\033[0;35m";
  };
  const char *reset = color ? "\033[0m" : "";
  std::fprintf(out, "[%lld] Reporting Status\n", t);
  std::fprintf(out, "%8sTimestep:  %lld/%lld (%.0f%%)\n", "", t, tend, tend > 0 ? 100.0 * (double)t / (double)tend : 0.0);
  std::fprintf(out, "%8sSim Time:  %.2gs\n", "", time);
  std::fprintf(out, "%8s%-16s%11s%11s\n", "", "", "Min", "Max");
  for (const VarStatus &s : st)
    std::fprintf(out, "%8s%-16s%s%11.2e%s%s%11.2e%s\n", "", s.name.c_str(), col(s.min), s.min, reset, col(s.max), s.max,
                 reset);
  std::fflush(out);
  return st;
}

}  // namespace fiesta_hip
