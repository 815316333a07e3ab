// Host-side rk_func interface over the C-ABI (include/fiesta_hip.h).
//
// Mirrors the reference's module interface so host code written against it
// keeps its shape:
//   class rk_func            src/rkfunction.hpp:36-69  (compute, preStep,
//                            postStep, preSim, postSim; var, varx, dvar,
//                            tmp1, grid; varNames, varxNames; timers)
//   cart3d_func/cart2d_func  src/cart3d.cpp:44-378, src/cart2d.cpp:30-311
//   gen3d_func/gen2d_func    src/gen3d.cpp:39-370, src/gen2d.cpp:22-333
//   rkAdvance(cf, f)         src/rk.cpp:28-72
//   applyBCs(cf, f)          src/bc.cpp:179-249
//   Fiesta::step             src/fiesta.cpp:333-340
// The device owns every array; var/varx/dvar/tmp1 here are host mirrors
// (LayoutLeft, ghosts included) filled by pull() and sent back by push(), the
// explicit copies the reference's writers and readers make with
// Kokkos::deep_copy (src/block.cpp:106-109, src/reader.cpp:33-87).  Errors of
// the C-ABI are turned into fiesta_hip::Error exceptions at this layer.
#pragma once
#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "fiesta_hip.h"

namespace fiesta_hip {

class Error : public std::runtime_error {
 public:
  Error(const std::string &what, int code) : std::runtime_error(what), code(code) {}
  int code;
};

// Host mirror of one device array: (ngi, ngj, ngk, nvar), i fastest.
class HostArray {
 public:
  HostArray() = default;
  HostArray(int ni, int nj, int nk, int nv) { resize(ni, nj, nk, nv); }
  void resize(int ni, int nj, int nk, int nv) {
    ext_[0] = ni; ext_[1] = nj; ext_[2] = nk; ext_[3] = nv;
    data_.assign((size_t)ni * nj * nk * nv, 0.0);
  }
  double &operator()(int i, int j, int k, int v) {
    return data_[(size_t)i + (size_t)ext_[0] * ((size_t)j + (size_t)ext_[1] * ((size_t)k + (size_t)ext_[2] * v))];
  }
  double operator()(int i, int j, int k, int v) const {
    return data_[(size_t)i + (size_t)ext_[0] * ((size_t)j + (size_t)ext_[1] * ((size_t)k + (size_t)ext_[2] * v))];
  }
  int extent(int d) const { return ext_[d]; }
  double *data() { return data_.data(); }
  const double *data() const { return data_.data(); }
  size_t size() const { return data_.size(); }

 private:
  int ext_[4] = {0, 0, 0, 0};
  std::vector<double> data_;
};

// Elapsed device time of the calls bracketed by reset()/accumulate()
// (Timer::fiestaTimer, src/timer.hpp:52-119), read from fst_last_step_ms.
struct Timer {
  std::string description;
  double ms = 0.0;
};

class rk_func {
 public:
  explicit rk_func(const fst_config &cf);
  virtual ~rk_func();
  rk_func(const rk_func &) = delete;
  rk_func &operator=(const rk_func &) = delete;

  virtual void compute() = 0;
  virtual void preStep() = 0;
  virtual void postStep() = 0;
  virtual void preSim() = 0;
  virtual void postSim() = 0;

  HostArray var, varx, dvar, tmp1;
  HostArray grid;  // node coordinates (ni, nj, nk, 3); generalized grids only
  std::vector<std::string> varNames, varxNames;
  std::map<std::string, Timer> timers;

  // device <-> host copies of one array (FST_VAR, FST_VARX, FST_DVAR, FST_TMP1,
  // and FST_GRID on generalized grids)
  void pull(int array);
  void push(int array);
  void sync();
  fst_ctx *context() const { return ctx_; }
  const fst_config &config() const { return cf_; }
  const fst_dims &dims() const { return d_; }

 protected:
  void check(int rc, const char *what) const;
  fst_config cf_;
  fst_dims d_{};
  fst_ctx *ctx_ = nullptr;

 private:
  HostArray *mirror(int array);
};

// cart3d_func: compute() applies the boundary conditions, then the RHS
// (src/cart3d.cpp:220-298); preStep does nothing (:187-188).
class cart3d_func : public rk_func {
 public:
  explicit cart3d_func(const fst_config &cf, const std::vector<std::string> &speciesName = {});
  void compute() override;
  void preStep() override;
  void postStep() override;
  void preSim() override;
  void postSim() override;
};

// cart2d_func: boundary conditions once per step in preStep
// (src/cart2d.cpp:147-149), RHS in compute() (:151-237).
class cart2d_func : public rk_func {
 public:
  explicit cart2d_func(const fst_config &cf, const std::vector<std::string> &speciesName = {});
  void compute() override;
  void preStep() override;
  void postStep() override;
  void preSim() override;
  void postSim() override;
};

// gen3d_func / gen2d_func: generalized (curvilinear) grids.  preSim computes
// the metrics of `grid` (push(FST_GRID) first) and their mirrored ghosts
// (src/gen3d.cpp:131-304, src/gen2d.cpp:187-331); compute() evaluates the EOS,
// contravariant velocity, WENO fluxes and metric pressure gradient
// (src/gen3d.cpp:341-369, src/gen2d.cpp:100-131).  Neither module applies
// boundary conditions; preStep does nothing.
class gen3d_func : public rk_func {
 public:
  explicit gen3d_func(const fst_config &cf, const std::vector<std::string> &speciesName = {});
  void compute() override;
  void preStep() override;
  void postStep() override;
  void preSim() override;
  void postSim() override;
};

class gen2d_func : public rk_func {
 public:
  explicit gen2d_func(const fst_config &cf, const std::vector<std::string> &speciesName = {});
  void compute() override;
  void preStep() override;
  void postStep() override;
  void preSim() override;
  void postSim() override;
};

// The module the reference builds for a configuration (src/fiesta.cpp: cart2d,
// cart3d, gen2d or gen3d from grid.ndim and grid.type).
rk_func *make_module(const fst_config &cf, const std::vector<std::string> &speciesName = {});

// rkAdvance (src/rk.cpp:28-72): compute, stage-1 update, compute, stage-2
// update; with rk_order = 3, the three compute + stage-update pairs of SSP-RK3.
void rkAdvance(rk_func &f);
// applyBCs (src/bc.cpp:179-249).
void applyBCs(rk_func &f);
// Fiesta::step (src/fiesta.cpp:333-340) for nsteps steps on the fused path
// (one call, results equal to preStep + rkAdvance + postStep without the
// postStep secondary-variable pass, which runs when called explicitly).
void step(rk_func &f, int nsteps);

// statusCheck (src/status.cpp:155-251): min and max of every variable of var
// and varx over the interior cells, reduced on the device (fst_status, so no
// array is copied to the host), printed to `out` (NULL: not printed) as the
// reference's table. isBad values (not normal and not 0, or beyond 1e200
// in magnitude) print red, isConcern values (beyond 1e16) yellow, others
// magenta, when `color` is set. varx holds what the last postStep computed.
struct VarStatus {
  std::string name;
  double min = 0.0, max = 0.0;
};
bool statusIsBad(double v);
bool statusIsConcern(double v);
std::vector<VarStatus> statusCheck(rk_func &f, std::FILE *out, bool color, long long t, long long tend, double time);

}  // namespace fiesta_hip
