// fiesta_run: time steps through the host rk_func interface.
//
//   fiesta_run CONFIG STATE OUT NSTEPS [rk|fused] [GRID]
//
// CONFIG is a raw fst_config (include/fiesta_hip.h), STATE the initial var
// (LayoutLeft doubles, nvt planes with ghosts), GRID the node coordinates of
// a generalized grid (LayoutLeft (ni,nj,nk,3) doubles; required when
// CONFIG.grid = 1), OUT receives var after NSTEPS
// steps.  "rk" runs the module sequence of Fiesta::step (src/fiesta.cpp:
// 333-340): preStep, rkAdvance (src/rk.cpp:28-72), postStep; "fused" runs
// fst_step.  Prints one JSON line with the step count and device time.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>

#include "fiesta_host.hpp"

using namespace fiesta_hip;

static bool read_file(const char *path, void *dst, size_t bytes) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.read(static_cast<char *>(dst), (std::streamsize)bytes);
  return (size_t)f.gcount() == bytes && f.peek() == EOF;
}

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s CONFIG STATE OUT NSTEPS [rk|fused] [GRID]\n", argv[0]);
    return 2;
  }
  const int nsteps = std::atoi(argv[4]);
  const std::string mode = argc > 5 ? argv[5] : "fused";
  if (nsteps < 0 || (mode != "rk" && mode != "fused")) {
    std::fprintf(stderr, "fiesta_run: bad NSTEPS or mode\n");
    return 2;
  }
  fst_config cf;
  if (!read_file(argv[1], &cf, sizeof(cf))) {
    std::fprintf(stderr, "fiesta_run: %s is not a %zu-byte fst_config\n", argv[1], sizeof(cf));
    return 2;
  }
  try {
    std::unique_ptr<rk_func> f(make_module(cf));
    if (!read_file(argv[2], f->var.data(), f->var.size() * sizeof(double))) {
      std::fprintf(stderr, "fiesta_run: %s does not hold %zu doubles\n", argv[2], f->var.size());
      return 2;
    }
    f->push(FST_VAR);
    if (cf.grid == 1) {
      if (argc < 7 || !read_file(argv[6], f->grid.data(), f->grid.size() * sizeof(double))) {
        std::fprintf(stderr, "fiesta_run: a generalized grid needs GRID with %zu doubles\n", f->grid.size());
        return 2;
      }
      f->push(FST_GRID);
    }
    f->preSim();
    if (mode == "fused") fst_set_fused(f->context(), 1);
    f->sync();
    const auto t0 = std::chrono::steady_clock::now();
    if (mode == "rk") {
      for (int t = 0; t < nsteps; ++t) {
        f->preStep();
        rkAdvance(*f);
        f->postStep();
      }
    } else {
      step(*f, nsteps);
    }
    f->sync();
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    f->pull(FST_VAR);
    f->postSim();
    std::ofstream o(argv[3], std::ios::binary);
    o.write(reinterpret_cast<const char *>(f->var.data()), (std::streamsize)(f->var.size() * sizeof(double)));
    if (!o) {
      std::fprintf(stderr, "fiesta_run: cannot write %s\n", argv[3]);
      return 1;
    }
    const fst_dims &d = f->dims();
    std::printf("{\"steps\": %d, \"mode\": \"%s\", \"cells\": %lld, \"wall_s\": %.6f, \"cell_stages_per_s\": %.6e, "
                "\"vars\": \"%s..%s\"}\n",
                nsteps, mode.c_str(), (long long)d.ncells, wall,
                wall > 0 ? 2.0 * (double)d.ncells * nsteps / wall : 0.0, f->varNames.front().c_str(),
                f->varNames.back().c_str());
  } catch (const Error &e) {
    std::fprintf(stderr, "fiesta_run: %s (code %d)\n", e.what(), e.code);
    return 1;
  }
  return 0;
}
