// fiesta_run: time steps through the host rk_func interface, with FIESTA's
// restart and solution files.
//
//   fiesta_run CONFIG STATE OUT NSTEPS [rk|fused] [GRID] [options]
//
// CONFIG is a raw fst_config (include/fiesta_hip.h), STATE the initial var
// (LayoutLeft doubles, nvt planes with ghosts; '-' with --restart-in), GRID
// the node coordinates of a generalized grid (LayoutLeft (ni,nj,nk,3)
// doubles; required when CONFIG.grid = 1 without --restart-in), OUT receives
// var after NSTEPS steps.  "rk" runs the module sequence of Fiesta::step
// (src/fiesta.cpp:333-340): preStep, rkAdvance (src/rk.cpp:28-72), postStep;
// "fused" runs fst_step (with the noise filter on, one step and one postStep
// at a time; otherwise all steps at once and one postStep when files are
// written, as the output step's postStep computes varx, src/cart3d.cpp:300-317).
//
// Options (the deck keys of src/input.cpp:130-207 and src/luaReader.cpp:208-312):
//   --species A,B,..          species names (variable names "<A> Density")
//   --title TEXT              title attribute
//   --tend N                  step-index width of file names: (int)log10(N) + 1 digits
//   --restart-in FILE         read var (and grid) from a restart file (readRestart)
//   --restart-out PATH NAME   after the steps, write the restart file PATH/NAME-<step>.h5
//   --view NAME PATH AVG      after the steps, write an IO view of the whole domain
//   --status                  after the steps, print statusCheck's table (--status-color: in colour)
//   --view-box NAME PATH AVG START LIMIT STRIDE
//                             an IO view of cells START..LIMIT (1-based, inclusive,
//                             comma lists as the deck's) every STRIDE cells
// Prints one JSON line with the step count, device time and the files written.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "fiesta_host.hpp"
#include "fiesta_io.hpp"

using namespace fiesta_hip;

static bool read_file(const char *path, void *dst, size_t bytes) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  f.read(static_cast<char *>(dst), (std::streamsize)bytes);
  return (size_t)f.gcount() == bytes && f.peek() == EOF;
}

namespace {

std::vector<std::string> split(const std::string &s) {
  std::vector<std::string> out;
  std::string cur;
  std::istringstream in(s);
  while (std::getline(in, cur, ',')) out.push_back(cur);
  return out;
}

bool to_sizes(const std::string &s, int n, long long sub, std::vector<size_t> &out) {
  const auto parts = split(s);
  if ((int)parts.size() != n) return false;
  out.clear();
  for (const auto &p : parts) {
    char *end = nullptr;
    const long long v = std::strtoll(p.c_str(), &end, 10);
    if (!end || *end || v - sub < 0) return false;
    out.push_back((size_t)(v - sub));
  }
  return true;
}

struct ViewArg {
  std::string name, path;
  bool avg = true, box = false;
  std::string start, limit, stride;
};

int usage(const char *prog) {
  std::fprintf(stderr,
               "usage: %s CONFIG STATE OUT NSTEPS [rk|fused] [GRID] [--species A,B] [--title T] [--tend N]\n"
               "       [--restart-in FILE] [--restart-out PATH NAME] [--view NAME PATH AVG] [--status[-color]]\n"
               "       [--view-box NAME PATH AVG START LIMIT STRIDE]\n",
               prog);
  return 2;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<std::string> pos;
  std::vector<std::string> species;
  std::vector<ViewArg> views;
  IoConfig io;
  std::string restart_in, rpath, rname;
  bool status = false, status_color = false;
  for (int a = 1; a < argc; ++a) {
    const std::string k = argv[a];
    auto need = [&](int n) { return a + n < argc; };
    if (k.rfind("--", 0) != 0) {
      pos.push_back(k);
    } else if (k == "--species" && need(1)) {
      species = split(argv[++a]);
    } else if (k == "--title" && need(1)) {
      io.title = argv[++a];
    } else if (k == "--tend" && need(1)) {
      io.tend = std::atoll(argv[++a]);
    } else if (k == "--status" || k == "--status-color") {
      status = true;
      status_color = k == "--status-color";
    } else if (k == "--restart-in" && need(1)) {
      restart_in = argv[++a];
    } else if (k == "--restart-out" && need(2)) {
      rpath = argv[a + 1];
      rname = argv[a + 2];
      a += 2;
    } else if (k == "--view" && need(3)) {
      ViewArg v;
      v.name = argv[a + 1];
      v.path = argv[a + 2];
      v.avg = std::atoi(argv[a + 3]) != 0;
      views.push_back(v);
      a += 3;
    } else if (k == "--view-box" && need(6)) {
      ViewArg v;
      v.name = argv[a + 1];
      v.path = argv[a + 2];
      v.avg = std::atoi(argv[a + 3]) != 0;
      v.box = true;
      v.start = argv[a + 4];
      v.limit = argv[a + 5];
      v.stride = argv[a + 6];
      views.push_back(v);
      a += 6;
    } else {
      std::fprintf(stderr, "fiesta_run: unknown or incomplete option %s\n", k.c_str());
      return usage(argv[0]);
    }
  }
  if (pos.size() < 4) return usage(argv[0]);
  const int nsteps = std::atoi(pos[3].c_str());
  const std::string mode = pos.size() > 4 ? pos[4] : "fused";
  if (nsteps < 0 || (mode != "rk" && mode != "fused")) {
    std::fprintf(stderr, "fiesta_run: bad NSTEPS or mode\n");
    return 2;
  }
  fst_config cf;
  if (!read_file(pos[0].c_str(), &cf, sizeof(cf))) {
    std::fprintf(stderr, "fiesta_run: %s is not a %zu-byte fst_config\n", pos[0].c_str(), sizeof(cf));
    return 2;
  }
  if (pos[1] == "-" && restart_in.empty()) {
    std::fprintf(stderr, "fiesta_run: STATE '-' needs --restart-in\n");
    return 2;
  }
  try {
    std::unique_ptr<rk_func> f(make_module(cf, species));
    // the writers, built (and their views checked) before any step runs
    std::vector<blockWriter<float>> writers;
    for (const ViewArg &v : views) {
      if (!v.box) {
        writers.emplace_back(*f, io, v.name, v.path, v.avg, 1, true);
        continue;
      }
      std::vector<size_t> st, li, sd;
      if (!to_sizes(v.start, cf.ndim, 1, st) || !to_sizes(v.limit, cf.ndim, 1, li) ||
          !to_sizes(v.stride, cf.ndim, 0, sd)) {
        std::fprintf(stderr, "fiesta_run: view '%s': START, LIMIT and STRIDE need %d comma-separated integers\n",
                     v.name.c_str(), cf.ndim);
        return 2;
      }
      writers.emplace_back(*f, io, v.name, v.path, v.avg, 1, st, li, sd, true);
    }
    int tstart = 0;
    double time0 = 0.0;
    if (!restart_in.empty()) {
      const RestartInfo r = readRestart(*f, restart_in, false);
      tstart = r.tstart;
      time0 = r.time;
    } else {
      if (!read_file(pos[1].c_str(), f->var.data(), f->var.size() * sizeof(double))) {
        std::fprintf(stderr, "fiesta_run: %s does not hold %zu doubles\n", pos[1].c_str(), f->var.size());
        return 2;
      }
      f->push(FST_VAR);
      if (cf.grid == 1) {
        if (pos.size() < 6 || !read_file(pos[5].c_str(), f->grid.data(), f->grid.size() * sizeof(double))) {
          std::fprintf(stderr, "fiesta_run: a generalized grid needs GRID with %zu doubles\n", f->grid.size());
          return 2;
        }
        f->push(FST_GRID);
      }
    }
    f->preSim();
    if (mode == "fused") {
      if (int rc = fst_set_fused(f->context(), 1)) {
        std::fprintf(stderr, "fiesta_run: fused mode: %s (code %d); use \"rk\"\n", fst_last_error(f->context()), rc);
        return 2;
      }
    }
    f->sync();
    const auto t0 = std::chrono::steady_clock::now();
    if (mode == "rk") {
      for (int t = 0; t < nsteps; ++t) {
        f->preStep();
        rkAdvance(*f);
        f->postStep();
      }
    } else if (cf.noise) {
      // the noise filter is part of every step's postStep (Fiesta::step,
      // src/fiesta.cpp:333-340), so the fused steps go one at a time
      for (int t = 0; t < nsteps; ++t) {
        step(*f, 1);
        f->postStep();
      }
    } else {
      step(*f, nsteps);
    }
    f->sync();
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const bool files = !writers.empty() || !rpath.empty();
    // without the filter postStep only computes varx, which the output step needs
    if (mode == "fused" && !cf.noise && (files || status)) f->postStep();
    // checkIO at the last step (src/fiesta.cpp:194-272): the restart, then the views
    const int tdx = tstart + nsteps;
    const double time = time0 + nsteps * cf.dt;
    std::vector<std::string> written;
    if (!rpath.empty()) {
      blockWriter<double> restart(*f, io, rname, rpath, false, 0, true);
      written.push_back(restart.write(*f, tdx, time));
    }
    for (auto &w : writers) written.push_back(w.write(*f, tdx, time));
    if (status) statusCheck(*f, stdout, status_color, tdx, io.tend, time);
    f->pull(FST_VAR);
    f->postSim();
    std::ofstream o(pos[2], std::ios::binary);
    o.write(reinterpret_cast<const char *>(f->var.data()), (std::streamsize)(f->var.size() * sizeof(double)));
    if (!o) {
      std::fprintf(stderr, "fiesta_run: cannot write %s\n", pos[2].c_str());
      return 1;
    }
    const fst_dims &d = f->dims();
    std::string fl;
    for (const auto &w : written) fl += (fl.empty() ? "\"" : ", \"") + w + "\"";
    std::printf("{\"steps\": %d, \"mode\": \"%s\", \"cells\": %lld, \"wall_s\": %.6f, \"cell_stages_per_s\": %.6e, "
                "\"vars\": \"%s..%s\", \"tstart\": %d, \"time\": %.17g, \"files\": [%s]}\n",
                nsteps, mode.c_str(), (long long)d.ncells, wall,
                wall > 0 ? 2.0 * (double)d.ncells * nsteps / wall : 0.0, f->varNames.front().c_str(),
                f->varNames.back().c_str(), tstart, time, fl.c_str());
  } catch (const Error &e) {
    std::fprintf(stderr, "fiesta_run: %s (code %d)\n", e.what(), e.code);
    return 1;
  }
  return 0;
}
