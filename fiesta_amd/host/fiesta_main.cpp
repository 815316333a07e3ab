// fiesta: run a FIESTA input deck (Lua) on one GPU, as the reference's main
// program does (src/main.cpp:40-85, src/fiesta.cpp:140-340):
//
//   fiesta DECK [--mode fused|rk] [--status-color] [--out FILE] [--threads N]
//
// The deck is read by the library's own Lua interpreter (include/fiesta_deck.h);
// then, without a restart, the grid (generalized grids: fiesta.initialize_grid)
// and the initial state (fiesta.initial_conditions) are evaluated, or a
// restart file is read; preSim; the time loop runs t = tstart .. tend with
// checkIO (status table, restart file, IO views at their frequencies) before
// each step, and each step is preStep, rkAdvance, postStep.  As in the
// reference the loop runs tend - tstart + 1 steps and the last checkIO is at
// t = tend, before the last step.
//   --mode fused   the fused 3D stage kernels where the configuration allows
//                  them (default), else the module's kernel sequence
//   --mode rk      the module's compute()/rk_stage sequence (rkAdvance)
//   --out FILE     write the final var (LayoutLeft doubles, ghosts included)
//   --threads N    host threads evaluating the deck's callbacks (0: all)
// Prints one JSON line: steps, mode, device time, cell-stages per second,
// the files written.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "fiesta_deck.h"
#include "fiesta_host.hpp"
#include "fiesta_io.hpp"

using namespace fiesta_hip;

namespace {

int usage(const char *prog) {
  std::fprintf(stderr, "usage: %s DECK [--mode fused|rk] [--status-color] [--out FILE] [--threads N]\n", prog);
  return 2;
}

struct DeckHandle {
  fdk_deck *d = nullptr;
  ~DeckHandle() {
    if (d) fdk_close(d);
  }
};

}  // namespace

int main(int argc, char **argv) {
  std::string deck, mode = "fused", out;
  bool color = false;
  int threads = 0;
  for (int a = 1; a < argc; ++a) {
    const std::string k = argv[a];
    if (k == "--mode" && a + 1 < argc) {
      mode = argv[++a];
    } else if (k == "--status-color") {
      color = true;
    } else if (k == "--out" && a + 1 < argc) {
      out = argv[++a];
    } else if (k == "--threads" && a + 1 < argc) {
      threads = std::atoi(argv[++a]);
    } else if (k.rfind("--", 0) == 0 || !deck.empty()) {
      return usage(argv[0]);
    } else {
      deck = k;
    }
  }
  if (deck.empty() || (mode != "fused" && mode != "rk")) return usage(argv[0]);
  DeckHandle dh;
  if (int rc = fdk_open(deck.c_str(), &dh.d)) return rc == FDK_ERR_IO ? 2 : 1;
  fst_config cf;
  fdk_run run;
  if (fdk_read_config(dh.d, &cf, &run)) {
    std::fprintf(stderr, "fiesta: %s\n", fdk_last_error(dh.d));
    return 1;
  }
  if (cf.grid == 2) {
    std::fprintf(stderr, "fiesta: grid.type = \"terrain\" has no module in the reference (nor here)\n");
    return 1;
  }
  if (run.write_freq > 0)
    std::fprintf(stderr, "fiesta: write_frequency is the reference's retired writer; use ioviews (ignored)\n");
  std::vector<std::string> species((size_t)cf.ns);
  for (int s = 0; s < cf.ns; ++s) {
    char buf[256];
    fdk_species_name(dh.d, s, buf, sizeof buf);
    species[(size_t)s] = buf;
  }
  IoConfig io;
  io.title = run.title;
  io.metadata = run.metadata;
  io.tend = run.tend;
  io.chunkable = run.hdf5_chunk != 0;
  io.compressible = run.hdf5_compress != 0;
  std::vector<std::string> written;
  try {
    std::unique_ptr<rk_func> f(make_module(cf, species));
    long long tstart = run.tstart;
    double time = run.time0;
    // initializeSimulation, src/fiesta.cpp:157-190
    if (!run.restart) {
      if (cf.grid == 1) {
        if (fdk_grid(dh.d, &cf, f->grid.data(), f->grid.size(), threads)) {
          std::fprintf(stderr, "fiesta: %s\n", fdk_last_error(dh.d));
          return 1;
        }
        f->push(FST_GRID);
      }
      if (fdk_initial_conditions(dh.d, &cf, cf.grid == 1 ? f->grid.data() : nullptr, f->var.data(), f->var.size(),
                                 threads)) {
        std::fprintf(stderr, "fiesta: %s\n", fdk_last_error(dh.d));
        return 1;
      }
      f->push(FST_VAR);
    } else {
      const RestartInfo r = readRestart(*f, run.restart_name, run.restart_reset != 0);
      if (!run.restart_reset) {
        tstart = r.tstart;
        time = r.time;
      }
    }
    const long long tend = run.tinterval ? tstart + run.nt : run.tend;
    std::filesystem::create_directories(run.restart_path);
    blockWriter<double> restartview(*f, io, run.auto_restart_name, run.restart_path, false, (size_t)run.restart_freq,
                                    !run.auto_restart);
    std::vector<blockWriter<float>> views;
    for (int i = 0; i < run.n_ioviews; ++i) {
      fdk_ioview v;
      fdk_get_ioview(dh.d, i, &v);
      if (v.whole) {
        views.emplace_back(*f, io, v.name, v.path, v.average != 0, (size_t)v.frequency, true);
      } else {
        std::vector<size_t> st, li, sd;
        for (int a = 0; a < cf.ndim; ++a) {
          st.push_back((size_t)v.start[a]);
          li.push_back((size_t)v.limit[a]);
          sd.push_back((size_t)v.stride[a]);
        }
        views.emplace_back(*f, io, v.name, v.path, v.average != 0, (size_t)v.frequency, st, li, sd, true);
      }
    }
    f->preSim();
    bool fused = mode == "fused" && fst_set_fused(f->context(), 1) == FST_OK;
    if (!fused) fst_set_fused(f->context(), 0);
    f->sync();
    const auto t0 = std::chrono::steady_clock::now();
    double device_ms = 0.0;
    long long steps = 0;
    for (long long t = tstart; t < tend + 1; ++t) {
      // checkIO, src/fiesta.cpp:194-272
      if (run.out_freq > 0 && t % run.out_freq == 0)
        std::printf("[%lld] Timestep %lld of %lld. Simulation Time: %.2es\n", t, t, tend, time);
      if (run.stat_freq > 0 && t % run.stat_freq == 0) statusCheck(*f, stdout, color, t, tend, time);
      if (run.restart_freq > 0 && t > tstart && t % run.restart_freq == 0)
        written.push_back(restartview.write(*f, (int)t, time));
      for (auto &v : views)
        if (v.frq() > 0 && t % (long long)v.frq() == 0) written.push_back(v.write(*f, (int)t, time));
      // Fiesta::step, src/fiesta.cpp:333-340
      if (fused) {
        step(*f, 1);  // preStep + rkAdvance on the fused path
        double ms = 0.0;
        if (fst_last_step_ms(f->context(), &ms) == FST_OK) device_ms += ms;
      } else {
        f->preStep();
        rkAdvance(*f);
      }
      f->postStep();
      time += cf.dt;
      ++steps;
    }
    f->sync();
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    f->postSim();
    if (!out.empty()) {
      f->pull(FST_VAR);
      std::ofstream o(out, std::ios::binary);
      o.write(reinterpret_cast<const char *>(f->var.data()), (std::streamsize)(f->var.size() * sizeof(double)));
      if (!o) {
        std::fprintf(stderr, "fiesta: cannot write %s\n", out.c_str());
        return 1;
      }
    }
    const fst_dims &d = f->dims();
    std::string fl;
    for (const auto &w : written) fl += (fl.empty() ? "\"" : ", \"") + w + "\"";
    const int nst = cf.rk_order == 3 ? 3 : 2;
    std::printf("{\"deck\": \"%s\", \"steps\": %lld, \"tstart\": %lld, \"tend\": %lld, \"mode\": \"%s\", "
                "\"cells\": %lld, \"wall_s\": %.6f, \"fused_device_ms\": %.3f, \"cell_stages_per_s\": %.6e, "
                "\"time\": %.17g, \"files\": [%s]}\n",
                deck.c_str(), steps, tstart, tend, fused ? "fused" : "rk", (long long)d.ncells, wall, device_ms,
                wall > 0 ? (double)nst * (double)d.ncells * (double)steps / wall : 0.0, time, fl.c_str());
  } catch (const Error &e) {
    std::fprintf(stderr, "fiesta: %s (code %d)\n", e.what(), e.code);
    return 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "fiesta: %s\n", e.what());
    return 1;
  }
  return 0;
}
