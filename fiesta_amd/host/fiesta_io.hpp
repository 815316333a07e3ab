// Solution and restart writers and the restart reader over the host rk_func
// layer, in FIESTA's HDF5 + XDMF file format (the C-ABI and the file layout
// are in include/fiesta_io.h).
//
//   blockWriter<T>   src/block.hpp:31-103, src/block.cpp:44-454
//   readRestart      src/reader.cpp:33-119
//
// T = float for IO views (src/luaReader.cpp:208-312), double for the restart
// writer (src/fiesta.cpp:184).  One process owns the whole domain.
#pragma once
#include <string>
#include <type_traits>
#include <vector>

#include "fiesta_host.hpp"
#include "fiesta_io.h"

namespace fiesta_hip {

// Deck values the writers read from inputConfig.
struct IoConfig {
  std::string title;
  std::string metadata = "none";    // src/input.cpp:162
  long long tend = 1;               // step-index width (src/block.cpp:72)
  bool chunkable = false;           // hdf5.chunk (src/input.cpp:202)
  bool compressible = false;        // hdf5.compress, needs chunk (src/input.cpp:203-207)
};

template <typename T>
class blockWriter {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "float or double datasets");

 public:
  // every interior cell (src/block.cpp:44-98)
  blockWriter(const rk_func &f, const IoConfig &io, std::string name, std::string path, bool avg, size_t freq,
              bool appStep);
  // cells start..end (inclusive, 0-based) every stride (src/block.cpp:100-280)
  blockWriter(const rk_func &f, const IoConfig &io, std::string name, std::string path, bool avg, size_t freq,
              std::vector<size_t> start, std::vector<size_t> end, std::vector<size_t> stride, bool appStep);

  // src/block.cpp:282-401; returns the path of the .h5 file
  std::string write(rk_func &f, int tdx, double time);
  size_t frq() const { return freq_; }
  const fio_extents &extents() const { return ext_; }

 private:
  void init(const rk_func &f);
  IoConfig io_;
  std::string name_, path_;
  size_t freq_;
  fio_block b_{};
  fio_extents ext_{};
};

struct RestartInfo {
  int tstart = 0;
  double time = 0.0;
  std::string title;
};

// readRestart (src/reader.cpp:33-119) into f's device arrays; with reset the
// returned index and time are left at 0 (the deck's values apply).
RestartInfo readRestart(rk_func &f, const std::string &fname, bool reset);

}  // namespace fiesta_hip
