/* fiesta_io.h -- solution and restart files in FIESTA's HDF5 + XDMF format,
 * written from and read into a device context (include/fiesta_hip.h).
 *
 * Host code (libfiesta_host.so, fiesta_amd/host/fiesta_io.cpp); the device
 * work is fst_pack (blockWriter::dataPack on the device) and fst_upload.  It
 * replaces, for one process owning the whole domain:
 *   blockWriter<T> constructors  src/block.cpp:44-98 (whole domain),
 *                                src/block.cpp:100-280 (start/limit/stride view)
 *   blockWriter<T>::write        src/block.cpp:282-401
 *   h5Writer<T>                  src/h5.cpp:33-352
 *   writeXMF                     src/xdmf.cpp:40-178
 *   readRestart                  src/reader.cpp:33-119
 * The C++ classes over the host rk_func layer are in fiesta_amd/host/fiesta_io.hpp.
 *
 * File layout (what the reference writes and reads):
 *   <path>/<name>-<step, zero padded>.h5   (<name>.h5 without append_step)
 *     /Grid/Dimension<d>    node coordinates, generalized grids only
 *     /Solution/<var name>  one dataset per variable of var (and of varx for
 *                           start/limit/stride views), dims reversed (C order
 *                           (k, j, i)), float (views) or double (restarts)
 *     /Properties           attributes time_index (int), time (double), title,
 *                           metadata, fiesta_version, fiesta_options,
 *                           fiesta_btime (strings), restart_file_version (int, 2)
 *   <path>/<same stem>.xmf  XDMF 3 descriptor of the .h5 file
 *
 * Every function returns 0 or an FST_ERR_* code and writes a message into
 * err (err_len bytes, may be 0) on failure; no exception crosses the ABI. */
#ifndef FIESTA_IO_H
#define FIESTA_IO_H

#include <stddef.h>
#include <stdint.h>

#include "fiesta_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FIO_RESTART_VERSION 2 /* FIESTA_RESTART_VERSION, src/fiesta.hpp:31 */

/* One writer: a restart writer (whole = 1, f64 = 1, average = 0) or an IO
 * view (src/luaReader.cpp:208-312: float, average defaults to 1; start,
 * limit and stride present -> whole = 0). */
typedef struct fio_block {
  const char *name;      /* file-name stem */
  const char *path;      /* directory; created when missing */
  int whole;             /* 1: every interior cell (src/block.cpp:44-98) */
  int64_t start[3];      /* whole = 0: first cell, 0-based (deck value - 1) */
  int64_t end[3];        /* whole = 0: last cell, inclusive, 0-based */
  int64_t stride[3];     /* whole = 0: 1..4 */
  int average;           /* pack the mean of each stride block */
  int append_step;       /* file name carries the step index */
  int f64;               /* 1: double datasets, 0: float */
  int chunk, compress;   /* cf.chunkable / cf.compressible (src/input.cpp:202-207) */
  int64_t tend;          /* cf.tend: the step index has (int)log10(tend) + 1 digits */
  const char *title;     /* cf.title */
  const char *metadata;  /* cf.metadata ("none" by default) */
} fio_block;

/* Extents a writer uses (single process: local = global, offset 0). */
typedef struct fio_extents {
  int64_t g_ext[3];    /* dataset extent, cells */
  int64_t g_ext_g[3];  /* dataset extent, grid nodes */
  int64_t l_start[3];  /* first packed cell */
  int64_t l_end[3];    /* last packed cell */
  int64_t l_end_g[3];  /* last packed node */
  int64_t l_ext[3];    /* packed extent, cells */
  int64_t l_ext_g[3];  /* packed extent, nodes */
  double origin[3];    /* XDMF origin: start * d */
  double spacing[3];   /* XDMF spacing: d * stride */
  int write_varx;      /* views also write varx */
  int chunk;           /* chunking in effect (never for views) */
} fio_extents;

/* The extents blockWriter's constructors compute, with their argument checks
 * (start and end inside the domain, start <= end, stride <= 4, stride not
 * larger than a non-empty range).  No device work. */
int fio_block_extents(const fst_config *cf, const fio_block *b, fio_extents *out, char *err, size_t err_len);

/* blockWriter::write: pack each variable on the device, write the .h5 and
 * .xmf files.  var_names has dims.nvt entries (the module's varNames),
 * varx_names up to dims.nvx (written by views only).  h5_path receives the
 * path of the file written (may be NULL). */
int fio_write(fst_ctx *ctx, const fst_config *cf, const fio_block *b, const char *const *var_names, int n_var,
              const char *const *varx_names, int n_varx, int tdx, double time, char *h5_path, size_t h5_path_len,
              char *err, size_t err_len);

/* readRestart: /Solution/<var_names[v]> into the interior of var (ghosts
 * zero, as a fresh host mirror copied to the device leaves them), the grid of
 * a generalized grid from /Grid/Dimension<d>; checks extents and
 * restart_file_version.  Unless reset, *tdx and *time receive the file's
 * time_index and time; title (title_len bytes) receives its title. */
int fio_read_restart(fst_ctx *ctx, const fst_config *cf, const char *fname, const char *const *var_names, int n_var,
                     int reset, int *tdx, double *time, char *title, size_t title_len, char *err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* FIESTA_IO_H */
