/*
 * fiesta_deck.h -- C-ABI of the Lua input deck reader (host code, no GPU).
 *
 * FIESTA is configured by a Lua script that fills the root table `fiesta`
 * (src/luaReader.cpp:36-73 creates it with its sub-tables) and defines the
 * callbacks fiesta.initial_conditions(x, y, z, v) and
 * fiesta.initialize_grid(i, j, k, v).  This library runs such a script with
 * an interpreter of its own (the Lua subset decks use: numbers, strings,
 * booleans, tables and constructors, local and global variables, functions
 * and closures, if / while / repeat / numeric and generic for, multiple
 * assignment, the arithmetic, comparison, logical, concatenation and length
 * operators, the math library and print / type / tostring / tonumber /
 * pairs / ipairs / error / assert / string.format) and maps it onto the
 * library's fst_config exactly as executeConfiguration does.
 *
 * Reference interfaces replaced (paths relative to the FIESTA source tree):
 *   luaReader (constructor, get, getSpeciesData, getIOBlock, call)
 *                                   src/luaReader.cpp:36-345, src/luaReader.hpp
 *   executeConfiguration            src/input.cpp:130-443
 *   loadInitialConditions           src/input.cpp:445-516
 *   loadGrid                        src/input.cpp:527-557
 *
 * Every function returns 0 on success and a nonzero FDK_ERR_* code otherwise,
 * with a message from fdk_last_error(); the reference prints and exits.
 */
#ifndef FIESTA_DECK_H
#define FIESTA_DECK_H

#include <stddef.h>
#include <stdint.h>

#include "fiesta_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum fdk_err {
  FDK_OK = 0,
  FDK_ERR_ARG = 1,     /* bad argument, or a deck value of the wrong kind / out of range */
  FDK_ERR_MISSING = 2, /* a required key is absent (luaReader::get without default) */
  FDK_ERR_SCRIPT = 3,  /* the script does not parse or raises an error */
  FDK_ERR_IO = 4       /* the file cannot be read */
};

typedef struct fdk_deck fdk_deck;

/* The deck's settings off the right-hand-side path (the rest of
 * inputConfig that executeConfiguration fills). */
typedef struct fdk_run {
  char title[256];
  char metadata[256];
  int64_t nt, tstart, tend; /* time.nt or time.tend - time.start_index */
  int32_t tinterval;        /* 1: time.nt given, 0: time.tend given */
  double time0;             /* time.start_time */
  int64_t out_freq;         /* progress.frequency */
  int64_t stat_freq;        /* status.frequency */
  int64_t write_freq;       /* write_frequency */
  int32_t restart, restart_reset, auto_restart;
  int64_t restart_freq, restart_time_remaining;
  char restart_name[512];   /* restart.name (auto restart: path/auto_name.h5) */
  char restart_path[512];   /* restart.path */
  char auto_restart_name[256];
  int32_t hdf5_chunk, hdf5_compress;
  char mpi_type[32];
  int32_t procs[3];
  char terrain_name[256];
  double tdx, tdy, h;       /* terrain grids */
  double beta, betae;       /* ceq.beta, ceq.betae */
  int64_t st;               /* ceq.st */
  int32_t n_ioviews;
} fdk_run;

/* One entry of fiesta.ioviews (getIOBlock, src/luaReader.cpp:208-312):
 * whole = 1 when start, limit and stride are all absent (the whole-domain
 * writer); start/limit are 0-based cell indices (the deck's are 1-based). */
typedef struct fdk_ioview {
  char name[256];
  char path[512];
  int64_t frequency;
  int32_t average;
  int32_t whole;
  int64_t start[3], limit[3], stride[3];
} fdk_ioview;

/* Run the script at `path` (luaReader's constructor). */
int fdk_open(const char *path, fdk_deck **out);
int fdk_close(fdk_deck *d);
const char *fdk_last_error(const fdk_deck *d);

/* executeConfiguration: fills *cf (fst_config_init defaults first) and *run.
 * Extensions of this build, read with defaults: fiesta.precision ("fp64",
 * "fp32" or "fp64_fast"), fiesta.ceq.learned (false), fiesta.time.integrator
 * ("rk2" or "ssp-rk3").  mpi.procs is required as in the reference; it is
 * recorded, and a decomposition into several ranks is run by this one
 * process on one GPU (a note goes to stderr). */
int fdk_read_config(fdk_deck *d, fst_config *cf, fdk_run *run);

/* Name of species s (getSpeciesData), NUL-terminated into buf. */
int fdk_species_name(fdk_deck *d, int s, char *buf, size_t len);

/* IO view i of fiesta.ioviews (0-based). */
int fdk_get_ioview(fdk_deck *d, int i, fdk_ioview *out);

/* loadGrid: node coordinates of a generalized grid from
 * fiesta.initialize_grid(i, j, k, v) for v < ndim, into `grid` (LayoutLeft
 * (ni, nj, nk, 3) doubles, `count` of them; the v >= ndim planes are
 * zeroed).  Evaluated on up to `threads` host threads (0: the host's
 * count), each running its own copy of the script. */
int fdk_grid(fdk_deck *d, const fst_config *cf, double *grid, size_t count, int threads);

/* loadInitialConditions: var(i, j, k, v) = fiesta.initial_conditions(x, y,
 * z, v) for every interior cell and v < nv, at the cell centre (cartesian:
 * dx*(i - ng) + dx/2, ...; generalized: the mean of the cell's 8 (2D: 4)
 * nodes of `grid`, which may be NULL on cartesian grids).  `var` holds nvt
 * planes with ghosts (count values); ghosts and the C-equation variables
 * are set to 0. */
int fdk_initial_conditions(fdk_deck *d, const fst_config *cf, const double *grid, double *var, size_t count,
                           int threads);

/* luaReader::call: fiesta.<fname>(args...) must return a number. */
int fdk_call(fdk_deck *d, const char *fname, int nargs, const double *args, double *out);

/* luaReader::get: the value at a dotted key below the root table
 * ("time.dt", "species.1.gamma"; integer parts index arrays).  kind: 0 nil,
 * 1 boolean (num = 0/1), 2 number, 3 string (into str), 4 table, 5 function. */
int fdk_get(fdk_deck *d, const char *key, int *kind, double *num, char *str, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* FIESTA_DECK_H */
