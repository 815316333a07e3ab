// FIESTA's solution/restart files (HDF5 + XDMF) from and into a device
// context; the C-ABI is include/fiesta_io.h, the C++ classes fiesta_io.hpp.
//
// Only the packed values cross to the host: each dataset is packed on the
// device by fst_pack (the reference copies the whole var/varx/grid arrays to
// host mirrors first, src/block.cpp:333-347) and written with one H5Dwrite.
#include "fiesta_io.hpp"

#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <utility>
#include <vector>

namespace fiesta_hip {
namespace {

// ------------------------------------------------------------------ HDF5 handles

// An HDF5 identifier closed by its own close function when it leaves scope.
class Hid {
 public:
  using Closer = herr_t (*)(hid_t);
  Hid(hid_t id, Closer c, const std::string &what) : id_(id), close_(c) {
    if (id_ < 0) throw Error("HDF5: " + what + " failed", FST_ERR_ARG);
  }
  ~Hid() {
    if (id_ >= 0 && close_) close_(id_);
  }
  Hid(const Hid &) = delete;
  Hid &operator=(const Hid &) = delete;
  operator hid_t() const { return id_; }

 private:
  hid_t id_;
  Closer close_;
};

void h5check(herr_t rc, const std::string &what) {
  if (rc < 0) throw Error("HDF5: " + what + " failed", FST_ERR_ARG);
}

template <typename T>
hid_t h5type() {
  return std::is_same<T, float>::value ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
}

// h5Writer::write (src/h5.cpp:69-138) for one process: the dataset spans the
// global extent, this process writes its local box at `off`.  Extents are
// given i first and reversed into HDF5's C order.
template <typename T>
void write_dataset(hid_t group, const std::string &name, int ndim, const int64_t *gext, const int64_t *lext,
                   const int64_t *off, const T *data, bool chunk, bool compress) {
  hsize_t g[3], l[3], o[3];
  for (int a = 0; a < ndim; ++a) {
    g[a] = (hsize_t)gext[ndim - 1 - a];
    l[a] = (hsize_t)lext[ndim - 1 - a];
    o[a] = (hsize_t)off[ndim - 1 - a];
  }
  hid_t dcpl_id = H5P_DEFAULT;
  if (chunk) {
    dcpl_id = H5Pcreate(H5P_DATASET_CREATE);
    if (dcpl_id < 0) throw Error("HDF5: dataset creation property list failed", FST_ERR_ARG);
  }
  Hid dcpl(dcpl_id, chunk ? H5Pclose : nullptr, "H5Pcreate");
  if (chunk) {
    h5check(H5Pset_chunk(dcpl, ndim, l), "H5Pset_chunk");
    if (compress) {
      // szip with nearest-neighbour coding, 32 pixels per block (src/h5.cpp:100-104)
      if (H5Zfilter_avail(H5Z_FILTER_SZIP) <= 0)
        throw Error("HDF5: compression uses the szip filter, which this HDF5 library lacks", FST_ERR_ARG);
      h5check(H5Pset_szip(dcpl, H5_SZIP_NN_OPTION_MASK, 32), "H5Pset_szip");
    }
  }
  Hid fspace(H5Screate_simple(ndim, g, nullptr), H5Sclose, "H5Screate_simple");
  Hid dset(H5Dcreate2(group, name.c_str(), h5type<T>(), fspace, H5P_DEFAULT, dcpl, H5P_DEFAULT), H5Dclose,
           "H5Dcreate2(" + name + ")");
  Hid mspace(H5Screate_simple(ndim, l, nullptr), H5Sclose, "H5Screate_simple");
  h5check(H5Sselect_hyperslab(fspace, H5S_SELECT_SET, o, nullptr, l, nullptr), "H5Sselect_hyperslab");
  h5check(H5Dwrite(dset, h5type<T>(), mspace, fspace, H5P_DEFAULT, data), "H5Dwrite(" + name + ")");
}

// h5Writer::writeAttribute (src/h5.cpp:140-198): scalar attributes of a group
void attr_num(hid_t group, const char *name, hid_t type, const void *v) {
  Hid space(H5Screate(H5S_SCALAR), H5Sclose, "H5Screate");
  Hid att(H5Acreate2(group, name, type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, std::string("H5Acreate2 ") + name);
  h5check(H5Awrite(att, type, v), std::string("H5Awrite ") + name);
}

// fixed-length, null-terminated-padding strings of the text's length; an
// empty text is stored as one NUL (HDF5 has no zero-size string type)
void attr_str(hid_t group, const char *name, const std::string &s) {
  Hid space(H5Screate(H5S_SCALAR), H5Sclose, "H5Screate");
  Hid type(H5Tcopy(H5T_C_S1), H5Tclose, "H5Tcopy");
  h5check(H5Tset_size(type, s.empty() ? 1 : s.size()), "H5Tset_size");
  h5check(H5Tset_strpad(type, H5T_STR_NULLTERM), "H5Tset_strpad");
  Hid att(H5Acreate2(group, name, type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, std::string("H5Acreate2 ") + name);
  const std::string buf = s.empty() ? std::string(1, '\0') : s;
  h5check(H5Awrite(att, type, buf.c_str()), std::string("H5Awrite ") + name);
}

// h5Writer::read + checkDataDimensions (src/h5.cpp:200-242, 327-349), one
// process: the whole dataset, whose rank and extents must be the expected ones
template <typename T>
void read_dataset(hid_t file, const std::string &path, int ndim, const int64_t *gext, T *data) {
  Hid dset(H5Dopen2(file, path.c_str(), H5P_DEFAULT), H5Dclose, "H5Dopen2(" + path + ")");
  Hid fspace(H5Dget_space(dset), H5Sclose, "H5Dget_space");
  const int rank = H5Sget_simple_extent_ndims(fspace);
  if (rank != ndim)
    throw Error("Expected " + std::to_string(ndim) + "D restart but found " + std::to_string(rank) + "D restart.",
                FST_ERR_ARG);
  hsize_t dims[3] = {0, 0, 0};
  H5Sget_simple_extent_dims(fspace, dims, nullptr);
  for (int a = 0; a < ndim; ++a) {
    const hsize_t want = (hsize_t)gext[ndim - 1 - a];
    if (dims[a] != want)
      throw Error("Extents of restart file different than expected. (Got " + std::to_string(dims[a]) +
                      " but expected " + std::to_string(want) + " in direction " + std::to_string(a) + ".)",
                  FST_ERR_ARG);
  }
  h5check(H5Dread(dset, h5type<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "H5Dread(" + path + ")");
}

// h5Writer::readAttribute (src/h5.cpp:244-325): attributes of /Properties
template <typename T>
T read_attr(hid_t file, const char *name) {
  Hid group(H5Gopen2(file, "Properties", H5P_DEFAULT), H5Gclose, "H5Gopen2(/Properties)");
  Hid att(H5Aopen(group, name, H5P_DEFAULT), H5Aclose, std::string("H5Aopen ") + name);
  T v{};
  h5check(H5Aread(att, std::is_same<T, int>::value ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE, &v),
          std::string("H5Aread ") + name);
  return v;
}

std::string read_attr_str(hid_t file, const char *name) {
  Hid group(H5Gopen2(file, "Properties", H5P_DEFAULT), H5Gclose, "H5Gopen2(/Properties)");
  Hid att(H5Aopen(group, name, H5P_DEFAULT), H5Aclose, std::string("H5Aopen ") + name);
  Hid ftype(H5Aget_type(att), H5Tclose, "H5Aget_type");
  if (H5Tget_class(ftype) != H5T_STRING || H5Tis_variable_str(ftype) > 0)
    throw Error(std::string("attribute ") + name + " is not a fixed-length string", FST_ERR_ARG);
  Hid mtype(H5Tget_native_type(ftype, H5T_DIR_ASCEND), H5Tclose, "H5Tget_native_type");
  std::vector<char> buf(H5Tget_size(ftype) + 1, '\0');
  h5check(H5Aread(att, mtype, buf.data()), std::string("H5Aread ") + name);
  return std::string(buf.data());
}

// ------------------------------------------------------------------ writer extents

int64_t cells(const fst_config &cf, int a) { return a == 0 ? cf.nci : a == 1 ? cf.ncj : cf.nck; }
double spacing(const fst_config &cf, int a) { return a == 0 ? cf.dx : a == 1 ? cf.dy : cf.dz; }

// The two blockWriter constructors (src/block.cpp:44-98 and 100-280) with the
// subdomain offset 0 and the local cell dimensions equal to the global ones.
fio_extents block_extents(const fst_config &cf, const fio_block &b) {
  fio_extents e{};
  const int nd = cf.ndim;
  for (int a = nd; a < 3; ++a) {  // unused directions: one point
    e.l_ext[a] = e.l_ext_g[a] = e.g_ext[a] = e.g_ext_g[a] = 1;
  }
  if (b.whole) {
    for (int a = 0; a < nd; ++a) {
      const int64_t n = cells(cf, a);
      e.g_ext[a] = n;
      e.g_ext_g[a] = n + 1;
      e.l_start[a] = 0;
      e.l_end[a] = n - 1;
      e.l_end_g[a] = n;  // nodes 0..n
      e.l_ext[a] = n;
      e.l_ext_g[a] = n + 1;
      e.origin[a] = 0.0;
      e.spacing[a] = spacing(cf, a);
    }
    e.write_varx = 0;
    e.chunk = b.chunk;
    return e;
  }
  const std::string view = b.name ? b.name : "";
  for (int a = 0; a < nd; ++a) {
    const int64_t n = cells(cf, a), s = b.stride[a];
    int64_t gs = b.start[a], ge = b.end[a];
    if (gs < 0 || gs > n) throw Error("Starting index of ioview '" + view + "' exceeds domain bounds.", FST_ERR_ARG);
    if (ge < 0 || ge > n) throw Error("Ending index of ioview '" + view + "' exceeds domain bounds.", FST_ERR_ARG);
    if (gs > ge) throw Error("Sarting index of ioview '" + view + "' exceeds ending index.", FST_ERR_ARG);
    if (s < 1 || s > 4) throw Error("Stride of ioview '" + view + "' exceeds 4.", FST_ERR_ARG);
    if (ge > gs && ge - gs < s) throw Error("Stride of ioview '" + view + "' exceeds view dimensions.", FST_ERR_ARG);
    // the global end falls on a stride; one process holds the whole view
    ge -= (ge - gs) % s;
    int64_t gext = (ge - gs) / s + 1;
    int64_t ls = gs, le = ge < n ? ge : n - 1, leg = n - 1;
    le -= (le - ls) % s;
    if (n > ge) {
      // keep the last stride block (and the node after it) inside the domain
      if ((n - 1) - le < s - 1) {
        le -= s;
        ge -= s;
        gext -= 1;
      }
      leg = le + s;
    }
    if (le < ls || gext < 1)
      throw Error("ioview '" + view + "' holds no cells once its end is aligned to the stride", FST_ERR_ARG);
    e.g_ext[a] = gext;
    e.g_ext_g[a] = gext + 1;
    e.l_start[a] = ls;
    e.l_end[a] = le;
    e.l_end_g[a] = leg;
    e.l_ext[a] = (le - ls) / s + 1;
    e.l_ext_g[a] = (leg - ls) / s + 1;
    e.origin[a] = (double)gs * spacing(cf, a);
    e.spacing[a] = spacing(cf, a) * (double)s;
  }
  e.write_varx = 1;
  e.chunk = 0;  // chunking is not supported for strided views (src/block.cpp:278-279)
  return e;
}

// ------------------------------------------------------------------ XDMF

// writeXMF (src/xdmf.cpp:40-178): the same descriptor text, cell-centred
// attributes over a co-rectilinear (cartesian) or curvilinear (grid) mesh.
void write_xmf(const std::string &fname, const std::string &hname, int gridType, double time, int ndim,
               const int64_t *g_ext, const double *origin, const double *dx, const std::vector<std::string> &vn,
               int nvt, bool varx, const std::vector<std::string> &vxn) {
  size_t d[3] = {0, 0, 0};
  for (int a = 0; a < ndim; ++a) d[a] = (size_t)g_ext[ndim - 1 - a];
  std::FILE *x = std::fopen(fname.c_str(), "w");
  if (!x) throw Error("cannot write " + fname, FST_ERR_ARG);
  auto dims = [&](const size_t *e, int n) {
    std::string s;
    for (int a = 0; a < n; ++a) s += (a ? " " : "") + std::to_string(e[a]);
    return s;
  };
  const size_t dn[3] = {d[0] + 1, d[1] + 1, d[2] + 1};
  auto item = [&](const std::string &path, const size_t *e) {
    std::fprintf(x, "       <DataItem Dimensions=\"%s\" NumberType=\"Float\" Precision=\"4\" Format=\"HDF\">\n",
                 dims(e, ndim).c_str());
    std::fprintf(x, "        %s\n", path.c_str());
    std::fprintf(x, "       </DataItem>\n");
  };
  // a vector attribute joins the first three datasets of the list (also in 2D,
  // where the JOIN reads the first two)
  auto vector_attr = [&](const char *label, const std::vector<std::string> &names) {
    std::fprintf(x, "     <Attribute Name=\"%s\" AttributeType=\"Vector\" Center=\"Cell\">\n", label);
    if (ndim == 2)
      std::fprintf(x, "      <DataItem Dimensions=\"%zu %zu 2\" Function=\"JOIN($0,$1)\" ItemType=\"Function\">\n",
                   d[0], d[1]);
    else
      std::fprintf(x,
                   "      <DataItem Dimensions=\"%zu %zu %zu 3\" Function=\"JOIN($0,$1,$2)\" ItemType=\"Function\">\n",
                   d[0], d[1], d[2]);
    for (int c = 0; c < 3; ++c) item(hname + ":/Solution/" + names[c], d);
    std::fprintf(x, "      </DataItem>\n");
    std::fprintf(x, "     </Attribute>\n");
  };
  auto scalar_attr = [&](const std::string &name) {
    std::fprintf(x, "     <Attribute Name=\"%s\" AttributeType=\"Scalar\" Center=\"Cell\">\n", name.c_str());
    item(hname + ":/Solution/" + name, d);
    std::fprintf(x, "     </Attribute>\n");
  };
  std::fprintf(x, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf Version=\"3.0\">\n");
  std::fprintf(x, " <Domain>\n   <Grid Name=\"mesh1\" GridType=\"Uniform\">\n");
  std::fprintf(x, "     <Time Value=\"%e\" />\n", time);
  if (gridType == 0) {
    const char *sfx = ndim == 2 ? "DXDY" : "DXDYDZ";
    std::fprintf(x, "     <Topology TopologyType=\"%dDCORECTMesh\" Dimensions=\"%s\"/>\n", ndim, dims(dn, ndim).c_str());
    std::fprintf(x, "     <Geometry GeometryType=\"ORIGIN_%s\">\n", sfx);
    const char *lbl[2] = {"Origin", "Spacing"};
    const double *val[2] = {origin, dx};
    for (int q = 0; q < 2; ++q) {
      std::fprintf(x, "       <DataItem Name=\"%s\" Dimensions=\"%d\" NumberType=\"Float\" Precision=\"4\" Format=\"XML\">\n",
                   lbl[q], ndim);
      std::fprintf(x, "           ");
      for (int a = ndim - 1; a >= 0; --a) std::fprintf(x, " %f", val[q][a]);
      std::fprintf(x, "\n       </DataItem>\n");
    }
    std::fprintf(x, "     </Geometry>\n");
  } else {
    std::fprintf(x, "     <Topology TopologyType=\"%dDSMesh\" NumberOfElements=\"%s\"/>\n", ndim, dims(dn, ndim).c_str());
    std::fprintf(x, "     <Geometry GeometryType=\"%s\">\n", ndim == 2 ? "X_Y" : "X_Y_Z");
    for (int a = 0; a < ndim; ++a) item(hname + ":/Grid/Dimension" + std::to_string(a), dn);
    std::fprintf(x, "     </Geometry>\n");
  }
  vector_attr("Momentum", vn);
  for (int v = ndim; v < nvt; ++v) scalar_attr(vn[v]);
  if (varx) {
    vector_attr("Velocity", vxn);
    for (size_t v = ndim; v < vxn.size(); ++v) scalar_attr(vxn[v]);
  }
  std::fprintf(x, "   </Grid>\n </Domain>\n</Xdmf>\n");
  if (std::fclose(x) != 0) throw Error("cannot write " + fname, FST_ERR_ARG);
}

// ------------------------------------------------------------------ write / read

template <typename T>
void pack(fst_ctx *ctx, int ndim, int array, int vn, const fio_extents &e, const fio_block &b, bool nodes,
          std::vector<T> &out) {
  int64_t start[3], end[3], stride[3];
  size_t n = 1;
  for (int a = 0; a < 3; ++a) {
    start[a] = e.l_start[a];
    end[a] = nodes ? e.l_end_g[a] : e.l_end[a];
    stride[a] = b.whole || a >= ndim ? 1 : b.stride[a];
    n *= (size_t)(nodes ? e.l_ext_g[a] : e.l_ext[a]);
  }
  out.resize(n);
  const int avg = nodes ? 0 : b.average;
  if (int rc = fst_pack(ctx, array, vn, start, end, stride, avg, std::is_same<T, float>::value, out.data(), n))
    throw Error(std::string("fst_pack: ") + fst_last_error(ctx), rc);
}

template <typename T>
std::string write_block(fst_ctx *ctx, const fst_config &cf, const fio_block &b, const fio_extents &e,
                        const std::vector<std::string> &vn, const std::vector<std::string> &vxn, int tdx, double time) {
  fst_dims d{};
  if (int rc = fst_config_dims(&cf, &d)) throw Error("invalid fst_config", rc);
  if ((int)vn.size() != d.nvt) throw Error("var_names must name every variable of var", FST_ERR_ARG);
  if (e.write_varx && (int)vxn.size() > d.nvx) throw Error("more varx names than varx variables", FST_ERR_ARG);
  if (e.write_varx && vxn.size() < 3) throw Error("a view writes the velocity vector: >= 3 varx names", FST_ERR_ARG);
  const int nd = cf.ndim;
  // src/block.cpp:284-295: "<name>-<step with (int)log10(tend) + 1 digits>"
  const int pad = (int)std::log10((double)(b.tend > 0 ? b.tend : 1)) + 1;
  std::string stem = b.name ? b.name : "";
  if (b.append_step) {
    char idx[32];
    std::snprintf(idx, sizeof(idx), "%0*d", pad, tdx);
    stem += std::string("-") + idx;
  }
  const std::string dir = b.path && *b.path ? b.path : ".";
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  const std::string hname = stem + ".h5", hpath = dir + "/" + hname, xpath = dir + "/" + stem + ".xmf";
  const int64_t zero[3] = {0, 0, 0};
  {
    Hid file(H5Fcreate(hpath.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "H5Fcreate(" + hpath + ")");
    std::vector<T> buf;
    if (cf.grid > 0) {
      Hid g(H5Gcreate2(file, "/Grid", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "H5Gcreate2(/Grid)");
      for (int v = 0; v < nd; ++v) {
        pack(ctx, nd, FST_GRID, v, e, b, true, buf);
        write_dataset(g, "Dimension" + std::to_string(v), nd, e.g_ext_g, e.l_ext_g, zero, buf.data(), e.chunk,
                      b.compress && e.chunk);
      }
    }
    {
      Hid g(H5Gcreate2(file, "/Solution", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "H5Gcreate2(/Solution)");
      for (int v = 0; v < d.nvt; ++v) {
        pack(ctx, nd, FST_VAR, v, e, b, false, buf);
        write_dataset(g, vn[v], nd, e.g_ext, e.l_ext, zero, buf.data(), e.chunk, b.compress && e.chunk);
      }
      if (e.write_varx)
        for (size_t v = 0; v < vxn.size(); ++v) {
          pack(ctx, nd, FST_VARX, (int)v, e, b, false, buf);
          write_dataset(g, vxn[v], nd, e.g_ext, e.l_ext, zero, buf.data(), e.chunk, b.compress && e.chunk);
        }
    }
    Hid g(H5Gcreate2(file, "/Properties", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose, "H5Gcreate2(/Properties)");
    const int version = FIO_RESTART_VERSION;
    attr_num(g, "time_index", H5T_NATIVE_INT, &tdx);
    attr_num(g, "time", H5T_NATIVE_DOUBLE, &time);
    attr_str(g, "title", b.title ? b.title : "");
    attr_str(g, "metadata", b.metadata ? b.metadata : "none");
    attr_str(g, "fiesta_version", "fiesta_amd-1.0");
    attr_str(g, "fiesta_options", "hip-gfx950");
    attr_str(g, "fiesta_btime", __DATE__ " " __TIME__);
    attr_num(g, "restart_file_version", H5T_NATIVE_INT, &version);
  }
  write_xmf(xpath, hname, cf.grid, time, nd, e.g_ext, e.origin, e.spacing, vn, d.nvt, e.write_varx != 0, vxn);
  return hpath;
}

RestartInfo read_restart(fst_ctx *ctx, const fst_config &cf, const std::string &fname,
                         const std::vector<std::string> &vn, bool reset) {
  fst_dims d{};
  if (int rc = fst_config_dims(&cf, &d)) throw Error("invalid fst_config", rc);
  if ((int)vn.size() != d.nvt) throw Error("var_names must name every variable of var", FST_ERR_ARG);
  std::error_code ec;
  if (!std::filesystem::exists(fname, ec)) throw Error("Restart file '" + fname + "' does not exist.", FST_ERR_ARG);
  Hid file(H5Fopen(fname.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, "H5Fopen(" + fname + ")");
  // the version first: nothing reaches the device from a file of another version
  const int version = read_attr<int>(file, "restart_file_version");
  if (version != FIO_RESTART_VERSION)
    throw Error("Cannot read restart file '" + fname + "' with version " + std::to_string(version) + ". Only version " +
                    std::to_string(FIO_RESTART_VERSION) + " files are readable.",
                FST_ERR_ARG);
  const int nd = cf.ndim;
  const int64_t nc[3] = {cf.nci, cf.ncj, nd == 3 ? cf.nck : 1};
  if (cf.grid == 1) {
    const int64_t nn[3] = {nc[0] + 1, nc[1] + 1, nd == 3 ? nc[2] + 1 : 1};
    const size_t plane = (size_t)(nn[0] * nn[1] * nn[2]);
    std::vector<double> grid(plane * 3, 0.0), buf(plane);
    for (int v = 0; v < nd; ++v) {
      read_dataset(file, "/Grid/Dimension" + std::to_string(v), nd, nn, buf.data());
      std::memcpy(grid.data() + plane * v, buf.data(), plane * sizeof(double));
    }
    if (int rc = fst_upload(ctx, FST_GRID, grid.data(), grid.size()))
      throw Error(std::string("fst_upload(grid): ") + fst_last_error(ctx), rc);
  }
  // the interior of a zeroed var (src/reader.cpp:62-79; k from ng in 3D, 0 in 2D)
  const size_t N = (size_t)d.ncells_ghost;
  std::vector<double> var(N * d.nvt, 0.0), buf((size_t)(nc[0] * nc[1] * nc[2]));
  const int ko = nd == 3 ? cf.ng : 0, g = cf.ng;
  for (int v = 0; v < d.nvt; ++v) {
    read_dataset(file, "/Solution/" + vn[v], nd, nc, buf.data());
    double *dst = var.data() + N * v;
    for (int64_t k = 0; k < nc[2]; ++k)
      for (int64_t j = 0; j < nc[1]; ++j)
        for (int64_t i = 0; i < nc[0]; ++i)
          dst[(size_t)(i + g) + (size_t)d.ngi * ((size_t)(j + g) + (size_t)d.ngj * (size_t)(k + ko))] =
              buf[(size_t)(i + nc[0] * (j + nc[1] * k))];
  }
  if (int rc = fst_upload(ctx, FST_VAR, var.data(), var.size()))
    throw Error(std::string("fst_upload(var): ") + fst_last_error(ctx), rc);
  RestartInfo r;
  r.title = read_attr_str(file, "title");
  if (!reset) {
    r.tstart = read_attr<int>(file, "time_index");
    r.time = read_attr<double>(file, "time");
  }
  return r;
}

std::vector<std::string> name_list(const char *const *names, int n) {
  std::vector<std::string> v;
  for (int i = 0; i < n; ++i) {
    if (!names || !names[i]) throw Error("null variable name", FST_ERR_ARG);
    v.emplace_back(names[i]);
  }
  return v;
}

int report(const Error &e, char *err, size_t len) {
  if (err && len) std::snprintf(err, len, "%s", e.what());
  return e.code ? e.code : FST_ERR_ARG;
}

void copy_out(const std::string &s, char *dst, size_t len) {
  if (dst && len) std::snprintf(dst, len, "%s", s.c_str());
}

}  // namespace

// ------------------------------------------------------------------ C++ classes

template <typename T>
blockWriter<T>::blockWriter(const rk_func &f, const IoConfig &io, std::string name, std::string path, bool avg,
                            size_t freq, bool appStep)
    : io_(io), name_(std::move(name)), path_(std::move(path)), freq_(freq) {
  b_.whole = 1;
  b_.average = avg;
  b_.append_step = appStep;
  init(f);
}

template <typename T>
blockWriter<T>::blockWriter(const rk_func &f, const IoConfig &io, std::string name, std::string path, bool avg,
                            size_t freq, std::vector<size_t> start, std::vector<size_t> end,
                            std::vector<size_t> stride, bool appStep)
    : io_(io), name_(std::move(name)), path_(std::move(path)), freq_(freq) {
  const int nd = f.config().ndim;
  if ((int)start.size() < nd || (int)end.size() < nd || (int)stride.size() < nd)
    throw Error("ioview '" + name_ + "': start, limit and stride need one entry per dimension", FST_ERR_ARG);
  b_.whole = 0;
  for (int a = 0; a < nd; ++a) {
    b_.start[a] = (int64_t)start[a];
    b_.end[a] = (int64_t)end[a];
    b_.stride[a] = (int64_t)stride[a];
  }
  b_.average = avg;
  b_.append_step = appStep;
  init(f);
}

template <typename T>
void blockWriter<T>::init(const rk_func &f) {
  b_.name = name_.c_str();
  b_.path = path_.c_str();
  b_.f64 = std::is_same<T, double>::value;
  b_.chunk = io_.chunkable;
  b_.compress = io_.chunkable && io_.compressible;
  b_.tend = io_.tend;
  b_.title = io_.title.c_str();
  b_.metadata = io_.metadata.c_str();
  ext_ = block_extents(f.config(), b_);
}

template <typename T>
std::string blockWriter<T>::write(rk_func &f, int tdx, double time) {
  b_.name = name_.c_str();
  b_.path = path_.c_str();
  b_.title = io_.title.c_str();
  b_.metadata = io_.metadata.c_str();
  return write_block<T>(f.context(), f.config(), b_, ext_, f.varNames, f.varxNames, tdx, time);
}

template class blockWriter<float>;
template class blockWriter<double>;

RestartInfo readRestart(rk_func &f, const std::string &fname, bool reset) {
  return read_restart(f.context(), f.config(), fname, f.varNames, reset);
}

}  // namespace fiesta_hip

// ------------------------------------------------------------------ C-ABI

using namespace fiesta_hip;

extern "C" int fio_block_extents(const fst_config *cf, const fio_block *b, fio_extents *out, char *err,
                                 size_t err_len) {
  if (!cf || !b || !out) return FST_ERR_ARG;
  try {
    fst_dims d{};
    if (int rc = fst_config_dims(cf, &d)) throw Error("invalid fst_config", rc);
    *out = block_extents(*cf, *b);
    return FST_OK;
  } catch (const Error &e) {
    return report(e, err, err_len);
  }
}

extern "C" int fio_write(fst_ctx *ctx, const fst_config *cf, const fio_block *b, const char *const *var_names,
                         int n_var, const char *const *varx_names, int n_varx, int tdx, double time, char *h5_path,
                         size_t h5_path_len, char *err, size_t err_len) {
  if (!ctx || !cf || !b) return FST_ERR_ARG;
  try {
    fst_dims d{};
    if (int rc = fst_config_dims(cf, &d)) throw Error("invalid fst_config", rc);
    const fio_extents e = block_extents(*cf, *b);
    const auto vn = name_list(var_names, n_var), vxn = name_list(varx_names, n_varx);
    const std::string p = b->f64 ? write_block<double>(ctx, *cf, *b, e, vn, vxn, tdx, time)
                                 : write_block<float>(ctx, *cf, *b, e, vn, vxn, tdx, time);
    copy_out(p, h5_path, h5_path_len);
    return FST_OK;
  } catch (const Error &e) {
    return report(e, err, err_len);
  }
}

extern "C" int fio_read_restart(fst_ctx *ctx, const fst_config *cf, const char *fname, const char *const *var_names,
                                int n_var, int reset, int *tdx, double *time, char *title, size_t title_len, char *err,
                                size_t err_len) {
  if (!ctx || !cf || !fname) return FST_ERR_ARG;
  try {
    const RestartInfo r = read_restart(ctx, *cf, fname, name_list(var_names, n_var), reset != 0);
    if (!reset) {
      if (tdx) *tdx = r.tstart;
      if (time) *time = r.time;
    }
    copy_out(r.title, title, title_len);
    return FST_OK;
  } catch (const Error &e) {
    return report(e, err, err_len);
  }
}
