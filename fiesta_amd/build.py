"""Build the HIP library (gfx950) in-tree, and the oracle for tests.

`python -m fiesta_amd.build` compiles every HIP source under fiesta_amd/csrc
into fiesta_amd/libfiesta_hip.so.  FP64 parity needs -ffp-contract=off (no
FMA contraction: the reference CPU build has none); packed-FP32 instructions
are switched off for all device code.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fiesta_amd")
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfiesta_hip.so")
ARCH = os.environ.get("FST_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build_hip(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = []
    cmds = []
    for f in srcs:
        o = os.path.join(OBJ, f[:-4] + ".o")
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


HOST_DIR = os.path.join(PKG, "host")
RUN_BIN = os.path.join(PKG, "fiesta_run")
MAIN_BIN = os.path.join(PKG, "fiesta")  # runs a Lua input deck (the reference's main program)
HOST_LIB = os.path.join(PKG, "libfiesta_host.so")
# HDF5 (serial, C API) for the solution/restart files.  The library is linked
# through a directory holding only a link to it, so the run-time search path
# of the host code names no directory of unrelated shared libraries.
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")
HDF5_SONAME = "libhdf5.so.103"
HDF5_LINK_DIR = os.path.join(PKG, "_hdf5")


def _hdf5_link():
    inc = os.path.join(HDF5_DIR, "include")
    lib = os.path.join(HDF5_DIR, "lib", HDF5_SONAME)
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(lib)):
        raise RuntimeError("HDF5 not found under %s (set HDF5_DIR): the writers need hdf5.h and %s" %
                           (HDF5_DIR, HDF5_SONAME))
    os.makedirs(HDF5_LINK_DIR, exist_ok=True)
    link = os.path.join(HDF5_LINK_DIR, HDF5_SONAME)
    if os.path.lexists(link) and os.readlink(link) != lib:
        os.remove(link)
    if not os.path.lexists(link):
        os.symlink(lib, link)
    return inc, link


def build_host():
    """The host rk_func layer, the HDF5/XDMF writers and restart reader and
    the Lua deck reader (C++, no HIP code) as libfiesta_host.so, and the
    fiesta_run and fiesta drivers;
    linked against libfiesta_hip.so (run-time path: the package directory)."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    inc, h5 = _hdf5_link()
    common = ["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-I" + os.path.join(ROOT, "include"), "-I" + HOST_DIR]
    rpath = ["-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/_hdf5", "-Wl,-rpath-link," + os.path.join(rocm, "lib"),
             "-Wl,-rpath," + os.path.join(rocm, "lib")]
    srcs = [os.path.join(HOST_DIR, f) for f in ("fiesta_host.cpp", "fiesta_io.cpp", "fiesta_lua.cpp", "fiesta_deck.cpp")]
    tmp = HOST_LIB + ".tmp"
    _run(common + ["-fPIC", "-shared", "-isystem", inc] + srcs + [h5, "-L" + PKG, "-lfiesta_hip"] + rpath +
         ["-o", tmp])
    os.replace(tmp, HOST_LIB)
    for src, dst in (("fiesta_run.cpp", RUN_BIN), ("fiesta_main.cpp", MAIN_BIN)):
        tmp = dst + ".tmp"
        _run(common + [os.path.join(HOST_DIR, src), "-L" + PKG, "-lfiesta_host", "-lfiesta_hip"] + rpath +
             ["-o", tmp])
        os.replace(tmp, dst)
    return RUN_BIN


def build_oracle():
    d = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "-C", d, "liborc.so"])
    return os.path.join(d, "liborc.so")


def build_all(verbose=False):
    lib = build_hip(verbose)
    run = build_host()
    orc = build_oracle()
    return lib, run, orc


if __name__ == "__main__":
    print(build_all(verbose="-v" in sys.argv))
