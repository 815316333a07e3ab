// fst_step_static: the fused time step with no device allocation.  The
// stage state, primitive planes, partial right-hand side and stage maxima
// that a context allocates live here in one static device array per GPU.
#include <cstdio>
#include <mutex>
#include <string>

#include "fst_common.hpp"
#include "fst_fused.hpp"
#include "fst_kernels.hpp"

using namespace fst;

namespace {

constexpr size_t kScratchBytes = (size_t)FST_STATIC_SCRATCH_MB << 20;
constexpr int kMaxDevices = 64;

// Device memory of every fst_step_static call on a GPU.  Host code never
// passes this name to a kernel or a copy: it stands for a host-side shadow.
// The device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {};

// The scratch address on the current device (looked up on its first call).
int scratch_on_device(unsigned char **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FST_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "fst_step_static: hipGetSymbolAddress: %s\n", hipGetErrorString(e));
      return FST_ERR_HIP;
    }
    g_scratch[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch[dev];
  return FST_OK;
}

// Zero fill (a kernel, not hipMemsetAsync, so that a captured graph replays it).
__global__ void k_zero_words(unsigned long long *__restrict__ p, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    p[t] = 0ull;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Scratch layout: U* (nvt planes), U2 (nvt planes, SSP-RK3 only), two
// primitive sets (6 + ns planes each), the partial RHS (nv planes), two
// Scalars records.
struct Layout {
  size_t stage, stage2, prim, dvar, sc, total;
};

Layout layout(const Geo &g, int ns, bool rk3) {
  const size_t plane = (size_t)g.N * sizeof(double);
  Layout L;
  L.stage = 0;
  L.stage2 = align256(L.stage + plane * g.nvt);
  L.prim = rk3 ? align256(L.stage2 + plane * g.nvt) : L.stage2;
  L.dvar = align256(L.prim + plane * 2 * (6 + ns));
  L.sc = align256(L.dvar + plane * g.nv);
  L.total = align256(L.sc + 2 * sizeof(Scalars));
  return L;
}

}  // namespace

extern "C" int fst_step_static(const fst_config *cfg, double *var, int nsteps, void *stream) {
  if (!cfg || !var || nsteps < 0) return FST_ERR_ARG;
  fst_dims d;
  std::string why;
  if (module_check(*cfg, &d, &why) != FST_OK) {
    fprintf(stderr, "fst_step_static: %s\n", why.c_str());
    return FST_ERR_ARG;
  }
  fst_config cf;
  Geo g;
  Phys ph;
  module_setup(*cfg, d, &cf, &g, &ph);
  if (cf.learned_c || !fused_supported(cf)) {
    fprintf(stderr, "fst_step_static: runs the fused 3D cartesian path of the plain forms only (WENO5, ns <= 3, "
                    "no learned C-indicator)\n");
    return FST_ERR_ARG;
  }
  const Layout L = layout(g, cf.ns, rk_stages(cf) == 3);
  if (L.total > kScratchBytes) {
    fprintf(stderr, "fst_step_static: this grid needs %zu bytes of scratch; the static buffer holds %zu (%d MiB)\n",
            L.total, kScratchBytes, (int)FST_STATIC_SCRATCH_MB);
    return FST_ERR_ARG;
  }
  unsigned char *base = nullptr;
  if (int rc = scratch_on_device(&base)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t words = L.total / sizeof(unsigned long long);
  const unsigned blocks = (unsigned)std::min<size_t>((words + 255) / 256, 4096);
  k_zero_words<<<blocks, 256, 0, s>>>(reinterpret_cast<unsigned long long *>(base), words);
  FusedBufs fb;
  fb.stage = reinterpret_cast<double *>(base + L.stage);
  fb.stage2 = rk_stages(cf) == 3 ? reinterpret_cast<double *>(base + L.stage2) : nullptr;
  fb.prim = reinterpret_cast<double *>(base + L.prim);
  fb.sc = reinterpret_cast<Scalars *>(base + L.sc);
  Bufs b;
  b.var = var;
  b.dvar = reinterpret_cast<double *>(base + L.dvar);
  std::string err;
  const int rc = fused_steps(g, ph, cf, b, fb, nullptr, nsteps, s, &err);
  if (rc) {
    fprintf(stderr, "fst_step_static: %s\n", err.c_str());
    return rc;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "fst_step_static: launch failed: %s\n", hipGetErrorString(e));
    return FST_ERR_HIP;
  }
  return FST_OK;
}
