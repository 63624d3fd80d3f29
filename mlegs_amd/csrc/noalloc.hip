// An entry point that allocates no device memory: svv_filter with its
// reduction scratch in a static device array.
//
// mlx_svv_filter keeps the partial sums of its reduction and the per-field
// results in buffers the context allocated.  mlx_svv_filter_noalloc runs the
// same three kernels with that scratch in `scratch` below, a file-scope
// device array of fixed size that every GPU the library runs on holds until
// the process ends.  Host code never passes the array's name to a kernel
// (on the host it is only a stand-in symbol): noalloc_setup looks up each
// device's own address of it once, with that device current, and the entry
// point uses that address.
//
// Every call on a GPU shares that GPU's array, which keeps the last call's
// contents: the partial-sum slots are all written by the reduction before
// they are read, and the gain slot is written from the caller's value at
// the start of each call.  Calls that overlap on one GPU (two host threads,
// or two streams) race on the array.
#include <cstring>
#include <mutex>

#include "mlx_internal.h"

namespace {

// partial sums: 2 per workgroup of the reduction, at most kSvvParts
// workgroups (one field); results: 3; gain: 1 (rounded up to 8)
constexpr size_t kScratchParts = 2 * (size_t)mlx::kSvvParts;
constexpr size_t kScratchOut = kScratchParts;
constexpr size_t kScratchGain = kScratchOut + 4;
constexpr size_t kScratchDoubles = kScratchParts + 8;

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double *g_ptr[kMaxDevices] = {nullptr};

__global__ void k_set_gain(double *g, double v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) g[0] = v;
}

}  // namespace

namespace mlx {

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) MLX_THROW(MLX_E_ARG, "noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_ptr[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    MLX_THROW(MLX_E_HIP, std::string("noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_ptr[device] = static_cast<double *>(p);
}

}  // namespace mlx

using namespace mlx;

extern "C" int mlx_svv_filter_noalloc(mlx_ctx *c, mlx_field *f, double *gain) {
  if (!c) return MLX_E_ARG;
  try {
    if (!gain) MLX_THROW(MLX_E_ARG, "svv_filter_noalloc: gain is a null pointer");
    if (!c->p.is_svv) return MLX_OK;
    if (!c->kit.ready) MLX_THROW(MLX_E_ARG, "transformation kit is not initialized");
    if (std::strncmp(f->space, "FFF", 3) != 0) MLX_THROW(MLX_E_LAYOUT, "svv_filter: scalar must be in FFF space");
    if (c->device < 0 || c->device >= kMaxDevices || !g_ptr[c->device])
      MLX_THROW(MLX_E_ARG, "svv_filter_noalloc: the device's scratch array was not set up");
    // the reduction takes up to kSvvParts workgroups: one per slice and
    // column range, at least one per slice
    if ((size_t)c->kit.nzdim > (size_t)kSvvParts)
      MLX_THROW(MLX_E_ARG, "svv_filter_noalloc: nz above 4096 exceeds the static scratch (4096 partial sums)");
    double *base = g_ptr[c->device];
    double *g = base + kScratchGain;
    hipLaunchKernelGGL(k_set_gain, dim3(1), dim3(64), 0, c->stream, g, *gain);
    k_svv(c, f->e, g, nullptr, nullptr, nullptr, -1, 1, 0, base, base + kScratchOut);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->hsmall, g, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *gain = c->hsmall[0];
    return MLX_OK;
  } catch (const MlxError &e) {
    c->err = e.msg;
    return e.code;
  }
}
