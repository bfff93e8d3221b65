// pdipc_toi_noalloc: the minimum ACCD time of impact over the current
// candidate records (what pdipc_debug_toi returns) for callers that cannot
// allocate device memory -- code under HIP graph capture, tight serving loops.
// Positions and the result stay on the device, the work is only enqueued, and
// the per-workgroup minima live in `scratch`, a file-scope device array each
// GPU holds for the life of the process.  Its device address comes from
// hipGetSymbolAddress once per device (pdipc_init, with the handle's device
// current); host code passes that address, never the array's name.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "handle.hpp"

namespace pdipc {

namespace {

constexpr int kRecsPerBlock = 256;     // k_ccd_blockmin's workgroup size
constexpr int kMaxBlocks = 65536;      // 512 KiB of minima: at most 16,777,216 candidate records
constexpr int kMaxDevices = 64;
constexpr int kReduceThreads = 1024;

static __device__ __align__(256) double scratch[kMaxBlocks];

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// one workgroup: min over the n block minima, then the reference's safety
// scale (getImpactOfTimeHost, PdIpcGpuFunc.cu:2698-2699): t < 1 -> t * scale
__global__ void __launch_bounds__(kReduceThreads) k_toi_reduce(const double* __restrict__ blockMin, int n,
                                                               double scale, double* __restrict__ out) {
  __shared__ double part[kReduceThreads];
  double m = 1.0;
  for (int i = threadIdx.x; i < n; i += kReduceThreads) {
    double v = blockMin[i];
    m = v < m ? v : m;
  }
  part[threadIdx.x] = m;
  __syncthreads();
  for (int o = kReduceThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      double v = part[threadIdx.x + o];
      part[threadIdx.x] = v < part[threadIdx.x] ? v : part[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0] < 1.0 ? part[0] * scale : 1.0;
}

}  // namespace

// With `device` current: looks up this device's copy of `scratch` once.
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(PDIPC_E_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return PDIPC_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(PDIPC_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = (double*)p;
  return PDIPC_OK;
}

}  // namespace pdipc

using namespace pdipc;

int pdipc_toi_noalloc(pdipc_handle* h, const double* x0, const double* x1, double* toi, void* stream) {
  int rc = require_device(h);
  if (rc) return rc;
  if (!x0 || !x1 || !toi) return fail(PDIPC_E_ARG, "null argument");
  const int nRecs = h->lastRecs;
  const int nBlocks = (nRecs + kRecsPerBlock - 1) / kRecsPerBlock;
  if (nBlocks > kMaxBlocks)
    return fail(PDIPC_E_CAPACITY, "pdipc_toi_noalloc: " + std::to_string(nRecs) +
                                      " candidate records exceed its limit of 16777216 (use pdipc_debug_toi)");
  double* scr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    scr = g_scratch[h->device];
  }
  if (!scr) return fail(PDIPC_E_ARG, "pdipc_toi_noalloc: no scratch address for the handle's device");
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  const pdipc_params& p = h->prm;
  launch_ccd_blockmin(h->m, nRecs, x0, x1, p.ccd_filter, p.ccd_iters, p.ccd_step, p.ccd_gap, p.touch_policy, scr, s);
  hipLaunchKernelGGL(k_toi_reduce, dim3(1), dim3(kReduceThreads), 0, s, scr, nBlocks, p.toi_scale, toi);
  return check_launch(h);
}
