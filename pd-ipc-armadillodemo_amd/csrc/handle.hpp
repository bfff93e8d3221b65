// The library handle shared by the step driver (pdipc_api.hip) and the adjoint
// (adjoint.hip), plus the host helpers both use.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pdipc.h"
#include "device.hpp"
#include "host_setup.hpp"

namespace pdipc {

int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return ::pdipc::fail(PDIPC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct AdjointState;                 // adjoint.hip
void adjoint_free(AdjointState* a);  // releases its device memory

}  // namespace pdipc

using pdipc::fail;

struct pdipc_handle {
  using HostSetup = pdipc::HostSetup;
  using SellLayout = pdipc::SellLayout;
  using DeviceMesh = pdipc::DeviceMesh;
  using ModelInput = pdipc::ModelInput;
  pdipc_params prm;
  std::vector<ModelInput> sims, statics;
  bool hostReady = false, deviceReady = false;
  HostSetup hs;
  SellLayout sell;
  DeviceMesh m;
  hipStream_t stream = nullptr;
  int device = 0;  // the HIP device current at pdipc_init
  // local/global loop captured once per (collision flag, precision) and replayed
  hipGraphExec_t lgGraph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  std::vector<void*> allocs;
  // host mirrors
  void* scalHost = nullptr;  // pinned: residual, toiBits, counters[10]
  void* scalDev = nullptr;
  std::vector<double> stage;
  pdipc_counters cnt;
  int lastRecs = 0, lastPairs = 0;
  double* trackToi = nullptr;  // per-record TOIs when track_toi_argmin is on
  int trackCap = 0;

  template <class T>
  int alloc(T** p, size_t n) {
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) return fail(PDIPC_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(*p);
    return PDIPC_OK;
  }
  template <class T>
  int upload(T** p, const std::vector<T>& v) {
    int rc = alloc(p, v.size());
    if (rc) return rc;
    if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return PDIPC_OK;
  }
  void release(void* p) {
    if (!p) return;
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) allocs.erase(it);
    (void)hipFree(p);
  }
  pdipc::AdjointState* adj = nullptr;
  ~pdipc_handle() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();  // the executable graphs go before the buffers their nodes point to
    delete m.timer;
    pdipc::adjoint_free(adj);
    for (void* p : allocs) (void)hipFree(p);
    if (scalHost) (void)hipHostFree(scalHost);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void drop_graphs() {
    for (auto& row : lgGraph)
      for (auto& g : row)
        if (g) {
          (void)hipGraphExecDestroy(g);
          g = nullptr;
        }
  }
  double* hResidual() { return (double*)scalHost; }
  unsigned long long* hToi() { return (unsigned long long*)((char*)scalHost + 8); }
  int* hCounters() { return (int*)((char*)scalHost + 16); }
};


namespace pdipc {
int put_soa(pdipc_handle* h, double* dst, const double* aos, int npts);
int put_soa_prefix(pdipc_handle* h, double* dst, const double* aos, int npts);
int get_soa(pdipc_handle* h, const double* src, double* aos, int npts);
int check_launch(pdipc_handle* h);
int read_scalars(pdipc_handle* h);
int broad_phase(pdipc_handle* h, const double* x, const double* dx);
int collision(pdipc_handle* h);
int require_device(pdipc_handle* h);
int advance(pdipc_handle* h);
int adjoint_refresh_values(pdipc_handle* h);  // adjoint.hip: after a stiffness change
int noalloc_setup(int device);                 // noalloc.hip: with `device` current
}  // namespace pdipc
