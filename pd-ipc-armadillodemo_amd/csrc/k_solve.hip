// The global step: Chebyshev-accelerated two-step weighted Jacobi on the fixed
// PD system (chebyshevJacobiR2SolverHost, Simulator/PD-IPC/PdIpcGpuFunc.cu:1245-1346).
//
//   M^2 = ((1-w) I + w D^{-1} R)^2         (the "R2" row operators, :899-974)
//   c   = w^2 D^{-1} R D^{-1} b + (2-w) w D^{-1} b          (:1041-1074)
//   x1  = M^2 x0 + c                                        (:1076-1129)
//   x_{k+1} = W_k (M^2 x_k + c - x_{k-1}) + x_{k-1}         (:1185-1242)
//   with r = rho^2, W_1 = 2/(2-r), W_{k+1} = 4/(4 - r W_k)  (:1318-1340)
//
// M^2 is stored in a SELL layout (slices of 64/G rows, G lanes per row, every
// 64-entry column of a slice loaded by one wavefront), so the matrix stream --
// the dominant HBM traffic of the whole step -- is perfectly coalesced.  Each
// matrix row is read once for all three coordinates (the reference re-reads it
// per coordinate).  The operator update rebuilds the M^2 values from the
// 2-path lists once per outer iteration.
//
// FP32-split: the sweeps run in FP32 on the correction delta = x - x0, driven
// by the FP64 residual of the fixed-point map r0 = M^2 x0 + c - x0:
//   delta_1 = r0,  delta_{k+1} = W_k (M^2 delta_k + r0 - delta_{k-1}) + delta_{k-1}
// which is algebraically the same iterate; FP32 only rounds the (small)
// correction, never the absolute coordinates.
#include "device.hpp"

namespace pdipc {
namespace {

constexpr int kBlock = 256;

// M^2_ij = (w^2 sum_k R1_ik R1_kj / D_k + 2 w (1 - w) R1_ij) / D_i + (1 - w)^2 [i == j]
// (updateR1/R2/R2RowOperatorsKernel, PdIpcGpuFunc.cu:899-974), recomputed only
// for rows whose value can have changed: a row depends on D_i and on D_k of its
// R1 neighbours k, so rows with no changed diagonal within one R1 hop keep
// their (bit-identical) values.
__global__ void k_dirty_rows(int n, const int* __restrict__ off, const int* __restrict__ col,
                             const uint8_t* __restrict__ changed, uint8_t* __restrict__ dirty) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t d = changed[i];
  for (int q = off[i]; q < off[i + 1] && !d; q++) d |= changed[col[q]];
  dirty[i] = d;
}

__global__ void k_operators(int64_t nE, double w, const int* __restrict__ rowFlag, const uint8_t* __restrict__ dirty,
                            const double* __restrict__ r1, const int64_t* __restrict__ pOff,
                            const int* __restrict__ pK, const double* __restrict__ pA,
                            const double* __restrict__ D, double* __restrict__ val, float* __restrict__ valF) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  int rf = rowFlag[e];
  int row = rf & 0x7fffffff;
  if (!dirty[row]) return;
  double s = 0.0;
  for (int64_t p = pOff[e]; p < pOff[e + 1]; p++) s += pA[p] / D[pK[p]];
  double v = (w * w * s + 2.0 * (1.0 - w) * w * r1[e]) / D[row];
  if (rf < 0) v += (1.0 - w) * (1.0 - w);
  val[e] = v;
  valF[e] = (float)v;
}

template <int G, class V>
__device__ inline void group_reduce(V acc[3]) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; k++) acc[k] += __shfl_xor(acc[k], o, kWave);
}

struct FirstTermArgs {  // c = (w^2 sum_j R1_ij b_j / D_j + (2 - w) w b_i) / D_i   (PdIpcGpuFunc.cu:1041-1074)
  const int* off;
  const int* col;
  const double* val;
  const double* D;
  const double* b;
  double w;
};

// MODE 0: x1 = M^2 x0 + c, computing c on the fly (the first term is fused
// into the first sweep: the G lanes of a row split its R1 entries) and
// storing it for the later sweeps;  MODE 1: Chebyshev step;  MODE 2:
// FP32-split start: c as in MODE 0, then r0 = M^2 x0 + c - x0, delta_0 = 0,
// delta_1 = r0.
template <int G, int MODE>
__global__ void __launch_bounds__(kBlock) k_sweep64(int n, int nSlices, int ld, double W,
                                                    const int64_t* __restrict__ sliceOff,
                                                    const int* __restrict__ sliceW, const int* __restrict__ col,
                                                    const double* __restrict__ val, const double* __restrict__ xcur,
                                                    const double* __restrict__ xpre, double* __restrict__ cvec,
                                                    double* __restrict__ xnext, float* __restrict__ dpreF,
                                                    float* __restrict__ dnextF, float* __restrict__ r0F,
                                                    FirstTermArgs ft) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int slice = gt / kWave, lane = gt % kWave;
  if (slice >= nSlices) return;  // whole wavefront
  constexpr int S = kWave / G;
  int row = slice * S + lane / G;
  int sub = lane % G;
  int64_t off = sliceOff[slice] + lane;
  int width = sliceW[slice];
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 4
  for (int c = 0; c < width; c++) {
    int64_t e = off + (int64_t)c * kWave;
    int j = col[e];
    double v = val[e];
    acc[0] += v * xcur[j];
    acc[1] += v * xcur[ld + j];
    acc[2] += v * xcur[2 * ld + j];
  }
  group_reduce<G>(acc);
  double cr[3] = {0.0, 0.0, 0.0};
  if (MODE != 1) {
    if (row < n)
      for (int q = ft.off[row] + sub; q < ft.off[row + 1]; q += G) {
        int j = ft.col[q];
        double f = ft.val[q] / ft.D[j];
#pragma unroll
        for (int k = 0; k < 3; k++) cr[k] += f * ft.b[k * ld + j];
      }
    group_reduce<G>(cr);
  }
  if (sub != 0 || row >= n) return;
  double di = MODE != 1 ? ft.D[row] : 1.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int idx = k * ld + row;
    double c;
    if (MODE != 1) {
      c = (ft.w * ft.w * cr[k] + (2.0 - ft.w) * ft.w * ft.b[idx]) / di;
      cvec[idx] = c;
    } else {
      c = cvec[idx];
    }
    double y = acc[k] + c;
    if (MODE == 0) {
      xnext[idx] = y;
    } else if (MODE == 1) {
      double p = xpre[idx];
      xnext[idx] = W * (y - p) + p;
    } else {
      float r = (float)(y - xcur[idx]);
      r0F[idx] = r;
      dpreF[idx] = 0.0f;
      dnextF[idx] = r;
    }
  }
}

// FP32 Chebyshev step on the correction; LAST also writes x = x0 + delta
template <int G, int LAST>
__global__ void __launch_bounds__(kBlock) k_sweep32(int n, int nSlices, int ld, float W,
                                                    const int64_t* __restrict__ sliceOff,
                                                    const int* __restrict__ sliceW, const int* __restrict__ col,
                                                    const float* __restrict__ val, const float* __restrict__ dcur,
                                                    const float* __restrict__ dpre, const float* __restrict__ r0,
                                                    float* __restrict__ dnext, double* __restrict__ X) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int slice = gt / kWave, lane = gt % kWave;
  if (slice >= nSlices) return;
  constexpr int S = kWave / G;
  int row = slice * S + lane / G;
  int64_t off = sliceOff[slice] + lane;
  int width = sliceW[slice];
  float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll 4
  for (int c = 0; c < width; c++) {
    int64_t e = off + (int64_t)c * kWave;
    int j = col[e];
    float v = val[e];
    acc[0] += v * dcur[j];
    acc[1] += v * dcur[ld + j];
    acc[2] += v * dcur[2 * ld + j];
  }
  group_reduce<G>(acc);
  if ((lane % G) != 0 || row >= n) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int idx = k * ld + row;
    float p = dpre[idx];
    float d = W * (acc[k] + r0[idx] - p) + p;
    if (LAST) X[idx] = X[idx] + (double)d;
    else dnext[idx] = d;
  }
}

__global__ void k_apply_delta(int n, int ld, const float* __restrict__ d, double* __restrict__ X) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 3; k++) X[k * ld + i] += (double)d[k * ld + i];
}

template <int G>
void solve_impl(const DeviceMesh& m, double w, double rho, int iters, int precision, hipStream_t s) {
  const int grid = (m.nSlices * kWave + kBlock - 1) / kBlock;
  const FirstTermArgs ft{m.r1Off, m.r1Col, m.r1Val, m.sysDiag, m.sysRhs, w};
  const FirstTermArgs none{nullptr, nullptr, nullptr, nullptr, nullptr, 0.0};
  double r = rho * rho;
  double W = 2.0 / (2.0 - r);
  if (precision == 0) {
    double *pre = m.Pre, *cur = m.X, *next = m.Next;
    hipLaunchKernelGGL((k_sweep64<G, 0>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, 1.0, m.sliceOff,
                       m.sliceW, m.m2Col, m.m2Val, cur, pre, m.firstTerm, next, nullptr, nullptr, nullptr, ft);
    double* t = pre; pre = cur; cur = next; next = t;
    for (int it = 1; it < iters; it++) {
      hipLaunchKernelGGL((k_sweep64<G, 1>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, W, m.sliceOff,
                         m.sliceW, m.m2Col, m.m2Val, cur, pre, m.firstTerm, next, nullptr, nullptr, nullptr, none);
      W = 4.0 / (4.0 - r * W);
      t = pre; pre = cur; cur = next; next = t;
    }
    if (cur != m.X) launch_copy(m.X, cur, 3 * (int64_t)m.ld, s);
  } else {
    float *pre = m.dPre, *cur = m.dCur, *next = m.dNext;
    // delta_0 = 0 -> pre, delta_1 = r0 -> cur
    hipLaunchKernelGGL((k_sweep64<G, 2>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, 1.0, m.sliceOff,
                       m.sliceW, m.m2Col, m.m2Val, m.X, nullptr, m.firstTerm, nullptr, pre, cur, m.r0, ft);
    if (iters <= 1) {
      hipLaunchKernelGGL(k_apply_delta, dim3((m.nSim + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m.nSim, m.ld, cur,
                         m.X);
      return;
    }
    for (int it = 1; it < iters; it++) {
      if (it == iters - 1)
        hipLaunchKernelGGL((k_sweep32<G, 1>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, (float)W,
                           m.sliceOff, m.sliceW, m.m2Col, m.m2ValF, cur, pre, m.r0, next, m.X);
      else
        hipLaunchKernelGGL((k_sweep32<G, 0>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, (float)W,
                           m.sliceOff, m.sliceW, m.m2Col, m.m2ValF, cur, pre, m.r0, next, m.X);
      W = 4.0 / (4.0 - r * W);
      float* t = pre; pre = cur; cur = next; next = t;
    }
  }
}

template <int G>
void one_sweep_impl(const DeviceMesh& m, int precision, hipStream_t s) {
  const int grid = (m.nSlices * kWave + kBlock - 1) / kBlock;
  const double W = 1.9;
  if (precision == 0)
    hipLaunchKernelGGL((k_sweep64<G, 1>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, W, m.sliceOff,
                       m.sliceW, m.m2Col, m.m2Val, m.X, m.Pre, m.firstTerm, m.Next, nullptr, nullptr, nullptr,
                       FirstTermArgs{nullptr, nullptr, nullptr, nullptr, nullptr, 0.0});
  else
    hipLaunchKernelGGL((k_sweep32<G, 0>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, (float)W,
                       m.sliceOff, m.sliceW, m.m2Col, m.m2ValF, m.dCur, m.dPre, m.r0, m.dNext, m.X);
}

}  // namespace

void launch_one_sweep(const DeviceMesh& m, int precision, hipStream_t s) {
  switch (m.sellLanes) {
    case 1: one_sweep_impl<1>(m, precision, s); break;
    case 2: one_sweep_impl<2>(m, precision, s); break;
    case 8: one_sweep_impl<8>(m, precision, s); break;
    default: one_sweep_impl<4>(m, precision, s); break;
  }
}

void launch_operators(const DeviceMesh& m, double omega, int precision, hipStream_t s) {
  (void)precision;  // both the FP64 and the FP32 copies are kept current
  int64_t nE = m.sellEntries;
  hipLaunchKernelGGL(k_dirty_rows, dim3((m.nSim + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m.nSim, m.r1Off, m.r1Col,
                     m.dChanged, m.dDirty);
  hipLaunchKernelGGL(k_operators, dim3((unsigned)((nE + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, nE, omega,
                     m.m2RowFlag, m.dDirty, m.m2R1, m.p2Off, m.p2K, m.p2A, m.sysDiag, m.m2Val, m.m2ValF);
}

void launch_solve(const DeviceMesh& m, double omega, double rho, int iters, int precision, hipStream_t s) {
  switch (m.sellLanes) {
    case 1: solve_impl<1>(m, omega, rho, iters, precision, s); break;
    case 2: solve_impl<2>(m, omega, rho, iters, precision, s); break;
    case 8: solve_impl<8>(m, omega, rho, iters, precision, s); break;
    default: solve_impl<4>(m, omega, rho, iters, precision, s); break;
  }
}

}  // namespace pdipc
