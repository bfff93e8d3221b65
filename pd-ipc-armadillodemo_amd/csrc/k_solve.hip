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

// The recorded solve of the adjoint's forward pass: the FP64 recurrence with
// every iterate x_1 .. x_iters written to its own buffer (x_0 = X is kept).
template <int G>
void solve_taped_impl(const DeviceMesh& m, double w, double rho, int iters, double* const* outs, hipStream_t s) {
  const int grid = (m.nSlices * kWave + kBlock - 1) / kBlock;
  const FirstTermArgs ft{m.r1Off, m.r1Col, m.r1Val, m.sysDiag, m.sysRhs, w};
  const FirstTermArgs none{nullptr, nullptr, nullptr, nullptr, nullptr, 0.0};
  double r = rho * rho;
  double W = 2.0 / (2.0 - r);
  hipLaunchKernelGGL((k_sweep64<G, 0>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, 1.0, m.sliceOff,
                     m.sliceW, m.m2Col, m.m2Val, m.X, m.X, m.firstTerm, outs[1], nullptr, nullptr, nullptr, ft);
  for (int it = 1; it < iters; it++) {
    const double* pre = it == 1 ? m.X : outs[it - 1];
    hipLaunchKernelGGL((k_sweep64<G, 1>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, W, m.sliceOff,
                       m.sliceW, m.m2Col, m.m2Val, outs[it], pre, m.firstTerm, outs[it + 1], nullptr, nullptr, nullptr,
                       none);
    W = 4.0 / (4.0 - r * W);
  }
  // the sweeps write the simulated rows only: copy those back (the static rows of X stay)
  launch_axpby(m, 1.0, outs[iters], 0.0, nullptr, m.X, s);
}

// Reverse sweep of the Chebyshev recurrence in the scaled adjoint u = D^{-1} xbar
// (M^T = D M D^{-1} because R1 is symmetric, so M^T xbar = D M u):
//   u_out = a M u_in + b u_oth,   Gbar_e += a u_in[row] . x_k[col],   uc[row] += a u_in[row]
// Gbar accumulates D^{-1}-scaled adjoints of the M^2 entries, uc the scaled
// adjoint of the first term c.
template <int G>
__global__ void __launch_bounds__(kBlock) k_adj_sweep(int n, int nSlices, int ld, double a, double b,
                                                      const int64_t* __restrict__ sliceOff,
                                                      const int* __restrict__ sliceW, const int* __restrict__ col,
                                                      const double* __restrict__ val, const double* __restrict__ uin,
                                                      const double* __restrict__ uoth, const double* __restrict__ xk,
                                                      double* __restrict__ uout, double* __restrict__ Gbar,
                                                      double* __restrict__ uc) {
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int slice = gt / kWave, lane = gt % kWave;
  if (slice >= nSlices) return;
  constexpr int S = kWave / G;
  int row = slice * S + lane / G;
  int sub = lane % G;
  int64_t off = sliceOff[slice] + lane;
  int width = sliceW[slice];
  double ui[3] = {0.0, 0.0, 0.0};
  if (row < n) { ui[0] = uin[row]; ui[1] = uin[ld + row]; ui[2] = uin[2 * ld + row]; }
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll 4
  for (int c = 0; c < width; c++) {
    int64_t e = off + (int64_t)c * kWave;
    int j = col[e];
    double v = val[e];
    double u0 = uin[j], u1 = uin[ld + j], u2 = uin[2 * ld + j];
    acc[0] += v * u0;
    acc[1] += v * u1;
    acc[2] += v * u2;
    Gbar[e] += a * (ui[0] * xk[j] + ui[1] * xk[ld + j] + ui[2] * xk[2 * ld + j]);
  }
  group_reduce<G>(acc);
  if (sub != 0 || row >= n) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int idx = k * ld + row;
    uout[idx] = a * acc[k] + (b != 0.0 ? b * uoth[idx] : 0.0);
    uc[idx] += a * ui[k];
  }
}

template <int G>
void adj_sweep_impl(const DeviceMesh& m, double a, double b, const double* uin, const double* uoth, const double* xk,
                    double* uout, double* Gbar, double* uc, hipStream_t s) {
  const int grid = (m.nSlices * kWave + kBlock - 1) / kBlock;
  hipLaunchKernelGGL((k_adj_sweep<G>), dim3(grid), dim3(kBlock), 0, s, m.nSim, m.nSlices, m.ld, a, b, m.sliceOff,
                     m.sliceW, m.m2Col, m.m2Val, uin, uoth, xk, uout, Gbar, uc);
}

// Adjoint of the first term c_i = (w^2 sum_j R1_ij b_j / D_j + (2 - w) w b_i) / D_i
// given its scaled adjoint uc = D^{-1} cbar:  bbar = w^2 D^{-1} R1 uc + (2 - w) w uc,
// and the D and R1 parts of the derivative accumulated into Dbar and R1bar.
__global__ void k_adj_first(int n, int ld, double w, const int* __restrict__ off, const int* __restrict__ col,
                            const double* __restrict__ r1, const double* __restrict__ D, const double* __restrict__ b,
                            const double* __restrict__ uc, double* __restrict__ bbar, double* __restrict__ Dbar,
                            double* __restrict__ R1bar) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ui[3] = {uc[i], uc[ld + i], uc[2 * ld + i]};
  double s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
  const double w2 = w * w;
  for (int q = off[i]; q < off[i + 1]; q++) {
    int j = col[q];
    double a = r1[q], dj = D[j];
    double bj[3] = {b[j] / dj, b[ld + j] / dj, b[2 * ld + j] / dj};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s1[k] += a * uc[k * ld + j];
      s2[k] += a * bj[k];
    }
    R1bar[q] += w2 * (ui[0] * bj[0] + ui[1] * bj[1] + ui[2] * bj[2]);
  }
  double di = D[i];
  double db = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double bi = b[k * ld + i];
    double c = (w2 * s2[k] + (2.0 - w) * w * bi) / di;
    bbar[k * ld + i] = w2 * s1[k] / di + (2.0 - w) * w * ui[k];
    db -= ui[k] * c + w2 * bi * s1[k] / (di * di);
  }
  Dbar[i] += db;
}

// Adjoint of the operator update: from Gbar (M^2-entry adjoints scaled by D_row^{-1})
// to the diagonal D and the R1 values (through the 2-path products and R1_ij).
__global__ void k_adj_operators(int64_t nE, double w, const int* __restrict__ rowFlag, const double* __restrict__ r1e,
                                const int* __restrict__ r1Idx, const int64_t* __restrict__ pOff,
                                const int* __restrict__ pK, const double* __restrict__ pA, const int* __restrict__ pIa,
                                const int* __restrict__ pIb, const double* __restrict__ r1Val,
                                const double* __restrict__ D, const double* __restrict__ Gbar,
                                double* __restrict__ Dbar, double* __restrict__ R1bar) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nE) return;
  double g = Gbar[e];
  if (g == 0.0) return;
  int row = rowFlag[e] & 0x7fffffff;
  const double w2 = w * w;
  double s = 0.0;
  for (int64_t p = pOff[e]; p < pOff[e + 1]; p++) {
    int k = pK[p];
    double dk = D[k];
    s += pA[p] / dk;
    atomicAdd(Dbar + k, -g * w2 * pA[p] / (dk * dk));
    double abar = g * w2 / dk;
    atomicAdd(R1bar + pIa[p], abar * r1Val[pIb[p]]);
    atomicAdd(R1bar + pIb[p], abar * r1Val[pIa[p]]);
  }
  atomicAdd(Dbar + row, -g * (w2 * s + 2.0 * (1.0 - w) * w * r1e[e]) / D[row]);
  int q = r1Idx[e];
  if (q >= 0) atomicAdd(R1bar + q, g * 2.0 * (1.0 - w) * w);
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

void launch_solve_taped(const DeviceMesh& m, double omega, double rho, int iters, double* const* outs, hipStream_t s) {
  switch (m.sellLanes) {
    case 1: solve_taped_impl<1>(m, omega, rho, iters, outs, s); break;
    case 2: solve_taped_impl<2>(m, omega, rho, iters, outs, s); break;
    case 8: solve_taped_impl<8>(m, omega, rho, iters, outs, s); break;
    default: solve_taped_impl<4>(m, omega, rho, iters, outs, s); break;
  }
}

void launch_adj_sweep(const DeviceMesh& m, double a, double b, const double* uin, const double* uoth, const double* xk,
                      double* uout, double* Gbar, double* uc, hipStream_t s) {
  switch (m.sellLanes) {
    case 1: adj_sweep_impl<1>(m, a, b, uin, uoth, xk, uout, Gbar, uc, s); break;
    case 2: adj_sweep_impl<2>(m, a, b, uin, uoth, xk, uout, Gbar, uc, s); break;
    case 8: adj_sweep_impl<8>(m, a, b, uin, uoth, xk, uout, Gbar, uc, s); break;
    default: adj_sweep_impl<4>(m, a, b, uin, uoth, xk, uout, Gbar, uc, s); break;
  }
}

void launch_adj_first(const DeviceMesh& m, double omega, const double* b, const double* uc, double* bbar, double* Dbar,
                      double* R1bar, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_first, dim3((m.nSim + kBlock - 1) / kBlock), dim3(kBlock), 0, s, m.nSim, m.ld, omega,
                     m.r1Off, m.r1Col, m.r1Val, m.sysDiag, b, uc, bbar, Dbar, R1bar);
}

void launch_adj_operators(const DeviceMesh& m, const AdjMesh& a, double omega, const double* Gbar, double* Dbar,
                          double* R1bar, hipStream_t s) {
  int64_t nE = m.sellEntries;
  hipLaunchKernelGGL(k_adj_operators, dim3((unsigned)((nE + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, nE, omega,
                     m.m2RowFlag, m.m2R1, a.sellR1Idx, m.p2Off, m.p2K, m.p2A, a.pIa, a.pIb, m.r1Val, m.sysDiag, Gbar,
                     Dbar, R1bar);
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
