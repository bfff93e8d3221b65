// Broad phase (gpuCcdCulling, Simulator/PD-IPC/PdIpcSimulator.cpp:1799-1869), rebuilt
// MI355X-first:
//   1. swept, inflated patch boxes (updateCcdPatchBboxesKernel, PdIpcGpuFunc.cu:1407-1473);
//   2. a Karras LBVH over the patch boxes, rebuilt every step on the device:
//      30-bit Morton codes of the box centroids normalised to the scene box
//      (lbvh/morton_code.cuh:9-40, bvh.cuh:197-221), a stable LSD radix sort
//      (8-bit digits; in-block stable ranking with wavefront ballots), the
//      (code << 32 | index) tie-break, Karras determine_range / find_split
//      (bvh.cuh:69-163) and a bottom-up refit with arrival flags;
//   3. one lane per patch traverses the tree and reports every patch q > p
//      whose box overlaps (checkOverlap, PdIpcGpuFunc.cu:1615-1626, closed boxes):
//      a count pass, an exclusive scan and a fill pass, so the pair list is
//      ordered by p and deterministic;
//   4. every overlapping patch pair expands to its point-triangle and
//      edge-edge records in the reference's order (getCcdPotentialICPairsListKernel,
//      PdIpcGpuFunc.cu:1838-1948), one lane per record.
// The pair set equals the reference heap-BVH's (both are "all overlapping
// pairs"); the heap topology itself is not reproduced.
#include "device.hpp"

namespace pdipc {
namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 4;                       // per thread
constexpr int kScanTile = kBlock * kScanItems;      // 1024
constexpr int kRadixItems = 8;
constexpr int kRadixTile = kBlock * kRadixItems;    // 2048
constexpr int kStack = 64;

__device__ inline double dmin(double a, double b) { return a < b ? a : b; }
__device__ inline double dmax(double a, double b) { return a > b ? a : b; }

// ---------------- patch boxes ----------------
__global__ void k_patch_boxes(int P, int ld, const int* __restrict__ pfOff, const int* __restrict__ pfList,
                              const int3* __restrict__ faces, const double* __restrict__ X,
                              const double* __restrict__ DX, int nSim, double infl, double* __restrict__ box) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double lo[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
  double hi[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
  for (int q = pfOff[p]; q < pfOff[p + 1]; q++) {
    int3 f = faces[pfList[q]];
    int v[3] = {f.x, f.y, f.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double a0 = X[k * ld + v[0]], a1 = X[k * ld + v[1]], a2 = X[k * ld + v[2]];
      double b0 = a0 + (v[0] < nSim ? DX[k * ld + v[0]] : 0.0);
      double b1 = a1 + (v[1] < nSim ? DX[k * ld + v[1]] : 0.0);
      double b2 = a2 + (v[2] < nSim ? DX[k * ld + v[2]] : 0.0);
      double mn = dmin(dmin(dmin(a0, a1), dmin(a2, b0)), dmin(b1, b2)) - infl;
      double mx = dmax(dmax(dmax(a0, a1), dmax(a2, b0)), dmax(b1, b2)) + infl;
      lo[k] = dmin(lo[k], mn);
      hi[k] = dmax(hi[k], mx);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    box[k * P + p] = lo[k];
    box[(3 + k) * P + p] = hi[k];
  }
}

// ---------------- scene box of the leaf boxes (lbvh's aabb_whole) ----------------
__global__ void k_scene_partial(int P, const double* __restrict__ box, double* __restrict__ part) {
  __shared__ double red[6][kBlock / kWave];
  double v[6];
#pragma unroll
  for (int k = 0; k < 3; k++) { v[k] = 1.7976931348623157e308; v[3 + k] = -1.7976931348623157e308; }
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = dmin(v[k], box[k * P + p]);
      v[3 + k] = dmax(v[3 + k], box[(3 + k) * P + p]);
    }
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = dmin(v[k], __shfl_xor(v[k], o, kWave));
      v[3 + k] = dmax(v[3 + k], __shfl_xor(v[3 + k], o, kWave));
    }
  int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 6; k++) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double a = red[k][0], b = red[3 + k][0];
      for (int q = 1; q < kBlock / kWave; q++) { a = dmin(a, red[k][q]); b = dmax(b, red[3 + k][q]); }
      part[6 * blockIdx.x + k] = a;
      part[6 * blockIdx.x + 3 + k] = b;
    }
  }
}

__device__ inline uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void k_scene_final(int nPart, const double* __restrict__ part, double* __restrict__ scene) {
  __shared__ double red[6][kBlock / kWave];
  double v[6];
#pragma unroll
  for (int k = 0; k < 3; k++) { v[k] = 1.7976931348623157e308; v[3 + k] = -1.7976931348623157e308; }
  for (int q = threadIdx.x; q < nPart; q += blockDim.x)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = dmin(v[k], part[6 * q + k]);
      v[3 + k] = dmax(v[3 + k], part[6 * q + 3 + k]);
    }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = dmin(v[k], __shfl_xor(v[k], o, kWave));
      v[3 + k] = dmax(v[3 + k], __shfl_xor(v[3 + k], o, kWave));
    }
  int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 6; k++) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; k++) {
      double a = red[k][0], b = red[3 + k][0];
      for (int q = 1; q < kBlock / kWave; q++) { a = dmin(a, red[k][q]); b = dmax(b, red[3 + k][q]); }
      scene[k] = a;
      scene[3 + k] = b;
    }
}

// morton_code(double4) of lbvh/morton_code.cuh:34-44 applied to the box
// centroid normalised to the scene box (default_morton_code_calculator, bvh.cuh:197-221)
__global__ void k_morton(int P, const double* __restrict__ scene, const double* __restrict__ box,
                         uint32_t* __restrict__ code, int* __restrict__ idx) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint32_t e[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double c = (box[(3 + k) * P + p] + box[k * P + p]) * 0.5;
    c -= scene[k];
    c /= (scene[3 + k] - scene[k]);
    double r = fmin(fmax(c * 1024.0, 0.0), 1023.0);
    e[k] = expand_bits((uint32_t)r);
  }
  code[p] = e[0] * 4 + e[1] * 2 + e[2];
  idx[p] = p;
}

// ---------------- stable LSD radix sort (8-bit digits) ----------------
__global__ void k_radix_hist(int n, int shift, const uint32_t* __restrict__ keys, int* __restrict__ hist) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int base = blockIdx.x * kRadixTile;
#pragma unroll
  for (int j = 0; j < kRadixItems; j++) {
    int i = base + j * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// keys of a tile are visited in order (round j covers tile positions j*256 + t);
// ranks among equal digits come from 8 wavefront ballots, then from the waves
// before this one, then from the earlier rounds -- a stable scatter.
__global__ void k_radix_scatter(int n, int shift, const uint32_t* __restrict__ keysIn, const int* __restrict__ idxIn,
                                const int* __restrict__ histScan, uint32_t* __restrict__ keysOut,
                                int* __restrict__ idxOut) {
  constexpr int NW = kBlock / kWave;
  __shared__ int waveCnt[NW][256];
  __shared__ int digitBase[256];
  __shared__ int globalBase[256];
  int t = threadIdx.x, lane = t % kWave, w = t / kWave;
  digitBase[t] = 0;
  globalBase[t] = histScan[t * gridDim.x + blockIdx.x];
#pragma unroll
  for (int q = 0; q < NW; q++) waveCnt[q][t] = 0;
  __syncthreads();
  int base = blockIdx.x * kRadixTile;
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int j = 0; j < kRadixItems; j++) {
    int i = base + j * kBlock + t;
    bool valid = i < n;
    uint32_t key = valid ? keysIn[i] : 0u;
    int val = valid ? idxIn[i] : 0;
    int d = (int)((key >> shift) & 255u);
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      unsigned long long bb = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? bb : ~bb;
    }
    int rank = __popcll(peers & lt);
    int leader = valid ? (__ffsll((long long)peers) - 1) : -1;
    if (valid && lane == leader) waveCnt[w][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = digitBase[d] + rank;
      for (int q = 0; q < w; q++) pos += waveCnt[q][d];
      int dst = globalBase[d] + pos;
      keysOut[dst] = key;
      idxOut[dst] = val;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) { add += waveCnt[q][t]; waveCnt[q][t] = 0; }
    digitBase[t] += add;
    __syncthreads();
  }
}

// ---------------- exclusive scan of int counts ----------------
__device__ inline int block_exclusive_scan(int v, int* sh, int* total) {
  // inclusive scan within waves via shuffles, then across waves
  int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  int x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) sh[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int q = 0; q < kBlock / kWave; q++) { int c = sh[q]; sh[q] = s; s += c; }
    sh[kBlock / kWave] = s;
  }
  __syncthreads();
  int r = x - v + sh[w];
  *total = sh[kBlock / kWave];
  __syncthreads();
  return r;
}

__global__ void k_scan_tiles(int n, const int* __restrict__ in, int* __restrict__ out, int* __restrict__ blockSums) {
  __shared__ int sh[kBlock / kWave + 1];
  int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems], s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int tot;
  int pre = block_exclusive_scan(s, sh, &tot);
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    if (base + k < n) out[base + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) blockSums[blockIdx.x] = tot;
}

__global__ void k_scan_sums(int nb, int* __restrict__ sums) {
  __shared__ int sh[kBlock / kWave + 1];
  int carry = 0;
  for (int base = 0; base < nb; base += kScanTile) {
    int idx = base + threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      v[k] = (idx + k < nb) ? sums[idx + k] : 0;
      s += v[k];
    }
    int tot;
    int pre = block_exclusive_scan(s, sh, &tot) + carry;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      if (idx + k < nb) sums[idx + k] = pre;
      pre += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void k_scan_add(int n, int nb, const int* __restrict__ sums, int* __restrict__ out, int* __restrict__ total) {
  int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int add = sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; k++)
    if (base + k < n) out[base + k] += add;
  if (blockIdx.x == 0 && threadIdx.x == 0 && total) *total = sums[nb];
}

// ---------------- Karras LBVH ----------------
__device__ inline int cub64(uint64_t a, uint64_t b) { return __clzll((long long)(a ^ b)); }

__global__ void k_make_keys(int P, const uint32_t* __restrict__ code, const int* __restrict__ idx,
                            uint64_t* __restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P) keys[i] = ((uint64_t)code[i] << 32) | (uint32_t)idx[i];
}

__global__ void k_karras(int P, const uint64_t* __restrict__ keys, int* __restrict__ left, int* __restrict__ right,
                         int* __restrict__ parent, int2* __restrict__ range) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P - 1) return;
  int first, last;
  if (idx == 0) {
    first = 0;
    last = P - 1;
  } else {
    uint64_t self = keys[idx];
    int L = cub64(self, keys[idx - 1]);
    int R = cub64(self, keys[idx + 1]);
    int d = R > L ? 1 : -1;
    int dmin_ = L < R ? L : R;
    int lmax = 2, delta = -1;
    int it = idx + d * lmax;
    if (0 <= it && it < P) delta = cub64(self, keys[it]);
    while (delta > dmin_) {
      lmax <<= 1;
      it = idx + d * lmax;
      delta = -1;
      if (0 <= it && it < P) delta = cub64(self, keys[it]);
    }
    int l = 0, t = lmax >> 1;
    while (t > 0) {
      it = idx + (l + t) * d;
      delta = -1;
      if (0 <= it && it < P) delta = cub64(self, keys[it]);
      if (delta > dmin_) l += t;
      t >>= 1;
    }
    int jdx = idx + l * d;
    first = d < 0 ? jdx : idx;
    last = d < 0 ? idx : jdx;
  }
  // find_split
  uint64_t fc = keys[first], lc = keys[last];
  int gamma;
  if (fc == lc) {
    gamma = (first + last) >> 1;
  } else {
    int dn = cub64(fc, lc);
    int split = first, stride = last - first;
    do {
      stride = (stride + 1) >> 1;
      int middle = split + stride;
      if (middle < last) {
        int dd = cub64(fc, keys[middle]);
        if (dd > dn) split = middle;
      }
    } while (stride > 1);
    gamma = split;
  }
  int lc_ = gamma, rc_ = gamma + 1;
  if (first == gamma) lc_ += P - 1;
  if (last == gamma + 1) rc_ += P - 1;
  left[idx] = lc_;
  right[idx] = rc_;
  range[idx] = make_int2(first, last);
  parent[lc_] = idx;
  parent[rc_] = idx;
  if (idx == 0) parent[0] = -1;
}

// Node boxes without inter-block synchronisation: every internal node covers a
// contiguous leaf range [first, last] (Karras), so its box is the union of the
// leaf boxes over that range, read from a sparse table of power-of-two window
// unions (level j holds the union of leaves [i, i + 2^j)).  Min / max are exact
// and order-free, so the result equals the bottom-up refit bit for bit.
__global__ void k_table_level0(int P, const int* __restrict__ sortedIdx, const double* __restrict__ pbox,
                               double* __restrict__ tab) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  int obj = sortedIdx[i];
#pragma unroll
  for (int k = 0; k < 6; k++) tab[(int64_t)k * P + i] = pbox[(int64_t)k * P + obj];
}

__global__ void k_table_level(int P, int half, const double* __restrict__ prev, double* __restrict__ cur) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 2 * half > P) return;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    cur[(int64_t)k * P + i] = dmin(prev[(int64_t)k * P + i], prev[(int64_t)k * P + i + half]);
    cur[(int64_t)(3 + k) * P + i] = dmax(prev[(int64_t)(3 + k) * P + i], prev[(int64_t)(3 + k) * P + i + half]);
  }
}

__global__ void k_node_boxes(int P, const int2* __restrict__ range, const double* __restrict__ tab,
                             double* __restrict__ nbox) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int NN = 2 * P - 1;
  if (n >= NN) return;
  if (n >= P - 1) {
    int i = n - (P - 1);
#pragma unroll
    for (int k = 0; k < 6; k++) nbox[(int64_t)k * NN + n] = tab[(int64_t)k * P + i];
    return;
  }
  int2 r = range[n];
  int len = r.y - r.x + 1;
  int j = 31 - __clz(len);
  const double* t = tab + (int64_t)j * 6 * P;
  int a = r.x, b = r.y - (1 << j) + 1;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    nbox[(int64_t)k * NN + n] = dmin(t[(int64_t)k * P + a], t[(int64_t)k * P + b]);
    nbox[(int64_t)(3 + k) * NN + n] = dmax(t[(int64_t)(3 + k) * P + a], t[(int64_t)(3 + k) * P + b]);
  }
}

__device__ inline bool overlap(const double* a, const double* b) {
  if (a[3] < b[0] || b[3] < a[0]) return false;
  if (a[4] < b[1] || b[4] < a[1]) return false;
  if (a[5] < b[2] || b[5] < a[2]) return false;
  return true;
}

// One lane per patch p walks the tree and appends every patch q > p whose box
// overlaps; appends are aggregated per wavefront (one atomic per wave per hit
// round).  Pair order is therefore not fixed; the pair set is.  Static patches
// follow all simulated ones in patch order, so with skipStatic a static query
// has nothing to report and exits at once.
__global__ void __launch_bounds__(kBlock) k_traverse(int P, const double* __restrict__ pbox,
                                                     const double* __restrict__ nbox, const int* __restrict__ left,
                                                     const int* __restrict__ right, const int* __restrict__ sortedIdx,
                                                     const uint8_t* __restrict__ patchStatic, int skipStatic,
                                                     int* __restrict__ counter, int2* __restrict__ pairs, int cap,
                                                     int* __restrict__ overflow) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  int lane = threadIdx.x % kWave;
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  bool alive = p < P && !(skipStatic && patchStatic[p]);
  if (!alive) return;
  const int NN = 2 * P - 1;
  double q[6];
#pragma unroll
  for (int k = 0; k < 6; k++) q[k] = pbox[(int64_t)k * P + p];
  int stack[kStack];
  int sp = 0;
  stack[sp++] = 0;  // root: node 0 (the single leaf when P == 1)
  while (sp > 0) {
    int node = stack[--sp];
    double b[6];
#pragma unroll
    for (int k = 0; k < 6; k++) b[k] = nbox[(int64_t)k * NN + node];
    bool hit = false;
    int obj = -1;
    if (overlap(q, b)) {
      if (node >= P - 1) {
        obj = sortedIdx[node - (P - 1)];
        // every pair p < q is emitted once: by p, or -- when static queries are
        // skipped -- by the simulated side of a static/simulated pair
        hit = obj > p || (skipStatic && obj < p && patchStatic[obj]);
      } else if (sp + 2 <= kStack) {
        stack[sp++] = right[node];  // right first: the left subtree is visited first
        stack[sp++] = left[node];
      } else {
        atomicOr(overflow, 1);  // reported to the host as an error, never silently dropped
      }
    }
    unsigned long long m = __ballot(hit);
    if (m) {
      int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(counter, (int)__popcll(m));
      base = __shfl(base, leader, kWave);
      if (hit) {
        int pos = base + (int)__popcll(m & lt);
        if (pos < cap) pairs[pos] = make_int2(min(p, obj), max(p, obj));
      }
    }
  }
}

__global__ void k_rec_count(int Q, const int2* __restrict__ pairs, const int* __restrict__ ppOff,
                            const int* __restrict__ pfOff, const int* __restrict__ peOff, int* __restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Q) return;
  int2 pr = pairs[i];
  int pa = ppOff[pr.x + 1] - ppOff[pr.x], fa = pfOff[pr.x + 1] - pfOff[pr.x], ea = peOff[pr.x + 1] - peOff[pr.x];
  int pb = ppOff[pr.y + 1] - ppOff[pr.y], fb = pfOff[pr.y + 1] - pfOff[pr.y], eb = peOff[pr.y + 1] - peOff[pr.y];
  count[i] = pa * fb + pb * fa + ea * eb;
}

__device__ inline int last_le(const int* __restrict__ off, int lo, int hi, int r) {
  while (lo < hi) {  // last i in [lo, hi] with off[i] <= r
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(kBlock) k_rec_fill(int Q, int K, const int2* __restrict__ pairs,
                                                     const int* __restrict__ recOff, const int* __restrict__ ppOff,
                                                     const int* __restrict__ ppList, const int* __restrict__ pfOff,
                                                     const int* __restrict__ pfList, const int* __restrict__ peOff,
                                                     const int* __restrict__ peList, const int3* __restrict__ faces,
                                                     const int2* __restrict__ edges, int4* __restrict__ recs) {
  __shared__ int sRange[2];
  __shared__ int sOff[kBlock + 1];
  int r0 = blockIdx.x * kBlock;
  int r1 = min(r0 + kBlock, K) - 1;
  if (threadIdx.x == 0) sRange[0] = last_le(recOff, 0, Q - 1, r0);
  if (threadIdx.x == 1) sRange[1] = last_le(recOff, 0, Q - 1, r1);
  __syncthreads();
  int lo = sRange[0], hi = sRange[1];
  bool local = hi - lo + 1 <= kBlock;
  if (local && threadIdx.x <= hi - lo) sOff[threadIdx.x] = recOff[lo + threadIdx.x];
  __syncthreads();
  int r = r0 + threadIdx.x;
  if (r >= K) return;
  int pi = local ? lo + last_le(sOff, 0, hi - lo, r) : last_le(recOff, lo, hi, r);
  int2 pr = pairs[pi];
  int id = r - recOff[pi];
  int A = pr.x, B = pr.y;
  int pa = ppOff[A + 1] - ppOff[A], fa = pfOff[A + 1] - pfOff[A];
  int pb = ppOff[B + 1] - ppOff[B], fb = pfOff[B + 1] - pfOff[B], eb = peOff[B + 1] - peOff[B];
  int VaFb = pa * fb, VbFa = pb * fa;
  int4 out;
  if (id < VaFb) {
    int3 f = faces[pfList[pfOff[B] + id % fb]];
    out = make_int4(ppList[ppOff[A] + id / fb], f.x, f.y, f.z);
  } else if ((id -= VaFb) < VbFa) {
    int3 f = faces[pfList[pfOff[A] + id % fa]];
    out = make_int4(ppList[ppOff[B] + id / fa], f.x, f.y, f.z);
  } else {
    id -= VbFa;
    int2 e1 = edges[peList[peOff[A] + id / eb]];
    int2 e2 = edges[peList[peOff[B] + id % eb]];
    out = make_int4(~e1.x, e1.y, e2.x, e2.y);
  }
  recs[r] = out;
}

}  // namespace

void launch_exclusive_scan(int* in, int* out, int n, int* total_out, int* blockScratch, hipStream_t s) {
  if (n <= 0) return;
  int nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(kBlock), 0, s, n, in, out, blockScratch);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, nb, blockScratch);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kBlock), 0, s, n, nb, blockScratch, out, total_out);
}

void launch_patch_boxes(const DeviceMesh& m, const double* x, const double* dx, double inflation, hipStream_t s) {
  int P = m.nPatches;
  if (P <= 0) return;
  hipLaunchKernelGGL(k_patch_boxes, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, m.ld, m.pfOff, m.pfList,
                     m.bFaces, x, dx, m.nSim, inflation, m.patchBox);
}

void launch_lbvh(DeviceMesh& m, hipStream_t s) {
  int P = m.nPatches;
  if (P <= 0) return;
  int nPart = std::min(256, (P + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_scene_partial, dim3(nPart), dim3(kBlock), 0, s, P, m.patchBox, m.sceneBoxPart);
  hipLaunchKernelGGL(k_scene_final, dim3(1), dim3(kBlock), 0, s, nPart, m.sceneBoxPart, m.sceneBox);
  hipLaunchKernelGGL(k_morton, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, m.sceneBox, m.patchBox,
                     m.mortonA, m.idxA);
  int nb = (P + kRadixTile - 1) / kRadixTile;
  uint32_t *kin = m.mortonA, *kout = m.mortonB;
  int *iin = m.idxA, *iout = m.idxB;
  for (int shift = 0; shift < 30; shift += 8) {
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kBlock), 0, s, P, shift, kin, m.radixHist);
    launch_exclusive_scan(m.radixHist, m.radixHist, 256 * nb, nullptr, m.radixBlockSums, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kBlock), 0, s, P, shift, kin, iin, m.radixHist, kout, iout);
    std::swap(kin, kout);
    std::swap(iin, iout);
  }
  // 4 passes: the sorted data is back in the A buffers
  hipLaunchKernelGGL(k_make_keys, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, kin, iin, m.keys);
  if (P > 1)
    hipLaunchKernelGGL(k_karras, dim3((P - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, m.keys, m.nodeLeft,
                       m.nodeRight, m.nodeParent, m.nodeRange);
  hipLaunchKernelGGL(k_table_level0, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, iin, m.patchBox,
                     m.nodeTable);
  for (int j = 1; (1 << j) <= P; j++)
    hipLaunchKernelGGL(k_table_level, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, 1 << (j - 1),
                       m.nodeTable + (int64_t)(j - 1) * 6 * P, m.nodeTable + (int64_t)j * 6 * P);
  hipLaunchKernelGGL(k_node_boxes, dim3((2 * P - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, m.nodeRange,
                     m.nodeTable, m.nodeBox);
}

void launch_pairs(const DeviceMesh& m, int skipStatic, hipStream_t s) {
  int P = m.nPatches;
  if (P <= 0) return;
  hipLaunchKernelGGL(k_traverse, dim3((P + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P, m.patchBox, m.nodeBox,
                     m.nodeLeft, m.nodeRight, m.idxA, m.patchStatic, skipStatic, m.counters + 0, m.pairs, m.pairCap,
                     m.counters + 3);
}

void launch_rec_count(const DeviceMesh& m, int nPairs, hipStream_t s) {
  if (nPairs <= 0) return;
  hipLaunchKernelGGL(k_rec_count, dim3((nPairs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nPairs, m.pairs, m.ppOff,
                     m.pfOff, m.peOff, m.recCount);
  launch_exclusive_scan(m.recCount, m.recOff, nPairs, m.counters + 1, m.scanBlock, s);
}

void launch_rec_fill(const DeviceMesh& m, int nPairs, int nRecs, hipStream_t s) {
  if (nRecs <= 0 || nPairs <= 0) return;
  hipLaunchKernelGGL(k_rec_fill, dim3((nRecs + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nPairs, nRecs, m.pairs,
                     m.recOff, m.ppOff, m.ppList, m.pfOff, m.pfList, m.peOff, m.peList, m.bFaces, m.bEdges, m.recs);
}

}  // namespace pdipc
