// Ordered stream compaction for gfx950: the indices e in [0, n) of each
// batch b whose element x[b * n + e] satisfies a predicate, listed in
// increasing order (so work driven by the list keeps the original order).
// Three launches: per-block counts, an exclusive scan of the counts per
// batch, and the scatter.
#pragma once

#include "common.h"

namespace totri_hip {
namespace {

constexpr int kCompactBlock = 1024;  // elements per counting block

struct NonZero {
  __device__ bool operator()(float v) const { return v != 0.f; }
};
struct Positive {
  __device__ bool operator()(float v) const { return v > 0.f; }
};

template <class Pred>
__global__ void compact_count_kernel(const float* __restrict__ x, int n,
                                     int* __restrict__ counts, int nblocks, Pred pred) {
  const int b = blockIdx.y;
  const int base = blockIdx.x * kCompactBlock;
  int c = 0;
  for (int e = threadIdx.x; e < kCompactBlock; e += blockDim.x) {
    const int v = base + e;
    if (v < n && pred(x[(size_t)b * n + v])) ++c;
  }
  __shared__ int red[256];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[(size_t)b * nblocks + blockIdx.x] = red[0];
}

// Exclusive scan of the per-block counts of one batch (one 256-lane
// workgroup per b): each lane scans a contiguous run of counts, the run
// totals are scanned in LDS, and each lane writes its run's offsets.
constexpr int kCompactScanLanes = 256;

__global__ __launch_bounds__(kCompactScanLanes) void compact_scan_kernel(
    int* __restrict__ counts, int nblocks, int* __restrict__ totals) {
  __shared__ int run_tot[kCompactScanLanes];
  const int b = blockIdx.x;
  int* c = counts + (size_t)b * nblocks;
  const int per = (nblocks + kCompactScanLanes - 1) / kCompactScanLanes;
  const int lo = min((int)threadIdx.x * per, nblocks), hi = min(lo + per, nblocks);
  int sum = 0;
  for (int k = lo; k < hi; ++k) sum += c[k];
  run_tot[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < kCompactScanLanes; ++k) {
      const int v = run_tot[k];
      run_tot[k] = acc;
      acc += v;
    }
    totals[b] = acc;
  }
  __syncthreads();
  int acc = run_tot[threadIdx.x];
  for (int k = lo; k < hi; ++k) {
    const int v = c[k];
    c[k] = acc;
    acc += v;
  }
}

template <class Pred>
__global__ void compact_scatter_kernel(const float* __restrict__ x, int n,
                                       const int* __restrict__ offsets, int nblocks,
                                       int* __restrict__ list, Pred pred) {
  const int b = blockIdx.y;
  const int base = blockIdx.x * kCompactBlock;
  __shared__ int wave_tot[4];
  int out = offsets[(size_t)b * nblocks + blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int e0 = 0; e0 < kCompactBlock; e0 += blockDim.x) {
    const int v = base + e0 + threadIdx.x;
    const bool keep = v < n && pred(x[(size_t)b * n + v]);
    const unsigned long long m = __ballot(keep);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = __popcll(m);
    __syncthreads();
    int wbase = 0;
    for (int k = 0; k < wave; ++k) wbase += wave_tot[k];
    const int step = wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
    if (keep) list[(size_t)b * n + out + wbase + before] = v;
    out += step;
    __syncthreads();
  }
}

// Workspace layout: counts [B][nblocks] | totals [B] | list [B][n], each
// 256-byte aligned.
struct CompactLayout {
  int nblocks;
  size_t counts_off, totals_off, list_off, end;
};

inline CompactLayout compact_layout(int64_t B, int64_t n, size_t base) {
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  CompactLayout c;
  c.nblocks = (int)((n + kCompactBlock - 1) / kCompactBlock);
  c.counts_off = align(base);
  c.totals_off = c.counts_off + align((size_t)B * c.nblocks * sizeof(int));
  c.list_off = c.totals_off + align((size_t)B * sizeof(int));
  c.end = c.list_off + align((size_t)B * n * sizeof(int));
  return c;
}

// Launches the three kernels (n >= 1); `ws` holds the layout.
template <class Pred>
int launch_compact(const float* x, int B, int n, const CompactLayout& c, void* ws, Pred pred,
                   hipStream_t st) {
  char* w = (char*)ws;
  int* counts = (int*)(w + c.counts_off);
  int* totals = (int*)(w + c.totals_off);
  int* list = (int*)(w + c.list_off);
  hipLaunchKernelGGL(compact_count_kernel<Pred>, dim3(c.nblocks, (unsigned)B), dim3(256), 0,
                     st, x, n, counts, c.nblocks, pred);
  int rc = check_launch("compact_count_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(compact_scan_kernel, dim3((unsigned)B), dim3(kCompactScanLanes), 0, st,
                     counts, c.nblocks, totals);
  rc = check_launch("compact_scan_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(compact_scatter_kernel<Pred>, dim3(c.nblocks, (unsigned)B), dim3(256), 0,
                     st, x, n, counts, c.nblocks, list, pred);
  return check_launch("compact_scatter_kernel");
}

}  // namespace
}  // namespace totri_hip
