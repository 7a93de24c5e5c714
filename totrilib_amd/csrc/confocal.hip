// Confocal volume render (R1) and its exact transpose (A1) for gfx950.
//
// Reference: src/render/volume_cuda.cu:11-186 (kernels),
//            src/render/volume.cpp:7-41 (shape checks).
//
// A "sample" is one (scan point i, bin t, column (xi, yi)) triple.  Its
// discrete decisions -- d2 > 0 and zi = floor(z_hat) with 0 <= zi <= Z-2 --
// are taken exactly as the reference's FP32 formula takes them
// (volume_cuda.cu:23,35,38,40-49): products and sums rounded one by one and
// a correctly rounded square root.  The fast path uses the hardware square
// root and a fused z_hat; it is only trusted when z_hat lies farther than a
// proven error bound `delta` from an integer, otherwise the sample is
// recomputed with the exact formula.  Weights (alpha, f2) are FP32 and agree
// with the reference formula to a few ulps.
//
// R1 (gather into bins): a wave owns one scan point and 64*K bins, one bin
// per lane and window, accumulating in registers.  The workgroup's waves
// (eight scan points) share volume chunks that are streamed through LDS with
// double buffering; per chunk each wave culls columns whose active bin range
// misses its windows.  Column groups split the volume across workgroups when
// there are too few scan points to fill the GPU; their partial transients are
// summed in a fixed order by a second kernel (deterministic).
//
// A1 (scatter into voxels): a wave owns an 8x8 tile of columns (one column
// per lane) and a z slab of ZS voxels whose accumulators live in LDS, one
// row per voxel and one slot per lane, so no two lanes ever touch the same
// address and updates are plain read-modify-writes (no atomics).  Bins are uniform across the wave; their
// per-bin factors come from a small precomputed table through scalar loads.
// Scan-point ranges split the work when tiles x slabs are too few; partial
// volumes are summed in a fixed order.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <math.h>
#include <string.h>

#include "common.h"

namespace totri_hip {
namespace {

constexpr int kWaves = 8;            // scan points per R1 workgroup
constexpr int kChunkVoxels = 4096;   // voxels per R1 LDS chunk buffer (32 KiB of pairs)
constexpr int kSlab = 32;            // A1 z-slab depth
constexpr int kFlush = 32;           // A1 scan points per first-level sum

struct ConfGeom {
  int X, Y, Z, N, T;
  float ox, oy, oz;
  float szx, szy, szz;
  float w, off;
  float s;       // (float)Z / size.z, volume_cuda.cu:30
  float tc_max;  // max |tc| over all bins
};

// Guard half-width on frac(z_hat): an upper bound (with a factor-4 margin)
// on |z_hat_fast - z_hat_exact| over samples with z_hat in [-1, Z], where
// z_hat_fast = fma(sqrt_fast(d2), s, c) and z_hat_exact follows the
// reference's rounding sequence.  Terms of the bound, in units of 2^-24:
//   4 s tc_max       v_sqrt_f32 within 2 ulp of sqrt_rn (depth <= tc_max), times s
//   s|oz| + 2(Z+1)   roundings of z = wz + depth and of (z - oz) * s
//   2 s|wz - oz|     rounding of the sum inside z, scaled by s
//   3Z + |c0| + |c_shift| + 3
//                    the final -0.5 and the roundings of the precomputed
//                    constant c0 = (wz - oz) * s - 0.5 - shift
// The sum B is multiplied by 4 (2^-22 instead of 2^-24) for margin.
__device__ __forceinline__ float guard_delta(const ConfGeom& g, float wz, float c0,
                                             float c_shift) {
  const float s = fabsf(g.s);
  float b = 4.f * s * g.tc_max + s * fabsf(g.oz) + 2.f * (float)(g.Z + 1) +
            2.f * s * fabsf(wz - g.oz) + 3.f * (float)g.Z + fabsf(c0) +
            fabsf(c_shift) + 3.f;
  return fminf(b * 2.384185791015625e-07f /* 2^-22 */, 0.5f);
}

// Conservative range of bins t whose one-way radius |tc| lies in
// [sqrt(rho2 + dlo^2), sqrt(rho2 + dhi^2)] (both branches of the sign of tc).
// Returns lo > hi when no bin can be active.
__device__ __forceinline__ void bin_range(const ConfGeom& g, float rho2, float dlo,
                                          float dhi, int* lo, int* hi) {
  if (!(dhi > 0.f)) {
    *lo = 1;
    *hi = 0;
    return;
  }
  if (!(g.w > 0.f) || !isfinite(g.w) || !isfinite(g.off)) {
    *lo = 0;
    *hi = g.T - 1;
    return;
  }
  const float dl = fmaxf(dlo, 0.f);
  const float r_lo = sqrtf(rho2 + dl * dl) * (1.f - 1e-5f) - 1e-6f;
  const float r_hi = sqrtf(rho2 + dhi * dhi) * (1.f + 1e-5f) + 1e-6f;
  const float top = (float)g.T + 1.f;
  // Positive radii tc in [r_lo, r_hi] and negative radii tc in [-r_hi, -r_lo]
  // give two bin intervals; each is clipped to [0, T-1] on its own and the
  // result is the hull of the non-empty ones.
  const float pa = fminf(fmaxf((r_lo - g.off) / g.w - 2.f, -1.f), top);
  const float pb = fminf(fmaxf((r_hi - g.off) / g.w + 2.f, -1.f), top);
  const float na = fminf(fmaxf((-r_hi - g.off) / g.w - 2.f, -1.f), top);
  const float nb = fminf(fmaxf((-r_lo - g.off) / g.w + 2.f, -1.f), top);
  int plo = max((int)floorf(pa), 0), phi = min((int)ceilf(pb), g.T - 1);
  int nlo = max((int)floorf(na), 0), nhi = min((int)ceilf(nb), g.T - 1);
  const bool pe = plo <= phi, ne = nlo <= nhi;
  if (pe && ne) {
    *lo = min(plo, nlo);
    *hi = max(phi, nhi);
  } else if (pe) {
    *lo = plo;
    *hi = phi;
  } else {
    *lo = nlo;
    *hi = nhi;
  }
}

// Depth interval [dlo, dhi) that keeps z_hat in [zlo, zhi) for a wall point
// at height wz, widened by a relative margin (culling only).
__device__ __forceinline__ void depth_range(const ConfGeom& g, float wz, float zlo,
                                            float zhi, float* dlo, float* dhi) {
  const float is = 1.f / g.s;
  const float margin = 1e-4f * (fabsf(g.oz) + fabsf(wz) + ((float)g.Z + 2.f) * fabsf(is)) + 1e-6f;
  *dlo = ((zlo + 0.5f) * is + g.oz) - wz - margin;
  *dhi = ((zhi + 0.5f) * is + g.oz) - wz + margin;
}

// Fast part of the decision: z_hat = fma(sqrt_fast(d2), s, cfast) gives the
// candidate floor (returned) and alpha.  `near` tells whether the exact
// sequence must confirm it (see confocal_decide).
__device__ __forceinline__ int decide_fast(const ConfGeom& g, float d2, float cfast,
                                           float delta, float one_m_delta, float* alpha,
                                           bool* near) {
  const float zh = fmaf(sqrt_fast(d2), g.s, cfast);
  const float al = __builtin_amdgcn_fractf(zh);
  *alpha = al;
  *near = d2 > 0.f && (al < delta || al > one_m_delta);
  return floor_to_int(zh);
}

// Exact reference sequence (volume_cuda.cu:42-51) for a sample flagged `near`.
__device__ __forceinline__ int decide_exact(const ConfGeom& g, float d2, float wz,
                                            int zshift, float* alpha) {
  const float de = sqrt_rn(d2);
  const float z = wz + de;
  const float zs = (z - g.oz) * g.s;
  const float zx = zs - 0.5f;
  const float fx = floorf(zx);
  *alpha = zx - fx;
  return (int)fx - zshift;
}

// Discrete decision of one sample with depth^2 = d2 (already computed with
// the reference rounding).  Returns floor(z_hat) - zshift and writes the
// interpolation weight alpha = z_hat - floor(z_hat).  `cfast` is
// (wz - oz) * s - 0.5 - zshift; the fast z_hat = fma(sqrt_fast(d2), s, cfast)
// is trusted only when its fractional part is more than `delta` away from an
// integer, otherwise z_hat is recomputed with the reference's exact sequence
// (volume_cuda.cu:42-47).  For d2 <= 0 the result is unspecified; callers
// test d2 > 0 themselves.
__device__ __forceinline__ int confocal_decide(const ConfGeom& g, float d2, float wz,
                                               float cfast, int zshift, float delta,
                                               float one_m_delta, float* alpha) {
  const float depth = sqrt_fast(d2);
  const float zh = fmaf(depth, g.s, cfast);
  // alpha = zh - floor(zh) (exact; v_fract only clamps results that would
  // round to 1.0, which happens for zh in (-1, 0), a rejected range anyway).
  float al = __builtin_amdgcn_fractf(zh);
  int r = floor_to_int(zh);
  if (d2 > 0.f && (al < delta || al > one_m_delta)) {
    const float de = sqrt_rn(d2);
    const float z = wz + de;
    const float zs = (z - g.oz) * g.s;
    const float zx = zs - 0.5f;
    const float fx = floorf(zx);
    al = zx - fx;
    r = (int)fx - zshift;
  }
  *alpha = al;
  return r;
}

__device__ __forceinline__ float sq(float v) { return v * v; }

__device__ __forceinline__ float lane_value(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// ---------------------------------------------------------------------------
// R1
// ---------------------------------------------------------------------------
// One R1 work item: scan points sgrp*8 .. sgrp*8+7 (one per wave), bins
// t_base .. t_base + 64K - 1, column group grp.  Every thread of the
// workgroup calls it (it synchronises the workgroup).  LDS holds two chunk
// buffers of (v[z], v[z+1] - v[z]) pairs, so that one 8-byte read and one
// FMA give the interpolated albedo.
template <int K>
__device__ __forceinline__ void r1_block(
    const float* __restrict__ vol, const float* __restrict__ mp,
    float* __restrict__ dst,  // [n_groups][N][T] partial sums, or [N][T]
    const ConfGeom& g, int t_base, int grp, int sgrp, int n_groups, int chunk_cols,
    int chunks_per_group, float2* __restrict__ lds2) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int i = sgrp * kWaves + wave;
  const bool scan_ok = i < g.N;
  const int n_cols = g.X * g.Y;
  const int chunk_vox = chunk_cols * g.Z;

  // Per-lane bin constants for each window.
  float tc2[K], itc4[K], a[K], acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int t = t_base + 64 * k + lane;
    const float tc = bin_distance(t, g.w, g.off);
    const float q = tc * tc;
    tc2[k] = (t < g.T) ? q : -1.f;  // -1 makes d2 negative: lane inactive
    itc4[k] = 1.f / (q * q);
    acc[k] = 0.f;
    a[k] = 0.f;
  }

  const int ii = scan_ok ? i : 0;
  const float wx = uniform_f32(mp[ii]);
  const float wy = uniform_f32(mp[g.N + ii]);
  const float wz = uniform_f32(mp[2 * g.N + ii]);
  float c0 = (wz - g.oz) * g.s;
  c0 = c0 - 0.5f;
  const float delta = guard_delta(g, wz, c0, 0.f);
  const float one_m_delta = 1.f - delta;
  float dlo, dhi;
  depth_range(g, wz, 0.f, (float)(g.Z - 1), &dlo, &dhi);

  const int chunk_first = grp * chunks_per_group;
  const int total_chunks = (n_cols + chunk_cols - 1) / chunk_cols;
  const int chunk_end = min(chunk_first + chunks_per_group, total_chunks);
  const int nthreads = kWaves * 64;
  const int vol_vox = n_cols * g.Z;

  // Pair (v[z], v[z+1] - v[z]) of global voxel index ge (z = Z-1 pairs with 0).
  auto load_pair = [&](int ge) -> float2 {
    if (ge >= vol_vox) return make_float2(0.f, 0.f);
    const float v0 = vol[ge];
    const bool last = (ge % g.Z) == g.Z - 1;
    const float v1 = last ? v0 : vol[ge + 1];
    return make_float2(v0, v1 - v0);
  };

  if (chunk_first < chunk_end) {
    const int base = chunk_first * chunk_vox;
    for (int e = threadIdx.x; e < chunk_vox; e += nthreads) lds2[e] = load_pair(base + e);
  }
  __syncthreads();

  int cached_xi = -1;
  constexpr int kPref = kChunkVoxels / (kWaves * 64);  // voxels per thread
  for (int c = chunk_first; c < chunk_end; ++c) {
    const int buf = (c - chunk_first) & 1;
    const float2* cur = lds2 + buf * chunk_vox;
    float2* nxt = lds2 + (buf ^ 1) * chunk_vox;
    const bool has_next = c + 1 < chunk_end;
    float2 pref[kPref];
    if (has_next) {
      const int base = (c + 1) * chunk_vox;
#pragma unroll
      for (int r = 0; r < kPref; ++r) {
        const int e = threadIdx.x + r * nthreads;
        pref[r] = e < chunk_vox ? load_pair(base + e) : make_float2(0.f, 0.f);
      }
    }

    if (scan_ok) {
      // Lane-parallel column setup: lane j describes chunk column j.
      const int col = c * chunk_cols + lane;
      const bool cvalid = lane < chunk_cols && col < n_cols;
      const int colc = cvalid ? col : 0;
      const int xi = colc / g.Y;
      const int yi = colc - xi * g.Y;
      const float x = cell_centre(xi, g.ox, g.szx, g.X);
      const float y = cell_centre(yi, g.oy, g.szy, g.Y);
      const float dx = wx - x, dy = wy - y;
      const float dx2 = dx * dx, dy2 = dy * dy;
      int tlo, thi;
      bin_range(g, dx2 + dy2, dlo, dhi, &tlo, &thi);
      const int klo = (max(tlo, t_base) - t_base) >> 6;
      const int khi = (min(thi, t_base + 64 * K - 1) - t_base) >> 6;
      const bool any = cvalid && thi >= t_base && tlo <= t_base + 64 * K - 1 && tlo <= thi;
      // Bit k set <=> window k may hold an active sample of this column.
      const unsigned packed =
          any ? (((2u << khi) - 1u) & ~((1u << klo) - 1u)) : 0u;
      unsigned long long mask = __ballot(any);
      // Two columns per pass: their samples are independent, which gives the
      // scheduler two dependency chains to interleave.  Windows are the union
      // of both columns' windows; a sample outside its column's range simply
      // fails the exact activity test.
      while (mask) {
        const int j1 = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int j2 = mask ? __builtin_ctzll(mask) : j1;
        const bool two = mask != 0;
        if (two) mask &= mask - 1;
        const unsigned wm1 = (unsigned)__builtin_amdgcn_readlane((int)packed, j1);
        const unsigned wm2 = two ? (unsigned)__builtin_amdgcn_readlane((int)packed, j2) : 0u;
        const unsigned wmask = wm1 | wm2;
        const int xi1 = __builtin_amdgcn_readlane(xi, j1);
        const int xi2 = __builtin_amdgcn_readlane(xi, j2);
        if (xi1 != cached_xi) {
          // tc^2 - dx^2 (volume_cuda.cu:40, first subtraction) depends on the
          // column's x only; recompute it when x changes.
          const float cdx2 = __builtin_bit_cast(
              float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, dx2), j1));
#pragma unroll
          for (int k = 0; k < K; ++k) a[k] = tc2[k] - cdx2;
          cached_xi = xi1;
        }
        // Second column's tc^2 - dx^2 (equal to a[] when it shares x).
        const float cdx2b = __builtin_bit_cast(
            float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, dx2), j2));
        const bool same_x = xi2 == cached_xi;
        const float cdy2a = __builtin_bit_cast(
            float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, dy2), j1));
        const float cdy2b = two ? __builtin_bit_cast(
            float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, dy2), j2)) : INFINITY;
        const float2* col1 = cur + j1 * g.Z;
        const float2* col2 = cur + j2 * g.Z;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (!((wmask >> k) & 1u)) continue;
          const float d2a = a[k] - cdy2a;  // volume_cuda.cu:40
          const float d2b = (same_x ? a[k] : tc2[k] - cdx2b) - cdy2b;
          float ala, alb;
          bool na, nb;
          int za = decide_fast(g, d2a, c0, delta, one_m_delta, &ala, &na);
          int zb = decide_fast(g, d2b, c0, delta, one_m_delta, &alb, &nb);
          if (__ballot(na || nb)) {  // rare: confirm near-boundary samples exactly
            if (na) za = decide_exact(g, d2a, wz, 0, &ala);
            if (nb) zb = decide_exact(g, d2b, wz, 0, &alb);
          }
          const bool acta = d2a > 0.f && (unsigned)za <= (unsigned)(g.Z - 2);
          const bool actb = d2b > 0.f && (unsigned)zb <= (unsigned)(g.Z - 2);
          // Inactive lanes read slot 0 and keep their accumulator unchanged
          // (the discarded candidate may be NaN or inf).  Selects measure
          // faster here than exec-masked updates.
          const float2 va = col1[acta ? za : 0];
          const float2 vb = col2[actb ? zb : 0];
          const float fa = fmaf(ala, va.y, va.x);
          const float fb = fmaf(alb, vb.y, vb.x);
          const float qa = d2a * itc4[k];
          const float qb = d2b * itc4[k];
          const float ca = fmaf(fa, qa * qa, acc[k]);
          const float cb = fmaf(fb, qb * qb, acta ? ca : acc[k]);
          acc[k] = actb ? cb : (acta ? ca : acc[k]);
        }
      }
    }

    if (has_next) {
#pragma unroll
      for (int r = 0; r < kPref; ++r) {
        const int e = threadIdx.x + r * nthreads;
        if (e < chunk_vox) nxt[e] = pref[r];
      }
    }
    __syncthreads();
  }

  if (scan_ok) {
    float* row = dst + ((size_t)grp * g.N + i) * g.T;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int t = t_base + 64 * k + lane;
      if (t < g.T) row[t] = acc[k];
    }
  }
}

template <int K>
__global__ __launch_bounds__(kWaves * 64) void r1_forward_kernel(
    const float* __restrict__ vol, const float* __restrict__ mp,
    float* __restrict__ dst, ConfGeom g, int t_base, int n_groups, int chunk_cols,
    int chunks_per_group) {
  extern __shared__ float2 lds2[];
  r1_block<K>(vol, mp, dst, g, t_base, (int)(blockIdx.x % n_groups),
              (int)(blockIdx.x / n_groups), n_groups, chunk_cols, chunks_per_group, lds2);
}

// ---------------------------------------------------------------------------
// A6: gradient of <G, R1(v; w)> w.r.t. the wall points w (volume_cuda.cu:188-265)
// ---------------------------------------------------------------------------
// Same sample set and streaming as R1 (one scan point per wave, one bin per
// lane and window).  Per active sample, with dx = wx - x, f1 the interpolated
// albedo, dv = v[zi+1] - v[zi] and f2 = (depth / tc^2)^4:
//   d/dwx = -dx (4 f1 depth^2 / tc^8 + s dv f2 / depth)   (likewise y)
//   d/dwz = s dv f2
// which is the reference's f1 * df2/dw + df1/dw * f2 with its
// tmp = 4 (-depth / tc^(8/3))^3 = -4 depth^3 / tc^8 written out.  Each lane
// sums its bins over columns, multiplies by G[i][t] once, and the wave
// reduces to one (gx, gy, gz) per scan point, column group and bin pass.
template <int K, typename GT>
__global__ __launch_bounds__(kWaves * 64) void a6_kernel(
    const float* __restrict__ vol, const float* __restrict__ mp,
    const GT* __restrict__ G, float* __restrict__ part,  // [parts][3][N]
    ConfGeom g, int t_base, int pass, int n_groups, int chunk_cols,
    int chunks_per_group) {
  extern __shared__ float2 lds2[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int grp = blockIdx.x % n_groups;
  const int sgrp = blockIdx.x / n_groups;
  const int i = sgrp * kWaves + wave;
  const bool scan_ok = i < g.N;
  const int n_cols = g.X * g.Y;
  const int chunk_vox = chunk_cols * g.Z;

  float tc2[K], itc4[K], a[K], gx[K], gy[K], gz[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int t = t_base + 64 * k + lane;
    const float tc = bin_distance(t, g.w, g.off);
    const float q = tc * tc;
    tc2[k] = (t < g.T) ? q : -1.f;
    itc4[k] = 1.f / (q * q);
    a[k] = 0.f;
    gx[k] = gy[k] = gz[k] = 0.f;
  }
  const int ii = scan_ok ? i : 0;
  const float wx = uniform_f32(mp[ii]);
  const float wy = uniform_f32(mp[g.N + ii]);
  const float wz = uniform_f32(mp[2 * g.N + ii]);
  float c0 = (wz - g.oz) * g.s;
  c0 = c0 - 0.5f;
  const float delta = guard_delta(g, wz, c0, 0.f);
  const float one_m_delta = 1.f - delta;
  float dlo, dhi;
  depth_range(g, wz, 0.f, (float)(g.Z - 1), &dlo, &dhi);

  const int chunk_first = grp * chunks_per_group;
  const int total_chunks = (n_cols + chunk_cols - 1) / chunk_cols;
  const int chunk_end = min(chunk_first + chunks_per_group, total_chunks);
  const int nthreads = kWaves * 64;
  const int vol_vox = n_cols * g.Z;
  auto load_pair = [&](int ge) -> float2 {
    if (ge >= vol_vox) return make_float2(0.f, 0.f);
    const float v0 = vol[ge];
    const bool last = (ge % g.Z) == g.Z - 1;
    const float v1 = last ? v0 : vol[ge + 1];
    return make_float2(v0, v1 - v0);
  };
  for (int c = chunk_first; c < chunk_end; ++c) {
    // Single-buffered: A6 is not on the benchmark path.
    __syncthreads();
    const int base = c * chunk_vox;
    for (int e = threadIdx.x; e < chunk_vox; e += nthreads) lds2[e] = load_pair(base + e);
    __syncthreads();
    if (!scan_ok) continue;
    const int col = c * chunk_cols + lane;
    const bool cvalid = lane < chunk_cols && col < n_cols;
    const int colc = cvalid ? col : 0;
    const int xi = colc / g.Y;
    const int yi = colc - xi * g.Y;
    const float x = cell_centre(xi, g.ox, g.szx, g.X);
    const float y = cell_centre(yi, g.oy, g.szy, g.Y);
    const float dx = wx - x, dy = wy - y;
    const float dx2 = dx * dx, dy2 = dy * dy;
    int tlo, thi;
    bin_range(g, dx2 + dy2, dlo, dhi, &tlo, &thi);
    const int klo = (max(tlo, t_base) - t_base) >> 6;
    const int khi = (min(thi, t_base + 64 * K - 1) - t_base) >> 6;
    const bool any = cvalid && thi >= t_base && tlo <= t_base + 64 * K - 1 && tlo <= thi;
    const unsigned packed = any ? (((2u << khi) - 1u) & ~((1u << klo) - 1u)) : 0u;
    unsigned long long mask = __ballot(any);
    while (mask) {
      const int j = __builtin_ctzll(mask);
      mask &= mask - 1;
      const unsigned wmask = (unsigned)__builtin_amdgcn_readlane((int)packed, j);
      const float cdx = lane_value(dx, j), cdy = lane_value(dy, j);
      const float cdx2 = lane_value(dx2, j), cdy2 = lane_value(dy2, j);
#pragma unroll
      for (int k = 0; k < K; ++k) a[k] = tc2[k] - cdx2;
      const float2* column = lds2 + j * g.Z;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (!((wmask >> k) & 1u)) continue;
        const float d2 = a[k] - cdy2;  // volume_cuda.cu:222
        float al;
        const int zi = confocal_decide(g, d2, wz, c0, 0, delta, one_m_delta, &al);
        if (d2 > 0.f && (unsigned)zi <= (unsigned)(g.Z - 2)) {
          const float2 v = column[zi];
          const float f1 = fmaf(al, v.y, v.x);
          const float i8 = itc4[k] * itc4[k];
          const float f2 = (d2 * d2) * i8;
          const float depth = sqrt_fast(d2);
          const float sdv = g.s * v.y;
          const float cxy = 4.f * f1 * d2 * i8 + sdv * f2 / depth;
          gx[k] = fmaf(-cdx, cxy, gx[k]);
          gy[k] = fmaf(-cdy, cxy, gy[k]);
          gz[k] = fmaf(sdv, f2, gz[k]);
        }
      }
    }
  }
  if (!scan_ok) return;
  float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int t = t_base + 64 * k + lane;
    const float gt = t < g.T ? load_elem(G + (size_t)i * g.T + t) : 0.f;
    sx = fmaf(gx[k], gt, sx);
    sy = fmaf(gy[k], gt, sy);
    sz = fmaf(gz[k], gt, sz);
  }
  sx = wave_sum_f32(sx);
  sy = wave_sum_f32(sy);
  sz = wave_sum_f32(sz);
  if (lane == 0) {
    float* o = part + ((size_t)pass * n_groups + grp) * 3 * g.N;
    o[i] = sx;
    o[g.N + i] = sy;
    o[2 * g.N + i] = sz;
  }
}

// Sum of column-group partials in a fixed order, converted to the storage type.
template <typename OutT>
__global__ void sum_partials_kernel(const float* __restrict__ part, int n_parts,
                                    size_t n, OutT* __restrict__ out) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    float v = part[e];
    for (int p = 1; p < n_parts; ++p) v += part[(size_t)p * n + e];
    store_elem(out + e, v);
  }
}

template <typename T>
__global__ void fill_zero_kernel(T* __restrict__ out, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    store_elem(out + e, 0.f);
}

// ---------------------------------------------------------------------------
// A1
// ---------------------------------------------------------------------------

// Per-bin tables for A1: tc2[t] (exact reference rounding) and, per (i, t),
// Kf[i][t] = G[i][t] / tc^8 so that a sample's weight is d2^2 * Kf.  In the
// `safe` form (tiny radii, where 1/tc^8 could overflow) Kf = G and
// itc4[t] = 1/tc^4 is used as (d2 * itc4)^2 * G instead.
template <typename GT>
__device__ __forceinline__ void a1_table_elem(const GT* __restrict__ G, const ConfGeom& g,
                                              int safe, int t_stride, size_t e,
                                              float* __restrict__ tc2,
                                              float* __restrict__ itc4,
                                              float* __restrict__ Kf) {
  const int t = (int)(e % t_stride);
  const size_t i = e / t_stride;
  const bool in = t < g.T;
  const float tc = bin_distance(t, g.w, g.off);
  const float q = tc * tc;
  const float i4 = 1.f / (q * q);
  if (i == 0) {
    tc2[t] = in ? q : -1.f;
    itc4[t] = in ? i4 : 0.f;
  }
  const float gv = in ? load_elem(G + i * g.T + t) : 0.f;
  Kf[e] = safe ? gv : gv * (i4 * i4);
}

template <typename GT>
__global__ void a1_tables_kernel(const GT* __restrict__ G, ConfGeom g, int safe,
                                 int t_stride, float* __restrict__ tc2,
                                 float* __restrict__ itc4, float* __restrict__ Kf) {
  const size_t n = (size_t)g.N * t_stride;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    a1_table_elem(G, g, safe, t_stride, e, tc2, itc4, Kf);
}

constexpr int kA1Waves = 4;  // independent waves (work items) per A1 workgroup


// One A1 work item, run by one wave on its own LDS accumulator block `acc`
// ((kSlab + 2) rows of 64 floats): the 8x8 column tile (tx, ty), z slab
// `slab` and scan-point range `range` of n_ranges.  `item` enumerates
// (tx, ty, slab, range) with tx fastest.  No workgroup barrier is used.
template <bool SAFE>
__device__ __forceinline__ void a1_item(
    const float* __restrict__ Kf, const float* __restrict__ tc2t,
    const float* __restrict__ itc4t, const float* __restrict__ mp,
    float* __restrict__ dst,  // [R][X][Y][Z] partials, or [X][Y][Z]
    const ConfGeom& g, int item, int tiles_x, int tiles_y, int n_slabs, int n_ranges,
    int t_stride, float* __restrict__ acc) {
  const int lane = threadIdx.x & 63;
  const int tx = item % tiles_x;
  item /= tiles_x;
  const int ty = item % tiles_y;
  item /= tiles_y;
  const int slab = item % n_slabs;
  const int range = item / n_slabs;

  const int xi = tx * 8 + (lane >> 3);
  const int yi = ty * 8 + (lane & 7);
  const bool cvalid = xi < g.X && yi < g.Y;
  const float x = cell_centre(xi, g.ox, g.szx, g.X);
  const float y = cell_centre(yi, g.oy, g.szy, g.Y);

  const int z0 = slab * kSlab;
  // Row r of the accumulator holds voxel z0 - 1 + r; samples with
  // zi in [z0-1, z0+kSlab-1] touch rows zi-z0+1 and zi-z0+2.
  const int rmin = (z0 == 0) ? 1 : 0;                   // zi >= 0
  const int rmax = min(kSlab, (g.Z - 2) - (z0 - 1));    // zi <= Z-2
  const int rspan = rmax - rmin;
  const int zshift = z0 - 1 + rmin;                     // zi = row' + zshift
  const float zlo = (float)max(z0 - 1, 0);
  const float zhi = (float)min(z0 + kSlab, g.Z - 1);

  for (int r = 0; r < kSlab + 2; ++r) acc[r * 64 + lane] = 0.f;
  // Only this lane touches its slots, so no barrier is needed.
  // Two-level sums: the LDS rows hold the sum over a block of kFlush scan
  // points, which is then added to a per-lane register total and cleared,
  // so no voxel's sum is one long FP32 chain over every scan point.  (Rows
  // 0 and kSlab + 1 belong to the neighbouring slabs and are never read.)
  float tot[kSlab];
#pragma unroll
  for (int r = 0; r < kSlab; ++r) tot[r] = 0.f;

  const int per = (g.N + n_ranges - 1) / n_ranges;
  const int i_begin = range * per;
  const int i_end = min(g.N, i_begin + per);
  float* my = acc + rmin * 64 + lane;

  for (int i = i_begin; i < i_end; ++i) {
    if (i > i_begin && (i - i_begin) % kFlush == 0) {
#pragma unroll
      for (int r = 0; r < kSlab; ++r) {
        tot[r] += acc[(r + 1) * 64 + lane];
        acc[(r + 1) * 64 + lane] = 0.f;
      }
    }
    const float wx = mp[i], wy = mp[g.N + i], wz = mp[2 * g.N + i];
    float c0 = (wz - g.oz) * g.s;
    c0 = c0 - 0.5f;
    const float cp = c0 - (float)zshift;
    const float delta = guard_delta(g, wz, c0, cp);
    const float one_m_delta = 1.f - delta;
    const float dx = wx - x, dy = wy - y;
    const float dx2 = cvalid ? dx * dx : INFINITY, dy2 = dy * dy;
    float dlo, dhi;
    depth_range(g, wz, zlo, zhi, &dlo, &dhi);
    int tlo, thi;
    bin_range(g, dx2 + dy2, dlo, dhi, &tlo, &thi);
    if (!cvalid || rspan < 0 || tlo > thi) {
      tlo = 0x7fffffff;
      thi = -1;
    }
    const int t_begin = uniform_i32(wave_min_i32(tlo));
    const int t_end = uniform_i32(wave_max_i32(thi));
    const float* Ki = Kf + (size_t)i * t_stride;
    // Bins are visited in batches of 64: one coalesced vector load per table
    // brings a batch into registers (lane j holds bin tb + j), and each bin's
    // value is broadcast with readlane.  Tables are padded to a multiple of
    // 64 bins (K = 0, tc2 = -1 beyond T).
    for (int tb = t_begin & ~63; tb <= t_end; tb += 64) {
      const float kreg = Ki[tb + lane];
      const float treg = tc2t[tb + lane];
      const float ireg = SAFE ? itc4t[tb + lane] : 0.f;
      const int u0 = max(t_begin - tb, 0);
      const int u1 = min(t_end - tb, 63);
      // Two bins per pass: the geometry of both is computed before either
      // accumulator update (their updates may touch the same rows, so the
      // read-modify-writes stay in order).
      for (int u = u0; u <= u1; u += 2) {
        const int ub = min(u + 1, u1);
        const float kta = lane_value(kreg, u);
        const float ktb = (u + 1 <= u1) ? lane_value(kreg, ub) : 0.f;
        const float d2a = (lane_value(treg, u) - dx2) - dy2;
        const float d2b = (lane_value(treg, ub) - dx2) - dy2;
        float ala, alb;
        bool na, nb;
        int ra = decide_fast(g, d2a, cp, delta, one_m_delta, &ala, &na);
        int rb = decide_fast(g, d2b, cp, delta, one_m_delta, &alb, &nb);
        if (__ballot(na || nb)) {  // rare: confirm near-boundary samples exactly
          if (na) ra = decide_exact(g, d2a, wz, zshift, &ala);
          if (nb) rb = decide_exact(g, d2b, wz, zshift, &alb);
        }
        // dx2 is +inf for lanes without a column, so d2 > 0 also checks that.
        const bool acta = d2a > 0.f && (unsigned)ra <= (unsigned)rspan;
        const bool actb = d2b > 0.f && (unsigned)rb <= (unsigned)rspan;
        // Each lane owns its slots and a wave's LDS accesses complete in
        // order, so plain read-modify-writes are exact.  (LDS float atomics
        // measure ~25x slower than this on gfx950.)  Updates run under the
        // activity mask (exec).
        if (acta) {
          const float wa = SAFE ? sq(d2a * lane_value(ireg, u)) * kta : (d2a * d2a) * kta;
          float* pa = my + ra * 64;
          const float a0 = pa[0], a1 = pa[64];
          pa[0] = fmaf(wa, 1.f - ala, a0);
          pa[64] = fmaf(wa, ala, a1);
        }
        if (actb) {
          const float wb = SAFE ? sq(d2b * lane_value(ireg, ub)) * ktb : (d2b * d2b) * ktb;
          float* pb = my + rb * 64;
          const float b0 = pb[0], b1 = pb[64];
          pb[0] = fmaf(wb, 1.f - alb, b0);
          pb[64] = fmaf(wb, alb, b1);
        }
      }
    }
  }

  if (cvalid) {
    float* out = dst + (((size_t)range * g.X + xi) * g.Y + yi) * g.Z;
#pragma unroll
    for (int r = 0; r < kSlab; ++r) {
      const int z = z0 + r;
      if (z < g.Z) out[z] = tot[r] + acc[(r + 1) * 64 + lane];
    }
  }
}

template <bool SAFE>
__global__ __launch_bounds__(kA1Waves * 64) void a1_transpose_kernel(
    const float* __restrict__ Kf, const float* __restrict__ tc2t,
    const float* __restrict__ itc4t, const float* __restrict__ mp,
    float* __restrict__ dst, ConfGeom g, int tiles_x, int tiles_y, int n_slabs,
    int n_ranges, int t_stride, int n_items) {
  __shared__ float acc_all[kA1Waves][(kSlab + 2) * 64];
  const int wave = threadIdx.x >> 6;
  const int item = blockIdx.x * kA1Waves + wave;
  if (item >= n_items) return;  // whole wave exits; no barrier follows
  a1_item<SAFE>(Kf, tc2t, itc4t, mp, dst, g, item, tiles_x, tiles_y, n_slabs, n_ranges,
                t_stride, acc_all[wave]);
}

int launch_grid(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (int)b;
}

int check_common(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                 const float* origin, const float* size, float w, float off,
                 int dtype) {
  TOTRI_CHECK_ARG(X >= 0 && Y >= 0 && Z >= 0 && N >= 0 && T >= 0,
                  "negative dimension");
  TOTRI_CHECK_ARG(X < (1ll << 31) && Y < (1ll << 31) && Z < (1ll << 31) &&
                      N < (1ll << 31) && T < (1ll << 31),
                  "dimension exceeds 32-bit indexing");
  TOTRI_CHECK_ARG(X * Y < (1ll << 31) && X * Y * Z < (1ll << 31),
                  "volume has more than 2^31 elements");
  TOTRI_CHECK_ARG(N * T < (1ll << 31), "transient has more than 2^31 elements");
  TOTRI_CHECK_ARG(origin != nullptr && size != nullptr, "null geometry");
  TOTRI_CHECK_ARG(dtype == TOTRI_F32 || dtype == TOTRI_BF16 || dtype == TOTRI_F16,
                  "unsupported transient dtype %d", dtype);
  (void)w;
  (void)off;
  return TOTRI_OK;
}

ConfGeom make_geom(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                   const float* origin, const float* size, float w, float off) {
  ConfGeom g;
  g.X = (int)X;
  g.Y = (int)Y;
  g.Z = (int)Z;
  g.N = (int)N;
  g.T = (int)T;
  g.ox = origin[0];
  g.oy = origin[1];
  g.oz = origin[2];
  g.szx = size[0];
  g.szy = size[1];
  g.szz = size[2];
  g.w = w;
  g.off = off;
  g.s = (float)Z / size[2];
  // Largest |tc| over bins, with the reference rounding.
  float a = fabsf(off), b = fabsf((float)(T > 0 ? T - 1 : 0) * w + off);
  g.tc_max = a > b ? a : b;
  return g;
}

template <typename OutT>
int zero_fill(OutT* out, size_t n, hipStream_t st) {
  if (n == 0) return TOTRI_OK;
  hipLaunchKernelGGL(fill_zero_kernel<OutT>, dim3(launch_grid(n)), dim3(256), 0, st,
                     out, n);
  return check_launch("fill_zero_kernel");
}

struct R1Plan {
  int K, passes, n_groups, chunk_cols, chunks_per_group, total_chunks;
  size_t ws;
};

R1Plan plan_r1(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T) {
  R1Plan p;
  const int64_t bins_pass = T <= 64 ? 64 : T <= 128 ? 128 : T <= 256 ? 256 : T <= 512 ? 512 : 1024;
  p.K = (int)(bins_pass / 64);
  p.passes = (int)((T + bins_pass - 1) / bins_pass);
  p.chunk_cols = (int)(Z > 0 ? (kChunkVoxels / Z < 64 ? kChunkVoxels / Z : 64) : 64);
  if (p.chunk_cols < 1) p.chunk_cols = 1;
  const int64_t n_cols = X * Y;
  p.total_chunks = (int)((n_cols + p.chunk_cols - 1) / p.chunk_cols);
  const int64_t scan_groups = (N + kWaves - 1) / kWaves;
  const int64_t target = 2 * (int64_t)device_cu_count();
  int64_t G = 1;
  if (scan_groups < target) G = (target + scan_groups - 1) / scan_groups;
  if (G > p.total_chunks) G = p.total_chunks;
  if (G < 1) G = 1;
  p.chunks_per_group = (int)((p.total_chunks + G - 1) / G);
  p.n_groups = (int)((p.total_chunks + p.chunks_per_group - 1) / p.chunks_per_group);
  // FP32 staging of the partial sums (also needed with one group when the
  // transient is stored in a 16-bit type).
  p.ws = (size_t)p.n_groups * N * T * sizeof(float);
  return p;
}

size_t r1_workspace(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T) {
  if (X <= 0 || Y <= 0 || Z < 2 || N <= 0 || T <= 0) return 0;
  return plan_r1(X, Y, Z, N, T).ws;
}

template <int K>
void launch_r1(const R1Plan& p, const float* vol, const float* mp, float* dst,
               const ConfGeom& g, int t_base, hipStream_t st) {
  const int scan_groups = (g.N + kWaves - 1) / kWaves;
  const int blocks = scan_groups * p.n_groups;
  const size_t lds = 2 * (size_t)p.chunk_cols * g.Z * sizeof(float2);
  hipLaunchKernelGGL(r1_forward_kernel<K>, dim3(blocks), dim3(kWaves * 64), lds, st,
                     vol, mp, dst, g, t_base, p.n_groups, p.chunk_cols,
                     p.chunks_per_group);
}

template <typename OutT>
int finish_r1(const R1Plan& p, const float* staging, size_t n, OutT* out,
              hipStream_t st) {
  hipLaunchKernelGGL(sum_partials_kernel<OutT>, dim3(launch_grid(n)), dim3(256), 0, st,
                     staging, p.n_groups, n, out);
  return check_launch("sum_partials_kernel");
}

}  // namespace
}  // namespace totri_hip

using namespace totri_hip;

extern "C" size_t totri_volume_confocal_forward_workspace(int64_t X, int64_t Y,
                                                          int64_t Z, int64_t N,
                                                          int64_t T) {
  return r1_workspace(X, Y, Z, N, T);
}

extern "C" int totri_volume_confocal_forward(
    const float* volume, int64_t X, int64_t Y, int64_t Z,
    const float* measurement_points, int64_t N, void* transient, int64_t T,
    int transient_dtype, const float volume_origin[3], const float volume_size[3],
    float bin_width, float bin_offset, void* workspace, size_t workspace_bytes,
    void* stream) {
  int rc = check_common(X, Y, Z, N, T, volume_origin, volume_size, bin_width,
                        bin_offset, transient_dtype);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const size_t n = (size_t)N * T;
  if (n == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(transient != nullptr && measurement_points != nullptr,
                  "null transient or measurement points");
  if (X == 0 || Y == 0 || Z < 2) {
    // No sample can satisfy zi >= 0 && zi + 1 < Z (volume_cuda.cu:49).
    switch (transient_dtype) {
      case TOTRI_F32: return zero_fill((float*)transient, n, st);
      case TOTRI_BF16: return zero_fill((__hip_bfloat16*)transient, n, st);
      default: return zero_fill((__half*)transient, n, st);
    }
  }
  TOTRI_CHECK_ARG(volume != nullptr, "null volume");
  TOTRI_CHECK_ARG(Z <= kChunkVoxels, "volume depth Z=%lld exceeds %d", (long long)Z,
                  kChunkVoxels);
  const R1Plan p = plan_r1(X, Y, Z, N, T);
  const size_t need = r1_workspace(X, Y, Z, N, T);
  const bool direct = p.n_groups == 1 && transient_dtype == TOTRI_F32;
  if (!direct && workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu",
                need, workspace_bytes);
  if (!direct) TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  const ConfGeom g = make_geom(X, Y, Z, N, T, volume_origin, volume_size,
                               bin_width, bin_offset);
  float* dst = direct ? (float*)transient : (float*)workspace;
  for (int pass = 0; pass < p.passes; ++pass) {
    const int t_base = pass * 64 * p.K;
    switch (p.K) {
      case 1: launch_r1<1>(p, volume, measurement_points, dst, g, t_base, st); break;
      case 2: launch_r1<2>(p, volume, measurement_points, dst, g, t_base, st); break;
      case 4: launch_r1<4>(p, volume, measurement_points, dst, g, t_base, st); break;
      case 8: launch_r1<8>(p, volume, measurement_points, dst, g, t_base, st); break;
      default: launch_r1<16>(p, volume, measurement_points, dst, g, t_base, st); break;
    }
    rc = check_launch("r1_forward_kernel");
    if (rc) return rc;
  }
  if (direct) return TOTRI_OK;
  switch (transient_dtype) {
    case TOTRI_F32: return finish_r1(p, dst, n, (float*)transient, st);
    case TOTRI_BF16: return finish_r1(p, dst, n, (__hip_bfloat16*)transient, st);
    default: return finish_r1(p, dst, n, (__half*)transient, st);
  }
}

namespace totri_hip {
namespace {

struct A1Plan {
  int tiles_x, tiles_y, n_slabs, n_ranges, t_stride;
  bool safe;
  size_t tables_bytes, partial_bytes;
};

A1Plan plan_a1(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T, float w,
               float off) {
  A1Plan p;
  p.tiles_x = (int)((X + 7) / 8);
  p.tiles_y = (int)((Y + 7) / 8);
  p.n_slabs = (int)((Z + kSlab - 1) / kSlab);
  const int64_t items = (int64_t)p.tiles_x * p.tiles_y * p.n_slabs;
  const int64_t target = 16 * (int64_t)device_cu_count();
  int64_t R = 1;
  if (items < target) R = (target + items - 1) / items;
  if (R > N) R = N;
  if (R < 1) R = 1;
  p.n_ranges = (int)R;
  // Smallest nonzero |tc| decides whether 1/tc^8 may overflow.
  float tmin = INFINITY;
  for (int64_t t = 0; t < T && t < 4; ++t) {
    const float tc = fabsf((float)t * w + off);
    if (tc > 0.f && tc < tmin) tmin = tc;
  }
  if (w != 0.f) {
    // Where the sign of tc changes, |tc| may be small anywhere in the range.
    const float t0 = -off / w;
    if (t0 > 0.f && t0 < (float)T) {
      const int64_t tl = (int64_t)floorf(t0);
      for (int64_t t = tl - 1; t <= tl + 2; ++t)
        if (t >= 0 && t < T) {
          const float tc = fabsf((float)t * w + off);
          if (tc > 0.f && tc < tmin) tmin = tc;
        }
    }
  }
  // K = G / tc^8 stays finite for |G| < 3e38 * tc^8: with tc >= 1e-3 m that
  // is |G| < 3e14, far above any transient gradient.
  p.safe = !(tmin >= 1e-3f);
  p.t_stride = (int)((T + 63) & ~(int64_t)63);
  p.tables_bytes = ((size_t)2 * p.t_stride + (size_t)N * p.t_stride) * sizeof(float);
  p.tables_bytes = (p.tables_bytes + 255) & ~(size_t)255;
  p.partial_bytes = p.n_ranges > 1 ? (size_t)p.n_ranges * X * Y * Z * sizeof(float) : 0;
  return p;
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_volume_confocal_grad_volume_workspace(int64_t X, int64_t Y,
                                                              int64_t Z, int64_t N,
                                                              int64_t T) {
  if (X <= 0 || Y <= 0 || Z < 2 || N <= 0 || T <= 0) return 0;
  // The plan's size does not depend on the geometry, only on the safe flag's
  // table layout, which is the same in both forms.
  A1Plan p = plan_a1(X, Y, Z, N, T, 1.f, 1.f);
  return p.tables_bytes + p.partial_bytes;
}

extern "C" int totri_volume_confocal_grad_volume(
    float* volume, int64_t X, int64_t Y, int64_t Z, const float* measurement_points,
    int64_t N, const void* transient, int64_t T, int transient_dtype,
    const float volume_origin[3], const float volume_size[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_common(X, Y, Z, N, T, volume_origin, volume_size, bin_width,
                        bin_offset, transient_dtype);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const size_t nv = (size_t)X * Y * Z;
  if (nv == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr, "null volume");
  if (N == 0 || T == 0 || Z < 2) return zero_fill(volume, nv, st);
  TOTRI_CHECK_ARG(transient != nullptr && measurement_points != nullptr,
                  "null transient or measurement points");
  const A1Plan p = plan_a1(X, Y, Z, N, T, bin_width, bin_offset);
  const size_t need = p.tables_bytes + p.partial_bytes;
  if (workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu",
                need, workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  const ConfGeom g = make_geom(X, Y, Z, N, T, volume_origin, volume_size,
                               bin_width, bin_offset);
  float* tc2 = (float*)workspace;
  float* itc4 = tc2 + p.t_stride;
  float* Kf = itc4 + p.t_stride;
  float* partial = (float*)((char*)workspace + p.tables_bytes);
  const size_t ntab = (size_t)N * p.t_stride;
  switch (transient_dtype) {
    case TOTRI_F32:
      hipLaunchKernelGGL(a1_tables_kernel<float>, dim3(launch_grid(ntab)), dim3(256), 0,
                         st, (const float*)transient, g, (int)p.safe, p.t_stride, tc2, itc4, Kf);
      break;
    case TOTRI_BF16:
      hipLaunchKernelGGL(a1_tables_kernel<__hip_bfloat16>, dim3(launch_grid(ntab)),
                         dim3(256), 0, st, (const __hip_bfloat16*)transient, g,
                         (int)p.safe, p.t_stride, tc2, itc4, Kf);
      break;
    default:
      hipLaunchKernelGGL(a1_tables_kernel<__half>, dim3(launch_grid(ntab)), dim3(256), 0,
                         st, (const __half*)transient, g, (int)p.safe, p.t_stride, tc2, itc4, Kf);
      break;
  }
  rc = check_launch("a1_tables_kernel");
  if (rc) return rc;
  float* dst = p.n_ranges > 1 ? partial : volume;
  const int items = p.tiles_x * p.tiles_y * p.n_slabs * p.n_ranges;
  const int blocks = (items + kA1Waves - 1) / kA1Waves;
  if (p.safe)
    hipLaunchKernelGGL(a1_transpose_kernel<true>, dim3(blocks), dim3(kA1Waves * 64), 0, st,
                       Kf, tc2, itc4, measurement_points, dst, g, p.tiles_x, p.tiles_y,
                       p.n_slabs, p.n_ranges, p.t_stride, items);
  else
    hipLaunchKernelGGL(a1_transpose_kernel<false>, dim3(blocks), dim3(kA1Waves * 64), 0,
                       st, Kf, tc2, itc4, measurement_points, dst, g, p.tiles_x,
                       p.tiles_y, p.n_slabs, p.n_ranges, p.t_stride, items);
  rc = check_launch("a1_transpose_kernel");
  if (rc) return rc;
  if (p.n_ranges > 1) {
    hipLaunchKernelGGL(sum_partials_kernel<float>, dim3(launch_grid(nv)), dim3(256), 0,
                       st, partial, p.n_ranges, nv, volume);
    rc = check_launch("sum_partials_kernel");
  }
  return rc;
}

// ---------------------------------------------------------------------------
// Batched step: many R1 / A1 jobs in one set of persistent launches
// ---------------------------------------------------------------------------
namespace totri_hip {
namespace {

constexpr int kBatchMax = TOTRI_BATCH_MAX_JOBS;
constexpr int kClassA1 = 5;  // classes 0..4: R1 with K = 1, 2, 4, 8, 16
constexpr int kClasses = 7;  // 5: A1, 6: A1 in the safe (tiny radius) form
constexpr unsigned kBatchMagic = 0x74427431u;

int r1_class(int K) { return K == 1 ? 0 : K == 2 ? 1 : K == 4 ? 2 : K == 8 ? 3 : 4; }

// One job as the kernels see it: geometry, plan and buffer addresses.
struct BatchJob {
  ConfGeom g;
  int kind;
  int K, passes, scan_groups, n_groups, chunk_cols, chunks_per_group;
  int tiles_x, tiles_y, n_slabs, n_ranges, t_stride, safe;
  const float* vol_in;
  const float* mp;
  const float* tr_in;
  float* r1_dst;
  float* tc2;
  float* itc4;
  float* kf;
  float* a1_dst;
  const float* fin_src;
  float* fin_dst;
  long long fin_n;
  int fin_parts;
  int pad_;
};

// One step.  The work-queue heads come first, in a 256-byte block of their
// own, and are zero in every table that batch_prepare writes, so uploading
// the table resets them.  Per class: the jobs with work and the prefix sums
// of their item counts; likewise the A1 table elements and the partial-sum
// elements of the finishing pass.
struct BatchTable {
  int counters[kClasses];
  unsigned magic;
  int n_jobs;
  int pad_[64 - kClasses - 2];
  int cls_n[kClasses];
  int cls_jobs[kClasses][kBatchMax];
  int cls_prefix[kClasses][kBatchMax + 1];
  int tab_n;
  int tab_jobs[kBatchMax];
  long long tab_prefix[kBatchMax + 1];
  int fin_n;
  int fin_jobs[kBatchMax];
  long long fin_prefix[kBatchMax + 1];
  BatchJob jobs[kBatchMax];
};

// Largest slot j in [0, n) with prefix[j] <= e (prefix is non-decreasing and
// every listed slot is non-empty, so this is the slot that holds e).
template <typename IT>
__device__ __forceinline__ int find_slot(const IT* prefix, int n, IT e) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Persistent R1 over the class-`cls` jobs (all with K windows).  Each
// workgroup takes the next item from the class queue until it is empty; the
// item index goes through LDS slot 0, which every thread has read before
// r1_block refills the chunk buffers.
// The table is read through a read-only view (job fields become scalar
// loads) and the queue heads are updated through `counters`, which points at
// the same table's counter block (the view never reads it).
template <int K>
__global__ __launch_bounds__(kWaves * 64) void r1_batch_kernel(
    const BatchTable* __restrict__ tab, int* __restrict__ counters, int cls) {
  extern __shared__ float2 lds2[];
  const int n = tab->cls_n[cls];
  const int total = tab->cls_prefix[cls][n];
  int* slot0 = reinterpret_cast<int*>(lds2);
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) slot0[0] = atomicAdd(&counters[cls], 1);
    __syncthreads();
    const int item = uniform_i32(slot0[0]);
    __syncthreads();
    if (item >= total) break;
    const int j = find_slot(tab->cls_prefix[cls], n, item);
    const BatchJob* jb = &tab->jobs[tab->cls_jobs[cls][j]];
    const ConfGeom g = jb->g;
    const int n_groups = jb->n_groups, scan_groups = jb->scan_groups;
    int local = item - tab->cls_prefix[cls][j];
    const int grp = local % n_groups;
    local /= n_groups;
    const int sgrp = local % scan_groups;
    const int pass = local / scan_groups;
    r1_block<K>(jb->vol_in, jb->mp, jb->r1_dst, g, pass * 64 * K, grp, sgrp, n_groups,
                jb->chunk_cols, jb->chunks_per_group, lds2);
  }
}

// Persistent A1 over the class-`cls` jobs: every wave takes its own items.
template <bool SAFE>
__global__ __launch_bounds__(kA1Waves * 64) void a1_batch_kernel(
    const BatchTable* __restrict__ tab, int* __restrict__ counters, int cls) {
  __shared__ float acc_all[kA1Waves][(kSlab + 2) * 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int n = tab->cls_n[cls];
  const int total = tab->cls_prefix[cls][n];
  for (;;) {
    int item = 0;
    if (lane == 0) item = atomicAdd(&counters[cls], 1);
    item = uniform_i32(__shfl(item, 0, 64));
    if (item >= total) break;
    const int j = find_slot(tab->cls_prefix[cls], n, item);
    const BatchJob* jb = &tab->jobs[tab->cls_jobs[cls][j]];
    const ConfGeom g = jb->g;
    a1_item<SAFE>(jb->kf, jb->tc2, jb->itc4, jb->mp, jb->a1_dst, g,
                  item - tab->cls_prefix[cls][j], jb->tiles_x, jb->tiles_y, jb->n_slabs,
                  jb->n_ranges, jb->t_stride, acc_all[wave]);
  }
}

__global__ __launch_bounds__(256) void a1_tables_batch_kernel(const BatchTable* tab) {
  const int n = tab->tab_n;
  const long long total = tab->tab_prefix[n];
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int j = find_slot(tab->tab_prefix, n, e);
    const BatchJob* jb = &tab->jobs[tab->tab_jobs[j]];
    a1_table_elem(jb->tr_in, jb->g, jb->safe, jb->t_stride,
                  (size_t)(e - tab->tab_prefix[j]), jb->tc2, jb->itc4, jb->kf);
  }
}

// Partial sums in the same fixed order as sum_partials_kernel.
__global__ __launch_bounds__(256) void finish_batch_kernel(const BatchTable* tab) {
  const int n = tab->fin_n;
  const long long total = tab->fin_prefix[n];
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int j = find_slot(tab->fin_prefix, n, e);
    const BatchJob* jb = &tab->jobs[tab->fin_jobs[j]];
    const size_t le = (size_t)(e - tab->fin_prefix[j]);
    const size_t m = (size_t)jb->fin_n;
    float v = jb->fin_src[le];
    for (int p = 1; p < jb->fin_parts; ++p) v += jb->fin_src[(size_t)p * m + le];
    jb->fin_dst[le] = v;
  }
}

// Validates one job and fills its kernel-side description; *need receives
// the workspace it needs.  With check_buffers false only the sizes are
// validated (the workspace query, made before buffers exist).
int plan_job(const totri_confocal_job& j, BatchJob* out, size_t* need, bool check_buffers) {
  *need = 0;
  TOTRI_CHECK_ARG(j.kind == TOTRI_JOB_RENDER || j.kind == TOTRI_JOB_BACKPROJECT,
                  "unknown job kind %d", j.kind);
  int rc = check_common(j.X, j.Y, j.Z, j.N, j.T, j.volume_origin, j.volume_size,
                        j.bin_width, j.bin_offset, TOTRI_F32);
  if (rc) return rc;
  TOTRI_CHECK_ARG(j.X >= 1 && j.Y >= 1 && j.Z >= 2 && j.N >= 1 && j.T >= 1,
                  "batched jobs need X, Y, N, T >= 1 and Z >= 2");
  TOTRI_CHECK_ARG(j.Z <= kChunkVoxels, "volume depth Z=%lld exceeds %d", (long long)j.Z,
                  kChunkVoxels);
  if (check_buffers)
    TOTRI_CHECK_ARG(j.volume && j.measurement_points && j.transient, "null job buffer");
  BatchJob b;
  memset(&b, 0, sizeof(b));
  b.g = make_geom(j.X, j.Y, j.Z, j.N, j.T, j.volume_origin, j.volume_size, j.bin_width,
                  j.bin_offset);
  b.kind = j.kind;
  b.mp = j.measurement_points;
  float* ws = (float*)j.workspace;
  if (j.kind == TOTRI_JOB_RENDER) {
    const R1Plan p = plan_r1(j.X, j.Y, j.Z, j.N, j.T);
    b.K = p.K;
    b.passes = p.passes;
    b.scan_groups = (int)((j.N + kWaves - 1) / kWaves);
    b.n_groups = p.n_groups;
    b.chunk_cols = p.chunk_cols;
    b.chunks_per_group = p.chunks_per_group;
    b.vol_in = j.volume;
    *need = p.n_groups > 1 ? p.ws : 0;
    b.r1_dst = p.n_groups > 1 ? ws : j.transient;
    if (p.n_groups > 1) {
      b.fin_src = ws;
      b.fin_dst = j.transient;
      b.fin_n = (long long)j.N * j.T;
      b.fin_parts = p.n_groups;
    }
  } else {
    const A1Plan p = plan_a1(j.X, j.Y, j.Z, j.N, j.T, j.bin_width, j.bin_offset);
    b.tiles_x = p.tiles_x;
    b.tiles_y = p.tiles_y;
    b.n_slabs = p.n_slabs;
    b.n_ranges = p.n_ranges;
    b.t_stride = p.t_stride;
    b.safe = p.safe ? 1 : 0;
    b.tr_in = j.transient;
    *need = p.tables_bytes + p.partial_bytes;
    b.tc2 = ws;
    b.itc4 = ws ? ws + p.t_stride : nullptr;
    b.kf = ws ? ws + 2 * (size_t)p.t_stride : nullptr;
    float* partial = ws ? (float*)((char*)ws + p.tables_bytes) : nullptr;
    b.a1_dst = p.n_ranges > 1 ? partial : j.volume;
    if (p.n_ranges > 1) {
      b.fin_src = partial;
      b.fin_dst = j.volume;
      b.fin_n = (long long)j.X * j.Y * j.Z;
      b.fin_parts = p.n_ranges;
    }
  }
  if (*need > 0 && check_buffers) {
    TOTRI_CHECK_ARG(j.workspace != nullptr, "null job workspace");
    if (j.workspace_bytes < *need)
      return fail(TOTRI_EWORKSPACE, "job workspace too small: need %zu bytes, got %zu",
                  *need, j.workspace_bytes);
  }
  *out = b;
  return TOTRI_OK;
}

long long job_items(const BatchJob& b) {
  if (b.kind == TOTRI_JOB_RENDER) return (long long)b.passes * b.scan_groups * b.n_groups;
  return (long long)b.tiles_x * b.tiles_y * b.n_slabs * b.n_ranges;
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_confocal_batch_table_bytes(void) { return sizeof(BatchTable); }

extern "C" size_t totri_confocal_job_workspace(const totri_confocal_job* job) {
  if (!job) return 0;
  BatchJob b;
  size_t need = 0;
  if (plan_job(*job, &b, &need, false) != TOTRI_OK) return 0;
  return need;
}

extern "C" int totri_confocal_batch_prepare(const totri_confocal_job* jobs, int n_jobs,
                                            void* host_table, size_t table_bytes) {
  TOTRI_CHECK_ARG(host_table != nullptr, "null host table");
  TOTRI_CHECK_ARG(table_bytes >= sizeof(BatchTable), "table needs %zu bytes, got %zu",
                  sizeof(BatchTable), table_bytes);
  TOTRI_CHECK_ARG(n_jobs >= 0 && n_jobs <= kBatchMax, "n_jobs=%d outside [0, %d]", n_jobs,
                  kBatchMax);
  TOTRI_CHECK_ARG(n_jobs == 0 || jobs != nullptr, "null jobs");
  BatchTable* t = (BatchTable*)host_table;
  memset(t, 0, sizeof(BatchTable));
  for (int k = 0; k < n_jobs; ++k) {
    size_t need = 0;
    int rc = plan_job(jobs[k], &t->jobs[k], &need, true);
    if (rc) {
      memset(t, 0, sizeof(BatchTable));
      return rc;
    }
  }
  long long cls_total[kClasses] = {0};
  for (int k = 0; k < n_jobs; ++k) {
    const BatchJob& b = t->jobs[k];
    const long long items = job_items(b);
    const int c = b.kind == TOTRI_JOB_RENDER ? r1_class(b.K) : kClassA1 + b.safe;
    if (items > 0) {
      TOTRI_CHECK_ARG(cls_total[c] + items < (1ll << 30), "too many work items in one step");
      const int m = t->cls_n[c]++;
      t->cls_jobs[c][m] = k;
      t->cls_prefix[c][m] = (int)cls_total[c];
      cls_total[c] += items;
      t->cls_prefix[c][m + 1] = (int)cls_total[c];
    }
    if (b.kind == TOTRI_JOB_BACKPROJECT) {
      const int m = t->tab_n++;
      t->tab_jobs[m] = k;
      t->tab_prefix[m + 1] = t->tab_prefix[m] + (long long)b.g.N * b.t_stride;
    }
    if (b.fin_parts > 1) {
      const int m = t->fin_n++;
      t->fin_jobs[m] = k;
      t->fin_prefix[m + 1] = t->fin_prefix[m] + b.fin_n;
    }
  }
  t->n_jobs = n_jobs;
  t->magic = kBatchMagic;
  return TOTRI_OK;
}

extern "C" int totri_confocal_batch_run(const void* host_table, void* device_table,
                                        size_t table_bytes, void* stream) {
  TOTRI_CHECK_ARG(host_table != nullptr && device_table != nullptr, "null table");
  TOTRI_CHECK_ARG(table_bytes >= sizeof(BatchTable), "table needs %zu bytes, got %zu",
                  sizeof(BatchTable), table_bytes);
  TOTRI_CHECK_ARG(((const BatchTable*)host_table)->magic == kBatchMagic,
                  "host table was not written by totri_confocal_batch_prepare");
  hipStream_t st = (hipStream_t)stream;
  TOTRI_CHECK_HIP(hipMemcpyAsync(device_table, host_table, sizeof(BatchTable),
                                 hipMemcpyHostToDevice, st));
  const BatchTable* tab = (const BatchTable*)device_table;
  int* ctr = ((BatchTable*)device_table)->counters;
  const int cu = device_cu_count();
  const size_t r1_lds = 2 * (size_t)kChunkVoxels * sizeof(float2);
  hipLaunchKernelGGL(a1_tables_batch_kernel, dim3(4 * cu), dim3(256), 0, st, tab);
  int rc = check_launch("a1_tables_batch_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(r1_batch_kernel<16>, dim3(2 * cu), dim3(kWaves * 64), r1_lds, st, tab, ctr, 4);
  hipLaunchKernelGGL(r1_batch_kernel<8>, dim3(2 * cu), dim3(kWaves * 64), r1_lds, st, tab, ctr, 3);
  hipLaunchKernelGGL(r1_batch_kernel<4>, dim3(2 * cu), dim3(kWaves * 64), r1_lds, st, tab, ctr, 2);
  hipLaunchKernelGGL(r1_batch_kernel<2>, dim3(2 * cu), dim3(kWaves * 64), r1_lds, st, tab, ctr, 1);
  hipLaunchKernelGGL(r1_batch_kernel<1>, dim3(2 * cu), dim3(kWaves * 64), r1_lds, st, tab, ctr, 0);
  rc = check_launch("r1_batch_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(a1_batch_kernel<false>, dim3(4 * cu), dim3(kA1Waves * 64), 0, st, tab,
                     ctr, kClassA1);
  hipLaunchKernelGGL(a1_batch_kernel<true>, dim3(4 * cu), dim3(kA1Waves * 64), 0, st, tab,
                     ctr, kClassA1 + 1);
  rc = check_launch("a1_batch_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(finish_batch_kernel, dim3(4 * cu), dim3(256), 0, st, tab);
  return check_launch("finish_batch_kernel");
}

namespace totri_hip {
namespace {

__global__ void confocal_samples_kernel(const float* __restrict__ mp,
                                        const int* __restrict__ idx, int n, ConfGeom g,
                                        int* __restrict__ zi_out,
                                        float* __restrict__ alpha_out,
                                        float* __restrict__ d2_out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = idx[4 * k], t = idx[4 * k + 1], xi = idx[4 * k + 2], yi = idx[4 * k + 3];
  if (i < 0 || i >= g.N || t < 0 || t >= g.T || xi < 0 || xi >= g.X || yi < 0 ||
      yi >= g.Y) {
    zi_out[k] = -2000000;
    alpha_out[k] = 0.f;
    d2_out[k] = 0.f;
    return;
  }
  const float wx = mp[i], wy = mp[g.N + i], wz = mp[2 * g.N + i];
  float c0 = (wz - g.oz) * g.s;
  c0 = c0 - 0.5f;
  const float delta = guard_delta(g, wz, c0, 0.f);
  const float tc = bin_distance(t, g.w, g.off);
  const float x = cell_centre(xi, g.ox, g.szx, g.X);
  const float y = cell_centre(yi, g.oy, g.szy, g.Y);
  const float dx = wx - x, dy = wy - y;
  const float a = tc * tc - dx * dx;
  const float d2 = a - dy * dy;
  float al;
  const int zi = confocal_decide(g, d2, wz, c0, 0, delta, 1.f - delta, &al);
  d2_out[k] = d2;
  zi_out[k] = d2 > 0.f ? zi : -1000000;
  alpha_out[k] = d2 > 0.f ? al : 0.f;
}

}  // namespace
}  // namespace totri_hip

extern "C" int totri_debug_confocal_samples(
    const float* measurement_points, int64_t N, const int32_t* idx, int64_t n,
    int64_t X, int64_t Y, int64_t Z, int64_t T, const float volume_origin[3],
    const float volume_size[3], float bin_width, float bin_offset, int32_t* zi_out,
    float* alpha_out, float* d2_out, void* stream) {
  int rc = check_common(X, Y, Z, N, T, volume_origin, volume_size, bin_width,
                        bin_offset, TOTRI_F32);
  if (rc) return rc;
  TOTRI_CHECK_ARG(n >= 0 && n < (1ll << 31), "bad sample count");
  if (n == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(measurement_points && idx && zi_out && alpha_out && d2_out,
                  "null pointer");
  // Every listed bin index must be < T (T fixes the guard's bound on |tc|).
  ConfGeom g = make_geom(X, Y, Z, N, T, volume_origin, volume_size, bin_width,
                         bin_offset);
  hipLaunchKernelGGL(confocal_samples_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, (hipStream_t)stream, measurement_points, (const int*)idx, (int)n, g,
                     (int*)zi_out, alpha_out, d2_out);
  return check_launch("confocal_samples_kernel");
}

namespace totri_hip {
namespace {

template <int K, typename GT>
void launch_a6(const R1Plan& p, const float* vol, const float* mp, const GT* G, float* part,
               const ConfGeom& g, int t_base, int pass, hipStream_t st) {
  const int scan_groups = (g.N + kWaves - 1) / kWaves;
  const size_t lds = (size_t)p.chunk_cols * g.Z * sizeof(float2);
  hipLaunchKernelGGL((a6_kernel<K, GT>), dim3(scan_groups * p.n_groups), dim3(kWaves * 64),
                     lds, st, vol, mp, G, part, g, t_base, pass, p.n_groups, p.chunk_cols,
                     p.chunks_per_group);
}

struct A6Plan {
  R1Plan r;
  int K, passes;
  size_t ws;
};

A6Plan plan_a6(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T) {
  A6Plan p;
  p.r = plan_r1(X, Y, Z, N, T);
  p.K = p.r.K < 8 ? p.r.K : 8;
  p.passes = (int)((T + 64 * p.K - 1) / (64 * p.K));
  p.ws = (size_t)p.passes * p.r.n_groups * 3 * N * sizeof(float);
  return p;
}

template <typename GT>
int run_a6(const A6Plan& p, const float* vol, const float* mp, const void* Gv, float* part,
           const ConfGeom& g, hipStream_t st) {
  const GT* G = (const GT*)Gv;
  for (int pass = 0; pass < p.passes; ++pass) {
    const int t_base = pass * 64 * p.K;
    switch (p.K) {
      case 1: launch_a6<1, GT>(p.r, vol, mp, G, part, g, t_base, pass, st); break;
      case 2: launch_a6<2, GT>(p.r, vol, mp, G, part, g, t_base, pass, st); break;
      case 4: launch_a6<4, GT>(p.r, vol, mp, G, part, g, t_base, pass, st); break;
      default: launch_a6<8, GT>(p.r, vol, mp, G, part, g, t_base, pass, st); break;
    }
    int rc = check_launch("a6_kernel");
    if (rc) return rc;
  }
  return TOTRI_OK;
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_volume_confocal_grad_measurement_points_workspace(
    int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T) {
  if (X <= 0 || Y <= 0 || Z < 2 || N <= 0 || T <= 0) return 0;
  return plan_a6(X, Y, Z, N, T).ws;
}

extern "C" int totri_volume_confocal_grad_measurement_points(
    float* measurement_points_grad, const float* volume, int64_t X, int64_t Y, int64_t Z,
    const float* measurement_points, int64_t N, const void* transient_grad, int64_t T,
    int transient_dtype, const float volume_origin[3], const float volume_size[3],
    float bin_width, float bin_offset, void* workspace, size_t workspace_bytes,
    void* stream) {
  int rc = check_common(X, Y, Z, N, T, volume_origin, volume_size, bin_width, bin_offset,
                        transient_dtype);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (N == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(measurement_points_grad != nullptr, "null gradient output");
  // volume_cuda.cu:285 zero-fills the gradient before accumulating.
  if (T == 0 || X == 0 || Y == 0 || Z < 2) return zero_fill(measurement_points_grad, (size_t)3 * N, st);
  TOTRI_CHECK_ARG(volume && measurement_points && transient_grad, "null input");
  TOTRI_CHECK_ARG(Z <= kChunkVoxels, "volume depth Z=%lld exceeds %d", (long long)Z,
                  kChunkVoxels);
  const A6Plan p = plan_a6(X, Y, Z, N, T);
  if (workspace_bytes < p.ws)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", p.ws,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  const ConfGeom g = make_geom(X, Y, Z, N, T, volume_origin, volume_size, bin_width,
                               bin_offset);
  float* part = (float*)workspace;
  switch (transient_dtype) {
    case TOTRI_F32: rc = run_a6<float>(p, volume, measurement_points, transient_grad, part, g, st); break;
    case TOTRI_BF16: rc = run_a6<__hip_bfloat16>(p, volume, measurement_points, transient_grad, part, g, st); break;
    default: rc = run_a6<__half>(p, volume, measurement_points, transient_grad, part, g, st); break;
  }
  if (rc) return rc;
  const size_t n = (size_t)3 * N;
  hipLaunchKernelGGL(sum_partials_kernel<float>, dim3(launch_grid(n)), dim3(256), 0, st, part,
                     p.passes * p.r.n_groups, n, measurement_points_grad);
  return check_launch("sum_partials_kernel");
}
