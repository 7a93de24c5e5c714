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

#include <type_traits>

#include "common.h"

namespace totri_hip {
namespace {

constexpr int kWaves = 8;            // scan points per R1 workgroup
constexpr int kChunkVoxels = 4096;   // voxels per R1 LDS chunk buffer (32 KiB of pairs)
constexpr int kSlab = 32;            // A1 z-slab depth
constexpr int kFlush = 32;           // A1 scan points per first-level sum
constexpr int kSlabQ = 8;            // xy-mirror and x-pair A1 row-walk z-slab depth
// General A1 row-walk z-slab depth.  Deeper slabs amortise a walk's start
// over more rows but measured slower at 128^3 / 64^2 / 1024 (16 rows 27.2 ms,
// 32 rows 27.7 ms with four times the scan ranges, against 26.2 ms).
constexpr int kSlabG = 8;

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
  // Hardware square root and reciprocal (a few ulp, i.e. a few parts in
  // 1e7): the relative margin of 1e-5 on the radii and the 2 bins on the
  // indices are far wider, up to T of order 1e7.
  const float dl = fmaxf(dlo, 0.f);
  const float r_lo = sqrt_fast(rho2 + dl * dl) * (1.f - 1e-5f) - 1e-6f;
  const float r_hi = sqrt_fast(rho2 + dhi * dhi) * (1.f + 1e-5f) + 1e-6f;
  const float iw = __builtin_amdgcn_rcpf(g.w);
  const float top = (float)g.T + 1.f;
  // Positive radii tc in [r_lo, r_hi] and negative radii tc in [-r_hi, -r_lo]
  // give two bin intervals; each is clipped to [0, T-1] on its own and the
  // result is the hull of the non-empty ones.
  const float pa = fminf(fmaxf((r_lo - g.off) * iw - 2.f, -1.f), top);
  const float pb = fminf(fmaxf((r_hi - g.off) * iw + 2.f, -1.f), top);
  const float na = fminf(fmaxf((-r_hi - g.off) * iw - 2.f, -1.f), top);
  const float nb = fminf(fmaxf((-r_lo - g.off) * iw + 2.f, -1.f), top);
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

// bin_range for the row walks' geometries (w > 0 and off >= 0, both finite):
// every bin's radius t w + off is >= 0, so only the positive-radius interval
// can hold an active bin, and it is the whole result.
__device__ __forceinline__ void bin_range_pos(const ConfGeom& g, float rho2, float dlo,
                                              float dhi, int* lo, int* hi) {
  if (!(dhi > 0.f)) {
    *lo = 1;
    *hi = 0;
    return;
  }
  const float dl = fmaxf(dlo, 0.f);
  const float r_lo = sqrt_fast(rho2 + dl * dl) * (1.f - 1e-5f) - 1e-6f;
  const float r_hi = sqrt_fast(rho2 + dhi * dhi) * (1.f + 1e-5f) + 1e-6f;
  const float iw = __builtin_amdgcn_rcpf(g.w);
  const float top = (float)g.T + 1.f;
  const float pa = fminf(fmaxf((r_lo - g.off) * iw - 2.f, -1.f), top);
  const float pb = fminf(fmaxf((r_hi - g.off) * iw + 2.f, -1.f), top);
  *lo = max((int)floorf(pa), 0);
  *hi = min((int)ceilf(pb), g.T - 1);
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
                                           float delta, float half_m_delta, float* alpha,
                                           bool* near) {
  const float zh = fmaf(sqrt_fast(d2), g.s, cfast);
  const float al = __builtin_amdgcn_fractf(zh);
  *alpha = al;
  // frac within delta of an integer: |frac - 1/2| > 1/2 - delta (one compare;
  // the rounding of frac - 1/2 is far inside delta's margin).  d2 <= 0 gives
  // a NaN frac, which compares false (such samples are inactive anyway).
  *near = fabsf(al - 0.5f) > half_m_delta;
  return floor_to_int(zh);
}

// The fast decision as a row key: z_hat itself, whose floor is the candidate
// row, with alpha and `near` as in decide_fast.  The row walks test "row <=
// rr" as key < rr + 1 and so need no float-to-int conversion per step.
__device__ __forceinline__ float decide_fast_key(const ConfGeom& g, float d2, float cfast,
                                                 float half_m_delta, float* alpha,
                                                 bool* near) {
  const float zh = fmaf(sqrt_fast(d2), g.s, cfast);
  const float al = __builtin_amdgcn_fractf(zh);
  *alpha = al;
  *near = fabsf(al - 0.5f) > half_m_delta;
  return zh;
}

// Row key of row r: floor(r + 1/2) == r for |r| < 2^23; larger rows keep a
// key above every row a walk compares with.
__device__ __forceinline__ float row_key(int r) { return (float)r + 0.5f; }

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
                                               float half_m_delta, float* alpha) {
  const float depth = sqrt_fast(d2);
  const float zh = fmaf(depth, g.s, cfast);
  // alpha = zh - floor(zh) (exact; v_fract only clamps results that would
  // round to 1.0, which happens for zh in (-1, 0), a rejected range anyway).
  float al = __builtin_amdgcn_fractf(zh);
  int r = floor_to_int(zh);
  if ((d2 > 0.f) & (fabsf(al - 0.5f) > half_m_delta)) {
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
template <int K, bool BELOW = false>
__device__ __forceinline__ void r1_block(
    const float* __restrict__ vol, const float* __restrict__ mp,
    float* __restrict__ dst,  // [n_groups][N][T] partial sums, or [N][T]
    const ConfGeom& g, int t_base, int grp, int sgrp, int n_groups, int chunk_cols,
    int chunks_per_group, float2* __restrict__ lds2) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int i = uniform_i32(sgrp * kWaves + wave);  // scalar registers
  const bool scan_ok = i < g.N;
  const int n_cols = g.X * g.Y;
  const int chunk_vox = chunk_cols * g.Z;

  // Per-lane bin constants for each window.  A sample adds v * d2^2 to its
  // bin's sum, which is scaled by 1/tc^4 twice at the end: the weight
  // (d2 / tc^4)^2 = (depth / tc^2)^4 of volume_cuda.cu:49-51 with the common
  // factor taken out of the sum (1/tc^4 is kept finite, so that a bin with
  // no active sample, tc = 0 among them, ends at 0).
  float tc2[K], itc4[K], a[K], acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int t = t_base + 64 * k + lane;
    const float tc = bin_distance(t, g.w, g.off);
    const float q = tc * tc;
    tc2[k] = (t < g.T) ? q : -1.f;  // -1 makes d2 negative: lane inactive
    itc4[k] = q > 0.f ? fminf(1.f / (q * q), 3.4028235e38f) : 0.f;
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
  const float half_m_delta = 0.5f - delta;
  float dlo, dhi;
  depth_range(g, wz, 0.f, (float)(g.Z - 1), &dlo, &dhi);

  const int chunk_first = grp * chunks_per_group;
  const int total_chunks = (n_cols + chunk_cols - 1) / chunk_cols;
  const int chunk_end = min(chunk_first + chunks_per_group, total_chunks);
  const int nthreads = kWaves * 64;
  const int vol_vox = n_cols * g.Z;

  // Pair (v[z], v[z+1] - v[z]) of global voxel index ge; z = Z-1 holds (0, 0):
  // no active sample reads it (that needs z <= Z-2), and BELOW reads it for
  // inactive samples.
  auto load_pair = [&](int ge) -> float2 {
    if (ge >= vol_vox) return make_float2(0.f, 0.f);
    const bool last = (ge % g.Z) == g.Z - 1;
    if (last) return make_float2(0.f, 0.f);
    const float v0 = vol[ge];
    return make_float2(v0, vol[ge + 1] - v0);
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
      // Two columns of the same x per pass: their samples are independent,
      // which gives the scheduler two dependency chains to interleave, and
      // they share tc^2 - dx^2.  Windows are the union of both columns'
      // windows; a sample outside its column's range simply fails the exact
      // activity test.  A column whose next one has another x goes alone.
      while (mask) {
        const int j1 = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int xi1 = __builtin_amdgcn_readlane(xi, j1);
        int j2 = j1;
        bool two = false;
        if (mask) {
          const int jn = __builtin_ctzll(mask);
          if (__builtin_amdgcn_readlane(xi, jn) == xi1) {
            j2 = jn;
            two = true;
            mask &= mask - 1;
          }
        }
        const unsigned wm1 = (unsigned)__builtin_amdgcn_readlane((int)packed, j1);
        const unsigned wm2 = two ? (unsigned)__builtin_amdgcn_readlane((int)packed, j2) : 0u;
        const unsigned wmask = wm1 | wm2;
        if (xi1 != cached_xi) {
          // tc^2 - dx^2 (volume_cuda.cu:40, first subtraction) depends on the
          // column's x only; recompute it when x changes.
          const float cdx2 = __builtin_bit_cast(
              float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, dx2), j1));
#pragma unroll
          for (int k = 0; k < K; ++k) a[k] = tc2[k] - cdx2;
          cached_xi = xi1;
        }
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
          const float d2b = a[k] - cdy2b;
          if (BELOW) {
            // The wall lies below the box: a sample with d2 <= 0 (clamped to
            // 0) decides a negative row, or after an exact decision a row of
            // weight 0; rows outside 0 .. Z-2 read the zero row Z-1.
            const float d2ca = fmaxf(d2a, 0.f), d2cb = fmaxf(d2b, 0.f);
            float ala, alb;
            bool na, nb;
            int za = decide_fast(g, d2ca, c0, delta, half_m_delta, &ala, &na);
            int zb = decide_fast(g, d2cb, c0, delta, half_m_delta, &alb, &nb);
            if (__ballot(na || nb)) {
              if (na) za = decide_exact(g, d2ca, wz, 0, &ala);
              if (nb) zb = decide_exact(g, d2cb, wz, 0, &alb);
            }
            const float2 va = col1[min((unsigned)za, (unsigned)(g.Z - 1))];
            const float2 vb = col2[min((unsigned)zb, (unsigned)(g.Z - 1))];
            const float fa = fmaf(ala, va.y, va.x);
            const float fb = fmaf(alb, vb.y, vb.x);
            acc[k] = fmaf(fb, d2cb * d2cb, fmaf(fa, d2ca * d2ca, acc[k]));
            continue;
          }
          float ala, alb;
          bool na, nb;
          int za = decide_fast(g, d2a, c0, delta, half_m_delta, &ala, &na);
          int zb = decide_fast(g, d2b, c0, delta, half_m_delta, &alb, &nb);
          if (__ballot(na || nb)) {  // rare: confirm near-boundary samples exactly
            if (na) za = decide_exact(g, d2a, wz, 0, &ala);
            if (nb) zb = decide_exact(g, d2b, wz, 0, &alb);
          }
          const bool acta = (d2a > 0.f) & ((unsigned)za <= (unsigned)(g.Z - 2));
          const bool actb = (d2b > 0.f) & ((unsigned)zb <= (unsigned)(g.Z - 2));
          // Inactive lanes read slot 0 and keep their accumulator unchanged
          // (the discarded candidate may be NaN or inf).  Selects measure
          // faster here than exec-masked updates.
          const float2 va = col1[acta ? za : 0];
          const float2 vb = col2[actb ? zb : 0];
          const float fa = fmaf(ala, va.y, va.x);
          const float fb = fmaf(alb, vb.y, vb.x);
          const float ca = fmaf(fa, d2a * d2a, acc[k]);
          const float cb = fmaf(fb, d2b * d2b, acta ? ca : acc[k]);
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
      if (t < g.T) row[t] = (acc[k] * itc4[k]) * itc4[k];
    }
  }
}

// True when the wall points of all of the workgroup's scan points (sgrp * 8
// .. sgrp * 8 + 7) lie below the first voxel centre, c0 = (wz - oz) s - 1/2
// < 0.  The same value in every thread.
__device__ __forceinline__ bool scans_below_box(const float* __restrict__ mp, int sgrp,
                                                const ConfGeom& g) {
  bool below = true;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const int i = sgrp * kWaves + w;
    if (i < g.N) {
      float c0 = (mp[2 * g.N + i] - g.oz) * g.s;
      c0 = c0 - 0.5f;
      below = below && (c0 < 0.f);
    }
  }
  return __builtin_amdgcn_readfirstlane((int)below) != 0;
}

// General R1 item; workgroups whose wall points lie below the box take the
// select-free activity test.
template <int K>
__device__ __forceinline__ void r1_item(const float* __restrict__ vol,
                                        const float* __restrict__ mp,
                                        float* __restrict__ dst, const ConfGeom& g,
                                        int t_base, int grp, int sgrp, int n_groups,
                                        int chunk_cols, int chunks_per_group,
                                        float2* __restrict__ lds2) {
  if (scans_below_box(mp, sgrp, g))
    r1_block<K, true>(vol, mp, dst, g, t_base, grp, sgrp, n_groups, chunk_cols,
                      chunks_per_group, lds2);
  else
    r1_block<K, false>(vol, mp, dst, g, t_base, grp, sgrp, n_groups, chunk_cols,
                       chunks_per_group, lds2);
}

// `skip` (optional): a device flag; when it is set the launch does nothing
// (the mirror-symmetric kernel below has done the work).
template <int K>
__global__ __launch_bounds__(kWaves * 64) void r1_forward_kernel(
    const float* __restrict__ vol, const float* __restrict__ mp,
    float* __restrict__ dst, ConfGeom g, int t_base, int n_groups, int chunk_cols,
    int chunks_per_group, const int* __restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ float2 lds2[];
  r1_item<K>(vol, mp, dst, g, t_base, (int)(blockIdx.x % n_groups),
             (int)(blockIdx.x / n_groups), n_groups, chunk_cols, chunks_per_group, lds2);
}

// ---------------------------------------------------------------------------
// Mirror-symmetric R1
// ---------------------------------------------------------------------------
// When the voxel grid is antisymmetric in x and y (cell centre X-1-xi is
// exactly -centre(xi), likewise in y) and the scan points come in quads
//   s0 = (wx, wy, wz), s1 = (-wx, wy, wz), s2 = (wx, -wy, wz), s3 = (-wx, -wy, wz)
// (exact float values), the samples (s_m, column c_(m^f)) for m = 0..3 share
// dx^2 and dy^2 bit for bit, where c_0 = (x, y), c_1 = (-x, y), c_2 = (x, -y),
// c_3 = (-x, -y) are the mirror images of a column in the quadrant
// xi < X/2, yi < Y/2.  Their decisions (zi, alpha) and weights are therefore
// identical, so one decision serves four samples: family f of a column
// quad has dx^2 = (wx - x_(f&1))^2 and dy^2 = (wy - y_(f&2))^2.
//
// The quads are found on the device (scan_mirror_kernel, scan_quads_kernel),
// which writes a flag; the symmetric kernel runs when it is set, the general
// kernel otherwise.  Results equal the general kernel's up to the FP32
// summation order over columns.

// For each scan point, the index of its x-, y- and xy-mirror image: the
// smallest matching index, or N and above if there is none (partner[3][N]
// starts at INT_MAX, see mirror_init_kernel).  Work is split into wave items
// (block of 64 points, segment of the candidates); a wave compares its
// points with every candidate of its segment (uniform, scalar loads) and
// merges its first matches with atomicMin.
constexpr int kMirrorPts = 64;

__host__ __device__ inline int mirror_segments(int N) {
  const int blocks = (N + kMirrorPts - 1) / kMirrorPts;
  int segs = 4096 / (blocks > 0 ? blocks : 1);
  return segs < 1 ? 1 : (segs > 64 ? 64 : segs);
}

__device__ __forceinline__ void scan_mirror_item(const float* __restrict__ mp, int N,
                                                 int* __restrict__ partner, int item) {
  const int segs = mirror_segments(N);
  const int blk = item / segs, seg = item - blk * segs;
  const int per = (N + segs - 1) / segs;
  const int j0 = seg * per, j1 = min(N, j0 + per);
  const int i = blk * kMirrorPts + (int)(threadIdx.x & 63);
  const int ii = i < N ? i : 0;
  const float wx = mp[ii], wy = mp[N + ii], wz = mp[2 * N + ii];
  int px = 0x7fffffff, py = 0x7fffffff, pxy = 0x7fffffff;
  for (int j = j0; j < j1; ++j) {
    const float x = mp[j], y = mp[N + j], z = mp[2 * N + j];
    const bool zs = z == wz;
    const bool xm = x == -wx, xs = x == wx, ym = y == -wy, ys = y == wy;
    if (zs && xm && ys) px = min(px, j);
    if (zs && xs && ym) py = min(py, j);
    if (zs && xm && ym) pxy = min(pxy, j);
  }
  if (i < N) {
    if (px != 0x7fffffff) atomicMin(&partner[i], px);
    if (py != 0x7fffffff) atomicMin(&partner[N + i], py);
    if (pxy != 0x7fffffff) atomicMin(&partner[2 * N + i], pxy);
  }
}

__global__ void mirror_init_kernel(int* __restrict__ partner, int n) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    partner[e] = 0x7fffffff;
}

__global__ __launch_bounds__(64) void scan_mirror_kernel(const float* __restrict__ mp, int N,
                                                         int* __restrict__ partner) {
  scan_mirror_item(mp, N, partner, (int)blockIdx.x);
}

// Inclusive scan of one int per thread over a 1024-thread workgroup.
__device__ __forceinline__ int block_scan_1024(int v, int* s_cnt) {
  const int tid = threadIdx.x;
  __syncthreads();
  s_cnt[tid] = v;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int u = tid >= off ? s_cnt[tid - off] : 0;
    __syncthreads();
    s_cnt[tid] += u;
    __syncthreads();
  }
  return s_cnt[tid];
}

// One workgroup: checks whether the mirror relation splits all N scan
// points into quads of four distinct points, and whether the x-mirror alone
// splits them into pairs, and lists them in order of their smallest index:
// quads (s0, s1, s2, s3), pairs (s0, s1).
//   info[0] = 1 when every point is in a quad, info[1] = number of quads,
//   info[2] = 1 when every point is in an x-pair, info[3] = number of pairs.
__device__ __forceinline__ void scan_quads_block(const int* __restrict__ partner, int N,
                                                 int* __restrict__ quads,
                                                 int* __restrict__ pairs,
                                                 int* __restrict__ info) {
  __shared__ int s_cnt[1024];
  __shared__ int s_bad[2];
  const int tid = threadIdx.x;
  if (tid < 2) s_bad[tid] = 0;
  __syncthreads();
  const int per = (N + 1023) / 1024;
  const int b0 = min(N, tid * per), b1 = min(N, b0 + per);
  const int* PX = partner;
  const int* PY = partner + N;
  const int* PXY = partner + 2 * N;
  // an index outside [0, N) means "no mirror image"
  auto quad_ok = [&](int i) {
    const unsigned px = PX[i], py = PY[i], pxy = PXY[i];
    if (px >= (unsigned)N || py >= (unsigned)N || pxy >= (unsigned)N) return false;
    return (int)px != i && (int)py != i && (int)pxy != i && px != py && px != pxy &&
           py != pxy && PX[px] == i && PY[py] == i && PX[pxy] == (int)py &&
           PY[pxy] == (int)px;
  };
  auto pair_ok = [&](int i) {
    const unsigned px = PX[i];
    return px < (unsigned)N && (int)px != i && PX[px] == i;
  };
  int qc = 0, pc = 0, qbad = 0, pbad = 0;
  for (int i = b0; i < b1; ++i) {
    if (!quad_ok(i)) qbad = 1;
    else if (i < PX[i] && i < PY[i] && i < PXY[i]) ++qc;
    if (!pair_ok(i)) pbad = 1;
    else if (i < PX[i]) ++pc;
  }
  if (qbad) atomicOr(&s_bad[0], 1);
  if (pbad) atomicOr(&s_bad[1], 1);
  const int q_incl = block_scan_1024(qc, s_cnt);
  const int n_quads = s_cnt[1023];
  const int p_incl = block_scan_1024(pc, s_cnt);
  const int n_pairs = s_cnt[1023];
  const bool quads_ok = s_bad[0] == 0;
  const bool pairs_ok = s_bad[1] == 0;
  if (quads_ok) {
    int q = q_incl - qc;
    for (int i = b0; i < b1; ++i) {
      if (i < PX[i] && i < PY[i] && i < PXY[i]) {
        quads[4 * q + 0] = i;
        quads[4 * q + 1] = PX[i];
        quads[4 * q + 2] = PY[i];
        quads[4 * q + 3] = PXY[i];
        ++q;
      }
    }
  }
  if (pairs_ok) {
    int q = p_incl - pc;
    for (int i = b0; i < b1; ++i) {
      if (i < PX[i]) {
        pairs[2 * q + 0] = i;
        pairs[2 * q + 1] = PX[i];
        ++q;
      }
    }
  }
  if (tid == 0) {
    info[0] = (quads_ok && 4 * n_quads == N) ? 1 : 0;
    info[1] = n_quads;
    info[2] = (pairs_ok && 2 * n_pairs == N) ? 1 : 0;
    info[3] = n_pairs;
  }
}

__global__ __launch_bounds__(1024) void scan_quads_kernel(const int* __restrict__ partner,
                                                          int N, int* __restrict__ quads,
                                                          int* __restrict__ pairs,
                                                          int* __restrict__ info) {
  scan_quads_block(partner, N, quads, pairs, info);
}

// True when the wall points of all of the workgroup's quads (M = 4) or
// x-pairs (M = 2) lie below the first voxel centre (c0 = (wz - oz) s - 1/2
// < 0): a sample with d2 <= 0 then decides a negative row.  The same value in
// every thread.
template <int M = 4>
__device__ __forceinline__ bool quads_below_box(const float* __restrict__ mp,
                                                const int* __restrict__ quads, int n_quads,
                                                int qgrp, const ConfGeom& g) {
  bool below = true;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const int q = qgrp * kWaves + w;
    if (q < n_quads) {
      const float wz = mp[2 * g.N + quads[M * q]];
      float c0 = (wz - g.oz) * g.s;
      c0 = c0 - 0.5f;
      below = below && (c0 < 0.f);
    }
  }
  return __builtin_amdgcn_readfirstlane((int)below) != 0;
}

// One workgroup item of the symmetric R1: scan quads qgrp*8 .. qgrp*8+7 (one
// per wave), bins t_base .. t_base + 64K - 1, column-quad group grp.  The
// chunk buffers hold, per quadrant column, its four mirror columns:
// lds2[((j * 4) + m) * Z + z] for column quad j and mirror m.
// BELOW (quads_below_box holds): the activity test is the row range alone.
// M = 2: the same walk for x-pairs of scan points (s0, s1 = x-mirror of s0)
// over the half grid x < X/2 and every y: a column and its x-mirror make a
// column pair, families f = 0, 1 share dy^2, and one decision serves two
// samples (scan s_m at column c_(m^f)).
template <int K, bool BELOW = false, int M = 4>
__device__ __forceinline__ void r1_sym_block(
    const float* __restrict__ vol, const float* __restrict__ mp,
    const int* __restrict__ quads, int n_quads, float* __restrict__ dst,
    const ConfGeom& g, int t_base, int grp, int qgrp, int n_groups, int reps_per_chunk,
    int chunks_per_group, float2* __restrict__ lds2) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  static_assert(M == 4 || M == 2, "quads or x-pairs");
  const int q = uniform_i32(qgrp * kWaves + wave);  // scalar registers
  const bool quad_ok = q < n_quads;
  const int HX = g.X / 2, HY = M == 4 ? g.Y / 2 : g.Y;
  const int n_reps = HX * HY;
  // A chunk holds reps_per_chunk column quads (pairs) as [rep][z][m] pairs
  // with z = 0 .. Z, where rows Z - 1 and Z are (0, 0) (no active sample
  // reads row Z - 1: it needs z <= Z - 2): a sample's M mirror pairs sit side
  // by side, and an inactive sample reads a zero row.
  const int zrows = g.Z + 1;
  const int rep_pairs = M * zrows;
  const int chunk_vox = reps_per_chunk * rep_pairs;  // pairs
  const int chunk_items = reps_per_chunk * zrows;    // (rep, z) rows of M pairs

  float tc2[K], itc4[K], acc[M][K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int t = t_base + 64 * k + lane;
    const float tc = bin_distance(t, g.w, g.off);
    const float qq = tc * tc;
    tc2[k] = (t < g.T) ? qq : -1.f;
    // finite, so that a bin with no active sample (tc = 0 among them) ends at 0;
    // the sums take v * d2^2 and are scaled by 1/tc^4 twice at the end
    itc4[k] = qq > 0.f ? fminf(1.f / (qq * qq), 3.4028235e38f) : 0.f;
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m][k] = 0.f;
  }
  int sidx[M];
#pragma unroll
  for (int m = 0; m < M; ++m) sidx[m] = quad_ok ? quads[M * q + m] : 0;
  const int s0 = uniform_i32(sidx[0]);
  const float wx = uniform_f32(mp[s0]);
  const float wy = uniform_f32(mp[g.N + s0]);
  const float wz = uniform_f32(mp[2 * g.N + s0]);
  float c0 = (wz - g.oz) * g.s;
  c0 = c0 - 0.5f;
  const float delta = guard_delta(g, wz, c0, 0.f);
  const float half_m_delta = 0.5f - delta;
  float dlo, dhi;
  depth_range(g, wz, 0.f, (float)(g.Z - 1), &dlo, &dhi);

  const int total_chunks = (n_reps + reps_per_chunk - 1) / reps_per_chunk;
  const int chunk_first = grp * chunks_per_group;
  const int chunk_end = min(chunk_first + chunks_per_group, total_chunks);
  const int nthreads = kWaves * 64;

  // Row `it` of chunk c: the pairs (v[z], v[z+1] - v[z]) of the M mirror
  // columns m = 0..M-1 of one column quad (pair) at one z, as float4 (m = 0, 1
  // and m = 2, 3).  Consecutive rows are consecutive z of the same columns,
  // so a wave's loads run along M columns.
  auto load_row = [&](int c, int it, float4* lo, float4* hi) {
    const int jr = it / zrows;
    const int z = it - jr * zrows;
    const int r = c * reps_per_chunk + jr;
    float2 p[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) p[m] = make_float2(0.f, 0.f);
    if (r < n_reps && z < g.Z - 1) {
      const int rx = r / HY, ry = r - rx * HY;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int xi = (m & 1) ? g.X - 1 - rx : rx;
        const int yi = (m & 2) ? g.Y - 1 - ry : ry;
        const float* col = vol + ((size_t)xi * g.Y + yi) * g.Z;
        const float v0 = col[z];
        const float v1 = col[z + 1];
        p[m] = make_float2(v0, v1 - v0);
      }
    }
    *lo = make_float4(p[0].x, p[0].y, p[1].x, p[1].y);
    *hi = make_float4(p[2].x, p[2].y, p[3].x, p[3].y);
  };

  auto store_row = [&](float4* base, int it, float4 lo, float4 hi) {
    if (M == 4) {
      base[2 * it] = lo;
      base[2 * it + 1] = hi;
    } else {
      base[it] = lo;
    }
  };
  if (chunk_first < chunk_end) {
    float4* dst4 = reinterpret_cast<float4*>(lds2);
    for (int it = threadIdx.x; it < chunk_items; it += nthreads) {
      float4 lo, hi;
      load_row(chunk_first, it, &lo, &hi);
      store_row(dst4, it, lo, hi);
    }
  }
  __syncthreads();

  constexpr int kPref = kChunkVoxels / (M * kWaves * 64);  // rows per thread
  for (int c = chunk_first; c < chunk_end; ++c) {
    const int buf = (c - chunk_first) & 1;
    const float2* cur = lds2 + buf * chunk_vox;
    float4* nxt4 = reinterpret_cast<float4*>(lds2 + (buf ^ 1) * chunk_vox);
    const bool has_next = c + 1 < chunk_end;
    // The next chunk goes to the other buffer right away (every wave left
    // it at the previous barrier); no registers stay held during the walk.
    if (has_next) {
#pragma unroll
      for (int r = 0; r < kPref; ++r) {
        const int it = threadIdx.x + r * nthreads;
        if (it < chunk_items) {
          float4 lo, hi;
          load_row(c + 1, it, &lo, &hi);
          store_row(nxt4, it, lo, hi);
        }
      }
    }

    if (quad_ok) {
      // Lane j describes column quad j of the chunk.
      const int r = c * reps_per_chunk + lane;
      const bool rvalid = lane < reps_per_chunk && r < n_reps;
      const int rr = rvalid ? r : 0;
      const int rx = rr / HY, ry = rr - rx * HY;
      const float x0 = cell_centre(rx, g.ox, g.szx, g.X);
      const float x1 = cell_centre(g.X - 1 - rx, g.ox, g.szx, g.X);
      const float y0 = cell_centre(ry, g.oy, g.szy, g.Y);
      const float y1 = cell_centre(g.Y - 1 - ry, g.oy, g.szy, g.Y);
      const float dxa = wx - x0, dxb = wx - x1, dya = wy - y0, dyb = wy - y1;
      const float A = dxa * dxa, B = dxb * dxb, C = dya * dya, D = dyb * dyb;
      unsigned packed = 0u;
      bool any = false;
#pragma unroll
      for (int f = 0; f < M; ++f) {
        const float dx2 = (f & 1) ? B : A;
        const float dy2 = (f & 2) ? D : C;
        int tlo, thi;
        bin_range(g, dx2 + dy2, dlo, dhi, &tlo, &thi);
        const bool hit = rvalid && thi >= t_base && tlo <= t_base + 64 * K - 1 && tlo <= thi;
        const int klo = (max(tlo, t_base) - t_base) >> 6;
        const int khi = (min(thi, t_base + 64 * K - 1) - t_base) >> 6;
        const unsigned bits = hit ? (((2u << khi) - 1u) & ~((1u << klo) - 1u)) : 0u;
        packed |= bits << (8 * f);
        any = any || hit;
      }
      unsigned long long mask = __ballot(any);
      while (mask) {
        const int j = __builtin_ctzll(mask);
        mask &= mask - 1;
        const unsigned wm = (unsigned)__builtin_amdgcn_readlane((int)packed, j);
        const float cA = lane_value(A, j), cB = lane_value(B, j);
        const float cC = lane_value(C, j), cD = M == 4 ? lane_value(D, j) : 0.f;
        const float2* quad = cur + j * rep_pairs;
        // Families 2h and 2h+1 share dy^2; both are evaluated for a window
        // where either is active (two independent decisions per pass; a
        // sample outside its family's range fails the activity test).  The
        // first subtraction tc^2 - dx^2 is shared by h = 0 and h = 1.
        const unsigned wmh0 = (wm | (wm >> 8)) & 0xffu;
        const unsigned wmh1 = M == 4 ? ((wm >> 16) | (wm >> 24)) & 0xffu : 0u;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (!(((wmh0 | wmh1) >> k) & 1u)) continue;
          const float pA = tc2[k] - cA, pB = tc2[k] - cB;  // volume_cuda.cu:40
#pragma unroll
          for (int h = 0; h < M / 2; ++h) {
            const int fa = 2 * h, fb = 2 * h + 1;
            if (!(((h ? wmh1 : wmh0) >> k) & 1u)) continue;
            const float dy2 = h ? cD : cC;
            const float d2a = pA - dy2;
            const float d2b = pB - dy2;
            // Decisions and weights use d2 clamped at 0: identical for an
            // active sample (d2 > 0), finite for an inactive one, which then
            // reads the zero row: its term is 0 * 0 + 0 times a finite weight.
            const float d2ca = fmaxf(d2a, 0.f), d2cb = fmaxf(d2b, 0.f);
            float ala, alb;
            bool na, nb;
            int za = decide_fast(g, d2ca, c0, delta, half_m_delta, &ala, &na);
            int zb = decide_fast(g, d2cb, c0, delta, half_m_delta, &alb, &nb);
            if (__ballot(na || nb)) {
              if (na) za = decide_exact(g, d2ca, wz, 0, &ala);
              if (nb) zb = decide_exact(g, d2cb, wz, 0, &alb);
            }
            int ra, rb;
            if (BELOW) {
              // rows past Z - 2 and negative rows (among them every d2 <= 0
              // sample, or, after the exact decision, a row of weight 0)
              // read the zero row Z - 1
              ra = (int)min((unsigned)za, (unsigned)(g.Z - 1));
              rb = (int)min((unsigned)zb, (unsigned)(g.Z - 1));
            } else {
              const bool acta = (d2a > 0.f) & ((unsigned)za <= (unsigned)(g.Z - 2));
              const bool actb = (d2b > 0.f) & ((unsigned)zb <= (unsigned)(g.Z - 2));
              ra = acta ? za : g.Z;
              rb = actb ? zb : g.Z;
            }
            const float2* pa = quad + M * ra;
            const float2* pb = quad + M * rb;
            // the common factor 1/tc^8 is applied to the sums at the end
            const float wa = d2ca * d2ca, wb = d2cb * d2cb;
#pragma unroll
            for (int m = 0; m < M; ++m) {
              const float2 va = pa[m ^ fa];
              const float2 vb = pb[m ^ fb];
              const float fva = fmaf(ala, va.y, va.x);
              const float fvb = fmaf(alb, vb.y, vb.x);
              acc[m][k] = fmaf(fvb, wb, fmaf(fva, wa, acc[m][k]));
            }
          }
        }
      }
    }

    __syncthreads();
  }

  if (quad_ok) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      float* row = dst + ((size_t)grp * g.N + sidx[m]) * g.T;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int t = t_base + 64 * k + lane;
        if (t < g.T) row[t] = (acc[m][k] * itc4[k]) * itc4[k];
      }
    }
  }
}

// One workgroup item of the symmetric A6: the walk of r1_sym_block (same
// staging, culling and decisions), accumulating per scan point s_m of the
// wave's quad and bin the three partial derivatives of A6 (a6_kernel).  The
// sample of family f for s_m reads column c_(m^f) and has offsets
// (sigma_x(m) dx_f, sigma_y(m) dy_f) with dx_f = wx - x_f, dy_f = wy - y_f
// (sigma = -1 on a mirrored axis), so the shared decision serves four scan
// points whose gradients differ only in those signs and in the gathered
// pair.  The wave then contracts its bins with G[s_m][t] and writes one
// (gx, gy, gz) per scan point to part[pass][grp][3][N].  M = 2: the same walk
// over x-pairs of scan points (as r1_sym_block<K, BELOW, 2>).
template <int K, typename GT, int M = 4>
__device__ __forceinline__ void a6_sym_block(
    const float* __restrict__ vol, const float* __restrict__ mp,
    const int* __restrict__ quads, int n_quads, const GT* __restrict__ G,
    float* __restrict__ part, const ConfGeom& g, int t_base, int pass, int grp, int qgrp,
    int n_groups, int reps_per_chunk, int chunks_per_group, float2* __restrict__ lds2) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  static_assert(M == 4 || M == 2, "quads or x-pairs");
  const int q = uniform_i32(qgrp * kWaves + wave);  // scalar registers
  const bool quad_ok = q < n_quads;
  const int HX = g.X / 2, HY = M == 4 ? g.Y / 2 : g.Y;
  const int n_reps = HX * HY;
  // A chunk holds reps_per_chunk column quads (pairs) as [rep][z][m] pairs
  // with z = 0 .. Z, where row z = Z is (0, 0): a sample's M mirror pairs sit
  // side by side, and an inactive sample reads the zero row.
  const int zrows = g.Z + 1;
  const int rep_pairs = M * zrows;
  const int chunk_vox = reps_per_chunk * rep_pairs;  // pairs
  const int chunk_items = reps_per_chunk * zrows;    // (rep, z) rows of M pairs

  float tc2[K], itc4[K], gx[M][K], gy[M][K], gz[M][K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int t = t_base + 64 * k + lane;
    const float tc = bin_distance(t, g.w, g.off);
    const float qq = tc * tc;
    tc2[k] = (t < g.T) ? qq : -1.f;
    // finite, so that an inactive sample (d2 clamped to 0) has weight 0
    itc4[k] = qq > 0.f ? fminf(1.f / (qq * qq), 3.4028235e38f) : 0.f;
#pragma unroll
    for (int m = 0; m < M; ++m) gx[m][k] = gy[m][k] = gz[m][k] = 0.f;
  }
  int sidx[M];
#pragma unroll
  for (int m = 0; m < M; ++m) sidx[m] = quad_ok ? quads[M * q + m] : 0;
  const int s0 = uniform_i32(sidx[0]);
  const float wx = uniform_f32(mp[s0]);
  const float wy = uniform_f32(mp[g.N + s0]);
  const float wz = uniform_f32(mp[2 * g.N + s0]);
  float c0 = (wz - g.oz) * g.s;
  c0 = c0 - 0.5f;
  const float delta = guard_delta(g, wz, c0, 0.f);
  const float half_m_delta = 0.5f - delta;
  float dlo, dhi;
  depth_range(g, wz, 0.f, (float)(g.Z - 1), &dlo, &dhi);

  const int total_chunks = (n_reps + reps_per_chunk - 1) / reps_per_chunk;
  const int chunk_first = grp * chunks_per_group;
  const int chunk_end = min(chunk_first + chunks_per_group, total_chunks);
  const int nthreads = kWaves * 64;

  // Row `it` of chunk c: the pairs (v[z], v[z+1] - v[z]) of the four mirror
  // columns m = 0..3 of one column quad at one z, as two float4 (m = 0, 1 and
  // m = 2, 3).  Consecutive rows are consecutive z of the same columns, so a
  // wave's loads run along four columns.
  auto load_row = [&](int c, int it, float4* lo, float4* hi) {
    const int jr = it / zrows;
    const int z = it - jr * zrows;
    const int r = c * reps_per_chunk + jr;
    float2 p[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) p[m] = make_float2(0.f, 0.f);
    if (r < n_reps && z < g.Z) {
      const int rx = r / HY, ry = r - rx * HY;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int xi = (m & 1) ? g.X - 1 - rx : rx;
        const int yi = (m & 2) ? g.Y - 1 - ry : ry;
        const float* col = vol + ((size_t)xi * g.Y + yi) * g.Z;
        const float v0 = col[z];
        const float v1 = (z == g.Z - 1) ? v0 : col[z + 1];
        p[m] = make_float2(v0, v1 - v0);
      }
    }
    *lo = make_float4(p[0].x, p[0].y, p[1].x, p[1].y);
    *hi = make_float4(p[2].x, p[2].y, p[3].x, p[3].y);
  };

  auto store_row = [&](float4* base, int it, float4 lo, float4 hi) {
    if (M == 4) {
      base[2 * it] = lo;
      base[2 * it + 1] = hi;
    } else {
      base[it] = lo;
    }
  };
  if (chunk_first < chunk_end) {
    float4* dst4 = reinterpret_cast<float4*>(lds2);
    for (int it = threadIdx.x; it < chunk_items; it += nthreads) {
      float4 lo, hi;
      load_row(chunk_first, it, &lo, &hi);
      store_row(dst4, it, lo, hi);
    }
  }
  __syncthreads();

  constexpr int kPref = kChunkVoxels / (M * kWaves * 64);  // rows per thread
  for (int c = chunk_first; c < chunk_end; ++c) {
    const int buf = (c - chunk_first) & 1;
    const float2* cur = lds2 + buf * chunk_vox;
    float4* nxt4 = reinterpret_cast<float4*>(lds2 + (buf ^ 1) * chunk_vox);
    const bool has_next = c + 1 < chunk_end;
    // The next chunk goes to the other buffer right away (every wave left
    // it at the previous barrier); no registers stay held during the walk.
    if (has_next) {
#pragma unroll
      for (int r = 0; r < kPref; ++r) {
        const int it = threadIdx.x + r * nthreads;
        if (it < chunk_items) {
          float4 lo, hi;
          load_row(c + 1, it, &lo, &hi);
          store_row(nxt4, it, lo, hi);
        }
      }
    }

    if (quad_ok) {
      // Lane j describes column quad j of the chunk.
      const int r = c * reps_per_chunk + lane;
      const bool rvalid = lane < reps_per_chunk && r < n_reps;
      const int rr = rvalid ? r : 0;
      const int rx = rr / HY, ry = rr - rx * HY;
      const float x0 = cell_centre(rx, g.ox, g.szx, g.X);
      const float x1 = cell_centre(g.X - 1 - rx, g.ox, g.szx, g.X);
      const float y0 = cell_centre(ry, g.oy, g.szy, g.Y);
      const float y1 = cell_centre(g.Y - 1 - ry, g.oy, g.szy, g.Y);
      const float dxa = wx - x0, dxb = wx - x1, dya = wy - y0, dyb = wy - y1;
      const float A = dxa * dxa, B = dxb * dxb, C = dya * dya, D = dyb * dyb;
      unsigned packed = 0u;
      bool any = false;
#pragma unroll
      for (int f = 0; f < M; ++f) {
        const float dx2 = (f & 1) ? B : A;
        const float dy2 = (f & 2) ? D : C;
        int tlo, thi;
        bin_range(g, dx2 + dy2, dlo, dhi, &tlo, &thi);
        const bool hit = rvalid && thi >= t_base && tlo <= t_base + 64 * K - 1 && tlo <= thi;
        const int klo = (max(tlo, t_base) - t_base) >> 6;
        const int khi = (min(thi, t_base + 64 * K - 1) - t_base) >> 6;
        const unsigned bits = hit ? (((2u << khi) - 1u) & ~((1u << klo) - 1u)) : 0u;
        packed |= bits << (8 * f);
        any = any || hit;
      }
      unsigned long long mask = __ballot(any);
      while (mask) {
        const int j = __builtin_ctzll(mask);
        mask &= mask - 1;
        const unsigned wm = (unsigned)__builtin_amdgcn_readlane((int)packed, j);
        const float cA = lane_value(A, j), cB = lane_value(B, j);
        const float cC = lane_value(C, j), cD = M == 4 ? lane_value(D, j) : 0.f;
        const float cxa = lane_value(dxa, j), cxb = lane_value(dxb, j);
        const float cya = lane_value(dya, j), cyb = M == 4 ? lane_value(dyb, j) : 0.f;
        const float2* quad = cur + j * rep_pairs;
        // Families 2h and 2h+1 share dy^2; both are evaluated for a window
        // where either is active (two independent decisions per pass; a
        // sample outside its family's range fails the activity test).
#pragma unroll
        for (int h = 0; h < M / 2; ++h) {
          const int fa = 2 * h, fb = 2 * h + 1;
          const unsigned wmask = ((wm >> (8 * fa)) | (wm >> (8 * fb))) & 0xffu;
          const float dy2 = h ? cD : cC;
          const float cy = h ? cyb : cya;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            if (!((wmask >> k) & 1u)) continue;
            const float d2a = (tc2[k] - cA) - dy2;  // volume_cuda.cu:40
            const float d2b = (tc2[k] - cB) - dy2;
            // Decisions and weights use d2 clamped at 0: identical for an
            // active sample (d2 > 0), finite for an inactive one, which then
            // reads the zero row: its term is 0 * 0 + 0 times a finite weight.
            const float d2ca = fmaxf(d2a, 0.f), d2cb = fmaxf(d2b, 0.f);
            float ala, alb;
            bool na, nb;
            int za = decide_fast(g, d2ca, c0, delta, half_m_delta, &ala, &na);
            int zb = decide_fast(g, d2cb, c0, delta, half_m_delta, &alb, &nb);
            if (__ballot(na || nb)) {
              if (na) za = decide_exact(g, d2ca, wz, 0, &ala);
              if (nb) zb = decide_exact(g, d2cb, wz, 0, &alb);
            }
            const bool acta = (d2a > 0.f) & ((unsigned)za <= (unsigned)(g.Z - 2));
            const bool actb = (d2b > 0.f) & ((unsigned)zb <= (unsigned)(g.Z - 2));
            const float2* pa = quad + M * (acta ? za : g.Z);
            const float2* pb = quad + M * (actb ? zb : g.Z);
            // Shared per decision, with u = d2 / tc^8: f2 = (depth / tc^2)^4
            // = u d2, the first term's factor 4 u, and f2 / depth (an
            // inactive sample has d2 = 0 and a finite 1 / max(depth, tiny)).
            const float i8 = fminf(itc4[k] * itc4[k], 3.4028235e38f);
            const float ua = d2ca * i8, ub = d2cb * i8;
            const float f2a = ua * d2ca, f2b = ub * d2cb;
            const float Aa = 4.f * ua, Ab = 4.f * ub;
            const float Ba = f2a * __builtin_amdgcn_rcpf(fmaxf(sqrt_fast(d2ca), 1e-30f));
            const float Bb = f2b * __builtin_amdgcn_rcpf(fmaxf(sqrt_fast(d2cb), 1e-30f));
#pragma unroll
            for (int m = 0; m < M; ++m) {
              // d/dw of scan s_m (a6_kernel): -dx (4 f1 u + s dv f2 / depth),
              // likewise y, and s dv f2 for z; sample offsets carry m's signs
              const float sx = (m & 1) ? 1.f : -1.f, sy = (m & 2) ? 1.f : -1.f;
              const float2 va = pa[m ^ fa];
              const float2 vb = pb[m ^ fb];
              const float sdva = g.s * va.y, sdvb = g.s * vb.y;
              const float cxya = fmaf(fmaf(ala, va.y, va.x), Aa, sdva * Ba);
              const float cxyb = fmaf(fmaf(alb, vb.y, vb.x), Ab, sdvb * Bb);
              gx[m][k] = fmaf(sx * cxb, cxyb, fmaf(sx * cxa, cxya, gx[m][k]));
              gy[m][k] = fmaf(sy * cy, cxyb, fmaf(sy * cy, cxya, gy[m][k]));
              gz[m][k] = fmaf(sdvb, f2b, fmaf(sdva, f2a, gz[m][k]));
            }
          }
        }
      }
    }

    __syncthreads();
  }

  if (quad_ok) {
    float* o = part + ((size_t)pass * n_groups + grp) * 3 * g.N;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      float sxs = 0.f, sys = 0.f, szs = 0.f;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int t = t_base + 64 * k + lane;
        const float gt = t < g.T ? load_elem(G + (size_t)sidx[m] * g.T + t) : 0.f;
        sxs = fmaf(gx[m][k], gt, sxs);
        sys = fmaf(gy[m][k], gt, sys);
        szs = fmaf(gz[m][k], gt, szs);
      }
      sxs = wave_sum_f32(sxs);
      sys = wave_sum_f32(sys);
      szs = wave_sum_f32(szs);
      if (lane == 0) {
        o[sidx[m]] = sxs;
        o[g.N + sidx[m]] = sys;
        o[2 * g.N + sidx[m]] = szs;
      }
    }
  }
}

// Two workgroups per CU (their chunk buffers take 128 KiB of LDS), so the
// compiler is held to 128 VGPRs.  blockIdx.y counts the bin passes from
// the last one.
template <int K>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(4)))
void r1_sym_kernel(
    const float* __restrict__ vol, const float* __restrict__ mp,
    const int* __restrict__ quads, const int* __restrict__ info, float* __restrict__ dst,
    ConfGeom g, int n_groups, int reps_per_chunk, int chunks_per_group) {
  if (!info[0]) return;
  extern __shared__ float2 lds2[];
  const int n_quads = info[1];
  const int qgrp = (int)(blockIdx.x / n_groups);
  if (qgrp * kWaves >= n_quads) return;  // whole workgroup idle
  // later bin passes first: they usually reach more of the volume (the first
  // bins end short of a box placed away from the wall), so the costlier
  // workgroups start in the first round
  const int t_base = (int)(gridDim.y - 1 - blockIdx.y) * 64 * K;
  const int grp = (int)(blockIdx.x % n_groups);
  if (quads_below_box(mp, quads, n_quads, qgrp, g))
    r1_sym_block<K, true>(vol, mp, quads, n_quads, dst, g, t_base, grp, qgrp, n_groups,
                          reps_per_chunk, chunks_per_group, lds2);
  else
    r1_sym_block<K, false>(vol, mp, quads, n_quads, dst, g, t_base, grp, qgrp, n_groups,
                           reps_per_chunk, chunks_per_group, lds2);
}

// x-pair R1: the mirror R1 over x-pairs of scan points (info[3] of them,
// found on the device), for scans whose points all have x-mirrors (info[2])
// and that the xy-mirror kernel does not take (quad_on and info[0]).
template <int K>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(4)))
void r1_pair_kernel(
    const float* __restrict__ vol, const float* __restrict__ mp,
    const int* __restrict__ pairs, const int* __restrict__ info, float* __restrict__ dst,
    ConfGeom g, int n_groups, int reps_per_chunk, int chunks_per_group, int quad_on) {
  if (!info[2] || (quad_on && info[0])) return;
  extern __shared__ float2 lds2[];
  const int n_pairs = info[3];
  const int qgrp = (int)(blockIdx.x / n_groups);
  if (qgrp * kWaves >= n_pairs) return;  // whole workgroup idle
  const int t_base = (int)(gridDim.y - 1 - blockIdx.y) * 64 * K;  // later passes first
  const int grp = (int)(blockIdx.x % n_groups);
  if (quads_below_box<2>(mp, pairs, n_pairs, qgrp, g))
    r1_sym_block<K, true, 2>(vol, mp, pairs, n_pairs, dst, g, t_base, grp, qgrp, n_groups,
                             reps_per_chunk, chunks_per_group, lds2);
  else
    r1_sym_block<K, false, 2>(vol, mp, pairs, n_pairs, dst, g, t_base, grp, qgrp, n_groups,
                              reps_per_chunk, chunks_per_group, lds2);
}

// Symmetric A6, all bin passes in one launch (blockIdx.y counts them from the
// last one, as in r1_sym_kernel).
template <int K, typename GT>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(4)))
void a6_sym_kernel(
    const float* __restrict__ vol, const float* __restrict__ mp,
    const int* __restrict__ quads, const int* __restrict__ info, const GT* __restrict__ G,
    float* __restrict__ part, ConfGeom g, int n_groups, int reps_per_chunk,
    int chunks_per_group) {
  if (!info[0]) return;
  extern __shared__ float2 lds2[];
  const int n_quads = info[1];
  const int qgrp = (int)(blockIdx.x / n_groups);
  if (qgrp * kWaves >= n_quads) return;  // whole workgroup idle
  const int pass = (int)(gridDim.y - 1 - blockIdx.y);
  a6_sym_block<K, GT>(vol, mp, quads, n_quads, G, part, g, pass * 64 * K, pass,
                      (int)(blockIdx.x % n_groups), qgrp, n_groups, reps_per_chunk,
                      chunks_per_group, lds2);
}

// x-pair A6, all bin passes in one launch, for the scans r1_pair_kernel takes.
template <int K, typename GT>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(4)))
void a6_pair_kernel(
    const float* __restrict__ vol, const float* __restrict__ mp,
    const int* __restrict__ pairs, const int* __restrict__ info, const GT* __restrict__ G,
    float* __restrict__ part, ConfGeom g, int n_groups, int reps_per_chunk,
    int chunks_per_group, int quad_on) {
  if (!info[2] || (quad_on && info[0])) return;
  extern __shared__ float2 lds2[];
  const int n_pairs = info[3];
  const int qgrp = (int)(blockIdx.x / n_groups);
  if (qgrp * kWaves >= n_pairs) return;  // whole workgroup idle
  const int pass = (int)(gridDim.y - 1 - blockIdx.y);
  a6_sym_block<K, GT, 2>(vol, mp, pairs, n_pairs, G, part, g, pass * 64 * K, pass,
                         (int)(blockIdx.x % n_groups), qgrp, n_groups, reps_per_chunk,
                         chunks_per_group, lds2);
}

// Sum of partials in a fixed order where the number of parts depends on
// which path ran: the xy-mirror one (info[0] set and n_sym given), else the
// x-pair one (info[2] set and n_pair given), else the general one.
template <typename OutT>
__global__ void sum_partials_flag_kernel(const float* __restrict__ part,
                                         const int* __restrict__ info, int n_sym, int n_pair,
                                         int n_gen, size_t n, OutT* __restrict__ out) {
  // n_sym (n_pair) is 0 when the xy-mirror (x-pair) kernel was not launched
  const int n_parts = (n_sym && info[0]) ? n_sym : (n_pair && info[2]) ? n_pair : n_gen;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    float v = part[e];
    for (int p = 1; p < n_parts; ++p) v += part[(size_t)p * n + e];
    store_elem(out + e, v);
  }
}

// Boundary rows of a row walk for element e of an [X][Y][Z] volume of n
// voxels: with slabs of slab_h rows, voxel z = k * slab_h (k >= 1) receives
// the alpha part of slab k - 1's top sample row, summed over the n_parts
// scan ranges in order.
__device__ __forceinline__ float a1_boundary(const float* __restrict__ bnd, size_t e,
                                             size_t n, int Z, int n_slabs, int n_parts,
                                             int slab_h) {
  const int z = (int)(e % (size_t)Z);
  if (z == 0 || z % slab_h) return 0.f;
  const size_t ncol = n / (size_t)Z, col = e / (size_t)Z;
  const int sb = z / slab_h - 1;
  float b = bnd[col * n_slabs + sb];
  for (int p = 1; p < n_parts; ++p) b += bnd[((size_t)p * ncol + col) * n_slabs + sb];
  return b;
}

// A1 partial sums: the xy-mirror kernel ran when quad_on and info[0] (its
// boundary rows are added too), else the x-mirror kernel when info[2] (the
// pair row walk with its boundary rows bnd_pair when given, else the window
// walk), else the general kernel (its boundary rows added when bnd_gen is
// given, the row walk); each has its own partial count.  info may be null:
// no mirror path.  The mirror row walks' boundary rows use n_slabs =
// ceil(Z / kSlabQ), the general row walk's n_slabs_gen = ceil(Z / kSlabG).
__global__ void sum_partials_a1_kernel(const float* __restrict__ part,
                                       const float* __restrict__ bnd,
                                       const int* __restrict__ info, int quad_on, int n_quad,
                                       int n_pair, int n_gen, size_t n, int Z, int n_slabs,
                                       int n_slabs_gen, const float* __restrict__ bnd_gen,
                                       const float* __restrict__ bnd_pair,
                                       float* __restrict__ out) {
  const bool quad = info && quad_on && info[0];
  const bool pair = info && !quad && info[2];
  const int n_parts = quad ? n_quad : pair ? n_pair : n_gen;
  const float* b = quad ? bnd : pair ? bnd_pair : bnd_gen;
  const int slab_h = (quad || pair) ? kSlabQ : kSlabG;
  const int slabs = (quad || pair) ? n_slabs : n_slabs_gen;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    float v = part[e];
    for (int p = 1; p < n_parts; ++p) v += part[(size_t)p * n + e];
    if (b) v += a1_boundary(b, e, n, Z, slabs, n_parts, slab_h);
    out[e] = v;
  }
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
    int chunks_per_group, const int* __restrict__ skip) {
  if (skip && *skip) return;  // the mirror kernel has done the work
  extern __shared__ float2 lds2[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int grp = blockIdx.x % n_groups;
  const int sgrp = blockIdx.x / n_groups;
  const int i = uniform_i32(sgrp * kWaves + wave);  // scalar registers
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
  const float half_m_delta = 0.5f - delta;
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
        const int zi = confocal_decide(g, d2, wz, c0, 0, delta, half_m_delta, &al);
        if ((d2 > 0.f) & ((unsigned)zi <= (unsigned)(g.Z - 2))) {
          const float2 v = column[zi];
          const float f1 = fmaf(al, v.y, v.x);
          const float i8 = itc4[k] * itc4[k];
          const float f2 = (d2 * d2) * i8;
          const float depth = sqrt_fast(d2);
          const float sdv = g.s * v.y;
          // f2 / depth with the hardware reciprocal (about 1 ulp) instead of
          // an IEEE division; the term only enters a gradient sum
          const float cxy = 4.f * f1 * d2 * i8 + sdv * (f2 * __builtin_amdgcn_rcpf(depth));
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
    const float half_m_delta = 0.5f - delta;
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
        int ra = decide_fast(g, d2a, cp, delta, half_m_delta, &ala, &na);
        int rb = decide_fast(g, d2b, cp, delta, half_m_delta, &alb, &nb);
        if (__ballot(na || nb)) {  // rare: confirm near-boundary samples exactly
          if (na) ra = decide_exact(g, d2a, wz, zshift, &ala);
          if (nb) rb = decide_exact(g, d2b, wz, zshift, &alb);
        }
        // dx2 is +inf for lanes without a column, so d2 > 0 also checks that.
        const bool acta = (d2a > 0.f) & ((unsigned)ra <= (unsigned)rspan);
        const bool actb = (d2b > 0.f) & ((unsigned)rb <= (unsigned)rspan);
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

// `skip` (optional): a device flag; when it is set the launch does nothing
// (the x-mirror kernel below has done the work).
template <bool SAFE>
__global__ __launch_bounds__(kA1Waves * 64) void a1_transpose_kernel(
    const float* __restrict__ Kf, const float* __restrict__ tc2t,
    const float* __restrict__ itc4t, const float* __restrict__ mp,
    float* __restrict__ dst, ConfGeom g, int tiles_x, int tiles_y, int n_slabs,
    int n_ranges, int t_stride, int n_items, const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ float acc_all[kA1Waves][(kSlab + 2) * 64];
  const int wave = threadIdx.x >> 6;
  const int item = blockIdx.x * kA1Waves + wave;
  if (item >= n_items) return;  // whole wave exits; no barrier follows
  a1_item<SAFE>(Kf, tc2t, itc4t, mp, dst, g, item, tiles_x, tiles_y, n_slabs, n_ranges,
                t_stride, acc_all[wave]);
}

// ---------------------------------------------------------------------------
// x-mirror A1
// ---------------------------------------------------------------------------
// When the voxel grid is antisymmetric in x and the scan points come in
// x-mirror pairs s0 = (wx, wy, wz), s1 = (-wx, wy, wz), the samples
// (s_m, c_(m^f)) for m = 0, 1 share one decision, where c_0 = (x, y) and
// c_1 = (-x, y): family f has dx^2 = (wx - x_f)^2.  A lane owns the column
// pair (c_0, c_1) of an 8x8 tile in the half xi < X/2 and keeps both columns'
// accumulators as float2 rows in LDS, so one decision updates two voxels
// pairs: component n of the row gets the weight of scan s_(n^f).
constexpr int kSlabS = 16;  // x-mirror A1 z-slab depth

template <bool SAFE>
__device__ __forceinline__ void a1_sym_item(
    const float* __restrict__ Kf, const float* __restrict__ tc2t,
    const float* __restrict__ itc4t, const float* __restrict__ mp,
    const int* __restrict__ pairs, int n_pairs,
    float* __restrict__ dst,  // [R][X][Y][Z] partials, or [X][Y][Z]
    const ConfGeom& g, int item, int tiles_x, int tiles_y, int n_slabs, int n_ranges,
    int t_stride, float2* __restrict__ acc) {
  const int lane = threadIdx.x & 63;
  const int tx = item % tiles_x;
  item /= tiles_x;
  const int ty = item % tiles_y;
  item /= tiles_y;
  const int slab = item % n_slabs;
  const int range = item / n_slabs;

  const int HX = g.X / 2;
  const int xi = tx * 8 + (lane >> 3);
  const int yi = ty * 8 + (lane & 7);
  const bool cvalid = xi < HX && yi < g.Y;
  const int xm = g.X - 1 - xi;
  const float x0 = cell_centre(xi, g.ox, g.szx, g.X);
  const float x1 = cell_centre(xm, g.ox, g.szx, g.X);
  const float y = cell_centre(yi, g.oy, g.szy, g.Y);

  const int z0 = slab * kSlabS;
  const int rmin = (z0 == 0) ? 1 : 0;
  const int rmax = min(kSlabS, (g.Z - 2) - (z0 - 1));
  const int rspan = rmax - rmin;
  const int zshift = z0 - 1 + rmin;
  const float zlo = (float)max(z0 - 1, 0);
  const float zhi = (float)min(z0 + kSlabS, g.Z - 1);

  for (int r = 0; r < kSlabS + 2; ++r) acc[r * 64 + lane] = make_float2(0.f, 0.f);
  float2 tot[kSlabS];
#pragma unroll
  for (int r = 0; r < kSlabS; ++r) tot[r] = make_float2(0.f, 0.f);

  const int per = (n_pairs + n_ranges - 1) / n_ranges;
  const int p_begin = range * per;
  const int p_end = min(n_pairs, p_begin + per);
  float2* my = acc + rmin * 64 + lane;

  for (int pi = p_begin; pi < p_end; ++pi) {
    if (pi > p_begin && (pi - p_begin) % (kFlush / 2) == 0) {
#pragma unroll
      for (int r = 0; r < kSlabS; ++r) {
        const float2 a = acc[(r + 1) * 64 + lane];
        tot[r].x += a.x;
        tot[r].y += a.y;
        acc[(r + 1) * 64 + lane] = make_float2(0.f, 0.f);
      }
    }
    const int s0 = uniform_i32(pairs[2 * pi]);
    const int s1 = uniform_i32(pairs[2 * pi + 1]);
    const float wx = mp[s0], wy = mp[g.N + s0], wz = mp[2 * g.N + s0];
    float c0 = (wz - g.oz) * g.s;
    c0 = c0 - 0.5f;
    const float cp = c0 - (float)zshift;
    const float delta = guard_delta(g, wz, c0, cp);
    const float half_m_delta = 0.5f - delta;
    const float dxa = wx - x0, dxb = wx - x1, dy = wy - y;
    const float dy2 = dy * dy;
    float dlo, dhi;
    depth_range(g, wz, zlo, zhi, &dlo, &dhi);
    const float* K0 = Kf + (size_t)s0 * t_stride;
    const float* K1 = Kf + (size_t)s1 * t_stride;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      const float dx2 = cvalid ? (f ? dxb * dxb : dxa * dxa) : INFINITY;
      int tlo, thi;
      bin_range(g, dx2 + dy2, dlo, dhi, &tlo, &thi);
      if (!cvalid || rspan < 0 || tlo > thi) {
        tlo = 0x7fffffff;
        thi = -1;
      }
      const int t_begin = uniform_i32(wave_min_i32(tlo));
      const int t_end = uniform_i32(wave_max_i32(thi));
      for (int tb = t_begin & ~63; tb <= t_end; tb += 64) {
        // component n of a row takes scan s_(n^f)
        const float kreg0 = (f ? K1 : K0)[tb + lane];
        const float kreg1 = (f ? K0 : K1)[tb + lane];
        const float treg = tc2t[tb + lane];
        const float ireg = SAFE ? itc4t[tb + lane] : 0.f;
        const int u0 = max(t_begin - tb, 0);
        const int u1 = min(t_end - tb, 63);
        for (int u = u0; u <= u1; u += 2) {
          const int ub = min(u + 1, u1);
          const bool hb = u + 1 <= u1;
          const float d2a = (lane_value(treg, u) - dx2) - dy2;
          const float d2b = (lane_value(treg, ub) - dx2) - dy2;
          float ala, alb;
          bool na, nb;
          int ra = decide_fast(g, d2a, cp, delta, half_m_delta, &ala, &na);
          int rb = decide_fast(g, d2b, cp, delta, half_m_delta, &alb, &nb);
          if (__ballot(na || nb)) {
            if (na) ra = decide_exact(g, d2a, wz, zshift, &ala);
            if (nb) rb = decide_exact(g, d2b, wz, zshift, &alb);
          }
          const bool acta = (d2a > 0.f) & ((unsigned)ra <= (unsigned)rspan);
          const bool actb = hb & (d2b > 0.f) & ((unsigned)rb <= (unsigned)rspan);
          if (acta) {
            const float base = SAFE ? sq(d2a * lane_value(ireg, u)) : d2a * d2a;
            const float w0 = base * lane_value(kreg0, u), w1 = base * lane_value(kreg1, u);
            float2* pa = my + ra * 64;
            float2 a0 = pa[0], a1 = pa[64];
            const float om = 1.f - ala;
            a0.x = fmaf(w0, om, a0.x);
            a0.y = fmaf(w1, om, a0.y);
            a1.x = fmaf(w0, ala, a1.x);
            a1.y = fmaf(w1, ala, a1.y);
            pa[0] = a0;
            pa[64] = a1;
          }
          if (actb) {
            const float base = SAFE ? sq(d2b * lane_value(ireg, ub)) : d2b * d2b;
            const float w0 = base * lane_value(kreg0, ub), w1 = base * lane_value(kreg1, ub);
            float2* pb = my + rb * 64;
            float2 b0 = pb[0], b1 = pb[64];
            const float om = 1.f - alb;
            b0.x = fmaf(w0, om, b0.x);
            b0.y = fmaf(w1, om, b0.y);
            b1.x = fmaf(w0, alb, b1.x);
            b1.y = fmaf(w1, alb, b1.y);
            pb[0] = b0;
            pb[64] = b1;
          }
        }
      }
    }
  }

  if (cvalid) {
    float* out0 = dst + (((size_t)range * g.X + xi) * g.Y + yi) * g.Z;
    float* out1 = dst + (((size_t)range * g.X + xm) * g.Y + yi) * g.Z;
#pragma unroll
    for (int r = 0; r < kSlabS; ++r) {
      const int z = z0 + r;
      if (z < g.Z) {
        const float2 a = acc[(r + 1) * 64 + lane];
        out0[z] = tot[r].x + a.x;
        out1[z] = tot[r].y + a.y;
      }
    }
  }
}

constexpr int kA1SymWaves = 1;  // waves per x-mirror A1 workgroup (independent items)

template <bool SAFE>
__global__ __launch_bounds__(kA1SymWaves * 64) void a1_sym_kernel(
    const float* __restrict__ Kf, const float* __restrict__ tc2t,
    const float* __restrict__ itc4t, const float* __restrict__ mp,
    const int* __restrict__ pairs, const int* __restrict__ info, float* __restrict__ dst,
    ConfGeom g, int tiles_x, int tiles_y, int n_slabs, int n_ranges, int t_stride,
    int n_items, int quad_on) {
  if (!info[2] || (quad_on && info[0])) return;
  __shared__ float2 acc_all[kA1SymWaves][(kSlabS + 2) * 64];
  const int wave = threadIdx.x >> 6;
  const int item = blockIdx.x * kA1SymWaves + wave;
  if (item >= n_items) return;  // whole wave exits; no barrier follows
  a1_sym_item<SAFE>(Kf, tc2t, itc4t, mp, pairs, info[3], dst, g, item, tiles_x, tiles_y,
                    n_slabs, n_ranges, t_stride, acc_all[wave]);
}

// ---------------------------------------------------------------------------
// xy-mirror A1, row-synchronous
// ---------------------------------------------------------------------------
// When the grid is antisymmetric in x and y and the scan points come in
// mirror quads s_m (m = bit0: x mirror, bit1: y mirror), the samples
// (s_m, c_(m^f)) share one decision, where c_n is the lane's column mirrored
// by n: family f has dx^2 = (wx - x_(f&1))^2 and dy^2 = (wy - y_(f>>1))^2.
// A lane owns the four mirror columns of one column of the quadrant
// x < X/2, y < Y/2 and keeps their accumulators as float4 rows; component n
// takes scan s_(n^f).
//
// With tc >= 0 (w > 0, off >= 0) and s > 0 the row floor(z_hat) of a lane's
// bins never decreases with the bin index.  The wave walks the slab's rows in
// order; in each row every lane walks its own bins until its row moves on, so
// lanes do not share a bin window.  The row's two partial sums stay in
// registers; the lower one joins the LDS row once per row, the upper one is
// carried into the next row.  Table values for the lane's next bin are
// loaded one step ahead.

// True when `p` holds on any active lane of the wave.
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// base[off / 4] for a wave-uniform base and a 32-bit byte offset (lets the
// compiler use the scalar-base load form).
__device__ __forceinline__ float load_at(const float* __restrict__ base, uint32_t off) {
  return *(const float*)((const char*)base + off);
}

// Row stride of the interleaved quad tables: bins 0 .. T-1 and a zero entry
// at T, which a lane may read one step past its last bin (no clamp needed).
__host__ __device__ __forceinline__ size_t kq_stride(int T) { return (size_t)T + 1; }

// The row walks' tc^2 table in LDS: tc^2 = (t w + off)^2 of bins 0 .. T-1 at
// the walks' rounding, shared by a workgroup's waves, for T up to
// kQuadTabMaxT (xy-mirror walk) or kRowTabMaxT (general walk).  A step then
// takes tc^2 with one LDS read instead of three VALU operations.  The table
// sits at the start of the kernel's one static LDS array (so its address
// folds into the read's offset) and the accumulator rows follow it.  A step
// reads the entry of the bin after the next one, up to bin T + 1: past T - 1
// that is an unwritten entry or an accumulator row of the same array, and
// feeds only a decision that a stopped lane discards.
// At 128^3 / 64^2 / 1024 the table took the xy-mirror A1 from 8.91 to 8.60 ms
// and the general A1 from 27.6 to 27.0 ms (2-wave workgroups; 8-wave ones,
// whose LDS stays held until their slowest item ends, measured 28.3 ms).  The
// x-pair walk measured 1 % slower with it (more VGPRs) and computes tc^2.
constexpr int kQuadWaves = 4;
constexpr int kRowWaves = 2;
constexpr int kQuadTabMaxT = 1024;
constexpr int kRowTabMaxT = 1024;

__device__ __forceinline__ void fill_tc2_table(float* __restrict__ tab, const ConfGeom& g) {
  for (int e = threadIdx.x; e < g.T; e += blockDim.x) {
    const float tc = bin_distance(e, g.w, g.off);
    tab[e] = tc * tc;
  }
  __syncthreads();
}

// Component n of v ^ f: (v[0^f], v[1^f], v[2^f], v[3^f]) for a wave-uniform f.
__device__ __forceinline__ float4 xor_perm4(float4 v, int f) {
  if (f & 1) v = make_float4(v.y, v.x, v.w, v.z);
  if (f & 2) v = make_float4(v.z, v.w, v.x, v.y);
  return v;
}

// Per-lane accumulator vectors of the mirror walks: component n for the
// lane's column n (M = 4: float4, M = 2: float2).
template <int M> struct MirrorVecT;
template <> struct MirrorVecT<4> { using type = float4; };
template <> struct MirrorVecT<2> { using type = float2; };
template <int M> using MirrorVec = typename MirrorVecT<M>::type;

template <int M>
__device__ __forceinline__ MirrorVec<M> mv_zero() {
  if constexpr (M == 4) return make_float4(0.f, 0.f, 0.f, 0.f);
  else return make_float2(0.f, 0.f);
}
__device__ __forceinline__ void mv_add(float4& a, float4 b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  a.w += b.w;
}
__device__ __forceinline__ void mv_add(float2& a, float2 b) {
  a.x += b.x;
  a.y += b.y;
}
// a += L - H per component
__device__ __forceinline__ void mv_add_diff(float4& a, float4 L, float4 H) {
  a.x += L.x - H.x;
  a.y += L.y - H.y;
  a.z += L.z - H.z;
  a.w += L.w - H.w;
}
__device__ __forceinline__ void mv_add_diff(float2& a, float2 L, float2 H) {
  a.x += L.x - H.x;
  a.y += L.y - H.y;
}
// a += k * s per component
__device__ __forceinline__ void mv_fma(float4& a, float4 k, float s) {
  a.x = fmaf(k.x, s, a.x);
  a.y = fmaf(k.y, s, a.y);
  a.z = fmaf(k.z, s, a.z);
  a.w = fmaf(k.w, s, a.w);
}
__device__ __forceinline__ void mv_fma(float2& a, float2 k, float s) {
  a.x = fmaf(k.x, s, a.x);
  a.y = fmaf(k.y, s, a.y);
}
__device__ __forceinline__ float4 mv_perm(float4 v, int f) { return xor_perm4(v, f); }
__device__ __forceinline__ float2 mv_perm(float2 v, int f) {
  return (f & 1) ? make_float2(v.y, v.x) : v;
}
__device__ __forceinline__ float mv_get(float4 v, int n) {
  return n == 0 ? v.x : n == 1 ? v.y : n == 2 ? v.z : v.w;
}
__device__ __forceinline__ float mv_get(float2 v, int n) { return n == 0 ? v.x : v.y; }

// The walk for M = 4 (xy-mirror quads of scan points, a lane's column and
// its three mirror columns) and M = 2 (x-mirror pairs, a lane's column and
// its x-mirror).  The group's M table rows are read interleaved (Kq[q][t],
// one 16- or 8-byte load per step) and summed in table order (component m
// for scan s_m); the sums are permuted into the family's order when a row is
// flushed.
template <bool SAFE, int M = 4, bool TAB = false>
__device__ __forceinline__ void a1_quad_item(
    const MirrorVec<M>* __restrict__ Kq, const float* __restrict__ tc2t,
    const float* __restrict__ itc4t, const float* __restrict__ mp,
    const int* __restrict__ quads, int n_quads,
    float* __restrict__ dst,  // [R][X][Y][Z] partial volumes
    float* __restrict__ bnd,  // [R][X][Y][n_slabs] boundary rows
    const ConfGeom& g, int item, int tiles_x, int tiles_y, int n_slabs, int n_ranges,
    int t_stride, MirrorVec<M>* __restrict__ acc, const float* __restrict__ tab = nullptr) {
  using V = MirrorVec<M>;
  constexpr uint32_t kVB = sizeof(V);  // table bytes per bin
  const int lane = threadIdx.x & 63;
  const int tx = item % tiles_x;
  item /= tiles_x;
  const int ty = item % tiles_y;
  item /= tiles_y;
  const int slab = item % n_slabs;
  const int range = item / n_slabs;

  const int xi = tx * 8 + (lane >> 3);
  const int yi = ty * 8 + (lane & 7);
  const bool cvalid = xi < g.X / 2 && yi < (M == 4 ? g.Y / 2 : g.Y);
  const int xm = g.X - 1 - xi, ym = g.Y - 1 - yi;
  const float x0 = cell_centre(xi, g.ox, g.szx, g.X);
  const float x1 = cell_centre(xm, g.ox, g.szx, g.X);
  const float y0 = cell_centre(yi, g.oy, g.szy, g.Y);
  const float y1 = cell_centre(ym, g.oy, g.szy, g.Y);

  // The slab walks the sample rows floor(z_hat) = z0 .. z0 + rspan (at most
  // kSlabQ of them, never past Z - 2).  LDS row k holds voxel z0 + k; row
  // kSlabQ takes the alpha part of the slab's top sample row, which belongs
  // to the next slab's first voxel: it goes to the boundary buffer and the
  // partial-sum pass adds it there, so no slab walks its neighbour's rows.
  const int z0 = slab * kSlabQ;
  const int rspan = min(kSlabQ, (g.Z - 1) - z0) - 1;
  const int zshift = z0;
  const float zlo = (float)z0;
  const float zhi = (float)min(z0 + kSlabQ, g.Z - 1);

  const V zero4 = mv_zero<M>();
  for (int r = 0; r <= kSlabQ; ++r) acc[r * 64 + lane] = zero4;
  V tot[kSlabQ + 1];
#pragma unroll
  for (int r = 0; r <= kSlabQ; ++r) tot[r] = zero4;

  const int per = (n_quads + n_ranges - 1) / n_ranges;
  const int q_begin = range * per;
  const int q_end = min(n_quads, q_begin + per);
  V* my = acc + lane;

  for (int qi = q_begin; qi < q_end; ++qi) {
    if (qi > q_begin && (qi - q_begin) % (kFlush / 4) == 0) {
#pragma unroll
      for (int r = 0; r <= kSlabQ; ++r) {
        mv_add(tot[r], acc[r * 64 + lane]);
        acc[r * 64 + lane] = zero4;
      }
    }
    // the group's points share wz; s0 gives (wx, wy) of family 0
    const int s0 = uniform_i32(quads[M * qi]);
    const float wx = mp[s0], wy = mp[g.N + s0], wz = mp[2 * g.N + s0];
    float c0 = (wz - g.oz) * g.s;
    c0 = c0 - 0.5f;
    const float cp = c0 - (float)zshift;
    const float delta = guard_delta(g, wz, c0, cp);
    const float half_m_delta = 0.5f - delta;
    const float dxa = wx - x0, dxb = wx - x1, dya = wy - y0, dyb = wy - y1;
    float dlo, dhi;
    depth_range(g, wz, zlo, zhi, &dlo, &dhi);
// Each family's walk is compiled separately (f a constant), so the
    // permutations between the sums' order and the LDS rows' order cost no
    // instructions.
    auto walk_family = [&](auto fc) {
      constexpr int f = decltype(fc)::value;
      const float dx2 = cvalid ? ((f & 1) ? dxb * dxb : dxa * dxa) : INFINITY;
      const float dy2 = (M == 4 && (f & 2)) ? dyb * dyb : dya * dya;
      int tlo, thi;
      bin_range_pos(g, dx2 + dy2, dlo, dhi, &tlo, &thi);
      if (!cvalid || rspan < 0 || tlo > thi) {
        tlo = 0;
        thi = -1;
      }
      if (!wave_any(tlo <= thi)) return;
      // A lane's bins past thi have rows past the slab (the bin range is
      // conservative), so a walk ends by itself there unless thi was clipped
      // at the last bins.  When every lane ends at least two bins before
      // the last (thi <= T - 3), the walk needs no per-step stop test and
      // reads at most bin thi + 1 <= T - 2 of its rows; otherwise each lane
      // stops at thi.
      auto walk = [&](auto sc) {
        constexpr bool STOP = decltype(sc)::value;
        // component n of the LDS rows takes scan s_(n^f); the sums' component m
        // takes s_m
        const char* kqp = (const char*)(Kq + (size_t)qi * kq_stride(g.T));
        // decision of bin tt (unspecified for d2 <= 0; the walk never meets
        // such a bin).  tc^2 is recomputed with the tables' rounding.
        auto decide_bin = [&](int tt, float* d2o, float* alo) -> int {
          const float tc = bin_distance(tt, g.w, g.off);
          const float d2 = (tc * tc - dx2) - dy2;
          bool nr;
          int r = decide_fast(g, d2, cp, delta, half_m_delta, alo, &nr);
          if (nr) r = decide_exact(g, d2, wz, zshift, alo);
          *d2o = d2;
          return r;
        };
        // the same for the bin index held as a float (tf == (float)tt)
        auto decide_bin_f = [&](float tf, float* d2o, float* alo) -> float {
          float tc = tf * g.w;
          tc = tc + g.off;
          const float d2 = (tc * tc - dx2) - dy2;
          bool nr;
          float zk = decide_fast_key(g, d2, cp, half_m_delta, alo, &nr);
          // a divergent branch, skipped when no lane is near (fewer scalar
          // instructions than a ballot around it)
          if (nr) zk = row_key(decide_exact(g, d2, wz, zshift, alo));
          *d2o = d2;
          return zk;
        };
        // the same from d2 = (tc^2 - dx^2) - dy^2 with tc^2 from the LDS
        // table (TAB)
        auto decide_d2 = [&](float d2, float* d2o, float* alo) -> float {
          bool nr;
          float zk = decide_fast_key(g, d2, cp, half_m_delta, alo, &nr);
          // a divergent branch, skipped when no lane is near (fewer scalar
          // instructions than a ballot around it)
          if (nr) zk = row_key(decide_exact(g, d2, wz, zshift, alo));
          *d2o = d2;
          return zk;
        };
        // The walk starts at the lane's first bin with d2 > 0 and row >= 0:
        // estimated from the real-valued inverse (depth of row 0's start), then
        // settled by exact decisions (backward while the previous bin also
        // qualifies, forward while this one does not).  Every later bin then
        // has d2 > 0 and a row no lower.
        const float rho2 = dx2 + dy2;
        const float depth0 = fmaxf(((float)zshift + 0.5f) / g.s + g.oz - wz, 0.f);
        const float te = (sqrt_fast(rho2 + depth0 * depth0) - g.off) * __builtin_amdgcn_rcpf(g.w);
        int t = tlo;
        if (te > (float)tlo + 1.f) t = te < (float)thi ? (int)te - 1 : thi;
        float d2, al;
        int r = decide_bin(t, &d2, &al);
        bool back = (t <= thi) & (t > tlo) & (d2 > 0.f) & (r >= 0);
        while (back) {  // divergent: each lane leaves at its own bin
          float d2p, alp;
          const int rp = decide_bin(t - 1, &d2p, &alp);
          back = (d2p > 0.f) & (rp >= 0);
          if (back) {
            --t;
            r = rp;
            al = alp;
            d2 = d2p;
            back = t > tlo;
          }
        }
        bool fwd = (t <= thi) & !((d2 > 0.f) & (r >= 0));
        while (fwd) {  // divergent: each lane leaves at its own bin
          ++t;
          r = decide_bin(t, &d2, &al);
          fwd = (t <= thi) & !((d2 > 0.f) & (r >= 0));
        }
        // A lane with no bin left carries the row 0x7fffffff (a key above
        // every row), so "still walking and on row rr" is the single test
        // key < rr + 1.
        if (!(t <= thi)) r = 0x7fffffff;
        int rr = uniform_i32(wave_min_i32(r));
        // from here on the row is held as a key (row_key; key < rr + 1 <=> row <= rr)
        float zk = row_key(r);
        V carry = zero4;
        // At the start of row rr every live lane's row is >= rr, so a lane
        // that is live with row <= rr is on row rr.
        // The table values of the lane's current bin are loaded one step ahead
        // (right after the previous bin used them), so a whole step covers
        // their latency.  off16 is the bin's byte offset in the lane's Kq row;
        // the bin index runs as a float (exact below 2^24), so tc = t w + off
        // needs no conversion.  The walk starts at t <= thi + 1 <= T and steps
        // only from bins t <= thi, so every read is at a bin <= T: the Kq rows
        // (stride T + 1) and the itc4 table (t_stride >= T + 1 entries) hold a
        // zero there, and no clamp is needed.
        uint32_t off16 = (uint32_t)t * kVB;
        V kq = *(const V*)(kqp + off16);
        float ic = SAFE ? load_at(itc4t, off16 / (kVB / 4)) : 0.f;
        float tf = (float)t;
        const float thif = (float)thi;
        // TAB: tc^2 of the next bin, read from the LDS table one step ahead
        // (entries past T - 1 are read only for decisions a stopped lane
        // discards), and the stop test on byte offsets (thi >= -1).
        float tcn = TAB ? tab[t + 1] : 0.f;
        const int thib = thi * (int)kVB;
        auto live = [&]() -> bool { return TAB ? (int)off16 <= thib : tf <= thif; };
        // Per row: L = carry + sum of K d2^2, H = sum of K d2^2 alpha; the row's
        // voxel gets L - H (its (1 - alpha) parts plus the carry) and H is
        // carried into the next row.
        while (rr <= rspan) {
          if (!wave_any(live())) break;
          V L = carry, H = zero4;
          const float rr1 = (float)(rr + 1);
          // A divergent loop: each lane leaves it at its row's end (the
          // compiler keeps one exec mask update per step, where a uniform loop
          // around an `if` costs several scalar instructions and branches)
          while (zk < rr1) {
            // this bin's weights, then the next bin's decision in place (past
            // thi the lane stops and the decision is unused)
            float base = SAFE ? sq(d2 * ic) : d2 * d2;
            float wh = base * al;
            // keep both products ahead of the next decision, so that d2 and al
            // are free for it (no register copies); emits no instruction
            asm volatile("" : "+v"(base), "+v"(wh));
            float rn;
            if (TAB) {
              // the next bin's decision from tcn, which then takes tc^2 of
              // the bin after it (a whole step ahead of its use)
              const float d2n = (tcn - dx2) - dy2;
              tcn = load_at(tab, off16 / (kVB / 4) + 8u);
              rn = decide_d2(d2n, &d2, &al);
            } else {
              tf = tf + 1.f;
              rn = decide_bin_f(tf, &d2, &al);
            }
            mv_fma(L, kq, base);
            mv_fma(H, kq, wh);
            off16 += kVB;
            kq = *(const V*)(kqp + off16);
            if (SAFE) ic = load_at(itc4t, off16 / (kVB / 4));
            zk = (!STOP || live()) ? rn : INFINITY;
          }
          V* row = my + rr * 64;
          V a = *row;
          L = mv_perm(L, f);
          H = mv_perm(H, f);
          mv_add_diff(a, L, H);
          *row = a;
          carry = mv_perm(H, f);  // back to the sums' order (the permutation is an involution)
          ++rr;
        }
        // the alpha part of the last walked row; when no row was walked (rr
        // started past the slab, or at INT_MAX with no live lane) the carry is
        // zero and rr may lie outside the LDS block, so nothing is written
        if (rr <= rspan + 1) {
          V* row = my + rr * 64;
          V a = *row;
          carry = mv_perm(carry, f);
          mv_add(a, carry);
          *row = a;
        }
      };
      if (wave_any(thi > g.T - 3))
        walk(std::true_type{});
      else
        walk(std::false_type{});
    };
    walk_family(std::integral_constant<int, 0>{});
    walk_family(std::integral_constant<int, 1>{});
    if constexpr (M == 4) {
      walk_family(std::integral_constant<int, 2>{});
      walk_family(std::integral_constant<int, 3>{});
    }
  }

  if (cvalid) {
    // column n of the group: x mirrored when n & 1, y mirrored when n & 2
    size_t col[M];
#pragma unroll
    for (int n = 0; n < M; ++n)
      col[n] = (size_t)((n & 1) ? xm : xi) * g.Y + ((n & 2) ? ym : yi);
#pragma unroll
    for (int r = 0; r < kSlabQ; ++r) {
      const int z = z0 + r;
      if (z < g.Z) {
        V a = acc[r * 64 + lane];
        mv_add(a, tot[r]);
#pragma unroll
        for (int n = 0; n < M; ++n)
          dst[((size_t)range * g.X * g.Y + col[n]) * g.Z + z] = mv_get(a, n);
      }
    }
    if (z0 + kSlabQ < g.Z) {
      V a = acc[kSlabQ * 64 + lane];
      mv_add(a, tot[kSlabQ]);
      float* b = bnd + (size_t)range * g.X * g.Y * n_slabs + slab;
#pragma unroll
      for (int n = 0; n < M; ++n) b[col[n] * n_slabs] = mv_get(a, n);
    }
  }
}

// Persistent: one single-wave workgroup per wave slot takes items from
// `counter` (zeroed by a1_quad_order_kernel) in the cost order, so no slot
// waits for the dispatcher between items.
template <bool SAFE, bool TAB>
__global__ __launch_bounds__(kQuadWaves * 64) void a1_quad_kernel(
    const float4* __restrict__ Kq, const float* __restrict__ tc2t,
    const float* __restrict__ itc4t, const float* __restrict__ mp,
    const int* __restrict__ quads, const int* __restrict__ info, float* __restrict__ dst,
    float* __restrict__ bnd, ConfGeom g, int tiles_x, int tiles_y, int n_slabs, int n_ranges,
    int t_stride, int n_items, const int* __restrict__ order, int* __restrict__ counter) {
  if (!info[0]) return;
  __shared__ float4 lds4[kQuadTabMaxT / 4 + kQuadWaves * (kSlabQ + 1) * 64];
  const int lane = threadIdx.x & 63;
  float* tab = (float*)lds4;
  float4* acc = lds4 + kQuadTabMaxT / 4 + (threadIdx.x >> 6) * (kSlabQ + 1) * 64;
  if (TAB) fill_tc2_table(tab, g);
  for (;;) {
    int k = 0;
    if (lane == 0) k = atomicAdd(counter, 1);
    k = uniform_i32(__shfl(k, 0, 64));
    if (k >= n_items) break;
    a1_quad_item<SAFE, 4, TAB>(Kq, tc2t, itc4t, mp, quads, info[1], dst, bnd, g, order[k],
                               tiles_x, tiles_y, n_slabs, n_ranges, t_stride, acc, tab);
  }
}

// ---------------------------------------------------------------------------
// General A1, row-synchronous
// ---------------------------------------------------------------------------
// The walk of a1_quad_item for scans without mirror symmetry: one scan point
// at a time, one column per lane (an 8x8 tile of the whole grid), a scalar
// table row Kf[i][t] (one 4-byte load per step; the rows have at least T + 1
// entries, zero past T-1, so the one-step-ahead load needs no clamp), and
// per-row sums L and H in registers.  The same conditions as the mirror walk
// (w > 0, off >= 0, a positive z size: rows never decrease with the bin) and
// the same boundary rows.  It replaces the window-synchronous a1_item, whose
// lanes share one bin window and update LDS once per sample.
template <bool SAFE, bool TAB = false>
__device__ __forceinline__ void a1_row_item(
    const float* __restrict__ Kf, const float* __restrict__ itc4t,
    const float* __restrict__ mp,
    float* __restrict__ dst,  // [R][X][Y][Z] partial volumes
    float* __restrict__ bnd,  // [R][X][Y][n_slabs] boundary rows
    const ConfGeom& g, int item, int tiles_x, int tiles_y, int n_slabs, int n_ranges,
    int t_stride, float* __restrict__ acc, const float* __restrict__ tab = nullptr) {
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

  // Slab rows as in a1_quad_item: sample rows z0 .. z0 + rspan, LDS row k for
  // voxel z0 + k, row kSlabG for the next slab's first voxel (boundary).
  const int z0 = slab * kSlabG;
  const int rspan = min(kSlabG, (g.Z - 1) - z0) - 1;
  const int zshift = z0;
  const float zlo = (float)z0;
  const float zhi = (float)min(z0 + kSlabG, g.Z - 1);

  for (int r = 0; r <= kSlabG; ++r) acc[r * 64 + lane] = 0.f;
  float tot[kSlabG + 1];
#pragma unroll
  for (int r = 0; r <= kSlabG; ++r) tot[r] = 0.f;

  const int per = (g.N + n_ranges - 1) / n_ranges;
  const int i_begin = range * per;
  const int i_end = min(g.N, i_begin + per);
  float* my = acc + lane;

  for (int i = i_begin; i < i_end; ++i) {
    if (i > i_begin && (i - i_begin) % kFlush == 0) {
#pragma unroll
      for (int r = 0; r <= kSlabG; ++r) {
        tot[r] += acc[r * 64 + lane];
        acc[r * 64 + lane] = 0.f;
      }
    }
    const float wx = mp[i], wy = mp[g.N + i], wz = mp[2 * g.N + i];
    float c0 = (wz - g.oz) * g.s;
    c0 = c0 - 0.5f;
    const float cp = c0 - (float)zshift;
    const float delta = guard_delta(g, wz, c0, cp);
    const float half_m_delta = 0.5f - delta;
    const float dx = wx - x, dy = wy - y;
    const float dx2 = cvalid ? dx * dx : INFINITY, dy2 = dy * dy;
    float dlo, dhi;
    depth_range(g, wz, zlo, zhi, &dlo, &dhi);
    int tlo, thi;
    bin_range_pos(g, dx2 + dy2, dlo, dhi, &tlo, &thi);
    if (!cvalid || rspan < 0 || tlo > thi) {
      tlo = 0;
      thi = -1;
    }
    if (!wave_any(tlo <= thi)) continue;
    // Per-step stop test only where some lane's range reaches the last bins
    // (see a1_quad_item).
    // OFF0: a zero bin offset, where t w + 0 rounds to t w, so the step's
    // radius needs no addition.
    auto walk = [&](auto sc, auto oc) {
      constexpr bool STOP = decltype(sc)::value;
      constexpr bool OFF0 = decltype(oc)::value;
      const float* kp = Kf + (size_t)i * t_stride;
      auto decide_bin = [&](int tt, float* d2o, float* alo) -> int {
        const float tc = bin_distance(tt, g.w, g.off);
        const float d2 = (tc * tc - dx2) - dy2;
        bool nr;
        int r = decide_fast(g, d2, cp, delta, half_m_delta, alo, &nr);
        if (nr) r = decide_exact(g, d2, wz, zshift, alo);
        *d2o = d2;
        return r;
      };
      auto decide_bin_f = [&](float tf, float* d2o, float* alo) -> float {
        float tc = tf * g.w;
        if (!OFF0) tc = tc + g.off;
        const float d2 = (tc * tc - dx2) - dy2;
        bool nr;
        float zk = decide_fast_key(g, d2, cp, half_m_delta, alo, &nr);
        if (nr) zk = row_key(decide_exact(g, d2, wz, zshift, alo));
        *d2o = d2;
        return zk;
      };
      auto decide_d2 = [&](float d2, float* d2o, float* alo) -> float {
        bool nr;
        float zk = decide_fast_key(g, d2, cp, half_m_delta, alo, &nr);
        if (nr) zk = row_key(decide_exact(g, d2, wz, zshift, alo));
        *d2o = d2;
        return zk;
      };
      // Exact walk start (see a1_quad_item).
      const float rho2 = dx2 + dy2;
      const float depth0 = fmaxf(((float)zshift + 0.5f) / g.s + g.oz - wz, 0.f);
      const float te = (sqrt_fast(rho2 + depth0 * depth0) - g.off) * __builtin_amdgcn_rcpf(g.w);
      int t = tlo;
      if (te > (float)tlo + 1.f) t = te < (float)thi ? (int)te - 1 : thi;
      float d2, al;
      int r = decide_bin(t, &d2, &al);
      bool back = (t <= thi) & (t > tlo) & (d2 > 0.f) & (r >= 0);
      while (back) {  // divergent: each lane leaves at its own bin
        float d2p, alp;
        const int rp = decide_bin(t - 1, &d2p, &alp);
        back = (d2p > 0.f) & (rp >= 0);
        if (back) {
          --t;
          r = rp;
          al = alp;
          d2 = d2p;
          back = t > tlo;
        }
      }
      bool fwd = (t <= thi) & !((d2 > 0.f) & (r >= 0));
      while (fwd) {  // divergent: each lane leaves at its own bin
        ++t;
        r = decide_bin(t, &d2, &al);
        fwd = (t <= thi) & !((d2 > 0.f) & (r >= 0));
      }
      if (!(t <= thi)) r = 0x7fffffff;
      int rr = uniform_i32(wave_min_i32(r));
      float zk = row_key(r);  // 0x7fffffff: a key above every row
      float carry = 0.f;
      // t <= thi + 1 <= T < t_stride: every table read is in the row
      uint32_t off4 = (uint32_t)t * 4u;
      float k = load_at(kp, off4);
      float ic = SAFE ? load_at(itc4t, off4) : 0.f;
      float tf = (float)t;
      const float thif = (float)thi;
      // TAB: the next bin's tc^2 from the LDS table, one step ahead (see
      // a1_quad_item); the table and the K row share the byte offset.
      float tcn = TAB ? tab[t + 1] : 0.f;
      const int thib = thi * 4;
      auto live = [&]() -> bool { return TAB ? (int)off4 <= thib : tf <= thif; };
      while (rr <= rspan) {
        if (!wave_any(live())) break;
        float L = carry, H = 0.f;
        const float rr1 = (float)(rr + 1);
        while (zk < rr1) {  // divergent, as in a1_quad_item
          float base = SAFE ? sq(d2 * ic) : d2 * d2;
          float wh = base * al;
          asm volatile("" : "+v"(base), "+v"(wh));  // see a1_quad_item
          float rn;
          if (TAB) {  // as in a1_quad_item
            const float d2n = (tcn - dx2) - dy2;
            tcn = load_at(tab, off4 + 8u);
            rn = decide_d2(d2n, &d2, &al);
          } else {
            tf = tf + 1.f;
            rn = decide_bin_f(tf, &d2, &al);
          }
          L = fmaf(k, base, L);
          H = fmaf(k, wh, H);
          off4 += 4u;
          k = load_at(kp, off4);
          if (SAFE) ic = load_at(itc4t, off4);
          zk = (!STOP || live()) ? rn : INFINITY;
        }
        float* row = my + rr * 64;
        *row += L - H;
        carry = H;
        ++rr;
      }
      if (rr <= rspan + 1) my[rr * 64] += carry;
    };
    const bool stop = wave_any(thi > g.T - 3);
    if (TAB) {  // the table holds tc^2 for either offset
      if (stop)
        walk(std::true_type{}, std::false_type{});
      else
        walk(std::false_type{}, std::false_type{});
    } else if (g.off == 0.f) {
      if (stop)
        walk(std::true_type{}, std::true_type{});
      else
        walk(std::false_type{}, std::true_type{});
    } else {
      if (stop)
        walk(std::true_type{}, std::false_type{});
      else
        walk(std::false_type{}, std::false_type{});
    }
  }

  if (cvalid) {
    float* out = dst + (((size_t)range * g.X + xi) * g.Y + yi) * g.Z;
#pragma unroll
    for (int r = 0; r < kSlabG; ++r) {
      const int z = z0 + r;
      if (z < g.Z) out[z] = tot[r] + acc[r * 64 + lane];
    }
    if (z0 + kSlabG < g.Z)
      bnd[(((size_t)range * g.X + xi) * g.Y + yi) * n_slabs + slab] =
          tot[kSlabG] + acc[kSlabG * 64 + lane];
  }
}

// One item per wave, kRowWaves waves per workgroup (they share the tc^2
// table when TAB).  `skip` (optional): a device flag; when it is set the
// launch does nothing (a mirror kernel has done the work).
template <bool SAFE, bool TAB>
__global__ __launch_bounds__(kRowWaves * 64) __attribute__((amdgpu_waves_per_eu(8)))
void a1_row_kernel(
    const float* __restrict__ Kf, const float* __restrict__ itc4t,
    const float* __restrict__ mp, float* __restrict__ dst, float* __restrict__ bnd,
    ConfGeom g, int tiles_x, int tiles_y, int n_slabs, int n_ranges, int t_stride,
    int n_items, const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ float lds1[(TAB ? kRowTabMaxT : 0) + kRowWaves * (kSlabG + 1) * 64];
  float* tab = lds1;
  float* acc = lds1 + (TAB ? kRowTabMaxT : 0) + (threadIdx.x >> 6) * (kSlabG + 1) * 64;
  if (TAB) fill_tc2_table(tab, g);
  // wave-uniform (scalar registers): the scan loop, its row pointers and
  // the wall-point loads derived from it stay scalar
  const int item = uniform_i32((int)blockIdx.x * kRowWaves + (int)(threadIdx.x >> 6));
  if (item >= n_items) return;  // whole wave; no barrier follows
  a1_row_item<SAFE, TAB>(Kf, itc4t, mp, dst, bnd, g, item, tiles_x, tiles_y, n_slabs,
                         n_ranges, t_stride, acc, tab);
}

// Kq[q][t] = (K[s0][t], K[s1][t], K[s2][t], K[s3][t]) for the device-found
// quads q < info[1] (s_m = quads[4q + m]): the four rows one xy-mirror walk
// reads, interleaved.
__global__ __launch_bounds__(256) void a1_quad_table_kernel(
    const float* __restrict__ Kf, const int* __restrict__ quads, const int* __restrict__ info,
    int t_stride, int T, int max_quads, float4* __restrict__ Kq) {
  if (!info[0]) return;
  const size_t ks = kq_stride(T);
  const size_t n = (size_t)min(info[1], max_quads) * ks;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t qi = e / ks, t = e - qi * ks;
    const int* qs = quads + 4 * qi;
    Kq[e] = t < (size_t)T
                ? make_float4(Kf[(size_t)qs[0] * t_stride + t], Kf[(size_t)qs[1] * t_stride + t],
                              Kf[(size_t)qs[2] * t_stride + t], Kf[(size_t)qs[3] * t_stride + t])
                : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Kp[p][t] = (K[s0][t], K[s1][t]) for the device-found x-pairs p < info[3]
// (s_m = pairs[2p + m]), built only when the x-pair walk runs (pairs found,
// and not the xy-mirror walk).
__global__ __launch_bounds__(256) void a1_pair_table_kernel(
    const float* __restrict__ Kf, const int* __restrict__ pairs, const int* __restrict__ info,
    int quad_on, int t_stride, int T, int max_pairs, float2* __restrict__ Kp) {
  if (!info[2] || (quad_on && info[0])) return;
  const size_t ks = kq_stride(T);
  const size_t n = (size_t)min(info[3], max_pairs) * ks;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t pi = e / ks, t = e - pi * ks;
    const int* ps = pairs + 2 * pi;
    Kp[e] = t < (size_t)T ? make_float2(Kf[(size_t)ps[0] * t_stride + t],
                                        Kf[(size_t)ps[1] * t_stride + t])
                          : make_float2(0.f, 0.f);
  }
}

// x-pair A1 row walk: one single-wave workgroup per item (a tile of 8x8
// columns of the half grid x < 0, one 8-row slab, one pair range).
template <bool SAFE>
__global__ __launch_bounds__(64) void a1_pair_kernel(
    const float2* __restrict__ Kp, const float* __restrict__ tc2t,
    const float* __restrict__ itc4t, const float* __restrict__ mp,
    const int* __restrict__ pairs, const int* __restrict__ info, float* __restrict__ dst,
    float* __restrict__ bnd, ConfGeom g, int tiles_x, int tiles_y, int n_slabs, int n_ranges,
    int t_stride, int quad_on) {
  if (!info[2] || (quad_on && info[0])) return;
  __shared__ float2 acc[(kSlabQ + 1) * 64];
  a1_quad_item<SAFE, 2>(Kp, tc2t, itc4t, mp, pairs, info[3], dst, bnd, g, (int)blockIdx.x,
                        tiles_x, tiles_y, n_slabs, n_ranges, t_stride, acc);
}

// Dispatch order of the xy-mirror A1 items.  Item costs differ by a factor of
// several (deep slabs see more bins per voxel row than shallow ones; tiles
// near the scan's centre more than those at its edge), and workgroups start
// in index order as slots free up, so the costliest items go first
// (longest-processing-time order).  The order changes scheduling only: every
// item writes its own part of the partial volumes.
//
// Cost estimate of one item, in bins: for 16 quads spread over the item's
// scan range and each family, the widest of the slab's bin ranges at the
// tile's four corner columns.
__global__ __launch_bounds__(64) void a1_quad_cost_kernel(
    const float* __restrict__ mp, const int* __restrict__ quads, const int* __restrict__ info,
    ConfGeom g, int tiles_x, int tiles_y, int n_slabs, int n_ranges, int n_items,
    float* __restrict__ cost) {
  if (!info[0]) return;
  int item = blockIdx.x;
  if (item >= n_items) return;
  const int lane = threadIdx.x;
  const int tx = item % tiles_x;
  item /= tiles_x;
  const int ty = item % tiles_y;
  item /= tiles_y;
  const int slab = item % n_slabs;
  const int range = item / n_slabs;
  const int n_quads = info[1];
  const int per = (n_quads + n_ranges - 1) / n_ranges;
  const int q_begin = range * per;
  const int cnt = min(n_quads, q_begin + per) - q_begin;
  const int k = lane >> 2, f = lane & 3;
  const int samples = min(cnt, 16);
  float bins = 0.f;
  if (k < samples) {
    const int q = q_begin + (int)(((long long)k * cnt) / samples);
    const int s0 = quads[4 * q];
    const float wx = mp[s0], wy = mp[g.N + s0], wz = mp[2 * g.N + s0];
    const int z0 = slab * kSlabQ;
    float dlo, dhi;
    depth_range(g, wz, (float)z0, (float)min(z0 + kSlabQ, g.Z - 1), &dlo, &dhi);
    int widest = 0;
    for (int c = 0; c < 4; ++c) {
      const int xi = min(tx * 8 + ((c & 1) ? 7 : 0), g.X / 2 - 1);
      const int yi = min(ty * 8 + ((c & 2) ? 7 : 0), g.Y / 2 - 1);
      const float x = cell_centre((f & 1) ? g.X - 1 - xi : xi, g.ox, g.szx, g.X);
      const float y = cell_centre((f & 2) ? g.Y - 1 - yi : yi, g.oy, g.szy, g.Y);
      int tlo, thi;
      bin_range(g, (wx - x) * (wx - x) + (wy - y) * (wy - y), dlo, dhi, &tlo, &thi);
      widest = max(widest, thi - tlo + 1);
    }
    bins = (float)widest;
  }
  const float total = wave_sum_f32(bins);
  if (lane == 0) cost[blockIdx.x] = samples > 0 ? total * ((float)cnt / (float)samples) : 0.f;
}

// Counting sort of the items into 256 cost buckets, costliest first (one
// workgroup; the order within a bucket is immaterial).
__global__ __launch_bounds__(1024) void a1_quad_order_kernel(const float* __restrict__ cost,
                                                             const int* __restrict__ info,
                                                             int n, int* __restrict__ order,
                                                             int* __restrict__ counter) {
  if (!info[0]) return;
  __shared__ int hist[256];
  __shared__ float red[1024];
  const int tid = threadIdx.x;
  if (tid == 0) *counter = 0;  // work queue head of the A1 kernel that follows
  float m = 0.f;
  for (int i = tid; i < n; i += 1024) m = fmaxf(m, cost[i]);
  red[tid] = m;
  if (tid < 256) hist[tid] = 0;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if (tid < h) red[tid] = fmaxf(red[tid], red[tid + h]);
    __syncthreads();
  }
  const float scale = red[0] > 0.f ? 255.f / red[0] : 0.f;
  auto bucket = [&](int i) { return 255 - min(255, max(0, (int)(cost[i] * scale))); };
  for (int i = tid; i < n; i += 1024) atomicAdd(&hist[bucket(i)], 1);
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int b = 0; b < 256; ++b) {
      const int c = hist[b];
      hist[b] = run;
      run += c;
    }
  }
  __syncthreads();
  for (int i = tid; i < n; i += 1024) order[atomicAdd(&hist[bucket(i)], 1)] = i;
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

// Workspace of the scan-mirror detection: partner indices [3][N], quads
// [N/4][4], x-pairs [N/2][2] and the info block, from byte `base`.
struct MirrorLayout {
  size_t partner_off, quads_off, pairs_off, info_off, end;
};

MirrorLayout mirror_layout(int64_t N, size_t base) {
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  MirrorLayout m;
  m.partner_off = align(base);
  m.quads_off = m.partner_off + align(3 * (size_t)N * sizeof(int));
  m.pairs_off = m.quads_off + align((size_t)N * sizeof(int));
  m.info_off = m.pairs_off + align((size_t)N * sizeof(int));
  m.end = m.info_off + 256;
  return m;
}

// Plan of the mirror-symmetric R1 (used when the device finds scan quads).
struct R1SymPlan {
  bool possible;  // sizes allow it (X, Y even, N a multiple of 4, Z small enough)
  int K, passes, n_groups, reps_per_chunk, chunks_per_group;
  MirrorLayout ml;
  size_t staging_off, ws;
};

// Cell centres of one axis, computed as the kernels compute them, are
// antisymmetric: centre(n-1-i) == -centre(i) for every i.
bool axis_antisymmetric(float o, float sz, int64_t n) {
  if (n < 2 || (n & 1)) return false;
  for (int64_t i = 0; i < n / 2; ++i) {
    float a = ((float)i + 0.5f) * sz;
    a = a / (float)n;
    a = o + a;
    float b = ((float)(n - 1 - i) + 0.5f) * sz;
    b = b / (float)n;
    b = o + b;
    if (!(a == -b)) return false;
  }
  return true;
}

// wg_per_cu: workgroups per CU to aim for when the scan points alone are too
// few (column groups make up the rest).
R1Plan plan_r1(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T, int wg_per_cu = 2) {
  R1Plan p;
  const int64_t bins_pass = T <= 64 ? 64 : T <= 128 ? 128 : T <= 256 ? 256 : T <= 512 ? 512 : 1024;
  p.K = (int)(bins_pass / 64);
  p.passes = (int)((T + bins_pass - 1) / bins_pass);
  p.chunk_cols = (int)(Z > 0 ? (kChunkVoxels / Z < 64 ? kChunkVoxels / Z : 64) : 64);
  if (p.chunk_cols < 1) p.chunk_cols = 1;
  const int64_t n_cols = X * Y;
  p.total_chunks = (int)((n_cols + p.chunk_cols - 1) / p.chunk_cols);
  const int64_t scan_groups = (N + kWaves - 1) / kWaves;
  const int64_t target = wg_per_cu * (int64_t)device_cu_count();
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

R1SymPlan plan_r1_sym(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                      const R1Plan& gen) {
  R1SymPlan p;
  memset(&p, 0, sizeof(p));
  p.possible = X >= 2 && Y >= 2 && !(X & 1) && !(Y & 1) && N >= 4 && !(N & 3) &&
               Z >= 2 && 4 * (Z + 1) <= kChunkVoxels && N <= (1 << 24);
  if (!p.possible) return p;
  p.K = gen.K < 8 ? gen.K : 8;
  p.passes = (int)((T + 64 * p.K - 1) / (64 * p.K));
  const int64_t n_reps = (X / 2) * (Y / 2);
  int64_t rpc = kChunkVoxels / (4 * (Z + 1));  // column quads plus a zero row each
  if (rpc > 64) rpc = 64;
  p.reps_per_chunk = (int)rpc;
  const int64_t total_chunks = (n_reps + rpc - 1) / rpc;
  const int64_t quad_blocks = (N / 4 + kWaves - 1) / kWaves;
  // four workgroups per CU and pass (two run at a time): smaller items
  // shorten the tail of the last wave of workgroups
  const int64_t target = 4 * (int64_t)device_cu_count();
  int64_t G = 1;
  if (quad_blocks < target) G = (target + quad_blocks - 1) / quad_blocks;
  if (G > total_chunks) G = total_chunks;
  p.chunks_per_group = (int)((total_chunks + G - 1) / G);
  p.n_groups = (int)((total_chunks + p.chunks_per_group - 1) / p.chunks_per_group);
  p.ml = mirror_layout(N, 0);
  p.staging_off = p.ml.end;
  const int parts = p.n_groups > gen.n_groups ? p.n_groups : gen.n_groups;
  p.ws = p.staging_off + (size_t)parts * N * T * sizeof(float);
  return p;
}

// Plan of the x-pair R1 (used when the device finds x-pairs of scan points and
// no quads the xy-mirror R1 takes): the mirror plan over column pairs of the
// half grid x < X/2.
R1SymPlan plan_r1_pair(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                       const R1Plan& gen) {
  R1SymPlan p;
  memset(&p, 0, sizeof(p));
  p.possible = X >= 2 && !(X & 1) && Y >= 1 && N >= 2 && !(N & 1) && Z >= 2 &&
               2 * (Z + 1) <= kChunkVoxels && N <= (1 << 24);
  if (!p.possible) return p;
  p.K = gen.K < 8 ? gen.K : 8;
  p.passes = (int)((T + 64 * p.K - 1) / (64 * p.K));
  const int64_t n_reps = (X / 2) * Y;
  int64_t rpc = kChunkVoxels / (2 * (Z + 1));  // column pairs plus a zero row each
  if (rpc > 64) rpc = 64;
  p.reps_per_chunk = (int)rpc;
  const int64_t total_chunks = (n_reps + rpc - 1) / rpc;
  const int64_t pair_blocks = (N / 2 + kWaves - 1) / kWaves;
  const int64_t target = 4 * (int64_t)device_cu_count();
  int64_t G = 1;
  if (pair_blocks < target) G = (target + pair_blocks - 1) / pair_blocks;
  if (G > total_chunks) G = total_chunks;
  p.chunks_per_group = (int)((total_chunks + G - 1) / G);
  p.n_groups = (int)((total_chunks + p.chunks_per_group - 1) / p.chunks_per_group);
  p.ml = mirror_layout(N, 0);
  p.staging_off = p.ml.end;
  const int parts = p.n_groups > gen.n_groups ? p.n_groups : gen.n_groups;
  p.ws = p.staging_off + (size_t)parts * N * T * sizeof(float);
  return p;
}

size_t r1_workspace(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T) {
  if (X <= 0 || Y <= 0 || Z < 2 || N <= 0 || T <= 0) return 0;
  const R1Plan p = plan_r1(X, Y, Z, N, T);
  const R1SymPlan q = plan_r1_sym(X, Y, Z, N, T, p);
  const R1SymPlan pr = plan_r1_pair(X, Y, Z, N, T, p);
  size_t ws = p.ws;
  if (q.possible && q.ws > ws) ws = q.ws;
  if (pr.possible && pr.ws > ws) ws = pr.ws;
  return ws;
}

template <int K>
void launch_r1(const R1Plan& p, const float* vol, const float* mp, float* dst,
               const ConfGeom& g, int t_base, hipStream_t st, const int* skip) {
  const int scan_groups = (g.N + kWaves - 1) / kWaves;
  const int blocks = scan_groups * p.n_groups;
  const size_t lds = 2 * (size_t)p.chunk_cols * g.Z * sizeof(float2);
  hipLaunchKernelGGL(r1_forward_kernel<K>, dim3(blocks), dim3(kWaves * 64), lds, st,
                     vol, mp, dst, g, t_base, p.n_groups, p.chunk_cols,
                     p.chunks_per_group, skip);
}

template <int K>
void launch_r1_pair(const R1SymPlan& p, const float* vol, const float* mp, const int* pairs,
                    const int* info, float* dst, const ConfGeom& g, int quad_on,
                    hipStream_t st) {
  const int pair_blocks = (g.N / 2 + kWaves - 1) / kWaves;
  const int blocks = pair_blocks * p.n_groups;
  const size_t lds = 2 * (size_t)p.reps_per_chunk * 2 * (g.Z + 1) * sizeof(float2);
  hipLaunchKernelGGL(r1_pair_kernel<K>, dim3(blocks, p.passes), dim3(kWaves * 64), lds, st,
                     vol, mp, pairs, info, dst, g, p.n_groups, p.reps_per_chunk,
                     p.chunks_per_group, quad_on);
}

template <int K>
void launch_r1_sym(const R1SymPlan& p, const float* vol, const float* mp, const int* quads,
                   const int* info, float* dst, const ConfGeom& g, hipStream_t st) {
  const int quad_blocks = (g.N / 4 + kWaves - 1) / kWaves;
  const int blocks = quad_blocks * p.n_groups;
  const size_t lds = 2 * (size_t)p.reps_per_chunk * 4 * (g.Z + 1) * sizeof(float2);
  hipLaunchKernelGGL(r1_sym_kernel<K>, dim3(blocks, p.passes), dim3(kWaves * 64), lds, st,
                     vol, mp, quads, info, dst, g, p.n_groups, p.reps_per_chunk,
                     p.chunks_per_group);
}

// Finds the scan-point quads and x-pairs on the device (no host
// synchronisation).
int launch_mirror_detect(const MirrorLayout& ml, const float* mp, int N, void* ws,
                         hipStream_t st) {
  int* partner = (int*)((char*)ws + ml.partner_off);
  int* quads = (int*)((char*)ws + ml.quads_off);
  int* pairs = (int*)((char*)ws + ml.pairs_off);
  int* info = (int*)((char*)ws + ml.info_off);
  hipLaunchKernelGGL(mirror_init_kernel, dim3(launch_grid((size_t)3 * N)), dim3(256), 0, st,
                     partner, 3 * N);
  const int items = ((N + kMirrorPts - 1) / kMirrorPts) * mirror_segments(N);
  hipLaunchKernelGGL(scan_mirror_kernel, dim3(items), dim3(64), 0, st, mp, N, partner);
  int rc = check_launch("scan_mirror_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(scan_quads_kernel, dim3(1), dim3(1024), 0, st, partner, N, quads, pairs,
                     info);
  return check_launch("scan_quads_kernel");
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
  const R1SymPlan q = plan_r1_sym(X, Y, Z, N, T, p);
  const R1SymPlan pr = plan_r1_pair(X, Y, Z, N, T, p);
  const size_t need = r1_workspace(X, Y, Z, N, T);
  // The xy-mirror path needs the volume grid antisymmetric in x and y, the
  // x-pair path in x; the scan-point quads and x-pairs are then looked for
  // on the device.  (Every grid that allows the first allows the second.)
  const bool xanti = axis_antisymmetric(volume_origin[0], volume_size[0], X);
  const bool sym = q.possible && xanti &&
                   axis_antisymmetric(volume_origin[1], volume_size[1], Y);
  const bool xsym = pr.possible && xanti;
  const bool direct = !sym && !xsym && p.n_groups == 1 && transient_dtype == TOTRI_F32;
  if (!direct && workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu",
                need, workspace_bytes);
  if (!direct) TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  const ConfGeom g = make_geom(X, Y, Z, N, T, volume_origin, volume_size,
                               bin_width, bin_offset);
  const int* info = nullptr;
  const int* skip = nullptr;
  float* dst = direct ? (float*)transient : (float*)workspace;
  if (xsym) {
    rc = launch_mirror_detect(pr.ml, measurement_points, (int)N, workspace, st);
    if (rc) return rc;
    info = (const int*)((char*)workspace + pr.ml.info_off);
    // the general kernel exits when every point has an x-mirror (info[2]):
    // one of the mirror kernels below does the work (quads imply x-pairs)
    skip = info + 2;
    dst = (float*)((char*)workspace + pr.staging_off);
  }
  if (sym) {
    const int* quads = (const int*)((char*)workspace + q.ml.quads_off);
    switch (q.K) {
      case 1: launch_r1_sym<1>(q, volume, measurement_points, quads, info, dst, g, st); break;
      case 2: launch_r1_sym<2>(q, volume, measurement_points, quads, info, dst, g, st); break;
      case 4: launch_r1_sym<4>(q, volume, measurement_points, quads, info, dst, g, st); break;
      default: launch_r1_sym<8>(q, volume, measurement_points, quads, info, dst, g, st); break;
    }
    rc = check_launch("r1_sym_kernel");
    if (rc) return rc;
  }
  if (xsym) {
    const int* pairs = (const int*)((char*)workspace + pr.ml.pairs_off);
    const int qo = sym ? 1 : 0;
    switch (pr.K) {
      case 1: launch_r1_pair<1>(pr, volume, measurement_points, pairs, info, dst, g, qo, st); break;
      case 2: launch_r1_pair<2>(pr, volume, measurement_points, pairs, info, dst, g, qo, st); break;
      case 4: launch_r1_pair<4>(pr, volume, measurement_points, pairs, info, dst, g, qo, st); break;
      default: launch_r1_pair<8>(pr, volume, measurement_points, pairs, info, dst, g, qo, st); break;
    }
    rc = check_launch("r1_pair_kernel");
    if (rc) return rc;
  }
  for (int pass = 0; pass < p.passes; ++pass) {
    const int t_base = pass * 64 * p.K;
    switch (p.K) {
      case 1: launch_r1<1>(p, volume, measurement_points, dst, g, t_base, st, skip); break;
      case 2: launch_r1<2>(p, volume, measurement_points, dst, g, t_base, st, skip); break;
      case 4: launch_r1<4>(p, volume, measurement_points, dst, g, t_base, st, skip); break;
      case 8: launch_r1<8>(p, volume, measurement_points, dst, g, t_base, st, skip); break;
      default: launch_r1<16>(p, volume, measurement_points, dst, g, t_base, st, skip); break;
    }
    rc = check_launch("r1_forward_kernel");
    if (rc) return rc;
  }
  if (direct) return TOTRI_OK;
  if (xsym) {
    const dim3 grid(launch_grid(n)), block(256);
    const int nq = sym ? q.n_groups : 0;
    switch (transient_dtype) {
      case TOTRI_F32:
        hipLaunchKernelGGL(sum_partials_flag_kernel<float>, grid, block, 0, st, dst, info,
                           nq, pr.n_groups, p.n_groups, n, (float*)transient);
        break;
      case TOTRI_BF16:
        hipLaunchKernelGGL(sum_partials_flag_kernel<__hip_bfloat16>, grid, block, 0, st, dst,
                           info, nq, pr.n_groups, p.n_groups, n, (__hip_bfloat16*)transient);
        break;
      default:
        hipLaunchKernelGGL(sum_partials_flag_kernel<__half>, grid, block, 0, st, dst, info,
                           nq, pr.n_groups, p.n_groups, n, (__half*)transient);
        break;
    }
    return check_launch("sum_partials_flag_kernel");
  }
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
  // general row walk (a1_row_kernel): 8x8 column tiles, kSlabG-voxel slabs
  int row_tiles_x, row_tiles_y, row_slabs, row_ranges;
  bool safe;
  size_t tables_bytes, partial_bytes;
  size_t row_bnd_bytes;  // boundary rows [row_ranges][X][Y][row_slabs]
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
  // at least T + 1 entries (zero past T): the row walks read one bin ahead
  p.t_stride = (int)((T + 64) & ~(int64_t)63);
  p.tables_bytes = ((size_t)2 * p.t_stride + (size_t)N * p.t_stride) * sizeof(float);
  p.tables_bytes = (p.tables_bytes + 255) & ~(size_t)255;
  p.row_tiles_x = (int)((X + 7) / 8);
  p.row_tiles_y = (int)((Y + 7) / 8);
  p.row_slabs = (int)((Z + kSlabG - 1) / kSlabG);
  {
    const int64_t items = (int64_t)p.row_tiles_x * p.row_tiles_y * p.row_slabs;
    const int64_t target = 64 * (int64_t)device_cu_count();
    int64_t R = 1;
    if (items < target) R = (target + items - 1) / items;
    if (R > N / 4) R = N / 4;
    if (R < 1) R = 1;
    p.row_ranges = (int)R;
  }
  // the row walk always writes partial volumes (its boundary rows are added
  // in the partial-sum pass)
  const int64_t parts = p.n_ranges > p.row_ranges ? p.n_ranges : p.row_ranges;
  p.partial_bytes = (size_t)parts * X * Y * Z * sizeof(float);
  p.row_bnd_bytes = (((size_t)p.row_ranges * X * Y * p.row_slabs * sizeof(float)) + 255) &
                    ~(size_t)255;
  return p;
}

// The row walks need rows that never decrease with the bin index.  They
// count bins in a float (tc = t w + off without a conversion), which is exact
// only below 2^24, so longer transients take the window walks.
constexpr int64_t kRowWalkMaxBins = (1 << 24) - 2;
bool a1_row_geometry(const float* size, int64_t T, float w, float off) {
  return T <= kRowWalkMaxBins && w > 0.f && isfinite(w) && off >= 0.f && isfinite(off) &&
         size[2] > 0.f && isfinite(size[2]);
}

}  // namespace
}  // namespace totri_hip

namespace totri_hip {
namespace {

// Plan of the x-mirror A1 (used when the device finds x-pairs of scan
// points).  Workspace: tables | partial volumes (the larger count of the two
// paths) | mirror detection.
// Plan of the xy-mirror A1 (scan points in mirror quads).
struct A1QuadPlan {
  bool possible;  // sizes allow it (X, Y even, N a multiple of 4)
  int tiles_x, tiles_y, n_slabs, n_ranges;
};

A1QuadPlan plan_a1_quad(int64_t X, int64_t Y, int64_t Z, int64_t N) {
  A1QuadPlan p;
  memset(&p, 0, sizeof(p));
  p.possible = X >= 2 && Y >= 2 && !(X & 1) && !(Y & 1) && N >= 4 && !(N & 3) &&
               N <= (1 << 24);
  if (!p.possible) return p;
  p.tiles_x = (int)((X / 2 + 7) / 8);
  p.tiles_y = (int)((Y / 2 + 7) / 8);
  p.n_slabs = (int)((Z + kSlabQ - 1) / kSlabQ);
  const int64_t items = (int64_t)p.tiles_x * p.tiles_y * p.n_slabs;
  // four single-wave workgroups per wave slot: item costs differ by several
  // times with the geometry, and with only one or two rounds the slots that
  // draw two costly items set the kernel's time.  Partial volumes then take
  // at most 64 x CU x 2048 voxels of floats (128 MiB on 256 CUs) at any
  // volume size.  Ranges keep at least 4 quads.
  const int64_t target = 64 * (int64_t)device_cu_count();
  int64_t R = 1;
  if (items < target) R = (target + items - 1) / items;
  if (R > N / 16) R = N / 16;
  if (R < 1) R = 1;
  p.n_ranges = (int)R;
  return p;
}

// Plan of the x-pair row walk: tiles of 8x8 columns of the half grid
// x < 0, 8-row slabs, pair ranges as in plan_a1_quad (at least 4 pairs each).
A1QuadPlan plan_a1_pair_rows(int64_t X, int64_t Y, int64_t Z, int64_t N) {
  A1QuadPlan p;
  memset(&p, 0, sizeof(p));
  p.possible = X >= 2 && !(X & 1) && N >= 2 && !(N & 1) && N <= (1 << 24);
  if (!p.possible) return p;
  p.tiles_x = (int)((X / 2 + 7) / 8);
  p.tiles_y = (int)((Y + 7) / 8);
  p.n_slabs = (int)((Z + kSlabQ - 1) / kSlabQ);
  const int64_t items = (int64_t)p.tiles_x * p.tiles_y * p.n_slabs;
  const int64_t target = 64 * (int64_t)device_cu_count();
  int64_t R = 1;
  if (items < target) R = (target + items - 1) / items;
  if (R > N / 8) R = N / 8;
  if (R < 1) R = 1;
  p.n_ranges = (int)R;
  return p;
}

// The xy-mirror walk needs each lane's row to grow with the bin index:
// tc = t w + off >= 0 for every bin and a positive z scale.
bool a1_quad_geometry(const float* origin, const float* size, int64_t Y, int64_t T, float w,
                      float off) {
  return a1_row_geometry(size, T, w, off) && axis_antisymmetric(origin[1], size[1], Y);
}

struct A1SymPlan {
  bool possible;  // sizes allow it (X even, N even)
  int tiles_x, tiles_y, n_slabs, n_ranges;
  A1QuadPlan quad;
  size_t partial_bytes;
  MirrorLayout ml;
  size_t cost_off, order_off;  // xy-mirror item costs (float) and dispatch order (int)
  size_t counter_off;          // xy-mirror work-queue head
  size_t bnd_off;              // xy-mirror boundary rows [R][X][Y][n_slabs]
  size_t kq_off;               // xy-mirror interleaved quad tables [N/4][T+1] float4
  A1QuadPlan prow;             // x-pair row walk
  size_t pbnd_off;             // x-pair row walk boundary rows [R][X][Y][n_slabs]
  size_t kp_off;               // x-pair row walk interleaved pair tables [N/2][T+1] float2
  size_t ws;
};

A1SymPlan plan_a1_sym(int64_t X, int64_t Y, int64_t Z, int64_t N, const A1Plan& gen) {
  A1SymPlan p;
  memset(&p, 0, sizeof(p));
  p.possible = X >= 2 && !(X & 1) && N >= 2 && !(N & 1) && N <= (1 << 24);
  if (!p.possible) return p;
  p.tiles_x = (int)((X / 2 + 7) / 8);
  p.tiles_y = (int)((Y + 7) / 8);
  p.n_slabs = (int)((Z + kSlabS - 1) / kSlabS);
  const int64_t items = (int64_t)p.tiles_x * p.tiles_y * p.n_slabs;
  const int64_t target = 16 * (int64_t)device_cu_count();
  int64_t R = 1;
  if (items < target) R = (target + items - 1) / items;
  if (R > N / 2) R = N / 2;
  if (R < 1) R = 1;
  p.n_ranges = (int)R;
  p.quad = plan_a1_quad(X, Y, Z, N);
  p.prow = plan_a1_pair_rows(X, Y, Z, N);
  int parts = p.n_ranges > gen.n_ranges ? p.n_ranges : gen.n_ranges;
  if (gen.row_ranges > parts) parts = gen.row_ranges;
  if (p.quad.possible && p.quad.n_ranges > parts) parts = p.quad.n_ranges;
  if (p.prow.possible && p.prow.n_ranges > parts) parts = p.prow.n_ranges;
  // the row walks always write partial volumes (their boundary rows are
  // added in the partial-sum pass)
  p.partial_bytes = (parts > 1 || p.quad.possible || p.prow.possible)
                        ? (size_t)parts * X * Y * Z * sizeof(float)
                        : 0;
  p.ml = mirror_layout(N, gen.tables_bytes + p.partial_bytes);
  p.ws = p.ml.end;
  if (p.quad.possible) {
    const size_t items = (size_t)p.quad.tiles_x * p.quad.tiles_y * p.quad.n_slabs *
                         p.quad.n_ranges;
    const size_t bytes = (items * sizeof(int) + 255) & ~(size_t)255;
    p.cost_off = p.ws;
    p.order_off = p.cost_off + bytes;
    p.counter_off = p.order_off + bytes;
    p.bnd_off = p.counter_off + 256;
    p.kq_off = p.bnd_off + (((size_t)p.quad.n_ranges * X * Y * p.quad.n_slabs * sizeof(float) +
                             255) & ~(size_t)255);
    // kq_stride(T) = T + 1 <= t_stride + 1
    p.ws = p.kq_off + (((size_t)(N / 4) * (gen.t_stride + 1) * sizeof(float4) + 255) &
                       ~(size_t)255);
  }
  if (p.prow.possible) {
    p.pbnd_off = p.ws;
    p.kp_off = p.pbnd_off + (((size_t)p.prow.n_ranges * X * Y * p.prow.n_slabs * sizeof(float) +
                              255) & ~(size_t)255);
    // kq_stride(T) = T + 1 <= t_stride + 1
    p.ws = p.kp_off + (((size_t)(N / 2) * (gen.t_stride + 1) * sizeof(float2) + 255) &
                       ~(size_t)255);
  }
  return p;
}

// Which A1 walks a geometry allows; the device's mirror detection then picks
// among the mirror ones.
struct A1Walks {
  bool sym;   // x-mirror paths: sizes fit and the grid is antisymmetric in x
  bool quad;  // xy-mirror row walk: also antisymmetric in y, monotone rows
  bool row;   // general row walk: monotone rows, T below the float-counter limit
  bool prow;  // x-pair row walk: x-mirror sizes and the row walk's conditions
};

A1Walks choose_a1_walks(const A1SymPlan& q, int64_t X, int64_t Y, int64_t T,
                        const float* origin, const float* size, float w, float off) {
  A1Walks k;
  k.sym = q.possible && axis_antisymmetric(origin[0], size[0], X);
  k.quad = k.sym && q.quad.possible && a1_quad_geometry(origin, size, Y, T, w, off);
  k.row = a1_row_geometry(size, T, w, off);
  k.prow = k.sym && q.prow.possible && k.row;
  return k;
}

// Offset of the general row walk's boundary rows: after everything else.
size_t a1_row_bnd_off(const A1Plan& p, const A1SymPlan& q) {
  const size_t gen = p.tables_bytes + p.partial_bytes;
  const size_t end = q.possible && q.ws > gen ? q.ws : gen;
  return (end + 255) & ~(size_t)255;
}

size_t a1_workspace(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T) {
  if (X <= 0 || Y <= 0 || Z < 2 || N <= 0 || T <= 0) return 0;
  // The size does not depend on the geometry, only on the safe flag's
  // table layout, which is the same in both forms.
  const A1Plan p = plan_a1(X, Y, Z, N, T, 1.f, 1.f);
  const A1SymPlan q = plan_a1_sym(X, Y, Z, N, p);
  return a1_row_bnd_off(p, q) + p.row_bnd_bytes;
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_volume_confocal_grad_volume_workspace(int64_t X, int64_t Y,
                                                              int64_t Z, int64_t N,
                                                              int64_t T) {
  return a1_workspace(X, Y, Z, N, T);
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
  const A1SymPlan q = plan_a1_sym(X, Y, Z, N, p);
  const A1Walks wk = choose_a1_walks(q, X, Y, T, volume_origin, volume_size, bin_width,
                                     bin_offset);
  const bool sym = wk.sym, quad = wk.quad;
  const size_t need = a1_workspace(X, Y, Z, N, T);
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
  const int* info = nullptr;
  const bool row = wk.row;
  float* row_bnd = (float*)((char*)workspace + a1_row_bnd_off(p, q));
  const bool prow = wk.prow;
  const bool staged = p.n_ranges > 1 || (sym && q.n_ranges > 1) || quad || row || prow;
  float* dst = staged ? partial : volume;
  if (sym) {
    rc = launch_mirror_detect(q.ml, measurement_points, (int)N, workspace, st);
    if (rc) return rc;
    info = (const int*)((char*)workspace + q.ml.info_off);
    const int* pairs = (const int*)((char*)workspace + q.ml.pairs_off);
    if (quad) {
      const int* quads = (const int*)((char*)workspace + q.ml.quads_off);
      float* cost = (float*)((char*)workspace + q.cost_off);
      int* order = (int*)((char*)workspace + q.order_off);
      float* bnd = (float*)((char*)workspace + q.bnd_off);
      int* counter = (int*)((char*)workspace + q.counter_off);
      float4* Kq = (float4*)((char*)workspace + q.kq_off);
      hipLaunchKernelGGL(a1_quad_table_kernel,
                         dim3(launch_grid((size_t)(N / 4) * kq_stride((int)T))), dim3(256), 0, st, Kf,
                         quads, info, p.t_stride, (int)T, (int)(N / 4), Kq);
      const A1QuadPlan& qd = q.quad;
      const int items = qd.tiles_x * qd.tiles_y * qd.n_slabs * qd.n_ranges;
      hipLaunchKernelGGL(a1_quad_cost_kernel, dim3(items), dim3(64), 0, st, measurement_points,
                         quads, info, g, qd.tiles_x, qd.tiles_y, qd.n_slabs, qd.n_ranges, items,
                         cost);
      hipLaunchKernelGGL(a1_quad_order_kernel, dim3(1), dim3(1024), 0, st, cost, info, items,
                         order, counter);
      // one persistent wave per wave slot (16 per CU: four workgroups of
      // kQuadWaves waves, each with 40 KiB of LDS)
      const int waves = items < 16 * device_cu_count() ? items : 16 * device_cu_count();
      const int blocks = (waves + kQuadWaves - 1) / kQuadWaves;
      const bool tab = T <= kQuadTabMaxT;
      auto quad_kernel = p.safe ? (tab ? a1_quad_kernel<true, true> : a1_quad_kernel<true, false>)
                                : (tab ? a1_quad_kernel<false, true> : a1_quad_kernel<false, false>);
      hipLaunchKernelGGL(quad_kernel, dim3(blocks), dim3(kQuadWaves * 64), 0, st, Kq, tc2,
                         itc4, measurement_points, quads, info, dst, bnd, g, qd.tiles_x,
                         qd.tiles_y, qd.n_slabs, qd.n_ranges, p.t_stride, items, order,
                         counter);
      rc = check_launch("a1_quad_kernel");
      if (rc) return rc;
    }
    if (prow) {
      const A1QuadPlan& pr = q.prow;
      float2* Kp = (float2*)((char*)workspace + q.kp_off);
      float* pbnd = (float*)((char*)workspace + q.pbnd_off);
      hipLaunchKernelGGL(a1_pair_table_kernel,
                         dim3(launch_grid((size_t)(N / 2) * kq_stride((int)T))), dim3(256), 0, st, Kf,
                         pairs, info, (int)quad, p.t_stride, (int)T, (int)(N / 2), Kp);
      const int items = pr.tiles_x * pr.tiles_y * pr.n_slabs * pr.n_ranges;
      if (p.safe)
        hipLaunchKernelGGL(a1_pair_kernel<true>, dim3(items), dim3(64), 0, st, Kp, tc2, itc4,
                           measurement_points, pairs, info, dst, pbnd, g, pr.tiles_x,
                           pr.tiles_y, pr.n_slabs, pr.n_ranges, p.t_stride, (int)quad);
      else
        hipLaunchKernelGGL(a1_pair_kernel<false>, dim3(items), dim3(64), 0, st, Kp, tc2, itc4,
                           measurement_points, pairs, info, dst, pbnd, g, pr.tiles_x,
                           pr.tiles_y, pr.n_slabs, pr.n_ranges, p.t_stride, (int)quad);
      rc = check_launch("a1_pair_kernel");
      if (rc) return rc;
    } else {
      const int items = q.tiles_x * q.tiles_y * q.n_slabs * q.n_ranges;
      const int blocks = (items + kA1SymWaves - 1) / kA1SymWaves;
      if (p.safe)
        hipLaunchKernelGGL(a1_sym_kernel<true>, dim3(blocks), dim3(kA1SymWaves * 64), 0, st,
                           Kf, tc2, itc4, measurement_points, pairs, info, dst, g, q.tiles_x,
                           q.tiles_y, q.n_slabs, q.n_ranges, p.t_stride, items, (int)quad);
      else
        hipLaunchKernelGGL(a1_sym_kernel<false>, dim3(blocks), dim3(kA1SymWaves * 64), 0, st,
                           Kf, tc2, itc4, measurement_points, pairs, info, dst, g, q.tiles_x,
                           q.tiles_y, q.n_slabs, q.n_ranges, p.t_stride, items, (int)quad);
      rc = check_launch("a1_sym_kernel");
      if (rc) return rc;
    }
  }
  const int* skip = info ? info + 2 : nullptr;
  if (row) {
    const int items = p.row_tiles_x * p.row_tiles_y * p.row_slabs * p.row_ranges;
    const int blocks = (items + kRowWaves - 1) / kRowWaves;
    const bool tab = T <= kRowTabMaxT;
    auto row_kernel = p.safe ? (tab ? a1_row_kernel<true, true> : a1_row_kernel<true, false>)
                             : (tab ? a1_row_kernel<false, true> : a1_row_kernel<false, false>);
    hipLaunchKernelGGL(row_kernel, dim3(blocks), dim3(kRowWaves * 64), 0, st, Kf, itc4,
                       measurement_points, dst, row_bnd, g, p.row_tiles_x, p.row_tiles_y,
                       p.row_slabs, p.row_ranges, p.t_stride, items, skip);
    rc = check_launch("a1_row_kernel");
  } else {
    const int items = p.tiles_x * p.tiles_y * p.n_slabs * p.n_ranges;
    const int blocks = (items + kA1Waves - 1) / kA1Waves;
    if (p.safe)
      hipLaunchKernelGGL(a1_transpose_kernel<true>, dim3(blocks), dim3(kA1Waves * 64), 0, st,
                         Kf, tc2, itc4, measurement_points, dst, g, p.tiles_x, p.tiles_y,
                         p.n_slabs, p.n_ranges, p.t_stride, items, skip);
    else
      hipLaunchKernelGGL(a1_transpose_kernel<false>, dim3(blocks), dim3(kA1Waves * 64), 0,
                         st, Kf, tc2, itc4, measurement_points, dst, g, p.tiles_x,
                         p.tiles_y, p.n_slabs, p.n_ranges, p.t_stride, items, skip);
    rc = check_launch("a1_transpose_kernel");
  }
  if (rc) return rc;
  if (!staged) return TOTRI_OK;
  if (sym || row) {
    // info[0] and info[2] say which path ran (no info: the general one)
    hipLaunchKernelGGL(sum_partials_a1_kernel, dim3(launch_grid(nv)), dim3(256), 0, st,
                       partial, sym ? (const float*)((char*)workspace + q.bnd_off) : nullptr,
                       info, (int)quad, q.quad.n_ranges, prow ? q.prow.n_ranges : q.n_ranges,
                       row ? p.row_ranges : p.n_ranges, nv, (int)Z,
                       (int)((Z + kSlabQ - 1) / kSlabQ), p.row_slabs,
                       row ? row_bnd : nullptr,
                       prow ? (const float*)((char*)workspace + q.pbnd_off) : nullptr, volume);
    return check_launch("sum_partials_a1_kernel");
  }
  hipLaunchKernelGGL(sum_partials_kernel<float>, dim3(launch_grid(nv)), dim3(256), 0, st,
                     partial, p.n_ranges, nv, volume);
  return check_launch("sum_partials_kernel");
}

#include "confocal_batch.inc"

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
  const int zi = confocal_decide(g, d2, wz, c0, 0, delta, 0.5f - delta, &al);
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
               const ConfGeom& g, int t_base, int pass, hipStream_t st, const int* skip) {
  const int scan_groups = (g.N + kWaves - 1) / kWaves;
  const size_t lds = (size_t)p.chunk_cols * g.Z * sizeof(float2);
  hipLaunchKernelGGL((a6_kernel<K, GT>), dim3(scan_groups * p.n_groups), dim3(kWaves * 64),
                     lds, st, vol, mp, G, part, g, t_base, pass, p.n_groups, p.chunk_cols,
                     p.chunks_per_group, skip);
}

struct A6Plan {
  R1Plan r;
  int K, passes;
  size_t ws;
};

A6Plan plan_a6(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T) {
  A6Plan p;
  // two rounds of workgroups: A6's per-scan-point work is uneven, and a
  // second round shortens the tail
  p.r = plan_r1(X, Y, Z, N, T, 4);
  p.K = p.r.K < 8 ? p.r.K : 8;
  p.passes = (int)((T + 64 * p.K - 1) / (64 * p.K));
  p.ws = (size_t)p.passes * p.r.n_groups * 3 * N * sizeof(float);
  return p;
}

// Plan of the symmetric A6: the symmetric R1's grid (quad blocks x column
// groups, the same chunks) with K <= 4 windows (three accumulators per scan
// point and window).  Workspace: mirror detection | partials of the larger
// of the two paths.
struct A6SymPlan {
  bool possible;
  int K, passes, n_groups, reps_per_chunk, chunks_per_group;
  MirrorLayout ml;
  size_t part_off, ws;
};

A6SymPlan plan_a6_sym(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                      const A6Plan& gen) {
  A6SymPlan p;
  memset(&p, 0, sizeof(p));
  const R1SymPlan r = plan_r1_sym(X, Y, Z, N, T, gen.r);
  p.possible = r.possible;
  if (!p.possible) return p;
  p.K = r.K < 4 ? r.K : 4;
  p.passes = (int)((T + 64 * p.K - 1) / (64 * p.K));
  p.n_groups = r.n_groups;
  p.reps_per_chunk = r.reps_per_chunk;
  p.chunks_per_group = r.chunks_per_group;
  p.ml = mirror_layout(N, 0);
  p.part_off = (p.ml.end + 255) & ~(size_t)255;
  const size_t parts_s = (size_t)p.passes * p.n_groups;
  const size_t parts_g = (size_t)gen.passes * gen.r.n_groups;
  p.ws = p.part_off + (parts_s > parts_g ? parts_s : parts_g) * 3 * N * sizeof(float);
  return p;
}

// Plan of the x-pair A6: the x-pair R1's grid with K <= 4 windows.
A6SymPlan plan_a6_pair(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                       const A6Plan& gen) {
  A6SymPlan p;
  memset(&p, 0, sizeof(p));
  const R1SymPlan r = plan_r1_pair(X, Y, Z, N, T, gen.r);
  p.possible = r.possible;
  if (!p.possible) return p;
  p.K = r.K < 4 ? r.K : 4;
  p.passes = (int)((T + 64 * p.K - 1) / (64 * p.K));
  p.n_groups = r.n_groups;
  p.reps_per_chunk = r.reps_per_chunk;
  p.chunks_per_group = r.chunks_per_group;
  p.ml = mirror_layout(N, 0);
  p.part_off = (p.ml.end + 255) & ~(size_t)255;
  const size_t parts_s = (size_t)p.passes * p.n_groups;
  const size_t parts_g = (size_t)gen.passes * gen.r.n_groups;
  p.ws = p.part_off + (parts_s > parts_g ? parts_s : parts_g) * 3 * N * sizeof(float);
  return p;
}

size_t a6_workspace(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T) {
  const A6Plan p = plan_a6(X, Y, Z, N, T);
  const A6SymPlan q = plan_a6_sym(X, Y, Z, N, T, p);
  const A6SymPlan pr = plan_a6_pair(X, Y, Z, N, T, p);
  size_t ws = p.ws;
  if (q.possible && q.ws > ws) ws = q.ws;
  if (pr.possible && pr.ws > ws) ws = pr.ws;
  return ws;
}

template <typename GT>
int run_a6(const A6Plan& p, const float* vol, const float* mp, const void* Gv, float* part,
           const ConfGeom& g, hipStream_t st, const int* skip) {
  const GT* G = (const GT*)Gv;
  for (int pass = 0; pass < p.passes; ++pass) {
    const int t_base = pass * 64 * p.K;
    switch (p.K) {
      case 1: launch_a6<1, GT>(p.r, vol, mp, G, part, g, t_base, pass, st, skip); break;
      case 2: launch_a6<2, GT>(p.r, vol, mp, G, part, g, t_base, pass, st, skip); break;
      case 4: launch_a6<4, GT>(p.r, vol, mp, G, part, g, t_base, pass, st, skip); break;
      default: launch_a6<8, GT>(p.r, vol, mp, G, part, g, t_base, pass, st, skip); break;
    }
    int rc = check_launch("a6_kernel");
    if (rc) return rc;
  }
  return TOTRI_OK;
}

template <typename GT>
int run_a6_sym(const A6SymPlan& q, const float* vol, const float* mp, const int* quads,
               const int* info, const void* Gv, float* part, const ConfGeom& g,
               hipStream_t st) {
  const GT* G = (const GT*)Gv;
  const int quad_blocks = (g.N / 4 + kWaves - 1) / kWaves;
  const dim3 grid(quad_blocks * q.n_groups, q.passes);
  const size_t lds = 2 * (size_t)q.reps_per_chunk * 4 * (g.Z + 1) * sizeof(float2);
  switch (q.K) {
    case 1:
      hipLaunchKernelGGL((a6_sym_kernel<1, GT>), grid, dim3(kWaves * 64), lds, st, vol, mp,
                         quads, info, G, part, g, q.n_groups, q.reps_per_chunk,
                         q.chunks_per_group);
      break;
    case 2:
      hipLaunchKernelGGL((a6_sym_kernel<2, GT>), grid, dim3(kWaves * 64), lds, st, vol, mp,
                         quads, info, G, part, g, q.n_groups, q.reps_per_chunk,
                         q.chunks_per_group);
      break;
    default:
      hipLaunchKernelGGL((a6_sym_kernel<4, GT>), grid, dim3(kWaves * 64), lds, st, vol, mp,
                         quads, info, G, part, g, q.n_groups, q.reps_per_chunk,
                         q.chunks_per_group);
      break;
  }
  return check_launch("a6_sym_kernel");
}

template <typename GT>
int run_a6_pair(const A6SymPlan& q, const float* vol, const float* mp, const int* pairs,
                const int* info, const void* Gv, float* part, const ConfGeom& g, int quad_on,
                hipStream_t st) {
  const GT* G = (const GT*)Gv;
  const int pair_blocks = (g.N / 2 + kWaves - 1) / kWaves;
  const dim3 grid(pair_blocks * q.n_groups, q.passes);
  const size_t lds = 2 * (size_t)q.reps_per_chunk * 2 * (g.Z + 1) * sizeof(float2);
  switch (q.K) {
    case 1:
      hipLaunchKernelGGL((a6_pair_kernel<1, GT>), grid, dim3(kWaves * 64), lds, st, vol, mp,
                         pairs, info, G, part, g, q.n_groups, q.reps_per_chunk,
                         q.chunks_per_group, quad_on);
      break;
    case 2:
      hipLaunchKernelGGL((a6_pair_kernel<2, GT>), grid, dim3(kWaves * 64), lds, st, vol, mp,
                         pairs, info, G, part, g, q.n_groups, q.reps_per_chunk,
                         q.chunks_per_group, quad_on);
      break;
    default:
      hipLaunchKernelGGL((a6_pair_kernel<4, GT>), grid, dim3(kWaves * 64), lds, st, vol, mp,
                         pairs, info, G, part, g, q.n_groups, q.reps_per_chunk,
                         q.chunks_per_group, quad_on);
      break;
  }
  return check_launch("a6_pair_kernel");
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_volume_confocal_grad_measurement_points_workspace(
    int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T) {
  if (X <= 0 || Y <= 0 || Z < 2 || N <= 0 || T <= 0) return 0;
  return a6_workspace(X, Y, Z, N, T);
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
  const A6SymPlan q = plan_a6_sym(X, Y, Z, N, T, p);
  const A6SymPlan pr = plan_a6_pair(X, Y, Z, N, T, p);
  // As for R1: the xy-mirror path needs the voxel grid antisymmetric in x
  // and y, the x-pair path in x; the scan quads and x-pairs are then looked
  // for on the device.
  const bool xanti = axis_antisymmetric(volume_origin[0], volume_size[0], X);
  const bool sym = q.possible && xanti &&
                   axis_antisymmetric(volume_origin[1], volume_size[1], Y);
  const bool xsym = pr.possible && xanti;
  const size_t need = a6_workspace(X, Y, Z, N, T);
  if (workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", need,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  const ConfGeom g = make_geom(X, Y, Z, N, T, volume_origin, volume_size, bin_width,
                               bin_offset);
  float* part = (float*)((char*)workspace + (xsym ? pr.part_off : 0));
  const int* info = nullptr;
  const int* skip = nullptr;
  if (xsym) {  // (sym implies xsym; both use the same detection layout)
    rc = launch_mirror_detect(pr.ml, measurement_points, (int)N, workspace, st);
    if (rc) return rc;
    info = (const int*)((char*)workspace + pr.ml.info_off);
    skip = info + 2;  // a mirror kernel takes every scan whose points all have x-mirrors
    const int* pairs = (const int*)((char*)workspace + pr.ml.pairs_off);
    const int qo = sym ? 1 : 0;
    switch (transient_dtype) {
      case TOTRI_F32: rc = run_a6_pair<float>(pr, volume, measurement_points, pairs, info, transient_grad, part, g, qo, st); break;
      case TOTRI_BF16: rc = run_a6_pair<__hip_bfloat16>(pr, volume, measurement_points, pairs, info, transient_grad, part, g, qo, st); break;
      default: rc = run_a6_pair<__half>(pr, volume, measurement_points, pairs, info, transient_grad, part, g, qo, st); break;
    }
    if (rc) return rc;
  }
  if (sym) {
    const int* quads = (const int*)((char*)workspace + q.ml.quads_off);
    switch (transient_dtype) {
      case TOTRI_F32: rc = run_a6_sym<float>(q, volume, measurement_points, quads, info, transient_grad, part, g, st); break;
      case TOTRI_BF16: rc = run_a6_sym<__hip_bfloat16>(q, volume, measurement_points, quads, info, transient_grad, part, g, st); break;
      default: rc = run_a6_sym<__half>(q, volume, measurement_points, quads, info, transient_grad, part, g, st); break;
    }
    if (rc) return rc;
  }
  switch (transient_dtype) {
    case TOTRI_F32: rc = run_a6<float>(p, volume, measurement_points, transient_grad, part, g, st, skip); break;
    case TOTRI_BF16: rc = run_a6<__hip_bfloat16>(p, volume, measurement_points, transient_grad, part, g, st, skip); break;
    default: rc = run_a6<__half>(p, volume, measurement_points, transient_grad, part, g, st, skip); break;
  }
  if (rc) return rc;
  const size_t n = (size_t)3 * N;
  if (xsym) {
    // info[0] and info[2] say which path ran
    hipLaunchKernelGGL(sum_partials_flag_kernel<float>, dim3(launch_grid(n)), dim3(256), 0, st,
                       part, info, sym ? q.passes * q.n_groups : 0, pr.passes * pr.n_groups,
                       p.passes * p.r.n_groups, n, measurement_points_grad);
    return check_launch("sum_partials_flag_kernel");
  }
  hipLaunchKernelGGL(sum_partials_kernel<float>, dim3(launch_grid(n)), dim3(256), 0, st, part,
                     p.passes * p.r.n_groups, n, measurement_points_grad);
  return check_launch("sum_partials_kernel");
}

// Host-side: the A1 walks the planner allows for a geometry, as bits
// (1 x-mirror paths, 2 xy-mirror row walk, 4 general row walk, 8 x-pair row
// walk); no device work.
extern "C" int totri_debug_a1_walks(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                                    const float volume_origin[3], const float volume_size[3],
                                    float bin_width, float bin_offset) {
  int rc = check_common(X, Y, Z, N, T, volume_origin, volume_size, bin_width, bin_offset,
                        TOTRI_F32);
  if (rc) return -rc;
  const A1Plan p = plan_a1(X, Y, Z, N, T, bin_width, bin_offset);
  const A1SymPlan q = plan_a1_sym(X, Y, Z, N, p);
  const A1Walks k = choose_a1_walks(q, X, Y, T, volume_origin, volume_size, bin_width,
                                    bin_offset);
  return (k.sym ? 1 : 0) | (k.quad ? 2 : 0) | (k.row ? 4 : 0) | (k.prow ? 8 : 0);
}

extern "C" int totri_debug_r1_paths(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                                    const float volume_origin[3], const float volume_size[3]) {
  int rc = check_common(X, Y, Z, N, T, volume_origin, volume_size, 1.f, 0.f, TOTRI_F32);
  if (rc) return -rc;
  if (X <= 0 || Y <= 0 || Z < 2 || N <= 0 || T <= 0) return 0;
  const R1Plan p = plan_r1(X, Y, Z, N, T);
  const R1SymPlan q = plan_r1_sym(X, Y, Z, N, T, p);
  const R1SymPlan pr = plan_r1_pair(X, Y, Z, N, T, p);
  const bool xanti = axis_antisymmetric(volume_origin[0], volume_size[0], X);
  const bool sym = q.possible && xanti && axis_antisymmetric(volume_origin[1], volume_size[1], Y);
  const bool xsym = pr.possible && xanti;
  return (sym ? 1 : 0) | (xsym ? 2 : 0);
}

extern "C" size_t totri_debug_mirror_workspace(int64_t N) {
  if (N <= 0) return 0;
  return mirror_layout(N, 0).end;
}

extern "C" int totri_debug_mirror_info(const float* measurement_points, int64_t N,
                                       void* workspace, size_t workspace_bytes,
                                       int32_t* info_out, void* stream) {
  TOTRI_CHECK_ARG(N >= 1 && N <= (1 << 24), "N=%lld outside [1, 2^24]", (long long)N);
  TOTRI_CHECK_ARG(measurement_points && workspace && info_out, "null argument");
  const MirrorLayout ml = mirror_layout(N, 0);
  if (workspace_bytes < ml.end)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", ml.end,
                workspace_bytes);
  hipStream_t st = (hipStream_t)stream;
  int rc = launch_mirror_detect(ml, measurement_points, (int)N, workspace, st);
  if (rc) return rc;
  TOTRI_CHECK_HIP(hipMemcpyAsync(info_out, (char*)workspace + ml.info_off, 4 * sizeof(int),
                                 hipMemcpyDeviceToDevice, st));
  return TOTRI_OK;
}
