// Exhaustive (non-confocal) volume render R2 and its adjoint A5 for gfx950.
//
// Reference: src/render/volume_cuda.cu:297-466 (R2 kernel and the A5 launch,
// whose kernel body is empty), src/render/volume.cpp:83-111.
//
// R2 spreads each occupied voxel's occupancy over the bins of the path
// lengths |p_s - q| + |p_l - q| between its lower and upper z faces
// (volume_cuda.cu:330-375).  The reference scatters with global float
// atomics from one thread per (column, scan point, laser point).  Here one
// workgroup owns one (scan point, laser point) pair and its whole time
// histogram: every wave accumulates into a private copy of the histogram in
// LDS, the copies are summed and written once.  The histograms hold 64-bit
// fixed-point sums (integer LDS atomics, which run about three times as fast
// as float ones here): each weight is truncated to a multiple of 2^-e, with
// e chosen per batch so that no bin can overflow, and integer sums do not
// depend on the order of the atomics, so the result is the same on every
// run.  Occupied voxels come from a compacted list built in index order.
//
// A5 is written as R2's exact transpose with unit occupancy: one lane per
// voxel gathers, for every (scan, laser) pair, the weighted bins of the
// transposed upstream gradient staged in LDS.
//
// Decisions (bin span, clipping) follow the reference's FP32 formula with
// IEEE correctly rounded square roots and divisions.
#include <math.h>

#include "common.h"
#include "compact.h"

namespace totri_hip {
namespace {

constexpr int kR2Waves = 4;
// Bins per R2 histogram tile: the waves' private 64-bit histograms fill 64
// KiB of LDS.  Longer transients take several tiles (grid z), each pass
// recomputing the voxel spans and keeping the bins of its tile.
constexpr int kR2Tile = 64 * 1024 / (kR2Waves * (int)sizeof(long long));
constexpr int kStage = 8192;  // floats of staged gradient rows (32 KiB)

struct ExGeom {
  int B, D, H, W, T;
  float sx, sy, sz, ex, ey, ez;
  float w, off;
};

// Voxel geometry (volume_cuda.cu:316-319,335-338): column centre (xf, yf)
// and the z of the voxel's lower and upper faces.
__device__ __forceinline__ void voxel_faces(const ExGeom& g, int xi, int yi, int zi,
                                            float* xf, float* yf, float* zf0, float* zf1) {
  float a = ((float)xi + 0.5f) / (float)g.W;
  *xf = g.sx + a * (g.ex - g.sx);
  a = ((float)yi + 0.5f) / (float)g.H;
  *yf = g.sy + a * (g.ey - g.sy);
  a = ((float)zi + 0.5f) / (float)g.D;
  const float zf = g.sz + a * (g.ez - g.sz);
  const float h = (0.5f / (float)g.D) * (g.ez - g.sz);
  *zf0 = zf - h;
  *zf1 = zf + h;
}

struct Span {
  int first, last;      // inclusive bin span; first > last when empty
  float wf, wm, wl;     // weights of the first, inner and last bins
};

// Bin span and weights of one voxel for one (scan, laser) pair
// (volume_cuda.cu:339-374).  With first == last the whole occupancy goes to
// that bin (wf = occ).
// MID: every radicand lies in [2^-64, 2^64] (both points keep a distance
// from the box, sqrt_mid_range), so the short correctly rounded root serves.
template <bool MID>
__device__ __forceinline__ float ex_dist(float ax, float ay, float az, float bx, float by,
                                         float bz) {
  return MID ? sqrt_rn_mid(sqdist3(ax, ay, az, bx, by, bz)) : dist3(ax, ay, az, bx, by, bz);
}

template <bool MID>
__device__ __forceinline__ Span voxel_span(const ExGeom& g, float occ, float xf, float yf,
                                           float zf0, float zf1, float spx, float spy,
                                           float spz, float lpx, float lpy, float lpz) {
  Span r;
  float d0 = ex_dist<MID>(spx, spy, spz, xf, yf, zf0) + ex_dist<MID>(lpx, lpy, lpz, xf, yf, zf0);
  float d1 = ex_dist<MID>(spx, spy, spz, xf, yf, zf1) + ex_dist<MID>(lpx, lpy, lpz, xf, yf, zf1);
  if (d1 < d0) {
    const float t = d0;
    d0 = d1;
    d1 = t;
  }
  float bf0 = (d0 - g.off) / g.w;
  float bf1 = (d1 - g.off) / g.w;
  const float range = bf1 - bf0;
  int fb = (int)floorf(bf0), lb = (int)floorf(bf1);
  r.first = 1;
  r.last = 0;
  r.wf = r.wm = r.wl = 0.f;
  if (fb < 0) {
    if (lb < 0) return r;
    bf0 = 0.f;
    fb = 0;
  }
  if (lb >= g.T) {
    if (fb >= g.T) return r;
    bf1 = (float)(g.T - 1);
    lb = g.T - 1;
  }
  r.first = fb;
  r.last = lb;
  if (fb == lb) {
    r.wf = occ;
    return r;
  }
  const float frac = occ / range;
  r.wf = (1.f - (bf0 - (float)fb)) * frac;
  r.wm = frac;
  r.wl = (bf1 - (float)lb) * frac;
  return r;
}

// Flag of each point of a [B][3][N] array: 1 where every squared distance
// from it to a point of the volume box lies in [2^-64, 2^64] (it is at least
// 2^-20 from the box and at most 2^30 from its far corner), so that the
// exhaustive kernels may take the short correctly rounded root for it.
__device__ __forceinline__ bool ex_point_mid(const ExGeom& g, float px, float py, float pz) {
  const float lo = box_sqdist(px, py, pz, g.sx, g.sy, g.sz, g.ex, g.ey, g.ez);
  float hi = 0.f;
  for (int c = 0; c < 8; ++c)
    hi = fmaxf(hi, sqdist3(px, py, pz, (c & 1) ? g.ex : g.sx, (c & 2) ? g.ey : g.sy,
                           (c & 4) ? g.ez : g.sz));
  return lo >= 0x1p-40f && hi <= 0x1p60f;  // false for NaN
}

__global__ __launch_bounds__(256) void ex_point_flags_kernel(const float* __restrict__ pts,
                                                             int N, ExGeom g,
                                                             int* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.B * N) return;
  const int b = i / N, k = i - b * N;
  const float* p = pts + (size_t)b * 3 * N;
  flags[i] = ex_point_mid(g, p[k], p[N + k], p[2 * N + k]) ? 1 : 0;
}

// ---- R2 -------------------------------------------------------------------
// The occupied voxels' geometry in list order, computed once per call:
// geo[k] = {xf, yf, zf0, zf1} and occ[k] = the occupancy of list entry k;
// amax[b] collects the largest |occupancy| of batch b as float bits (zeroed
// before the launch; for non-negative floats the bit order is the value
// order, and NaN sorts above +inf).
__global__ __launch_bounds__(256) void r2_voxels_kernel(const float* __restrict__ vol,
                                                        const int* __restrict__ list,
                                                        const int* __restrict__ totals,
                                                        ExGeom g, float4* __restrict__ geo,
                                                        float* __restrict__ occ,
                                                        unsigned* __restrict__ amax) {
  const int b = blockIdx.y;
  const int nvox = g.D * g.H * g.W;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned bits = 0u;
  if (k < totals[b]) {
    const size_t bn = (size_t)b * nvox;
    const int v = list[bn + k];
    const int xi = v % g.W, yi = (v / g.W) % g.H, zi = v / (g.W * g.H);
    float xf, yf, zf0, zf1;
    voxel_faces(g, xi, yi, zi, &xf, &yf, &zf0, &zf1);
    geo[bn + k] = make_float4(xf, yf, zf0, zf1);
    const float o = vol[bn + v];
    occ[bn + k] = o;
    bits = __float_as_uint(fabsf(o));
  }
  // one atomic per wave
  for (int m = 32; m >= 1; m >>= 1) bits = max(bits, (unsigned)__shfl_xor((int)bits, m, 64));
  if ((threadIdx.x & 63) == 0 && bits) atomicMax(amax + b, bits);
}

__global__ void zero_u32_kernel(unsigned* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// trunc(x * 2^e) as an integer, for |x| * 2^e < 2^62: the float's mantissa
// shifted into place (zero for zeros and subnormals, which lie far below
// the resolution 2^-e, since the largest occupancy is normal on this path).
__device__ __forceinline__ long long r2_fixed(float x, int e) {
  const unsigned u = __float_as_uint(x);
  const int ex = (int)((u >> 23) & 255u);
  const unsigned long long m = (unsigned long long)((u & 0x7fffffu) | 0x800000u);
  const int sh = ex - 150 + e;  // x = m * 2^(ex - 150)
  unsigned long long q = sh >= 0 ? (m << min(sh, 63)) : (sh > -64 ? (m >> -sh) : 0ull);
  q = ex == 0 ? 0ull : q;
  return (u >> 31) ? -(long long)q : (long long)q;
}

// Fixed-point exponent of batch b: n occupied voxels of largest |occupancy|
// amax.  Every weight is at most |occupancy|, and a voxel adds to one bin at
// most once per pair, so a bin's sum stays below n * amax < 2^61 / 2^e.
__device__ __forceinline__ int r2_exponent(float amax, int n) {
  const int la = ilogbf(amax) + 1;    // amax < 2^la
  const int ln = 32 - __clz(n);       // n < 2^ln
  return 61 - la - ln;
}

// One workgroup per (scan point, laser point) pair and bin tile.  Lane j of
// the workgroup (0 <= j < 256) walks its own run of c = ceil(n / 256)
// consecutive occupied voxels, entries j * c .. j * c + c - 1, loading the
// next entry while it bins the current one.  So the 64 lanes of a wave work
// on voxels spread over the whole list at each step.  Batches whose
// occupancy holds an infinity or a NaN take float histograms instead (LDS
// float atomics), so that those values propagate as in the reference.
__global__ __launch_bounds__(kR2Waves * 64) void r2_forward_kernel(
    const float4* __restrict__ geo, const float* __restrict__ occ,
    const int* __restrict__ totals, const unsigned* __restrict__ amax,
    const float* __restrict__ scan,                                   // [B][3][S]
    const float* __restrict__ laser,                                  // [B][3][L]
    float* __restrict__ transient,                                    // [B][T][S][L]
    ExGeom g, int S, int L, int tile) {
  extern __shared__ long long hist64[];  // [kR2Waves][tile]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int b = blockIdx.y;
  const int pair = blockIdx.x;
  const int s = pair / L, l = pair - s * L;
  const int nvox = g.D * g.H * g.W;
  const int t0 = (int)blockIdx.z * tile;  // this block's bins t0 .. t0 + tn - 1
  const int tn = min(tile, g.T - t0);
  const int n = totals[b];
  const float am = __uint_as_float(amax[b]);
  // fixed point unless the largest |occupancy| is subnormal, inf or NaN
  // (or zero: no occupied voxel)
  const bool fixed = am >= 1.17549435e-38f && am < INFINITY;
  const int e = fixed ? r2_exponent(am, n) : 0;
  long long* h = hist64 + wave * tile;
  float* hf = (float*)hist64 + wave * tile;
  for (int t = lane; t < tn; t += 64) {
    if (fixed)
      h[t] = 0;
    else
      hf[t] = 0.f;
  }
  const float spx = scan[(size_t)b * 3 * S + s], spy = scan[(size_t)b * 3 * S + S + s],
              spz = scan[(size_t)b * 3 * S + 2 * S + s];
  const float lpx = laser[(size_t)b * 3 * L + l], lpy = laser[(size_t)b * 3 * L + L + l],
              lpz = laser[(size_t)b * 3 * L + 2 * L + l];
  const float4* gb = geo + (size_t)b * nvox;
  const float* ob = occ + (size_t)b * nvox;
  const bool mid = ex_point_mid(g, spx, spy, spz) && ex_point_mid(g, lpx, lpy, lpz);
  const int c = (n + kR2Waves * 64 - 1) / (kR2Waves * 64);
  const int k0 = (int)threadIdx.x * c, k1 = min(k0 + c, n);
  float4 f = k0 < k1 ? gb[k0] : make_float4(0.f, 0.f, 0.f, 0.f);
  float o = k0 < k1 ? ob[k0] : 0.f;
  for (int j = 0; j < c; ++j) {
    const int k = k0 + j;
    const float4 fn = k + 1 < k1 ? gb[k + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float on = k + 1 < k1 ? ob[k + 1] : 0.f;
    if (k < k1) {
      const Span sp = mid ? voxel_span<true>(g, o, f.x, f.y, f.z, f.w, spx, spy, spz, lpx, lpy, lpz)
                          : voxel_span<false>(g, o, f.x, f.y, f.z, f.w, spx, spy, spz, lpx, lpy,
                                              lpz);
      // The span's bins inside this tile, in increasing order: wf at the
      // first bin, wl at the last (when it differs), wm between.
      const int lo = max(sp.first, t0), hi = min(sp.last, t0 + tn - 1);
      if (fixed) {
        const long long qf = r2_fixed(sp.wf, e), qm = r2_fixed(sp.wm, e),
                        ql = r2_fixed(sp.wl, e);
        for (int bn = lo; bn <= hi; ++bn) {
          const long long q = bn == sp.first ? qf : (bn == sp.last ? ql : qm);
          atomicAdd((unsigned long long*)(h + (bn - t0)), (unsigned long long)q);
        }
      } else {
        for (int bn = lo; bn <= hi; ++bn) {
          const float wgt = bn == sp.first ? sp.wf : (bn == sp.last ? sp.wl : sp.wm);
          atomicAdd(hf + (bn - t0), wgt);
        }
      }
    }
    f = fn;
    o = on;
  }
  __syncthreads();
  // Sum of the wave histograms; one write per bin.
  float* out = transient + (size_t)b * g.T * S * L;
  for (int t = threadIdx.x; t < tn; t += blockDim.x) {
    float v;
    if (fixed) {
      long long q = hist64[t];
      for (int w = 1; w < kR2Waves; ++w) q += hist64[w * tile + t];
      v = (float)ldexp((double)q, -e);
    } else {
      const float* hv = (const float*)hist64;
      v = hv[t];
      for (int w = 1; w < kR2Waves; ++w) v += hv[w * tile + t];
    }
    out[((size_t)(t0 + t) * S + s) * L + l] = v;
  }
}

// ---- A5 -------------------------------------------------------------------
// Float4s per thread of one stage copy (kStage floats, 256 lanes).
constexpr int kA5Copy4 = kStage / 4 / 256;

// Row stride of the transposed gradient: T rounded up to whole float4s.
__host__ __device__ __forceinline__ int a5_stride(int T) { return (T + 3) & ~3; }

// One lane per voxel.  gT holds one row of a5_stride(T) floats per
// (scan point, laser point) pair.  A workgroup stages R rows in LDS
// (STAGED), loading the next R into registers while it works on the current
// ones, or reads them from global memory (rows longer than the stage).
template <bool STAGED>
__global__ __launch_bounds__(256) void r2_adjoint_kernel(
    const float* __restrict__ gT,      // [B][S*L][a5_stride(T)]
    const float* __restrict__ scan, const float* __restrict__ laser,
    const int* __restrict__ sflag, const int* __restrict__ lflag,  // [B][S], [B][L]
    float* __restrict__ grad,          // [B][D][H][W]
    ExGeom g, int S, int L, int R) {
  extern __shared__ float4 stage4[];
  const int b = blockIdx.y;
  const int nvox = g.D * g.H * g.W;
  const int P = a5_stride(g.T);
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = v < nvox;
  const int vv = ok ? v : 0;
  const int xi = vv % g.W, yi = (vv / g.W) % g.H, zi = vv / (g.W * g.H);
  float xf, yf, zf0, zf1;
  voxel_faces(g, xi, yi, zi, &xf, &yf, &zf0, &zf1);
  const float* sb = scan + (size_t)b * 3 * S;
  const float* lb = laser + (size_t)b * 3 * L;
  const int* sf = sflag + (size_t)b * S;
  const int* lf = lflag + (size_t)b * L;
  const int nrows = S * L;
  const float* gb = gT + (size_t)b * nrows * P;
  float4 buf[kA5Copy4];
  if (STAGED) {
    const int n4 = min(R, nrows) * P / 4;
#pragma unroll
    for (int i = 0; i < kA5Copy4; ++i) {
      const int e = i * 256 + (int)threadIdx.x;
      buf[i] = e < n4 ? ((const float4*)gb)[e] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  float acc = 0.f;
  for (int r0 = 0; r0 < nrows; r0 += R) {
    const int nr = min(R, nrows - r0);
    const float* src = gb + (size_t)r0 * P;
    if (STAGED) {
      __syncthreads();  // the previous rows are no longer read
      const int n4 = nr * P / 4;
#pragma unroll
      for (int i = 0; i < kA5Copy4; ++i) {
        const int e = i * 256 + (int)threadIdx.x;
        if (e < n4) stage4[e] = buf[i];
      }
      __syncthreads();
      if (r0 + R < nrows) {
        const int n4n = min(R, nrows - r0 - R) * P / 4;
        const float4* nx = (const float4*)(gb + (size_t)(r0 + R) * P);
#pragma unroll
        for (int i = 0; i < kA5Copy4; ++i) {
          const int e = i * 256 + (int)threadIdx.x;
          buf[i] = e < n4n ? nx[e] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
      src = (const float*)stage4;
    }
    for (int k = 0; k < nr; ++k) {
      const int r = r0 + k;
      const int s = r / L, l = r - s * L;
      const Span sp =
          sf[s] & lf[l]
              ? voxel_span<true>(g, 1.f, xf, yf, zf0, zf1, sb[s], sb[S + s], sb[2 * S + s],
                                 lb[l], lb[L + l], lb[2 * L + l])
              : voxel_span<false>(g, 1.f, xf, yf, zf0, zf1, sb[s], sb[S + s], sb[2 * S + s],
                                  lb[l], lb[L + l], lb[2 * L + l]);
      if (sp.first > sp.last) continue;
      const float* row = src + k * P;
      if (sp.first == sp.last) {
        acc += sp.wf * row[sp.first];
      } else {
        acc += sp.wf * row[sp.first];
        for (int bn = sp.first + 1; bn < sp.last; ++bn) acc += sp.wm * row[bn];
        acc += sp.wl * row[sp.last];
      }
    }
  }
  if (ok) grad[(size_t)b * nvox + v] = acc;
}

ExGeom ex_geom(int64_t B, int64_t D, int64_t H, int64_t W, int64_t T, const float* vs,
               const float* ve, float w, float off) {
  ExGeom g;
  g.B = (int)B;
  g.D = (int)D;
  g.H = (int)H;
  g.W = (int)W;
  g.T = (int)T;
  g.sx = vs[0];
  g.sy = vs[1];
  g.sz = vs[2];
  g.ex = ve[0];
  g.ey = ve[1];
  g.ez = ve[2];
  g.w = w;
  g.off = off;
  return g;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int check_sizes(int64_t B, int64_t D, int64_t H, int64_t W, int64_t T, int64_t S, int64_t L) {
  for (int64_t d : {B, D, H, W, T, S, L})
    TOTRI_CHECK_ARG(d >= 0 && d < (1ll << 31), "dimension out of range");
  TOTRI_CHECK_ARG(B * D * H * W < (1ll << 31) && B * T * S * L < (1ll << 31),
                  "tensor has more than 2^31 elements");
  return TOTRI_OK;
}

}  // namespace
}  // namespace totri_hip

using namespace totri_hip;

namespace totri_hip {
namespace {

// Workspace of R2: the compaction, then the occupied voxels' geometry
// (float4) and occupancy, [B][D*H*W] each, and the largest |occupancy| of
// each batch.
struct R2Layout {
  CompactLayout cl;
  size_t geo_off, occ_off, amax_off, end;
};

R2Layout r2_layout(int64_t B, int64_t nvox) {
  R2Layout l;
  l.cl = compact_layout(B, nvox, 0);
  l.geo_off = align256(l.cl.end);
  l.occ_off = align256(l.geo_off + (size_t)B * nvox * sizeof(float4));
  l.amax_off = align256(l.occ_off + (size_t)B * nvox * sizeof(float));
  l.end = align256(l.amax_off + (size_t)B * sizeof(unsigned));
  return l;
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_volume_render_exhaustive_forward_workspace(int64_t B, int64_t D,
                                                                   int64_t H, int64_t W) {
  return r2_layout(B, D * H * W).end;
}

extern "C" int totri_volume_render_exhaustive_forward(
    const float* volume, int64_t B, int64_t D, int64_t H, int64_t W, const float* scan_points,
    int64_t S, const float* laser_points, int64_t L, float* transient, int64_t T,
    const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_sizes(B, D, H, W, T, S, L);
  if (rc) return rc;
  TOTRI_CHECK_ARG(volume_start && volume_end, "null volume box");
  hipStream_t st = (hipStream_t)stream;
  if (B == 0 || T == 0 || S == 0 || L == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(transient && scan_points && laser_points, "null input");
  const size_t nvox = (size_t)D * H * W;
  const ExGeom g = ex_geom(B, D, H, W, T, volume_start, volume_end, bin_width, bin_offset);
  const size_t need = totri_volume_render_exhaustive_forward_workspace(B, D, H, W);
  if (workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", need,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  TOTRI_CHECK_ARG(nvox == 0 || volume != nullptr, "null volume");
  TOTRI_CHECK_ARG(((uintptr_t)workspace & 15) == 0, "workspace must be 16-byte aligned");
  const R2Layout lay = r2_layout(B, (int64_t)nvox);
  const CompactLayout& cl = lay.cl;
  char* ws = (char*)workspace;
  const int* totals = (const int*)(ws + cl.totals_off);
  const int* list = (const int*)(ws + cl.list_off);
  float4* geo = (float4*)(ws + lay.geo_off);
  float* occ = (float*)(ws + lay.occ_off);
  unsigned* amax = (unsigned*)(ws + lay.amax_off);
  TOTRI_CHECK_ARG(B <= 65535, "B=%lld too large", (long long)B);
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, amax,
                     (int)B);
  rc = check_launch("zero_u32_kernel");
  if (rc) return rc;
  if (nvox > 0) {
    rc = launch_compact(volume, (int)B, (int)nvox, cl, workspace, NonZero(), st);
    if (rc) return rc;
    hipLaunchKernelGGL(r2_voxels_kernel, dim3((unsigned)((nvox + 255) / 256), (unsigned)B),
                       dim3(256), 0, st, volume, list, totals, g, geo, occ, amax);
    rc = check_launch("r2_voxels_kernel");
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL(compact_scan_kernel, dim3((unsigned)B), dim3(kCompactScanLanes), 0, st,
                       (int*)(ws + cl.counts_off), 0, (int*)(ws + cl.totals_off));
    rc = check_launch("compact_scan_kernel");
    if (rc) return rc;
  }
  const int tile = T < kR2Tile ? (int)T : kR2Tile;
  const int64_t tiles = (T + tile - 1) / tile;
  TOTRI_CHECK_ARG(S * L < (1ll << 31) && tiles <= 65535,
                  "launch grid too large (S*L=%lld, %lld bin tiles)", (long long)(S * L),
                  (long long)tiles);
  hipLaunchKernelGGL(r2_forward_kernel, dim3((unsigned)(S * L), (unsigned)B, (unsigned)tiles),
                     dim3(kR2Waves * 64), (size_t)kR2Waves * tile * sizeof(long long), st, geo,
                     occ, totals, amax, scan_points, laser_points, transient, g, (int)S, (int)L,
                     tile);
  return check_launch("r2_forward_kernel");
}

extern "C" size_t totri_volume_render_exhaustive_grad_volume_workspace(int64_t B, int64_t T,
                                                                       int64_t S, int64_t L) {
  if (B <= 0 || T < 0 || S <= 0 || L <= 0) return 0;
  // the transposed gradient rows, then the point flags of the scan and laser points
  return align256((size_t)B * S * L * a5_stride((int)T) * sizeof(float)) +
         align256((size_t)B * (S + L) * sizeof(int));
}

extern "C" int totri_volume_render_exhaustive_grad_volume(
    const float* transient_grad, int64_t B, int64_t T, int64_t S, int64_t L,
    const float* scan_points, const float* laser_points, float* volume_grad, int64_t D,
    int64_t H, int64_t W, const float volume_start[3], const float volume_end[3],
    float bin_width, float bin_offset, void* workspace, size_t workspace_bytes,
    void* stream) {
  int rc = check_sizes(B, D, H, W, T, S, L);
  if (rc) return rc;
  TOTRI_CHECK_ARG(volume_start && volume_end, "null volume box");
  TOTRI_CHECK_ARG(T < (1 << 30), "T=%lld too large", (long long)T);
  hipStream_t st = (hipStream_t)stream;
  const size_t nvox = (size_t)D * H * W;
  if (B == 0 || nvox == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume_grad != nullptr, "null volume_grad");
  const ExGeom g = ex_geom(B, D, H, W, T, volume_start, volume_end, bin_width, bin_offset);
  const dim3 grid((unsigned)((nvox + 255) / 256), (unsigned)B);
  if (T == 0 || S == 0 || L == 0) {  // no bins or no pairs: a zero gradient
    hipLaunchKernelGGL(r2_adjoint_kernel<false>, grid, dim3(256), 0, st, (const float*)nullptr,
                       scan_points, laser_points, (const int*)nullptr, (const int*)nullptr,
                       volume_grad, g, 0, 1, 1);
    return check_launch("r2_adjoint_kernel");
  }
  TOTRI_CHECK_ARG(transient_grad && scan_points && laser_points, "null input");
  const size_t need = totri_volume_render_exhaustive_grad_volume_workspace(B, T, S, L);
  if (workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", need,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr && ((uintptr_t)workspace & 15) == 0,
                  "workspace must be a non-null, 16-byte aligned device pointer");
  float* gT = (float*)workspace;
  const int P = a5_stride((int)T);
  int* sflag = (int*)((char*)workspace + align256((size_t)B * S * L * P * sizeof(float)));
  int* lflag = sflag + B * S;
  // [B][T][S*L] -> [B][S*L][P]
  rc = transpose_strided(transient_grad, gT, (int)B, (int)T, (int)(S * L), P, 0, st);
  if (rc) return rc;
  hipLaunchKernelGGL(ex_point_flags_kernel, dim3((unsigned)((B * S + 255) / 256)), dim3(256), 0,
                     st, scan_points, (int)S, g, sflag);
  rc = check_launch("ex_point_flags_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(ex_point_flags_kernel, dim3((unsigned)((B * L + 255) / 256)), dim3(256), 0,
                     st, laser_points, (int)L, g, lflag);
  rc = check_launch("ex_point_flags_kernel");
  if (rc) return rc;
  // rows per LDS stage; rows longer than the stage are read from global memory
  const int R = (int)std::min<int64_t>(S * L, kStage / P);
  if (R > 0) {
    hipLaunchKernelGGL(r2_adjoint_kernel<true>, grid, dim3(256), (size_t)R * P * sizeof(float),
                       st, gT, scan_points, laser_points, sflag, lflag, volume_grad, g, (int)S,
                       (int)L, R);
  } else {
    hipLaunchKernelGGL(r2_adjoint_kernel<false>, grid, dim3(256), 0, st, gT, scan_points,
                       laser_points, sflag, lflag, volume_grad, g, (int)S, (int)L, (int)(S * L));
  }
  return check_launch("r2_adjoint_kernel");
}

namespace totri_hip {
namespace {

__global__ void sqrt_mid_check_kernel(uint32_t lo, uint32_t hi,
                                      unsigned long long* __restrict__ count) {
  unsigned long long bad = 0;
  const uint64_t n = (uint64_t)hi - lo + 1;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const float x = __uint_as_float((uint32_t)(lo + i));
    bad += __float_as_uint(sqrt_rn_mid(x)) != __float_as_uint(sqrt_rn(x)) ? 1 : 0;
  }
  for (int m = 32; m >= 1; m >>= 1) bad += __shfl_xor(bad, m, 64);
  if ((threadIdx.x & 63) == 0 && bad) atomicAdd(count, bad);
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ull; }

}  // namespace
}  // namespace totri_hip

extern "C" int totri_debug_sqrt_mid_mismatches(uint32_t lo_bits, uint32_t hi_bits,
                                               unsigned long long* mismatches, void* stream) {
  TOTRI_CHECK_ARG(mismatches != nullptr && lo_bits <= hi_bits, "invalid arguments");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, st, mismatches);
  int rc = check_launch("zero_u64_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(sqrt_mid_check_kernel, dim3(device_cu_count() * 16), dim3(256), 0, st,
                     lo_bits, hi_bits, mismatches);
  return check_launch("sqrt_mid_check_kernel");
}
