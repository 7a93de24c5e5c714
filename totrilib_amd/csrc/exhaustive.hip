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
// LDS, the copies are summed in a fixed order and written once.  Occupied
// voxels come from a compacted list built in index order, so a wave's lanes
// are all doing useful work and the result is reproducible.
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
// Bins per R2 histogram tile: the waves' private histograms fill 64 KiB of
// LDS.  Longer transients take several tiles (grid z), each pass recomputing
// the voxel spans and keeping the bins of its tile.
constexpr int kR2Tile = 64 * 1024 / (kR2Waves * (int)sizeof(float));
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
__device__ __forceinline__ Span voxel_span(const ExGeom& g, float occ, float xf, float yf,
                                           float zf0, float zf1, float spx, float spy,
                                           float spz, float lpx, float lpy, float lpz) {
  Span r;
  float d0 = dist3(spx, spy, spz, xf, yf, zf0) + dist3(lpx, lpy, lpz, xf, yf, zf0);
  float d1 = dist3(spx, spy, spz, xf, yf, zf1) + dist3(lpx, lpy, lpz, xf, yf, zf1);
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

// ---- compaction of occupied voxels (index order) -------------------------
// ---- R2 -------------------------------------------------------------------
__global__ __launch_bounds__(kR2Waves * 64) void r2_forward_kernel(
    const float* __restrict__ vol, const int* __restrict__ list,
    const int* __restrict__ totals, const float* __restrict__ scan,  // [B][3][S]
    const float* __restrict__ laser,                                  // [B][3][L]
    float* __restrict__ transient,                                    // [B][T][S][L]
    ExGeom g, int S, int L, int tile) {
  extern __shared__ float hist[];  // [kR2Waves][tile]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int b = blockIdx.y;
  const int pair = blockIdx.x;
  const int s = pair / L, l = pair - s * L;
  const int nvox = g.D * g.H * g.W;
  const int t0 = (int)blockIdx.z * tile;  // this block's bins t0 .. t0 + tn - 1
  const int tn = min(tile, g.T - t0);
  float* h = hist + wave * tile;
  for (int t = lane; t < tn; t += 64) h[t] = 0.f;
  const float spx = scan[(size_t)b * 3 * S + s], spy = scan[(size_t)b * 3 * S + S + s],
              spz = scan[(size_t)b * 3 * S + 2 * S + s];
  const float lpx = laser[(size_t)b * 3 * L + l], lpy = laser[(size_t)b * 3 * L + L + l],
              lpz = laser[(size_t)b * 3 * L + 2 * L + l];
  const int n = totals[b];
  const int* lst = list + (size_t)b * nvox;
  const float* vb = vol + (size_t)b * nvox;
  for (int k0 = wave * 64; k0 < n; k0 += kR2Waves * 64) {
    const int k = k0 + lane;
    if (k < n) {
      const int v = lst[k];
      const int xi = v % g.W, yi = (v / g.W) % g.H, zi = v / (g.W * g.H);
      float xf, yf, zf0, zf1;
      voxel_faces(g, xi, yi, zi, &xf, &yf, &zf0, &zf1);
      const Span sp = voxel_span(g, vb[v], xf, yf, zf0, zf1, spx, spy, spz, lpx, lpy, lpz);
      // The span's bins inside this tile, in increasing order: wf at the
      // first bin, wl at the last (when it differs), wm between.  Lanes of a
      // wave may hit the same bins: LDS atomics on the wave's private
      // histogram.
      const int lo = max(sp.first, t0), hi = min(sp.last, t0 + tn - 1);
      for (int bn = lo; bn <= hi; ++bn) {
        const float wgt = bn == sp.first ? sp.wf : (bn == sp.last ? sp.wl : sp.wm);
        atomicAdd(h + (bn - t0), wgt);
      }
    }
  }
  __syncthreads();
  // Fixed-order reduction of the wave histograms; one write per bin.
  float* out = transient + (size_t)b * g.T * S * L;
  for (int t = threadIdx.x; t < tn; t += blockDim.x) {
    float v = hist[t];
    for (int w = 1; w < kR2Waves; ++w) v += hist[w * tile + t];
    out[((size_t)(t0 + t) * S + s) * L + l] = v;
  }
}

// ---- A5 -------------------------------------------------------------------
__global__ __launch_bounds__(256) void r2_adjoint_kernel(
    const float* __restrict__ gT,      // [B][S][L][T] transposed upstream gradient
    const float* __restrict__ scan, const float* __restrict__ laser,
    float* __restrict__ grad,          // [B][D][H][W]
    ExGeom g, int S, int L, int rows_per_stage) {
  extern __shared__ float rows[];
  const int b = blockIdx.y;
  const int nvox = g.D * g.H * g.W;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = v < nvox;
  const int vv = ok ? v : 0;
  const int xi = vv % g.W, yi = (vv / g.W) % g.H, zi = vv / (g.W * g.H);
  float xf, yf, zf0, zf1;
  voxel_faces(g, xi, yi, zi, &xf, &yf, &zf0, &zf1);
  const float* sb = scan + (size_t)b * 3 * S;
  const float* lb = laser + (size_t)b * 3 * L;
  const float* gb = gT + (size_t)b * S * L * g.T;
  const int nrows = S * L;
  float acc = 0.f;
  // rows_per_stage == 0 (rows longer than the LDS stage): read the rows from
  // global memory; the choice is uniform over the launch.
  const bool staged = rows_per_stage > 0;
  const int step = staged ? rows_per_stage : 1;
  for (int r0 = 0; r0 < nrows; r0 += step) {
    const int nr = min(step, nrows - r0);
    const float* src = gb + (size_t)r0 * g.T;
    if (staged) {
      __syncthreads();
      for (int e = threadIdx.x; e < nr * g.T; e += blockDim.x) rows[e] = src[e];
      __syncthreads();
      src = rows;
    }
    for (int k = 0; k < nr; ++k) {
      const int r = r0 + k;
      const int s = r / L, l = r - s * L;
      const Span sp = voxel_span(g, 1.f, xf, yf, zf0, zf1, sb[s], sb[S + s], sb[2 * S + s],
                                 lb[l], lb[L + l], lb[2 * L + l]);
      if (sp.first > sp.last) continue;
      const float* row = src + k * g.T;
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

extern "C" size_t totri_volume_render_exhaustive_forward_workspace(int64_t B, int64_t D,
                                                                   int64_t H, int64_t W) {
  return compact_layout(B, D * H * W, 0).end;
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
  const CompactLayout cl = compact_layout(B, (int64_t)nvox, 0);
  char* ws = (char*)workspace;
  const int* totals = (const int*)(ws + cl.totals_off);
  const int* list = (const int*)(ws + cl.list_off);
  if (nvox > 0) {
    rc = launch_compact(volume, (int)B, (int)nvox, cl, workspace, NonZero(), st);
    if (rc) return rc;
  } else {
    hipLaunchKernelGGL(compact_scan_kernel, dim3((unsigned)B), dim3(64), 0, st,
                       (int*)(ws + cl.counts_off), 0, (int*)(ws + cl.totals_off));
    rc = check_launch("compact_scan_kernel");
    if (rc) return rc;
  }
  const int tile = T < kR2Tile ? (int)T : kR2Tile;
  const int64_t tiles = (T + tile - 1) / tile;
  TOTRI_CHECK_ARG(S * L < (1ll << 31) && B <= 65535 && tiles <= 65535,
                  "launch grid too large (S*L=%lld, B=%lld, %lld bin tiles)",
                  (long long)(S * L), (long long)B, (long long)tiles);
  hipLaunchKernelGGL(r2_forward_kernel, dim3((unsigned)(S * L), (unsigned)B, (unsigned)tiles),
                     dim3(kR2Waves * 64), (size_t)kR2Waves * tile * sizeof(float), st, volume,
                     list, totals, scan_points, laser_points, transient, g, (int)S, (int)L,
                     tile);
  return check_launch("r2_forward_kernel");
}

extern "C" size_t totri_volume_render_exhaustive_grad_volume_workspace(int64_t B, int64_t T,
                                                                       int64_t S, int64_t L) {
  return (size_t)B * T * S * L * sizeof(float);
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
  hipStream_t st = (hipStream_t)stream;
  const size_t nvox = (size_t)D * H * W;
  if (B == 0 || nvox == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume_grad != nullptr, "null volume_grad");
  const ExGeom g = ex_geom(B, D, H, W, T, volume_start, volume_end, bin_width, bin_offset);
  const dim3 grid((unsigned)((nvox + 255) / 256), (unsigned)B);
  if (T == 0 || S == 0 || L == 0) {
    hipLaunchKernelGGL(r2_adjoint_kernel, grid, dim3(256), 0, st, (const float*)nullptr,
                       scan_points, laser_points, volume_grad, g, 0, 1, 1);
    return check_launch("r2_adjoint_kernel");
  }
  TOTRI_CHECK_ARG(transient_grad && scan_points && laser_points, "null input");
  const size_t need = totri_volume_render_exhaustive_grad_volume_workspace(B, T, S, L);
  if (workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", need,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  float* gT = (float*)workspace;
  // [B][T][S*L] -> [B][S*L][T]
  rc = transpose_batched(transient_grad, gT, (int)B, (int)T, (int)(S * L), st);
  if (rc) return rc;
  // rows longer than the LDS stage are read from global memory (rows = 0)
  const int rows = T <= kStage ? (int)(kStage / T) : 0;
  hipLaunchKernelGGL(r2_adjoint_kernel, grid, dim3(256), (size_t)rows * T * sizeof(float), st,
                     gT, scan_points, laser_points, volume_grad, g, (int)S, (int)L, rows);
  return check_launch("r2_adjoint_kernel");
}
