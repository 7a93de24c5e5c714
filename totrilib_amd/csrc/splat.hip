// Gaussian RBF -> volume splat (S1) and its parameter gradient (S2), gfx950.
//
// Reference: src/util/rbf2volume_cuda.cu:15-232, src/util/rbf2volume.cpp:7-27.
//
//   g_i(q)   = exp(-0.5 |q - c_i|^2 / sigma_i^2)             (:15-29)
//   V0(q)    = sum_i g_i(q)
//   Va(q)    = sum_i g_i(q) a_{i,a} / (V0 < 1e-8 ? 1 : V0)    (:49-93)
// on the inclusive voxel grid from `start` to `end`.  The reference sums all
// N RBFs for every voxel.  Here a truncation radius (in units of sigma) may
// be given: RBFs with |q - c|^2 > (trunc sigma)^2 are skipped (trunc <= 0
// keeps every RBF, which reproduces the reference).
//
// S1: a workgroup owns a brick of 8 x 8 x 4 voxels (one per lane).  RBFs are
// streamed through LDS in chunks; each chunk is culled against the brick's
// bounding box and the survivors are compacted in index order, so every
// voxel sums its RBFs in the reference's order.
//
// S2: one wave per RBF walks the voxels of the RBF's support box (the whole
// volume without truncation) and reduces its gradient over the wave in a
// fixed order; the reference instead launches one block per (16 columns x
// 4 slices, RBF) and adds per-block partial sums with float atomics.
#include <math.h>
#include <string.h>

#include "common.h"

namespace totri_hip {
namespace {

constexpr int kMaxAttr = 8;  // rbf2volume_cuda.cu:13
constexpr int kBX = 8, kBY = 8, kBZ = 4;
constexpr int kChunk = 256;

struct SplatGeom {
  int B, N, A, D, H, W;
  float sx, sy, sz, ex, ey, ez;
  float trunc;
};

// rbf2volume_cuda.cu:58-60: (end - start) * i / (n - 1) + start, or the box
// centre when n == 1.
__device__ __forceinline__ float splat_coord(int i, int n, float a, float b) {
  if (n == 1) return 0.5f * (a + b);
  float t = (b - a) * (float)i;
  t = t / (float)(n - 1);
  return a + t;
}

// |q - c|^2 in the reference's order (rbf2volume_cuda.cu:21-24).
__device__ __forceinline__ float rbf_r2(float x, float y, float z, float xr, float yr,
                                        float zr) {
  float r2 = (xr - x) * (xr - x);
  r2 = r2 + (yr - y) * (yr - y);
  r2 = r2 + (zr - z) * (zr - z);
  return r2;
}

// The Gaussian exp(-r2 / (2 sigma^2)) (rbf2volume_cuda.cu:15-29, written
// there as expf(-0.5 r2 / sigma^2)) as 2^(r2 * k), with the RBF's
// k = -log2(e) / (2 sigma^2) computed once per RBF: the hardware base-2
// exponential (within 1 ulp) of a product rounded once, two instructions
// per (voxel, RBF) instead of a division and expf's expansion.  The result
// differs from the reference's by a relative |r2 k| 2^-23 or so, far inside
// the splat's 1e-5 bar (its terms with |r2 k| above 30 are below 1e-9 of the
// peak).
__device__ __forceinline__ float gauss_k(float sg) { return -0.72134752f / (sg * sg); }

__device__ __forceinline__ float gauss_exp2(float r2, float k) {
  return __builtin_amdgcn_exp2f(r2 * k);
}

// Truncation test: keep the RBF when r2 <= (trunc * sigma)^2.
__device__ __forceinline__ bool within(float r2, float sg, float trunc) {
  if (!(trunc > 0.f)) return true;
  const float lim = trunc * sg;
  return !(r2 > lim * lim);
}

// Binned lists (truncated splats): per (batch, brick) the RBFs whose support
// may reach the brick, found with the brick-cull test of the full scan, so a
// brick sums the same RBFs in the same (index) order either way.  Each brick
// has `cap` list slots; a brick reached by more RBFs scans all of them.
struct SplatBins {
  int* count;  // [B][bricks] RBFs that may reach the brick (the fill cursor)
  int* list;   // [B][bricks][cap] the first cap of their indices, in fill order
  int cap;
};

constexpr int kSlotMax = 512;  // list slots per brick, at most (LDS sort buffers)
constexpr int kSlotMin = 64;

// Conservative brick cull (the full scan's test): the squared distance from
// the RBF centre to the brick's sample box against (trunc sigma)^2.
__device__ __forceinline__ bool brick_may_hold(float xr, float yr, float zr, float sg,
                                               float trunc, float lox, float hix, float loy,
                                               float hiy, float loz, float hiz) {
  const float dx = fmaxf(fmaxf(lox - xr, xr - hix), 0.f);
  const float dy = fmaxf(fmaxf(loy - yr, yr - hiy), 0.f);
  const float dz = fmaxf(fmaxf(loz - zr, zr - hiz), 0.f);
  const float lim = trunc * fabsf(sg);
  return !((dx * dx + dy * dy + dz * dz) > lim * lim * 1.0001f + 1e-12f);
}

// The sample box of voxels [x0, x1] x [y0, y1] x [z0, z1] (inclusive, inside
// the grid).
__device__ __forceinline__ void voxel_box(const SplatGeom& g, int x0, int x1, int y0, int y1,
                                          int z0, int z1, float* lox, float* hix, float* loy,
                                          float* hiy, float* loz, float* hiz) {
  const float bxa = splat_coord(x0, g.W, g.sx, g.ex), bxb = splat_coord(x1, g.W, g.sx, g.ex);
  const float bya = splat_coord(y0, g.H, g.sy, g.ey), byb = splat_coord(y1, g.H, g.sy, g.ey);
  const float bza = splat_coord(z0, g.D, g.sz, g.ez), bzb = splat_coord(z1, g.D, g.sz, g.ez);
  *lox = fminf(bxa, bxb);
  *hix = fmaxf(bxa, bxb);
  *loy = fminf(bya, byb);
  *hiy = fmaxf(bya, byb);
  *loz = fminf(bza, bzb);
  *hiz = fmaxf(bza, bzb);
}

// The sample box of brick (bx, by, bz): the voxels it really holds.
__device__ __forceinline__ void brick_box(const SplatGeom& g, int bx, int by, int bz,
                                          float* lox, float* hix, float* loy, float* hiy,
                                          float* loz, float* hiz) {
  const int x0 = bx * kBX, x1 = min(bx * kBX + kBX, g.W) - 1;
  const int y0 = by * kBY, y1 = min(by * kBY + kBY, g.H) - 1;
  const int z0 = bz * kBZ, z1 = min(bz * kBZ + kBZ, g.D) - 1;
  voxel_box(g, x0, x1, y0, y1, z0, z1, lox, hix, loy, hiy, loz, hiz);
}

// Brick index range [lo, hi] along an axis whose bricks hold `bs` samples
// each, covering the samples within [c - r, c + r] (one sample of margin; the
// exact test is brick_may_hold).
__device__ __forceinline__ void brick_range(float c, float r, int n, int bs, float a, float b,
                                            int* lo, int* hi) {
  const int nb = (n + bs - 1) / bs;
  if (n == 1) {
    *lo = 0;
    *hi = nb - 1;
    return;
  }
  const float step = (b - a) / (float)(n - 1);
  float u0 = (c - r - a) / step, u1 = (c + r - a) / step;
  if (u0 > u1) {
    const float t = u0;
    u0 = u1;
    u1 = t;
  }
  u0 = fminf(fmaxf(u0 - 1.f, -1.f), (float)n);
  u1 = fminf(fmaxf(u1 + 1.f, -1.f), (float)n);
  *lo = max((int)floorf(u0), 0) / bs;
  *hi = min((int)ceilf(u1), n - 1) / bs;
  if (!(r < INFINITY)) {
    *lo = 0;
    *hi = nb - 1;
  }
}

// One pass: per brick, the RBFs that may reach it take the next slot of the
// brick's list (the count past cap makes the brick scan instead).  One wave
// per RBF, its lanes over the bricks of the RBF's support box.
constexpr int kBinWaves = 4;

__global__ __launch_bounds__(kBinWaves * 64) void splat_bin_kernel(
    const float* __restrict__ rbf, SplatGeom g, int bricks_x, int bricks_y, int bricks_z,
    SplatBins bins) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * kBinWaves + (threadIdx.x >> 6);
  const int b = blockIdx.y;
  if (i >= g.N) return;  // whole wave; no barrier follows
  const int C = 4 + g.A;
  const float* p = rbf + ((size_t)b * g.N + i) * C;
  const float xr = p[0], yr = p[1], zr = p[2], sg = p[3];
  const float r = g.trunc * fabsf(sg) * 1.0001f + 1e-6f;
  int xa, xb, ya, yb, za, zb;
  brick_range(xr, r, g.W, kBX, g.sx, g.ex, &xa, &xb);
  brick_range(yr, r, g.H, kBY, g.sy, g.ey, &ya, &yb);
  brick_range(zr, r, g.D, kBZ, g.sz, g.ez, &za, &zb);
  const int nx = xb - xa + 1, ny = yb - ya + 1, nz = zb - za + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  const int total = nx * ny * nz;
  const size_t nb = (size_t)bricks_x * bricks_y * bricks_z;
  for (int e = lane; e < total; e += 64) {
    const int bx = xa + e % nx;
    const int by = ya + (e / nx) % ny;
    const int bz = za + e / (nx * ny);
    float lox, hix, loy, hiy, loz, hiz;
    brick_box(g, bx, by, bz, &lox, &hix, &loy, &hiy, &loz, &hiz);
    if (!brick_may_hold(xr, yr, zr, sg, g.trunc, lox, hix, loy, hiy, loz, hiz)) continue;
    const size_t k = b * nb + ((size_t)bz * bricks_y + by) * bricks_x + bx;
    const int at = atomicAdd(bins.count + k, 1);
    if (at < bins.cap) bins.list[k * bins.cap + at] = i;
  }
}

__global__ __launch_bounds__(256) void splat_forward_kernel(
    const float* __restrict__ rbf,  // [B][N][4+A]
    float* __restrict__ vol,        // [B][1+A][D][H][W]
    SplatGeom g, int bricks_x, int bricks_y, SplatBins bins, int binned) {
  __shared__ float prm[kChunk * (4 + kMaxAttr)];
  __shared__ float pk[kChunk];  // gauss_k of the staged RBFs
  __shared__ int keep[kSlotMax > kChunk ? kSlotMax : kChunk];
  __shared__ int fill[kSlotMax];
  __shared__ int wave_cnt[4];
  const int C = 4 + g.A;
  const int b = blockIdx.y;
  int brick = blockIdx.x;
  const int bx = brick % bricks_x;
  brick /= bricks_x;
  const int by = brick % bricks_y;
  const int bz = brick / bricks_y;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  // Each wave holds a 4x4x4 quarter of the brick (quarter `wave` of its 8x8
  // footprint, all 4 z), so that it can skip the RBFs that miss its quarter.
  static_assert(kBX == 8 && kBY == 8 && kBZ == 4, "four 4x4x4 quarters per brick");
  const int qx0 = bx * kBX + (wave & 1) * 4, qy0 = by * kBY + (wave >> 1) * 4;
  const int xi = qx0 + (lane & 3);
  const int yi = qy0 + ((lane >> 2) & 3);
  const int zi = bz * kBZ + (lane >> 4);
  const bool ok = xi < g.W && yi < g.H && zi < g.D;
  const float x = splat_coord(xi, g.W, g.sx, g.ex);
  const float y = splat_coord(yi, g.H, g.sy, g.ey);
  const float z = splat_coord(zi, g.D, g.sz, g.ez);
  // Brick bounding box over the voxels it really holds.
  float lox, hix, loy, hiy, loz, hiz;
  brick_box(g, bx, by, bz, &lox, &hix, &loy, &hiy, &loz, &hiz);

  float sum = 0.f;
  float attr[kMaxAttr];
#pragma unroll
  for (int a = 0; a < kMaxAttr; ++a) attr[a] = 0.f;
  const float* rb = rbf + (size_t)b * g.N * C;

  // Binned path: this brick's list, sorted into index order in LDS, then
  // evaluated in chunks.  A brick whose count passed its slots scans instead.
  int n_list = -1;
  if (binned) {
    const size_t k = (size_t)b * gridDim.x + blockIdx.x;
    const int cnt = bins.count[k];
    if (cnt <= bins.cap) n_list = cnt;
    if (n_list >= 0) {
      const int* src = bins.list + k * bins.cap;
      for (int e = tid; e < n_list; e += blockDim.x) fill[e] = src[e];
      __syncthreads();
      // Rank sort (the indices are distinct): an entry's place is the number
      // of smaller entries; every lane reads fill[j] in step (broadcast).
      for (int e = tid; e < n_list; e += blockDim.x) {
        const int v = fill[e];
        int rank = 0;
        for (int j = 0; j < n_list; ++j) rank += fill[j] < v ? 1 : 0;
        keep[rank] = v;
      }
    }
  }
  if (n_list >= 0) {
    // The wave's quarter box (empty past the grid's edge): an RBF that fails
    // the brick test on it reaches none of the wave's voxels, so skipping it
    // leaves every voxel's sum, taken in index order, unchanged.
    const int qx1 = min(qx0 + 3, g.W - 1), qy1 = min(qy0 + 3, g.H - 1);
    const int qz0 = bz * kBZ, qz1 = min(qz0 + kBZ, g.D) - 1;
    const bool qnonempty = qx0 <= qx1 && qy0 <= qy1 && qz0 <= qz1;
    float qlox = 0.f, qhix = 0.f, qloy = 0.f, qhiy = 0.f, qloz = 0.f, qhiz = 0.f;
    if (qnonempty)
      voxel_box(g, qx0, qx1, qy0, qy1, qz0, qz1, &qlox, &qhix, &qloy, &qhiy, &qloz, &qhiz);
    for (int c0 = 0; c0 < n_list; c0 += kChunk) {
      const int nc = min(kChunk, n_list - c0);
      __syncthreads();
      for (int e = tid; e < nc * C; e += blockDim.x) {
        const int r = e / C;
        prm[e] = rb[(size_t)keep[c0 + r] * C + (e - r * C)];
      }
      for (int e = tid; e < nc; e += blockDim.x) pk[e] = gauss_k(rb[(size_t)keep[c0 + e] * C + 3]);
      __syncthreads();
      for (int k0 = 0; k0 < nc; k0 += 64) {
        const int kk = k0 + lane;
        bool pass = false;
        if (qnonempty && kk < nc) {
          const float* p = prm + kk * C;
          pass = brick_may_hold(p[0], p[1], p[2], p[3], g.trunc, qlox, qhix, qloy, qhiy, qloz,
                                qhiz);
        }
        unsigned long long m = __ballot(pass);
        while (m) {  // the wave's RBFs of this group, in index order
          const int k = k0 + __builtin_ctzll(m);
          m &= m - 1;
          const float* p = prm + k * C;
          const float r2 = rbf_r2(x, y, z, p[0], p[1], p[2]);
          if (!within(r2, p[3], g.trunc)) continue;
          const float wgt = gauss_exp2(r2, pk[k]);           // :77
#pragma unroll
          for (int a = 0; a < kMaxAttr; ++a)
            if (a < g.A) attr[a] += wgt * p[4 + a];          // :78-80
          sum += wgt;                                        // :81
        }
      }
    }
  }

  // Untruncated (the reference's default): every RBF reaches every voxel, so
  // there is nothing to cull; the chunk is walked directly, four RBFs per
  // step (their loads overlap), summed in index order.
  for (int c0 = 0; n_list < 0 && !(g.trunc > 0.f) && c0 < g.N; c0 += kChunk) {
    const int nc = min(kChunk, g.N - c0);
    __syncthreads();
    for (int e = tid; e < nc * C; e += blockDim.x) prm[e] = rb[(size_t)c0 * C + e];
    for (int e = tid; e < nc; e += blockDim.x) pk[e] = gauss_k(rb[(size_t)(c0 + e) * C + 3]);
    __syncthreads();
    int k = 0;
    for (; k + 4 <= nc; k += 4) {
      float w4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float* p = prm + (k + u) * C;
        w4[u] = gauss_exp2(rbf_r2(x, y, z, p[0], p[1], p[2]), pk[k + u]);  // :77
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float* p = prm + (k + u) * C;
#pragma unroll
        for (int a = 0; a < kMaxAttr; ++a)
          if (a < g.A) attr[a] += w4[u] * p[4 + a];          // :78-80
        sum += w4[u];                                        // :81
      }
    }
    for (; k < nc; ++k) {
      const float* p = prm + k * C;
      const float wgt = gauss_exp2(rbf_r2(x, y, z, p[0], p[1], p[2]), pk[k]);
#pragma unroll
      for (int a = 0; a < kMaxAttr; ++a)
        if (a < g.A) attr[a] += wgt * p[4 + a];
      sum += wgt;
    }
  }

  for (int c0 = 0; n_list < 0 && g.trunc > 0.f && c0 < g.N; c0 += kChunk) {
    const int nc = min(kChunk, g.N - c0);
    __syncthreads();
    for (int e = tid; e < nc * C; e += blockDim.x) prm[e] = rb[(size_t)c0 * C + e];
    for (int e = tid; e < nc; e += blockDim.x) pk[e] = gauss_k(rb[(size_t)(c0 + e) * C + 3]);
    __syncthreads();
    // Cull the chunk against the brick (conservative: squared distance from
    // the centre to the box, against the truncation radius plus a margin).
    bool pass = false;
    if (tid < nc)
      pass = brick_may_hold(prm[tid * C], prm[tid * C + 1], prm[tid * C + 2], prm[tid * C + 3],
                            g.trunc, lox, hix, loy, hiy, loz, hiz);
    const unsigned long long m = __ballot(pass);
    if (lane == 0) wave_cnt[wave] = __popcll(m);
    __syncthreads();
    int base = 0;
    for (int k = 0; k < wave; ++k) base += wave_cnt[k];
    const int nkeep = wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
    if (pass) keep[base + __popcll(m & ((1ull << lane) - 1ull))] = tid;
    __syncthreads();
    for (int k = 0; k < nkeep; ++k) {
      const float* p = prm + keep[k] * C;
      const float r2 = rbf_r2(x, y, z, p[0], p[1], p[2]);
      if (!within(r2, p[3], g.trunc)) continue;
      const float wgt = gauss_exp2(r2, pk[keep[k]]);         // :77
#pragma unroll
      for (int a = 0; a < kMaxAttr; ++a)
        if (a < g.A) attr[a] += wgt * p[4 + a];              // :78-80
      sum += wgt;                                            // :81
    }
  }
  if (!ok) return;
  const size_t V = (size_t)g.D * g.H * g.W;
  const size_t vi = ((size_t)zi * g.H + yi) * g.W + xi;
  float* ob = vol + (size_t)b * (1 + g.A) * V;
  ob[vi] = sum;                                              // :83
  const float den = sum < 1.e-8f ? 1.f : sum;                // :85-87
#pragma unroll
  for (int a = 0; a < kMaxAttr; ++a)
    if (a < g.A) ob[(size_t)(1 + a) * V + vi] = attr[a] / den;  // :89-91
}

// Index range of grid samples within [c - r, c + r] along one axis (widened
// by one sample; the exact test is done per voxel).
__device__ __forceinline__ void axis_range(float c, float r, int n, float a, float b,
                                           int* lo, int* hi) {
  if (n == 1 || !(r < INFINITY)) {
    *lo = 0;
    *hi = n - 1;
    return;
  }
  const float step = (b - a) / (float)(n - 1);
  float u0 = (c - r - a) / step, u1 = (c + r - a) / step;
  if (u0 > u1) {
    const float t = u0;
    u0 = u1;
    u1 = t;
  }
  u0 = fminf(fmaxf(u0 - 1.f, -1.f), (float)n);
  u1 = fminf(fmaxf(u1 + 1.f, -1.f), (float)n);
  *lo = max((int)floorf(u0), 0);
  *hi = min((int)ceilf(u1), n - 1);
}

// Waves (RBFs) per workgroup of the truncated gradient.  Support boxes
// differ by up to ~70x in voxels, and a workgroup holds its slot until its
// slowest wave ends, so each RBF gets a workgroup of its own (0.28 ms at
// 20,000 RBFs -> 128^3, against 0.32 ms with two and 0.37 ms with four).
constexpr int kS2Waves = 1;

__global__ __launch_bounds__(kS2Waves * 64) void splat_grad_kernel(
    const float* __restrict__ rbf, const float* __restrict__ vol,
    const float* __restrict__ vgrad, float* __restrict__ rbf_grad, SplatGeom g) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * kS2Waves + wave;
  const int b = blockIdx.y;
  if (i >= g.N) return;
  const int C = 4 + g.A;
  const float* p = rbf + ((size_t)b * g.N + i) * C;
  const float xr = p[0], yr = p[1], zr = p[2], sg = p[3];
  float av[kMaxAttr];
#pragma unroll
  for (int a = 0; a < kMaxAttr; ++a) av[a] = a < g.A ? p[4 + a] : 0.f;
  const float rad = g.trunc > 0.f ? g.trunc * fabsf(sg) * 1.0001f + 1e-6f : INFINITY;
  int xa, xb, ya, yb, za, zb;
  axis_range(xr, rad, g.W, g.sx, g.ex, &xa, &xb);
  axis_range(yr, rad, g.H, g.sy, g.ey, &ya, &yb);
  axis_range(zr, rad, g.D, g.sz, g.ez, &za, &zb);
  const int nx = xb - xa + 1, ny = yb - ya + 1, nz = zb - za + 1;
  const size_t V = (size_t)g.D * g.H * g.W;
  const float* v0b = vol + (size_t)b * (1 + g.A) * V;
  const float* g0b = vgrad + (size_t)b * (1 + g.A) * V;
  float acc[4 + kMaxAttr];
#pragma unroll
  for (int c = 0; c < 4 + kMaxAttr; ++c) acc[c] = 0.f;
  const float si = 1.f / sg;                                   // :38
  const float kg = gauss_k(sg);
  // Voxel coordinates of the support box, one table per axis per wave when
  // each side has at most 64 voxels (written and read by the same wave).
  __shared__ float ctab[kS2Waves][3][64];
  const bool tabbed = nx <= 64 && ny <= 64 && nz <= 64;
  if (tabbed) {
    ctab[wave][0][lane] = splat_coord(xa + min(lane, nx - 1), g.W, g.sx, g.ex);
    ctab[wave][1][lane] = splat_coord(ya + min(lane, ny - 1), g.H, g.sy, g.ey);
    ctab[wave][2][lane] = splat_coord(za + min(lane, nz - 1), g.D, g.sz, g.ez);
  }
  if (nx > 0 && ny > 0 && nz > 0) {
    const int total = nx * ny * nz;
    // lane offsets (xo, yo, zo) of entry e = xo + nx (yo + ny zo), advanced
    // by 64 entries per step as a mixed-radix addition
    int xo = lane % nx, yo = (lane / nx) % ny, zo = lane / (nx * ny);
    const int r1 = 64 % nx, q1 = 64 / nx, r2 = q1 % ny, q2 = q1 / ny;
    for (int e = lane; e < total; e += 64) {
      const int xi = xa + xo, yi = ya + yo, zi = za + zo;
      const float x = tabbed ? ctab[wave][0][xo] : splat_coord(xi, g.W, g.sx, g.ex);
      const float y = tabbed ? ctab[wave][1][yo] : splat_coord(yi, g.H, g.sy, g.ey);
      const float z = tabbed ? ctab[wave][2][zo] : splat_coord(zi, g.D, g.sz, g.ez);
      const float r2v = rbf_r2(x, y, z, xr, yr, zr);
      if (within(r2v, sg, g.trunc)) {
        const float wgt = gauss_exp2(r2v, kg);                 // :150
        const size_t vi = ((size_t)zi * g.H + yi) * g.W + xi;
        const float dl0 = g0b[vi];                             // :151
        const float dxr = (((x - xr) * si) * si) * wgt;        // :39-41
        const float dyr = (((y - yr) * si) * si) * wgt;
        const float dzr = (((z - zr) * si) * si) * wgt;
        const float dsr = (((r2v * si) * si) * si) * wgt;      // :42-46
        if (g.A == 0) {                                        // :159-163
          acc[0] += dl0 * dxr;
          acc[1] += dl0 * dyr;
          acc[2] += dl0 * dzr;
          acc[3] += dl0 * dsr;
        } else {
          float v0 = v0b[vi];                                  // :165
          if (v0 < 1.e-8f) v0 = 1.f;                           // :167-169
          const float inv = 1.f / v0;                          // :170
          float sum = 0.f;
#pragma unroll
          for (int a = 0; a < kMaxAttr; ++a) {
            if (a < g.A) {
              const float at = v0b[(size_t)(1 + a) * V + vi];
              const float dla = g0b[(size_t)(1 + a) * V + vi];
              sum += dla * (av[a] - at);                       // :176
              acc[4 + a] += (dla * wgt) * inv;                 // :178
            }
          }
          sum = sum * inv;                                     // :180
          acc[0] += dxr * (dl0 + sum);                         // :181-184
          acc[1] += dyr * (dl0 + sum);
          acc[2] += dzr * (dl0 + sum);
          acc[3] += dsr * (dl0 + sum);
        }
      }
      xo += r1;
      const int c1 = xo >= nx ? 1 : 0;
      xo -= c1 * nx;
      yo += r2 + c1;
      const int c2 = yo >= ny ? 1 : 0;
      yo -= c2 * ny;
      zo += q2 + c2;
    }
  }
  float* out = rbf_grad + ((size_t)b * g.N + i) * C;
#pragma unroll
  for (int c = 0; c < 4 + kMaxAttr; ++c) {
    if (c < C) {
      const float s = wave_sum_f32(acc[c]);
      if (lane == 0) out[c] = s;
    }
  }
}

// S2 without truncation: every RBF sees every voxel, so a wave per RBF would
// read the whole volume once per RBF.  Here a workgroup's waves own
// kGradRpw RBFs each and share voxel tiles staged in LDS: per voxel the tile
// holds its coordinates, dL/dV0, 1/V0 (guarded as in the forward), the sum
// over attributes of dL/dVa * Va and dL/dVa, structure of arrays.  Each lane
// sums its voxels in tile order, then the wave reduces in a fixed order.
constexpr int kGradWaves = 8;        // waves per workgroup
constexpr int kGradTileFloats = 14336;

// AT: the attribute count when known at compile time (0 or 1), or -1 (any,
// up to kMaxAttr, read from g.A).  RPW: RBFs per wave.
template <int AT, int RPW>
__global__ __launch_bounds__(kGradWaves * 64) void splat_grad_all_kernel(
    const float* __restrict__ rbf, const float* __restrict__ vol,
    const float* __restrict__ vgrad, float* __restrict__ rbf_grad, SplatGeom g, int tv) {
  constexpr int AN = AT >= 0 ? AT : kMaxAttr;  // attribute slots
  extern __shared__ float tile[];  // [6 + A][tv]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int b = blockIdx.y;
  const int A = AT >= 0 ? AT : g.A;
  const int C = 4 + A;
  const int i0 = uniform_i32((blockIdx.x * kGradWaves + wave) * RPW);  // scalar registers
  // this wave's RBFs (indices past N carry sigma 1 and are not written)
  float xr[RPW], yr[RPW], zr[RPW], si2[RPW], si3[RPW], nh[RPW], av[RPW][AN > 0 ? AN : 1];
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int i = i0 + r;
    const float* p = rbf + ((size_t)b * g.N + (i < g.N ? i : 0)) * C;
    xr[r] = p[0];
    yr[r] = p[1];
    zr[r] = p[2];
    const float sg = i < g.N ? p[3] : 1.f;
    const float si = 1.f / sg;                                 // :38
    si2[r] = si * si;
    si3[r] = si2[r] * si;
    nh[r] = gauss_k(sg);
#pragma unroll
    for (int a = 0; a < AN; ++a) av[r][a] = a < A ? p[4 + a] : 0.f;
  }
  float acc[RPW][4 + AN];
#pragma unroll
  for (int r = 0; r < RPW; ++r)
#pragma unroll
    for (int c = 0; c < 4 + AN; ++c) acc[r][c] = 0.f;

  const int V = g.D * g.H * g.W;
  const float* v0b = vol + (size_t)b * (1 + A) * V;
  const float* g0b = vgrad + (size_t)b * (1 + A) * V;
  float* tx = tile;
  float* ty = tile + tv;
  float* tz = tile + 2 * tv;
  float* tg = tile + 3 * tv;
  float* ti = tile + 4 * tv;
  float* tp = tile + 5 * tv;
  float* ta = tile + 6 * tv;  // [A][tv]
  for (int v0 = 0; v0 < V; v0 += tv) {
    const int nt = min(tv, V - v0);
    __syncthreads();
    for (int e = threadIdx.x; e < nt; e += blockDim.x) {
      const int v = v0 + e;
      const int xi = v % g.W, yi = (v / g.W) % g.H, zi = v / (g.W * g.H);
      tx[e] = splat_coord(xi, g.W, g.sx, g.ex);
      ty[e] = splat_coord(yi, g.H, g.sy, g.ey);
      tz[e] = splat_coord(zi, g.D, g.sz, g.ez);
      tg[e] = g0b[v];
      if (A > 0) {
        float w0 = v0b[v];
        if (w0 < 1.e-8f) w0 = 1.f;                             // :167-169
        ti[e] = 1.f / w0;                                      // :170
        float pa = 0.f;
        for (int a = 0; a < A; ++a) {
          const float dla = g0b[(size_t)(1 + a) * V + v];
          pa += dla * v0b[(size_t)(1 + a) * V + v];
          ta[a * tv + e] = dla;
        }
        tp[e] = pa;
      }
    }
    __syncthreads();
    for (int e = lane; e < nt; e += 64) {
      const float x = tx[e], y = ty[e], z = tz[e], dl0 = tg[e];
      float inv = 0.f, pa = 0.f, dla[AN > 0 ? AN : 1];
      if (AN > 0) {
        inv = ti[e];
        pa = tp[e];
#pragma unroll
        for (int a = 0; a < AN; ++a) dla[a] = (AT >= 0 || a < A) ? ta[a * tv + e] : 0.f;
      }
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        const float dx = x - xr[r], dy = y - yr[r], dz = z - zr[r];
        float r2 = dx * dx;
        r2 = r2 + dy * dy;
        r2 = r2 + dz * dz;
        const float wgt = gauss_exp2(r2, nh[r]);               // :15-29, :150
        const float ws = si2[r] * wgt;
        const float dxr = dx * ws, dyr = dy * ws, dzr = dz * ws; // :39-41
        const float dsr = (r2 * si3[r]) * wgt;                   // :42-46
        float t = dl0;
        if (AN > 0) {
          float sa = 0.f;
#pragma unroll
          for (int a = 0; a < AN; ++a) {
            sa += dla[a] * av[r][a];
            acc[r][4 + a] += (dla[a] * wgt) * inv;             // :178
          }
          t = dl0 + (sa - pa) * inv;                           // :176-180
        }
        acc[r][0] += dxr * t;                                  // :181-184
        acc[r][1] += dyr * t;
        acc[r][2] += dzr * t;
        acc[r][3] += dsr * t;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int i = i0 + r;
    float* out = rbf_grad + ((size_t)b * g.N + (i < g.N ? i : 0)) * C;
#pragma unroll
    for (int c = 0; c < 4 + AN; ++c) {
      if (c < C) {
        const float sum = wave_sum_f32(acc[r][c]);
        if (lane == 0 && i < g.N) out[c] = sum;
      }
    }
  }
}

template <int AT, int RPW>
int launch_grad_all(const float* rbf, const float* volume, const float* volume_grad,
                    float* rbf_grad, const SplatGeom& g, hipStream_t st) {
  const int per = kGradWaves * RPW;
  const int tv = (kGradTileFloats / (6 + g.A)) & ~63;
  hipLaunchKernelGGL((splat_grad_all_kernel<AT, RPW>),
                     dim3((unsigned)((g.N + per - 1) / per), (unsigned)g.B),
                     dim3(kGradWaves * 64), (size_t)(6 + g.A) * tv * sizeof(float), st, rbf,
                     volume, volume_grad, rbf_grad, g, tv);
  return check_launch("splat_grad_all_kernel");
}

__global__ void fill_zero_i32_kernel(int* __restrict__ p, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    p[e] = 0;
}

SplatGeom splat_geom(int64_t B, int64_t N, int64_t A, int64_t D, int64_t H, int64_t W,
                     const float* vs, const float* ve, float trunc) {
  SplatGeom g;
  g.B = (int)B;
  g.N = (int)N;
  g.A = (int)A;
  g.D = (int)D;
  g.H = (int)H;
  g.W = (int)W;
  g.sx = vs[0];
  g.sy = vs[1];
  g.sz = vs[2];
  g.ex = ve[0];
  g.ey = ve[1];
  g.ez = ve[2];
  g.trunc = trunc;
  return g;
}

int check_splat(int64_t B, int64_t N, int64_t A, int64_t D, int64_t H, int64_t W,
                const float* vs, const float* ve) {
  for (int64_t d : {B, N, D, H, W})
    TOTRI_CHECK_ARG(d >= 0 && d < (1ll << 31), "dimension out of range");
  TOTRI_CHECK_ARG(A >= 0 && A <= kMaxAttr, "attributes per RBF must be in [0, %d], got %lld",
                  kMaxAttr, (long long)A);
  TOTRI_CHECK_ARG(B * (1 + A) * D * H * W < (1ll << 31) && B * N * (4 + A) < (1ll << 31),
                  "tensor has more than 2^31 elements");
  TOTRI_CHECK_ARG(vs && ve, "null volume box");
  return TOTRI_OK;
}

}  // namespace
}  // namespace totri_hip

using namespace totri_hip;

namespace totri_hip {
namespace {

struct BinLayout {
  size_t nb;     // bricks per batch entry
  size_t count;  // byte offset of the [B][nb] counts
  size_t list;   // byte offset of the [B][nb][cap] lists
  int cap;       // list slots per brick
  size_t end;
};

// Workspace of the binned splat (truncated splats only): a count per brick
// and `cap` list slots per brick, kSlotMax at most and fewer for large grids
// (at most 2^24 slots in all, at least kSlotMin per brick).  About 20-60
// RBFs reach a brick at the usual sigma / voxel ratios; a brick reached by
// more than cap scans all RBFs.
BinLayout bin_layout(int64_t B, int64_t D, int64_t H, int64_t W) {
  BinLayout l;
  const size_t bx = (size_t)((W + kBX - 1) / kBX), by = (size_t)((H + kBY - 1) / kBY),
               bz = (size_t)((D + kBZ - 1) / kBZ);
  l.nb = bx * by * bz;
  const size_t bricks = (size_t)B * l.nb;
  size_t cap = ((size_t)1 << 24) / (bricks > 0 ? bricks : 1);
  if (cap > (size_t)kSlotMax) cap = kSlotMax;
  if (cap < (size_t)kSlotMin) cap = kSlotMin;
  l.cap = (int)cap;
  l.count = 0;
  l.list = (bricks * sizeof(int) + 255) & ~(size_t)255;
  l.end = l.list + ((bricks * cap * sizeof(int) + 255) & ~(size_t)255);
  return l;
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_rbf2volume_gaussian_forward_workspace(int64_t B, int64_t N, int64_t A,
                                                              int64_t D, int64_t H, int64_t W,
                                                              float truncation) {
  (void)A;
  if (!(truncation > 0.f) || B <= 0 || N <= 0 || D <= 0 || H <= 0 || W <= 0) return 0;
  return bin_layout(B, D, H, W).end;
}

extern "C" int totri_rbf2volume_gaussian_forward(const float* rbf, int64_t B, int64_t N,
                                                 int64_t A, float* volume, int64_t D,
                                                 int64_t H, int64_t W,
                                                 const float volume_start[3],
                                                 const float volume_end[3],
                                                 float truncation, void* workspace,
                                                 size_t workspace_bytes, void* stream) {
  int rc = check_splat(B, N, A, D, H, W, volume_start, volume_end);
  if (rc) return rc;
  const size_t V = (size_t)D * H * W;
  if (B == 0 || V == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr && (N == 0 || rbf != nullptr), "null pointer");
  hipStream_t st = (hipStream_t)stream;
  const SplatGeom g = splat_geom(B, N, A, D, H, W, volume_start, volume_end, truncation);
  const int bx = (int)((W + kBX - 1) / kBX), by = (int)((H + kBY - 1) / kBY),
            bz = (int)((D + kBZ - 1) / kBZ);
  SplatBins bins;
  memset(&bins, 0, sizeof(bins));
  const size_t need = totri_rbf2volume_gaussian_forward_workspace(B, N, A, D, H, W, truncation);
  const int binned = need > 0 ? 1 : 0;
  if (binned) {
    if (workspace_bytes < need)
      return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", need,
                  workspace_bytes);
    TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
    const BinLayout l = bin_layout(B, D, H, W);
    char* w = (char*)workspace;
    bins.count = (int*)(w + l.count);
    bins.list = (int*)(w + l.list);
    bins.cap = l.cap;
    const size_t n = (size_t)B * l.nb;
    TOTRI_CHECK_ARG(n < (1u << 31), "too many bricks");
    hipLaunchKernelGGL(fill_zero_i32_kernel, dim3((unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536)),
                       dim3(256), 0, st, bins.count, n);
    rc = check_launch("fill_zero_i32_kernel");
    if (rc) return rc;
    const dim3 grid((unsigned)((N + kBinWaves - 1) / kBinWaves), (unsigned)B);
    hipLaunchKernelGGL(splat_bin_kernel, grid, dim3(kBinWaves * 64), 0, st, rbf, g, bx, by, bz,
                       bins);
    rc = check_launch("splat_bin_kernel");
    if (rc) return rc;
  }
  hipLaunchKernelGGL(splat_forward_kernel, dim3((unsigned)(bx * by * bz), (unsigned)B),
                     dim3(256), 0, st, rbf, volume, g, bx, by, bins, binned);
  return check_launch("splat_forward_kernel");
}

extern "C" int totri_rbf2volume_gaussian_grad(const float* rbf, int64_t B, int64_t N,
                                              int64_t A, const float* volume,
                                              const float* volume_grad, float* rbf_grad,
                                              int64_t D, int64_t H, int64_t W,
                                              const float volume_start[3],
                                              const float volume_end[3], float truncation,
                                              void* stream) {
  int rc = check_splat(B, N, A, D, H, W, volume_start, volume_end);
  if (rc) return rc;
  if (B == 0 || N == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(rbf && volume && volume_grad && rbf_grad, "null pointer");
  const SplatGeom g = splat_geom(B, N, A, D, H, W, volume_start, volume_end, truncation);
  if (!(truncation > 0.f)) {
    // every RBF reaches every voxel: shared voxel tiles
    hipStream_t st = (hipStream_t)stream;
    // two RBFs per wave: more per wave share the tile reads but hold more
    // registers (8 per wave: 151 VGPRs, two waves per SIMD, 61 ms at
    // 20,000 RBFs -> 128^3; 4: 38.9 ms; 2: 34.9 ms; 1: 45.6 ms)
    if (A == 0) return launch_grad_all<0, 2>(rbf, volume, volume_grad, rbf_grad, g, st);
    if (A == 1) return launch_grad_all<1, 2>(rbf, volume, volume_grad, rbf_grad, g, st);
    return launch_grad_all<-1, 2>(rbf, volume, volume_grad, rbf_grad, g, st);
  }
  hipLaunchKernelGGL(splat_grad_kernel,
                     dim3((unsigned)((N + kS2Waves - 1) / kS2Waves), (unsigned)B),
                     dim3(kS2Waves * 64), 0, (hipStream_t)stream, rbf, volume, volume_grad,
                     rbf_grad, g);
  return check_launch("splat_grad_kernel");
}
