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

__device__ __forceinline__ float gaussian(float r2, float sg) {
  float e = -0.5f * r2;  // :20
  e = e / (sg * sg);     // :25-27
  return expf(e);
}

// Truncation test: keep the RBF when r2 <= (trunc * sigma)^2.
__device__ __forceinline__ bool within(float r2, float sg, float trunc) {
  if (!(trunc > 0.f)) return true;
  const float lim = trunc * sg;
  return !(r2 > lim * lim);
}

__global__ __launch_bounds__(256) void splat_forward_kernel(
    const float* __restrict__ rbf,  // [B][N][4+A]
    float* __restrict__ vol,        // [B][1+A][D][H][W]
    SplatGeom g, int bricks_x, int bricks_y) {
  __shared__ float prm[kChunk * (4 + kMaxAttr)];
  __shared__ int keep[kChunk];
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
  const int xi = bx * kBX + (tid % kBX);
  const int yi = by * kBY + ((tid / kBX) % kBY);
  const int zi = bz * kBZ + tid / (kBX * kBY);
  const bool ok = xi < g.W && yi < g.H && zi < g.D;
  const float x = splat_coord(xi, g.W, g.sx, g.ex);
  const float y = splat_coord(yi, g.H, g.sy, g.ey);
  const float z = splat_coord(zi, g.D, g.sz, g.ez);
  // Brick bounding box over the voxels it really holds.
  const int x0 = bx * kBX, x1 = min(bx * kBX + kBX, g.W) - 1;
  const int y0 = by * kBY, y1 = min(by * kBY + kBY, g.H) - 1;
  const int z0 = bz * kBZ, z1 = min(bz * kBZ + kBZ, g.D) - 1;
  const float bxa = splat_coord(x0, g.W, g.sx, g.ex), bxb = splat_coord(x1, g.W, g.sx, g.ex);
  const float bya = splat_coord(y0, g.H, g.sy, g.ey), byb = splat_coord(y1, g.H, g.sy, g.ey);
  const float bza = splat_coord(z0, g.D, g.sz, g.ez), bzb = splat_coord(z1, g.D, g.sz, g.ez);
  const float lox = fminf(bxa, bxb), hix = fmaxf(bxa, bxb);
  const float loy = fminf(bya, byb), hiy = fmaxf(bya, byb);
  const float loz = fminf(bza, bzb), hiz = fmaxf(bza, bzb);

  float sum = 0.f;
  float attr[kMaxAttr];
#pragma unroll
  for (int a = 0; a < kMaxAttr; ++a) attr[a] = 0.f;
  const float* rb = rbf + (size_t)b * g.N * C;

  for (int c0 = 0; c0 < g.N; c0 += kChunk) {
    const int nc = min(kChunk, g.N - c0);
    __syncthreads();
    for (int e = tid; e < nc * C; e += blockDim.x) prm[e] = rb[(size_t)c0 * C + e];
    __syncthreads();
    // Cull the chunk against the brick (conservative: squared distance from
    // the centre to the box, against the truncation radius plus a margin).
    bool pass = false;
    if (tid < nc) {
      if (g.trunc > 0.f) {
        const float xr = prm[tid * C], yr = prm[tid * C + 1], zr = prm[tid * C + 2];
        const float sg = prm[tid * C + 3];
        const float dx = fmaxf(fmaxf(lox - xr, xr - hix), 0.f);
        const float dy = fmaxf(fmaxf(loy - yr, yr - hiy), 0.f);
        const float dz = fmaxf(fmaxf(loz - zr, zr - hiz), 0.f);
        const float lim = g.trunc * fabsf(sg);
        pass = !((dx * dx + dy * dy + dz * dz) > lim * lim * 1.0001f + 1e-12f);
      } else {
        pass = true;
      }
    }
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
      const float wgt = gaussian(r2, p[3]);                  // :77
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

__global__ __launch_bounds__(256) void splat_grad_kernel(
    const float* __restrict__ rbf, const float* __restrict__ vol,
    const float* __restrict__ vgrad, float* __restrict__ rbf_grad, SplatGeom g) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * 4 + wave;
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
  if (nx > 0 && ny > 0 && nz > 0) {
    const int total = nx * ny * nz;
    for (int e = lane; e < total; e += 64) {
      const int xi = xa + e % nx;
      const int yi = ya + (e / nx) % ny;
      const int zi = za + e / (nx * ny);
      const float x = splat_coord(xi, g.W, g.sx, g.ex);
      const float y = splat_coord(yi, g.H, g.sy, g.ey);
      const float z = splat_coord(zi, g.D, g.sz, g.ez);
      const float r2 = rbf_r2(x, y, z, xr, yr, zr);
      if (!within(r2, sg, g.trunc)) continue;
      const float wgt = gaussian(r2, sg);                     // :150
      const size_t vi = ((size_t)zi * g.H + yi) * g.W + xi;
      const float dl0 = g0b[vi];                               // :151
      const float dxr = (((x - xr) * si) * si) * wgt;          // :39-41
      const float dyr = (((y - yr) * si) * si) * wgt;
      const float dzr = (((z - zr) * si) * si) * wgt;
      const float dsr = (((r2 * si) * si) * si) * wgt;         // :42-46
      if (g.A == 0) {                                          // :159-163
        acc[0] += dl0 * dxr;
        acc[1] += dl0 * dyr;
        acc[2] += dl0 * dzr;
        acc[3] += dl0 * dsr;
      } else {
        float v0 = v0b[vi];                                    // :165
        if (v0 < 1.e-8f) v0 = 1.f;                             // :167-169
        const float inv = 1.f / v0;                            // :170
        float sum = 0.f;
#pragma unroll
        for (int a = 0; a < kMaxAttr; ++a) {
          if (a < g.A) {
            const float at = v0b[(size_t)(1 + a) * V + vi];
            const float dla = g0b[(size_t)(1 + a) * V + vi];
            sum += dla * (av[a] - at);                         // :176
            acc[4 + a] += (dla * wgt) * inv;                   // :178
          }
        }
        sum = sum * inv;                                       // :180
        acc[0] += dxr * (dl0 + sum);                           // :181-184
        acc[1] += dyr * (dl0 + sum);
        acc[2] += dzr * (dl0 + sum);
        acc[3] += dsr * (dl0 + sum);
      }
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

extern "C" int totri_rbf2volume_gaussian_forward(const float* rbf, int64_t B, int64_t N,
                                                 int64_t A, float* volume, int64_t D,
                                                 int64_t H, int64_t W,
                                                 const float volume_start[3],
                                                 const float volume_end[3],
                                                 float truncation, void* stream) {
  int rc = check_splat(B, N, A, D, H, W, volume_start, volume_end);
  if (rc) return rc;
  const size_t V = (size_t)D * H * W;
  if (B == 0 || V == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr && (N == 0 || rbf != nullptr), "null pointer");
  const SplatGeom g = splat_geom(B, N, A, D, H, W, volume_start, volume_end, truncation);
  const int bx = (int)((W + kBX - 1) / kBX), by = (int)((H + kBY - 1) / kBY),
            bz = (int)((D + kBZ - 1) / kBZ);
  hipLaunchKernelGGL(splat_forward_kernel, dim3((unsigned)(bx * by * bz), (unsigned)B),
                     dim3(256), 0, (hipStream_t)stream, rbf, volume, g, bx, by);
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
  hipLaunchKernelGGL(splat_grad_kernel, dim3((unsigned)((N + 3) / 4), (unsigned)B), dim3(256),
                     0, (hipStream_t)stream, rbf, volume, volume_grad, rbf_grad, g);
  return check_launch("splat_grad_kernel");
}
