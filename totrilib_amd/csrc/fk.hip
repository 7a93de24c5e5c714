// f-k migration (Lindell et al. 2019) around a library FFT, gfx950.
//
// Reference: totri/reconstruct/fk_migration.py:5-120 (pure PyTorch, no CUDA
// source).  For a confocal transient t[D][H][W]:
//   P[d][h][w] = sqrt(t[d][h][w]) * lin(d; 0, 1, D)  for d < D, 0 for d >= D   (:108-112)
//   F = fftn(P)                                   (:114; the host layer calls hipFFT)
//   G(k, h, w) = wgt * trilinear(fftshift(F), (x, y, zz))                       (:41-49,66-85)
//       x = lin(w; -1, 1, W), y = lin(h; -1, 1, H), z = lin(k; -1, 1, 2D)
//       zz = sqrt(x^2 A + y^2 B + z^2), A = (W depth / (2 D width))^2, B likewise
//       wgt = (z > 0) |z| / max(zz, 1e-6)
//       trilinear = torch grid_sample, align_corners=False, zero padding
//   V = ifftn(ifftshift(G))                       (:118; hipFFT)
//   out[d][h][w] = |V[d][h][w]|, d < D           (:119)
//
// The reference builds the sample grid and the weights as tensors and runs
// grid_sample twice (real, imaginary) before a weighting pass and both
// shifts as separate copies.  Here one kernel computes the coordinates per
// output element, gathers the eight complex corners once and writes the
// weighted result; both shifts are folded into its indexing (the fftshift
// into the reads, the ifftshift into the writes), which is exact for odd
// sizes too.  Half of G is zero (z <= 0): in unshifted order that half is
// depth rows D..2D-1, exactly where P holds its zero padding, so the host
// layer writes G over P and the kernel skips the zero half.
#include <math.h>

#include "common.h"

namespace totri_hip {
namespace {

constexpr int kFkBlock = 256;

// torch.linspace(a, b, n)[i] as computed on the device for float32: one
// step (b - a) / (n - 1), then a + step * i for the first half and
// b - step * (n - 1 - i) for the second half; n == 1 gives a.
__device__ __forceinline__ float fk_lin(int i, int n, float a, float b) {
  if (n == 1) return a;
  const float step = (b - a) / (float)(n - 1);
  if (i < n / 2) return a + step * (float)i;
  return b - step * (float)(n - 1 - i);
}

// Position in the unshifted array of index i of the fftshift-ed one:
// fftshift(x)[i] = x[(i - n/2) mod n], and ifftshift(y)[j] = y[(j + n/2) mod n],
// so the element a kernel computes at shifted index i is stored at the same
// (i - n/2) mod n.
__device__ __forceinline__ int fk_unshift(int i, int n) {
  const int j = i - n / 2;
  return j < 0 ? j + n : j;
}

// grid_sample's coordinate unnormalisation for align_corners=False:
// ((c + 1) * size - 1) / 2, rounded after every operation.
__device__ __forceinline__ float fk_unnormalize(float c, int size) {
  float a = c + 1.f;
  a = a * (float)size;
  a = a - 1.f;
  return a / 2.f;
}

struct FkStoltGeom {
  int C, D2, H, W;
  float ax, ay;  // (W depth / (D2 width))^2, (H depth / (D2 height))^2
  int fold;      // 1: read raw fftn order, write ifftn order
  int k0;        // first depth index written (D2 / 2 when the zero half is skipped)
};

__global__ __launch_bounds__(kFkBlock) void fk_prepare_kernel(
    const float* __restrict__ t,  // [D][HW]
    float2* __restrict__ P,       // [2D][HW]
    int D, int HW, int blocks_per_row) {
  const int d = blockIdx.x / blocks_per_row;
  const int e = (blockIdx.x - d * blocks_per_row) * kFkBlock + threadIdx.x;
  if (e >= HW) return;
  const size_t i = (size_t)d * HW + e;
  float v = 0.f;
  if (d < D) v = sqrtf(t[i]) * fk_lin(d, D, 0.f, 1.f);
  P[i] = make_float2(v, 0.f);
}

__global__ __launch_bounds__(kFkBlock) void fk_stolt_kernel(
    const float2* __restrict__ F,  // [C][D2][H][W]
    float2* __restrict__ G,        // [C][D2][H][W]
    FkStoltGeom g, int n_out) {
  const int idx = blockIdx.x * kFkBlock + threadIdx.x;
  if (idx >= n_out) return;
  const int c = blockIdx.y;
  const int w = idx % g.W;
  const int r = idx / g.W;
  const int h = r % g.H;
  const int k = g.k0 + r / g.H;

  const float x = fk_lin(w, g.W, -1.f, 1.f);
  const float y = fk_lin(h, g.H, -1.f, 1.f);
  const float z = fk_lin(k, g.D2, -1.f, 1.f);
  const size_t plane = (size_t)g.H * g.W;
  const float2* Fc = F + (size_t)c * g.D2 * plane;
  float re = 0.f, im = 0.f;
  if (z > 0.f) {
    float zz = (x * x) * g.ax;
    zz = zz + (y * y) * g.ay;
    zz = zz + z * z;
    zz = sqrtf(zz);
    const float wgt = fabsf(z) / fmaxf(zz, 1e-6f);
    const float ix = fk_unnormalize(x, g.W);
    const float iy = fk_unnormalize(y, g.H);
    const float iz = fk_unnormalize(zz, g.D2);
    // zz >= 0 keeps iz >= (D2 - 1) / 2; past D2 every corner is padding.
    if (iz < (float)g.D2) {
      const float fx = floorf(ix), fy = floorf(iy), fz = floorf(iz);
      const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
      const float wx[2] = {(fx + 1.f) - ix, ix - fx};
      const float wy[2] = {(fy + 1.f) - iy, iy - fy};
      const float wz[2] = {(fz + 1.f) - iz, iz - fz};
#pragma unroll
      for (int dz = 0; dz < 2; ++dz) {
        const int zi = z0 + dz;
        if (zi < 0 || zi >= g.D2) continue;
        const size_t zoff = (size_t)(g.fold ? fk_unshift(zi, g.D2) : zi) * plane;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
          const int yi = y0 + dy;
          if (yi < 0 || yi >= g.H) continue;
          const size_t yoff = zoff + (size_t)(g.fold ? fk_unshift(yi, g.H) : yi) * g.W;
#pragma unroll
          for (int dx = 0; dx < 2; ++dx) {
            const int xi = x0 + dx;
            if (xi < 0 || xi >= g.W) continue;
            const float2 v = Fc[yoff + (g.fold ? fk_unshift(xi, g.W) : xi)];
            const float cw = (wx[dx] * wy[dy]) * wz[dz];
            re = re + v.x * cw;
            im = im + v.y * cw;
          }
        }
      }
    }
    re = re * wgt;
    im = im * wgt;
  }
  const int ko = g.fold ? fk_unshift(k, g.D2) : k;
  const int ho = g.fold ? fk_unshift(h, g.H) : h;
  const int wo = g.fold ? fk_unshift(w, g.W) : w;
  G[(size_t)c * g.D2 * plane + (size_t)ko * plane + (size_t)ho * g.W + wo] =
      make_float2(re, im);
}

__global__ __launch_bounds__(kFkBlock) void fk_magnitude_kernel(
    const float2* __restrict__ V,  // [C][2D][HW]
    float* __restrict__ out,       // [C][D][HW]
    int D, int HW, int blocks_per_row) {
  const int d = blockIdx.x / blocks_per_row;
  const int e = (blockIdx.x - d * blocks_per_row) * kFkBlock + threadIdx.x;
  if (e >= HW) return;
  const size_t c = blockIdx.y;
  const float2 v = V[(c * 2 * D + d) * HW + e];
  out[(c * D + d) * HW + e] = hypotf(v.x, v.y);
}

int check_fk_dims(int64_t C, int64_t D2, int64_t H, int64_t W) {
  TOTRI_CHECK_ARG(C >= 0 && D2 >= 0 && H >= 0 && W >= 0, "negative size");
  TOTRI_CHECK_ARG(C <= 65535, "at most 65535 channels (got %lld)", (long long)C);
  TOTRI_CHECK_ARG(D2 * H * W < ((int64_t)1 << 31) - kFkBlock,
                  "one channel must hold fewer than 2^31 elements");
  return TOTRI_OK;
}

}  // namespace
}  // namespace totri_hip

using namespace totri_hip;

extern "C" int totri_fk_prepare(const float* transient, int64_t D, int64_t H, int64_t W,
                                void* padded, void* stream) {
  int rc = check_fk_dims(1, 2 * D, H, W);
  if (rc) return rc;
  const int64_t HW = H * W;
  if (D == 0 || HW == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(transient != nullptr && padded != nullptr, "null pointer");
  const int64_t bpr = (HW + kFkBlock - 1) / kFkBlock;
  TOTRI_CHECK_ARG(2 * D * bpr < ((int64_t)1 << 31), "transient too large");
  hipLaunchKernelGGL(fk_prepare_kernel, dim3((unsigned)(2 * D * bpr)), dim3(kFkBlock), 0,
                     (hipStream_t)stream, transient, (float2*)padded, (int)D, (int)HW,
                     (int)bpr);
  return check_launch("fk_prepare_kernel");
}

extern "C" int totri_fk_stolt(const void* spectrum, int64_t C, int64_t D2, int64_t H,
                              int64_t W, double depth, double height, double width,
                              int layout, int out_half_zeroed, void* out, void* stream) {
  int rc = check_fk_dims(C, D2, H, W);
  if (rc) return rc;
  TOTRI_CHECK_ARG(D2 % 2 == 0, "the depth size must be even (got %lld)", (long long)D2);
  TOTRI_CHECK_ARG(layout == TOTRI_FK_SHIFTED || layout == TOTRI_FK_UNSHIFTED,
                  "unknown layout %d", layout);
  TOTRI_CHECK_ARG(isfinite(depth) && isfinite(height) && isfinite(width) && height > 0 &&
                      width > 0,
                  "depth must be finite and height, width finite and positive");
  if (C == 0 || D2 == 0 || H == 0 || W == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(spectrum != nullptr && out != nullptr, "null pointer");
  TOTRI_CHECK_ARG(spectrum != out, "spectrum and out must not overlap");
  // fk_migration.py:44-47: the factors are formed in double and squared there;
  // the products with the float32 coordinates round them to float.
  const double D = (double)(D2 / 2);
  const double a = ((double)W * depth) / (2.0 * D * width);
  const double b = ((double)H * depth) / (2.0 * D * height);
  FkStoltGeom g;
  g.C = (int)C;
  g.D2 = (int)D2;
  g.H = (int)H;
  g.W = (int)W;
  g.ax = (float)(a * a);
  g.ay = (float)(b * b);
  g.fold = layout == TOTRI_FK_UNSHIFTED;
  g.k0 = out_half_zeroed ? (int)(D2 / 2) : 0;
  const int64_t n_out = (D2 - g.k0) * H * W;
  hipLaunchKernelGGL(fk_stolt_kernel,
                     dim3((unsigned)((n_out + kFkBlock - 1) / kFkBlock), (unsigned)C),
                     dim3(kFkBlock), 0, (hipStream_t)stream, (const float2*)spectrum,
                     (float2*)out, g, (int)n_out);
  return check_launch("fk_stolt_kernel");
}

extern "C" int totri_fk_magnitude(const void* volume, int64_t C, int64_t D, int64_t H,
                                  int64_t W, float* out, void* stream) {
  int rc = check_fk_dims(C, 2 * D, H, W);
  if (rc) return rc;
  const int64_t HW = H * W;
  if (C == 0 || D == 0 || HW == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr && out != nullptr, "null pointer");
  const int64_t bpr = (HW + kFkBlock - 1) / kFkBlock;
  TOTRI_CHECK_ARG(D * bpr < ((int64_t)1 << 31), "volume too large");
  hipLaunchKernelGGL(fk_magnitude_kernel, dim3((unsigned)(D * bpr), (unsigned)C),
                     dim3(kFkBlock), 0, (hipStream_t)stream, (const float2*)volume, out,
                     (int)D, (int)HW, (int)bpr);
  return check_launch("fk_magnitude_kernel");
}
