// Compatibility forms of the confocal backprojection for gfx950.
//
// a1_window_kernel -- the reference's own window form of the albedo gradient
// (src/render/volume_cuda.cu:95-160), for callers of apply_transposed that
// need its exact result rather than the exact transpose of the render.  The
// two differ on the top voxel slice z = Z-1, which the window form fills
// with samples (z_hat in [Z-1, Z)) that the render never reads, and by
// rounding elsewhere: the window form decides a voxel's bins by ceilings of
// path-length quotients and takes depth^2 = tc^2 - (dx^2 + dy^2).
//
// Mapping: one lane per voxel, the 64 lanes of a wave along z of one column,
// so the wall distance in the plane is uniform across the wave; scan points
// are visited in order and each lane walks its own window of bins, in the
// reference's order, so the FP32 sum is the reference's, operation for
// operation (IEEE square roots and quotients, no contraction).
#include <math.h>

#include "common.h"

namespace totri_hip {
namespace {

constexpr int kWindowWaves = 4;  // independent columns per workgroup

template <typename GT>
__global__ __launch_bounds__(kWindowWaves * 64) void a1_window_kernel(
    float* __restrict__ vol, const float* __restrict__ mp, const GT* __restrict__ G,
    int X, int Y, int Z, int N, int T, float ox, float oy, float oz, float szx, float szy,
    float szz, float w, float off, int zblocks) {
  const int lane = threadIdx.x & 63;
  const int item = blockIdx.x * kWindowWaves + (threadIdx.x >> 6);
  const int ncol = X * Y;
  if (item >= ncol * zblocks) return;  // whole wave; no barrier follows
  const int col = item / zblocks;
  const int zi = (item - col * zblocks) * 64 + lane;
  const int xi = col / Y, yi = col - (col / Y) * Y;
  const bool valid = zi < Z;
  // volume_cuda.cu:107-114
  const float x = cell_centre(xi, ox, szx, X);
  const float y = cell_centre(yi, oy, szy, Y);
  const float z = cell_centre(valid ? zi : 0, oz, szz, Z);
  const float s = (float)Z / szz;
  const float dz = szz / (float)Z;
  const float zz = zi > 0 ? dz : 0.f;  // :123
  float value = 0.f;
  for (int i = 0; i < N; ++i) {
    const float wx = mp[i], wy = mp[N + i], wz = mp[2 * N + i];
    // :124-127 distances to the previous, this and the next voxel centre
    float dx = x - wx, dy = y - wy;
    float dps = dx * dx;
    dps = dps + dy * dy;
    const float a = (z - zz) - wz, b = z - wz, c = (z + dz) - wz;
    const float ds = sqrt_rn(dps + a * a);
    const float dm = sqrt_rn(dps + b * b);
    const float de = sqrt_rn(dps + c * c);
    // :129-134 bin window [ts, te) with the lower part below tm
    int ts = (int)ceilf((ds - off) / w);
    int tm = (int)ceilf((dm - off) / w);
    int te = (int)ceilf((de - off) / w);
    ts = min(max(ts, 0), T);
    tm = min(max(tm, 0), T);
    te = min(max(te, 0), T);
    if (!valid) te = ts;
    const GT* g = G + (size_t)i * T;
    for (int ti = ts; ti < te; ++ti) {  // :135-156
      const float tc = bin_distance(ti, w, off);
      const float t2 = tc * tc;
      const float d2 = t2 - dps;
      if (!(d2 > 0.f)) continue;
      const float depth = sqrt_rn(d2);
      const float zc = wz + depth;
      float zh = (zc - oz) * s;
      zh = zh - 0.5f;
      const float zf = floorf(zh);
      const float alpha = zh - zf;
      const float df1 = ti >= tm ? 1.f - alpha : alpha;
      const float q = -depth / t2;
      const float q2 = q * q;
      const float f2 = q2 * q2;
      value = value + (df1 * f2) * load_elem(g + ti);
    }
  }
  if (valid) vol[((size_t)xi * Y + yi) * Z + zi] = value;  // written, not accumulated
}

template <typename GT>
int launch_window(float* volume, int X, int Y, int Z, const float* mp, int N, const GT* G,
                  int T, const float* o, const float* sz, float w, float off,
                  hipStream_t st) {
  const int zblocks = (Z + 63) / 64;
  const size_t items = (size_t)X * Y * zblocks;
  const size_t blocks = (items + kWindowWaves - 1) / kWindowWaves;
  TOTRI_CHECK_ARG(blocks <= 0x7fffffff, "volume too large for one launch");
  hipLaunchKernelGGL(a1_window_kernel<GT>, dim3((unsigned)blocks), dim3(kWindowWaves * 64), 0,
                     st, volume, mp, G, X, Y, Z, N, T, o[0], o[1], o[2], sz[0], sz[1], sz[2], w,
                     off, zblocks);
  return check_launch("a1_window_kernel");
}

}  // namespace
}  // namespace totri_hip

extern "C" int totri_volume_confocal_grad_volume_window(
    float* volume, int64_t X, int64_t Y, int64_t Z, const float* measurement_points,
    int64_t N, const void* transient, int64_t T, int transient_dtype,
    const float volume_origin[3], const float volume_size[3], float bin_width,
    float bin_offset, void* stream) {
  using namespace totri_hip;
  TOTRI_CHECK_ARG(X >= 0 && Y >= 0 && Z >= 0 && N >= 0 && T >= 0, "negative dimension");
  TOTRI_CHECK_ARG(X < (1ll << 31) && Y < (1ll << 31) && Z < (1ll << 31) && N < (1ll << 31) &&
                      T < (1ll << 31),
                  "dimension exceeds 32-bit indexing");
  TOTRI_CHECK_ARG(X * Y < (1ll << 31) && X * Y * Z < (1ll << 31),
                  "volume has more than 2^31 elements");
  TOTRI_CHECK_ARG(N * T < (1ll << 31), "transient has more than 2^31 elements");
  TOTRI_CHECK_ARG(volume_origin != nullptr && volume_size != nullptr, "null geometry");
  TOTRI_CHECK_ARG(transient_dtype == TOTRI_F32 || transient_dtype == TOTRI_BF16 ||
                      transient_dtype == TOTRI_F16,
                  "unsupported transient dtype %d", transient_dtype);
  if (X * Y * Z == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr, "null volume");
  TOTRI_CHECK_ARG(N == 0 || T == 0 || (transient != nullptr && measurement_points != nullptr),
                  "null transient or measurement points");
  hipStream_t st = (hipStream_t)stream;
  const float* mp = measurement_points;
  const int n = (int)N, t = (int)T;
  switch (transient_dtype) {
    case TOTRI_F32:
      return launch_window(volume, (int)X, (int)Y, (int)Z, mp, n, (const float*)transient, t,
                           volume_origin, volume_size, bin_width, bin_offset, st);
    case TOTRI_BF16:
      return launch_window(volume, (int)X, (int)Y, (int)Z, mp, n,
                           (const __hip_bfloat16*)transient, t, volume_origin, volume_size,
                           bin_width, bin_offset, st);
    default:
      return launch_window(volume, (int)X, (int)Y, (int)Z, mp, n, (const __half*)transient, t,
                           volume_origin, volume_size, bin_width, bin_offset, st);
  }
}
