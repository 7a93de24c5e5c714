// Double-precision forms of the confocal render (R1), its adjoint (A1, exact
// transpose and the reference's window form) and the confocal Velten
// backprojection (A2), for float64 tensors: the reference dispatches these
// kernels on float and double (AT_DISPATCH_FLOATING_TYPES,
// src/render/volume_cuda.cu:85,175, src/reconstruct/backprojection_cuda.cu:215).
//
// With scalar_t = double the reference keeps some float steps, restated here:
//   - geometry arguments are float; the bin radius t * w + off is a float
//     expression, widened (volume_cuda.cu:23,137);
//   - floorf(z_hat) and ceilf(tf) narrow their double argument to float
//     (:48,:132-134,:146);
//   - A2 converts the points to float3 and measures float distances; the bin
//     arithmetic and the sum are double (backprojection_cuda.cu:167-190).
// pow(x, 2) and pow(x, 4) are written as products.
//
// These are plain one-thread-per-output kernels in the reference's loop
// order (columns, scan points and bins in order), so each sum is the
// reference's sequence of double operations; FP64 is the accuracy form, the
// FP32 kernels are the fast path.
#include <math.h>

#include "common.h"

namespace totri_hip {
namespace {

__device__ __forceinline__ double cell_centre_d(int i, float o, float sz, int n) {
  double a = ((double)i + 0.5) * (double)sz;  // volume_cuda.cu:35
  a = a / (double)n;
  return (double)o + a;
}

struct SampleD {
  bool active;
  int zi;
  double alpha, f2;
};

// One R1 sample (volume_cuda.cu:40-58 with scalar_t = double).
__device__ __forceinline__ SampleD sample_d(double tc, double wx, double wy, double wz,
                                            double x, double y, double o, double s, int Z) {
  SampleD r;
  r.active = false;
  r.zi = 0;
  r.alpha = 0.0;
  r.f2 = 0.0;
  const double t2 = tc * tc;
  const double dx = wx - x, dy = wy - y;
  double d2 = t2 - dx * dx;
  d2 = d2 - dy * dy;
  if (!(d2 > 0.0)) return r;
  const double depth = sqrt(d2);
  const double z = wz + depth;
  double zh = (z - o) * s;
  zh = zh - 0.5;
  const int zi = (int)floorf((float)zh);  // floorf narrows to float
  if (!(zi >= 0 && zi + 1 < Z)) return r;
  r.active = true;
  r.zi = zi;
  r.alpha = zh - (double)zi;
  const double q = -depth / t2;
  const double q2 = q * q;
  r.f2 = q2 * q2;
  return r;
}

struct GeomD {
  int X, Y, Z, N, T;
  float ox, oy, oz, sx, sy, sz, w, off;
};

// R1: one thread per (scan point, bin); columns in order (volume_cuda.cu:33-63).
__global__ __launch_bounds__(256) void r1_f64_kernel(const double* __restrict__ vol,
                                                     const double* __restrict__ mp,
                                                     double* __restrict__ tr, GeomD g) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)g.N * g.T) return;
  const int i = (int)(e / g.T), ti = (int)(e - (long long)i * g.T);
  const double wx = mp[i], wy = mp[g.N + i], wz = mp[2 * g.N + i];
  const double o = (double)g.oz;
  const double s = (double)g.Z / (double)g.sz;
  const double tc = (double)bin_distance(ti, g.w, g.off);
  double value = 0.0;
  for (int xi = 0; xi < g.X; ++xi) {
    const double x = cell_centre_d(xi, g.ox, g.sx, g.X);
    for (int yi = 0; yi < g.Y; ++yi) {
      const double y = cell_centre_d(yi, g.oy, g.sy, g.Y);
      const SampleD sm = sample_d(tc, wx, wy, wz, x, y, o, s, g.Z);
      if (!sm.active) continue;
      const double* col = vol + ((size_t)xi * g.Y + yi) * g.Z;
      const double f1 = (1.0 - sm.alpha) * col[sm.zi] + sm.alpha * col[sm.zi + 1];
      value = value + f1 * sm.f2;
    }
  }
  tr[e] = value;
}

// A1 as the exact transpose of r1_f64_kernel: one thread per voxel; over scan
// points and bins in order, a sample flooring to z adds (1 - alpha) f2 G and
// one flooring to z - 1 adds alpha f2 G.  Only bins whose radius may put
// z_hat within two voxels of [z - 1, z + 1) are visited (the decision itself
// is taken exactly for each of them).
__global__ __launch_bounds__(256) void a1_f64_kernel(double* __restrict__ vol,
                                                     const double* __restrict__ mp,
                                                     const double* __restrict__ G, GeomD g) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)g.X * g.Y * g.Z) return;
  const int z = (int)(e % g.Z);
  const int col = (int)(e / g.Z);
  const int xi = col / g.Y, yi = col - (col / g.Y) * g.Y;
  const double x = cell_centre_d(xi, g.ox, g.sx, g.X);
  const double y = cell_centre_d(yi, g.oy, g.sy, g.Y);
  const double o = (double)g.oz;
  const double s = (double)g.Z / (double)g.sz;
  const bool windowed = g.w > 0.f && g.off >= 0.f && g.sz > 0.f;
  double value = 0.0;
  for (int i = 0; i < g.N; ++i) {
    const double wx = mp[i], wy = mp[g.N + i], wz = mp[2 * g.N + i];
    int t0 = 0, t1 = g.T - 1;
    if (windowed) {
      const double dx = wx - x, dy = wy - y;
      const double rho2 = dx * dx + dy * dy;
      const double dlo = fmax(((double)z - 3.0 + 0.5) / s + o - wz, 0.0);
      const double dhi = ((double)z + 3.0 + 0.5) / s + o - wz;
      if (!(dhi > 0.0)) continue;
      const double rlo = sqrt(rho2 + dlo * dlo) * (1.0 - 1e-6);
      const double rhi = sqrt(rho2 + dhi * dhi) * (1.0 + 1e-6);
      const double ta = floor((rlo - (double)g.off) / (double)g.w) - 2.0;
      const double tb = ceil((rhi - (double)g.off) / (double)g.w) + 2.0;
      t0 = ta < 0.0 ? 0 : (ta > (double)g.T ? g.T : (int)ta);
      t1 = tb > (double)(g.T - 1) ? g.T - 1 : (tb < -1.0 ? -1 : (int)tb);
    }
    const double* gi = G + (size_t)i * g.T;
    for (int ti = t0; ti <= t1; ++ti) {
      const double tc = (double)bin_distance(ti, g.w, g.off);
      const SampleD sm = sample_d(tc, wx, wy, wz, x, y, o, s, g.Z);
      if (!sm.active) continue;
      if (sm.zi == z) {
        value += (1.0 - sm.alpha) * (sm.f2 * gi[ti]);
      } else if (sm.zi + 1 == z) {
        value += sm.alpha * (sm.f2 * gi[ti]);
      }
    }
  }
  vol[e] = value;
}

// The reference's window-form A1 in double (volume_cuda.cu:95-160): one
// thread per voxel, scan points and window bins in order.
__global__ __launch_bounds__(256) void a1_window_f64_kernel(double* __restrict__ vol,
                                                            const double* __restrict__ mp,
                                                            const double* __restrict__ G,
                                                            GeomD g) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)g.X * g.Y * g.Z) return;
  const int zi = (int)(e % g.Z);
  const int col = (int)(e / g.Z);
  const int xi = col / g.Y, yi = col - (col / g.Y) * g.Y;
  const double x = cell_centre_d(xi, g.ox, g.sx, g.X);  // :107-109
  const double y = cell_centre_d(yi, g.oy, g.sy, g.Y);
  const double z = cell_centre_d(zi, g.oz, g.sz, g.Z);
  const double o = (double)g.oz;
  const double s = (double)g.Z / (double)g.sz;          // :112
  const double dz = (double)g.sz / (double)g.Z;         // :114
  double value = 0.0;
  for (int i = 0; i < g.N; ++i) {
    const double wx = mp[i], wy = mp[g.N + i], wz = mp[2 * g.N + i];
    const double zz = zi > 0 ? dz : 0.0;                // :123
    const double ax = x - wx, ay = y - wy;
    const double dps = ax * ax + ay * ay;               // :124
    const double a = (z - zz) - wz, b = z - wz, c = (z + dz) - wz;
    const double ds = sqrt(dps + a * a);                // :125-127
    const double dm = sqrt(dps + b * b);
    const double de = sqrt(dps + c * c);
    int ts = (int)ceilf((float)((ds - (double)g.off) / (double)g.w));  // :129-134
    int tm = (int)ceilf((float)((dm - (double)g.off) / (double)g.w));
    int te = (int)ceilf((float)((de - (double)g.off) / (double)g.w));
    ts = min(max(ts, 0), g.T);
    tm = min(max(tm, 0), g.T);
    te = min(max(te, 0), g.T);
    const double* gi = G + (size_t)i * g.T;
    for (int ti = ts; ti < te; ++ti) {                  // :135-156
      const double tc = (double)bin_distance(ti, g.w, g.off);
      const double t2 = tc * tc;
      const double d2 = t2 - dps;
      if (!(d2 > 0.0)) continue;
      const double depth = sqrt(d2);
      const double zc = wz + depth;
      double zh = (zc - o) * s;
      zh = zh - 0.5;
      const int zf = (int)floorf((float)zh);
      const double alpha = zh - (double)zf;
      const double df1 = ti >= tm ? 1.0 - alpha : alpha;
      const double q = -depth / t2;
      const double q2 = q * q;
      const double f2 = q2 * q2;
      value = value + (df1 * f2) * gi[ti];
    }
  }
  vol[e] = value;
}

// A2 in double (backprojection_cuda.cu:152-194): one thread per voxel of
// [B][D][H][W]; float points and distances, double bin arithmetic and sum.
__global__ __launch_bounds__(256) void velten_f64_kernel(
    const double* __restrict__ tr, const double* __restrict__ scan,
    const double* __restrict__ orig, double* __restrict__ vol, int B, int T, int S, int So,
    int D, int H, int W, float sx, float sy, float sz, float ex, float ey, float ez, float w,
    float off) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nv = (long long)D * H * W;
  if (e >= (long long)B * nv) return;
  const int b = (int)(e / nv);
  const int v = (int)(e - (long long)b * nv);
  const int x = v % W, y = (v / W) % H, z = v / (W * H);
  const float qx = inclusive_coord(x, W, sx, ex);
  const float qy = inclusive_coord(y, H, sy, ey);
  const float qz = inclusive_coord(z, D, sz, ez);
  const double* sb = scan + (size_t)b * S * 3;
  const double* ob = orig + (size_t)b * So * 3;
  const double* tb = tr + (size_t)b * T * S;
  double value = 0.0;
  for (int s = 0; s < S; ++s) {
    const float px = (float)sb[3 * s], py = (float)sb[3 * s + 1], pz = (float)sb[3 * s + 2];
    const double dsv = (double)dist3(px, py, pz, qx, qy, qz);                 // :177
    const double dso = So > 0 ? (double)dist3(px, py, pz, (float)ob[3 * s],
                                              (float)ob[3 * s + 1], (float)ob[3 * s + 2])
                              : 0.0;                                          // :178-183
    const double dist = 2.0 * dso + 2.0 * dsv;                               // :185
    const double tf = (dist - (double)off) / (double)w + 0.5;
    const int ti = (int)floorf((float)tf);                                   // :187
    if (ti >= 0 && ti < T) value += tb[(size_t)ti * S + s];                  // :188-190
  }
  vol[e] = value;                                                            // :192
}

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

int check_conf_f64(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T, const float* o,
                   const float* sz) {
  TOTRI_CHECK_ARG(X >= 0 && Y >= 0 && Z >= 0 && N >= 0 && T >= 0, "negative dimension");
  TOTRI_CHECK_ARG(X < (1ll << 31) && Y < (1ll << 31) && Z < (1ll << 31) && N < (1ll << 31) &&
                      T < (1ll << 31),
                  "dimension exceeds 32-bit indexing");
  TOTRI_CHECK_ARG(X * Y * Z < (1ll << 31) && N * T < (1ll << 31),
                  "tensor has more than 2^31 elements");
  TOTRI_CHECK_ARG(o != nullptr && sz != nullptr, "null geometry");
  return TOTRI_OK;
}

GeomD geom_d(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T, const float* o,
             const float* sz, float w, float off) {
  GeomD g;
  g.X = (int)X;
  g.Y = (int)Y;
  g.Z = (int)Z;
  g.N = (int)N;
  g.T = (int)T;
  g.ox = o[0];
  g.oy = o[1];
  g.oz = o[2];
  g.sx = sz[0];
  g.sy = sz[1];
  g.sz = sz[2];
  g.w = w;
  g.off = off;
  return g;
}

}  // namespace
}  // namespace totri_hip

using namespace totri_hip;

extern "C" int totri_volume_confocal_forward_f64(
    const double* volume, int64_t X, int64_t Y, int64_t Z, const double* measurement_points,
    int64_t N, double* transient, int64_t T, const float volume_origin[3],
    const float volume_size[3], float bin_width, float bin_offset, void* stream) {
  int rc = check_conf_f64(X, Y, Z, N, T, volume_origin, volume_size);
  if (rc) return rc;
  if (N * T == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(transient && measurement_points && (X * Y * Z == 0 || volume),
                  "null pointer");
  const GeomD g = geom_d(X, Y, Z, N, T, volume_origin, volume_size, bin_width, bin_offset);
  hipLaunchKernelGGL(r1_f64_kernel, dim3(blocks_for(N * T)), dim3(256), 0,
                     (hipStream_t)stream, volume, measurement_points, transient, g);
  return check_launch("r1_f64_kernel");
}

extern "C" int totri_volume_confocal_grad_volume_f64(
    double* volume, int64_t X, int64_t Y, int64_t Z, const double* measurement_points,
    int64_t N, const double* transient, int64_t T, const float volume_origin[3],
    const float volume_size[3], float bin_width, float bin_offset, int window_form,
    void* stream) {
  int rc = check_conf_f64(X, Y, Z, N, T, volume_origin, volume_size);
  if (rc) return rc;
  const long long nv = X * Y * Z;
  if (nv == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume && (N * T == 0 || (measurement_points && transient)), "null pointer");
  const GeomD g = geom_d(X, Y, Z, N, T, volume_origin, volume_size, bin_width, bin_offset);
  if (window_form) {
    hipLaunchKernelGGL(a1_window_f64_kernel, dim3(blocks_for(nv)), dim3(256), 0,
                       (hipStream_t)stream, volume, measurement_points, transient, g);
    return check_launch("a1_window_f64_kernel");
  }
  hipLaunchKernelGGL(a1_f64_kernel, dim3(blocks_for(nv)), dim3(256), 0, (hipStream_t)stream,
                     volume, measurement_points, transient, g);
  return check_launch("a1_f64_kernel");
}

extern "C" int totri_backprojection_confocal_velten_forward_f64(
    const double* transient, int64_t B, int64_t T, int64_t S, const double* scan_points,
    const double* scan_origin, int64_t So, double* volume, int64_t D, int64_t H, int64_t W,
    const float volume_start[3], const float volume_end[3], float bin_width, float bin_offset,
    void* stream) {
  for (int64_t d : {B, T, S, So, D, H, W})
    TOTRI_CHECK_ARG(d >= 0 && d < (1ll << 31), "dimension out of range");
  TOTRI_CHECK_ARG(B * T * S < (1ll << 31) && B * D * H * W < (1ll << 31),
                  "tensor has more than 2^31 elements");
  TOTRI_CHECK_ARG(So == 0 || So == S, "scan_origin must have 0 or S points");
  TOTRI_CHECK_ARG(volume_start && volume_end, "null volume box");
  const long long n = B * D * H * W;
  if (n == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume && (T * S == 0 || (transient && scan_points)), "null pointer");
  TOTRI_CHECK_ARG(So == 0 || scan_origin, "null scan_origin");
  hipLaunchKernelGGL(velten_f64_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream,
                     transient, scan_points, scan_origin, volume, (int)B, (int)T, (int)S,
                     (int)So, (int)D, (int)H, (int)W, volume_start[0], volume_start[1],
                     volume_start[2], volume_end[0], volume_end[1], volume_end[2], bin_width,
                     bin_offset);
  return check_launch("velten_f64_kernel");
}
