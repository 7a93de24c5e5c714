// Backprojections for gfx950: Velten confocal (A2), Velten exhaustive (A3)
// and Arellano ellipsoid splatting (A4).
//
// Reference: src/reconstruct/backprojection_cuda.cu:10-322,
//            src/reconstruct/backprojection.cpp:7-49.
//
// Every discrete decision (the nearest bin of a path length, the voxel an
// ellipsoid hit falls into) follows the reference's FP32 formula with IEEE
// correctly rounded square roots and divisions, no contraction, and
// norm3df(d) restated as sqrt((dx*dx + dy*dy) + dz*dz).
//
// A2/A3 (gathers): one lane per voxel, flat over [D][H][W] so that volume
// writes are coalesced.  The transient is first transposed so that each scan
// point's (and laser point's) bins are contiguous; a workgroup then stages the
// rows of a group of scan points in LDS and every lane gathers its bin from
// there.  Sums run over scan (and laser) points in the reference's order.
//
// A4 (scatter): the positive bins of the transient are first listed in
// (scan, laser, bin) order (stream compaction) and the list is split into
// chunks.  A wave owns 64 volume columns (x, y) and one chunk; the columns'
// D voxels are accumulated in LDS slots owned by their lanes (row-major
// [D][64]), so the reference's global atomics become plain read-modify-
// writes; list entries are uniform across the wave.  Chunks write partial
// volumes that are summed in chunk order.
#include <math.h>

#include <mutex>

#include "common.h"
#include "compact.h"

namespace totri_hip {
namespace {

constexpr int kStageFloats = 8192;  // LDS floats of staged transient rows (32 KiB)

int grid1d(size_t n, int block = 256) {
  size_t b = (n + block - 1) / block;
  if (b > 65535 * 4) b = 65535 * 4;
  return (int)(b < 1 ? 1 : b);
}

// out[b][c][r] = in[b][r][c] for a [B][R][C] array (tiled through LDS).
__global__ void transpose_kernel(const float* __restrict__ in, float* __restrict__ out,
                                 int R, int C) {
  __shared__ float tile[32][33];
  const int b = blockIdx.z;
  const float* src = in + (size_t)b * R * C;
  float* dst = out + (size_t)b * R * C;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + threadIdx.x;
    if (r < R && c < C) tile[k][threadIdx.x] = src[(size_t)r * C + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + threadIdx.x;
    if (r < R && c < C) dst[(size_t)c * R + r] = tile[threadIdx.x][k];
  }
}

int launch_transpose(const float* in, float* out, int B, int R, int C, hipStream_t st) {
  return transpose_batched(in, out, B, R, C, st);
}

struct BoxGeom {
  int B, D, H, W, T;
  float sx, sy, sz, ex, ey, ez;
  float w, off;
};

// Nearest bin of a path length (backprojection_cuda.cu:187,279):
// floorf((dist - offset) / width + 0.5f).
__device__ __forceinline__ int nearest_bin(float dist, float w, float off) {
  float tf = (dist - off) / w;
  tf = tf + 0.5f;
  return (int)floorf(tf);
}

// ---------------------------------------------------------------------------
// A2: vol[b][z][y][x] = sum_s tr[b][t(s)][s],
// t = nearest_bin(2 |p_s - o_s| + 2 |p_s - q|)       (backprojection_cuda.cu:152-194)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void velten_confocal_kernel(
    const float* __restrict__ trT,    // [B][S][T]
    const float* __restrict__ scan,   // [B][S][3]
    const float* __restrict__ orig,   // [B][So][3], So in {0, S}
    float* __restrict__ vol,          // [B][D][H][W]
    BoxGeom g, int S, int So, int rows_per_stage) {
  extern __shared__ float rows[];
  const int b = blockIdx.y;
  const int nvox = g.D * g.H * g.W;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = v < nvox;
  const int vv = ok ? v : 0;
  const int x = vv % g.W;
  const int y = (vv / g.W) % g.H;
  const int z = vv / (g.W * g.H);
  const float qx = inclusive_coord(x, g.W, g.sx, g.ex);
  const float qy = inclusive_coord(y, g.H, g.sy, g.ey);
  const float qz = inclusive_coord(z, g.D, g.sz, g.ez);
  const float* sb = scan + (size_t)b * S * 3;
  const float* ob = orig + (size_t)b * So * 3;
  const float* tb = trT + (size_t)b * S * g.T;
  float value = 0.f;
  // rows_per_stage == 0 (rows longer than the LDS stage): read the rows from
  // global memory; the choice is uniform over the launch.
  const bool staged = rows_per_stage > 0;
  const int step = staged ? rows_per_stage : 1;
  for (int s0 = 0; s0 < S; s0 += step) {
    const int ns = min(step, S - s0);
    const float* src = tb + (size_t)s0 * g.T;
    if (staged) {
      __syncthreads();
      for (int e = threadIdx.x; e < ns * g.T; e += blockDim.x) rows[e] = src[e];
      __syncthreads();
      src = rows;
    }
    for (int k = 0; k < ns; ++k) {
      const int s = s0 + k;
      const float px = sb[3 * s], py = sb[3 * s + 1], pz = sb[3 * s + 2];
      const float dsv = dist3(px, py, pz, qx, qy, qz);                     // :177
      const float dso = So > 0 ? dist3(px, py, pz, ob[3 * s], ob[3 * s + 1], ob[3 * s + 2])
                               : 0.f;                                        // :178-183
      const float dist = 2.f * dso + 2.f * dsv;                              // :185
      const int ti = nearest_bin(dist, g.w, g.off);                          // :187
      if (ti >= 0 && ti < g.T) value += src[k * g.T + ti];                   // :188-190
    }
  }
  if (ok) vol[(size_t)b * nvox + v] = value;                                 // :192
}

// ---------------------------------------------------------------------------
// A3: vol = sum_{s,l} tr[b][t(s,l)][s][l],
// t = nearest_bin(|p_s - o_s| + |p_s - q| + |p_l - q| + |p_l - o_l|)
// (backprojection_cuda.cu:228-287)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void velten_exhaustive_kernel(
    const float* __restrict__ trT,     // [B][S][L][T]
    const float* __restrict__ scan,    // [B][S][3]
    const float* __restrict__ laser,   // [B][L][3]
    const float* __restrict__ sorig,   // [B][So][3]
    const float* __restrict__ lorig,   // [B][Lo][3]
    float* __restrict__ vol, BoxGeom g, int S, int L, int So, int Lo,
    int rows_per_stage) {
  extern __shared__ float rows[];
  const int b = blockIdx.y;
  const int nvox = g.D * g.H * g.W;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = v < nvox;
  const int vv = ok ? v : 0;
  const int x = vv % g.W;
  const int y = (vv / g.W) % g.H;
  const int z = vv / (g.W * g.H);
  const float qx = inclusive_coord(x, g.W, g.sx, g.ex);
  const float qy = inclusive_coord(y, g.H, g.sy, g.ey);
  const float qz = inclusive_coord(z, g.D, g.sz, g.ez);
  const float* sb = scan + (size_t)b * S * 3;
  const float* lb = laser + (size_t)b * L * 3;
  const float* sob = sorig + (size_t)b * So * 3;
  const float* lob = lorig + (size_t)b * Lo * 3;
  const float* tb = trT + (size_t)b * S * L * g.T;
  const int nrows = S * L;
  float value = 0.f;
  // rows_per_stage == 0: rows from global memory (as in A2)
  const bool staged = rows_per_stage > 0;
  const int step = staged ? rows_per_stage : 1;
  for (int r0 = 0; r0 < nrows; r0 += step) {
    const int nr = min(step, nrows - r0);
    const float* src = tb + (size_t)r0 * g.T;
    if (staged) {
      __syncthreads();
      for (int e = threadIdx.x; e < nr * g.T; e += blockDim.x) rows[e] = src[e];
      __syncthreads();
      src = rows;
    }
    for (int k = 0; k < nr; ++k) {
      const int r = r0 + k;
      const int s = r / L, l = r - s * L;
      const float px = sb[3 * s], py = sb[3 * s + 1], pz = sb[3 * s + 2];
      const float dsv = dist3(px, py, pz, qx, qy, qz);                     // :255
      const float dso = So > 0 ? dist3(px, py, pz, sob[3 * s], sob[3 * s + 1], sob[3 * s + 2])
                               : 0.f;                                        // :256-261
      const float lx = lb[3 * l], ly = lb[3 * l + 1], lz = lb[3 * l + 2];
      const float dlv = dist3(lx, ly, lz, qx, qy, qz);                     // :268
      const float dlo = Lo > 0 ? dist3(lx, ly, lz, lob[3 * l], lob[3 * l + 1], lob[3 * l + 2])
                               : 0.f;                                        // :269-274
      float dist = dso + dsv;                                                // :276-277
      dist = dist + dlv;
      dist = dist + dlo;
      const int ti = nearest_bin(dist, g.w, g.off);                          // :279
      if (ti >= 0 && ti < g.T) value += src[k * g.T + ti];                   // :280-282
    }
  }
  if (ok) vol[(size_t)b * nvox + v] = value;                                 // :285
}

// ---------------------------------------------------------------------------
// A4: Arellano ellipsoid backprojection (backprojection_cuda.cu:10-115).
// ---------------------------------------------------------------------------
// Geometry of one scan/laser pair (backprojection_cuda.cu:51-74), shared by
// all of the pair's bins.
struct EllPair {
  float sd, ld, len, nx, ny, nz, cx, cy, cz;
};

__device__ __forceinline__ EllPair ell_pair(const float* sb, const float* lb,
                                            const float* sob, const float* lob, int S,
                                            int L, int So, int Lo, int s, int l) {
  EllPair e;
  const float spx = sb[s], spy = sb[S + s], spz = sb[2 * S + s];
  const float lpx = lb[l], lpy = lb[L + l], lpz = lb[2 * L + l];
  e.sd = So > 0 ? dist3(spx, spy, spz, sob[s], sob[So + s], sob[2 * So + s]) : 0.f;
  e.ld = Lo > 0 ? dist3(lpx, lpy, lpz, lob[l], lob[Lo + l], lob[2 * Lo + l]) : 0.f;
  e.cx = 0.5f * (spx + lpx);
  e.cy = 0.5f * (spy + lpy);
  e.cz = 0.5f * (spz + lpz);
  float nx = spx - lpx, ny = spy - lpy, nz = spz - lpz;                      // :62-66
  float ls = nx * nx;
  ls = ls + ny * ny;
  ls = ls + nz * nz;
  e.len = sqrt_rn(ls);                                                       // :67
  if (e.len > 1.e-6f) {                                                      // :68-70
    const float inv = 1.f / e.len;
    nx = inv * nx;
    ny = inv * ny;
    nz = inv * nz;
  }
  e.nx = nx;
  e.ny = ny;
  e.nz = nz;
  return e;
}

// One wave: columns tile*64 .. tile*64+63 of batch b, list chunk `chunk` of
// n_chunks.  dst is the chunk's partial volume [B][D][H][W] (written).
template <bool IN_LDS>
__global__ __launch_bounds__(64) void arellano_kernel(
    const float* __restrict__ trT,     // [B][S][L][T]
    const int* __restrict__ list,      // [B][S*L*T] positive bins, in order
    const int* __restrict__ totals,    // [B]
    const float* __restrict__ scan,    // [B][3][S]
    const float* __restrict__ laser,   // [B][3][L]
    const float* __restrict__ sorig,   // [B][3][So]
    const float* __restrict__ lorig,   // [B][3][Lo]
    float* __restrict__ dst,           // [n_chunks][B][D][H][W]
    BoxGeom g, int S, int L, int So, int Lo, int n_chunks) {
  extern __shared__ float colacc[];    // [D][64] when IN_LDS
  const int lane = threadIdx.x;
  const int chunk = blockIdx.y;
  const int b = blockIdx.z;
  const int ncol = g.H * g.W;
  const int v = blockIdx.x * 64 + lane;
  const bool ok = v < ncol;
  const int vv = ok ? v : 0;
  const int xi = vv % g.W, yi = vv / g.W;                                    // :27-28
  const float xf = inclusive_coord(xi, g.W, g.sx, g.ex);                     // :29-30
  const float yf = inclusive_coord(yi, g.H, g.sy, g.ey);                     // :31-32
  float* gcol = dst + ((size_t)chunk * g.B + b) * g.D * ncol + vv;           // stride ncol
  if (IN_LDS) {
    for (int z = 0; z < g.D; ++z) colacc[z * 64 + lane] = 0.f;
  } else if (ok) {
    for (int z = 0; z < g.D; ++z) gcol[(size_t)z * ncol] = 0.f;
  }
  const int n_elem = S * L * g.T;
  const int total = totals[b];
  const int per = (total + n_chunks - 1) / n_chunks;
  const int k0 = chunk * per, k1 = min(total, k0 + per);
  const int* lb_list = list + (size_t)b * n_elem;
  const float* tb = trT + (size_t)b * n_elem;
  const float* sb = scan + (size_t)b * 3 * S;
  const float* lb = laser + (size_t)b * 3 * L;
  const float* sob = sorig + (size_t)b * 3 * So;
  const float* lob = lorig + (size_t)b * 3 * Lo;
  int cur_pair = -1;
  EllPair e;
  float rsx = 0.f, rsy = 0.f, rsz = 0.f, rex = 0.f, rey = 0.f, rez = 0.f, dps = 0.f, dpe = 0.f;
  for (int k = k0; k < k1; ++k) {
    const int idx = uniform_i32(lb_list[k]);
    const int pair = idx / g.T;                                              // s * L + l
    const int ti = idx - pair * g.T;
    if (pair != cur_pair) {
      cur_pair = pair;
      const int s = pair / L, l = pair - s * L;
      e = ell_pair(sb, lb, sob, lob, S, L, So, Lo, s, l);
      rsx = xf - e.cx;                                                       // :72
      rsy = yf - e.cy;
      rsz = g.sz - e.cz;
      rex = xf - e.cx;                                                       // :74
      rey = yf - e.cy;
      rez = g.ez - e.cz;
      dps = rsx * e.nx;
      dps = dps + rsy * e.ny;
      dps = dps + rsz * e.nz;
      dpe = rex * e.nx;
      dpe = dpe + rey * e.ny;
      dpe = dpe + rez * e.nz;
    }
    const float value = tb[idx];                                             // > 0 (:77)
    float tc = (float)ti * g.w;                                              // :79-80
    tc = tc + g.off;
    tc = tc - e.sd;
    tc = tc - e.ld;
    if (!(tc > e.len)) continue;                                             // :81
    const float r0 = tc / 2.f;                                               // :83
    float r1s = r0 * r0;
    r1s = r1s - (e.len * e.len) / 4.f;                                       // :84
    const float r1 = sqrt_rn(r1s);
    const float sc = r1 / r0;                                                // :85
    const float ks = (sc - 1.f) * dps;                                       // :87
    const float stx = rsx + ks * e.nx, sty = rsy + ks * e.ny, stz = rsz + ks * e.nz;
    const float ke = (sc - 1.f) * dpe;                                       // :88
    const float etx = rex + ke * e.nx, ety = rey + ke * e.ny, etz = rez + ke * e.nz;
    float dx = etx - stx, dy = ety - sty, dz = etz - stz;                    // :89
    float rl = dx * dx;
    rl = rl + dy * dy;
    rl = rl + dz * dz;
    const float rlen = sqrt_rn(rl);                                          // :90
    const float ir = 1.f / rlen;                                             // :91
    dx = ir * dx;
    dy = ir * dy;
    dz = ir * dz;
    float v0 = dx * stx;                                                     // :93
    v0 = v0 + dy * sty;
    v0 = v0 + dz * stz;
    float ss = stx * stx;
    ss = ss + sty * sty;
    ss = ss + stz * stz;
    float v1 = v0 * v0 - (ss - r1 * r1);                                     // :94
    if (!(v1 > 0.f) || !ok) continue;                                        // :95
    v1 = sqrt_rn(v1);
    float d = -v0 - v1;
    if (d < 0.f) d = -v0 + v1;                                               // :97-100
    float bf = d / rlen;
    bf = bf * (float)g.D;
    bf = bf + 0.5f;
    const int bini = (int)bf;                                                // :101-102
    // :103-109: the interpolation weight is int - int = 0, so only bini
    // receives the value.
    if (bini >= 0 && bini < g.D) {
      if (IN_LDS)
        colacc[bini * 64 + lane] += value;
      else
        gcol[(size_t)bini * ncol] += value;
    }
  }
  if (IN_LDS && ok)
    for (int z = 0; z < g.D; ++z) gcol[(size_t)z * ncol] = colacc[z * 64 + lane];
}

__global__ void zero_kernel(float* __restrict__ out, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    out[e] = 0.f;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// out[e] = sum over parts p of part[p * n + e], in part order.
__global__ void sum_parts_kernel(const float* __restrict__ part, int n_parts, size_t n,
                                 float* __restrict__ out) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    float v = part[e];
    for (int p = 1; p < n_parts; ++p) v += part[(size_t)p * n + e];
    out[e] = v;
  }
}

BoxGeom box_geom(int64_t B, int64_t D, int64_t H, int64_t W, int64_t T,
                 const float* vs, const float* ve, float w, float off) {
  BoxGeom g;
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

int check_dims(std::initializer_list<int64_t> dims, int64_t elems_limit_a, int64_t elems_limit_b) {
  for (int64_t d : dims)
    TOTRI_CHECK_ARG(d >= 0 && d < (1ll << 31), "dimension out of range");
  TOTRI_CHECK_ARG(elems_limit_a < (1ll << 31) && elems_limit_b < (1ll << 31),
                  "tensor has more than 2^31 elements");
  return TOTRI_OK;
}

}  // namespace

int transpose_batched(const float* in, float* out, int B, int R, int C, hipStream_t st) {
  dim3 grid((C + 31) / 32, (R + 31) / 32, B);
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, st, in, out, R, C);
  return check_launch("transpose_kernel");
}

}  // namespace totri_hip

using namespace totri_hip;

extern "C" size_t totri_backprojection_confocal_velten_workspace(int64_t B, int64_t T,
                                                                 int64_t S) {
  return (size_t)B * T * S * sizeof(float);
}

namespace totri_hip {
namespace {

// A2 with its scratch (the transposed transient, B*T*S floats) supplied by
// the caller: the workspace of the entry point below, or the library's own
// array for the allocation-free variant.
int velten_confocal(const float* transient, int64_t B, int64_t T, int64_t S,
                    const float* scan_points, const float* scan_origin, int64_t So,
                    float* volume, int64_t D, int64_t H, int64_t W,
                    const float volume_start[3], const float volume_end[3], float bin_width,
                    float bin_offset, void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_dims({B, T, S, So, D, H, W}, B * T * S, B * D * H * W);
  if (rc) return rc;
  TOTRI_CHECK_ARG(So == 0 || So == S, "scan_origin must have 0 or S points");
  TOTRI_CHECK_ARG(volume_start && volume_end, "null volume box");
  hipStream_t st = (hipStream_t)stream;
  const size_t nvox = (size_t)D * H * W;
  if (B == 0 || nvox == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr, "null volume");
  const BoxGeom g = box_geom(B, D, H, W, T, volume_start, volume_end, bin_width, bin_offset);
  if (T == 0 || S == 0) {
    hipLaunchKernelGGL(velten_confocal_kernel, dim3(grid1d(nvox), (unsigned)B), dim3(256), 0,
                       st, (const float*)nullptr, scan_points, scan_origin, volume, g, 0,
                       (int)So, 1);
    return check_launch("velten_confocal_kernel");
  }
  TOTRI_CHECK_ARG(transient && scan_points, "null input");
  TOTRI_CHECK_ARG(So == 0 || scan_origin, "null scan_origin");
  const size_t need = totri_backprojection_confocal_velten_workspace(B, T, S);
  if (workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", need,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  float* trT = (float*)workspace;
  rc = launch_transpose(transient, trT, (int)B, (int)T, (int)S, st);
  if (rc) return rc;
  // rows longer than the LDS stage are read from global memory (rows = 0)
  const int rows = T <= kStageFloats ? (int)(kStageFloats / T) : 0;
  hipLaunchKernelGGL(velten_confocal_kernel, dim3(grid1d(nvox), (unsigned)B), dim3(256),
                     (size_t)rows * T * sizeof(float), st, trT, scan_points, scan_origin,
                     volume, g, (int)S, (int)So, rows);
  return check_launch("velten_confocal_kernel");
}

// Scratch of the allocation-free A2: the transposed transient [B][S][T].
// Every device that runs the library holds its own copy of this array, at
// its own address, until the process ends.  Host code never passes the
// array's name to a kernel (on the host it only stands for the symbol): it
// uses the device address that hipGetSymbolAddress returns, kept per device.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
__device__ __align__(256) float scratch[kNoAllocScratchBytes / sizeof(float)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];  // per device, resolved on its first call

// The current device's address of `scratch`, resolved once per device with
// that device current.
int scratch_address(void** out) {
  int dev = 0;
  TOTRI_CHECK_HIP(hipGetDevice(&dev));
  TOTRI_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d supported", dev,
                  kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(TOTRI_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TOTRI_OK;
}

}  // namespace
}  // namespace totri_hip

extern "C" int totri_backprojection_confocal_velten_forward(
    const float* transient, int64_t B, int64_t T, int64_t S, const float* scan_points,
    const float* scan_origin, int64_t So, float* volume, int64_t D, int64_t H, int64_t W,
    const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream) {
  return totri_hip::velten_confocal(transient, B, T, S, scan_points, scan_origin, So, volume, D,
                                    H, W, volume_start, volume_end, bin_width, bin_offset,
                                    workspace, workspace_bytes, stream);
}

extern "C" size_t totri_backprojection_confocal_velten_noalloc_max_bytes(void) {
  return totri_hip::kNoAllocScratchBytes;
}

extern "C" int totri_backprojection_confocal_velten_forward_noalloc(
    const float* transient, int64_t B, int64_t T, int64_t S, const float* scan_points,
    const float* scan_origin, int64_t So, float* volume, int64_t D, int64_t H, int64_t W,
    const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* stream) {
  using namespace totri_hip;
  TOTRI_CHECK_ARG(B >= 0 && T >= 0 && S >= 0, "negative dimension");
  const size_t need = totri_backprojection_confocal_velten_workspace(B, T, S);
  if (need > kNoAllocScratchBytes)
    return fail(TOTRI_EINVAL,
                "transient of %zu bytes (B=%lld T=%lld S=%lld) exceeds the %zu-byte scratch of "
                "the allocation-free backprojection; use "
                "totri_backprojection_confocal_velten_forward with a workspace",
                need, (long long)B, (long long)T, (long long)S, kNoAllocScratchBytes);
  void* ws = nullptr;
  int rc = scratch_address(&ws);
  if (rc) return rc;
  return velten_confocal(transient, B, T, S, scan_points, scan_origin, So, volume, D, H, W,
                         volume_start, volume_end, bin_width, bin_offset, ws,
                         kNoAllocScratchBytes, stream);
}

extern "C" size_t totri_backprojection_exhaustive_velten_workspace(int64_t B, int64_t T,
                                                                   int64_t S, int64_t L) {
  return (size_t)B * T * S * L * sizeof(float);
}

extern "C" int totri_backprojection_exhaustive_velten_forward(
    const float* transient, int64_t B, int64_t T, int64_t S, int64_t L,
    const float* scan_points, const float* laser_points, const float* scan_origin,
    int64_t So, const float* laser_origin, int64_t Lo, float* volume, int64_t D, int64_t H,
    int64_t W, const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_dims({B, T, S, L, So, Lo, D, H, W}, B * T * S * L, B * D * H * W);
  if (rc) return rc;
  TOTRI_CHECK_ARG(So == 0 || So == S, "scan_origin must have 0 or S points");
  TOTRI_CHECK_ARG(Lo == 0 || Lo == L, "laser_origin must have 0 or L points");
  TOTRI_CHECK_ARG(volume_start && volume_end, "null volume box");
  hipStream_t st = (hipStream_t)stream;
  const size_t nvox = (size_t)D * H * W;
  if (B == 0 || nvox == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr, "null volume");
  const BoxGeom g = box_geom(B, D, H, W, T, volume_start, volume_end, bin_width, bin_offset);
  if (T == 0 || S == 0 || L == 0) {
    hipLaunchKernelGGL(velten_exhaustive_kernel, dim3(grid1d(nvox), (unsigned)B), dim3(256),
                       0, st, (const float*)nullptr, scan_points, laser_points, scan_origin,
                       laser_origin, volume, g, 0, 1, (int)So, (int)Lo, 1);
    return check_launch("velten_exhaustive_kernel");
  }
  TOTRI_CHECK_ARG(transient && scan_points && laser_points, "null input");
  TOTRI_CHECK_ARG((So == 0 || scan_origin) && (Lo == 0 || laser_origin), "null origin");
  const size_t need = totri_backprojection_exhaustive_velten_workspace(B, T, S, L);
  if (workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", need,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  float* trT = (float*)workspace;
  rc = launch_transpose(transient, trT, (int)B, (int)T, (int)(S * L), st);
  if (rc) return rc;
  const int rows = T <= kStageFloats ? (int)(kStageFloats / T) : 0;
  hipLaunchKernelGGL(velten_exhaustive_kernel, dim3(grid1d(nvox), (unsigned)B), dim3(256),
                     (size_t)rows * T * sizeof(float), st, trT, scan_points, laser_points,
                     scan_origin, laser_origin, volume, g, (int)S, (int)L, (int)So, (int)Lo,
                     rows);
  return check_launch("velten_exhaustive_kernel");
}

namespace totri_hip {
namespace {

// Arellano plan: chunks of the positive-bin list per column tile, enough
// waves to fill the device, partial volumes capped at 1 GiB.  Workspace:
// transposed transient | compaction | partial volumes (when chunks > 1).
struct ArellanoPlan {
  int chunks;
  size_t tr_bytes;
  CompactLayout cl;
  size_t part_off, ws;
};

ArellanoPlan plan_arellano(int64_t B, int64_t T, int64_t S, int64_t L, int64_t D, int64_t H,
                           int64_t W) {
  ArellanoPlan p;
  const int64_t n = T * S * L;
  const int64_t V = D * H * W;
  const int64_t tiles = ((H * W + 63) / 64) * (B > 0 ? B : 1);
  const int64_t target = 8 * (int64_t)device_cu_count();
  int64_t C = tiles > 0 ? (target + tiles - 1) / tiles : 1;
  if (C > 64) C = 64;
  const int64_t cap = (int64_t)1 << 28;  // floats of partial volumes
  if (B * V > 0 && C * B * V > cap) C = cap / (B * V);
  if (C < 1) C = 1;
  p.chunks = (int)C;
  p.tr_bytes = align256((size_t)B * n * sizeof(float));
  p.cl = compact_layout(B, n, p.tr_bytes);
  p.part_off = p.cl.end;
  p.ws = p.part_off + (C > 1 ? align256((size_t)C * B * V * sizeof(float)) : 0);
  return p;
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_backprojection_exhaustive_arellano_workspace(
    int64_t B, int64_t T, int64_t S, int64_t L, int64_t D, int64_t H, int64_t W) {
  if (B <= 0 || T <= 0 || S <= 0 || L <= 0 || D <= 0 || H <= 0 || W <= 0) return 0;
  return plan_arellano(B, T, S, L, D, H, W).ws;
}

extern "C" int totri_backprojection_exhaustive_arellano_forward(
    const float* transient, int64_t B, int64_t T, int64_t S, int64_t L,
    const float* scan_points, const float* laser_points, const float* scan_origin,
    int64_t So, const float* laser_origin, int64_t Lo, float* volume, int64_t D, int64_t H,
    int64_t W, const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_dims({B, T, S, L, So, Lo, D, H, W}, B * T * S * L, B * D * H * W);
  if (rc) return rc;
  TOTRI_CHECK_ARG(So == 0 || So == S, "scan_origin must have 0 or S points");
  TOTRI_CHECK_ARG(Lo == 0 || Lo == L, "laser_origin must have 0 or L points");
  TOTRI_CHECK_ARG(volume_start && volume_end, "null volume box");
  hipStream_t st = (hipStream_t)stream;
  const size_t ncol = (size_t)H * W;
  const size_t nv = (size_t)B * D * ncol;
  if (nv == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr, "null volume");
  if (T == 0 || S == 0 || L == 0) {
    // nothing to backproject: the written volume is zero
    hipLaunchKernelGGL(zero_kernel, dim3(grid1d(nv)), dim3(256), 0, st, volume, nv);
    return check_launch("zero_kernel");
  }
  TOTRI_CHECK_ARG(transient && scan_points && laser_points, "null input");
  TOTRI_CHECK_ARG((So == 0 || scan_origin) && (Lo == 0 || laser_origin), "null origin");
  const ArellanoPlan p = plan_arellano(B, T, S, L, D, H, W);
  if (workspace_bytes < p.ws)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", p.ws,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr, "null workspace");
  const BoxGeom g = box_geom(B, D, H, W, T, volume_start, volume_end, bin_width, bin_offset);
  float* trT = (float*)workspace;
  rc = launch_transpose(transient, trT, (int)B, (int)T, (int)(S * L), st);
  if (rc) return rc;
  rc = launch_compact(trT, (int)B, (int)(T * S * L), p.cl, workspace, Positive(), st);
  if (rc) return rc;
  const int* totals = (const int*)((char*)workspace + p.cl.totals_off);
  const int* list = (const int*)((char*)workspace + p.cl.list_off);
  float* dst = p.chunks > 1 ? (float*)((char*)workspace + p.part_off) : volume;
  const dim3 grid((unsigned)((ncol + 63) / 64), (unsigned)p.chunks, (unsigned)B);
  if (D <= 256) {
    hipLaunchKernelGGL(arellano_kernel<true>, grid, dim3(64), (size_t)D * 64 * sizeof(float),
                       st, trT, list, totals, scan_points, laser_points, scan_origin,
                       laser_origin, dst, g, (int)S, (int)L, (int)So, (int)Lo, p.chunks);
  } else {
    hipLaunchKernelGGL(arellano_kernel<false>, grid, dim3(64), 0, st, trT, list, totals,
                       scan_points, laser_points, scan_origin, laser_origin, dst, g, (int)S,
                       (int)L, (int)So, (int)Lo, p.chunks);
  }
  rc = check_launch("arellano_kernel");
  if (rc || p.chunks == 1) return rc;
  hipLaunchKernelGGL(sum_parts_kernel, dim3(grid1d(nv)), dim3(256), 0, st, dst, p.chunks, nv,
                     volume);
  return check_launch("sum_parts_kernel");
}
