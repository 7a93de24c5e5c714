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

// Blocks of a grid-stride loop over n elements (capped).
int grid1d(size_t n, int block = 256) {
  size_t b = (n + block - 1) / block;
  if (b > 65535 * 4) b = 65535 * 4;
  return (int)(b < 1 ? 1 : b);
}

// Blocks of a one-lane-per-element launch over n < 2^31 elements (uncapped).
unsigned blocks_for(size_t n, int block = 256) {
  const size_t b = (n + block - 1) / block;
  return (unsigned)(b < 1 ? 1 : b);
}

// out[b][c][ooff + r] = in[b][r][c] for a [B][R][C] array, out rows of
// stride ldo (tiled through LDS).
__global__ void transpose_kernel(const float* __restrict__ in, float* __restrict__ out,
                                 int R, int C, int ldo, int ooff) {
  __shared__ float tile[32][33];
  const int b = blockIdx.z;
  const float* src = in + (size_t)b * R * C;
  float* dst = out + (size_t)b * C * ldo + ooff;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + threadIdx.x;
    if (r < R && c < C) tile[k][threadIdx.x] = src[(size_t)r * C + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + threadIdx.x;
    if (r < R && c < C) dst[(size_t)c * ldo + r] = tile[threadIdx.x][k];
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
//
// For one scan point, t is a non-decreasing function of the squared distance
// e = (dx*dx + dy*dy) + dz*dz: every step of the formula (correctly rounded
// square root, doubling, the sum with 2 |p_s - o_s|, the subtraction of the
// offset, the division by a positive width, + 0.5, floor, saturating
// conversion) is monotone.  So t(e) >= k exactly when e >= E_k, with
// E_k = min{e : t(e) >= k}.  velten_bins_kernel finds E_0 .. E_T per scan
// point by bisection over the float bit patterns, evaluating the exact
// formula.  The backprojection then takes a cheap estimate a of t (hardware
// square root, one FMA with precomputed constants), within one bin of t, and
// corrects it with two table entries:
//   t = a - [e < E_a] + [e >= E_(a+1)].
// That is the exact decision, with no correctly rounded root or division
// per (voxel, scan point) pair.
// ---------------------------------------------------------------------------

// The exact bin of squared distance e for a scan point whose origin term is
// dso2 = 2 |p_s - o_s| (0 without an origin): backprojection_cuda.cu:177-187.
__device__ __forceinline__ int velten_bin(float e, float dso2, float w, float off) {
  const float dsv = sqrt_rn(e);
  const float dist = dso2 + 2.f * dsv;
  return nearest_bin(dist, w, off);
}

// Per-scan-point constants of the estimate: the point and
// c = ((dso2 - off) * (1/w)) + 0.5, so that a = floor(fma(sqrt(e), 2/w, c)).
// c is NaN where the estimate's error bound does not hold (below); those
// scan points take the exact formula per pair.
struct VeltenPoint {
  float px, py, pz, c;
};

// One row of floats per scan point, all parts 16-byte aligned:
//   [0, 4)             the point's constants (VeltenPoint)
//   [4, 4 + nv)        its transient bins, vals[1 + t] = bin t, zeros at 0
//                      and T + 1 (nv = T + 2 rounded up to a multiple of 4)
//   [4 + nv, 4 + nv + nt)  its thresholds { -inf, E_0, ..., E_T, NaN }
//                      (nt = T + 3 rounded up to a multiple of 4)
// so that a group of consecutive scan points is one contiguous block.
struct VeltenRow {
  int nv, nt, len;
};

__host__ __device__ __forceinline__ VeltenRow velten_row(int T) {
  VeltenRow r;
  r.nv = (T + 5) & ~3;
  r.nt = (T + 6) & ~3;
  r.len = 4 + r.nv + r.nt;
  return r;
}

// Estimate error: the exact chain rounds five times (root, sum, difference,
// quotient, + 0.5) and the estimate about eight times (hardware root within
// 1 ulp, 1/w, 2/w, the constant's three roundings, the FMA), each a relative
// error of at most 2^-24 (2^-23 for the root) of a term no larger than M bins,
// M = (2 d_max + |dso2| + |off|) / w + 1 with d_max the largest distance from
// the scan point to the box's corners.  The two differ by under 16 * 2^-24 * M
// bins; M < 2^16 keeps that below 1/16 of a bin, so the estimate is within
// one bin of the exact value.
constexpr float kVeltenMaxBins = 65536.f;

// Thread (b, s, j), j in [0, T + 3): threshold j of scan point s, that is
// -inf, E_0, ..., E_T, NaN, where E_k = +inf when no finite e reaches bin k.
// For an estimate a in [-1, T] the rows' threshold part gives E_a at index
// a + 1 and E_(a+1) at a + 2 (no e is below -inf, and no e compares >= NaN).
// Thread j = 0 also writes the point's constants and the zero pads of its
// bins.
__global__ __launch_bounds__(256) void velten_bins_kernel(
    const float* __restrict__ scan, const float* __restrict__ orig, BoxGeom g, int S, int So,
    float rw, float* __restrict__ rows) {
  const int T = g.T;
  const VeltenRow vr = velten_row(T);
  const int n = T + 3;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)g.B * S * n) return;
  const int j = (int)(id % n);
  const size_t bs = id / n;  // b * S + s
  const int b = (int)(bs / S), s = (int)(bs - (size_t)b * S);
  float* row = rows + bs * vr.len;
  const float* p = scan + bs * 3;
  const float px = p[0], py = p[1], pz = p[2];
  float dso = 0.f;
  if (So > 0) {
    const float* o = orig + ((size_t)b * So + s) * 3;
    dso = dist3(px, py, pz, o[0], o[1], o[2]);
  }
  const float dso2 = 2.f * dso;
  float e;
  if (j == 0) {
    e = -INFINITY;
  } else if (j == n - 1) {
    e = __builtin_nanf("");
  } else {
    // E_k: the smallest e >= 0 (as a bit pattern; monotone for non-negative
    // floats) with velten_bin(e) >= k.  A guess from the bin's edge in
    // double, ((k - 1/2) w + off - dso2)^2 / 4, is a few ulps from it: walk
    // from there while the neighbour agrees, in at most 64 steps, else bisect
    // over all patterns.  +inf (0x7f800000) always qualifies.
    const int k = j - 1;
    const double dk = ((double)k - 0.5) * (double)g.w + (double)g.off - (double)dso2;
    const double gd = dk > 0.0 ? 0.25 * dk * dk : 0.0;
    const float gf = gd < 3.0e38 ? (float)gd : INFINITY;
    unsigned u = __float_as_uint(gf);  // gf >= 0 or NaN (NaN: bisect below)
    bool found = false;
    if (gf == gf) {
      if (velten_bin(__uint_as_float(u), dso2, g.w, g.off) >= k) {
        for (int it = 0; it < 64; ++it) {  // step down to the first pattern that qualifies
          if (u == 0u || velten_bin(__uint_as_float(u - 1u), dso2, g.w, g.off) < k) {
            found = true;
            break;
          }
          --u;
        }
      } else {
        for (int it = 0; it < 64 && u < 0x7f800000u; ++it) {
          ++u;
          if (velten_bin(__uint_as_float(u), dso2, g.w, g.off) >= k) {
            found = true;
            break;
          }
        }
      }
    }
    if (!found) {
      unsigned lo = 0u, hi = 0x7f800000u;
      if (velten_bin(0.f, dso2, g.w, g.off) >= k) {
        hi = 0u;
      } else {
        while (hi - lo > 1u) {
          const unsigned mid = lo + (hi - lo) / 2u;
          if (velten_bin(__uint_as_float(mid), dso2, g.w, g.off) >= k)
            hi = mid;
          else
            lo = mid;
        }
      }
      u = hi;
    }
    e = __uint_as_float(u);
  }
  row[4 + vr.nv + j] = e;
  if (j == 0) {
    // the largest distance to the box is at a corner
    float dmax = 0.f;
    for (int c = 0; c < 8; ++c)
      dmax = fmaxf(dmax, dist3(px, py, pz, (c & 1) ? g.ex : g.sx, (c & 2) ? g.ey : g.sy,
                               (c & 4) ? g.ez : g.sz));
    const float m = (2.f * dmax + fabsf(dso2) + fabsf(g.off)) / g.w + 1.f;
    const bool est = g.w > 0.f && m < kVeltenMaxBins;  // false for NaN or inf
    *(float4*)row = make_float4(px, py, pz, est ? (dso2 - g.off) * rw + 0.5f
                                                : __builtin_nanf(""));
    row[4] = 0.f;
    row[4 + T + 1] = 0.f;
  }
}

// Bin of squared distance e for the scan point of `row` (stage or global):
// the corrected estimate, or the exact formula where the point's c is NaN.
// The result is clamped to [-1, T], whose ends read the zero pads.
template <bool EXACT>
__device__ __forceinline__ int velten_take_bin(const float* row, const VeltenRow& vr, float e,
                                               float c, float k2, int T, float dso2,
                                               const BoxGeom& g) {
  if (!EXACT) {
    int a = (int)floorf(__builtin_fmaf(sqrt_fast(e), k2, c));
    a = min(max(a, -1), T);
    const float* E = row + 4 + vr.nv + a + 1;
    return a + (e < E[0] ? -1 : 0) + (e >= E[1] ? 1 : 0);
  }
  const int t = velten_bin(e, dso2, g.w, g.off);                             // :177,185-187
  return min(max(t, -1), T);
}

__device__ __forceinline__ float velten_sqdist(const float4& p, float qx, float qy, float qz) {
  const float dx = p.x - qx, dy = p.y - qy, dz = p.z - qz;
  float e = dx * dx;
  e = e + dy * dy;
  e = e + dz * dz;
  return e;
}

// Scan points per unrolled group of the backprojection loop.
constexpr int kVeltenUnroll = 8;

// LDS floats of one stage of scan-point rows (32 KiB, four workgroups per
// CU), and the float4s per thread that copy one stage with a 256-lane
// workgroup.
constexpr int kVeltenStage = kStageFloats;
constexpr int kVeltenCopy4 = kVeltenStage / 4 / 256;

// One lane per voxel, flat over [D][H][W]; rows as laid out above.  A
// workgroup stages the rows of R scan points in LDS (STAGED), loading the
// next group into registers while it works on the current one, or reads
// them from global memory (rows too long for the stage).  Sums run over scan
// points in order, as the reference's loop.
template <bool STAGED>
__global__ __launch_bounds__(256) void velten_confocal_kernel(
    const float* __restrict__ rows, const float* __restrict__ orig, float* __restrict__ vol,
    BoxGeom g, int S, int So, int R, float k2) {
  extern __shared__ float4 lds4[];
  const int T = g.T;
  const VeltenRow vr = velten_row(T);
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
  const size_t bS = (size_t)b * S;
  const float* ob = orig + bS * 3;
  float* stage = (float*)lds4;
  // the rows of the next group of points, loaded into registers
  float4 buf[kVeltenCopy4];
  if (STAGED) {
    const int n4 = min(R, S) * vr.len / 4;
    const float4* src = (const float4*)(rows + bS * vr.len);
#pragma unroll
    for (int i = 0; i < kVeltenCopy4; ++i) {
      const int e = i * 256 + (int)threadIdx.x;
      buf[i] = e < n4 ? src[e] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  float value = 0.f;
  for (int s0 = 0; s0 < S; s0 += R) {
    const int ns = min(R, S - s0);
    const float* base = rows + (bS + s0) * vr.len;
    if (STAGED) {
      __syncthreads();  // the previous group is no longer read
      const int n4 = ns * vr.len / 4;
#pragma unroll
      for (int i = 0; i < kVeltenCopy4; ++i) {
        const int e = i * 256 + (int)threadIdx.x;
        if (e < n4) lds4[e] = buf[i];
      }
      __syncthreads();
      if (s0 + R < S) {
        const int n4n = min(R, S - s0 - R) * vr.len / 4;
        const float4* src = (const float4*)(rows + (bS + s0 + R) * vr.len);
#pragma unroll
        for (int i = 0; i < kVeltenCopy4; ++i) {
          const int e = i * 256 + (int)threadIdx.x;
          buf[i] = e < n4n ? src[e] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
      base = stage;
    }
    // kVeltenUnroll points at a time (the last group of a stage may have
    // fewer, nu; its missing slots repeat the last point and add nothing):
    // their loads and decisions overlap, their values are added in point
    // order.  A group with a point whose c is NaN takes the per-point path
    // (the choice is uniform).
    for (int k = 0; k < ns; k += kVeltenUnroll) {
      const int nu = min(kVeltenUnroll, ns - k);
      float4 p[kVeltenUnroll];
      bool est = true;
#pragma unroll
      for (int u = 0; u < kVeltenUnroll; ++u) {
        p[u] = *(const float4*)(base + (k + min(u, nu - 1)) * vr.len);
        est = est && p[u].w == p[u].w;
      }
      if (est) {
        float a[kVeltenUnroll];
#pragma unroll
        for (int u = 0; u < kVeltenUnroll; ++u) {
          const float* r = base + (k + min(u, nu - 1)) * vr.len;
          const float e = velten_sqdist(p[u], qx, qy, qz);
          a[u] = r[4 + velten_take_bin<false>(r, vr, e, p[u].w, k2, T, 0.f, g) + 1];
        }
        if (nu == kVeltenUnroll) {
#pragma unroll
          for (int u = 0; u < kVeltenUnroll; ++u) value += a[u];            // :188-190
        } else {
#pragma unroll
          for (int u = 0; u < kVeltenUnroll; ++u)
            if (u < nu) value += a[u];
        }
      } else {
        for (int u = 0; u < nu; ++u) {
          const float* r = base + (k + u) * vr.len;
          const float4 q = *(const float4*)r;
          const float e = velten_sqdist(q, qx, qy, qz);
          int t;
          if (q.w == q.w) {
            t = velten_take_bin<false>(r, vr, e, q.w, k2, T, 0.f, g);
          } else {
            const int s = s0 + k + u;
            const float dso = So > 0 ? dist3(q.x, q.y, q.z, ob[3 * s], ob[3 * s + 1],
                                             ob[3 * s + 2])
                                     : 0.f;                                  // :178-183
            t = velten_take_bin<true>(r, vr, e, q.w, k2, T, 2.f * dso, g);
          }
          value += r[4 + t + 1];  // bins -1 and T read the zero pads (:188-190)
        }
      }
    }
  }
  if (ok) vol[(size_t)b * nvox + v] = value;                                 // :192
}

// ---------------------------------------------------------------------------
// A3: vol = sum_{s,l} tr[b][t(s,l)][s][l],
// t = nearest_bin(|p_s - o_s| + |p_s - q| + |p_l - q| + |p_l - o_l|)
// (backprojection_cuda.cu:228-287)
//
// The path length D is computed as the reference does (correctly rounded
// roots, sums in its order).  Its bin floor((D - off) / w + 0.5) is a
// non-decreasing function of D, the same for every pair, so one table of
// thresholds D_k = min{D : bin(D) >= k} (bisection over the float bit
// patterns) turns the division into an estimate a = floor(fma(D, 1/w,
// 0.5 - off/w)), within one bin where D/w stays below 2^16 bins, and its
// correction a - [D < D_a] + [D >= D_(a+1)].  The pair's constant terms
// (the points and their origin distances) are computed once per pair.
// ---------------------------------------------------------------------------

// One row of floats per (scan point, laser point) pair r = s * L + l, 16-byte
// aligned parts:
//   [0, 4)    p_s, |p_s - o_s|        [4, 8)  p_l, |p_l - o_l|
//   [8, 12)   1 where the estimate's bound holds for the pair (else 0),
//             1 where both points keep every squared distance to the box in
//             [2^-64, 2^64] (the short correctly rounded root serves), 0, 0
//   [12, 12 + nv)  its bins, vals[1 + t] = bin t, zeros at 0 and T + 1
//                  (nv = T + 2 rounded up to a multiple of 4)
struct ExRow {
  int nv, len;
};

__host__ __device__ __forceinline__ ExRow ex_row(int T) {
  ExRow r;
  r.nv = (T + 5) & ~3;
  r.len = 12 + r.nv;
  return r;
}

// Thread (b, r): the pair's constants and the zero pads of its bins.  The
// estimate's error: the reference chain rounds the difference, the quotient
// and + 0.5, the estimate 1/w, off/w, the constant and the FMA; under
// 8 * 2^-24 * M bins for M = (D_max + |off|) / w + 1, D_max bounding D over
// the box (origin terms plus the largest distances to the box's corners).
// M < 2^16 keeps that far below one bin.
__global__ __launch_bounds__(256) void velten_ex_rows_kernel(
    const float* __restrict__ scan, const float* __restrict__ laser,
    const float* __restrict__ sorig, const float* __restrict__ lorig, BoxGeom g, int S,
    int L, int So, int Lo, float* __restrict__ rows) {
  const ExRow xr = ex_row(g.T);
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)g.B * S * L) return;
  const int b = (int)(id / ((size_t)S * L));
  const int r = (int)(id - (size_t)b * S * L);
  const int s = r / L, l = r - s * L;
  const float* ps = scan + ((size_t)b * S + s) * 3;
  const float* pl = laser + ((size_t)b * L + l) * 3;
  const float px = ps[0], py = ps[1], pz = ps[2];
  const float lx = pl[0], ly = pl[1], lz = pl[2];
  float dso = 0.f, dlo = 0.f;
  if (So > 0) {
    const float* o = sorig + ((size_t)b * So + s) * 3;
    dso = dist3(px, py, pz, o[0], o[1], o[2]);                               // :256-261
  }
  if (Lo > 0) {
    const float* o = lorig + ((size_t)b * Lo + l) * 3;
    dlo = dist3(lx, ly, lz, o[0], o[1], o[2]);                               // :269-274
  }
  float ds = 0.f, dl = 0.f;
  for (int c = 0; c < 8; ++c) {
    const float cx = (c & 1) ? g.ex : g.sx, cy = (c & 2) ? g.ey : g.sy,
                cz = (c & 4) ? g.ez : g.sz;
    ds = fmaxf(ds, dist3(px, py, pz, cx, cy, cz));
    dl = fmaxf(dl, dist3(lx, ly, lz, cx, cy, cz));
  }
  const float m = (dso + dlo + ds + dl + fabsf(g.off)) / g.w + 1.f;
  const bool est = g.w > 0.f && m < kVeltenMaxBins;  // false for NaN or inf
  // at least 2^-20 from the box and at most 2^30 from its far corners
  const float bs = box_sqdist(px, py, pz, g.sx, g.sy, g.sz, g.ex, g.ey, g.ez);
  const float bl = box_sqdist(lx, ly, lz, g.sx, g.sy, g.sz, g.ex, g.ey, g.ez);
  const bool mid = bs >= 0x1p-40f && bl >= 0x1p-40f && ds <= 0x1p30f && dl <= 0x1p30f;
  float4* h = (float4*)(rows + id * xr.len);
  h[0] = make_float4(px, py, pz, dso);
  h[1] = make_float4(lx, ly, lz, dlo);
  h[2] = make_float4(est ? 1.f : 0.f, mid ? 1.f : 0.f, 0.f, 0.f);
  rows[id * xr.len + 12] = 0.f;
  rows[id * xr.len + 12 + g.T + 1] = 0.f;
}

// Thread j in [0, T + 3): the path-length thresholds
// {-inf, D_0, ..., D_T, NaN}, D_k = +inf when no finite D reaches bin k.
__global__ __launch_bounds__(256) void velten_ex_table_kernel(BoxGeom g,
                                                              float* __restrict__ tab) {
  const int n = g.T + 3;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float d;
  if (j == 0) {
    d = -INFINITY;
  } else if (j == n - 1) {
    d = __builtin_nanf("");
  } else {
    const int k = j - 1;
    unsigned lo = 0u, hi = 0x7f800000u;
    if (nearest_bin(0.f, g.w, g.off) >= k) {
      hi = 0u;
    } else {
      while (hi - lo > 1u) {
        const unsigned mid = lo + (hi - lo) / 2u;
        if (nearest_bin(__uint_as_float(mid), g.w, g.off) >= k)
          hi = mid;
        else
          lo = mid;
      }
    }
    d = __uint_as_float(hi);
  }
  tab[j] = d;
}

// Bin of one (voxel, pair): D as the reference computes it (MID: with the
// short correctly rounded root, the same values), then the corrected estimate
// (EST) or the reference's division; clamped to [-1, T].
template <bool EST, bool MID>
__device__ __forceinline__ int velten_ex_bin(const float* row, const float* tab, float qx,
                                             float qy, float qz, float rw, float c0,
                                             const BoxGeom& g) {
  const float4 hs = *(const float4*)row, hl = *(const float4*)(row + 4);
  const float dsv = MID ? sqrt_rn_mid(sqdist3(hs.x, hs.y, hs.z, qx, qy, qz))
                        : dist3(hs.x, hs.y, hs.z, qx, qy, qz);                // :255
  const float dlv = MID ? sqrt_rn_mid(sqdist3(hl.x, hl.y, hl.z, qx, qy, qz))
                        : dist3(hl.x, hl.y, hl.z, qx, qy, qz);                // :268
  float dist = hs.w + dsv;                                                   // :276-277
  dist = dist + dlv;
  dist = dist + hl.w;
  if (EST) {
    int a = (int)floorf(__builtin_fmaf(dist, rw, c0));
    a = min(max(a, -1), g.T);
    return a + (dist < tab[a + 1] ? -1 : 0) + (dist >= tab[a + 2] ? 1 : 0);
  }
  const int t = nearest_bin(dist, g.w, g.off);                               // :279
  return min(max(t, -1), g.T);
}

// One lane per voxel, flat over [D][H][W]; rows as laid out above, the
// threshold table in LDS.  Rows are staged in LDS as in A2 (STAGED), or read
// from global memory (long rows).  Sums run over pairs in the reference's
// order (scan point outer, laser point inner).
template <bool STAGED>
__global__ __launch_bounds__(256) void velten_exhaustive_kernel(
    const float* __restrict__ rows, const float* __restrict__ gtab, float* __restrict__ vol,
    BoxGeom g, int NR, int R, float rw, float c0) {
  extern __shared__ float4 lds4[];
  const int T = g.T;
  const ExRow xr = ex_row(T);
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
  const size_t bR = (size_t)b * NR;
  // LDS: the threshold table (T + 3 floats, rounded to float4s), then the stage
  const int ntab4 = (T + 6) / 4;
  float* tab = (float*)lds4;
  float* stage = (float*)(lds4 + ntab4);
  const float* tb = gtab;
  if (STAGED) {
    for (int e = threadIdx.x; e < T + 3; e += blockDim.x) tab[e] = gtab[e];
    tb = tab;
  }
  float4 buf[kVeltenCopy4];
  if (STAGED) {
    const int n4 = min(R, NR) * xr.len / 4;
    const float4* src = (const float4*)(rows + bR * xr.len);
#pragma unroll
    for (int i = 0; i < kVeltenCopy4; ++i) {
      const int e = i * 256 + (int)threadIdx.x;
      buf[i] = e < n4 ? src[e] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  float value = 0.f;
  for (int r0 = 0; r0 < NR; r0 += R) {
    const int nr = min(R, NR - r0);
    const float* base = rows + (bR + r0) * xr.len;
    if (STAGED) {
      __syncthreads();  // the previous group is no longer read
      const int n4 = nr * xr.len / 4;
#pragma unroll
      for (int i = 0; i < kVeltenCopy4; ++i) {
        const int e = i * 256 + (int)threadIdx.x;
        if (e < n4) ((float4*)stage)[e] = buf[i];
      }
      __syncthreads();
      if (r0 + R < NR) {
        const int n4n = min(R, NR - r0 - R) * xr.len / 4;
        const float4* src = (const float4*)(rows + (bR + r0 + R) * xr.len);
#pragma unroll
        for (int i = 0; i < kVeltenCopy4; ++i) {
          const int e = i * 256 + (int)threadIdx.x;
          buf[i] = e < n4n ? src[e] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
      base = stage;
    }
    // groups of kVeltenUnroll pairs, as in A2
    for (int k = 0; k < nr; k += kVeltenUnroll) {
      const int nu = min(kVeltenUnroll, nr - k);
      bool est = true;
#pragma unroll
      for (int u = 0; u < kVeltenUnroll; ++u) {
        const float* r = base + (k + min(u, nu - 1)) * xr.len;
        est = est && r[8] != 0.f && r[9] != 0.f;
      }
      if (est) {
        float a[kVeltenUnroll];
#pragma unroll
        for (int u = 0; u < kVeltenUnroll; ++u) {
          const float* r = base + (k + min(u, nu - 1)) * xr.len;
          a[u] = r[12 + velten_ex_bin<true, true>(r, tb, qx, qy, qz, rw, c0, g) + 1];
        }
        if (nu == kVeltenUnroll) {
#pragma unroll
          for (int u = 0; u < kVeltenUnroll; ++u) value += a[u];            // :280-282
        } else {
#pragma unroll
          for (int u = 0; u < kVeltenUnroll; ++u)
            if (u < nu) value += a[u];
        }
      } else {
        for (int u = 0; u < nu; ++u) {
          const float* r = base + (k + u) * xr.len;
          const int t = r[8] != 0.f ? velten_ex_bin<true, false>(r, tb, qx, qy, qz, rw, c0, g)
                                    : velten_ex_bin<false, false>(r, tb, qx, qy, qz, rw, c0, g);
          value += r[12 + t + 1];  // bins -1 and T read the zero pads (:280-282)
        }
      }
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

int transpose_strided(const float* in, float* out, int B, int R, int C, int ldo, int ooff,
                      hipStream_t st) {
  dim3 grid((C + 31) / 32, (R + 31) / 32, B);
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, st, in, out, R, C, ldo, ooff);
  return check_launch("transpose_kernel");
}

int transpose_batched(const float* in, float* out, int B, int R, int C, hipStream_t st) {
  return transpose_strided(in, out, B, R, C, R, 0, st);
}

}  // namespace totri_hip

using namespace totri_hip;

namespace totri_hip {
namespace {

// Workspace of A2: one row per scan point (velten_row), [B][S][len] floats.
size_t velten_bytes(int64_t B, int64_t T, int64_t S) {
  return align256((size_t)B * S * velten_row((int)T).len * sizeof(float));
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_backprojection_confocal_velten_workspace(int64_t B, int64_t T,
                                                                 int64_t S) {
  if (B <= 0 || T < 0 || S <= 0) return 0;
  return totri_hip::velten_bytes(B, T, S);
}

namespace totri_hip {
namespace {

// A2 with its scratch (the transient's value rows, the bin thresholds and the
// per-point constants) supplied by the caller: the workspace of the entry
// point below, or the library's own array for the allocation-free variant.
int velten_confocal(const float* transient, int64_t B, int64_t T, int64_t S,
                    const float* scan_points, const float* scan_origin, int64_t So,
                    float* volume, int64_t D, int64_t H, int64_t W,
                    const float volume_start[3], const float volume_end[3], float bin_width,
                    float bin_offset, void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_dims({B, T, S, So, D, H, W}, B * T * S, B * D * H * W);
  if (rc) return rc;
  TOTRI_CHECK_ARG(So == 0 || So == S, "scan_origin must have 0 or S points");
  TOTRI_CHECK_ARG(volume_start && volume_end, "null volume box");
  TOTRI_CHECK_ARG(T < (1 << 30), "T=%lld too large", (long long)T);
  hipStream_t st = (hipStream_t)stream;
  const size_t nvox = (size_t)D * H * W;
  if (B == 0 || nvox == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr, "null volume");
  const BoxGeom g = box_geom(B, D, H, W, T, volume_start, volume_end, bin_width, bin_offset);
  if (T == 0 || S == 0) {  // no bins or no scan points: a zero volume
    hipLaunchKernelGGL(zero_kernel, dim3(grid1d(B * nvox)), dim3(256), 0, st, volume,
                       (size_t)B * nvox);
    return check_launch("zero_kernel");
  }
  TOTRI_CHECK_ARG(transient && scan_points, "null input");
  TOTRI_CHECK_ARG(So == 0 || scan_origin, "null scan_origin");
  const size_t need = velten_bytes(B, T, S);
  if (workspace_bytes < need)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", need,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr && ((uintptr_t)workspace & 15) == 0,
                  "workspace must be a non-null, 16-byte aligned device pointer");
  float* rows = (float*)workspace;
  const VeltenRow vr = velten_row((int)T);
  rc = transpose_strided(transient, rows, (int)B, (int)T, (int)S, vr.len, 5, st);
  if (rc) return rc;
  const float rw = 1.f / bin_width;
  const size_t nthr = (size_t)B * S * (T + 3);
  hipLaunchKernelGGL(velten_bins_kernel, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st,
                     scan_points, scan_origin, g, (int)S, (int)So, rw, rows);
  rc = check_launch("velten_bins_kernel");
  if (rc) return rc;
  // Scan points per LDS stage.  Where fewer than four rows fit (T above
  // about 1020 bins) the stage's barriers would come every few points: the
  // rows are then read from global memory, through the caches.
  int R = (int)std::min<int64_t>(S, kVeltenStage / vr.len);
  if (R < 4 && R < S) R = 0;
  const float k2 = 2.f * rw;
  const dim3 grid(blocks_for(nvox), (unsigned)B);
  if (R > 0) {
    hipLaunchKernelGGL(velten_confocal_kernel<true>, grid, dim3(256),
                       (size_t)R * vr.len * sizeof(float), st, rows, scan_origin, volume, g,
                       (int)S, (int)So, R, k2);
  } else {
    hipLaunchKernelGGL(velten_confocal_kernel<false>, grid, dim3(256), 0, st, rows,
                       scan_origin, volume, g, (int)S, (int)So, (int)S, k2);
  }
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

namespace totri_hip {
namespace {

// Workspace of A3: the pair rows [B][S*L][len], then the threshold table.
struct ExLayout {
  size_t tab_off, end;
};

ExLayout ex_layout(int64_t B, int64_t T, int64_t S, int64_t L) {
  ExLayout l;
  l.tab_off = align256((size_t)B * S * L * ex_row((int)T).len * sizeof(float));
  l.end = align256(l.tab_off + (size_t)(T + 3) * sizeof(float));
  return l;
}

}  // namespace
}  // namespace totri_hip

extern "C" size_t totri_backprojection_exhaustive_velten_workspace(int64_t B, int64_t T,
                                                                   int64_t S, int64_t L) {
  if (B <= 0 || T < 0 || S <= 0 || L <= 0) return 0;
  return totri_hip::ex_layout(B, T, S, L).end;
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
  TOTRI_CHECK_ARG(T < (1 << 30), "T=%lld too large", (long long)T);
  hipStream_t st = (hipStream_t)stream;
  const size_t nvox = (size_t)D * H * W;
  if (B == 0 || nvox == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(volume != nullptr, "null volume");
  const BoxGeom g = box_geom(B, D, H, W, T, volume_start, volume_end, bin_width, bin_offset);
  if (T == 0 || S == 0 || L == 0) {  // no bins or no pairs: a zero volume
    hipLaunchKernelGGL(zero_kernel, dim3(grid1d(B * nvox)), dim3(256), 0, st, volume,
                       (size_t)B * nvox);
    return check_launch("zero_kernel");
  }
  TOTRI_CHECK_ARG(transient && scan_points && laser_points, "null input");
  TOTRI_CHECK_ARG((So == 0 || scan_origin) && (Lo == 0 || laser_origin), "null origin");
  const ExLayout lay = ex_layout(B, T, S, L);
  if (workspace_bytes < lay.end)
    return fail(TOTRI_EWORKSPACE, "workspace too small: need %zu bytes, got %zu", lay.end,
                workspace_bytes);
  TOTRI_CHECK_ARG(workspace != nullptr && ((uintptr_t)workspace & 15) == 0,
                  "workspace must be a non-null, 16-byte aligned device pointer");
  float* rows = (float*)workspace;
  float* tab = (float*)((char*)workspace + lay.tab_off);
  const ExRow xr = ex_row((int)T);
  const int NR = (int)(S * L);
  rc = transpose_strided(transient, rows, (int)B, (int)T, NR, xr.len, 13, st);
  if (rc) return rc;
  const size_t npairs = (size_t)B * NR;
  hipLaunchKernelGGL(velten_ex_rows_kernel, dim3((unsigned)((npairs + 255) / 256)), dim3(256),
                     0, st, scan_points, laser_points, scan_origin, laser_origin, g, (int)S,
                     (int)L, (int)So, (int)Lo, rows);
  rc = check_launch("velten_ex_rows_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(velten_ex_table_kernel, dim3((unsigned)((T + 3 + 255) / 256)), dim3(256), 0,
                     st, g, tab);
  rc = check_launch("velten_ex_table_kernel");
  if (rc) return rc;
  const float rw = 1.f / bin_width;
  const float c0 = -bin_offset * rw + 0.5f;
  // pairs per LDS stage (beside the table); with fewer than four per stage
  // the rows and the table are read from global memory
  int R = (int)std::min<int64_t>(NR, kVeltenStage / xr.len);
  if (R < 4 && R < NR) R = 0;
  const dim3 grid(blocks_for(nvox), (unsigned)B);
  if (R > 0) {
    const size_t lds = ((size_t)(T + 6) / 4 * 4 + (size_t)R * xr.len) * sizeof(float);
    hipLaunchKernelGGL(velten_exhaustive_kernel<true>, grid, dim3(256), lds, st, rows, tab,
                       volume, g, NR, R, rw, c0);
  } else {
    hipLaunchKernelGGL(velten_exhaustive_kernel<false>, grid, dim3(256), 0, st, rows, tab,
                       volume, g, NR, NR, rw, c0);
  }
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
