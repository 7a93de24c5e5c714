// f-k migration (Lindell et al. 2019) around a library FFT, gfx950.
//
// Reference: totri/reconstruct/fk_migration.py:5-120 (pure PyTorch, no CUDA
// source).  For a confocal transient t[D][H][W]:
//   P[d][h][w] = sqrt(t[d][h][w]) * lin(d; 0, 1, D)  for d < D, 0 for d >= D   (:108-112)
//   F = fftn(P)                                   (:114; the host layer calls hipFFT)
//       P is real, so F[kz][ky][kx] = conj F[-kz][-ky][-kx]: the host layer
//       may pass the one-sided transform (kx <= W/2, rfftn) instead
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
// depth rows D..2D-1.  Like P's padding rows, it can be left zero in a
// buffer kept from call to call, so the pre-pass and the Stolt kernel write
// only the other half.
#include <math.h>

#include <algorithm>

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

// Workgroup order that keeps neighbouring work on one XCD: workgroups are
// dealt to the 8 XCDs in turn, so workgroup b runs on XCD b mod 8; the
// first 8 * floor(n / 8) are renumbered so that each XCD gets a contiguous
// range (the rest keep their numbers).
__device__ __forceinline__ int fk_xcd_order(int b, int n) {
  const int full = n & ~7;
  if (b >= full) return b;
  return (b & 7) * (full >> 3) + (b >> 3);
}

struct FkStoltGeom {
  int C, D2, H, W;
  float ax, ay;  // (W depth / (D2 width))^2, (H depth / (D2 height))^2
  int fold;      // 1: read raw fftn order, write ifftn order
  int onesided;  // 1: the spectrum holds kx <= W/2 only (rfftn of a real input)
  int k0;        // first depth index written (D2 / 2 when the zero half is skipped)
};

__global__ __launch_bounds__(kFkBlock) void fk_prepare_kernel(
    const float* __restrict__ t,  // [D][HW]
    float* __restrict__ P,        // [2D][HW] (rows < D, or all rows)
    int D, int HW, int blocks_per_row) {
  const int d = blockIdx.x / blocks_per_row;
  const int e = (blockIdx.x - d * blocks_per_row) * kFkBlock + threadIdx.x;
  if (e >= HW) return;
  const size_t i = (size_t)d * HW + e;
  float v = 0.f;
  if (d < D) v = sqrtf(t[i]) * fk_lin(d, D, 0.f, 1.f);
  P[i] = v;
}

// One wave per output row (depth index k, wall row h), lanes along the wall
// row.  y, z and the two y corners are wave-uniform; each lane gathers its
// eight corners at clamped (always valid) addresses and zeroes the ones
// outside the spectrum, so the loads issue back to back.  Corners are added
// in grid_sample's order: z outer, then y, then x.
__global__ __launch_bounds__(kFkBlock) void fk_stolt_kernel(
    const float2* __restrict__ F,  // [C][D2][H][W]
    float2* __restrict__ G,        // [C][D2][H][W]
    FkStoltGeom g, int n_rows) {
  const int row = __builtin_amdgcn_readfirstlane(blockIdx.x * (kFkBlock / 64) +
                                                 (int)(threadIdx.x >> 6));
  if (row >= n_rows) return;
  const int lane = threadIdx.x & 63;
  const int kr = row / g.H;
  const int h = row - kr * g.H;
  const int k = g.k0 + kr;
  const size_t plane = (size_t)g.H * g.W;
  const float2* Fc = F + (size_t)blockIdx.y * g.D2 * g.H *
                            (size_t)(g.onesided ? g.W / 2 + 1 : g.W);
  float2* Gc = G + (size_t)blockIdx.y * g.D2 * plane;

  const float y = fk_lin(h, g.H, -1.f, 1.f);
  const float z = fk_lin(k, g.D2, -1.f, 1.f);
  const float y2 = (y * y) * g.ay;
  const float z2 = z * z;
  const float iy = fk_unnormalize(y, g.H);
  const float fy = floorf(iy);
  const int y0 = (int)fy;
  const float wy[2] = {(fy + 1.f) - iy, iy - fy};
  // Spectrum rows hold Wf = W (full) or W/2 + 1 (one-sided) elements.  A
  // one-sided corner with kx > W/2 is read as conj F[-kz][-ky][W - kx]: the
  // y and z offsets come in a direct and a mirrored form.
  const int Wf = g.onesided ? g.W / 2 + 1 : g.W;
  const size_t fplane = (size_t)g.H * Wf;
  bool yok[2];
  size_t yoff[2], yoffm[2];
#pragma unroll
  for (int dy = 0; dy < 2; ++dy) {
    const int yi = y0 + dy;
    yok[dy] = yi >= 0 && yi < g.H;
    const int yc = min(max(yi, 0), g.H - 1);
    const int uy = g.fold ? fk_unshift(yc, g.H) : yc;
    yoff[dy] = (size_t)uy * Wf;
    yoffm[dy] = (size_t)(uy == 0 ? 0 : g.H - uy) * Wf;
  }
  const size_t orow = (size_t)(g.fold ? fk_unshift(k, g.D2) : k) * plane +
                      (size_t)(g.fold ? fk_unshift(h, g.H) : h) * g.W;

  for (int w = lane; w < g.W; w += 64) {
    float re = 0.f, im = 0.f;
    if (z > 0.f) {
      const float x = fk_lin(w, g.W, -1.f, 1.f);
      float zz = (x * x) * g.ax;
      zz = zz + y2;
      zz = zz + z2;
      zz = sqrtf(zz);
      const float wgt = fabsf(z) / fmaxf(zz, 1e-6f);
      const float ix = fk_unnormalize(x, g.W);
      const float iz = fk_unnormalize(zz, g.D2);
      // zz >= 0 keeps iz >= (D2 - 1) / 2; past D2 (or not finite) every
      // corner is padding.
      const bool zin = iz < (float)g.D2;
      const float fx = floorf(ix), fz = zin ? floorf(iz) : 0.f;
      const int x0 = (int)fx, z0 = (int)fz;
      const float wx[2] = {(fx + 1.f) - ix, ix - fx};
      const float wz[2] = {(fz + 1.f) - iz, iz - fz};
      bool xok[2], zok[2], xm[2];
      int xo[2];
      size_t zo[2], zom[2];
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const int xi = x0 + d, zi = z0 + d;
        xok[d] = xi >= 0 && xi < g.W;
        zok[d] = zin && zi < g.D2;
        const int xc = min(max(xi, 0), g.W - 1), zc = min(max(zi, 0), g.D2 - 1);
        const int ux = g.fold ? fk_unshift(xc, g.W) : xc;
        const int uz = g.fold ? fk_unshift(zc, g.D2) : zc;
        xm[d] = g.onesided && ux > g.W / 2;
        xo[d] = xm[d] ? g.W - ux : ux;
        zo[d] = (size_t)uz * fplane;
        zom[d] = (size_t)(uz == 0 ? 0 : g.D2 - uz) * fplane;
      }
      float2 v[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int dz = c >> 2, dy = (c >> 1) & 1, dx = c & 1;
        v[c] = xm[dx] ? Fc[zom[dz] + yoffm[dy] + xo[dx]] : Fc[zo[dz] + yoff[dy] + xo[dx]];
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int dz = c >> 2, dy = (c >> 1) & 1, dx = c & 1;
        const bool ok = zok[dz] && yok[dy] && xok[dx];
        const float cw = (wx[dx] * wy[dy]) * wz[dz];
        const float vi = xm[dx] ? -v[c].y : v[c].y;
        re = re + (ok ? v[c].x : 0.f) * cw;
        im = im + (ok ? vi : 0.f) * cw;
      }
      re = re * wgt;
      im = im * wgt;
    }
    Gc[orow + (g.fold ? fk_unshift(w, g.W) : w)] = make_float2(re, im);
  }
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

// ---------------------------------------------------------------------------
// Native FFT path (power-of-two sizes).  The whole migration in five kernels:
//   fk_slice_fwd_kernel  pre-pass and 2-D FFT of each of the D nonzero depth
//                        slices; keeps kx <= W/2 (the input is real)
//   fk_depth_fwd_kernel  length-2D FFT of each (ky, kx) column; the padding
//                        rows d >= D are zeros in LDS, never read
//   fk_stolt_kernel      one-sided layout, nonzero half only
//   fk_depth_inv_kernel  inverse length-2D FFT of each (ky, kx) column from
//                        its D nonzero rows, keeping the D output rows used
//   fk_slice_inv_kernel  inverse 2-D FFT of each kept slice and |.| / N
// Each FFT runs in LDS as radix-2 decimation in frequency (natural order in,
// bit-reversed order out, read back through the bit reversal), three stages
// per pass in registers, with twiddles exp(-2 pi i k / N) computed in double.
// ---------------------------------------------------------------------------
constexpr int kFkMaxPlane = 16384;    // H * W of a slice held in LDS
constexpr int kFkMaxDepth2 = 4096;    // 2D of a column held in LDS
constexpr int kFkColumnLds = 8704;    // float2 of columns per depth workgroup (68 KiB: two per CU)
constexpr int kFkSliceThreads = 512;
constexpr int kFkDepthThreads = 1024;

__device__ __forceinline__ int fk_bitrev(int i, int logn) {
  return logn == 0 ? 0 : (int)(__builtin_bitreverse32((unsigned)i) >> (32 - logn));
}

// Multiply by W_8^r = exp(-2 pi i r / 8) (conjugated for the inverse), r < 4.
template <bool INV, int R>
__device__ __forceinline__ float2 fk_rot8(float2 b) {
  constexpr float c = 0.70710678118654752f;
  if (R == 0) return b;
  if (R == 2) return INV ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
  if (R == 1)
    return INV ? make_float2(c * (b.x - b.y), c * (b.x + b.y))
               : make_float2(c * (b.x + b.y), c * (b.y - b.x));
  // R == 3: W_8^3 = W_8 * W_8^2
  const float2 t = fk_rot8<INV, 1>(b);
  return INV ? make_float2(-t.y, t.x) : make_float2(t.y, -t.x);
}

// The K stages of one pass on a thread's 2^K registers (see fk_dif_pass).
// Stage st = lq + l pairs v[m] and v[m + 2^l] (bit l of m clear); its twiddle
// W_2S^(i mod S), i = j + m q, S = 2^st, is b_l * W_(2^(l+1))^(m mod 2^l) with
// b_l = W_2S^j = tw[j << (logn - 1 - st)]: one table read per stage, the
// rest multiplications by eighth roots of unity.
template <bool INV, int K>
__device__ __forceinline__ void fk_dif_stages(float2 (&v)[1 << K], int j, int lq, int logn,
                                              const float2* tw) {
  constexpr int G = 1 << K;
#pragma unroll
  for (int l = K - 1; l >= 0; --l) {
    const int h = 1 << l;
    float2 b = tw[j << (logn - 1 - (lq + l))];
    if (INV) b.y = -b.y;
#pragma unroll
    for (int m = 0; m < G; ++m) {
      if (m & h) continue;
      const int r = (m & (h - 1)) << (2 - l);  // eighths: W_(2h)^(m mod h) = W_8^(r)
      float2 w;
      if (l == 0 || r == 0) w = b;
      else if (r == 1) w = fk_rot8<INV, 1>(b);
      else if (r == 2) w = fk_rot8<INV, 2>(b);
      else w = fk_rot8<INV, 3>(b);
      const float2 a = v[m], c = v[m + h];
      const float dr = a.x - c.x, di = a.y - c.y;
      v[m] = make_float2(a.x + c.x, a.y + c.y);
      v[m + h] = make_float2(dr * w.x - di * w.y, dr * w.y + di * w.x);
    }
  }
}

// One LDS pass of K consecutive radix-2 decimation-in-frequency stages
// (spans 2^s, 2^(s-1), ..., 2^(s-K+1)).  Stage span S pairs elements i and
// i + S with i mod 2S < S: a - c is multiplied by W_2S^(i mod S) =
// tw[(i mod S) * N / 2S].  The 2^K elements base + j + m q (m < 2^K,
// q = 2^(s-K+1), j < q, base a multiple of 2^(s+1)) only meet each other in
// these K stages, so a thread loads them, runs the K stages in registers and
// stores them back.  tf_fast puts consecutive lanes on consecutive
// transforms (for transforms that are adjacent in memory).
template <bool INV, int K>
__device__ __forceinline__ void fk_dif_pass(float2* buf, int logn, int s, int n_tf,
                                            int log_ntf, int tf_stride, int el_stride,
                                            const float2* tw, bool tf_fast) {
  constexpr int G = 1 << K;
  const int lq = s - K + 1;
  const int q = 1 << lq;
  const int lgpt = logn - K;  // log2 of the groups per transform
  const int total = n_tf << lgpt;
  const int step = q * el_stride;
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    int t, gg;
    if (tf_fast) {
      t = b & ((1 << log_ntf) - 1);
      gg = b >> log_ntf;
    } else {
      t = b >> lgpt;
      gg = b & ((1 << lgpt) - 1);
    }
    const int j = gg & (q - 1);
    const int base = (gg >> lq) << (s + 1);
    float2* p = buf + t * tf_stride + (base + j) * el_stride;
    float2 v[G];
#pragma unroll
    for (int m = 0; m < G; ++m) v[m] = p[m * step];
    fk_dif_stages<INV, K>(v, j, lq, logn, tw);
#pragma unroll
    for (int m = 0; m < G; ++m) p[m * step] = v[m];
  }
}

// In-place radix-2 DIF FFT of n_tf transforms of length 2^logn in LDS:
// element e of transform t is buf[t * tf_stride + e * el_stride].  The result
// is in bit-reversed order.  tw[k] = exp(-2 pi i k / 2^logn), k < 2^(logn-1),
// conjugated for the inverse.  Three stages per LDS pass; every pass ends
// with a barrier.  tf_fast needs n_tf = 2^log_ntf.  s_first >= 0 starts at
// the stage of span 2^s_first (the earlier stages done by the caller).
template <bool INV>
__device__ void fk_fft_dif(float2* buf, int logn, int n_tf, int tf_stride, int el_stride,
                           const float2* tw, bool tf_fast = false, int log_ntf = 0,
                           int s_first = -1) {
  int s = s_first < 0 ? logn - 1 : s_first;
  while (s >= 0) {
    if (s >= 2) {
      fk_dif_pass<INV, 3>(buf, logn, s, n_tf, log_ntf, tf_stride, el_stride, tw, tf_fast);
      s -= 3;
    } else if (s == 1) {
      fk_dif_pass<INV, 2>(buf, logn, s, n_tf, log_ntf, tf_stride, el_stride, tw, tf_fast);
      s -= 2;
    } else {
      fk_dif_pass<INV, 1>(buf, logn, s, n_tf, log_ntf, tf_stride, el_stride, tw, tf_fast);
      s -= 1;
    }
    __syncthreads();
  }
}

// tw[k] = exp(-2 pi i k / N) for the three tables (W, H and 2D) laid end to end.
__global__ __launch_bounds__(256) void fk_twiddle_kernel(float2* tw, int nW, int nH, int nZ,
                                                         int W, int H, int Z) {
  int i = blockIdx.x * 256 + threadIdx.x;
  int N;
  float2* out = tw;
  if (i < nW) {
    N = W;
  } else if (i < nW + nH) {
    i -= nW;
    out += nW;
    N = H;
  } else if (i < nW + nH + nZ) {
    i -= nW + nH;
    out += nW + nH;
    N = Z;
  } else {
    return;
  }
  double sn, cs;
  sincospi(-2.0 * (double)i / (double)N, &sn, &cs);
  out[i] = make_float2((float)cs, (float)sn);
}

__device__ __forceinline__ void fk_stage_twiddles(float2* dst, const float2* src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__global__ __launch_bounds__(kFkSliceThreads) void fk_slice_fwd_kernel(
    const float* __restrict__ t,    // [D][H][W]
    float2* __restrict__ S,         // [D][H][W/2+1]
    const float2* __restrict__ twW, const float2* __restrict__ twH, int D, int logH,
    int logW) {
  extern __shared__ float2 lds2[];
  const int H = 1 << logH, W = 1 << logW, ls = W + 1, HW = H * W, Wh = W / 2 + 1;
  float2* buf = lds2;
  float2* tw_w = buf + H * ls;
  float2* tw_h = tw_w + (W / 2 + 1);
  fk_stage_twiddles(tw_w, twW, W / 2);
  fk_stage_twiddles(tw_h, twH, H / 2);
  const int d = blockIdx.x;
  const float ramp = fk_lin(d, D, 0.f, 1.f);
  const float* td = t + (size_t)d * HW;
  for (int e = threadIdx.x; e < HW; e += blockDim.x)
    buf[(e >> logW) * ls + (e & (W - 1))] = make_float2(sqrtf(td[e]) * ramp, 0.f);
  __syncthreads();
  fk_fft_dif<false>(buf, logW, H, ls, 1, tw_w);  // rows: over w
  fk_fft_dif<false>(buf, logH, W, 1, ls, tw_h, true, logW);  // columns: over h
  float2* Sd = S + (size_t)d * H * Wh;
  for (int e = threadIdx.x; e < H * Wh; e += blockDim.x) {
    const int ky = e / Wh, kx = e - ky * Wh;
    Sd[e] = buf[fk_bitrev(ky, logH) * ls + fk_bitrev(kx, logW)];
  }
}

// Columns c0 .. c0 + nb - 1 of src [Z/2][ncol] (rows Z/2 .. Z-1 are zero),
// one length-Z transform each, column c at buf[c * (Z + 1) + e].  For Z >= 8
// the first three stages run on the registers loaded from global memory
// (the zero half is never stored), then the rest in LDS.
template <bool INV>
__device__ void fk_depth_transform(const float2* __restrict__ src, int ncol, int c0, int nb,
                                   float2* buf, int logZ, const float2* tw) {
  const int Z = 1 << logZ, D = Z / 2, ts = Z + 1;
  if (logZ >= 3) {
    const int q = Z >> 3;
    for (int b = threadIdx.x; b < nb * q; b += blockDim.x) {
      const int j = b / nb, c = b - j * nb;
      float2 v[8];
#pragma unroll
      for (int m = 0; m < 4; ++m) v[m] = src[(size_t)(j + m * q) * ncol + c0 + c];
#pragma unroll
      for (int m = 4; m < 8; ++m) v[m] = make_float2(0.f, 0.f);
      fk_dif_stages<INV, 3>(v, j, logZ - 3, logZ, tw);
      float2* p = buf + c * ts + j;
#pragma unroll
      for (int m = 0; m < 8; ++m) p[m * q] = v[m];
    }
    __syncthreads();
    if (logZ > 3) fk_fft_dif<INV>(buf, logZ, nb, ts, 1, tw, false, 0, logZ - 4);
  } else {
    for (int e = threadIdx.x; e < nb * D; e += blockDim.x) {
      const int d = e / nb, c = e - d * nb;
      buf[c * ts + d] = src[(size_t)d * ncol + c0 + c];
      buf[c * ts + D + d] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    fk_fft_dif<INV>(buf, logZ, nb, ts, 1, tw);
  }
}

__global__ __launch_bounds__(kFkDepthThreads) void fk_depth_fwd_kernel(
    const float2* __restrict__ S,   // [D][ncol]
    float2* __restrict__ F,         // [2D][ncol]
    const float2* __restrict__ twZ, int logZ, int ncol, int cb) {
  extern __shared__ float2 lds2[];
  const int Z = 1 << logZ;
  const int c0 = fk_xcd_order(blockIdx.x, gridDim.x) * cb, nb = min(cb, ncol - c0);
  float2* buf = lds2;
  float2* tw = buf + cb * (Z + 1);
  fk_stage_twiddles(tw, twZ, Z / 2);
  __syncthreads();
  fk_depth_transform<false>(S, ncol, c0, nb, buf, logZ, tw);
#pragma unroll 4
  for (int e = threadIdx.x; e < nb * Z; e += blockDim.x) {
    const int kz = e / nb, c = e - kz * nb;
    F[(size_t)kz * ncol + c0 + c] = buf[c * (Z + 1) + fk_bitrev(kz, logZ)];
  }
}

__global__ __launch_bounds__(kFkDepthThreads) void fk_depth_inv_kernel(
    const float2* __restrict__ G,   // [D][ncol]: rows 0 .. D-1 of the spectrum, the rest zero
    float2* __restrict__ V,         // [D][ncol]: output rows 0 .. D-1
    const float2* __restrict__ twZ, int logZ, int ncol, int cb) {
  extern __shared__ float2 lds2[];
  const int Z = 1 << logZ, D = Z / 2;
  const int c0 = fk_xcd_order(blockIdx.x, gridDim.x) * cb, nb = min(cb, ncol - c0);
  float2* buf = lds2;
  float2* tw = buf + cb * (Z + 1);
  fk_stage_twiddles(tw, twZ, Z / 2);
  __syncthreads();
  fk_depth_transform<true>(G, ncol, c0, nb, buf, logZ, tw);
#pragma unroll 4
  for (int e = threadIdx.x; e < nb * D; e += blockDim.x) {
    const int n = e / nb, c = e - n * nb;
    V[(size_t)n * ncol + c0 + c] = buf[c * (Z + 1) + fk_bitrev(n, logZ)];
  }
}

__global__ __launch_bounds__(kFkSliceThreads) void fk_slice_inv_kernel(
    const float2* __restrict__ V,   // [D][H][W]
    float* __restrict__ out,        // [D][H][W]
    const float2* __restrict__ twW, const float2* __restrict__ twH, int logH, int logW,
    float scale) {
  extern __shared__ float2 lds2[];
  const int H = 1 << logH, W = 1 << logW, ls = W + 1, HW = H * W;
  float2* buf = lds2;
  float2* tw_w = buf + H * ls;
  float2* tw_h = tw_w + (W / 2 + 1);
  fk_stage_twiddles(tw_w, twW, W / 2);
  fk_stage_twiddles(tw_h, twH, H / 2);
  const size_t off = (size_t)blockIdx.x * HW;
  for (int e = threadIdx.x; e < HW; e += blockDim.x)
    buf[(e >> logW) * ls + (e & (W - 1))] = V[off + e];
  __syncthreads();
  fk_fft_dif<true>(buf, logH, W, 1, ls, tw_h, true, logW);   // columns: over ky
  fk_fft_dif<true>(buf, logW, H, ls, 1, tw_w);   // rows: over kx
  for (int e = threadIdx.x; e < HW; e += blockDim.x) {
    const float2 v = buf[fk_bitrev(e >> logW, logH) * ls + fk_bitrev(e & (W - 1), logW)];
    out[off + e] = hypotf(v.x * scale, v.y * scale);
  }
}

int fk_log2(int64_t n) {
  int l = 0;
  while (((int64_t)1 << l) < n) ++l;
  return ((int64_t)1 << l) == n ? l : -1;
}

struct FkNativePlan {
  bool ok;
  int logD, logH, logW, logZ, cb_fwd, cb_inv;
  size_t tw_off, a_off, b_off, bytes;
  size_t slice_lds, depth_lds;
};

FkNativePlan plan_fk_native(int64_t D, int64_t H, int64_t W) {
  FkNativePlan p{};
  p.logD = fk_log2(D);
  p.logH = fk_log2(H);
  p.logW = fk_log2(W);
  p.ok = D >= 1 && H >= 1 && W >= 1 && p.logD >= 0 && p.logH >= 0 && p.logW >= 0 &&
         H * W <= kFkMaxPlane && 2 * D <= kFkMaxDepth2;
  if (!p.ok) return p;
  p.logZ = p.logD + 1;
  const int64_t Z = 2 * D, Wh = W / 2 + 1;
  const int cb = (int)std::max<int64_t>(1, std::min<int64_t>(16, kFkColumnLds / (Z + 1)));
  p.cb_fwd = p.cb_inv = cb;
  auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t ntw = (size_t)(W / 2 + H / 2 + Z / 2);
  p.tw_off = 0;
  p.a_off = align(ntw * sizeof(float2));
  const size_t a_elems = (size_t)D * H * std::max<int64_t>(Wh, W);       // S, then G
  p.b_off = p.a_off + align(a_elems * sizeof(float2));
  const size_t b_elems = std::max<size_t>((size_t)Z * H * Wh, (size_t)D * H * W);  // F, then V
  p.bytes = p.b_off + align(b_elems * sizeof(float2));
  p.slice_lds = ((size_t)H * (W + 1) + (W / 2 + 1) + (H / 2 + 1)) * sizeof(float2);
  p.depth_lds = ((size_t)cb * (Z + 1) + Z / 2) * sizeof(float2);
  return p;
}

template <typename K>
int fk_allow_lds(K kernel, size_t bytes, const char* what) {
  if (bytes <= 65536) return TOTRI_OK;
  hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess)
    return fail(TOTRI_EHIP, "%s: %zu bytes of LDS: %s", what, bytes, hipGetErrorString(e));
  return TOTRI_OK;
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
                                float* padded, int pad_zeroed, void* stream) {
  int rc = check_fk_dims(1, 2 * D, H, W);
  if (rc) return rc;
  const int64_t HW = H * W;
  if (D == 0 || HW == 0) return TOTRI_OK;
  TOTRI_CHECK_ARG(transient != nullptr && padded != nullptr, "null pointer");
  const int64_t bpr = (HW + kFkBlock - 1) / kFkBlock;
  TOTRI_CHECK_ARG(2 * D * bpr < ((int64_t)1 << 31), "transient too large");
  const int64_t rows = pad_zeroed ? D : 2 * D;
  hipLaunchKernelGGL(fk_prepare_kernel, dim3((unsigned)(rows * bpr)), dim3(kFkBlock), 0,
                     (hipStream_t)stream, transient, padded, (int)D, (int)HW, (int)bpr);
  return check_launch("fk_prepare_kernel");
}

extern "C" int totri_fk_stolt(const void* spectrum, int64_t C, int64_t D2, int64_t H,
                              int64_t W, double depth, double height, double width,
                              int layout, int out_half_zeroed, void* out, void* stream) {
  int rc = check_fk_dims(C, D2, H, W);
  if (rc) return rc;
  TOTRI_CHECK_ARG(D2 % 2 == 0, "the depth size must be even (got %lld)", (long long)D2);
  TOTRI_CHECK_ARG(layout == TOTRI_FK_SHIFTED || layout == TOTRI_FK_UNSHIFTED ||
                      layout == TOTRI_FK_UNSHIFTED_ONESIDED,
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
  g.fold = layout != TOTRI_FK_SHIFTED;
  g.onesided = layout == TOTRI_FK_UNSHIFTED_ONESIDED;
  g.k0 = out_half_zeroed ? (int)(D2 / 2) : 0;
  const int64_t n_rows = (D2 - g.k0) * H;
  const int64_t rows_per_block = kFkBlock / 64;
  hipLaunchKernelGGL(fk_stolt_kernel,
                     dim3((unsigned)((n_rows + rows_per_block - 1) / rows_per_block),
                          (unsigned)C),
                     dim3(kFkBlock), 0, (hipStream_t)stream, (const float2*)spectrum,
                     (float2*)out, g, (int)n_rows);
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

extern "C" size_t totri_fk_migration_workspace(int64_t D, int64_t H, int64_t W) {
  const FkNativePlan p = plan_fk_native(D, H, W);
  return p.ok ? p.bytes : 0;
}

extern "C" int totri_fk_migration(const float* transient, int64_t D, int64_t H, int64_t W,
                                  double depth, double height, double width, float* out,
                                  void* workspace, size_t workspace_bytes, void* stream) {
  const FkNativePlan p = plan_fk_native(D, H, W);
  TOTRI_CHECK_ARG(p.ok,
                  "sizes D=%lld H=%lld W=%lld: the native path needs powers of two with "
                  "H*W <= %d and 2D <= %d",
                  (long long)D, (long long)H, (long long)W, kFkMaxPlane, kFkMaxDepth2);
  TOTRI_CHECK_ARG(isfinite(depth) && isfinite(height) && isfinite(width) && height > 0 &&
                      width > 0,
                  "depth must be finite and height, width finite and positive");
  TOTRI_CHECK_ARG(transient != nullptr && out != nullptr && workspace != nullptr,
                  "null pointer");
  if (workspace_bytes < p.bytes)
    return fail(TOTRI_EWORKSPACE, "workspace of %zu bytes, %zu needed", workspace_bytes,
                p.bytes);
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  float2* tw = (float2*)(ws + p.tw_off);
  float2* A = (float2*)(ws + p.a_off);
  float2* B = (float2*)(ws + p.b_off);
  const int Hi = (int)H, Wi = (int)W, Di = (int)D, Z = 2 * Di;
  const int nW = Wi / 2, nH = Hi / 2, nZ = Z / 2;
  float2* twW = tw;
  float2* twH = tw + nW;
  float2* twZ = tw + nW + nH;
  int rc;
  if ((rc = fk_allow_lds(fk_slice_fwd_kernel, p.slice_lds, "fk_slice_fwd_kernel"))) return rc;
  if ((rc = fk_allow_lds(fk_slice_inv_kernel, p.slice_lds, "fk_slice_inv_kernel"))) return rc;
  if ((rc = fk_allow_lds(fk_depth_fwd_kernel, p.depth_lds, "fk_depth_fwd_kernel"))) return rc;
  if ((rc = fk_allow_lds(fk_depth_inv_kernel, p.depth_lds, "fk_depth_inv_kernel"))) return rc;

  hipLaunchKernelGGL(fk_twiddle_kernel, dim3((unsigned)((nW + nH + nZ + 255) / 256)),
                     dim3(256), 0, st, tw, nW, nH, nZ, Wi, Hi, Z);
  if ((rc = check_launch("fk_twiddle_kernel"))) return rc;
  hipLaunchKernelGGL(fk_slice_fwd_kernel, dim3((unsigned)Di), dim3(kFkSliceThreads),
                     p.slice_lds, st, transient, A, twW, twH, Di, p.logH, p.logW);
  if ((rc = check_launch("fk_slice_fwd_kernel"))) return rc;
  const int ncol_half = Hi * (Wi / 2 + 1), ncol = Hi * Wi;
  hipLaunchKernelGGL(fk_depth_fwd_kernel, dim3((unsigned)((ncol_half + p.cb_fwd - 1) / p.cb_fwd)),
                     dim3(kFkDepthThreads), p.depth_lds, st, A, B, twZ, p.logZ, ncol_half,
                     p.cb_fwd);
  if ((rc = check_launch("fk_depth_fwd_kernel"))) return rc;
  // Stolt: one-sided spectrum B -> nonzero half of G (rows 0 .. D-1 unshifted) in A
  const double a = ((double)W * depth) / (2.0 * (double)D * width);
  const double b = ((double)H * depth) / (2.0 * (double)D * height);
  FkStoltGeom g;
  g.C = 1;
  g.D2 = Z;
  g.H = Hi;
  g.W = Wi;
  g.ax = (float)(a * a);
  g.ay = (float)(b * b);
  g.fold = 1;
  g.onesided = 1;
  g.k0 = Di;
  const int n_rows = Di * Hi;
  hipLaunchKernelGGL(fk_stolt_kernel, dim3((unsigned)((n_rows + 3) / 4), 1u), dim3(kFkBlock),
                     0, st, (const float2*)B, A, g, n_rows);
  if ((rc = check_launch("fk_stolt_kernel"))) return rc;
  hipLaunchKernelGGL(fk_depth_inv_kernel, dim3((unsigned)((ncol + p.cb_inv - 1) / p.cb_inv)),
                     dim3(kFkDepthThreads), p.depth_lds, st, (const float2*)A, B, twZ, p.logZ,
                     ncol, p.cb_inv);
  if ((rc = check_launch("fk_depth_inv_kernel"))) return rc;
  const float scale = ldexpf(1.f, -(p.logZ + p.logH + p.logW));
  hipLaunchKernelGGL(fk_slice_inv_kernel, dim3((unsigned)Di), dim3(kFkSliceThreads),
                     p.slice_lds, st, (const float2*)B, out, twW, twH, p.logH, p.logW, scale);
  return check_launch("fk_slice_inv_kernel");
}
