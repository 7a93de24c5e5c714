// Shared device helpers and host error plumbing for the totri HIP library.
//
// Build flags (see totrilib_amd/build.py): -O3 --offload-arch=gfx950
// -ffp-contract=off, packed FP32 disabled.  Contraction is off so that every
// `a * b + c` in the decision math rounds twice, as the reference source text
// is written; fused multiply-adds are requested explicitly with fmaf() only
// in weight arithmetic, never in a decision.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/totri_hip.h"

namespace totri_hip {

// ---------------------------------------------------------------------------
// Host-side error state (one message per host thread).
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
int check_launch(const char* what);

#define TOTRI_CHECK_ARG(cond, ...)                                 \
  do {                                                             \
    if (!(cond)) return ::totri_hip::fail(TOTRI_EINVAL, __VA_ARGS__); \
  } while (0)

#define TOTRI_CHECK_HIP(expr)                                                   \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      return ::totri_hip::fail(TOTRI_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// Number of compute units of the current device (cached per device).
int device_cu_count();

// out[b][c][r] = in[b][r][c] for B arrays of R x C floats, on stream st.
int transpose_batched(const float* in, float* out, int B, int R, int C, hipStream_t st);
// out[b][c][ooff + r] = in[b][r][c], out rows of stride ldo.
int transpose_strided(const float* in, float* out, int B, int R, int C, int ldo, int ooff,
                      hipStream_t st);

// ---------------------------------------------------------------------------
// Device math
// ---------------------------------------------------------------------------

// IEEE correctly rounded square root (the compiler expands sqrtf to the
// v_sqrt_f32 + residual fix-up sequence under its default
// -fhip-fp32-correctly-rounded-divide-sqrt).
__device__ __forceinline__ float sqrt_rn(float x) { return sqrtf(x); }

// Correctly rounded square root for x in [2^-64, 2^64] (sqrt_mid_range):
// the hardware root r, within 1 ulp, then the same neighbour tests as the
// compiler's sqrtf expansion (with r- and r+ the floats 1 ulp below and
// above r: x - r- * r <= 0 picks r-, then x - r+ * r > 0 picks r+, each
// product-difference one FMA), without the scaling of tiny inputs and the
// special-value handling that inputs in this range do not need: 9
// instructions instead of 16.  Equal to sqrt_rn on every float of the range
// (totri_debug_sqrt_mid_mismatches checks all 1.07e9 of them on the GPU).
__device__ __forceinline__ bool sqrt_mid_range(float x) {
  return x >= 0x1p-64f && x <= 0x1p64f;
}

__device__ __forceinline__ float sqrt_rn_mid(float x) {
  const float r = __builtin_amdgcn_sqrtf(x);
  const float rm = __uint_as_float(__float_as_uint(r) - 1u);
  const float rp = __uint_as_float(__float_as_uint(r) + 1u);
  float y = __builtin_fmaf(-rm, r, x) <= 0.f ? rm : r;
  y = __builtin_fmaf(-rp, r, x) > 0.f ? rp : y;
  return y;
}

// Hardware square root, at most 1 ulp from the correctly rounded result.
// Used only on the fast path, behind a guard that falls back to sqrt_rn
// wherever the 1-ulp difference could change a discrete decision.
__device__ __forceinline__ float sqrt_fast(float x) {
  return __builtin_amdgcn_sqrtf(x);
}

// tc = t * w + off, rounded after the product and after the sum
// (src/render/volume_cuda.cu:23).
__device__ __forceinline__ float bin_distance(int t, float w, float off) {
  float a = (float)t * w;
  return a + off;
}

// Cell-centred coordinate origin + (i + 0.5) * size / n
// (src/render/volume_cuda.cu:35,38,107-109): product, then quotient, then sum.
__device__ __forceinline__ float cell_centre(int i, float o, float size, int n) {
  float a = ((float)i + 0.5f) * size;
  a = a / (float)n;
  return o + a;
}

// Inclusive-grid coordinate start + (i / (n-1)) * (end - start)
// (src/reconstruct/backprojection_cuda.cu:29-32,167-170).
__device__ __forceinline__ float inclusive_coord(int i, int n, float a, float b) {
  float t = (float)i / (float)(n - 1);
  return a + t * (b - a);
}

// (dx*dx + dy*dy) + dz*dz for d = a - b, the radicand of dist3.
__device__ __forceinline__ float sqdist3(float ax, float ay, float az, float bx, float by,
                                         float bz) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  float s = dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return s;
}

// Squared distance from p to the nearest point of the box [lo, hi] (0 inside).
__device__ __forceinline__ float box_sqdist(float px, float py, float pz, float lx, float ly,
                                            float lz, float hx, float hy, float hz) {
  const float dx = fmaxf(fmaxf(lx - px, px - hx), 0.f);
  const float dy = fmaxf(fmaxf(ly - py, py - hy), 0.f);
  const float dz = fmaxf(fmaxf(lz - pz, pz - hz), 0.f);
  return dx * dx + dy * dy + dz * dz;
}

// norm3df restated as sqrt((dx*dx + dy*dy) + dz*dz), correctly rounded.
__device__ __forceinline__ float dist3(float ax, float ay, float az, float bx,
                                      float by, float bz) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  float s = dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return sqrt_rn(s);
}

// Wave-wide integer reductions over all 64 lanes.
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// floor(x) converted to int in one instruction (v_cvt_flr_i32_f32).
__device__ __forceinline__ int floor_to_int(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

__device__ __forceinline__ float uniform_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ int uniform_i32(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// Transient element stores for the supported storage types.
__device__ __forceinline__ void store_elem(float* p, float v) { *p = v; }
__device__ __forceinline__ void store_elem(__hip_bfloat16* p, float v) {
  *p = __float2bfloat16(v);
}
__device__ __forceinline__ void store_elem(__half* p, float v) { *p = __float2half(v); }
__device__ __forceinline__ float load_elem(const float* p) { return *p; }
__device__ __forceinline__ float load_elem(const __hip_bfloat16* p) {
  return __bfloat162float(*p);
}
__device__ __forceinline__ float load_elem(const __half* p) { return __half2float(*p); }

}  // namespace totri_hip
