/*
 * totri_hip.h -- C ABI of the MI355X (gfx950) volumetric transient library.
 *
 * Every entry point takes plain device pointers, sizes and scalars, never
 * allocates its outputs (the caller owns every buffer, as in the reference,
 * where Python allocates and C++ writes in place), launches on the given
 * HIP stream (`stream` is a hipStream_t; NULL means the default stream) and
 * returns a status code instead of throwing:
 *
 *   TOTRI_OK        success (kernels are enqueued, not finished)
 *   TOTRI_EINVAL    invalid argument: shapes, sizes or geometry
 *                   (the reference throws std::range_error -> ValueError)
 *   TOTRI_EHIP      HIP runtime or launch error (-> RuntimeError)
 *   TOTRI_EWORKSPACE  the workspace is smaller than *_workspace() reports
 *
 * totri_last_error() returns the message of the last failure on the calling
 * host thread.
 *
 * Layouts are dense row-major (C-contiguous) with 32-bit element counts per
 * dimension, as the reference's packed_accessor32.  All geometry is FP32.
 * Outputs are written, not accumulated, unless an entry says otherwise.
 *
 * Scratch memory: entry points that need scratch take a device `workspace`
 * of at least the size returned by the matching *_workspace() query; the
 * library itself never calls hipMalloc.
 */
#ifndef TOTRI_HIP_H
#define TOTRI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOTRI_OK 0
#define TOTRI_EINVAL 1
#define TOTRI_EHIP 2
#define TOTRI_EWORKSPACE 3

/* Storage types of transient buffers (volumes are always FP32). */
#define TOTRI_F32 0
#define TOTRI_BF16 1
#define TOTRI_F16 2

/* Message of the last failure on this host thread ("" if none). */
const char* totri_last_error(void);

/* ABI version (major * 100 + minor). */
int totri_abi_version(void);

/*
 * R1 -- confocal volume render.
 * Replaces totri_cpp.render.volume_confocal_forward
 * (src/render/volume.h:10-16, src/render/volume.cpp:7-23,
 *  kernel src/render/volume_cuda.cu:11-93).
 *   volume              [X][Y][Z] FP32 albedo (z fastest)
 *   measurement_points  [3][N]    FP32 wall points (row 0 = x, 1 = y, 2 = z)
 *   transient           [N][T]    output, written; storage type transient_dtype
 *   volume_origin/size  box corner and extent (metres)
 *   bin_width/offset    one-way radius of bin t is t * bin_width + bin_offset
 */
size_t totri_volume_confocal_forward_workspace(int64_t X, int64_t Y, int64_t Z,
                                               int64_t N, int64_t T);
int totri_volume_confocal_forward(const float* volume, int64_t X, int64_t Y,
                                  int64_t Z, const float* measurement_points,
                                  int64_t N, void* transient, int64_t T,
                                  int transient_dtype,
                                  const float volume_origin[3],
                                  const float volume_size[3], float bin_width,
                                  float bin_offset, void* workspace,
                                  size_t workspace_bytes, void* stream);

/*
 * A1 -- albedo gradient / confocal backprojection, the exact transpose of R1.
 * Replaces totri_cpp.render.volume_confocal_grad_volume
 * (src/render/volume.h:18-24, src/render/volume.cpp:25-41,
 *  kernel src/render/volume_cuda.cu:95-186).
 *   volume     [X][Y][Z] FP32 output, written
 *   transient  [N][T]    input (upstream gradient), storage transient_dtype
 * The reference kernel evaluates a per-voxel bin window that differs from
 * R1's transpose on the top z slice; this entry point is the exact
 * transpose of totri_volume_confocal_forward.
 */
size_t totri_volume_confocal_grad_volume_workspace(int64_t X, int64_t Y,
                                                   int64_t Z, int64_t N,
                                                   int64_t T);
int totri_volume_confocal_grad_volume(float* volume, int64_t X, int64_t Y,
                                      int64_t Z, const float* measurement_points,
                                      int64_t N, const void* transient,
                                      int64_t T, int transient_dtype,
                                      const float volume_origin[3],
                                      const float volume_size[3],
                                      float bin_width, float bin_offset,
                                      void* workspace, size_t workspace_bytes,
                                      void* stream);

/*
 * A6 -- gradient of sum(G * R1(volume; w)) with respect to the wall points w.
 * Replaces totri_cpp.render.volume_confocal_grad_measurement_points
 * (src/render/volume.h:26-33, src/render/volume.cpp:43-81,
 *  kernel src/render/volume_cuda.cu:188-295).
 *   measurement_points_grad [3][N] FP32 output, written
 *   transient_grad          [N][T] upstream gradient G, storage transient_dtype
 * Sums are reduced in a fixed order (the reference uses float atomics).
 */
size_t totri_volume_confocal_grad_measurement_points_workspace(int64_t X, int64_t Y,
                                                               int64_t Z, int64_t N,
                                                               int64_t T);
int totri_volume_confocal_grad_measurement_points(
    float* measurement_points_grad, const float* volume, int64_t X, int64_t Y,
    int64_t Z, const float* measurement_points, int64_t N,
    const void* transient_grad, int64_t T, int transient_dtype,
    const float volume_origin[3], const float volume_size[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream);

/*
 * Test hook with no counterpart in the reference: evaluates the per-sample
 * decision function shared by the R1 and A1 kernels for n samples listed as
 * idx[n][4] = (scan point i, bin t < T, xi, yi) and writes
 *   zi_out[k]    floor(z_hat), or -1000000 when depth^2 <= 0
 *   alpha_out[k] z_hat - floor(z_hat)
 *   d2_out[k]    depth^2
 * so that bin decisions can be compared bit for bit with a CPU restatement.
 */
int totri_debug_confocal_samples(const float* measurement_points, int64_t N,
                                 const int32_t* idx, int64_t n, int64_t X,
                                 int64_t Y, int64_t Z, int64_t T,
                                 const float volume_origin[3],
                                 const float volume_size[3], float bin_width,
                                 float bin_offset, int32_t* zi_out,
                                 float* alpha_out, float* d2_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TOTRI_HIP_H */
