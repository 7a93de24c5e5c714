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
 * A1 in the reference's own window form, for callers of apply_transposed
 * that need the reference's result (src/render/volume_cuda.cu:95-160, the
 * kernel totri_cpp.render.volume_confocal_grad_volume runs there).  Each voxel
 * sums, over scan points, the bins whose radius lies between its
 * neighbours' centres, weighted (1 - alpha) from its own centre's bin on and
 * alpha below it.  It differs from totri_volume_confocal_grad_volume on the
 * top slice z = Z-1 (samples the render never reads) and by rounding at the
 * window edges.  Same arguments; no workspace.
 */
int totri_volume_confocal_grad_volume_window(
    float* volume, int64_t X, int64_t Y, int64_t Z, const float* measurement_points,
    int64_t N, const void* transient, int64_t T, int transient_dtype,
    const float volume_origin[3], const float volume_size[3], float bin_width,
    float bin_offset, void* stream);

/*
 * Double precision: R1, A1 and A2 for float64 tensors, as the reference
 * dispatches float and double (AT_DISPATCH_FLOATING_TYPES,
 * src/render/volume_cuda.cu:85,175, src/reconstruct/backprojection_cuda.cu:215).
 * They follow the reference's double semantics: float geometry, the bin
 * radius t * w + off in float, floorf/ceilf narrowing to float, and for A2
 * float points and distances with double bin arithmetic and sums.
 * window_form = 1 selects the reference's window form of A1, 0 the exact
 * transpose of totri_volume_confocal_forward_f64.  No workspace.
 */
int totri_volume_confocal_forward_f64(
    const double* volume, int64_t X, int64_t Y, int64_t Z, const double* measurement_points,
    int64_t N, double* transient, int64_t T, const float volume_origin[3],
    const float volume_size[3], float bin_width, float bin_offset, void* stream);
int totri_volume_confocal_grad_volume_f64(
    double* volume, int64_t X, int64_t Y, int64_t Z, const double* measurement_points,
    int64_t N, const double* transient, int64_t T, const float volume_origin[3],
    const float volume_size[3], float bin_width, float bin_offset, int window_form,
    void* stream);
int totri_backprojection_confocal_velten_forward_f64(
    const double* transient, int64_t B, int64_t T, int64_t S, const double* scan_points,
    const double* scan_origin, int64_t So, double* volume, int64_t D, int64_t H, int64_t W,
    const float volume_start[3], const float volume_end[3], float bin_width, float bin_offset,
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
 * A2 -- confocal Velten backprojection.
 * Replaces totri_cpp.reconstruct.backprojection_confocal_velten_forward
 * (src/reconstruct/backprojection.h:21-28, backprojection.cpp:22-34,
 *  kernel src/reconstruct/backprojection_cuda.cu:152-226).
 *   transient    [B][T][S]      FP32
 *   scan_points  [B][S][3]      wall points (also the laser points)
 *   scan_origin  [B][So][3]     So = 0 ignores the origin-to-wall distance
 *   volume       [B][D][H][W]   FP32 output, written; inclusive grid from
 *                               volume_start to volume_end
 *   bin t collects the round trip 2|o - p| + 2|p - q| in [t w + off -/+ w/2)
 */
size_t totri_backprojection_confocal_velten_workspace(int64_t B, int64_t T, int64_t S);
int totri_backprojection_confocal_velten_forward(
    const float* transient, int64_t B, int64_t T, int64_t S, const float* scan_points,
    const float* scan_origin, int64_t So, float* volume, int64_t D, int64_t H, int64_t W,
    const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream);

/*
 * A2 without allocation -- no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same computation and
 * kernels as totri_backprojection_confocal_velten_forward, with the scratch
 * (one row per scan point: the transposed transient, the point's bin
 * thresholds and constants; totri_backprojection_confocal_velten_workspace()
 * bytes, about 8*B*S*(T+4)) in a file-scope device array of
 * totri_backprojection_confocal_velten_noalloc_max_bytes() bytes (64 MiB)
 * that each GPU the library runs on holds until the process ends.  Larger
 * transients are rejected with TOTRI_EINVAL.  The call allocates nothing; the
 * caller passes inputs and output as for the entry point above.  All calls on
 * one GPU share that GPU's array: calls that overlap on one GPU (different
 * streams, or host threads) race and must be serialised by the caller.
 */
size_t totri_backprojection_confocal_velten_noalloc_max_bytes(void);
int totri_backprojection_confocal_velten_forward_noalloc(
    const float* transient, int64_t B, int64_t T, int64_t S, const float* scan_points,
    const float* scan_origin, int64_t So, float* volume, int64_t D, int64_t H, int64_t W,
    const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* stream);

/*
 * A3 -- exhaustive Velten backprojection.
 * Replaces totri_cpp.reconstruct.backprojection_exhaustive_velten_forward
 * (src/reconstruct/backprojection.h:30-39, backprojection.cpp:36-47,
 *  kernel src/reconstruct/backprojection_cuda.cu:228-322).
 *   transient [B][T][S][L]; scan_points [B][S][3]; laser_points [B][L][3];
 *   scan_origin [B][So][3], laser_origin [B][Lo][3] (So in {0,S}, Lo in {0,L});
 *   volume [B][D][H][W] written.
 */
size_t totri_backprojection_exhaustive_velten_workspace(int64_t B, int64_t T, int64_t S,
                                                        int64_t L);
int totri_backprojection_exhaustive_velten_forward(
    const float* transient, int64_t B, int64_t T, int64_t S, int64_t L,
    const float* scan_points, const float* laser_points, const float* scan_origin,
    int64_t So, const float* laser_origin, int64_t Lo, float* volume, int64_t D, int64_t H,
    int64_t W, const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream);

/*
 * A4 -- Arellano ellipsoid backprojection.
 * Replaces totri_cpp.reconstruct.backprojection_exhaustive_arellano_forward
 * (src/reconstruct/backprojection.h:10-19, backprojection.cpp:7-20,
 *  kernel src/reconstruct/backprojection_cuda.cu:10-150).
 *   transient [B][T][S][L]; scan_points [B][3][S]; laser_points [B][3][L];
 *   scan_origin [B][3][So], laser_origin [B][3][Lo]; volume [B][D][H][W]
 *   written (the reference accumulates into a zeroed volume).  As in the
 *   reference, the interpolation weight between neighbouring voxels is 0.
 */
size_t totri_backprojection_exhaustive_arellano_workspace(int64_t B, int64_t T,
                                                         int64_t S, int64_t L, int64_t D,
                                                         int64_t H, int64_t W);
int totri_backprojection_exhaustive_arellano_forward(
    const float* transient, int64_t B, int64_t T, int64_t S, int64_t L,
    const float* scan_points, const float* laser_points, const float* scan_origin,
    int64_t So, const float* laser_origin, int64_t Lo, float* volume, int64_t D, int64_t H,
    int64_t W, const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream);

/*
 * R2 -- exhaustive (non-confocal) volume render.
 * Replaces totri_cpp.render.volume_render_exhaustive_forward
 * (src/render/volume.h:35-44, src/render/volume.cpp:83-96,
 *  kernel src/render/volume_cuda.cu:297-412).
 *   volume        [B][D][H][W] FP32 occupancy
 *   scan_points   [B][3][S], laser_points [B][3][L]
 *   transient     [B][T][S][L] FP32 output, written (bin t holds path
 *                 lengths |p_s - q| + |p_l - q| in [t w + off, (t+1) w + off))
 * The reference also takes scan/laser origins and ignores them
 * (volume_cuda.cu:312-313); they are not part of this entry point.
 */
size_t totri_volume_render_exhaustive_forward_workspace(int64_t B, int64_t D, int64_t H,
                                                        int64_t W);
int totri_volume_render_exhaustive_forward(
    const float* volume, int64_t B, int64_t D, int64_t H, int64_t W, const float* scan_points,
    int64_t S, const float* laser_points, int64_t L, float* transient, int64_t T,
    const float volume_start[3], const float volume_end[3], float bin_width,
    float bin_offset, void* workspace, size_t workspace_bytes, void* stream);

/*
 * A5 -- gradient of R2 with respect to the volume (R2's exact transpose).
 * Replaces totri_cpp.render.volume_render_exhaustive_grad_volume
 * (src/render/volume.h:46-55, src/render/volume.cpp:98-111), whose kernel in
 * the reference is an empty stub (src/render/volume_cuda.cu:414-431).
 *   transient_grad [B][T][S][L]; volume_grad [B][D][H][W] written.
 */
size_t totri_volume_render_exhaustive_grad_volume_workspace(int64_t B, int64_t T, int64_t S,
                                                            int64_t L);
int totri_volume_render_exhaustive_grad_volume(
    const float* transient_grad, int64_t B, int64_t T, int64_t S, int64_t L,
    const float* scan_points, const float* laser_points, float* volume_grad, int64_t D,
    int64_t H, int64_t W, const float volume_start[3], const float volume_end[3],
    float bin_width, float bin_offset, void* workspace, size_t workspace_bytes,
    void* stream);

/*
 * S1 -- Gaussian RBF to volume splat.
 * Replaces totri_cpp.util.rbf2volume_gaussian_forward
 * (src/util/rbf2volume.h:10-14, src/util/rbf2volume.cpp:7-15,
 *  kernel src/util/rbf2volume_cuda.cu:49-120).
 *   rbf     [B][N][4+A] (x, y, z, sigma, A attributes), 0 <= A <= 8
 *   volume  [B][1+A][D][H][W] output, written: channel 0 the summed
 *           Gaussians, channels 1..A the weighted attribute means
 *   truncation  > 0 skips RBFs farther than truncation * sigma from a voxel;
 *               <= 0 sums all RBFs (the reference's behaviour)
 *   workspace   truncated splats list each brick's RBFs first; the size
 *               comes from the query (0 without truncation)
 */
size_t totri_rbf2volume_gaussian_forward_workspace(int64_t B, int64_t N, int64_t A,
                                                   int64_t D, int64_t H, int64_t W,
                                                   float truncation);
int totri_rbf2volume_gaussian_forward(const float* rbf, int64_t B, int64_t N, int64_t A,
                                      float* volume, int64_t D, int64_t H, int64_t W,
                                      const float volume_start[3], const float volume_end[3],
                                      float truncation, void* workspace,
                                      size_t workspace_bytes, void* stream);

/*
 * S2 -- gradient of the splat with respect to the RBF parameters.
 * Replaces totri_cpp.util.rbf2volume_gaussian_grad
 * (src/util/rbf2volume.h:16-22, src/util/rbf2volume.cpp:17-27,
 *  kernel src/util/rbf2volume_cuda.cu:123-232).
 *   volume, volume_grad [B][1+A][D][H][W] (forward output and its gradient)
 *   rbf_grad            [B][N][4+A] output, written
 */
int totri_rbf2volume_gaussian_grad(const float* rbf, int64_t B, int64_t N, int64_t A,
                                   const float* volume, const float* volume_grad,
                                   float* rbf_grad, int64_t D, int64_t H, int64_t W,
                                   const float volume_start[3], const float volume_end[3],
                                   float truncation, void* stream);

/*
 * Batched confocal step (serving).  No counterpart in the reference, whose
 * volume_confocal_forward / volume_confocal_grad_volume each run one request
 * (src/render/volume.h:10-24); this runs up to TOTRI_BATCH_MAX_JOBS of them,
 * renders (R1) and backprojections (A1) of different sizes, in one step of
 * persistent kernels that take work items from a queue.
 *
 * Each job is computed with the same plan and the same per-item code as the
 * single-request entry points, so its result is bit-identical to
 * totri_volume_confocal_forward / _grad_volume (FP32 transient) on the same
 * inputs, whatever else shares the step.
 *
 *   totri_confocal_job_workspace(job)   device scratch bytes the job needs
 *   totri_confocal_batch_table_bytes()  size of one step table
 *   totri_confocal_batch_prepare()      host only: validates the jobs and
 *                                       writes a step table into host memory
 *   totri_confocal_batch_run()          copies the table (host, should be
 *                                       pinned) to device_table and enqueues
 *                                       the step on `stream`
 *
 * batch_run's launches depend only on its arguments, not on the table
 * contents, so one call can be captured into a HIP graph and replayed after
 * each batch_prepare into the same host table.  The host table must not be
 * rewritten until the previous step on it has finished its copy (wait for the
 * stream).  Job buffers must not overlap each other or the table.
 *
 * Slices: a job with slices = S > 1 runs over several steps, slices 0 .. S-1
 * in that order, a step running slice_run consecutive slices from `slice`
 * (later steps must not start before earlier ones finish, e.g. the same
 * stream), with the same description and untouched workspace between them.
 * Each slice is a consecutive 1/S of the job's work items; the mirror
 * detection and tables run with slice 0 and the output is complete after
 * slice S-1, bit-identical to the unsliced job.  So a long job can share its
 * steps with requests that arrive while it runs.
 */
#define TOTRI_JOB_RENDER 0      /* R1: volume (input) -> transient (output)      */
#define TOTRI_JOB_BACKPROJECT 1 /* A1: transient (input) -> volume (output)      */
#define TOTRI_BATCH_MAX_JOBS 64

typedef struct totri_confocal_job {
  int kind;                         /* TOTRI_JOB_RENDER or TOTRI_JOB_BACKPROJECT */
  int slice;                        /* this step's slice, 0 <= slice < slices */
  float* volume;                    /* [X][Y][Z] FP32 */
  const float* measurement_points;  /* [3][N] FP32 */
  float* transient;                 /* [N][T] FP32 */
  int64_t X, Y, Z, N, T;
  float volume_origin[3];
  float volume_size[3];
  float bin_width, bin_offset;
  int slices;                       /* 0 or 1: the whole job in this step */
  int slice_run;                    /* slices this step runs, from `slice`: 0 or 1
                                       for one, up to slices - slice */
  void* workspace;                  /* >= totri_confocal_job_workspace(job) */
  size_t workspace_bytes;
} totri_confocal_job;

size_t totri_confocal_job_workspace(const totri_confocal_job* job);
size_t totri_confocal_batch_table_bytes(void);
int totri_confocal_batch_prepare(const totri_confocal_job* jobs, int n_jobs,
                                 void* host_table, size_t table_bytes);
int totri_confocal_batch_run(const void* host_table, void* device_table,
                             size_t table_bytes, void* stream);

/*
 * f-k migration.  The reference's totri.reconstruct.FkMigration is pure
 * PyTorch (totri/reconstruct/fk_migration.py:5-120); these three entry points
 * are its non-FFT steps, and the host layer runs the two 3-D FFTs (hipFFT)
 * between them.  Complex buffers are interleaved complex64 (re, im).
 *
 * totri_fk_prepare replaces fk_migration.py:108-112:
 *   transient [D][H][W] FP32;  padded [2D][H][W] FP32 (the reference's
 *   complex input has a zero imaginary part), written:
 *   sqrt(transient[d]) * linspace(0, 1, D)[d] for d < D, 0 for d >= D.
 *   pad_zeroed  nonzero: rows d >= D already hold zeros and are not written
 */
int totri_fk_prepare(const float* transient, int64_t D, int64_t H, int64_t W,
                     float* padded, int pad_zeroed, void* stream);

/*
 * totri_fk_stolt replaces StoltInterpolation (fk_migration.py:21-85) and, with
 * layout TOTRI_FK_UNSHIFTED, the fftshift / ifftshift around it (:114-118).
 *   spectrum, out  [C][D2][H][W] complex, D2 even; they must not overlap
 *   depth, height, width  StoltInterpolation's constructor arguments
 *                         (height, width > 0)
 *   layout  TOTRI_FK_SHIFTED: spectrum and out in fftshift order (what
 *           StoltInterpolation.forward takes and returns);
 *           TOTRI_FK_UNSHIFTED: spectrum as fftn returns it and out in the
 *           order ifftn takes;
 *           TOTRI_FK_UNSHIFTED_ONESIDED: as TOTRI_FK_UNSHIFTED, but the
 *           spectrum is [C][D2][H][W/2+1], the rfftn of a real input (the
 *           other half follows from F[-k] = conj F[k])
 *   out_half_zeroed  nonzero: out already holds zeros where the weights are
 *           zero (depth index < D2/2 in fftshift order, >= D2/2 unshifted),
 *           and that half is not written
 */
#define TOTRI_FK_SHIFTED 0
#define TOTRI_FK_UNSHIFTED 1
#define TOTRI_FK_UNSHIFTED_ONESIDED 2
int totri_fk_stolt(const void* spectrum, int64_t C, int64_t D2, int64_t H, int64_t W,
                   double depth, double height, double width, int layout,
                   int out_half_zeroed, void* out, void* stream);

/*
 * totri_fk_magnitude replaces fk_migration.py:119:
 *   volume [C][2D][H][W] complex;  out [C][D][H][W] FP32, written:
 *   |volume[c][d][h][w]| for d < D.
 */
int totri_fk_magnitude(const void* volume, int64_t C, int64_t D, int64_t H, int64_t W,
                       float* out, void* stream);

/*
 * totri_fk_migration replaces FkMigration.forward (fk_migration.py:93-120)
 * for one transient, FFTs included, when D, H and W are powers of two with
 * H * W <= 16384 and 2D <= 4096:
 *   transient [D][H][W] FP32;  out [D][H][W] FP32, written (the reference's
 *   volume[0] for a batch of one)
 *   depth, height, width  as totri_fk_stolt
 * totri_fk_migration_workspace returns the scratch bytes it needs, or 0 for
 * sizes it does not take (the host layer then runs the three kernels above
 * around hipFFT).  The workspace needs no initial contents.
 */
size_t totri_fk_migration_workspace(int64_t D, int64_t H, int64_t W);
int totri_fk_migration(const float* transient, int64_t D, int64_t H, int64_t W,
                       double depth, double height, double width, float* out,
                       void* workspace, size_t workspace_bytes, void* stream);

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

/*
 * Test hook with no counterpart in the reference: runs the scan-mirror
 * detection that selects the mirror-symmetric R1 / A1 kernels and writes
 * four ints to info_out (device memory):
 *   [0] 1 if all N points form quads (wx, wy), (-wx, wy), (wx, -wy), (-wx, -wy)
 *       with equal wz -- R1 shares decisions when the voxel grid is also
 *       antisymmetric in x and y
 *   [1] number of quads
 *   [2] 1 if all points form x-mirror pairs -- A1 shares decisions when the
 *       voxel grid is antisymmetric in x
 *   [3] number of pairs
 * workspace: at least totri_debug_mirror_workspace(N) bytes of device memory.
 */
size_t totri_debug_mirror_workspace(int64_t N);
int totri_debug_mirror_info(const float* measurement_points, int64_t N,
                            void* workspace, size_t workspace_bytes,
                            int32_t* info_out, void* stream);

/*
 * Testing aid with no counterpart in the reference: the A1 walks that the
 * planner of totri_volume_confocal_grad_volume (and of the batched step)
 * allows for a geometry, computed on the host without device work.  Bits:
 *   1  x-mirror paths (X even, N even, grid antisymmetric in x)
 *   2  xy-mirror row walk (also antisymmetric in y, rows monotone in the bin)
 *   4  general row walk (bin_width > 0, bin_offset >= 0, z size > 0 and
 *      T < 2^24 - 1: the walks count bins in a float)
 *   8  x-pair row walk (bits 1 and 4)
 * Returns a negative status code for invalid sizes.
 */
int totri_debug_a1_walks(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                         const float volume_origin[3], const float volume_size[3],
                         float bin_width, float bin_offset);

/*
 * Testing aid with no counterpart in the reference: the mirror R1 kernels
 * that totri_volume_confocal_forward (and the batched step) launch for a
 * geometry, computed on the host without device work; the device's mirror
 * detection then picks among them.  Bits:
 *   1  xy-mirror R1 (X, Y even, N a multiple of 4, grid antisymmetric in x
 *      and y)
 *   2  x-pair R1 (X even, N even, grid antisymmetric in x)
 * Returns a negative status code for invalid sizes.
 */
int totri_debug_r1_paths(int64_t X, int64_t Y, int64_t Z, int64_t N, int64_t T,
                         const float volume_origin[3], const float volume_size[3]);

/*
 * Testing aid with no counterpart in the reference: counts the floats x
 * whose bit patterns lie in [lo_bits, hi_bits] for which the library's
 * short correctly rounded square root (used by the exhaustive backprojection,
 * render and adjoint where every radicand lies in [2^-64, 2^64]) differs
 * from the compiler's sqrtf, into *mismatches (device memory, 8 bytes,
 * overwritten).  Asynchronous on `stream`.
 */
int totri_debug_sqrt_mid_mismatches(uint32_t lo_bits, uint32_t hi_bits,
                                    unsigned long long* mismatches, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TOTRI_HIP_H */
