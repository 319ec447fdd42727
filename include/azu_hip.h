/*
 * azu_hip.h - C ABI of the MI355X (gfx950) KinectFusion frame loop.
 *
 * This is the drop-in boundary below the host classes include/tsdf/TSDFVolume.h and
 * include/tracking/ICPTracker.h. Plain pointers and sizes only; every call returns an
 * azu_status and never throws; azu_last_error() gives the message of the calling
 * thread's last failure. Pointers named d_* are device pointers (azu_malloc or any HIP
 * allocation on the current device); everything else is host memory. Calls on one
 * handle must be serialised by the caller; work is enqueued on the given stream
 * (NULL = the default stream) and is asynchronous unless the comment says it
 * synchronises.
 *
 * Conventions shared with the reference (iphands/azu):
 *   - poses are 4x4 column-major float[16] (Eigen's default storage), world-from-camera;
 *   - intrinsics K = {fx, fy, cx, cy} in pixels;
 *   - vectors (vertices, normals) are packed xyz float triples (Eigen::Vector3f / float3);
 *   - a volume index is z*res*res + y*res + x, voxel corner at origin + (x,y,z)*voxel_size;
 *   - the empty voxel is tsdf = 1, weight = 0.
 */
#ifndef AZU_HIP_H
#define AZU_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AZU_HIP_ABI_VERSION 1

typedef enum azu_status {
    AZU_OK = 0,
    AZU_ERR_ARG = 1,   /* invalid argument (null handle, bad size, unsupported format) */
    AZU_ERR_HIP = 2,   /* a HIP runtime call or kernel launch failed */
    AZU_ERR_ALLOC = 3, /* device or pinned allocation failed */
    AZU_ERR_STATE = 4  /* call not valid in the object's current state */
} azu_status;

/* Voxel storage forms. F32: float tsdf + float weight (8 B/voxel). F16_U16: binary16
 * tsdf + uint16 weight (4 B/voxel). The reference stores a 20 B AoS voxel with colour
 * (include/tsdf/TSDFVolume.h:55-62); colour is not stored here. */
typedef enum azu_format { AZU_FMT_F32 = 0, AZU_FMT_F16_U16 = 1 } azu_format;

/* Kernel arithmetic. EXACT: the FP32 parity kernels (IEEE division, no contraction,
 * the reference's incremental x-step), used for either storage form; FAST: hardware
 * reciprocals and FMA, about 20% faster integration and raycast, results within
 * rounding of EXACT but not bit-identical (so an ICP that is not converged within its
 * iteration budget can land elsewhere). AUTO is EXACT. */
typedef enum azu_mode { AZU_MODE_AUTO = 0, AZU_MODE_EXACT = 1, AZU_MODE_FAST = 2 } azu_mode;

typedef void* azu_stream; /* hipStream_t */

/* Mirrors kfusion::tsdf::TSDFParams (include/tsdf/TSDFVolume.h:39-53). */
typedef struct azu_tsdf_params {
    int resolution;
    float voxel_size;
    float truncation;
    float max_weight;
    float origin[3];
    float min_depth;
    float max_depth;
} azu_tsdf_params;

/* Mirrors kfusion::tracking::ICPParams (include/tracking/ICPTracker.h:19-25). */
typedef struct azu_icp_params {
    int max_iterations[3]; /* per pyramid level 0, 1, 2 */
    float dist_threshold;  /* m */
    float angle_threshold; /* degrees, sanitised to [0, 85] (NaN -> 30) */
    float min_depth;
    float max_depth;
} azu_icp_params;

/* Mirrors kfusion::tracking::ICPResult (include/tracking/ICPTracker.h:27-47). */
typedef struct azu_icp_result {
    float pose[16];
    float error;
    int inliers;
    int converged;
    int tracking_ok;
    float final_step;
    float information[36]; /* row-major, undamped J^T W J of the last solve */
    int valid_live_points;
    int valid_model_points;
    int projected_points;
    int dist_filtered;
    int angle_filtered;
    int iterations; /* accepted steps, one byte per level (level l at bit 8*l) */
} azu_icp_result;

/* Frame-level tracking grade of kfusion::tracking::TrackQuality (TrackingPolicy.h:36). */
typedef enum azu_track_quality { AZU_TRACK_GOOD = 0, AZU_TRACK_POOR = 1, AZU_TRACK_FAILED = 2 } azu_track_quality;

/* Mirrors kfusion::tracking::TrackingPolicy (include/tracking/TrackingPolicy.h:40-48). */
typedef struct azu_tracking_policy {
    int min_inliers;             /* below this: Failed */
    int min_good_inliers;        /* Good needs this much support */
    float min_fit_ratio;         /* inliers / valid_model_points */
    float max_rms_m;             /* sqrt(error) */
    float max_frame_translation; /* m between consecutive frames */
    float max_frame_rotation;    /* rad */
    int failures_before_lost;
} azu_tracking_policy;

/* One graded ICP result: what the pipeline does with the frame. */
typedef struct azu_track_decision {
    int quality;         /* azu_track_quality */
    int degenerate_dofs; /* directions keepObservableMotion dropped (0 when it is off) */
    int integrate;       /* 1 for Good: update the pose and integrate */
    float pose[16];      /* the pose to continue from: the graded pose, or prev_pose when Failed */
} azu_track_decision;

typedef struct azu_volume azu_volume;
typedef struct azu_tracker azu_tracker;

/* ---- library ---------------------------------------------------------------- */
const char* azu_last_error(void);
int azu_abi_version(void);
/* Current device id, its name (name_len bytes incl. NUL) and HBM size. */
int azu_device_info(int* device, char* name, size_t name_len, size_t* hbm_bytes);

/* ---- device memory and streams (thin wrappers, for callers without HIP headers) ---- */
int azu_malloc(void** d_ptr, size_t bytes);
int azu_free(void* d_ptr);
int azu_host_alloc(void** ptr, size_t bytes); /* pinned host memory */
int azu_host_free(void* ptr);
int azu_memcpy_h2d(void* d_dst, const void* src, size_t bytes, azu_stream s);
int azu_memcpy_d2h(void* dst, const void* d_src, size_t bytes, azu_stream s);
int azu_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, azu_stream s);
int azu_stream_create(azu_stream* s);
int azu_stream_destroy(azu_stream s);
int azu_stream_sync(azu_stream s); /* synchronises */

/* ---- TSDF volume (kfusion::tsdf::TSDFVolume, include/tsdf/TSDFVolume.h:69-190) ---- */
/* Allocates res^3 voxels in the given format and resets them to the empty state
 * (TSDFVolume::TSDFVolume, src/tsdf/TSDFVolume.cpp:96-102). Synchronises. */
int azu_vol_create(const azu_tsdf_params* p, int format, azu_volume** out);
int azu_vol_destroy(azu_volume* v);
/* TSDFVolume::reset (TSDFVolume.cpp:112-122). */
int azu_vol_reset(azu_volume* v, azu_stream s);
/* TSDFVolume::setParams (TSDFVolume.cpp:124-146): any field difference clears the
 * volume; a resolution change reallocates (the new storage first: if that fails, the
 * call returns AZU_ERR_ALLOC and the volume keeps its parameters and contents).
 * Synchronises. */
int azu_vol_set_params(azu_volume* v, const azu_tsdf_params* p);
int azu_vol_get_params(const azu_volume* v, azu_tsdf_params* p, int* format);
int azu_vol_set_mode(azu_volume* v, int mode);
/* Empty-space skipping in the raycast (on by default; results are the same either way).
 * Off, every sample is taken and no brick map is read or rebuilt: for callers that write
 * voxels through azu_vol_storage and do not want a map rebuild per raycast. No
 * counterpart in the reference. */
int azu_vol_set_skip(azu_volume* v, int on);
/* Diagnostic: whether the EXACT kernels divide by the voxel size / the truncation with
 * a reciprocal and two exact corrections (1) or with an IEEE division (0), and take the
 * camera depth's reciprocal by Newton-Raphson (1) or IEEE (0). Each fast form is chosen
 * only when an exhaustive device check (at create / setParams; once per device for the
 * reciprocal) found it equal to the IEEE result over the inputs that matter (every float
 * |x| < 2^32; every finite float); results are the same either way. Any pointer may be
 * NULL. No counterpart in the reference. */
int azu_vol_division_check(const azu_volume* v, int* voxel_size_fast, int* truncation_fast, int* reciprocal_fast);
/* Whole-volume copies to / from host float arrays of res^3 (voxelData, syncFromGPU /
 * syncToGPU, TSDFVolume.h:134, 161-162). They wait for all work queued on the device
 * (any stream) first, then synchronise. */
int azu_vol_download(azu_volume* v, float* tsdf, float* weight);
int azu_vol_upload(azu_volume* v, const float* tsdf, const float* weight);
/* voxelAt (TSDFVolume.h:117-118) without a whole-volume copy: count voxels from linear
 * index `first` (z*res*res + y*res + x) read to / written from host floats. F16_U16
 * volumes store the tsdf rounded to binary16 and need integer weights in [0, 65535].
 * A write is seen by every later call (the next raycast first rebuilds its empty-space
 * map from the voxels). Both wait for all work queued on the device, then synchronise. */
int azu_vol_read_voxels(azu_volume* v, size_t first, size_t count, float* tsdf, float* weight);
int azu_vol_write_voxels(azu_volume* v, size_t first, size_t count, const float* tsdf, const float* weight);
/* TSDFVolume::usageFraction's count (TSDFVolume.cpp:563-569): voxels with weight > 0.
 * Waits for the device, synchronises. */
int azu_vol_usage(azu_volume* v, unsigned long long* observed);
/* Raw device storage (format-dependent words, index as above). The raycast skips empty
 * space with a brick map that integrate / reset / upload keep; once the storage has been
 * handed out, writes through it are possible and every later raycast of this volume
 * first rebuilds the map from the voxels (a full read of the volume). */
int azu_vol_storage(azu_volume* v, void** d_ptr, size_t* bytes);
/* Voxels updated / newly observed by integrate since the last reset of the counters
 * (reset != 0 clears them after reading). Synchronises. */
int azu_vol_counters(azu_volume* v, unsigned long long* updated, unsigned long long* newly, int reset);
int azu_vol_integrated_frames(const azu_volume* v, int* frames);
/* (With s NULL, this and azu_mc_extract first wait for all work queued on the device.)
 * TSDFVolume::extractGlobalPointCloud (TSDFVolume.cpp:571-632; its GPU form
 * compactPointsKernel, TSDFVolume_cuda.cu:409-447): every voxel with weight > 1 and
 * |tsdf| < 0.2, as origin + (x, y, z) * voxel_size, xyz float triplets in linear voxel
 * order (x fastest). Writes min(count, capacity) points to d_points (device memory for
 * capacity * 3 floats; NULL with capacity 0 queries the count) and the full count to
 * *count. Colour is not stored, so no colours are produced. Synchronises. */
int azu_vol_point_cloud(azu_volume* v, float* d_points, unsigned long long capacity, unsigned long long* count,
                        azu_stream s);

/* TSDFVolume::integrate / integrateGPU (TSDFVolume.cpp:148-346, TSDFVolume.h:89-96,
 * 169-176): d_depth_m is W*H float metres on the device. */
int azu_integrate(azu_volume* v, const float* d_depth_m, int W, int H, const float K[4], const float pose[16],
                  float min_depth, float max_depth, azu_stream s);
/* No counterpart in the reference: azu_integrate for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and the same result bit for bit; it calls no hipMalloc /
 * hipFree: the per-frame depth-bounds scratch of the exact kernels is a library-wide
 * 4 MiB device array (one per GPU, held until the process ends) instead of the volume's
 * own buffer. Limit: frames whose bounds need more than 4 MiB (about 1.57 Mpx: 1440 x
 * 1080 fits, 1600 x 1200 does not) are rejected with AZU_ERR_ARG. Every call on one GPU shares
 * that array: calls that may overlap (other streams, other threads) race and must be
 * serialised by the caller. The scratch is fully rewritten before it is read, so
 * nothing needs clearing between calls. */
int azu_integrate_noalloc(azu_volume* v, const float* d_depth_m, int W, int H, const float K[4], const float pose[16],
                          float min_depth, float max_depth, azu_stream s);
/* Same, pose read from device memory (16 floats) when the kernel runs. */
int azu_integrate_dpose(azu_volume* v, const float* d_depth_m, int W, int H, const float K[4], const float* d_pose,
                        float min_depth, float max_depth, azu_stream s);

/* TSDFVolume::raycast / raycastGPU (TSDFVolume.cpp:348-476, TSDFVolume.h:99-104,
 * 177-181): W*H packed xyz vertices and normals on the device; a miss is all zeros.
 * d_t_hit (optional, may be NULL) receives the ray parameter of each hit (NaN on a miss).
 * Samples that the volume's brick map proves to be far (unobserved, or tsdf >= 0.999)
 * are not read; results are the same as taking them. Raycasts of one volume share its
 * skip scratch: issue them on one stream (or synchronise between them). */
int azu_raycast(azu_volume* v, const float pose[16], const float K[4], int W, int H, float* d_vertices,
                float* d_normals, azu_stream s);
int azu_raycast_dpose(azu_volume* v, const float* d_pose, const float K[4], int W, int H, float* d_vertices,
                      float* d_normals, float* d_t_hit, azu_stream s);

/* Raycast-depth loss and its adjoint (no counterpart in the reference; march and secant
 * as TSDFVolume::raycast, TSDFVolume.cpp:384-445, with a uniform step of step_voxels
 * voxels on a lattice anchored at min_depth |ray|):
 *   L = sum_pixels (d_ray - d_obs)^2 over pixels with a hit and d_obs > 0 (metres).
 * out[0] = L, out[1..6] = dL/dxi for the pose perturbation pose <- pose * [R(w) | t],
 * xi = [t; w] (the ICP increment convention, ICPTracker.cpp:100-121), out[7] = the
 * number of contributing pixels. backward == 0 computes L only. d_grad_tsdf (res^3
 * floats, may be NULL) receives += dL/dtsdf (float atomics: summation order not fixed).
 * d_depth_out (W*H, may be NULL) receives the ray depths (0 on a miss). step_voxels must
 * be in [0.05, 8]. Synchronises. */
int azu_raycast_depth_loss(azu_volume* v, const float pose[16], const float K[4], int W, int H,
                           const float* d_depth_obs, float step_voxels, int backward, float* d_grad_tsdf,
                           float* d_depth_out, double out[8], azu_stream s);

/* ---- marching cubes (kfusion::meshing::MarchingCubes, include/meshing/MarchingCubes.h:35-121) ---- */
typedef struct azu_mesh azu_mesh;
/* MarchingCubes::extract (src/meshing/MarchingCubes.cpp:174-382) on the device: the same
 * vertices, normals and triangle indices, in the same order, as the CPU mesher (colour is
 * not stored by this volume, so no colours are produced). max_triangles is the triangle
 * cap (MarchingCubes.h:49-62, default 2000000); a capped mesh reports truncated = 1.
 * Synchronises. */
int azu_mc_extract(azu_volume* v, unsigned long long max_triangles, azu_mesh** out, azu_stream s);
int azu_mesh_info(const azu_mesh* m, unsigned long long* vertices, unsigned long long* triangles, int* truncated);
/* Host copies: positions / normals vertices*3 floats, indices triangles*3. */
int azu_mesh_download(const azu_mesh* m, float* positions, float* normals, uint32_t* indices);
int azu_mesh_device(const azu_mesh* m, const float** d_positions, const float** d_normals,
                    const uint32_t** d_indices);
int azu_mesh_destroy(azu_mesh* m);

/* ---- ICP tracker (kfusion::tracking::ICPTracker, include/tracking/ICPTracker.h:87-134) ---- */
/* W x H is the live frame size; K its intrinsics (also the camera of the model image). */
int azu_tracker_create(int W, int H, const float K[4], const azu_icp_params* p, azu_tracker** out);
int azu_tracker_destroy(azu_tracker* t);
int azu_tracker_set_params(azu_tracker* t, const azu_icp_params* p);
/* ICPTracker::params (ICPTracker.h:128): the parameters as sanitised. */
int azu_tracker_get_params(const azu_tracker* t, azu_icp_params* out);
/* ICPTracker::sanitizeParams (src/tracking/ICPTracker.cpp:24-34), what the constructor
 * and setParams apply: the angle threshold clamps to [0, 85] degrees and a non-finite
 * one becomes 30; no other field changes. Host computation only. */
int azu_icp_params_sanitize(const azu_icp_params* in, azu_icp_params* out);
/* ICPResult{} (ICPTracker.h:27-47): identity pose, final_step +inf, everything else 0.
 * Host computation only. */
int azu_icp_result_default(azu_icp_result* out);
/* ICPTracker::prepareLiveGPU (ICPTracker.h:108) with the CPU semantics of buildFrameData
 * + computeNormals + buildFramePyramid (src/sensor/FrameData.cpp:113-345): depth in
 * uint16 millimetres (x 0.001 m) or float metres, band-gated to [min_depth, max_depth]. */
int azu_tracker_prepare_mm(azu_tracker* t, const uint16_t* d_depth_mm, azu_stream s);
int azu_tracker_prepare_m(azu_tracker* t, const float* d_depth_m, azu_stream s);
/* Same from raw Kinect v1 11-bit codes (the reference's sensor boundary, cpuDepthMeters,
 * include/sensor/DepthValidity.h:55-93, feeding buildFrameData, FrameData.cpp:137-194):
 * raw 0 and raw >= 2047 are invalid, metres = 1 / (raw * -0.0030711016 + 3.3309495161)
 * in float, then the same [min_depth, max_depth] band (which also rejects the codes
 * around and beyond the curve's pole at raw ~1084.6). */
int azu_tracker_prepare_raw(azu_tracker* t, const uint16_t* d_raw, azu_stream s);
/* ICPTracker::track takes a host FramePyramid (ICPTracker.h:91-94); these load one
 * level of it from device buffers instead (copies; d_depth may be NULL for levels the
 * ICP reads only as vertices / normals), and rebuild levels 1 and 2 from level 0 with
 * the buildFramePyramid downsample (FrameData.cpp:262-345). */
int azu_tracker_load_level(azu_tracker* t, int level, const float* d_depth, const float* d_vertices,
                           const float* d_normals, azu_stream s);
int azu_tracker_downsample(azu_tracker* t, azu_stream s);
/* Device buffers of pyramid level 0..2 (depth metres, packed vertices / normals). */
int azu_tracker_level(azu_tracker* t, int level, const float** d_depth, const float** d_vertices,
                      const float** d_normals, int* W, int* H);
/* ICPTracker::track / trackPreparedGPU (ICPTracker.cpp:40-134) against a model image
 * (MW x MH packed world-space vertices / normals on the device, raycast at ref_pose).
 * Enqueues every iteration; read the outcome with azu_tracker_result. */
int azu_tracker_track(azu_tracker* t, const float* d_model_v, const float* d_model_n, int MW, int MH,
                      const float pose_estimate[16], const float ref_pose[16], azu_stream s);
/* Same, estimate and reference pose read from device memory when the first kernel runs. */
int azu_tracker_track_dpose(azu_tracker* t, const float* d_model_v, const float* d_model_n, int MW, int MH,
                            const float* d_pose_estimate, const float* d_ref_pose, azu_stream s);
/* Sets the tracker's device pose (the one azu_tracker_pose_device points at) without
 * tracking, e.g. the first frame's pose before any model exists. */
int azu_tracker_set_pose(azu_tracker* t, const float pose[16], azu_stream s);
/* The reference pipeline's constant-velocity estimate (src/app/PipelineController.cpp:
 * 1162-1187): predicted = current * (last^-1 * current), then last = current, where
 * current is the tracker's device pose. *d_predicted receives the device pointer of the
 * prediction (16 floats), to pass as the estimate of azu_tracker_track_dpose. */
int azu_tracker_predict(azu_tracker* t, const float** d_predicted, azu_stream s);
/* Sets the tracked pose (the one azu_tracker_pose_device points at and
 * azu_tracker_predict extrapolates from) without touching the motion model's previous
 * pose: after a graded track, the pose the pipeline continues from
 * (PipelineController.cpp:1500-1590 sets current_pose_ only). */
int azu_tracker_set_tracked_pose(azu_tracker* t, const float pose[16], azu_stream s);
/* The reference's TrackingPolicy defaults (TrackingPolicy.h:40-48): 100, 2000, 0.40,
 * 0.015 m, 0.15 m, 0.52 rad, 3. */
int azu_tracking_policy_defaults(azu_tracking_policy* p);
/* Grades one ICP result the way the reference pipeline does while tracking
 * (PipelineController.cpp:1370-1406): when degeneracy_rel > 0 the pose is first replaced
 * by keepObservableMotion(prev_pose, r->pose, r->information, degeneracy_rel)
 * (TrackingPolicy.h:126-161; the pipeline's default ratio is 5e-3), then graded by
 * classifyTracking against prev_pose (TrackingPolicy.h:93-99). p NULL: the defaults.
 * Host computation only. */
int azu_track_grade(const azu_icp_result* r, const float prev_pose[16], const azu_tracking_policy* p,
                    float degeneracy_rel, azu_track_decision* out);
/* weakestDirectionRatio (TrackingPolicy.h:104-112): smallest over largest eigenvalue of
 * a 6x6 information matrix, 0 for an empty, non-finite or unobserved direction. */
int azu_track_weakest_ratio(const float information[36], float* out);
/* Device pointer to the tracked pose (16 floats), valid once the enqueued track ran. */
int azu_tracker_pose_device(azu_tracker* t, const float** d_pose);
/* Synchronises the stream and returns the result of the last track. */
int azu_tracker_result(azu_tracker* t, azu_icp_result* out, azu_stream s);
/* Diagnostic, summed over the ICP iterations that ran (not skipped after a level's
 * stop) since the last reset: out[0] device clock ms of block 0's cross-block FP64 sum
 * (every block computes the same one), out[1] of its solve + pose update, out[2] the
 * number of such iterations,
 * out[3] interior live pixels they visited, out[4] their in-bounds model projections.
 * Synchronises; reset != 0 clears. */
int azu_tracker_phase_times(azu_tracker* t, double out[5], int reset);

/* ---- depth conditioning (kfusion::sensor::SignalConditioner, include/sensor/SignalConditioner.h:29-47) ----
 * The depth half of SignalConditioner::process (src/sensor/SignalConditioner_omp.cpp:
 * 337-349) over raw Kinect v1 11-bit codes (uint16, W x H, row-major): a 3x3 median
 * (denoiseDepthSpatial, :442-478), then, unless the handle is in minimal mode, the raw-0
 * hole fill (fillDepthHoles, :578-635), the 9x9 luma-guided filter (guidedDepthFilter,
 * :637-716) and the temporal EMA (applyDepthEma, :494-576), whose per-pixel history the
 * handle keeps. The RGB half (bilateral, median blur, CAS sharpening) is not here: the
 * guidance luma it produces (SignalConditioner_omp.cpp:380-385, one float in [0, 1] per
 * pixel) is an input; NULL stands for the all-zero guidance of a cleared conditioner
 * (:283). Codes are the reference's bit for bit; the guided filter's Gaussians are the
 * correctly rounded float exponential (the reference calls the C library's expf). */
typedef struct azu_conditioner azu_conditioner;
enum {
    AZU_COND_MEDIAN = 1, /* denoiseDepthSpatial */
    AZU_COND_HOLES = 2,  /* fillDepthHoles */
    AZU_COND_GUIDED = 4, /* guidedDepthFilter */
    AZU_COND_EMA = 8,    /* applyDepthEma */
    AZU_COND_ALL = 15
};
/* SignalConditioner() (SignalConditioner_omp.cpp:233-245) for W x H frames, W, H in
 * [4, 8192]; the EMA history starts at 0 (no history). Synchronises. */
int azu_cond_create(int width, int height, azu_conditioner** out);
int azu_cond_destroy(azu_conditioner* c);
/* resetEMA() (SignalConditioner_omp.cpp:304-311): every history slot to +0.0f. The
 * handle holds no other depth state, so this is also reset()'s depth part (:273-302). */
int azu_cond_reset_ema(azu_conditioner* c, azu_stream s);
/* setMinimalDepth (SignalConditioner.h:37): process() runs the median only. */
int azu_cond_set_minimal_depth(azu_conditioner* c, int minimal);
/* process()'s depth stages in the reference's order, band [min_m, max_m] metres.
 * d_out may equal d_in. d_luma: W x H floats or NULL. */
int azu_cond_process(azu_conditioner* c, const uint16_t* d_in, const float* d_luma, uint16_t* d_out,
                     float min_m, float max_m, azu_stream s);
/* Any subset of the stages (AZU_COND_* bits), still in the order median, holes,
 * guided, EMA: the reference's single-stage test seams fillDepthHolesForTests /
 * applyDepthEmaForTests (SignalConditioner.h:201-206) are AZU_COND_HOLES / AZU_COND_EMA. */
int azu_cond_run_stages(azu_conditioner* c, const uint16_t* d_in, const float* d_luma, uint16_t* d_out,
                        float min_m, float max_m, int stages, azu_stream s);
/* emaStateForTests (SignalConditioner.h:208): the W x H history in metres, 0 = none, to
 * dst (device or host memory; a host dst is complete once the stream has synchronised),
 * and the inverse, for callers that restore a conditioner's state. */
int azu_cond_read_ema(azu_conditioner* c, float* dst, azu_stream s);
int azu_cond_write_ema(azu_conditioner* c, const float* src, azu_stream s);

/* ---- per-kernel timing (HIP events on the launch stream) ---- */
/* While enabled, every kernel launch is bracketed by events; azu_profile_read
 * synchronises and returns, per kernel name, launches and total milliseconds.
 * names: n_max entries of 48 chars. Returns the number of entries in *n. */
int azu_profile_enable(int on);
int azu_profile_reset(void);
int azu_profile_read(char* names, int n_max, long long* launches, double* total_ms, int* n);

#ifdef __cplusplus
}
#endif

#endif /* AZU_HIP_H */
