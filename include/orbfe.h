/*
 * liborbfe -- MI355X (gfx950) ORB front-end: pyramid, FAST-style corner
 * score, the reference's corner culling, rotated BRIEF, and Hamming
 * matching, as HIP kernels behind a plain C ABI.
 *
 * Conventions (the C form of the reference's `std::optional<std::string>`
 * error convention, compute/include/compute/pass.hpp:24-37):
 *   - every entry point returns ORBFE_OK (0) or a negative ORBFE_E* code;
 *   - on failure orbfe_last_error(ctx) returns the message (owned by ctx,
 *     valid until the next call on ctx);
 *   - "_async" entry points take DEVICE pointers and only enqueue work on
 *     `stream`; the others take HOST pointers and return after the results
 *     are in host memory;
 *   - calls on one context must be issued from one host thread at a time
 *     and, for the async forms, on one stream (the context owns scratch
 *     memory that consecutive calls reuse).
 *
 * Layouts (bit-compatible with the reference):
 *   - orbfe_gpu_feature == gpumodels::Feature (wslam/src/models.hpp:31-69);
 *   - a "feature array" is gpumodels::FeatureArray (models.hpp:72-77):
 *     `levels` blocks of {uint32 count; uint32 pad; orbfe_gpu_feature
 *     v[cap]}, block stride 8 + 48*cap bytes.  count is the number of corner
 *     survivors at that level (it can exceed cap, as the reference's
 *     atomicAdd counter can); v[0 .. min(count, cap)) hold the features in
 *     raster order (y, then x) and the remaining slots are zero;
 *   - orbfe_feature == wslam::Feature (models.hpp:86-111);
 *   - a pyramid buffer holds levels 1..levels-1 as tightly packed row-major
 *     float32 images, back to back (level 0 is the input frame / 255).
 */
#ifndef ORBFE_H
#define ORBFE_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORBFE_OK 0
#define ORBFE_EINVAL (-1)
#define ORBFE_EHIP (-2)
#define ORBFE_ENOMEM (-3)
#define ORBFE_ERANGE (-4)

#define ORBFE_MAX_LEVELS 8

/* Tunables of the reference's passes. */
typedef struct orbfe_config {
    uint32_t frame_width;   /* W0: WSLAM_FRAME_WIDTH, wslam/include/wslam/common.hpp:11-13 */
    uint32_t frame_height;  /* H0: WSLAM_FRAME_HEIGHT, common.hpp:15-17.  Frame
                               sizes (here and in batches) run from 1 to 65535
                               per side, with each frame's pyramid, smoothed
                               images and survivor words under 4 GiB (32-bit
                               offsets; about 19000 x 19000 at 6 levels) and
                               (height - 1) * pitch + width below 2^32;
                               ORBFE_ERANGE otherwise */
    uint32_t levels;        /* GPUConst::levels_of_detail = 6, common.hpp:43 */
    double scale_factor;    /* GPUConst::lod_scale_factor = 1.2, common.hpp:51 */
    float fast_threshold;   /* THRESHOLD override, detect_corners.wgsl:22 (0.15) */
    uint32_t fast_min_run;  /* N_SIMILLAR_MIN override, detect_corners.wgsl:23 (9) */
    int32_t orientation_radius; /* ORIENTATION_R, generate_features.wgsl:46 (3) */
    uint32_t features_per_level; /* FEATURE_BLOCK_SZ; 0 = W0*H0/250, common.hpp:47 */
    uint32_t border_zero;   /* textureLoad out of range: 0 = Tint clamp
                               (robustness.cc:344-355), 1 = read zero */
    float lowe_ratio;       /* kLoweTestThreshold, match_features_cpu.cpp:75 (0.85);
                               must lie in (0, 1] (ORBFE_ERANGE otherwise): above
                               1 a tie at the best distance could pass the
                               ratio test, and which tied candidate the
                               reference keeps depends on its unstable sort */
    /* Launch shapes (no counterpart in the reference; results do not depend
       on them).  0 = default.  detect_tile_rows: rows of a detect tile, 8, 16
       (default) or 32.  describe_lanes: lanes per keypoint in the describe
       kernel, 8, 16 or 32 (0: 16 for batches, 32 for a single frame).
       match_lanes: lanes per query in the
       pipeline matcher, 16, 32 or 64 (0: 16 for batches, 64 for a single
       frame pair).  uni_slam_amd/autotune.py picks them by timing the
       variants on a workload. */
    uint32_t detect_tile_rows;
    uint32_t describe_lanes;
    uint32_t match_lanes;
} orbfe_config;

typedef struct orbfe_gpu_feature {
    float x;
    float y;
    uint32_t strength;
    float orientation;
    uint32_t descriptor[8];
} orbfe_gpu_feature;

typedef struct orbfe_feature {
    uint32_t x;
    uint32_t y;
    uint32_t lod;
    uint32_t strength;
    float orientation;
    uint32_t descriptor[8];
} orbfe_feature;

/* One frame of a batched detect call (device pointers). */
typedef struct orbfe_frame {
    const uint8_t* image;   /* u8 grayscale, height rows of `pitch` bytes */
    size_t pitch;
    uint32_t width;
    uint32_t height;
    uint32_t cap;           /* features per level; 0 = width*height/250 */
    uint32_t reserved;
    float* pyramid;         /* nullable: levels 1.. output */
    void* features;         /* feature array output, stride 8 + 48*cap */
} orbfe_frame;

/* One frame pair of a batched pipeline-match call (device pointers). */
typedef struct orbfe_match_pair {
    const void* prev_features; /* feature array of the reference frame */
    const void* curr_features; /* feature array of the current frame */
    int32_t* prev_match;       /* levels*cap entries, see orbfe_match_pipeline_async */
    uint32_t width;            /* frame extent both arrays were detected at */
    uint32_t height;
    uint32_t cap;              /* cap both arrays were detected with */
    uint32_t reserved;
} orbfe_match_pair;

/* Defaults of the reference for a W0 x H0 frame. */
int orbfe_config_default(orbfe_config* cfg, uint32_t width, uint32_t height);

/* Creates a context: builds the per-resolution plan (level extents, the
 * cull mask of cull_corners.wgsl) and allocates scratch for one frame.
 * Replaces the initialize() chain of FillPyramidPass / PassDetectCorners /
 * CullCornersPass / GenerateFeaturesPass
 * (wslam/src/feature_detect/fill_pyramid.cpp, detect_corners.cpp:27-58,
 * cull_corners.cpp:38-60, generate_features.cpp:71-101). */
int orbfe_create(const orbfe_config* cfg, int device, struct orbfe_ctx** out);
void orbfe_destroy(struct orbfe_ctx* ctx);
const char* orbfe_last_error(const struct orbfe_ctx* ctx);

/* Level extents (texture sizes of wslam/src/common.cpp:38-45). */
int orbfe_level_dims(const struct orbfe_ctx* ctx, uint32_t* widths, uint32_t* heights);
/* Bytes of one feature array for this context (6 * (8 + 48*cap) by default). */
size_t orbfe_feature_array_bytes(const struct orbfe_ctx* ctx);
/* Floats of one pyramid buffer (levels 1..L-1). */
size_t orbfe_pyramid_floats(const struct orbfe_ctx* ctx);
uint32_t orbfe_cap(const struct orbfe_ctx* ctx);

/* Feature-detect stage of one frame: FillPyramidPass::prepareExecute
 * (fill_pyramid.cpp:206-314), PassDetectCorners::prepareExecute
 * (detect_corners.cpp:355-385), CullCornersPass::prepareExecute
 * (cull_corners.cpp:292-353), GenerateFeaturesPass::prepareExecute
 * (generate_features.cpp:495-527), enqueued on `stream`.  image is the u8
 * frame (the reference uploads it as float(u8)/255, provider_base.cpp:121-124;
 * the conversion happens in-kernel).  pyramid may be NULL. */
int orbfe_detect_async(struct orbfe_ctx* ctx, const uint8_t* image, size_t pitch,
                       float* pyramid, void* features, hipStream_t stream);

/* Host-memory form of orbfe_detect_async plus the readback of
 * LoadDataCPUPass::loadFeatures (load_features_cpu.cpp:164-188): copies the
 * frame in, runs the stage, copies the feature array (and the pyramid when
 * pyramid != NULL) out. */
int orbfe_detect(struct orbfe_ctx* ctx, const uint8_t* image, size_t pitch,
                 float* pyramid, void* features);

/* The same stage for many frames of arbitrary sizes, one launch per kernel
 * for the whole batch.  Frames may differ in size and cap; all share the
 * context's other tunables.  Results are identical to per-frame
 * orbfe_detect_async.  The frame table is uploaded through a small ring of
 * pinned staging slots, so this form suits a serving loop that forms a new
 * batch every tick; it is not capturable into a HIP graph. */
int orbfe_detect_batch_async(struct orbfe_ctx* ctx, const orbfe_frame* frames,
                             uint32_t n_frames, hipStream_t stream);

/* A prepared batch: the frame table and scratch are set up once (this call
 * synchronizes), after which orbfe_batch_detect_async only enqueues the
 * kernels, so it can be captured into a HIP graph and replayed.  The
 * batch's buffers belong to the caller and must stay valid. */
int orbfe_batch_create(struct orbfe_ctx* ctx, const orbfe_frame* frames,
                       uint32_t n_frames, struct orbfe_batch** out);
int orbfe_batch_detect_async(struct orbfe_ctx* ctx, struct orbfe_batch* batch,
                             hipStream_t stream);
void orbfe_batch_destroy(struct orbfe_batch* batch);

/* FromGpuModel over a host feature array (models.hpp:113-120,
 * load_features_cpu.cpp:180-185): writes the FeatureSet flattened level by
 * level into out (room for levels*cap entries) and counts[l] = min(count,
 * cap).  Pure host code. */
int orbfe_load_features(const void* features, uint32_t levels, uint32_t cap,
                        orbfe_feature* out, uint32_t* counts);

/* Zero-copy readback: replaces GPU::readBuffer of the feature slab
 * (compute/src/gpu.cpp:892-929) and the FromGpuModel loop of
 * LoadDataCPUPass::loadFeatures (load_features_cpu.cpp:164-188).  A kernel
 * on `stream` writes the FeatureSet of a device feature array of this
 * context's frame size straight into pinned host memory: out gets the
 * orbfe_feature (wslam::Feature) records of every level back to back,
 * min(count, cap) per level, and counts[l] = min(count, cap).  out and
 * counts must come from hipHostMalloc (room for levels*cap records and
 * levels counts; ORBFE_EINVAL otherwise); they hold the result once the
 * stream has passed this call's work.  Only the kept records cross the bus,
 * not the cap-sized slab. */
int orbfe_export_features_async(struct orbfe_ctx* ctx, const void* features,
                                orbfe_feature* out, uint32_t* counts, hipStream_t stream);

/* FindMatches(prev, curr) of match_features_cpu.cpp:152-207 on two device
 * feature arrays of this context's frame size: LOD+-1 candidate pools, the
 * 10% spatial window, the Lowe ratio, the mutual check.  prev_match gets
 * levels*cap int32: entry [l*cap + i] is, for prev feature i of level l, the
 * matched current feature as (lod << 24 | index), or -1.  This is the
 * MatchResult of the reference (result[l] maps that current feature to prev
 * feature i). */
int orbfe_match_pipeline_async(struct orbfe_ctx* ctx, const void* prev_features,
                               const void* curr_features, int32_t* prev_match,
                               hipStream_t stream);

/* Batched forms: many frame pairs, possibly of different sizes, one launch
 * per kernel; dynamic (pinned staging ring) and prepared (capturable). */
int orbfe_match_pipeline_batch_async(struct orbfe_ctx* ctx,
                                     const orbfe_match_pair* pairs,
                                     uint32_t n_pairs, hipStream_t stream);
int orbfe_match_batch_create(struct orbfe_ctx* ctx, const orbfe_match_pair* pairs,
                             uint32_t n_pairs, struct orbfe_match_batch** out);
int orbfe_match_batch_async(struct orbfe_ctx* ctx, struct orbfe_match_batch* batch,
                            hipStream_t stream);
void orbfe_match_batch_destroy(struct orbfe_match_batch* batch);

/* Host-memory form of orbfe_match_pipeline_async (MatchFeaturesCPU::execute,
 * match_features_cpu.cpp:217-268): host feature arrays in, prev_match out;
 * n_pairs (nullable) gets the number of matches. */
int orbfe_match_pipeline(struct orbfe_ctx* ctx, const void* prev_features,
                         const void* curr_features, int32_t* prev_match,
                         uint32_t* n_pairs);

/* Brute-force Hamming search of nq query against nt train 256-bit
 * descriptors (8 uint32 words each, device memory): for each query the
 * lowest-index train at the minimum distance (idx, d1) and the second
 * smallest distance d2 over all trains (0xFFFFFFFF when nt < 2; idx =
 * 0xFFFFFFFF and d1 = 0xFFFFFFFF when nt == 0).  The exact distance of
 * HammingDistanceBounded (match_features_cpu.cpp:26-55) without a window.
 * nt must be < 2^23. */
int orbfe_match_bf_async(struct orbfe_ctx* ctx, const uint32_t* query, uint32_t nq,
                         const uint32_t* train, uint32_t nt, uint32_t* idx,
                         uint32_t* d1, uint32_t* d2, hipStream_t stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the search of orbfe_match_bf_async, with identical
 * results, whose per-split partial results live in a static device array
 * of ORBFE_BF_NOALLOC_SCRATCH_BYTES that the library holds on every device
 * it runs on until the process ends, instead of context memory grown with
 * hipMalloc.  The call allocates nothing.  Limit: nq * ceil(nt / 8192) *
 * 8 bytes must fit the array (nq * ceil(nt / 8192) <= 2^21, e.g. 65536
 * queries against up to 262144 train descriptors); larger searches fail
 * with ORBFE_ERANGE.  Every call on one device shares that device's array:
 * calls that overlap on one GPU (on different streams, or from different
 * contexts) race; calls ordered on one stream do not. */
#define ORBFE_BF_NOALLOC_SCRATCH_BYTES (16u << 20)
int orbfe_match_bf_noalloc_async(struct orbfe_ctx* ctx, const uint32_t* query, uint32_t nq,
                                 const uint32_t* train, uint32_t nt, uint32_t* idx,
                                 uint32_t* d1, uint32_t* d2, hipStream_t stream);

/* Device scratch (bytes) that one frame of width x height with `cap`
 * features per level takes in a dynamic batch of this context (detect and
 * match stages, pyramid not exported): what a serving loop admits against
 * its HBM budget.  Creates the frame size's plan if needed. */
int orbfe_frame_scratch_bytes(struct orbfe_ctx* ctx, uint32_t width, uint32_t height,
                              uint32_t cap, size_t* bytes);

/* ---------------------------------------------------------------------------
 * Served form: continuous batching of frames from many camera streams on one
 * GPU.  The reference runs one frame per stage submission from one host loop
 * (wslam/src/main.cpp:64-135, feature_detect.cpp:43-86); this is that loop
 * when many streams share the device.  Per stream a ring of
 * 2 * frames_per_stream + 2 slots stays resident in HBM (upload buffer,
 * feature array, match table; frame k of a stream uses slot k mod slots), so
 * each stream's previous-frame descriptors stay on the device.  Each launch
 * admits the oldest pending frames (FIFO over all streams, at most
 * frames_per_stream consecutive frames of one stream, each stream in order)
 * within max_batch, the HBM budget (orbfe_frame_scratch_bytes per frame) and
 * the free in-flight slots, uploads them from pinned host memory and runs
 * one batched detect and one batched match (each frame against its stream's
 * previous frame) on that slot's HIP stream; a batch's uploads overlap the
 * kernels of the batch before it.  Every frame's features and match table
 * equal the plain per-frame path's (orbfe_detect_async +
 * orbfe_match_pipeline_async) bit for bit.  Create allocates every buffer and
 * warms every plan, arena and staging slot; serving allocates no device
 * memory.  One host thread drives a server.
 * ------------------------------------------------------------------------- */
typedef struct orbfe_stream_spec {
    uint32_t width;
    uint32_t height;
    uint32_t cap;       /* features per level; 0 = width*height/250 */
    uint32_t reserved;
} orbfe_stream_spec;

typedef struct orbfe_server_config {
    orbfe_config detect;        /* tunables and launch shapes of every stream
                                   (frame size and cap fields are per stream) */
    size_t hbm_budget;          /* device scratch of the frames in flight; 0 = 1 GiB */
    uint32_t max_batch;         /* frames per batch; 0 = 64 */
    uint32_t inflight;          /* batches on the device at once (one HIP
                                   stream each), 1..8; 0 = 2 */
    uint32_t frames_per_stream; /* consecutive frames of one stream per batch; 0 = 3 */
    uint32_t log_batches;       /* nonzero: record every batch's frames
                                   (orbfe_server_batch_log) */
} orbfe_server_config;

typedef struct orbfe_arrival {
    double t;          /* arrival time, seconds from the start of the schedule */
    uint32_t stream;
    uint32_t seq;      /* frame number within the stream (picks the pool frame) */
} orbfe_arrival;

typedef struct orbfe_completion {
    uint32_t stream;
    uint32_t seq;      /* frame number within the stream, in submission order */
    double t_arrival;  /* seconds on the server clock */
    double t_done;     /* when the server saw the frame's batch complete */
} orbfe_completion;

int orbfe_server_config_default(orbfe_server_config* cfg);
int orbfe_server_create(const orbfe_server_config* cfg, const orbfe_stream_spec* streams,
                        uint32_t n_streams, int device, struct orbfe_server** out);
void orbfe_server_destroy(struct orbfe_server* srv);
const char* orbfe_server_last_error(const struct orbfe_server* srv);
/* Seconds since the server was created (steady clock): the time base of
 * t_arrival and t_done. */
double orbfe_server_clock(const struct orbfe_server* srv);
/* Queues one frame of `stream` (width*height u8 in pinned host memory, valid
 * until the frame completes).  features_out / match_out (nullable, pinned
 * host memory) receive the frame's feature array and its match table against
 * the stream's previous frame (seq > 0) by the time it completes.  *seq (nullable)
 * gets the frame's number within the stream. */
int orbfe_server_submit(struct orbfe_server* srv, uint32_t stream, const uint8_t* frame,
                        double t_arrival, void* features_out, int32_t* match_out,
                        uint32_t* seq);
/* Retires finished batches and launches as many new ones as the in-flight
 * slots allow; hands out up to max_done completions (n_done of them) and the
 * number of batches launched.  Never blocks. */
int orbfe_server_pump(struct orbfe_server* srv, orbfe_completion* done, uint32_t max_done,
                      uint32_t* n_done, uint32_t* launched);
/* 1 when nothing is pending, in flight or waiting to be handed out. */
int orbfe_server_idle(const struct orbfe_server* srv);
int orbfe_server_stats(const struct orbfe_server* srv, uint64_t* batches, uint32_t* max_batch);
/* With log_batches: the frames of every launched batch, in launch order
 * (batch number, stream, seq, submission number); *n gets the total. */
typedef struct orbfe_batch_entry {
    uint32_t batch;
    uint32_t stream;
    uint32_t seq;
    uint32_t order;
} orbfe_batch_entry;
int orbfe_server_batch_log(const struct orbfe_server* srv, orbfe_batch_entry* out, size_t max,
                           size_t* n);
/* Replays a schedule: arrival i (sorted by t, each stream's in seq order) is
 * submitted at t * time_scale seconds after the call starts (time_scale 0:
 * every frame waiting at once), with frame pool[stream * pool_per_stream +
 * seq % pool_per_stream]; pumps until every frame completed.  done gets n
 * completions in completion order; *wall_s the elapsed seconds. */
int orbfe_server_run(struct orbfe_server* srv, const orbfe_arrival* arrivals, uint32_t n,
                     const uint8_t* const* pool, uint32_t pool_per_stream, double time_scale,
                     orbfe_completion* done, double* wall_s);

/* RANSAC hypothesis scoring, the per-iteration inner loop of
 * RansacCPU::execute (ClassifyInliers + SampsonDistance,
 * wslam/src/pose_estimate_cpu/ransac_cpu.cpp:178-209), for k hypotheses at
 * once.  F: k fundamental matrices, 9 row-major doubles each; prev, curr:
 * n correspondences as (x, y) double pairs in LOD-0 pixels (16-byte
 * aligned), prev the first image's point (F's column side).  For each
 * hypothesis h, counts[h] gets the number of correspondences whose Sampson
 * distance (f64, the reference's operation order, +inf for a denominator
 * below 1e-12) is below `threshold`; with mask != NULL, mask[h*n + i] = 1
 * for each inlier and 0 otherwise.  Device pointers, k <= 65535. */
int orbfe_ransac_score_async(struct orbfe_ctx* ctx, const double* F, uint32_t k,
                             const double* prev, const double* curr, uint32_t n,
                             double threshold, uint32_t* counts, uint8_t* mask,
                             hipStream_t stream);

/* Kernel timing: when enabled, every kernel launch of the context is
 * bracketed by HIP events on its stream; orbfe_timing_read synchronizes and
 * returns, for kernel `name`, the launch count and summed duration (ms). */
int orbfe_timing_enable(struct orbfe_ctx* ctx, int enable);
int orbfe_timing_read(struct orbfe_ctx* ctx, const char* name, uint32_t* launches,
                      double* total_ms);
int orbfe_timing_reset(struct orbfe_ctx* ctx);

/* Sustained VALU issue rate of the context's device, in wave64 instructions
 * per second: a probe kernel of independent 32-bit integer xor/add chains on
 * every CU, timed with HIP events on `stream` (synchronizes).  No
 * counterpart in the reference: the peak that the integer rooflines of the
 * matchers are quoted against. */
int orbfe_probe_valu_rate(struct orbfe_ctx* ctx, hipStream_t stream,
                          double* wave_instructions_per_s);

/* Sustained int8 matrix-core rate of the context's device, in int8
 * operations (2 per multiply-add) per second: a probe kernel of independent
 * v_mfma_i32_32x32x32_i8 chains (the instruction of the brute-force
 * matcher) on every CU, timed with HIP events on `stream` (synchronizes).
 * No counterpart in the reference: the measured ceiling beside the published
 * 5.0 POP/s that the brute-force matcher is quoted against. */
int orbfe_probe_mfma_rate(struct orbfe_ctx* ctx, hipStream_t stream, double* int8_ops_per_s);

#ifdef __cplusplus
}
#endif

#endif
