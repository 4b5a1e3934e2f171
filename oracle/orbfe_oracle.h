/*
 * orbfe oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C++ CPU restatement of the reference ORB front-end path
 * (resources/{create_pyramid,detect_corners,cull_corners,generate_features}.wgsl
 * with the host dispatch/override semantics of
 * wslam/src/feature_detect/{fill_pyramid,detect_corners,cull_corners,
 * generate_features}.cpp, plus the CPU matcher of
 * wslam/src/pose_estimate_cpu/match_features_cpu.cpp).
 *
 * It is the checker for the HIP path and the CPU baseline of the benchmark.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links or calls it.
 *
 * Every function names the reference file:line it restates.  Operations the
 * WebGPU spec leaves implementation-defined are fixed by this decision
 * register (shared with the HIP path, which must match bit for bit):
 *   D1 bilinear filtering: u = uv*dim, u' = u - 0.5, i0 = floor(u'),
 *      a = u' - i0, taps clamped to [0, dim-1] independently, lerp written
 *      a + t*(b - a) with separate multiply and add, x first then y;
 *   D2 f32 division correctly rounded; no FMA contraction anywhere except the
 *      explicit fmaf() calls inside D3;
 *   D3 atan2/sin/cos: own f32 routines (Cody-Waite reduction, Taylor
 *      polynomials) defined in orbfe_oracle.cpp;
 *   D4 out-of-bounds textureLoad: Tint clamp min(u32(c), dim-1) (negative
 *      coordinates reach the far edge) by default, zero mode as a flag;
 *   D5 the corners[lod].width race resolves to width = texture width;
 *   D6 features are emitted in raster order (y, x) per level and the first
 *      `cap` are kept; the block count is the number of survivors (uncapped),
 *      as the reference's atomicAdd counter is.
 */
#ifndef ORBFE_ORACLE_H
#define ORBFE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_LEVELS 8

typedef struct oracle_params {
    uint32_t frame_width;        /* W0 */
    uint32_t frame_height;       /* H0 */
    uint32_t levels;             /* 6 */
    double scale_factor;         /* 1.2 */
    float fast_threshold;        /* THRESHOLD, 0.15f in the reference */
    uint32_t fast_min_run;       /* N_SIMILLAR_MIN, 9 */
    int32_t orientation_radius;  /* ORIENTATION_R, 3 */
    uint32_t cap;                /* FEATURE_BLOCK_SZ per level */
    uint32_t border_zero;        /* 0 = Tint clamp (D4), 1 = zero */
} oracle_params;

/* gpumodels::Feature, wslam/src/models.hpp:31-69 (48 bytes). */
typedef struct oracle_gpu_feature {
    float x;
    float y;
    uint32_t strength;
    float orientation;
    uint32_t descriptor[8];
} oracle_gpu_feature;

/* wslam::Feature, wslam/src/models.hpp:86-111 (52 bytes). */
typedef struct oracle_feature {
    uint32_t x;
    uint32_t y;
    uint32_t lod;
    uint32_t strength;
    float orientation;
    uint32_t descriptor[8];
} oracle_feature;

/* Level dimensions, wslam/src/common.cpp:38-45 (texture extents). */
void oracle_level_dims(uint32_t W0, uint32_t H0, uint32_t levels, double scale,
                       uint32_t* w, uint32_t* h);

/* u8 -> f32 = float(u8)/255, data/src/provider_base.cpp:121-124. */
void oracle_lod0(const uint8_t* img, size_t pitch, uint32_t W0, uint32_t H0,
                 float* out);

/* create_pyramid.wgsl:5-14 for one level from LOD0. */
void oracle_pyramid_level(const float* lod0, uint32_t W0, uint32_t H0,
                          uint32_t w, uint32_t h, float* out);

/* detect_corners.wgsl:34-89 for one level; S has w*h entries. */
void oracle_detect(const float* lvl, uint32_t w, uint32_t h, float threshold,
                   uint32_t min_run, uint32_t border_zero, uint32_t* S);

/* cull_corners.wgsl:21-114 with the dispatch grids of
 * cull_corners.cpp:332-353.  S_all / O_all hold `levels` blocks of W0*H0
 * entries each (CORNER_BLOCK_SZ); dims are the level extents. */
void oracle_cull(const uint32_t* S_all, uint32_t W0, uint32_t H0,
                 uint32_t levels, const uint32_t* dims_w,
                 const uint32_t* dims_h, uint32_t* O_all);

/* generate_features.wgsl:48-172 for one level, survivors in raster order.
 * Writes min(count, cap) features, returns the survivor count. */
uint32_t oracle_generate(const float* lvl, uint32_t w, uint32_t h,
                         const uint32_t* O, uint32_t cap,
                         int32_t orientation_radius, uint32_t border_zero,
                         oracle_gpu_feature* out);

/* Whole feature-detect stage for one frame.  pyramid_out (nullable) gets
 * levels 1..L-1 packed back to back; feature_array_out gets the
 * gpumodels::FeatureArray layout: per level {u32 count; u32 pad;
 * Feature v[cap]} (models.hpp:72-77), unused slots zero. */
int oracle_detect_frame(const uint8_t* img, size_t pitch,
                        const oracle_params* p, float* pyramid_out,
                        void* feature_array_out);

/* LoadDataCPUPass::loadFeatures + FromGpuModel
 * (load_features_cpu.cpp:164-188, models.hpp:113-120).  Writes the
 * FeatureSet flattened level by level; counts[l] = min(count, cap). */
void oracle_load_features(const void* feature_array, uint32_t levels,
                          uint32_t cap, oracle_feature* out, uint32_t* counts);

/* Own f32 math of decision D3. */
float oracle_atan2f(float y, float x);
void oracle_sincosf(float x, float* s, float* c);

/* HammingDistanceBounded, match_features_cpu.cpp:42-55. */
uint32_t oracle_hamming_bounded(const uint32_t* a, const uint32_t* b,
                                uint32_t cap);

/* FindMatches (match_features_cpu.cpp:152-207) on two FeatureSets given
 * flattened with per-level counts.  For every prev feature (level l, index
 * i, flattened) writes the matched curr feature as (lod << 24 | index) or
 * -1.  Also writes the number of result pairs. */
int oracle_find_matches(const oracle_feature* prev, const uint32_t* prev_counts,
                        const oracle_feature* curr, const uint32_t* curr_counts,
                        uint32_t levels, uint32_t W0, uint32_t H0, float ratio,
                        int32_t* prev_match, uint32_t* n_pairs);

/* Brute-force best/second-best search: for each query the lowest-index
 * train with minimum Hamming distance, the distance, and the second
 * smallest distance over the multiset (UINT32_MAX when nt < 2). */
void oracle_match_bf(const uint32_t* q, uint32_t nq, const uint32_t* t,
                     uint32_t nt, uint32_t* idx, uint32_t* d1, uint32_t* d2);

/* SampsonDistance + ClassifyInliers (ransac_cpu.cpp:178-209) for k
 * fundamental matrices (9 row-major doubles each) over n correspondences
 * (prev, curr as (x, y) double pairs): counts[h] = inliers of hypothesis h;
 * mask (nullable) gets mask[h*n + i] = 1 for each inlier. */
void oracle_sampson_counts(const double* F, uint32_t k, const double* prev, const double* curr,
                           uint32_t n, double threshold, uint32_t* counts, uint8_t* mask);

/* FitFundamentalMatrix (ransac_cpu.cpp:111-175) on n >= 8 correspondences:
 * writes F (row-major, unit-norm null vector before denormalisation) and
 * returns 1, or returns 0 when degenerate.  The null vector comes from a
 * cyclic Jacobi eigendecomposition of A^T A. */
int oracle_fit_fundamental(const double* prev, const double* curr, uint32_t n, double* F);

/* RansacCPU::execute (ransac_cpu.cpp:262-370) over `sets` consecutive sets
 * of correspondences (sizes[s] each, prev and curr concatenated), one
 * generator seeded with 0xC0FFEE1234 carried across the sets, threshold
 * 3.84, confidence 0.99, min_matches 8.  Per set: stats[4*s..] = total,
 * iterations run, inliers, model found; F[9*s..]; mask (concatenated like
 * the sets) = 1 for the final inliers. */
void oracle_ransac(const uint32_t* sizes, uint32_t sets, const double* prev, const double* curr,
                   uint32_t max_iterations, uint64_t* stats, double* F, uint8_t* mask);

#ifdef __cplusplus
}
#endif

#endif
