// Shared definitions of the orbfe HIP kernels and their host launchers.
//
// Arithmetic contract (the HIP path and the CPU restatement used as its
// checker agree on every bit of it):
//   * bilinear sampling (textureSampleLevel with a linear clamp-to-edge
//     sampler): u = uv*dim, u' = u - 0.5, i0 = floor(u'), a = u' - i0, both
//     taps clamped to [0, dim-1], lerp a + t*(b - a) as separate multiply and
//     add, x then y;
//   * f32 division correctly rounded; no FMA contraction (the library is
//     compiled with -ffp-contract=off) except the explicit fmaf calls of the
//     trigonometric routines below;
//   * textureLoad out of range: Tint's robustness clamp min(u32(c), dim-1)
//     (vendor/dawn/src/tint/lang/core/ir/transform/robustness.cc:344-355), so
//     a negative coordinate reads the far edge, or zero when border_zero.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orbfe {

constexpr uint32_t kMaxLevels = 8;

// Detect tile: 64 columns (one wave per row segment) x tile_h rows (8, 16
// or 32), 256 threads, each thread scoring tile_h / 4 rows.  The halo of 3
// covers the radius-3 FAST ring.
constexpr int kTileW = 64;
constexpr int kTileHDefault = 16;
constexpr int kHalo = 3;
constexpr int kLdsW = kTileW + 2 * kHalo;  // 70
constexpr int kDetectThreads = 256;

// k_describe: desc_lanes lanes per keypoint (8, 16 or 32), so 256 /
// desc_lanes keypoints per 256-thread workgroup.
constexpr uint32_t kDescLanesDefault = 16;  // batches
constexpr uint32_t kDescLanesFrame = 32;    // single-frame launches (latency-bound)
// k_match_dir: match_lanes lanes per query (16, 32 or 64), so 256 /
// match_lanes queries per 256-thread workgroup.  By default a batch uses 16
// (many queries keep the device full) and a single frame pair 64 (its
// queries alone do not).
constexpr uint32_t kMatchLanesBatch = 16;
constexpr uint32_t kMatchLanesPair = 64;
// Query list of the mutual matcher pass: kListSegs segments, each with its
// count (kListStride words apart: no two counters share a 128-byte line) and
// room for the queries of every direction block b with b % kListSegs == s
// (blocks of at most 16 queries; 256 / QPB queries per block for QPB >= 4).
constexpr uint32_t kListSegs = 64;
constexpr uint32_t kListStride = 32;
__host__ __device__ constexpr uint32_t query_list_words(uint32_t L, uint32_t cap) {
    return kListSegs * kListStride + 2u * (L * cap + 64u * L + 64u * kListSegs);
}
__host__ __device__ inline uint32_t query_segment_capacity(uint32_t blocks, uint32_t qpb) {
    return (blocks + kListSegs - 1) / kListSegs * qpb;
}

// Matcher cell index: LOD0 plane cut into kCellsX x kCellsY cells of a
// quarter of the search window's width and half its height (0.025*W0 x
// 0.05*H0), covering 1.2x the frame; a query's candidates in one cell row
// span 1.125x the window's width.
constexpr uint32_t kCellsX = 48;
constexpr uint32_t kCellsY = 24;
constexpr uint32_t kCells = kCellsX * kCellsY;
// Candidate runs of one query: up to 3 pool levels x kCellsY cell rows.
constexpr uint32_t kMaxRanges = 3 * kCellsY;

// Key packing of the pipeline matcher: distance (<= 256) above a 22-bit
// candidate id (2 bits pool level slot, 20 bits index).
constexpr uint32_t kMatchKeyShift = 22;
constexpr uint32_t kMatchIdxBits = 20;
// Brute-force matcher keys: distance above a 23-bit train index.
constexpr uint32_t kBfKeyShift = 23;
// Static partial-result scratch of the allocation-free brute-force matcher
// (ORBFE_BF_NOALLOC_SCRATCH_BYTES of orbfe.h): 8 bytes per query and split.
constexpr size_t kBfScratchBytes = size_t(16) << 20;

// The smoothed image of a level holds G(P), the 9-tap Gaussian of
// generate_features.wgsl:149-172, for every integer point P of an extended
// grid [0, w+kSmoothExt) x [0, h+kSmoothExt).  For P >= w+3 (and for the
// huge values negative coordinates wrap to) every sampler tap clamps to the
// last column, so G(P) = G(w+3); the same holds for rows.
constexpr uint32_t kSmoothExt = 4;

// Layout of a level's smoothed image: bands of 4 grid rows, each band
// column-interleaved (the value at (x, y) is float 4 * x + y % 4 of band
// y / 4), bands smooth_band_floats(w) floats apart (the extended width
// rounded up to 8 columns).  A 128-byte line then holds an 8 x 4 block of
// the grid, which a rotated BRIEF pattern's scattered points share more
// often than 32 x 1 runs of one row, and the 4 rows of one column a detect
// wave computes are one 16-byte store.
__host__ __device__ inline uint32_t smooth_band_floats(uint32_t w) {
    return 4u * ((w + kSmoothExt + 7u) & ~7u);
}
__host__ __device__ inline uint32_t smooth_floats(uint32_t w, uint32_t h) {
    return smooth_band_floats(w) * ((h + kSmoothExt + 3u) / 4u);
}

// Sampler taps of one coordinate along one axis: i0 | i1 << 16 and the
// fraction, as axis_taps computes them.
struct TapEntry {
    uint32_t i01;
    float frac;
};

struct LevelGeom {
    uint32_t w, h;
    uint32_t wpr;        // 64-pixel words per row, ceil(w / 64)
    uint32_t tiles_x, tiles_y;  // detect tiles over the extended smooth grid
    uint32_t tile_base;  // first detect tile of this level within the frame
    uint32_t pyr_off;    // float offset in the pyramid buffer (levels >= 1)
    uint32_t word_off;   // 64-bit word offset in the cull mask / survivor map
    uint32_t feat_off;   // byte offset of the level block in the feature array
    uint32_t kp_off;     // entry offset of the level's keypoint list
    uint32_t sm_off;     // float offset of the level's smoothed image
    float lod0_scale;    // product of l factors 1.2f (toLod0)
    float texel_x, texel_y;  // 1.0f / w, 1.0f / h (generate_features.wgsl:151)
    // Entry offset of the level's tap tables in Plan::taps: level texel
    // column cx (w entries) and row cy (h entries) sampled from level 0 at
    // (c + 0.5) / n, then the 9-tap Gaussian's column taps [3][w + 4] and
    // row taps [3][h + 4] at offsets -1.2, 0, 1.2 texels, then the smoothing
    // tap codes (smooth_tap_code) of every row [h + 4].
    uint32_t tap_off;
    uint32_t pad_;
};

struct Plan {
    uint32_t W0, H0, L, cap;
    uint32_t tile_h;      // detect tile rows
    uint32_t desc_lanes;  // describe lanes per keypoint
    uint32_t tiles;       // detect tiles of one frame
    uint32_t words;       // 64-bit survivor words of one frame
    uint32_t pyr_floats;  // pyramid floats (levels >= 1)
    uint32_t feat_stride; // 8 + 48*cap
    uint32_t desc_groups; // describe blocks per level = ceil(cap / (256 / desc_lanes))
    float max_dx, max_dy; // 0.1f*W0, 0.1f*H0 (match_features_cpu.cpp:76-77)
    uint32_t sm_floats;   // smoothed-image floats of one frame
    float cell_inv_x, cell_inv_y;  // 1 / cell size (cells of max_dx/4 x max_dy/2)
    // Cell index of one side (u32 words): cell starts [L][kCells + 1], then
    // the records of every level in cell order as arrays over the global
    // record number g = level * cap + k: heads (uint4: LOD0 x and y as float
    // bits, record index, 0) and the low and high descriptor halves (uint4
    // each), so a candidate is three 16-byte loads.
    uint32_t idx_words;   // words of one side's index
    uint32_t idx_head, idx_dlo, idx_dhi;  // word offsets of the arrays
    const uint64_t* mask; // static cull mask, `words` words
    const TapEntry* taps; // sampler tap tables of all levels
    LevelGeom lv[kMaxLevels];
};

// Logical block of dispatch index i in a grid of n: workgroups are dealt to
// the 8 XCDs round-robin (i % 8), so XCD x is given the contiguous logical
// range of its share; the blocks of one frame then share an L2.
__host__ __device__ inline uint32_t xcd_block(uint32_t i, uint32_t n) {
    constexpr uint32_t kXcds = 8;
    const uint32_t q = n / kXcds, r = n % kXcds;
    const uint32_t x = i % kXcds, k = i / kXcds;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// One detect tile of a plan, filled on the host with the tile decode the
// kernel would otherwise repeat: level | tile_x << 4 | tile_y << 16, then the
// source rectangle of level 0 its staged texels sample (first dword column
// cs, first row rowmin, nd dwords per row, nrows rows) and whether it fits
// the LDS source buffer of the kernel's two forms (bit 31: the full buffer;
// bit 30: the small one of single-frame launches).
struct TileDesc {
    uint32_t lxy;
    uint32_t cs;
    uint32_t rowmin;
    uint32_t nd_rows;  // nd | nrows << 16 | fits_small << 30 | fits << 31
};

// Dwords of the detect kernel's LDS source buffer for tiles of th rows: at
// level scales up to 2.5 (six levels of 1.2) a tile's source rectangle spans
// at most (th+6)*2.5 + 4 rows of 45 dwords (70 texels * 2.5 plus the
// bilinear neighbour and the dword alignment).
__host__ __device__ constexpr uint32_t detect_src_words(uint32_t th) {
    return ((th + 6) * 5u / 2u + 4u) * 45u;
}

// The small source buffer of single-frame launches: level scales up to 1.75
// (the first four levels of 1.2), (th+6)*1.75 + 4 rows of 33 dwords; tiles
// of higher levels read level 0 from memory.  A 16-row tile's workgroup then
// takes 14.9 KB of LDS instead of 20 KB.  Occupancy is the same (the wave
// slots, 8 per SIMD, cap both forms at 8 workgroups per CU), but a single
// 1280x720 frame's detect runs about 1.2 us faster with it.
__host__ __device__ constexpr uint32_t detect_src_words_small(uint32_t th) {
    return ((th + 6) * 7u / 4u + 4u) * 33u;
}

struct Job {
    const Plan* plan;
    const TileDesc* tiles;  // the plan's tile descriptors
    const uint8_t* img;
    uint64_t pitch;
    float* pyr;        // nullable
    uint8_t* feat;     // feature array
    uint64_t* surv;    // survivor bitmap scratch, plan->words, then the survivor
                       // count of each detect wave (u32, 4 per tile: surv_counts)
    uint32_t* kp;      // keypoint list scratch, L*cap (x | y << 16)
    float* smooth;     // smoothed-image scratch, plan->sm_floats
    // copies of plan fields, so a describe block finds its level's records
    // without waiting for the plan
    uint32_t desc_groups, cap, feat_stride, pad_;
};

// Survivor counts of the detect waves of a frame: wave w of detect tile t
// (frame-wide tile index, LevelGeom::tile_base + ty * tiles_x + tx) writes
// the survivors of its TH / 4 rows at 4 * t + w, after the survivor words.
__host__ __device__ inline uint32_t* surv_counts(uint64_t* surv, uint32_t words) {
    return reinterpret_cast<uint32_t*>(surv + words);
}

struct MatchJob {
    const Plan* plan;
    const uint8_t* prev;  // feature arrays
    const uint8_t* curr;
    int32_t* prev_match;  // L*cap
    uint32_t* q_list;     // scratch, query_list_words(L, cap): per segment a
                          // count, then entries (b, a): the curr features
                          // b = level << 24 | slot with a direction match a
    // copies of plan fields, so a matcher block reaches its feature counts
    // and query record without waiting for the plan
    uint32_t cap, feat_stride, levels, pad_;
    uint32_t* index_prev; // scratch: cell index of prev, plan->idx_words
    uint32_t* index_curr; // scratch: cell index of curr
};

struct Params {
    float threshold;
    uint32_t min_run;
    int32_t orient_r;
    uint32_t border_zero;
    float ratio;
    uint32_t pad[3];
};

// Batch tables in device memory: jobs and exclusive prefix sums of the
// per-job block counts of each kernel.
struct BatchView {
    const Job* jobs;
    const uint32_t* tile_prefix;  // n+1
    const uint32_t* desc_prefix;  // n+1
    const uint16_t* tile_job;     // job of every detect tile (even-padded)
    const uint16_t* desc_job;     // job of every describe block (even-padded)
    uint32_t n;
};

struct MatchBatchView {
    const MatchJob* jobs;
    const uint32_t* block_prefix;  // n+1
    const uint16_t* block_job;     // job of every matcher block (even-padded)
    uint32_t n;
};

// ---------------------------------------------------------------------------
// Own f32 math: Cody & Waite (1980) reduction by pi/2 with a three-part
// split of pi/2 and Taylor polynomials of sin and cos on [-pi/4, pi/4];
// atan2 through atan on [0, 1], using atan(t) = pi/4 + atan((t-1)/(t+1))
// above tan(pi/8) and the Taylor series of atan below it.  Each constant is
// the f32 rounding of the real number named beside it.
// ---------------------------------------------------------------------------
namespace fm {
constexpr float kTwoOverPi = 0x1.45f306p-1f;  // 2/pi
constexpr float kPio2Hi = 0x1.921fb6p+0f;     // pi/2
constexpr float kPio2Mid = -0x1.777a5cp-25f;  // pi/2 - kPio2Hi
constexpr float kPio2Lo = -0x1.ee59dap-50f;   // pi/2 - kPio2Hi - kPio2Mid
constexpr float kPio4 = 0x1.921fb6p-1f;       // pi/4
constexpr float kPi = 0x1.921fb6p+1f;         // pi
constexpr float kTanPi8 = 0x1.a8279ap-2f;     // tan(pi/8) = sqrt(2) - 1
}  // namespace fm

__host__ __device__ inline float atan_taylor(float t) {
    // sum_{k=0..10} (-1)^k t^(2k+1) / (2k+1), Horner in t^2
    const float t2 = t * t;
    float p = 0x1.861862p-5f;               //  1/21
    p = __builtin_fmaf(p, t2, -0x1.af286cp-5f);  // -1/19
    p = __builtin_fmaf(p, t2, 0x1.e1e1e2p-5f);   //  1/17
    p = __builtin_fmaf(p, t2, -0x1.111112p-4f);  // -1/15
    p = __builtin_fmaf(p, t2, 0x1.3b13b2p-4f);   //  1/13
    p = __builtin_fmaf(p, t2, -0x1.745d18p-4f);  // -1/11
    p = __builtin_fmaf(p, t2, 0x1.c71c72p-4f);   //  1/9
    p = __builtin_fmaf(p, t2, -0x1.24924ap-3f);  // -1/7
    p = __builtin_fmaf(p, t2, 0x1.99999ap-3f);   //  1/5
    p = __builtin_fmaf(p, t2, -0x1.555556p-2f);  // -1/3
    const float t3 = t2 * t;
    return __builtin_fmaf(p, t3, t);
}

__host__ __device__ inline float atan2_own(float y, float x) {
    const float ax = __builtin_fabsf(x);
    const float ay = __builtin_fabsf(y);
    const bool swap = ay > ax;
    const float mx = swap ? ay : ax;
    const float mn = swap ? ax : ay;
    // Both divisions and one polynomial on every lane, no divergent branch
    // (lanes of one wave take either side); each lane's operations are the
    // ones of its side.  mx == 0 gives t = 0 whatever the quotient.
    const float q = mn / (mx == 0.0f ? 1.0f : mx);
    const float t = mx == 0.0f ? 0.0f : q;
    const bool reduce = t > fm::kTanPi8;
    const float u = (t - 1.0f) / (t + 1.0f);
    float a = atan_taylor(reduce ? u : t);
    if (reduce) a = fm::kPio4 + a;
    if (swap) a = fm::kPio2Hi - a;
    if (__builtin_signbit(x)) a = fm::kPi - a;
    if (__builtin_signbit(y)) a = -a;
    return a;
}

__host__ __device__ inline void sincos_own(float x, float* s, float* c) {
    const float k = __builtin_rintf(x * fm::kTwoOverPi);
    float r = __builtin_fmaf(-k, fm::kPio2Hi, x);
    r = __builtin_fmaf(-k, fm::kPio2Mid, r);
    r = __builtin_fmaf(-k, fm::kPio2Lo, r);
    const float r2 = r * r;
    // sin: r - r^3/3! + r^5/5! - r^7/7! + r^9/9!
    float ps = __builtin_fmaf(0x1.71de3ap-19f, r2, -0x1.a01a02p-13f);
    ps = __builtin_fmaf(ps, r2, 0x1.111112p-7f);
    ps = __builtin_fmaf(ps, r2, -0x1.555556p-3f);
    const float r3 = r2 * r;
    const float sn = __builtin_fmaf(ps, r3, r);
    // cos: 1 - r^2/2! + r^4/4! - r^6/6! + r^8/8! - r^10/10!
    float pc = __builtin_fmaf(-0x1.27e4fcp-22f, r2, 0x1.a01a02p-16f);
    pc = __builtin_fmaf(pc, r2, -0x1.6c16c2p-10f);
    pc = __builtin_fmaf(pc, r2, 0x1.555556p-5f);
    pc = __builtin_fmaf(pc, r2, -0x1p-1f);
    const float cs = __builtin_fmaf(pc, r2, 1.0f);
    const int q = static_cast<int>(k) & 3;
    const float s0 = (q & 1) ? cs : sn;
    const float c0 = (q & 1) ? sn : cs;
    *s = (q & 2) ? -s0 : s0;
    *c = ((q + 1) & 2) ? -c0 : c0;
}

// 1 + the longest run of set bits of a 16-bit ring mask (no wrap): the
// `streak` of detect_corners.wgsl:67-79 (curr = 1, streak = 1, then per
// ring test curr = diff ? curr + 1 : 1, streak = max(streak, curr)).
// r_k marks the positions where k set bits start (r_2k = r_k & r_k >> k);
// the run is built greedily from 8, 4, 2 and 1: a run of n bits from i
// followed by one of k bits from i + n is a run of n + k bits.
__host__ __device__ inline uint32_t ring_streak(uint32_t ring) {
    const uint32_t r1 = ring & 0xFFFFu;
    const uint32_t r2 = r1 & (r1 >> 1), r4 = r2 & (r2 >> 2), r8 = r4 & (r4 >> 4);
    uint32_t cur = 0xFFFFu, n = 0;  // cur: starts of runs of n set bits
    const uint32_t rk[4] = {r8, r4, r2, r1};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t c = cur & (rk[j] >> n);
        if (c != 0u) {
            cur = c;
            n += 8u >> j;
        }
    }
    return 1u + (r1 == 0xFFFFu ? 16u : n);
}

// One axis of the linear clamp-to-edge sampler.
struct AxisTaps {
    uint32_t i0, i1;
    float frac;
};

__host__ __device__ inline AxisTaps axis_taps(float coord, uint32_t n) {
    const float u = coord * static_cast<float>(n);
    const float uu = u - 0.5f;
    const float f = __builtin_floorf(uu);
    AxisTaps t;
    t.frac = uu - f;
    const float hi = static_cast<float>(n - 1);
    const float f1 = f + 1.0f;
    const float c0 = f < 0.0f ? 0.0f : (f > hi ? hi : f);
    const float c1 = f1 < 0.0f ? 0.0f : (f1 > hi ? hi : f1);
    t.i0 = static_cast<uint32_t>(c0);
    t.i1 = static_cast<uint32_t>(c1);
    return t;
}

// Tap tables of one level (LevelGeom::tap_off), filled on the host with the
// operations the kernels would otherwise repeat per tile.
__host__ __device__ inline void level_tap_offsets(uint32_t base, uint32_t w, uint32_t h,
                                                  uint32_t* lvl_col, uint32_t* lvl_row,
                                                  uint32_t* sm_col, uint32_t* sm_row) {
    *lvl_col = base;
    *lvl_row = base + w;
    *sm_col = base + w + h;
    *sm_row = base + w + h + 3 * (w + kSmoothExt);
}

__host__ __device__ inline uint32_t level_tap_entries(uint32_t w, uint32_t h) {
    return w + h + 3 * (w + kSmoothExt) + 4 * (h + kSmoothExt);
}

// Entry offset of the smoothing tap codes of a level (after its four
// tables): one per row of the extended grid.
__host__ __device__ inline uint32_t level_row_code_offset(uint32_t base, uint32_t w, uint32_t h) {
    return base + w + h + 3 * (w + kSmoothExt) + 3 * (h + kSmoothExt);
}


__host__ __device__ inline float lerp2(float t00, float t10, float t01, float t11,
                                       float a, float b) {
    const float top = t00 + a * (t10 - t00);
    const float bot = t01 + a * (t11 - t01);
    return top + b * (bot - top);
}

// Tint's textureLoad clamp of one coordinate (int in, index out).
__host__ __device__ inline uint32_t tint_clamp(int32_t c, uint32_t n) {
    const uint32_t u = static_cast<uint32_t>(c);
    return u < n - 1 ? u : n - 1;
}

// Smoothing tap code of row p of a level of n rows, from the table entries
// (i0 | i1 << 16) of its three taps at offsets -1.2, 0, 1.2: the kernels'
// tile staging holds the texel of tint_clamp(g) at staged row g (0 outside
// the level with border_zero), so a tap whose entry equals the pattern for
// sample floor f reads exactly the staged rows f, f + 1.  Bit 0: the floors
// are p - 2, p or p - 1, p + 1 under tint_clamp; bit 1: the middle tap's
// floor is p - 1 (bit 0 set); bits 2, 3: the same with border_zero (floors
// inside the level).
enum : uint32_t { kCodeFast = 1u, kCodeMidLow = 2u, kCodeFastBz = 4u, kCodeMidLowBz = 8u };

__host__ __device__ inline uint32_t smooth_tap_code(uint32_t e0, uint32_t e1, uint32_t e2,
                                                    int32_t p, uint32_t n) {
    const auto pat = [n](int32_t f) { return tint_clamp(f, n) | (tint_clamp(f + 1, n) << 16); };
    const auto pat_bz = [n](int32_t f) -> uint32_t {
        return f >= 0 && f + 1 < static_cast<int32_t>(n)
                   ? static_cast<uint32_t>(f) | (static_cast<uint32_t>(f + 1) << 16)
                   : 0xFFFFFFFFu;  // never a table entry
    };
    uint32_t code = 0;
    const bool a = e1 == pat(p), b = e1 == pat(p - 1);
    if (e0 == pat(p - 2) && e2 == pat(p + 1) && (a || b)) code |= kCodeFast | (a ? 0u : kCodeMidLow);
    const bool az = e1 == pat_bz(p), bzz = e1 == pat_bz(p - 1);
    if (e0 == pat_bz(p - 2) && e2 == pat_bz(p + 1) && (az || bzz))
        code |= kCodeFastBz | (az ? 0u : kCodeMidLowBz);
    return code;
}

}  // namespace orbfe
