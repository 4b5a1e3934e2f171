// HIP kernels of the ORB front-end for gfx950 (CDNA4, wave64).
//
//   k_detect    pyramid levels, smoothed image, FAST-style score and cull
//               mask over one 64 x TH tile per workgroup (TH = 8/16/32),
//               LDS-staged halo; survivors leave as a bitmap built with wave
//               ballots (one 64-bit word per 64-pixel row segment).
//   k_compact   per (frame, level): raster-order prefix sum over the survivor
//               bitmap, first `cap` survivors to a keypoint list, count.
//   k_describe  8/16/32 lanes per keypoint: corner strength, intensity-
//               centroid angle, 256 rotated BRIEF tests looked up in the
//               smoothed image, 48-byte feature record.
//   k_match_index  per (pair, side, level) counting sort of the records by
//               LOD0 cell.
//   k_match_dir pipeline matcher, one wave per query: LOD+-1 pool, candidates
//               from the cells under the search window, XOR/popcount, best/
//               second-best keys reduced across the wave.
//   k_bf_mfma / k_bf_merge  brute-force best/second-best Hamming search:
//               distances as int8 sign-vector dot products on the matrix
//               cores (v_mfma_i32_32x32x32_i8), best two keys per train
//               split, merged across splits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "brief_pattern.h"
#include "orbfe_device.h"

namespace orbfe {

// f32(v) / 255 for every byte v (provider_base.cpp:121-124), correctly
// rounded on the host; workgroups copy it into their LDS lookup tables.
__constant__ float c_u8_unit[256];

// The describe kernel evaluates the 256 BRIEF tests (brief_pattern.h) in two
// groups of 128.  The 512 test points name only 375 distinct pattern
// positions, and a test's two points always rotate together, so each group
// gathers every distinct position of its tests once (at most kDescPts, ids
// past a group's count repeat a position no test reads).  The groups are
// chosen on the host (build_desc_tables) so that every aligned run of 8
// tests puts 4 in each (each lane of a keypoint group evaluates TPL / 2
// tests of either group) and the larger group has as few positions as a
// local search finds: 189 and 189, against 220 and 221 for the tests of
// each parity, so a keypoint gathers 378 values instead of 512.
// A keypoint's sample of point id p goes to slot desc_slot(p) = p + p / 16
// of its LDS buffer (one pad word per 16 against bank conflicts).
//   c_desc_ppt8[s]       pattern positions in slot s of both groups, int8 x
//                        and y of group 0 in bytes 0-1, of group 1 in 2-3
//                        (0 in pad slots);
//   c_desc_tests[g][m]   the group's tests in increasing order: byte offset
//                        of A's slot | byte offset of B's slot << 16 |
//                        (test mod 32) << 26;
//   g_desc_order[v][..]  per lanes-per-keypoint variant v (8, 16, 32), angle
//                        bin, keypoint alignment (x mod 8 | (y mod 4) << 3),
//                        group and lane: the slots of its PPL =
//                        kDescPts / LANES points (u8, in rows of
//                        desc_order_row bytes) in load order.
constexpr uint32_t kDescPts = 192;
constexpr uint32_t kDescSlots = kDescPts + kDescPts / 16;

__host__ __device__ constexpr uint32_t desc_slot(uint32_t p) { return p + p / 16; }

constexpr uint32_t kAngleBins = 64;
constexpr float kAngleBinScale = 10.185916357881302f;  // kAngleBins / (2 pi)
// Keypoint positions modulo the 8 x 4 blocks of the smoothed layout.
constexpr uint32_t kOrderAligns = 32;

// Bytes per lane row of the gather order (PPL rounded up to whole u32).
template <uint32_t LANES>
constexpr uint32_t desc_order_row() {
    return ((kDescPts / LANES) + 3u) & ~3u;
}

constexpr size_t kDescOrderBytes =
    size_t(kAngleBins) * kOrderAligns * 2 * 32 * desc_order_row<32>();  // the largest (1 MiB)
__constant__ uint32_t c_desc_ppt8[kDescSlots];
__constant__ uint32_t c_desc_tests[2][128];
__device__ uint8_t g_desc_order[3][kDescOrderBytes];

template <uint32_t LANES>
__device__ inline const uint8_t* desc_order() {
    return g_desc_order[LANES == 8 ? 0 : (LANES == 16 ? 1 : 2)];
}

namespace {

// detect_corners.wgsl:31-32
constexpr int8_t c_circle_x[16] = {0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1};
constexpr int8_t c_circle_y[16] = {-3, -3, -2, -1, 0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3};

__device__ inline uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ inline uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

// acc + popcount(x) as one v_bcnt_u32_b32 (the compiler would otherwise
// split popcount chains into bcnt(x, 0) and add3 trees).  ALU only.
__device__ inline uint32_t bcnt_acc(uint32_t x, uint32_t acc) {
    uint32_t r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}

// Middle value of three (v_med3_u32).  ALU only.
__device__ inline uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Hamming distance of two 256-bit descriptors (a: 8 words; b: two halves).
__device__ inline uint32_t hamming256(const uint32_t* a, const uint4& lo, const uint4& hi) {
    uint32_t d = bcnt_acc(a[0] ^ lo.x, 0u);
    d = bcnt_acc(a[1] ^ lo.y, d);
    d = bcnt_acc(a[2] ^ lo.z, d);
    d = bcnt_acc(a[3] ^ lo.w, d);
    d = bcnt_acc(a[4] ^ hi.x, d);
    d = bcnt_acc(a[5] ^ hi.y, d);
    d = bcnt_acc(a[6] ^ hi.z, d);
    return bcnt_acc(a[7] ^ hi.w, d);
}

// (min1, min2) <- the two smallest of (min1, min2, key), given min1 <= min2.
__device__ inline void keep_two_smallest(uint32_t& min1, uint32_t& min2, uint32_t key) {
    min2 = med3_u32(min1, min2, key);
    min1 = umin(min1, key);
}

// Orders the LDS accesses of the lanes of one wave around this point.  The
// LDS executes one wave's DS instructions in issue order, so only the
// compiler must be kept from moving memory operations across; global
// stores in flight are not waited for.
__device__ inline void wave_sync() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// Plans, job tables and prefix tables do not change while a kernel runs and
// are read at workgroup-uniform addresses: reading them through the
// constant address space lets the compiler use scalar loads.
#define ORBFE_CONST __attribute__((address_space(4)))

template <typename T>
__device__ inline const ORBFE_CONST T* const_ptr(const T* p) {
    return reinterpret_cast<const ORBFE_CONST T*>(reinterpret_cast<uintptr_t>(p));
}

template <typename T>
__device__ inline const ORBFE_CONST T& as_const(const T* p) {
    return *const_ptr(p);
}

// f32 at byte offset `off` of a global buffer whose base is uniform: the
// load takes the base from SGPRs and a 32-bit offset from one VGPR.
#define ORBFE_GLOBAL __attribute__((address_space(1)))

__device__ inline float load_f32(const uint8_t* base, uint32_t off) {
    const ORBFE_GLOBAL uint8_t* g =
        reinterpret_cast<const ORBFE_GLOBAL uint8_t*>(reinterpret_cast<uintptr_t>(base));
    return *reinterpret_cast<const ORBFE_GLOBAL float*>(g + off);
}

// Stores an f32 at byte offset `off` of a global buffer whose base is
// uniform (SGPR base, one VGPR offset; no 64-bit address per lane), as a
// nontemporal store: the kernels use it for data that later launches read,
// not the storing one.
__device__ inline void store_f32_nt(float* base, uint32_t off, float v) {
    ORBFE_GLOBAL uint8_t* g = reinterpret_cast<ORBFE_GLOBAL uint8_t*>(reinterpret_cast<uintptr_t>(base));
    __builtin_nontemporal_store(v, reinterpret_cast<ORBFE_GLOBAL float*>(g + off));
}

// The global-address-space view of a device pointer.  Accesses through it
// are global_* instructions (uniform base in SGPRs, 32-bit VGPR offsets)
// instead of flat ones, which need 64-bit addresses per lane and also count
// against the counter LDS reads wait on.
template <typename T>
__device__ inline ORBFE_GLOBAL T* gptr(T* p) {
    return reinterpret_cast<ORBFE_GLOBAL T*>(reinterpret_cast<uintptr_t>(p));
}

// Native vector types for loads and stores through the global address space.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16-byte global load / store (native vectors: the HIP vector classes have
// no operators for objects outside the generic address space).
__device__ inline uint4 gload_u4(const ORBFE_GLOBAL uint32_t* p) {
    const u32x4 v = *reinterpret_cast<const ORBFE_GLOBAL u32x4*>(p);
    return make_uint4(v.x, v.y, v.z, v.w);
}
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ inline uint2 gload_u2(const ORBFE_GLOBAL uint32_t* p) {
    const u32x2 v = *reinterpret_cast<const ORBFE_GLOBAL u32x2*>(p);
    return make_uint2(v.x, v.y);
}
__device__ inline void gstore_u4(ORBFE_GLOBAL uint32_t* p, uint4 v) {
    *reinterpret_cast<ORBFE_GLOBAL u32x4*>(p) = u32x4{v.x, v.y, v.z, v.w};
}

// Smoothed values of rows y .. y + RPT - 1 of column x (y a multiple of
// min(RPT, 4)) into a level's smoothed image at float offset `level_off`
// (smooth_band_floats layout): one 16-byte store per band of 4 rows, 8 bytes
// for RPT = 2, for each band whose first row is a grid row (< sh).  Rows of
// such a band past the grid are stored too (the band holds them; nothing
// reads them).  The stores are nontemporal: the describe launch reads the
// image after the whole detect grid (and the compaction) has run, and
// keeping the detect writes out of L2 measured faster for both kernels in
// batches and for detect on single frames.
template <int RPT>
__device__ inline void store_smooth_rows(float* sm, uint32_t level_off, uint32_t band_floats,
                                         uint32_t x, uint32_t y, uint32_t sh,
                                         const float (&v)[RPT]) {
    static_assert(RPT == 2 || RPT % 4 == 0, "whole or half bands");
    ORBFE_GLOBAL uint32_t* g = reinterpret_cast<ORBFE_GLOBAL uint32_t*>(
        reinterpret_cast<uintptr_t>(sm + level_off));
    if constexpr (RPT == 2) {
        if (y < sh)
            __builtin_nontemporal_store(
                u32x2{__float_as_uint(v[0]), __float_as_uint(v[1])},
                reinterpret_cast<ORBFE_GLOBAL u32x2*>(g + (y >> 2) * band_floats + 4 * x + (y & 3u)));
    } else {
#pragma unroll
        for (int b = 0; b < RPT / 4; ++b)
            if (y + 4 * b < sh)
                __builtin_nontemporal_store(
                    u32x4{__float_as_uint(v[4 * b]), __float_as_uint(v[4 * b + 1]),
                          __float_as_uint(v[4 * b + 2]), __float_as_uint(v[4 * b + 3])},
                    reinterpret_cast<ORBFE_GLOBAL u32x4*>(g + ((y >> 2) + b) * band_floats + 4 * x));
    }
}

// Element i of a global array with a uniform base (32-bit offsets).
__device__ inline uint32_t load_u32(const uint32_t* base, uint32_t i) {
    return reinterpret_cast<const ORBFE_GLOBAL uint32_t*>(reinterpret_cast<uintptr_t>(base))[i];
}

__device__ inline TapEntry load_tap(const TapEntry* t, uint32_t i) {
    const ORBFE_GLOBAL uint32_t* g =
        reinterpret_cast<const ORBFE_GLOBAL uint32_t*>(reinterpret_cast<uintptr_t>(t)) + 2 * i;
    return TapEntry{g[0], __uint_as_float(g[1])};
}

// Copy of a table entry, dword by dword (the loads stay scalar).
template <typename T>
__device__ inline T copy_const(const ORBFE_CONST T* p) {
    static_assert(sizeof(T) % 4 == 0, "dword-sized entries");
    T out;
    const ORBFE_CONST uint32_t* src = reinterpret_cast<const ORBFE_CONST uint32_t*>(p);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; ++i) dst[i] = src[i];
    return out;
}

using ConstPlan = const ORBFE_CONST Plan&;

// Job of block b from the batch's u16 block-job table (one scalar load).
__device__ inline uint32_t job_of(const uint16_t* table, uint32_t b) {
    const uint32_t w = as_const(reinterpret_cast<const uint32_t*>(table) + (b >> 1));
    return (w >> ((b & 1u) * 16u)) & 0xFFFFu;
}

// Level-l texel (x, y) computed from the u8 frame: level 0 is float(u8)/255
// (data/src/provider_base.cpp:121-124); level l >= 1 is the bilinear sample of
// level 0 at the texel centre (create_pyramid.wgsl:5-14).  x, y in range.
__device__ inline float level_texel_from_frame(const Job& J, ConstPlan P,
                                               const LevelGeom& G, uint32_t l,
                                               const float* lut, uint32_t x,
                                               uint32_t y) {
    const ORBFE_GLOBAL uint8_t* img = gptr(J.img);
    if (l == 0) return lut[img[y * J.pitch + x]];
    const float u = (static_cast<float>(x) + 0.5f) / static_cast<float>(G.w);
    const float v = (static_cast<float>(y) + 0.5f) / static_cast<float>(G.h);
    const AxisTaps tx = axis_taps(u, P.W0);
    const AxisTaps ty = axis_taps(v, P.H0);
    const ORBFE_GLOBAL uint8_t* r0 = img + ty.i0 * J.pitch;
    const ORBFE_GLOBAL uint8_t* r1 = img + ty.i1 * J.pitch;
    return lerp2(lut[r0[tx.i0]], lut[r0[tx.i1]], lut[r1[tx.i0]], lut[r1[tx.i1]],
                 tx.frac, ty.frac);
}

// float(u8) / 255, the value the LDS table of the kernels holds.
__device__ inline float u8_unit(uint32_t v) { return static_cast<float>(v) / 255.0f; }

// level_texel_from_frame without the LDS table (usable before it is filled).
__device__ inline float level_texel_direct(const Job& J, ConstPlan P, const LevelGeom& G,
                                           uint32_t l, uint32_t x, uint32_t y) {
    const ORBFE_GLOBAL uint8_t* img = gptr(J.img);
    if (l == 0) return u8_unit(img[y * J.pitch + x]);
    const float u = (static_cast<float>(x) + 0.5f) / static_cast<float>(G.w);
    const float v = (static_cast<float>(y) + 0.5f) / static_cast<float>(G.h);
    const AxisTaps tx = axis_taps(u, P.W0);
    const AxisTaps ty = axis_taps(v, P.H0);
    const ORBFE_GLOBAL uint8_t* r0 = img + ty.i0 * J.pitch;
    const ORBFE_GLOBAL uint8_t* r1 = img + ty.i1 * J.pitch;
    return lerp2(u8_unit(r0[tx.i0]), u8_unit(r0[tx.i1]), u8_unit(r1[tx.i0]),
                 u8_unit(r1[tx.i1]), tx.frac, ty.frac);
}

}  // namespace

// ---------------------------------------------------------------------------
// k_detect
// ---------------------------------------------------------------------------
// One LDS buffer holds first the u8 source region of a level >= 1 tile and
// then, once the level tile is staged, the corner scores and candidate lists
// of the FAST pass.  At level scales up to 2.5 (six levels of 1.2) the region
// spans at most (TH+6)*2.5 + 4 rows of 45 dwords (70 texels * 2.5 plus the
// bilinear neighbour and the dword alignment): 7.0 KB, 10.6 KB, 17.8 KB for
// TH = 8, 16, 32.  Regions that do not fit (levels past the sixth; with the
// small buffer of single-frame launches, levels past the fourth) fall back to
// per-texel gathers.
template <int TH, bool kSmallSrc>
constexpr uint32_t stage_words() {
    return kSmallSrc ? detect_src_words_small(TH) : detect_src_words(TH);
}

// Byte i of the staged source region, read as a dword and shifted, so the
// load stays a 32-bit LDS read.
__device__ inline uint32_t src_byte(const uint32_t* src, uint32_t i) {
    return (src[i >> 2] >> ((i & 3u) * 8u)) & 0xFFu;
}

// Smoothed value at (x, y) evaluated directly from the frame (tiles whose
// sampler taps fall outside the staged region: only at the far corner of
// the extended grid).
__device__ float smooth_direct(const Job& J, ConstPlan P, const LevelGeom& G, uint32_t l,
                               const float* lut, uint32_t x, uint32_t y) {
    const float offs[3] = {-1.2f, 0.0f, 1.2f};
    const float wts[3] = {0.3125f, 0.375f, 0.3125f};
    const float texel_x = 1.0f / static_cast<float>(G.w);
    const float texel_y = 1.0f / static_cast<float>(G.h);
    const float uvx = (static_cast<float>(x) + 0.5f) * texel_x;
    const float uvy = (static_cast<float>(y) + 0.5f) * texel_y;
    float sum = 0.0f;
    for (int jj = 0; jj < 3; ++jj) {
        const AxisTaps ay = axis_taps(uvy + offs[jj] * texel_y, G.h);
        for (int i = 0; i < 3; ++i) {
            const AxisTaps ax = axis_taps(uvx + offs[i] * texel_x, G.w);
            const float sv = lerp2(level_texel_from_frame(J, P, G, l, lut, ax.i0, ay.i0),
                                   level_texel_from_frame(J, P, G, l, lut, ax.i1, ay.i0),
                                   level_texel_from_frame(J, P, G, l, lut, ax.i0, ay.i1),
                                   level_texel_from_frame(J, P, G, l, lut, ax.i1, ay.i1),
                                   ax.frac, ay.frac);
            sum = sum + sv * wts[i] * wts[jj];
        }
    }
    return sum;
}

// Everything a workgroup needs to know about one detect tile, from scalar
// loads: the frame's job, the tile's descriptor (level, position and the
// source rectangle of level 0 its staged texels sample, precomputed with the
// plan), the level geometry and the tap table offsets.  The source row loads
// need only the job and the descriptor, so they are issued while the level
// geometry is still loading.
struct DetectTile {
    Job J;
    LevelGeom G;
    uint32_t l, tile_x, tile_y;
    int32_t x0, y0;
    uint32_t t_lc, t_lr, t_sc, t_sr;
    uint32_t c_first, r_first, c_last, r_last;
    uint32_t rowmin, cs, nd, nrows;
    bool use_src;
};

template <int TH, bool kSmallSrc>
__device__ inline DetectTile decode_tile(const BatchView& bv, const Job& job0, const Params& prm,
                                         uint32_t b) {
    constexpr int LH = TH + 2 * kHalo;
    DetectTile T;
    uint32_t j = 0;
    if (bv.n) {
        j = job_of(bv.tile_job, b);
        b -= as_const(bv.tile_prefix + j);
    }
    T.J = bv.n ? copy_const(const_ptr(bv.jobs + j)) : job0;
    const TileDesc td = copy_const(const_ptr(T.J.tiles + b));
    const uint32_t l = td.lxy & 0xFu;
    T.l = l;
    T.tile_x = (td.lxy >> 4) & 0xFFFu;
    T.tile_y = td.lxy >> 16;
    T.rowmin = td.rowmin;
    T.cs = td.cs;
    T.nd = td.nd_rows & 0xFFFFu;
    T.nrows = (td.nd_rows >> 16) & 0x3FFFu;
    T.use_src = ((td.nd_rows >> (kSmallSrc ? 30 : 31)) & 1u) != 0 && (T.J.pitch & 3u) == 0 &&
                (reinterpret_cast<uintptr_t>(T.J.img) & 3u) == 0;
    ConstPlan P = as_const(T.J.plan);
    T.G = copy_const(&P.lv[l]);  // by value: stores cannot alias it
    const LevelGeom& G = T.G;
    const int32_t x0 = static_cast<int32_t>(T.tile_x) * kTileW;
    const int32_t y0 = static_cast<int32_t>(T.tile_y) * TH;
    T.x0 = x0;
    T.y0 = y0;
    const uint32_t w = G.w, h = G.h;
    level_tap_offsets(G.tap_off, w, h, &T.t_lc, &T.t_lr, &T.t_sc, &T.t_sr);
    T.c_first = x0 >= kHalo ? 0u : static_cast<uint32_t>(kHalo - x0);
    T.r_first = y0 >= kHalo ? 0u : static_cast<uint32_t>(kHalo - y0);
    T.c_last = kLdsW - 1;
    T.r_last = LH - 1;
    if (prm.border_zero) {
        const int32_t cl_ = static_cast<int32_t>(w) - 1 - x0 + kHalo;
        const int32_t rl_ = static_cast<int32_t>(h) - 1 - y0 + kHalo;
        T.c_last = cl_ < kLdsW - 1 ? static_cast<uint32_t>(cl_ < 0 ? 0 : cl_) : kLdsW - 1;
        T.r_last = rl_ < LH - 1 ? static_cast<uint32_t>(rl_ < 0 ? 0 : rl_) : LH - 1;
    }
    return T;
}

// Set-up items of a tile, one per thread: e < kLdsW column taps, then LH row
// taps, then 3 x TH row taps of the 9-tap Gaussian, then TH cull-mask words
// (row-0 neighbour texels of row-0 tiles follow; they are not prefetched).
template <int TH>
struct SetupItems {
    static constexpr uint32_t kTap = kLdsW + TH + 2 * kHalo;
    static constexpr uint32_t kSrow = kTap + 3 * TH;
    static constexpr uint32_t kMask = kSrow + TH;
};

// Direct-to-LDS loads (global_load_lds_dword) of a tile's u8 source
// rectangle: dword e of the row-major rectangle (nd dwords per row) goes to
// src[e], lane e % 64 of round e / 64, rounds dealt to the waves.  They run
// in the background; the consumer waits for them (vmcnt) and a barrier.
// Tiles whose rectangle does not fit the buffer load nothing (their texels
// are read from the frame directly).
template <int TH>
__device__ inline void issue_source(const DetectTile& T, uint32_t* src, uint32_t tid) {
    if (!T.use_src) return;  // block-uniform
    constexpr uint32_t kWaves = kDetectThreads / 64;
    const uint32_t lane = tid & 63, wave = tid >> 6;
    const uint32_t total = T.nd * T.nrows;  // <= the buffer's words (the tile's fit bit)
    // row = floor(e / nd) through a reciprocal: e < 2^13 and nd <= 45 keep
    // (e + 0.5) / nd at least 1/90 from an integer, far above the relative
    // rounding error (~2^-23) of the product, so the floor is exact.
    const float inv_nd = 1.0f / static_cast<float>(T.nd);
    const ORBFE_GLOBAL uint8_t* img = reinterpret_cast<const ORBFE_GLOBAL uint8_t*>(
        reinterpret_cast<uintptr_t>(T.J.img));
    for (uint32_t r = wave; r * 64 < total; r += kWaves) {
        const uint32_t e = r * 64 + lane;
        if (e < total) {
            const uint32_t row = static_cast<uint32_t>((static_cast<float>(e) + 0.5f) * inv_nd);
            const uint32_t col = e - row * T.nd;
            const ORBFE_GLOBAL uint8_t* g =
                img + (static_cast<uint64_t>(T.rowmin + row) * T.J.pitch + T.cs + 4u * col);
            __builtin_amdgcn_global_load_lds(
                reinterpret_cast<const ORBFE_GLOBAL void*>(g),
                reinterpret_cast<__attribute__((address_space(3))) void*>(
                    reinterpret_cast<uintptr_t>(src + r * 64)),
                4, 0, 0);
        }
    }
}

// Waits for this wave's direct-to-LDS loads.
__device__ inline void wait_lds_dma() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// The set-up item e = tid of a tile (raw dwords): a column or row level tap,
// a row tap of the 9-tap Gaussian or a cull-mask word.
template <int TH>
__device__ inline uint2 setup_item(const DetectTile& T, const Params& prm, uint32_t tid) {
    using SI = SetupItems<TH>;
    ConstPlan P = as_const(T.J.plan);
    const uint32_t e = tid;
    const uint32_t w = T.G.w, h = T.G.h;
    // Every item is one 8-byte entry (a tap entry or a cull-mask word) or a
    // value computed here; the branches below only pick the entry, and one
    // load per thread follows (entry 0 of the tap table for threads with
    // nothing to load), so a wave holding several kinds of item makes one
    // round trip, not one per kind.
    const ORBFE_GLOBAL uint2* base = reinterpret_cast<const ORBFE_GLOBAL uint2*>(gptr(P.taps));
    uint32_t idx = 0;
    bool load = false;
    uint2 item = make_uint2(0u, 0u);
    if (e < SI::kTap) {
        const bool is_col = e < kLdsW;
        const uint32_t c = is_col ? e : e - kLdsW;
        const int32_t g = (is_col ? T.x0 : T.y0) - kHalo + static_cast<int32_t>(c);
        const uint32_t n = is_col ? w : h;
        const bool inside = g >= 0 && g < static_cast<int32_t>(n);
        const uint32_t cl = prm.border_zero ? (inside ? static_cast<uint32_t>(g) : 0u)
                                            : tint_clamp(g, n);
        if (T.l > 0) {
            idx = (is_col ? T.t_lc : T.t_lr) + cl;
            load = true;
        } else {
            item = make_uint2(cl | (cl << 16), 0u);
        }
    } else if (e < SI::kSrow) {
        const uint32_t q = e - SI::kTap;
        const uint32_t jj = q / TH, rr = q - jj * TH;
        const uint32_t y = static_cast<uint32_t>(T.y0) + rr;
        const uint32_t sh = h + kSmoothExt;
        if (y < sh) {  // rows past the extended grid are not smoothed
            idx = T.t_sr + jj * sh + y;
            load = true;
        }
    } else if (e < SI::kMask) {
        const uint32_t rr = e - SI::kSrow;
        const uint32_t y = static_cast<uint32_t>(T.y0) + rr;
        if (T.tile_x < T.G.wpr && y < h) {
            base = reinterpret_cast<const ORBFE_GLOBAL uint2*>(gptr(P.mask));
            idx = T.G.word_off + y * T.G.wpr + T.tile_x;
            load = true;
        }
    }
    const uint2 v = gload_u2(reinterpret_cast<const ORBFE_GLOBAL uint32_t*>(base + idx));
    return load ? v : item;
}

// One workgroup per tile.  kSmallSrc: the small LDS source buffer
// (single-frame launches).
template <int TH, bool kSmallSrc = false>
__global__ __launch_bounds__(kDetectThreads) void k_detect(BatchView bv, Job job0,
                                                            Params prm) {
    constexpr int LH = TH + 2 * kHalo;                     // staged rows
    constexpr int RPT = TH / (kDetectThreads / 64);        // rows per thread
    constexpr uint32_t kSrcWords = stage_words<TH, kSmallSrc>();
    constexpr uint32_t kWaves = kDetectThreads / 64;
    using SI = SetupItems<TH>;
    static_assert(SI::kMask <= kDetectThreads, "one set-up item per thread");
    __shared__ float lut[256];
    __shared__ float tile[LH][kLdsW];
    // u8 source region while staging, then the corner scores and candidate
    // lists of the FAST pass (per wave RPT x 64 bytes and RPT x 64 u16).
    __shared__ uint32_t stage[kSrcWords];
    static_assert(kDetectThreads / 64 * RPT * 64 * 3 <= kSrcWords * 4, "FAST scratch fits");
    uint32_t* const src = stage;
    // taps as i0 | i1 << 16 and fraction
    __shared__ uint32_t col_i01[kLdsW];
    __shared__ float col_a[kLdsW];
    __shared__ uint32_t row_i01[LH];
    __shared__ float row_b[LH];
    __shared__ uint32_t col_ok[kLdsW], row_ok[LH];
    __shared__ uint32_t srow_i01[3][TH];
    __shared__ float srow_b[3][TH];
    __shared__ uint32_t row0_score[kTileW + 4];
    __shared__ float row0_ring[4][17];      // row-0 neighbours: 16 ring texels + centre
    __shared__ uint64_t mask_lds[TH];       // cull-mask word of each tile row

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const float wts[3] = {0.3125f, 0.375f, 0.3125f};
    for (uint32_t i = tid; i < 256; i += kDetectThreads) lut[i] = c_u8_unit[i];

    const DetectTile T =
        decode_tile<TH, kSmallSrc>(bv, job0, prm,
                                   bv.n ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x);
    issue_source<TH>(T, src, tid);
    const uint2 item = setup_item<TH>(T, prm, tid);
    const Job& J = T.J;
    ConstPlan P = as_const(J.plan);
    const LevelGeom& G = T.G;
    const uint32_t l = T.l, tile_x = T.tile_x;
    const int32_t x0 = T.x0, y0 = T.y0;
    const uint32_t w = G.w, h = G.h;
    const uint32_t sw = w + kSmoothExt, sh = h + kSmoothExt;
    const uint32_t t_sc = T.t_sc;
    const TapEntry* taps = P.taps;
    const uint32_t c_first = T.c_first, r_first = T.r_first;
    const uint32_t c_last = T.c_last, r_last = T.r_last;
    const uint32_t rowmin = T.rowmin, cs = T.cs, nd = T.nd;
    const bool use_src = T.use_src;

    // Set-up items and the row-0 neighbour texels into LDS.
    const bool has_row0 = y0 == 0 && tile_x < G.wpr;
    {
        const uint32_t e = tid;
        if (e < SI::kTap) {
            const bool is_col = e < kLdsW;
            const uint32_t c = is_col ? e : e - kLdsW;
            const int32_t g = (is_col ? x0 : y0) - kHalo + static_cast<int32_t>(c);
            const uint32_t n = is_col ? w : h;
            const bool inside = g >= 0 && g < static_cast<int32_t>(n);
            const float frac = __uint_as_float(item.y);
            const uint32_t ok = (!prm.border_zero || inside) ? 1u : 0u;
            if (is_col) {
                col_i01[c] = item.x; col_a[c] = frac; col_ok[c] = ok;
            } else {
                row_i01[c] = item.x; row_b[c] = frac; row_ok[c] = ok;
            }
        } else if (e < SI::kSrow) {
            const uint32_t q = e - SI::kTap;
            const uint32_t jj = q / TH, rr = q - jj * TH;
            srow_i01[jj][rr] = item.x;
            srow_b[jj][rr] = __uint_as_float(item.y);
        } else if (e < SI::kMask) {
            mask_lds[e - SI::kSrow] = static_cast<uint64_t>(item.x) |
                                      (static_cast<uint64_t>(item.y) << 32);
        }
    }
    if (has_row0) {
        for (uint32_t q = tid; q < 4 * 17; q += kDetectThreads) {
            const uint32_t k = q / 17, i = q - k * 17;
            const int32_t xe = k < 2 ? x0 - 2 + static_cast<int32_t>(k)
                                     : x0 + kTileW + static_cast<int32_t>(k) - 2;
            float v = 0.0f;
            if (xe >= 0 && xe < static_cast<int32_t>(w)) {
                if (i == 16) {
                    v = level_texel_direct(J, P, G, l, static_cast<uint32_t>(xe), 0);
                } else {
                    const int32_t gx = xe + c_circle_x[i], gy = c_circle_y[i];
                    if (prm.border_zero) {
                        if (gx >= 0 && gy >= 0 && gx < static_cast<int32_t>(w) &&
                            gy < static_cast<int32_t>(h))
                            v = level_texel_direct(J, P, G, l, gx, gy);
                    } else {
                        v = level_texel_direct(J, P, G, l, tint_clamp(gx, w), tint_clamp(gy, h));
                    }
                }
            }
            row0_ring[k][i] = v;
        }
    }
    // The source rectangle's direct-to-LDS loads land: this wave's, then
    // everybody's.
    wait_lds_dma();
    __syncthreads();
    const uint32_t spitch = nd * 4;
    const uint8_t* srcb = reinterpret_cast<const uint8_t*>(src);

    // One texel of the level tile from its row and column taps: level 0 is
    // the LUT of the source byte, levels >= 1 the bilinear sample of level 0
    // (create_pyramid.wgsl:5-14).  Source bytes come from the staged region
    // (byte offsets >= 0 inside it) or, outside it, from the frame.
    const auto texel = [&](uint32_t ri0, uint32_t ri1, float rb, bool row_in, uint32_t ci0,
                           uint32_t ci1, float ca, bool col_in) -> float {
        if (row_in && col_in) {
            const uint32_t o0 = (ri0 - rowmin) * spitch, a0 = ci0 - cs;
            if (l == 0) return lut[srcb[o0 + a0]];
            const uint32_t o1 = (ri1 - rowmin) * spitch, a1 = ci1 - cs;
            return lerp2(lut[srcb[o0 + a0]], lut[srcb[o0 + a1]], lut[srcb[o1 + a0]],
                         lut[srcb[o1 + a1]], ca, rb);
        }
        const ORBFE_GLOBAL uint8_t* s0 = gptr(J.img) + ri0 * J.pitch;
        if (l == 0) return lut[s0[ci0]];
        const ORBFE_GLOBAL uint8_t* s1 = gptr(J.img) + ri1 * J.pitch;
        return lerp2(lut[s0[ci0]], lut[s0[ci1]], lut[s1[ci0]], lut[s1[ci1]], ca, rb);
    };

    // Stage the level tile ((TH + 6) x 70 texels).  Lane c stages column c
    // of rows wave, wave + 4, ... with its column taps in registers and the
    // row taps wave-uniform; columns 64..69 follow, one texel per thread.
    // Tiles whose every texel lies in the staged source region (no
    // border_zero, not at the level's top or left edge) take a branch-free
    // form: the source bytes of all the lane's rows are read first, then
    // their LUT values, then the lerps.
    constexpr uint32_t kStageRows = (LH + kWaves - 1) / kWaves;
    const bool stage_fast = use_src && !prm.border_zero && c_first == 0 && r_first == 0 &&
                            c_last == kLdsW - 1 && r_last == LH - 1;
    if (stage_fast) {
        const uint32_t c = lane;
        const uint32_t ci01 = col_i01[c];
        const uint32_t a0 = (ci01 & 0xFFFFu) - cs, a1 = (ci01 >> 16) - cs;
        const float ca = col_a[c];
        uint32_t o0[kStageRows], o1[kStageRows];
        float rb[kStageRows];
#pragma unroll
        for (uint32_t k = 0; k < kStageRows; ++k) {
            const uint32_t r = umin(wave + k * kWaves, LH - 1);  // rows past LH: recomputed, not stored
            const uint32_t e = __builtin_amdgcn_readfirstlane(row_i01[r]);
            o0[k] = ((e & 0xFFFFu) - rowmin) * spitch;
            o1[k] = ((e >> 16) - rowmin) * spitch;
            rb[k] = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(row_b[r])));
        }
        if (l == 0) {
            uint32_t b[kStageRows];
#pragma unroll
            for (uint32_t k = 0; k < kStageRows; ++k) b[k] = srcb[o0[k] + a0];
#pragma unroll
            for (uint32_t k = 0; k < kStageRows; ++k) {
                const uint32_t r = wave + k * kWaves;
                if (r < LH) tile[r][c] = lut[b[k]];
            }
        } else {
            uint32_t b00[kStageRows], b10[kStageRows], b01[kStageRows], b11[kStageRows];
#pragma unroll
            for (uint32_t k = 0; k < kStageRows; ++k) {
                b00[k] = srcb[o0[k] + a0];
                b10[k] = srcb[o0[k] + a1];
                b01[k] = srcb[o1[k] + a0];
                b11[k] = srcb[o1[k] + a1];
            }
            float t[kStageRows];
#pragma unroll
            for (uint32_t k = 0; k < kStageRows; ++k)
                t[k] = lerp2(lut[b00[k]], lut[b10[k]], lut[b01[k]], lut[b11[k]], ca, rb[k]);
#pragma unroll
            for (uint32_t k = 0; k < kStageRows; ++k) {
                const uint32_t r = wave + k * kWaves;
                if (r < LH) tile[r][c] = t[k];
            }
        }
        constexpr uint32_t kExtra = (kLdsW - kTileW) * LH;
        static_assert(kExtra <= kDetectThreads, "one extra texel per thread");
        if (tid < kExtra) {
            const uint32_t r = tid / (kLdsW - kTileW);
            const uint32_t ce = kTileW + (tid - r * (kLdsW - kTileW));
            const uint32_t ri01 = row_i01[r], ce01 = col_i01[ce];
            tile[r][ce] = texel(ri01 & 0xFFFFu, ri01 >> 16, row_b[r], true, ce01 & 0xFFFFu,
                                ce01 >> 16, col_a[ce], true);
        }
    } else {
        const uint32_t c = lane;
        const uint32_t ci01 = col_i01[c];
        const uint32_t ci0 = ci01 & 0xFFFFu, ci1 = ci01 >> 16;
        const float ca = col_a[c];
        const bool cok = col_ok[c] != 0;
        const bool cin = use_src && c >= c_first && c <= c_last;
#pragma unroll
        for (uint32_t k = 0; k < (LH + kWaves - 1) / kWaves; ++k) {
            const uint32_t r = wave + k * kWaves;
            if (r < LH) {
                const uint32_t ri01 = __builtin_amdgcn_readfirstlane(row_i01[r]);
                const uint32_t ri0 = ri01 & 0xFFFFu, ri1 = ri01 >> 16;
                const float rb = __uint_as_float(
                    __builtin_amdgcn_readfirstlane(__float_as_uint(row_b[r])));
                const bool rok = __builtin_amdgcn_readfirstlane(row_ok[r]) != 0;
                const bool rin = use_src && r >= r_first && r <= r_last;
                tile[r][c] = (cok && rok) ? texel(ri0, ri1, rb, rin, ci0, ci1, ca, cin) : 0.0f;
            }
        }
        constexpr uint32_t kExtra = (kLdsW - kTileW) * LH;
        static_assert(kExtra <= kDetectThreads, "one extra texel per thread");
        if (tid < kExtra) {
            const uint32_t r = tid / (kLdsW - kTileW);
            const uint32_t ce = kTileW + (tid - r * (kLdsW - kTileW));
            const bool in = use_src && ce >= c_first && ce <= c_last && r >= r_first &&
                            r <= r_last;
            const uint32_t ri01 = row_i01[r], ci01 = col_i01[ce];
            tile[r][ce] = (col_ok[ce] & row_ok[r])
                              ? texel(ri01 & 0xFFFFu, ri01 >> 16, row_b[r], in, ci01 & 0xFFFFu,
                                      ci01 >> 16, col_a[ce], in)
                              : 0.0f;
        }
    }
    __syncthreads();

    const uint32_t x = static_cast<uint32_t>(x0) + lane;
    // the wave's first output row (a multiple of RPT) and the level's band
    // stride (smooth_band_floats)
    const uint32_t y_wave = static_cast<uint32_t>(y0) + wave * RPT;
    const uint32_t band_floats = smooth_band_floats(w);

    // Smoothed image over the extended grid (generate_features.wgsl:149-172
    // at every integer point): per point, 3 x 3 bilinear samples of the
    // level at offsets -1.2, 0, 1.2 texels, weighted, in the shader's order.
    // A sample is the vertical lerp of two horizontal lerps of staged texels.
    //
    // Staged column (row) g - x0 + kHalo holds the level texel at
    // tint_clamp(g) (0 outside the level when border_zero).  A sampler tap
    // names texels inside the level (clamp to edge), and for such a texel g
    // that staged column holds g itself, so each lane reads its three
    // column taps at the staged columns their table entries name, at the
    // level's edges too.  Rows: where the plan's row tap codes
    // (smooth_tap_code) confirm the floors y - 2, y or y - 1, and y + 1 for
    // all of a wave's output rows (all but the first two rows of a level),
    // the horizontal lerps of its rows y - 2 .. y + RPT + 1 sit in a
    // register window shared by its RPT outputs; other waves evaluate each
    // sample from the staged texels the tables name.  A middle row tap with
    // fraction 0 (most rows) yields its first operand exactly (no value here
    // is -0, infinite or NaN), so its lerp is skipped.
    const bool slow = x0 > static_cast<int32_t>(w) + 2 || y0 > static_cast<int32_t>(h) + 2;
    if (!slow) {
        if (x < sw) {
            const uint32_t t_rc = level_row_code_offset(G.tap_off, w, h);
            TapEntry tp[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) tp[i] = load_tap(taps, t_sc + i * sw + x);
            const bool bz = prm.border_zero != 0;
            const uint32_t fast_bit = bz ? kCodeFastBz : kCodeFast;
            const uint32_t low_bit = bz ? kCodeMidLowBz : kCodeMidLow;
            // Staged columns of the lane's column taps: a tap names level
            // texels g in [0, w) (the sampler clamps to the edge), and staged
            // column g - x0 + kHalo holds texel g for every such g the tile's
            // halo reaches, in both border modes, at the level's edges too.
            uint32_t ca[3], cb[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                ca[i] = (tp[i].i01 & 0xFFFFu) - x0 + kHalo;
                cb[i] = (tp[i].i01 >> 16) - x0 + kHalo;
            }
            // wave-uniform (scalar loads): rows of this wave's outputs
            bool row_fast = true;
            uint32_t row_b1 = 0;  // bit k: middle tap of output k has floor y - 1
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const uint32_t y = static_cast<uint32_t>(y0) + wave * RPT + k;
                const uint32_t rcode = copy_const(const_ptr(taps + t_rc + umin(y, sh - 1))).i01;
                if (y >= sh) continue;
                row_fast = row_fast && (rcode & fast_bit) != 0;
                if (rcode & low_bit) row_b1 |= 1u << k;
            }
            const float fr0 = tp[0].frac, fr1 = tp[1].frac, fr2 = tp[2].frac;
            if (row_fast) {  // wave-uniform
                const float frs[3] = {fr0, fr1, fr2};
                float hv[3][RPT + 4];
                float out[RPT];
#pragma unroll
                for (int m = 0; m < RPT + 4; ++m) {
                    const float* tr = tile[wave * RPT + 1 + m];  // row y - 2 + m
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        const float t0 = tr[ca[i]];
                        hv[i][m] = t0 + frs[i] * (tr[cb[i]] - t0);
                    }
                }
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const uint32_t rr = wave * RPT + k;
                    const uint32_t y = static_cast<uint32_t>(y0) + rr;
                    out[k] = 0.0f;
                    if (y >= sh) continue;
                    const bool b1 = (row_b1 >> k) & 1u;
                    // The shader's sum starts at 0; every term is >= +0 (a
                    // lerp of non-negative texels times positive weights,
                    // never -0), so 0 + t is t and the first term starts it.
                    float sum = 0.0f;
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj) {
                        const float bb = __uint_as_float(
                            __builtin_amdgcn_readfirstlane(__float_as_uint(srow_b[jj][rr])));
                        if (jj == 1 && bb == 0.0f) {  // wave-uniform
#pragma unroll
                            for (int i = 0; i < 3; ++i) {
                                const float top = b1 ? hv[i][k + 1] : hv[i][k + 2];
                                sum = sum + top * wts[i] * wts[jj];
                            }
                            continue;
                        }
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            float top, bot;
                            if (jj == 0) {
                                top = hv[i][k];
                                bot = hv[i][k + 1];
                            } else if (jj == 1) {
                                top = b1 ? hv[i][k + 1] : hv[i][k + 2];
                                bot = b1 ? hv[i][k + 2] : hv[i][k + 3];
                            } else {
                                top = hv[i][k + 3];
                                bot = hv[i][k + 4];
                            }
                            const float sv = top + bb * (bot - top);
                            const float t = sv * wts[i] * wts[jj];
                            sum = jj == 0 && i == 0 ? t : sum + t;  // 0 + t == t
                        }
                    }
                    out[k] = sum;
                }
                store_smooth_rows<RPT>(J.smooth, G.sm_off, band_floats, x, y_wave, sh, out);
            } else {
                const float frs[3] = {fr0, fr1, fr2};
                float out[RPT];
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const uint32_t rr = wave * RPT + k;
                    const uint32_t y = static_cast<uint32_t>(y0) + rr;
                    out[k] = 0.0f;
                    if (y >= sh) continue;
                    float sum = 0.0f;
#pragma unroll
                    for (int jj = 0; jj < 3; ++jj) {
                        const uint32_t e = __builtin_amdgcn_readfirstlane(srow_i01[jj][rr]);
                        const uint32_t ra = (e & 0xFFFFu) - y0 + kHalo;
                        const uint32_t rb = (e >> 16) - y0 + kHalo;
                        const float bb = __uint_as_float(
                            __builtin_amdgcn_readfirstlane(__float_as_uint(srow_b[jj][rr])));
#pragma unroll
                        for (int i = 0; i < 3; ++i) {
                            const uint32_t c0 = (tp[i].i01 & 0xFFFFu) - x0 + kHalo;
                            const uint32_t c1 = (tp[i].i01 >> 16) - x0 + kHalo;
                            const float t0 = tile[ra][c0];
                            const float top = t0 + frs[i] * (tile[ra][c1] - t0);
                            const float t1 = tile[rb][c0];
                            const float bot = t1 + frs[i] * (tile[rb][c1] - t1);
                            const float sv = top + bb * (bot - top);
                            const float t = sv * wts[i] * wts[jj];
                            sum = jj == 0 && i == 0 ? t : sum + t;  // 0 + t == t, as above
                        }
                    }
                    out[k] = sum;
                }
                store_smooth_rows<RPT>(J.smooth, G.sm_off, band_floats, x, y_wave, sh, out);
            }
        }
    } else if (x < sw) {
        for (int k = 0; k < RPT; ++k) {
            const uint32_t y = y_wave + k;
            if (y < sh)
                store_f32_nt(J.smooth, 4 * (G.sm_off + (y >> 2) * band_floats + 4 * x + (y & 3u)),
                             smooth_direct(J, P, G, l, lut, x, y));
        }
    }

    // Corner scores (detect_corners.wgsl:58-81) for the pixels the cull can
    // keep: masked pixels of rows >= 1 and the 9-aligned windows of row 0.
    // A run of >= 9 ring differences without wrap covers ring positions 7
    // and 8, so for min_run >= 9 a pixel failing either of those two tests
    // scores 0; the rest are gathered into a per-wave list and get the full
    // ring.  Scores land in score_lds (0 elsewhere).
    // The source region is dead since the tile staging: its words become
    // the per-wave FAST scratch.
    uint8_t* wscore = reinterpret_cast<uint8_t*>(stage) + wave * RPT * 64;
    uint16_t* wlist = reinterpret_cast<uint16_t*>(reinterpret_cast<uint8_t*>(stage) +
                                                  kDetectThreads * RPT) + wave * RPT * 64;
    const bool quick = prm.min_run >= 9;
    // The texels and mask words of all RPT rows are read first (every index
    // lies in the staged tile), then each row's test is formed without
    // branches.
    float cen[RPT], q7[RPT], q8[RPT];
    uint64_t mrow[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const uint32_t r = wave * RPT + k;
        cen[k] = tile[r + kHalo][lane + kHalo];
        q7[k] = tile[r + kHalo + c_circle_y[7]][lane + kHalo + c_circle_x[7]];
        q8[k] = tile[r + kHalo + c_circle_y[8]][lane + kHalo + c_circle_x[8]];
        mrow[k] = mask_lds[r];
    }
    uint32_t nlist = 0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const uint32_t r = wave * RPT + k;
        const uint32_t y = static_cast<uint32_t>(y0) + r;
        wscore[k * 64 + lane] = 0;
        const bool inb = x < w && y < h;
        if (inb && l > 0 && J.pyr)  // nontemporal, as the smoothed image (read by later launches)
            store_f32_nt(J.pyr, 4 * (G.pyr_off + y * w + x), cen[k]);
        bool need = inb && tile_x < G.wpr &&
                    (y == 0 ? (x % 9) < 3 : ((mrow[k] >> lane) & 1ull) != 0);
        if (quick)
            need = need & (__builtin_fabsf(cen[k] - q7[k]) > prm.threshold) &
                   (__builtin_fabsf(cen[k] - q8[k]) > prm.threshold);
        const uint64_t bits = __ballot(need);
        if (need) {
            const uint32_t pos = nlist + __popcll(bits & ((1ull << lane) - 1ull));
            wlist[pos] = static_cast<uint16_t>((k << 6) | lane);
        }
        nlist += __popcll(bits);
    }
    wave_sync();
    for (uint32_t e0 = 0; e0 < nlist; e0 += 64) {
        const uint32_t e = e0 + lane;
        if (e < nlist) {
            const uint32_t v = wlist[e];
            const uint32_t k = v >> 6, c = v & 63u;
            const uint32_t r = wave * RPT + k;
            const float center = tile[r + kHalo][c + kHalo];
            uint32_t curr = 1, streak = 1;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const float counter = tile[r + kHalo + c_circle_y[i]][c + kHalo + c_circle_x[i]];
                const bool diff = __builtin_fabsf(center - counter) > prm.threshold;
                curr = diff ? curr + 1 : 1;
                streak = curr > streak ? curr : streak;
            }
            wscore[k * 64 + c] = static_cast<uint8_t>(streak > prm.min_run ? streak : 0u);
        }
    }
    wave_sync();
    uint32_t score[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) score[k] = wscore[k * 64 + lane];

    // Row 0 of the level keeps only the first maximum of each 3-wide window
    // starting at a multiple of 9 (cull_corners.wgsl:36-56 with x = 9*gid.x).
    if (has_row0) {
        if (wave == 0) row0_score[2 + lane] = score[0];
        if (tid >= 64 && tid < 68) {
            // detect_corners.wgsl:58-81 over the neighbour's staged ring.
            const uint32_t k = tid - 64;
            const float center = row0_ring[k][16];
            uint32_t curr = 1, streak = 1;
            for (int i = 0; i < 16; ++i) {
                const bool diff = __builtin_fabsf(center - row0_ring[k][i]) > prm.threshold;
                curr = diff ? curr + 1 : 1;
                streak = curr > streak ? curr : streak;
            }
            const int32_t xe = k < 2 ? x0 - 2 + static_cast<int32_t>(k)
                                     : x0 + kTileW + static_cast<int32_t>(k) - 2;
            const bool valid = xe >= 0 && xe < static_cast<int32_t>(w);
            row0_score[k < 2 ? k : kTileW + k] = valid && streak > prm.min_run ? streak : 0u;
        }
        __syncthreads();  // row0_score is read across waves (block-uniform branch)
    }

    // The wave's survivor count (its rows' popcounts; 0 for extension-only
    // tiles and rows past the level) goes after the survivor words, so a
    // compaction part finds the survivors before its rows without
    // recounting words.
    uint32_t wave_count = 0;  // wave-uniform
    if (tile_x < G.wpr) {  // extension-only tiles have no corners
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const uint32_t r = wave * RPT + k;
        const uint32_t y = static_cast<uint32_t>(y0) + r;
        if (y >= h) continue;  // wave-uniform
        bool keep = x < w && ((mask_lds[r] >> lane) & 1ull);
        if (keep) {
            if (y == 0) {
                const uint32_t s = x - x % 9;
                const uint32_t e = s + 3 < w ? s + 3 : w;
                uint32_t m = s;
                uint32_t best = row0_score[s - x0 + 2];
                for (uint32_t q = s + 1; q < e; ++q) {
                    const uint32_t v = row0_score[q - x0 + 2];
                    if (v > best) { best = v; m = q; }
                }
                keep = (m == x) && best > 0;
            } else {
                keep = score[k] > 0;
            }
        }
        const uint64_t bits = __ballot(keep);
        if (lane == 0) gptr(J.surv)[G.word_off + y * G.wpr + tile_x] = bits;
        wave_count += __popcll(bits);
    }
    }
    if (lane == 0)
        gptr(surv_counts(J.surv, P.words))[4 * (G.tile_base + T.tile_y * G.tiles_x + tile_x) + wave] =
            wave_count;
}

// ---------------------------------------------------------------------------
// k_compact: `parts` workgroups of 16 waves per (frame, level), part p
// taking the p-th share of the level's detect tile rows (their survivor
// words): 8 for a single frame (the level's work spread over 8 CUs), 1 for
// batches (enough workgroups already).  A part first sums the survivor
// counts the detect waves of the tile rows before its share wrote (one
// load per thread for a 1280x720 level) for its base rank, so the parts
// need nothing from each other.  Its slice of words is cut into
// chunks of 64 words (one per lane, read coalesced); in each round of up to
// kRoundChunks chunks, chunk c goes to wave c % 16.  A round counts its
// chunks (words kept in registers), a block scan over the chunk counts gives
// every chunk its base rank, then every wave ranks the survivors of its
// chunks whose base is below the cap and writes their keypoints, all waves
// at once.  The last part writes the level's survivor count.
// ---------------------------------------------------------------------------
constexpr uint32_t kCompactThreads = 1024;
constexpr uint32_t kCompactWaves = kCompactThreads / 64;
constexpr uint32_t kCompactChunks = 16;                            // per wave and round
constexpr uint32_t kRoundChunks = kCompactWaves * kCompactChunks;  // 256 (16384 words)

// v plus the value of the lane the DPP control names (0 where it names none
// or its row is masked off): one VALU with a DPP source, no LDS round trip.
template <int kCtrl, int kRowMask = 0xf>
__device__ inline uint32_t dpp_add(uint32_t v) {
    return v + static_cast<uint32_t>(
                   __builtin_amdgcn_update_dpp(0, static_cast<int>(v), kCtrl, kRowMask, 0xf, false));
}

// Inclusive sum over the aligned group of W (16, 32 or 64) lanes: shifts
// 1, 2, 4, 8 within each 16-lane row, then lane 15 of a row into the next
// row (W >= 32) and lane 31 into the upper half (W = 64).  Every lane of a
// group must be active.
template <uint32_t W>
__device__ inline uint32_t group_scan_dpp(uint32_t v) {
    static_assert(W == 16 || W == 32 || W == 64, "group of 16, 32 or 64 lanes");
    v = dpp_add<0x111>(v);  // row_shr:1
    v = dpp_add<0x112>(v);  // row_shr:2
    v = dpp_add<0x114>(v);  // row_shr:4
    v = dpp_add<0x118>(v);  // row_shr:8
    if (W >= 32) v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
    if (W == 64) v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
    return v;
}

__device__ inline uint32_t wave_inclusive_sum(uint32_t v) { return group_scan_dpp<64>(v); }

// The value of the lane the DPP control names (every lane of the row valid).
template <int kCtrl>
__device__ inline uint32_t dpp_mov(uint32_t v) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), kCtrl, 0xf, 0xf, false));
}

// Butterfly partners within a 16-lane row: lane ^ 1, lane ^ 2 (quad_perm),
// the mirrored quad of the 8-lane half (row_half_mirror) and the mirrored
// half of the row (row_mirror); after the four steps every lane has met
// every other lane of its row.
constexpr int kDppQuadXor1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;

// Sum over the aligned group of W (16, 32 or 64) lanes, in every lane.
template <uint32_t W>
__device__ inline uint32_t group_sum_dpp(uint32_t v) {
    v += dpp_mov<kDppQuadXor1>(v);
    v += dpp_mov<kDppQuadXor2>(v);
    v += dpp_mov<kDppHalfMirror>(v);
    v += dpp_mov<kDppRowMirror>(v);
#pragma unroll
    for (uint32_t off = 16; off < W; off <<= 1) v += __shfl_xor(v, static_cast<int>(off), W);
    return v;
}

// Sum over the wave (every lane active), wave-uniform.
__device__ inline uint32_t wave_total(uint32_t v) {
    return __builtin_amdgcn_readlane(group_scan_dpp<64>(v), 63);
}

__device__ inline void emit_word(uint64_t v, uint32_t k, uint32_t rank, uint32_t cap,
                                 uint32_t wpr, ORBFE_GLOBAL uint32_t* kp) {
    if (!v || rank >= cap) return;
    const uint32_t y = k / wpr;
    const uint32_t xb = (k - y * wpr) * 64;
    while (v && rank < cap) {
        const uint32_t bit = __ffsll(static_cast<unsigned long long>(v)) - 1;
        v &= v - 1;
        kp[rank++] = (xb + bit) | (y << 16);
    }
}

__global__ __launch_bounds__(kCompactThreads) void k_compact(BatchView bv, Job job0,
                                                              uint32_t levels, uint32_t parts) {
    __shared__ uint32_t cbase[kRoundChunks];
    __shared__ uint32_t wsum[kCompactWaves];
    __shared__ uint32_t round_total;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const uint32_t per_frame = levels * parts;
    const uint32_t blk = bv.n ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t j = blk / per_frame;
    const uint32_t rem = blk - j * per_frame;
    const uint32_t l = rem / parts;
    const uint32_t part = rem - l * parts;
    const Job J = bv.n ? copy_const(const_ptr(bv.jobs + j)) : job0;
    ConstPlan P = as_const(J.plan);
    const LevelGeom G = copy_const(&P.lv[l]);  // by value: stores below cannot alias it
    const uint32_t words = G.h * G.wpr;
    // Part p takes whole detect tile rows (tile_h level rows each): words
    // [w0, w1) of the level, within the 64-word chunks [c0, c1).
    const uint32_t th = P.tile_h;
    const uint32_t trows = (G.h + th - 1) / th;
    const uint32_t part_rows = (trows + parts - 1) / parts;
    const uint32_t tr0 = umin(part * part_rows, trows);
    const uint32_t tr1 = umin(tr0 + part_rows, trows);
    const uint32_t w0 = tr0 * th * G.wpr;
    const uint32_t w1 = umin(tr1 * th, G.h) * G.wpr;
    const uint32_t c0 = w0 / 64, c1 = (w1 + 63) / 64;
    const ORBFE_GLOBAL uint64_t* surv = gptr(J.surv) + G.word_off;
    ORBFE_GLOBAL uint32_t* kp = gptr(J.kp) + G.kp_off;
    const uint32_t cap = P.cap;
    const bool last = part == parts - 1;

    // Survivors before the slice: the detect waves' counts of the tile rows
    // [0, tr0) (4 per tile, tiles_x tiles per tile row, contiguous).
    uint32_t carry = 0;  // block-uniform
    if (tr0 > 0) {
        const ORBFE_GLOBAL uint32_t* cnt =
            gptr(surv_counts(J.surv, P.words)) + 4 * G.tile_base;
        const uint32_t n = 4 * tr0 * G.tiles_x;
        uint32_t pc = 0;
        constexpr uint32_t kU = 4;  // loads in flight per thread
        for (uint32_t k0 = 0; k0 < n; k0 += kU * kCompactThreads) {
            // loads at clamped indices, unconditional, so all kU are in
            // flight at once (a load under a branch is waited for inside it)
            uint32_t v[kU];
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u) v[u] = cnt[umin(k0 + u * kCompactThreads + tid, n - 1)];
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u)
                if (k0 + u * kCompactThreads + tid >= n) v[u] = 0u;
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u) pc += v[u];
        }
        const uint32_t t = wave_total(pc);
        if (lane == 0) wsum[wave] = t;
        __syncthreads();
#pragma unroll
        for (uint32_t w = 0; w < kCompactWaves; ++w) carry += wsum[w];
        if (carry >= cap && !last) return;  // block-uniform: nothing of the slice is kept
    }

    for (uint32_t r0 = c0; r0 < c1; r0 += kRoundChunks) {
        uint64_t wv[kCompactChunks];
        uint32_t cc[kCompactChunks];
#pragma unroll
        for (uint32_t c = 0; c < kCompactChunks; ++c)
            wv[c] = surv[umin((r0 + wave + kCompactWaves * c) * 64 + lane, words - 1)];
#pragma unroll
        for (uint32_t c = 0; c < kCompactChunks; ++c) {
            const uint32_t chunk = r0 + wave + kCompactWaves * c;
            const uint32_t k = chunk * 64 + lane;
            if (!(chunk < c1 && k >= w0 && k < w1)) wv[c] = 0ull;
            cc[c] = __popcll(wv[c]);
        }
#pragma unroll
        for (uint32_t c = 0; c < kCompactChunks; ++c) {
            const uint32_t t = wave_total(cc[c]);
            if (lane == 0) cbase[wave + kCompactWaves * c] = t;
        }
        __syncthreads();
        if (wave == 0) {  // exclusive scan of the round's chunk counts, kPer per lane
            constexpr uint32_t kPer = kRoundChunks / 64;
            uint32_t v[kPer];
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t k = 0; k < kPer; ++k) {
                v[k] = cbase[kPer * lane + k];
                sum += v[k];
            }
            const uint32_t incl = wave_inclusive_sum(sum);
            uint32_t run = carry + incl - sum;
#pragma unroll
            for (uint32_t k = 0; k < kPer; ++k) {
                cbase[kPer * lane + k] = run;
                run += v[k];
            }
            if (lane == 63) round_total = incl;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t c = 0; c < kCompactChunks; ++c) {
            const uint32_t cl = wave + kCompactWaves * c;
            const uint32_t base = cbase[cl];
            if (r0 + cl < c1 && base < cap) {  // wave-uniform
                const uint32_t incl = wave_inclusive_sum(cc[c]);
                emit_word(wv[c], (r0 + cl) * 64 + lane, base + incl - cc[c], cap, G.wpr, kp);
            }
        }
        carry += round_total;
        __syncthreads();  // cbase and round_total are rewritten by the next round
    }
    if (last && tid == 0) {
        ORBFE_GLOBAL uint32_t* hdr = reinterpret_cast<ORBFE_GLOBAL uint32_t*>(gptr(J.feat) + G.feat_off);
        hdr[0] = carry;
        hdr[1] = 0;
    }
}

// ---------------------------------------------------------------------------
// k_describe: one wave per keypoint slot.
// ---------------------------------------------------------------------------
namespace {

struct LevelSrc {
    const ORBFE_GLOBAL uint8_t* img;  // level 0
    uint64_t pitch;
    const ORBFE_GLOBAL float* lvl;    // levels >= 1
    const float* lut;
    uint32_t w, h;
    bool base;
    __device__ float at(uint32_t x, uint32_t y) const {
        return base ? lut[img[y * pitch + x]] : lvl[y * w + x];
    }
    __device__ float load(int32_t x, int32_t y, uint32_t border_zero) const {
        if (border_zero) {
            if (x < 0 || y < 0 || x >= static_cast<int32_t>(w) || y >= static_cast<int32_t>(h))
                return 0.0f;
            return at(x, y);
        }
        return at(tint_clamp(x, w), tint_clamp(y, h));
    }
    // load() in two steps, so that several loads can be in flight: raw()
    // reads the texel at the clamped coordinates (the u8 for level 0, the f32
    // bits otherwise), value() turns it into load()'s result.  In range both
    // clamps are the identity; outside, border_zero gives 0.
    template <bool kBase>
    __device__ uint32_t raw(int32_t x, int32_t y) const {
        const uint32_t cx = tint_clamp(x, w), cy = tint_clamp(y, h);
        if constexpr (kBase) return img[cy * pitch + cx];
        return __float_as_uint(lvl[cy * w + cx]);
    }
    __device__ float value(uint32_t r, int32_t x, int32_t y, uint32_t border_zero) const {
        const bool out =
            x < 0 || y < 0 || x >= static_cast<int32_t>(w) || y >= static_cast<int32_t>(h);
        if (border_zero && out) return 0.0f;
        return base ? lut[r] : __uint_as_float(r);
    }
};

}  // namespace

// LANES lanes per keypoint (8, 16 or 32; 256 / LANES keypoints per
// workgroup): lane s of a group scores ring pixels s, s + LANES, ... and
// evaluates tests TPL*s .. TPL*s + TPL - 1, TPL = 256 / LANES.  Every
// variant computes each record with the same operations in the same order;
// kAllLoads only changes how many gather loads are in flight at once.
constexpr uint32_t kDescribeThreads = 256;

template <uint32_t LANES, bool kAllLoads>
__global__ __launch_bounds__(kDescribeThreads)
__attribute__((amdgpu_waves_per_eu(LANES == 16 ? 8 : 1))) void k_describe(BatchView bv, Job job0,
                                                                Params prm) {
    constexpr uint32_t KPB = kDescribeThreads / LANES;  // keypoints per block
    constexpr uint32_t TPL = 256 / LANES;               // tests per lane
    constexpr uint32_t LPW = LANES / 8;                 // lanes per descriptor word
    constexpr uint32_t RING = LANES < 16 ? 16 / LANES : 1;
    constexpr uint32_t OUTS = (12 + LANES - 1) / LANES; // record words per lane
    constexpr uint64_t GMASK = LANES == 64 ? ~0ull : ((1ull << LANES) - 1ull);
    // The 7 x 7 neighbourhood of a keypoint covers the FAST ring and, for
    // orientation radii up to 3, the moment disc.
    constexpr int32_t kBoxR = 3;
    constexpr uint32_t kBoxSide = 2 * kBoxR + 1, kBoxN = kBoxSide * kBoxSide;
    // Sample points of one test group: the point in slot q of group `half`
    // is at the pattern position in bytes 2 * half, 2 * half + 1 of
    // c_desc_ppt8[q]; its sample sits in slot q of the keypoint's buffer.
    // The tables sit in LDS (one load per thread to fill them), so the
    // texture-address path, which bounds this kernel in batches, serves only
    // the frame's gathers.
    constexpr uint32_t PPL = kDescPts / LANES;          // points per lane and group
    constexpr uint32_t ROW = desc_order_row<LANES>();   // gather-order bytes per lane
    constexpr uint32_t kSlots = kDescSlots;
    static_assert(PPL <= ROW && TPL % 2 == 0, "gather rows hold PPL ids; TPL / 2 tests per group");
    __shared__ float lut[256];
    // Pattern positions of both groups' slots: packed int8 (4 B per slot)
    // where LDS is tight (8 and 16 lanes: 16 or 32 keypoints per block),
    // float pairs per group (no conversion before the gathers) at 32 lanes.
    constexpr bool kPackedPts = LANES < 32;
    __shared__ uint32_t ppt[kPackedPts ? kSlots : 4 * kSlots];
    float2 (*const pptf)[kSlots] = reinterpret_cast<float2 (*)[kSlots]>(ppt);
    __shared__ uint32_t tests[2][128];   // c_desc_tests
    __shared__ float box[KPB][kBoxN];    // per keypoint: level texels around it
    __shared__ float samp[KPB][kSlots];  // per keypoint: smoothed value of each point
    const uint32_t tid = threadIdx.x;
    // The block's level, its survivor count and keypoint entries are loaded
    // first (from the job alone), then the tables are filled while they
    // arrive.  Every XCD takes a contiguous share of the blocks, single
    // frames included: consecutive keypoints of a level lie in the same image
    // rows, so an XCD's gathers share the lines its L2 fetched (the smoothed
    // image reaches describe from memory; detect stores it nontemporally).
    uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    uint32_t j = 0;
    if (bv.n) {
        j = job_of(bv.desc_job, b);
        b -= as_const(bv.desc_prefix + j);
    }
    const Job J = bv.n ? copy_const(const_ptr(bv.jobs + j)) : job0;
    const uint32_t l = b / J.desc_groups;
    const uint32_t g = b - l * J.desc_groups;
    const uint32_t lane = tid & 63;
    const uint32_t sub = lane & (LANES - 1);   // lane within the keypoint group
    const uint32_t gbase = lane & ~(LANES - 1);  // first lane of the group
    const uint32_t slot = g * KPB + tid / LANES;
    const bool in_cap = slot < J.cap;
    ORBFE_GLOBAL uint8_t* block = gptr(J.feat) + static_cast<size_t>(l) * J.feat_stride;  // the level's FeatureArray
    const uint32_t count = reinterpret_cast<const ORBFE_GLOBAL uint32_t*>(block)[0];
    const uint32_t kp_entry = in_cap ? gptr(J.kp)[l * J.cap + slot] : 0u;
    ConstPlan P = as_const(J.plan);
    const LevelGeom G = copy_const(&P.lv[l]);  // by value: stores below cannot alias it
    // This lane's TPL / 2 tests of each group (tests TPL * sub .. + TPL - 1).
    static_assert(kDescribeThreads == 256, "one lookup entry and one test entry per thread");
    lut[tid] = c_u8_unit[tid];
    tests[tid >> 7][tid & 127u] = c_desc_tests[tid >> 7][tid & 127u];
    if (tid < kSlots) {
        const uint32_t v = c_desc_ppt8[tid];
        if constexpr (kPackedPts) {
            ppt[tid] = v;
        } else {
            const auto f = [v](int b) {
                return static_cast<float>(static_cast<int8_t>((v >> (8 * b)) & 0xFFu));
            };
            pptf[0][tid] = make_float2(f(0), f(1));
            pptf[1][tid] = make_float2(f(2), f(3));
        }
    }
    __syncthreads();

    const uint32_t kept = count < J.cap ? count : J.cap;
    const bool active = slot < kept;
    ORBFE_GLOBAL uint32_t* out = reinterpret_cast<ORBFE_GLOBAL uint32_t*>(block + 8 + 48ull * slot);
    if (__ballot(active) == 0) {  // whole wave empty: zero its slots
#pragma unroll
        for (uint32_t o = 0; o < OUTS; ++o) {
            const uint32_t e = sub + o * LANES;
            if (in_cap && e < 12) out[e] = 0u;
        }
        return;
    }
    const uint32_t kpv = active ? kp_entry : 0u;
    const uint32_t x = kpv & 0xFFFFu;
    const uint32_t y = kpv >> 16;

    LevelSrc S;
    S.img = gptr(J.img);
    S.pitch = J.pitch;
    S.lvl = gptr(J.pyr) + G.pyr_off;
    S.lut = lut;
    S.w = G.w;
    S.h = G.h;
    S.base = l == 0;

    // The group's lanes load the keypoint's 7 x 7 neighbourhood once (with
    // the textureLoad robustness rule); the ring and the moment disc read it
    // from LDS.  Radii above 3 read the level directly.
    const int32_t r = prm.orient_r;
    const bool boxed = r <= kBoxR;
    float* bx = box[tid / LANES];
    if (boxed && kAllLoads) {  // single frames: one entry at a time measured faster
        for (uint32_t e = sub; e < kBoxN; e += LANES) {
            const int32_t dy = static_cast<int32_t>(e / kBoxSide) - kBoxR;
            const int32_t dx = static_cast<int32_t>(e % kBoxSide) - kBoxR;
            bx[e] = S.load(static_cast<int32_t>(x) + dx, static_cast<int32_t>(y) + dy,
                           prm.border_zero);
        }
        wave_sync();
    } else if (boxed && !S.base && x >= kBoxR && y >= kBoxR && x + kBoxR + 1 < S.w &&
               y + kBoxR < S.h) {
        // A pyramid level's keypoint whose box and the column after it lie
        // inside the level: lanes 2r and 2r + 1 load row r as two 16-byte
        // runs (columns x - 3 .. x and x + 1 .. x + 4, the last unused):
        // ceil(14 / LANES) loads per lane instead of ceil(49 / LANES); the
        // values are the in-range texels the per-entry form reads.
        for (uint32_t q = sub; q < 2 * kBoxSide; q += LANES) {
            const uint32_t r = q >> 1, half = q & 1u;
            const uint4 v = gload_u4(reinterpret_cast<const ORBFE_GLOBAL uint32_t*>(
                S.lvl + (y - kBoxR + r) * S.w + x - kBoxR + 4 * half));
            float* row = bx + r * kBoxSide + 4 * half;
            row[0] = __uint_as_float(v.x);
            row[1] = __uint_as_float(v.y);
            row[2] = __uint_as_float(v.z);
            if (half == 0) row[3] = __uint_as_float(v.w);
        }
        wave_sync();
    } else if (boxed) {
        // every lane's loads first (entries past the box repeat the last),
        // then the values
        constexpr uint32_t kPer = (kBoxN + LANES - 1) / LANES;
        uint32_t raw[kPer];
        const auto load_all = [&](auto base) {
#pragma unroll
            for (uint32_t k = 0; k < kPer; ++k) {
                const uint32_t e = umin(sub + k * LANES, kBoxN - 1);
                const int32_t dy = static_cast<int32_t>(e / kBoxSide) - kBoxR;
                const int32_t dx = static_cast<int32_t>(e % kBoxSide) - kBoxR;
                raw[k] = S.template raw<decltype(base)::value>(static_cast<int32_t>(x) + dx,
                                                               static_cast<int32_t>(y) + dy);
            }
        };
        if (S.base)  // block-uniform: one loop of loads either way
            load_all(std::true_type{});
        else
            load_all(std::false_type{});
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t e = sub + k * LANES;
            const int32_t dy = static_cast<int32_t>(e / kBoxSide) - kBoxR;
            const int32_t dx = static_cast<int32_t>(e % kBoxSide) - kBoxR;
            if (e < kBoxN)
                bx[e] = S.value(raw[k], static_cast<int32_t>(x) + dx,
                                static_cast<int32_t>(y) + dy, prm.border_zero);
        }
        wave_sync();
    }

    // Corner strength (detect_corners.wgsl:58-81): ring bit i from lane
    // i % LANES.
    const float center = boxed ? bx[kBoxR * kBoxSide + kBoxR] : S.at(x, y);
    uint32_t ring = 0;
#pragma unroll
    for (uint32_t q = 0; q < RING; ++q) {
        const uint32_t i = sub + q * LANES;
        bool diff = false;
        if (i < 16) {
            const float counter =
                boxed ? bx[(kBoxR + c_circle_y[i]) * kBoxSide + kBoxR + c_circle_x[i]]
                      : S.load(static_cast<int32_t>(x) + c_circle_x[i],
                               static_cast<int32_t>(y) + c_circle_y[i], prm.border_zero);
            diff = __builtin_fabsf(center - counter) > prm.threshold;
        }
        ring |= static_cast<uint32_t>((__ballot(diff) >> gbase) & GMASK) << (q * LANES);
    }
    const uint32_t streak = ring_streak(ring);
    const uint32_t strength = streak > prm.min_run ? streak : 0u;

    // Orientation (generate_features.wgsl:80-99): sequential moments.
    float m10 = 0.0f, m01 = 0.0f;
    if (r == kBoxR) {
        // The shader's radius: its 29-point disc unrolled, in the shader's
        // order, texels at constant offsets in the box.  A zero weight's
        // term is +0 (texels are finite and >= 0), and neither sum is ever
        // -0 (both start at +0, and a round-to-nearest sum is -0 only when
        // both operands are), so adding that term changes nothing and it is
        // left out.
#pragma unroll
        for (int32_t dy = -kBoxR; dy <= kBoxR; ++dy) {
#pragma unroll
            for (int32_t dx = -kBoxR; dx <= kBoxR; ++dx) {
                if (dx * dx + dy * dy > kBoxR * kBoxR) continue;
                const float pixel = bx[(kBoxR + dy) * static_cast<int32_t>(kBoxSide) + kBoxR + dx];
                if (dx != 0) m10 += static_cast<float>(dx) * pixel;
                if (dy != 0) m01 += static_cast<float>(dy) * pixel;
            }
        }
    } else {
        for (int32_t dy = -r; dy <= r; ++dy) {
            for (int32_t dx = -r; dx <= r; ++dx) {
                if (dx * dx + dy * dy > r * r) continue;
                const float pixel =
                    boxed ? bx[(kBoxR + dy) * static_cast<int32_t>(kBoxSide) + kBoxR + dx]
                          : S.load(static_cast<int32_t>(x) + dx, static_cast<int32_t>(y) + dy,
                                   prm.border_zero);
                m10 += static_cast<float>(dx) * pixel;
                m01 += static_cast<float>(dy) * pixel;
            }
        }
    }
    const float theta = atan2_own(m01, m10);
    float sn, cs;
    sincos_own(theta, &sn, &cs);

    // Descriptor (generate_features.wgsl:101-144): G at the wrapped u32
    // coordinates (:116-129) is looked up in the smoothed image, where
    // coordinates past the extension (including wrapped negatives) share the
    // value at w+3 / h+3.  The tests run in two groups: each group's
    // distinct sample points (c_desc_ppt8, at most kDescPts) are gathered into
    // LDS, then every lane evaluates its tests of the group from there.  Two
    // tests that name the same pattern position read the same rotated point,
    // so gathering it once changes no value.  The gather order comes from a
    // table per angle bin and keypoint alignment (x mod 8, y mod 4) that
    // groups the points by the 8 x 4 block of the grid (one 128-byte line of
    // the smoothed layout) their position rotated to the bin's centre falls
    // in, so the lanes of one load touch few lines; it only orders the
    // loads, every value is the one at the point's own exact rotated
    // coordinates.
    //
    // round() is half-to-even.  For |v| < 2^22, v + 1.5*2^23 lies in
    // [2^23, 2^24), where the float spacing is 1, so the addition rounds v
    // to the nearest integer, ties to even, and leaves that integer in the
    // mantissa: i32(round(v)) = bits(v + 1.5*2^23) - bits(1.5*2^23).  The
    // pattern coordinates are at most 13 in magnitude, so |v| <= 19.  The
    // subtraction is folded into the keypoint coordinate (mod 2^32); columns
    // are clamped as byte offsets within a band (16 bytes per column),
    // 16 * (x + ax) staying above 16 * (sw - 1) for the wrapped negatives.
    constexpr float kRound = 12582912.0f;  // 1.5 * 2^23
    constexpr uint32_t kRoundBits = 0x4B400000u;
    const uint32_t sw = G.w + kSmoothExt, sh = G.h + kSmoothExt;
    const uint8_t* sm = reinterpret_cast<const uint8_t*>(J.smooth + G.sm_off);
    const uint32_t yb = y - kRoundBits;
    const uint32_t xb16 = 16 * (x - kRoundBits), xmax16 = 16 * (sw - 1);
    const uint32_t band_bytes = 4 * smooth_band_floats(G.w);
    const uint32_t bin = umin(static_cast<uint32_t>((theta + fm::kPi) * kAngleBinScale),
                              kAngleBins - 1);
    const uint32_t align = (x & 7u) | (y & 3u) << 3;
    const uint32_t* order = reinterpret_cast<const uint32_t*>(
        desc_order<LANES>() +
        ((static_cast<size_t>(bin) * kOrderAligns + align) * 2 * LANES + sub) * ROW);
    float* ks = samp[tid / LANES];
    uint32_t bits = 0;
    // Byte offset in the smoothed image of the point in slot q of a group,
    // rotated by the keypoint's angle (band index and band stride below
    // 2^24: 24-bit factors).
    const auto point_offset = [&](uint32_t half, uint32_t q) -> uint32_t {
        float2 pp;
        if constexpr (kPackedPts) {
            const uint32_t pv = ppt[q] >> (16 * half);
            pp = make_float2(static_cast<float>(static_cast<int8_t>(pv & 0xFFu)),
                             static_cast<float>(static_cast<int8_t>((pv >> 8) & 0xFFu)));
        } else {
            pp = pptf[half][q];
        }
        const uint32_t ax16 =
            umin(16 * __float_as_uint((pp.x * cs - pp.y * sn) + kRound) + xb16, xmax16);
        const uint32_t ay = umin(__float_as_uint((pp.x * sn + pp.y * cs) + kRound) + yb, sh - 1);
        return __umul24(ay >> 2, band_bytes) + (ax16 | (ay & 3u) << 2);
    };
    // The lane's tests of a group, once the group's samples are in ks: entry
    // e names the byte offsets of A's and B's samples in ks and the test t,
    // which sets bit t - TPL * sub = t mod TPL of the lane's bits.
    const char* ksb = reinterpret_cast<const char*>(ks);
    const auto group_tests = [&](uint32_t half) {
        wave_sync();
#pragma unroll
        for (uint32_t m = 0; m < TPL / 2; ++m) {
            const uint32_t e = tests[half][TPL / 2 * sub + m];
            const float sa = *reinterpret_cast<const float*>(ksb + (e & 0xFFFFu));
            const float sb = *reinterpret_cast<const float*>(ksb + ((e >> 16) & 0x3FFu));
            bits |= (sa > sb ? 1u : 0u) << ((e >> 26) & (TPL - 1));
        }
        wave_sync();
    };
    const auto point_slot = [&](const uint32_t* ids, uint32_t k) -> uint32_t {
        return (ids[k / 4] >> (8 * (k % 4))) & 0xFFu;
    };
    if constexpr (kAllLoads) {
        // Single-frame launches (latency-bound): both groups' loads in flight
        // at once, then each group through LDS.
        float v[2][PPL];
        uint32_t qs[2][PPL];
#pragma unroll
        for (uint32_t half = 0; half < 2; ++half) {
            uint32_t ids[ROW / 4];
#pragma unroll
            for (uint32_t k = 0; k < ROW / 4; ++k) ids[k] = order[half * LANES * ROW / 4 + k];
#pragma unroll
            for (uint32_t k = 0; k < PPL; ++k) {
                qs[half][k] = point_slot(ids, k);
                v[half][k] = load_f32(sm, point_offset(half, qs[half][k]));
            }
        }
#pragma unroll
        for (uint32_t half = 0; half < 2; ++half) {
#pragma unroll
            for (uint32_t k = 0; k < PPL; ++k) ks[qs[half][k]] = v[half][k];
            group_tests(half);
        }
    } else {
#pragma unroll
        for (uint32_t half = 0; half < 2; ++half) {
            uint32_t ids[ROW / 4];
#pragma unroll
            for (uint32_t k = 0; k < ROW / 4; ++k) ids[k] = order[half * LANES * ROW / 4 + k];
            // loads in flight per batch (a divisor of PPL)
            constexpr uint32_t kB = PPL <= 12 ? PPL : (PPL % 8 == 0 ? 8 : PPL / 2);
            static_assert(PPL % kB == 0, "whole batches");
#pragma unroll
            for (uint32_t k0 = 0; k0 < PPL; k0 += kB) {
                float v[kB];
                uint32_t qs[kB];
#pragma unroll
                for (uint32_t u = 0; u < kB; ++u) {
                    qs[u] = point_slot(ids, k0 + u);
                    v[u] = load_f32(sm, point_offset(half, qs[u]));
                }
#pragma unroll
                for (uint32_t u = 0; u < kB; ++u) ks[qs[u]] = v[u];
            }
            group_tests(half);
        }
    }
    // Record word e: x, y, strength, angle, then descriptor word e - 4 =
    // bits of lanes LPW*(e-4) .. LPW*(e-4) + LPW - 1, low lane first.
#pragma unroll
    for (uint32_t o = 0; o < OUTS; ++o) {
        const uint32_t e = sub + o * LANES;
        const uint32_t wsel = e >= 4 && e < 12 ? e - 4 : 0;
        uint32_t word = 0;
#pragma unroll
        for (uint32_t q = 0; q < LPW; ++q)
            word |= __shfl(bits, static_cast<int>(gbase + wsel * LPW + q)) << (q * TPL);
        if (in_cap && e < 12) {
            uint32_t v;
            if (!active) {
                v = 0u;
            } else {
                switch (e) {
                    case 0: v = __float_as_uint(static_cast<float>(x)); break;
                    case 1: v = __float_as_uint(static_cast<float>(y)); break;
                    case 2: v = strength; break;
                    case 3: v = __float_as_uint(theta); break;
                    default: v = word; break;
                }
            }
            out[e] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Pipeline matcher (FindMatchesInLod / FindMatches).
// ---------------------------------------------------------------------------
namespace {

struct FeatView {
    const uint8_t* base;
    uint32_t stride;  // feature array block stride
    __device__ const ORBFE_GLOBAL uint32_t* rec(uint32_t lod, uint32_t i) const {
        return reinterpret_cast<const ORBFE_GLOBAL uint32_t*>(gptr(base) + stride * lod + 8 +
                                                              48ull * i);
    }
    __device__ uint32_t count(uint32_t lod, uint32_t cap) const {
        const uint32_t c =
            *reinterpret_cast<const ORBFE_GLOBAL uint32_t*>(gptr(base) + stride * lod);
        return c < cap ? c : cap;
    }
};

}  // namespace

constexpr uint32_t kMatchThreads = 256;
// Candidates a lane loads per round (their loads are issued together; more
// would cost occupancy).
constexpr uint32_t kCandPerLane = 2;

// LOD0 position of a record (toLod0, match_features_cpu.cpp:58-64).
__device__ inline void rec_lod0(const uint32_t* r, float scale, float* x, float* y) {
    *x = static_cast<float>(static_cast<uint32_t>(__uint_as_float(r[0]))) * scale;
    *y = static_cast<float>(static_cast<uint32_t>(__uint_as_float(r[1]))) * scale;
}

// Cell coordinate of v on an axis of n cells of width 1 / inv (clamped).
__device__ inline uint32_t cell_of(float v, float inv, uint32_t n) {
    const float c = __builtin_floorf((v > 0.0f ? v : 0.0f) * inv);
    return c < static_cast<float>(n - 1) ? static_cast<uint32_t>(c) : n - 1;
}

// k_match_index: per (pair, side, level) a counting sort of the level's
// records by LOD0 cell.  The index of one side holds the cell starts of
// every level ([L][kCells + 1]) and, for every level's records in cell
// order, heads (LOD0 position, record index) and the two descriptor halves
// as 16-byte arrays over a global record number g = level * cap + k, so a
// cell range of candidates is read with contiguous loads.  Cell order is
// monotone in the coordinates, so every record inside a search window lies
// in the cells the window's (margin-widened) corners map to.
// 1024 threads; each thread loads its first kIndexHeld records once (all
// loads issued together) and keeps them in registers across the histogram
// and its scan; records past 1024 * kIndexHeld (caps above 4096) are
// reloaded in the scatter pass.
constexpr uint32_t kIndexThreads = 1024;
constexpr uint32_t kIndexHeld = 4;

__global__ __launch_bounds__(kIndexThreads) void k_match_index(MatchBatchView bv, MatchJob job0) {
    __shared__ uint32_t hist[kCells + 1];
    const uint32_t tid = threadIdx.x;
    const MatchJob M = bv.n ? copy_const(const_ptr(bv.jobs)) : job0;  // layout constants only
    const uint32_t L = as_const(M.plan).L;
    const uint32_t blk = bv.n ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t j = blk / (2 * L);
    const uint32_t side = (blk / L) & 1;
    const uint32_t l = blk % L;
    const MatchJob J = bv.n ? copy_const(const_ptr(bv.jobs + j)) : job0;
    ConstPlan P = as_const(J.plan);
    const FeatView A{side ? J.curr : J.prev, P.feat_stride};
    ORBFE_GLOBAL uint32_t* base = gptr(side ? J.index_curr : J.index_prev);
    ORBFE_GLOBAL uint32_t* starts = base + l * (kCells + 1);
    const uint32_t n = A.count(l, P.cap);
    const float sc = P.lv[l].lod0_scale;
    const float inv_x = P.cell_inv_x, inv_y = P.cell_inv_y;
    // Held records as named variables (an array indexed in a loop would be
    // placed in scratch or LDS before the loop is unrolled).
    struct Held {
        uint2 head;  // the record's x, y
        uint4 lo, hi;
        uint32_t cell;
        float x, y;
    };
    static_assert(kIndexHeld == 4, "four held records");
    // Unconditional loads at a clamped slot (slot 0 of an empty level is
    // still inside the level's record block), so all held records' loads
    // are in flight together; slots past lim are never used.
    const auto load = [&](uint32_t i, uint32_t lim) -> Held {
        Held h{};
        const auto* r = A.rec(l, umin(i, (lim ? lim : 1u) - 1u));
        h.head = gload_u2(r);  // only x and y (no dead halves to be reused early)
        h.lo = gload_u4(r + 4);
        h.hi = gload_u4(r + 8);
        return h;
    };
    // A single pair (latency-bound) loads the level's slots up to the cap
    // without waiting for its count (slots past the count are read, never
    // used); batches load only the counted records.
    Held h0, h1, h2, h3;
    if (bv.n == 0) {
        h0 = load(tid, P.cap);
        h1 = load(tid + kIndexThreads, P.cap);
        h2 = load(tid + 2 * kIndexThreads, P.cap);
        h3 = load(tid + 3 * kIndexThreads, P.cap);
    } else {
        h0 = load(tid, n);
        h1 = load(tid + kIndexThreads, n);
        h2 = load(tid + 2 * kIndexThreads, n);
        h3 = load(tid + 3 * kIndexThreads, n);
    }
    // The mutual pass writes only the matched entries of prev_match; the
    // direction pass appends to the query list.
    if (side == 0)
        for (uint32_t i = tid; i < P.cap; i += kIndexThreads) gptr(J.prev_match)[l * P.cap + i] = -1;
    if (side == 0 && l == 0 && tid < kListSegs) gptr(J.q_list)[tid * kListStride] = 0;
    for (uint32_t c = tid; c <= kCells; c += kIndexThreads) hist[c] = 0;
    // LOD0 position (toLod0, match_features_cpu.cpp:58-64) and cell
    const auto pos_of = [&](uint32_t rx, uint32_t ry, float* x, float* y) -> uint32_t {
        *x = static_cast<float>(static_cast<uint32_t>(__uint_as_float(rx))) * sc;
        *y = static_cast<float>(static_cast<uint32_t>(__uint_as_float(ry))) * sc;
        return cell_of(*y, inv_y, kCellsY) * kCellsX + cell_of(*x, inv_x, kCellsX);
    };
    __syncthreads();
    const auto count = [&](Held h, uint32_t i) -> Held {
        if (i < n) {
            float x, y;
            h.cell = pos_of(h.head.x, h.head.y, &x, &y);
            h.x = x;
            h.y = y;
            atomicAdd(&hist[h.cell], 1u);
        }
        return h;
    };
    h0 = count(h0, tid);
    h1 = count(h1, tid + kIndexThreads);
    h2 = count(h2, tid + 2 * kIndexThreads);
    h3 = count(h3, tid + 3 * kIndexThreads);
    for (uint32_t i = tid + kIndexHeld * kIndexThreads; i < n; i += kIndexThreads) {
        const auto* r = A.rec(l, i);
        float x, y;
        atomicAdd(&hist[pos_of(r[0], r[1], &x, &y)], 1u);
    }
    __syncthreads();
    if (tid < 64) {  // exclusive scan of kCells counts by one wave, 9 per lane
        constexpr uint32_t kPer = kCells / 64;
        uint32_t v[kPer];
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            v[k] = hist[tid * kPer + k];
            sum += v[k];
        }
        const uint32_t incl = wave_inclusive_sum(sum);
        uint32_t run = incl - sum;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            hist[tid * kPer + k] = run;
            starts[tid * kPer + k] = run;
            run += v[k];
        }
        if (tid == 63) starts[kCells] = run;
    }
    __syncthreads();
    const size_t g0 = static_cast<size_t>(l) * P.cap;
    // 16-byte entries, addressed in dwords
    ORBFE_GLOBAL uint32_t* shead = base + P.idx_head + 4 * g0;
    ORBFE_GLOBAL uint32_t* sdlo = base + P.idx_dlo + 4 * g0;
    ORBFE_GLOBAL uint32_t* sdhi = base + P.idx_dhi + 4 * g0;
    const auto scatter = [&](Held h, uint32_t i) {
        if (i < n) {
            const uint32_t k = atomicAdd(&hist[h.cell], 1u);
            gstore_u4(shead + 4 * k, make_uint4(__float_as_uint(h.x), __float_as_uint(h.y), i, 0u));
            gstore_u4(sdlo + 4 * k, h.lo);
            gstore_u4(sdhi + 4 * k, h.hi);
        }
    };
    scatter(h0, tid);
    scatter(h1, tid + kIndexThreads);
    scatter(h2, tid + 2 * kIndexThreads);
    scatter(h3, tid + 3 * kIndexThreads);
    for (uint32_t i = tid + kIndexHeld * kIndexThreads; i < n; i += kIndexThreads) {
        const auto* r = A.rec(l, i);
        float x, y;
        const uint32_t k = atomicAdd(&hist[pos_of(r[0], r[1], &x, &y)], 1u);
        gstore_u4(shead + 4 * k, make_uint4(__float_as_uint(x), __float_as_uint(y), i, 0u));
        gstore_u4(sdlo + 4 * k, gload_u4(r + 4));
        gstore_u4(sdhi + 4 * k, gload_u4(r + 8));
    }
}

namespace {

__device__ inline uint4 load_u4(const ORBFE_GLOBAL uint32_t* base, uint32_t i) {
    const u32x4 v = reinterpret_cast<const ORBFE_GLOBAL u32x4*>(reinterpret_cast<uintptr_t>(base))[i];
    return make_uint4(v.x, v.y, v.z, v.w);
}

}  // namespace

// Pipeline matcher, LANES lanes per query (256 / LANES queries per block).
// kMutual = false: slot = curr feature b; query b, pool from prev: its match
//   a; every b with a match is appended as the entry (b, a) to segment
//   (block % kListSegs) of the query list (one atomic per wave on that
//   segment's counter).
// kMutual = true:  block b takes entries (b / kListSegs) * QPB + group of
//   segment b % kListSegs; an entry names b and its match a; query a,
//   pool from curr; prev_match[a] = b when a's best match is b.  This is the
//   reference's mutual check (match_features_cpu.cpp:196-205: a matches
//   iff its best match b has b_to_a[b] == a) evaluated only for the prev
//   features some b points at; every other entry keeps the -1 written by
//   k_match_index.  The list's order varies from run to run; each entry's
//   result does not depend on it.
//
// A query's candidates are the records of its pool levels (l, l-1, l+1, the
// reference's concatenation order) in the cells under its search window:
// one contiguous run of the index per (pool level, cell row).  The group
// flattens the runs with a prefix sum (kept in LDS), then lane s takes
// flattened candidates s, s + LANES, ... and walks forward through the runs.
// Keys (d << 22 | pool slot << 20 | record index) order candidates like the
// reference's scan with strict '<' (match_features_cpu.cpp:133-140).
template <uint32_t LANES, bool kMutual>
__global__ __launch_bounds__(kMatchThreads) void k_match_dir(MatchBatchView bv, MatchJob job0,
                                                              float ratio) {
    constexpr uint32_t QPB = kMatchThreads / LANES;
    __shared__ uint32_t r_end[QPB][kMaxRanges];
    __shared__ uint32_t r_gb[QPB][kMaxRanges];
    const uint32_t tid = threadIdx.x;
    const uint32_t sub = tid & (LANES - 1);
    // a query per wave at 64 lanes: its slot arithmetic stays scalar
    const uint32_t grp = LANES == 64 ? __builtin_amdgcn_readfirstlane(tid / LANES) : tid / LANES;
    uint32_t b = bv.n ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    uint32_t j = 0;
    if (bv.n) {
        j = job_of(bv.block_job, b);
        b -= as_const(bv.block_prefix + j);
    }
    const MatchJob M = bv.n ? copy_const(const_ptr(bv.jobs + j)) : job0;
    ConstPlan P = as_const(M.plan);
    const FeatView Q{kMutual ? M.prev : M.curr, M.feat_stride};
    const FeatView C{kMutual ? M.curr : M.prev, M.feat_stride};
    const ORBFE_GLOBAL uint32_t* cidx = gptr(kMutual ? M.index_curr : M.index_prev);
    uint32_t l, slot;  // the curr feature b
    uint32_t ab_entry = 0;  // mutual: b's direction match a, from its list entry
    if (!kMutual) {
        const uint32_t groups = (M.cap + QPB - 1) / QPB;  // blocks per level
        l = b / groups;
        slot = (b - l * groups) * QPB + grp;
        if (slot >= M.cap) return;
    } else {
        const uint32_t seg = b % kListSegs;
        const uint32_t t = (b / kListSegs) * QPB + grp;
        const uint32_t blocks = M.levels * ((M.cap + QPB - 1) / QPB);
        // Entry t of the segment (t < the segment's capacity for every block
        // of the grid) is loaded together with the segment's count, which
        // the direction pass wrote.
        const uint2 e = gload_u2(gptr(M.q_list) + kListSegs * kListStride +
                                 2 * (seg * query_segment_capacity(blocks, QPB) + t));
        if (t >= gptr(M.q_list)[seg * kListStride]) return;
        l = e.x >> 24;
        slot = e.x & 0xFFFFFFu;
        ab_entry = e.y;
    }
    // Query: feature (ql, qs) of the query side.
    uint32_t ql = l, qs = slot;
    if (kMutual) {
        ql = ab_entry >> 24;
        qs = ab_entry & 0xFFFFFFu;
    }

    // Pool levels in the reference's concatenation order: ql, ql-1, ql+1.
    const uint32_t p0 = ql;
    const uint32_t p1 = ql > 0 ? ql - 1 : ql + 1;
    const uint32_t p2 = ql + 1;
    // The query's record (qs < cap: inside its level block), the pool
    // counts and, in the direction pass, the query level's own count are
    // loaded together before any of them is used.
    const auto* qr = Q.rec(ql, qs);
    const uint4 qw0 = gload_u4(qr);
    const uint4 qw1 = gload_u4(qr + 4);
    const uint4 qw2 = gload_u4(qr + 8);
    const uint32_t c0 = C.count(ql, M.cap);
    const uint32_t c1 = C.count(ql > 0 ? ql - 1 : ql, M.cap);
    const uint32_t c2 = C.count(ql + 1 < M.levels ? ql + 1 : ql, M.cap);
    if (!kMutual && slot >= FeatView{M.curr, M.feat_stride}.count(l, M.cap)) return;
    uint32_t npool = 1, total_pool = c0;
    if (ql > 0) { total_pool += c1; ++npool; }
    if (ql + 1 < M.levels) { total_pool += c2; ++npool; }
    const auto pool_level = [&](uint32_t s) { return s == 0 ? p0 : (s == 1 ? p1 : p2); };

    int32_t result = -1;
    if (total_pool >= 2) {
        float qx, qy;
        {
            const uint32_t head[2] = {qw0.x, qw0.y};
            rec_lod0(head, P.lv[ql].lod0_scale, &qx, &qy);
        }
        const uint32_t qd[8] = {qw1.x, qw1.y, qw1.z, qw1.w, qw2.x, qw2.y, qw2.z, qw2.w};
        const float lo_x = qx - P.max_dx;
        const float hi_x = qx + P.max_dx;
        const float dy = P.max_dy;
        // Cells of the window widened by one pixel (covers the rounding of
        // the exact gates below).
        const uint32_t cx0 = cell_of(lo_x - 1.0f, P.cell_inv_x, kCellsX);
        const uint32_t cx1 = cell_of(hi_x + 1.0f, P.cell_inv_x, kCellsX);
        const uint32_t cy0 = cell_of(qy - dy - 1.0f, P.cell_inv_y, kCellsY);
        const uint32_t cy1 = cell_of(qy + dy + 1.0f, P.cell_inv_y, kCellsY);
        const uint32_t rows = cy1 - cy0 + 1;
        const uint32_t nranges = npool * rows;  // <= kMaxRanges

        // Runs (pool slot s, cell row): r_end = flattened end offset,
        // r_gb = global record number minus flattened offset.
        uint32_t total = 0;
        for (uint32_t k0 = 0; k0 < nranges; k0 += LANES) {
            const uint32_t k = k0 + sub;
            uint32_t len = 0, gstart = 0;
            if (k < nranges) {
                const uint32_t sk = k / rows;
                const uint32_t lv = pool_level(sk);
                const ORBFE_GLOBAL uint32_t* st =
                    cidx + lv * (kCells + 1) + (cy0 + (k - sk * rows)) * kCellsX;
                const uint32_t a = st[cx0];
                len = st[cx1 + 1] - a;
                gstart = lv * M.cap + a;
            }
            const uint32_t incl = total + group_scan_dpp<LANES>(len);
            if (k < nranges) {
                r_end[grp][k] = incl;
                r_gb[grp][k] = gstart - (incl - len);
            }
            total += group_sum_dpp<LANES>(len);
        }
        wave_sync();

        const ORBFE_GLOBAL uint32_t* chead = cidx + P.idx_head;
        const ORBFE_GLOBAL uint32_t* cdlo = cidx + P.idx_dlo;
        const ORBFE_GLOBAL uint32_t* cdhi = cidx + P.idx_dhi;
        uint32_t min1 = 0xFFFFFFFFu, min2 = 0xFFFFFFFFu;
        uint32_t k = 0;
        uint32_t kend = total ? r_end[grp][0] : 0u, kgb = total ? r_gb[grp][0] : 0u;
        uint32_t kslot = 0;  // pool slot of run k (k / rows)
        for (uint32_t t0 = 0; t0 < total; t0 += kCandPerLane * LANES) {
            uint32_t g[kCandPerLane], sl[kCandPerLane];
            bool ok[kCandPerLane];
#pragma unroll
            for (uint32_t u = 0; u < kCandPerLane; ++u) {
                const uint32_t t = t0 + sub + u * LANES;
                ok[u] = t < total;
                g[u] = 0;
                sl[u] = 0;
                if (ok[u]) {
                    while (t >= kend) {  // ends at the last run: its end is total
                        ++k;
                        kend = r_end[grp][k];
                        kgb = r_gb[grp][k];
                        kslot = (k >= rows ? 1u : 0u) + (k >= 2 * rows ? 1u : 0u);
                    }
                    g[u] = kgb + t;
                    sl[u] = kslot;
                }
            }
            // Branch-free: every lane loads (record 0 past the runs' end) and
            // scores its candidates, and one outside the window or past the
            // end offers the neutral key; with no loads or waits under
            // branches, the loads of a round stay in flight together.
            uint4 head[kCandPerLane], dlo[kCandPerLane], dhi[kCandPerLane];
#pragma unroll
            for (uint32_t u = 0; u < kCandPerLane; ++u) {
                head[u] = load_u4(chead, g[u]);
                dlo[u] = load_u4(cdlo, g[u]);
                dhi[u] = load_u4(cdhi, g[u]);
            }
#pragma unroll
            for (uint32_t u = 0; u < kCandPerLane; ++u) {
                const float cx = __uint_as_float(head[u].x), cy = __uint_as_float(head[u].y);
                const bool in =
                    ok[u] && !(cx < lo_x || cx > hi_x || __builtin_fabsf(qy - cy) > dy);
                const uint32_t d = hamming256(qd, dlo[u], dhi[u]);
                const uint32_t key = (d << kMatchKeyShift) | (sl[u] << kMatchIdxBits) | head[u].z;
                keep_two_smallest(min1, min2, in ? key : 0xFFFFFFFFu);
            }
        }
        // The two smallest keys of the group (each lane holds the two
        // smallest of a disjoint candidate subset).
        const auto merge = [&](uint32_t o1, uint32_t o2) {
            const uint32_t n1 = umin(min1, o1);
            const uint32_t n2 = umin(umax(min1, o1), umin(min2, o2));
            min1 = n1;
            min2 = n2;
        };
        merge(dpp_mov<kDppQuadXor1>(min1), dpp_mov<kDppQuadXor1>(min2));
        merge(dpp_mov<kDppQuadXor2>(min1), dpp_mov<kDppQuadXor2>(min2));
        merge(dpp_mov<kDppHalfMirror>(min1), dpp_mov<kDppHalfMirror>(min2));
        merge(dpp_mov<kDppRowMirror>(min1), dpp_mov<kDppRowMirror>(min2));
#pragma unroll
        for (uint32_t off = 16; off < LANES; off <<= 1)
            merge(__shfl_xor(min1, static_cast<int>(off), LANES),
                  __shfl_xor(min2, static_cast<int>(off), LANES));
        if (min1 != 0xFFFFFFFFu) {
            const uint32_t d1 = min1 >> kMatchKeyShift;
            const uint32_t d2 = min2 == 0xFFFFFFFFu ? 0xFFFFFFFFu : (min2 >> kMatchKeyShift);
            if (static_cast<float>(d1) < ratio * static_cast<float>(d2)) {
                const uint32_t s = (min1 >> kMatchIdxBits) & 3u;
                const uint32_t idx = min1 & ((1u << kMatchIdxBits) - 1u);
                result = static_cast<int32_t>((pool_level(s) << 24) | idx);
            }
        }
    }
    if (!kMutual) {
        // append b to the query list: one atomic per wave
        const uint32_t lane = tid & 63;
        const bool claim = sub == 0 && result >= 0;
        const uint64_t m = __ballot(claim);
        if (m) {
            const uint32_t seg = b % kListSegs;
            const uint32_t blocks = M.levels * ((M.cap + QPB - 1) / QPB);
            const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(m))) - 1;
            uint32_t base = 0;
            if (lane == leader)
                base = atomicAdd(M.q_list + seg * kListStride, static_cast<uint32_t>(__popcll(m)));
            base = __builtin_amdgcn_readlane(base, leader);
            if (claim) {  // the entry: b and its match a
                ORBFE_GLOBAL uint32_t* ent =
                    gptr(M.q_list) + kListSegs * kListStride +
                    2 * (seg * query_segment_capacity(blocks, QPB) + base +
                         __popcll(m & ((1ull << lane) - 1ull)));
                *reinterpret_cast<ORBFE_GLOBAL u32x2*>(ent) =
                    u32x2{(l << 24) | slot, static_cast<uint32_t>(result)};
            }
        }
    } else if (sub == 0 && result == static_cast<int32_t>((l << 24) | slot)) {
        gptr(M.prev_match)[ql * M.cap + qs] = result;
    }
}

template __global__ void k_match_dir<16, false>(MatchBatchView, MatchJob, float);
template __global__ void k_match_dir<16, true>(MatchBatchView, MatchJob, float);
template __global__ void k_match_dir<32, false>(MatchBatchView, MatchJob, float);
template __global__ void k_match_dir<32, true>(MatchBatchView, MatchJob, float);
template __global__ void k_match_dir<64, false>(MatchBatchView, MatchJob, float);
template __global__ void k_match_dir<64, true>(MatchBatchView, MatchJob, float);

// ---------------------------------------------------------------------------
// Brute-force matcher: merge of the per-split best/second-best keys.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bf_merge(const uint2* __restrict__ part, uint32_t nq,
                                                   uint32_t splits, uint32_t* __restrict__ idx,
                                                   uint32_t* __restrict__ d1,
                                                   uint32_t* __restrict__ d2) {
    const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    uint32_t m1 = 0xFFFFFFFFu, m2 = 0xFFFFFFFFu;
    for (uint32_t s = 0; s < splits; ++s) {
        const uint2 p = part[static_cast<size_t>(s) * nq + qi];
        const uint32_t n1 = umin(m1, p.x);
        const uint32_t n2 = umin(umax(m1, p.x), umin(m2, p.y));
        m1 = n1;
        m2 = n2;
    }
    const uint32_t mask = (1u << kBfKeyShift) - 1u;
    idx[qi] = m1 == 0xFFFFFFFFu ? 0xFFFFFFFFu : (m1 & mask);
    d1[qi] = m1 == 0xFFFFFFFFu ? 0xFFFFFFFFu : (m1 >> kBfKeyShift);
    d2[qi] = m2 == 0xFFFFFFFFu ? 0xFFFFFFFFu : (m2 >> kBfKeyShift);
}

// ---------------------------------------------------------------------------
// Brute-force matcher on the matrix cores.  A 256-bit descriptor is read as
// 256 signs, one int8 per bit: train descriptors as +64 for a 0 bit and -64
// for a 1 bit, queries the other way round.  The dot product of the two
// vectors is then 4096 * (2d - 256) = 8192 d - 2^20 for Hamming distance d,
// exact in int32.  v_mfma_i32_32x32x32_i8 forms it for 32 train rows x 32
// queries, 32 bits per step (8 steps per descriptor); its accumulator input
// carries each train row's index t in the split (t < 8192), so every output
// is the key 8192 d + t - 2^20, which orders as (d, t): per query, the two
// smallest keys are the best and second-best train descriptors, ties to the
// lower index.  Rows past the split end get a bias of 2^30 and never rank
// below a real row.
//
// Accumulator layout of the 32x32 products (lane l, register r): query
// column l % 32, train row 8 (r / 4) + 4 (l / 32) + r % 4.  Operand lane l
// holds train row (A) or query (B) l % 32 and 16 bits of the step's 32-bit
// word (bm_signs16); both operands use the same bit order, so which bit
// lands in which byte does not change the dot product.
// ---------------------------------------------------------------------------
typedef int32_t bm_i32x4 __attribute__((ext_vector_type(4)));
typedef int32_t bm_i32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t kBmThreads = 256;  // 4 waves, each its own queries
constexpr uint32_t kBmWindow = 8192;  // train rows per split: 13-bit index in the key
constexpr int32_t kBmBias = 1 << 20;
constexpr int32_t kBmPastEnd = 1 << 30;

// 16 bits of a 32-bit word to 16 int8 signs, +64 (0x40) for a 0 bit and
// -64 (0xC0) for a 1 bit.  `v` is the word shifted left by 4 (l / 32): lane
// half 0 takes bits 4..7 of every byte, half 1 bits 0..3; int32 j of the
// result holds bit 7 - j of each byte of `v`, moved to bit 7 of the byte.
__device__ inline bm_i32x4 bm_signs16(uint32_t v) {
    bm_i32x4 r;
    r.x = static_cast<int32_t>((v & 0x80808080u) | 0x40404040u);
    r.y = static_cast<int32_t>(((v << 1) & 0x80808080u) | 0x40404040u);
    r.z = static_cast<int32_t>(((v << 2) & 0x80808080u) | 0x40404040u);
    r.w = static_cast<int32_t>(((v << 3) & 0x80808080u) | 0x40404040u);
    return r;
}

// Middle value of three, written as max(min(a, b), min(max(a, b), c)),
// the form the backend folds into one v_med3_i32.  Plain code rather than
// inline asm: the reads of a matrix-core result stay visible to the
// compiler, whose hazard recognizer inserts the wait states such a read
// needs.
__device__ inline int32_t med3_i32(int32_t a, int32_t b, int32_t c) {
    const int32_t lo = a < b ? a : b, hi = a < b ? b : a;
    const int32_t m = hi < c ? hi : c;
    return lo < m ? m : lo;
}

// Key relative to the tile at `tb` (8192 d + t - tb - 2^20) to the key the
// merge compares across splits (d << kBfKeyShift | t); kept keys of rows
// past the end (or none) become 0xFFFFFFFF.
__device__ inline uint32_t bm_global_key(int32_t k, uint32_t tb, uint32_t t0) {
    if (k >= kBmPastEnd - kBmBias - static_cast<int32_t>(kBmWindow)) return 0xFFFFFFFFu;
    const uint32_t u = static_cast<uint32_t>(k + kBmBias) + (tb - t0);  // 8192 d + t - t0
    return ((u >> 13) << kBfKeyShift) | (t0 + (u & (kBmWindow - 1)));
}

template <uint32_t Q>
struct BmQueries {
    bm_i32x4 b[Q][8];   // sign operands of the wave's 32 Q queries
    int32_t m1[Q], m2[Q];  // the two smallest keys per query (this lane's rows)
};

// Moves the kept keys from one tile's base to the next (32 rows on).
template <uint32_t Q>
__device__ __forceinline__ void bm_shift(BmQueries<Q>& st) {
#pragma unroll
    for (uint32_t g = 0; g < Q; ++g) {
        st.m1[g] -= 32;
        st.m2[g] -= 32;
    }
}

// One tile of 32 train rows against the wave's queries: the train operands
// of the tile's 8 steps from LDS (`sa`), the products into the two smallest
// keys per query.  `crow` is the accumulator input (each row's offset in the
// tile); rows at offsets >= `valid` (a split's ragged last tile) get
// kBmPastEnd added before the fold.
template <uint32_t Q>
__device__ __forceinline__ void bm_tile_lds(BmQueries<Q>& st, const bm_i32x4 (*sa)[64],
                                            uint32_t lane, const bm_i32x16& crow,
                                            uint32_t valid) {
    bm_i32x16 acc[Q];
#pragma unroll
    for (uint32_t s = 0; s < 8; ++s) {
        const bm_i32x4 a = sa[s][lane];
#pragma unroll
        for (uint32_t g = 0; g < Q; ++g)
            acc[g] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, st.b[g][s], s ? acc[g] : crow, 0,
                                                           0, 0);
    }
    if (valid < 32) {  // workgroup-uniform
#pragma unroll
        for (uint32_t g = 0; g < Q; ++g)
#pragma unroll
            for (uint32_t r = 0; r < 16; ++r)
                acc[g][r] += static_cast<uint32_t>(crow[r]) >= valid ? kBmPastEnd : 0;
    }
#pragma unroll
    for (uint32_t g = 0; g < Q; ++g) {
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r) {
            const int32_t k = acc[g][r];
            const int32_t n1 = st.m1[g] < k ? st.m1[g] : k;
            st.m2[g] = med3_i32(st.m1[g], st.m2[g], k);
            st.m1[g] = n1;
        }
    }
}

// Workgroup b: split b % splits (consecutive workgroups go to different
// XCDs, so each XCD's L2 holds the rows of few splits), query block
// b / splits of 4 waves x 32 Q queries.
template <uint32_t Q>
__global__ __launch_bounds__(kBmThreads) void k_bf_mfma(const uint32_t* __restrict__ query,
                                                         uint32_t nq,
                                                         const uint32_t* __restrict__ train,
                                                         uint32_t nt, uint32_t split_len,
                                                         uint32_t splits,
                                                         uint2* __restrict__ part) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const uint32_t col = lane & 31;
    const uint32_t half = lane >> 5;
    const uint32_t split = blockIdx.x % splits;
    const uint32_t qb = blockIdx.x / splits;
    const uint32_t t0 = split * split_len;
    const uint32_t t1 = umin(t0 + split_len, nt);
    const uint32_t q0 = (qb * (kBmThreads / 64) + wave) * 32 * Q;
    // Waves past the last query still stage their share of every tile (the
    // barriers below are workgroup-wide); their results are not written.
    // Train operands of a tile, double-buffered: wave w expands steps 2w and
    // 2w + 1 of the tile's 32 rows, every wave reads all 8.
    __shared__ bm_i32x4 sa[2][8][64];

    BmQueries<Q> st;
#pragma unroll
    for (uint32_t g = 0; g < Q; ++g) {
        const uint32_t qi = umin(q0 + 32 * g + col, nq - 1);
        const uint4* src = reinterpret_cast<const uint4*>(query + 8ull * qi);
        const uint4 lo = src[0], hi = src[1];
        const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s) st.b[g][s] = bm_signs16(~w[s] << (4 * half));
        st.m1[g] = st.m2[g] = 0x7FFFFFFF;
    }

    // The running keys are kept relative to the current tile: the
    // accumulator input is the row's offset in the tile, and moving to the
    // next tile lowers the kept keys by 32 (a uniform shift keeps their
    // order; at most 256 tiles, so the initial maximum cannot wrap).  Rows
    // past the split end (last tile only) carry kBmPastEnd.
    uint32_t tb = t0;
    if (t0 < t1) {
        bm_i32x16 crow;
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r)  // + (8192 d - 2^20) from the products
            crow[r] = static_cast<int32_t>(8 * (r >> 2) + 4 * half + (r & 3));
        const uint2* tr = reinterpret_cast<const uint2*>(train);  // 4 word pairs per row
        uint32_t ti = umin(t0 + col, t1 - 1);
        uint2 wp = tr[4ull * ti + wave];  // words 2 wave, 2 wave + 1 of this lane's row
        uint32_t buf = 0;
        for (;; tb += 32) {
            sa[buf][2 * wave][lane] = bm_signs16(wp.x << (4 * half));
            sa[buf][2 * wave + 1][lane] = bm_signs16(wp.y << (4 * half));
            const bool more = tb + 32 < t1;  // workgroup-uniform
            if (more) {
                ti = umin(tb + 32 + col, t1 - 1);  // next tile (clamped: the last row again)
                wp = tr[4ull * ti + wave];
            }
            // The buffer written here was last read two tiles ago, before
            // every wave passed the previous tile's barrier.
            __syncthreads();
            bm_tile_lds<Q>(st, sa[buf], lane, crow, more ? 32u : t1 - tb);
            if (!more) break;
            bm_shift<Q>(st);
            buf ^= 1u;
        }
    }
#pragma unroll
    for (uint32_t g = 0; g < Q; ++g) {
        const int32_t o1 = __shfl_xor(st.m1[g], 32);
        const int32_t o2 = __shfl_xor(st.m2[g], 32);
        const int32_t n1 = st.m1[g] < o1 ? st.m1[g] : o1;
        const int32_t hi1 = st.m1[g] < o1 ? o1 : st.m1[g];
        const int32_t lo2 = st.m2[g] < o2 ? st.m2[g] : o2;
        const int32_t n2 = hi1 < lo2 ? hi1 : lo2;
        const uint32_t qi = q0 + 32 * g + col;
        if (half == 0 && qi < nq)
            part[static_cast<size_t>(split) * nq + qi] =
                make_uint2(bm_global_key(n1, tb, t0), bm_global_key(n2, tb, t0));
    }
}

// ---------------------------------------------------------------------------
// k_export: the FeatureSet of a device feature array (the FromGpuModel loop
// of LoadDataCPUPass::loadFeatures, load_features_cpu.cpp:164-188) written
// by the GPU straight into pinned host memory: per level min(count, cap)
// wslam::Feature records of 13 dwords {x, y, lod, strength, orientation,
// descriptor[8]}, levels back to back, and the per-level counts.  Block
// (level, chunk) writes the records [256 chunk, 256 chunk + 256) of its
// level; its threads store consecutive dwords of that output span, so
// every store instruction covers 1 KB of contiguous host memory.
// ---------------------------------------------------------------------------
constexpr uint32_t kExportThreads = 256;
constexpr uint32_t kFeatureDwords = 13;  // sizeof(wslam::Feature) / 4

__global__ __launch_bounds__(kExportThreads) void k_export(const uint8_t* __restrict__ feat,
                                                            uint32_t levels, uint32_t cap,
                                                            uint32_t stride, uint32_t chunks,
                                                            uint32_t* __restrict__ out,
                                                            uint32_t* __restrict__ counts) {
    const uint32_t l = blockIdx.x / chunks;
    const uint32_t chunk = blockIdx.x - l * chunks;
    // Records of the levels before l: the output offset of this level.
    uint32_t base = 0, kept = 0;
    for (uint32_t k = 0; k <= l; ++k) {
        const uint32_t c = umin(load_u32(reinterpret_cast<const uint32_t*>(feat + k * stride), 0),
                                cap);
        if (k < l) base += c;
        else kept = c;
    }
    if (blockIdx.x == 0 && threadIdx.x < levels)
        counts[threadIdx.x] =
            umin(load_u32(reinterpret_cast<const uint32_t*>(feat + threadIdx.x * stride), 0), cap);
    const uint32_t r0 = chunk * kExportThreads;
    if (r0 >= kept) return;
    const uint32_t n = umin(kept - r0, kExportThreads);
    const uint8_t* src = feat + l * stride + 8 + 48u * r0;
    uint32_t* dst = out + static_cast<size_t>(base + r0) * kFeatureDwords;
    for (uint32_t e = threadIdx.x; e < n * kFeatureDwords; e += kExportThreads) {
        const uint32_t rec = e / kFeatureDwords, field = e - rec * kFeatureDwords;
        // source dword of the 48-byte record: x, y, strength, orientation,
        // descriptor[8]; field 2 (lod) has none
        const uint32_t sw = field < 2 ? field : (field == 2 ? 0 : field - 1);
        const uint32_t v = load_u32(reinterpret_cast<const uint32_t*>(src + 48u * rec), sw);
        uint32_t o;
        if (field < 2) o = static_cast<uint32_t>(__uint_as_float(v));  // u32(f32), models.hpp:113-120
        else if (field == 2) o = l;
        else o = v;
        dst[e] = o;
    }
}

// ---------------------------------------------------------------------------
// k_ransac_score: RANSAC hypothesis scoring (RansacCPU::execute's
// ClassifyInliers with SampsonDistance, wslam/src/pose_estimate_cpu/
// ransac_cpu.cpp:178-209) for many fundamental matrices at once.  Workgroup
// h scores hypothesis h against every correspondence: the Sampson distance
// of (prev, curr) in f64 with the reference's operation order (F x_prev and
// F^T x_curr as left-to-right row sums, x_curr . F x_prev, the four squares
// summed left to right; a denominator below 1e-12 gives +inf), compared
// with the threshold; the inlier count is reduced over the workgroup and
// written once.  With a mask, mask[h * n + i] = 1 for each inlier.
// ---------------------------------------------------------------------------
constexpr uint32_t kRansacThreads = 1024;

__device__ inline bool sampson_inlier(const double* f, double px, double py, double cx,
                                      double cy, double thr) {
    const double fx0 = f[0] * px + f[1] * py + f[2] * 1.0;
    const double fx1 = f[3] * px + f[4] * py + f[5] * 1.0;
    const double fx2 = f[6] * px + f[7] * py + f[8] * 1.0;
    const double ft0 = f[0] * cx + f[3] * cy + f[6] * 1.0;
    const double ft1 = f[1] * cx + f[4] * cy + f[7] * 1.0;
    const double num = cx * fx0 + cy * fx1 + 1.0 * fx2;
    const double denom = fx0 * fx0 + fx1 * fx1 + ft0 * ft0 + ft1 * ft1;
    if (denom < 1e-12) return false;  // distance +inf
    return (num * num) / denom < thr;
}

__global__ __launch_bounds__(kRansacThreads) void k_ransac_score(
    const double* __restrict__ F, const double2* __restrict__ prev,
    const double2* __restrict__ curr, uint32_t n, double thr, uint32_t* __restrict__ counts,
    uint8_t* __restrict__ mask) {
    __shared__ uint32_t wsum[kRansacThreads / 64];
    const uint32_t h = blockIdx.x;
    double f[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i] = F[9 * h + i];
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kRansacThreads) {
        const double2 p = prev[i], q = curr[i];
        const bool in = sampson_inlier(f, p.x, p.y, q.x, q.y, thr);
        c += in ? 1u : 0u;
        if (mask) mask[static_cast<size_t>(h) * n + i] = in ? 1 : 0;
    }
    const uint32_t t = wave_total(c);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
#pragma unroll
        for (uint32_t w = 0; w < kRansacThreads / 64; ++w) total += wsum[w];
        counts[h] = total;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from orbfe_api.cpp).
// ---------------------------------------------------------------------------
// Tables of the describe kernel (c_desc_ppt8, c_desc_tests, g_desc_order).
struct DescTables {
    int8_t pts[2][kDescPts][2];      // position of point id p of group g
    uint32_t ppt8[kDescSlots];        // the device form, by slot (c_desc_ppt8)
    uint32_t tests[2][128];
    uint8_t order[3][kDescOrderBytes];
};

// Split of the 256 tests into two groups of 128, 4 of every aligned run of 8
// in each, with few distinct pattern positions per group: starting from the
// tests of each parity, swap a random test of group 0 with one of group 1 in
// the same run of 8 whenever that does not raise (larger group's positions,
// total positions); a fixed xorshift seed makes the split the same on every
// build.  group[t] is the group of test t; returns each group's position
// count.
static void split_tests(uint8_t group[256], uint32_t npos[2]) {
    // pattern position (x, y in [-13, 13]) -> key in [0, 27 * 27)
    const auto key = [](int x, int y) { return static_cast<uint32_t>((y + 13) * 27 + x + 13); };
    uint32_t pa[256], pb[256];
    for (uint32_t t = 0; t < 256; ++t) {
        pa[t] = key(kOrbfeBriefPattern[t][0], kOrbfeBriefPattern[t][1]);
        pb[t] = key(kOrbfeBriefPattern[t][2], kOrbfeBriefPattern[t][3]);
    }
    static uint16_t cnt[2][27 * 27];
    std::fill(&cnt[0][0], &cnt[0][0] + 2 * 27 * 27, 0);
    npos[0] = npos[1] = 0;
    const auto add = [&](uint32_t g, uint32_t k, int d) {
        if (d > 0 && cnt[g][k]++ == 0) ++npos[g];
        if (d < 0 && --cnt[g][k] == 0) --npos[g];
    };
    const auto put = [&](uint32_t t, uint32_t g, int d) {
        add(g, pa[t], d);
        add(g, pb[t], d);
    };
    for (uint32_t t = 0; t < 256; ++t) {
        group[t] = static_cast<uint8_t>(t & 1u);
        put(t, group[t], 1);
    }
    const auto cost = [&]() { return std::max(npos[0], npos[1]) * 1024u + npos[0] + npos[1]; };
    uint64_t st = 0x9E3779B97F4A7C15ull;
    const auto rnd = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return st;
    };
    for (uint32_t it = 0; it < 300000; ++it) {
        const uint32_t run = static_cast<uint32_t>(rnd() % 32) * 8;
        uint32_t in0[4], in1[4], n0 = 0, n1 = 0;
        for (uint32_t t = run; t < run + 8; ++t) (group[t] ? in1[n1++] : in0[n0++]) = t;
        const uint32_t a = in0[rnd() % 4], b = in1[rnd() % 4];
        const uint32_t before = cost();
        put(a, 0, -1); put(a, 1, 1); put(b, 1, -1); put(b, 0, 1);
        if (cost() > before) {
            put(a, 1, -1); put(a, 0, 1); put(b, 0, -1); put(b, 1, 1);
        } else {
            group[a] = 1;
            group[b] = 0;
        }
    }
}

// Gather order of one lanes-per-keypoint variant: for each angle bin,
// keypoint alignment a (x mod 8 = a & 7, y mod 4 = a >> 3) and group, the
// group's kDescPts point ids sorted by the 8 x 4 block of the smoothed
// layout (band, then block column) that their position rotated to the bin's
// centre angle reaches from a keypoint of that alignment, then by row and
// column; lane s takes points s, s + LANES, ... of that order (one load of
// the group's lanes: consecutive points, each run of 16 transposed so that
// the 4 lanes of a quad take every fourth point of it, which measured
// faster), stored as their slots.  The order only decides which lane loads
// which point, never a value.
static void build_desc_order(const int8_t (*pts)[kDescPts][2], uint8_t* out, uint32_t lanes,
                             uint32_t row) {
    const uint32_t ppl = kDescPts / lanes;
    for (uint32_t bin = 0; bin < kAngleBins; ++bin) {
        const double th = -3.14159265358979323846 + (bin + 0.5) * (2.0 * 3.14159265358979323846 / kAngleBins);
        const double c = std::cos(th), sn = std::sin(th);
        for (uint32_t a = 0; a < kOrderAligns; ++a) {
            for (uint32_t g = 0; g < 2; ++g) {
                uint32_t seq[kDescPts];
                long key[kDescPts];
                for (uint32_t q = 0; q < kDescPts; ++q) {
                    const double px = pts[g][q][0], py = pts[g][q][1];
                    const long rx = std::lround(px * c - py * sn), ry = std::lround(px * sn + py * c);
                    // grid position from a keypoint at (64 + a mod 8, 64 + a / 8)
                    const long gx = 64 + (a & 7u) + rx, gy = 64 + (a >> 3) + ry;
                    key[q] = (((gy >> 2) * 32 + (gx >> 3)) * 256 + (ry + 64)) * 256 + (rx + 64);
                    seq[q] = q;
                }
                std::stable_sort(seq, seq + kDescPts,
                                 [&](uint32_t u, uint32_t v) { return key[u] < key[v]; });
                // each run of 16 (8) consecutive ids transposed as a 4 x 4
                // (4 x 2) block: lane 4i + j of a load takes point i + 4j
                // (i + 2j) of the run
                uint32_t tr[kDescPts];
                const uint32_t run = lanes < 16 ? lanes : 16, per = run / 4;
                for (uint32_t i = 0; i < kDescPts; ++i)
                    tr[i] = seq[i - i % run + (i % run) % 4 * per + (i % run) / 4];
                std::copy(tr, tr + kDescPts, seq);
                uint8_t* dst = out + ((static_cast<size_t>(bin) * kOrderAligns + a) * 2 + g) * lanes * row;
                for (uint32_t s = 0; s < lanes; ++s)
                    for (uint32_t k = 0; k < row; ++k)
                        dst[s * row + k] = k < ppl ? static_cast<uint8_t>(desc_slot(seq[lanes * k + s])) : 0;
            }
        }
    }
}

// Returns false if a group needs more than kDescPts points (the pattern in
// brief_pattern.h needs 189 each).
static bool build_desc_tables(DescTables* d) {
    uint8_t group[256];
    uint32_t npos[2];
    split_tests(group, npos);
    if (npos[0] > kDescPts || npos[1] > kDescPts) return false;
    for (uint32_t g = 0; g < 2; ++g) {
        int id_of[27 * 27];
        std::fill(id_of, id_of + 27 * 27, -1);
        uint32_t n = 0, m = 0;
        const auto id = [&](int x, int y) -> uint32_t {
            int& e = id_of[(y + 13) * 27 + x + 13];
            if (e < 0) {
                e = static_cast<int>(n);
                d->pts[g][n][0] = static_cast<int8_t>(x);
                d->pts[g][n][1] = static_cast<int8_t>(y);
                ++n;
            }
            return static_cast<uint32_t>(e);
        };
        for (uint32_t t = 0; t < 256; ++t) {
            if (group[t] != g) continue;
            const int8_t* pt = kOrbfeBriefPattern[t];
            const uint32_t qa = id(pt[0], pt[1]), qb = id(pt[2], pt[3]);
            d->tests[g][m++] = 4 * desc_slot(qa) | 4 * desc_slot(qb) << 16 | (t & 31u) << 26;
        }
        if (m != 128) return false;
        for (uint32_t q = n; q < kDescPts; ++q) {  // unread slots: a valid position
            d->pts[g][q][0] = d->pts[g][0][0];
            d->pts[g][q][1] = d->pts[g][0][1];
        }
    }
    for (uint32_t q = 0; q < kDescSlots; ++q) d->ppt8[q] = 0;
    for (uint32_t p = 0; p < kDescPts; ++p) {
        uint32_t v = 0;
        for (uint32_t g = 0; g < 2; ++g)
            for (uint32_t c = 0; c < 2; ++c)
                v |= static_cast<uint32_t>(static_cast<uint8_t>(d->pts[g][p][c])) << (8 * (2 * g + c));
        d->ppt8[desc_slot(p)] = v;
    }
    build_desc_order(d->pts, d->order[0], 8, desc_order_row<8>());
    build_desc_order(d->pts, d->order[1], 16, desc_order_row<16>());
    build_desc_order(d->pts, d->order[2], 32, desc_order_row<32>());
    return true;
}

hipError_t upload_pattern() {
    static DescTables tables;
    static int built = 0;  // 1 built, -1 failed
    if (!built) built = build_desc_tables(&tables) ? 1 : -1;
    if (built < 0) return hipErrorInvalidValue;
    float unit[256];
    for (uint32_t v = 0; v < 256; ++v) unit[v] = static_cast<float>(v) / 255.0f;
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_u8_unit), unit, sizeof(unit));
    if (e == hipSuccess)
        e = hipMemcpyToSymbol(HIP_SYMBOL(c_desc_ppt8), tables.ppt8, sizeof(tables.ppt8));
    if (e == hipSuccess)
        e = hipMemcpyToSymbol(HIP_SYMBOL(c_desc_tests), tables.tests, sizeof(tables.tests));
    if (e == hipSuccess)
        e = hipMemcpyToSymbol(HIP_SYMBOL(g_desc_order), tables.order, sizeof(tables.order));
    return e;
}

hipError_t launch_detect(uint32_t tile_h, uint32_t blocks, BatchView bv, const Job& job0,
                         const Params& prm, hipStream_t s) {
    switch (tile_h) {
        case 8: hipLaunchKernelGGL(k_detect<8>, dim3(blocks), dim3(kDetectThreads), 0, s, bv, job0, prm); break;
        case 16:
            // single frames: the small source buffer (measured faster there);
            // batches: the full buffer (no tile reads level 0 from memory)
            if (bv.n == 0)
                hipLaunchKernelGGL((k_detect<16, true>), dim3(blocks), dim3(kDetectThreads), 0, s,
                                   bv, job0, prm);
            else
                hipLaunchKernelGGL((k_detect<16, false>), dim3(blocks), dim3(kDetectThreads), 0, s,
                                   bv, job0, prm);
            break;
        case 32: hipLaunchKernelGGL(k_detect<32>, dim3(blocks), dim3(kDetectThreads), 0, s, bv, job0, prm); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}


hipError_t launch_compact(uint32_t frames, BatchView bv, const Job& job0, uint32_t levels,
                          hipStream_t s) {
    const uint32_t parts = frames == 1 ? 8u : 1u;
    hipLaunchKernelGGL(k_compact, dim3(frames * levels * parts), dim3(kCompactThreads), 0, s, bv,
                       job0, levels, parts);
    return hipGetLastError();
}

// Single-frame launches (no batch table) use the variant with both test
// groups' gathers in flight (latency-bound: a few waves per SIMD); batches
// gather one group at a time (fewer VGPRs, more resident waves).
template <uint32_t LANES>
void launch_describe_t(uint32_t blocks, BatchView bv, const Job& job0, const Params& prm,
                       hipStream_t s) {
    if (bv.n == 0)
        hipLaunchKernelGGL((k_describe<LANES, true>), dim3(blocks), dim3(kDescribeThreads), 0, s,
                           bv, job0, prm);
    else
        hipLaunchKernelGGL((k_describe<LANES, false>), dim3(blocks), dim3(kDescribeThreads), 0, s,
                           bv, job0, prm);
}

hipError_t launch_describe(uint32_t lanes, uint32_t blocks, BatchView bv, const Job& job0,
                           const Params& prm, hipStream_t s) {
    switch (lanes) {
        case 8: launch_describe_t<8>(blocks, bv, job0, prm, s); break;
        case 16: launch_describe_t<16>(blocks, bv, job0, prm, s); break;
        case 32: launch_describe_t<32>(blocks, bv, job0, prm, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_match_index(uint32_t blocks, MatchBatchView bv, const MatchJob& job0,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_match_index, dim3(blocks), dim3(kIndexThreads), 0, s, bv, job0);
    return hipGetLastError();
}

template <uint32_t LANES>
void launch_match_dir_t(bool mutual, uint32_t blocks, MatchBatchView bv, const MatchJob& job0,
                        float ratio, hipStream_t s) {
    if (mutual)
        hipLaunchKernelGGL((k_match_dir<LANES, true>), dim3(blocks), dim3(kMatchThreads), 0, s, bv,
                           job0, ratio);
    else
        hipLaunchKernelGGL((k_match_dir<LANES, false>), dim3(blocks), dim3(kMatchThreads), 0, s,
                           bv, job0, ratio);
}

hipError_t launch_match_dir(uint32_t lanes, bool mutual, uint32_t blocks, MatchBatchView bv,
                            const MatchJob& job0, float ratio, hipStream_t s) {
    switch (lanes) {
        case 16: launch_match_dir_t<16>(mutual, blocks, bv, job0, ratio, s); break;
        case 32: launch_match_dir_t<32>(mutual, blocks, bv, job0, ratio, s); break;
        case 64: launch_match_dir_t<64>(mutual, blocks, bv, job0, ratio, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// VALU issue-rate probe: 8 independent xor/add chains per lane, kProbeIters
// iterations (16 VALU instructions each); the result only feeds an
// impossible comparison so the chains stay live.
constexpr int kProbeIters = 4096;
constexpr uint32_t kProbeVaLU = 16 * kProbeIters;  // VALU instructions per wave

__global__ __launch_bounds__(256) void k_valu_probe(uint32_t* sink, uint32_t seed) {
    uint32_t a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = seed + threadIdx.x * 8 + k;
    for (int i = 0; i < kProbeIters; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = (a[k] ^ 0x9e3779b9u) + 0x7f4a7c15u;
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) s ^= a[k];
    if (s == 0x12345678u) sink[blockIdx.x] = s;
}

hipError_t launch_valu_probe(uint32_t* sink, uint32_t blocks, hipStream_t s,
                             double* instr_per_launch) {
    hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, s, sink, 1u);
    *instr_per_launch = static_cast<double>(blocks) * 4.0 * kProbeVaLU;
    return hipGetLastError();
}

// Matrix-core rate probe: two independent v_mfma_i32_32x32x32_i8 chains per
// wave (65536 int8 ops each), the form the brute-force matcher issues.
constexpr uint32_t kProbeMfmaIters = 4096;

__global__ __launch_bounds__(256) void k_mfma_probe(uint32_t* sink, uint32_t seed) {
    const int32_t v = static_cast<int32_t>(seed + threadIdx.x);
    const bm_i32x4 a = {v, v ^ 0x5a5a5a5a, 3, 4};
    const bm_i32x4 b = {5, v, 7, 8};
    bm_i32x16 c0 = {}, c1 = {};
    for (uint32_t i = 0; i < kProbeMfmaIters; ++i) {
        c0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(b, a, c1, 0, 0, 0);
    }
    int32_t x = 0;
#pragma unroll
    for (uint32_t r = 0; r < 16; ++r) x ^= c0[r] + c1[r];
    if (x == 0x12345678) sink[blockIdx.x] = static_cast<uint32_t>(x);
}

hipError_t launch_mfma_probe(uint32_t* sink, uint32_t blocks, hipStream_t s,
                             double* ops_per_launch) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(blocks), dim3(256), 0, s, sink, 1u);
    *ops_per_launch = static_cast<double>(blocks) * 4.0 * 2.0 * kProbeMfmaIters * 65536.0;
    return hipGetLastError();
}

constexpr uint32_t kBmGroups = 4;  // 32-query groups per wave
uint32_t bf_mfma_queries_per_block() { return kBmThreads / 64 * 32 * kBmGroups; }
uint32_t bf_mfma_window() { return kBmWindow; }

hipError_t launch_bf_mfma(const uint32_t* q, uint32_t nq, const uint32_t* t, uint32_t nt,
                          uint32_t splits, uint32_t split_len, uint2* part, hipStream_t s) {
    const uint32_t qb = (nq + bf_mfma_queries_per_block() - 1) / bf_mfma_queries_per_block();
    hipLaunchKernelGGL(k_bf_mfma<kBmGroups>, dim3(qb * splits), dim3(kBmThreads), 0, s, q, nq, t,
                       nt, split_len, splits, part);
    return hipGetLastError();
}

hipError_t launch_bf_merge(const uint2* part, uint32_t nq, uint32_t splits, uint32_t* idx,
                           uint32_t* d1, uint32_t* d2, hipStream_t s) {
    hipLaunchKernelGGL(k_bf_merge, dim3((nq + 255) / 256), dim3(256), 0, s, part, nq, splits, idx,
                       d1, d2);
    return hipGetLastError();
}

hipError_t launch_export(const uint8_t* feat, uint32_t levels, uint32_t cap, uint32_t stride,
                         uint32_t* out, uint32_t* counts, hipStream_t s) {
    const uint32_t chunks = (cap + kExportThreads - 1) / kExportThreads;
    hipLaunchKernelGGL(k_export, dim3(levels * chunks), dim3(kExportThreads), 0, s, feat, levels,
                       cap, stride, chunks, out, counts);
    return hipGetLastError();
}

hipError_t launch_ransac_score(const double* F, uint32_t k, const double* prev,
                               const double* curr, uint32_t n, double thr, uint32_t* counts,
                               uint8_t* mask, hipStream_t s) {
    hipLaunchKernelGGL(k_ransac_score, dim3(k), dim3(kRansacThreads), 0, s, F,
                       reinterpret_cast<const double2*>(prev),
                       reinterpret_cast<const double2*>(curr), n, thr, counts, mask);
    return hipGetLastError();
}

// Partial results of orbfe_match_bf_noalloc_async: a static device array
// (one copy per GPU, held until the process ends) instead of context memory
// grown with hipMalloc.  Host code never passes `scratch` itself to a
// kernel: in host code the name is only a stand-in, and the device address
// comes from bf_scratch_address.
static __device__ __align__(256) uint2 scratch[kBfScratchBytes / sizeof(uint2)];

// Device address of `scratch` on the current device.
hipError_t bf_scratch_address(void** out) {
    return hipGetSymbolAddress(out, HIP_SYMBOL(scratch));
}

}  // namespace orbfe
