// liborbfe host side: plans (level extents and the static cull mask),
// scratch memory, batch tables, kernel timing and the C entry points of
// include/orbfe.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/orbfe.h"
#include "orbfe_device.h"

namespace orbfe {
hipError_t upload_pattern();
hipError_t launch_detect(uint32_t tile_h, uint32_t blocks, BatchView bv, const Job& job0,
                         const Params& prm, hipStream_t s);
hipError_t launch_compact(uint32_t frames, BatchView bv, const Job& job0, uint32_t levels,
                          hipStream_t s);
hipError_t launch_describe(uint32_t lanes, uint32_t blocks, BatchView bv, const Job& job0,
                           const Params& prm, hipStream_t s);
hipError_t launch_match_index(uint32_t blocks, MatchBatchView bv, const MatchJob& job0,
                              hipStream_t s);
hipError_t launch_match_dir(uint32_t lanes, bool mutual, uint32_t blocks, MatchBatchView bv,
                            const MatchJob& job0, float ratio, hipStream_t s);
uint32_t bf_mfma_queries_per_block();
uint32_t bf_mfma_window();
hipError_t launch_bf_mfma(const uint32_t* q, uint32_t nq, const uint32_t* t, uint32_t nt,
                          uint32_t splits, uint32_t split_len, uint2* part, hipStream_t s);
hipError_t launch_bf_merge(const uint2* part, uint32_t nq, uint32_t splits, uint32_t* idx,
                           uint32_t* d1, uint32_t* d2, hipStream_t s);
hipError_t bf_scratch_address(void** out);
hipError_t launch_ransac_score(const double* F, uint32_t k, const double* prev,
                               const double* curr, uint32_t n, double thr, uint32_t* counts,
                               uint8_t* mask, hipStream_t s);
hipError_t launch_export(const uint8_t* feat, uint32_t levels, uint32_t cap, uint32_t stride,
                         uint32_t* out, uint32_t* counts, hipStream_t s);
hipError_t launch_valu_probe(uint32_t* sink, uint32_t blocks, hipStream_t s,
                             double* instr_per_launch);
hipError_t launch_mfma_probe(uint32_t* sink, uint32_t blocks, hipStream_t s,
                             double* ops_per_launch);
}  // namespace orbfe

using orbfe::BatchView;
using orbfe::Job;
using orbfe::LevelGeom;
using orbfe::MatchBatchView;
using orbfe::MatchJob;
using orbfe::Params;
using orbfe::Plan;

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t v, size_t a = kAlign) { return (v + a - 1) / a * a; }

uint32_t ceil_div(uint64_t n, uint64_t d) { return static_cast<uint32_t>((n + d - 1) / d); }

// Level extents as the reference creates its textures:
// width = uint32(W0 / pow(scale, lod)) (wslam/src/common.cpp:38-45).
void level_dims(uint32_t W0, uint32_t H0, uint32_t L, double scale, uint32_t* w, uint32_t* h) {
    for (uint32_t lod = 0; lod < L; ++lod) {
        const double factor = lod == 0 ? 1.0 : std::pow(scale, static_cast<double>(lod));
        w[lod] = static_cast<uint32_t>(W0 / factor);
        h[lod] = static_cast<uint32_t>(H0 / factor);
    }
}

// Whether a W0 x H0 frame's per-frame buffers stay within the 32-bit byte
// offsets the kernels address them with: the u8 frame (W0 * H0), the
// pyramid (f32, levels >= 1), the smoothed images (f32, smooth_floats per
// level) and the survivor words (u64, ceil(w / 64) per row).
bool frame_fits_offsets(uint32_t W0, uint32_t H0, uint32_t L, double scale) {
    uint32_t dw[ORBFE_MAX_LEVELS], dh[ORBFE_MAX_LEVELS];
    level_dims(W0, H0, L, scale, dw, dh);
    uint64_t pyr = 0, sm = 0, words = 0;
    for (uint32_t l = 0; l < L; ++l) {
        if (l > 0) pyr += static_cast<uint64_t>(dw[l]) * dh[l];
        sm += static_cast<uint64_t>(4u * ((dw[l] + orbfe::kSmoothExt + 7u) / 8u * 8u)) *
              ((dh[l] + orbfe::kSmoothExt + 3u) / 4u);
        words += static_cast<uint64_t>(dh[l]) * ((dw[l] + 63u) / 64u);
    }
    constexpr uint64_t kMax = UINT32_MAX;
    return static_cast<uint64_t>(W0) * H0 <= kMax && 4 * pyr <= kMax && 4 * sm <= kMax &&
           8 * words <= kMax;
}

// Whether the kernels' 32-bit offsets y * pitch + x reach every byte of a
// frame's rows (pitch >= width, checked first).
bool pitch_fits_offsets(size_t pitch, uint32_t w, uint32_t h) {
    if (h == 0) return true;  // an empty frame is refused with its plan
    return pitch <= UINT32_MAX &&
           static_cast<uint64_t>(pitch) * (h - 1) + w <= static_cast<uint64_t>(UINT32_MAX);
}

// Static part of the corner cull of cull_corners.wgsl for one level, as a
// bitmap in the survivor-word layout (bit x&63 of word y*wpr + x/64).
//
// main_horizontal (cull_corners.wgsl:21-57) runs over gid in
// [0, 64*ceil(ceil(W0/3)/64)) x [0, H0) (cull_corners.cpp:332-341) and
// starts at s = 9*gid.x + gid.y*w.  For s >= w its loop is empty and it
// copies S[s] alone; for s < w it keeps the first maximum of S over
// [s, min(s+3, w)).  main_vertical (cull_corners.wgsl:60-98) runs over
// [0, 64*ceil(W0/64)) x [0, ceil(H0/3)) (cull_corners.cpp:343-350), starts at
// s = 3*gid.x + 3*gid.y*w and copies T at s and at s + k*w for
// k < min((W0*H0 - s)/w, 3) (getVerticalEnd, cull_corners.wgsl:101-109).
// A survivor is a position p < w*h with O[p] > 0, so for p >= w the static
// condition is "p is a horizontal start and p is vertically copied", and for
// p < w it is "p lies in a 3-wide window at a multiple of 9 and p is
// vertically copied"; the window maximum is evaluated on the scores.
void build_level_mask(uint32_t W0, uint32_t H0, uint32_t w, uint32_t h, uint32_t wpr,
                      uint64_t* words) {
    const uint64_t N0 = static_cast<uint64_t>(W0) * H0;
    const uint64_t n = static_cast<uint64_t>(w) * h;
    std::vector<uint8_t> hor(n, 0), ver(n, 0);
    const uint64_t hor_gx = 64ull * ceil_div(ceil_div(W0, 3), 64);
    for (uint64_t gy = 0; gy < H0; ++gy) {
        for (uint64_t gx = 0; gx < hor_gx; ++gx) {
            const uint64_t s = 9 * gx + gy * w;
            if (s >= n) break;
            hor[s] = 1;
        }
    }
    const uint64_t ver_gx = 64ull * ceil_div(W0, 64);
    const uint64_t ver_gy = ceil_div(H0, 3);
    for (uint64_t gy = 0; gy < ver_gy; ++gy) {
        for (uint64_t gx = 0; gx < ver_gx; ++gx) {
            const uint64_t s = 3 * gx + 3 * gy * w;
            if (s >= n) break;
            ver[s] = 1;
            const uint64_t steps = std::min<uint64_t>((N0 - s) / w, 3);
            for (uint64_t k = 0; k < steps; ++k)
                if (s + k * w < n) ver[s + k * w] = 1;
        }
    }
    std::fill(words, words + static_cast<size_t>(h) * wpr, 0ull);
    for (uint64_t p = 0; p < n; ++p) {
        const bool keep = p < w ? (ver[p] && (p % 9) < 3) : (ver[p] && hor[p]);
        if (!keep) continue;
        const uint64_t y = p / w, x = p % w;
        words[y * wpr + x / 64] |= 1ull << (x % 64);
    }
}

orbfe::TapEntry tap_entry(const orbfe::AxisTaps& t) {
    return {t.i0 | (t.i1 << 16), t.frac};
}

// Tap tables of one level (layout: level_tap_offsets).  The host computes
// them with the same f32 operations as the device code (correctly rounded
// division, no contraction), so the entries are the ones the kernels would
// compute: level texels sample level 0 at uv = (c + 0.5) / n
// (create_pyramid.wgsl:5-14); the 9-tap Gaussian samples the level at
// uv = (c + 0.5) * texel + off * texel (generate_features.wgsl:149-172).
void build_level_taps(const Plan& P, const LevelGeom& G, orbfe::TapEntry* out) {
    uint32_t lc, lr, sc, sr;
    orbfe::level_tap_offsets(0, G.w, G.h, &lc, &lr, &sc, &sr);
    for (uint32_t c = 0; c < G.w; ++c)
        out[lc + c] = tap_entry(orbfe::axis_taps(
            (static_cast<float>(c) + 0.5f) / static_cast<float>(G.w), P.W0));
    for (uint32_t c = 0; c < G.h; ++c)
        out[lr + c] = tap_entry(orbfe::axis_taps(
            (static_cast<float>(c) + 0.5f) / static_cast<float>(G.h), P.H0));
    const float offs[3] = {-1.2f, 0.0f, 1.2f};
    const uint32_t sw = G.w + orbfe::kSmoothExt, sh = G.h + orbfe::kSmoothExt;
    for (uint32_t i = 0; i < 3; ++i) {
        for (uint32_t x = 0; x < sw; ++x) {
            const float uv = (static_cast<float>(x) + 0.5f) * G.texel_x;
            out[sc + i * sw + x] = tap_entry(orbfe::axis_taps(uv + offs[i] * G.texel_x, G.w));
        }
        for (uint32_t y = 0; y < sh; ++y) {
            const float uv = (static_cast<float>(y) + 0.5f) * G.texel_y;
            out[sr + i * sh + y] = tap_entry(orbfe::axis_taps(uv + offs[i] * G.texel_y, G.h));
        }
    }
    const uint32_t rc = orbfe::level_row_code_offset(0, G.w, G.h);
    for (uint32_t y = 0; y < sh; ++y)
        out[rc + y] = {orbfe::smooth_tap_code(out[sr + y].i01, out[sr + sh + y].i01,
                                              out[sr + 2 * sh + y].i01,
                                              static_cast<int32_t>(y), G.h),
                       0.0f};
}

// Tile descriptors of one level (TileDesc): the source rectangle of level 0
// that a tile's staged texels (70 columns x TH + 6 rows from (x0 - 3, y0 -
// 3), coordinates clamped by the textureLoad rule) sample.  Staged columns
// and rows with g >= 0 map monotonically onto the frame, so the rectangle's
// corners are the taps of the first and last staged column and row;
// negative g, which Tint sends to the far edge, lie outside it and are
// gathered by the kernel.
void build_level_tiles(const Plan& P, const LevelGeom& G, uint32_t l,
                       const orbfe::TapEntry* taps, bool border_zero, orbfe::TileDesc* out) {
    const int32_t th = static_cast<int32_t>(P.tile_h);
    const int32_t lh = th + 2 * orbfe::kHalo;
    const int32_t w = static_cast<int32_t>(G.w), h = static_cast<int32_t>(G.h);
    uint32_t lc, lr, sc, sr;
    orbfe::level_tap_offsets(G.tap_off, G.w, G.h, &lc, &lr, &sc, &sr);
    const auto clamp_coord = [&](int32_t g, uint32_t n) -> uint32_t {
        const bool inside = g >= 0 && g < static_cast<int32_t>(n);
        return border_zero ? (inside ? static_cast<uint32_t>(g) : 0u) : orbfe::tint_clamp(g, n);
    };
    for (uint32_t ty = 0; ty < G.tiles_y; ++ty) {
        for (uint32_t tx = 0; tx < G.tiles_x; ++tx) {
            const int32_t x0 = static_cast<int32_t>(tx) * orbfe::kTileW;
            const int32_t y0 = static_cast<int32_t>(ty) * th;
            const int32_t c_first = x0 >= orbfe::kHalo ? 0 : orbfe::kHalo - x0;
            const int32_t r_first = y0 >= orbfe::kHalo ? 0 : orbfe::kHalo - y0;
            int32_t c_last = orbfe::kLdsW - 1, r_last = lh - 1;
            if (border_zero) {
                const int32_t cl = w - 1 - x0 + orbfe::kHalo, rl = h - 1 - y0 + orbfe::kHalo;
                c_last = cl < orbfe::kLdsW - 1 ? (cl < 0 ? 0 : cl) : orbfe::kLdsW - 1;
                r_last = rl < lh - 1 ? (rl < 0 ? 0 : rl) : lh - 1;
            }
            uint32_t colmin = clamp_coord(x0 - orbfe::kHalo + c_first, G.w);
            uint32_t colmax = clamp_coord(x0 - orbfe::kHalo + c_last, G.w);
            uint32_t rowmin = clamp_coord(y0 - orbfe::kHalo + r_first, G.h);
            uint32_t rowmax = clamp_coord(y0 - orbfe::kHalo + r_last, G.h);
            if (l > 0) {
                colmin = taps[lc + colmin].i01 & 0xFFFFu;
                colmax = taps[lc + colmax].i01 >> 16;
                rowmin = taps[lr + rowmin].i01 & 0xFFFFu;
                rowmax = taps[lr + rowmax].i01 >> 16;
            }
            const uint32_t cs = colmin & ~3u;
            const uint32_t nd = colmax >= cs ? (colmax - cs) / 4 + 1 : 0;
            const uint32_t nrows = rowmax >= rowmin ? rowmax - rowmin + 1 : 0;
            const bool fits = colmax >= colmin && rowmax >= rowmin && nd <= 64 &&
                              nd * nrows <= orbfe::detect_src_words(P.tile_h);
            const bool fits_small = fits && nd * nrows <= orbfe::detect_src_words_small(P.tile_h);
            orbfe::TileDesc& d = out[ty * G.tiles_x + tx];
            d.lxy = l | (tx << 4) | (ty << 16);
            d.cs = cs;
            d.rowmin = rowmin;
            d.nd_rows = fits ? (nd | (nrows << 16) | (fits_small ? 1u << 30 : 0u) | (1u << 31)) : 0u;
        }
    }
}

struct PlanEntry {
    Plan host{};
    Plan* dev = nullptr;
    uint64_t* mask = nullptr;
    orbfe::TapEntry* taps = nullptr;
    orbfe::TileDesc* tiles = nullptr;
    orbfe::TileDesc* tiles_single = nullptr;  // dispatch order of single-frame launches
    // per-frame scratch sizes (bytes), 256-aligned
    size_t surv_bytes = 0, kp_bytes = 0, pyr_bytes = 0, feat_bytes = 0, match_bytes = 0;
    size_t sm_bytes = 0, index_bytes = 0;
    size_t pair_bytes = 0;  // per pair matcher scratch: the query list

    PlanEntry() = default;
    PlanEntry(const PlanEntry&) = delete;
    PlanEntry& operator=(const PlanEntry&) = delete;
    // Frees the plan's device buffers (on the current device: the context
    // that owns the plan selects its device first).
    ~PlanEntry() {
        for (void* p : {static_cast<void*>(dev), static_cast<void*>(mask),
                        static_cast<void*>(taps), static_cast<void*>(tiles),
                        static_cast<void*>(tiles_single)})
            if (p) (void)hipFree(p);
    }
};

struct Event2 {
    std::string name;
    hipEvent_t a, b;
};

struct StageSlot {
    void* host = nullptr;
    void* dev = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};

}  // namespace

struct orbfe_ctx {
    int device = 0;
    orbfe_config cfg{};
    Params prm{};
    uint32_t match_lanes_batch = orbfe::kMatchLanesBatch;  // matcher lanes per query
    uint32_t match_lanes_pair = orbfe::kMatchLanesPair;
    uint32_t desc_lanes_batch = orbfe::kDescLanesDefault;  // describe lanes per keypoint
    uint32_t desc_lanes_frame = orbfe::kDescLanesFrame;
    std::string err;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, std::unique_ptr<PlanEntry>> plans;
    PlanEntry* main = nullptr;
    hipStream_t stream = nullptr;  // for the host-memory entry points
    // single-frame scratch
    uint64_t* surv = nullptr;
    uint32_t* kp = nullptr;
    float* pyr_scratch = nullptr;
    float* smooth = nullptr;
    uint32_t* q_list = nullptr;  // the pipeline matcher's query list
    uint32_t* index[2] = {nullptr, nullptr};
    // host-memory entry points: device staging
    uint8_t* d_img = nullptr;
    float* d_pyr = nullptr;
    uint8_t* d_feat[2] = {nullptr, nullptr};
    int32_t* d_match = nullptr;
    // brute-force partials
    uint2* bf_part = nullptr;
    size_t bf_part_bytes = 0;
    uint2* bf_static = nullptr;  // the device's static scratch (bf_scratch_address)
    // dynamic batches
    StageSlot ring[4];
    uint32_t ring_next = 0;
    uint8_t* arena = nullptr;
    size_t arena_bytes = 0;
    hipStream_t arena_stream = nullptr;
    // timing
    bool timing = false;
    std::vector<Event2> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<uint32_t, double>> totals;
};

struct orbfe_batch {
    orbfe_ctx* ctx = nullptr;
    uint32_t n = 0;
    uint32_t tiles = 0, desc_blocks = 0;
    void* table = nullptr;  // device: jobs, tile prefix, desc prefix
    uint8_t* arena = nullptr;
    BatchView view{};
};

struct orbfe_match_batch {
    orbfe_ctx* ctx = nullptr;
    uint32_t n = 0;
    uint32_t blocks = 0;
    void* table = nullptr;
    uint8_t* arena = nullptr;
    MatchBatchView view{};
};

namespace {

int fail(orbfe_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define HIP_TRY(ctx, expr)                                                          \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return fail((ctx), ORBFE_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(orbfe_ctx* ctx, T** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = kAlign;
    void* v = nullptr;
    const hipError_t e = hipMalloc(&v, bytes);
    if (e != hipSuccess)
        return fail(ctx, ORBFE_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    *p = static_cast<T*>(v);
    return ORBFE_OK;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Device address of the static brute-force scratch of each device, looked
// up once per device (with that device current) on the first orbfe_create
// there.
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;

int scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it == g_scratch_addr.end()) {
        void* p = nullptr;
        if (orbfe::bf_scratch_address(&p) != hipSuccess || !p) return ORBFE_EHIP;
        it = g_scratch_addr.emplace(device, p).first;
    }
    *out = it->second;
    return ORBFE_OK;
}

// Train splits of a brute-force search: at most 8192 rows each (the index
// field of the matrix-core keys), and more when the query blocks alone are
// too few to cover the 256 CUs several times over; no split is empty.
// Partial results take splits * nq * 8 bytes.
void bf_split_plan(uint32_t nq, uint32_t nt, uint32_t* splits_out, uint32_t* len_out) {
    const uint32_t qblocks = ceil_div(nq, orbfe::bf_mfma_queries_per_block());
    uint32_t splits = 1;
    if (nt > 0) {
        const uint32_t want = std::max<uint32_t>(1, 2048 / qblocks);
        const uint32_t max_splits = std::max<uint32_t>(1, nt / 256);
        splits = std::max(ceil_div(nt, orbfe::bf_mfma_window()), std::min(want, max_splits));
    }
    const uint32_t split_len = nt > 0 ? ceil_div(nt, splits) : 0;
    if (nt > 0) splits = ceil_div(nt, split_len);
    *splits_out = splits;
    *len_out = split_len;
}

uint32_t default_cap(uint32_t W0, uint32_t H0) {
    // GPUConst::max_features_per_lod = W*H/250 (wslam/include/wslam/common.hpp:47)
    return static_cast<uint32_t>(static_cast<uint64_t>(W0) * H0 / 250);
}

int get_plan(orbfe_ctx* ctx, uint32_t W0, uint32_t H0, uint32_t cap, PlanEntry** out) {
    if (cap == 0) cap = default_cap(W0, H0);
    const auto key = std::make_tuple(W0, H0, cap);
    auto it = ctx->plans.find(key);
    if (it != ctx->plans.end()) {
        *out = it->second.get();
        return ORBFE_OK;
    }
    const uint32_t L = ctx->cfg.levels;
    if (W0 == 0 || H0 == 0 || W0 > 65535 || H0 > 65535)
        return fail(ctx, ORBFE_ERANGE, "frame %ux%u outside 1..65535", W0, H0);
    if (!frame_fits_offsets(W0, H0, L, ctx->cfg.scale_factor))
        return fail(ctx, ORBFE_ERANGE, "frame %ux%u: per-frame buffers past 32-bit offsets", W0,
                    H0);
    if (cap == 0 || cap >= (1u << orbfe::kMatchIdxBits))
        return fail(ctx, ORBFE_ERANGE, "cap %u outside 1..%u", cap,
                    (1u << orbfe::kMatchIdxBits) - 1);
    uint32_t dw[ORBFE_MAX_LEVELS], dh[ORBFE_MAX_LEVELS];
    level_dims(W0, H0, L, ctx->cfg.scale_factor, dw, dh);
    auto pe = std::make_unique<PlanEntry>();
    Plan& P = pe->host;
    P.W0 = W0;
    P.H0 = H0;
    P.L = L;
    P.cap = cap;
    P.tile_h = ctx->cfg.detect_tile_rows;
    P.desc_lanes = ctx->cfg.describe_lanes;
    P.feat_stride = 8 + 48 * cap;
    P.desc_groups = ceil_div(cap, 256 / P.desc_lanes);
    P.max_dx = 0.1f * static_cast<float>(W0);
    P.max_dy = 0.1f * static_cast<float>(H0);
    P.cell_inv_x = 4.0f / P.max_dx;
    P.cell_inv_y = 2.0f / P.max_dy;
    {
        const auto al4 = [](uint32_t v) { return (v + 3u) & ~3u; };  // 16-byte aligned
        const uint32_t recs = L * cap;
        P.idx_head = al4(L * (orbfe::kCells + 1));
        P.idx_dlo = P.idx_head + 4 * recs;
        P.idx_dhi = P.idx_dlo + 4 * recs;
        P.idx_words = P.idx_dhi + 4 * recs;
    }
    uint32_t tiles = 0, words = 0, pyr = 0, sm = 0, ntaps = 0;
    float scale = 1.0f;
    for (uint32_t l = 0; l < L; ++l) {
        if (dw[l] == 0 || dh[l] == 0)
            return fail(ctx, ORBFE_ERANGE, "level %u of a %ux%u frame is empty", l, W0, H0);
        LevelGeom& G = P.lv[l];
        G.w = dw[l];
        G.h = dh[l];
        G.wpr = ceil_div(G.w, 64);
        G.tiles_x = ceil_div(G.w + orbfe::kSmoothExt, orbfe::kTileW);
        G.tiles_y = ceil_div(G.h + orbfe::kSmoothExt, P.tile_h);
        G.tile_base = tiles;
        tiles += G.tiles_x * G.tiles_y;
        G.pyr_off = l == 0 ? 0 : pyr;
        if (l > 0) pyr += G.w * G.h;
        G.word_off = words;
        words += G.h * G.wpr;
        G.feat_off = l * P.feat_stride;
        G.kp_off = l * cap;
        G.sm_off = sm;
        sm += orbfe::smooth_floats(G.w, G.h);
        G.lod0_scale = scale;  // toLod0 (match_features_cpu.cpp:58-64)
        G.texel_x = 1.0f / static_cast<float>(G.w);
        G.texel_y = 1.0f / static_cast<float>(G.h);
        G.tap_off = ntaps;
        ntaps += orbfe::level_tap_entries(G.w, G.h);
        scale *= static_cast<float>(1.2);
    }
    P.tiles = tiles;
    P.words = words;
    P.pyr_floats = pyr;
    P.sm_floats = sm;
    std::vector<uint64_t> mask(words ? words : 1, 0ull);
    for (uint32_t l = 0; l < L; ++l)
        build_level_mask(W0, H0, P.lv[l].w, P.lv[l].h, P.lv[l].wpr,
                         mask.data() + P.lv[l].word_off);
    int rc = dalloc(ctx, &pe->mask, mask.size() * 8);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpy(pe->mask, mask.data(), mask.size() * 8, hipMemcpyHostToDevice));
    P.mask = pe->mask;
    std::vector<orbfe::TapEntry> taps(ntaps);
    for (uint32_t l = 0; l < L; ++l) build_level_taps(P, P.lv[l], taps.data() + P.lv[l].tap_off);
    rc = dalloc(ctx, &pe->taps, taps.size() * sizeof(orbfe::TapEntry));
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpy(pe->taps, taps.data(), taps.size() * sizeof(orbfe::TapEntry),
                           hipMemcpyHostToDevice));
    P.taps = pe->taps;
    std::vector<orbfe::TileDesc> td(tiles ? tiles : 1);
    for (uint32_t l = 0; l < L; ++l)
        build_level_tiles(P, P.lv[l], l, taps.data(), ctx->prm.border_zero != 0,
                          td.data() + P.lv[l].tile_base);
    rc = dalloc(ctx, &pe->tiles, td.size() * sizeof(orbfe::TileDesc));
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpy(pe->tiles, td.data(), td.size() * sizeof(orbfe::TileDesc),
                           hipMemcpyHostToDevice));
    // Single-frame launches: dispatch index i takes tile S[xcd_block(i, n)],
    // S = the tiles of band 0 of every level, then band 1, ... (band x: tile
    // rows [x * ty / 8, (x + 1) * ty / 8) of a level), so XCD x works on
    // about band x of each level (balanced across levels, one L2 per
    // region).
    {
        std::vector<uint32_t> order;
        order.reserve(tiles);
        for (uint32_t x = 0; x < 8; ++x)
            for (uint32_t l = 0; l < L; ++l) {
                const LevelGeom& G = P.lv[l];
                for (uint32_t ty = x * G.tiles_y / 8; ty < (x + 1) * G.tiles_y / 8; ++ty)
                    for (uint32_t tx = 0; tx < G.tiles_x; ++tx)
                        order.push_back(G.tile_base + ty * G.tiles_x + tx);
            }
        if (order.size() != tiles) return fail(ctx, ORBFE_EHIP, "tile order covers %zu of %u tiles",
                                               order.size(), tiles);
        std::vector<orbfe::TileDesc> ts(td.size());
        for (uint32_t i = 0; i < tiles; ++i) ts[i] = td[order[orbfe::xcd_block(i, tiles)]];
        rc = dalloc(ctx, &pe->tiles_single, ts.size() * sizeof(orbfe::TileDesc));
        if (rc) return rc;
        HIP_TRY(ctx, hipMemcpy(pe->tiles_single, ts.data(), ts.size() * sizeof(orbfe::TileDesc),
                               hipMemcpyHostToDevice));
    }
    rc = dalloc(ctx, &pe->dev, sizeof(Plan));
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpy(pe->dev, &P, sizeof(Plan), hipMemcpyHostToDevice));
    // survivor words, then one survivor count per detect wave (4 per tile)
    pe->surv_bytes = align_up(static_cast<size_t>(words) * 8 + static_cast<size_t>(tiles) * 16);
    pe->kp_bytes = align_up(static_cast<size_t>(L) * cap * 4);
    pe->pyr_bytes = align_up(static_cast<size_t>(pyr) * 4);
    pe->feat_bytes = align_up(static_cast<size_t>(L) * P.feat_stride);
    pe->match_bytes = align_up(static_cast<size_t>(L) * cap * 4);
    pe->sm_bytes = align_up(static_cast<size_t>(sm) * 4);
    pe->index_bytes = align_up(static_cast<size_t>(P.idx_words) * 4);
    pe->pair_bytes = align_up(static_cast<size_t>(orbfe::query_list_words(L, cap)) * 4);
    *out = pe.get();
    ctx->plans.emplace(key, std::move(pe));
    return ORBFE_OK;
}

template <typename F>
hipError_t timed(orbfe_ctx* ctx, const char* name, hipStream_t s, F&& launch) {
    if (!ctx->timing) return launch();
    hipEvent_t ev[2];
    for (auto& e : ev) {
        if (!ctx->pool.empty()) {
            e = ctx->pool.back();
            ctx->pool.pop_back();
        } else {
            const hipError_t err = hipEventCreate(&e);
            if (err != hipSuccess) return err;
        }
    }
    hipError_t err = hipEventRecord(ev[0], s);
    if (err != hipSuccess) return err;
    err = launch();
    if (err != hipSuccess) return err;
    err = hipEventRecord(ev[1], s);
    ctx->pending.push_back({name, ev[0], ev[1]});
    return err;
}

int launch_detect_stage(orbfe_ctx* ctx, const BatchView& bv, const Job& job0, uint32_t tiles,
                        uint32_t frames, uint32_t desc_lanes, uint32_t desc_blocks,
                        hipStream_t s) {
    const Params prm = ctx->prm;
    HIP_TRY(ctx, timed(ctx, "k_detect", s, [&] {
                return orbfe::launch_detect(ctx->cfg.detect_tile_rows, tiles, bv, job0, prm, s);
            }));
    HIP_TRY(ctx, timed(ctx, "k_compact", s, [&] {
                return orbfe::launch_compact(frames, bv, job0, ctx->cfg.levels, s);
            }));
    HIP_TRY(ctx, timed(ctx, "k_describe", s, [&] {
                return orbfe::launch_describe(desc_lanes, desc_blocks, bv, job0, prm, s);
            }));
    return ORBFE_OK;
}

// Matcher blocks of one frame pair of `plan` with `lanes` lanes per query.
uint32_t match_blocks(const Plan& plan, uint32_t lanes) {
    return plan.L * ceil_div(plan.cap, 256 / lanes);
}

int launch_match_stage(orbfe_ctx* ctx, const MatchBatchView& bv, const MatchJob& job0,
                       uint32_t lanes, uint32_t blocks, hipStream_t s) {
    const float ratio = ctx->prm.ratio;
    const uint32_t pairs = bv.n ? bv.n : 1;
    HIP_TRY(ctx, timed(ctx, "k_match_index", s, [&] {
                return orbfe::launch_match_index(pairs * 2 * ctx->cfg.levels, bv, job0, s);
            }));
    HIP_TRY(ctx, timed(ctx, "k_match_dir", s, [&] {
                return orbfe::launch_match_dir(lanes, false, blocks, bv, job0,
                                               ratio, s);
            }));
    HIP_TRY(ctx, timed(ctx, "k_match_mutual", s, [&] {
                return orbfe::launch_match_dir(lanes, true, blocks, bv, job0,
                                               ratio, s);
            }));
    return ORBFE_OK;
}

// Job index of every block of a batched launch (u16 per block), so a block
// finds its frame with one scalar load.
constexpr uint32_t kMaxBatchJobs = 65535;

void fill_block_jobs(const std::vector<uint32_t>& prefix, std::vector<uint16_t>& out) {
    const size_t n = prefix.size() - 1;
    out.assign(prefix[n] + 1, 0);  // even count padding: read as dwords
    for (size_t j = 0; j < n; ++j)
        std::fill(out.begin() + prefix[j], out.begin() + prefix[j + 1], static_cast<uint16_t>(j));
}

// Layout of a detect batch: table (jobs, tile prefix, desc prefix, tile and
// describe block job indices) and per frame scratch in one arena.
struct DetectLayout {
    std::vector<Job> jobs;
    std::vector<uint32_t> tile_prefix, desc_prefix;
    std::vector<uint16_t> tile_job, desc_job;
    size_t table_bytes = 0, arena_bytes = 0;
    size_t off_tiles = 0, off_desc = 0, off_tile_job = 0, off_desc_job = 0;
};

int layout_detect(orbfe_ctx* ctx, const orbfe_frame* frames, uint32_t n, DetectLayout& lay,
                  uint8_t* arena) {
    if (n > kMaxBatchJobs) return fail(ctx, ORBFE_ERANGE, "batch of %u frames > %u", n, kMaxBatchJobs);
    lay.jobs.resize(n);
    lay.tile_prefix.assign(n + 1, 0);
    lay.desc_prefix.assign(n + 1, 0);
    size_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const orbfe_frame& f = frames[i];
        if (!f.image || !f.features)
            return fail(ctx, ORBFE_EINVAL, "frame %u: null image or features", i);
        if (f.pitch < f.width)
            return fail(ctx, ORBFE_EINVAL, "frame %u: pitch %zu < width %u", i, f.pitch, f.width);
        if (!pitch_fits_offsets(f.pitch, f.width, f.height))
            return fail(ctx, ORBFE_ERANGE, "frame %u: pitch %zu past 32-bit offsets", i, f.pitch);
        PlanEntry* pe = nullptr;
        const int rc = get_plan(ctx, f.width, f.height, f.cap, &pe);
        if (rc) return rc;
        Job& J = lay.jobs[i];
        J.plan = pe->dev;
        J.tiles = pe->tiles;
        J.desc_groups = pe->host.desc_groups;
        J.cap = pe->host.cap;
        J.feat_stride = pe->host.feat_stride;
        J.img = f.image;
        J.pitch = f.pitch;
        J.feat = static_cast<uint8_t*>(f.features);
        J.surv = reinterpret_cast<uint64_t*>(arena + off);
        off += pe->surv_bytes;
        J.kp = reinterpret_cast<uint32_t*>(arena + off);
        off += pe->kp_bytes;
        J.smooth = reinterpret_cast<float*>(arena + off);
        off += pe->sm_bytes;
        if (f.pyramid) {
            J.pyr = f.pyramid;
        } else {
            J.pyr = reinterpret_cast<float*>(arena + off);
            off += pe->pyr_bytes;
        }
        lay.tile_prefix[i + 1] = lay.tile_prefix[i] + pe->host.tiles;
        lay.desc_prefix[i + 1] = lay.desc_prefix[i] + pe->host.L * pe->host.desc_groups;
    }
    lay.arena_bytes = off;
    fill_block_jobs(lay.tile_prefix, lay.tile_job);
    fill_block_jobs(lay.desc_prefix, lay.desc_job);
    lay.off_tiles = align_up(sizeof(Job) * n);
    lay.off_desc = lay.off_tiles + align_up(4 * (n + 1));
    lay.off_tile_job = lay.off_desc + align_up(4 * (n + 1));
    lay.off_desc_job = lay.off_tile_job + align_up(2 * lay.tile_job.size());
    lay.table_bytes = lay.off_desc_job + align_up(2 * lay.desc_job.size());
    return ORBFE_OK;
}

void pack_detect_table(const DetectLayout& lay, uint8_t* host) {
    std::memcpy(host, lay.jobs.data(), sizeof(Job) * lay.jobs.size());
    std::memcpy(host + lay.off_tiles, lay.tile_prefix.data(), 4 * lay.tile_prefix.size());
    std::memcpy(host + lay.off_desc, lay.desc_prefix.data(), 4 * lay.desc_prefix.size());
    std::memcpy(host + lay.off_tile_job, lay.tile_job.data(), 2 * lay.tile_job.size());
    std::memcpy(host + lay.off_desc_job, lay.desc_job.data(), 2 * lay.desc_job.size());
}

BatchView detect_view(const DetectLayout& lay, const uint8_t* dev) {
    BatchView v;
    v.jobs = reinterpret_cast<const Job*>(dev);
    v.tile_prefix = reinterpret_cast<const uint32_t*>(dev + lay.off_tiles);
    v.desc_prefix = reinterpret_cast<const uint32_t*>(dev + lay.off_desc);
    v.tile_job = reinterpret_cast<const uint16_t*>(dev + lay.off_tile_job);
    v.desc_job = reinterpret_cast<const uint16_t*>(dev + lay.off_desc_job);
    v.n = static_cast<uint32_t>(lay.jobs.size());
    return v;
}

struct MatchLayout {
    std::vector<MatchJob> jobs;
    std::vector<uint32_t> prefix;
    std::vector<uint16_t> block_job;
    size_t table_bytes = 0, arena_bytes = 0, off_prefix = 0, off_block_job = 0;
};

int layout_match(orbfe_ctx* ctx, const orbfe_match_pair* pairs, uint32_t n, MatchLayout& lay,
                 uint8_t* arena) {
    if (n > kMaxBatchJobs) return fail(ctx, ORBFE_ERANGE, "batch of %u pairs > %u", n, kMaxBatchJobs);
    lay.jobs.resize(n);
    lay.prefix.assign(n + 1, 0);
    size_t off = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const orbfe_match_pair& p = pairs[i];
        if (!p.prev_features || !p.curr_features || !p.prev_match)
            return fail(ctx, ORBFE_EINVAL, "pair %u: null pointer", i);
        PlanEntry* pe = nullptr;
        const int rc = get_plan(ctx, p.width, p.height, p.cap, &pe);
        if (rc) return rc;
        MatchJob& M = lay.jobs[i];
        M.plan = pe->dev;
        M.prev = static_cast<const uint8_t*>(p.prev_features);
        M.curr = static_cast<const uint8_t*>(p.curr_features);
        M.prev_match = p.prev_match;
        M.q_list = reinterpret_cast<uint32_t*>(arena + off);
        M.cap = pe->host.cap;
        M.feat_stride = pe->host.feat_stride;
        M.levels = pe->host.L;
        off += pe->pair_bytes;
        M.index_prev = reinterpret_cast<uint32_t*>(arena + off);
        off += pe->index_bytes;
        M.index_curr = reinterpret_cast<uint32_t*>(arena + off);
        off += pe->index_bytes;
        lay.prefix[i + 1] = lay.prefix[i] + match_blocks(pe->host, ctx->match_lanes_batch);
    }
    lay.arena_bytes = off;
    fill_block_jobs(lay.prefix, lay.block_job);
    lay.off_prefix = align_up(sizeof(MatchJob) * n);
    lay.off_block_job = lay.off_prefix + align_up(4 * (n + 1));
    lay.table_bytes = lay.off_block_job + align_up(2 * lay.block_job.size());
    return ORBFE_OK;
}

void pack_match_table(const MatchLayout& lay, uint8_t* host) {
    std::memcpy(host, lay.jobs.data(), sizeof(MatchJob) * lay.jobs.size());
    std::memcpy(host + lay.off_prefix, lay.prefix.data(), 4 * lay.prefix.size());
    std::memcpy(host + lay.off_block_job, lay.block_job.data(), 2 * lay.block_job.size());
}

MatchBatchView match_view(const MatchLayout& lay, const uint8_t* dev) {
    MatchBatchView v;
    v.jobs = reinterpret_cast<const MatchJob*>(dev);
    v.block_prefix = reinterpret_cast<const uint32_t*>(dev + lay.off_prefix);
    v.block_job = reinterpret_cast<const uint16_t*>(dev + lay.off_block_job);
    v.n = static_cast<uint32_t>(lay.jobs.size());
    return v;
}

// Dynamic batches: the arena grows on demand (after the stream that used it
// has drained) and tables go through a ring of pinned staging slots.
int ensure_arena(orbfe_ctx* ctx, size_t bytes, hipStream_t s) {
    if (bytes <= ctx->arena_bytes) return ORBFE_OK;
    if (ctx->arena) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->arena_stream));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        dfree(ctx->arena);
    }
    const size_t want = align_up(std::max(bytes, ctx->arena_bytes * 2), 1 << 20);
    const int rc = dalloc(ctx, &ctx->arena, want);
    if (rc) return rc;
    ctx->arena_bytes = want;
    ctx->arena_stream = s;
    return ORBFE_OK;
}

int stage_table(orbfe_ctx* ctx, size_t bytes, hipStream_t s, StageSlot** out) {
    StageSlot& slot = ctx->ring[ctx->ring_next];
    ctx->ring_next = (ctx->ring_next + 1) % 4;
    if (slot.pending) {
        HIP_TRY(ctx, hipEventSynchronize(slot.ev));
        slot.pending = false;
    }
    if (slot.bytes < bytes) {
        if (slot.host) (void)hipHostFree(slot.host);
        dfree(slot.dev);
        slot.host = nullptr;
        const size_t want = align_up(std::max<size_t>(bytes, 64 << 10), 64 << 10);
        HIP_TRY(ctx, hipHostMalloc(&slot.host, want, hipHostMallocDefault));
        const int rc = dalloc(ctx, &slot.dev, want);
        if (rc) return rc;
        slot.bytes = want;
    }
    if (!slot.ev) HIP_TRY(ctx, hipEventCreateWithFlags(&slot.ev, hipEventDisableTiming));
    (void)s;
    *out = &slot;
    return ORBFE_OK;
}

int commit_table(orbfe_ctx* ctx, StageSlot* slot, size_t bytes, hipStream_t s) {
    HIP_TRY(ctx, hipMemcpyAsync(slot->dev, slot->host, bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipEventRecord(slot->ev, s));
    slot->pending = true;
    return ORBFE_OK;
}

}  // namespace

extern "C" {

int orbfe_config_default(orbfe_config* cfg, uint32_t width, uint32_t height) {
    if (!cfg) return ORBFE_EINVAL;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->frame_width = width;
    cfg->frame_height = height;
    cfg->levels = 6;
    cfg->scale_factor = 1.2;
    cfg->fast_threshold = 0.15f;
    cfg->fast_min_run = 9;
    cfg->orientation_radius = 3;
    cfg->features_per_level = 0;
    cfg->border_zero = 0;
    cfg->lowe_ratio = 0.85f;
    return ORBFE_OK;
}

const char* orbfe_last_error(const orbfe_ctx* ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

void orbfe_destroy(orbfe_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    (void)hipDeviceSynchronize();
    ctx->plans.clear();  // PlanEntry frees its device buffers
    dfree(ctx->surv);
    dfree(ctx->kp);
    dfree(ctx->pyr_scratch);
    dfree(ctx->smooth);
    dfree(ctx->q_list);
    dfree(ctx->index[0]);
    dfree(ctx->index[1]);
    dfree(ctx->d_img);
    dfree(ctx->d_pyr);
    dfree(ctx->d_feat[0]);
    dfree(ctx->d_feat[1]);
    dfree(ctx->d_match);
    dfree(ctx->bf_part);
    dfree(ctx->arena);
    for (auto& slot : ctx->ring) {
        if (slot.host) (void)hipHostFree(slot.host);
        dfree(slot.dev);
        if (slot.ev) (void)hipEventDestroy(slot.ev);
    }
    for (auto& e : ctx->pending) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto e : ctx->pool) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int orbfe_create(const orbfe_config* cfg, int device, orbfe_ctx** out) {
    if (!cfg || !out) return ORBFE_EINVAL;
    *out = nullptr;
    if (cfg->levels == 0 || cfg->levels > ORBFE_MAX_LEVELS) return ORBFE_ERANGE;
    if (!(cfg->scale_factor > 1.0)) return ORBFE_ERANGE;
    // A ratio above 1 lets a tie at the best distance pass the ratio test,
    // and the reference then keeps whichever tied candidate its unstable
    // sort by x puts first (match_features_cpu.cpp:101-146): refused.  NaN
    // fails the comparison as well.
    if (!(cfg->lowe_ratio > 0.0f && cfg->lowe_ratio <= 1.0f)) return ORBFE_ERANGE;
    if (cfg->orientation_radius < 0 || cfg->orientation_radius > 15) return ORBFE_ERANGE;
    if (cfg->fast_min_run > 16) return ORBFE_ERANGE;
    const uint32_t th = cfg->detect_tile_rows ? cfg->detect_tile_rows : orbfe::kTileHDefault;
    const uint32_t dl = cfg->describe_lanes;
    const uint32_t ml = cfg->match_lanes;
    if ((th != 8 && th != 16 && th != 32) || (dl != 0 && dl != 8 && dl != 16 && dl != 32) ||
        (ml != 0 && ml != 16 && ml != 32 && ml != 64))
        return ORBFE_ERANGE;
    // the main frame size (checked again, with every other size, per plan)
    if (cfg->frame_width == 0 || cfg->frame_height == 0 || cfg->frame_width > 65535 ||
        cfg->frame_height > 65535 ||
        !frame_fits_offsets(cfg->frame_width, cfg->frame_height, cfg->levels, cfg->scale_factor))
        return ORBFE_ERANGE;
    auto ctx = std::make_unique<orbfe_ctx>();
    ctx->device = device;
    ctx->cfg = *cfg;
    ctx->cfg.detect_tile_rows = th;  // every plan of a context shares the launch shapes
    ctx->desc_lanes_batch = dl ? dl : orbfe::kDescLanesDefault;
    ctx->desc_lanes_frame = dl ? dl : orbfe::kDescLanesFrame;
    ctx->cfg.describe_lanes = ctx->desc_lanes_batch;  // the plans' (batch) describe shape
    ctx->match_lanes_batch = ml ? ml : orbfe::kMatchLanesBatch;
    ctx->match_lanes_pair = ml ? ml : orbfe::kMatchLanesPair;
    ctx->prm.threshold = cfg->fast_threshold;
    ctx->prm.min_run = cfg->fast_min_run;
    ctx->prm.orient_r = cfg->orientation_radius;
    ctx->prm.border_zero = cfg->border_zero ? 1u : 0u;
    ctx->prm.ratio = cfg->lowe_ratio;
    orbfe_ctx* c = ctx.get();
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        // Report through the return code only: there is no context to own text.
        return ORBFE_EHIP;
    }
    if (device < 0 || device >= ndev) return ORBFE_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return ORBFE_EHIP;
    if (orbfe::upload_pattern() != hipSuccess) return ORBFE_EHIP;
    if (hipStreamCreate(&c->stream) != hipSuccess) return ORBFE_EHIP;
    if (scratch_address(device, reinterpret_cast<void**>(&c->bf_static)) != ORBFE_OK) {
        orbfe_destroy(ctx.release());
        return ORBFE_EHIP;
    }
    // From here on every failure goes through orbfe_destroy, which releases
    // the stream, the plans and whatever scratch was allocated.
    PlanEntry* pe = nullptr;
    if (const int rc = get_plan(c, cfg->frame_width, cfg->frame_height,
                                cfg->features_per_level, &pe)) {
        orbfe_destroy(ctx.release());
        return rc;
    }
    c->main = pe;
    if (dalloc(c, &c->surv, pe->surv_bytes) || dalloc(c, &c->kp, pe->kp_bytes) ||
        dalloc(c, &c->pyr_scratch, pe->pyr_bytes) || dalloc(c, &c->smooth, pe->sm_bytes) ||
        dalloc(c, &c->q_list, pe->pair_bytes) || dalloc(c, &c->index[0], pe->index_bytes) ||
        dalloc(c, &c->index[1], pe->index_bytes) ||
        dalloc(c, &c->d_img, align_up(static_cast<size_t>(cfg->frame_width) * cfg->frame_height)) ||
        dalloc(c, &c->d_pyr, pe->pyr_bytes) || dalloc(c, &c->d_feat[0], pe->feat_bytes) ||
        dalloc(c, &c->d_feat[1], pe->feat_bytes) || dalloc(c, &c->d_match, pe->match_bytes)) {
        orbfe_destroy(ctx.release());
        return ORBFE_ENOMEM;
    }
    *out = ctx.release();
    return ORBFE_OK;
}

int orbfe_level_dims(const orbfe_ctx* ctx, uint32_t* widths, uint32_t* heights) {
    if (!ctx || !widths || !heights) return ORBFE_EINVAL;
    for (uint32_t l = 0; l < ctx->cfg.levels; ++l) {
        widths[l] = ctx->main->host.lv[l].w;
        heights[l] = ctx->main->host.lv[l].h;
    }
    return ORBFE_OK;
}

size_t orbfe_feature_array_bytes(const orbfe_ctx* ctx) {
    return ctx ? static_cast<size_t>(ctx->cfg.levels) * ctx->main->host.feat_stride : 0;
}

size_t orbfe_pyramid_floats(const orbfe_ctx* ctx) {
    return ctx ? ctx->main->host.pyr_floats : 0;
}

uint32_t orbfe_cap(const orbfe_ctx* ctx) { return ctx ? ctx->main->host.cap : 0; }

int orbfe_detect_async(orbfe_ctx* ctx, const uint8_t* image, size_t pitch, float* pyramid,
                       void* features, hipStream_t stream) {
    if (!ctx) return ORBFE_EINVAL;
    if (!image || !features) return fail(ctx, ORBFE_EINVAL, "null image or features");
    const Plan& P = ctx->main->host;
    if (pitch < P.W0) return fail(ctx, ORBFE_EINVAL, "pitch %zu < width %u", pitch, P.W0);
    if (!pitch_fits_offsets(pitch, P.W0, P.H0))
        return fail(ctx, ORBFE_ERANGE, "pitch %zu past 32-bit offsets", pitch);
    Job job0{};
    job0.plan = ctx->main->dev;
    job0.tiles = ctx->main->tiles_single;
    // single-frame launches: their own describe shape (desc_lanes_frame)
    job0.desc_groups = ceil_div(P.cap, 256 / ctx->desc_lanes_frame);
    job0.cap = P.cap;
    job0.feat_stride = P.feat_stride;
    job0.img = image;
    job0.pitch = pitch;
    job0.pyr = pyramid ? pyramid : ctx->pyr_scratch;
    job0.feat = static_cast<uint8_t*>(features);
    job0.surv = ctx->surv;
    job0.kp = ctx->kp;
    job0.smooth = ctx->smooth;
    const BatchView bv{nullptr, nullptr, nullptr, nullptr, nullptr, 0};
    return launch_detect_stage(ctx, bv, job0, P.tiles, 1, ctx->desc_lanes_frame,
                               P.L * job0.desc_groups, stream);
}

int orbfe_detect(orbfe_ctx* ctx, const uint8_t* image, size_t pitch, float* pyramid,
                 void* features) {
    if (!ctx) return ORBFE_EINVAL;
    if (!image || !features) return fail(ctx, ORBFE_EINVAL, "null image or features");
    const Plan& P = ctx->main->host;
    if (pitch < P.W0) return fail(ctx, ORBFE_EINVAL, "pitch %zu < width %u", pitch, P.W0);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpy2DAsync(ctx->d_img, P.W0, image, pitch, P.W0, P.H0,
                                  hipMemcpyHostToDevice, s));
    const int rc = orbfe_detect_async(ctx, ctx->d_img, P.W0, pyramid ? ctx->d_pyr : nullptr,
                                      ctx->d_feat[0], s);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(features, ctx->d_feat[0],
                                static_cast<size_t>(P.L) * P.feat_stride,
                                hipMemcpyDeviceToHost, s));
    if (pyramid)
        HIP_TRY(ctx, hipMemcpyAsync(pyramid, ctx->d_pyr, static_cast<size_t>(P.pyr_floats) * 4,
                                    hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return ORBFE_OK;
}

int orbfe_detect_batch_async(orbfe_ctx* ctx, const orbfe_frame* frames, uint32_t n,
                             hipStream_t stream) {
    if (!ctx) return ORBFE_EINVAL;
    if (n == 0) return ORBFE_OK;
    if (!frames) return fail(ctx, ORBFE_EINVAL, "null frames");
    DetectLayout lay;
    int rc = layout_detect(ctx, frames, n, lay, nullptr);
    if (rc) return rc;
    rc = ensure_arena(ctx, lay.arena_bytes, stream);
    if (rc) return rc;
    rc = layout_detect(ctx, frames, n, lay, ctx->arena);
    if (rc) return rc;
    StageSlot* slot = nullptr;
    rc = stage_table(ctx, lay.table_bytes, stream, &slot);
    if (rc) return rc;
    pack_detect_table(lay, static_cast<uint8_t*>(slot->host));
    rc = commit_table(ctx, slot, lay.table_bytes, stream);
    if (rc) return rc;
    const BatchView bv = detect_view(lay, static_cast<uint8_t*>(slot->dev));
    return launch_detect_stage(ctx, bv, Job{}, lay.tile_prefix[n], n, ctx->desc_lanes_batch,
                               lay.desc_prefix[n], stream);
}

int orbfe_batch_create(orbfe_ctx* ctx, const orbfe_frame* frames, uint32_t n,
                       orbfe_batch** out) {
    if (!ctx || !out) return ORBFE_EINVAL;
    *out = nullptr;
    if (n == 0 || !frames) return fail(ctx, ORBFE_EINVAL, "empty batch");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    DetectLayout lay;
    int rc = layout_detect(ctx, frames, n, lay, nullptr);
    if (rc) return rc;
    auto b = std::make_unique<orbfe_batch>();
    rc = dalloc(ctx, &b->arena, lay.arena_bytes);
    if (rc) return rc;
    rc = layout_detect(ctx, frames, n, lay, b->arena);
    if (rc) { dfree(b->arena); return rc; }
    std::vector<uint8_t> host(lay.table_bytes);
    pack_detect_table(lay, host.data());
    uint8_t* table = nullptr;
    rc = dalloc(ctx, &table, lay.table_bytes);
    if (rc) { dfree(b->arena); return rc; }
    b->table = table;
    if (hipMemcpy(table, host.data(), lay.table_bytes, hipMemcpyHostToDevice) != hipSuccess) {
        dfree(b->arena);
        dfree(table);
        return fail(ctx, ORBFE_EHIP, "batch table upload failed");
    }
    b->ctx = ctx;
    b->n = n;
    b->tiles = lay.tile_prefix[n];
    b->desc_blocks = lay.desc_prefix[n];
    b->view = detect_view(lay, table);
    *out = b.release();
    return ORBFE_OK;
}

int orbfe_batch_detect_async(orbfe_ctx* ctx, orbfe_batch* batch, hipStream_t stream) {
    if (!ctx || !batch || batch->ctx != ctx) return fail(ctx, ORBFE_EINVAL, "bad batch");
    return launch_detect_stage(ctx, batch->view, Job{}, batch->tiles, batch->n,
                               ctx->desc_lanes_batch, batch->desc_blocks, stream);
}

void orbfe_batch_destroy(orbfe_batch* batch) {
    if (!batch) return;
    (void)hipDeviceSynchronize();
    dfree(batch->arena);
    if (batch->table) (void)hipFree(batch->table);
    delete batch;
}

int orbfe_load_features(const void* features, uint32_t levels, uint32_t cap, orbfe_feature* out,
                        uint32_t* counts) {
    if (!features || !out || !counts || levels == 0 || levels > ORBFE_MAX_LEVELS)
        return ORBFE_EINVAL;
    const size_t stride = 8 + 48ull * cap;
    const uint8_t* base = static_cast<const uint8_t*>(features);
    size_t n = 0;
    for (uint32_t l = 0; l < levels; ++l) {
        const uint8_t* block = base + stride * l;
        uint32_t count;
        std::memcpy(&count, block, 4);
        const uint32_t take = std::min(count, cap);
        for (uint32_t i = 0; i < take; ++i) {
            orbfe_gpu_feature g;
            std::memcpy(&g, block + 8 + 48ull * i, sizeof(g));
            orbfe_feature& f = out[n++];
            f.x = static_cast<uint32_t>(g.x);
            f.y = static_cast<uint32_t>(g.y);
            f.lod = l;
            f.strength = g.strength;
            f.orientation = g.orientation;
            std::memcpy(f.descriptor, g.descriptor, 32);
        }
        counts[l] = take;
    }
    return ORBFE_OK;
}

int orbfe_frame_scratch_bytes(orbfe_ctx* ctx, uint32_t width, uint32_t height, uint32_t cap,
                              size_t* bytes) {
    if (!ctx || !bytes) return ORBFE_EINVAL;
    PlanEntry* pe = nullptr;
    const int rc = get_plan(ctx, width, height, cap, &pe);
    if (rc) return rc;
    // the arrays layout_detect and layout_match place per frame / pair
    *bytes = pe->surv_bytes + pe->kp_bytes + pe->sm_bytes + pe->pyr_bytes + pe->pair_bytes +
             2 * pe->index_bytes;
    return ORBFE_OK;
}

int orbfe_export_features_async(orbfe_ctx* ctx, const void* features, orbfe_feature* out,
                                uint32_t* counts, hipStream_t stream) {
    if (!ctx) return ORBFE_EINVAL;
    if (!features || !out || !counts) return fail(ctx, ORBFE_EINVAL, "null pointer");
    // The kernel writes through the device addresses of the pinned buffers.
    void* d_out = nullptr;
    void* d_counts = nullptr;
    if (hipHostGetDevicePointer(&d_out, out, 0) != hipSuccess || !d_out ||
        hipHostGetDevicePointer(&d_counts, counts, 0) != hipSuccess || !d_counts) {
        (void)hipGetLastError();
        return fail(ctx, ORBFE_EINVAL, "out and counts must be pinned host memory (hipHostMalloc)");
    }
    const Plan& P = ctx->main->host;
    HIP_TRY(ctx, timed(ctx, "k_export", stream, [&] {
                return orbfe::launch_export(static_cast<const uint8_t*>(features), P.L, P.cap,
                                            P.feat_stride, static_cast<uint32_t*>(d_out),
                                            static_cast<uint32_t*>(d_counts), stream);
            }));
    return ORBFE_OK;
}

int orbfe_match_pipeline_async(orbfe_ctx* ctx, const void* prev_features,
                               const void* curr_features, int32_t* prev_match,
                               hipStream_t stream) {
    if (!ctx) return ORBFE_EINVAL;
    if (!prev_features || !curr_features || !prev_match)
        return fail(ctx, ORBFE_EINVAL, "null pointer");
    const Plan& P = ctx->main->host;
    MatchJob job0{};
    job0.plan = ctx->main->dev;
    job0.prev = static_cast<const uint8_t*>(prev_features);
    job0.curr = static_cast<const uint8_t*>(curr_features);
    job0.prev_match = prev_match;
    job0.q_list = ctx->q_list;
    job0.cap = P.cap;
    job0.feat_stride = P.feat_stride;
    job0.levels = P.L;
    job0.index_prev = ctx->index[0];
    job0.index_curr = ctx->index[1];
    const MatchBatchView bv{nullptr, nullptr, nullptr, 0};
    return launch_match_stage(ctx, bv, job0, ctx->match_lanes_pair,
                              match_blocks(P, ctx->match_lanes_pair), stream);
}

int orbfe_match_pipeline_batch_async(orbfe_ctx* ctx, const orbfe_match_pair* pairs, uint32_t n,
                                     hipStream_t stream) {
    if (!ctx) return ORBFE_EINVAL;
    if (n == 0) return ORBFE_OK;
    if (!pairs) return fail(ctx, ORBFE_EINVAL, "null pairs");
    MatchLayout lay;
    int rc = layout_match(ctx, pairs, n, lay, nullptr);
    if (rc) return rc;
    rc = ensure_arena(ctx, lay.arena_bytes, stream);
    if (rc) return rc;
    rc = layout_match(ctx, pairs, n, lay, ctx->arena);
    if (rc) return rc;
    StageSlot* slot = nullptr;
    rc = stage_table(ctx, lay.table_bytes, stream, &slot);
    if (rc) return rc;
    pack_match_table(lay, static_cast<uint8_t*>(slot->host));
    rc = commit_table(ctx, slot, lay.table_bytes, stream);
    if (rc) return rc;
    const MatchBatchView bv = match_view(lay, static_cast<uint8_t*>(slot->dev));
    return launch_match_stage(ctx, bv, MatchJob{}, ctx->match_lanes_batch, lay.prefix[n], stream);
}

int orbfe_match_batch_create(orbfe_ctx* ctx, const orbfe_match_pair* pairs, uint32_t n,
                             orbfe_match_batch** out) {
    if (!ctx || !out) return ORBFE_EINVAL;
    *out = nullptr;
    if (n == 0 || !pairs) return fail(ctx, ORBFE_EINVAL, "empty batch");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    MatchLayout lay;
    int rc = layout_match(ctx, pairs, n, lay, nullptr);
    if (rc) return rc;
    auto b = std::make_unique<orbfe_match_batch>();
    rc = dalloc(ctx, &b->arena, lay.arena_bytes);
    if (rc) return rc;
    rc = layout_match(ctx, pairs, n, lay, b->arena);
    if (rc) { dfree(b->arena); return rc; }
    std::vector<uint8_t> host(lay.table_bytes);
    pack_match_table(lay, host.data());
    uint8_t* table = nullptr;
    rc = dalloc(ctx, &table, lay.table_bytes);
    if (rc) { dfree(b->arena); return rc; }
    b->table = table;
    if (hipMemcpy(table, host.data(), lay.table_bytes, hipMemcpyHostToDevice) != hipSuccess) {
        dfree(b->arena);
        dfree(table);
        return fail(ctx, ORBFE_EHIP, "match table upload failed");
    }
    b->ctx = ctx;
    b->n = n;
    b->blocks = lay.prefix[n];
    b->view = match_view(lay, table);
    *out = b.release();
    return ORBFE_OK;
}

int orbfe_match_batch_async(orbfe_ctx* ctx, orbfe_match_batch* batch, hipStream_t stream) {
    if (!ctx || !batch || batch->ctx != ctx) return fail(ctx, ORBFE_EINVAL, "bad batch");
    return launch_match_stage(ctx, batch->view, MatchJob{}, ctx->match_lanes_batch, batch->blocks,
                              stream);
}

void orbfe_match_batch_destroy(orbfe_match_batch* batch) {
    if (!batch) return;
    (void)hipDeviceSynchronize();
    dfree(batch->arena);
    if (batch->table) (void)hipFree(batch->table);
    delete batch;
}

int orbfe_match_pipeline(orbfe_ctx* ctx, const void* prev_features, const void* curr_features,
                         int32_t* prev_match, uint32_t* n_pairs) {
    if (!ctx) return ORBFE_EINVAL;
    if (!prev_features || !curr_features || !prev_match)
        return fail(ctx, ORBFE_EINVAL, "null pointer");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const Plan& P = ctx->main->host;
    const size_t fb = static_cast<size_t>(P.L) * P.feat_stride;
    const size_t mb = static_cast<size_t>(P.L) * P.cap * 4;
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->d_feat[0], prev_features, fb, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->d_feat[1], curr_features, fb, hipMemcpyHostToDevice, s));
    const int rc = orbfe_match_pipeline_async(ctx, ctx->d_feat[0], ctx->d_feat[1], ctx->d_match, s);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(prev_match, ctx->d_match, mb, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (n_pairs) {
        uint32_t c = 0;
        for (size_t i = 0; i < mb / 4; ++i) c += prev_match[i] >= 0 ? 1u : 0u;
        *n_pairs = c;
    }
    return ORBFE_OK;
}

int orbfe_match_bf_async(orbfe_ctx* ctx, const uint32_t* query, uint32_t nq,
                         const uint32_t* train, uint32_t nt, uint32_t* idx, uint32_t* d1,
                         uint32_t* d2, hipStream_t stream) {
    if (!ctx) return ORBFE_EINVAL;
    if (nq == 0) return ORBFE_OK;
    if (!query || !idx || !d1 || !d2 || (nt > 0 && !train))
        return fail(ctx, ORBFE_EINVAL, "null pointer");
    if (nt >= (1u << orbfe::kBfKeyShift))
        return fail(ctx, ORBFE_ERANGE, "train count %u >= 2^%u", nt, orbfe::kBfKeyShift);
    uint32_t splits = 0, split_len = 0;
    bf_split_plan(nq, nt, &splits, &split_len);
    const size_t need = static_cast<size_t>(splits) * nq * sizeof(uint2);
    if (need > ctx->bf_part_bytes) {
        HIP_TRY(ctx, hipStreamSynchronize(stream));
        dfree(ctx->bf_part);
        const int rc = dalloc(ctx, &ctx->bf_part, need);
        if (rc) {
            ctx->bf_part_bytes = 0;
            return rc;
        }
        ctx->bf_part_bytes = need;
    }
    HIP_TRY(ctx, timed(ctx, "k_bf_mfma", stream, [&] {
                return orbfe::launch_bf_mfma(query, nq, train, nt, splits, split_len,
                                             ctx->bf_part, stream);
            }));
    HIP_TRY(ctx, timed(ctx, "k_bf_merge", stream, [&] {
                return orbfe::launch_bf_merge(ctx->bf_part, nq, splits, idx, d1, d2, stream);
            }));
    return ORBFE_OK;
}

int orbfe_match_bf_noalloc_async(orbfe_ctx* ctx, const uint32_t* query, uint32_t nq,
                                 const uint32_t* train, uint32_t nt, uint32_t* idx, uint32_t* d1,
                                 uint32_t* d2, hipStream_t stream) {
    if (!ctx) return ORBFE_EINVAL;
    if (nq == 0) return ORBFE_OK;
    if (!query || !idx || !d1 || !d2 || (nt > 0 && !train))
        return fail(ctx, ORBFE_EINVAL, "null pointer");
    if (nt >= (1u << orbfe::kBfKeyShift))
        return fail(ctx, ORBFE_ERANGE, "train count %u >= 2^%u", nt, orbfe::kBfKeyShift);
    if (!ctx->bf_static) return fail(ctx, ORBFE_EHIP, "static scratch address unknown");
    // The split plan of orbfe_match_bf_async.  Its occupancy term keeps
    // splits * nq <= 2^20 (at most 2048 / ceil(nq / 512) splits), so only
    // the ceil(nt / 8192) splits the key's index field needs can exceed the
    // scratch.
    static_assert(orbfe::kBfScratchBytes == ORBFE_BF_NOALLOC_SCRATCH_BYTES, "orbfe.h limit");
    constexpr size_t kEntries = orbfe::kBfScratchBytes / sizeof(uint2);
    uint32_t splits = 0, split_len = 0;
    bf_split_plan(nq, nt, &splits, &split_len);
    if (static_cast<size_t>(splits) * nq > kEntries)
        return fail(ctx, ORBFE_ERANGE,
                    "%u queries x %u train splits need %zu bytes of partial results; the static "
                    "scratch holds %zu (nq * ceil(nt / 8192) <= %zu)",
                    nq, splits, static_cast<size_t>(splits) * nq * sizeof(uint2),
                    orbfe::kBfScratchBytes, kEntries);
    HIP_TRY(ctx, timed(ctx, "k_bf_mfma", stream, [&] {
                return orbfe::launch_bf_mfma(query, nq, train, nt, splits, split_len,
                                             ctx->bf_static, stream);
            }));
    HIP_TRY(ctx, timed(ctx, "k_bf_merge", stream, [&] {
                return orbfe::launch_bf_merge(ctx->bf_static, nq, splits, idx, d1, d2, stream);
            }));
    return ORBFE_OK;
}

int orbfe_ransac_score_async(orbfe_ctx* ctx, const double* F, uint32_t k, const double* prev,
                             const double* curr, uint32_t n, double threshold, uint32_t* counts,
                             uint8_t* mask, hipStream_t stream) {
    if (!ctx) return ORBFE_EINVAL;
    if (k == 0) return ORBFE_OK;
    if (!F || !counts || (n > 0 && (!prev || !curr)))
        return fail(ctx, ORBFE_EINVAL, "null pointer");
    if (k > 65535) return fail(ctx, ORBFE_ERANGE, "%u hypotheses > 65535", k);
    if ((reinterpret_cast<uintptr_t>(prev) | reinterpret_cast<uintptr_t>(curr)) & 15u)
        return fail(ctx, ORBFE_EINVAL, "prev and curr must be 16-byte aligned");
    HIP_TRY(ctx, timed(ctx, "k_ransac_score", stream, [&] {
                return orbfe::launch_ransac_score(F, k, prev, curr, n, threshold, counts, mask,
                                                  stream);
            }));
    return ORBFE_OK;
}

// Times 5 launches of a probe kernel (after one warm-up) with HIP events on
// `stream`; *rate = work per launch / seconds per launch.
static int probe_rate(orbfe_ctx* ctx, hipStream_t stream, uint32_t blocks_per_cu,
                      const char* what,
                      hipError_t (*launch)(uint32_t*, uint32_t, hipStream_t, double*),
                      double* rate) {
    if (!ctx || !rate) return ORBFE_EINVAL;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int cus = 0;
    HIP_TRY(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const uint32_t blocks = static_cast<uint32_t>(cus) * blocks_per_cu;
    uint32_t* sink = nullptr;
    int rc = dalloc(ctx, &sink, 4ull * blocks);
    if (rc) return rc;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double per = 0.0;
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e == hipSuccess) e = launch(sink, blocks, stream, &per);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(ev[0], stream);
    for (int r = 0; r < 5 && e == hipSuccess; ++r) e = launch(sink, blocks, stream, &per);
    if (e == hipSuccess) e = hipEventRecord(ev[1], stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    for (auto x : ev)
        if (x) (void)hipEventDestroy(x);
    dfree(sink);
    if (e != hipSuccess) return fail(ctx, ORBFE_EHIP, "%s probe: %s", what, hipGetErrorString(e));
    *rate = 5.0 * per / (static_cast<double>(ms) * 1e-3);
    return ORBFE_OK;
}

int orbfe_probe_valu_rate(orbfe_ctx* ctx, hipStream_t stream, double* rate) {
    // 8 workgroups per CU, 4 rounds
    return probe_rate(ctx, stream, 32, "valu", orbfe::launch_valu_probe, rate);
}

int orbfe_probe_mfma_rate(orbfe_ctx* ctx, hipStream_t stream, double* rate) {
    // 8 workgroups of 4 waves per CU: 8 waves per SIMD, 2 chains each
    return probe_rate(ctx, stream, 8, "mfma", orbfe::launch_mfma_probe, rate);
}

int orbfe_timing_enable(orbfe_ctx* ctx, int enable) {
    if (!ctx) return ORBFE_EINVAL;
    ctx->timing = enable != 0;
    return ORBFE_OK;
}

int orbfe_timing_reset(orbfe_ctx* ctx) {
    if (!ctx) return ORBFE_EINVAL;
    for (auto& e : ctx->pending) {
        HIP_TRY(ctx, hipEventSynchronize(e.b));
        ctx->pool.push_back(e.a);
        ctx->pool.push_back(e.b);
    }
    ctx->pending.clear();
    ctx->totals.clear();
    return ORBFE_OK;
}

int orbfe_timing_read(orbfe_ctx* ctx, const char* name, uint32_t* launches, double* total_ms) {
    if (!ctx || !name) return ORBFE_EINVAL;
    for (auto& e : ctx->pending) {
        HIP_TRY(ctx, hipEventSynchronize(e.b));
        float ms = 0.0f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, e.a, e.b));
        auto& t = ctx->totals[e.name];
        t.first += 1;
        t.second += ms;
        ctx->pool.push_back(e.a);
        ctx->pool.push_back(e.b);
    }
    ctx->pending.clear();
    const auto it = ctx->totals.find(name);
    if (launches) *launches = it == ctx->totals.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == ctx->totals.end() ? 0.0 : it->second.second;
    return ORBFE_OK;
}

}  // extern "C"
