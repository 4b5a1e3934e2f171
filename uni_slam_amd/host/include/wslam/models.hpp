#pragma once

// Data layouts of the feature path, identical to the reference's
// wslam/src/models.hpp:14-129: the 48-byte device record, the 52-byte host
// feature, FeatureSet and MatchResult.  MatchResult keeps the reference's
// ordering (a map keyed by the current-frame feature, ordered by Feature's
// member-wise comparison); std::map stands in for C++23 std::flat_map.

#include <array>
#include <compare>
#include <cstddef>
#include <cstdint>
#include <map>
#include <vector>

namespace wslam {

inline constexpr uint32_t kLevelsOfDetail = 6;  // GPUConst::levels_of_detail

namespace gpumodels {
struct alignas(8) Feature {
    float x;
    float y;
    uint32_t strength;
    float orientation;
    std::array<uint32_t, 8> descriptor;
};
static_assert(sizeof(Feature) == 48);
static_assert(alignof(Feature) == 8);
static_assert(offsetof(Feature, strength) == 8);
static_assert(offsetof(Feature, orientation) == 12);
static_assert(offsetof(Feature, descriptor) == 16);

// One level block of the device feature array: {count; pad; Feature[cap]}.
inline constexpr size_t FeatureBlockBytes(uint32_t cap) { return 8 + sizeof(Feature) * cap; }
}  // namespace gpumodels

struct Feature {
    uint32_t x;
    uint32_t y;
    uint32_t lod;
    uint32_t strength;
    float orientation;
    std::array<uint32_t, 8> descriptor;

    static constexpr size_t kDescriptorBits = 256;
    [[nodiscard]] constexpr bool getBit(size_t idx) const noexcept {
        return ((descriptor[idx / 32] >> (idx % 32)) & 1U) != 0U;
    }
    constexpr auto operator<=>(const Feature&) const = default;
    constexpr bool operator==(const Feature&) const = default;
};
static_assert(sizeof(Feature) == 52);

constexpr Feature FromGpuModel(const gpumodels::Feature& f, uint32_t lod) {
    return Feature{static_cast<uint32_t>(f.x), static_cast<uint32_t>(f.y), lod, f.strength,
                   f.orientation, f.descriptor};
}

struct FeatureLess {
    bool operator()(const Feature& a, const Feature& b) const {
        return (a <=> b) == std::partial_ordering::less;
    }
};

using FeatureSet = std::array<std::vector<Feature>, kLevelsOfDetail>;
// Per prev-frame level: current-frame feature -> matched prev-frame feature.
using MatchResult = std::array<std::map<Feature, Feature, FeatureLess>, kLevelsOfDetail>;

// Output of the RANSAC pass (models.hpp:130-145 of the reference).
struct RansacStats {
    size_t total_matches = 0;
    size_t inlier_count = 0;
    size_t iterations_run = 0;
    bool model_found = false;
};

struct RansacResult {
    MatchResult inliers;
    std::array<std::array<double, 3>, 3> fundamental_matrix{};
    RansacStats stats;
};

}  // namespace wslam
