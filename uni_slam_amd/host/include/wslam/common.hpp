#pragma once

// Store keys of the feature path (wslam/include/wslam/common.hpp:114-159 of
// the reference) and the frame type handed to the feature-detect stage.

#include <cstdint>
#include <string>
#include <vector>

namespace wslam {

namespace ResourceIdentifier {
inline std::string GetFrameName(uint32_t frame_idx, uint32_t lod = 0) {
    return "res:frame:" + std::to_string(frame_idx) + ":lod:" + std::to_string(lod);
}
inline std::string GetFeatureSetName(size_t keyframes_ago) {
    return "gen:feat:-" + std::to_string(keyframes_ago);
}
inline const std::string MatchedFeaturesName = "gen:feat:match";
// Consume-once marker asking the loader to advance the reference feature set.
inline const std::string FeatureReferenceAdvanceName = "gen:feat:refadv";
inline const std::string RansacResultName = "gen:feat:ransac";
}  // namespace ResourceIdentifier

// A grayscale frame kept as u8 end to end (the reference converts to f32 on
// the host, data/src/provider_base.cpp:121-124; here the conversion happens
// in the detect kernel).
struct FrameU8 {
    uint32_t width = 0;
    uint32_t height = 0;
    uint64_t timestamp = 0;
    std::vector<uint8_t> pixels;  // row-major, pitch = width
};

namespace data {
// The reference's grayscale frame (data/include/data/provider_base.hpp:57-62),
// published by its sensor loader under res:frame:<i>:lod:0
// (wslam/src/feature_detect/sensor_loader.cpp:105): f32 pixels
// float(u8) / 255 (data/src/provider_base.cpp:121-124).
struct FrameBW {
    uint64_t timestamp;
    std::vector<float> pixels;
    uint16_t width;
    uint16_t height;
};
}  // namespace data

}  // namespace wslam
