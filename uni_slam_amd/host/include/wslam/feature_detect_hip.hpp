#pragma once

// The feature path of the SLAM pipeline on a HIP device, behind the
// pipeline's pass interface:
//
//   Feature detect stage (HipStage):
//     FeatureDetectHipPass  (HIP)  replaces FillPyramidPass, PassDetectCorners,
//                                  CullCornersPass and GenerateFeaturesPass
//                                  (wslam/src/feature_detect/feature_detect.cpp:43-86)
//   Pose estimate stage (HipStage), first passes:
//     LoadFeaturesHipPass   (CPU)  replaces LoadDataCPUPass
//                                  (wslam/src/pose_estimate_cpu/load_features_cpu.cpp:54-142)
//     MatchFeaturesHipPass  (HIP)  replaces MatchFeaturesCPU::execute's search
//                                  (match_features_cpu.cpp:217-268) on the GPU
//     PublishMatchesPass    (CPU)  builds MatchResult under gen:feat:match
//
// The passes share a FeatureBuffers object: the device feature arrays of the
// current set (gen:feat:-0), the reference set (gen:feat:-1) and the frame
// being detected, plus pinned staging (the frame, the exported FeatureSet,
// the match table).  Keys and layouts are the reference's.  Every pass
// enqueues the same work for the same buffer slots, so the stages may run
// with HIP graph capture (HipStage::enableGraphCapture): one graph per slot
// combination, the frame staged on the host before each replay.

#include <hip/hip_runtime_api.h>

#include <array>
#include <memory>
#include <optional>
#include <string>

#include "../../../../include/orbfe.h"
#include "compute/hip_stage.hpp"
#include "compute/pass.hpp"
#include "compute/store.hpp"
#include "wslam/common.hpp"
#include "wslam/models.hpp"

namespace wslam {

struct FeatureDetectOpts {
    uint32_t frame_width = 752;
    uint32_t frame_height = 480;
    uint32_t features_per_level = 0;  // 0 = W*H/250
    float fast_threshold = 0.15f;
    float lowe_ratio = 0.85f;
    uint32_t border_zero = 0;
};

class FeatureBuffers {
   public:
    FeatureBuffers(std::shared_ptr<compute::Device> device, FeatureDetectOpts opts);
    ~FeatureBuffers();
    FeatureBuffers(const FeatureBuffers&) = delete;
    FeatureBuffers& operator=(const FeatureBuffers&) = delete;

    [[nodiscard]] std::optional<std::string> initialize();

    orbfe_ctx* ctx() const { return ctx_; }
    const FeatureDetectOpts& opts() const { return opts_; }
    uint32_t cap() const { return cap_; }
    size_t featureBytes() const { return feature_bytes_; }

    // Device arrays by role; detect writes `next`.
    void* next() const { return arrays_[next_]; }
    void* current() const { return arrays_[cur_]; }
    void* reference() const { return arrays_[ref_]; }
    bool hasReference() const { return has_ref_; }
    bool hasCurrent() const { return has_cur_; }

    // The loader's shift (load_features_cpu.cpp:111-142) on the device side:
    // optionally the current array becomes the reference, then the freshly
    // detected array becomes current.
    void rotate(bool advance_reference);

    // Slot of each device array (0..2), for graph keys.
    int nextSlot() const { return next_; }
    int currentSlot() const { return cur_; }
    int referenceSlot() const { return ref_; }

    uint8_t* hostFrame() const { return host_frame_; }
    uint8_t* deviceFrame() const { return dev_frame_; }
    // FeatureSet of the last detected frame, written by the GPU into pinned
    // memory (orbfe_export_features_async): records level by level, counts.
    const Feature* hostSet() const { return reinterpret_cast<const Feature*>(host_set_); }
    const uint32_t* hostCounts() const { return host_counts_; }
    int32_t* deviceMatch() const { return dev_match_; }
    int32_t* hostMatch() const { return host_match_; }

   private:
    std::shared_ptr<compute::Device> device_;
    FeatureDetectOpts opts_;
    orbfe_ctx* ctx_ = nullptr;
    uint32_t cap_ = 0;
    size_t feature_bytes_ = 0;
    std::array<void*, 3> arrays_{};
    int cur_ = 0, ref_ = 1, next_ = 2;
    bool has_ref_ = false, has_cur_ = false;
    uint8_t* host_frame_ = nullptr;
    uint8_t* dev_frame_ = nullptr;
    orbfe_feature* host_set_ = nullptr;
    uint32_t* host_counts_ = nullptr;
    int32_t* dev_match_ = nullptr;
    int32_t* host_match_ = nullptr;
};

class FeatureDetectHipPass : public compute::HIPPass {
   public:
    FeatureDetectHipPass(std::shared_ptr<compute::Device> device, compute::Store& storage,
                         std::shared_ptr<FeatureBuffers> buffers);
    [[nodiscard]] std::string getId() const override;
    [[nodiscard]] std::optional<std::string> initialize() override;
    [[nodiscard]] std::optional<std::string> prepareHost() override;
    [[nodiscard]] std::optional<std::string> prepareExecute(hipStream_t stream) override;
    [[nodiscard]] uint64_t graphKey() const override;

   private:
    compute::Store& storage_;
    std::shared_ptr<FeatureBuffers> buffers_;
};

// Frame of the store to u8 pixels: FrameU8 as is, or the reference's f32
// FrameBW, whose pixels are exactly float(k) / 255 (k = round(v * 255),
// checked per pixel; other values are refused).  Returns the error text.
std::optional<std::string> StageFrameU8(const compute::Store& storage, const std::string& key,
                                        uint32_t width, uint32_t height, uint8_t* out);

class LoadFeaturesHipPass : public compute::Pass {
   public:
    LoadFeaturesHipPass(compute::Store& storage, std::shared_ptr<FeatureBuffers> buffers);
    [[nodiscard]] std::string getId() const override;
    [[nodiscard]] std::optional<std::string> initialize() override;
    [[nodiscard]] std::optional<std::string> execute() override;

   private:
    compute::Store& storage_;
    std::shared_ptr<FeatureBuffers> buffers_;
    size_t reference_age_frames_ = 0;
};

class MatchFeaturesHipPass : public compute::HIPPass {
   public:
    MatchFeaturesHipPass(std::shared_ptr<compute::Device> device,
                         std::shared_ptr<FeatureBuffers> buffers);
    [[nodiscard]] std::string getId() const override;
    [[nodiscard]] std::optional<std::string> initialize() override;
    [[nodiscard]] std::optional<std::string> prepareExecute(hipStream_t stream) override;
    [[nodiscard]] uint64_t graphKey() const override;

   private:
    std::shared_ptr<FeatureBuffers> buffers_;
};

class PublishMatchesPass : public compute::Pass {
   public:
    PublishMatchesPass(compute::Store& storage, std::shared_ptr<FeatureBuffers> buffers);
    [[nodiscard]] std::string getId() const override;
    [[nodiscard]] std::optional<std::string> initialize() override;
    [[nodiscard]] std::optional<std::string> execute() override;

   private:
    compute::Store& storage_;
    std::shared_ptr<FeatureBuffers> buffers_;
};

// Stage factories, mirroring CreateFeatureDetectStage
// (wslam/src/feature_detect/feature_detect.cpp:43-86) and the first half of
// the pose-estimate stage (wslam/src/pose_estimate_cpu/pose_estimate_cpu.cpp:15-50).
std::unique_ptr<compute::HipStage> CreateFeatureDetectStage(
    std::shared_ptr<compute::Device> device, compute::Store& storage,
    std::shared_ptr<FeatureBuffers> buffers);
// With `ransac`, the stage ends with RansacHipPass (the reference's RansacCPU,
// pose_estimate_cpu.cpp:30), its hypotheses scored on the GPU.
std::unique_ptr<compute::HipStage> CreateMatchStage(std::shared_ptr<compute::Device> device,
                                                    compute::Store& storage,
                                                    std::shared_ptr<FeatureBuffers> buffers,
                                                    bool ransac = false);

}  // namespace wslam
