#pragma once

// RANSAC over the frame's matches with the hypotheses scored on the GPU:
// the drop-in for the reference's RansacCPU pass
// (wslam/src/pose_estimate_cpu/ransac_cpu.{hpp,cpp}).
//
// The reference's loop draws an 8-sample, fits F with the normalised 8-point
// algorithm and scores it against every match (Sampson distance below the
// threshold), one hypothesis at a time, tightening the adaptive iteration
// bound whenever the best inlier count grows.  Here the samples are drawn
// with the same generator in the same order and fitted on the host, a round
// of hypotheses at a time, and each round is scored in one launch on the GPU
// (orbfe_ransac_score_async); the reference's loop then replays over the
// round's counts, so it stops at the same iteration with the same best
// model, and the generator is left where the reference's would be.  The
// refit on the best inlier set and its classification follow the reference.

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <random>
#include <string>
#include <vector>

#include "../../../../include/orbfe.h"
#include "compute/pass.hpp"
#include "compute/store.hpp"
#include "wslam/feature_detect_hip.hpp"
#include "wslam/models.hpp"

namespace wslam {

// RansacCPU::Opts (ransac_cpu.hpp:20-45), same defaults.
struct RansacOpts {
    uint32_t max_iterations = 500;
    double confidence = 0.99;
    double inlier_threshold = 3.84;  // squared LOD-0 pixels
    size_t min_matches = 8;
    uint64_t rng_seed = 0xC0FFEE'1234ULL;
    uint32_t round = 64;  // hypotheses scored per launch (no counterpart)
};

struct Point2 {
    double x, y;
};
using Mat3 = std::array<double, 9>;  // row-major

// Scores hypotheses: counts[h] = inliers of F[h] over the correspondences;
// with a mask (one hypothesis), mask[i] = 1 for each inlier.
using RansacScorer = std::function<std::optional<std::string>(
    const std::vector<Mat3>& F, std::vector<uint32_t>& counts, std::vector<uint8_t>* mask)>;

struct RansacOutcome {
    RansacStats stats;
    Mat3 F{};
    std::vector<size_t> inliers;  // indices into the flattened matches, ascending
};

// The normalised 8-point algorithm of FitFundamentalMatrix
// (ransac_cpu.cpp:111-175) on >= 8 correspondences; nullopt when degenerate.
std::optional<Mat3> FitFundamental(const std::vector<Point2>& prev,
                                   const std::vector<Point2>& curr);

// The reference's 8 distinct indices in [0, n) (SampleEight, ransac_cpu.cpp:212-225).
std::array<size_t, 8> SampleEightIndices(size_t n, std::mt19937_64& rng);

// RansacCPU::execute's model search (ransac_cpu.cpp:262-370) on flattened
// correspondences, hypotheses scored by `score` a round at a time.
std::optional<std::string> RunRansac(const std::vector<Point2>& prev,
                                     const std::vector<Point2>& curr, const RansacOpts& opts,
                                     std::mt19937_64& rng, const RansacScorer& score,
                                     RansacOutcome& out);

// Flattened correspondences of a MatchResult in the reference's order
// (FlattenMatches, ransac_cpu.cpp:48-72): levels in order, each map in key
// order; LOD-0 coordinates as double.
void FlattenMatchResult(const MatchResult& m, std::vector<Point2>& prev,
                        std::vector<Point2>& curr, std::vector<std::pair<Feature, Feature>>& feats);

// GPU scorer: correspondences uploaded once, F's per round, one launch each.
class GpuRansacScorer {
   public:
    GpuRansacScorer(orbfe_ctx* ctx, hipStream_t stream) : ctx_(ctx), stream_(stream) {}
    ~GpuRansacScorer();
    GpuRansacScorer(const GpuRansacScorer&) = delete;
    GpuRansacScorer& operator=(const GpuRansacScorer&) = delete;

    std::optional<std::string> upload(const std::vector<Point2>& prev,
                                      const std::vector<Point2>& curr, double threshold);
    std::optional<std::string> score(const std::vector<Mat3>& F, std::vector<uint32_t>& counts,
                                     std::vector<uint8_t>* mask);

   private:
    std::optional<std::string> reserve(size_t n, size_t k);
    orbfe_ctx* ctx_;
    hipStream_t stream_;
    double threshold_ = 0.0;
    size_t n_ = 0, cap_n_ = 0, cap_k_ = 0;
    double* d_prev_ = nullptr;
    double* d_curr_ = nullptr;
    double* d_F_ = nullptr;
    uint32_t* d_counts_ = nullptr;
    uint8_t* d_mask_ = nullptr;
    double* h_F_ = nullptr;  // pinned staging
    uint32_t* h_counts_ = nullptr;
    uint8_t* h_mask_ = nullptr;
};

// The pass: MatchResult (gen:feat:match) in, RansacResult (gen:feat:ransac)
// out, as RansacCPU::execute.  Its generator persists across frames.
class RansacHipPass : public compute::Pass {
   public:
    RansacHipPass(compute::Store& storage, std::shared_ptr<compute::Device> device,
                  std::shared_ptr<FeatureBuffers> buffers, RansacOpts opts = {});
    ~RansacHipPass() override;
    [[nodiscard]] std::string getId() const override;
    [[nodiscard]] std::optional<std::string> initialize() override;
    [[nodiscard]] std::optional<std::string> execute() override;

   private:
    compute::Store& storage_;
    std::shared_ptr<compute::Device> device_;
    std::shared_ptr<FeatureBuffers> buffers_;
    RansacOpts opts_;
    std::mt19937_64 rng_;
    std::unique_ptr<GpuRansacScorer> scorer_;  // on the buffers' context, from initialize()
};

}  // namespace wslam
