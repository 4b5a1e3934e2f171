// Feature path passes on a HIP device (see feature_detect_hip.hpp).
#include "wslam/feature_detect_hip.hpp"
#include "wslam/ransac_hip.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstddef>
#include <cstring>

namespace wslam {

// orbfe_feature is the C form of wslam::Feature: the exported records are
// the FeatureSet's elements byte for byte.
static_assert(sizeof(orbfe_feature) == sizeof(Feature));
static_assert(offsetof(orbfe_feature, lod) == offsetof(Feature, lod));
static_assert(offsetof(orbfe_feature, orientation) == offsetof(Feature, orientation));
static_assert(offsetof(orbfe_feature, descriptor) == offsetof(Feature, descriptor));

namespace {
std::string hip_msg(const char* what, hipError_t e) {
    return std::string(what) + ": " + hipGetErrorString(e);
}

std::string orbfe_msg(const char* what, orbfe_ctx* ctx, int rc) {
    return std::string(what) + " failed (" + std::to_string(rc) + "): " + orbfe_last_error(ctx);
}
}  // namespace

// --- FeatureBuffers --------------------------------------------------------------
FeatureBuffers::FeatureBuffers(std::shared_ptr<compute::Device> device, FeatureDetectOpts opts)
    : device_(std::move(device)), opts_(opts) {}

FeatureBuffers::~FeatureBuffers() {
    if (device_ && !device_->hostOnly()) (void)device_->synchronize();
    for (void* p : arrays_)
        if (p) (void)hipFree(p);
    if (dev_frame_) (void)hipFree(dev_frame_);
    if (dev_match_) (void)hipFree(dev_match_);
    if (host_frame_) (void)hipHostFree(host_frame_);
    if (host_set_) (void)hipHostFree(host_set_);
    if (host_counts_) (void)hipHostFree(host_counts_);
    if (host_match_) (void)hipHostFree(host_match_);
    if (ctx_) orbfe_destroy(ctx_);
}

std::optional<std::string> FeatureBuffers::initialize() {
    if (ctx_) return std::nullopt;
    orbfe_config cfg;
    orbfe_config_default(&cfg, opts_.frame_width, opts_.frame_height);
    if (cfg.levels != kLevelsOfDetail) return "level count must be " + std::to_string(kLevelsOfDetail);
    cfg.features_per_level = opts_.features_per_level;
    cfg.fast_threshold = opts_.fast_threshold;
    cfg.lowe_ratio = opts_.lowe_ratio;
    cfg.border_zero = opts_.border_zero;
    const int rc = orbfe_create(&cfg, device_->ordinal(), &ctx_);
    if (rc != ORBFE_OK) {
        ctx_ = nullptr;
        return "orbfe_create failed (" + std::to_string(rc) + ")";
    }
    cap_ = orbfe_cap(ctx_);
    feature_bytes_ = orbfe_feature_array_bytes(ctx_);
    const size_t frame_bytes = static_cast<size_t>(opts_.frame_width) * opts_.frame_height;
    const size_t match_bytes = static_cast<size_t>(kLevelsOfDetail) * cap_ * 4;
    hipError_t e;
    for (auto& p : arrays_) {
        if ((e = hipMalloc(&p, feature_bytes_)) != hipSuccess) return hip_msg("hipMalloc", e);
    }
    if ((e = hipMalloc(reinterpret_cast<void**>(&dev_frame_), frame_bytes)) != hipSuccess)
        return hip_msg("hipMalloc", e);
    if ((e = hipMalloc(reinterpret_cast<void**>(&dev_match_), match_bytes)) != hipSuccess)
        return hip_msg("hipMalloc", e);
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&host_frame_), frame_bytes,
                           hipHostMallocDefault)) != hipSuccess)
        return hip_msg("hipHostMalloc", e);
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&host_set_),
                           sizeof(orbfe_feature) * kLevelsOfDetail * cap_,
                           hipHostMallocDefault)) != hipSuccess)
        return hip_msg("hipHostMalloc", e);
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&host_counts_), 4 * kLevelsOfDetail,
                           hipHostMallocDefault)) != hipSuccess)
        return hip_msg("hipHostMalloc", e);
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&host_match_), match_bytes,
                           hipHostMallocDefault)) != hipSuccess)
        return hip_msg("hipHostMalloc", e);
    return std::nullopt;
}

void FeatureBuffers::rotate(bool advance_reference) {
    const bool shift = advance_reference && has_cur_;
    const int c = cur_, r = ref_, n = next_;
    if (shift) {
        cur_ = n;
        ref_ = c;
        next_ = r;
        has_ref_ = true;
    } else {
        cur_ = n;
        next_ = c;
    }
    has_cur_ = true;
}

// --- FeatureDetectHipPass ----------------------------------------------------------
FeatureDetectHipPass::FeatureDetectHipPass(std::shared_ptr<compute::Device> device,
                                           compute::Store& storage,
                                           std::shared_ptr<FeatureBuffers> buffers)
    : HIPPass(std::move(device)), storage_(storage), buffers_(std::move(buffers)) {}

std::string FeatureDetectHipPass::getId() const { return "[Feature detect HIP pass]"; }

std::optional<std::string> FeatureDetectHipPass::initialize() { return buffers_->initialize(); }

std::optional<std::string> StageFrameU8(const compute::Store& storage, const std::string& key,
                                        uint32_t width, uint32_t height, uint8_t* out) {
    const auto size_msg = [&](uint32_t w, uint32_t h) {
        return "frame is " + std::to_string(w) + "x" + std::to_string(h) +
               ", the stage was built for " + std::to_string(width) + "x" +
               std::to_string(height);
    };
    const size_t n = static_cast<size_t>(width) * height;
    if (auto u8 = storage.getPtr<FrameU8>(key)) {
        const FrameU8& f = **u8;
        if (f.width != width || f.height != height) return size_msg(f.width, f.height);
        if (f.pixels.size() != n) return "data is empty";
        std::memcpy(out, f.pixels.data(), n);
        return std::nullopt;
    }
    if (auto bw = storage.getPtr<data::FrameBW>(key)) {
        const data::FrameBW& f = **bw;
        if (f.width != width || f.height != height) return size_msg(f.width, f.height);
        if (f.pixels.size() != n) return "data is empty";
        // The loader's pixels are fl(k / 255) for a u8 k (provider_base.cpp:121-124):
        // k = round(v * 255) recovers k, and fl(k / 255) == v confirms it.
        float unit[256];
        for (int k = 0; k < 256; ++k) unit[k] = static_cast<float>(k) / 255.0f;
        bool ok = true;
        for (size_t i = 0; i < n; ++i) {
            const float v = f.pixels[i];
            const float r = std::nearbyint(v * 255.0f);
            const uint32_t k = r >= 0.0f && r <= 255.0f ? static_cast<uint32_t>(r) : 0u;
            ok &= unit[k] == v;
            out[i] = static_cast<uint8_t>(k);
        }
        if (!ok) return "frame pixels are not float(u8) / 255 values";
        return std::nullopt;
    }
    return "end of data (no frame under " + key + ")";
}

std::optional<std::string> FeatureDetectHipPass::prepareHost() {
    const auto& o = buffers_->opts();
    return StageFrameU8(storage_, ResourceIdentifier::GetFrameName(0), o.frame_width,
                        o.frame_height, buffers_->hostFrame());
}

std::optional<std::string> FeatureDetectHipPass::prepareExecute(hipStream_t stream) {
    const auto& o = buffers_->opts();
    const size_t bytes = static_cast<size_t>(o.frame_width) * o.frame_height;
    hipError_t e = hipMemcpyAsync(buffers_->deviceFrame(), buffers_->hostFrame(), bytes,
                                  hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_msg("upload frame", e);
    int rc = orbfe_detect_async(buffers_->ctx(), buffers_->deviceFrame(), o.frame_width, nullptr,
                                buffers_->next(), stream);
    if (rc != ORBFE_OK) return orbfe_msg("orbfe_detect_async", buffers_->ctx(), rc);
    // The FeatureSet straight into pinned memory (no slab copy, no CPU loop).
    rc = orbfe_export_features_async(
        buffers_->ctx(), buffers_->next(),
        reinterpret_cast<orbfe_feature*>(const_cast<Feature*>(buffers_->hostSet())),
        const_cast<uint32_t*>(buffers_->hostCounts()), stream);
    if (rc != ORBFE_OK) return orbfe_msg("orbfe_export_features_async", buffers_->ctx(), rc);
    return std::nullopt;
}

uint64_t FeatureDetectHipPass::graphKey() const {
    return static_cast<uint64_t>(buffers_->nextSlot());
}

// --- LoadFeaturesHipPass -------------------------------------------------------------
LoadFeaturesHipPass::LoadFeaturesHipPass(compute::Store& storage,
                                         std::shared_ptr<FeatureBuffers> buffers)
    : storage_(storage), buffers_(std::move(buffers)) {}

std::string LoadFeaturesHipPass::getId() const { return "[Load features HIP pass]"; }

std::optional<std::string> LoadFeaturesHipPass::initialize() { return std::nullopt; }

std::optional<std::string> LoadFeaturesHipPass::execute() {
    // loadFeatures (load_features_cpu.cpp:164-188): the GPU already wrote
    // min(count, cap) wslam::Feature records per level into pinned memory.
    FeatureSet result;
    const Feature* rec = buffers_->hostSet();
    const uint32_t* counts = buffers_->hostCounts();
    for (uint32_t lod = 0; lod < kLevelsOfDetail; ++lod) {
        if (counts[lod] > buffers_->cap()) return "exported level count exceeds the cap";
        result[lod].assign(rec, rec + counts[lod]);
        rec += counts[lod];
    }
    // shiftFeatureSets (load_features_cpu.cpp:111-142)
    const bool advance_requested =
        storage_.take<bool>(ResourceIdentifier::FeatureReferenceAdvanceName).value_or(false);
    const bool have_reference = storage_.has(ResourceIdentifier::GetFeatureSetName(1));
    bool shifted = false;
    if (have_reference && !advance_requested) {
        ++reference_age_frames_;
    } else {
        auto set = storage_.take<FeatureSet>(ResourceIdentifier::GetFeatureSetName(0));
        if (set) {
            reference_age_frames_ = 0;
            storage_.set(ResourceIdentifier::GetFeatureSetName(1), std::move(*set));
            shifted = true;
        }
    }
    buffers_->rotate(shifted);
    storage_.set(ResourceIdentifier::GetFeatureSetName(0), std::move(result));
    return std::nullopt;
}

// --- MatchFeaturesHipPass -------------------------------------------------------------
MatchFeaturesHipPass::MatchFeaturesHipPass(std::shared_ptr<compute::Device> device,
                                           std::shared_ptr<FeatureBuffers> buffers)
    : HIPPass(std::move(device)), buffers_(std::move(buffers)) {}

std::string MatchFeaturesHipPass::getId() const { return "[Match features HIP pass]"; }

std::optional<std::string> MatchFeaturesHipPass::initialize() { return std::nullopt; }

std::optional<std::string> MatchFeaturesHipPass::prepareExecute(hipStream_t stream) {
    if (!buffers_->hasReference()) return std::nullopt;  // nothing to match yet
    const int rc = orbfe_match_pipeline_async(buffers_->ctx(), buffers_->reference(),
                                              buffers_->current(), buffers_->deviceMatch(), stream);
    if (rc != ORBFE_OK) return orbfe_msg("orbfe_match_pipeline_async", buffers_->ctx(), rc);
    const hipError_t e =
        hipMemcpyAsync(buffers_->hostMatch(), buffers_->deviceMatch(),
                       static_cast<size_t>(kLevelsOfDetail) * buffers_->cap() * 4,
                       hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return hip_msg("read back matches", e);
    return std::nullopt;
}

uint64_t MatchFeaturesHipPass::graphKey() const {
    if (!buffers_->hasReference()) return 0;  // enqueues nothing
    return 1u | static_cast<uint64_t>(buffers_->referenceSlot()) << 4 |
           static_cast<uint64_t>(buffers_->currentSlot()) << 8;
}

// --- PublishMatchesPass -----------------------------------------------------------------
PublishMatchesPass::PublishMatchesPass(compute::Store& storage,
                                       std::shared_ptr<FeatureBuffers> buffers)
    : storage_(storage), buffers_(std::move(buffers)) {}

std::string PublishMatchesPass::getId() const { return "[Match features pass]"; }

std::optional<std::string> PublishMatchesPass::initialize() { return std::nullopt; }

std::optional<std::string> PublishMatchesPass::execute() {
    const auto prev = storage_.getPtr<FeatureSet>(ResourceIdentifier::GetFeatureSetName(1));
    if (!prev || !buffers_->hasReference()) {
        // No reference feature set yet (match_features_cpu.cpp:228-236).
        return compute::kStageStopExecution;
    }
    const auto curr = storage_.getPtr<FeatureSet>(ResourceIdentifier::GetFeatureSetName(0));
    if (!curr) return "no current feature set";
    const FeatureSet& A = **prev;
    const FeatureSet& B = **curr;
    const int32_t* pm = buffers_->hostMatch();
    const uint32_t cap = buffers_->cap();
    MatchResult result;
    for (uint32_t l = 0; l < kLevelsOfDetail; ++l) {
        for (size_t i = 0; i < A[l].size(); ++i) {
            const int32_t m = pm[l * cap + i];
            if (m < 0) continue;
            const uint32_t bl = static_cast<uint32_t>(m) >> 24;
            const uint32_t bi = static_cast<uint32_t>(m) & 0xFFFFFFu;
            if (bl >= kLevelsOfDetail || bi >= B[bl].size()) return "match index out of range";
            result[l].emplace(B[bl][bi], A[l][i]);
        }
    }
    storage_.set(ResourceIdentifier::MatchedFeaturesName, std::move(result));
    return std::nullopt;
}

// --- stage factories ----------------------------------------------------------------------
std::unique_ptr<compute::HipStage> CreateFeatureDetectStage(
    std::shared_ptr<compute::Device> device, compute::Store& storage,
    std::shared_ptr<FeatureBuffers> buffers) {
    auto stage = std::make_unique<compute::HipStage>("Feature detect", device);
    stage->add_pass(std::unique_ptr<compute::HIPPass>(
        new FeatureDetectHipPass(device, storage, buffers)));
    return stage;
}

std::unique_ptr<compute::HipStage> CreateMatchStage(std::shared_ptr<compute::Device> device,
                                                    compute::Store& storage,
                                                    std::shared_ptr<FeatureBuffers> buffers,
                                                    bool ransac) {
    auto stage = std::make_unique<compute::HipStage>("Pose estimate", device);
    stage->add_pass(std::unique_ptr<compute::Pass>(new LoadFeaturesHipPass(storage, buffers)));
    stage->add_pass(std::unique_ptr<compute::HIPPass>(new MatchFeaturesHipPass(device, buffers)));
    stage->add_pass(std::unique_ptr<compute::Pass>(new PublishMatchesPass(storage, buffers)));
    if (ransac)
        stage->add_pass(std::unique_ptr<compute::Pass>(new RansacHipPass(storage, device, buffers)));
    return stage;
}

}  // namespace wslam
