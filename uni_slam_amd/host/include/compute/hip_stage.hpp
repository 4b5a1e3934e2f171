#pragma once

// Stage that batches consecutive HIP passes (the reference's GpuStage,
// compute/include/compute/gpu_stage.hpp:22-47, for HIP): the passes'
// prepareExecute calls record onto the device stream, then one
// synchronisation; CPU passes between them flush the batch first, in
// insertion order.
//
// With graph capture enabled, each HIP batch is captured into a HIP graph
// the first time it runs under a given combination of its passes'
// graphKey()s, and that graph is replayed whenever the combination comes
// back.  The passes' prepareHost() runs before every execution, captured or
// replayed, outside the captured region; prepareExecute runs only while
// capturing.  resetGraphs() drops every graph.

#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <optional>
#include <string>
#include <variant>
#include <vector>

#include "compute/device.hpp"
#include "compute/pass.hpp"
#include "compute/stage.hpp"

namespace wslam::compute {

class HipStage : public Stage {
   public:
    using PassPtr = std::variant<std::unique_ptr<Pass>, std::unique_ptr<HIPPass>>;

    HipStage(std::string id, std::shared_ptr<Device> device);
    ~HipStage() override;

    void add_pass(std::unique_ptr<Pass> pass);
    void add_pass(std::unique_ptr<HIPPass> pass);

    void enableGraphCapture(bool enable) { capture_ = enable; }
    void resetGraphs();

    [[nodiscard]] std::optional<std::string> initialize() override;
    [[nodiscard]] std::optional<std::string> execute() override;

   private:
    [[nodiscard]] std::optional<std::string> runBatched();
    [[nodiscard]] std::optional<std::string> executeHIPBatch(std::vector<HIPPass*>& batch,
                                                             size_t batch_index);

    std::vector<PassPtr> passes_;
    std::shared_ptr<Device> device_;
    bool capture_ = false;
    // per HIP batch position: graph of each key combination (nullptr when
    // the batch enqueued nothing under it)
    std::vector<std::map<uint64_t, hipGraphExec_t>> graphs_;
};

}  // namespace wslam::compute
