#pragma once

// Pass interfaces of the pipeline engine, after the reference's
// compute/include/compute/pass.hpp:19-56.  Errors follow the reference's
// convention: std::nullopt is success, otherwise the message.
//
// HIPPass is the MI355X sibling of the reference's GPUPass: prepareExecute
// only enqueues work on the given HIP stream (it never waits), so that a
// HipStage can batch consecutive HIP passes into one stream sequence (or one
// HIP graph) followed by a single synchronisation.

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <string>

#include "compute/device.hpp"

namespace wslam::compute {

class Pass {
   public:
    Pass() = default;
    virtual ~Pass() = default;

    [[nodiscard]] virtual std::string getId() const = 0;
    [[nodiscard]] virtual std::optional<std::string> initialize() = 0;
    [[nodiscard]] virtual std::optional<std::string> execute() = 0;
};

class HIPPass {
   public:
    explicit HIPPass(std::shared_ptr<Device> device) : device_(std::move(device)) {}
    virtual ~HIPPass() = default;

    [[nodiscard]] virtual std::string getId() const = 0;
    [[nodiscard]] virtual std::optional<std::string> initialize() = 0;
    [[nodiscard]] virtual std::optional<std::string> prepareExecute(hipStream_t stream) = 0;

    // Host work of one execution that must happen every time and outside
    // any captured region (e.g. staging the frame into pinned memory).  The
    // stage calls it before prepareExecute, and before each replay of a
    // captured graph in place of prepareExecute.
    [[nodiscard]] virtual std::optional<std::string> prepareHost() { return std::nullopt; }
    // Names the device work prepareExecute would enqueue now (which buffers,
    // which sizes): equal keys mean identical work, so a HipStage with graph
    // capture replays the graph it captured under the same key.
    [[nodiscard]] virtual uint64_t graphKey() const { return 0; }

   protected:
    std::shared_ptr<Device> device_;
};

class CustomPass : public Pass {
   public:
    using Callback = std::function<std::optional<std::string>(CustomPass*)>;

    CustomPass(std::string id, Callback callback)
        : callback_(std::move(callback)), id_(std::move(id)) {}

    [[nodiscard]] std::optional<std::string> initialize() override { return std::nullopt; }
    [[nodiscard]] std::optional<std::string> execute() override { return callback_(this); }
    [[nodiscard]] std::string getId() const override { return id_; }

    Callback callback_;
    const std::string id_;
};

}  // namespace wslam::compute
