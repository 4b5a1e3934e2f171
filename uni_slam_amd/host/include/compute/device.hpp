#pragma once

// One GPU and the HIP stream the pipeline's HIP passes record on.  Plays the
// part the reference's GPU object (compute/include/compute/gpu.hpp:120-238)
// plays for its passes: passes hold a shared_ptr to it.

#include <hip/hip_runtime_api.h>

#include <memory>
#include <optional>
#include <string>

namespace wslam::compute {

class Device {
   public:
    // Selects GPU `ordinal` and creates its stream; nullptr + message on error.
    static std::shared_ptr<Device> create(int ordinal, std::string* error);
    // A device with no GPU behind it, for host-only tests of the stage
    // machinery: stream() is null and synchronize() does nothing.
    static std::shared_ptr<Device> createHostOnly();

    ~Device();
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;

    [[nodiscard]] hipStream_t stream() const { return stream_; }
    [[nodiscard]] int ordinal() const { return ordinal_; }
    [[nodiscard]] bool hostOnly() const { return host_only_; }

    // Waits for everything recorded on the stream.
    [[nodiscard]] std::optional<std::string> synchronize() const;

   private:
    Device() = default;
    int ordinal_ = -1;
    bool host_only_ = true;
    hipStream_t stream_ = nullptr;
};

}  // namespace wslam::compute
