#pragma once

// The pipeline engine: an ordered list of stages run once per execute()
// (compute/include/compute/compute.hpp:46-84 of the reference).  A stage
// returning kComputeStopExecution ends the iteration without error;
// finalize() runs the registered flushables.

#include <functional>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "compute/stage.hpp"
#include "compute/store.hpp"

namespace wslam::compute {

inline const std::string kComputeStopExecution = "COMP_STOP";

class Compute {
   public:
    using Flushable = std::function<std::optional<std::string>()>;

    void addStage(std::unique_ptr<Stage> stage) { stages_.emplace_back(std::move(stage)); }
    void addFlushable(Flushable f) { flushes_.emplace_back(std::move(f)); }

    [[nodiscard]] std::optional<std::string> initialize();
    [[nodiscard]] std::optional<std::string> execute();
    [[nodiscard]] std::optional<std::vector<std::string>> finalize();

    [[nodiscard]] Store& getStorage() { return storage_; }

   private:
    std::vector<std::unique_ptr<Stage>> stages_;
    std::vector<Flushable> flushes_;
    Store storage_;
};

}  // namespace wslam::compute
