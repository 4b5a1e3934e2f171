#pragma once

// A named, ordered sequence of CPU passes (compute/include/compute/stage.hpp:
// 18-38 of the reference).  A pass returning kStageStopExecution ends the
// stage early without error.

#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "compute/pass.hpp"

namespace wslam::compute {

inline const std::string kStageStopExecution = "STAGE_STOP";

class Stage {
   public:
    explicit Stage(std::string id) : id_(std::move(id)) {}
    Stage(const Stage&) = delete;
    Stage& operator=(const Stage&) = delete;
    virtual ~Stage() = default;

    [[nodiscard]] std::string getId() const { return id_; }
    [[nodiscard]] virtual std::optional<std::string> initialize();
    [[nodiscard]] virtual std::optional<std::string> execute();

    void add_pass(std::unique_ptr<Pass> pass) { passes_.emplace_back(std::move(pass)); }

   protected:
    std::string id_;

   private:
    std::vector<std::unique_ptr<Pass>> passes_;
};

}  // namespace wslam::compute
