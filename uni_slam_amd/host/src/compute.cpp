// Pipeline engine: Device, Stage, HipStage, Compute.
#include <hip/hip_runtime_api.h>

#include <sstream>

#include "compute/compute.hpp"
#include "compute/hip_stage.hpp"

namespace wslam::compute {

namespace {
std::string hip_message(const char* what, hipError_t e) {
    std::ostringstream os;
    os << what << ": " << hipGetErrorString(e);
    return os.str();
}

std::string pass_id(const HipStage::PassPtr& p) {
    return std::visit([](const auto& q) { return q->getId(); }, p);
}
}  // namespace

// --- Device ------------------------------------------------------------------
std::shared_ptr<Device> Device::create(int ordinal, std::string* error) {
    std::shared_ptr<Device> d(new Device());
    hipError_t e = hipSetDevice(ordinal);
    if (e != hipSuccess) {
        if (error) *error = hip_message("hipSetDevice", e);
        return nullptr;
    }
    // default priority stream
    e = hipStreamCreate(&d->stream_);
    if (e != hipSuccess) {
        if (error) *error = hip_message("hipStreamCreate", e);
        return nullptr;
    }
    d->ordinal_ = ordinal;
    d->host_only_ = false;
    return d;
}

std::shared_ptr<Device> Device::createHostOnly() { return std::shared_ptr<Device>(new Device()); }

Device::~Device() {
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

std::optional<std::string> Device::synchronize() const {
    if (host_only_) return std::nullopt;
    const hipError_t e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return hip_message("hipStreamSynchronize", e);
    return std::nullopt;
}

// --- Stage -------------------------------------------------------------------
std::optional<std::string> Stage::initialize() {
    for (auto& pass : passes_) {
        if (auto err = pass->initialize())
            return "initializing pass " + pass->getId() + ": " + *err;
    }
    return std::nullopt;
}

std::optional<std::string> Stage::execute() {
    for (auto& pass : passes_) {
        auto err = pass->execute();
        if (!err) continue;
        if (*err == kStageStopExecution) return std::nullopt;
        if (*err == kComputeStopExecution) return err;
        return pass->getId() + ": " + *err;
    }
    return std::nullopt;
}

// --- HipStage ----------------------------------------------------------------
HipStage::HipStage(std::string id, std::shared_ptr<Device> device)
    : Stage(std::move(id)), device_(std::move(device)) {}

HipStage::~HipStage() { resetGraphs(); }

void HipStage::add_pass(std::unique_ptr<Pass> pass) { passes_.emplace_back(std::move(pass)); }

void HipStage::add_pass(std::unique_ptr<HIPPass> pass) { passes_.emplace_back(std::move(pass)); }

void HipStage::resetGraphs() {
    for (auto& by_key : graphs_)
        for (auto& kv : by_key)
            if (kv.second) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
}

std::optional<std::string> HipStage::initialize() {
    for (auto& pass : passes_) {
        auto err = std::visit([](auto& p) { return p->initialize(); }, pass);
        if (err) return "initializing pass " + pass_id(pass) + ": " + *err;
    }
    return std::nullopt;
}

std::optional<std::string> HipStage::executeHIPBatch(std::vector<HIPPass*>& batch,
                                                     size_t batch_index) {
    if (batch.empty()) return std::nullopt;
    for (auto* pass : batch) {
        if (auto err = pass->prepareHost()) return "preparing pass " + pass->getId() + ": " + *err;
    }
    const hipStream_t s = device_->stream();
    const bool capture = capture_ && !device_->hostOnly();
    uint64_t key = 0xcbf29ce484222325ull;  // FNV-1a over the passes' keys
    if (capture) {
        for (auto* pass : batch) {
            const uint64_t k = pass->graphKey();
            for (int b = 0; b < 8; ++b) key = (key ^ ((k >> (8 * b)) & 0xFFu)) * 0x100000001b3ull;
        }
        if (graphs_.size() <= batch_index) graphs_.resize(batch_index + 1);
        const auto it = graphs_[batch_index].find(key);
        if (it != graphs_[batch_index].end()) {
            if (it->second) {
                const hipError_t e = hipGraphLaunch(it->second, s);
                if (e != hipSuccess) return hip_message("hipGraphLaunch", e);
            }
            return device_->synchronize();
        }
        // Relaxed: the passes' library calls may make non-stream HIP calls
        // (pointer queries) while the stream records.
        const hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) return hip_message("hipStreamBeginCapture", e);
    }
    std::optional<std::string> failure;
    for (auto* pass : batch) {
        if (auto err = pass->prepareExecute(s)) {
            failure = "preparing pass " + pass->getId() + ": " + *err;
            break;
        }
    }
    if (capture) {
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (failure) {
            if (graph) (void)hipGraphDestroy(graph);
            return failure;
        }
        if (e != hipSuccess) return hip_message("hipStreamEndCapture", e);
        size_t nodes = 0;
        hipError_t en = hipGraphGetNodes(graph, nullptr, &nodes);
        if (en != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return hip_message("hipGraphGetNodes", en);
        }
        hipGraphExec_t exec = nullptr;
        if (nodes > 0) {
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) return hip_message("hipGraphInstantiate", ei);
            const hipError_t el = hipGraphLaunch(exec, s);
            if (el != hipSuccess) {
                (void)hipGraphExecDestroy(exec);
                return hip_message("hipGraphLaunch", el);
            }
        } else {
            (void)hipGraphDestroy(graph);
        }
        graphs_[batch_index].emplace(key, exec);
    } else if (failure) {
        return failure;
    }
    return device_->synchronize();
}

std::optional<std::string> HipStage::runBatched() {
    static_assert(std::variant_size_v<PassPtr> == 2, "CPU and HIP passes only");
    std::vector<HIPPass*> batch;
    size_t batch_index = 0;
    for (auto& pass : passes_) {
        if (auto* hp = std::get_if<std::unique_ptr<HIPPass>>(&pass)) {
            batch.push_back(hp->get());
            continue;
        }
        if (auto err = executeHIPBatch(batch, batch_index)) return err;
        if (!batch.empty()) ++batch_index;
        batch.clear();
        auto& cpu = std::get<std::unique_ptr<Pass>>(pass);
        auto err = cpu->execute();
        if (!err) continue;
        if (*err == kStageStopExecution) return std::nullopt;
        if (*err == kComputeStopExecution) return err;
        return cpu->getId() + ": " + *err;
    }
    return executeHIPBatch(batch, batch_index);
}

std::optional<std::string> HipStage::execute() { return runBatched(); }

// --- Compute -----------------------------------------------------------------
std::optional<std::string> Compute::initialize() {
    for (auto& stage : stages_) {
        if (auto err = stage->initialize())
            return "initializing stage " + stage->getId() + ": " + *err;
    }
    return std::nullopt;
}

std::optional<std::string> Compute::execute() {
    for (auto& stage : stages_) {
        if (auto err = stage->execute()) {
            if (*err == kComputeStopExecution) return std::nullopt;
            return "executing stage " + stage->getId() + ": " + *err;
        }
    }
    return std::nullopt;
}

std::optional<std::vector<std::string>> Compute::finalize() {
    std::vector<std::string> errors;
    for (auto& f : flushes_)
        if (auto err = f()) errors.push_back(*err);
    if (errors.empty()) return std::nullopt;
    return errors;
}

}  // namespace wslam::compute
