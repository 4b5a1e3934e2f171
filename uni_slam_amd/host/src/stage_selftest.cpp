// Host-only checks of the stage machinery (no GPU): HIP passes are batched
// in insertion order and flushed before each CPU pass, kStageStopExecution
// ends a stage without error, kComputeStopExecution ends the iteration,
// other errors carry the pass id.  Exit status 0 when all checks hold.
#include <iostream>
#include <string>
#include <vector>

#include "compute/compute.hpp"
#include "compute/hip_stage.hpp"

using namespace wslam::compute;

namespace {
std::vector<std::string> g_log;

struct LogHip : HIPPass {
    LogHip(std::shared_ptr<Device> d, std::string id, std::string fail = "")
        : HIPPass(std::move(d)), id_(std::move(id)), fail_(std::move(fail)) {}
    std::string getId() const override { return id_; }
    std::optional<std::string> initialize() override { return std::nullopt; }
    std::optional<std::string> prepareExecute(hipStream_t) override {
        g_log.push_back("hip:" + id_);
        if (!fail_.empty()) return fail_;
        return std::nullopt;
    }
    std::string id_, fail_;
};

struct LogCpu : Pass {
    explicit LogCpu(std::string id, std::optional<std::string> ret = std::nullopt)
        : id_(std::move(id)), ret_(std::move(ret)) {}
    std::string getId() const override { return id_; }
    std::optional<std::string> initialize() override { return std::nullopt; }
    std::optional<std::string> execute() override {
        g_log.push_back("cpu:" + id_);
        return ret_;
    }
    std::string id_;
    std::optional<std::string> ret_;
};

int failures = 0;
void expect(bool ok, const std::string& what) {
    if (!ok) {
        std::cerr << "FAILED: " << what << "\n";
        ++failures;
    }
}
std::string joined() {
    std::string s;
    for (auto& e : g_log) s += e + ",";
    return s;
}
}  // namespace

int main() {
    auto dev = Device::createHostOnly();
    {
        HipStage st("s", dev);
        st.add_pass(std::unique_ptr<HIPPass>(new LogHip(dev, "a")));
        st.add_pass(std::unique_ptr<HIPPass>(new LogHip(dev, "b")));
        st.add_pass(std::unique_ptr<Pass>(new LogCpu("c")));
        st.add_pass(std::unique_ptr<HIPPass>(new LogHip(dev, "d")));
        g_log.clear();
        expect(!st.initialize(), "initialize");
        expect(!st.execute(), "execute ok");
        expect(joined() == "hip:a,hip:b,cpu:c,hip:d,", "order: " + joined());
    }
    {
        HipStage st("s", dev);
        st.add_pass(std::unique_ptr<Pass>(new LogCpu("stop", kStageStopExecution)));
        st.add_pass(std::unique_ptr<HIPPass>(new LogHip(dev, "never")));
        g_log.clear();
        expect(!st.execute(), "stage stop is not an error");
        expect(joined() == "cpu:stop,", "stage stop ends the stage: " + joined());
    }
    {
        Compute comp;
        auto s1 = std::make_unique<HipStage>("one", dev);
        s1->add_pass(std::unique_ptr<Pass>(new LogCpu("halt", kComputeStopExecution)));
        auto s2 = std::make_unique<HipStage>("two", dev);
        s2->add_pass(std::unique_ptr<Pass>(new LogCpu("never")));
        comp.addStage(std::move(s1));
        comp.addStage(std::move(s2));
        g_log.clear();
        expect(!comp.execute(), "compute stop is not an error");
        expect(joined() == "cpu:halt,", "compute stop ends the iteration: " + joined());
    }
    {
        HipStage st("s", dev);
        st.add_pass(std::unique_ptr<HIPPass>(new LogHip(dev, "bad", "boom")));
        auto err = st.execute();
        expect(err && err->find("bad") != std::string::npos && err->find("boom") != std::string::npos,
               "HIP pass errors name the pass");
    }
    {
        Store s;
        s.set("k", 5);
        expect(s.get<int>("k").value_or(0) == 5, "store get");
        expect(!s.get<float>("k"), "store type check");
        expect(s.take<int>("k").value_or(0) == 5 && !s.has("k"), "store take erases");
    }
    if (failures == 0) std::cout << "stage selftest ok\n";
    return failures == 0 ? 0 : 1;
}
