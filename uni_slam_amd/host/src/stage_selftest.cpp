// Host-only checks of the stage machinery (no GPU): HIP passes are batched
// in insertion order and flushed before each CPU pass, kStageStopExecution
// ends a stage without error, kComputeStopExecution ends the iteration,
// other errors carry the pass id.  Exit status 0 when all checks hold.
#include <iostream>
#include <string>
#include <vector>

#include "compute/compute.hpp"
#include "compute/hip_stage.hpp"
#include "wslam/feature_detect_hip.hpp"

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
    {
        // The reference's f32 frame (every u8 value as float(k) / 255)
        // stages to the same bytes as the u8 frame; a value that is not of
        // that form, a wrong size and a missing frame are refused.
        using namespace wslam;
        Store s;
        data::FrameBW bw{7, {}, 16, 16};
        for (int k = 0; k < 256; ++k) bw.pixels.push_back(static_cast<float>(k) / 255.0f);
        s.set("f", bw);
        uint8_t out[256];
        expect(!StageFrameU8(s, "f", 16, 16, out), "f32 frame stages");
        bool same = true;
        for (int k = 0; k < 256; ++k) same &= out[k] == k;
        expect(same, "f32 frame -> u8 recovers every value");
        bw.pixels[77] = 0.5f;  // 127.5 / 255: not a u8 value
        s.set("g", bw);
        expect(StageFrameU8(s, "g", 16, 16, out).has_value(), "non-u8 pixel refused");
        expect(StageFrameU8(s, "f", 32, 8, out).has_value(), "wrong size refused");
        expect(StageFrameU8(s, "missing", 16, 16, out).value_or("").find("end of data") == 0,
               "missing frame is end of data");
        FrameU8 u{16, 16, 0, std::vector<uint8_t>(256, 9)};
        s.set("u", u);
        expect(!StageFrameU8(s, "u", 16, 16, out) && out[200] == 9, "u8 frame stages");
    }
    if (failures == 0) std::cout << "stage selftest ok\n";
    return failures == 0 ? 0 : 1;
}
