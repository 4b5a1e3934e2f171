// Host-only checks of the stage machinery (no GPU): HIP passes are batched
// in insertion order and flushed before each CPU pass, kStageStopExecution
// ends a stage without error, kComputeStopExecution ends the iteration,
// other errors carry the pass id.  Exit status 0 when all checks hold.
#include <algorithm>
#include <cmath>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "compute/compute.hpp"
#include "compute/hip_stage.hpp"
#include "wslam/feature_detect_hip.hpp"
#include "wslam/ransac_hip.hpp"

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
    {
        // The 8-point fit of the RANSAC pass recovers a known fundamental
        // matrix from exact two-view correspondences: cameras K[I|0] and
        // K[R|t], F = K^-T [t]_x R K^-1, compared up to scale and sign.
        using namespace wslam;
        const double f = 900.0, cx = 640.0, cy = 360.0, a = 0.1;
        const double R[9] = {std::cos(a), 0, std::sin(a), 0, 1, 0, -std::sin(a), 0, std::cos(a)};
        const double t[3] = {1.0, 0.2, -0.1};
        std::mt19937_64 g(3);
        std::uniform_real_distribution<double> ux(-3, 3), uy(-2, 2), uz(4, 12);
        std::vector<Point2> p, c;
        for (int i = 0; i < 60; ++i) {
            const double X[3] = {ux(g), uy(g), uz(g)};
            p.push_back({f * X[0] / X[2] + cx, f * X[1] / X[2] + cy});
            double Y[3];
            for (int r = 0; r < 3; ++r) Y[r] = R[3 * r] * X[0] + R[3 * r + 1] * X[1] + R[3 * r + 2] * X[2] + t[r];
            c.push_back({f * Y[0] / Y[2] + cx, f * Y[1] / Y[2] + cy});
        }
        // E = [t]_x R, F = Ki^T E Ki with Ki = K^-1
        const double tx[9] = {0, -t[2], t[1], t[2], 0, -t[0], -t[1], t[0], 0};
        double E[9] = {0}, Ki[9] = {1 / f, 0, -cx / f, 0, 1 / f, -cy / f, 0, 0, 1}, T[9] = {0}, F[9] = {0};
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q)
                for (int k = 0; k < 3; ++k) E[3 * r + q] += tx[3 * r + k] * R[3 * k + q];
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q)
                for (int k = 0; k < 3; ++k) T[3 * r + q] += E[3 * r + k] * Ki[3 * k + q];
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q)
                for (int k = 0; k < 3; ++k) F[3 * r + q] += Ki[3 * k + r] * T[3 * k + q];
        const auto normed = [](const double* m, double* out) {
            double n = 0.0, big = 0.0;
            for (int i = 0; i < 9; ++i) {
                n += m[i] * m[i];
                if (std::fabs(m[i]) > std::fabs(big)) big = m[i];
            }
            for (int i = 0; i < 9; ++i) out[i] = m[i] / std::sqrt(n) * (big < 0 ? -1.0 : 1.0);
        };
        double want[9];
        normed(F, want);
        for (size_t n : {size_t(8), size_t(60)}) {
            const std::vector<Point2> pp(p.begin(), p.begin() + n), cc(c.begin(), c.begin() + n);
            const auto got = FitFundamental(pp, cc);
            expect(got.has_value(), "8-point fit on exact data");
            if (got) {
                double g2[9], err = 0.0;
                normed(got->data(), g2);
                for (int i = 0; i < 9; ++i) err = std::max(err, std::fabs(g2[i] - want[i]));
                expect(err < 1e-9, "8-point fit recovers F (" + std::to_string(n) + " points)");
            }
        }
        expect(!FitFundamental(std::vector<Point2>(p.begin(), p.begin() + 7),
                               std::vector<Point2>(c.begin(), c.begin() + 7)),
               "fewer than 8 correspondences do not fit");
    }
    if (failures == 0) std::cout << "stage selftest ok\n";
    return failures == 0 ? 0 : 1;
}
