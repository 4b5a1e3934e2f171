// wslam_pipeline: runs the feature path through the pipeline engine
// (Compute + HipStages + passes) over a sequence of frames and writes, per
// frame, the feature counts and the MatchResult in its map order.
//
//   wslam_pipeline FRAMES.raw W H N CAP THRESHOLD RATIO OUT.txt
//
// FRAMES.raw holds N frames of W*H u8 pixels back to back.  Every frame asks
// the loader to advance the reference set (frame-to-frame matching, as the
// reference does before its map origin is set).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "compute/compute.hpp"
#include "wslam/feature_detect_hip.hpp"

using namespace wslam;

int main(int argc, char** argv) {
    if (argc != 9) {
        std::cerr << "usage: wslam_pipeline FRAMES.raw W H N CAP THRESHOLD RATIO OUT.txt\n";
        return 2;
    }
    const std::string in = argv[1];
    const uint32_t W = std::stoul(argv[2]), H = std::stoul(argv[3]), N = std::stoul(argv[4]);
    FeatureDetectOpts opts;
    opts.frame_width = W;
    opts.frame_height = H;
    opts.features_per_level = std::stoul(argv[5]);
    opts.fast_threshold = std::stof(argv[6]);
    opts.lowe_ratio = std::stof(argv[7]);
    std::ifstream f(in, std::ios::binary);
    if (!f) {
        std::cerr << "cannot open " << in << "\n";
        return 1;
    }
    std::string err;
    auto device = compute::Device::create(0, &err);
    if (!device) {
        std::cerr << "device: " << err << "\n";
        return 1;
    }
    compute::Compute comp;
    auto buffers = std::make_shared<FeatureBuffers>(device, opts);
    comp.addStage(CreateFeatureDetectStage(device, comp.getStorage(), buffers));
    comp.addStage(CreateMatchStage(device, comp.getStorage(), buffers));
    if (auto e = comp.initialize()) {
        std::cerr << "initialize: " << *e << "\n";
        return 1;
    }
    std::ofstream out(argv[8]);
    auto& store = comp.getStorage();
    for (uint32_t t = 0; t < N; ++t) {
        FrameU8 frame;
        frame.width = W;
        frame.height = H;
        frame.timestamp = t;
        frame.pixels.resize(static_cast<size_t>(W) * H);
        f.read(reinterpret_cast<char*>(frame.pixels.data()), frame.pixels.size());
        if (!f) {
            std::cerr << "short read at frame " << t << "\n";
            return 1;
        }
        store.set(ResourceIdentifier::GetFrameName(0), std::move(frame));
        store.set(ResourceIdentifier::FeatureReferenceAdvanceName, true);
        if (auto e = comp.execute()) {
            std::cerr << "frame " << t << ": " << *e << "\n";
            return 1;
        }
        auto cur = store.getPtr<FeatureSet>(ResourceIdentifier::GetFeatureSetName(0));
        out << "frame " << t;
        for (const auto& lvl : **cur) out << " " << lvl.size();
        out << "\n";
        if (auto m = store.take<MatchResult>(ResourceIdentifier::MatchedFeaturesName)) {
            for (uint32_t l = 0; l < kLevelsOfDetail; ++l)
                for (const auto& [b, a] : (*m)[l])
                    out << "m " << t << " " << l << " " << b.x << " " << b.y << " " << b.lod << " "
                        << a.x << " " << a.y << "\n";
        }
    }
    if (auto errs = comp.finalize()) {
        for (auto& e : *errs) std::cerr << e << "\n";
        return 1;
    }
    return 0;
}
