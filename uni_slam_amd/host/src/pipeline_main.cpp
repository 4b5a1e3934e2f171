// wslam_pipeline: runs the feature path through the pipeline engine
// (Compute + HipStages + passes) over a sequence of frames and writes, per
// frame, the feature counts and the MatchResult in its map order.
//
//   wslam_pipeline FRAMES.raw W H N CAP THRESHOLD RATIO OUT.txt
//                  [--f32] [--graph] [--ransac] [--advance BITS]
//
// FRAMES.raw holds N frames of W*H u8 pixels back to back.
//   --f32      publish each frame as the reference's FrameBW (f32 pixels
//              u8 / 255, data/src/provider_base.cpp:121-124) instead of FrameU8;
//   --graph    run both stages with HIP graph capture (one graph per buffer
//              slot combination, replayed);
//   --ransac   end the match stage with RansacHipPass (hypotheses scored on
//              the GPU) and print "ransac T total iterations inliers found"
//              per frame that has a reference set;
//   --advance  BITS has one character per frame: '1' sets the loader's
//              advance marker for that frame (the reference set moves to the
//              previous frame), '0' leaves it unset (the reference keyframe is
//              held, load_features_cpu.cpp:111-142).  Default: all '1'
//              (frame-to-frame matching, as before the map origin is set).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "compute/compute.hpp"
#include "wslam/feature_detect_hip.hpp"

using namespace wslam;

int main(int argc, char** argv) {
    if (argc < 9) {
        std::cerr << "usage: wslam_pipeline FRAMES.raw W H N CAP THRESHOLD RATIO OUT.txt "
                     "[--f32] [--graph] [--ransac] [--advance BITS]\n";
        return 2;
    }
    const std::string in = argv[1];
    const uint32_t W = std::stoul(argv[2]), H = std::stoul(argv[3]), N = std::stoul(argv[4]);
    bool f32 = false, graph = false, ransac = false;
    std::string advance(N, '1');
    for (int i = 9; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--f32") {
            f32 = true;
        } else if (a == "--graph") {
            graph = true;
        } else if (a == "--ransac") {
            ransac = true;
        } else if (a == "--advance" && i + 1 < argc) {
            advance = argv[++i];
            if (advance.size() != N) {
                std::cerr << "--advance needs one character per frame\n";
                return 2;
            }
        } else {
            std::cerr << "unknown option " << a << "\n";
            return 2;
        }
    }
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
    auto detect = CreateFeatureDetectStage(device, comp.getStorage(), buffers);
    auto match = CreateMatchStage(device, comp.getStorage(), buffers, ransac);
    detect->enableGraphCapture(graph);
    match->enableGraphCapture(graph);
    comp.addStage(std::move(detect));
    comp.addStage(std::move(match));
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
        store.erase(ResourceIdentifier::GetFrameName(0));
        if (f32) {
            data::FrameBW bw;
            bw.timestamp = t;
            bw.width = static_cast<uint16_t>(W);
            bw.height = static_cast<uint16_t>(H);
            bw.pixels.resize(frame.pixels.size());
            for (size_t i = 0; i < frame.pixels.size(); ++i)
                bw.pixels[i] = static_cast<float>(frame.pixels[i]) / 255.0f;
            store.set(ResourceIdentifier::GetFrameName(0), std::move(bw));
        } else {
            store.set(ResourceIdentifier::GetFrameName(0), std::move(frame));
        }
        if (advance[t] == '1') store.set(ResourceIdentifier::FeatureReferenceAdvanceName, true);
        if (auto e = comp.execute()) {
            std::cerr << "frame " << t << ": " << *e << "\n";
            return 1;
        }
        auto cur = store.getPtr<FeatureSet>(ResourceIdentifier::GetFeatureSetName(0));
        out << "frame " << t;
        for (const auto& lvl : **cur) out << " " << lvl.size();
        out << "\n";
        if (auto r = store.take<RansacResult>(ResourceIdentifier::RansacResultName)) {
            out << "ransac " << t << " " << r->stats.total_matches << " "
                << r->stats.iterations_run << " " << r->stats.inlier_count << " "
                << (r->stats.model_found ? 1 : 0) << "\n";
            size_t kept = 0;
            for (const auto& lvl : r->inliers) kept += lvl.size();
            if (kept != r->stats.inlier_count) {
                std::cerr << "frame " << t << ": RansacResult holds " << kept << " of "
                          << r->stats.inlier_count << " inliers\n";
                return 1;
            }
        }
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
