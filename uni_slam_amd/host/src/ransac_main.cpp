// wslam_ransac: runs the RANSAC model search of the pose-estimate stage
// (RunRansac, hypotheses scored on the GPU) over sets of correspondences and
// prints what the reference's RansacResult would hold.
//
//   wslam_ransac SETS.bin ROUND [MAX_ITERATIONS]
//
// SETS.bin: uint32 number of sets, then per set uint32 n and n prev (x, y)
// then n curr (x, y) LOD-0 coordinates as doubles.  The sets are processed
// in order with one generator (seeded as RansacCPU's), as consecutive
// frames are.  ROUND is the number of hypotheses scored per launch.  Per
// set it prints
//   set S total T iterations I inliers K found B
//   idx i0 i1 ...        (inlier indices, ascending)
//   F f00 ... f22        (%.17g)
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "wslam/ransac_hip.hpp"

using namespace wslam;

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: wslam_ransac SETS.bin ROUND [MAX_ITERATIONS]\n";
        return 2;
    }
    std::ifstream f(argv[1], std::ios::binary);
    uint32_t sets = 0;
    f.read(reinterpret_cast<char*>(&sets), 4);
    RansacOpts opts;
    opts.round = std::stoul(argv[2]);
    if (argc > 3) opts.max_iterations = std::stoul(argv[3]);
    if (hipSetDevice(0) != hipSuccess) return 1;
    orbfe_config cfg;
    orbfe_config_default(&cfg, 64, 48);
    orbfe_ctx* ctx = nullptr;
    if (orbfe_create(&cfg, 0, &ctx) != ORBFE_OK) {
        std::cerr << "orbfe_create failed\n";
        return 1;
    }
    hipStream_t stream = nullptr;
    if (hipStreamCreate(&stream) != hipSuccess) return 1;
    int status = 0;
    {
        GpuRansacScorer scorer(ctx, stream);
        std::mt19937_64 rng(opts.rng_seed);
        for (uint32_t s = 0; s < sets && status == 0; ++s) {
            uint32_t n = 0;
            f.read(reinterpret_cast<char*>(&n), 4);
            std::vector<Point2> prev(n), curr(n);
            f.read(reinterpret_cast<char*>(prev.data()), 16 * n);
            f.read(reinterpret_cast<char*>(curr.data()), 16 * n);
            if (!f) {
                std::cerr << "short read in set " << s << "\n";
                status = 1;
                break;
            }
            if (auto e = scorer.upload(prev, curr, opts.inlier_threshold)) {
                std::cerr << *e << "\n";
                status = 1;
                break;
            }
            RansacOutcome out;
            const RansacScorer score = [&](const std::vector<Mat3>& F, std::vector<uint32_t>& c,
                                           std::vector<uint8_t>* mask) {
                return scorer.score(F, c, mask);
            };
            if (auto e = RunRansac(prev, curr, opts, rng, score, out)) {
                std::cerr << "set " << s << ": " << *e << "\n";
                status = 1;
                break;
            }
            std::printf("set %u total %zu iterations %zu inliers %zu found %d\nidx", s,
                        out.stats.total_matches, out.stats.iterations_run,
                        out.stats.inlier_count, out.stats.model_found ? 1 : 0);
            for (size_t i : out.inliers) std::printf(" %zu", i);
            std::printf("\nF");
            for (double v : out.F) std::printf(" %.17g", v);
            std::printf("\n");
        }
    }
    (void)hipStreamDestroy(stream);
    orbfe_destroy(ctx);
    return status;
}
