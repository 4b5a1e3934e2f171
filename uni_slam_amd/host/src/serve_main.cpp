// wslam_serve: drives liborbfe's served form from C++ and checks it against
// the plain per-frame path.
//
//   wslam_serve N THRESHOLD RATIO W H CAP FRAMES.raw [W H CAP FRAMES.raw ...]
//
// Each (W, H, CAP, FRAMES.raw) is one camera stream whose file holds N u8
// frames of W*H bytes.  All frames of all streams are submitted at once,
// interleaved round-robin over the streams, and the server is pumped until
// it is idle; every frame's feature array and match table (against the
// stream's previous frame) must equal what orbfe_detect and
// orbfe_match_pipeline give for that frame on its own.  Prints one summary
// line; exit status 0 when every frame agrees.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/orbfe.h"

namespace {

struct StreamIn {
    orbfe_stream_spec spec{};
    std::vector<uint8_t*> frames;          // pinned, N frames
    std::vector<uint8_t*> feats_out;       // pinned, per frame
    std::vector<int32_t*> match_out;       // pinned, per frame
    size_t feat_bytes = 0, match_bytes = 0;
};

template <typename T>
T* pinned(size_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return static_cast<T*>(p);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 8 || (argc - 4) % 4 != 0) {
        std::cerr << "usage: wslam_serve N THRESHOLD RATIO W H CAP FRAMES.raw [W H CAP FRAMES.raw ...]\n";
        return 2;
    }
    const uint32_t N = std::stoul(argv[1]);
    const float thr = std::stof(argv[2]);
    const float ratio = std::stof(argv[3]);
    std::vector<StreamIn> streams;
    for (int a = 4; a < argc; a += 4) {
        StreamIn st;
        st.spec.width = std::stoul(argv[a]);
        st.spec.height = std::stoul(argv[a + 1]);
        st.spec.cap = std::stoul(argv[a + 2]);
        const size_t fb = static_cast<size_t>(st.spec.width) * st.spec.height;
        st.feat_bytes = 6 * (8 + 48ull * st.spec.cap);
        st.match_bytes = 6ull * st.spec.cap * 4;
        std::ifstream f(argv[a + 3], std::ios::binary);
        for (uint32_t k = 0; k < N; ++k) {
            uint8_t* p = pinned<uint8_t>(fb);
            uint8_t* fo = pinned<uint8_t>(st.feat_bytes);
            int32_t* mo = pinned<int32_t>(st.match_bytes);
            if (!p || !fo || !mo) {
                std::cerr << "hipHostMalloc failed\n";
                return 1;
            }
            f.read(reinterpret_cast<char*>(p), fb);
            if (!f) {
                std::cerr << "short read in " << argv[a + 3] << "\n";
                return 1;
            }
            st.frames.push_back(p);
            st.feats_out.push_back(fo);
            st.match_out.push_back(mo);
        }
        streams.push_back(std::move(st));
    }
    orbfe_server_config cfg;
    orbfe_server_config_default(&cfg);
    cfg.detect.fast_threshold = thr;
    cfg.detect.lowe_ratio = ratio;
    cfg.max_batch = 8;  // several batches, streams mixed inside each
    std::vector<orbfe_stream_spec> specs;
    for (auto& st : streams) specs.push_back(st.spec);
    orbfe_server* srv = nullptr;
    int rc = orbfe_server_create(&cfg, specs.data(), static_cast<uint32_t>(specs.size()), 0, &srv);
    if (rc) {
        std::cerr << "orbfe_server_create: " << rc << " " << orbfe_server_last_error(srv) << "\n";
        return 1;
    }
    for (uint32_t k = 0; k < N; ++k)
        for (uint32_t s = 0; s < streams.size(); ++s) {
            uint32_t seq = 0;
            rc = orbfe_server_submit(srv, s, streams[s].frames[k], orbfe_server_clock(srv),
                                     streams[s].feats_out[k], streams[s].match_out[k], &seq);
            if (rc || seq != k) {
                std::cerr << "submit: " << orbfe_server_last_error(srv) << "\n";
                return 1;
            }
        }
    std::vector<orbfe_completion> done(64);
    size_t completed = 0;
    while (!orbfe_server_idle(srv)) {
        uint32_t n = 0, launched = 0;
        rc = orbfe_server_pump(srv, done.data(), static_cast<uint32_t>(done.size()), &n, &launched);
        if (rc) {
            std::cerr << "pump: " << orbfe_server_last_error(srv) << "\n";
            return 1;
        }
        completed += n;
        if (!launched && !n) std::this_thread::yield();
    }
    uint64_t batches = 0;
    uint32_t max_batch = 0;
    orbfe_server_stats(srv, &batches, &max_batch);
    orbfe_server_destroy(srv);
    if (completed != N * streams.size()) {
        std::cerr << completed << " of " << N * streams.size() << " frames completed\n";
        return 1;
    }
    // The plain per-frame path on every frame.
    size_t mismatches = 0;
    for (uint32_t s = 0; s < streams.size(); ++s) {
        const StreamIn& st = streams[s];
        orbfe_config c;
        orbfe_config_default(&c, st.spec.width, st.spec.height);
        c.features_per_level = st.spec.cap;
        c.fast_threshold = thr;
        c.lowe_ratio = ratio;
        orbfe_ctx* ctx = nullptr;
        if (orbfe_create(&c, 0, &ctx)) {
            std::cerr << "orbfe_create failed\n";
            return 1;
        }
        std::vector<uint8_t> prev(st.feat_bytes), cur(st.feat_bytes);
        std::vector<int32_t> pm(st.match_bytes / 4);
        for (uint32_t k = 0; k < N; ++k) {
            if (orbfe_detect(ctx, st.frames[k], st.spec.width, nullptr, cur.data())) return 1;
            if (std::memcmp(cur.data(), st.feats_out[k], st.feat_bytes) != 0) {
                std::cerr << "stream " << s << " frame " << k << ": features differ\n";
                ++mismatches;
            }
            if (k > 0) {
                if (orbfe_match_pipeline(ctx, prev.data(), cur.data(), pm.data(), nullptr))
                    return 1;
                if (std::memcmp(pm.data(), st.match_out[k], st.match_bytes) != 0) {
                    std::cerr << "stream " << s << " frame " << k << ": matches differ\n";
                    ++mismatches;
                }
            }
            prev.swap(cur);
        }
        orbfe_destroy(ctx);
    }
    std::cout << "served " << streams.size() << " streams x " << N << " frames in " << batches
              << " batches (largest " << max_batch << "): "
              << (mismatches ? "MISMATCH" : "every frame equals the plain path") << "\n";
    for (auto& st : streams) {
        for (auto* p : st.frames) (void)hipHostFree(p);
        for (auto* p : st.feats_out) (void)hipHostFree(p);
        for (auto* p : st.match_out) (void)hipHostFree(p);
    }
    return mismatches ? 1 : 0;
}
