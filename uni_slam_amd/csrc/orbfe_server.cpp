// liborbfe served form: a continuous-batching request loop for many camera
// streams on one GPU (orbfe_server_* of include/orbfe.h).
//
// Frames arrive from many streams at their own times.  Per stream the
// server keeps a ring of slots resident in HBM (upload buffer, feature
// array, match table; frame k uses slot k mod slots), so a stream's
// previous-frame descriptors stay on the device.  Each launch admits the
// oldest pending frames (FIFO over all streams, at most frames_per_stream
// consecutive frames of one stream, each stream in its own order) as far as
// the HBM budget, max_batch and the free in-flight slots allow, and on that
// slot's HIP stream uploads them from pinned host memory and launches one
// batched detect (any mix of sizes) and one batched match (each frame
// against its stream's previous frame).  A batch's uploads overlap the
// kernels of the batch before it; its kernels wait on the device for the
// batches holding its streams' previous frames.  Results are the plain
// per-frame path's bit for bit: the batched kernels compute every frame
// exactly as a per-frame launch does.
//
// The reference runs one frame per stage submission from one host loop
// (wslam/src/main.cpp:64-135, feature_detect.cpp:43-86); this is the form
// that loop takes when many streams share the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <queue>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/orbfe.h"

namespace {

constexpr uint32_t kLevels = 6;

struct Request {
    uint32_t stream;
    uint32_t seq;
    uint64_t order;  // submission number: the FIFO admission key
    double t_arrival;
    const uint8_t* host;
    void* features_out;
    int32_t* match_out;
};

// A batch's completion event; frames hold it while a later frame of their
// stream may have to wait for it.
struct BatchEvent {
    hipEvent_t ev = nullptr;
};

struct Stream {
    orbfe_stream_spec spec{};
    uint32_t slots = 0;
    std::vector<uint8_t*> upload;   // per slot, W*H
    std::vector<uint8_t*> feats;    // per slot, feature array
    std::vector<int32_t*> match;    // per slot, levels*cap
    std::vector<orbfe_frame> frame_c;
    std::vector<orbfe_match_pair> pair_c;
    size_t feat_bytes = 0, match_bytes = 0;
    size_t scratch = 0;             // device scratch of one frame in a batch
    uint32_t next_seq = 0;
    std::deque<Request> pending;
    std::unordered_map<uint32_t, std::shared_ptr<BatchEvent>> events;  // seq -> its batch
};

struct InFlight {
    bool busy = false;
    std::shared_ptr<BatchEvent> ev;
    std::vector<Request> reqs;
    size_t bytes = 0;
};

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct orbfe_server {
    int device = 0;
    orbfe_server_config cfg{};
    std::string err;
    std::chrono::steady_clock::time_point t0;
    std::vector<Stream> streams;
    std::vector<orbfe_ctx*> ctxs;          // one per in-flight slot
    std::vector<hipStream_t> hstreams;     // the contexts' streams
    std::vector<InFlight> inflight;
    std::vector<hipEvent_t> event_pool;
    // streams with pending frames, keyed by their oldest frame's submission
    // number (min-heap): popping in key order is FIFO over all frames
    std::priority_queue<std::pair<uint64_t, uint32_t>, std::vector<std::pair<uint64_t, uint32_t>>,
                        std::greater<>>
        ready;
    uint64_t order = 0;
    size_t in_flight_bytes = 0;
    std::vector<orbfe_completion> done;    // retired, not yet handed out
    // counters
    uint64_t batches = 0;
    uint32_t max_batch_seen = 0;
    std::vector<orbfe_batch_entry> log;  // with cfg.log_batches
    // per-launch scratch (kept to avoid reallocation)
    std::vector<Request> chosen;
    std::vector<orbfe_frame> frames_buf;
    std::vector<orbfe_match_pair> pairs_buf;
};

namespace {

int sfail(orbfe_server* s, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    return code;
}

#define SRV_HIP(s, expr)                                                            \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return sfail((s), ORBFE_EHIP, "%s: %s", #expr, hipGetErrorString(e_));  \
    } while (0)

std::shared_ptr<BatchEvent> take_event(orbfe_server* s) {
    hipEvent_t ev = nullptr;
    if (!s->event_pool.empty()) {
        ev = s->event_pool.back();
        s->event_pool.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        return nullptr;
    }
    orbfe_server* owner = s;
    return std::shared_ptr<BatchEvent>(new BatchEvent{ev}, [owner](BatchEvent* b) {
        owner->event_pool.push_back(b->ev);
        delete b;
    });
}

// FIFO admission: the oldest pending frames, at most frames_per_stream per
// stream (a stream's frames stay in order), up to max_batch, within the HBM
// budget (a frame is always admitted when nothing else is in flight).
size_t admit(orbfe_server* s) {
    s->chosen.clear();
    size_t bytes = 0;
    std::vector<std::pair<uint32_t, uint32_t>> taken;  // stream, count
    std::vector<uint32_t> deferred;
    const auto count_of = [&](uint32_t st) -> uint32_t& {
        for (auto& t : taken)
            if (t.first == st) return t.second;
        taken.push_back({st, 0});
        return taken.back().second;
    };
    while (!s->ready.empty() && s->chosen.size() < s->cfg.max_batch) {
        const uint32_t si = s->ready.top().second;
        Stream& st = s->streams[si];
        if (s->in_flight_bytes + bytes + st.scratch > s->cfg.hbm_budget &&
            (!s->chosen.empty() || s->in_flight_bytes > 0))
            break;
        s->ready.pop();
        s->chosen.push_back(st.pending.front());
        st.pending.pop_front();
        bytes += st.scratch;
        uint32_t& c = count_of(si);
        ++c;
        if (!st.pending.empty()) {
            if (c < s->cfg.frames_per_stream)
                s->ready.push({st.pending.front().order, si});
            else
                deferred.push_back(si);
        }
    }
    for (uint32_t si : deferred) s->ready.push({s->streams[si].pending.front().order, si});
    return bytes;
}

// Admits pending frames into free in-flight slots and launches them.
int launch_batches(orbfe_server* s, uint32_t* launched) {
    for (size_t slot = 0; slot < s->inflight.size(); ++slot) {
        if (s->ready.empty()) break;
        InFlight& fl = s->inflight[slot];
        if (fl.busy) continue;
        const size_t bytes = admit(s);
        if (s->chosen.empty()) break;
        orbfe_ctx* ctx = s->ctxs[slot];
        const hipStream_t hs = s->hstreams[slot];
        auto ev = take_event(s);
        if (!ev) return sfail(s, ORBFE_EHIP, "hipEventCreate failed");
        s->frames_buf.clear();
        s->pairs_buf.clear();
        std::vector<hipEvent_t> waited;
        const auto wait = [&](const std::shared_ptr<BatchEvent>& e) -> hipError_t {
            if (!e || std::find(waited.begin(), waited.end(), e->ev) != waited.end())
                return hipSuccess;
            waited.push_back(e->ev);
            return hipStreamWaitEvent(hs, e->ev, 0);
        };
        const auto event_of = [](Stream& st, int64_t seq) -> std::shared_ptr<BatchEvent> {
            if (seq < 0) return nullptr;
            auto it = st.events.find(static_cast<uint32_t>(seq));
            return it == st.events.end() ? nullptr : it->second;
        };
        // Uploads: each after the batches that last used its stream slot
        // (frame k - slots wrote it; frame k - slots + 1 read its features).
        for (const Request& r : s->chosen) {
            Stream& st = s->streams[r.stream];
            const uint32_t p = r.seq % st.slots;
            SRV_HIP(s, wait(event_of(st, static_cast<int64_t>(r.seq) - st.slots)));
            SRV_HIP(s, wait(event_of(st, static_cast<int64_t>(r.seq) - st.slots + 1)));
            SRV_HIP(s, hipMemcpyAsync(st.upload[p], r.host,
                                      static_cast<size_t>(st.spec.width) * st.spec.height,
                                      hipMemcpyHostToDevice, hs));
            s->frames_buf.push_back(st.frame_c[p]);
            if (r.seq > 0) s->pairs_buf.push_back(st.pair_c[p]);
        }
        // The batches holding the admitted frames' previous frames (their
        // features are this batch's match input), unless in this batch.
        for (const Request& r : s->chosen) {
            if (r.seq == 0) continue;
            bool in_batch = false;
            for (const Request& q : s->chosen)
                in_batch |= q.stream == r.stream && q.seq + 1 == r.seq;
            if (!in_batch)
                SRV_HIP(s, wait(event_of(s->streams[r.stream], static_cast<int64_t>(r.seq) - 1)));
        }
        int rc = orbfe_detect_batch_async(ctx, s->frames_buf.data(),
                                          static_cast<uint32_t>(s->frames_buf.size()), hs);
        if (rc) return sfail(s, rc, "batched detect: %s", orbfe_last_error(ctx));
        if (!s->pairs_buf.empty()) {
            rc = orbfe_match_pipeline_batch_async(ctx, s->pairs_buf.data(),
                                                  static_cast<uint32_t>(s->pairs_buf.size()), hs);
            if (rc) return sfail(s, rc, "batched match: %s", orbfe_last_error(ctx));
        }
        // Results the caller asked for, in stream order after the kernels.
        for (const Request& r : s->chosen) {
            Stream& st = s->streams[r.stream];
            const uint32_t p = r.seq % st.slots;
            if (r.features_out)
                SRV_HIP(s, hipMemcpyAsync(r.features_out, st.feats[p], st.feat_bytes,
                                          hipMemcpyDeviceToHost, hs));
            if (r.match_out && r.seq > 0)
                SRV_HIP(s, hipMemcpyAsync(r.match_out, st.match[p], st.match_bytes,
                                          hipMemcpyDeviceToHost, hs));
        }
        SRV_HIP(s, hipEventRecord(ev->ev, hs));
        for (const Request& r : s->chosen) {
            Stream& st = s->streams[r.stream];
            st.events[r.seq] = ev;
            if (r.seq >= st.slots) st.events.erase(r.seq - st.slots);  // no later frame waits for it
        }
        if (s->cfg.log_batches)
            for (const Request& r : s->chosen)
                s->log.push_back({static_cast<uint32_t>(s->batches), r.stream, r.seq,
                                  static_cast<uint32_t>(r.order)});
        fl.busy = true;
        fl.ev = ev;
        fl.reqs.assign(s->chosen.begin(), s->chosen.end());
        fl.bytes = bytes;
        s->in_flight_bytes += bytes;
        s->max_batch_seen = std::max<uint32_t>(s->max_batch_seen,
                                               static_cast<uint32_t>(s->chosen.size()));
        ++s->batches;
        if (launched) ++*launched;
    }
    return ORBFE_OK;
}

// Retires finished batches: their frames complete now.
int retire(orbfe_server* s) {
    for (InFlight& fl : s->inflight) {
        if (!fl.busy) continue;
        const hipError_t q = hipEventQuery(fl.ev->ev);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return sfail(s, ORBFE_EHIP, "batch: %s", hipGetErrorString(q));
        const double t = now_s(s->t0);
        for (const Request& r : fl.reqs) s->done.push_back({r.stream, r.seq, r.t_arrival, t});
        s->in_flight_bytes -= fl.bytes;
        fl.busy = false;
        fl.ev.reset();
        fl.reqs.clear();
    }
    return ORBFE_OK;
}

bool idle(const orbfe_server* s) {
    if (!s->ready.empty()) return false;
    for (const InFlight& fl : s->inflight)
        if (fl.busy) return false;
    return true;
}

int submit(orbfe_server* s, uint32_t stream, const uint8_t* frame, double t_arrival,
           void* features_out, int32_t* match_out, uint32_t* seq_out) {
    if (stream >= s->streams.size())
        return sfail(s, ORBFE_EINVAL, "stream %u of %zu", stream, s->streams.size());
    if (!frame) return sfail(s, ORBFE_EINVAL, "null frame");
    Stream& st = s->streams[stream];
    const Request r{stream, st.next_seq++, s->order++, t_arrival, frame, features_out, match_out};
    st.pending.push_back(r);
    if (st.pending.size() == 1) s->ready.push({r.order, stream});
    if (seq_out) *seq_out = r.seq;
    return ORBFE_OK;
}

}  // namespace

extern "C" {

int orbfe_server_config_default(orbfe_server_config* cfg) {
    if (!cfg) return ORBFE_EINVAL;
    std::memset(cfg, 0, sizeof(*cfg));
    orbfe_config_default(&cfg->detect, 0, 0);
    cfg->hbm_budget = size_t(1) << 30;
    cfg->max_batch = 64;
    cfg->inflight = 2;
    cfg->frames_per_stream = 3;
    return ORBFE_OK;
}

const char* orbfe_server_last_error(const orbfe_server* s) {
    return s ? s->err.c_str() : "null server";
}

void orbfe_server_destroy(orbfe_server* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    for (hipStream_t h : s->hstreams) (void)hipStreamSynchronize(h);
    for (InFlight& fl : s->inflight) fl.ev.reset();
    for (Stream& st : s->streams) {
        st.events.clear();
        for (auto* p : st.upload) (void)hipFree(p);
        for (auto* p : st.feats) (void)hipFree(p);
        for (auto* p : st.match) (void)hipFree(p);
    }
    for (orbfe_ctx* c : s->ctxs) orbfe_destroy(c);
    for (hipStream_t h : s->hstreams) (void)hipStreamDestroy(h);
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    s->event_pool.clear();
    delete s;
}

int orbfe_server_create(const orbfe_server_config* cfg, const orbfe_stream_spec* specs,
                        uint32_t n_streams, int device, orbfe_server** out) {
    if (!cfg || !specs || !out || n_streams == 0) return ORBFE_EINVAL;
    *out = nullptr;
    auto s = std::make_unique<orbfe_server>();
    s->device = device;
    s->cfg = *cfg;
    if (s->cfg.hbm_budget == 0) s->cfg.hbm_budget = size_t(1) << 30;
    if (s->cfg.max_batch == 0) s->cfg.max_batch = 64;
    if (s->cfg.inflight == 0) s->cfg.inflight = 2;
    if (s->cfg.frames_per_stream == 0) s->cfg.frames_per_stream = 3;
    if (s->cfg.inflight > 8 || s->cfg.frames_per_stream > 64) return ORBFE_ERANGE;
    s->t0 = std::chrono::steady_clock::now();
    orbfe_server* S = s.get();
    const auto bail = [&](int rc) {
        orbfe_server_destroy(s.release());
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return ORBFE_EHIP;
    // One context (own HIP stream, arena and staging ring) per in-flight slot.
    orbfe_config c0 = cfg->detect;
    c0.frame_width = specs[0].width;
    c0.frame_height = specs[0].height;
    c0.features_per_level = specs[0].cap;
    for (uint32_t i = 0; i < s->cfg.inflight; ++i) {
        orbfe_ctx* ctx = nullptr;
        const int rc = orbfe_create(&c0, device, &ctx);
        if (rc) return bail(rc);
        s->ctxs.push_back(ctx);
        hipStream_t h = nullptr;
        if (hipStreamCreate(&h) != hipSuccess) return bail(ORBFE_EHIP);
        s->hstreams.push_back(h);
    }
    s->inflight.resize(s->cfg.inflight);
    const uint32_t slots = 2 * s->cfg.frames_per_stream + 2;
    s->streams.resize(n_streams);
    for (uint32_t i = 0; i < n_streams; ++i) {
        Stream& st = s->streams[i];
        st.spec = specs[i];
        st.slots = slots;
        size_t scratch = 0;
        int rc = orbfe_frame_scratch_bytes(s->ctxs[0], specs[i].width, specs[i].height,
                                           specs[i].cap, &scratch);
        if (rc) {
            S->err = orbfe_last_error(s->ctxs[0]);
            return bail(rc);
        }
        st.scratch = scratch;
        const uint32_t cap = specs[i].cap ? specs[i].cap
                                          : static_cast<uint32_t>(uint64_t(specs[i].width) *
                                                                  specs[i].height / 250);
        st.feat_bytes = static_cast<size_t>(kLevels) * (8 + 48ull * cap);
        st.match_bytes = static_cast<size_t>(kLevels) * cap * 4;
        for (uint32_t p = 0; p < slots; ++p) {
            void *u = nullptr, *f = nullptr, *m = nullptr;
            if (hipMalloc(&u, static_cast<size_t>(specs[i].width) * specs[i].height) ||
                hipMalloc(&f, st.feat_bytes) || hipMalloc(&m, st.match_bytes))
                return bail(sfail(S, ORBFE_ENOMEM, "stream %u buffers", i));
            if (hipMemset(u, 0, static_cast<size_t>(specs[i].width) * specs[i].height) ||
                hipMemset(f, 0, st.feat_bytes) || hipMemset(m, 0xFF, st.match_bytes))
                return bail(sfail(S, ORBFE_EHIP, "stream %u buffers", i));
            st.upload.push_back(static_cast<uint8_t*>(u));
            st.feats.push_back(static_cast<uint8_t*>(f));
            st.match.push_back(static_cast<int32_t*>(m));
        }
        for (uint32_t p = 0; p < slots; ++p) {
            orbfe_frame fr{};
            fr.image = st.upload[p];
            fr.pitch = specs[i].width;
            fr.width = specs[i].width;
            fr.height = specs[i].height;
            fr.cap = cap;
            fr.pyramid = nullptr;  // not exported when serving
            fr.features = st.feats[p];
            st.frame_c.push_back(fr);
            orbfe_match_pair mp{};
            mp.prev_features = st.feats[(p + slots - 1) % slots];
            mp.curr_features = st.feats[p];
            mp.prev_match = st.match[p];
            mp.width = specs[i].width;
            mp.height = specs[i].height;
            mp.cap = cap;
            st.pair_c.push_back(mp);
        }
    }
    // Warm-up: every plan in every context, every arena and staging slot at
    // the largest batch admission can form (frames_per_stream frames of every
    // stream), so serving allocates nothing.  Stream state is then reset.
    {
        std::vector<orbfe_frame> frames;
        std::vector<orbfe_match_pair> pairs;
        const uint32_t k = std::min(s->cfg.frames_per_stream, slots - 1);
        for (Stream& st : s->streams)
            for (uint32_t p = 1; p <= k; ++p) {
                frames.push_back(st.frame_c[p]);
                pairs.push_back(st.pair_c[p]);
            }
        for (uint32_t c = 0; c < s->cfg.inflight; ++c) {
            for (int rep = 0; rep < 4; ++rep) {  // each slot of the staging ring
                int rc = orbfe_detect_batch_async(s->ctxs[c], frames.data(),
                                                  static_cast<uint32_t>(frames.size()),
                                                  s->hstreams[c]);
                if (!rc)
                    rc = orbfe_match_pipeline_batch_async(s->ctxs[c], pairs.data(),
                                                          static_cast<uint32_t>(pairs.size()),
                                                          s->hstreams[c]);
                if (rc) {
                    S->err = orbfe_last_error(s->ctxs[c]);
                    return bail(rc);
                }
            }
        }
        if (hipDeviceSynchronize() != hipSuccess) return bail(ORBFE_EHIP);
        for (Stream& st : s->streams)
            for (uint32_t p = 0; p < slots; ++p)
                if (hipMemset(st.feats[p], 0, st.feat_bytes) ||
                    hipMemset(st.match[p], 0xFF, st.match_bytes))
                    return bail(ORBFE_EHIP);
        if (hipDeviceSynchronize() != hipSuccess) return bail(ORBFE_EHIP);
        for (int e = 0; e < 64; ++e) {
            hipEvent_t ev = nullptr;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
                return bail(ORBFE_EHIP);
            s->event_pool.push_back(ev);
        }
    }
    *out = s.release();
    return ORBFE_OK;
}

double orbfe_server_clock(const orbfe_server* s) { return s ? now_s(s->t0) : 0.0; }

int orbfe_server_submit(orbfe_server* s, uint32_t stream, const uint8_t* frame, double t_arrival,
                        void* features_out, int32_t* match_out, uint32_t* seq) {
    if (!s) return ORBFE_EINVAL;
    return submit(s, stream, frame, t_arrival, features_out, match_out, seq);
}

int orbfe_server_pump(orbfe_server* s, orbfe_completion* done, uint32_t max_done,
                      uint32_t* n_done, uint32_t* launched) {
    if (!s) return ORBFE_EINVAL;
    if (launched) *launched = 0;
    int rc = retire(s);
    if (!rc) rc = launch_batches(s, launched);
    if (rc) return rc;
    const uint32_t n = static_cast<uint32_t>(std::min<size_t>(s->done.size(), max_done));
    if (n && done) std::memcpy(done, s->done.data(), n * sizeof(orbfe_completion));
    s->done.erase(s->done.begin(), s->done.begin() + (done ? n : 0));
    if (n_done) *n_done = done ? n : 0;
    return ORBFE_OK;
}

int orbfe_server_idle(const orbfe_server* s) { return s && idle(s) && s->done.empty() ? 1 : 0; }

int orbfe_server_stats(const orbfe_server* s, uint64_t* batches, uint32_t* max_batch) {
    if (!s) return ORBFE_EINVAL;
    if (batches) *batches = s->batches;
    if (max_batch) *max_batch = s->max_batch_seen;
    return ORBFE_OK;
}

int orbfe_server_batch_log(const orbfe_server* s, orbfe_batch_entry* out, size_t max, size_t* n) {
    if (!s) return ORBFE_EINVAL;
    if (out) std::memcpy(out, s->log.data(), std::min(max, s->log.size()) * sizeof(*out));
    if (n) *n = s->log.size();
    return ORBFE_OK;
}

int orbfe_server_run(orbfe_server* s, const orbfe_arrival* arrivals, uint32_t n,
                     const uint8_t* const* pool, uint32_t pool_per_stream, double time_scale,
                     orbfe_completion* done, double* wall_s) {
    if (!s) return ORBFE_EINVAL;
    if ((n && (!arrivals || !pool || !done)) || pool_per_stream == 0)
        return sfail(s, ORBFE_EINVAL, "null arrivals, pool or output");
    for (uint32_t i = 1; i < n; ++i)
        if (arrivals[i].t < arrivals[i - 1].t)
            return sfail(s, ORBFE_EINVAL, "arrivals not sorted by time");
    for (uint32_t i = 0; i < n; ++i)
        if (arrivals[i].stream >= s->streams.size())
            return sfail(s, ORBFE_EINVAL, "arrival %u: stream %u", i, arrivals[i].stream);
    s->done.clear();
    const double t0 = now_s(s->t0);
    uint32_t i = 0, n_done = 0;
    while (i < n || !idle(s)) {
        const double now = now_s(s->t0) - t0;
        while (i < n && arrivals[i].t * time_scale <= now) {
            const orbfe_arrival& a = arrivals[i];
            const uint8_t* f = pool[a.stream * pool_per_stream + a.seq % pool_per_stream];
            int rc = submit(s, a.stream, f, t0 + a.t * time_scale, nullptr, nullptr, nullptr);
            if (rc) return rc;
            ++i;
        }
        uint32_t launched = 0;
        int rc = retire(s);
        if (!rc) rc = launch_batches(s, &launched);
        if (rc) return rc;
        for (const orbfe_completion& c : s->done) {
            if (n_done < n) done[n_done] = c;
            ++n_done;
        }
        s->done.clear();
        if (!launched && (i < n || !idle(s))) {
            // nothing to launch: wait for the next arrival or a batch, briefly
            const double next = i < n ? arrivals[i].t * time_scale - (now_s(s->t0) - t0) : 0.0;
            if (next > 50e-6)
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            else
                std::this_thread::yield();
        }
    }
    if (wall_s) *wall_s = now_s(s->t0) - t0;
    if (n_done != n) return sfail(s, ORBFE_EHIP, "%u of %u frames completed", n_done, n);
    return ORBFE_OK;
}

}  // extern "C"
