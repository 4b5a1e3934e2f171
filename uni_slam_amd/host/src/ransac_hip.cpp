// RANSAC with GPU hypothesis scoring (see ransac_hip.hpp).
#include "wslam/ransac_hip.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numbers>
#include <unordered_set>

#include "wslam/common.hpp"

namespace wslam {

namespace {

// One-sided Jacobi SVD (Hestenes): orthogonalises the n columns of the m x n
// row-major matrix `a` (m may be below n) by plane rotations, accumulating
// them in V (n x n, row-major).  On return the columns of `a` are
// sigma_j u_j; sigma[j] = their norms.  A pair of columns counts as
// orthogonal once |a_p . a_q| <= 4 m eps |a_p| |a_q| (the rounding level of
// the dot products); returns false if some pair is still rotated after 60
// sweeps.
bool jacobi_svd(std::vector<double>& a, size_t m, size_t n, std::vector<double>& v,
                std::vector<double>& sigma) {
    v.assign(n * n, 0.0);
    for (size_t i = 0; i < n; ++i) v[i * n + i] = 1.0;
    const double kEps = 4.0 * static_cast<double>(std::max<size_t>(m, 1)) *
                        std::numeric_limits<double>::epsilon();
    // Columns whose squared norm falls below this share of the matrix's are
    // null directions (m < n leaves at least one): they are not rotated.
    double frob = 0.0;
    for (double x : a) frob += x * x;
    const double tiny = frob * std::numeric_limits<double>::epsilon() *
                        std::numeric_limits<double>::epsilon();
    bool converged = false;
    for (int sweep = 0; sweep < 60 && !converged; ++sweep) {
        converged = true;
        for (size_t p = 0; p + 1 < n; ++p) {
            for (size_t q = p + 1; q < n; ++q) {
                double alpha = 0.0, beta = 0.0, gamma = 0.0;
                for (size_t r = 0; r < m; ++r) {
                    const double ap = a[r * n + p], aq = a[r * n + q];
                    alpha += ap * ap;
                    beta += aq * aq;
                    gamma += ap * aq;
                }
                if (alpha <= tiny || beta <= tiny || gamma == 0.0 ||
                    std::fabs(gamma) <= kEps * std::sqrt(alpha * beta))
                    continue;
                converged = false;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) /
                                 (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (size_t r = 0; r < m; ++r) {
                    const double ap = a[r * n + p], aq = a[r * n + q];
                    a[r * n + p] = c * ap - s * aq;
                    a[r * n + q] = s * ap + c * aq;
                }
                for (size_t r = 0; r < n; ++r) {
                    const double vp = v[r * n + p], vq = v[r * n + q];
                    v[r * n + p] = c * vp - s * vq;
                    v[r * n + q] = s * vp + c * vq;
                }
            }
        }
    }
    sigma.assign(n, 0.0);
    for (size_t j = 0; j < n; ++j) {
        double s2 = 0.0;
        for (size_t r = 0; r < m; ++r) s2 += a[r * n + j] * a[r * n + j];
        sigma[j] = std::sqrt(s2);
    }
    return converged;
}

// Right singular vector of the smallest singular value (column of V).
bool smallest_right_vector(std::vector<double> a, size_t m, size_t n, std::vector<double>& out) {
    std::vector<double> v, sigma;
    if (!jacobi_svd(a, m, n, v, sigma)) return false;
    size_t jmin = 0;
    for (size_t j = 1; j < n; ++j)
        if (sigma[j] < sigma[jmin]) jmin = j;
    out.resize(n);
    for (size_t r = 0; r < n; ++r) out[r] = v[r * n + jmin];
    return true;
}

// Hartley-Zisserman normalising similarity (NormalisingTransform,
// ransac_cpu.cpp:78-109): centroid to the origin, mean distance to sqrt(2).
Mat3 normalising_transform(const std::vector<Point2>& pts) {
    double cx = 0.0, cy = 0.0;
    for (const auto& p : pts) {
        cx += p.x;
        cy += p.y;
    }
    cx /= static_cast<double>(pts.size());
    cy /= static_cast<double>(pts.size());
    double mean = 0.0;
    for (const auto& p : pts) {
        const double dx = p.x - cx, dy = p.y - cy;
        mean += std::sqrt(dx * dx + dy * dy);
    }
    mean /= static_cast<double>(pts.size());
    const double s = mean > 1e-12 ? std::numbers::sqrt2 / mean : 1.0;
    return {s, 0.0, -s * cx, 0.0, s, -s * cy, 0.0, 0.0, 1.0};
}

Point2 apply(const Mat3& t, const Point2& p) {
    return {t[0] * p.x + t[1] * p.y + t[2], t[3] * p.x + t[4] * p.y + t[5]};
}

}  // namespace

std::optional<Mat3> FitFundamental(const std::vector<Point2>& prev,
                                   const std::vector<Point2>& curr) {
    if (prev.size() != curr.size() || prev.size() < 8) return std::nullopt;
    const Mat3 tp = normalising_transform(prev), tc = normalising_transform(curr);
    const size_t m = prev.size();
    std::vector<double> a(m * 9);
    for (size_t i = 0; i < m; ++i) {
        const Point2 p = apply(tp, prev[i]), q = apply(tc, curr[i]);
        double* r = &a[i * 9];
        r[0] = q.x * p.x;
        r[1] = q.x * p.y;
        r[2] = q.x;
        r[3] = q.y * p.x;
        r[4] = q.y * p.y;
        r[5] = q.y;
        r[6] = p.x;
        r[7] = p.y;
        r[8] = 1.0;
    }
    std::vector<double> f;
    if (!smallest_right_vector(std::move(a), m, 9, f)) return std::nullopt;
    // Rank 2: F - (F v3) v3^T removes the smallest singular direction.
    std::vector<double> fn(f.begin(), f.end()), v3;
    if (!smallest_right_vector(fn, 3, 3, v3)) return std::nullopt;
    Mat3 f2;
    for (int r = 0; r < 3; ++r) {
        const double fv = fn[r * 3 + 0] * v3[0] + fn[r * 3 + 1] * v3[1] + fn[r * 3 + 2] * v3[2];
        for (int c = 0; c < 3; ++c) f2[r * 3 + c] = fn[r * 3 + c] - fv * v3[c];
    }
    // Denormalise: t_curr^T F2 t_prev.
    Mat3 tmp{}, out{};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) tmp[r * 3 + c] += f2[r * 3 + k] * tp[k * 3 + c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) out[r * 3 + c] += tc[k * 3 + r] * tmp[k * 3 + c];
    for (double x : out)
        if (!std::isfinite(x)) return std::nullopt;
    return out;
}

std::array<size_t, 8> SampleEightIndices(size_t n, std::mt19937_64& rng) {
    std::array<size_t, 8> out{};
    std::unordered_set<size_t> picked;
    std::uniform_int_distribution<size_t> dist(0, n - 1);
    size_t filled = 0;
    while (filled < out.size()) {
        const size_t s = dist(rng);
        if (picked.insert(s).second) out[filled++] = s;
    }
    return out;
}

std::optional<std::string> RunRansac(const std::vector<Point2>& prev,
                                     const std::vector<Point2>& curr, const RansacOpts& opts,
                                     std::mt19937_64& rng, const RansacScorer& score,
                                     RansacOutcome& out) {
    out = RansacOutcome{};
    const size_t n = prev.size();
    out.stats.total_matches = n;
    if (n < opts.min_matches || n < 8) return std::nullopt;
    size_t best = 0;
    Mat3 best_f{};
    uint32_t iter = 0, needed = opts.max_iterations;
    const uint32_t round = std::max<uint32_t>(1, opts.round);
    std::vector<Mat3> fits;
    std::vector<int> slot;  // per drawn sample: index in fits, or -1 (degenerate)
    std::vector<uint32_t> counts;
    std::vector<Point2> sp(8), sc(8);
    while (iter < needed) {
        const std::mt19937_64 saved = rng;
        const uint32_t draws = std::min(round, needed - iter);
        fits.clear();
        slot.clear();
        for (uint32_t j = 0; j < draws; ++j) {
            const auto s = SampleEightIndices(n, rng);
            for (size_t i = 0; i < 8; ++i) {
                sp[i] = prev[s[i]];
                sc[i] = curr[s[i]];
            }
            const auto f = FitFundamental(sp, sc);
            slot.push_back(f ? static_cast<int>(fits.size()) : -1);
            if (f) fits.push_back(*f);
        }
        counts.assign(fits.size(), 0);
        if (!fits.empty())
            if (auto e = score(fits, counts, nullptr)) return e;
        // The reference's iterations over this round's draws.
        uint32_t j = 0;
        for (; j < draws && iter < needed; ++j, ++iter) {
            if (slot[j] < 0) continue;  // degenerate sample
            const size_t c = counts[static_cast<size_t>(slot[j])];
            if (c <= best) continue;
            best = c;
            best_f = fits[static_cast<size_t>(slot[j])];
            const double ratio = static_cast<double>(best) / static_cast<double>(n);
            const double p_all = std::pow(ratio, 8.0);
            if (p_all >= 1.0 - 1e-12) {
                needed = iter + 1;
            } else if (p_all > 0.0) {
                const double bound = std::log(1.0 - opts.confidence) / std::log(1.0 - p_all);
                needed = static_cast<uint32_t>(
                    std::min<double>(opts.max_iterations, std::ceil(bound)));
            }
        }
        if (j < draws) {  // stopped inside the round: the generator as after j draws
            rng = saved;
            for (uint32_t t = 0; t < j; ++t) (void)SampleEightIndices(n, rng);
        }
    }
    out.stats.iterations_run = iter;
    if (best < opts.min_matches) return std::nullopt;
    // The best model's inliers, the refit on them and its classification.
    std::vector<uint8_t> mask;
    const auto inliers_of = [&](const Mat3& f, std::vector<size_t>& idx) -> std::optional<std::string> {
        std::vector<uint32_t> c(1);
        if (auto e = score({f}, c, &mask)) return e;
        idx.clear();
        for (size_t i = 0; i < n; ++i)
            if (mask[i]) idx.push_back(i);
        if (idx.size() != c[0]) return std::string("inlier mask and count disagree");
        return std::nullopt;
    };
    std::vector<size_t> best_in;
    if (auto e = inliers_of(best_f, best_in)) return e;
    std::vector<Point2> rp, rc;
    rp.reserve(best_in.size());
    rc.reserve(best_in.size());
    for (size_t i : best_in) {
        rp.push_back(prev[i]);
        rc.push_back(curr[i]);
    }
    if (auto refined = FitFundamental(rp, rc)) {
        std::vector<size_t> ref_in;
        if (auto e = inliers_of(*refined, ref_in)) return e;
        if (ref_in.size() >= best_in.size()) {
            best_f = *refined;
            best_in = std::move(ref_in);
        }
    }
    out.F = best_f;
    out.inliers = std::move(best_in);
    out.stats.inlier_count = out.inliers.size();
    out.stats.model_found = true;
    return std::nullopt;
}

void FlattenMatchResult(const MatchResult& m, std::vector<Point2>& prev,
                        std::vector<Point2>& curr,
                        std::vector<std::pair<Feature, Feature>>& feats) {
    const auto lod0 = [](const Feature& f) {
        double s = 1.0;  // LodScale: the product of lod factors 1.2 (ransac_cpu.cpp:35-46)
        for (uint32_t i = 0; i < f.lod; ++i) s *= 1.2;
        return Point2{static_cast<double>(f.x) * s, static_cast<double>(f.y) * s};
    };
    prev.clear();
    curr.clear();
    feats.clear();
    for (const auto& level : m)
        for (const auto& [c, p] : level) {
            prev.push_back(lod0(p));
            curr.push_back(lod0(c));
            feats.emplace_back(c, p);
        }
}

// --- GpuRansacScorer -----------------------------------------------------------
GpuRansacScorer::~GpuRansacScorer() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : {static_cast<void*>(d_prev_), static_cast<void*>(d_curr_),
                    static_cast<void*>(d_F_), static_cast<void*>(d_counts_),
                    static_cast<void*>(d_mask_)})
        if (p) (void)hipFree(p);
    for (void* p : {static_cast<void*>(h_F_), static_cast<void*>(h_counts_),
                    static_cast<void*>(h_mask_)})
        if (p) (void)hipHostFree(p);
}

std::optional<std::string> GpuRansacScorer::reserve(size_t n, size_t k) {
    const auto fail = [](const char* what, hipError_t e) {
        return std::string(what) + ": " + hipGetErrorString(e);
    };
    hipError_t e;
    if (n > cap_n_) {
        for (void* p : {static_cast<void*>(d_prev_), static_cast<void*>(d_curr_),
                        static_cast<void*>(d_mask_)})
            if (p) (void)hipFree(p);
        if (h_mask_) (void)hipHostFree(h_mask_);
        d_prev_ = d_curr_ = nullptr;
        d_mask_ = nullptr;
        h_mask_ = nullptr;
        const size_t want = std::max<size_t>(n, 4096);
        if ((e = hipMalloc(reinterpret_cast<void**>(&d_prev_), 16 * want)) != hipSuccess ||
            (e = hipMalloc(reinterpret_cast<void**>(&d_curr_), 16 * want)) != hipSuccess ||
            (e = hipMalloc(reinterpret_cast<void**>(&d_mask_), want)) != hipSuccess ||
            (e = hipHostMalloc(reinterpret_cast<void**>(&h_mask_), want, hipHostMallocDefault)) !=
                hipSuccess)
            return fail("scorer buffers", e);
        cap_n_ = want;
    }
    if (k > cap_k_) {
        if (d_F_) (void)hipFree(d_F_);
        if (d_counts_) (void)hipFree(d_counts_);
        if (h_F_) (void)hipHostFree(h_F_);
        if (h_counts_) (void)hipHostFree(h_counts_);
        const size_t want = std::max<size_t>(k, 64);
        if ((e = hipMalloc(reinterpret_cast<void**>(&d_F_), 72 * want)) != hipSuccess ||
            (e = hipMalloc(reinterpret_cast<void**>(&d_counts_), 4 * want)) != hipSuccess ||
            (e = hipHostMalloc(reinterpret_cast<void**>(&h_F_), 72 * want, hipHostMallocDefault)) !=
                hipSuccess ||
            (e = hipHostMalloc(reinterpret_cast<void**>(&h_counts_), 4 * want,
                               hipHostMallocDefault)) != hipSuccess)
            return fail("scorer buffers", e);
        cap_k_ = want;
    }
    return std::nullopt;
}

std::optional<std::string> GpuRansacScorer::upload(const std::vector<Point2>& prev,
                                                   const std::vector<Point2>& curr,
                                                   double threshold) {
    static_assert(sizeof(Point2) == 16);
    if (auto err = reserve(prev.size(), 0)) return err;
    n_ = prev.size();
    threshold_ = threshold;
    if (n_ == 0) return std::nullopt;
    hipError_t e = hipMemcpyAsync(d_prev_, prev.data(), 16 * n_, hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_curr_, curr.data(), 16 * n_, hipMemcpyHostToDevice, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return std::string("upload correspondences: ") + hipGetErrorString(e);
    return std::nullopt;
}

std::optional<std::string> GpuRansacScorer::score(const std::vector<Mat3>& F,
                                                  std::vector<uint32_t>& counts,
                                                  std::vector<uint8_t>* mask) {
    if (mask && F.size() != 1) return std::string("a mask is kept for one hypothesis");
    if (auto err = reserve(n_, F.size())) return err;
    std::memcpy(h_F_, F.data(), 72 * F.size());
    hipError_t e = hipMemcpyAsync(d_F_, h_F_, 72 * F.size(), hipMemcpyHostToDevice, stream_);
    if (e != hipSuccess) return std::string("upload hypotheses: ") + hipGetErrorString(e);
    const int rc = orbfe_ransac_score_async(ctx_, d_F_, static_cast<uint32_t>(F.size()), d_prev_,
                                            d_curr_, static_cast<uint32_t>(n_), threshold_,
                                            d_counts_, mask ? d_mask_ : nullptr, stream_);
    if (rc != ORBFE_OK)
        return std::string("orbfe_ransac_score_async: ") + orbfe_last_error(ctx_);
    e = hipMemcpyAsync(h_counts_, d_counts_, 4 * F.size(), hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess && mask)
        e = hipMemcpyAsync(h_mask_, d_mask_, n_, hipMemcpyDeviceToHost, stream_);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) return std::string("score hypotheses: ") + hipGetErrorString(e);
    counts.assign(h_counts_, h_counts_ + F.size());
    if (mask) mask->assign(h_mask_, h_mask_ + n_);
    return std::nullopt;
}

// --- RansacHipPass ---------------------------------------------------------------
RansacHipPass::RansacHipPass(compute::Store& storage, std::shared_ptr<compute::Device> device,
                             std::shared_ptr<FeatureBuffers> buffers, RansacOpts opts)
    : storage_(storage),
      device_(std::move(device)),
      buffers_(std::move(buffers)),
      opts_(opts),
      rng_(opts.rng_seed) {}

RansacHipPass::~RansacHipPass() = default;

std::string RansacHipPass::getId() const { return "[RANSAC Features HIP pass]"; }

std::optional<std::string> RansacHipPass::initialize() {
    if (auto e = buffers_->initialize()) return e;
    scorer_ = std::make_unique<GpuRansacScorer>(buffers_->ctx(), device_->stream());
    return std::nullopt;
}

std::optional<std::string> RansacHipPass::execute() {
    const auto m = storage_.getPtr<MatchResult>(ResourceIdentifier::MatchedFeaturesName);
    if (!m) {  // no matches this frame (ransac_cpu.cpp:269-277)
        storage_.set(ResourceIdentifier::RansacResultName, RansacResult{});
        return std::nullopt;
    }
    std::vector<Point2> prev, curr;
    std::vector<std::pair<Feature, Feature>> feats;
    FlattenMatchResult(**m, prev, curr, feats);
    RansacResult result;
    result.stats.total_matches = prev.size();
    if (prev.size() >= opts_.min_matches) {
        if (!scorer_) return std::string("not initialized");
        if (auto e = scorer_->upload(prev, curr, opts_.inlier_threshold)) return e;
        RansacOutcome out;
        const RansacScorer score = [this](const std::vector<Mat3>& F, std::vector<uint32_t>& c,
                                          std::vector<uint8_t>* mask) {
            return scorer_->score(F, c, mask);
        };
        if (auto e = RunRansac(prev, curr, opts_, rng_, score, out)) return e;
        result.stats = out.stats;
        if (out.stats.model_found) {
            // BuildFilteredMatchResult (ransac_cpu.cpp:227-240): keyed by the
            // current feature, in its level.
            for (size_t i : out.inliers) {
                const auto& [c, p] = feats[i];
                if (c.lod < kLevelsOfDetail) result.inliers[c.lod].emplace(c, p);
            }
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    result.fundamental_matrix[r][c] = out.F[r * 3 + c];
        }
    }
    storage_.set(ResourceIdentifier::RansacResultName, std::move(result));
    return std::nullopt;
}

}  // namespace wslam
