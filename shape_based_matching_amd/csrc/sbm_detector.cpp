// line2Dup::Detector on MI355X: the template bank (host model and packed HBM copy), training, and the
// entry points that hand images to the device pipelines.
// Reference: line2Dup.cpp:1027-1121 (Detector, match), 1270-1398 (training, getters).
#include "sbm_detector.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <optional>

#pragma clang fp contract(off)

namespace sbm {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(kEHip, std::string(what) + ": " + hipGetErrorString(e));
}

void DevMem::ensure(size_t n) {
    if (n <= bytes && p) return;
    release();
    size_t want = std::max<size_t>(n, 256);
    hip_check(hipMalloc(&p, want), "hipMalloc");
    bytes = want;
}
void DevMem::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
void HostPinned::ensure(size_t n) {
    if (n <= bytes && p) return;
    release();
    size_t want = std::max<size_t>(n, 4096);
    hip_check(hipHostMalloc(&p, want, hipHostMallocDefault), "hipHostMalloc");
    bytes = want;
}
void HostPinned::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

Detector::Detector() {}
Detector::Detector(const std::vector<int>& T) : pyramid_levels((int)T.size()), T_at_level(T) {}
Detector::Detector(int nf, const std::vector<int>& T, float weak, float strong)
    : weak_threshold(weak), num_features(nf), strong_threshold(strong), pyramid_levels((int)T.size()), T_at_level(T) {}

Detector::~Detector() {
    DeviceScope ds(device_);
    batch_pipes_.clear();
    pipe_.reset();
}

void Detector::ensure_device() {
    if (pipe_) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        throw Error(kENoDevice, "no HIP device is visible: the matching path runs only on the GPU");
    hip_check(hipGetDevice(&device_), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&num_cu_, hipDeviceAttributeMultiprocessorCount, device_), "attr");
    pipe_.reset(new Pipeline(device_, num_cu_));
}

Pipeline& Detector::pipeline() {
    ensure_device();
    return *pipe_;
}
hipStream_t Detector::stream() { return pipeline().stream(); }

const StageTimes& Detector::last_times() const { return last_pipe() ? last_pipe()->last_times : empty_times_; }
// the pipeline of the last finished match: the detector's own, or the batch pipeline of match_batch's last launch
const Pipeline* Detector::last_pipe() const { return last_pipe_ ? last_pipe_ : pipe_.get(); }
const std::vector<std::string>& Detector::kernel_names() const {
    return last_pipe() ? last_pipe()->kernel_names : empty_names_;
}
const std::vector<float>& Detector::kernel_ms() const { return last_pipe() ? last_pipe()->kernel_ms : empty_ms_; }

// Detector::match asserts (line2Dup.cpp:612, 681, 724-725): every level's rows*cols % 16 and rows, cols % T.
void Detector::check_sizes(int rows, int cols) const {
    if (rows <= 0 || cols <= 0) throw Error(kESize, "empty image");
    int r = rows, c = cols;
    for (int l = 0; l < pyramid_levels; ++l) {
        if (l > 0) { r /= 2; c /= 2; }
        const int T = T_at_level[l];
        if (T < 1 || T > 16) throw Error(kEInval, "T must be in [1, 16]");
        if (((long long)r * c) % 16 != 0 || r % T != 0 || c % T != 0 || r == 0 || c == 0)
            throw Error(kESize, "image size must be divisible by 16 and by T at every pyramid level "
                                "(line2Dup.cpp:612, 724-725)");
        // linear-memory offsets are 32-bit: planes*rows*cols plus the guard must stay below 2^31
        if (lm_level_bytes(orientation_bins, r, c, T) >= (1LL << 31))
            throw Error(kESize, "image too large for one match (more than 2^31 bytes of linear memories)");
    }
}

void Detector::check_channels(int channels) const {
    if (channels != 1 && channels != 3) throw Error(kEInval, "channels must be 1 or 3");
    if (orientation_bins == 16 && channels != 3)
        throw Error(kEInval, "the 16-orientation variant reads three interleaved channels "
                             "(line2Dup_16bit_ori.cpp:319-367)");
}

void Detector::ensure_bank() {
    if (!bank_dirty_) return;
    ensure_device();
    const int L = pyramid_levels;
    Bank& b = bank_;
    if (orientation_bins != 8 && orientation_bins != 16) throw Error(kEInval, "orientation_bins must be 8 or 16");
    b.L = L;
    b.T = T_at_level;
    b.weak = weak_threshold;
    b.bins = orientation_bins;
    b.g_class.clear();
    b.g_tid.clear();
    b.names.clear();
    b.h_tl.clear();
    std::vector<uint2> feats;
    int ci = 0;
    for (const auto& kv : class_templates) {
        b.names.push_back(kv.first);
        for (size_t t = 0; t < kv.second.size(); ++t) {
            const TemplatePyramid& tp = kv.second[t];
            if ((int)tp.size() != L)
                throw Error(kEInval, "template pyramid of class '" + kv.first + "' has " + std::to_string(tp.size()) +
                                         " levels, the detector has " + std::to_string(L));
            for (int l = 0; l < L; ++l) {
                const Template& te = tp[l];
                TemplLevelDev d;
                d.width = te.width;
                d.height = te.height;
                d.nfeat = (int)te.features.size();
                d.off = (int)feats.size();
                // line2Dup.cpp:1166, 1231 raise at 8192; the 16-orientation variant takes up to 16383
                // (line2Dup_16bit_ori.cpp:1166) -- 4 * 16383 still fits the 16-bit sums
                if (d.nfeat >= (orientation_bins == 16 ? 16384 : 8192))
                    throw Error(kEFeatures, "feature size too large (line2Dup.cpp:1166, 1231)");
                for (const Feature& f : te.features) {
                    if (f.label < 0 || f.label >= orientation_bins)
                        throw Error(kEInval, "Invalid value of quantized parameter");
                    if (f.x < -32768 || f.x > 32767 || f.y < -32768 || f.y > 32767)
                        throw Error(kEInval, "feature coordinate outside the int16 range");
                    uint2 u;
                    u.x = (uint32_t)(uint16_t)(int16_t)f.x | ((uint32_t)(uint16_t)(int16_t)f.y << 16);
                    u.y = (uint32_t)f.label;
                    feats.push_back(u);
                }
                if (feats.size() > (size_t)INT_MAX) throw Error(kEInval, "bank too large");
                b.h_tl.push_back(d);
            }
            b.g_class.push_back(ci);
            b.g_tid.push_back((int)t);
        }
        ++ci;
    }
    const int G = b.size();
    b.sorted_all.resize(G);
    std::iota(b.sorted_all.begin(), b.sorted_all.end(), 0);
    std::stable_sort(b.sorted_all.begin(), b.sorted_all.end(), [&](int a, int c) {
        return b.h_tl[(size_t)a * L + L - 1].nfeat > b.h_tl[(size_t)c * L + L - 1].nfeat;
    });
    hipStream_t s = pipe_->stream();
    hip_check(hipStreamSynchronize(s), "sync before bank upload");
    for (auto& p : batch_pipes_) hip_check(hipStreamSynchronize(p->stream()), "sync before bank upload");
    b.d_tl.ensure(std::max<size_t>(b.h_tl.size(), 1) * sizeof(TemplLevelDev));
    b.d_gclass.ensure(std::max<size_t>(b.g_class.size(), 1) * sizeof(int));
    b.d_gtid.ensure(std::max<size_t>(b.g_tid.size(), 1) * sizeof(int));
    if (G > 0) {
        hip_check(hipMemcpyAsync(b.d_gclass.p, b.g_class.data(), G * sizeof(int), hipMemcpyHostToDevice, s), "H2D class");
        hip_check(hipMemcpyAsync(b.d_gtid.p, b.g_tid.data(), G * sizeof(int), hipMemcpyHostToDevice, s), "H2D tid");
    }
    b.d_feats.ensure(std::max<size_t>(feats.size(), 1) * sizeof(uint2) + 256);
    if (!b.h_tl.empty())
        hip_check(hipMemcpyAsync(b.d_tl.p, b.h_tl.data(), b.h_tl.size() * sizeof(TemplLevelDev), hipMemcpyHostToDevice, s),
                  "H2D tl");
    if (!feats.empty())
        hip_check(hipMemcpyAsync(b.d_feats.p, feats.data(), feats.size() * sizeof(uint2), hipMemcpyHostToDevice, s),
                  "H2D feats");
    hip_check(hipStreamSynchronize(s), "bank upload");
    b.version = ++bank_versions_;
    bank_dirty_ = false;
}

void Detector::release_match_buffers() {
    DeviceScope ds(device_);
    if (pipe_) pipe_->release();
    for (auto& p : batch_pipes_) p->release();
}

size_t Detector::device_bytes() const {
    size_t s = bank_.d_tl.bytes + bank_.d_feats.bytes + bank_.d_gclass.bytes + bank_.d_gtid.bytes;
    if (pipe_) s += pipe_->device_bytes();
    for (const auto& p : batch_pipes_) s += p->device_bytes();
    return s;
}

std::vector<std::string> Detector::class_names() const {
    std::vector<std::string> v;
    for (const auto& kv : class_templates) v.push_back(kv.first);
    return v;
}

// Templates to match, as global indices: all classes in std::map order, or the listed ones in the
// given order (line2Dup.cpp:1095-1111).  `sorted` orders them by lowest-level feature count, largest
// first, for the persistent scheduler.
void Detector::selection(const std::vector<std::string>& class_ids, bool sorted, std::vector<int>& out) const {
    out.clear();
    const auto& names = bank_.names;
    std::vector<char> take(names.size(), 0);
    std::vector<int> cls_order;
    if (class_ids.empty()) {
        for (size_t i = 0; i < names.size(); ++i) cls_order.push_back((int)i);
    } else {
        for (const auto& c : class_ids) {
            auto it = std::lower_bound(names.begin(), names.end(), c);
            if (it != names.end() && *it == c) {
                const int i = (int)(it - names.begin());
                if (!take[i]) cls_order.push_back(i);  // a repeated class yields only exact duplicates
                take[i] = 1;
            }
        }
    }
    std::fill(take.begin(), take.end(), 0);
    for (int c : cls_order) take[c] = 1;
    if (sorted) {
        for (int g : bank_.sorted_all)
            if (take[bank_.g_class[g]]) out.push_back(g);
    } else {
        for (int c : cls_order)
            for (int g = 0; g < bank_.size(); ++g)
                if (bank_.g_class[g] == c) out.push_back(g);
    }
}

std::vector<MatchRec> Detector::match(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                                      const std::vector<std::string>& class_ids, const uint8_t* mask, int mode,
                                      bool device_input) {
    if (mode != kModePlain && mode != kModeBatched) throw Error(kEInval, "unknown match mode");
    check_channels(channels);
    if (stride < cols * channels) throw Error(kEInval, "stride smaller than a row");
    check_sizes(rows, cols);
    ensure_device();
    DeviceScope ds(device_);
    ensure_bank();
    std::vector<int> sel;
    selection(class_ids, mode == kModeBatched, sel);
    ImageArgs im;
    im.img = img;
    im.rows = rows;
    im.cols = cols;
    im.stride = stride;
    im.channels = channels;
    im.mask = mask;
    im.device = device_input;
    pipe_->tuning = tuning;
    pipe_->timing_enabled = timing_enabled;
    pipe_->graph_timing = graph_timing;
    last_pipe_ = pipe_.get();
    if (mode == kModeBatched) return pipe_->match_reuse(bank_, im, threshold, sel);
    return pipe_->match(bank_, im, threshold, sel, mode);
}

std::vector<MatchRec> Detector::match_nms(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                                          const std::vector<std::string>& class_ids, const uint8_t* mask,
                                          bool device_input, float score_threshold, float nms_threshold, int top_k,
                                          std::vector<MatchRec>* all) {
    check_channels(channels);
    if (stride < cols * channels) throw Error(kEInval, "stride smaller than a row");
    check_sizes(rows, cols);
    ensure_device();
    DeviceScope ds(device_);
    ensure_bank();
    std::vector<int> sel;
    selection(class_ids, true, sel);
    ImageArgs im;
    im.img = img;
    im.rows = rows;
    im.cols = cols;
    im.stride = stride;
    im.channels = channels;
    im.mask = mask;
    im.device = device_input;
    pipe_->tuning = tuning;
    pipe_->timing_enabled = timing_enabled;
    last_pipe_ = pipe_.get();
    return pipe_->match_nms(bank_, im, threshold, sel, score_threshold, nms_threshold, top_k, all);
}

std::vector<std::vector<MatchRec>> Detector::match_batch(const std::vector<ImageArgs>& images, float threshold,
                                                         const std::vector<std::string>& class_ids, int slots) {
    for (const auto& im : images) {
        check_channels(im.channels);
        if (im.stride < im.cols * im.channels) throw Error(kEInval, "stride smaller than a row");
        check_sizes(im.rows, im.cols);
    }
    ensure_device();
    DeviceScope ds(device_);
    ensure_bank();
    std::vector<int> sel;
    selection(class_ids, true, sel);
    slots = std::max(1, std::min(slots, 16));
    while ((int)batch_pipes_.size() < slots) batch_pipes_.emplace_back(new Pipeline(device_, num_cu_));
    // size buckets (rows, cols, channels, mask), in order of first appearance; each bucket goes out in
    // launches of up to `per` images, every stage of a launch covering all of its images
    std::vector<std::vector<int>> buckets;
    for (size_t i = 0; i < images.size(); ++i) {
        const ImageArgs& im = images[i];
        size_t b = 0;
        for (; b < buckets.size(); ++b) {
            const ImageArgs& f = images[buckets[b][0]];
            if (f.rows == im.rows && f.cols == im.cols && f.channels == im.channels &&
                (f.mask != nullptr) == (im.mask != nullptr))
                break;
        }
        if (b == buckets.size()) buckets.emplace_back();
        buckets[b].push_back((int)i);
    }
    std::vector<std::vector<int>> launches;
    for (const auto& bk : buckets) {
        const int m = (int)bk.size();
        // default: two launches per bucket in turn, so one runs while the host finishes the other
        int per = images_per_launch > 0 ? images_per_launch : (m + 1) / 2;
        per = std::max(1, std::min(per, kMaxBatch));
        for (int k = 0; k < m; k += per)
            launches.emplace_back(bk.begin() + k, bk.begin() + std::min(m, k + per));
    }
    std::vector<std::vector<MatchRec>> results(images.size());
    std::vector<int> inflight(slots, -1);
    std::vector<ImageArgs> ims;
    auto finish = [&](int s) {
        auto out = batch_pipes_[s]->graph_finish_batch(bank_);
        last_pipe_ = batch_pipes_[s].get();
        const auto& idx = launches[inflight[s]];
        for (size_t k = 0; k < idx.size(); ++k) results[idx[k]] = std::move(out[k]);
        inflight[s] = -1;
    };
    try {
        for (size_t li = 0; li < launches.size(); ++li) {
            const int s = (int)(li % slots);
            Pipeline& p = *batch_pipes_[s];
            if (inflight[s] >= 0) finish(s);
            p.tuning = tuning;
            p.timing_enabled = timing_enabled;  // per-kernel events: the launch runs its stages eagerly
            ims.clear();
            for (int i : launches[li]) ims.push_back(images[i]);
            p.graph_launch_batch(bank_, ims.data(), (int)ims.size(), threshold, sel);
            inflight[s] = (int)li;
        }
        for (int s = 0; s < slots; ++s)
            if (inflight[s] >= 0) finish(s);
    } catch (...) {
        for (auto& p : batch_pipes_)
            if (p->graph_pending()) {
                try {
                    (void)p->graph_finish_batch(bank_);
                } catch (...) {
                }
            }
        throw;
    }
    return results;
}

// Launch-shape search on a prepared workload: the coarse stage on the path the device takes for this image
// (the gather kernel's shapes, or the nonzero-response scatter's), then the refinement.  Every shape must
// return the same matches; the fastest of each stage (HIP-event time of its launches) is kept in `tuning`.
Tuning Detector::autotune(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                          const std::vector<std::string>& class_ids, bool device_input, int iters,
                          std::vector<TuneRow>* trace, const std::vector<ImageArgs>* batch) {
    if (batch)
        for (const ImageArgs& b : *batch)
            if (!b.img || b.rows != rows || b.cols != cols || b.channels != channels || b.mask)
                throw Error(kEInval, "autotune batch images must have the image's size and channels, and no mask");
    const bool was_timing = timing_enabled;
    const Tuning saved = tuning;
    timing_enabled = true;
    std::vector<MatchRec> ref;
    std::vector<TuneRow> rows_out;
    auto same = [](const std::vector<MatchRec>& m, const std::vector<MatchRec>& r) {
        return m.size() == r.size() && std::equal(m.begin(), m.end(), r.begin(), [](const MatchRec& a, const MatchRec& b) {
                   return a.x == b.x && a.y == b.y && a.similarity == b.similarity && a.class_index == b.class_index &&
                          a.template_id == b.template_id;
               });
    };
    // launches of n images when match_batch makes multi-image launches (the path in auto mode and the best
    // shapes depend on the launch's image count): the batch's images in turn, or copies of the image
    const int n = std::max(1, std::min(images_per_launch, kMaxBatch));
    const bool distinct = n > 1 && batch && !batch->empty();
    std::vector<ImageArgs> copies;
    std::vector<std::vector<MatchRec>> refs;  // the matches of copies[i] (distinct images)
    std::vector<int> sel;
    std::optional<DeviceScope> ds;
    if (n > 1) {
        check_channels(channels);
        check_sizes(rows, cols);
        ensure_device();
        ds.emplace(device_);
        ensure_bank();
        selection(class_ids, true, sel);
        while (batch_pipes_.size() < 2) batch_pipes_.emplace_back(new Pipeline(device_, num_cu_));
        ImageArgs im;
        im.img = img;
        im.rows = rows;
        im.cols = cols;
        im.stride = stride;
        im.channels = channels;
        im.device = device_input;
        copies.assign((size_t)n, im);
        if (distinct)
            for (int i = 0; i < n; ++i) copies[(size_t)i] = (*batch)[(size_t)i % batch->size()];
    }
    auto ref_of = [&](size_t i) -> const std::vector<MatchRec>& { return distinct ? refs[i] : ref; };
    // one n-image launch with shape t: its results (each must equal ref) and its device span
    auto launch_n = [&](const Tuning& t, bool check) {
        Pipeline& p = *batch_pipes_[0];
        p.tuning = t;
        p.timing_enabled = false;
        p.graph_timing = true;
        p.graph_launch_batch(bank_, copies.data(), n, threshold, sel);
        auto out = p.graph_finish_batch(bank_);
        p.graph_timing = false;
        if (check)
            for (size_t i = 0; i < out.size(); ++i)
                if (!same(out[i], ref_of(i))) throw Error(kEInval, "launch shapes disagree: autotune aborted");
        return p.last_times;
    };
    // n-image launches as match_batch streams them: two pipelines in turn, two launches in flight, so that a shape
    // is timed beside the other launch's kernels as it runs in a stream; host milliseconds per launch over ten
    // launches, after one checked launch on each pipeline (which captures its graph)
    auto stream_ms = [&](const Tuning& t) {
        Pipeline* pp[2] = {batch_pipes_[0].get(), batch_pipes_[1].get()};
        for (Pipeline* p : pp) {
            p->tuning = t;
            p->timing_enabled = false;
            p->graph_timing = false;
            p->graph_launch_batch(bank_, copies.data(), n, threshold, sel);
        }
        for (Pipeline* p : pp) {
            const auto out = p->graph_finish_batch(bank_);
            for (size_t i = 0; i < out.size(); ++i)
                if (!same(out[i], ref_of(i))) throw Error(kEInval, "launch shapes disagree: autotune aborted");
        }
        constexpr int kLaunches = 10;
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < kLaunches; ++i) {
            Pipeline* p = pp[i & 1];
            if (i >= 2) (void)p->graph_finish_batch(bank_);
            p->graph_launch_batch(bank_, copies.data(), n, threshold, sel);
        }
        for (Pipeline* p : pp) (void)p->graph_finish_batch(bank_);
        const std::chrono::duration<double, std::milli> dt = std::chrono::steady_clock::now() - t0;
        return (float)(dt.count() / kLaunches);
    };
    // device time of one stage's kernels (one image), or host time per n-image launch in a stream (multi: the
    // fastest of `iters`)
    auto timed = [&](const Tuning& t, int stage, bool multi) {
        if (multi) {
            float best = 1e30f;
            for (int it = 0; it < std::max(1, iters); ++it) best = std::min(best, stream_ms(t));
            rows_out.push_back({stage, t, best});
            return best;
        }
        tuning = t;
        float tot = 0.f;
        for (int it = 0; it < std::max(1, iters); ++it) {
            auto m = match(img, rows, cols, stride, channels, threshold, class_ids, nullptr, kModeBatched, device_input);
            if (!same(m, ref)) throw Error(kEInval, "launch shapes disagree: autotune aborted");
            const auto& names = kernel_names();
            const auto& ms = kernel_ms();
            for (size_t k = 0; k < names.size(); ++k) {
                const std::string& nm = names[k];
                const bool coarse = nm == "coarse_batched" || (stage == 0 && nm.rfind("pack_nibbles", 0) == 0);
                // the refinement's stage: its launches and the banded copies they read
                const bool refine = nm.rfind("refine_l", 0) == 0 || nm.rfind("band_l", 0) == 0;
                if ((stage < 2 && coarse) || (stage == 2 && refine)) tot += ms[k];
            }
        }
        const float avg = tot / (float)std::max(1, iters);
        rows_out.push_back({stage, t, avg});
        return avg;
    };
    // Stream timings drift over a search (clocks, the other launch in flight) by about as much as the best shapes
    // differ: the four fastest rows of a stage recorded since row `from` are timed twice more, in turn, and the
    // fastest of each shape's timings decides.
    auto playoff = [&](size_t from, int stage, Tuning& out) {
        std::vector<std::pair<float, Tuning>> c;
        for (size_t i = from; i < rows_out.size(); ++i)
            if (rows_out[i].stage == stage) c.push_back({rows_out[i].ms, rows_out[i].tuning});
        std::stable_sort(c.begin(), c.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        if (c.size() > 4) c.resize(4);
        for (int r = 0; r < 2; ++r)
            for (auto& e : c) e.first = std::min(e.first, stream_ms(e.second));
        const auto it = std::min_element(c.begin(), c.end(),
                                         [](const auto& x, const auto& y) { return x.first < y.first; });
        out = it->second;
        return it->first;
    };
    Tuning best = saved;
    try {
        ref = match(img, rows, cols, stride, channels, threshold, class_ids, nullptr, kModeBatched, device_input);
        const bool single_list = last_times().coarse_sparse != 0;
        if (distinct)
            for (const ImageArgs& c : copies)
                refs.push_back(match(c.img, c.rows, c.cols, c.stride, c.channels, threshold, class_ids, nullptr,
                                     kModeBatched, c.device));
        const bool list_path = n > 1 ? launch_n(saved, true).coarse_sparse != 0 : single_list;
        float best_ms = 1e30f;
        if (list_path) {
            const size_t list_from = rows_out.size();
            // The items stay on the work counter (list_order 0): fixed XCD shares give each XCD one image of an
            // eight-image launch, which balances on the copies timed here but not on distinct scenes.
            for (int th : {1024, 512, 256})
                for (int seg : {16384, 8192, 4096})
                    for (int bpc : {1, 2, 4, 8}) {
                        // a CU holds 2048 threads and 160 KB of LDS
                        if (bpc * th > 2048 || bpc * kern::sparse_block_lds(seg) > 160 * 1024) continue;
                        Tuning t = saved;
                        t.list_threads = th;
                        t.list_segment = seg;
                        t.list_bpc = bpc;
                        t.list_order = 0;
                        const float v = timed(t, 1, n > 1);
                        if (v < best_ms) { best_ms = v; best = t; }
                    }
            // the pruned scatter's split on the best shape (the shapes above ran with the detector's)
            const Tuning shape = best;
            for (int split : {100, 60, 50, 40}) {
                if (split == shape.list_split) continue;
                Tuning t = shape;
                t.list_split = split;
                const float v = timed(t, 1, n > 1);
                if (v < best_ms) { best_ms = v; best = t; }
            }
            if (n > 1) best_ms = playoff(list_from, 1, best);
        }
        // the gather kernel's shapes (coarse mode 0 while timed), on the launches that take it: the n-image
        // launches, or single images when only the n-image launches take the list.  When the n-image launches
        // take the list, the gather kernel is also timed on them, as whole launches in a stream like the list
        // shapes were: if it is faster there (short slices of the list: a small scene's planes hold few nonzero
        // responses each), the workload keeps coarse mode 0.
        auto gather_search = [&](const Tuning& base, bool multi, Tuning& out) {
            const size_t from = rows_out.size();
            float bm = 1e30f;
            for (int gather : {0, 1, 2, 3})
                for (int ppl : {4, 8, 16, 32}) {
                    if ((gather == 3) != (ppl == 32)) continue;  // the nibble gather owns 32 positions per lane
                    for (int split : {1, 2, 4, 8})
                        for (int wpb : {1, 2, 4}) {
                            // a split block is `split` wavefronts; its waves_per_block is recorded as min(split, 4)
                            if (split > 1 && wpb != std::min(split, 4)) continue;
                            for (int bpc : {2, 4, 8, 16}) {
                                Tuning t = base;
                                t.gather = gather;
                                t.ppl = ppl;
                                t.waves_per_block = wpb;
                                t.blocks_per_cu = bpc;
                                t.split = split;
                                t.coarse_mode = 0;  // the shapes are those of the gather kernel
                                const float v = timed(t, 0, multi);
                                if (v < bm) { bm = v; out = t; }
                            }
                        }
                }
            if (multi) bm = playoff(from, 0, out);
            return bm;
        };
        bool gather_wins = false;
        if (list_path && n > 1) {
            // one gather shape first (the detector's): the full search only when it comes within a quarter of
            // the list's best (large banks on large scenes are several times faster on the list)
            Tuning probe = best;
            probe.ppl = saved.ppl;
            probe.waves_per_block = saved.waves_per_block;
            probe.blocks_per_cu = saved.blocks_per_cu;
            probe.gather = saved.gather;
            probe.split = saved.split;
            probe.coarse_mode = 0;
            if (timed(probe, 0, true) < 1.25f * best_ms) {
                Tuning g = best;
                float gather_ms = gather_search(best, true, g);
                float list_ms = best_ms;
                // within a tenth of each other: three more stream timings of each, in turn, and the fastest of each
                if (gather_ms < 1.1f * list_ms && list_ms < 1.1f * gather_ms)
                    for (int r = 0; r < 3; ++r) {
                        list_ms = std::min(list_ms, stream_ms(best));
                        gather_ms = std::min(gather_ms, stream_ms(g));
                    }
                if (gather_ms < list_ms) {  // g.coarse_mode is 0
                    best = g;
                    best_ms = gather_ms;
                    gather_wins = true;
                }
            }
        }
        if (!gather_wins && (!list_path || !single_list)) {
            Tuning g = best;
            (void)gather_search(best, n > 1 && !list_path, g);
            best = g;
            best.coarse_mode = saved.coarse_mode;
        }
        if (bank_.L >= 2) {  // the refinement's shapes, with the coarse stage's best
            best_ms = 1e30f;
            const size_t refine_from = rows_out.size();
            const Tuning base = best;
            // (wavefronts per candidate, blocks per CU): 8, 16 or 32 wavefronts per CU (16 or 32 with 16-wave blocks)
            static const int kShapes[][2] = {{1, 8}, {1, 16}, {1, 32}, {2, 4}, {2, 8}, {2, 16}, {4, 2},
                                             {4, 4}, {4, 8}, {8, 1}, {8, 2}, {8, 4}, {16, 1}, {16, 2}};
            for (int layout : {0, 1, 2, 3, 4})
                for (const auto& wb : kShapes)
                    for (int group : {8, 16}) {
                        Tuning t = base;
                        t.refine_waves = wb[0];
                        t.refine_bpc = wb[1];
                        t.refine_group = group;
                        t.refine_layout = layout;
                        const float v = timed(t, 2, n > 1);
                        if (v < best_ms) { best_ms = v; best = t; }
                    }
            if (n > 1) best_ms = playoff(refine_from, 2, best);
        }
    } catch (...) {
        for (auto& p : batch_pipes_)  // a launch still in flight when a shape failed
            if (p->graph_pending()) {
                try {
                    (void)p->graph_finish_batch(bank_);
                } catch (...) {
                }
            }
        tuning = saved;
        timing_enabled = was_timing;
        throw;
    }
    tuning = best;
    timing_enabled = was_timing;
    last_autotune = rows_out;
    if (trace) *trace = std::move(rows_out);
    return best;
}

// ---------------------------------------------------------------------------------------------
// Bank editing (line2Dup.cpp:1270-1362) and getters (line2Dup.cpp:1363-1398)
// ---------------------------------------------------------------------------------------------
int Detector::add_template_pyramid(const std::string& class_id, const TemplatePyramid& tp) {
    auto& v = class_templates[class_id];
    v.push_back(tp);
    bank_dirty_ = true;
    return (int)v.size() - 1;
}

void Detector::gradient(const uint8_t* img, int rows, int cols, int stride, int channels, uint8_t* q, float* mag,
                        float* angle) {
    ensure_device();
    DeviceScope ds(device_);
    check_channels(channels);
    hipStream_t s = pipe_->stream();
    const size_t qbytes = (size_t)rows * cols * (size_t)(orientation_bins / 8);  // 16-bit values with 16 bins
    DevMem dimg, dq, dmag, dang;
    dimg.ensure((size_t)rows * cols * channels);
    dq.ensure(qbytes);
    dmag.ensure((size_t)rows * cols * 4);
    dang.ensure((size_t)rows * cols * 4);
    hip_check(hipMemcpy2DAsync(dimg.p, (size_t)cols * channels, img, (size_t)stride, (size_t)cols * channels, rows,
                               hipMemcpyHostToDevice, s), "H2D");
    hip_check(kern::gradient_quantize(dimg.as<uint8_t>(), rows, cols, cols * channels, channels, nullptr, weak_threshold,
                                      dq.as<uint8_t>(), dmag.as<float>(), dang.as<float>(), s, nullptr, 0,
                                      orientation_bins),
              "gradient");
    hip_check(hipMemcpyAsync(q, dq.p, qbytes, hipMemcpyDeviceToHost, s), "D2H");
    if (mag) hip_check(hipMemcpyAsync(mag, dmag.p, (size_t)rows * cols * 4, hipMemcpyDeviceToHost, s), "D2H");
    if (angle) hip_check(hipMemcpyAsync(angle, dang.p, (size_t)rows * cols * 4, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "sync");
}

// Detector::addTemplate: gradients on the GPU, candidate selection on the host.
int Detector::add_template(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask,
                           const std::string& class_id, int nf_override) {
    std::vector<TemplatePyramid>& tps = class_templates[class_id];  // created even on failure, as the reference
    const int template_id = (int)tps.size();
    ensure_device();
    DeviceScope ds(device_);
    check_channels(channels);
    hipStream_t s = pipe_->stream();
    if ((rows >> (pyramid_levels - 1)) < 1 || (cols >> (pyramid_levels - 1)) < 1)
        throw Error(kESize, "image too small for the pyramid (cv::pyrDown of an empty size)");
    TemplatePyramid tp(pyramid_levels);
    int nf = nf_override > 0 ? nf_override : num_features;
    std::vector<uint8_t> cur((size_t)rows * cols * channels), curmask;
    for (int y = 0; y < rows; ++y) memcpy(&cur[(size_t)y * cols * channels], img + (size_t)y * stride, (size_t)cols * channels);
    if (mask) curmask.assign(mask, mask + (size_t)rows * cols);
    int r = rows, c = cols;
    for (int l = 0; l < pyramid_levels; ++l) {
        if (l > 0) {
            nf /= 2;  // ColorGradientPyramid::pyrDown, line2Dup.cpp:400
            DevMem dsrc, ddst, dm, dmd;
            dsrc.ensure(cur.size());
            ddst.ensure((size_t)(r / 2) * (c / 2) * channels);
            hip_check(hipMemcpyAsync(dsrc.p, cur.data(), cur.size(), hipMemcpyHostToDevice, s), "H2D");
            if (mask) {
                dm.ensure(curmask.size());
                dmd.ensure((size_t)(r / 2) * (c / 2));
                hip_check(hipMemcpyAsync(dm.p, curmask.data(), curmask.size(), hipMemcpyHostToDevice, s), "H2D");
            }
            hip_check(kern::pyr_down(dsrc.as<uint8_t>(), r, c, c * channels, channels, ddst.as<uint8_t>(),
                                     mask ? dm.as<uint8_t>() : nullptr, mask ? dmd.as<uint8_t>() : nullptr, s),
                      "pyr_down");
            r /= 2;
            c /= 2;
            cur.resize((size_t)r * c * channels);
            hip_check(hipMemcpyAsync(cur.data(), ddst.p, cur.size(), hipMemcpyDeviceToHost, s), "D2H");
            if (mask) {
                curmask.resize((size_t)r * c);
                hip_check(hipMemcpyAsync(curmask.data(), dmd.p, curmask.size(), hipMemcpyDeviceToHost, s), "D2H");
            }
            hip_check(hipStreamSynchronize(s), "sync");
        }
        std::vector<float> mag((size_t)r * c), ang((size_t)r * c);
        if (orientation_bins == 16) {  // line2Dup_16bit_ori.cpp:1301-1335
            std::vector<uint16_t> q16((size_t)r * c);
            gradient(cur.data(), r, c, c * channels, channels, reinterpret_cast<uint8_t*>(q16.data()), mag.data(),
                     nullptr);
            if (!extract_template16(mag.data(), q16.data(), mask ? curmask.data() : nullptr, r, c, strong_threshold, nf,
                                    tp[l].features))
                return -1;
        } else {
            std::vector<uint8_t> q((size_t)r * c);
            gradient(cur.data(), r, c, c * channels, channels, q.data(), mag.data(), ang.data());
            if (!extract_template(mag.data(), q.data(), ang.data(), mask ? curmask.data() : nullptr, r, c,
                                  strong_threshold, nf, tp[l].features))
                return -1;
        }
        tp[l].width = -1;
        tp[l].height = -1;
        tp[l].pyramid_level = l;
    }
    crop_templates(tp);
    tps.push_back(tp);
    bank_dirty_ = true;
    return template_id;
}

// addTemplate_rotate with rotate2d / rotatePoint (line2Dup.cpp:1306-1362)
int Detector::add_template_rotate(const std::string& class_id, int zero_id, float theta, float cx, float cy) {
    if (orientation_bins == 16)  // line2Dup_16bit_ori.cpp defines no addTemplate_rotate, and keeps no angles
        throw Error(kEInval, "addTemplate_rotate is not part of the 16-orientation variant");
    std::vector<TemplatePyramid>& tps = class_templates[class_id];
    const int template_id = (int)tps.size();
    if (zero_id < 0 || zero_id >= (int)tps.size()) throw Error(kENotFound, "zero_id out of range");
    const TemplatePyramid src = tps[zero_id];
    TemplatePyramid tp(pyramid_levels);
    const double ang = (double)(-theta / 180) * 3.1415926535897932384626433832795;
    const double ca = std::cos(ang), sa = std::sin(ang);
    for (int l = 0; l < pyramid_levels; ++l) {
        if (l > 0) { cx /= 2; cy /= 2; }
        for (const Feature& f : src[l].features) {
            const float px = (float)(f.x + src[l].tl_x), py = (float)(f.y + src[l].tl_y);
            const float ix = px - cx, iy = py - cy;
            float ox = (float)(ca * ix - sa * iy);
            float oy = (float)(sa * ix + ca * iy);
            ox = ox + cx;
            oy = oy + cy;
            Feature g;
            g.x = (int)(ox + 0.5f);
            g.y = (int)(oy + 0.5f);
            g.theta = f.theta - theta;
            while (g.theta > 360) g.theta -= 360;
            while (g.theta < 0) g.theta += 360;
            g.label = ((int)(g.theta * 16 / 360 + 0.5f)) & 7;
            tp[l].features.push_back(g);
        }
        tp[l].pyramid_level = l;
    }
    crop_templates(tp);
    tps.push_back(tp);
    bank_dirty_ = true;
    return template_id;
}

const std::vector<Template>& Detector::get_templates(const std::string& class_id, int template_id) const {
    auto it = class_templates.find(class_id);
    if (it == class_templates.end()) throw Error(kENotFound, "class not found: " + class_id);
    if (template_id < 0 || (size_t)template_id >= it->second.size()) throw Error(kENotFound, "template_id out of range");
    return it->second[template_id];
}

int Detector::num_templates() const {
    int n = 0;
    for (const auto& kv : class_templates) n += (int)kv.second.size();
    return n;
}
int Detector::num_templates(const std::string& class_id) const {
    auto it = class_templates.find(class_id);
    return it == class_templates.end() ? 0 : (int)it->second.size();
}

// ---------------------------------------------------------------------------------------------
// Debug / parity helpers: intermediate buffers of the default pipeline
// ---------------------------------------------------------------------------------------------
void Detector::prepare_image(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask) {
    check_channels(channels);
    check_sizes(rows, cols);
    ensure_device();
    DeviceScope ds(device_);
    ensure_bank();
    ImageArgs im;
    im.img = img;
    im.rows = rows;
    im.cols = cols;
    im.stride = stride;
    im.channels = channels;
    im.mask = mask;
    pipe_->tuning = tuning;
    pipe_->prepare(bank_, im);
}

void Detector::read_level(int level, int what, uint8_t* out) {
    if (!pipe_) throw Error(kEInval, "no prepared image");
    DeviceScope ds(device_);
    pipe_->read_level(level, what, out);
}

static int global_index(const Bank& b, const std::string& cls, int tid) {
    auto it = std::lower_bound(b.names.begin(), b.names.end(), cls);
    if (it == b.names.end() || *it != cls) throw Error(kENotFound, "class not found: " + cls);
    const int ci = (int)(it - b.names.begin());
    for (int g = 0; g < b.size(); ++g)
        if (b.g_class[g] == ci && b.g_tid[g] == tid) return g;
    throw Error(kENotFound, "template_id out of range");
}

void Detector::similarity_map(const std::string& class_id, int template_id, uint16_t* out) {
    ensure_device();
    DeviceScope ds(device_);
    ensure_bank();
    const int L = pyramid_levels;
    const int g = global_index(bank_, class_id, template_id);
    const TemplLevelDev& t = bank_.h_tl[(size_t)g * L + L - 1];
    const LevelGeom gl = pipe_->geom(bank_, L - 1);
    if (!gl.lm) throw Error(kEInval, "no prepared image");
    hipStream_t s = pipe_->stream();
    DevMem d;
    d.ensure((size_t)gl.hw * 2);
    hip_check(kern::similarity_map(gl, nullptr, bank_.d_feats.as<uint2>(), t.width, t.height, t.nfeat, t.off,
                                   d.as<uint16_t>(), s), "simmap");
    hip_check(hipMemcpyAsync(out, d.p, (size_t)gl.hw * 2, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "sync");
}

void Detector::similarity_local(const std::string& class_id, int template_id, int level, int cx, int cy, uint16_t* out) {
    ensure_device();
    DeviceScope ds(device_);
    ensure_bank();
    const int L = pyramid_levels;
    if (level < 0 || level >= L) throw Error(kEInval, "level out of range");
    const int g = global_index(bank_, class_id, template_id);
    const TemplLevelDev& t = bank_.h_tl[(size_t)g * L + level];
    const LevelGeom gl = pipe_->geom(bank_, level);
    if (!gl.lm) throw Error(kEInval, "no prepared image");
    hipStream_t s = pipe_->stream();
    DevMem d;
    d.ensure(512);
    hip_check(kern::similarity_local(gl, bank_.d_feats.as<uint2>(), t.nfeat, t.off, cx, cy, d.as<uint16_t>(), s),
              "simlocal");
    hip_check(hipMemcpyAsync(out, d.p, 512, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "sync");
}

int Detector::division_check() {
    ensure_device();
    DeviceScope ds(device_);
    hipStream_t s = pipe_->stream();
    DevMem d;
    d.ensure(sizeof(int));
    hip_check(hipMemsetAsync(d.p, 0, sizeof(int), s), "memset");
    hip_check(kern::division_check(d.as<int>(), s), "division_check");
    int n = -1;
    hip_check(hipMemcpyAsync(&n, d.p, sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "sync");
    return n;
}

void Detector::bin_table(int lo, int hi, uint8_t* out) {
    ensure_device();
    DeviceScope ds(device_);
    hipStream_t s = pipe_->stream();
    const size_t n = (size_t)(hi - lo + 1) * (hi - lo + 1);
    DevMem d;
    d.ensure(n);
    hip_check(kern::bin_table(lo, hi, d.as<uint8_t>(), s), "bin_table");
    hip_check(hipMemcpyAsync(out, d.p, n, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "sync");
}

}  // namespace sbm
