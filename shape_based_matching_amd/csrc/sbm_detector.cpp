// line2Dup::Detector on MI355X: bank packing, the per-image device pipeline and the host glue.
// Reference: line2Dup.cpp:1027-1121 (Detector, match), 1131-1268 (matchClass), 1270-1362 (training).
#include "sbm_detector.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

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

// score = (raw * 100.f) / (4 * n)   (line2Dup.cpp:1177, 1244)
float score_of(int raw, int n) {
    const float num = (float)raw * 100.f;
    return num / (float)(4 * n);
}

// Smallest raw in [0, 4n] whose coarse score is > thr (line2Dup.cpp:1179); INT_MAX if none.
static int raw_min_coarse(int n, float thr) {
    if (n <= 0) return INT_MAX;  // 0/0 is NaN and never passes
    int lo = 0, hi = 4 * n + 1;  // answer in [lo, hi); hi == 4n+1 means none
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (score_of(mid, n) > thr) hi = mid;
        else lo = mid + 1;
    }
    return lo <= 4 * n ? lo : INT_MAX;
}

// Smallest raw in [0, 4n] whose refined similarity is kept, i.e. !(sim < thr) (line2Dup.cpp:1261-1263).
// A window with no positive score keeps best_score = 0 (line2Dup.cpp:1236-1253).
static int raw_min_refine(int n, float thr) {
    auto keep = [&](int raw) {
        const float sim = raw > 0 ? score_of(raw, n) : 0.f;
        return !(sim < thr);
    };
    const int top = n > 0 ? 4 * n : 0;
    int lo = 0, hi = top + 1;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (keep(mid)) hi = mid;
        else lo = mid + 1;
    }
    return lo <= top ? lo : INT_MAX;
}

Detector::Detector() {}
Detector::Detector(const std::vector<int>& T) : pyramid_levels((int)T.size()), T_at_level(T) {}
Detector::Detector(int nf, const std::vector<int>& T, float weak, float strong)
    : weak_threshold(weak), num_features(nf), strong_threshold(strong), pyramid_levels((int)T.size()), T_at_level(T) {}

Detector::~Detector() {
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : kt_ev_)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

void Detector::ensure_device() {
    if (stream) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        throw Error(kENoDevice, "no HIP device is visible: the matching path runs only on the GPU");
    hip_check(hipGetDevice(&device), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : ev_) hip_check(hipEventCreate(&e), "hipEventCreate");
    hip_check(hipDeviceGetAttribute(&num_cu_, hipDeviceAttributeMultiprocessorCount, device), "attr");
}

void Detector::kt_begin(const std::string& name) {
    if (!timing_enabled) return;
    while ((int)kt_ev_.size() < 2 * (kt_used_ + 1)) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        kt_ev_.push_back(e);
    }
    if ((int)kt_names_.size() <= kt_used_) kt_names_.resize(kt_used_ + 1);
    kt_names_[kt_used_] = name;
    hip_check(hipEventRecord(kt_ev_[2 * kt_used_], stream), "event");
}

void Detector::kt_end() {
    if (!timing_enabled) return;
    hip_check(hipEventRecord(kt_ev_[2 * kt_used_ + 1], stream), "event");
    ++kt_used_;
}

// Called after the stream has been synchronized.
void Detector::kt_collect() {
    kernel_names.clear();
    kernel_ms.clear();
    for (int i = 0; i < kt_used_; ++i) {
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, kt_ev_[2 * i], kt_ev_[2 * i + 1]), "elapsed");
        kernel_names.push_back(kt_names_[i]);
        kernel_ms.push_back(ms);
    }
    kt_used_ = 0;
}

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
        // linear-memory offsets are 32-bit: 8*rows*cols plus the guard must stay below 2^31
        if (8LL * r * c + lm_guard_bytes(c / T, r / T) >= (1LL << 31))
            throw Error(kESize, "image too large for one match (more than 2^28 pixels)");
    }
}

LevelGeom Detector::geom(int l) const {
    LevelGeom g;
    g.rows = levels_[l].rows;
    g.cols = levels_[l].cols;
    g.T = T_at_level[l];
    g.W = g.cols / g.T;
    g.H = g.rows / g.T;
    g.hw = (long long)g.W * g.H;
    g.lm = levels_[l].lm.as<uint8_t>();
    return g;
}

void Detector::run_preprocess(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask,
                              bool device_input, int* zero_counters) {
    if (channels != 1 && channels != 3) throw Error(kEInval, "channels must be 1 or 3");
    if (stride < cols * channels) throw Error(kEInval, "stride smaller than a row");
    check_sizes(rows, cols);
    ensure_device();
    kt_used_ = 0;
    const int L = pyramid_levels;
    if ((int)levels_.size() != L) levels_.resize(L);
    int r = rows, c = cols;
    for (int l = 0; l < L; ++l) {
        if (l > 0) { r /= 2; c /= 2; }
        Level& lv = levels_[l];
        lv.rows = r;
        lv.cols = c;
        const int T = T_at_level[l];
        const long long W = c / T, H = r / T;
        const size_t lm_bytes = (size_t)8 * r * c + (size_t)lm_guard_bytes((int)W, (int)H);
        const bool grew = lm_bytes > lv.lm.bytes;
        lv.lm.ensure(lm_bytes);
        if (grew) hip_check(hipMemsetAsync(lv.lm.p, 0, lv.lm.bytes, stream), "memset lm");
        lv.q.ensure((size_t)r * c);
        if (l > 0 || !device_input) lv.img.ensure((size_t)r * c * channels);
        if (mask && (l > 0 || !device_input)) lv.mask.ensure((size_t)r * c);
    }
    const uint8_t* img0 = img;
    const uint8_t* mask0 = mask;
    int stride0 = stride;
    if (!device_input) {
        // Upload the image (and mask) through pinned staging.
        const size_t img_bytes = (size_t)rows * cols * channels;
        const size_t mask_bytes = mask ? (size_t)rows * cols : 0;
        h_stage_.ensure(img_bytes + mask_bytes);
        uint8_t* st = static_cast<uint8_t*>(h_stage_.p);
        hip_check(hipStreamSynchronize(stream), "sync before staging");
        for (int y = 0; y < rows; ++y)
            memcpy(st + (size_t)y * cols * channels, img + (size_t)y * stride, (size_t)cols * channels);
        if (mask) memcpy(st + img_bytes, mask, mask_bytes);
        hip_check(hipMemcpyAsync(levels_[0].img.p, st, img_bytes, hipMemcpyHostToDevice, stream), "H2D image");
        if (mask)
            hip_check(hipMemcpyAsync(levels_[0].mask.p, st + img_bytes, mask_bytes, hipMemcpyHostToDevice, stream), "H2D mask");
        img0 = levels_[0].img.as<uint8_t>();
        mask0 = mask ? levels_[0].mask.as<uint8_t>() : nullptr;
        stride0 = cols * channels;
    }
    if (timing_enabled) hip_check(hipEventRecord(ev_[0], stream), "event");
    const uint8_t* prev_img = img0;
    const uint8_t* prev_mask = mask0;
    int prev_stride = stride0;
    for (int l = 0; l < L; ++l) {
        Level& lv = levels_[l];
        const uint8_t* cur_img = img0;
        const uint8_t* cur_mask = mask0;
        int cur_stride = stride0;
        if (l > 0) {
            const Level& pv = levels_[l - 1];
            kt_begin("pyr_down_l" + std::to_string(l));
            hip_check(kern::pyr_down(prev_img, pv.rows, pv.cols, prev_stride, channels, lv.img.as<uint8_t>(),
                                     mask ? prev_mask : nullptr, mask ? lv.mask.as<uint8_t>() : nullptr, stream),
                      "pyr_down");
            kt_end();
            cur_img = lv.img.as<uint8_t>();
            cur_mask = mask ? lv.mask.as<uint8_t>() : nullptr;
            cur_stride = lv.cols * channels;
        }
        kt_begin("gradient_l" + std::to_string(l));
        hip_check(kern::gradient_quantize(cur_img, lv.rows, lv.cols, cur_stride, channels, cur_mask, weak_threshold,
                                          lv.q.as<uint8_t>(), nullptr, nullptr, stream,
                                          l == 0 ? (uint32_t*)zero_counters : nullptr, kCounterSlots),
                  "gradient");
        kt_end();
        kt_begin("spread_linearize_l" + std::to_string(l));
        hip_check(kern::spread_linearize(lv.q.as<uint8_t>(), lv.rows, lv.cols, T_at_level[l], lv.lm.as<uint8_t>(), stream),
                  "spread_linearize");
        kt_end();
        prev_img = cur_img;
        prev_mask = cur_mask;
        prev_stride = cur_stride;
    }
    if (timing_enabled) hip_check(hipEventRecord(ev_[1], stream), "event");
    cur_rows_ = rows;
    cur_cols_ = cols;
    cur_channels_ = channels;
    cur_device_input_ = device_input;
}

void Detector::ensure_bank() {
    if (!bank_dirty_) return;
    ensure_device();
    const int L = pyramid_levels;
    g_class_.clear();
    g_tid_.clear();
    g_names_.clear();
    h_tl_.clear();
    std::vector<uint2> feats;
    int ci = 0;
    for (const auto& kv : class_templates) {
        g_names_.push_back(kv.first);
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
                if (d.nfeat >= 8192) throw Error(kEFeatures, "feature size too large (line2Dup.cpp:1166, 1231)");
                for (const Feature& f : te.features) {
                    if (f.label < 0 || f.label > 7) throw Error(kEInval, "Invalid value of quantized parameter");
                    if (f.x < -32768 || f.x > 32767 || f.y < -32768 || f.y > 32767)
                        throw Error(kEInval, "feature coordinate outside the int16 range");
                    uint2 u;
                    u.x = (uint32_t)(uint16_t)(int16_t)f.x | ((uint32_t)(uint16_t)(int16_t)f.y << 16);
                    u.y = (uint32_t)f.label;
                    feats.push_back(u);
                }
                if (feats.size() > (size_t)INT_MAX) throw Error(kEInval, "bank too large");
                h_tl_.push_back(d);
            }
            g_class_.push_back(ci);
            g_tid_.push_back((int)t);
        }
        ++ci;
    }
    const int G = (int)g_class_.size();
    sorted_all_.resize(G);
    std::iota(sorted_all_.begin(), sorted_all_.end(), 0);
    std::stable_sort(sorted_all_.begin(), sorted_all_.end(), [&](int a, int b) {
        return h_tl_[(size_t)a * L + L - 1].nfeat > h_tl_[(size_t)b * L + L - 1].nfeat;
    });
    d_tl_.ensure(std::max<size_t>(h_tl_.size(), 1) * sizeof(TemplLevelDev));
    d_feats_.ensure(std::max<size_t>(feats.size(), 1) * sizeof(uint2) + 256);
    if (!h_tl_.empty())
        hip_check(hipMemcpyAsync(d_tl_.p, h_tl_.data(), h_tl_.size() * sizeof(TemplLevelDev), hipMemcpyHostToDevice, stream), "H2D tl");
    if (!feats.empty())
        hip_check(hipMemcpyAsync(d_feats_.p, feats.data(), feats.size() * sizeof(uint2), hipMemcpyHostToDevice, stream), "H2D feats");
    d_order_.ensure((size_t)std::max(G, 1) * sizeof(int));
    d_rawmin_c_.ensure((size_t)std::max(G, 1) * sizeof(int));
    d_rawmin_r_.ensure((size_t)std::max(G, 1) * L * sizeof(int));
    hip_check(hipStreamSynchronize(stream), "bank upload");
    thr_valid_ = false;
    sel_cached_.clear();
    bank_dirty_ = false;
}

size_t Detector::device_bytes() const {
    size_t s = d_tl_.bytes + d_feats_.bytes + d_order_.bytes + d_rawmin_c_.bytes + d_rawmin_r_.bytes + d_work_.bytes;
    for (const auto& lv : levels_) s += lv.img.bytes + lv.mask.bytes + lv.q.bytes + lv.lm.bytes;
    return s;
}

void Detector::thresholds(float threshold) {
    if (thr_valid_ && threshold == thr_cached_) return;
    const int L = pyramid_levels;
    const int G = (int)g_class_.size();
    std::vector<int> rc(G), rr((size_t)G * L, INT_MAX);
    for (int g = 0; g < G; ++g) {
        rc[g] = raw_min_coarse(h_tl_[(size_t)g * L + L - 1].nfeat, threshold);
        for (int l = 0; l < L - 1; ++l) rr[(size_t)g * L + l] = raw_min_refine(h_tl_[(size_t)g * L + l].nfeat, threshold);
    }
    if (G) {
        hip_check(hipMemcpyAsync(d_rawmin_c_.p, rc.data(), (size_t)G * sizeof(int), hipMemcpyHostToDevice, stream), "H2D rawmin");
        hip_check(hipMemcpyAsync(d_rawmin_r_.p, rr.data(), (size_t)G * L * sizeof(int), hipMemcpyHostToDevice, stream), "H2D rawmin");
        hip_check(hipStreamSynchronize(stream), "thresholds");
    }
    thr_cached_ = threshold;
    thr_valid_ = true;
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
    std::vector<char> take(g_names_.size(), 0);
    std::vector<int> cls_order;
    if (class_ids.empty()) {
        for (size_t i = 0; i < g_names_.size(); ++i) cls_order.push_back((int)i);
    } else {
        for (const auto& c : class_ids) {
            auto it = std::lower_bound(g_names_.begin(), g_names_.end(), c);
            if (it != g_names_.end() && *it == c) {
                const int i = (int)(it - g_names_.begin());
                if (!take[i]) cls_order.push_back(i);  // a repeated class yields only exact duplicates
                take[i] = 1;
            }
        }
    }
    std::fill(take.begin(), take.end(), 0);
    for (int c : cls_order) take[c] = 1;
    if (sorted) {
        for (int g : sorted_all_)
            if (take[g_class_[g]]) out.push_back(g);
    } else {
        for (int c : cls_order)
            for (int g = 0; g < (int)g_class_.size(); ++g)
                if (g_class_[g] == c) out.push_back(g);
    }
}

static bool match_less(const MatchRec& a, const MatchRec& b) {
    if (a.similarity != b.similarity) return a.similarity > b.similarity;  // Match::operator<, line2Dup.h:117-124
    if (a.template_id != b.template_id) return a.template_id < b.template_id;
    if (a.class_index != b.class_index) return a.class_index < b.class_index;
    if (a.y != b.y) return a.y < b.y;
    return a.x < b.x;
}

std::vector<MatchRec> Detector::match(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                                      const std::vector<std::string>& class_ids, const uint8_t* mask, int mode,
                                      bool device_input) {
    if (mode != kModePlain && mode != kModeBatched) throw Error(kEInval, "unknown match mode");
    check_sizes(rows, cols);
    ensure_device();
    ensure_bank();
    const int L = pyramid_levels;
    std::vector<int> sel;
    selection(class_ids, mode == kModeBatched, sel);
    thresholds(threshold);
    const int nsel = (int)sel.size();
    if (nsel > 0 && sel != sel_cached_) {  // the template order stays resident while the selection is unchanged
        d_order_.ensure(sel.size() * sizeof(int));
        hip_check(hipMemcpyAsync(d_order_.p, sel.data(), sel.size() * sizeof(int), hipMemcpyHostToDevice, stream),
                  "H2D order");
        sel_cached_ = sel;
    }
    ensure_work();
    // the first preprocessing launch also clears the queue head and the candidate counters
    run_preprocess(img, rows, cols, stride, channels, mask, device_input, counters());
    std::vector<MatchRec> out;
    if (nsel == 0) {
        hip_check(hipStreamSynchronize(stream), "sync");
        last_times = StageTimes();
        return out;
    }
    const int ppl = tuning.ppl;
    const int wpb = tuning.waves_per_block;
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (attempt > 0) {
            ensure_work();
            hip_check(kern::fill_u32((uint32_t*)counters(), 0u, kCounterSlots, stream), "zero counters");
        }
        int* counts = counters();  // [0] work queue, [1] coarse, [2..] refine levels
        Cand* final_buf = out_cands();
        Cand* cA = ping();
        Cand* cB = pong();

        const LevelGeom gL = geom(L - 1);
        if (timing_enabled) hip_check(hipEventRecord(ev_[2], stream), "event");
        kern::CoarseArgs ca;
        ca.geom = gL;
        ca.tl = d_tl_.as<TemplLevelDev>();
        ca.levels = L;
        ca.level = L - 1;
        ca.feats = d_feats_.as<uint2>();
        ca.ntiles = (int)((gL.hw + 64LL * ppl - 1) / (64LL * ppl));
        ca.raw_min = d_rawmin_c_.as<int>();
        ca.cand = L == 1 ? final_buf : cA;
        ca.cand_count = counts + 1;
        ca.cand_cap = (int)cand_cap_;
        kt_begin(mode == kModeBatched ? "coarse_batched" : "coarse_plain");
        if (mode == kModeBatched) {
            ca.order = d_order_.as<int>();
            ca.ntempl = nsel;
            ca.work_counter = counts;
            const long long items = (long long)nsel * ca.ntiles;
            long long blocks = (long long)std::max(num_cu_, 1) * tuning.blocks_per_cu;
            blocks = std::min(blocks, (items + wpb - 1) / wpb);
            hip_check(kern::coarse(ca, ppl, wpb, (int)std::max(1LL, blocks), stream), "coarse batched");
        } else {
            ca.ntempl = 1;
            ca.work_counter = nullptr;
            for (int k = 0; k < nsel; ++k) {  // one launch per template, as matchClass iterates them
                ca.order = d_order_.as<int>() + k;
                hip_check(kern::coarse(ca, ppl, wpb, 0, stream), "coarse plain");
            }
        }
        kt_end();
        if (timing_enabled) hip_check(hipEventRecord(ev_[3], stream), "event");

        const Cand* cur = ca.cand;
        int* cur_count = counts + 1;
        Cand* nxt = cB;
        for (int l = L - 2; l >= 0; --l) {
            kern::RefineArgs ra;
            ra.geom = geom(l);
            ra.tl = d_tl_.as<TemplLevelDev>();
            ra.levels = L;
            ra.level = l;
            ra.feats = d_feats_.as<uint2>();
            ra.cand_in = cur;
            ra.count_in = cur_count;
            ra.cap_in = (int)cand_cap_;
            ra.raw_min = d_rawmin_r_.as<int>();
            ra.cand_out = l == 0 ? final_buf : nxt;
            ra.count_out = counts + 2 + (L - 2 - l);
            ra.cap_out = (int)cand_cap_;
            const int blocks = std::max(num_cu_, 1) * 4;
            kt_begin("refine_l" + std::to_string(l));
            hip_check(kern::refine(ra, blocks, stream), "refine");
            kt_end();
            cur = ra.cand_out;
            cur_count = ra.count_out;
            nxt = (nxt == cB) ? cA : cB;
        }
        if (timing_enabled) hip_check(hipEventRecord(ev_[4], stream), "event");

        // counters and the first kPrefetchCands results come back in one copy
        const size_t head = kCounterSlots * sizeof(int) + std::min<size_t>(kPrefetchCands, cand_cap_) * sizeof(Cand);
        h_stage_.ensure(head);
        hip_check(hipMemcpyAsync(h_stage_.p, d_work_.p, head, hipMemcpyDeviceToHost, stream), "D2H results");
        hip_check(hipStreamSynchronize(stream), "sync results");
        const int* hc = static_cast<const int*>(h_stage_.p);
        const int coarse_n = hc[1];
        const int final_n = L >= 2 ? hc[2 + (L - 2)] : coarse_n;
        if ((size_t)coarse_n > cand_cap_) {  // overflow: grow and run the matching stages again
            while (cand_cap_ < (size_t)coarse_n) cand_cap_ *= 2;
            kt_used_ = 0;
            continue;
        }
        if (timing_enabled) {
            kt_collect();
            float a = 0, b = 0, c = 0;
            hip_check(hipEventElapsedTime(&a, ev_[0], ev_[1]), "elapsed");
            hip_check(hipEventElapsedTime(&b, ev_[2], ev_[3]), "elapsed");
            hip_check(hipEventElapsedTime(&c, ev_[3], ev_[4]), "elapsed");
            last_times.preprocess_ms = a;
            last_times.coarse_ms = b;
            last_times.refine_ms = c;
            last_times.total_ms = a + b + c;
        }
        last_times.coarse_candidates = coarse_n;
        last_times.final_candidates = final_n;
        std::vector<Cand> hcand(final_n);
        if (final_n) {
            const Cand* pre = reinterpret_cast<const Cand*>(static_cast<const char*>(h_stage_.p) + kCounterSlots * sizeof(int));
            if ((size_t)final_n <= kPrefetchCands) {
                memcpy(hcand.data(), pre, (size_t)final_n * sizeof(Cand));
            } else {
                h_stage_.ensure((size_t)final_n * sizeof(Cand));
                hip_check(hipMemcpyAsync(h_stage_.p, out_cands(), (size_t)final_n * sizeof(Cand), hipMemcpyDeviceToHost,
                                         stream), "D2H cand");
                hip_check(hipStreamSynchronize(stream), "sync cand");
                memcpy(hcand.data(), h_stage_.p, (size_t)final_n * sizeof(Cand));
            }
        }
        out.reserve(final_n);
        for (const Cand& cd : hcand) {
            MatchRec m;
            m.x = cd.x;
            m.y = cd.y;
            const int n0 = h_tl_[(size_t)cd.g * L + 0].nfeat;
            if (L == 1) m.similarity = score_of(cd.raw, n0);
            else m.similarity = cd.raw > 0 ? score_of(cd.raw, n0) : 0.f;
            m.class_index = g_class_[cd.g];
            m.template_id = g_tid_[cd.g];
            out.push_back(m);
        }
        std::sort(out.begin(), out.end(), match_less);
        auto eq = [](const MatchRec& a, const MatchRec& b) {  // Match::operator==, line2Dup.h:126-129
            return a.x == b.x && a.y == b.y && a.similarity == b.similarity && a.class_index == b.class_index;
        };
        out.erase(std::unique(out.begin(), out.end(), eq), out.end());
        return out;
    }
    throw Error(kEBudget, "candidate buffer could not be grown enough");
}

// Tile-shape search for the batched coarse kernel on a prepared workload: every candidate shape must
// return the same matches; the fastest (HIP-event time of the coarse launch) is kept in `tuning`.
Tuning Detector::autotune(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                          const std::vector<std::string>& class_ids, bool device_input, int iters,
                          std::vector<std::pair<Tuning, float>>* trace) {
    const bool was_timing = timing_enabled;
    const Tuning saved = tuning;
    timing_enabled = true;
    std::vector<MatchRec> ref;
    Tuning best = saved;
    float best_ms = 1e30f;
    bool first = true;
    try {
        for (int ppl : {4, 8, 16})
            for (int wpb : {1, 2, 4})
                for (int bpc : {2, 4, 8, 16}) {
                    tuning.ppl = ppl;
                    tuning.waves_per_block = wpb;
                    tuning.blocks_per_cu = bpc;
                    float tot = 0.f;
                    for (int it = 0; it < std::max(1, iters); ++it) {
                        auto m = match(img, rows, cols, stride, channels, threshold, class_ids, nullptr, kModeBatched,
                                       device_input);
                        if (first) { ref = m; first = false; }
                        else if (m.size() != ref.size() ||
                                 !std::equal(m.begin(), m.end(), ref.begin(), [](const MatchRec& a, const MatchRec& b) {
                                     return a.x == b.x && a.y == b.y && a.similarity == b.similarity &&
                                            a.class_index == b.class_index && a.template_id == b.template_id;
                                 }))
                            throw Error(kEInval, "tile shapes disagree: autotune aborted");
                        for (size_t k = 0; k < kernel_names.size(); ++k)
                            if (kernel_names[k] == "coarse_batched") tot += kernel_ms[k];
                    }
                    const float ms = tot / (float)std::max(1, iters);
                    if (trace) trace->push_back({tuning, ms});
                    if (ms < best_ms) { best_ms = ms; best = tuning; }
                }
    } catch (...) {
        tuning = saved;
        timing_enabled = was_timing;
        throw;
    }
    tuning = best;
    timing_enabled = was_timing;
    return best;
}

// Work memory: [counters][final candidates][ping][pong], each candidate area cand_cap_ entries.
void Detector::ensure_work() {
    if (cand_cap_ == 0) cand_cap_ = 1 << 16;
    const size_t need = kCounterSlots * sizeof(int) + 3 * cand_cap_ * sizeof(Cand);
    if (need > d_work_.bytes) {
        d_work_.ensure(need);
        hip_check(hipMemsetAsync(d_work_.p, 0, kCounterSlots * sizeof(int), stream), "memset counters");
    }
}
int* Detector::counters() const { return d_work_.as<int>(); }
Cand* Detector::out_cands() const { return reinterpret_cast<Cand*>(d_work_.as<char>() + kCounterSlots * sizeof(int)); }
Cand* Detector::ping() const { return out_cands() + cand_cap_; }
Cand* Detector::pong() const { return out_cands() + 2 * cand_cap_; }

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
    if (channels != 1 && channels != 3) throw Error(kEInval, "channels must be 1 or 3");
    DevMem dimg, dq, dmag, dang;
    dimg.ensure((size_t)rows * cols * channels);
    dq.ensure((size_t)rows * cols);
    dmag.ensure((size_t)rows * cols * 4);
    dang.ensure((size_t)rows * cols * 4);
    hip_check(hipMemcpy2DAsync(dimg.p, (size_t)cols * channels, img, (size_t)stride, (size_t)cols * channels, rows,
                               hipMemcpyHostToDevice, stream), "H2D");
    hip_check(kern::gradient_quantize(dimg.as<uint8_t>(), rows, cols, cols * channels, channels, nullptr, weak_threshold,
                                      dq.as<uint8_t>(), dmag.as<float>(), dang.as<float>(), stream), "gradient");
    hip_check(hipMemcpyAsync(q, dq.p, (size_t)rows * cols, hipMemcpyDeviceToHost, stream), "D2H");
    if (mag) hip_check(hipMemcpyAsync(mag, dmag.p, (size_t)rows * cols * 4, hipMemcpyDeviceToHost, stream), "D2H");
    if (angle) hip_check(hipMemcpyAsync(angle, dang.p, (size_t)rows * cols * 4, hipMemcpyDeviceToHost, stream), "D2H");
    hip_check(hipStreamSynchronize(stream), "sync");
}

// Detector::addTemplate: gradients on the GPU, candidate selection on the host.
int Detector::add_template(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask,
                           const std::string& class_id, int nf_override) {
    std::vector<TemplatePyramid>& tps = class_templates[class_id];  // created even on failure, as the reference
    const int template_id = (int)tps.size();
    ensure_device();
    if (channels != 1 && channels != 3) throw Error(kEInval, "channels must be 1 or 3");
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
            hip_check(hipMemcpyAsync(dsrc.p, cur.data(), cur.size(), hipMemcpyHostToDevice, stream), "H2D");
            if (mask) {
                dm.ensure(curmask.size());
                dmd.ensure((size_t)(r / 2) * (c / 2));
                hip_check(hipMemcpyAsync(dm.p, curmask.data(), curmask.size(), hipMemcpyHostToDevice, stream), "H2D");
            }
            hip_check(kern::pyr_down(dsrc.as<uint8_t>(), r, c, c * channels, channels, ddst.as<uint8_t>(),
                                     mask ? dm.as<uint8_t>() : nullptr, mask ? dmd.as<uint8_t>() : nullptr, stream),
                      "pyr_down");
            r /= 2;
            c /= 2;
            cur.resize((size_t)r * c * channels);
            hip_check(hipMemcpyAsync(cur.data(), ddst.p, cur.size(), hipMemcpyDeviceToHost, stream), "D2H");
            if (mask) {
                curmask.resize((size_t)r * c);
                hip_check(hipMemcpyAsync(curmask.data(), dmd.p, curmask.size(), hipMemcpyDeviceToHost, stream), "D2H");
            }
            hip_check(hipStreamSynchronize(stream), "sync");
        }
        std::vector<uint8_t> q((size_t)r * c);
        std::vector<float> mag((size_t)r * c), ang((size_t)r * c);
        gradient(cur.data(), r, c, c * channels, channels, q.data(), mag.data(), ang.data());
        if (!extract_template(mag.data(), q.data(), ang.data(), mask ? curmask.data() : nullptr, r, c, strong_threshold,
                              nf, tp[l].features))
            return -1;
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
// Debug / parity helpers: intermediate buffers of the device pipeline
// ---------------------------------------------------------------------------------------------
void Detector::prepare_image(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask) {
    run_preprocess(img, rows, cols, stride, channels, mask, false, nullptr);
    hip_check(hipStreamSynchronize(stream), "sync");
}

// what: 0 = quantized (masked) orientations, 1 = linear memories [8][T*T][H*W], 2 = level image
void Detector::read_level(int level, int what, uint8_t* out) {
    if (level < 0 || level >= (int)levels_.size() || cur_rows_ == 0) throw Error(kEInval, "no prepared level");
    const Level& lv = levels_[level];
    size_t n = 0;
    const void* src = nullptr;
    if (what == 0) { n = (size_t)lv.rows * lv.cols; src = lv.q.p; }
    else if (what == 1) { n = (size_t)lv.rows * lv.cols * 8; src = lv.lm.p; }
    else if (what == 2) {
        if (level == 0 && cur_device_input_) throw Error(kEInval, "level 0 image is the caller's device buffer");
        n = (size_t)lv.rows * lv.cols * cur_channels_;
        src = lv.img.p;
    }
    else throw Error(kEInval, "unknown buffer");
    hip_check(hipMemcpyAsync(out, src, n, hipMemcpyDeviceToHost, stream), "D2H");
    hip_check(hipStreamSynchronize(stream), "sync");
}

static int global_index(const std::vector<std::string>& names, const std::vector<int>& gc, const std::vector<int>& gt,
                        const std::string& cls, int tid) {
    auto it = std::lower_bound(names.begin(), names.end(), cls);
    if (it == names.end() || *it != cls) throw Error(kENotFound, "class not found: " + cls);
    const int ci = (int)(it - names.begin());
    for (size_t g = 0; g < gc.size(); ++g)
        if (gc[g] == ci && gt[g] == tid) return (int)g;
    throw Error(kENotFound, "template_id out of range");
}

void Detector::similarity_map(const std::string& class_id, int template_id, uint16_t* out) {
    ensure_bank();
    if (cur_rows_ == 0) throw Error(kEInval, "no prepared image");
    const int L = pyramid_levels;
    const int g = global_index(g_names_, g_class_, g_tid_, class_id, template_id);
    const TemplLevelDev& t = h_tl_[(size_t)g * L + L - 1];
    const LevelGeom gl = geom(L - 1);
    DevMem d;
    d.ensure((size_t)gl.hw * 2);
    hip_check(kern::similarity_map(gl, nullptr, d_feats_.as<uint2>(), t.width, t.height, t.nfeat, t.off, d.as<uint16_t>(), stream), "simmap");
    hip_check(hipMemcpyAsync(out, d.p, (size_t)gl.hw * 2, hipMemcpyDeviceToHost, stream), "D2H");
    hip_check(hipStreamSynchronize(stream), "sync");
}

void Detector::similarity_local(const std::string& class_id, int template_id, int level, int cx, int cy, uint16_t* out) {
    ensure_bank();
    if (cur_rows_ == 0) throw Error(kEInval, "no prepared image");
    const int L = pyramid_levels;
    if (level < 0 || level >= L) throw Error(kEInval, "level out of range");
    const int g = global_index(g_names_, g_class_, g_tid_, class_id, template_id);
    const TemplLevelDev& t = h_tl_[(size_t)g * L + level];
    DevMem d;
    d.ensure(512);
    hip_check(kern::similarity_local(geom(level), d_feats_.as<uint2>(), t.nfeat, t.off, cx, cy, d.as<uint16_t>(), stream), "simlocal");
    hip_check(hipMemcpyAsync(out, d.p, 512, hipMemcpyDeviceToHost, stream), "D2H");
    hip_check(hipStreamSynchronize(stream), "sync");
}

void Detector::bin_table(int lo, int hi, uint8_t* out) {
    ensure_device();
    const size_t n = (size_t)(hi - lo + 1) * (hi - lo + 1);
    DevMem d;
    d.ensure(n);
    hip_check(kern::bin_table(lo, hi, d.as<uint8_t>(), stream), "bin_table");
    hip_check(hipMemcpyAsync(out, d.p, n, hipMemcpyDeviceToHost, stream), "D2H");
    hip_check(hipStreamSynchronize(stream), "sync");
}

}  // namespace sbm
