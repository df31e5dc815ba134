// The per-image device pipeline of Detector::match (line2Dup.cpp:1049-1268): preprocessing of every
// pyramid level, the coarse gather-accumulate with threshold compaction, the refinement levels, and
// the return of the compacted candidates -- launched eagerly, or replayed from a captured graph.
#include <algorithm>
#include <climits>
#include <cstring>
#include <unordered_map>

#include "sbm_detector.h"
#include "sbm_reach.h"

#pragma clang fp contract(off)

namespace sbm {

namespace {
bool match_less(const MatchRec& a, const MatchRec& b) {
    if (a.similarity != b.similarity) return a.similarity > b.similarity;  // Match::operator<, line2Dup.h:117-124
    if (a.template_id != b.template_id) return a.template_id < b.template_id;
    if (a.class_index != b.class_index) return a.class_index < b.class_index;
    if (a.y != b.y) return a.y < b.y;
    return a.x < b.x;
}

uint64_t sel_signature(const std::vector<int>& sel) {
    uint64_t h = 1469598103934665603ull ^ (uint64_t)sel.size();
    for (int v : sel) h = (h ^ (uint64_t)(uint32_t)v) * 1099511628211ull;
    return h;
}

uint32_t float_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
}  // namespace

// score = (raw * 100.f) / (4 * n)   (line2Dup.cpp:1177, 1244)
float score_of(int raw, int n) {
    const float num = (float)raw * 100.f;
    return num / (float)(4 * n);
}

// Smallest raw in [0, 4n] whose coarse score is > thr (line2Dup.cpp:1179); INT_MAX if none.
int raw_min_coarse(int n, float thr) {
    if (n <= 0) return INT_MAX;  // 0/0 is NaN and never passes
    int lo = 0, hi = 4 * n + 1;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (score_of(mid, n) > thr) hi = mid;
        else lo = mid + 1;
    }
    return lo <= 4 * n ? lo : INT_MAX;
}

// Smallest raw in [0, 4n] whose refined similarity is kept, i.e. !(sim < thr) (line2Dup.cpp:1261-1263).
// A window with no positive score keeps best_score = 0 (line2Dup.cpp:1236-1253).
int raw_min_refine(int n, float thr) {
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

Pipeline::Pipeline(int device, int num_cu) : device_(device), num_cu_(std::max(num_cu, 1)) {
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : ev_) hip_check(hipEventCreate(&e), "hipEventCreate");
    hip_check(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "hipEventCreate");
}

Pipeline::~Pipeline() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    clear_graphs();
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : kt_ev_)
        if (e) (void)hipEventDestroy(e);
    if (done_) (void)hipEventDestroy(done_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Pipeline::clear_graphs() {
    for (auto& g : graphs_) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    graphs_.clear();
}

void Pipeline::kt_begin(const std::string& name) {
    if (!timing_enabled) return;
    while ((int)kt_ev_.size() < 2 * (kt_used_ + 1)) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        kt_ev_.push_back(e);
    }
    if ((int)kt_names_.size() <= kt_used_) kt_names_.resize(kt_used_ + 1);
    kt_names_[kt_used_] = name;
    hip_check(hipEventRecord(kt_ev_[2 * kt_used_], stream_), "event");
}

void Pipeline::kt_end() {
    if (!timing_enabled) return;
    hip_check(hipEventRecord(kt_ev_[2 * kt_used_ + 1], stream_), "event");
    ++kt_used_;
}

// Called after the stream has been synchronized.
void Pipeline::kt_collect() {
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

int* Pipeline::counters() const { return d_work_.as<int>(); }
Cand* Pipeline::out_cands() const { return reinterpret_cast<Cand*>(d_work_.as<char>() + kCounterSlots * sizeof(int)); }
Cand* Pipeline::ping() const { return out_cands() + cand_cap_; }
Cand* Pipeline::pong() const { return out_cands() + 2 * cand_cap_; }

LevelGeom Pipeline::geom(const Bank& bank, int l) const {
    LevelGeom g{};
    if (l < 0 || l >= (int)levels_.size() || cur_rows_ == 0) return g;  // nothing prepared: g.lm == nullptr
    g.rows = levels_[l].rows;
    g.cols = levels_[l].cols;
    g.T = bank.T[l];
    g.W = g.cols / g.T;
    g.H = g.rows / g.T;
    g.hw = (long long)g.W * g.H;
    g.lm = levels_[l].lm.as<uint8_t>();
    g.planes = bank.bins;
    g.lm_bytes = (int)lm_level_bytes(g.planes, g.rows, g.cols, g.T);
    g.lm_nib = levels_[l].lm_nib.p ? levels_[l].lm_nib.as<uint8_t>() : nullptr;
    g.nib_bytes = (int)(levels_[l].nib_b ? std::min<long long>(levels_[l].nib_b, levels_[l].lm_nib.bytes)
                                         : (long long)levels_[l].lm_nib.bytes);
    g.img_stride = levels_[l].lm_b;
    g.nib_stride = levels_[l].nib_b;
    return g;
}

namespace {
size_t a256(size_t v) { return (v + 255) & ~(size_t)255; }
}  // namespace

// Bytes of level l's banded copy (rows x cols at that level) when the refinement reads one: refined levels
// with refine_layout 1 or 3 (the banded copy) or 4 (the nibble-banded copy) whose copy stays below 2^31 bytes
// (32-bit offsets); else 0.
size_t Pipeline::band_bytes(const Bank& bank, int l, int rows, int cols) const {
    const int lay = tuning.refine_layout;
    if ((lay != 1 && lay != 3 && lay != 4) || l >= bank.L - 1) return 0;
    const int T = bank.T[l];
    const BandGeom bg = lay == 4 ? nib_band_geom(bank.bins, T, cols / T, rows / T) : band_geom(bank.bins, T, cols / T, rows / T);
    return bg.bytes < (1LL << 31) ? (size_t)bg.bytes : 0;
}

size_t Pipeline::bytes_needed(const Bank& bank, int rows, int cols, int channels, bool mask, size_t cand_cap) {
    size_t s = 0;
    int r = rows, c = cols;
    const size_t P = (size_t)bank.bins, qb = P / 8;  // planes; bytes per quantized pixel
    for (int l = 0; l < bank.L; ++l) {
        if (l > 0) { r /= 2; c /= 2; }
        const int T = bank.T[l];
        const size_t px = (size_t)r * c;
        s += a256(px * channels) + a256(px * qb) + (mask ? a256(px) : 0);          // image, q, mask
        const size_t lm1 = (size_t)lm_level_bytes((int)P, r, c, T);                 // linear memories
        s += a256(lm1);
        if (l == bank.L - 1)
            s += a256((size_t)lm_nibble_bytes((long long)lm1)) + a256(sparse_rows_bytes(r, c, T, (int)P)) +
                 a256(sparse_entries_bytes(r, c));
    }
    s += a256((size_t)rows * cols * channels) + (mask ? a256((size_t)rows * cols) : 0);  // host-input buffer
    s += kCounterSlots * sizeof(int) + 3 * std::max(cand_cap, kInitialCandCap) * sizeof(Cand) + 1024 * sizeof(int);
    s += (size_t)bank.size() * sizeof(int) * (2 + bank.L);                      // order, thresholds
    s += 256;                                                                   // input table
    return s;
}

// Device bytes this pipeline would hold after running a launch of nimg images of this size: every buffer only
// grows, so each one ends at the larger of what it holds and what the launch needs (sizes as ensure_levels
// makes them).
size_t Pipeline::bytes_after(const Bank& bank, int rows, int cols, int channels, bool mask, int nimg) const {
    const size_t n = (size_t)std::max(nimg, 1);
    auto grown = [](size_t have, size_t want) { return want > have ? std::max<size_t>(want, 256) : have; };
    const bool same_levels = (int)levels_.size() == bank.L;
    size_t s = 0;
    int r = rows, c = cols;
    for (int l = 0; l < bank.L; ++l) {
        if (l > 0) { r /= 2; c /= 2; }
        const Level* lv = same_levels ? &levels_[l] : nullptr;
        const int T = bank.T[l];
        const size_t px = (size_t)r * c;
        const size_t lmb = (size_t)lm_level_bytes(bank.bins, r, c, T);
        s += grown(lv ? lv->lm.bytes : 0, n * a256(lmb));
        s += l == bank.L - 1 ? grown(lv ? lv->lm_nib.bytes : 0, n * a256((size_t)lm_nibble_bytes((long long)lmb)))
                             : (lv ? lv->lm_nib.bytes : 0);
        s += l == bank.L - 1 ? grown(lv ? lv->sp_rows.bytes : 0, n * a256(sparse_rows_bytes(r, c, T, bank.bins))) +
                                   grown(lv ? lv->sp_entries.bytes : 0, n * a256(sparse_entries_bytes(r, c)))
                             : (lv ? lv->sp_rows.bytes + lv->sp_entries.bytes : 0);
        const size_t bb = band_bytes(bank, l, r, c);
        s += bb ? grown(lv ? lv->band.bytes : 0, n * a256(bb)) : (lv ? lv->band.bytes : 0);
        s += grown(lv ? lv->q.bytes : 0, n * a256(px * (size_t)(bank.bins / 8)));
        s += grown(lv ? lv->img.bytes : 0, n * a256(px * channels));
        s += mask ? grown(lv ? lv->mask.bytes : 0, n * a256(px)) : (lv ? lv->mask.bytes : 0);
    }
    const size_t px0 = (size_t)rows * cols;
    s += grown(d_in_.bytes, n * a256(px0 * channels));
    s += mask ? grown(d_in_mask_.bytes, n * a256(px0)) : d_in_mask_.bytes;
    s += grown(d_work_.bytes, kCounterSlots * sizeof(int) + 3 * std::max(cand_cap_, kInitialCandCap) * sizeof(Cand));
    s += grown(d_bins_.bytes, 1024 * sizeof(int));
    s += grown(d_src_.bytes, sizeof(SrcTable));
    const size_t G = (size_t)std::max(bank.size(), 1);
    s += grown(d_order_.bytes, G * sizeof(int)) + grown(d_rawmin_c_.bytes, G * sizeof(int)) +
         grown(d_rawmin_r_.bytes, G * bank.L * sizeof(int));
    return s;
}

size_t Pipeline::device_bytes() const {
    size_t s = d_work_.bytes + d_order_.bytes + d_rawmin_c_.bytes + d_rawmin_r_.bytes + d_in_.bytes + d_in_mask_.bytes +
               d_bins_.bytes + d_src_.bytes;
    for (const auto& lv : levels_)
        s += lv.img.bytes + lv.mask.bytes + lv.q.bytes + lv.lm.bytes + lv.lm_nib.bytes + lv.sp_rows.bytes +
             lv.sp_entries.bytes + lv.band.bytes;
    return s;
}

// Allocates the level buffers for nimg images of this size (each image's block 256-byte aligned, at a fixed
// stride); returns true when any buffer moved (cached graphs refer to the old addresses then).
bool Pipeline::ensure_levels(const Bank& bank, int rows, int cols, int channels, bool mask, int nimg) {
    const int L = bank.L;
    bool moved = false;
    if ((int)levels_.size() != L) {
        levels_.clear();
        levels_.resize(L);
        moved = true;
    }
    auto a256 = [](size_t v) { return (long long)((v + 255) & ~(size_t)255); };
    const size_t n = (size_t)std::max(nimg, 1);
    int r = rows, c = cols;
    for (int l = 0; l < L; ++l) {
        if (l > 0) { r /= 2; c /= 2; }
        Level& lv = levels_[l];
        lv.rows = r;
        lv.cols = c;
        lv.nimg = (int)n;
        const int T = bank.T[l];
        lv.T = T;
        const size_t px = (size_t)r * c;
        const size_t lm1 = (size_t)lm_level_bytes(bank.bins, r, c, T);
        lv.lm_b = a256(lm1);
        if (n * lv.lm_b > lv.lm.bytes) {
            lv.lm.ensure(n * lv.lm_b);
            hip_check(hipMemsetAsync(lv.lm.p, 0, lv.lm.bytes, stream_), "memset lm");  // zero guards
            moved = true;
        }
        if (l == L - 1) {
            lv.nib_b = a256((size_t)lm_nibble_bytes((long long)lm1));
            if (n * lv.nib_b > lv.lm_nib.bytes) {
                lv.lm_nib.ensure(n * lv.nib_b);
                hip_check(hipMemsetAsync(lv.lm_nib.p, 0, lv.lm_nib.bytes, stream_), "memset lm_nib");
                moved = true;
            }
            lv.rows_b = a256(sparse_rows_bytes(r, c, T, bank.bins)) / (long long)sizeof(int);
            lv.ent_b = a256(sparse_entries_bytes(r, c)) / (long long)sizeof(uint32_t);
            if (n * lv.rows_b * sizeof(int) > lv.sp_rows.bytes) { lv.sp_rows.ensure(n * lv.rows_b * sizeof(int)); moved = true; }
            if (n * lv.ent_b * sizeof(uint32_t) > lv.sp_entries.bytes) {
                lv.sp_entries.ensure(n * lv.ent_b * sizeof(uint32_t));
                moved = true;
            }
        } else {
            lv.nib_b = lv.rows_b = lv.ent_b = 0;
        }
        const size_t bb = band_bytes(bank, l, r, c);
        lv.band_b = bb ? a256(bb) : 0;
        lv.band_nib = tuning.refine_layout == 4;
        if (n * lv.band_b > lv.band.bytes) {
            lv.band.ensure(n * lv.band_b);
            hip_check(hipMemsetAsync(lv.band.p, 0, lv.band.bytes, stream_), "memset band");  // zeroed blocks
            moved = true;
        }
        lv.q_b = a256(px * (size_t)(bank.bins / 8));  // 16-bit quantized values with 16 orientations
        if (n * lv.q_b > lv.q.bytes) { lv.q.ensure(n * lv.q_b); moved = true; }
        lv.img_b = a256(px * channels);
        if (n * lv.img_b > lv.img.bytes) { lv.img.ensure(n * lv.img_b); moved = true; }
        lv.mask_b = a256(px);
        if (mask && n * lv.mask_b > lv.mask.bytes) { lv.mask.ensure(n * lv.mask_b); moved = true; }
    }
    return moved;
}

// Work memory: [counters][final candidates][ping][pong], each candidate area cand_cap_ entries.
void Pipeline::ensure_work() {
    if (cand_cap_ == 0) cand_cap_ = kInitialCandCap;
    if (!d_bins_.p) d_bins_.ensure(1024 * sizeof(int));  // candidate row bins
    const size_t need = kCounterSlots * sizeof(int) + 3 * cand_cap_ * sizeof(Cand);
    if (need > d_work_.bytes) {
        d_work_.ensure(need);
        hip_check(hipMemsetAsync(d_work_.p, 0, kCounterSlots * sizeof(int), stream_), "memset counters");
        clear_graphs();
    }
}

// Integer thresholds for this threshold and the template order for this selection stay resident
// while they do not change.
void Pipeline::ensure_static(const Bank& bank, float thr, const std::vector<int>& sel) {
    const int L = bank.L, G = bank.size();
    const bool new_bank = static_bank_ != bank.version;
    if (new_bank || !static_valid_ || float_bits(thr) != float_bits(static_thr_)) {
        std::vector<int> rc(std::max(G, 1)), rr((size_t)std::max(G, 1) * L, INT_MAX);
        for (int g = 0; g < G; ++g) {
            rc[g] = raw_min_coarse(bank.h_tl[(size_t)g * L + L - 1].nfeat, thr);
            for (int l = 0; l < L - 1; ++l) rr[(size_t)g * L + l] = raw_min_refine(bank.h_tl[(size_t)g * L + l].nfeat, thr);
        }
        const bool moved = rc.size() * sizeof(int) > d_rawmin_c_.bytes || rr.size() * sizeof(int) > d_rawmin_r_.bytes;
        d_rawmin_c_.ensure(rc.size() * sizeof(int));
        d_rawmin_r_.ensure(rr.size() * sizeof(int));
        if (moved) clear_graphs();
        hip_check(hipMemcpyAsync(d_rawmin_c_.p, rc.data(), rc.size() * sizeof(int), hipMemcpyHostToDevice, stream_), "H2D");
        hip_check(hipMemcpyAsync(d_rawmin_r_.p, rr.data(), rr.size() * sizeof(int), hipMemcpyHostToDevice, stream_), "H2D");
        hip_check(hipStreamSynchronize(stream_), "sync thresholds");
        static_thr_ = thr;
    }
    if (new_bank || !static_valid_ || sel != static_sel_) {
        if (sel.size() * sizeof(int) > d_order_.bytes) {
            d_order_.ensure(std::max<size_t>(sel.size(), 1) * sizeof(int));
            clear_graphs();
        }
        if (!sel.empty()) {
            hip_check(hipMemcpyAsync(d_order_.p, sel.data(), sel.size() * sizeof(int), hipMemcpyHostToDevice, stream_),
                      "H2D order");
            hip_check(hipStreamSynchronize(stream_), "sync order");
        }
        static_sel_ = sel;
        coarse_features_ = 0;
        for (int g : sel) coarse_features_ += bank.h_tl[(size_t)g * L + L - 1].nfeat;
    }
    static_bank_ = bank.version;
    static_valid_ = true;
}

// Per-level preprocessing of nimg images; the first launch also clears the match counters.  With `src`
// (device memory) the level-0 pixels, strides and masks come from that table; otherwise im is the one image.
void Pipeline::enqueue_preprocess(const Bank& bank, const ImageArgs& im, int nimg, const SrcTable* src) {
    const int L = bank.L, channels = im.channels;
    const bool has_mask = im.mask != nullptr;
    const uint8_t* img0 = im.img;
    const uint8_t* mask0 = im.mask;
    int stride0 = im.stride;
    if (!src && !im.device) {
        const size_t img_bytes = (size_t)im.rows * im.cols * channels;
        const size_t mask_bytes = im.mask ? (size_t)im.rows * im.cols : 0;
        h_stage_.ensure(img_bytes + mask_bytes);
        uint8_t* st = static_cast<uint8_t*>(h_stage_.p);
        hip_check(hipStreamSynchronize(stream_), "sync before staging");
        for (int y = 0; y < im.rows; ++y)
            memcpy(st + (size_t)y * im.cols * channels, im.img + (size_t)y * im.stride, (size_t)im.cols * channels);
        if (im.mask) memcpy(st + img_bytes, im.mask, mask_bytes);
        hip_check(hipMemcpyAsync(levels_[0].img.p, st, img_bytes, hipMemcpyHostToDevice, stream_), "H2D image");
        if (im.mask)
            hip_check(hipMemcpyAsync(levels_[0].mask.p, st + img_bytes, mask_bytes, hipMemcpyHostToDevice, stream_),
                      "H2D mask");
        img0 = levels_[0].img.as<uint8_t>();
        mask0 = im.mask ? levels_[0].mask.as<uint8_t>() : nullptr;
        stride0 = im.cols * channels;
    }
    if (timing_enabled) hip_check(hipEventRecord(ev_[0], stream_), "event");
    for (int l = 0; l < L; ++l) {
        Level& lv = levels_[l];
        const uint8_t* cur_img = img0;
        const uint8_t* cur_mask = mask0;
        int cur_stride = stride0;
        if (l > 0) {  // written by the previous level's gradient kernel (pyrDown fused into it)
            cur_img = lv.img.as<uint8_t>();
            cur_mask = has_mask ? lv.mask.as<uint8_t>() : nullptr;
            cur_stride = lv.cols * channels;
        }
        BatchStrides gb;
        if (l == 0) gb.src = src;
        else { gb.in_b = lv.img_b; gb.in_mask_b = lv.mask_b; }
        gb.out_b = lv.q_b;
        uint8_t* down = nullptr;
        uint8_t* down_mask = nullptr;
        if (l + 1 < L) {  // the next level's image (and mask)
            Level& nv = levels_[l + 1];
            down = nv.img.as<uint8_t>();
            gb.down_b = nv.img_b;
            if (has_mask) {
                down_mask = nv.mask.as<uint8_t>();
                gb.down_mask_b = nv.mask_b;
            }
        }
        kt_begin("gradient_l" + std::to_string(l));
        hip_check(kern::gradient_quantize(cur_img, lv.rows, lv.cols, cur_stride, channels, cur_mask, bank.weak,
                                          lv.q.as<uint8_t>(), nullptr, nullptr, stream_,
                                          l == 0 ? (uint32_t*)counters() : nullptr, kCounterSlots, bank.bins, nimg, gb,
                                          down, down_mask),
                  "gradient");
        kt_end();
        kt_begin("spread_linearize_l" + std::to_string(l));
        const bool want_nib = l == L - 1 && tuning.gather == 3;  // the coarse level, two positions per byte
        bool nib_done = false;
        BatchStrides sb;
        sb.in_b = lv.q_b;
        sb.out_b = lv.lm_b;
        sb.nib_b = lv.nib_b;
        if (bank.bins == 16)
            hip_check(kern::spread_linearize16(lv.q.as<uint16_t>(), lv.rows, lv.cols, bank.T[l], lv.lm.as<uint8_t>(),
                                               stream_, nimg, sb),
                      "spread_linearize16");
        else
            hip_check(kern::spread_linearize(lv.q.as<uint8_t>(), lv.rows, lv.cols, bank.T[l], lv.lm.as<uint8_t>(),
                                             stream_, want_nib ? lv.lm_nib.as<uint8_t>() : nullptr, &nib_done, nimg, sb),
                      "spread_linearize");
        kt_end();
        if (l == L - 1) {
            nib_valid_ = want_nib;
            nib_T_ = bank.T[l];
        }
        if (lv.band_b > 0) {
            kt_begin("band_l" + std::to_string(l));
            const auto copy = lv.band_nib ? kern::nib_band_copy : kern::band_copy;
            hip_check(copy(lv.lm.as<uint8_t>(), lm_level_bytes(bank.bins, lv.rows, lv.cols, bank.T[l]), bank.bins,
                           bank.T[l], lv.cols / bank.T[l], lv.rows / bank.T[l], lv.band.as<uint8_t>(), stream_, nimg,
                           lv.lm_b, lv.band_b),
                      "band copy");
            kt_end();
        }
        if (want_nib && !nib_done) {
            kt_begin("pack_nibbles_l" + std::to_string(l));
            const int T = bank.T[l];
            const long long lmb = lm_level_bytes(bank.bins, lv.rows, lv.cols, T);
            BatchStrides nb;
            nb.in_b = lv.lm_b;
            nb.out_b = lv.nib_b;
            hip_check(kern::pack_nibbles(lv.lm.as<uint8_t>(), lmb, lv.lm_nib.as<uint8_t>(), stream_, nimg, nb),
                      "pack_nibbles");
            kt_end();
        }
    }
    if (timing_enabled) hip_check(hipEventRecord(ev_[1], stream_), "event");
    cur_rows_ = im.rows;
    cur_cols_ = im.cols;
    cur_channels_ = channels;
    cur_bins_ = bank.bins;
    cur_device_input_ = im.device || src != nullptr;
}

// Coarse stage over the selection in d_order_ and the refinement levels; results land in out_cands().
void Pipeline::enqueue_match(const Bank& bank, int nsel, int mode, int nimg) {
    if (mode != kModeBatched && nimg != 1) throw Error(kEInval, "the plain path matches one image at a time");
    const int L = bank.L;
    const int ppl = tuning.ppl, wpb = tuning.waves_per_block;
    int* counts = counters();  // [0] work queue, [1] coarse, [2..] refine levels
    Cand* final_buf = out_cands();
    Cand* cA = ping();
    Cand* cB = pong();
    const LevelGeom gL = geom(bank, L - 1);
    if (timing_enabled) hip_check(hipEventRecord(ev_[2], stream_), "event");
    kern::CoarseArgs ca;
    ca.geom = gL;
    ca.tl = bank.d_tl.as<TemplLevelDev>();
    ca.levels = L;
    ca.level = L - 1;
    ca.feats = bank.d_feats.as<uint2>();
    ca.ntiles = kern::coarse_tiles(gL.hw, ppl, tuning.gather);
    ca.raw_min = d_rawmin_c_.as<int>();
    ca.cand = L == 1 ? final_buf : cA;
    ca.cand_count = counts + 1;
    ca.cand_cap = (int)cand_cap_;
    // auto mode: the list path's fixed cost (list kernels, candidate sort, few work items per image) pays off on
    // large coarse stages only: features x positions over every image of the launch
    const bool worth = (double)nimg * (double)coarse_features_ * (double)gL.hw >= kSparseMinWork;
    const int mode_eff = tuning.coarse_mode == 1 && !worth ? 0 : tuning.coarse_mode;
    // a feature's range starts below cells + hw; the scatter's decode needs that below 2^29
    const bool list_fits = sparse_list_fits(gL);
    const bool sparse = mode == kModeBatched && mode_eff != 0 && list_fits;
    if (sparse) {  // list the coarse level's nonzero responses; the device picks the path
        Level& top = levels_[L - 1];
        kt_begin("sparse_lists");
        hip_check(kern::sparse_lists(gL, top.sp_rows.as<int>(), top.sp_rows.as<int>() + kern::sparse_chunks(gL) + 1,
                                     top.sp_entries.as<uint32_t>(), top.ent_b - kern::kSparseListPad, kSparseDensity,
                                     mode_eff, counts + kSparseNnzSlot, counts + kSparseFlagSlot, stream_, nimg,
                                     top.rows_b, top.ent_b),
                  "sparse lists");
        kt_end();
        ca.dense_skip = counts + kSparseFlagSlot;
    }
    kt_begin(mode == kModeBatched ? "coarse_batched" : "coarse_plain");
    if (sparse) {
        kern::SparseCoarseArgs sa;
        sa.geom = gL;
        sa.tl = bank.d_tl.as<TemplLevelDev>();
        sa.levels = L;
        sa.level = L - 1;
        sa.feats = bank.d_feats.as<uint2>();
        sa.order = d_order_.as<int>();
        sa.ntempl = nsel;
        sa.seg = (int)std::min<long long>(gL.hw, tuning.list_segment);
        sa.nseg = (int)((gL.hw + sa.seg - 1) / sa.seg);
        sa.raw_min = d_rawmin_c_.as<int>();
        sa.work_counter = counts + kSparseQueueSlot;
        sa.chunkstart = levels_[L - 1].sp_rows.as<int>();
        sa.nchunks = kern::sparse_chunks(gL);
        sa.entries = levels_[L - 1].sp_entries.as<uint32_t>();
        sa.entries_bytes = (int)std::min<long long>(levels_[L - 1].ent_b * (long long)sizeof(uint32_t), 0x7FFFFFFF);
        sa.entries_read = reinterpret_cast<unsigned long long*>(counts + kSparseReadSlot);
        sa.flag = counts + kSparseFlagSlot;
        sa.nimg = nimg;
        sa.cs_b = levels_[L - 1].rows_b;
        sa.ent_b = levels_[L - 1].ent_b;
        sa.cand = ca.cand;
        sa.cand_count = ca.cand_count;
        sa.cand_cap = ca.cand_cap;
        const long long items = (long long)nsel * sa.nseg * nimg;
        const long long blocks = std::min<long long>((long long)num_cu_ * tuning.list_bpc, items);
        sa.static_order = tuning.list_order == 1 && blocks >= 8 ? 1 : 0;
        sa.split_pct = tuning.list_split;
        hip_check(kern::coarse_sparse(sa, (int)std::max(1LL, blocks), tuning.list_threads, tuning.list_segment,
                                      stream_), "coarse sparse");
    }
    if (mode == kModeBatched) {
        ca.order = d_order_.as<int>();
        ca.ntempl = nsel;
        ca.nimg = nimg;
        ca.work_counter = counts;
        const long long items = (long long)nsel * ca.ntiles * nimg;
        const int per_block = tuning.split == 1 ? wpb : 1;
        long long blocks = (long long)num_cu_ * tuning.blocks_per_cu;
        blocks = std::min(blocks, (items + per_block - 1) / per_block);
        const long long workers = std::max(1LL, blocks) * per_block;
        ca.static_order = items <= 4 * workers ? 1 : 0;  // few items each: no queue at all
        ca.grab = 2;
        hip_check(kern::coarse(ca, ppl, wpb, (int)std::max(1LL, blocks), tuning.gather, tuning.split, stream_),
                  "coarse batched");
    } else {
        ca.ntempl = 1;
        ca.work_counter = nullptr;
        for (int k = 0; k < nsel; ++k) {  // one launch per template, as matchClass iterates them
            ca.order = d_order_.as<int>() + k;
            hip_check(kern::coarse(ca, ppl, wpb, 0, tuning.gather, tuning.split, stream_), "coarse plain");
        }
    }
    kt_end();
    if (timing_enabled) hip_check(hipEventRecord(ev_[3], stream_), "event");
    const Cand* cur = ca.cand;
    int* cur_count = counts + 1;
    Cand* nxt = cB;
    if (sparse && L >= 2) {  // the list path emits by template: spatial order for the refinement's XCD shares
        kt_begin("sort_candidates");
        hip_check(kern::sort_cands_by_row(cA, counts + 1, (int)cand_cap_, gL.T, gL.H, d_bins_.as<int>(), cB,
                                          num_cu_ * 2, stream_, nimg), "sort candidates");
        kt_end();
        cur = cB;
        nxt = cA;
    }
    for (int l = L - 2; l >= 0; --l) {
        kern::RefineArgs ra;
        ra.geom = geom(bank, l);
        ra.tl = bank.d_tl.as<TemplLevelDev>();
        ra.levels = L;
        ra.level = l;
        ra.feats = bank.d_feats.as<uint2>();
        ra.cand_in = cur;
        ra.count_in = cur_count;
        ra.cap_in = (int)cand_cap_;
        ra.raw_min = d_rawmin_r_.as<int>();
        ra.cand_out = l == 0 ? final_buf : nxt;
        ra.count_out = counts + 2 + (L - 2 - l);
        ra.cap_out = (int)cand_cap_;
        if (l == 0) ra.feat_count = reinterpret_cast<unsigned long long*>(counts + kRefineFeatSlot);
        if (l < kRefineQueueLevels) ra.queue = counts + kRefineQueueSlot + 8 * l;
        ra.layout = tuning.refine_layout;
        if (levels_[l].band_b > 0) {  // the level's (nibble-)banded copy, made after its linear memories
            ra.band = levels_[l].band.as<uint8_t>();
            ra.band_stride = levels_[l].band_b;
            ra.bgeom = levels_[l].band_nib ? nib_band_geom(bank.bins, ra.geom.T, ra.geom.W, ra.geom.H)
                                           : band_geom(bank.bins, ra.geom.T, ra.geom.W, ra.geom.H);
        } else if (ra.layout == 1 || ra.layout == 3 || ra.layout == 4) {  // no copy (2^31 bytes or more): the
            ra.layout = ra.layout == 1 ? 0 : 2;                           // linear memories
        }
        kt_begin("refine_l" + std::to_string(l));
        hip_check(kern::refine(ra, num_cu_ * tuning.refine_bpc, stream_, tuning.refine_waves, tuning.refine_group),
                  "refine");
        kt_end();
        cur = ra.cand_out;
        cur_count = ra.count_out;
        nxt = (nxt == cB) ? cA : cB;  // (refine level outputs alternate between the two)
    }
    if (timing_enabled) hip_check(hipEventRecord(ev_[4], stream_), "event");
}

// The counters and the first kPrefetchCands results come back in one copy.
void Pipeline::enqueue_head_copy() {
    const size_t head = kCounterSlots * sizeof(int) + std::min(kPrefetchCands, cand_cap_) * sizeof(Cand);
    hip_check(hipMemcpyAsync(h_out_.p, d_work_.p, head, hipMemcpyDeviceToHost, stream_), "D2H results");
}

// Host side after the head copy has landed: the candidates of each image of the launch (the image in the
// template index's high bits), then per image the canonical sort + unique (line2Dup.cpp:1113-1116).
std::vector<std::vector<MatchRec>> Pipeline::decode_batch(const Bank& bank, bool* overflow, int nimg) {
    const int L = bank.L;
    const int* hc = static_cast<const int*>(h_out_.p);
    const int coarse_n = hc[1];
    const int final_n = L >= 2 ? hc[2 + (L - 2)] : coarse_n;
    std::vector<std::vector<MatchRec>> out((size_t)std::max(nimg, 1));
    *overflow = (size_t)coarse_n > cand_cap_;
    last_times.coarse_candidates = coarse_n;
    last_times.final_candidates = final_n;
    last_times.coarse_sparse = hc[kSparseFlagSlot];
    {
        unsigned long long rd = 0;
        memcpy(&rd, hc + kSparseReadSlot, sizeof(rd));
        last_times.coarse_entries = (long long)rd;
        memcpy(&rd, hc + kRefineFeatSlot, sizeof(rd));
        last_times.refine_features = (long long)rd;
    }
    last_times.coarse_nonzero = hc[kSparseFlagSlot] || hc[kSparseNnzSlot] ? hc[kSparseNnzSlot] : -1;
    if (*overflow) {
        while (cand_cap_ < (size_t)coarse_n) cand_cap_ *= 2;
        return out;
    }
    std::vector<Cand> hcand(final_n);
    if (final_n) {
        const Cand* pre = reinterpret_cast<const Cand*>(static_cast<const char*>(h_out_.p) + kCounterSlots * sizeof(int));
        if ((size_t)final_n <= kPrefetchCands) {
            memcpy(hcand.data(), pre, (size_t)final_n * sizeof(Cand));
        } else {
            h_stage_.ensure((size_t)final_n * sizeof(Cand));
            hip_check(hipMemcpyAsync(h_stage_.p, out_cands(), (size_t)final_n * sizeof(Cand), hipMemcpyDeviceToHost,
                                     stream_), "D2H cand");
            hip_check(hipStreamSynchronize(stream_), "sync cand");
            memcpy(hcand.data(), h_stage_.p, (size_t)final_n * sizeof(Cand));
        }
    }
    if (nimg <= 1) {
        out[0] = canonical_matches(bank, hcand.data(), final_n);
        return out;
    }
    std::vector<std::vector<Cand>> per((size_t)nimg);
    for (const Cand& c : hcand) {
        const int z = c.g >> kCandImageShift;
        if (z < 0 || z >= nimg) throw Error(kEHip, "candidate of an image outside the launch");
        Cand t = c;
        t.g &= kCandTemplMask;
        per[z].push_back(t);
    }
    for (int z = 0; z < nimg; ++z) out[z] = canonical_matches(bank, per[z].data(), (int)per[z].size());
    return out;
}

std::vector<MatchRec> Pipeline::decode(const Bank& bank, bool* overflow) {
    auto v = decode_batch(bank, overflow, 1);
    return std::move(v[0]);
}

// Scores, then the canonical sort + unique (line2Dup.cpp:1113-1116) of n refined candidates.
std::vector<MatchRec> canonical_matches(const Bank& bank, const Cand* cands, int n) {
    const int L = bank.L;
    std::vector<MatchRec> out;
    out.reserve(n);
    for (int i = 0; i < n; ++i) {
        const Cand& cd = cands[i];
        MatchRec m;
        m.x = cd.x;
        m.y = cd.y;
        const int n0 = bank.h_tl[(size_t)cd.g * L + 0].nfeat;
        if (L == 1) m.similarity = score_of(cd.raw, n0);
        else m.similarity = cd.raw > 0 ? score_of(cd.raw, n0) : 0.f;
        m.class_index = bank.g_class[cd.g];
        m.template_id = bank.g_tid[cd.g];
        out.push_back(m);
    }
    std::sort(out.begin(), out.end(), match_less);
    auto eq = [](const MatchRec& a, const MatchRec& b) {  // Match::operator==, line2Dup.h:126-129
        return a.x == b.x && a.y == b.y && a.similarity == b.similarity && a.class_index == b.class_index;
    };
    out.erase(std::unique(out.begin(), out.end(), eq), out.end());
    return out;
}

void Pipeline::prepare(const Bank& bank, const ImageArgs& im) {
    ensure_levels(bank, im.rows, im.cols, im.channels, im.mask != nullptr);
    ensure_work();
    kt_used_ = 0;
    enqueue_preprocess(bank, im);
    hip_check(hipStreamSynchronize(stream_), "sync");
    kt_used_ = 0;
}

std::vector<MatchRec> Pipeline::match(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel,
                                      int mode) {
    ensure_static(bank, thr, sel);
    ensure_levels(bank, im.rows, im.cols, im.channels, im.mask != nullptr);
    h_out_.ensure(kCounterSlots * sizeof(int) + kPrefetchCands * sizeof(Cand));
    kt_used_ = 0;
    for (int attempt = 0; attempt < 8; ++attempt) {
        ensure_work();
        if (attempt == 0) {
            enqueue_preprocess(bank, im);  // also clears the counters
        } else {
            hip_check(kern::fill_u32((uint32_t*)counters(), 0u, kCounterSlots, stream_), "zero counters");
        }
        if (!sel.empty()) enqueue_match(bank, (int)sel.size(), mode);
        enqueue_head_copy();
        hip_check(hipStreamSynchronize(stream_), "sync results");
        bool overflow = false;
        auto out = decode(bank, &overflow);
        if (overflow) {  // grow the candidate memory and run the matching stages again
            kt_used_ = 0;
            continue;
        }
        if (timing_enabled && !sel.empty()) {
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
        kt_used_ = 0;
        return out;
    }
    throw Error(kEBudget, "candidate buffer could not be grown enough");
}

std::vector<MatchRec> Pipeline::match_nms(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel,
                                          float score_threshold, float nms_threshold, int top_k,
                                          std::vector<MatchRec>* all) {
    const int L = bank.L;
    ensure_static(bank, thr, sel);
    ensure_levels(bank, im.rows, im.cols, im.channels, im.mask != nullptr);
    const size_t out_bytes = (size_t)kFinalMax * 6 * sizeof(int);
    h_out_.ensure(std::max(kCounterSlots * sizeof(int) + kPrefetchCands * sizeof(Cand), kCounterSlots * sizeof(int) + out_bytes));
    kt_used_ = 0;
    for (int attempt = 0; attempt < 8; ++attempt) {
        ensure_work();  // ping holds the list and the kept positions, pong the suppression matrix
        if (attempt == 0) {
            enqueue_preprocess(bank, im);
        } else {
            hip_check(kern::fill_u32((uint32_t*)counters(), 0u, kCounterSlots, stream_), "zero counters");
        }
        if (!sel.empty()) enqueue_match(bank, (int)sel.size(), kModeBatched);
        kern::FinalArgs fa;
        fa.cand = out_cands();
        fa.count_in = counters() + (L >= 2 ? 2 + (L - 2) : 1);
        fa.cap_in = (int)cand_cap_;
        fa.tl = bank.d_tl.as<TemplLevelDev>();
        fa.levels = L;
        fa.gclass = bank.d_gclass.as<int>();
        fa.gtid = bank.d_gtid.as<int>();
        fa.score_threshold = score_threshold;
        fa.nms_threshold = nms_threshold;
        fa.top_k = top_k;
        fa.out = reinterpret_cast<int*>(ping());
        fa.kept = fa.out + 5 * kFinalMax;
        fa.mask = reinterpret_cast<uint64_t*>(pong());
        fa.counters = counters();
        kt_begin("finalize_nms");
        hip_check(kern::finalize(fa, stream_), "finalize");
        kt_end();
        char* h = static_cast<char*>(h_out_.p);
        hip_check(hipMemcpyAsync(h, d_work_.p, kCounterSlots * sizeof(int), hipMemcpyDeviceToHost, stream_), "D2H counters");
        hip_check(hipMemcpyAsync(h + kCounterSlots * sizeof(int), ping(), out_bytes, hipMemcpyDeviceToHost, stream_),
                  "D2H list");
        hip_check(hipStreamSynchronize(stream_), "sync results");
        const int* hc = reinterpret_cast<const int*>(h);
        if ((size_t)hc[1] > cand_cap_) {  // coarse overflow: grow and run the matching stages again
            while (cand_cap_ < (size_t)hc[1]) cand_cap_ *= 2;
            kt_used_ = 0;
            continue;
        }
        if (timing_enabled) kt_collect();
        kt_used_ = 0;
        std::vector<MatchRec> list, kept;
        if (hc[kern::kFinalFlagSlot]) {  // longer than the device finalizes: the host's sort, unique and NMS
            enqueue_head_copy();
            hip_check(hipStreamSynchronize(stream_), "sync results");
            bool overflow = false;
            list = decode(bank, &overflow);
            std::unordered_map<uint64_t, int> global;  // (class, template id) -> global template
            for (size_t g = 0; g < bank.g_class.size(); ++g)
                global[((uint64_t)(uint32_t)bank.g_class[g] << 32) | (uint32_t)bank.g_tid[g]] = (int)g;
            std::vector<int32_t> boxes;
            std::vector<float> scores;
            for (const MatchRec& m : list) {
                const int g = global.at(((uint64_t)(uint32_t)m.class_index << 32) | (uint32_t)m.template_id);
                const TemplLevelDev& t0 = bank.h_tl[(size_t)g * L];
                boxes.insert(boxes.end(), {m.x, m.y, t0.width, t0.height});
                scores.push_back(m.similarity);
            }
            const std::vector<int> idx =
                nms_boxes(boxes.data(), scores.data(), (int)list.size(), score_threshold, nms_threshold, 1.f, top_k);
            for (int i : idx) kept.push_back(list[i]);
        } else {
            const int M = hc[kern::kFinalCountSlot], nk = hc[kern::kKeptCountSlot];
            const int* r = hc + kCounterSlots;
            list.resize(M);
            for (int i = 0; i < M; ++i) {
                MatchRec& m = list[i];
                m.x = r[5 * i];
                m.y = r[5 * i + 1];
                uint32_t bits = (uint32_t)r[5 * i + 2];
                memcpy(&m.similarity, &bits, 4);
                m.class_index = r[5 * i + 3];
                m.template_id = r[5 * i + 4];
            }
            const int* kp = r + 5 * kFinalMax;
            for (int k = 0; k < nk; ++k) kept.push_back(list[kp[k]]);
        }
        if (all) *all = std::move(list);
        return kept;
    }
    throw Error(kEBudget, "candidate buffer could not be grown enough");
}

Pipeline::GraphKey Pipeline::make_key(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel,
                                      int nimg) const {
    GraphKey key;
    key.bank_version = bank.version;
    key.rows = im.rows;
    key.cols = im.cols;
    key.channels = im.channels;
    key.nimg = nimg;
    key.aligned = im.device ? ((uintptr_t)im.img % 4 == 0 && im.stride % 4 == 0) : (im.cols * im.channels) % 4 == 0;
    key.mask = im.mask != nullptr;
    key.thr_bits = float_bits(thr);
    key.sel_sig = sel_signature(sel);
    key.tuning = tuning;
    key.cand_cap = std::max(cand_cap_, kInitialCandCap);
    return key;
}

// Batched match that replays a captured graph when this call repeats the previous one's size,
// threshold, selection and tuning (capturing on the second occurrence), else launches eagerly.
std::vector<MatchRec> Pipeline::match_reuse(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel) {
    if (timing_enabled) return match(bank, im, thr, sel, kModeBatched);  // replays record no per-kernel events
    const GraphKey key = make_key(bank, im, thr, sel);
    bool have = last_key_valid_ && key == last_key_;
    for (const auto& g : graphs_)
        if (g.key == key) have = true;
    if (have) {
        graph_launch(bank, im, thr, sel);
        return graph_finish(bank);
    }
    last_key_ = key;
    last_key_valid_ = true;
    return match(bank, im, thr, sel, kModeBatched);
}

void Pipeline::graph_launch(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel) {
    graph_launch_batch(bank, &im, 1, thr, sel);
}

void Pipeline::graph_launch_batch(const Bank& bank, const ImageArgs* ims, int n, float thr,
                                  const std::vector<int>& sel) {
    if (pending_) throw Error(kEInval, "a graph launch is already pending on this pipeline");
    if (n < 1 || n > kMaxBatch) throw Error(kEInval, "a launch takes 1 to " + std::to_string(kMaxBatch) + " images");
    const ImageArgs& im0 = ims[0];
    const bool mask = im0.mask != nullptr;
    for (int k = 1; k < n; ++k)
        if (ims[k].rows != im0.rows || ims[k].cols != im0.cols || ims[k].channels != im0.channels ||
            (ims[k].mask != nullptr) != mask)
            throw Error(kEInval, "the images of one launch share their size, channels and mask presence");
    ensure_static(bank, thr, sel);
    if (ensure_levels(bank, im0.rows, im0.cols, im0.channels, mask, n)) clear_graphs();
    ensure_work();
    h_out_.ensure(kCounterSlots * sizeof(int) + kPrefetchCands * sizeof(Cand));
    if (!d_src_.p) clear_graphs();  // graphs read the table at its address
    d_src_.ensure(sizeof(SrcTable));
    h_src_.ensure(sizeof(SrcTable));

    // level-0 inputs: images in device memory are read where they are; host images go through pinned
    // memory into the pipeline's input buffer, one DMA for all of them
    const size_t row = (size_t)im0.cols * im0.channels;
    const size_t in_bytes = row * im0.rows, mask_bytes = mask ? (size_t)im0.rows * im0.cols : 0;
    const size_t in_b = (in_bytes + 255) & ~(size_t)255, mask_b = (mask_bytes + 255) & ~(size_t)255;
    int nhost = 0;
    for (int k = 0; k < n; ++k) nhost += ims[k].device ? 0 : 1;
    if (nhost) {
        d_in_.ensure(n * in_b);
        if (mask) d_in_mask_.ensure(n * mask_b);
    }
    SrcTable tab{};
    bool aligned = true;
    for (int k = 0; k < n; ++k) {
        const ImageArgs& im = ims[k];
        if (im.device) {
            tab.img[k] = im.img;
            tab.mask[k] = im.mask;
            tab.stride[k] = im.stride;
        } else {
            tab.img[k] = d_in_.as<uint8_t>() + k * in_b;
            tab.mask[k] = mask ? d_in_mask_.as<uint8_t>() + k * mask_b : nullptr;
            tab.stride[k] = (int)row;
        }
        aligned &= (uintptr_t)tab.img[k] % 4 == 0 && tab.stride[k] % 4 == 0;
    }
    GraphKey key = make_key(bank, im0, thr, sel, n);
    key.aligned = aligned;
    // with per-kernel timing the stages run eagerly below (a replay records no per-kernel events)
    const bool eager = timing_enabled;
    GraphEntry* entry = nullptr;
    for (auto& g : graphs_)
        if (g.key == key) entry = &g;
    if (!entry && !eager) {
        if (graphs_.size() >= 24) {  // keep the most recent (size, image count) buckets
            (void)hipGraphExecDestroy(graphs_.front().exec);
            (void)hipGraphDestroy(graphs_.front().graph);
            graphs_.erase(graphs_.begin());
        }
        ImageArgs gi = im0;  // geometry only: the kernels take the pixels from the table
        gi.img = tab.img[0];
        gi.mask = tab.mask[0];
        gi.stride = tab.stride[0];
        gi.device = true;
        const bool timing = timing_enabled;
        timing_enabled = false;
        kt_used_ = 0;
        src_aligned_ = aligned;
        GraphEntry ge;
        ge.key = key;
        hip_check(hipStreamSynchronize(stream_), "sync before capture");
        hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            enqueue_preprocess(bank, gi, n, d_src_.as<SrcTable>());
            if (!sel.empty()) enqueue_match(bank, (int)sel.size(), kModeBatched, n);
            enqueue_head_copy();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            timing_enabled = timing;
            throw;
        }
        hip_check(hipStreamEndCapture(stream_, &ge.graph), "end capture");
        timing_enabled = timing;
        hip_check(hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0), "instantiate");
        graphs_.push_back(ge);
        ++graphs_captured_;
        entry = &graphs_.back();
    }
    if (graph_timing) hip_check(hipEventRecord(ev_[0], stream_), "event");
    if (nhost) {
        // the previous DMA from the pinned buffers is done before they are rewritten
        h_in_.ensure(n * (in_b + mask_b));
        hip_check(hipStreamSynchronize(stream_), "sync before staging");
        uint8_t* st = static_cast<uint8_t*>(h_in_.p);
        for (int k = 0; k < n; ++k) {
            const ImageArgs& im = ims[k];
            if (im.device) continue;
            uint8_t* d = st + k * in_b;
            if ((size_t)im.stride == row) memcpy(d, im.img, in_bytes);
            else
                for (int y = 0; y < im.rows; ++y) memcpy(d + (size_t)y * row, im.img + (size_t)y * im.stride, row);
            if (mask) memcpy(st + n * in_b + k * mask_b, im.mask, mask_bytes);
        }
        hip_check(hipMemcpyAsync(d_in_.p, st, n * in_b, hipMemcpyHostToDevice, stream_), "copy input");
        if (mask)
            hip_check(hipMemcpyAsync(d_in_mask_.p, st + n * in_b, n * mask_b, hipMemcpyHostToDevice, stream_), "copy mask");
    }
    memcpy(h_src_.p, &tab, sizeof(SrcTable));
    hip_check(hipMemcpyAsync(d_src_.p, h_src_.p, sizeof(SrcTable), hipMemcpyHostToDevice, stream_), "copy table");
    if (eager) {
        ImageArgs gi = im0;  // geometry only: the kernels take the pixels from the table
        gi.img = tab.img[0];
        gi.mask = tab.mask[0];
        gi.stride = tab.stride[0];
        gi.device = true;
        kt_used_ = 0;
        src_aligned_ = aligned;
        enqueue_preprocess(bank, gi, n, d_src_.as<SrcTable>());
        if (!sel.empty()) enqueue_match(bank, (int)sel.size(), kModeBatched, n);
        enqueue_head_copy();
    } else {
        hip_check(hipGraphLaunch(entry->exec, stream_), "graph launch");
    }
    eager_timed_ = eager;
    if (graph_timing) hip_check(hipEventRecord(ev_[1], stream_), "event");
    pending_timed_ = graph_timing;
    hip_check(hipEventRecord(done_, stream_), "event");
    pending_ = true;
    pending_n_ = n;
    pending_ims_.assign(ims, ims + n);
    pending_im_ = im0;
    pending_thr_ = thr;
    pending_sel_ = sel;
}

bool Pipeline::graph_ready() const {
    if (!pending_) return false;
    const hipError_t e = hipEventQuery(done_);
    if (e == hipErrorNotReady) return false;
    hip_check(e, "event query");
    return true;
}

void Pipeline::release() {
    if (pending_) throw Error(kEInval, "release of a pipeline with a pending launch");
    hip_check(hipStreamSynchronize(stream_), "sync");
    clear_graphs();
    levels_.clear();
    d_work_.release();
    d_order_.release();
    d_rawmin_c_.release();
    d_rawmin_r_.release();
    d_in_.release();
    d_in_mask_.release();
    d_src_.release();
    d_bins_.release();
    h_in_.release();
    cand_cap_ = 0;
    static_valid_ = false;
    cur_rows_ = cur_cols_ = 0;
}

std::vector<MatchRec> Pipeline::graph_finish(const Bank& bank) {
    auto v = graph_finish_batch(bank);
    return std::move(v[0]);
}

std::vector<std::vector<MatchRec>> Pipeline::graph_finish_batch(const Bank& bank) {
    if (!pending_) throw Error(kEInval, "no graph launch pending");
    pending_ = false;
    hip_check(hipEventSynchronize(done_), "wait graph");
    if (eager_timed_) {
        kt_collect();
        eager_timed_ = false;
    }
    if (pending_timed_) {
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "elapsed");
        last_times.total_ms = ms;
        pending_timed_ = false;
    }
    bool overflow = false;
    auto out = decode_batch(bank, &overflow, pending_n_);
    if (!overflow) return out;
    // more candidates than the captured buffers hold: grow (new graphs next time) and redo eagerly
    clear_graphs();
    std::vector<std::vector<MatchRec>> res;
    for (const ImageArgs& im : pending_ims_) res.push_back(match(bank, im, pending_thr_, pending_sel_, kModeBatched));
    return res;
}

// what: 0 = quantized (masked) orientations (2 bytes per pixel with 16 orientations), 1 = linear
// memories [planes][T*T][H*W], 2 = level image, 3 = nibble-packed coarse level, 4 = banded or nibble-banded copy
// of a refined level (zeroed block included)
void Pipeline::read_level(int level, int what, uint8_t* out) {
    if (level < 0 || level >= (int)levels_.size() || cur_rows_ == 0) throw Error(kEInval, "no prepared level");
    const Level& lv = levels_[level];
    size_t n = 0;
    const void* src = nullptr;
    if (what == 0) {
        n = (size_t)lv.rows * lv.cols * (size_t)(cur_bins_ / 8);
        src = lv.q.p;
    } else if (what == 1) {
        n = (size_t)lv.rows * lv.cols * (size_t)cur_bins_;
        src = lv.lm.p;
    } else if (what == 3) {  // nibble-packed copy of the coarse level, guard included
        if (level != (int)levels_.size() - 1 || !lv.lm_nib.p || !nib_valid_)
            throw Error(kEInval, "no nibble-packed level (coarse level with gather 3 only)");
        const int T = nib_T_;
        n = (size_t)lm_nibble_bytes(lm_level_bytes(cur_bins_, lv.rows, lv.cols, T));
        src = lv.lm_nib.p;
    } else if (what == 4) {
        if (lv.band_b == 0 || !lv.band.p)
            throw Error(kEInval, "no banded copy (refined levels with refine_layout 1, 3 or 4)");
        n = (size_t)(lv.band_nib ? nib_band_geom(cur_bins_, lv.T, lv.cols / lv.T, lv.rows / lv.T)
                                 : band_geom(cur_bins_, lv.T, lv.cols / lv.T, lv.rows / lv.T)).bytes;
        src = lv.band.p;
    } else if (what == 2) {
        if (level == 0 && cur_device_input_) throw Error(kEInval, "level 0 image is the caller's device buffer");
        n = (size_t)lv.rows * lv.cols * cur_channels_;
        src = lv.img.p;
    } else {
        throw Error(kEInval, "unknown buffer");
    }
    hip_check(hipMemcpyAsync(out, src, n, hipMemcpyDeviceToHost, stream_), "D2H");
    hip_check(hipStreamSynchronize(stream_), "sync");
}

}  // namespace sbm
