// Allocation-free batched match of an HBM-resident image (sbm_match_device_noalloc): the whole
// per-image pipeline of Detector::match (line2Dup.cpp:1049-1268) with every buffer carved from one
// fixed device array, for callers that cannot allocate device memory.
//
// The array `scratch` below is a file-scope device variable: each GPU holds its own copy, at its own
// address, from the code object's load until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is a stand-in, not a device address); it asks the runtime for the
// device address once per device, with that device current, and uses only that address afterwards.
//
// Every call on one GPU shares that GPU's array and leaves its contents behind.  What must start at
// zero is cleared by the call's own kernels: the counters by the first gradient launch, the guard after
// each level's linear memories (and the nibble copy's) by the spread kernels.  Calls that overlap on
// one GPU (two threads, or two detectors on different streams) race on the array: the caller
// serializes them.
#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>

#include "sbm_detector.h"

namespace sbm {

constexpr size_t kNoallocScratchBytes = 64ull << 20;  // 64 MiB per GPU
constexpr int kNoallocCands = 1 << 16;                 // candidates per stage (coarse, refine ping/pong)
constexpr int kNoallocMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoallocScratchBytes];

namespace {
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kNoallocMaxDevices] = {};

// Device address of this device's copy of `scratch`; the device must be current.
uint8_t* scratch_on(int device) {
    if (device < 0 || device >= kNoallocMaxDevices) throw Error(kEInval, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw Error(kEHip, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[device] = static_cast<uint8_t*>(p);
    }
    return g_scratch_ptr[device];
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Byte offsets of one call's buffers inside the scratch array.
struct Plan {
    size_t counters = 0, cands = 0, order = 0, rawmin_c = 0, rawmin_r = 0;
    struct Lvl {
        int rows = 0, cols = 0;
        size_t img = 0, mask = 0, q = 0, lm = 0, nib = 0;  // img / mask: levels above 0 only
        long long lm_bytes = 0, nib_bytes = 0;
    };
    std::vector<Lvl> lv;
    size_t total = 0;
};

Plan make_plan(const Bank& bank, int rows, int cols, int channels, bool mask, int nsel, bool nib) {
    Plan p;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at = align256(at + bytes);
        return o;
    };
    p.counters = take(Pipeline::kCounterSlots * sizeof(int));
    p.cands = take(3ull * kNoallocCands * sizeof(Cand));
    p.order = take((size_t)std::max(nsel, 1) * sizeof(int));
    p.rawmin_c = take((size_t)std::max(bank.size(), 1) * sizeof(int));
    p.rawmin_r = take((size_t)std::max(bank.size(), 1) * bank.L * sizeof(int));
    int r = rows, c = cols;
    p.lv.resize(bank.L);
    for (int l = 0; l < bank.L; ++l) {
        if (l > 0) { r /= 2; c /= 2; }
        Plan::Lvl& v = p.lv[l];
        v.rows = r;
        v.cols = c;
        const int T = bank.T[l];
        const size_t px = (size_t)r * c;
        if (l > 0) {
            v.img = take(px * channels);
            if (mask) v.mask = take(px);
        }
        v.q = take(px * (size_t)(bank.bins / 8));
        v.lm_bytes = lm_level_bytes(bank.bins, r, c, T);
        v.lm = take((size_t)v.lm_bytes);
        if (nib && l == bank.L - 1) {
            v.nib_bytes = lm_nibble_bytes(v.lm_bytes);
            v.nib = take((size_t)v.nib_bytes);
        }
    }
    p.total = at;
    return p;
}
}  // namespace

std::vector<MatchRec> Detector::match_noalloc(const uint8_t* d_img, int rows, int cols, int stride, int channels,
                                              float threshold, const std::vector<std::string>& class_ids,
                                              const uint8_t* d_mask) {
    check_channels(channels);
    if (stride < cols * channels) throw Error(kEInval, "stride smaller than a row");
    check_sizes(rows, cols);
    if (!bank_ready())
        throw Error(kEInval, "the allocation-free match needs the bank in HBM first: call sbm_upload_bank after the "
                             "last change to the bank");
    DeviceScope ds(device_);
    const Bank& bank = bank_;
    const int L = bank.L;
    std::vector<int> sel;
    selection(class_ids, true, sel);
    const Tuning tu = tuning;
    const bool nib = tu.gather == 3 && bank.bins == 8;
    const Plan plan = make_plan(bank, rows, cols, channels, d_mask != nullptr, (int)sel.size(), nib);
    if (plan.total > kNoallocScratchBytes)
        throw Error(kEBudget, "image and selection need " + std::to_string(plan.total) +
                                  " bytes of device scratch; the allocation-free match holds " +
                                  std::to_string(kNoallocScratchBytes));
    uint8_t* base = scratch_on(device_);
    hipStream_t st = pipe_->stream();

    // integer thresholds and the template order (large banks first), refreshed on every call: another
    // call on this GPU may have left its own in the array
    std::vector<int> rc(std::max(bank.size(), 1), INT_MAX), rr((size_t)std::max(bank.size(), 1) * L, INT_MAX);
    for (int g = 0; g < bank.size(); ++g) {
        rc[g] = raw_min_coarse(bank.h_tl[(size_t)g * L + L - 1].nfeat, threshold);
        for (int l = 0; l < L - 1; ++l) rr[(size_t)g * L + l] = raw_min_refine(bank.h_tl[(size_t)g * L + l].nfeat, threshold);
    }
    hip_check(hipMemcpyAsync(base + plan.rawmin_c, rc.data(), rc.size() * sizeof(int), hipMemcpyHostToDevice, st), "H2D");
    hip_check(hipMemcpyAsync(base + plan.rawmin_r, rr.data(), rr.size() * sizeof(int), hipMemcpyHostToDevice, st), "H2D");
    if (!sel.empty())
        hip_check(hipMemcpyAsync(base + plan.order, sel.data(), sel.size() * sizeof(int), hipMemcpyHostToDevice, st),
                  "H2D order");

    int* counts = reinterpret_cast<int*>(base + plan.counters);
    Cand* final_buf = reinterpret_cast<Cand*>(base + plan.cands);
    Cand* cA = final_buf + kNoallocCands;
    Cand* cB = cA + kNoallocCands;
    auto geom = [&](int l) {
        const Plan::Lvl& v = plan.lv[l];
        LevelGeom g{};
        g.rows = v.rows;
        g.cols = v.cols;
        g.T = bank.T[l];
        g.W = g.cols / g.T;
        g.H = g.rows / g.T;
        g.hw = (long long)g.W * g.H;
        g.lm = base + v.lm;
        g.planes = bank.bins;
        g.lm_bytes = (int)v.lm_bytes;
        g.lm_nib = v.nib_bytes ? base + v.nib : nullptr;
        g.nib_bytes = (int)v.nib_bytes;
        return g;
    };

    // preprocessing of every level; the first launch clears the counters, and each level's gradient kernel
    // also writes the next level's image (and mask)
    for (int l = 0; l < L; ++l) {
        const Plan::Lvl& v = plan.lv[l];
        const uint8_t* cur_img = d_img;
        const uint8_t* cur_mask = d_mask;
        int cur_stride = stride;
        if (l > 0) {
            cur_img = base + v.img;
            cur_mask = d_mask ? base + v.mask : nullptr;
            cur_stride = v.cols * channels;
        }
        uint8_t* down = l + 1 < L ? base + plan.lv[l + 1].img : nullptr;
        uint8_t* down_mask = l + 1 < L && d_mask ? base + plan.lv[l + 1].mask : nullptr;
        hip_check(kern::gradient_quantize(cur_img, v.rows, v.cols, cur_stride, channels, cur_mask, bank.weak, base + v.q,
                                          nullptr, nullptr, st, l == 0 ? (uint32_t*)counts : nullptr,
                                          Pipeline::kCounterSlots, bank.bins, 1, BatchStrides{}, down, down_mask),
                  "gradient");
        const bool want_nib = nib && l == L - 1;
        bool nib_done = false;
        if (bank.bins == 16)
            hip_check(kern::spread_linearize16(reinterpret_cast<uint16_t*>(base + v.q), v.rows, v.cols, bank.T[l],
                                               base + v.lm, st),
                      "spread_linearize16");
        else
            hip_check(kern::spread_linearize(base + v.q, v.rows, v.cols, bank.T[l], base + v.lm, st,
                                             want_nib ? base + v.nib : nullptr, &nib_done),
                      "spread_linearize");
        if (want_nib && !nib_done) {
            hip_check(kern::pack_nibbles(base + v.lm, v.lm_bytes, base + v.nib, st), "pack_nibbles");
            // the packing covers whole 16-byte groups; the copy's last few bytes past them stay zero
            hip_check(kern::fill_u32(reinterpret_cast<uint32_t*>(base + v.nib + v.nib_bytes - 16), 0u, 4, st),
                      "clear nibble tail");
        }
    }

    // coarse stage over the whole selection in one persistent launch (the gather kernel), then refinement
    const int nsel = (int)sel.size();
    if (nsel > 0) {
        const LevelGeom gL = geom(L - 1);
        kern::CoarseArgs ca;
        ca.geom = gL;
        ca.tl = bank.d_tl.as<TemplLevelDev>();
        ca.levels = L;
        ca.level = L - 1;
        ca.feats = bank.d_feats.as<uint2>();
        ca.order = reinterpret_cast<const int*>(base + plan.order);
        ca.ntempl = nsel;
        ca.ntiles = kern::coarse_tiles(gL.hw, tu.ppl, tu.gather);
        ca.raw_min = reinterpret_cast<const int*>(base + plan.rawmin_c);
        ca.work_counter = counts;
        ca.cand = L == 1 ? final_buf : cA;
        ca.cand_count = counts + 1;
        ca.cand_cap = kNoallocCands;
        const long long items = (long long)nsel * ca.ntiles;
        const int per_block = tu.split == 1 ? tu.waves_per_block : 1;
        long long blocks = (long long)num_cu_ * tu.blocks_per_cu;
        blocks = std::min(blocks, (items + per_block - 1) / per_block);
        const long long workers = std::max(1LL, blocks) * per_block;
        ca.static_order = items <= 4 * workers ? 1 : 0;
        ca.grab = 2;
        hip_check(kern::coarse(ca, tu.ppl, tu.waves_per_block, (int)std::max(1LL, blocks), tu.gather, tu.split, st),
                  "coarse");
        const Cand* cur = ca.cand;
        const int* cur_count = counts + 1;
        Cand* nxt = cB;
        for (int l = L - 2; l >= 0; --l) {
            kern::RefineArgs ra;
            ra.geom = geom(l);
            ra.tl = bank.d_tl.as<TemplLevelDev>();
            ra.levels = L;
            ra.level = l;
            ra.feats = bank.d_feats.as<uint2>();
            ra.cand_in = cur;
            ra.count_in = cur_count;
            ra.cap_in = kNoallocCands;
            ra.raw_min = reinterpret_cast<const int*>(base + plan.rawmin_r);
            ra.cand_out = l == 0 ? final_buf : nxt;
            ra.count_out = counts + 2 + (L - 2 - l);
            ra.cap_out = kNoallocCands;
            if (l < Pipeline::kRefineQueueLevels) ra.queue = counts + Pipeline::kRefineQueueSlot + 8 * l;
            // the banded copies (layouts 1, 3, 4) have no room here: the linear memories, with 16-byte loads for
            // 2, 3 and 4
            ra.layout = tu.refine_layout >= 2 ? 2 : 0;
            hip_check(kern::refine(ra, num_cu_ * tu.refine_bpc, st, tu.refine_waves, tu.refine_group), "refine");
            cur = ra.cand_out;
            cur_count = ra.count_out;
            nxt = (nxt == cB) ? cA : cB;
        }
    }

    // counters, then the refined candidates, into host memory (pageable: nothing is allocated for them)
    int hc[Pipeline::kCounterSlots];
    hip_check(hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, st), "D2H counters");
    hip_check(hipStreamSynchronize(st), "sync");
    if (nsel == 0) return {};
    const int coarse_n = hc[1];
    if (coarse_n > kNoallocCands)
        throw Error(kEBudget, "the coarse stage passed " + std::to_string(coarse_n) + " candidates; the allocation-free "
                              "match holds " + std::to_string(kNoallocCands) + " (raise the threshold, or use "
                              "sbm_match_device)");
    const int final_n = L >= 2 ? hc[2 + (L - 2)] : coarse_n;
    std::vector<Cand> hcand((size_t)final_n);
    if (final_n > 0) {
        hip_check(hipMemcpyAsync(hcand.data(), final_buf, (size_t)final_n * sizeof(Cand), hipMemcpyDeviceToHost, st),
                  "D2H candidates");
        hip_check(hipStreamSynchronize(st), "sync");
    }
    return canonical_matches(bank, hcand.data(), final_n);
}

}  // namespace sbm
