// Host-side model of line2Dup::Detector (line2Dup.h:13-205) and the device pipeline behind match().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sbm_internal.h"

namespace sbm {

enum Status {
    kOk = 0,
    kEInval = -1,
    kESize = -2,
    kEFeatures = -3,
    kEDupClass = -4,
    kEHip = -5,
    kEBudget = -6,
    kEIO = -7,
    kENotFound = -8,
    kETooFew = -9,
    kENoDevice = -10,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);

// Feature / Template (line2Dup.h:13-39).  theta is kept for addTemplate_rotate; it is not persisted.
struct Feature {
    int x = 0, y = 0, label = 0;
    float theta = 0.f;
};
struct Template {
    int width = 0, height = 0, tl_x = 0, tl_y = 0, pyramid_level = 0;
    std::vector<Feature> features;
};
using TemplatePyramid = std::vector<Template>;
using TemplatesMap = std::map<std::string, std::vector<TemplatePyramid>>;

// One result row (Match, line2Dup.h:108-141) with the class as an index into the sorted class names.
struct MatchRec {
    int x, y;
    float similarity;
    int class_index;
    int template_id;
};

// Device memory that grows but never shrinks.
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n);
    void release();
    ~DevMem() { release(); }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Pinned host staging memory (hipHostMalloc), grow-only.
struct HostPinned {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n);
    void release();
    ~HostPinned() { release(); }
};

enum MatchMode { kModePlain = 0, kModeBatched = 1 };

struct Tuning {
    int ppl = 16;              // positions per lane in the coarse kernel (4, 8, 16)
    int waves_per_block = 4;   // coarse block size in wavefronts
    int blocks_per_cu = 8;     // persistent grid = CUs * blocks_per_cu
    int gather = 2;            // coarse loads: 0 aligned dwords + byte align, 1 unaligned buffer vectors,
                               // 2 aligned vectors + DPP, 3 as 2 on nibble-packed memories (ppl 32)
    int split = 1;             // wavefronts sharing one work item's features (1, 2, 4, 8; a block each)
    int coarse_mode = 1;       // batched coarse stage: 0 dense gather, 1 sparse when the coarse level's
                               // responses are at most kSparseDensity nonzero, 2 sparse whenever the list fits
    int list_threads = 1024;   // nonzero-response scatter: threads per block (256, 512, 1024)
    int list_segment = 16384;  // ... score-map positions per work item (4096, 8192, 16384: 16, 32, 64 KB of LDS)
    int list_bpc = 1;          // ... persistent blocks per CU (1, 2, 4, 8 within 2048 threads and 160 KB of LDS)
    int list_order = 0;        // ... items from a work counter (0) or in fixed XCD shares (1)
    int list_split = 100;      // ... percent of a template's features scattered before the positions that can
                               // still pass are kept and the rest gathered there (100: scatter them all)
    int refine_waves = 8;      // refinement: wavefronts per candidate (1, 2, 4, 8, 16)
    int refine_group = 16;     // ... features' loads in flight per wavefront (8, 16)
    int refine_bpc = 4;        // ... blocks per CU (1..32)
    int refine_layout = 0;     // ... reads the linear memories with 8-byte loads (0), the level's banded copy
                               // (1, see band_geom), the linear memories with 16-byte loads (2), the banded
                               // copy with 16-byte column loads (3), or the nibble-banded copy with 16-byte slot
                               // loads (4, see nib_band_geom)
    bool operator==(const Tuning& o) const {
        return ppl == o.ppl && waves_per_block == o.waves_per_block && blocks_per_cu == o.blocks_per_cu &&
               gather == o.gather && split == o.split && coarse_mode == o.coarse_mode &&
               list_threads == o.list_threads && list_segment == o.list_segment && list_bpc == o.list_bpc &&
               list_order == o.list_order && list_split == o.list_split &&
               refine_waves == o.refine_waves && refine_group == o.refine_group && refine_bpc == o.refine_bpc &&
               refine_layout == o.refine_layout;
    }
};

struct StageTimes {
    float preprocess_ms = 0, coarse_ms = 0, refine_ms = 0, total_ms = 0;
    int coarse_candidates = 0;
    int final_candidates = 0;
    int coarse_sparse = 0;     // the coarse stage ran over the nonzero-response list
    int coarse_nonzero = -1;   // nonzero responses of the coarse level (-1: not listed)
    long long coarse_entries = 0;  // list entries the nonzero-response path read
    long long refine_features = 0; // features of the candidates the finest level's refinement scored, summed
};

float score_of(int raw, int n);
int raw_min_coarse(int n, float thr);
int raw_min_refine(int n, float thr);

// The detector's bank packed for the device; read-only while matches run.
struct Bank {
    int L = 0;
    std::vector<int> T;
    float weak = 30.f;
    int bins = 8;                     // orientations: 8, or 16 for the 16-orientation variant
    uint64_t version = 0;             // changes on every re-pack
    DevMem d_tl, d_feats;             // [g * L + l] records, packed features
    DevMem d_gclass, d_gtid;          // class rank and template id of each global template
    std::vector<TemplLevelDev> h_tl;  // host copy of the records
    std::vector<int> g_class, g_tid;  // class rank and template id of each global template
    std::vector<int> sorted_all;      // global templates by lowest-level feature count, largest first
    std::vector<std::string> names;   // sorted class names
    int size() const { return (int)g_class.size(); }
};

// Scores and the canonical order + unique of n refined candidates (line2Dup.cpp:1113-1116).
std::vector<MatchRec> canonical_matches(const Bank& bank, const Cand* cands, int n);

// An image to match: host or device pointers.
struct ImageArgs {
    const uint8_t* img = nullptr;
    int rows = 0, cols = 0, stride = 0, channels = 1;
    const uint8_t* mask = nullptr;
    bool device = false;
};

// Per-image device state and the stages of one match: a stream, the per-level images, quantized
// orientations and linear memories (kept resident between calls), the candidate work memory, and a
// cache of captured graphs of the whole batched pipeline, one per image size.
class Pipeline {
public:
    Pipeline(int device, int num_cu);
    ~Pipeline();
    Pipeline(const Pipeline&) = delete;
    Pipeline& operator=(const Pipeline&) = delete;

    hipStream_t stream() const { return stream_; }

    // Eager match (plain or batched mode).
    std::vector<MatchRec> match(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel, int mode);
    // Batched match through the captured graph for this image size: launch (asynchronous) and finish.
    void graph_launch(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel);
    // n images of one size (rows, cols, channels, mask or none) in one launch of every stage: the
    // preprocessing kernels cover all of them (one grid slice each), the coarse stage takes (image, tile,
    // template) items, the refinement every image's candidates; through one captured graph per size and n
    void graph_launch_batch(const Bank& bank, const ImageArgs* ims, int n, float thr, const std::vector<int>& sel);
    std::vector<std::vector<MatchRec>> graph_finish_batch(const Bank& bank);
    // Batched match, then the canonical list and cv_dnn::NMSBoxes on the device (host when the list is
    // longer than kFinalMax).  Returns the kept matches in NMS order; *all receives the match list.
    std::vector<MatchRec> match_nms(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel,
                                    float score_threshold, float nms_threshold, int top_k, std::vector<MatchRec>* all);
    std::vector<MatchRec> match_reuse(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel);
    std::vector<MatchRec> graph_finish(const Bank& bank);
    bool graph_pending() const { return pending_; }
    // true once the pending graph launch has completed on the device (non-blocking)
    bool graph_ready() const;
    int graphs_captured() const { return graphs_captured_; }
    // frees every device buffer and graph (an idle pipeline gives its memory back)
    void release();

    // Stage access (parity tests)
    void prepare(const Bank& bank, const ImageArgs& im);
    void read_level(int level, int what, uint8_t* out);
    LevelGeom geom(const Bank& bank, int l) const;

    size_t device_bytes() const;
    // device bytes held after running a launch of nimg images of this size (buffers only grow)
    size_t bytes_after(const Bank& bank, int rows, int cols, int channels, bool mask, int nimg = 1) const;
    // bytes of level l's banded copy for this tuning (0: the refinement reads the linear memories)
    size_t band_bytes(const Bank& bank, int l, int rows, int cols) const;
    // device bytes a pipeline holds for an image of this size (admission estimate)
    static size_t bytes_needed(const Bank& bank, int rows, int cols, int channels, bool mask, size_t cand_cap);

    Tuning tuning;
    bool timing_enabled = false;
    bool graph_timing = false;  // device span of each graph replay into last_times.total_ms
    StageTimes last_times;
    std::vector<std::string> kernel_names;
    std::vector<float> kernel_ms;
    // [0] work queue, [1] coarse count, [2..] refine counts, 16-47 refine queues, 50-51 features refined,
    // 54-55 list entries read, 56 list queue, 60-62 finalize; then per image of a batched launch: nonzero
    // count (64 + z) and list flag (80 + z)
    static constexpr int kCounterSlots = 64 + 2 * kMaxBatch;
    static constexpr size_t kPrefetchCands = 1024;
    static constexpr size_t kInitialCandCap = 1 << 16;
    static constexpr double kSparseDensity = 0.05;  // auto mode: sparse coarse path at or below this density
    // ... and when features x positions of the coarse stage, over the images of the launch, reach this (the
    // 1024^2 x 360-template stage, 1.9e8 per image, is faster on the gather kernel alone or four to a launch,
    // on the list from eight to a launch)
    static constexpr double kSparseMinWork = 1e9;
    static constexpr int kSparseQueueSlot = 56, kSparseNnzSlot = 64, kSparseFlagSlot = 64 + kMaxBatch;
    static constexpr int kSparseReadSlot = 54;  // and 55: 64-bit count of list entries read
    static constexpr int kRefineFeatSlot = 50;  // and 51: 64-bit sum of the features refined at level 0
    // 16 + 8 l .. 16 + 8 l + 7: the refinement's per-XCD candidate counters of level l < kRefineQueueLevels
    static constexpr int kRefineQueueSlot = 16, kRefineQueueLevels = 4;

private:
    struct Level {
        int rows = 0, cols = 0, T = 0;
        DevMem img, mask, q, lm, lm_nib;  // lm_nib: nibble-packed copy, top level only
        DevMem sp_rows, sp_entries;       // nonzero-response list of the top level (sparse coarse path)
        DevMem band;                      // banded (refine_layout 1, 3) or nibble-banded (4) copy of a refined level
        // bytes from one image of a batch to the next (sp_rows in ints, sp_entries in entries); band_b 0: no copy
        long long img_b = 0, mask_b = 0, q_b = 0, lm_b = 0, nib_b = 0, rows_b = 0, ent_b = 0, band_b = 0;
        bool band_nib = false;            // the copy is the nibble-banded one
        int nimg = 0;                     // images the buffers are laid out for
    };
    struct GraphKey {
        uint64_t bank_version = 0;
        int rows = 0, cols = 0, channels = 0, nimg = 1;
        bool aligned = false;  // every level-0 input 4-byte aligned with a stride divisible by 4
        bool mask = false;
        uint32_t thr_bits = 0;
        uint64_t sel_sig = 0;
        Tuning tuning;
        size_t cand_cap = 0;
        bool operator==(const GraphKey& o) const {
            return bank_version == o.bank_version && rows == o.rows && cols == o.cols && channels == o.channels &&
                   nimg == o.nimg && aligned == o.aligned &&
                   mask == o.mask && thr_bits == o.thr_bits && sel_sig == o.sel_sig && tuning == o.tuning &&
                   cand_cap == o.cand_cap;
        }
    };
    struct GraphEntry {
        GraphKey key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    GraphKey make_key(const Bank& bank, const ImageArgs& im, float thr, const std::vector<int>& sel, int nimg = 1) const;

    bool ensure_levels(const Bank& bank, int rows, int cols, int channels, bool mask, int nimg = 1);
    void ensure_work();
    void ensure_static(const Bank& bank, float thr, const std::vector<int>& sel);
    // nimg images; `src` (device) holds their level-0 pixels when set, else im.img is the one image
    void enqueue_preprocess(const Bank& bank, const ImageArgs& im, int nimg = 1, const SrcTable* src = nullptr);
    void enqueue_match(const Bank& bank, int nsel, int mode, int nimg = 1);
    void enqueue_head_copy();
    // the launch's candidates per image (nimg lists), canonical order
    std::vector<std::vector<MatchRec>> decode_batch(const Bank& bank, bool* overflow, int nimg);
    std::vector<MatchRec> decode(const Bank& bank, bool* overflow);
    void clear_graphs();
    void kt_begin(const std::string& name);
    void kt_end();
    void kt_collect();
    int* counters() const;
    Cand* out_cands() const;
    Cand* ping() const;
    Cand* pong() const;

    int device_ = 0, num_cu_ = 1;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t done_ = nullptr;
    std::vector<hipEvent_t> kt_ev_;
    std::vector<std::string> kt_names_;
    int kt_used_ = 0;

    std::vector<Level> levels_;
    int cur_rows_ = 0, cur_cols_ = 0, cur_channels_ = 1, cur_bins_ = 8;
    bool cur_device_input_ = false;
    DevMem d_work_, d_order_, d_rawmin_c_, d_rawmin_r_, d_in_, d_in_mask_, d_bins_;
    DevMem d_src_;        // SrcTable of the level-0 inputs of a graph launch
    HostPinned h_src_;
    bool src_aligned_ = false;  // the table's images allow aligned dword reads (set before a capture)
    size_t cand_cap_ = 0;
    HostPinned h_stage_, h_out_;
    HostPinned h_in_;  // pinned copy of a host input for a graph launch (one DMA instead of per-row copies)
    long long coarse_features_ = 0;  // coarse-level features of the selection (set with the order)
    bool pending_timed_ = false;  // the pending graph launch is bracketed by ev_[0], ev_[1]
    bool eager_timed_ = false;    // the pending launch ran its stages eagerly with per-kernel events
    bool nib_valid_ = false;  // the coarse level's nibble copy belongs to the last prepared image
    int nib_T_ = 0;
    // resident per-bank state (template order, integer thresholds)
    uint64_t static_bank_ = ~0ull;
    float static_thr_ = 0.f;
    std::vector<int> static_sel_;
    bool static_valid_ = false;
    // graphs
    std::vector<GraphEntry> graphs_;
    int graphs_captured_ = 0;
    GraphKey last_key_;
    bool last_key_valid_ = false;
    bool pending_ = false;
    int pending_n_ = 0;
    std::vector<ImageArgs> pending_ims_;
    ImageArgs pending_im_{};
    float pending_thr_ = 0.f;
    std::vector<int> pending_sel_;
};

class Detector {
public:
    Detector();
    explicit Detector(const std::vector<int>& T);
    Detector(int num_features, const std::vector<int>& T, float weak, float strong);
    ~Detector();
    Detector(const Detector&) = delete;
    Detector& operator=(const Detector&) = delete;

    // modality (ColorGradient, line2Dup.h:88-106)
    float weak_threshold = 30.f;
    int num_features = 63;
    float strong_threshold = 60.f;
    int pyramid_levels = 2;
    std::vector<int> T_at_level{4, 8};
    TemplatesMap class_templates;
    // 8 orientations (line2Dup.cpp), or 16 for the variant of test/ori_16bit_experiment/line2Dup_16bit_ori.cpp:
    // 16-bit quantization and spread, its compiled response table, 3-channel images only, and its own
    // feature selection; set before any template is added
    int orientation_bins = 8;

    int add_template(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask,
                     const std::string& class_id, int num_features_override);
    int add_template_rotate(const std::string& class_id, int zero_id, float theta, float cx, float cy);
    int add_template_pyramid(const std::string& class_id, const TemplatePyramid& tp);
    // addTemplate(shapes.src_of(info), class_id, shapes.mask_of(info), num_features[i]) for every
    // (angle, scale) info of a shapeInfo_producer source, in order: the warps and gradients run on
    // the GPU, feature selection on host threads.  Returns the template ids (-1: refused).
    std::vector<int> add_templates_transformed(const uint8_t* img, int rows, int cols, int stride, int channels,
                                               const uint8_t* mask, const std::string& class_id,
                                               const std::vector<float>& angles, const std::vector<float>& scales,
                                               const std::vector<int>& num_features);
    // shapeInfo_producer::transform (binary_mask: mask_of's "> 0" as 0/255)
    void shape_transform(const uint8_t* img, int rows, int cols, int stride, int channels, float angle, float scale,
                         bool binary_mask, uint8_t* out);

    // img / mask are host pointers, or device pointers when device_input is set (image resident in HBM)
    std::vector<MatchRec> match(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                                const std::vector<std::string>& class_ids, const uint8_t* mask, int mode,
                                bool device_input = false);
    // Several images, batched mode: images of one size bucket go out in launches of several images
    // (images_per_launch), each through the captured graph of its size and count; `slots` pipelines
    // (streams) keep that many launches in flight.
    std::vector<std::vector<MatchRec>> match_batch(const std::vector<ImageArgs>& images, float threshold,
                                                   const std::vector<std::string>& class_ids, int slots);
    std::vector<std::string> class_names() const;
    // match followed by cv_dnn::NMSBoxes over (x, y, templ[0].width, templ[0].height) with eta 1, as the
    // reference's demos do (test.cpp:428-451), on the device
    std::vector<MatchRec> match_nms(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                                    const std::vector<std::string>& class_ids, const uint8_t* mask, bool device_input,
                                    float score_threshold, float nms_threshold, int top_k, std::vector<MatchRec>* all);

    const std::vector<Template>& get_templates(const std::string& class_id, int template_id) const;
    int num_templates() const;
    int num_templates(const std::string& class_id) const;

    // persistence (line2Dup.cpp:1400-1509)
    std::string read_class_text(const std::string& text, const std::string& class_id_override);
    std::string write_class_text(const std::string& class_id) const;
    void read_classes(const std::vector<std::string>& class_ids, const std::string& format);
    void write_classes(const std::string& format) const;
    std::string write_detector_text() const;
    void read_detector_text(const std::string& text);
    // packed binary bank file (sbm_bankfile.cpp): the whole bank and the detector's parameters
    std::string write_bank_binary() const;
    int read_bank_binary(const std::string& data);  // returns the classes added

    // device-side debug / parity helpers
    void prepare_image(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask);
    void read_level(int level, int what, uint8_t* out);
    void similarity_map(const std::string& class_id, int template_id, uint16_t* out);
    void similarity_local(const std::string& class_id, int template_id, int level, int cx, int cy, uint16_t* out);
    void gradient(const uint8_t* img, int rows, int cols, int stride, int channels, uint8_t* q, float* mag,
                  float* angle);
    void bin_table(int lo, int hi, uint8_t* out);
    int division_check();  // mismatching quotients of the gradient kernels over their domain (0 expected)

    // sbm_match_device without any device or pinned-host allocation (sbm_noalloc.hip): every per-image
    // buffer lives in the library's fixed device scratch array; refuses an image or selection that does
    // not fit it, and a bank that is not uploaded yet
    std::vector<MatchRec> match_noalloc(const uint8_t* d_img, int rows, int cols, int stride, int channels,
                                        float threshold, const std::vector<std::string>& class_ids,
                                        const uint8_t* d_mask);

    // One timed shape of the autotuner: stage 0 the gather coarse kernel, 1 the nonzero-response scatter,
    // 2 the refinement; ms is that stage's mean device time.
    struct TuneRow {
        int stage;
        Tuning tuning;
        float ms;
    };
    // Searches the launch shapes of the coarse stage on the path the device takes for this image (the gather
    // kernel's 240 shapes, or the scatter's 23), then the refinement's 112, checking that every shape returns
    // the same matches; keeps the fastest in `tuning`.  With images_per_launch > 1 it times launches of that
    // many images (the launches match_batch makes) as a stream: the images of `batch` in turn when given (all of
    // the image's size and channels), else copies of the image; when those take the list and the image alone
    // the gather kernel, it searches both (the gather shapes on single images).
    Tuning autotune(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                    const std::vector<std::string>& class_ids, bool device_input, int iters,
                    std::vector<TuneRow>* trace, const std::vector<ImageArgs>* batch = nullptr);
    std::vector<TuneRow> last_autotune;  // the rows of the last autotune

    // tuning and instrumentation (applied to every pipeline of this detector)
    Tuning tuning;
    // images per launch in match_batch (1..16; 0: half of a size bucket, so that two launches alternate)
    int images_per_launch = 0;
    bool timing_enabled = false;
    bool graph_timing = false;  // device span of each graph replay into last_times.total_ms
    const StageTimes& last_times() const;
    const std::vector<std::string>& kernel_names() const;
    const std::vector<float>& kernel_ms() const;

    void ensure_device();
    void ensure_bank();
    const Bank& bank() const { return bank_; }
    // the packed bank is in HBM and matches the host bank (nothing to upload before a match)
    bool bank_ready() const { return pipe_ && !bank_dirty_; }
    Pipeline& pipeline();
    hipStream_t stream();
    size_t device_bytes() const;
    // device bytes held after running an image of this size (buffers only grow)
    size_t bytes_after(const Bank& bank, int rows, int cols, int channels, bool mask) const;
    // gives back the device buffers of the detector's own pipelines (the bank stays); they grow again on use
    void release_match_buffers();
    int device() const { return device_; }
    int num_cu() const { return num_cu_; }
    void check_sizes(int rows, int cols) const;
    void check_channels(int channels) const;  // 1 or 3; 3 only with 16 orientations
    // templates to match, as global indices (class order, or sorted by feature count)
    void selection(const std::vector<std::string>& class_ids, bool sorted, std::vector<int>& out) const;

private:
    Bank bank_;
    bool bank_dirty_ = true;
    uint64_t bank_versions_ = 0;
    std::unique_ptr<Pipeline> pipe_;
    std::vector<std::unique_ptr<Pipeline>> batch_pipes_;
    int device_ = -1, num_cu_ = 0;
    StageTimes empty_times_;
    std::vector<std::string> empty_names_;
    const Pipeline* last_pipe_ = nullptr;
    const Pipeline* last_pipe() const;
    std::vector<float> empty_ms_;
};

// YAML in OpenCV FileStorage layout (sbm_yaml.cpp)
namespace yaml {
struct Node {
    enum Kind { kNone, kScalar, kSeq, kMap } kind = kNone;
    std::string scalar;
    std::vector<Node> seq;
    std::vector<std::pair<std::string, Node>> map;
    const Node* get(const std::string& key) const;
    int as_int() const;
    double as_real() const;
};
Node parse(const std::string& text);
std::string read_file(const std::string& path);   // plain or gzip
void write_file(const std::string& path, const std::string& data);  // gzip when path ends in .gz
std::string quote_string(const std::string& s);
std::string real_to_string(double v);
}  // namespace yaml

// Host halves of training (line2Dup.cpp:103-200, 425-512)
bool extract_template(const float* mag, const uint8_t* angle, const float* angle_ori, const uint8_t* mask, int rows,
                      int cols, float strong, int num_features, std::vector<Feature>& out);
// The 16-orientation variant (line2Dup_16bit_ori.cpp:167-210, 421-498): 16-bit angles, no exhaustive
// mode, selection stops at exactly num_features
bool extract_template16(const float* mag, const uint16_t* angle, const uint8_t* mask, int rows, int cols, float strong,
                        int num_features, std::vector<Feature>& out);
void crop_templates(std::vector<Template>& tp);

// Makes `dev` current for the scope's lifetime.
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev) {
        if (dev >= 0 && hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceScope() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// shapeInfo_producer helpers (line2Dup.h:209-332; cv::getRotationMatrix2D and the inversion
// cv::warpAffine applies)
void rotation_matrix2d(float cx, float cy, double angle_deg, double scale, double M[6]);
void invert_affine(const double M[6], double inv[6]);
void produce_infos(const std::vector<float>& angle_range, const std::vector<float>& scale_range, float angle_step,
                   float scale_step, float eps, std::vector<float>& angles, std::vector<float>& scales);
std::string write_infos_text(const std::vector<float>& angles, const std::vector<float>& scales);
void read_infos_text(const std::string& text, std::vector<float>& angles, std::vector<float>& scales);

// NMS (test.cpp:31-121)
std::vector<int> nms_boxes(const int32_t* boxes, const float* scores, int n, float score_threshold,
                           float nms_threshold, float eta, int top_k);

}  // namespace sbm
