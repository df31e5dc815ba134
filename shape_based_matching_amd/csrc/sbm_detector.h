// Host-side model of line2Dup::Detector (line2Dup.h:13-205) and the device pipeline behind match().
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
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
};

struct StageTimes {
    float preprocess_ms = 0, coarse_ms = 0, refine_ms = 0, total_ms = 0;
    long long coarse_gather_bytes = 0;  // sum over templates of in-bounds features x template positions
    int coarse_candidates = 0;
    int final_candidates = 0;
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

    int add_template(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask,
                     const std::string& class_id, int num_features_override);
    int add_template_rotate(const std::string& class_id, int zero_id, float theta, float cx, float cy);
    int add_template_pyramid(const std::string& class_id, const TemplatePyramid& tp);

    // img / mask are host pointers, or device pointers when device_input is set (image resident in HBM)
    std::vector<MatchRec> match(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                                const std::vector<std::string>& class_ids, const uint8_t* mask, int mode,
                                bool device_input = false);
    std::vector<std::string> class_names() const;

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

    // device-side debug / parity helpers
    void prepare_image(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask);
    void read_level(int level, int what, uint8_t* out);
    void similarity_map(const std::string& class_id, int template_id, uint16_t* out);
    void similarity_local(const std::string& class_id, int template_id, int level, int cx, int cy, uint16_t* out);
    void gradient(const uint8_t* img, int rows, int cols, int stride, int channels, uint8_t* q, float* mag,
                  float* angle);
    void bin_table(int lo, int hi, uint8_t* out);

    Tuning tuning;
    StageTimes last_times;
    // per-kernel device times of the last match when timing is enabled (HIP events on `stream`)
    std::vector<std::string> kernel_names;
    std::vector<float> kernel_ms;
    bool timing_enabled = false;
    int device = -1;
    hipStream_t stream = nullptr;
    size_t hbm_budget = 0;  // 0 = unlimited

    // bank in HBM
    void ensure_bank();
    void invalidate_bank() { bank_dirty_ = true; }
    size_t device_bytes() const;

    void ensure_device();

    Tuning autotune(const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
                    const std::vector<std::string>& class_ids, bool device_input, int iters,
                    std::vector<std::pair<Tuning, float>>* trace);

private:
    struct Level {
        int rows = 0, cols = 0;
        DevMem img, mask, q, lm;
    };
    void check_sizes(int rows, int cols) const;
    void run_preprocess(const uint8_t* img, int rows, int cols, int stride, int channels, const uint8_t* mask,
                        bool device_input, int* zero_counters);
    void ensure_work();
    int* counters() const;
    Cand* out_cands() const;
    Cand* ping() const;
    Cand* pong() const;
    static constexpr int kCounterSlots = 64;
    static constexpr size_t kPrefetchCands = 1024;
    std::vector<int> sel_cached_;
    LevelGeom geom(int l) const;
    void selection(const std::vector<std::string>& class_ids, bool sorted, std::vector<int>& out) const;
    void thresholds(float threshold);
    void kt_begin(const std::string& name);
    void kt_end();
    void kt_collect();

    bool bank_dirty_ = true;
    std::vector<int> g_class_, g_tid_;           // per global template
    std::vector<std::string> g_names_;           // sorted class names of the packed bank
    std::vector<TemplLevelDev> h_tl_;            // [g * L + l]
    std::vector<int> sorted_all_;                // global templates by lowest-level feature count desc
    DevMem d_tl_, d_feats_, d_order_, d_rawmin_c_, d_rawmin_r_, d_work_;
    size_t cand_cap_ = 0;
    float thr_cached_ = 0.f;
    bool thr_valid_ = false;
    std::vector<Level> levels_;
    int cur_rows_ = 0, cur_cols_ = 0, cur_channels_ = 1;
    bool cur_device_input_ = false;
    DevMem d_src_;
    HostPinned h_stage_;
    hipEvent_t ev_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> kt_ev_;
    std::vector<std::string> kt_names_;
    int kt_used_ = 0;
    int num_cu_ = 0;
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
void crop_templates(std::vector<Template>& tp);

// NMS (test.cpp:31-121)
std::vector<int> nms_boxes(const int32_t* boxes, const float* scores, int n, float score_threshold,
                           float nms_threshold, float eta, int top_k);

float score_of(int raw, int n);

}  // namespace sbm
