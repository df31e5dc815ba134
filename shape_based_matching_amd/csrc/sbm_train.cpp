// Host halves of template training and the post-match box NMS.
// Reference: line2Dup.cpp:103-200 (cropTemplates, selectScatteredFeatures), 425-512 (extractTemplate);
// test.cpp:31-121 (cv_dnn::NMSBoxes).
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>

#include "sbm_detector.h"

#pragma clang fp contract(off)

namespace sbm {

namespace {
struct Candidate {
    Feature f;
    float score;
};

int label_of(int q) {  // getLabel, line2Dup.cpp:32-56
    switch (q) {
        case 1: return 0;
        case 2: return 1;
        case 4: return 2;
        case 8: return 3;
        case 16: return 4;
        case 32: return 5;
        case 64: return 6;
        case 128: return 7;
        default: throw Error(kEInval, "Invalid value of quantized parameter");
    }
}

int label16_of(int q) {  // getLabel of the 16-orientation variant, line2Dup_16bit_ori.cpp:32-72
    for (int b = 0; b < 16; ++b)
        if (q == (1 << b)) return b;
    throw Error(kEInval, "Invalid value of quantized parameter");
}

// selectScatteredFeatures of the 16-orientation variant (line2Dup_16bit_ori.cpp:167-210): passes over the
// candidates, relaxing the distance by one after each, until exactly num_features are taken.  It ends:
// the caller has at least num_features candidates at distinct pixels, and a pass at distance 1 takes
// every one not yet taken.
void select_scattered16(const std::vector<Candidate>& cands, std::vector<Feature>& features, size_t num_features,
                        float distance) {
    features.clear();
    float dsq = distance * distance;
    size_t i = 0;
    while (features.size() < num_features) {
        const Candidate& c = cands[i];
        bool keep = true;
        for (size_t j = 0; j < features.size() && keep; ++j) {
            const Feature& f = features[j];
            keep = (float)((c.f.x - f.x) * (c.f.x - f.x) + (c.f.y - f.y) * (c.f.y - f.y)) >= dsq;
        }
        if (keep) features.push_back(c.f);
        if (++i == cands.size()) {
            i = 0;
            distance -= 1.0f;
            dsq = distance * distance;
        }
    }
}

// selectScatteredFeatures (line2Dup.cpp:151-200)
void select_scattered(const std::vector<Candidate>& cands, std::vector<Feature>& features, size_t num_features,
                      float distance) {
    features.clear();
    float dsq = distance * distance;
    int i = 0;
    bool first = true;
    for (;;) {
        const Candidate& c = cands[i];
        bool keep = true;
        for (size_t j = 0; j < features.size() && keep; ++j) {
            const Feature& f = features[j];
            keep = (float)((c.f.x - f.x) * (c.f.x - f.x) + (c.f.y - f.y) * (c.f.y - f.y)) >= dsq;
        }
        if (keep) features.push_back(c.f);
        if (++i == (int)cands.size()) {
            const bool num_ok = features.size() >= num_features;
            if (first) {
                if (num_ok) {
                    features.clear();
                    i = 0;
                    distance += 1.0f;
                    dsq = distance * distance;
                    continue;
                }
                first = false;
            }
            i = 0;
            distance -= 1.0f;
            dsq = distance * distance;
            if (num_ok || distance < 3) break;
        }
    }
}
}  // namespace

// The candidate stage of extractTemplate (line2Dup.cpp:425-497): 5x5 magnitude NMS over the eroded mask,
// strong threshold, stable sort by score.  angle8 or angle16 (exactly one non-null) holds the quantized
// single-bit orientations.  The 16-orientation variant's scan (line2Dup_16bit_ori.cpp:421-478) lacks the
// outer "if(!is_max) break;", which changes nothing: once is_max is false the score stays 0.
static std::vector<Candidate> collect_candidates(const float* mag, const uint8_t* angle8, const uint16_t* angle16,
                                                 const float* angle_ori, const uint8_t* mask, int rows, int cols,
                                                 float strong) {
    std::vector<uint8_t> lmask;
    std::vector<Candidate> cands;
    if (mask) {  // erode(mask, local_mask, 3x3, BORDER_REPLICATE)
        lmask.resize((size_t)rows * cols);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) {
                uint8_t m = 255;
                for (int dr = -1; dr <= 1; ++dr)
                    for (int dc = -1; dc <= 1; ++dc) {
                        const int rr = std::min(std::max(r + dr, 0), rows - 1), cc = std::min(std::max(c + dc, 0), cols - 1);
                        m = std::min(m, mask[(size_t)rr * cols + cc]);
                    }
                lmask[(size_t)r * cols + c] = m;
            }
    }
    const float thr_sq = strong * strong;
    const int K = 5;
    std::vector<uint8_t> valid((size_t)rows * cols, 255);
    for (int r = K / 2; r < rows - K / 2; ++r)
        for (int c = K / 2; c < cols - K / 2; ++c) {
            if (mask && !lmask[(size_t)r * cols + c]) continue;
            float score = 0;
            if (valid[(size_t)r * cols + c] > 0) {
                score = mag[(size_t)r * cols + c];
                bool is_max = true;
                for (int ro = -K / 2; ro <= K / 2 && is_max; ++ro)
                    for (int co = -K / 2; co <= K / 2; ++co) {
                        if (ro == 0 && co == 0) continue;
                        if (score < mag[(size_t)(r + ro) * cols + c + co]) {
                            score = 0;
                            is_max = false;
                            break;
                        }
                    }
                if (is_max)
                    for (int ro = -K / 2; ro <= K / 2; ++ro)
                        for (int co = -K / 2; co <= K / 2; ++co) {
                            if (ro == 0 && co == 0) continue;
                            valid[(size_t)(r + ro) * cols + c + co] = 0;
                        }
            }
            const int a = angle16 ? (int)angle16[(size_t)r * cols + c] : (int)angle8[(size_t)r * cols + c];
            if (score > thr_sq && a > 0) {
                Candidate cd;
                cd.f.x = c;
                cd.f.y = r;
                cd.f.label = angle16 ? label16_of(a) : label_of(a);
                cd.f.theta = angle_ori ? angle_ori[(size_t)r * cols + c] : 0.f;
                cd.score = score;
                cands.push_back(cd);
            }
        }
    std::stable_sort(cands.begin(), cands.end(), [](const Candidate& a, const Candidate& b) { return a.score > b.score; });
    return cands;
}

// extractTemplate (line2Dup.cpp:425-512): candidates, then the scattered selection; fewer candidates than
// num_features still make a template ("exhaustive mode") unless there are at most four.
bool extract_template(const float* mag, const uint8_t* angle, const float* angle_ori, const uint8_t* mask, int rows,
                      int cols, float strong, int num_features, std::vector<Feature>& out) {
    out.clear();
    if (num_features <= 0) return false;  // the reference divides by it (line2Dup.cpp:498)
    const std::vector<Candidate> cands = collect_candidates(mag, angle, nullptr, angle_ori, mask, rows, cols, strong);
    if (cands.size() < (size_t)num_features) {
        if (cands.size() <= 4) return false;  // "too few features, abort"
    }
    const float distance = static_cast<float>(cands.size() / (size_t)num_features + 1);
    select_scattered(cands, out, (size_t)num_features, distance);
    return true;
}

// extractTemplate of the 16-orientation variant (line2Dup_16bit_ori.cpp:421-498): fewer candidates than
// num_features refuse the template; the selection takes exactly num_features.
bool extract_template16(const float* mag, const uint16_t* angle, const uint8_t* mask, int rows, int cols, float strong,
                        int num_features, std::vector<Feature>& out) {
    out.clear();
    if (num_features <= 0) return false;  // the reference divides by it (line2Dup_16bit_ori.cpp:484)
    const std::vector<Candidate> cands = collect_candidates(mag, nullptr, angle, nullptr, mask, rows, cols, strong);
    if (cands.size() < (size_t)num_features) return false;
    const float distance = static_cast<float>(cands.size() / (size_t)num_features + 1);
    select_scattered16(cands, out, (size_t)num_features, distance);
    return true;
}

// cropTemplates (line2Dup.cpp:103-149)
void crop_templates(std::vector<Template>& templates) {
    int min_x = INT_MAX, min_y = INT_MAX, max_x = INT_MIN, max_y = INT_MIN;
    for (const Template& t : templates)
        for (const Feature& f : t.features) {
            const int x = f.x << t.pyramid_level, y = f.y << t.pyramid_level;
            min_x = std::min(min_x, x);
            min_y = std::min(min_y, y);
            max_x = std::max(max_x, x);
            max_y = std::max(max_y, y);
        }
    if (min_x % 2 == 1) --min_x;
    if (min_y % 2 == 1) --min_y;
    for (Template& t : templates) {
        t.width = (max_x - min_x) >> t.pyramid_level;
        t.height = (max_y - min_y) >> t.pyramid_level;
        t.tl_x = min_x >> t.pyramid_level;
        t.tl_y = min_y >> t.pyramid_level;
        for (Feature& f : t.features) {
            f.x -= t.tl_x;
            f.y -= t.tl_y;
        }
    }
}

// cv_dnn::NMSBoxes / NMSFast_ / rectOverlap (test.cpp:44-119)
static float rect_overlap(const int32_t* a, const int32_t* b) {
    const int aa = a[2] * a[3], ab = b[2] * b[3];
    if (aa + ab <= 0) return 1.f - 0.f;
    const int x1 = std::max(a[0], b[0]), y1 = std::max(a[1], b[1]);
    const int w = std::min(a[0] + a[2], b[0] + b[2]) - x1, h = std::min(a[1] + a[3], b[1] + b[3]) - y1;
    const double inter = (w <= 0 || h <= 0) ? 0.0 : (double)(w * h);
    const double dist = 1.0 - inter / ((double)(aa + ab) - inter);
    return 1.f - (float)dist;
}

std::vector<int> nms_boxes(const int32_t* boxes, const float* scores, int n, float score_threshold, float nms_threshold,
                           float eta, int top_k) {
    std::vector<std::pair<float, int>> v;
    for (int i = 0; i < n; ++i)
        if (scores[i] > score_threshold) v.emplace_back(scores[i], i);
    std::stable_sort(v.begin(), v.end(), [](const std::pair<float, int>& a, const std::pair<float, int>& b) {
        return a.first > b.first;
    });
    if (top_k > 0 && top_k < (int)v.size()) v.resize(top_k);
    float adaptive = nms_threshold;
    std::vector<int> idx;
    for (const auto& p : v) {
        bool keep = true;
        for (size_t k = 0; k < idx.size() && keep; ++k) keep = rect_overlap(boxes + 4 * p.second, boxes + 4 * idx[k]) <= adaptive;
        if (keep) idx.push_back(p.second);
        if (keep && eta < 1 && adaptive > 0.5) adaptive *= eta;
    }
    return idx;
}

}  // namespace sbm
