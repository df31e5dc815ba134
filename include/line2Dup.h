// line2Dup.h -- drop-in C++ API for line2Dup's template matcher on MI355X.
//
// Mirrors the reference's public interface (line2Dup.h:10-205: Feature, Template, Match, Detector and
// its addTemplate / addTemplate_rotate / match / readClasses / writeClasses / getters; line2Dup.h:209-332:
// shape_based_matching::shapeInfo_producer) on top of the C ABI of libsbm_hip.so (sbm_hip.h).
// Header-only; link with -lsbm_hip.
//
// Differences a caller sees: images are passed as ImageView (8-bit, 1 or 3 channels, row stride)
// instead of cv::Mat (a cv::Mat overload is provided when OpenCV's headers are present), errors are
// thrown as line2Dup::Error carrying the library's status code, readClass/writeClass take file paths
// instead of cv::FileNode/FileStorage, and match() returns the canonical order (similarity desc,
// template_id asc, class asc, y asc, x asc) where the reference leaves ties in unspecified order.
//
// Define LINE2DUP_16BIT_ORI before including this header to get the 16-orientation variant
// (test/ori_16bit_experiment/line2Dup_16bit_ori.cpp), which the reference builds in place of line2Dup.cpp:
// the same classes, with the variant's constructors and matching.
#ifndef LINE2DUP_MI355X_H
#define LINE2DUP_MI355X_H

#include <stdint.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sbm_hip.h"

#if defined(__has_include)
#if __has_include(<opencv2/core/core.hpp>)
#include <opencv2/core/core.hpp>
#define LINE2DUP_HAVE_OPENCV 1
#endif
#endif

namespace line2Dup {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != SBM_OK) throw Error(rc, sbm_last_error());
}

// Feature{x, y, label, theta} (line2Dup.h:13-26)
struct Feature {
    int x;
    int y;
    int label;
    float theta;
    Feature() : x(0), y(0), label(0), theta(0.f) {}
    Feature(int x_, int y_, int label_) : x(x_), y(y_), label(label_), theta(0.f) {}
};

// Template{width, height, tl_x, tl_y, pyramid_level, features} (line2Dup.h:28-39)
struct Template {
    int width = 0;
    int height = 0;
    int tl_x = 0;
    int tl_y = 0;
    int pyramid_level = 0;
    std::vector<Feature> features;
};

// Match (line2Dup.h:108-141): ordering and equality as the reference.
struct Match {
    Match() {}
    Match(int x_, int y_, float similarity_, const std::string& class_id_, int template_id_)
        : x(x_), y(y_), similarity(similarity_), class_id(class_id_), template_id(template_id_) {}
    bool operator<(const Match& rhs) const {
        if (similarity != rhs.similarity) return similarity > rhs.similarity;
        return template_id < rhs.template_id;
    }
    bool operator==(const Match& rhs) const {
        return x == rhs.x && y == rhs.y && similarity == rhs.similarity && class_id == rhs.class_id;
    }
    int x = 0;
    int y = 0;
    float similarity = 0.f;
    std::string class_id;
    int template_id = 0;
};

// 8-bit image: rows x cols, 1 (gray) or 3 (BGR) channels, row stride in bytes.
struct ImageView {
    const uint8_t* data = nullptr;
    int rows = 0, cols = 0, stride = 0, channels = 1;
    ImageView() {}
    ImageView(const uint8_t* d, int r, int c, int ch = 1, int s = 0)
        : data(d), rows(r), cols(c), stride(s ? s : c * ch), channels(ch) {}
#ifdef LINE2DUP_HAVE_OPENCV
    ImageView(const cv::Mat& m)  // NOLINT: implicit, so cv::Mat callers compile unchanged
        : data(m.data), rows(m.rows), cols(m.cols), stride((int)m.step[0]), channels(m.channels()) {}
#endif
};

// Owning 8-bit image (what shapeInfo_producer returns in place of cv::Mat).
struct Image {
    int rows = 0, cols = 0, channels = 1;
    std::vector<uint8_t> data;
    Image() {}
    Image(int r, int c, int ch) : rows(r), cols(c), channels(ch), data((size_t)r * c * ch) {}
    explicit Image(ImageView v) : Image(v.rows, v.cols, v.channels) {
        for (int y = 0; y < rows; ++y)
            std::copy(v.data + (size_t)y * v.stride, v.data + (size_t)y * v.stride + (size_t)cols * channels,
                      data.begin() + (size_t)y * cols * channels);
    }
    bool empty() const { return data.empty(); }
    operator ImageView() const { return ImageView(data.data(), rows, cols, channels); }  // NOLINT
#ifdef LINE2DUP_HAVE_OPENCV
    operator cv::Mat() const {  // NOLINT
        return cv::Mat(rows, cols, channels == 3 ? CV_8UC3 : CV_8UC1, (void*)data.data()).clone();
    }
#endif
};

struct Point2f {
    float x = 0.f, y = 0.f;
    Point2f() {}
    Point2f(float x_, float y_) : x(x_), y(y_) {}
};

class Detector {
public:
#ifdef LINE2DUP_16BIT_ORI
    // The 16-orientation variant, built like the reference builds test/ori_16bit_experiment/line2Dup_16bit_ori.cpp
    // in place of line2Dup.cpp: Detector() T = {5, 8}, Detector(T), Detector(num_features, T); weak/strong are the
    // variant's fixed 10/55 (line2Dup_16bit_ori.cpp:500-512, 1036-1056)
    Detector() { check(sbm_detector_create16(0, nullptr, 0, &h_)); }
    explicit Detector(const std::vector<int>& T) { check(sbm_detector_create16(0, T.data(), (int)T.size(), &h_)); }
    Detector(int num_features, const std::vector<int>& T) {
        check(sbm_detector_create16(num_features, T.data(), (int)T.size(), &h_));
    }
#else
    // Detector(): T = {4, 8}, 63 features, weak/strong 30/60 (line2Dup.cpp:1027-1033)
    Detector() { check(sbm_detector_create(0, nullptr, 0, 30.f, 60.f, &h_)); }
    // Detector(T) (line2Dup.cpp:1035-1040)
    explicit Detector(const std::vector<int>& T) { check(sbm_detector_create(0, T.data(), (int)T.size(), 30.f, 60.f, &h_)); }
    // Detector(num_features, T, weak, strong) (line2Dup.cpp:1042-1047)
    Detector(int num_features, const std::vector<int>& T, float weak_thresh = 30.0f, float strong_thresh = 60.0f) {
        check(sbm_detector_create(num_features, T.data(), (int)T.size(), weak_thresh, strong_thresh, &h_));
    }
#endif
    ~Detector() {
        if (h_) sbm_detector_destroy(h_);
    }
    Detector(const Detector&) = delete;
    Detector& operator=(const Detector&) = delete;
    Detector(Detector&& o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    Detector& operator=(Detector&& o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }

    // match(sources, threshold, class_ids, masks) (line2Dup.h:154-156)
    std::vector<Match> match(ImageView sources, float threshold, const std::vector<std::string>& class_ids = {},
                             const uint8_t* mask = nullptr, int mode = SBM_MODE_BATCHED) const {
        std::vector<const char*> ids;
        for (const auto& c : class_ids) ids.push_back(c.c_str());
        sbm_matches* m = nullptr;
        check(sbm_match(h_, sources.data, sources.rows, sources.cols, sources.stride, sources.channels, threshold,
                        ids.empty() ? nullptr : ids.data(), (int)ids.size(), mask, mode, &m));
        std::vector<Match> out;
        int n = 0;
        sbm_matches_count(m, &n);
        std::vector<int32_t> xyct((size_t)n * 4 + 4);
        std::vector<float> sim((size_t)n + 1);
        int got = 0;
        sbm_matches_copy(m, xyct.data(), sim.data(), n, &got);
        std::vector<std::string> names;
        char buf[4096];
        for (int ci = 0; sbm_matches_class_id(m, ci, buf, sizeof(buf)) == SBM_OK; ++ci) names.push_back(buf);
        out.reserve(n);
        for (int i = 0; i < got; ++i)
            out.emplace_back(xyct[4 * i], xyct[4 * i + 1], sim[i], names.at(xyct[4 * i + 2]), xyct[4 * i + 3]);
        sbm_matches_free(m);
        return out;
    }

    // addTemplate(sources, class_id, object_mask, num_features) (line2Dup.h:158-159); -1 on "too few features"
    int addTemplate(ImageView sources, const std::string& class_id, const uint8_t* object_mask, int num_features = 0) {
        int id = -1;
        check(sbm_add_template(h_, sources.data, sources.rows, sources.cols, sources.stride, sources.channels,
                               object_mask, class_id.c_str(), num_features, &id));
        return id;
    }

    // the same with the mask as an image (what shapeInfo_producer::mask_of returns; empty = no mask)
    int addTemplate(ImageView sources, const std::string& class_id, ImageView object_mask, int num_features = 0) {
        if (object_mask.data && (object_mask.rows != sources.rows || object_mask.cols != sources.cols ||
                                 object_mask.channels != 1))
            throw Error(SBM_EINVAL, "mask must be one channel of the source's size");
        std::vector<uint8_t> m;
        if (object_mask.data) m = Image(object_mask).data;
        return addTemplate(sources, class_id, m.empty() ? nullptr : m.data(), num_features);
    }

    // addTemplate_rotate(class_id, zero_id, theta, center) (line2Dup.h:161)
    int addTemplate_rotate(const std::string& class_id, int zero_id, float theta, Point2f center) {
        int id = -1;
        check(sbm_add_template_rotate(h_, class_id.c_str(), zero_id, theta, center.x, center.y, &id));
        return id;
    }

    int getT(int pyramid_level) const {
        int v = 0;
        check(sbm_get_T(h_, pyramid_level, &v));
        return v;
    }
    int pyramidLevels() const {
        int v = 0;
        check(sbm_pyramid_levels(h_, &v));
        return v;
    }
    // getTemplates(class_id, template_id) (line2Dup.h:169); returned by value
    std::vector<Template> getTemplates(const std::string& class_id, int template_id) const {
        std::vector<Template> out(pyramidLevels());
        for (int l = 0; l < (int)out.size(); ++l) {
            Template& t = out[l];
            int n = 0;
            check(sbm_get_template(h_, class_id.c_str(), template_id, l, &t.width, &t.height, &t.tl_x, &t.tl_y,
                                   &t.pyramid_level, &n));
            std::vector<int32_t> xyl((size_t)n * 3 + 3);
            std::vector<float> th((size_t)n + 1);
            check(sbm_get_template_features(h_, class_id.c_str(), template_id, l, xyl.data(), th.data(), n));
            t.features.resize(n);
            for (int i = 0; i < n; ++i) {
                t.features[i] = Feature(xyl[3 * i], xyl[3 * i + 1], xyl[3 * i + 2]);
                t.features[i].theta = th[i];
            }
        }
        return out;
    }
    int numTemplates() const {
        int v = 0;
        check(sbm_num_templates(h_, nullptr, &v));
        return v;
    }
    int numTemplates(const std::string& class_id) const {
        int v = 0;
        check(sbm_num_templates(h_, class_id.c_str(), &v));
        return v;
    }
    int numClasses() const {
        int v = 0;
        check(sbm_num_classes(h_, &v));
        return v;
    }
    std::vector<std::string> classIds() const {
        std::vector<std::string> ids;
        char buf[4096];
        for (int i = 0, n = numClasses(); i < n; ++i) {
            check(sbm_class_id(h_, i, buf, sizeof(buf)));
            ids.push_back(buf);
        }
        return ids;
    }

    // readClasses / writeClasses (line2Dup.h:183-185)
    void readClasses(const std::vector<std::string>& class_ids, const std::string& format = "templates_%s.yml.gz") {
        std::vector<const char*> ids;
        for (const auto& c : class_ids) ids.push_back(c.c_str());
        check(sbm_read_classes(h_, ids.data(), (int)ids.size(), format.c_str()));
    }
    void writeClasses(const std::string& format = "templates_%s.yml.gz") const {
        check(sbm_write_classes(h_, format.c_str()));
    }
    // readClass / writeClass (line2Dup.h:180-181) on files
    std::string readClass(const std::string& path, const std::string& class_id_override = "") {
        char buf[4096];
        check(sbm_read_class_file(h_, path.c_str(), class_id_override.c_str(), buf, sizeof(buf)));
        return buf;
    }
    void writeClass(const std::string& class_id, const std::string& path) const {
        check(sbm_write_class_file(h_, class_id.c_str(), path.c_str()));
    }
    // read / write (line2Dup.h:177-178) on files
    void read(const std::string& path) { check(sbm_read_detector_file(h_, path.c_str())); }
    void write(const std::string& path) const { check(sbm_write_detector_file(h_, path.c_str())); }
    // No reference counterpart: the whole bank as one packed binary file, for a serving detector's start-up
    void saveBank(const std::string& path) const { check(sbm_save_bank(h_, path.c_str())); }
    int loadBank(const std::string& path) {
        int n = 0;
        check(sbm_load_bank(h_, path.c_str(), &n));
        return n;
    }

    sbm_detector* handle() const { return h_; }

private:
    sbm_detector* h_ = nullptr;
};

}  // namespace line2Dup

namespace shape_based_matching {

// shapeInfo_producer (line2Dup.h:209-332): the (angle, scale) grid of a training image and the
// rotated/scaled images and masks for it.  transform() runs on the GPU (bit-exact with OpenCV's
// fixed-point warpAffine); src_of / mask_of return owning images.
class shapeInfo_producer {
public:
    line2Dup::Image src;
    line2Dup::Image mask;

    std::vector<float> angle_range;
    std::vector<float> scale_range;

    float angle_step = 15;
    float scale_step = 0.5;
    float eps = 0.00001f;

    class Info {
    public:
        float angle;
        float scale;
        Info(float angle_, float scale_) : angle(angle_), scale(scale_) {}
    };
    std::vector<Info> infos;

    explicit shapeInfo_producer(line2Dup::ImageView src_, line2Dup::ImageView mask_ = line2Dup::ImageView())
        : src(src_) {
        if (!mask_.data) {  // "make sure we have masks"
            mask = line2Dup::Image(src.rows, src.cols, 1);
            std::fill(mask.data.begin(), mask.data.end(), (uint8_t)255);
        } else {
            mask = line2Dup::Image(mask_);
        }
    }

    static line2Dup::Image transform(line2Dup::ImageView src, float angle, float scale) {
        return warp(src, angle, scale, 0);
    }
    static void save_infos(std::vector<Info>& infos, std::string path = "infos.yaml") {
        std::vector<float> a, s;
        for (const Info& i : infos) {
            a.push_back(i.angle);
            s.push_back(i.scale);
        }
        line2Dup::check(sbm_save_infos(path.c_str(), a.data(), s.data(), (int)a.size()));
    }
    static std::vector<Info> load_infos(std::string path = "info.yaml") {
        int n = 0;
        line2Dup::check(sbm_load_infos(path.c_str(), nullptr, nullptr, 0, &n));
        std::vector<float> a(n), s(n);
        line2Dup::check(sbm_load_infos(path.c_str(), a.data(), s.data(), n, &n));
        std::vector<Info> out;
        for (int i = 0; i < n; ++i) out.emplace_back(a[i], s[i]);
        return out;
    }
    void produce_infos() {
        int n = 0;
        line2Dup::check(sbm_produce_infos(angle_range.data(), (int)angle_range.size(), scale_range.data(),
                                          (int)scale_range.size(), angle_step, scale_step, eps, nullptr, nullptr, 0, &n));
        std::vector<float> a(n), s(n);
        line2Dup::check(sbm_produce_infos(angle_range.data(), (int)angle_range.size(), scale_range.data(),
                                          (int)scale_range.size(), angle_step, scale_step, eps, a.data(), s.data(), n,
                                          &n));
        infos.clear();
        for (int i = 0; i < n; ++i) infos.emplace_back(a[i], s[i]);
    }
    line2Dup::Image src_of(const Info& info) { return transform(src, info.angle, info.scale); }
    // transform(mask) > 0 as 0/255
    line2Dup::Image mask_of(const Info& info) { return warp(mask, info.angle, info.scale, 1); }

private:
    static line2Dup::Image warp(line2Dup::ImageView v, float angle, float scale, int binary) {
        line2Dup::Image out(v.rows, v.cols, v.channels);
        line2Dup::check(sbm_shape_transform(nullptr, v.data, v.rows, v.cols, v.stride, v.channels, angle, scale, binary,
                                            out.data.data()));
        return out;
    }
};

}  // namespace shape_based_matching

#endif  // LINE2DUP_MI355X_H
