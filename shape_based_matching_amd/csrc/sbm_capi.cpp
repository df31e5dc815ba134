// extern "C" surface of libsbm_hip.so (declared in include/sbm_hip.h).
#include "../../include/sbm_hip.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "sbm_detector.h"
#include "sbm_server.h"

struct sbm_detector {
    sbm::Detector d;
    sbm_detector(int nf, const std::vector<int>& T, float weak, float strong) : d(nf, T, weak, strong) {}
};

struct sbm_image {
    sbm::DevMem img, mask;
    int rows = 0, cols = 0, channels = 1;
    bool has_mask = false;
};

struct sbm_server {
    std::unique_ptr<sbm::Server> s;
    sbm_detector* det = nullptr;
};

struct sbm_matches {
    std::vector<sbm::MatchRec> rows;
    std::vector<std::string> names;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const sbm::Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return SBM_EBUDGET;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return SBM_EINVAL;
    }
}

int copy_str(const std::string& s, char* buf, int cap) {
    if (!buf || cap <= 0) throw sbm::Error(SBM_EINVAL, "null buffer");
    if ((int)s.size() + 1 > cap) throw sbm::Error(SBM_EINVAL, "buffer too small");
    memcpy(buf, s.c_str(), s.size() + 1);
    return SBM_OK;
}

std::vector<std::string> id_list(const char* const* ids, int n) {
    std::vector<std::string> v;
    if (ids)
        for (int i = 0; i < n; ++i)
            if (ids[i]) v.emplace_back(ids[i]);
    return v;
}

void need(const void* p) {
    if (!p) throw sbm::Error(SBM_EINVAL, "null pointer argument");
}
}  // namespace

extern "C" {

const char* sbm_last_error(void) { return g_last_error.c_str(); }
const char* sbm_version(void) { return "sbm_hip 0.1 gfx950"; }

int sbm_detector_create(int num_features, const int* T, int num_levels, float weak, float strong, sbm_detector** out) {
    return guarded([&] {
        need(out);
        std::vector<int> t;
        if (T) t.assign(T, T + num_levels);
        else t = {4, 8};
        if (t.empty()) throw sbm::Error(SBM_EINVAL, "empty pyramid");
        if (t.size() > 8) throw sbm::Error(SBM_EINVAL, "at most 8 pyramid levels");  // the device counter slots
        for (int v : t)
            if (v < 1 || v > 16) throw sbm::Error(SBM_EINVAL, "T must be in [1, 16]");
        *out = new sbm_detector(num_features > 0 ? num_features : 63, t, weak, strong);
        return SBM_OK;
    });
}

int sbm_detector_create16(int num_features, const int* T, int num_levels, sbm_detector** out) {
    return guarded([&] {
        need(out);
        std::vector<int> t;
        if (T) t.assign(T, T + num_levels);
        else t = {5, 8};  // Detector() of the variant, line2Dup_16bit_ori.cpp:1036-1042
        if (t.empty()) throw sbm::Error(SBM_EINVAL, "empty pyramid");
        if (t.size() > 8) throw sbm::Error(SBM_EINVAL, "at most 8 pyramid levels");  // the device counter slots
        for (int v : t)
            if (v < 1 || v > 16) throw sbm::Error(SBM_EINVAL, "T must be in [1, 16]");
        // ColorGradient() and Detector(num_features, T): weak 10, strong 55 (line2Dup_16bit_ori.cpp:500-512, 1051-1056)
        std::unique_ptr<sbm_detector> d(new sbm_detector(num_features > 0 ? num_features : 63, t, 10.0f, 55.0f));
        d->d.orientation_bins = 16;
        *out = d.release();
        return SBM_OK;
    });
}

int sbm_orientation_bins(const sbm_detector* det, int* out) {
    return guarded([&] { need(det); need(out); *out = det->d.orientation_bins; return SBM_OK; });
}

int sbm_detector_destroy(sbm_detector* det) {
    return guarded([&] {
        delete det;
        return SBM_OK;
    });
}

int sbm_pyramid_levels(const sbm_detector* det, int* out) {
    return guarded([&] { need(det); need(out); *out = det->d.pyramid_levels; return SBM_OK; });
}
int sbm_get_T(const sbm_detector* det, int level, int* out) {
    return guarded([&] {
        need(det); need(out);
        if (level < 0 || level >= (int)det->d.T_at_level.size()) throw sbm::Error(SBM_EINVAL, "level out of range");
        *out = det->d.T_at_level[level];
        return SBM_OK;
    });
}
int sbm_get_modality(const sbm_detector* det, float* weak, int* nf, float* strong) {
    return guarded([&] {
        need(det);
        if (weak) *weak = det->d.weak_threshold;
        if (nf) *nf = det->d.num_features;
        if (strong) *strong = det->d.strong_threshold;
        return SBM_OK;
    });
}
int sbm_num_templates(const sbm_detector* det, const char* class_id, int* out) {
    return guarded([&] {
        need(det); need(out);
        *out = class_id ? det->d.num_templates(class_id) : det->d.num_templates();
        return SBM_OK;
    });
}
int sbm_num_classes(const sbm_detector* det, int* out) {
    return guarded([&] { need(det); need(out); *out = (int)det->d.class_templates.size(); return SBM_OK; });
}
int sbm_class_id(const sbm_detector* det, int index, char* buf, int cap) {
    return guarded([&] {
        need(det);
        const auto names = det->d.class_names();
        if (index < 0 || index >= (int)names.size()) throw sbm::Error(SBM_EINVAL, "class index out of range");
        return copy_str(names[index], buf, cap);
    });
}
int sbm_get_template(const sbm_detector* det, const char* class_id, int template_id, int level, int* width, int* height,
                     int* tl_x, int* tl_y, int* pyramid_level, int* num_features) {
    return guarded([&] {
        need(det); need(class_id);
        const auto& tp = det->d.get_templates(class_id, template_id);
        if (level < 0 || level >= (int)tp.size()) throw sbm::Error(SBM_EINVAL, "level out of range");
        const sbm::Template& t = tp[level];
        if (width) *width = t.width;
        if (height) *height = t.height;
        if (tl_x) *tl_x = t.tl_x;
        if (tl_y) *tl_y = t.tl_y;
        if (pyramid_level) *pyramid_level = t.pyramid_level;
        if (num_features) *num_features = (int)t.features.size();
        return SBM_OK;
    });
}
int sbm_get_template_features(const sbm_detector* det, const char* class_id, int template_id, int level, int32_t* xyl,
                              float* theta, int cap) {
    return guarded([&] {
        need(det); need(class_id);
        const auto& tp = det->d.get_templates(class_id, template_id);
        if (level < 0 || level >= (int)tp.size()) throw sbm::Error(SBM_EINVAL, "level out of range");
        const auto& f = tp[level].features;
        if ((int)f.size() > cap) throw sbm::Error(SBM_EINVAL, "buffer too small");
        for (size_t i = 0; i < f.size(); ++i) {
            if (xyl) { xyl[3 * i] = f[i].x; xyl[3 * i + 1] = f[i].y; xyl[3 * i + 2] = f[i].label; }
            if (theta) theta[i] = f[i].theta;
        }
        return SBM_OK;
    });
}

int sbm_add_template(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
                     const uint8_t* mask, const char* class_id, int num_features, int* template_id) {
    return guarded([&] {
        need(det); need(img); need(class_id); need(template_id);
        if (rows <= 0 || cols <= 0) throw sbm::Error(SBM_EINVAL, "empty image");
        *template_id = det->d.add_template(img, rows, cols, stride, channels, mask, class_id, num_features);
        return SBM_OK;
    });
}
int sbm_add_template_rotate(sbm_detector* det, const char* class_id, int zero_id, float theta, float cx, float cy,
                            int* template_id) {
    return guarded([&] {
        need(det); need(class_id); need(template_id);
        *template_id = det->d.add_template_rotate(class_id, zero_id, theta, cx, cy);
        return SBM_OK;
    });
}
int sbm_add_templates_transformed(sbm_detector* det, const uint8_t* src, int rows, int cols, int stride, int channels,
                                  const uint8_t* mask, const char* class_id, const float* angles, const float* scales,
                                  const int* num_features, int n, int* template_ids) {
    return guarded([&] {
        need(det); need(src); need(class_id);
        if (n < 0) throw sbm::Error(SBM_EINVAL, "negative count");
        if (n > 0) { need(angles); need(scales); need(template_ids); }
        if (stride < cols * channels) throw sbm::Error(SBM_EINVAL, "stride smaller than a row");
        std::vector<float> a(angles, angles + n), sc(scales, scales + n);
        std::vector<int> nf;
        if (num_features) nf.assign(num_features, num_features + n);
        const std::vector<int> ids = det->d.add_templates_transformed(src, rows, cols, stride, channels, mask, class_id, a, sc, nf);
        std::copy(ids.begin(), ids.end(), template_ids);
        return SBM_OK;
    });
}

int sbm_add_template_features(sbm_detector* det, const char* class_id, int num_levels, const int32_t* tinfo6,
                              const int32_t* xyl, const float* theta, int* template_id) {
    return guarded([&] {
        need(det); need(class_id); need(tinfo6); need(template_id);
        if (num_levels != det->d.pyramid_levels) throw sbm::Error(SBM_EINVAL, "pyramid depth differs from the detector's");
        sbm::TemplatePyramid tp(num_levels);
        for (int l = 0; l < num_levels; ++l) {
            const int32_t* t = tinfo6 + 6 * l;
            tp[l].width = t[0];
            tp[l].height = t[1];
            tp[l].tl_x = t[2];
            tp[l].tl_y = t[3];
            tp[l].pyramid_level = l;
            if (t[4] < 0) throw sbm::Error(SBM_EINVAL, "negative feature count");
            if (t[4] > 0) need(xyl);
            tp[l].features.resize(t[4]);
            for (int i = 0; i < t[4]; ++i) {
                const int32_t* f = xyl + 3 * (size_t)(t[5] + i);
                tp[l].features[i].x = f[0];
                tp[l].features[i].y = f[1];
                tp[l].features[i].label = f[2];
                tp[l].features[i].theta = theta ? theta[t[5] + i] : 0.f;
            }
        }
        *template_id = det->d.add_template_pyramid(class_id, tp);
        return SBM_OK;
    });
}

int sbm_match(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels, float threshold,
              const char* const* class_ids, int num_class_ids, const uint8_t* mask, int mode, sbm_matches** out) {
    return guarded([&] {
        need(det); need(img); need(out);
        auto* m = new sbm_matches();
        try {
            m->rows = det->d.match(img, rows, cols, stride, channels, threshold, id_list(class_ids, num_class_ids), mask, mode);
            m->names = det->d.class_names();
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return SBM_OK;
    });
}
int sbm_match_device(sbm_detector* det, const uint8_t* d_img, int rows, int cols, int stride, int channels,
                     float threshold, const char* const* class_ids, int n, const uint8_t* d_mask, int mode,
                     sbm_matches** out) {
    return guarded([&] {
        need(det); need(d_img); need(out);
        auto* m = new sbm_matches();
        try {
            m->rows = det->d.match(d_img, rows, cols, stride, channels, threshold, id_list(class_ids, n), d_mask, mode, true);
            m->names = det->d.class_names();
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return SBM_OK;
    });
}

int sbm_match_device_noalloc(sbm_detector* det, const uint8_t* d_img, int rows, int cols, int stride, int channels,
                             float threshold, const char* const* class_ids, int n, const uint8_t* d_mask,
                             int32_t* xyct, float* similarity, int cap, int* count) {
    return guarded([&] {
        need(det); need(d_img); need(count);
        if (cap < 0 || (cap > 0 && (!xyct || !similarity))) throw sbm::Error(SBM_EINVAL, "bad output arrays");
        const auto rows_out = det->d.match_noalloc(d_img, rows, cols, stride, channels, threshold,
                                                   id_list(class_ids, n), d_mask);
        const int k = std::min(cap, (int)rows_out.size());
        for (int i = 0; i < k; ++i) {
            const auto& r = rows_out[i];
            xyct[4 * i] = r.x;
            xyct[4 * i + 1] = r.y;
            xyct[4 * i + 2] = r.class_index;
            xyct[4 * i + 3] = r.template_id;
            similarity[i] = r.similarity;
        }
        *count = (int)rows_out.size();
        return SBM_OK;
    });
}

int sbm_set_images_per_launch(sbm_detector* det, int images_per_launch) {
    return guarded([&] {
        need(det);
        if (images_per_launch < 0 || images_per_launch > sbm::kMaxBatch)
            throw sbm::Error(SBM_EINVAL, "images per launch must be 0 (automatic) to 16");
        det->d.images_per_launch = images_per_launch;
        return SBM_OK;
    });
}

int sbm_image_create(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
                     const uint8_t* mask, sbm_image** out) {
    return guarded([&] {
        need(det); need(img); need(out);
        if (rows <= 0 || cols <= 0 || (channels != 1 && channels != 3) || stride < cols * channels)
            throw sbm::Error(SBM_EINVAL, "bad image geometry");
        det->d.ensure_device();
        auto* im = new sbm_image();
        try {
            im->rows = rows;
            im->cols = cols;
            im->channels = channels;
            im->img.ensure((size_t)rows * cols * channels);
            sbm::hip_check(hipMemcpy2D(im->img.p, (size_t)cols * channels, img, (size_t)stride, (size_t)cols * channels,
                                       rows, hipMemcpyHostToDevice), "H2D image");
            if (mask) {
                im->mask.ensure((size_t)rows * cols);
                sbm::hip_check(hipMemcpy(im->mask.p, mask, (size_t)rows * cols, hipMemcpyHostToDevice), "H2D mask");
                im->has_mask = true;
            }
        } catch (...) {
            delete im;
            throw;
        }
        *out = im;
        return SBM_OK;
    });
}
int sbm_image_destroy(sbm_image* im) {
    return guarded([&] { delete im; return SBM_OK; });
}
int sbm_match_image(sbm_detector* det, const sbm_image* im, float threshold, const char* const* class_ids, int n,
                    int mode, sbm_matches** out) {
    return guarded([&] {
        need(det); need(im); need(out);
        auto* m = new sbm_matches();
        try {
            m->rows = det->d.match(im->img.as<uint8_t>(), im->rows, im->cols, im->cols * im->channels, im->channels,
                                   threshold, id_list(class_ids, n), im->has_mask ? im->mask.as<uint8_t>() : nullptr,
                                   mode, true);
            m->names = det->d.class_names();
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return SBM_OK;
    });
}
namespace {
int match_nms_common(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
                     const uint8_t* mask, bool device, float threshold, const char* const* class_ids, int n,
                     float score_threshold, float nms_threshold, int top_k, sbm_matches** kept, sbm_matches** all) {
    need(det); need(img); need(kept);
    auto* k = new sbm_matches();
    sbm_matches* a = all ? new sbm_matches() : nullptr;
    try {
        k->rows = det->d.match_nms(img, rows, cols, stride, channels, threshold, id_list(class_ids, n), mask, device,
                                   score_threshold, nms_threshold, top_k, a ? &a->rows : nullptr);
        k->names = det->d.class_names();
        if (a) a->names = k->names;
    } catch (...) {
        delete k;
        delete a;
        throw;
    }
    *kept = k;
    if (all) *all = a;
    return SBM_OK;
}
}  // namespace

int sbm_match_nms(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
                  const uint8_t* mask, float threshold, const char* const* class_ids, int n, float score_threshold,
                  float nms_threshold, int top_k, sbm_matches** kept, sbm_matches** all) {
    return guarded([&] {
        return match_nms_common(det, img, rows, cols, stride, channels, mask, false, threshold, class_ids, n,
                                score_threshold, nms_threshold, top_k, kept, all);
    });
}
int sbm_match_image_nms(sbm_detector* det, const sbm_image* im, float threshold, const char* const* class_ids, int n,
                        float score_threshold, float nms_threshold, int top_k, sbm_matches** kept, sbm_matches** all) {
    return guarded([&] {
        need(im);
        return match_nms_common(det, im->img.as<uint8_t>(), im->rows, im->cols, im->cols * im->channels, im->channels,
                                im->has_mask ? im->mask.as<uint8_t>() : nullptr, true, threshold, class_ids, n,
                                score_threshold, nms_threshold, top_k, kept, all);
    });
}

int sbm_match_batch(sbm_detector* det, const sbm_image* const* images, int n, float threshold,
                    const char* const* class_ids, int n_ids, int slots, sbm_matches** outs) {
    return guarded([&] {
        need(det); need(outs);
        if (n < 0) throw sbm::Error(SBM_EINVAL, "negative image count");
        std::vector<sbm::ImageArgs> ims(n);
        for (int i = 0; i < n; ++i) {
            need(images);
            need(images[i]);
            const sbm_image* im = images[i];
            ims[i].img = im->img.as<uint8_t>();
            ims[i].rows = im->rows;
            ims[i].cols = im->cols;
            ims[i].stride = im->cols * im->channels;
            ims[i].channels = im->channels;
            ims[i].mask = im->has_mask ? im->mask.as<uint8_t>() : nullptr;
            ims[i].device = true;
        }
        auto res = det->d.match_batch(ims, threshold, id_list(class_ids, n_ids), slots);
        const auto names = det->d.class_names();
        for (int i = 0; i < n; ++i) {
            auto* m = new sbm_matches();
            m->rows = std::move(res[i]);
            m->names = names;
            outs[i] = m;
        }
        return SBM_OK;
    });
}

int sbm_match_batch_arrays(sbm_detector* det, const sbm_image* const* images, int n, float threshold,
                           const char* const* class_ids, int n_ids, int slots, int32_t* xyct, float* similarity,
                           int cap, int* counts) {
    return guarded([&] {
        need(det); need(counts);
        if (n < 0 || cap < 0 || (cap > 0 && (!xyct || !similarity))) throw sbm::Error(SBM_EINVAL, "bad arguments");
        std::vector<sbm::ImageArgs> ims(n);
        for (int i = 0; i < n; ++i) {
            need(images);
            need(images[i]);
            const sbm_image* im = images[i];
            ims[i].img = im->img.as<uint8_t>();
            ims[i].rows = im->rows;
            ims[i].cols = im->cols;
            ims[i].stride = im->cols * im->channels;
            ims[i].channels = im->channels;
            ims[i].mask = im->has_mask ? im->mask.as<uint8_t>() : nullptr;
            ims[i].device = true;
        }
        const auto res = det->d.match_batch(ims, threshold, id_list(class_ids, n_ids), slots);
        int at = 0;
        for (int i = 0; i < n; ++i) {
            counts[i] = (int)res[i].size();
            for (const auto& r : res[i]) {
                if (at >= cap) break;
                xyct[4 * at] = r.x;
                xyct[4 * at + 1] = r.y;
                xyct[4 * at + 2] = r.class_index;
                xyct[4 * at + 3] = r.template_id;
                similarity[at] = r.similarity;
                ++at;
            }
        }
        return SBM_OK;
    });
}

int sbm_match_batch_host_arrays(sbm_detector* det, const uint8_t* const* imgs, const int* rows, const int* cols,
                                const int* strides, const int* channels, const uint8_t* const* masks, int n,
                                float threshold, const char* const* class_ids, int n_ids, int slots, int32_t* xyct,
                                float* similarity, int cap, int* counts) {
    return guarded([&] {
        need(det); need(counts);
        if (n < 0 || cap < 0 || (cap > 0 && (!xyct || !similarity))) throw sbm::Error(SBM_EINVAL, "bad arguments");
        if (n > 0) { need(imgs); need(rows); need(cols); need(strides); need(channels); }
        std::vector<sbm::ImageArgs> ims(n);
        for (int i = 0; i < n; ++i) {
            need(imgs[i]);
            ims[i].img = imgs[i];
            ims[i].rows = rows[i];
            ims[i].cols = cols[i];
            ims[i].stride = strides[i];
            ims[i].channels = channels[i];
            ims[i].mask = masks ? masks[i] : nullptr;
            ims[i].device = false;
        }
        const auto res = det->d.match_batch(ims, threshold, id_list(class_ids, n_ids), slots);
        int at = 0;
        for (int i = 0; i < n; ++i) {
            counts[i] = (int)res[i].size();
            for (const auto& r : res[i]) {
                if (at >= cap) break;
                xyct[4 * at] = r.x;
                xyct[4 * at + 1] = r.y;
                xyct[4 * at + 2] = r.class_index;
                xyct[4 * at + 3] = r.template_id;
                similarity[at] = r.similarity;
                ++at;
            }
        }
        return SBM_OK;
    });
}

int sbm_synchronize(sbm_detector* det) {
    return guarded([&] {
        need(det);
        sbm::hip_check(hipStreamSynchronize(det->d.stream()), "sync");
        return SBM_OK;
    });
}

int sbm_matches_count(const sbm_matches* m, int* count) {
    return guarded([&] { need(m); need(count); *count = (int)m->rows.size(); return SBM_OK; });
}
int sbm_matches_get(const sbm_matches* m, int i, int* x, int* y, float* sim, int* ci, int* tid) {
    return guarded([&] {
        need(m);
        if (i < 0 || i >= (int)m->rows.size()) throw sbm::Error(SBM_EINVAL, "match index out of range");
        const auto& r = m->rows[i];
        if (x) *x = r.x;
        if (y) *y = r.y;
        if (sim) *sim = r.similarity;
        if (ci) *ci = r.class_index;
        if (tid) *tid = r.template_id;
        return SBM_OK;
    });
}
int sbm_matches_copy(const sbm_matches* m, int32_t* xyct, float* sim, int cap, int* copied) {
    return guarded([&] {
        need(m);
        const int n = std::min(cap, (int)m->rows.size());
        for (int i = 0; i < n; ++i) {
            const auto& r = m->rows[i];
            if (xyct) { xyct[4 * i] = r.x; xyct[4 * i + 1] = r.y; xyct[4 * i + 2] = r.class_index; xyct[4 * i + 3] = r.template_id; }
            if (sim) sim[i] = r.similarity;
        }
        if (copied) *copied = n;
        return SBM_OK;
    });
}
int sbm_matches_class_id(const sbm_matches* m, int ci, char* buf, int cap) {
    return guarded([&] {
        need(m);
        if (ci < 0 || ci >= (int)m->names.size()) throw sbm::Error(SBM_EINVAL, "class index out of range");
        return copy_str(m->names[ci], buf, cap);
    });
}
int sbm_matches_free(sbm_matches* m) {
    return guarded([&] { delete m; return SBM_OK; });
}

int sbm_server_create(sbm_detector* det, int slots, long long budget, sbm_server** out) {
    return guarded([&] {
        need(det); need(out);
        if (slots < 1 || slots > 16) throw sbm::Error(SBM_EINVAL, "slots must be in [1, 16]");
        if (budget <= 0) throw sbm::Error(SBM_EINVAL, "budget must be positive");
        auto* h = new sbm_server();
        try {
            h->s.reset(new sbm::Server(det->d, slots, (size_t)budget));
            h->det = det;
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return SBM_OK;
    });
}
int sbm_server_destroy(sbm_server* srv) {
    return guarded([&] { delete srv; return SBM_OK; });
}
int sbm_server_submit(sbm_server* srv, const uint8_t* img, int rows, int cols, int stride, int channels,
                      const uint8_t* mask, float threshold, const char* const* class_ids, int n, long long* id) {
    return guarded([&] {
        need(srv); need(img); need(id);
        *id = srv->s->submit(img, rows, cols, stride, channels, mask, threshold, id_list(class_ids, n));
        return SBM_OK;
    });
}
int sbm_server_wait(sbm_server* srv, long long id, int timeout_ms, sbm_matches** out, double* latency_ms) {
    return guarded([&] {
        need(srv);
        std::vector<sbm::MatchRec> rows;
        std::string err;
        const int rc = srv->s->wait(id, timeout_ms, &rows, latency_ms, &err);
        if (rc == sbm::kETimeout) throw sbm::Error(SBM_ETIMEOUT, "request not finished");
        if (rc != SBM_OK) throw sbm::Error(rc, err);
        if (out) {
            auto* m = new sbm_matches();
            m->rows = std::move(rows);
            m->names = srv->det->d.class_names();
            *out = m;
        }
        return SBM_OK;
    });
}
int sbm_server_launches(sbm_server* srv, long long* launches) {
    return guarded([&] {
        need(srv); need(launches);
        *launches = srv->s->stats().launches;
        return SBM_OK;
    });
}

int sbm_server_stats(sbm_server* srv, long long* submitted, long long* admitted, long long* completed,
                     long long* failed, long long* deferrals, long long* peak_bytes) {
    return guarded([&] {
        need(srv);
        const sbm::ServerStats st = srv->s->stats();
        if (submitted) *submitted = st.submitted;
        if (admitted) *admitted = st.admitted;
        if (completed) *completed = st.completed;
        if (failed) *failed = st.failed;
        if (deferrals) *deferrals = st.deferrals;
        if (peak_bytes) *peak_bytes = st.peak_bytes;
        return SBM_OK;
    });
}

int sbm_read_classes(sbm_detector* det, const char* const* class_ids, int n, const char* format) {
    return guarded([&] {
        need(det);
        det->d.read_classes(id_list(class_ids, n), format ? format : "templates_%s.yml.gz");
        return SBM_OK;
    });
}
int sbm_write_classes(const sbm_detector* det, const char* format) {
    return guarded([&] {
        need(det);
        det->d.write_classes(format ? format : "templates_%s.yml.gz");
        return SBM_OK;
    });
}
int sbm_shape_transform(sbm_detector* det, const uint8_t* src, int rows, int cols, int stride, int channels, float angle,
                        float scale, int binary_mask, uint8_t* dst) {
    return guarded([&] {
        need(src); need(dst);
        if (stride < cols * channels) throw sbm::Error(SBM_EINVAL, "stride smaller than a row");
        if (det) {
            det->d.shape_transform(src, rows, cols, stride, channels, angle, scale, binary_mask != 0, dst);
            return SBM_OK;
        }
        // no detector: a library-wide one on the current device does the warp
        static std::mutex mu;
        static sbm::Detector* shared = nullptr;
        std::lock_guard<std::mutex> lk(mu);
        if (!shared) shared = new sbm::Detector();
        shared->shape_transform(src, rows, cols, stride, channels, angle, scale, binary_mask != 0, dst);
        return SBM_OK;
    });
}
int sbm_produce_infos(const float* angle_range, int n_angle_range, const float* scale_range, int n_scale_range,
                      float angle_step, float scale_step, float eps, float* angles, float* scales, int cap, int* count) {
    return guarded([&] {
        need(count);
        if (n_angle_range < 0 || n_scale_range < 0) throw sbm::Error(SBM_EINVAL, "negative range size");
        if (n_angle_range) need(angle_range);
        if (n_scale_range) need(scale_range);
        std::vector<float> a, sc;
        sbm::produce_infos(std::vector<float>(angle_range, angle_range + n_angle_range),
                           std::vector<float>(scale_range, scale_range + n_scale_range), angle_step, scale_step, eps, a, sc);
        *count = (int)a.size();
        const int m = std::min(cap, (int)a.size());
        if (m > 0) {
            need(angles); need(scales);
            std::copy(a.begin(), a.begin() + m, angles);
            std::copy(sc.begin(), sc.begin() + m, scales);
        }
        return SBM_OK;
    });
}
int sbm_save_infos(const char* path, const float* angles, const float* scales, int n) {
    return guarded([&] {
        need(path);
        if (n < 0) throw sbm::Error(SBM_EINVAL, "negative count");
        if (n) { need(angles); need(scales); }
        sbm::yaml::write_file(path, sbm::write_infos_text(std::vector<float>(angles, angles + n),
                                                          std::vector<float>(scales, scales + n)));
        return SBM_OK;
    });
}
int sbm_load_infos(const char* path, float* angles, float* scales, int cap, int* count) {
    return guarded([&] {
        need(path); need(count);
        std::vector<float> a, sc;
        sbm::read_infos_text(sbm::yaml::read_file(path), a, sc);
        *count = (int)a.size();
        const int m = std::min(cap, (int)a.size());
        if (m > 0) {
            need(angles); need(scales);
            std::copy(a.begin(), a.begin() + m, angles);
            std::copy(sc.begin(), sc.begin() + m, scales);
        }
        return SBM_OK;
    });
}

int sbm_read_class_file(sbm_detector* det, const char* path, const char* override_id, char* out, int cap) {
    return guarded([&] {
        need(det); need(path);
        const std::string id = det->d.read_class_text(sbm::yaml::read_file(path), override_id ? override_id : "");
        if (out) copy_str(id, out, cap);
        return SBM_OK;
    });
}
int sbm_write_class_file(const sbm_detector* det, const char* class_id, const char* path) {
    return guarded([&] {
        need(det); need(class_id); need(path);
        sbm::yaml::write_file(path, det->d.write_class_text(class_id));
        return SBM_OK;
    });
}
int sbm_upload_bank(sbm_detector* det) {
    return guarded([&] {
        need(det);
        det->d.ensure_device();
        sbm::DeviceScope ds(det->d.device());
        det->d.ensure_bank();
        return SBM_OK;
    });
}
int sbm_save_bank(const sbm_detector* det, const char* path) {
    return guarded([&] { need(det); need(path); sbm::yaml::write_file(path, det->d.write_bank_binary()); return SBM_OK; });
}
int sbm_load_bank(sbm_detector* det, const char* path, int* num_classes) {
    return guarded([&] {
        need(det); need(path);
        const int n = det->d.read_bank_binary(sbm::yaml::read_file(path));
        if (num_classes) *num_classes = n;
        return SBM_OK;
    });
}
int sbm_write_detector_file(const sbm_detector* det, const char* path) {
    return guarded([&] { need(det); need(path); sbm::yaml::write_file(path, det->d.write_detector_text()); return SBM_OK; });
}
int sbm_read_detector_file(sbm_detector* det, const char* path) {
    return guarded([&] { need(det); need(path); det->d.read_detector_text(sbm::yaml::read_file(path)); return SBM_OK; });
}

int sbm_nms_boxes(const int32_t* boxes, const float* scores, int n, float score_threshold, float nms_threshold,
                  float eta, int top_k, int32_t* indices, int* count) {
    return guarded([&] {
        if (n > 0) { need(boxes); need(scores); need(indices); }
        need(count);
        const auto idx = sbm::nms_boxes(boxes, scores, n, score_threshold, nms_threshold, eta, top_k);
        for (size_t i = 0; i < idx.size(); ++i) indices[i] = idx[i];
        *count = (int)idx.size();
        return SBM_OK;
    });
}

int sbm_set_tuning(sbm_detector* det, int ppl, int wpb, int bpc, int gather, int split) {
    return guarded([&] {
        need(det);
        if (gather < 0 || gather > 3) throw sbm::Error(SBM_EINVAL, "gather mode must be 0, 1, 2 or 3");
        if (gather == 3 ? ppl != 32 : (ppl != 4 && ppl != 8 && ppl != 16))
            throw sbm::Error(SBM_EINVAL, "positions per lane must be 4, 8 or 16 (32 with gather mode 3)");
        if (wpb < 1 || wpb > 4) throw sbm::Error(SBM_EINVAL, "waves per block must be in [1, 4]");
        if (bpc < 1 || bpc > 64) throw sbm::Error(SBM_EINVAL, "blocks per CU must be in [1, 64]");
        if (split != 1 && split != 2 && split != 4 && split != 8)
            throw sbm::Error(SBM_EINVAL, "split must be 1, 2, 4 or 8");
        det->d.tuning.ppl = ppl;
        det->d.tuning.waves_per_block = wpb;
        det->d.tuning.blocks_per_cu = bpc;
        det->d.tuning.gather = gather;
        det->d.tuning.split = split;
        return SBM_OK;
    });
}
int sbm_set_coarse_mode(sbm_detector* det, int mode) {
    return guarded([&] {
        need(det);
        if (mode < 0 || mode > 2) throw sbm::Error(SBM_EINVAL, "coarse mode must be 0, 1 or 2");
        det->d.tuning.coarse_mode = mode;
        return SBM_OK;
    });
}
int sbm_get_coarse_mode(const sbm_detector* det, int* mode) {
    return guarded([&] { need(det); need(mode); *mode = det->d.tuning.coarse_mode; return SBM_OK; });
}
int sbm_last_coarse_path(const sbm_detector* det, int* sparse, int* nonzero) {
    return guarded([&] {
        need(det);
        const auto& t = det->d.last_times();
        if (sparse) *sparse = t.coarse_sparse;
        if (nonzero) *nonzero = t.coarse_nonzero;
        return SBM_OK;
    });
}
int sbm_last_refine_features(const sbm_detector* det, long long* features) {
    return guarded([&] {
        need(det); need(features);
        *features = det->d.last_times().refine_features;
        return SBM_OK;
    });
}

int sbm_last_coarse_entries(const sbm_detector* det, long long* entries) {
    return guarded([&] { need(det); need(entries); *entries = det->d.last_times().coarse_entries; return SBM_OK; });
}
int sbm_get_tuning(const sbm_detector* det, int* ppl, int* wpb, int* bpc, int* gather, int* split) {
    return guarded([&] {
        need(det);
        if (ppl) *ppl = det->d.tuning.ppl;
        if (wpb) *wpb = det->d.tuning.waves_per_block;
        if (bpc) *bpc = det->d.tuning.blocks_per_cu;
        if (gather) *gather = det->d.tuning.gather;
        if (split) *split = det->d.tuning.split;
        return SBM_OK;
    });
}
int sbm_autotune(sbm_detector* det, const sbm_image* im, float threshold, const char* const* class_ids, int n,
                 int iters, int* ppl, int* wpb, int* bpc, int* gather, int* split, int32_t* trace, int trace_cap,
                 int* trace_count) {
    return guarded([&] {
        need(det); need(im);
        std::vector<sbm::Detector::TuneRow> tr;
        const sbm::Tuning best = det->d.autotune(im->img.as<uint8_t>(), im->rows, im->cols, im->cols * im->channels,
                                                 im->channels, threshold, id_list(class_ids, n), true, iters, &tr);
        if (ppl) *ppl = best.ppl;
        if (wpb) *wpb = best.waves_per_block;
        if (bpc) *bpc = best.blocks_per_cu;
        if (gather) *gather = best.gather;
        if (split) *split = best.split;
        int k = 0;
        for (const auto& r : tr) {  // the gather kernel's rows
            if (r.stage != 0) continue;
            if (!trace || k >= trace_cap) break;
            trace[6 * k] = r.tuning.ppl;
            trace[6 * k + 1] = r.tuning.waves_per_block;
            trace[6 * k + 2] = r.tuning.blocks_per_cu;
            trace[6 * k + 3] = r.tuning.gather;
            trace[6 * k + 4] = r.tuning.split;
            trace[6 * k + 5] = (int32_t)(r.ms * 1000.f);
            ++k;
        }
        if (trace_count) *trace_count = k;
        return SBM_OK;
    });
}

int sbm_autotune_images(sbm_detector* det, const sbm_image* const* images, int n, float threshold,
                        const char* const* class_ids, int n_ids, int iters) {
    return guarded([&] {
        need(det); need(images);
        if (n < 1) throw sbm::Error(SBM_EINVAL, "autotune needs at least one image");
        std::vector<sbm::ImageArgs> ims(n);
        for (int i = 0; i < n; ++i) {
            need(images[i]);
            const sbm_image* im = images[i];
            ims[i].img = im->img.as<uint8_t>();
            ims[i].rows = im->rows;
            ims[i].cols = im->cols;
            ims[i].stride = im->cols * im->channels;
            ims[i].channels = im->channels;
            ims[i].mask = im->has_mask ? im->mask.as<uint8_t>() : nullptr;
            ims[i].device = true;
        }
        const sbm::ImageArgs& a = ims[0];
        (void)det->d.autotune(a.img, a.rows, a.cols, a.stride, a.channels, threshold, id_list(class_ids, n_ids), true,
                              iters, nullptr, &ims);
        return SBM_OK;
    });
}
int sbm_autotune_row(const sbm_detector* det, int index, int* stage, int32_t* params, float* us, int* count) {
    return guarded([&] {
        need(det);
        const auto& rows = det->d.last_autotune;
        if (count) *count = (int)rows.size();
        if (index < 0) return SBM_OK;
        if (index >= (int)rows.size()) throw sbm::Error(SBM_EINVAL, "autotune row out of range");
        const auto& r = rows[index];
        if (stage) *stage = r.stage;
        if (params) {
            const sbm::Tuning& t = r.tuning;
            const int v[3][5] = {{t.ppl, t.waves_per_block, t.blocks_per_cu, t.gather, t.split},
                                 {t.list_threads, t.list_segment, t.list_bpc, t.list_order, t.list_split},
                                 {t.refine_waves, t.refine_group, t.refine_bpc, t.refine_layout, 0}};
            for (int i = 0; i < 5; ++i) params[i] = v[r.stage][i];
        }
        if (us) *us = r.ms * 1000.f;
        return SBM_OK;
    });
}

namespace {
int* tuning_field(sbm::Tuning& t, const std::string& name) {
    if (name == "positions_per_lane") return &t.ppl;
    if (name == "waves_per_block") return &t.waves_per_block;
    if (name == "blocks_per_cu") return &t.blocks_per_cu;
    if (name == "gather") return &t.gather;
    if (name == "split") return &t.split;
    if (name == "list_threads") return &t.list_threads;
    if (name == "list_segment") return &t.list_segment;
    if (name == "list_blocks_per_cu") return &t.list_bpc;
    if (name == "list_order") return &t.list_order;
    if (name == "list_split") return &t.list_split;
    if (name == "refine_waves") return &t.refine_waves;
    if (name == "refine_group") return &t.refine_group;
    if (name == "refine_blocks_per_cu") return &t.refine_bpc;
    if (name == "refine_layout") return &t.refine_layout;
    throw sbm::Error(SBM_EINVAL, "unknown tuning parameter '" + name + "'");
}
bool tuning_value_ok(const std::string& name, int v) {
    if (name == "list_threads") return v == 256 || v == 512 || v == 1024;
    if (name == "list_segment") return v == 4096 || v == 8192 || v == 16384;
    if (name == "list_blocks_per_cu") return v == 1 || v == 2 || v == 4 || v == 8;
    if (name == "list_order") return v == 0 || v == 1;
    if (name == "list_split") return v >= 10 && v <= 100;
    if (name == "refine_waves") return v == 1 || v == 2 || v == 4 || v == 8 || v == 16;
    if (name == "refine_group") return v == 8 || v == 16;
    if (name == "refine_blocks_per_cu") return v >= 1 && v <= 32;
    if (name == "refine_layout") return v >= 0 && v <= 4;
    return true;  // the gather kernel's parameters are checked at launch, as with sbm_set_tuning
}
}  // namespace

int sbm_set_tuning_param(sbm_detector* det, const char* name, int value) {
    return guarded([&] {
        need(det); need(name);
        sbm::Tuning t = det->d.tuning;
        int* f = tuning_field(t, name);
        if (!tuning_value_ok(name, value)) throw sbm::Error(SBM_EINVAL, std::string("bad value for ") + name);
        *f = value;
        if (t.list_bpc * t.list_threads > 2048 || t.list_bpc * sbm::kern::sparse_block_lds(t.list_segment) > 160 * 1024)
            throw sbm::Error(SBM_EINVAL, "the scatter's blocks per CU exceed a CU's threads or LDS");
        det->d.tuning = t;
        return SBM_OK;
    });
}

int sbm_get_tuning_param(const sbm_detector* det, const char* name, int* value) {
    return guarded([&] {
        need(det); need(name); need(value);
        sbm::Tuning t = det->d.tuning;
        *value = *tuning_field(t, name);
        return SBM_OK;
    });
}

int sbm_set_timing(sbm_detector* det, int enabled) {
    return guarded([&] {
        need(det);
        det->d.timing_enabled = enabled == 1;
        det->d.graph_timing = enabled == 2;
        return SBM_OK;
    });
}
int sbm_last_device_ms(const sbm_detector* det, float* ms) {
    return guarded([&] { need(det); need(ms); *ms = det->d.last_times().total_ms; return SBM_OK; });
}
int sbm_last_timing(const sbm_detector* det, float* pre, float* coarse, float* refine, int* cc, int* fc) {
    return guarded([&] {
        need(det);
        const auto& t = det->d.last_times();
        if (pre) *pre = t.preprocess_ms;
        if (coarse) *coarse = t.coarse_ms;
        if (refine) *refine = t.refine_ms;
        if (cc) *cc = t.coarse_candidates;
        if (fc) *fc = t.final_candidates;
        return SBM_OK;
    });
}
int sbm_last_kernel_time(const sbm_detector* det, int index, char* name, int cap, float* ms, int* count) {
    return guarded([&] {
        need(det);
        const int n = (int)det->d.kernel_ms().size();
        if (count) *count = n;
        if (index >= 0 && index < n) {
            if (name) copy_str(det->d.kernel_names()[index], name, cap);
            if (ms) *ms = det->d.kernel_ms()[index];
        }
        return SBM_OK;
    });
}
int sbm_coarse_gather_bytes(sbm_detector* det, int rows, int cols, const char* const* class_ids, int n, long long* bytes) {
    return guarded([&] {
        need(det); need(bytes);
        const auto& d = det->d;
        const int L = d.pyramid_levels;
        int r = rows, c = cols;
        for (int l = 1; l < L; ++l) { r /= 2; c /= 2; }
        const int T = d.T_at_level[L - 1];
        const int W = c / T, H = r / T;
        const auto ids = id_list(class_ids, n);
        long long total = 0;
        for (const auto& kv : d.class_templates) {
            if (!ids.empty() && std::find(ids.begin(), ids.end(), kv.first) == ids.end()) continue;
            for (const auto& tp : kv.second) {
                const sbm::Template& t = tp[L - 1];
                const int wf = (t.width - 1) / T + 1, hf = (t.height - 1) / T + 1;
                const long long P = (long long)(H - hf) * W + (W - wf) + 1;
                if (P <= 0) continue;
                long long inb = 0;
                for (const auto& f : t.features) inb += (f.x >= 0 && f.x < c && f.y >= 0 && f.y < r) ? 1 : 0;
                total += inb * P;
            }
        }
        *bytes = total;
        return SBM_OK;
    });
}
int sbm_copy_probe(sbm_detector* det, long long bytes) {
    return guarded([&] {
        need(det);
        if (bytes <= 0 || (bytes & 15)) throw sbm::Error(SBM_EINVAL, "bytes must be a positive multiple of 16");
        det->d.ensure_device();
        sbm::DevMem a, b;
        a.ensure((size_t)bytes);
        b.ensure((size_t)bytes);
        sbm::hip_check(hipMemsetAsync(a.p, 1, (size_t)bytes, det->d.stream()), "memset");
        sbm::hip_check(sbm::kern::copy_probe(a.as<uint4>(), b.as<uint4>(), bytes / 16, det->d.stream()), "copy");
        sbm::hip_check(hipStreamSynchronize(det->d.stream()), "sync");
        return SBM_OK;
    });
}
int sbm_device_bytes(const sbm_detector* det, long long* bytes) {
    return guarded([&] { need(det); need(bytes); *bytes = (long long)det->d.device_bytes(); return SBM_OK; });
}

int sbm_prepare_image(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels,
                      const uint8_t* mask) {
    return guarded([&] { need(det); need(img); det->d.prepare_image(img, rows, cols, stride, channels, mask); return SBM_OK; });
}
int sbm_read_level(sbm_detector* det, int level, int what, uint8_t* out) {
    return guarded([&] { need(det); need(out); det->d.read_level(level, what, out); return SBM_OK; });
}
int sbm_similarity_map(sbm_detector* det, const char* class_id, int template_id, uint16_t* out) {
    return guarded([&] { need(det); need(class_id); need(out); det->d.similarity_map(class_id, template_id, out); return SBM_OK; });
}
int sbm_similarity_local(sbm_detector* det, const char* class_id, int template_id, int level, int cx, int cy,
                         uint16_t* out) {
    return guarded([&] {
        need(det); need(class_id); need(out);
        det->d.similarity_local(class_id, template_id, level, cx, cy, out);
        return SBM_OK;
    });
}
int sbm_gradient(sbm_detector* det, const uint8_t* img, int rows, int cols, int stride, int channels, uint8_t* q,
                 float* mag, float* ang) {
    return guarded([&] { need(det); need(img); need(q); det->d.gradient(img, rows, cols, stride, channels, q, mag, ang); return SBM_OK; });
}
int sbm_bin_table(sbm_detector* det, int lo, int hi, uint8_t* out) {
    return guarded([&] {
        need(det); need(out);
        if (hi < lo || lo < -1020 || hi > 1020) throw sbm::Error(SBM_EINVAL, "range");
        det->d.bin_table(lo, hi, out);
        return SBM_OK;
    });
}
int sbm_division_check(sbm_detector* det, int* mismatches) {
    return guarded([&] {
        need(det); need(mismatches);
        *mismatches = det->d.division_check();
        return SBM_OK;
    });
}

}  // extern "C"
