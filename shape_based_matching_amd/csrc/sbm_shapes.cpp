// shapeInfo_producer (line2Dup.h:209-332) and batched training over its (angle, scale) infos.
//
// transform() = cv::getRotationMatrix2D((cols/2, rows/2), angle, scale) + cv::warpAffine
// (INTER_LINEAR, zero border, same size).  The matrix and its inversion are computed here in double,
// exactly as OpenCV does; the per-pixel fixed-point warp runs on the GPU (sbm_warp.hip).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>

#include "sbm_detector.h"
#include "sbm_internal.h"

#pragma clang fp contract(off)

namespace sbm {

// cv::getRotationMatrix2D: counter-clockwise by angle_deg about (cx, cy), scaled.
void rotation_matrix2d(float cx, float cy, double angle_deg, double scale, double M[6]) {
    const double a = angle_deg * (3.1415926535897932384626433832795 / 180);
    const double alpha = std::cos(a) * scale, beta = std::sin(a) * scale;
    M[0] = alpha;
    M[1] = beta;
    M[2] = (1 - alpha) * (double)cx - beta * (double)cy;
    M[3] = -beta;
    M[4] = alpha;
    M[5] = beta * (double)cx + (1 - alpha) * (double)cy;
}

// The inverse map cv::warpAffine builds from a forward matrix.
void invert_affine(const double M[6], double inv[6]) {
    double D = M[0] * M[4] - M[1] * M[3];
    D = D != 0 ? 1. / D : 0;
    inv[0] = M[4] * D;
    inv[1] = M[1] * -D;
    inv[3] = M[3] * -D;
    inv[4] = M[0] * D;
    inv[2] = -inv[0] * M[2] - inv[1] * M[5];
    inv[5] = -inv[3] * M[2] - inv[4] * M[5];
}

// shapeInfo_producer::produce_infos (line2Dup.h:281-320), float accumulation as written there.
void produce_infos(const std::vector<float>& ar_in, const std::vector<float>& sr_in, float astep, float sstep, float eps,
                   std::vector<float>& angles, std::vector<float>& scales) {
    angles.clear();
    scales.clear();
    if (ar_in.size() > 2 || sr_in.size() > 2) throw Error(kEInval, "angle_range and scale_range hold at most 2 values");
    if (!(astep > eps * 10) || !(sstep > eps * 10)) throw Error(kEInval, "angle_step and scale_step must exceed 10*eps");
    std::vector<float> ar = ar_in, sr = sr_in;
    if (ar.empty()) ar.push_back(0);
    if (sr.empty()) sr.push_back(1);
    auto push = [&](float a, float s) {
        angles.push_back(a);
        scales.push_back(s);
    };
    if (ar.size() == 1 && sr.size() == 1) {
        push(ar[0], sr[0]);
    } else if (ar.size() == 1) {
        if (!(sr[1] > sr[0])) throw Error(kEInval, "scale_range must be increasing");
        for (float s = sr[0]; s <= sr[1] + eps; s += sstep) push(ar[0], s);
    } else if (sr.size() == 1) {
        if (!(ar[1] > ar[0])) throw Error(kEInval, "angle_range must be increasing");
        for (float a = ar[0]; a <= ar[1] + eps; a += astep) push(a, sr[0]);
    } else {
        if (!(sr[1] > sr[0]) || !(ar[1] > ar[0])) throw Error(kEInval, "ranges must be increasing");
        for (float s = sr[0]; s <= sr[1] + eps; s += sstep)
            for (float a = ar[0]; a <= ar[1] + eps; a += astep) push(a, s);
    }
}

// save_infos (line2Dup.h:255-268): FileStorage's layout for a sequence of {angle, scale} maps.
std::string write_infos_text(const std::vector<float>& angles, const std::vector<float>& scales) {
    std::string o = "%YAML:1.0\n---\ninfos:\n";
    for (size_t i = 0; i < angles.size(); ++i) {
        o += "   -\n";
        o += "      angle: " + yaml::real_to_string((double)angles[i]) + "\n";
        o += "      scale: " + yaml::real_to_string((double)scales[i]) + "\n";
    }
    if (angles.empty()) o = "%YAML:1.0\n---\ninfos:\n   []\n";
    return o;
}

// load_infos (line2Dup.h:269-279)
void read_infos_text(const std::string& text, std::vector<float>& angles, std::vector<float>& scales) {
    angles.clear();
    scales.clear();
    const yaml::Node root = yaml::parse(text);
    const yaml::Node* seq = root.get("infos");
    if (!seq) return;  // FileNode of a missing key iterates over nothing
    for (const yaml::Node& n : seq->seq) {
        const yaml::Node* a = n.get("angle");
        const yaml::Node* s = n.get("scale");
        angles.push_back(a ? (float)a->as_real() : 0.f);
        scales.push_back(s ? (float)s->as_real() : 0.f);
    }
}

void Detector::shape_transform(const uint8_t* img, int rows, int cols, int stride, int channels, float angle, float scale,
                               bool binary_mask, uint8_t* out) {
    if (channels != 1 && channels != 3) throw Error(kEInval, "channels must be 1 or 3");
    if (rows <= 0 || cols <= 0) throw Error(kESize, "empty image");
    ensure_device();
    DeviceScope ds(device_);
    hipStream_t s = pipe_->stream();
    double M[6], inv[6];
    rotation_matrix2d(cols / 2.0f, rows / 2.0f, (double)angle, (double)scale, M);
    invert_affine(M, inv);
    const size_t bytes = (size_t)rows * cols * channels;
    DevMem dsrc, ddst, dinv;
    dsrc.ensure(bytes);
    ddst.ensure(bytes);
    dinv.ensure(sizeof(inv));
    hip_check(hipMemcpy2DAsync(dsrc.p, (size_t)cols * channels, img, (size_t)stride, (size_t)cols * channels, rows,
                               hipMemcpyHostToDevice, s), "H2D");
    hip_check(hipMemcpyAsync(dinv.p, inv, sizeof(inv), hipMemcpyHostToDevice, s), "H2D");
    hip_check(kern::warp_affine(dsrc.as<uint8_t>(), rows, cols, cols * channels, channels, dinv.as<double>(), 1,
                                ddst.as<uint8_t>(), binary_mask, s), "warp_affine");
    hip_check(hipMemcpyAsync(out, ddst.p, bytes, hipMemcpyDeviceToHost, s), "D2H");
    hip_check(hipStreamSynchronize(s), "sync");
}

// Per template: level images (warped, then pyrDown), masks, and the gradient outputs the host
// selection needs, all for one chunk of infos at a time.
std::vector<int> Detector::add_templates_transformed(const uint8_t* img, int rows, int cols, int stride, int channels,
                                                     const uint8_t* mask, const std::string& class_id,
                                                     const std::vector<float>& angles, const std::vector<float>& scales,
                                                     const std::vector<int>& nfeat) {
    check_channels(channels);
    if (rows <= 0 || cols <= 0) throw Error(kESize, "empty image");
    if (angles.size() != scales.size() || (!nfeat.empty() && nfeat.size() != angles.size()))
        throw Error(kEInval, "angles, scales and num_features must have the same length");
    const int n = (int)angles.size();
    std::vector<TemplatePyramid>& tps = class_templates[class_id];  // created even when every info fails
    std::vector<int> ids(n, -1);
    if (n == 0) return ids;
    ensure_device();
    DeviceScope ds(device_);
    hipStream_t s = pipe_->stream();
    const int L = pyramid_levels;

    // level geometry: rows/2, cols/2 per level (ColorGradientPyramid::pyrDown, line2Dup.cpp:397-417)
    std::vector<int> lr(L), lc(L);
    std::vector<size_t> px(L), off(L);
    size_t per_tpl = 0;
    for (int l = 0; l < L; ++l) {
        lr[l] = l ? lr[l - 1] / 2 : rows;
        lc[l] = l ? lc[l - 1] / 2 : cols;
        if (lr[l] < 1 || lc[l] < 1) throw Error(kESize, "image too small for the pyramid (cv::pyrDown of an empty size)");
        px[l] = (size_t)lr[l] * lc[l];
        off[l] = per_tpl;
        per_tpl += px[l];
    }
    // per template on the device: image (channels), mask, q (2 bytes with 16 orientations), mag, angle
    const size_t qb = (size_t)(orientation_bins / 8);
    const size_t dev_per = per_tpl * (channels + 1 + qb + 8);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)n, ((size_t)512 << 20) / std::max<size_t>(dev_per, 1)));

    DevMem dsrc, dmask, dinv, dimg, dmsk, dq, dmag, dang;
    dsrc.ensure((size_t)rows * cols * channels);
    dmask.ensure((size_t)rows * cols);
    hip_check(hipMemcpy2DAsync(dsrc.p, (size_t)cols * channels, img, (size_t)stride, (size_t)cols * channels, rows,
                               hipMemcpyHostToDevice, s), "H2D");
    // shapeInfo_producer(src, mask): an empty mask becomes all 255 (line2Dup.h:236-243)
    if (mask) {
        hip_check(hipMemcpyAsync(dmask.p, mask, (size_t)rows * cols, hipMemcpyHostToDevice, s), "H2D");
    } else {
        hip_check(hipMemsetAsync(dmask.p, 255, (size_t)rows * cols, s), "memset");
    }
    dinv.ensure((size_t)chunk * 6 * sizeof(double));
    dimg.ensure((size_t)chunk * per_tpl * channels);
    dmsk.ensure((size_t)chunk * per_tpl);
    dq.ensure((size_t)chunk * per_tpl * qb);
    dmag.ensure((size_t)chunk * per_tpl * 4);
    dang.ensure((size_t)chunk * per_tpl * 4);
    std::vector<uint8_t> hq((size_t)chunk * per_tpl * qb), hm((size_t)chunk * per_tpl);
    std::vector<float> hmag((size_t)chunk * per_tpl), hang((size_t)chunk * per_tpl);
    std::vector<TemplatePyramid> built(n);
    std::vector<char> ok(n, 0);

    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int cn = std::min(chunk, n - c0);
        std::vector<double> inv((size_t)cn * 6);
        for (int i = 0; i < cn; ++i) {
            double M[6];
            rotation_matrix2d(cols / 2.0f, rows / 2.0f, (double)angles[c0 + i], (double)scales[c0 + i], M);
            invert_affine(M, &inv[(size_t)i * 6]);
        }
        hip_check(hipMemcpyAsync(dinv.p, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice, s), "H2D");
        // level-0 images and masks, written at stride per_tpl per template
        for (int i = 0; i < cn; ++i) {
            hip_check(kern::warp_affine(dsrc.as<uint8_t>(), rows, cols, cols * channels, channels, dinv.as<double>() + 6 * i,
                                        1, dimg.as<uint8_t>() + (size_t)i * per_tpl * channels, false, s),
                      "warp_affine");
            hip_check(kern::warp_affine(dmask.as<uint8_t>(), rows, cols, cols, 1, dinv.as<double>() + 6 * i, 1,
                                        dmsk.as<uint8_t>() + (size_t)i * per_tpl, true, s),
                      "warp_affine mask");
        }
        for (int i = 0; i < cn; ++i) {
            uint8_t* im = dimg.as<uint8_t>() + (size_t)i * per_tpl * channels;
            uint8_t* mk = dmsk.as<uint8_t>() + (size_t)i * per_tpl;
            for (int l = 0; l < L; ++l) {
                if (l > 0)
                    hip_check(kern::pyr_down(im + off[l - 1] * channels, lr[l - 1], lc[l - 1], lc[l - 1] * channels,
                                             channels, im + off[l] * channels, mk + off[l - 1], mk + off[l], s),
                              "pyr_down");
                const size_t o = (size_t)i * per_tpl + off[l];
                hip_check(kern::gradient_quantize(im + off[l] * channels, lr[l], lc[l], lc[l] * channels, channels, nullptr,
                                                  weak_threshold, dq.as<uint8_t>() + o * qb, dmag.as<float>() + o,
                                                  dang.as<float>() + o, s, nullptr, 0, orientation_bins),
                          "gradient");
            }
        }
        const size_t cb = (size_t)cn * per_tpl;
        hip_check(hipMemcpyAsync(hq.data(), dq.p, cb * qb, hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipMemcpyAsync(hm.data(), dmsk.p, cb, hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipMemcpyAsync(hmag.data(), dmag.p, cb * 4, hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipMemcpyAsync(hang.data(), dang.p, cb * 4, hipMemcpyDeviceToHost, s), "D2H");
        hip_check(hipStreamSynchronize(s), "sync");

        // feature selection, one template per task on host threads
        const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::mutex err_mu;
        std::exception_ptr err;
        auto work = [&](int t0) {
            for (int i = t0; i < cn; i += nthreads) {
                try {
                    const int g = c0 + i;
                    int nf = (!nfeat.empty() && nfeat[g] > 0) ? nfeat[g] : num_features;
                    TemplatePyramid tp(L);
                    bool good = true;
                    for (int l = 0; l < L && good; ++l) {
                        if (l > 0) nf /= 2;  // ColorGradientPyramid::pyrDown, line2Dup.cpp:400
                        const size_t o = (size_t)i * per_tpl + off[l];
                        if (orientation_bins == 16)
                            good = extract_template16(&hmag[o], reinterpret_cast<const uint16_t*>(&hq[o * qb]), &hm[o],
                                                      lr[l], lc[l], strong_threshold, nf, tp[l].features);
                        else
                            good = extract_template(&hmag[o], &hq[o], &hang[o], &hm[o], lr[l], lc[l], strong_threshold,
                                                    nf, tp[l].features);
                        tp[l].width = -1;
                        tp[l].height = -1;
                        tp[l].pyramid_level = l;
                    }
                    if (good) {
                        crop_templates(tp);
                        built[g] = std::move(tp);
                        ok[g] = 1;
                    }
                } catch (...) {
                    std::lock_guard<std::mutex> lk(err_mu);
                    if (!err) err = std::current_exception();
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < std::min(nthreads, cn); ++t) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
        if (err) std::rethrow_exception(err);
    }
    for (int g = 0; g < n; ++g) {
        if (!ok[g]) continue;
        ids[g] = (int)tps.size();
        tps.push_back(std::move(built[g]));
    }
    bank_dirty_ = true;
    return ids;
}

}  // namespace sbm
