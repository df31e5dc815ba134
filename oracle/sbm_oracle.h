/*
 * CPU oracle for line2Dup's matching path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference method (meiqua/shape_based_matching,
 * line2Dup.cpp / line2Dup.h / test.cpp).  Only tests, the smoke check and the CPU baseline
 * leg of the benchmark may load it, and only as the checker: the product library never
 * links, loads or calls it.
 *
 * Every function cites the reference lines it restates.  The OpenCV calls the reference makes
 * (GaussianBlur, Sobel, phase, convertTo, pyrDown, resize, erode) are restated from their
 * published algorithms for the parameters the reference passes.
 *
 * Parity status per function (how each one is pinned is in tests/test_oracle_*.py):
 *   orc_gaussian_blur7, orc_sobel3, orc_pyr_down, orc_mask_down   pinned (brute force / closed form)
 *   orc_fast_atan2_deg                                          pinned against exact atan2 (error bound,
 *                                                                boundary-only disagreement); the exact bits of
 *                                                                OpenCV's own build are "parity unpinned"
 *   orc_quantized_orientations, orc_spread, orc_response_maps,
 *   orc_linearize, orc_similarity, orc_similarity_local, orc_match pinned (brute force, invariants,
 *                                                                hand-computed micro cases)
 *   orc_nms_boxes                                               pinned (hand-computed cases)
 *   orc_extract_template, orc_crop_templates, orc_rotate_template,
 *   orc_add_template                                            pinned against the reference's trained fixture
 *                                                                banks (extraction from the training images,
 *                                                                rotation of the extracted theta)
 *   orc_warp_affine, orc_shape_transform, orc_produce_infos     pinned against closed forms (identity, exact
 *                                                                quarter turns, dyadic scales) and the fixture
 *                                                                banks trained on warped images
 */
#ifndef SBM_ORACLE_H
#define SBM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- per-image preprocessing (line2Dup.cpp:286-423, 556-750) ---- */
void orc_gaussian_blur7(const uint8_t* src, int rows, int cols, uint8_t* dst);
void orc_sobel3(const uint8_t* blurred, int rows, int cols, int32_t* dx, int32_t* dy);
float orc_fast_atan2_deg(float y, float x);
int orc_angle_bin(float angle_deg);
void orc_bin_table(int lo, int hi, uint8_t* out);
void orc_quantized_orientations(const uint8_t* src, int rows, int cols, float weak_threshold,
                                uint8_t* quantized, float* magnitude, float* angle_ori);
void orc_gradient_color(const uint8_t* bgr, int rows, int cols, float weak_threshold,
                        uint8_t* quantized, float* magnitude, float* angle_ori);
void orc_pyr_down(const uint8_t* src, int rows, int cols, int channels, uint8_t* dst);
void orc_mask_down(const uint8_t* src, int rows, int cols, uint8_t* dst);
void orc_apply_mask(uint8_t* quantized, const uint8_t* mask, int n);
void orc_spread(const uint8_t* q, int rows, int cols, int T, uint8_t* dst);
void orc_response_maps(const uint8_t* spread, int rows, int cols, uint8_t* maps);
uint8_t orc_similarity_lut(int ori, int idx);
void orc_linearize(const uint8_t* response, int rows, int cols, int T, uint8_t* lm);

/* Full per-level linear-memory pyramid: lm_out[l] must hold 8*rows_l*cols_l + guard bytes. */
int orc_lm_guard_bytes(int rows, int cols, int T);
int orc_build_lm_pyramid(const uint8_t* img, int rows, int cols, int channels, const uint8_t* mask,
                         float weak_threshold, int num_levels, const int* T, uint8_t** lm_out);

/* ---- similarity (line2Dup.cpp:755-1021) ---- */
/* features: int32 triples (x, y, label) */
int orc_similarity(const uint8_t* lm, int rows, int cols, int T, const int32_t* feats, int n,
                   int width, int height, uint16_t* dst);
int orc_similarity_local(const uint8_t* lm, int rows, int cols, int T, const int32_t* feats, int n,
                         int cx, int cy, uint16_t* dst);

/* ---- driver (line2Dup.cpp:1049-1268, line2Dup.h:108-141) ----
 * tinfo: per template, per level, 6 int32: width, height, tl_x, tl_y, num_features, feature_offset
 * feats: int32 triples indexed by feature_offset
 * tclass: class rank of each template (rank in sorted class-name order)
 * tid: template id of each template within its class
 * out_i32: 4 int32 per match (x, y, class_rank, template_id); out_sim: similarity
 * Returns the number of matches after canonical sort + unique, or -(needed) if max_out is too small,
 * or -1000000000 on an invalid argument (sizes not divisible, too many features). */
int orc_match(const uint8_t* img, int rows, int cols, int channels, const uint8_t* mask,
              float weak_threshold, int num_levels, const int* T, float threshold,
              int num_templates, const int32_t* tinfo, const int32_t* feats,
              const int32_t* tclass, const int32_t* tid,
              int max_out, int32_t* out_i32, float* out_sim, int nthreads, double* stage_seconds);
int orc_match_lm(uint8_t* const* lm, int rows, int cols, int num_levels, const int* T, float threshold,
                 int num_templates, const int32_t* tinfo, const int32_t* feats,
                 const int32_t* tclass, const int32_t* tid,
                 int max_out, int32_t* out_i32, float* out_sim, int nthreads);
float orc_score(int raw, int n);
int orc_sort_unique(int count, int32_t* i32, float* sim);

/* ---- NMS (test.cpp:31-121) ---- */
int orc_nms_boxes(int n, const int32_t* boxes /* x,y,w,h */, const float* scores, float score_threshold,
                  float nms_threshold, float eta, int top_k, int32_t* indices);

/* ---- training (line2Dup.cpp:103-200, 425-512, 1270-1362) ---- */
int orc_extract_template(const float* magnitude, const uint8_t* angle, const float* angle_ori,
                         const uint8_t* mask, int rows, int cols, float strong_threshold,
                         int num_features, int max_out, int32_t* out_xyl, float* out_theta);
/* The candidate stage alone (eroded mask, 5x5 NMS, strong threshold, stable sort by score):
 * returns the candidate count; writes at most max_out. */
int orc_extract_candidates(const float* magnitude, const uint8_t* angle, const float* angle_ori,
                           const uint8_t* mask, int rows, int cols, float strong_threshold, int max_out,
                           int32_t* out_xyl, float* out_theta, float* out_score);
void orc_crop_templates(int levels, int32_t* tinfo5 /* per level: width,height,tl_x,tl_y,level */,
                        const int* nfeat, int32_t** feats_xyl);
int orc_add_template(const uint8_t* img, int rows, int cols, int channels, const uint8_t* mask,
                     float weak_threshold, float strong_threshold, int num_features,
                     int num_levels, int max_feats,
                     int32_t* tinfo6, int32_t* feats_xyl, float* feats_theta);
void orc_rotate_template(int num_levels, const int32_t* tinfo6_src, const int32_t* feats_src,
                         const float* theta_src, float theta, float cx, float cy,
                         int32_t* tinfo6_dst, int32_t* feats_dst, float* theta_dst);

/* ---- shapeInfo_producer (line2Dup.h:209-332) ----
 * transform(): cv::getRotationMatrix2D(center=(cols/2, rows/2), angle, scale) then cv::warpAffine with
 * INTER_LINEAR and a constant zero border, same output size (OpenCV 3.4 imgproc, restated from its
 * published fixed-point algorithm: 10-bit affine steps, 5-bit sub-pixel table, 15-bit weights). */
void orc_rotation_matrix2d(float cx, float cy, double angle_deg, double scale, double M[6]);
void orc_invert_affine(const double M[6], double iM[6]);
void orc_warp_affine(const uint8_t* src, int rows, int cols, int channels, const double M[6], uint8_t* dst);
void orc_shape_transform(const uint8_t* src, int rows, int cols, int channels, float angle, float scale,
                         uint8_t* dst);
/* produce_infos(): angle/scale grid with float accumulation; returns the count, writes at most cap. */
int orc_produce_infos(int n_angle, const float* angle_range, int n_scale, const float* scale_range,
                      float angle_step, float scale_step, float eps, int cap, float* angles, float* scales);

#ifdef __cplusplus
}
#endif
#endif
