/*
 * CPU oracle for the 16-orientation variant of line2Dup (test/ori_16bit_experiment) -- TEST
 * INFRASTRUCTURE ONLY, like sbm_oracle.h: tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it, only as the checker.
 *
 * Parity status per function (how each one is pinned is in tests/test_oracle16_pins.py):
 *   orc16_similarity_lut                      pinned against the table compiled into line2Dup_16bit_ori.cpp
 *   orc16_generated_lut                       pinned against LUT16.txt (LUT_gen.cpp's table, not compiled
 *                                             into the variant; kept to show the two differ)
 *   orc16_angle_bin, orc16_gradient_color     pinned by closed forms: ramps at the 16 bin centres and their
 *                                             opposites, and a brute-force vote on random patches
 *   orc16_spread, orc16_response_maps         pinned by brute force (window OR) and by the table lookups
 *   orc16_similarity, orc16_similarity_local  pinned by brute force
 *   orc16_match                               pinned through the above plus hand-checked micro cases
 *   orc16_extract_template, orc16_add_template pinned by invariants (exactly num_features features, refusal
 *                                             below num_features candidates, selection order); the reference
 *                                             holds no trained 16-orientation bank, so the selection itself
 *                                             is otherwise parity unpinned
 */
#ifndef SBM_ORACLE16_H
#define SBM_ORACLE16_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int orc16_angle_bin(float angle_deg);
void orc16_bin_table(int lo, int hi, uint8_t* out);
void orc16_gradient_color(const uint8_t* bgr, int rows, int cols, float weak_threshold, uint16_t* quantized,
                          float* magnitude, float* angle_ori);
void orc16_apply_mask(uint16_t* q, const uint8_t* mask, int n);
void orc16_spread(const uint16_t* q, int rows, int cols, int T, uint16_t* dst);
uint8_t orc16_similarity_lut(int ori, int idx);
uint8_t orc16_generated_lut(int ori, int idx);
void orc16_response_maps(const uint16_t* spread, int rows, int cols, uint8_t* maps);
/* lm_out[l]: 16*rows_l*cols_l bytes plus orc_lm_guard_bytes */
int orc16_build_lm_pyramid(const uint8_t* img, int rows, int cols, int channels, const uint8_t* mask,
                           float weak_threshold, int num_levels, const int* T, uint8_t** lm_out);
int orc16_similarity(const uint8_t* lm, int rows, int cols, int T, const int32_t* feats, int n, int width,
                     int height, uint16_t* dst);
int orc16_similarity_local(const uint8_t* lm, int rows, int cols, int T, const int32_t* feats, int n, int cx,
                           int cy, uint16_t* dst);
int orc16_match_lm(uint8_t* const* lm, int rows, int cols, int num_levels, const int* T, float threshold,
                   int num_templates, const int32_t* tinfo, const int32_t* feats, const int32_t* tclass,
                   const int32_t* tid, int max_out, int32_t* out_i32, float* out_sim, int nthreads);
int orc16_match(const uint8_t* img, int rows, int cols, int channels, const uint8_t* mask, float weak_threshold,
                int num_levels, const int* T, float threshold, int num_templates, const int32_t* tinfo,
                const int32_t* feats, const int32_t* tclass, const int32_t* tid, int max_out, int32_t* out_i32,
                float* out_sim, int nthreads);
int orc16_extract_template(const float* mag, const uint16_t* angle, const uint8_t* mask, int rows, int cols,
                           float strong, int num_features, int max_out, int32_t* out_xyl);
int orc16_add_template(const uint8_t* img, int rows, int cols, int channels, const uint8_t* mask, float weak,
                       float strong, int num_features, int num_levels, int max_feats, int32_t* tinfo6,
                       int32_t* feats_xyl);

#ifdef __cplusplus
}
#endif
#endif
