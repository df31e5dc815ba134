/*
 * CPU oracle for the 16-orientation variant of line2Dup -- TEST INFRASTRUCTURE ONLY (the checker), like
 * sbm_oracle.c, and loaded the same way.
 *
 * Restates test/ori_16bit_experiment/line2Dup_16bit_ori.cpp (cited as "16bit:<line>") and the table
 * generator test/ori_16bit_experiment/LUT_gen.cpp.  What differs from the 8-orientation method:
 *   - quantization: phase * 32/360 saturated to 16 bits, "& 15", a 16-bin 3x3 vote, quantized value
 *     1 << bin in 16 bits (16bit:216-296)
 *   - gradients: the 3-channel code only; its loop reads three interleaved channels per pixel
 *     (16bit:298-373), so a single-channel image is outside what the variant defines
 *   - spread: OR of 16-bit values (16bit:542-573)
 *   - responses: the compiled SIMILARITY_LUT[1024] (16bit:575-609; values 0, 1 and 4) looked up with the
 *     four nibbles of the spread value, maximum of the four; the fourth nibble is taken as
 *     (s & (15 << 16)) >> 16 (16bit:639), which is 0 for a 16-bit s, so bits 12..15 of the spread
 *     value never respond (16bit:611-700)
 *   - similarity: up to 63 features accumulate in 8 bits, otherwise in 16 bits, for fewer than 16384
 *     features (16bit:760-1030); with responses of at most 4 neither sum can wrap
 *   - matchClass: a level with 16384 or more features calls neither similarity function, so its
 *     coarse map is empty and its refinement windows score 0 (16bit:1151-1296)
 *   - training: fewer candidates than num_features refuses the template (16bit:479-481), and
 *     selectScatteredFeatures stops at exactly num_features (16bit:167-210); no feature angle is kept
 *   - defaults: weak 10, strong 55, 63 features, T {5, 8} (16bit:500-512, 1036-1056)
 */
#include "sbm_oracle16.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc_internal.h"
#include "sbm_oracle.h"

/* angle.convertTo(q, CV_16U, 32.0/360.0): saturate_cast<ushort>(RN(a * (float)(32/360))), round half to
 * even (16bit:222-223). */
int orc16_angle_bin(float a) {
    const float scale = (float)(32.0 / 360.0);
    long q = lrintf(a * scale);
    if (q < 0) q = 0;
    if (q > 65535) q = 65535;
    return (int)q;
}

/* Exhaustive (dx,dy) -> 16-bin table for dx,dy in [lo,hi]: out[(dy-lo)*(hi-lo+1) + (dx-lo)] */
void orc16_bin_table(int lo, int hi, uint8_t* out) {
    int n = hi - lo + 1;
    for (int dy = lo; dy <= hi; ++dy)
        for (int dx = lo; dx <= hi; ++dx)
            out[(size_t)(dy - lo) * n + (dx - lo)] =
                (uint8_t)(orc16_angle_bin(orc_fast_atan2_deg((float)dy, (float)dx)) & 15);
}

/* hysteresisGradient, 16 bins                                                       16bit:216-296 */
static void hysteresis16(const float* mag, const float* ag, int rows, int cols, float threshold, uint16_t* out) {
    size_t n = (size_t)rows * cols;
    uint16_t* qu = (uint16_t*)malloc(n * 2);
    for (size_t i = 0; i < n; ++i) qu[i] = (uint16_t)orc16_angle_bin(ag[i]);
    for (int c = 0; c < cols; ++c) {
        qu[c] = 0;
        qu[(size_t)(rows - 1) * cols + c] = 0;
    }
    for (int r = 0; r < rows; ++r) {
        qu[(size_t)r * cols] = 0;
        qu[(size_t)r * cols + cols - 1] = 0;
    }
    for (int r = 1; r < rows - 1; ++r)
        for (int c = 1; c < cols - 1; ++c) qu[(size_t)r * cols + c] &= 15;
    memset(out, 0, n * 2);
    for (int r = 1; r < rows - 1; ++r)
        for (int c = 1; c < cols - 1; ++c) {
            if (!(mag[(size_t)r * cols + c] > threshold)) continue;
            int h[16] = {0};
            for (int dr = -1; dr <= 1; ++dr)
                for (int dc = -1; dc <= 1; ++dc) h[qu[(size_t)(r + dr) * cols + c + dc]]++;
            int max_votes = 0, index = -1;
            for (int i = 0; i < 16; ++i)
                if (max_votes < h[i]) {
                    index = i;
                    max_votes = h[i];
                }
            if (max_votes >= 5) out[(size_t)r * cols + c] = (uint16_t)(1u << index);
        }
    free(qu);
}

/* quantizedOrientations (3-channel code; the magnitude and phase are those of the 8-orientation
 * colour branch, line2Dup.cpp:304-374 == 16bit:298-373) */
void orc16_gradient_color(const uint8_t* bgr, int rows, int cols, float weak_threshold, uint16_t* quantized,
                          float* magnitude, float* angle_ori) {
    size_t n = (size_t)rows * cols;
    uint8_t* q8 = (uint8_t*)malloc(n);
    float* mag = (float*)malloc(n * 4);
    float* ag = (float*)malloc(n * 4);
    orc_gradient_color(bgr, rows, cols, weak_threshold, q8, mag, ag);
    hysteresis16(mag, ag, rows, cols, weak_threshold * weak_threshold, quantized);
    if (magnitude) memcpy(magnitude, mag, n * 4);
    if (angle_ori) memcpy(angle_ori, ag, n * 4);
    free(q8); free(mag); free(ag);
}

/* ColorGradientPyramid::quantize: dst = 0; angle.copyTo(dst, mask)                   16bit:415-419 */
void orc16_apply_mask(uint16_t* q, const uint8_t* mask, int n) {
    if (!mask) return;
    for (int i = 0; i < n; ++i)
        if (!mask[i]) q[i] = 0;
}

/* spread + orUnaligned16u: OR over the forward TxT window, clipped at the bottom/right edge
 *                                                                                     16bit:542-573 */
void orc16_spread(const uint16_t* q, int rows, int cols, int T, uint16_t* dst) {
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            uint16_t v = 0;
            for (int dr = 0; dr < T && r + dr < rows; ++dr)
                for (int dc = 0; dc < T && c + dc < cols; ++dc) v |= q[(size_t)(r + dr) * cols + c + dc];
            dst[(size_t)r * cols + c] = v;
        }
}

/* SIMILARITY_LUT[1024] as compiled into the variant (16bit:575-609), by its closed form: entry
 * 64*ori + 16*m + n is the maximum over the set bits b of the nibble n (orientation 4m + b) of
 *   4 when the circular distance on the ring of 16 orientations between ori and 4m + b is at most 2,
 *   1 when it is 3 or 4, else 0.
 * LUT16.txt beside it holds another table (8 minus the distance, the output of LUT_gen.cpp:11-60) that
 * the variant does not compile; orc16_generated_lut restates that one. */
static int ring_distance16(int a, int b) {
    int d = (a - b + 16) & 15;
    return 16 - d < d ? 16 - d : d;
}

uint8_t orc16_similarity_lut(int ori, int idx) {
    int m = idx >> 4, nib = idx & 15, best = 0;
    for (int bit = 0; bit < 4; ++bit) {
        if (!(nib & (1 << bit))) continue;
        int d = ring_distance16(ori, bit + 4 * m), v = d <= 2 ? 4 : (d <= 4 ? 1 : 0);
        if (v > best) best = v;
    }
    return (uint8_t)best;
}

/* LUT_gen.cpp:11-60: 0 for n == 0, else 8 - d with d the distance from ori to the nearest set bit of
 * n << 4m (its forward and backward walks stop at the first step that reaches one). */
uint8_t orc16_generated_lut(int ori, int idx) {
    int m = idx >> 4, nib = idx & 15, best = 16;
    if (nib == 0) return 0;
    for (int bit = 0; bit < 4; ++bit)
        if (nib & (1 << bit)) {
            int d = ring_distance16(ori, bit + 4 * m);
            if (d < best) best = d;
        }
    return (uint8_t)(8 - best);
}

/* computeResponseMaps (scalar branch, 16bit:676-697; the SSSE3 branch computes the same): maps holds
 * 16 planes of rows*cols bytes. */
void orc16_response_maps(const uint16_t* s, int rows, int cols, uint8_t* maps) {
    size_t n = (size_t)rows * cols;
    for (int ori = 0; ori < 16; ++ori)
        for (size_t i = 0; i < n; ++i) {
            const uint32_t v = s[i];
            const int n0 = (int)(v & 15u);
            const int n1 = (int)((v & (15u << 4)) >> 4);
            const int n2 = (int)((v & (15u << 8)) >> 8);
            const int n3 = (int)((v & (15u << 16)) >> 16); /* 16bit:639: always 0 for a 16-bit value */
            uint8_t a = orc16_similarity_lut(ori, n0), b = orc16_similarity_lut(ori, 16 + n1);
            uint8_t c = orc16_similarity_lut(ori, 32 + n2), d = orc16_similarity_lut(ori, 48 + n3);
            uint8_t m01 = a > b ? a : b, m23 = c > d ? c : d;
            maps[ori * n + i] = m01 > m23 ? m01 : m23;
        }
}

/* Detector::match, response-map construction with 16 linear memories per level     16bit:1058-1104
 * lm_out[l] holds 16*rows_l*cols_l bytes ([ori 16][T*T][H*W]) followed by a zeroed guard. */
int orc16_build_lm_pyramid(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, int L,
                           const int* T, uint8_t** lm_out) {
    if (ch != 3) return ORC_EINVAL;
    uint8_t* cur = (uint8_t*)malloc((size_t)rows * cols * 3);
    memcpy(cur, img, (size_t)rows * cols * 3);
    uint8_t* cmask = NULL;
    if (mask) {
        cmask = (uint8_t*)malloc((size_t)rows * cols);
        memcpy(cmask, mask, (size_t)rows * cols);
    }
    int r = rows, c = cols;
    for (int l = 0; l < L; ++l) {
        if (l > 0) {
            int nr = r / 2, nc = c / 2;
            uint8_t* nxt = (uint8_t*)malloc((size_t)nr * nc * 3);
            orc_pyr_down(cur, r, c, 3, nxt);
            free(cur);
            cur = nxt;
            if (cmask) {
                uint8_t* nm = (uint8_t*)malloc((size_t)nr * nc);
                orc_mask_down(cmask, r, c, nm);
                free(cmask);
                cmask = nm;
            }
            r = nr;
            c = nc;
        }
        size_t n = (size_t)r * c;
        if ((n % 16) != 0 || r % T[l] != 0 || c % T[l] != 0) {
            free(cur);
            free(cmask);
            return ORC_EINVAL;
        }
        uint16_t* q = (uint16_t*)malloc(n * 2);
        uint16_t* sp = (uint16_t*)malloc(n * 2);
        uint8_t* resp = (uint8_t*)malloc(n * 16);
        orc16_gradient_color(cur, r, c, weak, q, NULL, NULL);
        orc16_apply_mask(q, cmask, (int)n);
        orc16_spread(q, r, c, T[l], sp);
        orc16_response_maps(sp, r, c, resp);
        size_t hw = (size_t)(r / T[l]) * (c / T[l]);
        for (int ori = 0; ori < 16; ++ori)
            orc_linearize(resp + ori * n, r, c, T[l], lm_out[l] + (size_t)ori * T[l] * T[l] * hw);
        memset(lm_out[l] + 16 * n, 0, (size_t)orc_lm_guard_bytes(r, c, T[l]));
        free(q); free(sp); free(resp);
    }
    free(cur);
    free(cmask);
    return 0;
}

/* accessLinearMemory over the flat [ori 16][T*T][H*W] block                         16bit:735-758 */
static inline const uint8_t* access_lm16(const uint8_t* lm, int fx, int fy, int label, int T, int W, size_t hw) {
    const uint8_t* grid = lm + (size_t)label * T * T * hw;
    int gi = (fy % T) * T + fx % T;
    return grid + (size_t)gi * hw + (size_t)(fy / T) * W + fx / T;
}

/* similarity_64 (n <= 63: 8-bit sums) and similarity (n < 16384: 16-bit sums), as the 16-bit map that
 * matchClass reads (16bit:760-815, 873-956, 1183-1186).  Returns template_positions, or EINVAL when
 * n >= 16384 (neither function runs). */
int orc16_similarity(const uint8_t* lm, int rows, int cols, int T, const int32_t* feats, int n, int width, int height,
                     uint16_t* dst) {
    if (n >= 16384) return ORC_EINVAL;
    int W = cols / T, H = rows / T;
    size_t hw = (size_t)W * H;
    int wf = (width - 1) / T + 1, hf = (height - 1) / T + 1;
    int span_x = W - wf, span_y = H - hf;
    int P = span_y * W + span_x + 1;
    const int narrow = n < 64;
    memset(dst, 0, hw * 2);
    for (int i = 0; i < n; ++i) {
        int fx = feats[3 * i], fy = feats[3 * i + 1], lab = feats[3 * i + 2];
        if (fx < 0 || fx >= cols || fy < 0 || fy >= rows) continue;
        const uint8_t* p = access_lm16(lm, fx, fy, lab, T, W, hw);
        if (narrow)
            for (int j = 0; j < P; ++j) dst[j] = (uint8_t)(dst[j] + p[j]);
        else
            for (int j = 0; j < P; ++j) dst[j] = (uint16_t)(dst[j] + p[j]);
    }
    return P;
}

/* similarityLocal_64 / similarityLocal over the 16x16 window                     16bit:817-871, 958-1030 */
int orc16_similarity_local(const uint8_t* lm, int rows, int cols, int T, const int32_t* feats, int n, int cx, int cy,
                           uint16_t* dst) {
    if (n >= 16384) return ORC_EINVAL;
    int W = cols / T, H = rows / T;
    size_t hw = (size_t)W * H;
    int offx = (cx / T - 8) * T, offy = (cy / T - 8) * T;
    const int narrow = n < 64;
    memset(dst, 0, 256 * 2);
    for (int i = 0; i < n; ++i) {
        int fx = feats[3 * i] + offx, fy = feats[3 * i + 1] + offy, lab = feats[3 * i + 2];
        if (fx < 0 || fy < 0 || fx >= cols || fy >= rows) continue;
        const uint8_t* p = access_lm16(lm, fx, fy, lab, T, W, hw);
        for (int r = 0; r < 16; ++r)
            for (int c = 0; c < 16; ++c) {
                uint16_t* d = &dst[r * 16 + c];
                *d = narrow ? (uint8_t)(*d + p[(size_t)r * W + c]) : (uint16_t)(*d + p[(size_t)r * W + c]);
            }
    }
    return 0;
}

/* Detector::matchClass for one template pyramid                                     16bit:1140-1299 */
static void match_template16(uint8_t* const* lm, const int* lrows, const int* lcols, int L, const int* T,
                             float threshold, const int32_t* ti, const int32_t* feats, int cls, int tid, omvec* out) {
    const int lowest = L - 1;
    const int32_t* tl = ti + 6 * lowest;
    int nfeat = tl[4];
    int rows = lrows[lowest], cols = lcols[lowest], TL = T[lowest];
    int W = cols / TL, H = rows / TL;
    omvec cand = {0, 0, 0};
    if (nfeat < 16384) { /* else the similarity map stays empty: no candidates */
        uint16_t* sim = (uint16_t*)malloc((size_t)W * H * 2);
        orc16_similarity(lm[lowest], rows, cols, TL, feats + 3 * (size_t)tl[5], nfeat, tl[0], tl[1], sim);
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                int raw = sim[r * W + c];
                float score = (raw * 100.f) / (4 * nfeat);
                if (score > threshold) {
                    int offset = TL / 2 + (TL % 2 - 1);
                    omatch m = {c * TL + offset, r * TL + offset, cls, tid, score};
                    orc_i_omvec_push(&cand, m);
                }
            }
        free(sim);
    }
    uint16_t loc[256];
    for (int l = L - 2; l >= 0; --l) {
        const int32_t* tlv = ti + 6 * l;
        int Tl = T[l];
        int rl = lrows[l], cl = lcols[l];
        int border = 8 * Tl;
        int offset = Tl / 2 + (Tl % 2 - 1);
        int max_x = cl - tlv[0] - border;
        int max_y = rl - tlv[1] - border;
        int nf = tlv[4];
        for (int m = 0; m < cand.n; ++m) {
            omatch* mm = &cand.v[m];
            int x = mm->x * 2 + 1, y = mm->y * 2 + 1;
            x = x > border ? x : border;
            y = y > border ? y : border;
            x = x < max_x ? x : max_x;
            y = y < max_y ? y : max_y;
            float best = 0;
            int br = -1, bc = -1;
            if (nf < 16384) { /* else no local map: best_score stays 0 */
                orc16_similarity_local(lm[l], rl, cl, Tl, feats + 3 * (size_t)tlv[5], nf, x, y, loc);
                for (int r = 0; r < 16; ++r)
                    for (int c = 0; c < 16; ++c) {
                        float s = (loc[r * 16 + c] * 100.f) / (4 * nf);
                        if (s > best) { best = s; br = r; bc = c; }
                    }
            }
            mm->sim = best;
            mm->x = (x / Tl - 8 + bc) * Tl + offset;
            mm->y = (y / Tl - 8 + br) * Tl + offset;
        }
        int w = 0;
        for (int m = 0; m < cand.n; ++m)
            if (!(cand.v[m].sim < threshold)) cand.v[w++] = cand.v[m];
        cand.n = w;
    }
    for (int m = 0; m < cand.n; ++m) orc_i_omvec_push(out, cand.v[m]);
    free(cand.v);
}

int orc16_match_lm(uint8_t* const* lm, int rows, int cols, int L, const int* T, float threshold, int num_templates,
                   const int32_t* tinfo, const int32_t* feats, const int32_t* tclass, const int32_t* tid, int max_out,
                   int32_t* out_i32, float* out_sim, int nthreads) {
    int lrows[16], lcols[16];
    if (L < 1 || L > 16) return ORC_EINVAL;
    lrows[0] = rows; lcols[0] = cols;
    for (int l = 1; l < L; ++l) { lrows[l] = lrows[l - 1] / 2; lcols[l] = lcols[l - 1] / 2; }
    omvec* per = (omvec*)calloc((size_t)(num_templates > 0 ? num_templates : 1), sizeof(omvec));
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int t = 0; t < num_templates; ++t)
        match_template16(lm, lrows, lcols, L, T, threshold, tinfo + (size_t)6 * L * t, feats, tclass[t], tid[t], &per[t]);
    (void)nthreads;
    omvec all = {0, 0, 0};
    for (int t = 0; t < num_templates; ++t) {
        for (int i = 0; i < per[t].n; ++i) orc_i_omvec_push(&all, per[t].v[i]);
        free(per[t].v);
    }
    free(per);
    int m = orc_i_sort_unique(all.v, all.n);
    if (m > max_out) { free(all.v); return -m; }
    for (int i = 0; i < m; ++i) {
        out_i32[4 * i] = all.v[i].x; out_i32[4 * i + 1] = all.v[i].y;
        out_i32[4 * i + 2] = all.v[i].cls; out_i32[4 * i + 3] = all.v[i].tid;
        out_sim[i] = all.v[i].sim;
    }
    free(all.v);
    return m;
}

/* Detector::match                                                                  16bit:1058-1130 */
int orc16_match(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, int L, const int* T,
                float threshold, int num_templates, const int32_t* tinfo, const int32_t* feats, const int32_t* tclass,
                const int32_t* tid, int max_out, int32_t* out_i32, float* out_sim, int nthreads) {
    if (L < 1 || L > 16 || ch != 3) return ORC_EINVAL;
    uint8_t* lm[16];
    int r = rows, c = cols;
    for (int l = 0; l < L; ++l) {
        if (l > 0) { r /= 2; c /= 2; }
        if (T[l] <= 0) {
            for (int k = 0; k < l; ++k) free(lm[k]);
            return ORC_EINVAL;
        }
        lm[l] = (uint8_t*)malloc((size_t)r * c * 16 + (size_t)orc_lm_guard_bytes(r, c, T[l]));
    }
    int rc = orc16_build_lm_pyramid(img, rows, cols, ch, mask, weak, L, T, lm);
    if (rc == 0)
        rc = orc16_match_lm(lm, rows, cols, L, T, threshold, num_templates, tinfo, feats, tclass, tid, max_out, out_i32,
                            out_sim, nthreads);
    for (int l = 0; l < L; ++l) free(lm[l]);
    return rc;
}

/* selectScatteredFeatures: relax the distance after each pass until exactly num_features are taken
 *                                                                                     16bit:167-210 */
static int select_scattered16(const ocand* cands, int nc, int num_features, float distance, ocand* feats) {
    int nf = 0, i = 0;
    float dsq = distance * distance;
    while (nf < num_features) {
        ocand c = cands[i];
        int keep = 1;
        for (int j = 0; j < nf && keep; ++j) {
            int ddx = c.x - feats[j].x, ddy = c.y - feats[j].y;
            keep = (float)(ddx * ddx + ddy * ddy) >= dsq;
        }
        if (keep) feats[nf++] = c;
        if (++i == nc) {
            i = 0;
            distance -= 1.0f;
            dsq = distance * distance;
        }
    }
    return nf;
}

/* ColorGradientPyramid::extractTemplate (features only)                            16bit:421-498
 * Returns the feature count (always num_features), or -1 when the template is refused. */
int orc16_extract_template(const float* mag, const uint16_t* angle, const uint8_t* mask, int rows, int cols, float strong,
                           int num_features, int max_out, int32_t* out_xyl) {
    if (num_features <= 0) return -1; /* the reference divides by it (16bit:484) */
    ocand* cands = NULL;
    int nc = orc_i_extract_candidates(mag, NULL, angle, NULL, mask, rows, cols, strong, &cands);
    if (nc < num_features) { free(cands); return -1; }
    ocand* sel = (ocand*)malloc(sizeof(ocand) * (size_t)num_features);
    float distance = (float)(nc / num_features + 1);
    int nf = select_scattered16(cands, nc, num_features, distance, sel);
    int out = nf < max_out ? nf : max_out;
    for (int i = 0; i < out; ++i) {
        out_xyl[3 * i] = sel[i].x; out_xyl[3 * i + 1] = sel[i].y; out_xyl[3 * i + 2] = sel[i].label;
    }
    free(sel);
    free(cands);
    return nf;
}

/* Detector::addTemplate                                                            16bit:1301-1335
 * tinfo6 per level: width, height, tl_x, tl_y, num_features, feature_offset.  Returns the total feature
 * count written, or -1 when a level refuses the template. */
int orc16_add_template(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, float strong,
                       int num_features, int L, int max_feats, int32_t* tinfo6, int32_t* feats_xyl) {
    if (ch != 3 || L < 1) return -1;
    uint8_t* cur = (uint8_t*)malloc((size_t)rows * cols * 3);
    memcpy(cur, img, (size_t)rows * cols * 3);
    uint8_t* cm = NULL;
    if (mask) { cm = (uint8_t*)malloc((size_t)rows * cols); memcpy(cm, mask, (size_t)rows * cols); }
    int r = rows, c = cols, total = 0, nfl = num_features, ok = 1;
    int32_t* ti5 = (int32_t*)malloc(sizeof(int32_t) * 5 * (size_t)L);
    int* nfeat = (int*)calloc((size_t)L, sizeof(int));
    int32_t** fptr = (int32_t**)malloc(sizeof(int32_t*) * (size_t)L);
    for (int l = 0; l < L && ok; ++l) {
        if (l > 0) {
            nfl /= 2;
            int nr = r / 2, nc = c / 2;
            uint8_t* nx = (uint8_t*)malloc((size_t)nr * nc * 3);
            orc_pyr_down(cur, r, c, 3, nx);
            free(cur);
            cur = nx;
            if (cm) {
                uint8_t* nm = (uint8_t*)malloc((size_t)nr * nc);
                orc_mask_down(cm, r, c, nm);
                free(cm);
                cm = nm;
            }
            r = nr;
            c = nc;
        }
        size_t n = (size_t)r * c;
        uint16_t* q = (uint16_t*)malloc(n * 2);
        float* mag = (float*)malloc(n * 4);
        orc16_gradient_color(cur, r, c, weak, q, mag, NULL);
        int got = orc16_extract_template(mag, q, cm, r, c, strong, nfl, max_feats - total, feats_xyl + 3 * (size_t)total);
        free(q); free(mag);
        if (got < 0 || total + got > max_feats) { ok = 0; break; }
        fptr[l] = feats_xyl + 3 * (size_t)total;
        nfeat[l] = got;
        ti5[5 * l + 4] = l;
        tinfo6[6 * l + 4] = got;
        tinfo6[6 * l + 5] = total;
        total += got;
    }
    if (ok) {
        orc_crop_templates(L, ti5, nfeat, fptr);
        for (int l = 0; l < L; ++l)
            for (int k = 0; k < 4; ++k) tinfo6[6 * l + k] = ti5[5 * l + k];
    }
    free(cur); free(cm); free(ti5); free(nfeat); free(fptr);
    return ok ? total : -1;
}
