/*
 * CPU oracle for line2Dup's matching path -- TEST INFRASTRUCTURE ONLY (the checker).
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py may load this library.
 *
 * Plain C, built with -ffp-contract=off and without -ffast-math so that the float steps are
 * evaluated operation by operation in IEEE single precision, as the reference's SSE code does.
 * Citations are to the reference sources (line2Dup.cpp, line2Dup.h, test.cpp).
 */
#include "sbm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc_internal.h"

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------------------------
 * GaussianBlur(src, dst, Size(7,7), 0, 0, BORDER_REPLICATE)        line2Dup.cpp:291-293
 * sigma = 0 with ksize 7 selects OpenCV's fixed kernel [4,14,28,36,28,14,4]/128 per axis.
 * The 8-bit fixed-point path (kernel scaled by 256, 16 fractional bits, round half up) gives
 * floor((sum_ij k_i k_j p + 8192) / 16384) exactly.
 * ------------------------------------------------------------------------------------------ */
void orc_gaussian_blur7(const uint8_t* src, int rows, int cols, uint8_t* dst) {
    static const int k[7] = {4, 14, 28, 36, 28, 14, 4};
    int* tmp = (int*)malloc(sizeof(int) * (size_t)rows * cols);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            int s = 0;
            for (int j = 0; j < 7; ++j) s += k[j] * src[(size_t)r * cols + clampi(c + j - 3, 0, cols - 1)];
            tmp[(size_t)r * cols + c] = s;
        }
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            int s = 0;
            for (int i = 0; i < 7; ++i) s += k[i] * tmp[(size_t)clampi(r + i - 3, 0, rows - 1) * cols + c];
            dst[(size_t)r * cols + c] = (uint8_t)((s + 8192) >> 14);
        }
    free(tmp);
}

/* Sobel(smoothed, dx, CV_32F, 1, 0, 3, 1, 0, BORDER_REPLICATE) and the dy twin   line2Dup.cpp:297-298
 * Kernels [-1,0,1] x [1,2,1]; results are exact integers in [-1020, 1020]. */
void orc_sobel3(const uint8_t* b, int rows, int cols, int32_t* dx, int32_t* dy) {
    for (int r = 0; r < rows; ++r) {
        int ru = clampi(r - 1, 0, rows - 1), rd = clampi(r + 1, 0, rows - 1);
        for (int c = 0; c < cols; ++c) {
            int cl = clampi(c - 1, 0, cols - 1), cr = clampi(c + 1, 0, cols - 1);
#define B(rr, cc) ((int)b[(size_t)(rr) * cols + (cc)])
            int gx = (B(ru, cr) - B(ru, cl)) + 2 * (B(r, cr) - B(r, cl)) + (B(rd, cr) - B(rd, cl));
            int gy = (B(rd, cl) - B(ru, cl)) + 2 * (B(rd, c) - B(ru, c)) + (B(rd, cr) - B(ru, cr));
#undef B
            dx[(size_t)r * cols + c] = gx;
            dy[(size_t)r * cols + c] = gy;
        }
    }
}

/* phase(dx, dy, angle, true)                                              line2Dup.cpp:300
 * OpenCV 3.x fastAtan2: 7th-order odd polynomial in c = min/(max + DBL_EPSILON), degrees,
 * every operation rounded to float, no fused multiply-add. */
float orc_fast_atan2_deg(float y, float x) {
    const float s = (float)(180.0 / 3.1415926535897932384626433832795);
    const float p1 = 0.9997878412794807f * s;
    const float p3 = -0.3258083974640975f * s;
    const float p5 = 0.1555786518463281f * s;
    const float p7 = -0.04432655554792128f * s;
    const float eps = (float)DBL_EPSILON;
    float ax = fabsf(x), ay = fabsf(y), a, c, c2;
    if (ax >= ay) {
        c = ay / (ax + eps);
        c2 = c * c;
        a = (((p7 * c2 + p5) * c2 + p3) * c2 + p1) * c;
    } else {
        c = ax / (ay + eps);
        c2 = c * c;
        a = 90.f - (((p7 * c2 + p5) * c2 + p3) * c2 + p1) * c;
    }
    if (x < 0) a = 180.f - a;
    if (y < 0) a = 360.f - a;
    return a;
}

/* angle.convertTo(q, CV_8U, 16.0/360.0): saturate_cast<uchar>(a * (float)(16/360)), round half to
 * even, then "& 7" for the interior (line2Dup.cpp:213, 227-234). Returns the 16-bucket value. */
int orc_angle_bin(float a) {
    const float scale = (float)(16.0 / 360.0);
    float v = a * scale;
    long q = lrintf(v);
    if (q < 0) q = 0;
    if (q > 255) q = 255;
    return (int)q;
}

/* Exhaustive (dx,dy) -> 8-bin table for dx,dy in [lo,hi]: out[(dy-lo)*(hi-lo+1) + (dx-lo)]. */
void orc_bin_table(int lo, int hi, uint8_t* out) {
    int n = hi - lo + 1;
    for (int dy = lo; dy <= hi; ++dy)
        for (int dx = lo; dx <= hi; ++dx)
            out[(size_t)(dy - lo) * n + (dx - lo)] =
                (uint8_t)(orc_angle_bin(orc_fast_atan2_deg((float)dy, (float)dx)) & 7);
}

/* hysteresisGradient                                                       line2Dup.cpp:206-284 */
static void hysteresis(const float* mag, const float* ag, int rows, int cols, float threshold, uint8_t* out) {
    uint8_t* qu = (uint8_t*)malloc((size_t)rows * cols);
    for (size_t i = 0; i < (size_t)rows * cols; ++i) qu[i] = (uint8_t)orc_angle_bin(ag[i]);
    memset(qu, 0, (size_t)cols);
    memset(qu + (size_t)(rows - 1) * cols, 0, (size_t)cols);
    for (int r = 0; r < rows; ++r) {
        qu[(size_t)r * cols] = 0;
        qu[(size_t)r * cols + cols - 1] = 0;
    }
    for (int r = 1; r < rows - 1; ++r)
        for (int c = 1; c < cols - 1; ++c) qu[(size_t)r * cols + c] &= 7;
    memset(out, 0, (size_t)rows * cols);
    for (int r = 1; r < rows - 1; ++r)
        for (int c = 1; c < cols - 1; ++c) {
            if (!(mag[(size_t)r * cols + c] > threshold)) continue;
            int h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int dr = -1; dr <= 1; ++dr)
                for (int dc = -1; dc <= 1; ++dc) h[qu[(size_t)(r + dr) * cols + c + dc]]++;
            int max_votes = 0, index = -1;
            for (int i = 0; i < 8; ++i)
                if (max_votes < h[i]) {
                    index = i;
                    max_votes = h[i];
                }
            if (max_votes >= 5) out[(size_t)r * cols + c] = (uint8_t)(1 << index);
        }
    free(qu);
}

/* quantizedOrientations, single-channel branch                            line2Dup.cpp:286-302 */
void orc_quantized_orientations(const uint8_t* src, int rows, int cols, float weak_threshold,
                                uint8_t* quantized, float* magnitude, float* angle_ori) {
    size_t n = (size_t)rows * cols;
    uint8_t* bl = (uint8_t*)malloc(n);
    int32_t* dx = (int32_t*)malloc(n * 4);
    int32_t* dy = (int32_t*)malloc(n * 4);
    float* mag = (float*)malloc(n * 4);
    float* ag = (float*)malloc(n * 4);
    orc_gaussian_blur7(src, rows, cols, bl);
    orc_sobel3(bl, rows, cols, dx, dy);
    for (size_t i = 0; i < n; ++i) {
        float fx = (float)dx[i], fy = (float)dy[i];
        mag[i] = fx * fx + fy * fy;
        ag[i] = orc_fast_atan2_deg(fy, fx);
    }
    hysteresis(mag, ag, rows, cols, weak_threshold * weak_threshold, quantized);
    if (magnitude) memcpy(magnitude, mag, n * 4);
    if (angle_ori) memcpy(angle_ori, ag, n * 4);
    free(bl); free(dx); free(dy); free(mag); free(ag);
}

/* quantizedOrientations, 3-channel branch (BGR interleaved)                line2Dup.cpp:304-374
 * Per-channel blur and 16S Sobel; keep the channel with the largest dx^2+dy^2, ties to the lower
 * channel (the comparison chain at :343-360). */
void orc_gradient_color(const uint8_t* bgr, int rows, int cols, float weak_threshold,
                        uint8_t* quantized, float* magnitude, float* angle_ori) {
    size_t n = (size_t)rows * cols;
    uint8_t* plane = (uint8_t*)malloc(n);
    uint8_t* bl = (uint8_t*)malloc(n);
    int32_t* dx = (int32_t*)malloc(n * 4 * 3);
    int32_t* dy = (int32_t*)malloc(n * 4 * 3);
    for (int ch = 0; ch < 3; ++ch) {
        for (size_t i = 0; i < n; ++i) plane[i] = bgr[i * 3 + ch];
        orc_gaussian_blur7(plane, rows, cols, bl);
        orc_sobel3(bl, rows, cols, dx + ch * n, dy + ch * n);
    }
    float* fdx = (float*)malloc(n * 4);
    float* fdy = (float*)malloc(n * 4);
    float* mag = (float*)malloc(n * 4);
    float* ag = (float*)malloc(n * 4);
    for (size_t i = 0; i < n; ++i) {
        int m1 = dx[i] * dx[i] + dy[i] * dy[i];
        int m2 = dx[n + i] * dx[n + i] + dy[n + i] * dy[n + i];
        int m3 = dx[2 * n + i] * dx[2 * n + i] + dy[2 * n + i] * dy[2 * n + i];
        int ch, m;
        if (m1 >= m2 && m1 >= m3) { ch = 0; m = m1; }
        else if (m2 >= m1 && m2 >= m3) { ch = 1; m = m2; }
        else { ch = 2; m = m3; }
        fdx[i] = (float)dx[ch * n + i];
        fdy[i] = (float)dy[ch * n + i];
        mag[i] = (float)m;
        ag[i] = orc_fast_atan2_deg(fdy[i], fdx[i]);
    }
    hysteresis(mag, ag, rows, cols, weak_threshold * weak_threshold, quantized);
    if (magnitude) memcpy(magnitude, mag, n * 4);
    if (angle_ori) memcpy(angle_ori, ag, n * 4);
    free(plane); free(bl); free(dx); free(dy); free(fdx); free(fdy); free(mag); free(ag);
}

/* cv::pyrDown(src, dst, Size(cols/2, rows/2))                              line2Dup.cpp:404-406
 * 5x5 kernel [1,4,6,4,1]^2/256, round half up, BORDER_REFLECT_101, even source pixels. */
static inline int reflect101(int p, int n) {
    if (n == 1) return 0;
    while (p < 0 || p >= n) {
        if (p < 0) p = -p;
        if (p >= n) p = 2 * n - 2 - p;
    }
    return p;
}
void orc_pyr_down(const uint8_t* src, int rows, int cols, int ch, uint8_t* dst) {
    static const int w[5] = {1, 4, 6, 4, 1};
    int dr = rows / 2, dc = cols / 2;
    for (int y = 0; y < dr; ++y)
        for (int x = 0; x < dc; ++x)
            for (int k = 0; k < ch; ++k) {
                int s = 0;
                for (int i = 0; i < 5; ++i) {
                    int sr = reflect101(2 * y + i - 2, rows);
                    for (int j = 0; j < 5; ++j) {
                        int sc = reflect101(2 * x + j - 2, cols);
                        s += w[i] * w[j] * src[((size_t)sr * cols + sc) * ch + k];
                    }
                }
                dst[((size_t)y * dc + x) * ch + k] = (uint8_t)((s + 128) >> 8);
            }
}

/* resize(mask, next, size, 0, 0, INTER_NEAREST) with an exact factor 2     line2Dup.cpp:409-414 */
void orc_mask_down(const uint8_t* src, int rows, int cols, uint8_t* dst) {
    int dr = rows / 2, dc = cols / 2;
    for (int y = 0; y < dr; ++y)
        for (int x = 0; x < dc; ++x) dst[(size_t)y * dc + x] = src[(size_t)(2 * y) * cols + 2 * x];
}

/* ColorGradientPyramid::quantize: dst = 0; angle.copyTo(dst, mask)        line2Dup.cpp:419-423 */
void orc_apply_mask(uint8_t* q, const uint8_t* mask, int n) {
    if (!mask) return;
    for (int i = 0; i < n; ++i)
        if (!mask[i]) q[i] = 0;
}

/* spread + orUnaligned8u: OR over the forward TxT window, clipped at the bottom/right edge
 *                                                                          line2Dup.cpp:556-603 */
void orc_spread(const uint8_t* q, int rows, int cols, int T, uint8_t* dst) {
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            uint8_t v = 0;
            for (int dr = 0; dr < T && r + dr < rows; ++dr)
                for (int dc = 0; dc < T && c + dc < cols; ++dc) v |= q[(size_t)(r + dr) * cols + c + dc];
            dst[(size_t)r * cols + c] = v;
        }
}

/* SIMILARITY_LUT                                                           line2Dup.cpp:605-608
 * The 256-entry table restated by its closed form: entry 32*ori + (half ? 16 : 0) + nib is the
 * maximum over the set bits b of the nibble (orientation b + 4*half) of 4 if b == ori, 3 if b is
 * a circular neighbour of ori, else 0.  The test suite checks this against the literal table. */
uint8_t orc_similarity_lut(int ori, int idx) {
    int half = idx >= 16, nib = idx & 15, best = 0;
    for (int bit = 0; bit < 4; ++bit) {
        if (!(nib & (1 << bit))) continue;
        int b = bit + 4 * half, d = (b - ori + 8) & 7, v = 0;
        if (d == 0) v = 4;
        else if (d == 1 || d == 7) v = 3;
        if (v > best) best = v;
    }
    return (uint8_t)best;
}

/* computeResponseMaps                                                       line2Dup.cpp:610-720
 * maps: 8 planes of rows*cols bytes. */
void orc_response_maps(const uint8_t* s, int rows, int cols, uint8_t* maps) {
    size_t n = (size_t)rows * cols;
    for (int ori = 0; ori < 8; ++ori)
        for (size_t i = 0; i < n; ++i) {
            uint8_t lo = orc_similarity_lut(ori, s[i] & 15);
            uint8_t hi = orc_similarity_lut(ori, 16 + ((s[i] & 240) >> 4));
            maps[ori * n + i] = lo > hi ? lo : hi;
        }
}

/* linearize: response map -> T*T linear memories of (rows/T)*(cols/T) bytes   line2Dup.cpp:722-750 */
void orc_linearize(const uint8_t* resp, int rows, int cols, int T, uint8_t* lm) {
    int W = cols / T, H = rows / T;
    size_t hw = (size_t)W * H;
    int index = 0;
    for (int r0 = 0; r0 < T; ++r0)
        for (int c0 = 0; c0 < T; ++c0) {
            uint8_t* mem = lm + (size_t)index * hw;
            ++index;
            for (int r = r0; r < rows; r += T)
                for (int c = c0; c < cols; c += T) *mem++ = resp[(size_t)r * cols + c];
        }
}

int orc_lm_guard_bytes(int rows, int cols, int T) {
    /* room for any read the similarity functions can make past the last phase row */
    int W = cols / T, H = rows / T;
    return (H + 17) * W + 64;
}

/* Detector::match, response-map construction part                        line2Dup.cpp:1055-1091 */
int orc_build_lm_pyramid(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask,
                         float weak, int L, const int* T, uint8_t** lm_out) {
    uint8_t* cur = (uint8_t*)malloc((size_t)rows * cols * ch);
    memcpy(cur, img, (size_t)rows * cols * ch);
    uint8_t* cmask = NULL;
    if (mask) {
        cmask = (uint8_t*)malloc((size_t)rows * cols);
        memcpy(cmask, mask, (size_t)rows * cols);
    }
    int r = rows, c = cols;
    for (int l = 0; l < L; ++l) {
        if (l > 0) {
            int nr = r / 2, nc = c / 2;
            uint8_t* nxt = (uint8_t*)malloc((size_t)nr * nc * ch);
            orc_pyr_down(cur, r, c, ch, nxt);
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
        uint8_t* q = (uint8_t*)malloc(n);
        uint8_t* sp = (uint8_t*)malloc(n);
        uint8_t* resp = (uint8_t*)malloc(n * 8);
        if (ch == 1) orc_quantized_orientations(cur, r, c, weak, q, NULL, NULL);
        else orc_gradient_color(cur, r, c, weak, q, NULL, NULL);
        orc_apply_mask(q, cmask, (int)n);
        orc_spread(q, r, c, T[l], sp);
        orc_response_maps(sp, r, c, resp);
        size_t hw = (size_t)(r / T[l]) * (c / T[l]);
        for (int ori = 0; ori < 8; ++ori)
            orc_linearize(resp + ori * n, r, c, T[l], lm_out[l] + (size_t)ori * T[l] * T[l] * hw);
        memset(lm_out[l] + 8 * n, 0, (size_t)orc_lm_guard_bytes(r, c, T[l]));
        free(q); free(sp); free(resp);
    }
    free(cur);
    free(cmask);
    return 0;
}

/* accessLinearMemory                                                        line2Dup.cpp:755-778
 * lm is the flat [ori][T*T][H*W] block: each orientation's T*T rows are contiguous, as in the
 * reference's per-orientation cv::Mat, so reads past a phase row land in the next phase row. */
static inline const uint8_t* access_lm(const uint8_t* lm, int fx, int fy, int label, int T, int W, size_t hw) {
    const uint8_t* grid = lm + (size_t)label * T * T * hw;
    int gi = (fy % T) * T + fx % T;
    return grid + (size_t)gi * hw + (size_t)(fy / T) * W + fx / T;
}

/* similarity (n < 8192, uint16) and similarity_64 (n < 64, uint8)    line2Dup.cpp:780-831, 897-957
 * No add ever saturates or wraps (4*8191 < 32767, 4*63 < 255), so both are the exact integer sum. */
int orc_similarity(const uint8_t* lm, int rows, int cols, int T, const int32_t* feats, int n,
                   int width, int height, uint16_t* dst) {
    if (n >= 8192) return ORC_EINVAL;
    int W = cols / T, H = rows / T;
    size_t hw = (size_t)W * H;
    int wf = (width - 1) / T + 1, hf = (height - 1) / T + 1;
    int span_x = W - wf, span_y = H - hf;
    int P = span_y * W + span_x + 1;
    memset(dst, 0, hw * 2);
    for (int i = 0; i < n; ++i) {
        int fx = feats[3 * i], fy = feats[3 * i + 1], lab = feats[3 * i + 2];
        if (fx < 0 || fx >= cols || fy < 0 || fy >= rows) continue;
        const uint8_t* p = access_lm(lm, fx, fy, lab, T, W, hw);
        for (int j = 0; j < P; ++j) dst[j] = (uint16_t)(dst[j] + p[j]);
    }
    return P;
}

/* similarityLocal (uint16) and similarityLocal_64 (uint8)             line2Dup.cpp:833-895, 959-1021
 * The SSE/AVX2/scalar branch (:877-892) defines the result: dst[r][c] += lm_ptr[r*W + c]. */
int orc_similarity_local(const uint8_t* lm, int rows, int cols, int T, const int32_t* feats, int n,
                         int cx, int cy, uint16_t* dst) {
    if (n >= 8192) return ORC_EINVAL;
    int W = cols / T, H = rows / T;
    size_t hw = (size_t)W * H;
    int offx = (cx / T - 8) * T, offy = (cy / T - 8) * T;
    memset(dst, 0, 256 * 2);
    for (int i = 0; i < n; ++i) {
        int fx = feats[3 * i] + offx, fy = feats[3 * i + 1] + offy, lab = feats[3 * i + 2];
        if (fx < 0 || fy < 0 || fx >= cols || fy >= rows) continue;
        const uint8_t* p = access_lm(lm, fx, fy, lab, T, W, hw);
        for (int r = 0; r < 16; ++r)
            for (int c = 0; c < 16; ++c) dst[r * 16 + c] = (uint16_t)(dst[r * 16 + c] + p[(size_t)r * W + c]);
    }
    return 0;
}

/* score = (raw * 100.f) / (4 * n)                                     line2Dup.cpp:1177, 1244 */
float orc_score(int raw, int n) {
    float num = (float)raw * 100.f;
    return num / (float)(4 * n);
}

void orc_i_omvec_push(omvec* a, omatch m) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 16;
        a->v = (omatch*)realloc(a->v, sizeof(omatch) * (size_t)a->cap);
    }
    a->v[a->n++] = m;
}

/* Canonical total order: similarity desc, template_id asc, class asc, y asc, x asc.  The first two
 * keys are Match::operator< (line2Dup.h:117-124); the rest make the reference's unspecified order
 * of equal elements (std::sort, OpenMP concatenation at line2Dup.cpp:1137-1141) deterministic. */
static int cmp_canon(const void* pa, const void* pb) {
    const omatch* a = (const omatch*)pa;
    const omatch* b = (const omatch*)pb;
    if (a->sim != b->sim) return a->sim > b->sim ? -1 : 1;
    if (a->tid != b->tid) return a->tid < b->tid ? -1 : 1;
    if (a->cls != b->cls) return a->cls < b->cls ? -1 : 1;
    if (a->y != b->y) return a->y < b->y ? -1 : 1;
    if (a->x != b->x) return a->x < b->x ? -1 : 1;
    return 0;
}

/* std::sort + std::unique with Match::operator==       line2Dup.cpp:1114-1116, line2Dup.h:126-129 */
int orc_i_sort_unique(omatch* v, int n) {
    if (n == 0) return 0;
    qsort(v, (size_t)n, sizeof(omatch), cmp_canon);
    int w = 0;
    for (int i = 1; i < n; ++i) {
        omatch* a = &v[w];
        omatch* b = &v[i];
        int eq = a->x == b->x && a->y == b->y && a->sim == b->sim && a->cls == b->cls;
        if (!eq) v[++w] = *b;
    }
    return w + 1;
}

int orc_sort_unique(int count, int32_t* i32, float* sim) {
    omatch* v = (omatch*)malloc(sizeof(omatch) * (size_t)(count > 0 ? count : 1));
    for (int i = 0; i < count; ++i) {
        v[i].x = i32[4 * i]; v[i].y = i32[4 * i + 1]; v[i].cls = i32[4 * i + 2]; v[i].tid = i32[4 * i + 3];
        v[i].sim = sim[i];
    }
    int m = orc_i_sort_unique(v, count);
    for (int i = 0; i < m; ++i) {
        i32[4 * i] = v[i].x; i32[4 * i + 1] = v[i].y; i32[4 * i + 2] = v[i].cls; i32[4 * i + 3] = v[i].tid;
        sim[i] = v[i].sim;
    }
    free(v);
    return m;
}

/* Detector::matchClass for one template pyramid                          line2Dup.cpp:1143-1266 */
static int match_template(uint8_t* const* lm, const int* lrows, const int* lcols, int L, const int* T,
                          float threshold, const int32_t* ti, const int32_t* feats, int cls, int tid,
                          omvec* out) {
    const int lowest = L - 1;
    const int32_t* tl = ti + 6 * lowest;
    int nfeat = tl[4];
    if (nfeat >= 8192) return ORC_EINVAL;
    int rows = lrows[lowest], cols = lcols[lowest], TL = T[lowest];
    int W = cols / TL, H = rows / TL;
    uint16_t* sim = (uint16_t*)malloc((size_t)W * H * 2);
    orc_similarity(lm[lowest], rows, cols, TL, feats + 3 * (size_t)tl[5], nfeat, tl[0], tl[1], sim);
    omvec cand = {0, 0, 0};
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
        if (nf >= 8192) { free(cand.v); return ORC_EINVAL; }
        for (int m = 0; m < cand.n; ++m) {
            omatch* mm = &cand.v[m];
            int x = mm->x * 2 + 1, y = mm->y * 2 + 1;
            x = x > border ? x : border;
            y = y > border ? y : border;
            x = x < max_x ? x : max_x;
            y = y < max_y ? y : max_y;
            orc_similarity_local(lm[l], rl, cl, Tl, feats + 3 * (size_t)tlv[5], nf, x, y, loc);
            float best = 0;
            int br = -1, bc = -1;
            for (int r = 0; r < 16; ++r)
                for (int c = 0; c < 16; ++c) {
                    float s = (loc[r * 16 + c] * 100.f) / (4 * nf);
                    if (s > best) { best = s; br = r; bc = c; }
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
    return 0;
}

int orc_match_lm(uint8_t* const* lm, int rows, int cols, int L, const int* T, float threshold,
                 int num_templates, const int32_t* tinfo, const int32_t* feats,
                 const int32_t* tclass, const int32_t* tid,
                 int max_out, int32_t* out_i32, float* out_sim, int nthreads) {
    int lrows[16], lcols[16];
    if (L < 1 || L > 16) return ORC_EINVAL;
    lrows[0] = rows; lcols[0] = cols;
    for (int l = 1; l < L; ++l) { lrows[l] = lrows[l - 1] / 2; lcols[l] = lcols[l - 1] / 2; }
    omvec* per = (omvec*)calloc((size_t)(num_templates > 0 ? num_templates : 1), sizeof(omvec));
    int err = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(| : err)
#endif
    for (int t = 0; t < num_templates; ++t) {
        int e = match_template(lm, lrows, lcols, L, T, threshold, tinfo + (size_t)6 * L * t, feats, tclass[t],
                               tid[t], &per[t]);
        if (e) err |= 1;
    }
    (void)nthreads;
    if (err) {
        for (int t = 0; t < num_templates; ++t) free(per[t].v);
        free(per);
        return ORC_EINVAL;
    }
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

/* Detector::match                                                           line2Dup.cpp:1049-1121 */
int orc_match(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, int L,
              const int* T, float threshold, int num_templates, const int32_t* tinfo, const int32_t* feats,
              const int32_t* tclass, const int32_t* tid, int max_out, int32_t* out_i32, float* out_sim,
              int nthreads, double* stage_seconds) {
    if (L < 1 || L > 16) return ORC_EINVAL;
    uint8_t* lm[16];
    int r = rows, c = cols;
    for (int l = 0; l < L; ++l) {
        if (l > 0) { r /= 2; c /= 2; }
        if (T[l] <= 0) return ORC_EINVAL;
        lm[l] = (uint8_t*)malloc((size_t)r * c * 8 + (size_t)orc_lm_guard_bytes(r, c, T[l]));
    }
    double t0 = now_s();
    int rc = orc_build_lm_pyramid(img, rows, cols, ch, mask, weak, L, T, lm);
    double t1 = now_s();
    if (rc == 0)
        rc = orc_match_lm(lm, rows, cols, L, T, threshold, num_templates, tinfo, feats, tclass, tid, max_out,
                          out_i32, out_sim, nthreads);
    double t2 = now_s();
    if (stage_seconds) { stage_seconds[0] = t1 - t0; stage_seconds[1] = t2 - t1; }
    for (int l = 0; l < L; ++l) free(lm[l]);
    return rc;
}

/* ------------------------------------------------------------------------------------------
 * cv_dnn::NMSBoxes(boxes, scores, score_threshold, nms_threshold, indices, eta, top_k)
 *                                                                          test.cpp:31-121
 * ------------------------------------------------------------------------------------------ */
typedef struct { float s; int i; } score_idx;

static float rect_overlap(const int32_t* a, const int32_t* b) {
    int aa = a[2] * a[3], ab = b[2] * b[3];
    double dist;
    if ((aa + ab) <= 0) {
        dist = 0.0;
    } else {
        int x1 = a[0] > b[0] ? a[0] : b[0];
        int y1 = a[1] > b[1] ? a[1] : b[1];
        int x2 = (a[0] + a[2]) < (b[0] + b[2]) ? (a[0] + a[2]) : (b[0] + b[2]);
        int y2 = (a[1] + a[3]) < (b[1] + b[3]) ? (a[1] + a[3]) : (b[1] + b[3]);
        int w = x2 - x1, h = y2 - y1;
        int inter = (w <= 0 || h <= 0) ? 0 : w * h; /* Rect & Rect is empty when w or h <= 0 */
        double aab = (double)inter;
        dist = 1.0 - aab / ((double)(aa + ab) - aab);
    }
    return 1.f - (float)dist;
}

int orc_nms_boxes(int n, const int32_t* boxes, const float* scores, float score_threshold, float nms_threshold,
                  float eta, int top_k, int32_t* indices) {
    score_idx* v = (score_idx*)malloc(sizeof(score_idx) * (size_t)(n > 0 ? n : 1));
    int m = 0;
    for (int i = 0; i < n; ++i)
        if (scores[i] > score_threshold) { v[m].s = scores[i]; v[m].i = i; ++m; }
    /* stable sort by score descending (insertion-merge: a simple stable merge sort) */
    score_idx* tmp = (score_idx*)malloc(sizeof(score_idx) * (size_t)(m > 0 ? m : 1));
    for (int width = 1; width < m; width *= 2) {
        for (int lo = 0; lo < m; lo += 2 * width) {
            int mid = lo + width < m ? lo + width : m, hi = lo + 2 * width < m ? lo + 2 * width : m;
            int a = lo, b = mid, k = lo;
            while (a < mid && b < hi) tmp[k++] = (v[b].s > v[a].s) ? v[b++] : v[a++];
            while (a < mid) tmp[k++] = v[a++];
            while (b < hi) tmp[k++] = v[b++];
        }
        memcpy(v, tmp, sizeof(score_idx) * (size_t)m);
    }
    free(tmp);
    if (top_k > 0 && top_k < m) m = top_k;
    float adaptive = nms_threshold;
    int kept = 0;
    for (int i = 0; i < m; ++i) {
        int idx = v[i].i, keep = 1;
        for (int k = 0; k < kept && keep; ++k) keep = rect_overlap(boxes + 4 * idx, boxes + 4 * indices[k]) <= adaptive;
        if (keep) indices[kept++] = idx;
        if (keep && eta < 1 && adaptive > 0.5) adaptive *= eta;
    }
    free(v);
    return kept;
}

/* ------------------------------------------------------------------------------------------
 * Training: extractTemplate, selectScatteredFeatures, cropTemplates, addTemplate, addTemplate_rotate
 * ------------------------------------------------------------------------------------------ */
static int label_of(int q) {
    switch (q) {  /* getLabel, line2Dup.cpp:32-56 */
        case 1: return 0; case 2: return 1; case 4: return 2; case 8: return 3;
        case 16: return 4; case 32: return 5; case 64: return 6; case 128: return 7;
        default: return -1;
    }
}

/* getLabel of the 16-orientation variant: the single set bit of a 16-bit value
 *                                                     test/ori_16bit_experiment/line2Dup_16bit_ori.cpp:32-72 */
static int label16_of(int q) {
    for (int b = 0; b < 16; ++b)
        if (q == (1 << b)) return b;
    return -1;
}

/* selectScatteredFeatures                                                   line2Dup.cpp:151-200 */
static int select_scattered(const ocand* cands, int nc, int num_features, float distance, ocand* feats) {
    int nf = 0;
    float dsq = distance * distance;
    int i = 0, first = 1;
    for (;;) {
        ocand c = cands[i];
        int keep = 1;
        for (int j = 0; j < nf && keep; ++j) {
            int ddx = c.x - feats[j].x, ddy = c.y - feats[j].y;
            keep = (float)(ddx * ddx + ddy * ddy) >= dsq;
        }
        if (keep) feats[nf++] = c;
        if (++i == nc) {
            int num_ok = nf >= num_features;
            if (first) {
                if (num_ok) {
                    nf = 0;
                    i = 0;
                    distance += 1.0f;
                    dsq = distance * distance;
                    continue;
                } else {
                    first = 0;
                }
            }
            i = 0;
            distance -= 1.0f;
            dsq = distance * distance;
            if (num_ok || distance < 3) break;
        }
    }
    return nf;
}

/* The candidate stage of ColorGradientPyramid::extractTemplate: eroded mask, 5x5 magnitude NMS in
 * raster order, strong threshold, then std::stable_sort by score descending.   line2Dup.cpp:425-497
 * The 16-orientation variant (test/ori_16bit_experiment/line2Dup_16bit_ori.cpp:421-478) runs the same
 * scan with 16-bit angles; it lacks the outer "if(!is_max) break;" (:459), which changes nothing: once
 * is_max is false the score is 0 and stays 0.
 * Returns the candidate count (all of them). */
int orc_i_extract_candidates(const float* mag, const uint8_t* angle, const uint16_t* angle16, const float* angle_ori,
                             const uint8_t* mask, int rows, int cols, float strong, ocand** out) {
    uint8_t* lmask = NULL;
    if (mask) { /* erode(mask, local_mask, 3x3, BORDER_REPLICATE) */
        lmask = (uint8_t*)malloc((size_t)rows * cols);
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) {
                uint8_t m = 255;
                for (int dr = -1; dr <= 1; ++dr)
                    for (int dc = -1; dc <= 1; ++dc) {
                        uint8_t v = mask[(size_t)clampi(r + dr, 0, rows - 1) * cols + clampi(c + dc, 0, cols - 1)];
                        if (v < m) m = v;
                    }
                lmask[(size_t)r * cols + c] = m;
            }
    }
    float thr_sq = strong * strong;
    const int K = 5;
    uint8_t* valid = (uint8_t*)malloc((size_t)rows * cols);
    memset(valid, 255, (size_t)rows * cols);
    int nc = 0, capc = 1024;
    ocand* cands = (ocand*)malloc(sizeof(ocand) * (size_t)capc);
    for (int r = K / 2; r < rows - K / 2; ++r)
        for (int c = K / 2; c < cols - K / 2; ++c) {
            if (lmask && !lmask[(size_t)r * cols + c]) continue;
            float score = 0;
            if (valid[(size_t)r * cols + c] > 0) {
                score = mag[(size_t)r * cols + c];
                int is_max = 1;
                for (int ro = -K / 2; ro <= K / 2 && is_max; ++ro)
                    for (int co = -K / 2; co <= K / 2; ++co) {
                        if (ro == 0 && co == 0) continue;
                        if (score < mag[(size_t)(r + ro) * cols + c + co]) {
                            score = 0;
                            is_max = 0;
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
            int a = angle16 ? (int)angle16[(size_t)r * cols + c] : (int)angle[(size_t)r * cols + c];
            if (score > thr_sq && a > 0) {
                if (nc == capc) { capc *= 2; cands = (ocand*)realloc(cands, sizeof(ocand) * (size_t)capc); }
                ocand cd = {c, r, angle16 ? label16_of(a) : label_of(a), angle_ori ? angle_ori[(size_t)r * cols + c] : 0.f,
                            score};
                cands[nc++] = cd;
            }
        }
    free(valid);
    free(lmask);
    /* std::stable_sort by score descending (Candidate::operator<, line2Dup.h:62-65): bottom-up merge */
    ocand* tmp = (ocand*)malloc(sizeof(ocand) * (size_t)(nc > 0 ? nc : 1));
    for (int width = 1; width < nc; width *= 2) {
        for (int lo = 0; lo < nc; lo += 2 * width) {
            int mid = lo + width < nc ? lo + width : nc, hi = lo + 2 * width < nc ? lo + 2 * width : nc;
            int a = lo, b = mid, k = lo;
            while (a < mid && b < hi) tmp[k++] = (cands[b].score > cands[a].score) ? cands[b++] : cands[a++];
            while (a < mid) tmp[k++] = cands[a++];
            while (b < hi) tmp[k++] = cands[b++];
        }
        memcpy(cands, tmp, sizeof(ocand) * (size_t)nc);
    }
    free(tmp);
    *out = cands;
    return nc;
}

int orc_extract_candidates(const float* mag, const uint8_t* angle, const float* angle_ori, const uint8_t* mask,
                           int rows, int cols, float strong, int max_out, int32_t* out_xyl, float* out_theta,
                           float* out_score) {
    ocand* cands = NULL;
    int nc = orc_i_extract_candidates(mag, angle, NULL, angle_ori, mask, rows, cols, strong, &cands);
    for (int i = 0; i < nc && i < max_out; ++i) {
        out_xyl[3 * i] = cands[i].x; out_xyl[3 * i + 1] = cands[i].y; out_xyl[3 * i + 2] = cands[i].label;
        if (out_theta) out_theta[i] = cands[i].theta;
        if (out_score) out_score[i] = cands[i].score;
    }
    free(cands);
    return nc;
}

/* ColorGradientPyramid::extractTemplate (features only; width/height set by cropTemplates)
 *                                                                          line2Dup.cpp:425-512
 * Returns the feature count, or -1 for "too few features". */
int orc_extract_template(const float* mag, const uint8_t* angle, const float* angle_ori, const uint8_t* mask,
                         int rows, int cols, float strong, int num_features, int max_out, int32_t* out_xyl,
                         float* out_theta) {
    if (num_features <= 0) return -1; /* the reference divides by it (line2Dup.cpp:498) */
    ocand* cands = NULL;
    int nc = orc_i_extract_candidates(mag, angle, NULL, angle_ori, mask, rows, cols, strong, &cands);
    if (nc < num_features && nc <= 4) { free(cands); return -1; }
    ocand* sel = (ocand*)malloc(sizeof(ocand) * (size_t)(nc > 0 ? nc : 1));
    float distance = (float)(nc / num_features + 1);
    int nf = select_scattered(cands, nc, num_features, distance, sel);
    int out = nf < max_out ? nf : max_out;
    for (int i = 0; i < out; ++i) {
        out_xyl[3 * i] = sel[i].x; out_xyl[3 * i + 1] = sel[i].y; out_xyl[3 * i + 2] = sel[i].label;
        if (out_theta) out_theta[i] = sel[i].theta;
    }
    free(sel);
    free(cands);
    return nf;
}

/* cropTemplates                                                             line2Dup.cpp:103-149
 * tinfo5 per level: width, height, tl_x, tl_y, pyramid_level (the first four are outputs). */
void orc_crop_templates(int levels, int32_t* ti, const int* nfeat, int32_t** feats) {
    int min_x = 2147483647, min_y = 2147483647, max_x = -2147483647 - 1, max_y = -2147483647 - 1;
    for (int l = 0; l < levels; ++l) {
        int pl = ti[5 * l + 4];
        for (int j = 0; j < nfeat[l]; ++j) {
            int x = feats[l][3 * j] << pl, y = feats[l][3 * j + 1] << pl;
            if (x < min_x) min_x = x;
            if (y < min_y) min_y = y;
            if (x > max_x) max_x = x;
            if (y > max_y) max_y = y;
        }
    }
    if (min_x % 2 == 1) --min_x;
    if (min_y % 2 == 1) --min_y;
    for (int l = 0; l < levels; ++l) {
        int pl = ti[5 * l + 4];
        ti[5 * l + 0] = (max_x - min_x) >> pl;
        ti[5 * l + 1] = (max_y - min_y) >> pl;
        ti[5 * l + 2] = min_x >> pl;
        ti[5 * l + 3] = min_y >> pl;
        for (int j = 0; j < nfeat[l]; ++j) {
            feats[l][3 * j] -= ti[5 * l + 2];
            feats[l][3 * j + 1] -= ti[5 * l + 3];
        }
    }
}

/* Detector::addTemplate                                                     line2Dup.cpp:1270-1304
 * tinfo6 per level: width, height, tl_x, tl_y, num_features, feature_offset.
 * Returns the total feature count written, or -1 when a level has too few features. */
int orc_add_template(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, float strong,
                     int num_features, int L, int max_feats, int32_t* tinfo6, int32_t* feats_xyl,
                     float* feats_theta) {
    uint8_t* cur = (uint8_t*)malloc((size_t)rows * cols * ch);
    memcpy(cur, img, (size_t)rows * cols * ch);
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
            uint8_t* nx = (uint8_t*)malloc((size_t)nr * nc * ch);
            orc_pyr_down(cur, r, c, ch, nx);
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
        uint8_t* q = (uint8_t*)malloc(n);
        float* mag = (float*)malloc(n * 4);
        float* ag = (float*)malloc(n * 4);
        if (ch == 1) orc_quantized_orientations(cur, r, c, weak, q, mag, ag);
        else orc_gradient_color(cur, r, c, weak, q, mag, ag);
        int got = orc_extract_template(mag, q, ag, cm, r, c, strong, nfl, max_feats - total, feats_xyl + 3 * (size_t)total,
                                       feats_theta ? feats_theta + total : NULL);
        free(q); free(mag); free(ag);
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

/* Detector::addTemplate_rotate with rotate2d / rotatePoint               line2Dup.cpp:1306-1362
 * The destination keeps the source's per-level feature counts and offsets. */
void orc_rotate_template(int L, const int32_t* src6, const int32_t* fsrc, const float* thsrc, float theta, float cx,
                         float cy, int32_t* dst6, int32_t* fdst, float* thdst) {
    int32_t* ti5 = (int32_t*)malloc(sizeof(int32_t) * 5 * (size_t)L);
    int* nfeat = (int*)calloc((size_t)L, sizeof(int));
    int32_t** fptr = (int32_t**)malloc(sizeof(int32_t*) * (size_t)L);
    double ang = (double)(-theta / 180) * 3.1415926535897932384626433832795;
    double ca = cos(ang), sa = sin(ang);
    for (int l = 0; l < L; ++l) {
        if (l > 0) { cx /= 2; cy /= 2; }
        int n = src6[6 * l + 4], off = src6[6 * l + 5];
        for (int i = 0; i < n; ++i) {
            const int32_t* f = fsrc + 3 * (size_t)(off + i);
            float px = (float)(f[0] + src6[6 * l + 2]);
            float py = (float)(f[1] + src6[6 * l + 3]);
            float ix = px - cx, iy = py - cy;
            float ox = (float)(ca * ix - sa * iy);
            float oy = (float)(sa * ix + ca * iy);
            ox = ox + cx;
            oy = oy + cy;
            int32_t* g = fdst + 3 * (size_t)(off + i);
            g[0] = (int)(ox + 0.5f);
            g[1] = (int)(oy + 0.5f);
            float th = thsrc[off + i] - theta;
            while (th > 360) th -= 360;
            while (th < 0) th += 360;
            thdst[off + i] = th;
            g[2] = ((int)(th * 16 / 360 + 0.5f)) & 7;
        }
        fptr[l] = fdst + 3 * (size_t)off;
        nfeat[l] = n;
        ti5[5 * l + 4] = l;
        dst6[6 * l + 4] = n;
        dst6[6 * l + 5] = off;
    }
    orc_crop_templates(L, ti5, nfeat, fptr);
    for (int l = 0; l < L; ++l)
        for (int k = 0; k < 4; ++k) dst6[6 * l + k] = ti5[5 * l + k];
    free(ti5); free(nfeat); free(fptr);
}

/* ---------------------------------------------------------------------------------------------
 * shapeInfo_producer                                                          line2Dup.h:209-332
 * ------------------------------------------------------------------------------------------- */

/* cv::getRotationMatrix2D (OpenCV 3.4): angle in degrees, counter-clockwise, about `center`. */
void orc_rotation_matrix2d(float cx, float cy, double angle_deg, double scale, double M[6]) {
    double a = angle_deg * (3.1415926535897932384626433832795 / 180);
    double alpha = cos(a) * scale, beta = sin(a) * scale;
    M[0] = alpha;
    M[1] = beta;
    M[2] = (1 - alpha) * (double)cx - beta * (double)cy;
    M[3] = -beta;
    M[4] = alpha;
    M[5] = beta * (double)cx + (1 - alpha) * (double)cy;
}

/* The inversion cv::warpAffine applies to a forward matrix (no WARP_INVERSE_MAP). */
void orc_invert_affine(const double M[6], double iM[6]) {
    double D = M[0] * M[4] - M[1] * M[3];
    D = D != 0 ? 1. / D : 0;
    double A11 = M[4] * D, A22 = M[0] * D;
    iM[0] = A11;
    iM[1] = M[1] * -D;
    iM[3] = M[3] * -D;
    iM[4] = A22;
    iM[2] = -iM[0] * M[2] - iM[1] * M[5];
    iM[5] = -iM[3] * M[2] - iM[4] * M[5];
}

static inline int round_sat_int(double v) { /* saturate_cast<int>(double): round half to even */
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (-2147483647 - 1);
    return (int)lrint(v);
}
static inline int sat_short(int v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : v); }

/* cv::warpAffine(src, dst, M, src.size()) with INTER_LINEAR, BORDER_CONSTANT 0.
 * Inverse map in 10-bit fixed point per column step and per row, 5 sub-pixel bits; bilinear weights
 * (32-fy)(32-fx)*32 ... summing to 2^15; taps outside the source read 0; (sum + 2^14) >> 15. */
void orc_warp_affine(const uint8_t* src, int rows, int cols, int ch, const double M0[6], uint8_t* dst) {
    double M[6];
    orc_invert_affine(M0, M);
    const int AB_BITS = 10, AB_SCALE = 1 << AB_BITS, INTER_BITS = 5, TAB = 1 << INTER_BITS;
    const int round_delta = AB_SCALE / TAB / 2;
    int* adelta = (int*)malloc(sizeof(int) * (size_t)(cols > 0 ? cols : 1));
    int* bdelta = (int*)malloc(sizeof(int) * (size_t)(cols > 0 ? cols : 1));
    for (int x = 0; x < cols; ++x) {
        adelta[x] = round_sat_int(M[0] * x * AB_SCALE);
        bdelta[x] = round_sat_int(M[3] * x * AB_SCALE);
    }
    for (int y = 0; y < rows; ++y) {
        int X0 = round_sat_int((M[1] * y + M[2]) * AB_SCALE) + round_delta;
        int Y0 = round_sat_int((M[4] * y + M[5]) * AB_SCALE) + round_delta;
        for (int x = 0; x < cols; ++x) {
            int X = (X0 + adelta[x]) >> (AB_BITS - INTER_BITS);
            int Y = (Y0 + bdelta[x]) >> (AB_BITS - INTER_BITS);
            int sx = sat_short(X >> INTER_BITS), sy = sat_short(Y >> INTER_BITS);
            int fx = X & (TAB - 1), fy = Y & (TAB - 1);
            int w[4] = {(TAB - fy) * (TAB - fx) * 32, (TAB - fy) * fx * 32, fy * (TAB - fx) * 32, fy * fx * 32};
            for (int k = 0; k < ch; ++k) {
                int acc = 0;
                for (int t = 0; t < 4; ++t) {
                    int px = sx + (t & 1), py = sy + (t >> 1);
                    int v = (px >= 0 && px < cols && py >= 0 && py < rows) ? src[((size_t)py * cols + px) * ch + k] : 0;
                    acc += v * w[t];
                }
                int o = (acc + (1 << 14)) >> 15;
                dst[((size_t)y * cols + x) * ch + k] = (uint8_t)(o > 255 ? 255 : (o < 0 ? 0 : o));
            }
        }
    }
    free(adelta);
    free(bdelta);
}

/* shapeInfo_producer::transform                                                line2Dup.h:245-253 */
void orc_shape_transform(const uint8_t* src, int rows, int cols, int ch, float angle, float scale, uint8_t* dst) {
    double M[6];
    orc_rotation_matrix2d(cols / 2.0f, rows / 2.0f, (double)angle, (double)scale, M);
    orc_warp_affine(src, rows, cols, ch, M, dst);
}

/* shapeInfo_producer::produce_infos                                            line2Dup.h:281-320 */
int orc_produce_infos(int na, const float* ar, int ns, const float* sr, float astep, float sstep, float eps, int cap,
                      float* angles, float* scales) {
    float a0 = na > 0 ? ar[0] : 0.f, s0 = ns > 0 ? sr[0] : 1.f;
    int n = 0;
#define PUSH(A, S) do { if (n < cap) { angles[n] = (A); scales[n] = (S); } ++n; } while (0)
    if (na <= 1 && ns <= 1) {
        PUSH(a0, s0);
    } else if (na <= 1 && ns == 2) {
        for (float s = sr[0]; s <= sr[1] + eps; s += sstep) PUSH(a0, s);
    } else if (na == 2 && ns <= 1) {
        for (float a = ar[0]; a <= ar[1] + eps; a += astep) PUSH(a, s0);
    } else {
        for (float s = sr[0]; s <= sr[1] + eps; s += sstep)
            for (float a = ar[0]; a <= ar[1] + eps; a += astep) PUSH(a, s);
    }
#undef PUSH
    return n;
}
