/*
 * Vectorised, OpenMP-parallel CPU restatement of line2Dup's matching path -- TEST INFRASTRUCTURE ONLY:
 * the timed CPU baseline of bench.py.  It computes exactly what the scalar oracle (sbm_oracle.c) computes,
 * organised for SIMD units: every loop runs over contiguous rows so that the compiler's vectoriser takes it
 * (AVX2 or AVX-512, one build of this file per instruction set, chosen at run time by sbm_cpu_fast_dispatch.c),
 * and the templates are matched in parallel as the reference does (OpenMP over templates,
 * line2Dup.cpp:1137-1141).  Preprocessing stays on one thread, as in the reference (line2Dup.cpp:1061-1091).
 *
 * Each function cites the reference lines it restates; tests/test_cpu_fast.py holds every list it returns to
 * the scalar oracle's.  Nothing here is part of the product library.
 *
 * FAST_SUFFIX names the build (_avx2, _avx512, _generic); the file is compiled once per suffix.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc_internal.h"
#include "sbm_oracle.h"

#ifndef FAST_SUFFIX
#define FAST_SUFFIX _generic
#endif
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, FAST_SUFFIX)

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------------------------------
 * GaussianBlur 7x7 sigma 0 (kernel [4 14 28 36 28 14 4] per axis), BORDER_REPLICATE, then
 * (sum + 8192) >> 14                                                          line2Dup.cpp:291-293
 * Horizontal pass into uint16 (at most 128 * 255), vertical pass in int32, both over padded rows.
 * ------------------------------------------------------------------------------------------------ */
static void blur7(const uint8_t* src, int rows, int cols, uint16_t* tmp, uint8_t* pad, uint8_t* dst) {
    for (int r = 0; r < rows; ++r) {
        const uint8_t* s = src + (size_t)r * cols;
        for (int i = 0; i < 3; ++i) {
            pad[i] = s[0];
            pad[cols + 3 + i] = s[cols - 1];
        }
        memcpy(pad + 3, s, (size_t)cols);
        uint16_t* t = tmp + (size_t)r * cols;
        for (int c = 0; c < cols; ++c)
            t[c] = (uint16_t)(4 * (pad[c] + pad[c + 6]) + 14 * (pad[c + 1] + pad[c + 5]) +
                              28 * (pad[c + 2] + pad[c + 4]) + 36 * pad[c + 3]);
    }
    for (int r = 0; r < rows; ++r) {
        const uint16_t* t0 = tmp + (size_t)clampi(r - 3, 0, rows - 1) * cols;
        const uint16_t* t1 = tmp + (size_t)clampi(r - 2, 0, rows - 1) * cols;
        const uint16_t* t2 = tmp + (size_t)clampi(r - 1, 0, rows - 1) * cols;
        const uint16_t* t3 = tmp + (size_t)r * cols;
        const uint16_t* t4 = tmp + (size_t)clampi(r + 1, 0, rows - 1) * cols;
        const uint16_t* t5 = tmp + (size_t)clampi(r + 2, 0, rows - 1) * cols;
        const uint16_t* t6 = tmp + (size_t)clampi(r + 3, 0, rows - 1) * cols;
        uint8_t* d = dst + (size_t)r * cols;
        for (int c = 0; c < cols; ++c) {
            const int32_t s = 4 * ((int32_t)t0[c] + t6[c]) + 14 * ((int32_t)t1[c] + t5[c]) +
                              28 * ((int32_t)t2[c] + t4[c]) + 36 * (int32_t)t3[c];
            d[c] = (uint8_t)((s + 8192) >> 14);
        }
    }
}

/* ------------------------------------------------------------------------------------------------
 * Magnitude dx^2 + dy^2 and the fastAtan2 phase in degrees of n gradients, each float operation rounded as
 * OpenCV 3.x does (no contraction), then the 16-bucket bin (round half to even, saturated)
 *                                                   line2Dup.cpp:213, 299-300 (hysteresisGradient input)
 * Branch-free, so that it vectorises; the divisions and selects give the scalar oracle's bits.
 * ------------------------------------------------------------------------------------------------ */
static void phase_bins(const int32_t* gxs, const int32_t* gys, int n, float* mag, uint8_t* bin) {
    const float s = (float)(180.0 / 3.1415926535897932384626433832795);
    const float p1 = 0.9997878412794807f * s;
    const float p3 = -0.3258083974640975f * s;
    const float p5 = 0.1555786518463281f * s;
    const float p7 = -0.04432655554792128f * s;
    const float eps = (float)DBL_EPSILON;
    const float scale = (float)(16.0 / 360.0);
    for (int c = 0; c < n; ++c) {
        const float fx = (float)gxs[c], fy = (float)gys[c];
        const float fx2 = fx * fx;
        const float fy2 = fy * fy;
        if (mag) mag[c] = fx2 + fy2;
        const float ax = fabsf(fx), ay = fabsf(fy);
        const int xm = ax >= ay;
        const float num = xm ? ay : ax;
        const float den = xm ? ax : ay;
        const float den_e = den + eps;
        const float cc = num / den_e;
        const float c2 = cc * cc;
        float a = p7 * c2;
        a = a + p5;
        a = a * c2;
        a = a + p3;
        a = a * c2;
        a = a + p1;
        a = a * cc;
        const float a90 = 90.f - a;
        a = xm ? a : a90;
        const float a180 = 180.f - a;
        a = fx < 0 ? a180 : a;
        const float a360 = 360.f - a;
        a = fy < 0 ? a360 : a;
        const float v = a * scale;
        int q = (int)rintf(v);
        q = q < 0 ? 0 : (q > 255 ? 255 : q);
        bin[c] = (uint8_t)q;
    }
}

/* Sobel dx, dy (3x3, BORDER_REPLICATE) of the blurred image       line2Dup.cpp:297-298
 * One output row from three replicate-padded blurred rows (index c + 1 is column c). */
static void sobel_row(const uint8_t* up, const uint8_t* mid, const uint8_t* dn, int cols, int32_t* gx, int32_t* gy) {
    for (int c = 0; c < cols; ++c) {
        gx[c] = ((int)up[c + 2] - up[c]) + 2 * ((int)mid[c + 2] - mid[c]) + ((int)dn[c + 2] - dn[c]);
        gy[c] = ((int)dn[c] - up[c]) + 2 * ((int)dn[c + 1] - up[c + 1]) + ((int)dn[c + 2] - up[c + 2]);
    }
}

/* hysteresisGradient: border ring 0, interior "& 7", 3x3 vote with the first maximum, >= 5 votes, over
 * pixels with magnitude > weak^2                                                line2Dup.cpp:206-284
 * The nine votes of a pixel are counted in 4-bit fields of one word (at most 9 per bin). */
static void vote_rows(const uint8_t* qu, const float* mag, int rows, int cols, float thr, uint32_t* hrow,
                      uint8_t* out) {
    memset(out, 0, (size_t)rows * cols);
    for (int r = 1; r < rows - 1; ++r) {
        uint32_t* h[3];
        for (int k = 0; k < 3; ++k) {
            h[k] = hrow + (size_t)k * cols;
            const uint8_t* q = qu + (size_t)(r - 1 + k) * cols;
            for (int c = 1; c < cols - 1; ++c)
                h[k][c] = (1u << (4 * q[c - 1])) + (1u << (4 * q[c])) + (1u << (4 * q[c + 1]));
        }
        const float* m = mag + (size_t)r * cols;
        uint8_t* o = out + (size_t)r * cols;
        for (int c = 1; c < cols - 1; ++c) {
            const uint32_t hist = h[0][c] + h[1][c] + h[2][c];
            int maxv = 0, idx = 0;
            for (int b = 0; b < 8; ++b) {
                const int cnt = (int)((hist >> (4 * b)) & 15u);
                const int better = cnt > maxv;
                maxv = better ? cnt : maxv;
                idx = better ? b : idx;
            }
            const int keep = (m[c] > thr) & (maxv >= 5);
            o[c] = (uint8_t)(keep ? (1u << idx) : 0u);
        }
    }
}

/* quantizedOrientations, single-channel branch                             line2Dup.cpp:286-302 */
typedef struct {
    uint16_t* tmp;
    uint8_t* pad;
    uint8_t* bl;
    uint8_t* bpad;   /* three padded blurred rows */
    int32_t* gx;     /* one row of gradients */
    int32_t* gy;
    float* mag;
    uint8_t* qu;
    uint32_t* hrow;
} prep_buf;

static void quantize_gray(const uint8_t* src, int rows, int cols, float weak, prep_buf* b, uint8_t* q) {
    blur7(src, rows, cols, b->tmp, b->pad, b->bl);
    uint8_t* pr[3] = {b->bpad, b->bpad + (cols + 2), b->bpad + 2 * (size_t)(cols + 2)};
    for (int r = 0; r < rows; ++r) {
        for (int k = 0; k < 3; ++k) {
            const uint8_t* s = b->bl + (size_t)clampi(r - 1 + k, 0, rows - 1) * cols;
            pr[k][0] = s[0];
            memcpy(pr[k] + 1, s, (size_t)cols);
            pr[k][cols + 1] = s[cols - 1];
        }
        sobel_row(pr[0], pr[1], pr[2], cols, b->gx, b->gy);
        phase_bins(b->gx, b->gy, cols, b->mag + (size_t)r * cols, b->qu + (size_t)r * cols);
    }
    /* border ring 0, interior bins "& 7" (line2Dup.cpp:217-234) */
    for (int r = 0; r < rows; ++r) {
        uint8_t* qr = b->qu + (size_t)r * cols;
        if (r == 0 || r == rows - 1) {
            memset(qr, 0, (size_t)cols);
            continue;
        }
        for (int c = 0; c < cols; ++c) qr[c] &= 7;
        qr[0] = 0;
        qr[cols - 1] = 0;
    }
    vote_rows(b->qu, b->mag, rows, cols, weak * weak, b->hrow, q);
}

/* cv::pyrDown 5x5 [1 4 6 4 1]^2 / 256, round half up, BORDER_REFLECT_101     line2Dup.cpp:397-417
 * Separable: the horizontal taps at even columns into uint16 rows, then the vertical taps. */
static inline int reflect101(int p, int n) {
    if (n == 1) return 0;
    while (p < 0 || p >= n) {
        if (p < 0) p = -p;
        if (p >= n) p = 2 * n - 2 - p;
    }
    return p;
}

static void pyr_down(const uint8_t* src, int rows, int cols, int ch, uint16_t* hbuf, uint8_t* dst) {
    const int dr = rows / 2, dc = cols / 2;
    /* hbuf: rows x (dc * ch) */
    for (int r = 0; r < rows; ++r) {
        const uint8_t* s = src + (size_t)r * cols * ch;
        uint16_t* h = hbuf + (size_t)r * dc * ch;
        for (int x = 0; x < dc; ++x) {
            const int c0 = reflect101(2 * x - 2, cols), c1 = reflect101(2 * x - 1, cols), c2 = 2 * x,
                      c3 = reflect101(2 * x + 1, cols), c4 = reflect101(2 * x + 2, cols);
            for (int k = 0; k < ch; ++k)
                h[x * ch + k] = (uint16_t)(s[c0 * ch + k] + 4 * s[c1 * ch + k] + 6 * s[c2 * ch + k] +
                                           4 * s[c3 * ch + k] + s[c4 * ch + k]);
        }
    }
    const int n = dc * ch;
    for (int y = 0; y < dr; ++y) {
        const uint16_t* h0 = hbuf + (size_t)reflect101(2 * y - 2, rows) * n;
        const uint16_t* h1 = hbuf + (size_t)reflect101(2 * y - 1, rows) * n;
        const uint16_t* h2 = hbuf + (size_t)(2 * y) * n;
        const uint16_t* h3 = hbuf + (size_t)reflect101(2 * y + 1, rows) * n;
        const uint16_t* h4 = hbuf + (size_t)reflect101(2 * y + 2, rows) * n;
        uint8_t* d = dst + (size_t)y * n;
        for (int i = 0; i < n; ++i) {
            const int32_t v = (int32_t)h0[i] + 4 * (int32_t)h1[i] + 6 * (int32_t)h2[i] + 4 * (int32_t)h3[i] + h4[i];
            d[i] = (uint8_t)((v + 128) >> 8);
        }
    }
}

/* spread: OR over the forward TxT window clipped at the bottom and right edges  line2Dup.cpp:556-603
 * then the phase split of linearize (line2Dup.cpp:722-750): sp[phase][ly * W + lx] = S[ly*T + r0][lx*T + c0],
 * and the eight response planes from the closed form of SIMILARITY_LUT (line2Dup.cpp:605-720): 4 when the
 * orientation's bit is set, else 3 when a circular neighbour's is, else 0 -- per byte, contiguous.
 * lm: [ori 8][phase T*T][H*W] followed by the zeroed guard. */
static void spread_linearize(const uint8_t* q, int rows, int cols, int T, uint8_t* hor, uint8_t* sp, uint8_t* lm) {
    const int W = cols / T, H = rows / T;
    const size_t hw = (size_t)W * H;
    /* horizontal OR of T columns (zeros past the right edge) */
    for (int r = 0; r < rows; ++r) {
        const uint8_t* s = q + (size_t)r * cols;
        uint8_t* h = hor + (size_t)r * cols;
        memcpy(h, s, (size_t)cols);
        for (int d = 1; d < T; ++d)
            for (int c = 0; c + d < cols; ++c) h[c] |= s[c + d];
    }
    /* vertical OR of T rows (rows past the bottom ignored), straight into the phase planes */
    uint8_t* srow = (uint8_t*)malloc((size_t)cols);
    for (int r = 0; r < rows; ++r) {
        memcpy(srow, hor + (size_t)r * cols, (size_t)cols);
        for (int d = 1; d < T && r + d < rows; ++d) {
            const uint8_t* h = hor + (size_t)(r + d) * cols;
            for (int c = 0; c < cols; ++c) srow[c] |= h[c];
        }
        const int r0 = r % T, ly = r / T;
        for (int c0 = 0; c0 < T; ++c0) {
            uint8_t* dst = sp + (size_t)(r0 * T + c0) * hw + (size_t)ly * W;
            for (int lx = 0; lx < W; ++lx) dst[lx] = srow[lx * T + c0];
        }
    }
    free(srow);
    const size_t cells = (size_t)T * T * hw;
    for (int ori = 0; ori < 8; ++ori) {
        uint8_t* out = lm + (size_t)ori * cells;
        const int o1 = (ori + 1) & 7, o7 = (ori + 7) & 7;
        for (size_t i = 0; i < cells; ++i) {
            const uint8_t s = sp[i];
            const uint8_t b0 = (uint8_t)((s >> ori) & 1u);
            const uint8_t nb = (uint8_t)(((s >> o1) | (s >> o7)) & 1u);
            out[i] = (uint8_t)(b0 * 4u + (nb & (uint8_t)(b0 ^ 1u)) * 3u);
        }
    }
}

/* Detector::match, response-map construction (line2Dup.cpp:1055-1091): every level's linear memories.
 * Colour images take the scalar oracle's gradient (the baseline workloads are gray). */
static int build_pyramid(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, int L,
                         const int* T, uint8_t** lm) {
    const size_t n0 = (size_t)rows * cols;
    prep_buf b;
    b.tmp = (uint16_t*)malloc(n0 * 2 * (size_t)(ch > 1 ? ch : 1));
    b.pad = (uint8_t*)malloc((size_t)cols + 8);
    b.bl = (uint8_t*)malloc(n0);
    b.bpad = (uint8_t*)malloc(3 * ((size_t)cols + 2));
    b.gx = (int32_t*)malloc((size_t)cols * 4);
    b.gy = (int32_t*)malloc((size_t)cols * 4);
    b.mag = (float*)malloc(n0 * 4);
    b.qu = (uint8_t*)malloc(n0);
    b.hrow = (uint32_t*)malloc(3 * (size_t)cols * 4);
    uint8_t* q = (uint8_t*)malloc(n0);
    uint8_t* hor = (uint8_t*)malloc(n0);
    uint8_t* sp = (uint8_t*)malloc(n0);
    uint8_t* cur = (uint8_t*)malloc(n0 * ch);
    uint8_t* nxt = (uint8_t*)malloc(n0 * ch);
    uint8_t* cm = mask ? (uint8_t*)malloc(n0) : NULL;
    uint8_t* nm = mask ? (uint8_t*)malloc(n0) : NULL;
    memcpy(cur, img, n0 * ch);
    if (mask) memcpy(cm, mask, n0);
    int r = rows, c = cols, rc = 0;
    for (int l = 0; l < L; ++l) {
        if (l > 0) {
            pyr_down(cur, r, c, ch, b.tmp, nxt);
            uint8_t* t = cur; cur = nxt; nxt = t;
            if (mask) {
                orc_mask_down(cm, r, c, nm);
                t = cm; cm = nm; nm = t;
            }
            r /= 2;
            c /= 2;
        }
        const size_t n = (size_t)r * c;
        if ((n % 16) != 0 || r % T[l] != 0 || c % T[l] != 0) {
            rc = ORC_EINVAL;
            break;
        }
        if (ch == 1) quantize_gray(cur, r, c, weak, &b, q);
        else orc_gradient_color(cur, r, c, weak, q, NULL, NULL);
        if (mask)
            for (size_t i = 0; i < n; ++i) q[i] = cm[i] ? q[i] : 0;
        spread_linearize(q, r, c, T[l], hor, sp, lm[l]);
        memset(lm[l] + 8 * n, 0, (size_t)orc_lm_guard_bytes(r, c, T[l]));
    }
    free(b.tmp); free(b.pad); free(b.bl); free(b.bpad); free(b.gx); free(b.gy); free(b.mag); free(b.qu); free(b.hrow);
    free(q); free(hor); free(sp); free(cur); free(nxt); free(cm); free(nm);
    return rc;
}

/* score = (raw * 100.f) / (4 * n) (line2Dup.cpp:1177, 1244): smallest raw whose score is > thr (coarse) or
 * not below thr (refine); the score is non-decreasing in raw, so these integer thresholds select the same
 * positions as the float comparisons. */
static int raw_min(int n, float thr, int strict) {
    int lo = 0, hi = 4 * n + 1;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        const float s = orc_score(mid, n);
        const int pass = strict ? (s > thr) : !(s < thr);
        if (pass) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

static inline const uint8_t* access_lm(const uint8_t* lm, int fx, int fy, int label, int T, int W, size_t hw) {
    const uint8_t* grid = lm + (size_t)label * T * T * hw;
    const int gi = (fy % T) * T + fx % T;
    return grid + (size_t)gi * hw + (size_t)(fy / T) * W + fx / T;  /* accessLinearMemory, line2Dup.cpp:755-778 */
}

typedef struct {
    uint16_t* sim;
    omvec cand;
} tbuf;

/* Detector::matchClass for one template pyramid                           line2Dup.cpp:1143-1266
 * similarity (line2Dup.cpp:780-831): dst[j] += lm_f[j] over j < P, the reads past a phase row included;
 * similarityLocal (SSE/AVX2 branch, line2Dup.cpp:877-892): 16 rows of 16 bytes per feature. */
static int match_template(uint8_t* const* lm, const int* lrows, const int* lcols, int L, const int* T,
                          float threshold, const int32_t* ti, const int32_t* feats, int cls, int tid, tbuf* tb,
                          omvec* out) {
    const int lowest = L - 1;
    const int32_t* tl = ti + 6 * lowest;
    const int nfeat = tl[4];
    if (nfeat >= 8192) return ORC_EINVAL;
    const int rows = lrows[lowest], cols = lcols[lowest], TL = T[lowest];
    const int W = cols / TL, H = rows / TL;
    const size_t hw = (size_t)W * H;
    const int wf = (tl[0] - 1) / TL + 1, hf = (tl[1] - 1) / TL + 1;
    const int P = (H - hf) * W + (W - wf) + 1;
    uint16_t* sim = tb->sim;
    const int pos = P > 0 ? (P < (int)hw ? P : (int)hw) : 0;
    memset(sim, 0, hw * 2);
    const int32_t* F = feats + 3 * (size_t)tl[5];
    for (int i = 0; i < nfeat; ++i) {
        const int fx = F[3 * i], fy = F[3 * i + 1], lab = F[3 * i + 2];
        if (fx < 0 || fx >= cols || fy < 0 || fy >= rows) continue;
        const uint8_t* p = access_lm(lm[lowest], fx, fy, lab, TL, W, hw);
        for (int j = 0; j < pos; ++j) sim[j] = (uint16_t)(sim[j] + p[j]);
    }
    const int rmin = nfeat > 0 ? raw_min(nfeat, threshold, 1) : 1 << 30;
    omvec* cand = &tb->cand;
    cand->n = 0;
    const int offset = TL / 2 + (TL % 2 - 1);
    for (int j = 0; j < (int)hw; ++j) {
        const int raw = j < pos ? sim[j] : 0;
        if (raw >= rmin && nfeat > 0) {
            const int r = j / W, c = j - r * W;
            omatch m = {c * TL + offset, r * TL + offset, cls, tid, orc_score(raw, nfeat)};
            orc_i_omvec_push(cand, m);
        }
    }
    uint16_t loc[256];
    for (int l = L - 2; l >= 0; --l) {
        const int32_t* tlv = ti + 6 * l;
        const int Tl = T[l], rl = lrows[l], cl = lcols[l];
        const int Wl = cl / Tl, Hl = rl / Tl;
        const size_t hwl = (size_t)Wl * Hl;
        const int border = 8 * Tl, off = Tl / 2 + (Tl % 2 - 1);
        const int max_x = cl - tlv[0] - border, max_y = rl - tlv[1] - border;
        const int nf = tlv[4];
        if (nf >= 8192) return ORC_EINVAL;
        const int32_t* Fl = feats + 3 * (size_t)tlv[5];
        for (int m = 0; m < cand->n; ++m) {
            omatch* mm = &cand->v[m];
            int x = mm->x * 2 + 1, y = mm->y * 2 + 1;
            x = x > border ? x : border;
            y = y > border ? y : border;
            x = x < max_x ? x : max_x;
            y = y < max_y ? y : max_y;
            const int offx = (x / Tl - 8) * Tl, offy = (y / Tl - 8) * Tl;
            memset(loc, 0, sizeof(loc));
            for (int i = 0; i < nf; ++i) {
                const int fx = Fl[3 * i] + offx, fy = Fl[3 * i + 1] + offy, lab = Fl[3 * i + 2];
                if (fx < 0 || fy < 0 || fx >= cl || fy >= rl) continue;
                const uint8_t* p = access_lm(lm[l], fx, fy, lab, Tl, Wl, hwl);
                for (int r = 0; r < 16; ++r) {
                    const uint8_t* pr = p + (size_t)r * Wl;
                    uint16_t* lr = loc + 16 * r;
                    for (int c = 0; c < 16; ++c) lr[c] = (uint16_t)(lr[c] + pr[c]);
                }
            }
            /* first strict maximum in row-major order, from best = 0 (line2Dup.cpp:1236-1253) */
            int best = 0, bi = -1;
            for (int k = 0; k < 256; ++k)
                if (loc[k] > best) {
                    best = loc[k];
                    bi = k;
                }
            const int br = bi < 0 ? -1 : bi / 16, bc = bi < 0 ? -1 : bi % 16;
            mm->sim = best > 0 ? orc_score(best, nf) : 0.f;
            mm->x = (x / Tl - 8 + bc) * Tl + off;
            mm->y = (y / Tl - 8 + br) * Tl + off;
        }
        int w = 0;
        for (int m = 0; m < cand->n; ++m)
            if (!(cand->v[m].sim < threshold)) cand->v[w++] = cand->v[m];  /* line2Dup.cpp:1261-1263 */
        cand->n = w;
    }
    for (int m = 0; m < cand->n; ++m) orc_i_omvec_push(out, cand->v[m]);
    return 0;
}

/* Detector::match (line2Dup.cpp:1049-1121): the same arguments and results as orc_match.  stage_seconds
 * receives the reference's two timers: "construct response map" and "templ match" (:1093, :1118). */
int FN(orc_fast_match)(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak, int L,
                       const int* T, float threshold, int num_templates, const int32_t* tinfo, const int32_t* feats,
                       const int32_t* tclass, const int32_t* tid, int max_out, int32_t* out_i32, float* out_sim,
                       int nthreads, double* stage_seconds) {
    if (L < 1 || L > 16 || (ch != 1 && ch != 3)) return ORC_EINVAL;
    uint8_t* lm[16];
    int lrows[16], lcols[16];
    int r = rows, c = cols;
    for (int l = 0; l < L; ++l) {
        if (l > 0) { r /= 2; c /= 2; }
        if (T[l] <= 0) return ORC_EINVAL;
        lrows[l] = r;
        lcols[l] = c;
        lm[l] = (uint8_t*)malloc((size_t)r * c * 8 + (size_t)orc_lm_guard_bytes(r, c, T[l]));
    }
    const double t0 = now_s();
    int rc = build_pyramid(img, rows, cols, ch, mask, weak, L, T, lm);
    const double t1 = now_s();
    omvec* per = NULL;
    if (rc == 0) {
        per = (omvec*)calloc((size_t)(num_templates > 0 ? num_templates : 1), sizeof(omvec));
        const int TL = T[L - 1];
        const size_t hw = (size_t)(lcols[L - 1] / TL) * (lrows[L - 1] / TL);
        int err = 0;
        if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : err)
        {
            tbuf tb;
            tb.sim = (uint16_t*)malloc(hw * 2 + 64);
            tb.cand.v = NULL;
            tb.cand.n = tb.cand.cap = 0;
#pragma omp for schedule(dynamic, 1)
            for (int t = 0; t < num_templates; ++t)
                if (match_template(lm, lrows, lcols, L, T, threshold, tinfo + (size_t)6 * L * t, feats, tclass[t],
                                   tid[t], &tb, &per[t]))
                    err |= 1;
            free(tb.sim);
            free(tb.cand.v);
        }
        if (err) rc = ORC_EINVAL;
    }
    if (rc == 0) {
        omvec all = {0, 0, 0};
        for (int t = 0; t < num_templates; ++t)
            for (int i = 0; i < per[t].n; ++i) orc_i_omvec_push(&all, per[t].v[i]);
        const int m = orc_i_sort_unique(all.v, all.n);  /* line2Dup.cpp:1113-1116 */
        if (m > max_out) {
            rc = -m;
        } else {
            for (int i = 0; i < m; ++i) {
                out_i32[4 * i] = all.v[i].x;
                out_i32[4 * i + 1] = all.v[i].y;
                out_i32[4 * i + 2] = all.v[i].cls;
                out_i32[4 * i + 3] = all.v[i].tid;
                out_sim[i] = all.v[i].sim;
            }
            rc = m;
        }
        free(all.v);
    }
    const double t2 = now_s();
    if (stage_seconds) {
        stage_seconds[0] = t1 - t0;
        stage_seconds[1] = t2 - t1;
    }
    if (per) {
        for (int t = 0; t < num_templates; ++t) free(per[t].v);
        free(per);
    }
    for (int l = 0; l < L; ++l) free(lm[l]);
    return rc;
}

/* The linear memories of the fast preprocessing, for tests against orc_build_lm_pyramid. */
int FN(orc_fast_build_lm_pyramid)(const uint8_t* img, int rows, int cols, int ch, const uint8_t* mask, float weak,
                                  int L, const int* T, uint8_t** lm_out) {
    if (L < 1 || L > 16 || (ch != 1 && ch != 3)) return ORC_EINVAL;
    return build_pyramid(img, rows, cols, ch, mask, weak, L, T, lm_out);
}

/* The (dx, dy) -> 8-bin table of phase_bins for dx, dy in [lo, hi] (the layout of orc_bin_table), for tests. */
void FN(orc_fast_bin_table)(int lo, int hi, uint8_t* out) {
    const int n = hi - lo + 1;
    int32_t* gx = (int32_t*)malloc((size_t)n * 4);
    int32_t* gy = (int32_t*)malloc((size_t)n * 4);
    for (int i = 0; i < n; ++i) gx[i] = lo + i;
    for (int dy = lo; dy <= hi; ++dy) {
        for (int i = 0; i < n; ++i) gy[i] = dy;
        uint8_t* row = out + (size_t)(dy - lo) * n;
        phase_bins(gx, gy, n, NULL, row);
        for (int i = 0; i < n; ++i) row[i] &= 7;
    }
    free(gx);
    free(gy);
}
